// Device code of the EAM passes shared by the single-domain kernels
// (kernels_eam.hip) and the batched serving kernels (kernels_batch.hip): pair
// terms, spline evaluation and the per-block bodies of rho + df and force.
//
// Reference semantics (misa-md-hip): one result per local lattice site with
// type >= 0 (src/kernels/kernel_itl.hpp:69-105), neighbours from the parity
// stencil (:82-86), neighbours with type < 0 skipped (:48-57), pair included
// only for dist2 < rc^2 (src/kernels/kernel_pairs.hpp:21-32, :39-60), rho by
// the neighbour's type, del = x_i - x_j, results overwritten.
#ifndef MMD_EAM_BODY_H
#define MMD_EAM_BODY_H

#include "device_common.h"

namespace mmd {


#ifndef MMD_PAIR_UNROLL
#define MMD_PAIR_UNROLL 4
#endif
constexpr int kUnroll = MMD_PAIR_UNROLL;  // stencil entries whose loads are in flight together

// 1/sqrt(d2) for d2 > 0: hardware reciprocal-square-root seed refined by two
// Newton steps to full FP64 accuracy (a few ulp), instead of the IEEE sqrt and
// divide sequences.
__device__ __forceinline__ double rsqrt64(double d2) {
  double y = __builtin_amdgcn_rsq(d2);
  double t = fma(-d2 * y, y, 1.0);
  y = fma(0.5 * y, t, y);
  t = fma(-d2 * y, y, 1.0);
  return fma(0.5 * y, t, y);
}

// ------------------------------------------------------------ spline eval

struct Seg {
  int m;
  double p;
};

__device__ __forceinline__ Seg seg_f64(double x, double x0, double inv_dx, int n) {
  double p = (x - x0) * inv_dx;
  p = fmax(p, 0.0);
  int m = (int)p;
  m = min(m, n - 2);
  Seg s;
  s.p = fmin(p - (double)m, 1.0);
  s.m = m;
  return s;
}

struct SegS {
  int m;
  double q;
};

__device__ __forceinline__ SegS seg_s(double s2, const TabMixed &t) {
  double p = fma(s2, t.inv_ds, -t.s0 * t.inv_ds);
  p = fmax(p, 0.0);
  int m = min((int)p, t.n_s - 2);
  SegS r;
  r.m = m;
  r.q = fmin(p - (double)m, 1.0);
  return r;
}

// ------------------------------------------------------------- pair terms

// rho_{tj}(r) for r^2 = d2
template <int PREC>
__device__ __forceinline__ double rho_pair(const TabF64 &tf, const TabMixed &tm, int tj, double d2) {
  if (PREC == 1 && d2 < tm.s_tail && d2 >= tm.s_head) {
    const SegS sg = seg_s(d2, tm);
    return mix_value(tm.rho, (uint32_t)(tj * tm.n_s + sg.m), sg.q);
  }
  const double r = d2 * rsqrt64(d2);
  const Seg sg = seg_f64(r, tf.r0, tf.inv_dr, tf.n_r);
  return spl_value(tf.rho, (uint32_t)(tj * tf.n_r + sg.m), sg.p);
}

// fpair = -(df_i rho'_{tj} + df_j rho'_{ti} + phi'_{ti tj}) / r (the scalar that
// multiplies del = x_i - x_j); phi receives phi_{ti tj}(r) when WANT_E.
// rho'_{tj} and phi'_{ti tj} come from one combined per-type-pair record;
// rho'_{ti} is gathered separately, by unlike-type pairs only.
template <int PREC, bool WANT_E>
__device__ __forceinline__ double fpair(const TabF64 &tf, const TabMixed &tm, int ti, int tj, double d2, double dfi,
                                        double dfj, double &phi) {
  const int tp = ti * tf.nt + tj;
  if (PREC == 1 && d2 < tm.s_tail && d2 >= tm.s_head) {
    // r^2-keyed: d/ds tables, -(dpsi/dr)/r = -2 dpsi/ds
    const SegS sg = seg_s(d2, tm);
    const uint32_t rp = (uint32_t)(tp * tm.n_s + sg.m);
    const double2 b0 = ld(tm.fb, 2 * rp), b1 = ld(tm.fb, 2 * rp + 1);
    const float2 q2 = ld(tm.fq, rp);
    const double q = sg.q;
    const double rhojp = fma(q, fma(q, (double)q2.x, b0.y), b0.x);
    const double phip = fma(q, fma(q, (double)q2.y, b1.y), b1.x);
    double rhoip = rhojp;
    if (ti != tj) rhoip = mix_deriv(tm.rho, (uint32_t)(ti * tm.n_s + sg.m), q);
    if (WANT_E) phi = mix_value(tm.phi, rp, q);
    return -2.0 * (dfi * rhojp + dfj * rhoip + phip);
  }
  const double rinv = rsqrt64(d2);
  const double r = d2 * rinv;
  const Seg sg = seg_f64(r, tf.r0, tf.inv_dr, tf.n_r);
  const uint32_t rp = (uint32_t)(tp * tf.n_r + sg.m);
  const double2 c0 = ld(tf.fc, 3 * rp), c1 = ld(tf.fc, 3 * rp + 1), c2 = ld(tf.fc, 3 * rp + 2);
  const double p = sg.p;
  const double rhojp = (c0.x * p + c0.y) * p + c1.x;
  const double phip = (c1.y * p + c2.x) * p + c2.y;
  double rhoip = rhojp;
  if (ti != tj) rhoip = spl_deriv(tf.rho, (uint32_t)(ti * tf.n_r + sg.m), p);
  if (WANT_E) phi = spl_value(tf.phi, rp, p);
  return -(dfi * rhojp + dfj * rhoip + phip) * rinv;
}

// --------------------------------------------------------------- kernels
//
// One thread per local site, kUnroll stencil entries per trip: all neighbour
// loads of a trip are issued before any is used, pairs are evaluated without
// per-pair branches (out-of-cutoff and vacant neighbours are masked), and a
// trip whose pairs are all outside the cutoff for the whole wave is skipped.
// The stencil rows are padded with zero offsets past nk.  A neighbour is one
// 32-byte (x, y, z, type) record: two 16-byte loads.

template <int PREC, bool WANT_E>
__device__ __forceinline__ void rho_df_block(const Geom &g, int64_t lb, const int32_t *__restrict__ off,
                                             int32_t nk_even, int32_t nk_odd, const double4 *__restrict__ P,
                                             double *__restrict__ RHO, double *__restrict__ DF, const TabF64 &tf,
                                             const TabMixed &tm, double rc2, double *__restrict__ e_part,
                                             double *lds) {
  double e_loc = 0.0;
  Site st;
  if (map_site_lb(g, lb, st) && st.active) {
    const uint32_t i = (uint32_t)st.i;
    const double4 pi = ld(P, i);
    const int t0 = (int)pi.w;
    if (t0 >= 0) {
      const int32_t *o = off + st.s * kMaxStencil;
      const int nk = st.s ? nk_odd : nk_even;
      const double d2_safe = 0.25 * rc2;
      double acc = 0.0;
      for (int k = 0; k < nk; k += kUnroll) {
        int tj[kUnroll];
        double d2[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; u++) {
          const double4 pj = ld(P, i + (uint32_t)o[k + u]);
          tj[u] = (int)pj.w;
          const double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
          d2[u] = dx * dx + dy * dy + dz * dz;
        }
        bool any = false;
#pragma unroll
        for (int u = 0; u < kUnroll; u++) {
          const bool in = (k + u < nk) && tj[u] >= 0 && d2[u] < rc2;
          if (!in) d2[u] = -1.0;
          any |= in;
        }
        if (__ballot(any) == 0) continue;
#pragma unroll
        for (int u = 0; u < kUnroll; u++) {
          const bool in = d2[u] >= 0.0;
          const double v = rho_pair<PREC>(tf, tm, in ? tj[u] : 0, in ? d2[u] : d2_safe);
          acc += in ? v : 0.0;
        }
      }
      RHO[i] = acc;
      const Seg se = seg_f64(acc, tf.rho0, tf.inv_drho, tf.n_rho);
      const uint32_t re = (uint32_t)(t0 * tf.n_rho + se.m);
      DF[i] = spl_deriv(tf.emb, re, se.p);
      if (WANT_E) e_loc = spl_value(tf.emb, re, se.p);
    }
  }
  if (WANT_E) {
    const double b = block_sum(e_loc, lds);
    if (threadIdx.x == 0) e_part[lb] = b;
  }
}

template <int PREC, bool WANT_E>
__device__ __forceinline__ void force_block(const Geom &g, int64_t lb, const int32_t *__restrict__ off,
                                            int32_t nk_even, int32_t nk_odd, const double4 *__restrict__ P,
                                            const double *__restrict__ DF, double *__restrict__ FX,
                                            double *__restrict__ FY, double *__restrict__ FZ, const TabF64 &tf,
                                            const TabMixed &tm, double rc2, double *__restrict__ e_part,
                                            double *lds) {
  double e_loc = 0.0;
  Site st;
  if (map_site_lb(g, lb, st) && st.active) {
    const uint32_t i = (uint32_t)st.i;
    const double4 pi = ld(P, i);
    const int t0 = (int)pi.w;
    if (t0 >= 0) {
      const int32_t *o = off + st.s * kMaxStencil;
      const int nk = st.s ? nk_odd : nk_even;
      const double dfi = ld(DF, i);
      const double d2_safe = 0.25 * rc2;
      double ax = 0.0, ay = 0.0, az = 0.0, ep = 0.0;
      for (int k = 0; k < nk; k += kUnroll) {
        int tj[kUnroll];
        double dx[kUnroll], dy[kUnroll], dz[kUnroll], d2[kUnroll], dfj[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; u++) {
          const uint32_t j = i + (uint32_t)o[k + u];
          const double4 pj = ld(P, j);
          tj[u] = (int)pj.w;
          dx[u] = pi.x - pj.x;
          dy[u] = pi.y - pj.y;
          dz[u] = pi.z - pj.z;
          dfj[u] = ld(DF, j);
          d2[u] = dx[u] * dx[u] + dy[u] * dy[u] + dz[u] * dz[u];
        }
        bool any = false;
#pragma unroll
        for (int u = 0; u < kUnroll; u++) {
          const bool in = (k + u < nk) && tj[u] >= 0 && d2[u] < rc2;
          if (!in) d2[u] = -1.0;
          any |= in;
        }
        if (__ballot(any) == 0) continue;
#pragma unroll
        for (int u = 0; u < kUnroll; u++) {
          const bool in = d2[u] >= 0.0;
          double phi = 0.0;
          double fp = fpair<PREC, WANT_E>(tf, tm, t0, in ? tj[u] : t0, in ? d2[u] : d2_safe, dfi, dfj[u], phi);
          fp = in ? fp : 0.0;
          ax = fma(dx[u], fp, ax);
          ay = fma(dy[u], fp, ay);
          az = fma(dz[u], fp, az);
          if (WANT_E) ep += in ? phi : 0.0;
        }
      }
      FX[i] = ax;
      FY[i] = ay;
      FZ[i] = az;
      if (WANT_E) e_loc = 0.5 * ep;
    }
  }
  if (WANT_E) {
    const double b = block_sum(e_loc, lds);
    if (threadIdx.x == 0) e_part[lb] = b;
  }
}

}  // namespace mmd

#endif
