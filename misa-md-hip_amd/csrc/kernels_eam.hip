// gfx950 kernels of the EAM hot path: rho + fused dF/drho, standalone dF/drho
// and the pair + embedding force, in FP64 (reference arithmetic) and mixed
// precision (r^2-keyed splines, FP64 leading term + FP32 correction).
//
// Reference semantics (misa-md-hip): one result per local lattice site with
// type >= 0 (src/kernels/kernel_itl.hpp:69-105), neighbours from the parity
// stencil (:82-86), neighbours with type < 0 skipped (:48-57), pair included
// only for dist2 < rc^2 (src/kernels/kernel_pairs.hpp:21-32, :39-60), rho by
// the neighbour's type, del = x_i - x_j, results overwritten.
#include "device_common.h"
#include "kernels.h"

namespace mmd {

#ifndef MMD_PAIR_UNROLL
#define MMD_PAIR_UNROLL 4
#endif
constexpr int kUnroll = MMD_PAIR_UNROLL;  // stencil entries whose loads are in flight together

// Loads through a uniform base pointer with a 32-bit element index, so the
// address is an SGPR base + 32-bit VGPR offset (no 64-bit VALU arithmetic).
template <typename T>
__device__ __forceinline__ T ld(const T *__restrict__ base, uint32_t idx) {
  return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + idx * (uint32_t)sizeof(T));
}

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

// record = knot index (+ function offset); 8 doubles per record
__device__ __forceinline__ double val8(const double *__restrict__ tab, uint32_t rec, double p) {
  const double4 c = ld(reinterpret_cast<const double4 *>(tab), 2 * rec);
  return ((c.x * p + c.y) * p + c.z) * p + c.w;
}

__device__ __forceinline__ double der8(const double *__restrict__ tab, uint32_t rec, double p) {
  const double4 c = ld(reinterpret_cast<const double4 *>(tab), 2 * rec + 1);
  return (c.x * p + c.y) * p + c.z;
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

__device__ __forceinline__ double valS(const double4 *__restrict__ tab, uint32_t rec, double q) {
  const double4 c = ld(tab, rec);
  const float2 a23 = __builtin_bit_cast(float2, c.z);
  const float cc = fmaf((float)q, a23.y, a23.x);  // FP32 quadratic/cubic correction
  return fma(q, fma(q, (double)cc, c.y), c.x);
}

__device__ __forceinline__ double derS(const double4 *__restrict__ tab, uint32_t rec, double q) {
  const double4 c = ld(tab, rec);
  const float2 b2 = __builtin_bit_cast(float2, c.z);
  return fma(q, fma(q, (double)b2.x, c.y), c.x);
}

// ------------------------------------------------------------- pair terms

// rho_{tj}(r) for r^2 = d2
template <int PREC>
__device__ __forceinline__ double rho_pair(const TabF64 &tf, const TabMixed &tm, int tj, double d2) {
  if (PREC == 1 && d2 < tm.s_tail && d2 >= tm.s_head) {
    const SegS sg = seg_s(d2, tm);
    return valS(tm.rho_v, (uint32_t)(tj * tm.n_s + sg.m), sg.q);
  }
  const double r = d2 * rsqrt64(d2);
  const Seg sg = seg_f64(r, tf.r0, tf.inv_dr, tf.n_r);
  return val8(tf.rho, (uint32_t)(tj * tf.n_r + sg.m), sg.p);
}

// fpair = -(df_i rho'_{tj} + df_j rho'_{ti} + phi'_{ti tj}) / r (the scalar that
// multiplies del = x_i - x_j); phi receives phi_{ti tj}(r) when WANT_E.
template <int PREC, bool WANT_E>
__device__ __forceinline__ double fpair(const TabF64 &tf, const TabMixed &tm, int ti, int tj, double d2, double dfi,
                                        double dfj, double &phi) {
  if (PREC == 1 && d2 < tm.s_tail && d2 >= tm.s_head) {
    // r^2-keyed: d/ds tables, -(dpsi/dr)/r = -2 dpsi/ds
    const SegS sg = seg_s(d2, tm);
    const uint32_t rj = (uint32_t)(tj * tm.n_s + sg.m);
    const uint32_t rp = (uint32_t)((ti * tm.nt + tj) * tm.n_s + sg.m);
    const double rhojp = derS(tm.rho_d, rj, sg.q);
    double rhoip = rhojp;
    if (ti != tj) rhoip = derS(tm.rho_d, (uint32_t)(ti * tm.n_s + sg.m), sg.q);
    const double phip = derS(tm.phi_d, rp, sg.q);
    if (WANT_E) phi = valS(tm.phi_v, rp, sg.q);
    return -2.0 * (dfi * rhojp + dfj * rhoip + phip);
  }
  const double rinv = rsqrt64(d2);
  const double r = d2 * rinv;
  const Seg sg = seg_f64(r, tf.r0, tf.inv_dr, tf.n_r);
  const uint32_t rj = (uint32_t)(tj * tf.n_r + sg.m);
  const uint32_t rp = (uint32_t)((ti * tf.nt + tj) * tf.n_r + sg.m);
  const double rhojp = der8(tf.rho, rj, sg.p);
  double rhoip = rhojp;
  if (ti != tj) rhoip = der8(tf.rho, (uint32_t)(ti * tf.n_r + sg.m), sg.p);
  const double phip = der8(tf.phi, rp, sg.p);
  if (WANT_E) phi = val8(tf.phi, rp, sg.p);
  return -(dfi * rhojp + dfj * rhoip + phip) * rinv;
}

// --------------------------------------------------------------- kernels
//
// One thread per local site, kUnroll stencil entries per trip: all neighbour
// loads of a trip are issued before any is used, pairs are evaluated without
// per-pair branches (out-of-cutoff and vacant neighbours are masked), and a
// trip whose pairs are all outside the cutoff for the whole wave is skipped.
// The stencil rows are padded with zero offsets past nk.

template <int PREC, bool WANT_E>
__global__ __launch_bounds__(kBlock) void k_rho_df(Geom g, const int32_t *__restrict__ off, int32_t nk_even,
                                                     int32_t nk_odd, const double *__restrict__ X,
                                                     const double *__restrict__ Y, const double *__restrict__ Z,
                                                     const int8_t *__restrict__ T, double *__restrict__ RHO,
                                                     double *__restrict__ DF, TabF64 tf, TabMixed tm, double rc2,
                                                     double *__restrict__ e_part) {
  __shared__ double lds[kWavesPerBlock];
  double e_loc = 0.0;
  Site st;
  if (map_site(g, st) && st.active) {
    const uint32_t i = (uint32_t)st.i;
    const int t0 = ld(T, i);
    if (t0 >= 0) {
      const int32_t *o = off + st.s * kMaxStencil;
      const int nk = st.s ? nk_odd : nk_even;
      const double xi = ld(X, i), yi = ld(Y, i), zi = ld(Z, i);
      const double d2_safe = 0.25 * rc2;
      double acc = 0.0;
      for (int k = 0; k < nk; k += kUnroll) {
        int tj[kUnroll];
        double d2[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; u++) {
          const uint32_t j = i + (uint32_t)o[k + u];
          tj[u] = ld(T, j);
          const double dx = xi - ld(X, j), dy = yi - ld(Y, j), dz = zi - ld(Z, j);
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
      DF[i] = der8(tf.emb, re, se.p);
      if (WANT_E) e_loc = val8(tf.emb, re, se.p);
    }
  }
  if (WANT_E) {
    const double b = block_sum(e_loc, lds);
    if (threadIdx.x == 0) e_part[xcd_block()] = b;
  }
}

// standalone df (src/kernels/hip_kernels.hip.cpp:21-40)
__global__ __launch_bounds__(kBlock) void k_df(Geom g, const int8_t *__restrict__ T, const double *__restrict__ RHO,
                                                 double *__restrict__ DF, TabF64 tf) {
  Site st;
  if (!map_site(g, st) || !st.active) return;
  const uint32_t i = (uint32_t)st.i;
  const int t0 = ld(T, i);
  if (t0 < 0) return;
  const Seg se = seg_f64(ld(RHO, i), tf.rho0, tf.inv_drho, tf.n_rho);
  DF[i] = der8(tf.emb, (uint32_t)(t0 * tf.n_rho + se.m), se.p);
}

template <int PREC, bool WANT_E>
__global__ __launch_bounds__(kBlock) void k_force(Geom g, const int32_t *__restrict__ off, int32_t nk_even,
                                                    int32_t nk_odd, const double *__restrict__ X,
                                                    const double *__restrict__ Y, const double *__restrict__ Z,
                                                    const int8_t *__restrict__ T, const double *__restrict__ DF,
                                                    double *__restrict__ FX, double *__restrict__ FY,
                                                    double *__restrict__ FZ, TabF64 tf, TabMixed tm, double rc2,
                                                    double *__restrict__ e_part) {
  __shared__ double lds[kWavesPerBlock];
  double e_loc = 0.0;
  Site st;
  if (map_site(g, st) && st.active) {
    const uint32_t i = (uint32_t)st.i;
    const int t0 = ld(T, i);
    if (t0 >= 0) {
      const int32_t *o = off + st.s * kMaxStencil;
      const int nk = st.s ? nk_odd : nk_even;
      const double xi = ld(X, i), yi = ld(Y, i), zi = ld(Z, i), dfi = ld(DF, i);
      const double d2_safe = 0.25 * rc2;
      double ax = 0.0, ay = 0.0, az = 0.0, ep = 0.0;
      for (int k = 0; k < nk; k += kUnroll) {
        int tj[kUnroll];
        double dx[kUnroll], dy[kUnroll], dz[kUnroll], d2[kUnroll], dfj[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; u++) {
          const uint32_t j = i + (uint32_t)o[k + u];
          tj[u] = ld(T, j);
          dx[u] = xi - ld(X, j);
          dy[u] = yi - ld(Y, j);
          dz[u] = zi - ld(Z, j);
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
    if (threadIdx.x == 0) e_part[xcd_block()] = b;
  }
}

// ----------------------------------------------------------------- launch

int64_t site_blocks(const Geom &g) {
  int64_t nb = (g.n_waves + kWavesPerBlock - 1) / kWavesPerBlock;
  return (nb + kXcds - 1) / kXcds * kXcds;
}

hipError_t launch_rho_df(hipStream_t st, const Geom &g, const int32_t *off, int32_t nk_even, int32_t nk_odd, const double *X,
                         const double *Y, const double *Z, const int8_t *T, double *RHO, double *DF,
                         const TabF64 &tf, const TabMixed &tm, double rc2, int prec, double *e_part) {
  const dim3 grid((unsigned)site_blocks(g)), block(kBlock);
  if (prec == 0) {
    if (e_part)
      k_rho_df<0, true><<<grid, block, 0, st>>>(g, off, nk_even, nk_odd, X, Y, Z, T, RHO, DF, tf, tm, rc2, e_part);
    else
      k_rho_df<0, false><<<grid, block, 0, st>>>(g, off, nk_even, nk_odd, X, Y, Z, T, RHO, DF, tf, tm, rc2, e_part);
  } else {
    if (e_part)
      k_rho_df<1, true><<<grid, block, 0, st>>>(g, off, nk_even, nk_odd, X, Y, Z, T, RHO, DF, tf, tm, rc2, e_part);
    else
      k_rho_df<1, false><<<grid, block, 0, st>>>(g, off, nk_even, nk_odd, X, Y, Z, T, RHO, DF, tf, tm, rc2, e_part);
  }
  return hipGetLastError();
}

hipError_t launch_df(hipStream_t st, const Geom &g, const int8_t *T, const double *RHO, double *DF,
                     const TabF64 &tf) {
  k_df<<<dim3((unsigned)site_blocks(g)), dim3(kBlock), 0, st>>>(g, T, RHO, DF, tf);
  return hipGetLastError();
}

hipError_t launch_force(hipStream_t st, const Geom &g, const int32_t *off, int32_t nk_even, int32_t nk_odd, const double *X,
                        const double *Y, const double *Z, const int8_t *T, const double *DF, double *FX, double *FY,
                        double *FZ, const TabF64 &tf, const TabMixed &tm, double rc2, int prec, double *e_part) {
  const dim3 grid((unsigned)site_blocks(g)), block(kBlock);
  if (prec == 0) {
    if (e_part)
      k_force<0, true><<<grid, block, 0, st>>>(g, off, nk_even, nk_odd, X, Y, Z, T, DF, FX, FY, FZ, tf, tm, rc2, e_part);
    else
      k_force<0, false><<<grid, block, 0, st>>>(g, off, nk_even, nk_odd, X, Y, Z, T, DF, FX, FY, FZ, tf, tm, rc2, e_part);
  } else {
    if (e_part)
      k_force<1, true><<<grid, block, 0, st>>>(g, off, nk_even, nk_odd, X, Y, Z, T, DF, FX, FY, FZ, tf, tm, rc2, e_part);
    else
      k_force<1, false><<<grid, block, 0, st>>>(g, off, nk_even, nk_odd, X, Y, Z, T, DF, FX, FY, FZ, tf, tm, rc2, e_part);
  }
  return hipGetLastError();
}

}  // namespace mmd
