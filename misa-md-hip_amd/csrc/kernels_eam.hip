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

__device__ __forceinline__ double val8(const double *__restrict__ tab, Seg s) {
  const double4 c = *reinterpret_cast<const double4 *>(tab + (size_t)s.m * 8);
  return ((c.x * s.p + c.y) * s.p + c.z) * s.p + c.w;
}

__device__ __forceinline__ double der8(const double *__restrict__ tab, Seg s) {
  const double4 c = *reinterpret_cast<const double4 *>(tab + (size_t)s.m * 8 + 4);
  return (c.x * s.p + c.y) * s.p + c.z;
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

__device__ __forceinline__ double valS(const double4 *__restrict__ tab, SegS s) {
  const double4 rec = tab[s.m];
  const float2 a23 = __builtin_bit_cast(float2, rec.z);
  const float c = fmaf((float)s.q, a23.y, a23.x);  // FP32 quadratic/cubic correction
  return fma(s.q, fma(s.q, (double)c, rec.y), rec.x);
}

__device__ __forceinline__ double derS(const double4 *__restrict__ tab, SegS s) {
  const double4 rec = tab[s.m];
  const float2 b2 = __builtin_bit_cast(float2, rec.z);
  return fma(s.q, fma(s.q, (double)b2.x, rec.y), rec.x);
}

// ------------------------------------------------------------------ rho

template <int PREC>
__device__ __forceinline__ double rho_pair(const TabF64 &tf, const TabMixed &tm, int tj, double d2) {
  if (PREC == 0) {
    const Seg sg = seg_f64(sqrt(d2), tf.r0, tf.inv_dr, tf.n_r);
    return val8(tf.rho + (size_t)tj * tf.n_r * 8, sg);
  } else {
    if (d2 >= tm.s_tail || d2 < tm.s_head) return rho_pair<0>(tf, tm, tj, d2);
    const SegS sg = seg_s(d2, tm);
    return valS(tm.rho_v + (size_t)tj * tm.n_s, sg);
  }
}

template <int PREC, bool WANT_E>
__global__ __launch_bounds__(kBlock) void k_rho_df(Geom g, const int32_t *__restrict__ off, int32_t nk_even, int32_t nk_odd,
                                                     const double *__restrict__ X, const double *__restrict__ Y,
                                                     const double *__restrict__ Z, const int8_t *__restrict__ T,
                                                     double *__restrict__ RHO, double *__restrict__ DF, TabF64 tf,
                                                     TabMixed tm, double rc2, double *__restrict__ e_part) {
  __shared__ double lds[kWavesPerBlock];
  double e_loc = 0.0;
  Site st;
  if (map_site(g, st) && st.active) {
    const int64_t i = st.i;
    const int t0 = T[i];
    if (t0 >= 0) {
      const int32_t *o = off + st.s * kMaxStencil;
      const int nk = st.s ? nk_odd : nk_even;
      const double xi = X[i], yi = Y[i], zi = Z[i];
      double acc = 0.0;
      for (int k = 0; k < nk; k++) {
        const int64_t j = i + o[k];
        const int tj = T[j];
        const double dx = xi - X[j], dy = yi - Y[j], dz = zi - Z[j];
        const double d2 = dx * dx + dy * dy + dz * dz;
        if (tj >= 0 && d2 < rc2) acc += rho_pair<PREC>(tf, tm, tj, d2);
      }
      RHO[i] = acc;
      const Seg se = seg_f64(acc, tf.rho0, tf.inv_drho, tf.n_rho);
      const double *et = tf.emb + (size_t)t0 * tf.n_rho * 8;
      DF[i] = der8(et, se);
      if (WANT_E) e_loc = val8(et, se);
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
  const int t0 = T[st.i];
  if (t0 < 0) return;
  const Seg se = seg_f64(RHO[st.i], tf.rho0, tf.inv_drho, tf.n_rho);
  DF[st.i] = der8(tf.emb + (size_t)t0 * tf.n_rho * 8, se);
}

// ------------------------------------------------------------------ force

// fpair of one pair (centre type ti, neighbour tj); phi receives the pair
// energy when WANT_E.
template <int PREC, bool WANT_E>
__device__ __forceinline__ double fpair(const TabF64 &tf, const TabMixed &tm, int ti, int tj, double d2, double dfi,
                                        double dfj, double &phi) {
  if (PREC == 0) {
    const double r = sqrt(d2);
    const Seg sg = seg_f64(r, tf.r0, tf.inv_dr, tf.n_r);
    const size_t stride = (size_t)tf.n_r * 8;
    const double *pt = tf.phi + (size_t)(ti * tf.nt + tj) * stride;
    const double rhojp = der8(tf.rho + (size_t)tj * stride, sg);
    const double rhoip = der8(tf.rho + (size_t)ti * stride, sg);
    const double phip = der8(pt, sg);
    if (WANT_E) phi = val8(pt, sg);
    const double psip = dfi * rhojp + dfj * rhoip + phip;
    return -psip / r;
  } else {
    if (d2 >= tm.s_tail || d2 < tm.s_head) return fpair<0, WANT_E>(tf, tm, ti, tj, d2, dfi, dfj, phi);
    const SegS sg = seg_s(d2, tm);
    const size_t pr = (size_t)(ti * tm.nt + tj) * tm.n_s;
    const double rhojp = derS(tm.rho_d + (size_t)tj * tm.n_s, sg);
    const double rhoip = derS(tm.rho_d + (size_t)ti * tm.n_s, sg);
    const double phip = derS(tm.phi_d + pr, sg);
    if (WANT_E) phi = valS(tm.phi_v + pr, sg);
    // derivatives are d/ds with s = r^2:  -(dpsi/dr)/r = -2 dpsi/ds
    return -2.0 * (dfi * rhojp + dfj * rhoip + phip);
  }
}

template <int PREC, bool WANT_E>
__global__ __launch_bounds__(kBlock) void k_force(Geom g, const int32_t *__restrict__ off, int32_t nk_even, int32_t nk_odd,
                                                    const double *__restrict__ X, const double *__restrict__ Y,
                                                    const double *__restrict__ Z, const int8_t *__restrict__ T,
                                                    const double *__restrict__ DF, double *__restrict__ FX,
                                                    double *__restrict__ FY, double *__restrict__ FZ, TabF64 tf,
                                                    TabMixed tm, double rc2, double *__restrict__ e_part) {
  __shared__ double lds[kWavesPerBlock];
  double e_loc = 0.0;
  Site st;
  if (map_site(g, st) && st.active) {
    const int64_t i = st.i;
    const int t0 = T[i];
    if (t0 >= 0) {
      const int32_t *o = off + st.s * kMaxStencil;
      const int nk = st.s ? nk_odd : nk_even;
      const double xi = X[i], yi = Y[i], zi = Z[i], dfi = DF[i];
      double ax = 0.0, ay = 0.0, az = 0.0, ep = 0.0;
      for (int k = 0; k < nk; k++) {
        const int64_t j = i + o[k];
        const int tj = T[j];
        const double dx = xi - X[j], dy = yi - Y[j], dz = zi - Z[j];
        const double d2 = dx * dx + dy * dy + dz * dz;
        if (tj >= 0 && d2 < rc2) {
          double phi = 0.0;
          const double fp = fpair<PREC, WANT_E>(tf, tm, t0, tj, d2, dfi, DF[j], phi);
          ax += dx * fp;
          ay += dy * fp;
          az += dz * fp;
          if (WANT_E) ep += phi;
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
