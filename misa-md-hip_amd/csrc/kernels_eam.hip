// gfx950 kernels of the EAM hot path (single sub-domain): rho + fused dF/drho,
// standalone dF/drho and the pair + embedding force, in FP64 (reference
// arithmetic) and mixed precision (r^2-keyed splines, FP64 leading terms +
// FP32 corrections).  The per-block bodies live in eam_body.h.
#include "eam_body.h"
#include "kernels.h"

namespace mmd {

template <int PREC, bool WANT_E, bool LEAN>
__global__ __launch_bounds__(kBlock) void k_rho_df(Geom g, Halo h, const int32_t *__restrict__ off, int32_t nk_even,
                                                     int32_t nk_odd, const double4 *__restrict__ P,
                                                     double *__restrict__ RHO, double *__restrict__ DF, TabF64 tf,
                                                     TabMixed tm, double rc2, double *__restrict__ e_part,
                                                     int64_t lb0) {
  extern __shared__ double2 lds_tile[];
  __shared__ double red[kWavesPerBlock];
  rho_df_block<PREC, WANT_E, LEAN>(g, h, lb0 + xcd_block(), off, nk_even, nk_odd, P, RHO, DF, tf, tm, rc2, e_part,
                                   lds_tile, red);
}

template <int PREC, bool WANT_E>
__global__ __launch_bounds__(kBlock, kForceWavesPerSimd) void k_force(Geom g, Halo h, const int32_t *__restrict__ off, int32_t nk_even,
                                                    int32_t nk_odd, const double4 *__restrict__ P,
                                                    const double *__restrict__ DF, double *__restrict__ FX,
                                                    double *__restrict__ FY, double *__restrict__ FZ, TabF64 tf,
                                                    TabMixed tm, double rc2, double *__restrict__ e_part,
                                                    int64_t lb0) {
  extern __shared__ double2 lds_tile[];
  __shared__ double red[kWavesPerBlock];
  force_block<PREC, WANT_E>(g, h, lb0 + xcd_block(), off, nk_even, nk_odd, P, DF, FX, FY, FZ, tf, tm, rc2, e_part,
                            lds_tile, red);
}

// standalone df (src/kernels/hip_kernels.hip.cpp:21-40)
__global__ __launch_bounds__(kBlock) void k_df(Geom g, const int8_t *__restrict__ T, const double *__restrict__ RHO,
                                                 double *__restrict__ DF, TabF64 tf) {
  Site st;
  if (!map_site(g, st) || !st.active) return;
  const uint32_t i = (uint32_t)st.i;
  const int t0 = T[i];
  if (t0 < 0) return;
  const Seg se = seg_f64(RHO[i], tf.rho0, tf.inv_drho, tf.n_rho);
  DF[i] = spl_deriv(tf.emb, (uint32_t)(t0 * tf.n_rho + se.m), se.p);
}

// ----------------------------------------------------------------- launch

int64_t site_blocks(const Geom &g) {
  int64_t nb = (g.n_local + kBlock - 1) / kBlock;
  return (nb + kXcds - 1) / kXcds * kXcds;
}

int64_t tile_blocks(const Geom &g) { return (g.n_tiles + kXcds - 1) / kXcds * kXcds; }

hipError_t launch_rho_df(hipStream_t st, const Geom &g0, const Halo &h, const int32_t *off, int32_t nk_even,
                         int32_t nk_odd, const double4 *P, double *RHO, double *DF, const TabF64 &tf,
                         const TabMixed &tm, double rc2, int prec, double *e_part, bool lean, int64_t lb0,
                         int64_t lb1) {
  // tiles [lb0, lb1): the bodies take g.n_tiles as the end of the launch's range
  Geom g = g0;
  if (lb1 < 0 || lb1 > g0.n_tiles) lb1 = g0.n_tiles;
  if (lb0 >= lb1) return hipSuccess;
  g.n_tiles = lb1;
  const dim3 grid((unsigned)((lb1 - lb0 + kXcds - 1) / kXcds * kXcds)), block(kBlock);
  const size_t sh = tile_lds_bytes(h, false);
  if (prec == 0) {
    if (e_part)
      k_rho_df<0, true, false><<<grid, block, sh, st>>>(g, h, off, nk_even, nk_odd, P, RHO, DF, tf, tm, rc2, e_part, lb0);
    else
      k_rho_df<0, false, false><<<grid, block, sh, st>>>(g, h, off, nk_even, nk_odd, P, RHO, DF, tf, tm, rc2, e_part, lb0);
  } else {
    if (e_part && lean)
      k_rho_df<1, true, true><<<grid, block, sh, st>>>(g, h, off, nk_even, nk_odd, P, RHO, DF, tf, tm, rc2, e_part, lb0);
    else if (e_part)
      k_rho_df<1, true, false><<<grid, block, sh, st>>>(g, h, off, nk_even, nk_odd, P, RHO, DF, tf, tm, rc2, e_part, lb0);
    else if (lean)
      k_rho_df<1, false, true><<<grid, block, sh, st>>>(g, h, off, nk_even, nk_odd, P, RHO, DF, tf, tm, rc2, e_part, lb0);
    else
      k_rho_df<1, false, false><<<grid, block, sh, st>>>(g, h, off, nk_even, nk_odd, P, RHO, DF, tf, tm, rc2, e_part, lb0);
  }
  return hipGetLastError();
}

hipError_t launch_df(hipStream_t st, const Geom &g, const int8_t *T, const double *RHO, double *DF,
                     const TabF64 &tf) {
  k_df<<<dim3((unsigned)site_blocks(g)), dim3(kBlock), 0, st>>>(g, T, RHO, DF, tf);
  return hipGetLastError();
}

hipError_t launch_force(hipStream_t st, const Geom &g0, const Halo &h, const int32_t *off, int32_t nk_even,
                        int32_t nk_odd, const double4 *P, const double *DF, double *FX, double *FY, double *FZ,
                        const TabF64 &tf, const TabMixed &tm, double rc2, int prec, double *e_part, int64_t lb0,
                        int64_t lb1) {
  Geom g = g0;
  if (lb1 < 0 || lb1 > g0.n_tiles) lb1 = g0.n_tiles;
  if (lb0 >= lb1) return hipSuccess;
  g.n_tiles = lb1;
  const dim3 grid((unsigned)((lb1 - lb0 + kXcds - 1) / kXcds * kXcds)), block(kBlock);
  const size_t sh = tile_lds_bytes(h, true);
  if (prec == 0) {
    if (e_part)
      k_force<0, true><<<grid, block, sh, st>>>(g, h, off, nk_even, nk_odd, P, DF, FX, FY, FZ, tf, tm, rc2, e_part, lb0);
    else
      k_force<0, false><<<grid, block, sh, st>>>(g, h, off, nk_even, nk_odd, P, DF, FX, FY, FZ, tf, tm, rc2, e_part, lb0);
  } else {
    if (e_part)
      k_force<1, true><<<grid, block, sh, st>>>(g, h, off, nk_even, nk_odd, P, DF, FX, FY, FZ, tf, tm, rc2, e_part, lb0);
    else
      k_force<1, false><<<grid, block, sh, st>>>(g, h, off, nk_even, nk_odd, P, DF, FX, FY, FZ, tf, tm, rc2, e_part, lb0);
  }
  return hipGetLastError();
}

// Dynamic LDS beyond the default 64 KB must be enabled per kernel.
hipError_t pair_kernels_allow_lds(size_t bytes) {
  const void *fns[] = {reinterpret_cast<const void *>(k_rho_df<0, true, false>),
                       reinterpret_cast<const void *>(k_rho_df<0, false, false>),
                       reinterpret_cast<const void *>(k_rho_df<1, true, false>),
                       reinterpret_cast<const void *>(k_rho_df<1, false, false>),
                       reinterpret_cast<const void *>(k_rho_df<1, false, true>),
                       reinterpret_cast<const void *>(k_rho_df<1, true, true>),
                       reinterpret_cast<const void *>(k_force<0, true>), reinterpret_cast<const void *>(k_force<0, false>),
                       reinterpret_cast<const void *>(k_force<1, true>), reinterpret_cast<const void *>(k_force<1, false>)};
  for (const void *f : fns) {
    hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace mmd
