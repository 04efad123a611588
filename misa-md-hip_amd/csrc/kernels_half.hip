// Newton's-third-law half stencil for the periodic single-domain step: each
// lattice pair's force scalar is evaluated once, by the centre whose ideal
// displacement to the neighbour is lexicographically positive in (z, y, x).
//
// The reference's half-stencil mode (USE_NEWTONS_THIRD_LOW,
// src/hip_arch_acc_imp.cpp:118-130, src/kernels/kernel_pairs.hpp:28-31,
// :55-59) scatters the reaction into the neighbour's force with unordered
// updates; here the reaction is pulled instead, so every result is summed in
// a fixed order and needs no atomics:
//
//   forward pass  (k_force_fwd):  every local site walks its forward entries,
//     accumulates del * fpair into its own force (written) and stores fpair in
//     slot k of its record (slots[k][q], q = local site number; 0 for a pair
//     beyond the cutoff, a vacant neighbour or a vacant centre);
//   backward pass (k_force_back): every local site walks its backward entries,
//     reads the neighbour's slot for the entry that reaches back to it (the
//     neighbour's periodic image when it sits in the ghost shell) and adds
//     del * fpair, del = x_i - x_j as in the forward pass.
//
// The backward pass trusts the forward pass's cutoff decision (it reads the
// slot, it does not retest r), so each pair enters both atoms or neither.
// The ghost shell must hold the periodic images of the local sites (the
// periodic step refreshes it), so the mode exists only for that step.
#include "eam_body.h"
#include "kernels.h"

namespace mmd {

constexpr int kBackUnroll = 8;  // backward entries whose slot loads are in flight together
static_assert(kBackUnroll <= 8, "stencil padding too short for the unroll");

__device__ __forceinline__ int32_t wrap1(int32_t v, int32_t n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

// forward pass: force_block's HALF mode (forward entries, slot stores)
template <int PREC, bool WANT_E>
__global__ __launch_bounds__(kBlock, kForceWavesPerSimd) void k_force_fwd(Geom g, Halo h, const int32_t *__restrict__ img, int32_t nf_even,
                                                        int32_t nf_odd, const double4 *__restrict__ P,
                                                        const double *__restrict__ DF, double *__restrict__ FX,
                                                        double *__restrict__ FY, double *__restrict__ FZ,
                                                        double *__restrict__ SL, TabF64 tf, TabMixed tm, double rc2,
                                                        double *__restrict__ e_part) {
  extern __shared__ double2 lds_tile[];
  __shared__ double red[kWavesPerBlock];
  force_block<PREC, WANT_E, true>(g, h, xcd_block(), img + kHalfFwd, nf_even, nf_odd, P, DF, FX, FY, FZ, tf, tm, rc2,
                                  e_part, lds_tile, red, SL);
}

// LDS tile of the backward pass (n records): (x, y) [n], z [n], the local
// site number of the record's periodic image [n], type [n].
__global__ __launch_bounds__(kBlock) void k_force_back(Geom g, Halo h, const int32_t *__restrict__ img, int32_t nb_even,
                                                         int32_t nb_odd, const double4 *__restrict__ P,
                                                         double *__restrict__ FX, double *__restrict__ FY,
                                                         double *__restrict__ FZ, const double *__restrict__ SL) {
  extern __shared__ double2 lds_tile[];
  const int64_t lb = xcd_block();
  const TileSite t = map_tile(g, h, lb);
  double2 *lxy = lds_tile;
  double *lz = reinterpret_cast<double *>(lds_tile + h.n);
  uint32_t *lq = reinterpret_cast<uint32_t *>(lz + h.n);
  int8_t *lt = reinterpret_cast<int8_t *>(lq + h.n);
  if (lb < g.n_tiles) {
    for (int32_t idx = threadIdx.x; idx < h.n; idx += kBlock) {
      const int64_t src = tile_record_src(g, h, t, idx);
      double4 p = make_double4(0.0, 0.0, 0.0, -1.0);
      uint32_t qi = 0;
      if (src >= 0) {
        p = P[src];
        const int32_t lxh = idx % h.lx, r = idx / h.lx, s = r & 1, yz = r >> 1;
        const int32_t ly = yz % h.ly, lzi = yz / h.ly;
        qi = local_q(g, s, wrap1(t.x0 + lxh - h.rx, g.bxh), wrap1(t.y0 + ly - h.ry, g.by),
                     wrap1(t.z0 + lzi - h.rz, g.bz));
      }
      if (p.w < 0.0) p.x = p.y = p.z = kFarPos;
      lxy[idx] = make_double2(p.x, p.y);
      lz[idx] = p.z;
      lq[idx] = qi;
      lt[idx] = (int8_t)p.w;
    }
  }
  __syncthreads();
  if (!t.active || lt[t.c] < 0) return;
  const double2 cxy = lxy[t.c];
  const double cz = lz[t.c];
  const int32_t *o = img + kHalfBack + t.s * kMaxStencil;
  const int32_t *sk = img + kHalfSlot + t.s * kMaxStencil;
  const int nk = t.s ? nb_odd : nb_even;
  const size_t nl = (size_t)g.n_local;
  double ax = 0.0, ay = 0.0, az = 0.0;
  for (int k = 0; k < nk; k += kBackUnroll) {
    double f[kBackUnroll];
    int32_t j[kBackUnroll];
#pragma unroll
    for (int u = 0; u < kBackUnroll; u++) {
      j[u] = t.c + o[k + u];
      f[u] = SL[(size_t)sk[k + u] * nl + lq[j[u]]];
    }
#pragma unroll
    for (int u = 0; u < kBackUnroll; u++) {
      if (k + u < nk) {
        const double2 pxy = lxy[j[u]];
        ax = fma(cxy.x - pxy.x, f[u], ax);
        ay = fma(cxy.y - pxy.y, f[u], ay);
        az = fma(cz - lz[j[u]], f[u], az);
      }
    }
  }
  const uint32_t i = (uint32_t)t.i;
  FX[i] += ax;
  FY[i] += ay;
  FZ[i] += az;
}

size_t half_back_lds_bytes(const Halo &h) { return (size_t)h.n * 29; }

hipError_t launch_force_half(hipStream_t st, const Geom &g, const Halo &h, const int32_t *img, const int32_t nf[2],
                             const int32_t nb[2], const double4 *P, const double *DF, double *FX, double *FY,
                             double *FZ, double *SL, const TabF64 &tf, const TabMixed &tm, double rc2, int prec,
                             double *e_part) {
  const dim3 grid((unsigned)tile_blocks(g)), block(kBlock);
  const size_t sh = tile_lds_bytes(h, true);
  if (prec == 0) {
    if (e_part)
      k_force_fwd<0, true><<<grid, block, sh, st>>>(g, h, img, nf[0], nf[1], P, DF, FX, FY, FZ, SL, tf, tm, rc2, e_part);
    else
      k_force_fwd<0, false><<<grid, block, sh, st>>>(g, h, img, nf[0], nf[1], P, DF, FX, FY, FZ, SL, tf, tm, rc2, e_part);
  } else {
    if (e_part)
      k_force_fwd<1, true><<<grid, block, sh, st>>>(g, h, img, nf[0], nf[1], P, DF, FX, FY, FZ, SL, tf, tm, rc2, e_part);
    else
      k_force_fwd<1, false><<<grid, block, sh, st>>>(g, h, img, nf[0], nf[1], P, DF, FX, FY, FZ, SL, tf, tm, rc2, e_part);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  k_force_back<<<grid, block, half_back_lds_bytes(h), st>>>(g, h, img, nb[0], nb[1], P, FX, FY, FZ, SL);
  return hipGetLastError();
}

// Dynamic LDS beyond the default 64 KB must be enabled per kernel.
hipError_t half_kernels_allow_lds(size_t bytes) {
  const void *fns[] = {reinterpret_cast<const void *>(k_force_fwd<0, true>),
                       reinterpret_cast<const void *>(k_force_fwd<0, false>),
                       reinterpret_cast<const void *>(k_force_fwd<1, true>),
                       reinterpret_cast<const void *>(k_force_fwd<1, false>),
                       reinterpret_cast<const void *>(k_force_back)};
  for (const void *f : fns) {
    hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace mmd
