// Device-side geometry, table descriptors and small helpers shared by the
// gfx950 kernels.
//
// Device layout ("sub-lattice split rows"): the ghost-extended lattice is
// stored as rows of one BCC sub-lattice each.  Site (xe, ye, ze) of the
// reference layout (x doubled, parity s = xe & 1, cell xh = xe >> 1) lives at
//     ((ze * ext_y + ye) * 2 + s) * exh + xh,      exh = ext_x / 2.
// A wavefront takes 64 consecutive cells of one row, so every lane of a wave
// has the same sub-lattice parity: the neighbour stencil is wave-uniform
// (scalar loads) and the k-th neighbours of the 64 lanes are 64 consecutive
// sites of one row (fully coalesced 512-byte loads per coordinate).
#ifndef MMD_DEVICE_COMMON_H
#define MMD_DEVICE_COMMON_H

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mmd {

constexpr int kWave = 64;
constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr int kMaxStencil = 512;  // entries per parity
constexpr int kMaxTypes = 8;
constexpr int kXcds = 8;          // MI355X: 8 accelerator complex dies, blocks dispatched round-robin

struct Geom {
  int32_t ext_x, ext_y, ext_z;  // ext_x in doubled units
  int32_t exh;                  // ext_x / 2 (cells per row)
  int32_t gxh, gy, gz;          // ghost width (x in cells)
  int32_t bxh, by, bz;          // local box (x in cells)
  int32_t chunks;               // 64-lane chunks per local row
  int64_t n_ext;                // ext sites
  int64_t n_waves;              // waves covering the local box
};

// Per-knot coefficients of an FP64 r- or rho-keyed spline: [c3 c4 c5 c6 | c0 c1 c2 0]
// value = ((c3 p + c4) p + c5) p + c6, derivative = (c0 p + c1) p + c2.
struct TabF64 {
  const double *rho;  // [nt][n_r][8]
  const double *emb;  // [nt][n_rho][8]
  const double *phi;  // [nt*nt][n_r][8]
  int32_t nt, n_r, n_rho;
  double r0, inv_dr, rho0, inv_drho;
};

// Mixed-precision spline keyed by s = r^2 (no square root or division per
// pair).  Per knot, with q the fraction of the segment:
//   value = base + q * (a1 + q * c),   c = a2 + q * a3   (c in FP32)
//   d/ds  = dbase + q * (b1 + q * b2)                     (b2 stored as FP32)
// The FP64 base and linear terms carry the leading digits; the FP32 terms are
// O(ds^2) corrections, so their rounding stays below 1e-13 relative.
// Value record (32 B): {double base; double a1; float a2, a3; double 0}
// Deriv record (32 B): {double dbase; double b1; float b2, 0; double 0}
struct TabMixed {
  const double4 *rho_v;   // [nt][n_s] value records (rho by source type)
  const double4 *rho_d;   // [nt][n_s] derivative records
  const double4 *phi_d;   // [nt*nt][n_s]
  const double4 *phi_v;   // [nt*nt][n_s]
  int32_t nt, n_s;
  double s0, inv_ds;
  // The reference spline's first and last three segments use low-order end
  // slopes, which leave second-derivative kinks a uniform s grid cannot follow:
  // pairs with s < s_head or s >= s_tail evaluate the FP64 r-spline instead.
  double s_head, s_tail;
};

struct Masses {
  double inv_m[kMaxTypes];  // 1 / mass (amu^-1)
  double m[kMaxTypes];
};

__host__ __device__ inline int64_t dev_index(const Geom &g, int32_t s, int32_t xh, int32_t ye, int32_t ze) {
  return ((int64_t)(ze * (int64_t)g.ext_y + ye) * 2 + s) * g.exh + xh;
}

// Blocks are dispatched to the 8 XCDs round-robin.  Remap so that XCD k takes
// a contiguous range of logical blocks (contiguous z-slabs of the lattice):
// neighbour halos then stay in that XCD's L2.  gridDim.x is a multiple of 8.
__device__ inline int64_t xcd_block() {
  const int64_t per = gridDim.x / kXcds;
  return (int64_t)(blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
}

template <typename T>
__device__ inline T wave_sum(T v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

// Deterministic block sum (fixed tree): every thread of the block must call.
__device__ inline double block_sum(double v, double *lds4) {
  v = wave_sum(v);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) lds4[w] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) r = (lds4[0] + lds4[1]) + (lds4[2] + lds4[3]);
  __syncthreads();
  return r;
}

// One thread per local site: wave w of the launch takes 64 consecutive cells
// of local row w / chunks (rows ordered z, y, parity).

struct Site {
  int64_t i;    // device index
  int s;        // sub-lattice parity (wave-uniform)
  bool active;  // inside the local box
};

__device__ __forceinline__ bool map_site(const Geom &g, Site &st) {
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t w = xcd_block() * kWavesPerBlock + (threadIdx.x >> 6);
  if (w >= g.n_waves) return false;
  const int64_t chunk = w % g.chunks;
  const int64_t rl = w / g.chunks;
  const int s = (int)(rl & 1);
  const int64_t yz = rl >> 1;
  const int yl = (int)(yz % g.by), zl = (int)(yz / g.by);
  const int xh = (int)chunk * kWave + lane;
  st.active = xh < g.bxh;
  st.s = __builtin_amdgcn_readfirstlane(s);
  st.i = dev_index(g, st.s, xh + g.gxh, yl + g.gy, zl + g.gz);
  return true;
}

}  // namespace mmd

#endif
