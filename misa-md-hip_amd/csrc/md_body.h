// Device code of the streaming passes shared by the single-domain kernels
// (kernels_md.hip) and the batched serving kernels (kernels_batch.hip):
// ext-site coordinates, periodic ghost refresh and velocity Verlet.
#ifndef MMD_MD_BODY_H
#define MMD_MD_BODY_H

#include "device_common.h"

namespace mmd {

constexpr double kFtm2v = 9.64853321e-3;  // (eV/A/amu) -> A/fs^2
constexpr double kMvv2e = 1.0 / kFtm2v;   // amu A^2/fs^2 -> eV

struct ExtPos {
  int32_t s, xh, ye, ze;
};

// device index -> ext coordinates (32-bit arithmetic: n_ext < 2^31 is checked at set-up)
__device__ __forceinline__ ExtPos ext_of(const Geom &g, uint32_t idx) {
  ExtPos p;
  const uint32_t row = idx / (uint32_t)g.exh;
  p.xh = (int32_t)(idx - row * (uint32_t)g.exh);
  p.s = (int32_t)(row & 1u);
  const uint32_t yz = row >> 1;
  const uint32_t ze = yz / (uint32_t)g.ext_y;
  p.ye = (int32_t)(yz - ze * (uint32_t)g.ext_y);
  p.ze = (int32_t)ze;
  return p;
}

__device__ __forceinline__ int64_t ref_of(const Geom &g, const ExtPos &p) {
  return ((int64_t)p.ze * g.ext_y + p.ye) * g.ext_x + 2 * p.xh + p.s;
}

__device__ __forceinline__ int32_t wrap_img(int32_t v, int32_t n, int32_t &img) {
  if (v < 0) {
    img = -1;
    return v + n;
  }
  if (v >= n) {
    img = 1;
    return v - n;
  }
  img = 0;
  return v;
}

// periodic single-domain ghost refresh of ext site idx (no-op for local sites)
// Device index of the q-th ghost site, or -1 past the last: the ghost
// z-layers first (both sides), then the ghost y-rows of the local layers,
// then the ghost x-columns of the local rows (the first two are contiguous
// ranges of the device layout).
__device__ __forceinline__ int64_t ghost_index(const Geom &g, uint32_t q) {
  const uint32_t row = (uint32_t)g.ext_x;  // one y row: both parity rows, 2 exh sites
  const uint32_t layer = (uint32_t)g.ext_y * row;
  const uint32_t n_a = 2u * (uint32_t)g.gz * layer;
  if (q < n_a) return q < (uint32_t)g.gz * layer ? q : q + (uint32_t)g.bz * layer;
  q -= n_a;
  const uint32_t per_layer = 2u * (uint32_t)g.gy * row, n_b = (uint32_t)g.bz * per_layer;
  if (q < n_b) {
    const uint32_t zl = q / per_layer, r = q - zl * per_layer;
    const uint32_t off = r < (uint32_t)g.gy * row ? r : r + (uint32_t)g.by * row;
    return (int64_t)(zl + (uint32_t)g.gz) * layer + off;
  }
  q -= n_b;
  const uint32_t side = (uint32_t)g.gxh, per_row = 4u * side;  // 2 parities x 2 sides
  if (q >= (uint32_t)g.bz * (uint32_t)g.by * per_row) return -1;
  const uint32_t rl = q / per_row, r = q - rl * per_row;
  const uint32_t zl = rl / (uint32_t)g.by, yl = rl - zl * (uint32_t)g.by;
  const uint32_t s = r / (2u * side), c = r - s * 2u * side;
  const uint32_t xh = c < side ? c : c + (uint32_t)g.bxh;
  return dev_index(g, (int32_t)s, (int32_t)xh, (int32_t)yl + g.gy, (int32_t)zl + g.gz);
}

// periodic refresh of the k-th ghost site from its source in the local box
__device__ __forceinline__ void ghost_site(const Geom &g, uint32_t k, double Lx, double Ly, double Lz,
                                           double4 *__restrict__ P, int8_t *__restrict__ T, double *__restrict__ DF,
                                           uint32_t fields) {
  const int64_t gi = ghost_index(g, k);
  if (gi < 0) return;
  const uint32_t idx = (uint32_t)gi;
  const ExtPos p = ext_of(g, idx);
  const int32_t xl = p.xh - g.gxh, yl = p.ye - g.gy, zl = p.ze - g.gz;
  int32_t ix, iy, iz;
  const int32_t sx = wrap_img(xl, g.bxh, ix), sy = wrap_img(yl, g.by, iy), sz = wrap_img(zl, g.bz, iz);
  const int64_t src = dev_index(g, p.s, sx + g.gxh, sy + g.gy, sz + g.gz);
  if (fields & 0x01u) {
    T[idx] = T[src];
    double4 q = P[src];
    q.x += (double)ix * Lx;
    q.y += (double)iy * Ly;
    q.z += (double)iz * Lz;
    P[idx] = q;
  }
  if (fields & 0x10u) DF[idx] = DF[src];
}

// velocity Verlet on the sites of logical block lb: v += h a (+ x += dt v)
template <bool DRIFT, bool WANT_KE>
__device__ __forceinline__ void kick_block(const Geom &g, int64_t lb, const Masses &M, double h, double dt,
                                           double4 *__restrict__ P, double *__restrict__ VX, double *__restrict__ VY,
                                           double *__restrict__ VZ, const double *__restrict__ FX,
                                           const double *__restrict__ FY, const double *__restrict__ FZ,
                                           const int8_t *__restrict__ T, double *__restrict__ ke_part,
                                           double *lds) {
  double ke = 0.0;
  Site st;
  if (map_site_lb(g, lb, st) && st.active) {
    const int64_t i = st.i;
    const int t = T[i];
    if (t >= 0) {
      const double sc = h * kFtm2v * M.inv_m[t];
      const double vx = VX[i] + sc * FX[i], vy = VY[i] + sc * FY[i], vz = VZ[i] + sc * FZ[i];
      VX[i] = vx;
      VY[i] = vy;
      VZ[i] = vz;
      if (DRIFT) {
        double4 q = P[i];
        q.x += dt * vx;
        q.y += dt * vy;
        q.z += dt * vz;
        P[i] = q;
      }
      if (WANT_KE) ke = 0.5 * M.m[t] * (vx * vx + vy * vy + vz * vz) * kMvv2e;
    }
  }
  if (WANT_KE) {
    const double b = block_sum(ke, lds);
    if (threadIdx.x == 0) ke_part[lb] = b;
  }
}

// The closing half kick of one step fused with the opening half kick and
// drift of the next: v += h a (the completed step's kinetic energy from this
// v), then v += h a and x += dt v.  The same operations in the same order as
// the two separate passes, so the trajectory is bit-identical.
template <bool WANT_KE>
__device__ __forceinline__ void kick_fused_block(const Geom &g, int64_t lb, const Masses &M, double h, double dt,
                                                 double4 *__restrict__ P, double *__restrict__ VX,
                                                 double *__restrict__ VY, double *__restrict__ VZ,
                                                 const double *__restrict__ FX, const double *__restrict__ FY,
                                                 const double *__restrict__ FZ, const int8_t *__restrict__ T,
                                                 double *__restrict__ ke_part, double *lds) {
  double ke = 0.0;
  Site st;
  if (map_site_lb(g, lb, st) && st.active) {
    const int64_t i = st.i;
    const int t = T[i];
    if (t >= 0) {
      const double sc = h * kFtm2v * M.inv_m[t];
      const double fx = FX[i], fy = FY[i], fz = FZ[i];
      const double ux = VX[i] + sc * fx, uy = VY[i] + sc * fy, uz = VZ[i] + sc * fz;
      if (WANT_KE) ke = 0.5 * M.m[t] * (ux * ux + uy * uy + uz * uz) * kMvv2e;
      const double vx = ux + sc * fx, vy = uy + sc * fy, vz = uz + sc * fz;
      VX[i] = vx;
      VY[i] = vy;
      VZ[i] = vz;
      double4 q = P[i];
      q.x += dt * vx;
      q.y += dt * vy;
      q.z += dt * vz;
      P[i] = q;
    }
  }
  if (WANT_KE) {
    const double b = block_sum(ke, lds);
    if (threadIdx.x == 0) ke_part[lb] = b;
  }
}

}  // namespace mmd

#endif
