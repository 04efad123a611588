// gfx950 kernels around the EAM passes: layout conversion, host AoS record
// pack/unpack, periodic ghost-shell refresh, velocity Verlet and the
// deterministic energy reduction.  All are HBM-bound streaming kernels.
#include "kernels.h"
#include "md_body.h"

namespace mmd {

namespace {

template <typename T>
__global__ __launch_bounds__(kBlock) void k_ref2dev(Geom g, const T *__restrict__ ref, T *__restrict__ dev) {
  const uint32_t idx = blockIdx.x * kBlock + threadIdx.x;
  if (idx >= (uint32_t)g.n_ext) return;
  dev[idx] = ref[ref_of(g, ext_of(g, idx))];
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_dev2ref(Geom g, const T *__restrict__ dev, T *__restrict__ ref) {
  const uint32_t idx = blockIdx.x * kBlock + threadIdx.x;
  if (idx >= (uint32_t)g.n_ext) return;
  ref[ref_of(g, ext_of(g, idx))] = dev[idx];
}

__global__ __launch_bounds__(kBlock) void k_ref2dev_pos(Geom g, const double *__restrict__ rx,
                                                          const double *__restrict__ ry, const double *__restrict__ rz,
                                                          const int8_t *__restrict__ T, double4 *__restrict__ P,
                                                          bool set_xyz) {
  const uint32_t idx = blockIdx.x * kBlock + threadIdx.x;
  if (idx >= (uint32_t)g.n_ext) return;
  double4 p = P[idx];
  if (set_xyz) {
    const int64_t r = ref_of(g, ext_of(g, idx));
    p.x = rx[r], p.y = ry[r], p.z = rz[r];
  }
  p.w = (double)T[idx];
  P[idx] = p;
}

__global__ __launch_bounds__(kBlock) void k_dev2ref_pos(Geom g, const double4 *__restrict__ P, double *__restrict__ rx,
                                                          double *__restrict__ ry, double *__restrict__ rz) {
  const uint32_t idx = blockIdx.x * kBlock + threadIdx.x;
  if (idx >= (uint32_t)g.n_ext) return;
  const int64_t r = ref_of(g, ext_of(g, idx));
  const double4 p = P[idx];
  rx[r] = p.x, ry[r] = p.y, rz[r] = p.z;
}

// Types are checked here: a type with no table (>= n_types) is stored as a
// vacancy, so device state never indexes a table out of range, and the
// largest type read goes to max_type for the caller to reject the call: one
// atomic per block, spread over kMaxTypeSlots slots a cache line apart (one
// address for every block serialises the atomics).
__global__ __launch_bounds__(kBlock) void k_aos_unpack(Geom g, const char *__restrict__ aos, AosDesc d,
                                                         uint32_t fields, double4 *__restrict__ P,
                                                         int8_t *__restrict__ T, double *__restrict__ DF,
                                                         double *__restrict__ RHO, int32_t n_types,
                                                         int32_t *__restrict__ max_type, uint32_t i0, uint32_t i1) {
  const uint32_t idx = i0 + blockIdx.x * kBlock + threadIdx.x;
  const bool live = idx < i1;
  const char *rec = aos + (live ? ref_of(g, ext_of(g, idx)) * d.stride : 0);
  int32_t t = -1;
  if (live && (fields & 0x20u)) {
    if (d.type_bytes == 1)
      t = *reinterpret_cast<const int8_t *>(rec + d.off_type);
    else if (d.type_bytes == 2)
      t = *reinterpret_cast<const int16_t *>(rec + d.off_type);
    else
      t = *reinterpret_cast<const int32_t *>(rec + d.off_type);
  }
  if (fields & 0x20u) {
    __shared__ int32_t wmax[kWavesPerBlock];
    int32_t m = t;
    for (int o = 32; o > 0; o >>= 1) m = max(m, __shfl_xor(m, o, kWave));
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
      m = max(max(wmax[0], wmax[1]), max(wmax[2], wmax[3]));
      if (m >= 0) atomicMax(max_type + (blockIdx.x % kMaxTypeSlots) * kMaxTypeStride, m);
    }
  }
  if (!live) return;
  double4 p = P[idx];
  if (fields & 0x20u) {
    const int8_t t8 = (int8_t)(t < 0 || t >= n_types ? -1 : t);
    T[idx] = t8;
    p.w = (double)t8;
  }
  if (fields & 0x01u) {
    const double *x = reinterpret_cast<const double *>(rec + d.off_x);
    p.x = x[0], p.y = x[1], p.z = x[2];
  }
  if (fields & 0x21u) P[idx] = p;
  if (fields & 0x10u) DF[idx] = *reinterpret_cast<const double *>(rec + d.off_df);
  if (fields & 0x08u) RHO[idx] = *reinterpret_cast<const double *>(rec + d.off_rho);
}

// writes results of local sites with type >= 0 only (vacancies untouched)
// two doubles stored together at an 8-byte aligned record offset
typedef double DPair __attribute__((ext_vector_type(2), aligned(8)));

__global__ __launch_bounds__(kBlock) void k_aos_pack(Geom g, char *__restrict__ aos, AosDesc d, uint32_t fields,
                                                       const int8_t *__restrict__ T, const double *__restrict__ RHO,
                                                       const double *__restrict__ DF, const double *__restrict__ FX,
                                                       const double *__restrict__ FY, const double *__restrict__ FZ,
                                                       int64_t q0, int64_t q1) {
  Site st;
  const int64_t q = q0 + (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (q >= q1 || !map_site_q(g, q, st)) return;
  const int64_t i = st.i;
  if (T[i] < 0) return;
  ExtPos p;
  {
    const int64_t row = i / g.exh;
    p.xh = (int32_t)(i - row * g.exh);
    p.s = (int32_t)(row & 1);
    const int64_t yz = row >> 1;
    p.ze = (int32_t)(yz / g.ext_y);
    p.ye = (int32_t)(yz - (int64_t)p.ze * g.ext_y);
  }
  char *rec = aos + ref_of(g, p) * d.stride;
  // Adjacent output fields leave in one store each (rho and df as 16 bytes,
  // f as 16 + 8): over the host link every store of a lane is a transaction
  // of its own, so fewer, wider stores move the same bytes faster.
  if ((fields & 0x18u) == 0x18u && d.off_df == d.off_rho + 8) {
    *reinterpret_cast<DPair *>(rec + d.off_rho) = DPair{RHO[i], DF[i]};
  } else {
    if (fields & 0x08u) *reinterpret_cast<double *>(rec + d.off_rho) = RHO[i];
    if (fields & 0x10u) *reinterpret_cast<double *>(rec + d.off_df) = DF[i];
  }
  if (fields & 0x04u) {
    char *f = rec + d.off_f;
    *reinterpret_cast<DPair *>(f) = DPair{FX[i], FY[i]};
    *reinterpret_cast<double *>(f + 16) = FZ[i];
  }
}

__global__ __launch_bounds__(kBlock) void k_ghost(Geom g, double Lx, double Ly, double Lz, double4 *__restrict__ P,
                                                    int8_t *__restrict__ T, double *__restrict__ DF, uint32_t fields) {
  ghost_site(g, blockIdx.x * kBlock + threadIdx.x, Lx, Ly, Lz, P, T, DF, fields);
}

template <bool DRIFT, bool WANT_KE>
__global__ __launch_bounds__(kBlock) void k_kick(Geom g, Masses M, double h, double dt, double4 *__restrict__ P,
                                                   double *__restrict__ VX, double *__restrict__ VY,
                                                   double *__restrict__ VZ, const double *__restrict__ FX,
                                                   const double *__restrict__ FY, const double *__restrict__ FZ,
                                                   const int8_t *__restrict__ T, double *__restrict__ ke_part) {
  __shared__ double lds[kWavesPerBlock];
  kick_block<DRIFT, WANT_KE>(g, xcd_block(), M, h, dt, P, VX, VY, VZ, FX, FY, FZ, T, ke_part, lds);
}

template <bool WANT_KE>
__global__ __launch_bounds__(kBlock) void k_kick_fused(Geom g, Masses M, double h, double dt, double4 *__restrict__ P,
                                                         double *__restrict__ VX, double *__restrict__ VY,
                                                         double *__restrict__ VZ, const double *__restrict__ FX,
                                                         const double *__restrict__ FY, const double *__restrict__ FZ,
                                                         const int8_t *__restrict__ T, double *__restrict__ ke_part) {
  __shared__ double lds[kWavesPerBlock];
  kick_fused_block<WANT_KE>(g, xcd_block(), M, h, dt, P, VX, VY, VZ, FX, FY, FZ, T, ke_part, lds);
}

__device__ double ordered_sum(const double *__restrict__ p, int64_t n) {
  double s = 0.0;
  for (int64_t k = threadIdx.x; k < n; k += kBlock) s += p[k];
  return s;
}

__global__ __launch_bounds__(kBlock) void k_energy_finalize(const double *__restrict__ ep_a, int64_t na,
                                                              const double *__restrict__ ep_b, int64_t nb,
                                                              const double *__restrict__ ep_c, int64_t nc,
                                                              const double *__restrict__ ke, int64_t nk,
                                                              const double *__restrict__ ke_b, int64_t nkb,
                                                              double *__restrict__ out, int32_t *__restrict__ counter) {
  __shared__ double lds[kWavesPerBlock];
  double pe = 0.0;
  if (ep_a) pe += ordered_sum(ep_a, na);
  if (ep_b) pe += ordered_sum(ep_b, nb);
  if (ep_c) pe += ordered_sum(ep_c, nc);
  double kin = ke ? ordered_sum(ke, nk) : 0.0;
  if (ke_b) kin += ordered_sum(ke_b, nkb);
  const double PE = block_sum(pe, lds);
  const double KE = block_sum(kin, lds);
  if (threadIdx.x == 0) {
    const int32_t slot = counter[0];
    out[2 * slot + 0] = PE;
    out[2 * slot + 1] = KE;
    counter[0] = slot + 1;
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_fill(T *__restrict__ p, int64_t n, T v) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ __launch_bounds__(kBlock) void k_stream_copy(const double2 *__restrict__ src, double2 *__restrict__ dst,
                                                          int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) dst[i] = src[i];
}

inline dim3 grid_ext(const Geom &g) { return dim3((unsigned)((g.n_ext + kBlock - 1) / kBlock)); }

}  // namespace

hipError_t launch_ref2dev_pos(hipStream_t st, const Geom &g, const double *rx, const double *ry, const double *rz,
                              const int8_t *T, double4 *P, bool set_xyz) {
  k_ref2dev_pos<<<grid_ext(g), kBlock, 0, st>>>(g, rx, ry, rz, T, P, set_xyz);
  return hipGetLastError();
}
hipError_t launch_dev2ref_pos(hipStream_t st, const Geom &g, const double4 *P, double *rx, double *ry, double *rz) {
  k_dev2ref_pos<<<grid_ext(g), kBlock, 0, st>>>(g, P, rx, ry, rz);
  return hipGetLastError();
}
hipError_t launch_ref2dev_f64(hipStream_t st, const Geom &g, const double *ref, double *dev) {
  k_ref2dev<double><<<grid_ext(g), kBlock, 0, st>>>(g, ref, dev);
  return hipGetLastError();
}
hipError_t launch_dev2ref_f64(hipStream_t st, const Geom &g, const double *dev, double *ref) {
  k_dev2ref<double><<<grid_ext(g), kBlock, 0, st>>>(g, dev, ref);
  return hipGetLastError();
}
hipError_t launch_ref2dev_i8(hipStream_t st, const Geom &g, const int8_t *ref, int8_t *dev) {
  k_ref2dev<int8_t><<<grid_ext(g), kBlock, 0, st>>>(g, ref, dev);
  return hipGetLastError();
}
hipError_t launch_dev2ref_i8(hipStream_t st, const Geom &g, const int8_t *dev, int8_t *ref) {
  k_dev2ref<int8_t><<<grid_ext(g), kBlock, 0, st>>>(g, dev, ref);
  return hipGetLastError();
}

hipError_t launch_aos_unpack(hipStream_t st, const Geom &g, const char *aos, AosDesc d, uint32_t fields, double4 *P,
                             int8_t *T, double *DF, double *RHO, int32_t n_types, int32_t *max_type, int64_t i0,
                             int64_t i1) {
  if (i1 < 0 || i1 > g.n_ext) i1 = g.n_ext;
  if (i0 >= i1) return hipSuccess;
  k_aos_unpack<<<dim3((unsigned)((i1 - i0 + kBlock - 1) / kBlock)), kBlock, 0, st>>>(
      g, aos, d, fields, P, T, DF, RHO, n_types, max_type, (uint32_t)i0, (uint32_t)i1);
  return hipGetLastError();
}

hipError_t launch_aos_pack(hipStream_t st, const Geom &g, char *aos, AosDesc d, uint32_t fields, const int8_t *T,
                           const double *RHO, const double *DF, const double *FX, const double *FY,
                           const double *FZ, int64_t q0, int64_t q1) {
  if (q1 < 0 || q1 > g.n_local) q1 = g.n_local;
  if (q0 >= q1) return hipSuccess;
  k_aos_pack<<<dim3((unsigned)((q1 - q0 + kBlock - 1) / kBlock)), kBlock, 0, st>>>(g, aos, d, fields, T, RHO, DF, FX,
                                                                                   FY, FZ, q0, q1);
  return hipGetLastError();
}

hipError_t launch_ghost(hipStream_t st, const Geom &g, double Lx, double Ly, double Lz, double4 *P, int8_t *T,
                        double *DF, uint32_t fields) {
  if (g.n_ghost == 0) return hipSuccess;
  k_ghost<<<dim3((unsigned)((g.n_ghost + kBlock - 1) / kBlock)), kBlock, 0, st>>>(g, Lx, Ly, Lz, P, T, DF, fields);
  return hipGetLastError();
}

hipError_t launch_kick(hipStream_t st, const Geom &g, const Masses &M, double h, double dt, bool drift, double4 *P,
                       double *VX, double *VY, double *VZ, const double *FX, const double *FY, const double *FZ,
                       const int8_t *T, double *ke_part) {
  const dim3 grid((unsigned)site_blocks(g));
  if (drift) {
    if (ke_part)
      k_kick<true, true><<<grid, kBlock, 0, st>>>(g, M, h, dt, P, VX, VY, VZ, FX, FY, FZ, T, ke_part);
    else
      k_kick<true, false><<<grid, kBlock, 0, st>>>(g, M, h, dt, P, VX, VY, VZ, FX, FY, FZ, T, ke_part);
  } else {
    if (ke_part)
      k_kick<false, true><<<grid, kBlock, 0, st>>>(g, M, h, dt, P, VX, VY, VZ, FX, FY, FZ, T, ke_part);
    else
      k_kick<false, false><<<grid, kBlock, 0, st>>>(g, M, h, dt, P, VX, VY, VZ, FX, FY, FZ, T, ke_part);
  }
  return hipGetLastError();
}

hipError_t launch_kick_fused(hipStream_t st, const Geom &g, const Masses &M, double h, double dt, double4 *P,
                             double *VX, double *VY, double *VZ, const double *FX, const double *FY, const double *FZ,
                             const int8_t *T, double *ke_part) {
  const dim3 grid((unsigned)site_blocks(g));
  if (ke_part)
    k_kick_fused<true><<<grid, kBlock, 0, st>>>(g, M, h, dt, P, VX, VY, VZ, FX, FY, FZ, T, ke_part);
  else
    k_kick_fused<false><<<grid, kBlock, 0, st>>>(g, M, h, dt, P, VX, VY, VZ, FX, FY, FZ, T, ke_part);
  return hipGetLastError();
}

hipError_t launch_energy_finalize(hipStream_t st, const double *ep_a, int64_t na, const double *ep_b, int64_t nb,
                                  const double *ep_c, int64_t nc, const double *ke, int64_t nk, const double *ke_b,
                                  int64_t nkb, double *out, int32_t *counter) {
  k_energy_finalize<<<1, kBlock, 0, st>>>(ep_a, na, ep_b, nb, ep_c, nc, ke, nk, ke_b, nkb, out, counter);
  return hipGetLastError();
}

hipError_t launch_stream_copy(hipStream_t st, const void *src, void *dst, int64_t bytes) {
  const int64_t n = bytes / 16;
  k_stream_copy<<<dim3(8192), kBlock, 0, st>>>(static_cast<const double2 *>(src), static_cast<double2 *>(dst), n);
  return hipGetLastError();
}

hipError_t launch_fill_i32(hipStream_t st, int32_t *p, int64_t n, int32_t v) {
  if (n <= 0) return hipSuccess;
  k_fill<int32_t><<<dim3((unsigned)((n + kBlock - 1) / kBlock)), kBlock, 0, st>>>(p, n, v);
  return hipGetLastError();
}
hipError_t launch_fill_f64(hipStream_t st, double *p, int64_t n, double v) {
  if (n <= 0) return hipSuccess;
  k_fill<double><<<dim3((unsigned)((n + kBlock - 1) / kBlock)), kBlock, 0, st>>>(p, n, v);
  return hipGetLastError();
}
hipError_t launch_fill_i8(hipStream_t st, int8_t *p, int64_t n, int8_t v) {
  if (n <= 0) return hipSuccess;
  k_fill<int8_t><<<dim3((unsigned)((n + kBlock - 1) / kBlock)), kBlock, 0, st>>>(p, n, v);
  return hipGetLastError();
}

}  // namespace mmd
