// Batched kernels of the serving loop (see batch.h).  The work list is split
// into one contiguous range per XCD (blocks are dispatched to the 8 XCDs
// round-robin, so blockIdx.x % 8 names the XCD), so the blocks of an XCD take
// neighbouring blocks of the same replica and share its halo in L2.  The
// item counts are read from device memory, so a grid of fixed size serves
// any list up to the item capacity it was sized for: the launchers give one
// block per item of that capacity (the serving loop's step launches size it
// to the lists' capacities, and blocks past the count exit at once; its
// priming launches to the item count).
#include "batch.h"
#include "eam_body.h"
#include "kernels.h"
#include "md_body.h"

namespace mmd {

namespace {

// This block's item of a list of n: the XCD's contiguous range holds
// ceil(n / 8) items and the XCD gridDim.x / 8 >= ceil(n / 8) blocks (the
// launchers' grids cover the lists' capacities), so every item has a block
// and no block has two; -1 for a block past the count.
__device__ __forceinline__ int xcd_item(int n) {
  const int xcd = blockIdx.x % kXcds, per = (n + kXcds - 1) / kXcds;
  const int it = xcd * per + (int)(blockIdx.x / kXcds);
  return it < min(n, (xcd + 1) * per) ? it : -1;
}

// Work items and replica descriptors are read through the constant address
// space: the host writes them before a launch and no kernel writes them, so
// the compiler loads them with scalar loads.  Through plain global pointers
// it cannot prove them unwritten by the kernel's own stores and keeps the
// ~40 descriptor words in vector registers, which pushed the force bodies
// into register spills.
template <class T>
__device__ __forceinline__ T load_const(const T *p) {
  static_assert(sizeof(T) % 4 == 0, "read as 32-bit words");
  const uint32_t __attribute__((address_space(4))) *q = (const uint32_t __attribute__((address_space(4))) *)p;
  T out;
  uint32_t *o = reinterpret_cast<uint32_t *>(&out);
#pragma unroll
  for (int k = 0; k < (int)(sizeof(T) / 4); k++) o[k] = q[k];
  return out;
}

template <bool DRIFT, bool WANT_KE>
__global__ __launch_bounds__(kBlock) void kb_kick(BatchLists L, Masses M, double h, double dt) {
  __shared__ double lds[kWavesPerBlock];
  const int it = xcd_item(*L.n_site_items);
  if (it >= 0) {
    const int2 w = load_const(L.site_items + it);
    const BatchDom *d = L.dom + w.x;
    kick_block<DRIFT, WANT_KE>(load_const(&d->g), w.y, M, h, dt, load_const(&d->pos), load_const(&d->vx),
                               load_const(&d->vy), load_const(&d->vz), load_const(&d->fx), load_const(&d->fy),
                               load_const(&d->fz), load_const(&d->type), load_const(&d->e_ke), lds);
  }
}

template <bool WANT_KE>
__global__ __launch_bounds__(kBlock) void kb_kick_fused(BatchLists L, Masses M, double h, double dt) {
  __shared__ double lds[kWavesPerBlock];
  const int it = xcd_item(*L.n_site_items);
  if (it >= 0) {
    const int2 w = load_const(L.site_items + it);
    const BatchDom *d = L.dom + w.x;
    kick_fused_block<WANT_KE>(load_const(&d->g), w.y, M, h, dt, load_const(&d->pos), load_const(&d->vx),
                              load_const(&d->vy), load_const(&d->vz), load_const(&d->fx), load_const(&d->fy),
                              load_const(&d->fz), load_const(&d->type), load_const(&d->e_ke), lds);
  }
}

__global__ __launch_bounds__(kBlock) void kb_ghost(BatchLists L, uint32_t fields) {
  const int it = xcd_item(*L.n_ext_items);
  if (it >= 0) {
    const int2 w = load_const(L.ext_items + it);
    const BatchDom *d = L.dom + w.x;
    ghost_site(load_const(&d->g), (uint32_t)w.y * kBlock + threadIdx.x, load_const(&d->L[0]), load_const(&d->L[1]),
               load_const(&d->L[2]), load_const(&d->pos), load_const(&d->type), load_const(&d->df), fields);
  }
}

template <int PREC, bool WANT_E, bool LEAN>
__global__ __launch_bounds__(kBlock) void kb_rho_df(BatchLists L, Halo h, TabF64 tf, TabMixed tm, double rc2) {
  extern __shared__ double2 lds_tile[];
  __shared__ double red[kWavesPerBlock];
  const int it = xcd_item(*L.n_tile_items);
  if (it >= 0) {
    const int2 w = load_const(L.tile_items + it);
    const BatchDom *d = L.dom + w.x;
    rho_df_block<PREC, WANT_E, LEAN>(load_const(&d->g), h, w.y, load_const(&d->off), load_const(&d->nk_even),
                                     load_const(&d->nk_odd), load_const(&d->pos), load_const(&d->rho),
                                     load_const(&d->df), tf, tm, rc2, load_const(&d->e_rho), lds_tile, red);
  }
}

template <int PREC, bool WANT_E>
__global__ __launch_bounds__(kBlock, kForceWavesPerSimd) void kb_force(BatchLists L, Halo h, TabF64 tf, TabMixed tm, double rc2) {
  extern __shared__ double2 lds_tile[];
  __shared__ double red[kWavesPerBlock];
  const int it = xcd_item(*L.n_tile_items);
  if (it >= 0) {
    const int2 w = load_const(L.tile_items + it);
    const BatchDom *d = L.dom + w.x;
    force_block<PREC, WANT_E>(load_const(&d->g), h, w.y, load_const(&d->off), load_const(&d->nk_even),
                              load_const(&d->nk_odd), load_const(&d->pos), load_const(&d->df), load_const(&d->fx),
                              load_const(&d->fy), load_const(&d->fz), tf, tm, rc2, load_const(&d->e_force), lds_tile,
                              red);
  }
}

__device__ double ordered_sum_b(const double *__restrict__ p, int64_t n) {
  double s = 0.0;
  for (int64_t k = threadIdx.x; k < n; k += kBlock) s += p[k];
  return s;
}

// one block per replica: the same fixed-order reduction as the single-domain
// energy finalize; writes {potential, kinetic} to energy[2 * which]
__global__ __launch_bounds__(kBlock) void kb_energy(BatchLists L, int which) {
  __shared__ double lds[kWavesPerBlock];
  const int n = *L.n_reps;
  for (int q = blockIdx.x; q < n; q += gridDim.x) {
    const BatchDom &d = L.dom[L.reps[q]];
    double pe = ordered_sum_b(d.e_rho, d.n_tiles);
    pe += ordered_sum_b(d.e_force, d.n_tiles);
    const double kin = ordered_sum_b(d.e_ke, d.n_blocks);
    const double PE = block_sum(pe, lds);
    const double KE = block_sum(kin, lds);
    if (threadIdx.x == 0) {
      d.energy[2 * which] = PE;
      d.energy[2 * which + 1] = KE;
    }
  }
}

}  // namespace

hipError_t launch_batch_kick_fused(hipStream_t st, const BatchLists &L, int64_t items, const Masses &M, double h,
                                   double dt, bool want_ke) {
  const dim3 grid(batch_blocks(items)), block(kBlock);
  if (want_ke)
    kb_kick_fused<true><<<grid, block, 0, st>>>(L, M, h, dt);
  else
    kb_kick_fused<false><<<grid, block, 0, st>>>(L, M, h, dt);
  return hipGetLastError();
}

hipError_t launch_batch_kick(hipStream_t st, const BatchLists &L, int64_t items, const Masses &M, double h, double dt,
                             bool drift, bool want_ke) {
  const dim3 grid(batch_blocks(items)), block(kBlock);
  if (drift) {
    if (want_ke)
      kb_kick<true, true><<<grid, block, 0, st>>>(L, M, h, dt);
    else
      kb_kick<true, false><<<grid, block, 0, st>>>(L, M, h, dt);
  } else {
    if (want_ke)
      kb_kick<false, true><<<grid, block, 0, st>>>(L, M, h, dt);
    else
      kb_kick<false, false><<<grid, block, 0, st>>>(L, M, h, dt);
  }
  return hipGetLastError();
}

hipError_t launch_batch_ghost(hipStream_t st, const BatchLists &L, int64_t items, uint32_t fields) {
  kb_ghost<<<dim3(batch_blocks(items)), dim3(kBlock), 0, st>>>(L, fields);
  return hipGetLastError();
}

hipError_t launch_batch_rho_df(hipStream_t st, const BatchLists &L, int64_t items, const Halo &h, const TabF64 &tf,
                               const TabMixed &tm, double rc2, int prec, bool want_e, bool lean) {
  const dim3 grid(batch_blocks(items)), block(kBlock);
  const size_t sh = tile_lds_bytes(h, false);
  if (prec == 0) {
    if (want_e)
      kb_rho_df<0, true, false><<<grid, block, sh, st>>>(L, h, tf, tm, rc2);
    else
      kb_rho_df<0, false, false><<<grid, block, sh, st>>>(L, h, tf, tm, rc2);
  } else {
    if (want_e && lean)
      kb_rho_df<1, true, true><<<grid, block, sh, st>>>(L, h, tf, tm, rc2);
    else if (want_e)
      kb_rho_df<1, true, false><<<grid, block, sh, st>>>(L, h, tf, tm, rc2);
    else if (lean)
      kb_rho_df<1, false, true><<<grid, block, sh, st>>>(L, h, tf, tm, rc2);
    else
      kb_rho_df<1, false, false><<<grid, block, sh, st>>>(L, h, tf, tm, rc2);
  }
  return hipGetLastError();
}

hipError_t launch_batch_force(hipStream_t st, const BatchLists &L, int64_t items, const Halo &h, const TabF64 &tf,
                              const TabMixed &tm, double rc2, int prec, bool want_e) {
  const dim3 grid(batch_blocks(items)), block(kBlock);
  const size_t sh = tile_lds_bytes(h, true);
  if (prec == 0) {
    if (want_e)
      kb_force<0, true><<<grid, block, sh, st>>>(L, h, tf, tm, rc2);
    else
      kb_force<0, false><<<grid, block, sh, st>>>(L, h, tf, tm, rc2);
  } else {
    if (want_e)
      kb_force<1, true><<<grid, block, sh, st>>>(L, h, tf, tm, rc2);
    else
      kb_force<1, false><<<grid, block, sh, st>>>(L, h, tf, tm, rc2);
  }
  return hipGetLastError();
}

hipError_t batch_kernels_allow_lds(size_t bytes) {
  const void *fns[] = {reinterpret_cast<const void *>(kb_rho_df<0, true, false>),
                       reinterpret_cast<const void *>(kb_rho_df<0, false, false>),
                       reinterpret_cast<const void *>(kb_rho_df<1, true, false>),
                       reinterpret_cast<const void *>(kb_rho_df<1, false, false>),
                       reinterpret_cast<const void *>(kb_rho_df<1, false, true>),
                       reinterpret_cast<const void *>(kb_rho_df<1, true, true>),
                       reinterpret_cast<const void *>(kb_force<0, true>), reinterpret_cast<const void *>(kb_force<0, false>),
                       reinterpret_cast<const void *>(kb_force<1, true>), reinterpret_cast<const void *>(kb_force<1, false>)};
  for (const void *f : fns) {
    hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_batch_energy(hipStream_t st, const BatchLists &L, int which) {
  kb_energy<<<dim3(256), dim3(kBlock), 0, st>>>(L, which);
  return hipGetLastError();
}

}  // namespace mmd
