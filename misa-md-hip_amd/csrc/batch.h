// Batched (multi-replica) device descriptors and launchers used by the
// serving loop.  One launch covers every active replica: its blocks walk a
// device work list of (replica, block) items.  Each item runs exactly the
// per-block code of the single-domain kernels, and energies are reduced per
// replica in a fixed order, so a replica's results do not depend on which
// other replicas share the launch.
#ifndef MMD_BATCH_H
#define MMD_BATCH_H

#include <hip/hip_runtime.h>

#include <algorithm>

#include "device_common.h"

namespace mmd {

struct BatchDom {
  Geom g;
  double4 *pos;
  double *vx, *vy, *vz, *fx, *fy, *fz, *rho, *df;
  int8_t *type;
  const int32_t *off;  // stencil device image (kStencilImage words, shared by all replicas)
  int32_t nk_even, nk_odd;
  double *e_rho, *e_force;  // per pair-pass tile block
  double *e_ke;             // per local-site block
  double *energy;           // {potential, kinetic} after the last step, then at admission
  double L[3];
  int64_t n_blocks;  // local-site blocks, as the single-domain launch numbers them (site_blocks)
  int64_t n_tiles;   // pair-pass tile blocks (tile_blocks)
};

struct BatchLists {
  const BatchDom *dom;
  const int2 *site_items;  // (replica, local-site block): velocity Verlet
  const int *n_site_items;
  const int2 *tile_items;  // (replica, tile block): rho + df and force
  const int *n_tile_items;
  const int2 *ext_items;   // (replica, block of 256 ghost sites): ghost refresh
  const int *n_ext_items;
  const int *reps;         // replica indices for per-replica reductions
  const int *n_reps;
};

// blocks per launch: a multiple of the 8 XCDs, one block per item
inline int batch_blocks(int64_t items) { return (int)((std::max<int64_t>(items, 1) + 7) / 8 * 8); }

// The launchers take the number of items their grid covers, one block per
// item (batch_blocks(items) blocks): at least the count of the list they
// walk, which the kernels read from device memory.  Serving step launches
// pass the lists' capacities, so one captured graph serves every mix that
// fits them.

hipError_t launch_batch_kick(hipStream_t st, const BatchLists &L, int64_t items, const Masses &M, double h, double dt,
                             bool drift, bool want_ke);
// a step's closing half kick fused with the next step's opening kick and drift
hipError_t launch_batch_kick_fused(hipStream_t st, const BatchLists &L, int64_t items, const Masses &M, double h,
                                   double dt, bool want_ke);
hipError_t launch_batch_ghost(hipStream_t st, const BatchLists &L, int64_t items, uint32_t fields);
// lean (mixed, MD steps): rho and df from the quadratic records; with want_e the embedding energy
// from a second sum over the cubic value records
hipError_t launch_batch_rho_df(hipStream_t st, const BatchLists &L, int64_t items, const Halo &h, const TabF64 &tf,
                               const TabMixed &tm, double rc2, int prec, bool want_e, bool lean);
hipError_t launch_batch_force(hipStream_t st, const BatchLists &L, int64_t items, const Halo &h, const TabF64 &tf,
                              const TabMixed &tm, double rc2, int prec, bool want_e);
hipError_t batch_kernels_allow_lds(size_t bytes);
// per-replica {potential, kinetic} -> energy[2 * which], which = 0 (step) or 1 (admission)
hipError_t launch_batch_energy(hipStream_t st, const BatchLists &L, int which);

}  // namespace mmd

#endif
