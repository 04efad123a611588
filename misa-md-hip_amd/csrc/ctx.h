// The library context (one sub-domain on one device) and the internal entry
// points shared between translation units.
#ifndef MMD_CTX_H
#define MMD_CTX_H

#include <hip/hip_runtime.h>

#include <vector>

#include "device_common.h"
#include "mmd_eam.h"

struct mmd_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  // drop-in passes on pinned records: the copy stream brings the records in
  // by z-chunks (one event per chunk) while `stream` computes
  hipStream_t copy_stream = nullptr;
  std::vector<hipEvent_t> aos_events;
  bool has_domain = false, has_stencil = false, has_tables = false, md_ready = false;
  mmd_domain dom{};
  mmd::Geom g{};

  // device-resident state over the ghost-extended lattice (device layout):
  // pos[i] = (x, y, z, type) so a neighbour costs two 16-byte loads
  double4 *pos = nullptr;
  double *vx = nullptr, *vy = nullptr, *vz = nullptr;
  double *fx = nullptr, *fy = nullptr, *fz = nullptr;
  double *rho = nullptr, *df = nullptr;
  int8_t *type = nullptr;
  double *stage = nullptr;  // 3 * n_ext doubles: reference-layout staging for upload/download

  // stencil: [2][kMaxStencil] LDS offsets within the pair-pass tile, and its halo
  int32_t *d_off = nullptr;
  int32_t nk_even = 0, nk_odd = 0;
  mmd::Halo halo{};
  // Newton's-third-law half stencil of the periodic step: device image
  // (kHalfImage words), forward / backward entries per parity, whether the
  // stencil allows it (symmetric), whether it is on, and the forward entries'
  // pair-force slots ([max nf][n_local])
  int32_t *d_half = nullptr;
  int32_t nf[2] = {0, 0}, nb[2] = {0, 0};
  bool half_ok = false, half = false;
  double *slots = nullptr;

  // tables
  double *d_tab = nullptr, *d_mixed = nullptr;
  mmd::TabF64 tf{};
  mmd::TabMixed tm{};
  mmd::Masses masses{};
  int32_t n_types = 0;
  int32_t max_type = -1;  // largest atom type uploaded (checked against n_types before every pass)
  double r_max = 0.0;
  int32_t prec = MMD_PREC_FP64;

  // energy partials (one per site block) and the per-step record
  double *e_rho = nullptr, *e_force = nullptr, *e_ke = nullptr;
  double *d_energy = nullptr;
  int64_t energy_cap = 0;
  int32_t *d_counter = nullptr;

  // MD
  mmd_md_params md{};
  // captured steps, by (opening kick fused with the previous step's closing
  // one, that step's energy recorded, this step's energy recorded)
  hipGraphExec_t graph[8] = {};
  bool capturing = false;

  // drop-in AoS staging
  char *aos = nullptr;
  size_t aos_bytes = 0;
  int32_t *d_maxtype = nullptr;  // largest type the last drop-in unpack read, in kMaxTypeSlots slots

  // timing
  struct Pending {
    int cls;
    hipEvent_t start, stop;
  };
  bool timing = false;
  std::vector<hipEvent_t> event_pool;
  std::vector<Pending> pending;
  double t_ms[MMD_K_COUNT] = {};
  int64_t t_n[MMD_K_COUNT] = {};

  // interstitials (off-lattice atoms), see itl.hip
  int32_t n_itl = 0, itl_cap = 0, n_wide = 0;
  int32_t itl_max_type = -1;  // largest interstitial type (checked with max_type)
  double *ix = nullptr, *iy = nullptr, *iz = nullptr;
  double *ivx = nullptr, *ivy = nullptr, *ivz = nullptr;
  double *ifx = nullptr, *ify = nullptr, *ifz = nullptr;
  double *irho = nullptr, *idf = nullptr;
  int8_t *itype = nullptr;
  int32_t *d_wide = nullptr;  // [n_wide][3] wide stencil W, parity 0
  double *e_itl = nullptr;    // [itl_part_n] energy partials, then kinetic partials
  int64_t itl_part_n = 0, itl_kick_n = 0;
  bool itl_stale = true;      // anchors / neighbour lists predate the current positions
  void *itl_state = nullptr;  // device work arrays (itl.hip)
};

namespace mmd {

// device geometry of a sub-domain (shared with the serving loop)
Geom make_geom(const mmd_domain &d);
// device-layout offset of stencil entry (dX doubled, dY, dZ) for centre parity p
int32_t stencil_dev_offset(int p, int64_t dX, int64_t dY, int64_t dZ, const mmd_domain &d);
// stencil reach and LDS tile of stencils st3[p] (nk[p] entries (dX doubled, dY, dZ) per parity)
Halo make_halo(const int32_t *st3[2], const int32_t nk[2]);
// LDS offset of stencil entry (dX doubled, dY, dZ) for centre parity p within a tile of halo h
int32_t stencil_lds_offset(int p, int64_t dX, int64_t dY, int64_t dZ, const Halo &h);
// the stencil's device image for the pair passes (layout at kStencilImage)
std::vector<int32_t> stencil_image(const int32_t *st3[2], const int32_t nk[2], const Halo &h);
// the half stencil's device image (layout at kHalfImage) and its forward /
// backward entry counts; empty when the stencil is not symmetric
std::vector<int32_t> half_stencil_image(const int32_t *st3[2], const int32_t nk[2], const Halo &h, int32_t nf[2],
                                        int32_t nb[2]);
// fails with MMD_E_RANGE when the tile of h does not fit the device's LDS
int check_tile_lds(int device, const Halo &h);

int ctx_bind(mmd_ctx *c);
void ctx_drop_graphs(mmd_ctx *c);
int timed_begin(mmd_ctx *c, int cls, hipEvent_t *ev);
int timed_end(mmd_ctx *c, int cls, hipEvent_t start);
int run_rho_df(mmd_ctx *c, double cutoff, bool want_e, bool lean);
// periodic: the caller is the periodic single-domain step (the ghost shell
// holds the local sites' images), where the half stencil may be used
int run_force(mmd_ctx *c, double cutoff, bool want_e, bool periodic = false);
int run_ghost(mmd_ctx *c, uint32_t fields);

// interstitial passes (itl.hip): added on top of the lattice passes
int itl_rho_df(mmd_ctx *c, double cutoff, bool want_e);
int itl_df(mmd_ctx *c);
int itl_force(mmd_ctx *c, double cutoff, bool want_e);
int itl_kick(mmd_ctx *c, double h, bool drift, bool want_ke);
int itl_kick_fused(mmd_ctx *c, double h, bool want_ke);
void itl_free(mmd_ctx *c);
int itl_check(mmd_ctx *c);
// sizes the interstitial bins and energy partials for the MD box (no-op without a list)
int itl_setup(mmd_ctx *c);
// the interstitial bins are sized for the current MD box (the passes then allocate nothing)
bool itl_bins_ready(mmd_ctx *c);

}  // namespace mmd

#endif
