// C ABI of the device-resident EAM path: context, domain, stencil, tables,
// data movement, the three hook passes, the drop-in host-array forms, the MD
// step and kernel timing.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <limits>
#include <mutex>
#include <new>
#include <utility>
#include <vector>

#include "mmd_eam.h"
#include "ctx.h"
#include "host_helpers.h"
#include "kernels.h"

using namespace mmd;

#define HIP_TRY(expr)                                                                                       \
  do {                                                                                                      \
    hipError_t e_ = (expr);                                                                                 \
    if (e_ != hipSuccess)                                                                                   \
      return fail(MMD_E_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); \
  } while (0)

#define GUARD(...)                                                         \
  try {                                                                    \
    __VA_ARGS__                                                            \
  } catch (const std::bad_alloc &) {                                       \
    return fail(MMD_E_NOMEM, "%s: host allocation failed", __func__);      \
  } catch (...) {                                                          \
    return fail(MMD_E_INVALID, "%s: unexpected C++ exception", __func__);  \
  }

namespace mmd {

static int64_t floor_div2(int64_t v) { return v >= 0 ? v / 2 : -((-v + 1) / 2); }

Geom make_geom(const mmd_domain &d) {
  Geom g{};
  g.ext_x = d.ext[0], g.ext_y = d.ext[1], g.ext_z = d.ext[2];
  g.exh = d.ext[0] / 2;
  g.gxh = d.ghost[0] / 2, g.gy = d.ghost[1], g.gz = d.ghost[2];
  g.bxh = d.box[0] / 2, g.by = d.box[1], g.bz = d.box[2];
  g.n_ext = (int64_t)d.ext[0] * d.ext[1] * d.ext[2];
  g.n_local = (int64_t)2 * g.bxh * g.by * g.bz;
  g.n_ghost = g.n_ext - g.n_local;
  g.ntx = (g.bxh + kTileX - 1) / kTileX;
  g.nty = (g.by + kTileY - 1) / kTileY;
  g.ntz = (g.bz + kTileZ - 1) / kTileZ;
  g.n_tiles = (int64_t)g.ntx * g.nty * g.ntz;
  return g;
}

Halo make_halo(const int32_t *st3[2], const int32_t nk[2]) {
  Halo h{};
  for (int p = 0; p < 2; p++)
    for (int32_t k = 0; k < nk[p]; k++) {
      const int32_t *e = st3[p] + 3 * (size_t)k;
      h.rx = std::max<int32_t>(h.rx, (int32_t)std::llabs(floor_div2(p + e[0])));
      h.ry = std::max<int32_t>(h.ry, std::abs(e[1]));
      h.rz = std::max<int32_t>(h.rz, std::abs(e[2]));
    }
  h.lx = kTileX + 2 * h.rx, h.ly = kTileY + 2 * h.ry, h.lz = kTileZ + 2 * h.rz;
  h.n = h.lx * h.ly * h.lz * 2;
  return h;
}

int32_t stencil_lds_offset(int p, int64_t dX, int64_t dY, int64_t dZ, const Halo &h) {
  const int64_t s1 = (p + dX) & 1;
  return (int32_t)(((dZ * h.ly + dY) * 2 + (s1 - p)) * h.lx + floor_div2(p + dX));
}

// Ideal displacement of stencil entry (dX, dY, dZ) from a parity-p centre, in
// units of a/2: the neighbour has parity p ^ (dX & 1), and a parity-1 site
// sits at (a/2)(1,1,1) from its cell's corner, so y and z also shift by the
// parity difference.
static void ideal_disp2(int p, int64_t dX, int64_t dY, int64_t dZ, int64_t out[3]) {
  const int64_t ds = ((p + dX) & 1) - p;  // parity of the neighbour minus the centre's
  out[0] = dX, out[1] = 2 * dY + ds, out[2] = 2 * dZ + ds;
}

std::vector<int32_t> stencil_image(const int32_t *st3[2], const int32_t nk[2], const Halo &h) {
  std::vector<int32_t> img(kStencilImage, 0);
  std::vector<std::array<int64_t, 3>> disp[2];
  for (int p = 0; p < 2; p++)
    for (int32_t k = 0; k < nk[p]; k++) {
      const int32_t *e = &st3[p][3 * (size_t)k];
      img[(size_t)p * kMaxStencil + k] = stencil_lds_offset(p, e[0], e[1], e[2], h);
      std::array<int64_t, 3> d;
      ideal_disp2(p, e[0], e[1], e[2], d.data());
      disp[p].push_back(d);
    }
  // symmetric: the neighbour reached by every entry sees the centre through
  // an entry of its own parity with the opposite displacement
  bool sym = true;
  for (int p = 0; p < 2 && sym; p++)
    for (int32_t k = 0; k < nk[p] && sym; k++) {
      const int q = (p + st3[p][3 * (size_t)k]) & 1;
      const std::array<int64_t, 3> neg = {-disp[p][k][0], -disp[p][k][1], -disp[p][k][2]};
      sym = std::find(disp[q].begin(), disp[q].end(), neg) != disp[q].end();
    }
  for (int p = 0; p < 2; p++)
    for (int32_t k = 0; k < nk[p]; k++) {
      const auto &d = disp[p][k];
      // lexicographically positive in (z, y, x): the side that counts the pair
      const bool first = d[2] > 0 || (d[2] == 0 && (d[1] > 0 || (d[1] == 0 && d[0] > 0)));
      if (!sym || first) img[kStencilBits + (size_t)p * (kMaxStencil / 32) + k / 32] |= (int32_t)(1u << (k % 32));
    }
  img[kStencilHalfFlag] = sym ? 1 : 0;
  return img;
}

std::vector<int32_t> half_stencil_image(const int32_t *st3[2], const int32_t nk[2], const Halo &h, int32_t nf[2],
                                        int32_t nb[2]) {
  std::vector<std::array<int64_t, 3>> disp[2];
  std::vector<int32_t> fidx[2];  // forward entry number, -1 for a backward entry
  for (int p = 0; p < 2; p++) {
    nf[p] = nb[p] = 0;
    for (int32_t k = 0; k < nk[p]; k++) {
      const int32_t *e = &st3[p][3 * (size_t)k];
      std::array<int64_t, 3> d;
      ideal_disp2(p, e[0], e[1], e[2], d.data());
      disp[p].push_back(d);
      const bool fwd = d[2] > 0 || (d[2] == 0 && (d[1] > 0 || (d[1] == 0 && d[0] > 0)));
      fidx[p].push_back(fwd ? nf[p]++ : -1);
    }
  }
  std::vector<int32_t> img(kHalfImage, 0);
  for (int p = 0; p < 2; p++)
    for (int32_t k = 0; k < nk[p]; k++) {
      const int32_t *e = &st3[p][3 * (size_t)k];
      const int32_t lds = stencil_lds_offset(p, e[0], e[1], e[2], h);
      if (fidx[p][k] >= 0) {
        img[kHalfFwd + (size_t)p * kMaxStencil + fidx[p][k]] = lds;
        continue;
      }
      // the neighbour (parity q) reaches back through its entry of opposite
      // displacement, which is forward
      const int q = (p + e[0]) & 1;
      const std::array<int64_t, 3> neg = {-disp[p][k][0], -disp[p][k][1], -disp[p][k][2]};
      const auto it = std::find(disp[q].begin(), disp[q].end(), neg);
      if (it == disp[q].end()) return {};
      const int32_t back = fidx[q][(size_t)(it - disp[q].begin())];
      if (back < 0) return {};
      img[kHalfBack + (size_t)p * kMaxStencil + nb[p]] = lds;
      img[kHalfSlot + (size_t)p * kMaxStencil + nb[p]] = back;
      nb[p]++;
    }
  return img;
}

int check_tile_lds(int device, const Halo &h) {
  const size_t need = tile_lds_bytes(h, true);
  int lim = 0;
  HIP_TRY(hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, device));
  if (need > (size_t)lim)
    return fail(MMD_E_RANGE, "stencil reach (%d, %d, %d) needs %zu bytes of LDS per tile, the device allows %d",
                h.rx, h.ry, h.rz, need, lim);
  if (need > 65536) {
    HIP_TRY(pair_kernels_allow_lds(need));
    HIP_TRY(half_kernels_allow_lds(std::max(need, half_back_lds_bytes(h))));
  }
  return MMD_OK;
}

int32_t stencil_dev_offset(int p, int64_t dX, int64_t dY, int64_t dZ, const mmd_domain &d) {
  const int64_t s1 = (p + dX) & 1;
  return (int32_t)(((dZ * d.ext[1] + dY) * 2 + (s1 - p)) * (d.ext[0] / 2) + floor_div2(p + dX));
}

int ctx_bind(mmd_ctx *c) {
  if (!c) return fail(MMD_E_INVALID, "null context");
  HIP_TRY(hipSetDevice(c->device));
  return MMD_OK;
}

template <typename T>
static int dalloc(T **p, size_t n) {
  *p = nullptr;
  if (n == 0) return MMD_OK;
  hipError_t e = hipMalloc((void **)p, n * sizeof(T));
  if (e != hipSuccess) {
    *p = nullptr;
    (void)hipGetLastError();
    return fail(MMD_E_NOMEM, "hipMalloc of %zu bytes failed: %s", n * sizeof(T), hipGetErrorString(e));
  }
  return MMD_OK;
}

template <typename T>
static void dfree(T *&p) {
  if (p) (void)hipFree((void *)p);
  p = nullptr;
}

static void free_state(mmd_ctx *c) {
  dfree(c->pos), dfree(c->vx), dfree(c->vy), dfree(c->vz);
  dfree(c->fx), dfree(c->fy), dfree(c->fz), dfree(c->rho), dfree(c->df), dfree(c->type);
  dfree(c->stage), dfree(c->e_rho), dfree(c->e_force), dfree(c->e_ke), dfree(c->e_itl);
  dfree(c->aos), c->aos_bytes = 0;
  dfree(c->slots), c->half = false;
}

void ctx_drop_graphs(mmd_ctx *c) {
  for (auto &g : c->graph)
    if (g) {
      (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
}

// --------------------------------------------------------------- timing

int timed_begin(mmd_ctx *c, int cls, hipEvent_t *ev) {
  *ev = nullptr;
  if (!c->timing || c->capturing) return MMD_OK;
  if (c->event_pool.empty()) {
    hipEvent_t e;
    HIP_TRY(hipEventCreate(&e));
    c->event_pool.push_back(e);
  }
  *ev = c->event_pool.back();
  c->event_pool.pop_back();
  HIP_TRY(hipEventRecord(*ev, c->stream));
  (void)cls;
  return MMD_OK;
}

int timed_end(mmd_ctx *c, int cls, hipEvent_t start) {
  if (!start) return MMD_OK;
  if (c->event_pool.empty()) {
    hipEvent_t e;
    HIP_TRY(hipEventCreate(&e));
    c->event_pool.push_back(e);
  }
  hipEvent_t stop = c->event_pool.back();
  c->event_pool.pop_back();
  HIP_TRY(hipEventRecord(stop, c->stream));
  c->pending.push_back({cls, start, stop});
  return MMD_OK;
}

static int timing_collect(mmd_ctx *c) {
  HIP_TRY(hipStreamSynchronize(c->stream));
  for (auto &p : c->pending) {
    float ms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, p.start, p.stop));
    c->t_ms[p.cls] += ms;
    c->t_n[p.cls] += 1;
    c->event_pool.push_back(p.start);
    c->event_pool.push_back(p.stop);
  }
  c->pending.clear();
  return MMD_OK;
}

// ---------------------------------------------------------- pass helpers

int run_rho_df(mmd_ctx *c, double cutoff, bool want_e, bool lean) {
  hipEvent_t ev;
  int rc = timed_begin(c, MMD_K_RHO, &ev);
  if (rc) return rc;
  HIP_TRY(launch_rho_df(c->stream, c->g, c->halo, c->d_off, c->nk_even, c->nk_odd, c->pos, c->rho, c->df, c->tf, c->tm,
                        cutoff * cutoff, c->prec, want_e ? c->e_rho : nullptr, lean));
  rc = timed_end(c, MMD_K_RHO, ev);
  if (rc) return rc;
  if (c->n_itl > 0) return itl_rho_df(c, cutoff, want_e);
  return MMD_OK;
}

int run_force(mmd_ctx *c, double cutoff, bool want_e, bool periodic) {
  hipEvent_t ev;
  int rc = timed_begin(c, MMD_K_FORCE, &ev);
  if (rc) return rc;
  if (periodic && c->half)
    HIP_TRY(launch_force_half(c->stream, c->g, c->halo, c->d_half, c->nf, c->nb, c->pos, c->df, c->fx, c->fy, c->fz,
                              c->slots, c->tf, c->tm, cutoff * cutoff, c->prec, want_e ? c->e_force : nullptr));
  else
    HIP_TRY(launch_force(c->stream, c->g, c->halo, c->d_off, c->nk_even, c->nk_odd, c->pos, c->df, c->fx, c->fy, c->fz,
                         c->tf, c->tm, cutoff * cutoff, c->prec, want_e ? c->e_force : nullptr));
  rc = timed_end(c, MMD_K_FORCE, ev);
  if (rc) return rc;
  if (c->n_itl > 0) return itl_force(c, cutoff, want_e);
  return MMD_OK;
}

int run_ghost(mmd_ctx *c, uint32_t fields) {
  hipEvent_t ev;
  int rc = timed_begin(c, MMD_K_GHOST, &ev);
  if (rc) return rc;
  HIP_TRY(launch_ghost(c->stream, c->g, c->md.box_len[0], c->md.box_len[1], c->md.box_len[2], c->pos, c->type,
                       c->df, fields));
  return timed_end(c, MMD_K_GHOST, ev);
}

static int run_kick(mmd_ctx *c, double h, bool drift, bool want_ke) {
  hipEvent_t ev;
  int rc = timed_begin(c, MMD_K_INTEGRATE, &ev);
  if (rc) return rc;
  HIP_TRY(launch_kick(c->stream, c->g, c->masses, h, c->md.dt, drift, c->pos, c->vx, c->vy, c->vz, c->fx, c->fy,
                      c->fz, c->type, want_ke ? c->e_ke : nullptr));
  rc = timed_end(c, MMD_K_INTEGRATE, ev);
  if (rc) return rc;
  if (c->n_itl > 0) return itl_kick(c, h, drift, want_ke);
  return MMD_OK;
}

static int run_kick_fused(mmd_ctx *c, double h, bool want_ke) {
  hipEvent_t ev;
  int rc = timed_begin(c, MMD_K_INTEGRATE, &ev);
  if (rc) return rc;
  HIP_TRY(launch_kick_fused(c->stream, c->g, c->masses, h, c->md.dt, c->pos, c->vx, c->vy, c->vz, c->fx, c->fy, c->fz,
                            c->type, want_ke ? c->e_ke : nullptr));
  rc = timed_end(c, MMD_K_INTEGRATE, ev);
  if (rc) return rc;
  if (c->n_itl > 0) return itl_kick_fused(c, h, want_ke);
  return MMD_OK;
}

static int run_finalize(mmd_ctx *c, double *out, int32_t *counter) {
  const int64_t nt = tile_blocks(c->g), nb = site_blocks(c->g);
  const bool itl = c->n_itl > 0;
  HIP_TRY(launch_energy_finalize(c->stream, c->e_rho, nt, c->e_force, nt, itl ? c->e_itl : nullptr,
                                 itl ? c->itl_part_n : 0, c->e_ke, nb, itl ? c->e_itl + c->itl_part_n : nullptr,
                                 itl ? c->itl_kick_n : 0, out, counter));
  return MMD_OK;
}

// One velocity-Verlet step up to its force pass: the opening half kick and
// drift (fused with the previous step's closing half kick when `fused`, which
// then also records that step's energy when prev_e), ghost x, rho + df, ghost
// df, force (with the energy partials when want_e).  The step's closing half
// kick comes with the next step, or from md_close at the end of a run.
// Energies go to d_energy[(*d_counter)++].
static int md_step(mmd_ctx *c, bool fused, bool prev_e, bool want_e) {
  int rc;
  const double h = 0.5 * c->md.dt;
  if (fused) {
    if ((rc = run_kick_fused(c, h, prev_e))) return rc;
    if (prev_e && (rc = run_finalize(c, c->d_energy, c->d_counter))) return rc;
  } else if ((rc = run_kick(c, h, true, false))) {
    return rc;
  }
  if ((rc = run_ghost(c, MMD_F_X))) return rc;
  // MD dynamics take rho and df from the lean records on every step (so a
  // trajectory does not depend on which steps record energies); energy steps
  // also sum the cubic records for the embedding energy
  if ((rc = run_rho_df(c, c->md.cutoff, want_e, true))) return rc;
  if ((rc = run_ghost(c, MMD_F_DF))) return rc;
  return run_force(c, c->md.cutoff, want_e, true);
}

// the last step's closing half kick (and its energy record)
static int md_close(mmd_ctx *c, bool want_e) {
  int rc;
  if ((rc = run_kick(c, 0.5 * c->md.dt, false, want_e))) return rc;
  if (want_e && (rc = run_finalize(c, c->d_energy, c->d_counter))) return rc;
  return MMD_OK;
}

static int ensure_energy_cap(mmd_ctx *c, int64_t n) {
  if (n <= c->energy_cap) return MMD_OK;
  ctx_drop_graphs(c);  // graphs hold the old pointer
  dfree(c->d_energy);
  c->energy_cap = 0;
  int rc = dalloc(&c->d_energy, (size_t)(2 * n));
  if (rc) return rc;
  c->energy_cap = n;
  return MMD_OK;
}

// ------------------------------------------- allocation-free MD run scratch
//
// mmd_md_run_noalloc keeps the per-block energy partials of the lattice and
// interstitial passes, the energy records and the record counter of one call
// in this array instead of the context's own buffers.  Every GPU holds its
// own copy for the life of the process, shared by all calls on that GPU; a
// call clears what it uses first, and calls that overlap on one GPU race.
// Host code never passes `scratch` itself to a kernel or a copy (its host-side
// name is not a device address): it uses the per-device address that
// hipGetSymbolAddress returned when the first context on that device was
// created.
constexpr size_t kScratchDoubles = (size_t)1 << 20;  // 8 MiB per GPU
constexpr int kScratchDevices = 64;
static __device__ __align__(256) double scratch[kScratchDoubles];
static std::mutex g_scratch_mu;
static double *g_scratch[kScratchDevices] = {};

// the calling thread's current device must be `device`
static int scratch_bind(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (device < 0 || device >= kScratchDevices)
    return fail(MMD_E_RANGE, "device %d beyond the %d the library's scratch table covers", device, kScratchDevices);
  if (g_scratch[device]) return MMD_OK;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return fail(MMD_E_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
  g_scratch[device] = static_cast<double *>(p);
  return MMD_OK;
}

static double *scratch_of(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return device >= 0 && device < kScratchDevices ? g_scratch[device] : nullptr;
}

// Points the context's energy partials, records and counter at the scratch
// array for the duration of one call (and turns event timing off, whose
// events are created on demand); restores the context's own on exit.
struct ScratchUse {
  mmd_ctx *c;
  double *e_rho, *e_force, *e_ke, *e_itl, *d_energy;
  int32_t *d_counter;
  bool timing;
  ScratchUse(mmd_ctx *cc, double *S, int64_t nt, int64_t nb, int64_t ni, int64_t n_rec)
      : c(cc), e_rho(cc->e_rho), e_force(cc->e_force), e_ke(cc->e_ke), e_itl(cc->e_itl), d_energy(cc->d_energy),
        d_counter(cc->d_counter), timing(cc->timing) {
    c->e_rho = S;
    c->e_force = S + nt;
    c->e_ke = S + 2 * nt;
    c->e_itl = ni > 0 ? S + 2 * nt + nb : c->e_itl;
    c->d_energy = S + 2 * nt + nb + ni;
    c->d_counter = reinterpret_cast<int32_t *>(c->d_energy + 2 * n_rec);
    c->timing = false;
  }
  ~ScratchUse() {
    c->e_rho = e_rho, c->e_force = e_force, c->e_ke = e_ke, c->e_itl = e_itl, c->d_energy = d_energy;
    c->d_counter = d_counter, c->timing = timing;
  }
};

}  // namespace mmd

// ================================================================== C ABI

extern "C" int mmd_ctx_create(int32_t device, mmd_ctx **out) {
  GUARD({
    clear_error();
    if (!out) return fail(MMD_E_INVALID, "mmd_ctx_create: null out");
    *out = nullptr;
    int n = 0;
    HIP_TRY(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(MMD_E_INVALID, "mmd_ctx_create: device %d of %d", device, n);
    HIP_TRY(hipSetDevice(device));
    if (int rs = scratch_bind(device)) return rs;
    mmd_ctx *c = new mmd_ctx();
    c->device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
      delete c;
      return fail(MMD_E_HIP, "hipStreamCreate: %s", hipGetErrorString(e));
    }
    int rc = dalloc(&c->d_counter, 4);
    if (rc) {
      (void)hipStreamDestroy(c->stream);
      delete c;
      return rc;
    }
    *out = c;
    return MMD_OK;
  })
}

extern "C" int mmd_device_count(int32_t *n) {
  clear_error();
  if (!n) return fail(MMD_E_INVALID, "mmd_device_count: null out");
  int k = 0;
  HIP_TRY(hipGetDeviceCount(&k));
  *n = k;
  return MMD_OK;
}

extern "C" int mmd_ctx_destroy(mmd_ctx *c) {
  GUARD({
    if (!c) return MMD_OK;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    ctx_drop_graphs(c);
    free_state(c);
    itl_free(c);
    dfree(c->d_off), dfree(c->d_half), dfree(c->d_tab), dfree(c->d_mixed), dfree(c->d_energy), dfree(c->d_counter);
    dfree(c->d_maxtype);
    for (auto &p : c->pending) c->event_pool.push_back(p.start), c->event_pool.push_back(p.stop);
    for (auto e : c->event_pool) (void)hipEventDestroy(e);
    for (auto e : c->aos_events) (void)hipEventDestroy(e);
    if (c->copy_stream) (void)hipStreamDestroy(c->copy_stream);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return MMD_OK;
  })
}

extern "C" int mmd_set_domain(mmd_ctx *c, const mmd_domain *d) {
  GUARD({
    clear_error();
    int rc = ctx_bind(c);
    if (rc) return rc;
    if (!d) return fail(MMD_E_INVALID, "mmd_set_domain: null domain");
    for (int k = 0; k < 3; k++) {
      if (d->box[k] < 1 || d->ghost[k] < 0 || d->ext[k] != d->box[k] + 2 * d->ghost[k])
        return fail(MMD_E_INVALID, "mmd_set_domain: dim %d needs box>=1, ghost>=0, ext=box+2*ghost", k);
      if (d->ghost[k] > d->box[k])
        return fail(MMD_E_INVALID, "mmd_set_domain: ghost wider than the box in dim %d", k);
    }
    if ((d->box[0] | d->ghost[0]) & 1)
      return fail(MMD_E_INVALID, "mmd_set_domain: box and ghost along x must be even (two sites per cell)");
    if (((d->start[0] - d->ghost[0]) % 2 + 2) % 2 != 0)
      return fail(MMD_E_INVALID, "mmd_set_domain: start_x - ghost_x must be even");
    const int64_t n_ext = (int64_t)d->ext[0] * d->ext[1] * d->ext[2];
    if (n_ext >= ((int64_t)1 << 31)) return fail(MMD_E_RANGE, "mmd_set_domain: %lld ext sites exceed 2^31", (long long)n_ext);
    HIP_TRY(hipStreamSynchronize(c->stream));
    ctx_drop_graphs(c);
    free_state(c);
    // the MD parameters (box lengths, cutoff) belong to the old domain: the
    // step needs mmd_md_setup again, whether or not the new state allocates
    c->has_domain = false;
    c->has_stencil = false;
    c->md_ready = false;
    c->n_itl = 0;
    c->g = make_geom(*d);
    const Geom &g = c->g;
    c->dom = *d;
    const size_t n = (size_t)n_ext;
    double **f64[] = {&c->vx, &c->vy, &c->vz, &c->fx, &c->fy, &c->fz, &c->rho, &c->df};
    for (double **p : f64) {
      if ((rc = dalloc(p, n))) {
        free_state(c);
        return rc;
      }
      HIP_TRY(launch_fill_f64(c->stream, *p, (int64_t)n, 0.0));
    }
    if ((rc = dalloc(&c->stage, 3 * n)) || (rc = dalloc(&c->pos, n))) {
      free_state(c);
      return rc;
    }
    HIP_TRY(launch_fill_f64(c->stream, reinterpret_cast<double *>(c->pos), (int64_t)(4 * n), 0.0));
    if ((rc = dalloc(&c->type, n))) {
      free_state(c);
      return rc;
    }
    HIP_TRY(launch_fill_i8(c->stream, c->type, (int64_t)n, (int8_t)-1));
    HIP_TRY(launch_ref2dev_pos(c->stream, c->g, nullptr, nullptr, nullptr, c->type, c->pos, false));
    const size_t nb = (size_t)site_blocks(g), nt = (size_t)tile_blocks(g);
    if ((rc = dalloc(&c->e_rho, nt)) || (rc = dalloc(&c->e_force, nt)) || (rc = dalloc(&c->e_ke, nb))) {
      free_state(c);
      return rc;
    }
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->has_domain = true;
    return MMD_OK;
  })
}

static int64_t centred_mod(int64_t v, int64_t n) {
  int64_t r = ((v % n) + n) % n;
  if (r >= n / 2) r -= n;
  return r;
}

extern "C" int mmd_set_stencil(mmd_ctx *c, const int64_t *even, int32_t n_even, const int64_t *odd, int32_t n_odd) {
  GUARD({
    clear_error();
    int rc = ctx_bind(c);
    if (rc) return rc;
    if (!c->has_domain) return fail(MMD_E_STATE, "mmd_set_stencil: set the domain first");
    // rows are padded with zero offsets so unrolled pair loops may read past nk
    const int32_t cap = kMaxStencil - 8;
    if (n_even < 0 || n_odd < 0 || n_even > cap || n_odd > cap || (n_even && !even) || (n_odd && !odd))
      return fail(MMD_E_RANGE, "mmd_set_stencil: sizes must be in [0, %d]", cap);
    const mmd_domain &d = c->dom;
    // decompose the reference linear offsets into (dX doubled, dY, dZ)
    std::vector<int32_t> e3[2];
    for (int p = 0; p < 2; p++) {
      const int64_t *src = p ? odd : even;
      const int32_t n = p ? n_odd : n_even;
      e3[p].resize(3 * (size_t)n);
      for (int32_t k = 0; k < n; k++) {
        const int64_t o = src[k];
        const int64_t dX = centred_mod(o, d.ext[0]);
        const int64_t o1 = (o - dX) / d.ext[0];
        const int64_t dY = centred_mod(o1, d.ext[1]);
        const int64_t dZ = (o1 - dY) / d.ext[1];
        if (std::llabs(dX) > d.ghost[0] || std::llabs(dY) > d.ghost[1] || std::llabs(dZ) > d.ghost[2] ||
            (!dX && !dY && !dZ))
          return fail(MMD_E_RANGE, "mmd_set_stencil: parity %d entry %d (offset %lld) leaves the ghost shell", p, k,
                      (long long)o);
        e3[p][3 * (size_t)k] = (int32_t)dX, e3[p][3 * (size_t)k + 1] = (int32_t)dY, e3[p][3 * (size_t)k + 2] = (int32_t)dZ;
      }
    }
    // the pair passes address neighbours inside the LDS tile
    const int32_t *st3[2] = {e3[0].data(), e3[1].data()};
    const int32_t nk[2] = {n_even, n_odd};
    const Halo h = make_halo(st3, nk);
    if ((rc = check_tile_lds(c->device, h))) return rc;
    const std::vector<int32_t> host = stencil_image(st3, nk, h);
    int32_t nf[2], nb[2];
    const std::vector<int32_t> half = half_stencil_image(st3, nk, h, nf, nb);
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (!c->d_off && (rc = dalloc(&c->d_off, kStencilImage))) return rc;
    if (!c->d_half && (rc = dalloc(&c->d_half, kHalfImage))) return rc;
    // from here on the context holds the new stencil, or none if a copy fails;
    // a new stencil turns the half stencil off (its slots follow the stencil)
    c->has_stencil = false;
    ctx_drop_graphs(c);
    dfree(c->slots), c->half = false;
    HIP_TRY(hipMemcpy(c->d_off, host.data(), host.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    // every entry needs its reverse at the neighbour's parity, as for the
    // pair energy's half masks (a one-sided pair would be counted differently)
    c->half_ok = !half.empty() && host[kStencilHalfFlag] == 1;
    if (c->half_ok) {
      HIP_TRY(hipMemcpy(c->d_half, half.data(), half.size() * sizeof(int32_t), hipMemcpyHostToDevice));
      c->nf[0] = nf[0], c->nf[1] = nf[1], c->nb[0] = nb[0], c->nb[1] = nb[1];
    }
    c->halo = h;
    c->nk_even = n_even;
    c->nk_odd = n_odd;
    c->has_stencil = true;
    return MMD_OK;
  })
}

// Host image of the device tables: every array 256-byte aligned in one buffer.
struct Arena {
  std::vector<char> buf;
  size_t put(const void *p, size_t bytes) {
    const size_t off = (buf.size() + 255) & ~(size_t)255;
    buf.resize(off + bytes);
    std::memcpy(buf.data() + off, p, bytes);
    return off;
  }
};

struct SplHost {
  size_t val, d01, d2;
};

// FP64 spline family: spline7 coefficients split by width (see SplF64).
static SplHost pack_f64(Arena &A, const double *knots, int nfun, int n, double dx) {
  std::vector<double> c7((size_t)n * 7), val((size_t)nfun * n * 4), d01((size_t)nfun * n * 2),
      d2((size_t)nfun * n);
  for (int f = 0; f < nfun; f++) {
    spline7(knots + (size_t)f * n, n, dx, c7.data());
    for (int m = 0; m < n; m++) {
      const double *k = &c7[(size_t)m * 7];
      const size_t r = (size_t)f * n + m;
      val[4 * r] = k[3], val[4 * r + 1] = k[4], val[4 * r + 2] = k[5], val[4 * r + 3] = k[6];
      d01[2 * r] = k[0], d01[2 * r + 1] = k[1];
      d2[r] = k[2];
    }
  }
  SplHost h;
  h.val = A.put(val.data(), val.size() * 8);
  h.d01 = A.put(d01.data(), d01.size() * 8);
  h.d2 = A.put(d2.data(), d2.size() * 8);
  return h;
}

// 16-byte derivative record: FP64 base, FP32 linear and quadratic terms
struct DerivRec {
  double base;
  float b1, b2;
};
static_assert(sizeof(DerivRec) == 16, "derivative record must be one 16-byte load");

struct MixHost {
  std::vector<double> vb;  // (base, a1) per record
  std::vector<float> vc;   // (a2, a3) per record
  std::vector<DerivRec> dp;
  std::vector<char> ok;  // per segment: every function of the family meets the accuracy bounds
};


// Accuracy bounds of the r^2-keyed tables, relative error against the FP64
// r-spline they resample, per pair term: values 1e-9; the pair-force term
// phi' 2e-8 (the force on an atom cancels ~50x over its ~60 neighbours, so
// the per-atom force error stays well inside 1e-6); rho' enters the force
// multiplied by F'(rho), so its bound is 2e-8 / max|F'|.
// Each bound is relative to |f| + kMixedFloor * max|f| (max over the table):
// near a zero of a function or of its derivative the relative bound alone
// would ask more than FP32 correction terms resolve (their absolute error
// there is ~1e-10), while the absolute error the floor admits per pair is
// orders below the per-atom force target.
constexpr double kMixedTolValue = 1e-9;
constexpr double kMixedTolForce = 2e-8;
constexpr double kMixedFloor = 3e-2;
// The lean rho records (passes that compute no energy) meet 2e-8 on rho: rho
// reaches those passes' forces only through df = F'(rho), whose error is
// F''(rho) times rho's, orders below the force bound; energies, which take
// F(rho) itself, are computed from the 1e-9 value records.
constexpr double kMixedTolLean = 2e-8;

// Where the mixed tables are used: the segments from the first to the last
// good one, minus at most kMixedHoles interior runs of failing segments (runs
// separated by the fewest good segments are merged until that many remain;
// the merged good segments then use the FP64 spline too).  The first and the
// last segment never count as good: lookups below the table, beyond it or
// clamped to its ends land on one of them.  Returns the per-segment verdict.
static std::vector<char> mixed_ranges(const std::vector<char> &ok_a, const std::vector<char> &ok_b, int n_s, double s0,
                                      double ds, TabMixed &tm) {
  auto good = [&](int m) { return m > 0 && m < n_s - 2 && ok_a[(size_t)m] && ok_b[(size_t)m]; };
  int lo = 0, hi = n_s - 1;
  while (lo < hi && !good(lo)) lo++;
  while (hi > lo && !good(hi - 1)) hi--;
  std::vector<std::pair<int, int>> runs;  // [first, last + 1) failing runs inside [lo, hi)
  for (int m = lo; m < hi; m++) {
    if (good(m)) continue;
    if (!runs.empty() && runs.back().second == m)
      runs.back().second = m + 1;
    else
      runs.push_back({m, m + 1});
  }
  while ((int)runs.size() > kMixedHoles) {
    size_t best = 0;
    for (size_t k = 1; k + 1 < runs.size(); k++)
      if (runs[k + 1].first - runs[k].second < runs[best + 1].first - runs[best].second) best = k;
    runs[best].second = runs[best + 1].second;
    runs.erase(runs.begin() + (long)best + 1);
  }
  tm.s_head = s0 + ds * (double)lo;
  tm.s_tail = s0 + ds * (double)hi;
  std::vector<char> use((size_t)n_s, 0);
  for (int m = lo; m < hi; m++) use[(size_t)m] = 1;
  for (int k = 0; k < kMixedHoles; k++) {
    const bool used = k < (int)runs.size();
    tm.hole_lo[k] = used ? s0 + ds * (double)runs[(size_t)k].first : 0.0;
    tm.hole_hi[k] = used ? s0 + ds * (double)runs[(size_t)k].second : 0.0;
    if (used)
      for (int m = runs[(size_t)k].first; m < runs[(size_t)k].second; m++) use[(size_t)m] = 0;
  }
  return use;
}

// Marks the segments the mixed tables do not serve: the FP32 cubic term of
// the value record and the quadratic term of the derivative record become
// NaN, so any evaluation there yields NaN and the pair passes take the FP64
// spline for it (a NaN test instead of comparing s against the ranges).
static void mixed_flag(MixHost &h, int nfun, int n_s, const std::vector<char> &use) {
  const float nan = std::numeric_limits<float>::quiet_NaN();
  for (int f = 0; f < nfun; f++)
    for (int m = 0; m < n_s; m++)
      if (!use[(size_t)m]) {
        const size_t r = (size_t)f * n_s + m;
        h.vc[2 * r + 1] = nan;
        h.dp[r].b2 = nan;
      }
}

// Mixed-precision r^2-keyed family resampled from the FP64 r-keyed spline:
// knot values and exact derivatives of g(sqrt(s)) at uniform s, cubic Hermite
// in s; FP64 base and linear terms, FP32 quadratic and cubic terms.
static MixHost pack_mixed(const double *knots, int nfun, int n_r, double r0, double dr, int n_s, double s0, double ds,
                          double tol_deriv) {
  std::vector<double> c7((size_t)n_r * 7), h(n_s), hp(n_s), vb((size_t)nfun * n_s * 2);
  std::vector<float> vc((size_t)nfun * n_s * 2);
  std::vector<DerivRec> dp((size_t)nfun * n_s);
  const double inv_dr = 1.0 / dr;
  for (int f = 0; f < nfun; f++) {
    spline7(knots + (size_t)f * n_r, n_r, dr, c7.data());
    for (int i = 0; i < n_s; i++) {
      const double s = s0 + ds * (double)i;
      const double r = std::sqrt(s);
      h[i] = spline7_value(c7.data(), n_r, r0, inv_dr, r);
      hp[i] = spline7_deriv(c7.data(), n_r, r0, inv_dr, r) / (2.0 * r);
    }
    for (int m = 0; m < n_s; m++) {
      double a1 = 0, a2 = 0, a3 = 0;
      if (m < n_s - 1) {
        const double D = h[m + 1] - h[m];
        a1 = hp[m] * ds;
        const double a1n = hp[m + 1] * ds;
        a2 = 3.0 * D - 2.0 * a1 - a1n;
        a3 = a1 + a1n - 2.0 * D;
      }
      const size_t r = (size_t)f * n_s + m;
      vb[2 * r] = h[m], vb[2 * r + 1] = a1;
      vc[2 * r] = (float)a2, vc[2 * r + 1] = (float)a3;
      dp[r].base = hp[m];
      dp[r].b1 = (float)(2.0 * a2 / ds);
      dp[r].b2 = (float)(3.0 * a3 / ds);
    }
  }
  // Dense check of every segment against the FP64 spline (the same FP32
  // rounding as the device).  Segments that miss the bounds fall back to the
  // FP64 r-spline (mixed_ranges): near r -> 0 the uniform s grid is coarse in
  // r, the end segments of the r-spline carry low-order kinks, and a function
  // that is only C1 at an interior point has a kink there.
  std::vector<char> ok((size_t)n_s, 1);
  ok[(size_t)n_s - 1] = 0;  // the last knot starts no segment
  for (int f = 0; f < nfun; f++) {
    spline7(knots + (size_t)f * n_r, n_r, dr, c7.data());
    double hmax = 0.0, gmax = 0.0;
    for (int i = 0; i < n_s; i++) {
      const double r = std::sqrt(s0 + ds * (double)i);
      hmax = std::max(hmax, std::fabs(spline7_value(c7.data(), n_r, r0, inv_dr, r)));
      gmax = std::max(gmax, std::fabs(spline7_deriv(c7.data(), n_r, r0, inv_dr, r) / (2.0 * r)));
    }
    for (int m = 0; m < n_s - 1; m++) {
      const size_t rec = (size_t)f * n_s + m;
      for (int k = 1; k < 8; k++) {
        const double q = (double)k / 8.0, sv = s0 + ds * ((double)m + q), r = std::sqrt(sv);
        const double h_ex = spline7_value(c7.data(), n_r, r0, inv_dr, r);
        const double g_ex = spline7_deriv(c7.data(), n_r, r0, inv_dr, r) / (2.0 * r);
        const float c = std::fma((float)q, vc[2 * rec + 1], vc[2 * rec]);
        const double h = std::fma(q, std::fma(q, (double)c, vb[2 * rec + 1]), vb[2 * rec]);
        const float gc = std::fma((float)q, dp[rec].b2, dp[rec].b1);
        const double g = std::fma(q, (double)gc, dp[rec].base);
        if (std::fabs(h - h_ex) > kMixedTolValue * (std::fabs(h_ex) + kMixedFloor * hmax) ||
            std::fabs(g - g_ex) > tol_deriv * (std::fabs(g_ex) + kMixedFloor * gmax))
          ok[(size_t)m] = 0;
      }
    }
  }
  MixHost o;
  o.ok = std::move(ok);
  o.vb = std::move(vb);
  o.vc = std::move(vc);
  o.dp = std::move(dp);
  return o;
}

// Lean rho records: per segment the quadratic through the cubic value record
// c(q) = base + a1 q + a2 q^2 + a3 q^3 at q = 0, 1/2 and 1, i.e. base +
// (a1 - a3/2) q + (a2 + 3 a3/2) q^2, whose distance from the cubic is
// |a3 q (q - 1/2) (q - 1)| <= 0.048 |a3|.  Checked densely against the FP64
// spline with the device's own FP32 steps; a segment the cubic does not serve
// or that misses kMixedTolLean gets a NaN quadratic term (the FP64 spline
// then serves its pairs).  n_miss counts served segments that missed.
static std::vector<DerivRec> pack_lean_rho(const MixHost &mh, const double *knots, int nt, int n_r, double r0,
                                           double dr, int n_s, double s0, double ds, const std::vector<char> &use,
                                           int64_t &n_miss) {
  std::vector<DerivRec> out((size_t)nt * n_s);
  std::vector<double> c7((size_t)n_r * 7);
  const double inv_dr = 1.0 / dr;
  const float nan = std::numeric_limits<float>::quiet_NaN();
  n_miss = 0;
  for (int f = 0; f < nt; f++) {
    spline7(knots + (size_t)f * n_r, n_r, dr, c7.data());
    double hmax = 0.0;
    for (int i = 0; i < n_s; i++)
      hmax = std::max(hmax, std::fabs(spline7_value(c7.data(), n_r, r0, inv_dr, std::sqrt(s0 + ds * (double)i))));
    for (int m = 0; m < n_s; m++) {
      const size_t r = (size_t)f * n_s + m;
      DerivRec &o = out[r];
      o.base = mh.vb[2 * r];
      bool good = use[(size_t)m] != 0;
      if (good) {
        const double a1 = mh.vb[2 * r + 1], a2 = mh.vc[2 * r], a3 = mh.vc[2 * r + 1];
        o.b1 = (float)(a1 - 0.5 * a3);
        o.b2 = (float)(a2 + 1.5 * a3);
        for (int k = 1; k < 8 && good; k++) {
          const double q = (double)k / 8.0;
          const double h_ex = spline7_value(c7.data(), n_r, r0, inv_dr, std::sqrt(s0 + ds * ((double)m + q)));
          const float c = std::fma((float)q, o.b2, o.b1);
          const double h = std::fma(q, (double)c, o.base);
          good = std::fabs(h - h_ex) <= kMixedTolLean * (std::fabs(h_ex) + kMixedFloor * hmax);
        }
        if (!good) n_miss++;
      }
      if (!good) o.b1 = 0.0f, o.b2 = nan;
    }
  }
  return out;
}

// FP64 force records per type pair (a, b): derivative coefficients of rho_b
// and phi_ab per knot, three double2 (see TabF64::fc).
static size_t pack_force_f64(Arena &A, const double *rho_r, const double *phi_r, int nt, int n, double dx) {
  std::vector<double> cr((size_t)n * 7), cp((size_t)n * 7), out((size_t)nt * nt * n * 6);
  for (int a = 0; a < nt; a++)
    for (int b = 0; b < nt; b++) {
      spline7(rho_r + (size_t)b * n, n, dx, cr.data());
      spline7(phi_r + ((size_t)a * nt + b) * n, n, dx, cp.data());
      for (int m = 0; m < n; m++) {
        double *o = &out[(((size_t)a * nt + b) * n + m) * 6];
        const double *r = &cr[(size_t)m * 7], *q = &cp[(size_t)m * 7];
        o[0] = r[0], o[1] = r[1], o[2] = r[2], o[3] = q[0], o[4] = q[1], o[5] = q[2];
      }
    }
  return A.put(out.data(), out.size() * 8);
}

extern "C" int mmd_set_tables(mmd_ctx *c, const mmd_eam_tables *t) {
  GUARD({
    clear_error();
    int rc = ctx_bind(c);
    if (rc) return rc;
    if (!t || t->n_types < 1 || t->n_types > kMaxTypes || t->n_r < 5 || t->n_rho < 5 || !(t->dr > 0.0) ||
        !(t->drho > 0.0) || !(t->r0 >= 0.0) || !t->rho_r || !t->emb || !t->phi_r || !t->mass)
      return fail(MMD_E_INVALID, "mmd_set_tables: bad table descriptor");
    const int nt = t->n_types, nr = t->n_r, nrho = t->n_rho;
    // pair-table records are addressed with 24-bit multiply-adds
    if ((int64_t)nt * nt * nr >= ((int64_t)1 << 24))
      return fail(MMD_E_RANGE, "mmd_set_tables: %d types x %d r knots exceed 2^24 pair-table records", nt, nr);
    for (int a = 0; a < nt; a++) {
      if (!(t->mass[a] > 0.0)) return fail(MMD_E_INVALID, "mmd_set_tables: mass of type %d must be > 0", a);
      for (int b = 0; b < nt; b++)
        if (std::memcmp(t->phi_r + ((size_t)a * nt + b) * nr, t->phi_r + ((size_t)b * nt + a) * nr,
                        sizeof(double) * nr) != 0)
          return fail(MMD_E_INVALID, "mmd_set_tables: phi_%d%d != phi_%d%d (pair tables must be symmetric)", a, b, b,
                      a);
    }
    Arena A;
    const SplHost h_rho = pack_f64(A, t->rho_r, nt, nr, t->dr);
    const SplHost h_emb = pack_f64(A, t->emb, nt, nrho, t->drho);
    const SplHost h_phi = pack_f64(A, t->phi_r, nt * nt, nr, t->dr);
    // mixed: r^2-keyed, n_s knots over [r0^2, rmax^2]; a quarter of the r
    // knots keeps the resampling error near 1e-8 of the forces
    const int n_s = std::max(256, (nr + 3) / 4);
    const double rmax = t->r0 + t->dr * (double)(nr - 1);
    const double s0 = t->r0 * t->r0, ds = (rmax * rmax - s0) / (double)(n_s - 1);
    // largest |F'| over the central 80% of the embedding tables
    double fpmax = 0.02;
    {
      std::vector<double> ce((size_t)nrho * 7);
      for (int a = 0; a < nt; a++) {
        spline7(t->emb + (size_t)a * nrho, nrho, t->drho, ce.data());
        for (int m = nrho / 10; m < nrho - nrho / 10; m++) fpmax = std::max(fpmax, std::fabs(ce[(size_t)m * 7 + 2]));
      }
    }
    MixHost m_rho = pack_mixed(t->rho_r, nt, nr, t->r0, t->dr, n_s, s0, ds, kMixedTolForce / fpmax);
    MixHost m_phi = pack_mixed(t->phi_r, nt * nt, nr, t->r0, t->dr, n_s, s0, ds, kMixedTolForce);
    TabMixed tm_new = {};
    const std::vector<char> use = mixed_ranges(m_rho.ok, m_phi.ok, n_s, s0, ds, tm_new);
    mixed_flag(m_rho, nt, n_s, use);
    mixed_flag(m_phi, nt * nt, n_s, use);
    // derivative records as the force pass uses them: -2 d/ds (exact scaling;
    // NaN flags stay NaN)
    for (MixHost *h : {&m_rho, &m_phi})
      for (DerivRec &d : h->dp) d.base *= -2.0, d.b1 *= -2.0f, d.b2 *= -2.0f;
    const std::vector<DerivRec> &rdp = m_rho.dp, &pdp = m_phi.dp;
    auto put_mixed = [&](const MixHost &h) {
      const size_t vb = A.put(h.vb.data(), h.vb.size() * 8);
      const size_t vc = A.put(h.vc.data(), h.vc.size() * 4);
      const size_t dp = A.put(h.dp.data(), h.dp.size() * sizeof(DerivRec));
      return std::array<size_t, 3>{vb, vc, dp};
    };
    const std::array<size_t, 3> o_rho = put_mixed(m_rho), o_phi = put_mixed(m_phi);
    int64_t lean_miss = 0;
    const std::vector<DerivRec> lean = pack_lean_rho(m_rho, t->rho_r, nt, nr, t->r0, t->dr, n_s, s0, ds, use, lean_miss);
    const size_t h_lean = A.put(lean.data(), lean.size() * sizeof(DerivRec));
    const size_t h_fc = pack_force_f64(A, t->rho_r, t->phi_r, nt, nr, t->dr);
    // mixed force records per type pair (a, b): rho_b and phi_ab derivative records side by side
    std::vector<DerivRec> fb((size_t)nt * nt * n_s * 2);
    for (int a = 0; a < nt; a++)
      for (int b = 0; b < nt; b++)
        for (int m = 0; m < n_s; m++) {
          const size_t r = ((size_t)a * nt + b) * n_s + m;
          fb[2 * r] = rdp[(size_t)b * n_s + m];
          fb[2 * r + 1] = pdp[r];
        }
    const size_t h_fb = A.put(fb.data(), fb.size() * sizeof(DerivRec));

    HIP_TRY(hipStreamSynchronize(c->stream));
    ctx_drop_graphs(c);
    // the old tables go first: until the new ones are in place the context has none
    c->has_tables = false;
    dfree(c->d_tab), dfree(c->d_mixed);
    char *dev = nullptr;
    if ((rc = dalloc(&dev, A.buf.size()))) return rc;
    c->d_tab = reinterpret_cast<double *>(dev);
    HIP_TRY(hipMemcpy(dev, A.buf.data(), A.buf.size(), hipMemcpyHostToDevice));
    auto F = [&](const SplHost &h) {
      SplF64 s;
      s.val = reinterpret_cast<const double2 *>(dev + h.val);
      s.d01 = reinterpret_cast<const double2 *>(dev + h.d01);
      s.d2 = reinterpret_cast<const double *>(dev + h.d2);
      return s;
    };
    auto M = [&](const std::array<size_t, 3> &h) {
      SplMixed s;
      s.vb = reinterpret_cast<const double2 *>(dev + h[0]);
      s.vc = reinterpret_cast<const float2 *>(dev + h[1]);
      s.dp = reinterpret_cast<const double2 *>(dev + h[2]);
      return s;
    };
    TabF64 &tf = c->tf;
    tf.rho = F(h_rho), tf.emb = F(h_emb), tf.phi = F(h_phi);
    tf.fc = reinterpret_cast<const double2 *>(dev + h_fc);
    tf.nt = nt, tf.n_r = nr, tf.n_rho = nrho;
    tf.r0 = t->r0, tf.inv_dr = 1.0 / t->dr, tf.rho0 = t->rho0, tf.inv_drho = 1.0 / t->drho, tf.dr = t->dr;
    TabMixed &tm = c->tm;
    tm = tm_new;
    tm.rho = M(o_rho), tm.phi = M(o_phi);
    tm.fb = reinterpret_cast<const double2 *>(dev + h_fb);
    tm.rho_lean = reinterpret_cast<const double2 *>(dev + h_lean);
    tm.nt = nt, tm.n_s = n_s, tm.s0 = s0, tm.inv_ds = 1.0 / ds;
    for (int a = 0; a < kMaxTypes; a++) {
      c->masses.m[a] = a < nt ? t->mass[a] : 1.0;
      c->masses.inv_m[a] = 1.0 / c->masses.m[a];
    }
    c->n_types = nt;
    c->r_max = rmax;
    c->has_tables = true;
    return MMD_OK;
  })
}

extern "C" int mmd_set_precision(mmd_ctx *c, int32_t mode) {
  int rc = ctx_bind(c);
  if (rc) return rc;
  if (mode != MMD_PREC_FP64 && mode != MMD_PREC_MIXED) return fail(MMD_E_INVALID, "mmd_set_precision: mode %d", mode);
  if (c->prec != mode) ctx_drop_graphs(c);
  c->prec = mode;
  return MMD_OK;
}

extern "C" int mmd_mixed_ranges(mmd_ctx *c, double out[6]) {
  int rc = ctx_bind(c);
  if (rc) return rc;
  if (!out) return fail(MMD_E_INVALID, "mmd_mixed_ranges: null out");
  if (!c->has_tables) return fail(MMD_E_STATE, "mmd_mixed_ranges: tables not set");
  out[0] = c->tm.s_head, out[1] = c->tm.s_tail;
  for (int k = 0; k < kMixedHoles; k++) out[2 + 2 * k] = c->tm.hole_lo[k], out[3 + 2 * k] = c->tm.hole_hi[k];
  return MMD_OK;
}

static int check_ready(mmd_ctx *c, const char *fn, bool need_tables = true) {
  if (!c->has_domain) return fail(MMD_E_STATE, "%s: domain not set", fn);
  if (!c->has_stencil) return fail(MMD_E_STATE, "%s: stencil not set", fn);
  if (need_tables && !c->has_tables) return fail(MMD_E_STATE, "%s: tables not set", fn);
  const int32_t mt = std::max(c->max_type, c->n_itl > 0 ? c->itl_max_type : -1);
  if (need_tables && mt >= c->n_types)
    return fail(MMD_E_INVALID, "%s: atom type %d has no table (n_types = %d)", fn, mt, c->n_types);
  return MMD_OK;
}

static int scan_types_i8(const int8_t *t, size_t n, int32_t *mx) {
  int32_t m = -1;
  for (size_t i = 0; i < n; i++) m = std::max<int32_t>(m, t[i]);
  *mx = m;
  return MMD_OK;
}

static int check_cutoff(mmd_ctx *c, double rc, const char *fn) {
  if (!(rc > 0.0) || !(rc <= c->r_max * (1.0 + 1e-12)))
    return fail(MMD_E_INVALID, "%s: cutoff %g outside (0, %g] covered by the tables", fn, rc, c->r_max);
  return MMD_OK;
}

// ------------------------------------------------------------ data movement

extern "C" int mmd_upload(mmd_ctx *c, uint32_t fields, const mmd_host_soa *h) {
  GUARD({
    clear_error();
    int rc = ctx_bind(c);
    if (rc) return rc;
    if (!c->has_domain) return fail(MMD_E_STATE, "mmd_upload: domain not set");
    if (!h || (fields & ~MMD_F_ALL)) return fail(MMD_E_INVALID, "mmd_upload: bad arguments");
    const size_t n = (size_t)c->g.n_ext;
    struct F {
      uint32_t bit;
      const double *src;
      double *dst;
    } fs[] = {{MMD_F_V, h->vx, c->vx},   {MMD_F_V, h->vy, c->vy},   {MMD_F_V, h->vz, c->vz},
              {MMD_F_F, h->fx, c->fx},   {MMD_F_F, h->fy, c->fy},   {MMD_F_F, h->fz, c->fz},
              {MMD_F_RHO, h->rho, c->rho}, {MMD_F_DF, h->df, c->df}};
    for (auto &f : fs) {
      if (!(fields & f.bit)) continue;
      if (!f.src) return fail(MMD_E_INVALID, "mmd_upload: host pointer missing for field mask 0x%x", f.bit);
      HIP_TRY(hipMemcpyAsync(c->stage, f.src, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
      HIP_TRY(launch_ref2dev_f64(c->stream, c->g, c->stage, f.dst));
    }
    if (fields & MMD_F_X) {
      if (!h->x || !h->y || !h->z) return fail(MMD_E_INVALID, "mmd_upload: position pointers missing");
      c->itl_stale = true;
    }
    if (fields & MMD_F_TYPE) {
      if (!h->type) return fail(MMD_E_INVALID, "mmd_upload: type pointer missing");
      int32_t mx;
      scan_types_i8(h->type, n, &mx);
      if (mx >= kMaxTypes) return fail(MMD_E_INVALID, "mmd_upload: atom type %d beyond %d", mx, kMaxTypes - 1);
      c->max_type = mx;
      int8_t *st = reinterpret_cast<int8_t *>(c->stage);
      HIP_TRY(hipMemcpyAsync(st, h->type, n, hipMemcpyHostToDevice, c->stream));
      HIP_TRY(launch_ref2dev_i8(c->stream, c->g, st, c->type));
    }
    if (fields & (MMD_F_X | MMD_F_TYPE)) {
      double *sx = c->stage, *sy = c->stage + n, *sz = c->stage + 2 * n;
      if (fields & MMD_F_X) {
        HIP_TRY(hipMemcpyAsync(sx, h->x, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
        HIP_TRY(hipMemcpyAsync(sy, h->y, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
        HIP_TRY(hipMemcpyAsync(sz, h->z, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
      }
      HIP_TRY(launch_ref2dev_pos(c->stream, c->g, sx, sy, sz, c->type, c->pos, (fields & MMD_F_X) != 0));
    }
    HIP_TRY(hipStreamSynchronize(c->stream));
    return MMD_OK;
  })
}

extern "C" int mmd_download(mmd_ctx *c, uint32_t fields, const mmd_host_soa *h) {
  GUARD({
    clear_error();
    int rc = ctx_bind(c);
    if (rc) return rc;
    if (!c->has_domain) return fail(MMD_E_STATE, "mmd_download: domain not set");
    if (!h || (fields & ~MMD_F_ALL)) return fail(MMD_E_INVALID, "mmd_download: bad arguments");
    const size_t n = (size_t)c->g.n_ext;
    struct F {
      uint32_t bit;
      double *dst;
      const double *src;
    } fs[] = {{MMD_F_V, h->vx, c->vx},   {MMD_F_V, h->vy, c->vy},   {MMD_F_V, h->vz, c->vz},
              {MMD_F_F, h->fx, c->fx},   {MMD_F_F, h->fy, c->fy},   {MMD_F_F, h->fz, c->fz},
              {MMD_F_RHO, h->rho, c->rho}, {MMD_F_DF, h->df, c->df}};
    for (auto &f : fs) {
      if (!(fields & f.bit)) continue;
      if (!f.dst) return fail(MMD_E_INVALID, "mmd_download: host pointer missing for field mask 0x%x", f.bit);
      HIP_TRY(launch_dev2ref_f64(c->stream, c->g, f.src, c->stage));
      HIP_TRY(hipMemcpyAsync(f.dst, c->stage, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
      HIP_TRY(hipStreamSynchronize(c->stream));
    }
    if (fields & MMD_F_X) {
      if (!h->x || !h->y || !h->z) return fail(MMD_E_INVALID, "mmd_download: position pointers missing");
      double *sx = c->stage, *sy = c->stage + n, *sz = c->stage + 2 * n;
      HIP_TRY(launch_dev2ref_pos(c->stream, c->g, c->pos, sx, sy, sz));
      HIP_TRY(hipMemcpyAsync(h->x, sx, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
      HIP_TRY(hipMemcpyAsync(h->y, sy, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
      HIP_TRY(hipMemcpyAsync(h->z, sz, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
      HIP_TRY(hipStreamSynchronize(c->stream));
    }
    if (fields & MMD_F_TYPE) {
      if (!h->type) return fail(MMD_E_INVALID, "mmd_download: type pointer missing");
      int8_t *st = reinterpret_cast<int8_t *>(c->stage);
      HIP_TRY(launch_dev2ref_i8(c->stream, c->g, c->type, st));
      HIP_TRY(hipMemcpyAsync(h->type, st, n, hipMemcpyDeviceToHost, c->stream));
    }
    HIP_TRY(hipStreamSynchronize(c->stream));
    return MMD_OK;
  })
}

// ------------------------------------------------------------ hook passes

extern "C" int mmd_eam_rho_df(mmd_ctx *c, double cutoff) {
  GUARD({
    clear_error();
    int rc = ctx_bind(c);
    if (rc || (rc = check_ready(c, "mmd_eam_rho_df")) || (rc = check_cutoff(c, cutoff, "mmd_eam_rho_df"))) return rc;
    return run_rho_df(c, cutoff, false, false);
  })
}

extern "C" int mmd_eam_df(mmd_ctx *c) {
  GUARD({
    clear_error();
    int rc = ctx_bind(c);
    if (rc || (rc = check_ready(c, "mmd_eam_df"))) return rc;
    HIP_TRY(launch_df(c->stream, c->g, c->type, c->rho, c->df, c->tf));
    if (c->n_itl > 0) return itl_df(c);
    return MMD_OK;
  })
}

extern "C" int mmd_eam_force(mmd_ctx *c, double cutoff) {
  GUARD({
    clear_error();
    int rc = ctx_bind(c);
    if (rc || (rc = check_ready(c, "mmd_eam_force")) || (rc = check_cutoff(c, cutoff, "mmd_eam_force"))) return rc;
    return run_force(c, cutoff, false);
  })
}

extern "C" int mmd_sync(mmd_ctx *c) {
  int rc = ctx_bind(c);
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(c->stream));
  return MMD_OK;
}

// ------------------------------------------------- drop-in host AoS forms

static int aos_desc(const mmd_aos_layout *lay, AosDesc *d) {
  if (!lay || lay->stride < 8 || (lay->type_bytes != 1 && lay->type_bytes != 2 && lay->type_bytes != 4))
    return fail(MMD_E_INVALID, "AoS layout: stride >= 8 and type_bytes in {1,2,4} required");
  const size_t offs[] = {lay->off_x + 24, lay->off_v + 24, lay->off_f + 24, lay->off_rho + 8, lay->off_df + 8,
                         lay->off_type + (size_t)lay->type_bytes};
  for (size_t o : offs)
    if (o > lay->stride) return fail(MMD_E_INVALID, "AoS layout: field offset beyond the record stride");
  d->stride = (int64_t)lay->stride, d->off_type = (int64_t)lay->off_type, d->off_x = (int64_t)lay->off_x;
  d->off_v = (int64_t)lay->off_v, d->off_f = (int64_t)lay->off_f, d->off_rho = (int64_t)lay->off_rho;
  d->off_df = (int64_t)lay->off_df, d->type_bytes = lay->type_bytes;
  return MMD_OK;
}

// Copies the whole ext record array in, unpacks `in` fields into the device
// state, runs the pass and packs `out` fields into the local records, as the
// reference's streamed passes do.  Pinned records (from mmd_host_alloc /
// hip_create_atoms_mem) take the outputs straight from the pack kernel
// through their device mapping, so only the output fields of occupied local
// records cross the bus back (16 or 24 bytes a record instead of whole
// 104-byte records of the local layers); other host memory gets the local
// z-layers copied back.  The inputs stay one whole-array copy: records are
// read from host memory in 64-byte lines or more, so the 32-byte field
// windows a pass reads cost nearly the whole record either way (strided
// copies of the windows and kernel reads of the mapped records both measured
// slower than the plain copy).
// The rho (fused df) or force pass on pinned records, pipelined by z-chunks of
// the ghost-extended lattice: the copy stream brings the records in chunk by
// chunk while the compute stream unpacks each chunk, runs the pass over the
// z-rows of tiles whose halo has arrived and writes their outputs straight
// into the records.  The outputs leave over the host link while later chunks
// come in (the link carries both directions at once); the records a pack
// writes lie in chunks already copied, so no copy reads a record being
// written.  A type with no table is found after the last chunk: the call
// fails, and the output fields of some occupied local records may already
// hold this pass's values (every other field is untouched).
static int aos_pass_chunked(mmd_ctx *c, const char *atoms, char *mapped, const AosDesc &d, size_t stride,
                            uint32_t in, uint32_t out, int pass, double cutoff) {
  const Geom &g = c->g;
  const size_t layer = (size_t)g.ext_x * g.ext_y * stride;  // bytes of one ext z-layer of records
  const int64_t dev_layer = (int64_t)g.ext_x * g.ext_y;     // device sites of one ext z-layer
  const int64_t ext_z = g.ext_z;
  // chunks of at least the stencil's z reach, about sixteen of them
  const int64_t cz = std::max<int64_t>(std::max<int64_t>(c->halo.rz, 1), (ext_z + 15) / 16);
  const int64_t nch = (ext_z + cz - 1) / cz;
  if (!c->copy_stream) HIP_TRY(hipStreamCreateWithFlags(&c->copy_stream, hipStreamNonBlocking));
  while ((int64_t)c->aos_events.size() < nch + 1) {
    hipEvent_t e;
    HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    c->aos_events.push_back(e);
  }
  // the copies start after the compute stream's earlier work
  HIP_TRY(hipEventRecord(c->aos_events[(size_t)nch], c->stream));
  HIP_TRY(hipStreamWaitEvent(c->copy_stream, c->aos_events[(size_t)nch], 0));
  for (int64_t k = 0; k < nch; k++) {
    const int64_t z0 = k * cz, z1 = std::min(z0 + cz, ext_z);
    HIP_TRY(hipMemcpyAsync(c->aos + (size_t)z0 * layer, atoms + (size_t)z0 * layer, (size_t)(z1 - z0) * layer,
                           hipMemcpyHostToDevice, c->copy_stream));
    HIP_TRY(hipEventRecord(c->aos_events[(size_t)k], c->copy_stream));
  }
  if (in & MMD_F_TYPE) HIP_TRY(launch_fill_i32(c->stream, c->d_maxtype, (int64_t)kMaxTypeSlots * kMaxTypeStride, -1));
  const int64_t tiles_per_tz = (int64_t)g.ntx * g.nty;       // tiles of one z-row of tiles
  const int64_t sites_per_z = (int64_t)g.by * 2 * g.bxh;     // local sites of one local z-plane
  const double rc2 = cutoff * cutoff;
  int64_t tz_done = 0;
  for (int64_t k = 0; k < nch; k++) {
    const int64_t z0 = k * cz, z1 = std::min(z0 + cz, ext_z);
    HIP_TRY(hipStreamWaitEvent(c->stream, c->aos_events[(size_t)k], 0));
    HIP_TRY(launch_aos_unpack(c->stream, g, c->aos, d, in, c->pos, c->type, c->df, c->rho, c->n_types, c->d_maxtype,
                              z0 * dev_layer, z1 * dev_layer));
    // tile row tz reads ext layers [gz + kTileZ tz - rz, gz + kTileZ (tz + 1) + rz): ready once they are in
    const int64_t tz_end =
        z1 >= ext_z ? g.ntz : std::min<int64_t>(g.ntz, std::max<int64_t>(0, (z1 - g.gz - c->halo.rz) / kTileZ));
    if (tz_end <= tz_done) continue;
    if (pass == 0)
      HIP_TRY(launch_rho_df(c->stream, g, c->halo, c->d_off, c->nk_even, c->nk_odd, c->pos, c->rho, c->df, c->tf,
                            c->tm, rc2, c->prec, nullptr, false, tz_done * tiles_per_tz, tz_end * tiles_per_tz));
    else
      HIP_TRY(launch_force(c->stream, g, c->halo, c->d_off, c->nk_even, c->nk_odd, c->pos, c->df, c->fx, c->fy,
                           c->fz, c->tf, c->tm, rc2, c->prec, nullptr, tz_done * tiles_per_tz,
                           tz_end * tiles_per_tz));
    HIP_TRY(launch_aos_pack(c->stream, g, mapped, d, out, c->type, c->rho, c->df, c->fx, c->fy, c->fz,
                            tz_done * kTileZ * sites_per_z,
                            std::min<int64_t>(tz_end * kTileZ, g.bz) * sites_per_z));
    tz_done = tz_end;
  }
  if (in & MMD_F_X) c->itl_stale = true;
  if (in & MMD_F_TYPE) {
    std::vector<int32_t> slots((size_t)kMaxTypeSlots * kMaxTypeStride);
    HIP_TRY(hipMemcpyAsync(slots.data(), c->d_maxtype, slots.size() * sizeof(int32_t), hipMemcpyDeviceToHost,
                           c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    int32_t mx = -1;
    for (int k = 0; k < kMaxTypeSlots; k++) mx = std::max(mx, slots[(size_t)k * kMaxTypeStride]);
    c->max_type = std::min(mx, c->n_types - 1);
    if (mx >= c->n_types) return fail(MMD_E_INVALID, "atom type %d has no table (n_types = %d)", mx, c->n_types);
  }
  HIP_TRY(hipStreamSynchronize(c->stream));
  return MMD_OK;
}

static int aos_pass(mmd_ctx *c, void *atoms, const mmd_aos_layout *lay, uint32_t in, uint32_t out, int pass,
                    double cutoff) {
  if (!atoms) return fail(MMD_E_INVALID, "null atom array");
  AosDesc d;
  int rc = aos_desc(lay, &d);
  if (rc) return rc;
  const size_t bytes = (size_t)c->g.n_ext * lay->stride;
  if (!c->d_maxtype && (rc = dalloc(&c->d_maxtype, (size_t)kMaxTypeSlots * kMaxTypeStride))) return rc;
  hipPointerAttribute_t attr{};
  char *mapped = nullptr;  // device address of the host records when they are pinned
  if (hipPointerGetAttributes(&attr, atoms) == hipSuccess && attr.type == hipMemoryTypeHost && attr.devicePointer)
    mapped = static_cast<char *>(attr.devicePointer);
  (void)hipGetLastError();  // pageable memory is not an error here
  if (bytes > c->aos_bytes) {
    HIP_TRY(hipStreamSynchronize(c->stream));
    dfree(c->aos);
    c->aos_bytes = 0;
    if ((rc = dalloc(&c->aos, bytes))) return rc;
    c->aos_bytes = bytes;
  }
  if (mapped && c->n_itl == 0 && (pass == 0 || pass == 2))
    return aos_pass_chunked(c, static_cast<const char *>(atoms), mapped, d, lay->stride, in, out, pass, cutoff);
  HIP_TRY(hipMemcpyAsync(c->aos, atoms, bytes, hipMemcpyHostToDevice, c->stream));
  if (in & MMD_F_TYPE) HIP_TRY(launch_fill_i32(c->stream, c->d_maxtype, (int64_t)kMaxTypeSlots * kMaxTypeStride, -1));
  HIP_TRY(launch_aos_unpack(c->stream, c->g, c->aos, d, in, c->pos, c->type, c->df, c->rho, c->n_types,
                            c->d_maxtype));
  if (in & MMD_F_X) c->itl_stale = true;
  if (in & MMD_F_TYPE) {  // types are checked on the device (a type with no table was stored as a vacancy)
    std::vector<int32_t> slots((size_t)kMaxTypeSlots * kMaxTypeStride);
    HIP_TRY(hipMemcpyAsync(slots.data(), c->d_maxtype, slots.size() * sizeof(int32_t), hipMemcpyDeviceToHost,
                           c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    int32_t mx = -1;
    for (int k = 0; k < kMaxTypeSlots; k++) mx = std::max(mx, slots[(size_t)k * kMaxTypeStride]);
    c->max_type = std::min(mx, c->n_types - 1);
    if (mx >= c->n_types) return fail(MMD_E_INVALID, "atom type %d has no table (n_types = %d)", mx, c->n_types);
  }
  if (pass == 0) rc = run_rho_df(c, cutoff, false, false);
  if (pass == 1) {
    HIP_TRY(launch_df(c->stream, c->g, c->type, c->rho, c->df, c->tf));
  }
  if (pass == 2) rc = run_force(c, cutoff, false);
  if (rc) return rc;
  if (mapped) {
    HIP_TRY(launch_aos_pack(c->stream, c->g, mapped, d, out, c->type, c->rho, c->df, c->fx, c->fy, c->fz));
  } else {
    HIP_TRY(launch_aos_pack(c->stream, c->g, c->aos, d, out, c->type, c->rho, c->df, c->fx, c->fy, c->fz));
    const size_t layer = (size_t)c->g.ext_x * c->g.ext_y * lay->stride;
    const size_t first = (size_t)c->g.gz * layer, count = (size_t)c->g.bz * layer;
    HIP_TRY(hipMemcpyAsync(static_cast<char *>(atoms) + first, c->aos + first, count, hipMemcpyDeviceToHost,
                           c->stream));
  }
  HIP_TRY(hipStreamSynchronize(c->stream));
  return MMD_OK;
}

extern "C" int mmd_eam_rho_calc_aos(mmd_ctx *c, void *atoms, const mmd_aos_layout *lay, double cutoff) {
  GUARD({
    clear_error();
    int rc = ctx_bind(c);
    if (rc || (rc = check_ready(c, "mmd_eam_rho_calc_aos")) || (rc = check_cutoff(c, cutoff, "mmd_eam_rho_calc_aos")))
      return rc;
    return aos_pass(c, atoms, lay, MMD_F_X | MMD_F_TYPE, MMD_F_RHO | MMD_F_DF, 0, cutoff);
  })
}

extern "C" int mmd_eam_df_calc_aos(mmd_ctx *c, void *atoms, const mmd_aos_layout *lay) {
  GUARD({
    clear_error();
    int rc = ctx_bind(c);
    if (rc || (rc = check_ready(c, "mmd_eam_df_calc_aos"))) return rc;
    return aos_pass(c, atoms, lay, MMD_F_TYPE | MMD_F_RHO, MMD_F_DF, 1, 0.0);
  })
}

extern "C" int mmd_eam_force_calc_aos(mmd_ctx *c, void *atoms, const mmd_aos_layout *lay, double cutoff) {
  GUARD({
    clear_error();
    int rc = ctx_bind(c);
    if (rc || (rc = check_ready(c, "mmd_eam_force_calc_aos")) ||
        (rc = check_cutoff(c, cutoff, "mmd_eam_force_calc_aos")))
      return rc;
    return aos_pass(c, atoms, lay, MMD_F_X | MMD_F_TYPE | MMD_F_DF, MMD_F_F, 2, cutoff);
  })
}

extern "C" int mmd_host_alloc(void **out, size_t bytes) {
  if (!out) return fail(MMD_E_INVALID, "mmd_host_alloc: null out");
  *out = nullptr;
  if (bytes == 0) return MMD_OK;
  HIP_TRY(hipHostMalloc(out, bytes, hipHostMallocDefault));
  return MMD_OK;
}

extern "C" int mmd_host_free(void *p) {
  if (!p) return MMD_OK;
  HIP_TRY(hipHostFree(p));
  return MMD_OK;
}

// --------------------------------------------------------------- MD step

extern "C" int mmd_md_setup(mmd_ctx *c, const mmd_md_params *p) {
  GUARD({
    clear_error();
    int rc = ctx_bind(c);
    if (rc || (rc = check_ready(c, "mmd_md_setup"))) return rc;
    if (!p || !(p->dt > 0.0) || p->energy_every < 0 || !(p->lattice_const > 0.0))
      return fail(MMD_E_INVALID, "mmd_md_setup: bad parameters");
    for (int k = 0; k < 3; k++)
      if (!(p->box_len[k] > 0.0)) return fail(MMD_E_INVALID, "mmd_md_setup: box length %d must be > 0", k);
    if ((rc = check_cutoff(c, p->cutoff, "mmd_md_setup"))) return rc;
    for (int k = 0; k < 3; k++)
      if (c->dom.start[k] != 0 || c->dom.ghost[k] > c->dom.box[k])
        return fail(MMD_E_INVALID, "mmd_md_setup: the periodic single-domain step needs start = 0");
    c->md = *p;
    c->md_ready = true;
    ctx_drop_graphs(c);
    // interstitial bins follow the box, lattice constant and cutoff: sized here, never inside a step
    return itl_setup(c);
  })
}

extern "C" int mmd_set_half_stencil(mmd_ctx *c, int32_t on) {
  GUARD({
    clear_error();
    int rc = ctx_bind(c);
    if (rc) return rc;
    if (!c->has_stencil) return fail(MMD_E_STATE, "mmd_set_half_stencil: set the stencil first");
    if (on && !c->half_ok)
      return fail(MMD_E_INVALID, "mmd_set_half_stencil: the stencil is not symmetric (a pair would be seen once)");
    HIP_TRY(hipStreamSynchronize(c->stream));
    ctx_drop_graphs(c);
    dfree(c->slots), c->half = false;
    if (!on) return MMD_OK;
    const size_t n = (size_t)std::max(std::max(c->nf[0], c->nf[1]), 1) * (size_t)c->g.n_local;
    if ((rc = dalloc(&c->slots, n))) return rc;
    c->half = true;
    return MMD_OK;
  })
}

extern "C" int mmd_ghost_refresh(mmd_ctx *c, uint32_t fields) {
  GUARD({
    clear_error();
    int rc = ctx_bind(c);
    if (rc) return rc;
    if (!c->md_ready) return fail(MMD_E_STATE, "mmd_ghost_refresh: call mmd_md_setup first");
    if ((rc = check_ready(c, "mmd_ghost_refresh", false))) return rc;
    if (fields & ~(MMD_F_X | MMD_F_DF)) return fail(MMD_E_INVALID, "mmd_ghost_refresh: only MMD_F_X and MMD_F_DF");
    return run_ghost(c, fields);
  })
}

extern "C" int mmd_md_prime(mmd_ctx *c) {
  GUARD({
    clear_error();
    int rc = ctx_bind(c);
    if (rc) return rc;
    if (!c->md_ready) return fail(MMD_E_STATE, "mmd_md_prime: call mmd_md_setup first");
    if ((rc = check_ready(c, "mmd_md_prime")) || (rc = check_cutoff(c, c->md.cutoff, "mmd_md_prime"))) return rc;
    if ((rc = run_ghost(c, MMD_F_X)) || (rc = run_rho_df(c, c->md.cutoff, false, true)) || (rc = run_ghost(c, MMD_F_DF)) ||
        (rc = run_force(c, c->md.cutoff, false, true)))
      return rc;
    return MMD_OK;
  })
}

extern "C" int mmd_md_run(mmd_ctx *c, int32_t nsteps, double *energies, int32_t use_graph) {
  GUARD({
    clear_error();
    int rc = ctx_bind(c);
    if (rc) return rc;
    if (!c->md_ready) return fail(MMD_E_STATE, "mmd_md_run: call mmd_md_setup first");
    if ((rc = check_ready(c, "mmd_md_run")) || (rc = check_cutoff(c, c->md.cutoff, "mmd_md_run"))) return rc;
    if (nsteps < 0) return fail(MMD_E_INVALID, "mmd_md_run: nsteps < 0");
    const int every = c->md.energy_every;
    const int64_t n_rec = every > 0 ? nsteps / every : 0;
    if ((rc = ensure_energy_cap(c, std::max<int64_t>(n_rec, 1)))) return rc;
    HIP_TRY(launch_fill_i32(c->stream, c->d_counter, 1, 0));
    auto records = [&](int32_t s) { return every > 0 && (s + 1) % every == 0; };
    for (int32_t s = 0; s < nsteps; s++) {
      const bool want_e = records(s), fused = s > 0, prev_e = fused && records(s - 1);
      if (use_graph) {
        hipGraphExec_t &ge = c->graph[(fused ? 4 : 0) + (prev_e ? 2 : 0) + (want_e ? 1 : 0)];
        if (!ge) {
          hipGraph_t gr;
          HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
          c->capturing = true;
          rc = md_step(c, fused, prev_e, want_e);
          c->capturing = false;
          hipError_t ee = hipStreamEndCapture(c->stream, &gr);
          if (rc) {
            if (ee == hipSuccess) (void)hipGraphDestroy(gr);
            return rc;
          }
          HIP_TRY(ee);
          hipError_t ie = hipGraphInstantiate(&ge, gr, nullptr, nullptr, 0);
          (void)hipGraphDestroy(gr);
          HIP_TRY(ie);
        }
        HIP_TRY(hipGraphLaunch(ge, c->stream));
      } else {
        if ((rc = md_step(c, fused, prev_e, want_e))) return rc;
      }
    }
    if (nsteps > 0 && (rc = md_close(c, records(nsteps - 1)))) return rc;
    if ((rc = itl_check(c))) return rc;
    if (n_rec > 0 && energies) {
      std::vector<double> buf((size_t)(2 * n_rec));
      HIP_TRY(hipMemcpyAsync(buf.data(), c->d_energy, buf.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
      HIP_TRY(hipStreamSynchronize(c->stream));
      for (int64_t r = 0; r < n_rec; r++) {
        const int64_t step = (r + 1) * every - 1;
        energies[2 * step] = buf[2 * r];
        energies[2 * step + 1] = buf[2 * r + 1];
      }
    }
    return MMD_OK;
  })
}

extern "C" int mmd_md_run_noalloc(mmd_ctx *c, int32_t nsteps, double *energies) {
  GUARD({
    clear_error();
    int rc = ctx_bind(c);
    if (rc) return rc;
    if (!c->md_ready) return fail(MMD_E_STATE, "mmd_md_run_noalloc: call mmd_md_setup first");
    if ((rc = check_ready(c, "mmd_md_run_noalloc")) || (rc = check_cutoff(c, c->md.cutoff, "mmd_md_run_noalloc")))
      return rc;
    if (nsteps < 0) return fail(MMD_E_INVALID, "mmd_md_run_noalloc: nsteps < 0");
    double *S = scratch_of(c->device);
    if (!S) return fail(MMD_E_STATE, "mmd_md_run_noalloc: no scratch address for device %d", c->device);
    const int every = c->md.energy_every;
    const int64_t n_rec = every > 0 ? nsteps / every : 0;
    const int64_t nt = tile_blocks(c->g), nb = site_blocks(c->g);
    const int64_t ni = c->n_itl > 0 ? c->itl_part_n + c->itl_kick_n : 0;
    // partials, records, then one double holding the record counter
    const int64_t need = 2 * nt + nb + ni + 2 * std::max<int64_t>(n_rec, 1) + 1;
    if (need > (int64_t)kScratchDoubles)
      return fail(MMD_E_RANGE,
                  "mmd_md_run_noalloc: %lld energy partials and records exceed the %zu doubles of scratch "
                  "(fewer energy records or a smaller domain)",
                  (long long)need, kScratchDoubles);
    if (c->n_itl > 0 && !itl_bins_ready(c))
      return fail(MMD_E_STATE, "mmd_md_run_noalloc: interstitial bins are not sized for this box (mmd_md_prime "
                               "sizes them)");
    ScratchUse use(c, S, nt, nb, ni, n_rec);
    // the previous call's contents: partials of blocks a grid does not reach and the counter start at zero
    HIP_TRY(launch_fill_f64(c->stream, S, need, 0.0));
    auto records = [&](int32_t s) { return every > 0 && (s + 1) % every == 0; };
    for (int32_t s = 0; s < nsteps; s++) {
      const bool want_e = records(s), fused = s > 0, prev_e = fused && records(s - 1);
      if ((rc = md_step(c, fused, prev_e, want_e))) return rc;
    }
    if (nsteps > 0 && (rc = md_close(c, records(nsteps - 1)))) return rc;
    if ((rc = itl_check(c))) return rc;
    if (energies)
      for (int64_t r = 0; r < n_rec; r++) {
        const int64_t step = (r + 1) * every - 1;
        HIP_TRY(hipMemcpyAsync(energies + 2 * step, c->d_energy + 2 * r, 2 * sizeof(double), hipMemcpyDeviceToHost,
                               c->stream));
      }
    HIP_TRY(hipStreamSynchronize(c->stream));
    return MMD_OK;
  })
}

extern "C" int mmd_energy(mmd_ctx *c, double *epot, double *ekin) {
  GUARD({
    clear_error();
    int rc = ctx_bind(c);
    if (rc) return rc;
    if (!c->md_ready) return fail(MMD_E_STATE, "mmd_energy: call mmd_md_setup first");
    if ((rc = check_ready(c, "mmd_energy")) || (rc = check_cutoff(c, c->md.cutoff, "mmd_energy"))) return rc;
    if ((rc = ensure_energy_cap(c, 1))) return rc;
    HIP_TRY(launch_fill_i32(c->stream, c->d_counter, 1, 0));
    if ((rc = run_ghost(c, MMD_F_X)) || (rc = run_rho_df(c, c->md.cutoff, true, true)) || (rc = run_ghost(c, MMD_F_DF)) ||
        (rc = run_force(c, c->md.cutoff, true, true)) || (rc = run_kick(c, 0.0, false, true)) ||
        (rc = run_finalize(c, c->d_energy, c->d_counter)))
      return rc;
    if ((rc = itl_check(c))) return rc;
    double e[2];
    HIP_TRY(hipMemcpyAsync(e, c->d_energy, sizeof(e), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (epot) *epot = e[0];
    if (ekin) *ekin = e[1];
    return MMD_OK;
  })
}

// ---------------------------------------------------------------- timing

extern "C" int mmd_timing_enable(mmd_ctx *c, int32_t on) {
  int rc = ctx_bind(c);
  if (rc) return rc;
  c->timing = on != 0;
  return MMD_OK;
}

extern "C" int mmd_timing_get(mmd_ctx *c, double *ms_total, int64_t *launches) {
  int rc = ctx_bind(c);
  if (rc) return rc;
  if ((rc = timing_collect(c))) return rc;
  for (int k = 0; k < MMD_K_COUNT; k++) {
    if (ms_total) ms_total[k] = c->t_ms[k];
    if (launches) launches[k] = c->t_n[k];
  }
  return MMD_OK;
}

extern "C" int mmd_timing_reset(mmd_ctx *c) {
  int rc = ctx_bind(c);
  if (rc) return rc;
  if ((rc = timing_collect(c))) return rc;
  for (int k = 0; k < MMD_K_COUNT; k++) c->t_ms[k] = 0.0, c->t_n[k] = 0;
  return MMD_OK;
}

extern "C" int mmd_stream_copy(mmd_ctx *c, size_t bytes, int32_t repeats, double *ms) {
  GUARD({
    clear_error();
    int rc = ctx_bind(c);
    if (rc) return rc;
    if (bytes == 0 || bytes % 16 || repeats < 1) return fail(MMD_E_INVALID, "mmd_stream_copy: bytes % 16 == 0, repeats >= 1");
    char *a = nullptr, *b = nullptr;
    if ((rc = dalloc(&a, bytes))) return rc;
    if ((rc = dalloc(&b, bytes))) {
      dfree(a);
      return rc;
    }
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(launch_fill_f64(c->stream, reinterpret_cast<double *>(a), (int64_t)(bytes / 8), 1.0));
    HIP_TRY(launch_stream_copy(c->stream, a, b, (int64_t)bytes));  // warm-up
    HIP_TRY(hipEventRecord(e0, c->stream));
    for (int32_t r = 0; r < repeats; r++) HIP_TRY(launch_stream_copy(c->stream, a, b, (int64_t)bytes));
    HIP_TRY(hipEventRecord(e1, c->stream));
    HIP_TRY(hipEventSynchronize(e1));
    float t = 0.0f;
    HIP_TRY(hipEventElapsedTime(&t, e0, e1));
    if (ms) *ms = t / repeats;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    dfree(a);
    dfree(b);
    return MMD_OK;
  })
}
