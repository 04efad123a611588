// Serving loop: continuous batching of independent MD replicas on one GPU.
//
// Every replica (a periodic cubic BCC box of a random alloy) lives in its own
// range of one device-memory arena taken from free HBM when the server is
// created: device-layout state, stencil, energy partials.  Admission is FIFO
// by batch slot and arena space (serve_sched.h).  Host worker threads generate a
// replica's initial conditions, convert them to the device layout and copy
// them in on their own streams while the GPU keeps stepping the others.
// Resident replicas are primed (forces and admission energies), then advance
// together: one batched step is the velocity-Verlet sequence of the
// single-domain path (kick + drift, ghost x, rho + df, ghost df, force, kick),
// each pass one launch over a device work list of (replica, block) items, one
// block per item with the grid sized to the list's capacity, replayed from
// one captured HIP graph whatever the mix (between mixes only the device work
// lists change; a graph is captured again only when a list grows).  A replica's
// blocks run the same per-block code with the same block numbering as the
// single-domain kernels and its energies are reduced in the same fixed order,
// so its trajectory is bit-identical alone, in any mix and on the
// single-domain path.
#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "batch.h"
#include "ctx.h"
#include "host_helpers.h"
#include "kernels.h"
#include "mmd_eam.h"
#include "serve_sched.h"

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

namespace {

constexpr int kMaxCells = 512;  // keeps every replica's ext lattice below 2^31 sites

// Byte layout of one replica inside the arena.  [0, fx) is written from the
// host image prepared by a worker; [fx, bytes) is zero-filled on the device.
struct RepLayout {
  mmd_domain d;
  Geom g;
  double L[3];
  int64_t nb, nt, n_ext_blocks;
  size_t pos, vx, vy, vz, type, fx, fy, fz, rho, df, e_rho, e_force, e_ke, energy, bytes;
};

int rep_layout(int32_t cells, int32_t ghost, double a, RepLayout &R) {
  if (cells < 1 || cells > kMaxCells) return fail(MMD_E_RANGE, "replica of %d cells outside [1, %d]", cells, kMaxCells);
  mmd_lattice_spec s{};
  s.cells[0] = s.cells[1] = s.cells[2] = cells;
  s.ghost_cells = ghost;
  s.lattice_const = a;
  int rc = mmd_lattice_domain(&s, &R.d, R.L);
  if (rc) return rc;
  R.g = make_geom(R.d);
  R.nb = site_blocks(R.g);
  R.nt = tile_blocks(R.g);
  R.n_ext_blocks = (R.g.n_ghost + kBlock - 1) / kBlock;  // ghost refresh blocks
  const size_t n = (size_t)R.g.n_ext;
  size_t o = 0;
  auto put = [&](size_t bytes) {
    const size_t at = o;
    o = ArenaAlloc::round(o + bytes);
    return at;
  };
  R.pos = put(32 * n);
  R.vx = put(8 * n), R.vy = put(8 * n), R.vz = put(8 * n);
  R.type = put(n);
  R.fx = put(8 * n), R.fy = put(8 * n), R.fz = put(8 * n);
  R.rho = put(8 * n), R.df = put(8 * n);
  R.e_rho = put(8 * (size_t)R.nt), R.e_force = put(8 * (size_t)R.nt), R.e_ke = put(8 * (size_t)R.nb);
  R.energy = put(4 * sizeof(double));
  R.bytes = o;
  return MMD_OK;
}

void fill_result(const Rep &r, mmd_replica_result &o) {
  std::memset(&o, 0, sizeof(o));
  o.id = r.id;
  o.status = r.state == kDone ? MMD_OK : (r.status ? r.status : MMD_E_STATE);
  o.n_atoms = (int32_t)r.atoms;
  o.t_arrival = r.arrival;
  o.t_admit = r.t_admit;
  o.t_ready = r.t_ready;
  o.t_done = r.t_done;
  o.epot = r.e[0], o.ekin = r.e[1], o.epot0 = r.e[2], o.ekin0 = r.e[3];
  o.slot = r.slot;
  o.arena_offset = r.slot >= 0 ? r.off : 0;
  o.arena_bytes = r.slot >= 0 ? ArenaAlloc::round(r.bytes) : 0;
}

void fill_stats(const std::vector<Rep> &R, const SchedStats &s, size_t arena, mmd_server_stats *o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  std::vector<double> lat;
  double wait = 0.0;
  for (const Rep &r : R)
    if (r.state == kDone) {
      lat.push_back(r.t_done - r.arrival);
      wait += r.t_admit - r.arrival;
    }
  o->wall_s = s.wall_s;
  o->atom_steps = s.atom_steps;
  o->atom_steps_per_s = s.wall_s > 0.0 ? s.atom_steps / s.wall_s : 0.0;
  o->latency_p50_s = percentile(lat, 0.50);
  o->latency_p99_s = percentile(lat, 0.99);
  double m = 0.0;
  for (double v : lat) m += v;
  o->latency_mean_s = lat.empty() ? 0.0 : m / (double)lat.size();
  o->wait_mean_s = lat.empty() ? 0.0 : wait / (double)lat.size();
  o->completed = s.completed;
  o->failed = s.failed;
  o->max_active = s.max_active;
  o->batched_steps = s.batched_steps;
  o->arena_bytes = (double)arena;
}

int check_specs(const mmd_replica_spec *specs, int32_t n, int32_t ghost) {
  if (n < 0 || (n > 0 && !specs)) return fail(MMD_E_INVALID, "serving: bad replica list");
  for (int32_t i = 0; i < n; i++) {
    const mmd_replica_spec &s = specs[i];
    if (s.cells < ghost || s.cells > kMaxCells || s.steps < 0 || !(s.temperature >= 0.0) ||
        !(s.arrival_s >= 0.0) || !std::isfinite(s.arrival_s))
      return fail(MMD_E_INVALID, "serving: replica %d: cells in [%d, %d], steps >= 0, temperature >= 0, arrival >= 0",
                  i, ghost, kMaxCells);
  }
  return MMD_OK;
}

int make_reps(const mmd_replica_spec *specs, int32_t n, int32_t ghost, double a, std::vector<Rep> &R) {
  R.assign((size_t)n, Rep{});
  for (int32_t i = 0; i < n; i++) {
    RepLayout L;
    int rc = rep_layout(specs[i].cells, ghost, a, L);
    if (rc) return rc;
    Rep &r = R[(size_t)i];
    r.id = specs[i].id;
    r.seed = specs[i].seed;
    r.cells = specs[i].cells;
    r.steps = specs[i].steps;
    r.temperature = specs[i].temperature;
    r.arrival = specs[i].arrival_s;
    r.atoms = 2 * (int64_t)r.cells * r.cells * r.cells;
    r.bytes = L.bytes;
  }
  return MMD_OK;
}

// ------------------------------------------------------ simulated executor

struct SimExec final : Executor {
  const mmd_sim_params &p;
  std::vector<Rep> &R;
  double t = 0.0;
  std::vector<double> wfree;
  std::vector<std::pair<double, int>> loading;  // (end time, replica)

  SimExec(const mmd_sim_params &pp, std::vector<Rep> &RR) : p(pp), R(RR), wfree((size_t)pp.workers, 0.0) {}
  double now() override { return t; }
  int start_load(int r) override {
    auto w = std::min_element(wfree.begin(), wfree.end());
    const double end = std::max(t, *w) + p.load_atom_s * (double)R[r].atoms;
    *w = end;
    loading.push_back({end, r});
    R[r].status = MMD_OK;
    return MMD_OK;
  }
  void collect(std::vector<int> &out) {
    std::sort(loading.begin(), loading.end());
    size_t k = 0;
    while (k < loading.size() && loading[k].first <= t) out.push_back(loading[k++].second);
    loading.erase(loading.begin(), loading.begin() + (ptrdiff_t)k);
  }
  int wait_loaded(std::vector<int> &out, double deadline) override {
    collect(out);
    if (!out.empty() || deadline <= t) return MMD_OK;
    double next = deadline;
    for (auto &l : loading) next = std::min(next, l.first);
    t = std::max(t, next);
    collect(out);
    return MMD_OK;
  }
  double cost(const std::vector<int> &reps) const {
    double atoms = 0.0;
    for (int r : reps) atoms += (double)R[r].atoms;
    return p.step_overhead_s + p.step_atom_s * atoms;
  }
  int prime(const std::vector<int> &reps) override {
    t += cost(reps);
    return MMD_OK;
  }
  int run(const std::vector<int> &active, int steps) override {
    t += steps * cost(active);
    return MMD_OK;
  }
  int finish(const std::vector<int> &) override { return MMD_OK; }
};

// ------------------------------------------------------ device work lists

struct ListBuf {
  int2 *h_site = nullptr, *d_site = nullptr;
  int2 *h_tile = nullptr, *d_tile = nullptr;
  int2 *h_ext = nullptr, *d_ext = nullptr;
  size_t cap_site = 0, cap_tile = 0, cap_ext = 0;
  int *h_reps = nullptr, *d_reps = nullptr;  // [max_replicas]
  int *h_n = nullptr, *d_n = nullptr;        // {n_site, n_tile, n_ext, n_reps}
  int64_t k_site = 0, k_tile = 0, k_ext = 0;  // items each launch covers, one block per item

  BatchLists lists(const BatchDom *dom) const {
    return {dom, d_site, d_n, d_tile, d_n + 1, d_ext, d_n + 2, d_reps, d_n + 3};
  }
};

template <typename T>
int grow(T **h, T **d, size_t *cap, size_t need) {
  if (need <= *cap) return MMD_OK;
  size_t c = std::max<size_t>(need, 2 * *cap);
  c = std::max<size_t>(c, 1024);
  if (*h) (void)hipHostFree(*h);
  if (*d) (void)hipFree(*d);
  *h = nullptr, *d = nullptr, *cap = 0;
  HIP_TRY(hipHostMalloc((void **)h, c * sizeof(T), hipHostMallocDefault));
  HIP_TRY(hipMalloc((void **)d, c * sizeof(T)));
  *cap = c;
  return MMD_OK;
}

void free_lists(ListBuf &b) {
  if (b.h_site) (void)hipHostFree(b.h_site);
  if (b.h_tile) (void)hipHostFree(b.h_tile);
  if (b.h_ext) (void)hipHostFree(b.h_ext);
  if (b.h_reps) (void)hipHostFree(b.h_reps);
  if (b.h_n) (void)hipHostFree(b.h_n);
  if (b.d_site) (void)hipFree(b.d_site);
  if (b.d_tile) (void)hipFree(b.d_tile);
  if (b.d_ext) (void)hipFree(b.d_ext);
  if (b.d_reps) (void)hipFree(b.d_reps);
  if (b.d_n) (void)hipFree(b.d_n);
  b = ListBuf{};
}

}  // namespace

// ------------------------------------------------------------- GPU server

struct mmd_server {
  mmd_server_opts o{};
  mmd_ctx *tab = nullptr;  // tables, masses and precision; its stream is the compute stream
  hipStream_t cs = nullptr;
  char *arena = nullptr;
  size_t arena_bytes = 0;
  std::vector<int32_t> st3[2];  // stencil (dx doubled, dy, dz) per centre parity
  int32_t nk[2] = {0, 0};
  Halo halo{};                  // pair-pass LDS tile of the stencil
  int32_t *d_off_lds = nullptr; // stencil device image (LDS offsets, energy masks), shared by all replicas
  BatchDom *h_dom = nullptr, *d_dom = nullptr;  // [max_replicas], indexed by slot
  ListBuf step, prime;
  std::vector<int> step_set;  // active replicas the step lists were built for
  // The step launches' grids follow the work lists' capacities (grown
  // geometrically), not the current item counts, which the kernels read from
  // device memory: blocks past the count exit at once, so one captured graph
  // serves every mix of replicas until a list has to grow (only the device
  // work lists change between mixes).
  hipGraphExec_t graph[4] = {};  // captured steps by (fused opening kick, energy partials)
  int64_t graph_captures = 0;
  double *h_e = nullptr;  // pinned [max_replicas][4]

  // worker pool
  std::vector<std::thread> workers;
  std::mutex mu;
  std::condition_variable cv_job, cv_done;
  std::deque<int> jobs;
  std::vector<int> done_q;
  int in_flight = 0;  // loads queued or running
  bool stop = false;

  // current run
  std::vector<Rep> *R = nullptr;
  const mmd_replica_spec *specs = nullptr;
  std::chrono::steady_clock::time_point t0;
};

namespace {

void drop_graphs(mmd_server *S) {
  for (auto &g : S->graph)
    if (g) {
      (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
}

// Worker: initial conditions of replica r -> device-layout host image -> arena.
int load_one(mmd_server *S, int r, hipStream_t ws, char *&stage, size_t &stage_cap, std::vector<double> &ref,
             std::vector<int8_t> &rtype) {
  const Rep &rep = (*S->R)[(size_t)r];
  RepLayout L;
  int rc = rep_layout(rep.cells, S->o.ghost_cells, S->o.lattice_const, L);
  if (rc) return rc;
  const size_t n = (size_t)L.g.n_ext;
  ref.resize(6 * n);
  rtype.resize(n);
  mmd_host_soa h{};
  h.x = ref.data(), h.y = h.x + n, h.z = h.y + n, h.vx = h.z + n, h.vy = h.vx + n, h.vz = h.vy + n;
  h.type = rtype.data();
  mmd_lattice_spec s{};
  s.cells[0] = s.cells[1] = s.cells[2] = rep.cells;
  s.ghost_cells = S->o.ghost_cells;
  s.lattice_const = S->o.lattice_const;
  s.n_types = S->tab->n_types;
  for (int t = 0; t < 8; t++) s.type_cum[t] = S->o.type_cum[t];
  for (int t = 0; t < s.n_types; t++) s.mass[t] = S->tab->masses.m[t];
  s.sigma = S->o.sigma;
  s.temperature = rep.temperature;
  s.seed = rep.seed;
  int32_t n_itl = 0;
  if ((rc = mmd_lattice_init(&s, &h, 0, &n_itl, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr)))
    return rc;
  if (L.fx > stage_cap) {
    if (stage) (void)hipHostFree(stage);
    stage = nullptr, stage_cap = 0;
    HIP_TRY(hipHostMalloc((void **)&stage, L.fx, hipHostMallocDefault));
    stage_cap = L.fx;
  }
  double4 *P = reinterpret_cast<double4 *>(stage + L.pos);
  double *VX = reinterpret_cast<double *>(stage + L.vx), *VY = reinterpret_cast<double *>(stage + L.vy),
         *VZ = reinterpret_cast<double *>(stage + L.vz);
  int8_t *T = reinterpret_cast<int8_t *>(stage + L.type);
  const Geom &g = L.g;
  for (int64_t ze = 0; ze < g.ext_z; ze++)
    for (int64_t ye = 0; ye < g.ext_y; ye++)
      for (int64_t xe = 0; xe < g.ext_x; xe++) {
        const size_t i = (size_t)((ze * g.ext_y + ye) * g.ext_x + xe);
        const size_t d = (size_t)dev_index(g, (int32_t)(xe & 1), (int32_t)(xe >> 1), (int32_t)ye, (int32_t)ze);
        P[d] = make_double4(h.x[i], h.y[i], h.z[i], (double)h.type[i]);
        VX[d] = h.vx[i], VY[d] = h.vy[i], VZ[d] = h.vz[i];
        T[d] = h.type[i];
      }
  char *base = S->arena + rep.off;
  HIP_TRY(hipMemcpyAsync(base, stage, L.fx, hipMemcpyHostToDevice, ws));
  HIP_TRY(launch_fill_f64(ws, reinterpret_cast<double *>(base + L.fx), (int64_t)((L.bytes - L.fx) / 8), 0.0));
  HIP_TRY(hipStreamSynchronize(ws));
  return MMD_OK;
}

void worker_main(mmd_server *S) {
  hipStream_t ws = nullptr;
  bool ok = hipSetDevice(S->o.device) == hipSuccess && hipStreamCreateWithFlags(&ws, hipStreamNonBlocking) == hipSuccess;
  char *stage = nullptr;
  size_t stage_cap = 0;
  std::vector<double> ref;
  std::vector<int8_t> rtype;
  for (;;) {
    int r;
    {
      std::unique_lock<std::mutex> lk(S->mu);
      S->cv_job.wait(lk, [&] { return S->stop || !S->jobs.empty(); });
      if (S->jobs.empty()) break;
      r = S->jobs.front();
      S->jobs.pop_front();
    }
    int status;
    try {
      clear_error();
      status = ok ? load_one(S, r, ws, stage, stage_cap, ref, rtype) : MMD_E_HIP;
    } catch (const std::bad_alloc &) {
      status = MMD_E_NOMEM;
    } catch (...) {
      status = MMD_E_INVALID;
    }
    {
      std::lock_guard<std::mutex> lk(S->mu);
      (*S->R)[(size_t)r].status = status;
      S->done_q.push_back(r);
      S->in_flight--;
    }
    S->cv_done.notify_all();
  }
  if (stage) (void)hipHostFree(stage);
  if (ws) (void)hipStreamDestroy(ws);
}

int fill_lists(mmd_server *S, ListBuf &b, const std::vector<int> &reps, bool is_step) {
  HIP_TRY(hipStreamSynchronize(S->cs));  // the pinned images may still feed an earlier copy
  size_t n_site = 0, n_tile = 0, n_ext = 0;
  std::vector<RepLayout> lay(reps.size());
  for (size_t k = 0; k < reps.size(); k++) {
    const Rep &r = (*S->R)[(size_t)reps[k]];
    int rc = rep_layout(r.cells, S->o.ghost_cells, S->o.lattice_const, lay[k]);
    if (rc) return rc;
    n_site += (size_t)lay[k].nb;
    n_tile += (size_t)lay[k].nt;
    n_ext += (size_t)lay[k].n_ext_blocks;
  }
  if (n_site >= (size_t)INT32_MAX || n_tile >= (size_t)INT32_MAX || n_ext >= (size_t)INT32_MAX)
    return fail(MMD_E_RANGE, "serving: work list too long");
  // graphs hold the list pointers: they go before any list is freed or
  // regrown, so a failed allocation leaves no graph on freed lists
  if (is_step && (n_site > b.cap_site || n_tile > b.cap_tile || n_ext > b.cap_ext)) drop_graphs(S);
  int rc = grow(&b.h_site, &b.d_site, &b.cap_site, n_site);
  if (!rc) rc = grow(&b.h_tile, &b.d_tile, &b.cap_tile, n_tile);
  if (!rc) rc = grow(&b.h_ext, &b.d_ext, &b.cap_ext, n_ext);
  if (rc) return rc;
  size_t is = 0, it = 0, ie = 0;
  for (size_t k = 0; k < reps.size(); k++) {
    const int slot = (*S->R)[(size_t)reps[k]].slot;
    for (int64_t lb = 0; lb < lay[k].nb; lb++) b.h_site[is++] = make_int2(slot, (int)lb);
    for (int64_t tb = 0; tb < lay[k].nt; tb++) b.h_tile[it++] = make_int2(slot, (int)tb);
    for (int64_t eb = 0; eb < lay[k].n_ext_blocks; eb++) b.h_ext[ie++] = make_int2(slot, (int)eb);
    b.h_reps[k] = slot;
  }
  b.h_n[0] = (int)n_site, b.h_n[1] = (int)n_tile, b.h_n[2] = (int)n_ext, b.h_n[3] = (int)reps.size();
  if (is_step) {  // grids of the lists' capacities: the captured graphs stay valid for any mix that fits them
    b.k_site = (int64_t)b.cap_site, b.k_tile = (int64_t)b.cap_tile, b.k_ext = (int64_t)b.cap_ext;
  } else {  // priming: direct launches over the items
    b.k_site = (int64_t)n_site, b.k_tile = (int64_t)n_tile, b.k_ext = (int64_t)n_ext;
  }
  if (n_site) HIP_TRY(hipMemcpyAsync(b.d_site, b.h_site, n_site * sizeof(int2), hipMemcpyHostToDevice, S->cs));
  if (n_tile) HIP_TRY(hipMemcpyAsync(b.d_tile, b.h_tile, n_tile * sizeof(int2), hipMemcpyHostToDevice, S->cs));
  if (n_ext) HIP_TRY(hipMemcpyAsync(b.d_ext, b.h_ext, n_ext * sizeof(int2), hipMemcpyHostToDevice, S->cs));
  if (!reps.empty()) HIP_TRY(hipMemcpyAsync(b.d_reps, b.h_reps, reps.size() * sizeof(int), hipMemcpyHostToDevice, S->cs));
  HIP_TRY(hipMemcpyAsync(b.d_n, b.h_n, 4 * sizeof(int), hipMemcpyHostToDevice, S->cs));
  return MMD_OK;
}

// One batched step up to its force pass: the opening half kick and drift,
// fused with the previous step's closing half kick when `fused` (the same
// arithmetic as the two separate passes, as on the single-domain path).  The
// chunk's last step closes with step_close, which also records the energies.
int step_passes(mmd_server *S, const BatchLists &L, const ListBuf &B, bool fused, bool want_e) {
  const double dt = S->o.dt, h = 0.5 * S->o.dt, rc2 = S->o.cutoff * S->o.cutoff;
  const Masses &M = S->tab->masses;
  const int prec = S->o.precision;
  if (fused)
    HIP_TRY(launch_batch_kick_fused(S->cs, L, B.k_site, M, h, dt, false));
  else
    HIP_TRY(launch_batch_kick(S->cs, L, B.k_site, M, h, dt, true, false));
  HIP_TRY(launch_batch_ghost(S->cs, L, B.k_ext, MMD_F_X));
  HIP_TRY(launch_batch_rho_df(S->cs, L, B.k_tile, S->halo, S->tab->tf, S->tab->tm, rc2, prec, want_e, true));
  HIP_TRY(launch_batch_ghost(S->cs, L, B.k_ext, MMD_F_DF));
  HIP_TRY(launch_batch_force(S->cs, L, B.k_tile, S->halo, S->tab->tf, S->tab->tm, rc2, prec, want_e));
  return MMD_OK;
}

// the chunk's closing half kick and the energies of its last step
int step_close(mmd_server *S, const BatchLists &L, const ListBuf &B) {
  HIP_TRY(launch_batch_kick(S->cs, L, B.k_site, S->tab->masses, 0.5 * S->o.dt, S->o.dt, false, true));
  HIP_TRY(launch_batch_energy(S->cs, L, 0));
  return MMD_OK;
}

struct GpuExec final : Executor {
  mmd_server *S;
  explicit GpuExec(mmd_server *s) : S(s) {}
  double now() override {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - S->t0).count();
  }
  int start_load(int r) override {
    Rep &rep = (*S->R)[(size_t)r];
    RepLayout L;
    int rc = rep_layout(rep.cells, S->o.ghost_cells, S->o.lattice_const, L);
    if (rc) return rc;
    char *base = S->arena + rep.off;
    BatchDom &d = S->h_dom[rep.slot];
    d.g = L.g;
    d.pos = reinterpret_cast<double4 *>(base + L.pos);
    d.vx = reinterpret_cast<double *>(base + L.vx);
    d.vy = reinterpret_cast<double *>(base + L.vy);
    d.vz = reinterpret_cast<double *>(base + L.vz);
    d.fx = reinterpret_cast<double *>(base + L.fx);
    d.fy = reinterpret_cast<double *>(base + L.fy);
    d.fz = reinterpret_cast<double *>(base + L.fz);
    d.rho = reinterpret_cast<double *>(base + L.rho);
    d.df = reinterpret_cast<double *>(base + L.df);
    d.type = reinterpret_cast<int8_t *>(base + L.type);
    d.off = S->d_off_lds;
    d.nk_even = S->nk[0], d.nk_odd = S->nk[1];
    d.e_rho = reinterpret_cast<double *>(base + L.e_rho);
    d.e_force = reinterpret_cast<double *>(base + L.e_force);
    d.e_ke = reinterpret_cast<double *>(base + L.e_ke);
    d.energy = reinterpret_cast<double *>(base + L.energy);
    for (int k = 0; k < 3; k++) d.L[k] = L.L[k];
    d.n_blocks = L.nb;
    d.n_tiles = L.nt;
    HIP_TRY(hipMemcpyAsync(S->d_dom + rep.slot, &d, sizeof(BatchDom), hipMemcpyHostToDevice, S->cs));
    {
      std::lock_guard<std::mutex> lk(S->mu);
      S->jobs.push_back(r);
      S->in_flight++;
    }
    S->cv_job.notify_one();
    return MMD_OK;
  }
  int wait_loaded(std::vector<int> &out, double deadline) override {
    std::unique_lock<std::mutex> lk(S->mu);
    if (S->done_q.empty() && deadline > now()) {
      auto ready = [&] { return !S->done_q.empty(); };
      if (std::isinf(deadline))
        S->cv_done.wait(lk, ready);
      else
        S->cv_done.wait_until(lk, S->t0 + std::chrono::duration_cast<std::chrono::steady_clock::duration>(
                                              std::chrono::duration<double>(deadline)),
                              ready);
    }
    out.insert(out.end(), S->done_q.begin(), S->done_q.end());
    S->done_q.clear();
    return MMD_OK;
  }
  int prime(const std::vector<int> &reps) override {
    int rc = fill_lists(S, S->prime, reps, false);
    if (rc) return rc;
    const BatchLists L = S->prime.lists(S->d_dom);
    const double rc2 = S->o.cutoff * S->o.cutoff;
    const int prec = S->o.precision;
    const ListBuf &B = S->prime;
    HIP_TRY(launch_batch_ghost(S->cs, L, B.k_ext, MMD_F_X));
    HIP_TRY(launch_batch_rho_df(S->cs, L, B.k_tile, S->halo, S->tab->tf, S->tab->tm, rc2, prec, true, true));
    HIP_TRY(launch_batch_ghost(S->cs, L, B.k_ext, MMD_F_DF));
    HIP_TRY(launch_batch_force(S->cs, L, B.k_tile, S->halo, S->tab->tf, S->tab->tm, rc2, prec, true));
    HIP_TRY(launch_batch_kick(S->cs, L, B.k_site, S->tab->masses, 0.0, S->o.dt, false, true));
    HIP_TRY(launch_batch_energy(S->cs, L, 1));
    return fetch_energies(reps);
  }
  int fetch_energies(const std::vector<int> &reps) {
    for (size_t k = 0; k < reps.size(); k++) {
      const Rep &r = (*S->R)[(size_t)reps[k]];
      HIP_TRY(hipMemcpyAsync(S->h_e + 4 * k, S->h_dom[r.slot].energy, 4 * sizeof(double), hipMemcpyDeviceToHost,
                             S->cs));
    }
    HIP_TRY(hipStreamSynchronize(S->cs));
    for (size_t k = 0; k < reps.size(); k++) {
      Rep &r = (*S->R)[(size_t)reps[k]];
      for (int j = 0; j < 4; j++) r.e[j] = S->h_e[4 * k + j];
    }
    return MMD_OK;
  }
  int run(const std::vector<int> &active, int steps) override {
    int rc;
    if (active != S->step_set) {
      if ((rc = fill_lists(S, S->step, active, true))) return rc;
      S->step_set = active;
    }
    const BatchLists L = S->step.lists(S->d_dom);
    for (int s = 0; s < steps; s++) {
      const bool want_e = s == steps - 1, fused = s > 0;
      if (S->o.use_graph) {
        hipGraphExec_t &ge = S->graph[(fused ? 2 : 0) + (want_e ? 1 : 0)];
        if (!ge) {
          hipGraph_t gr;
          HIP_TRY(hipStreamBeginCapture(S->cs, hipStreamCaptureModeThreadLocal));
          rc = step_passes(S, L, S->step, fused, want_e);
          hipError_t ee = hipStreamEndCapture(S->cs, &gr);
          if (rc) {
            if (ee == hipSuccess) (void)hipGraphDestroy(gr);
            return rc;
          }
          HIP_TRY(ee);
          hipError_t ie = hipGraphInstantiate(&ge, gr, nullptr, nullptr, 0);
          (void)hipGraphDestroy(gr);
          HIP_TRY(ie);
          S->graph_captures++;
        }
        HIP_TRY(hipGraphLaunch(ge, S->cs));
      } else if ((rc = step_passes(S, L, S->step, fused, want_e))) {
        return rc;
      }
    }
    if (steps > 0 && (rc = step_close(S, L, S->step))) return rc;
    HIP_TRY(hipStreamSynchronize(S->cs));
    return MMD_OK;
  }
  int finish(const std::vector<int> &reps) override {
    std::vector<int> stepped;
    for (int r : reps)
      if ((*S->R)[(size_t)r].steps > 0) stepped.push_back(r);
    // replicas that stepped: their last chunk recorded energy[0..1]; admission energies were fetched at prime
    for (size_t k = 0; k < stepped.size(); k++) {
      const Rep &r = (*S->R)[(size_t)stepped[k]];
      HIP_TRY(hipMemcpyAsync(S->h_e + 4 * k, S->h_dom[r.slot].energy, 2 * sizeof(double), hipMemcpyDeviceToHost,
                             S->cs));
    }
    HIP_TRY(hipStreamSynchronize(S->cs));
    for (size_t k = 0; k < stepped.size(); k++) {
      Rep &r = (*S->R)[(size_t)stepped[k]];
      r.e[0] = S->h_e[4 * k], r.e[1] = S->h_e[4 * k + 1];
    }
    for (int r : reps) {
      const mmd_host_soa *keep = S->specs[r].keep;
      if (keep) {
        int rc = download(r, keep);
        if (rc) return rc;
      }
    }
    return MMD_OK;
  }
  // final state of replica r -> reference-layout host arrays
  int download(int r, const mmd_host_soa *h) {
    const Rep &rep = (*S->R)[(size_t)r];
    const BatchDom &d = S->h_dom[rep.slot];
    const Geom &g = d.g;
    const size_t n = (size_t)g.n_ext;
    std::vector<double4> P(n);
    std::vector<double> F(9 * n);
    std::vector<int8_t> T(n);
    HIP_TRY(hipMemcpyAsync(P.data(), d.pos, n * sizeof(double4), hipMemcpyDeviceToHost, S->cs));
    const double *src[9] = {d.vx, d.vy, d.vz, d.fx, d.fy, d.fz, d.rho, d.df, nullptr};
    for (int k = 0; k < 8; k++)
      HIP_TRY(hipMemcpyAsync(F.data() + k * n, src[k], n * sizeof(double), hipMemcpyDeviceToHost, S->cs));
    HIP_TRY(hipMemcpyAsync(T.data(), d.type, n, hipMemcpyDeviceToHost, S->cs));
    HIP_TRY(hipStreamSynchronize(S->cs));
    double *dst[8] = {h->vx, h->vy, h->vz, h->fx, h->fy, h->fz, h->rho, h->df};
    for (int64_t ze = 0; ze < g.ext_z; ze++)
      for (int64_t ye = 0; ye < g.ext_y; ye++)
        for (int64_t xe = 0; xe < g.ext_x; xe++) {
          const size_t i = (size_t)((ze * g.ext_y + ye) * g.ext_x + xe);
          const size_t j = (size_t)dev_index(g, (int32_t)(xe & 1), (int32_t)(xe >> 1), (int32_t)ye, (int32_t)ze);
          if (h->x) h->x[i] = P[j].x;
          if (h->y) h->y[i] = P[j].y;
          if (h->z) h->z[i] = P[j].z;
          for (int k = 0; k < 8; k++)
            if (dst[k]) dst[k][i] = F[k * n + j];
          if (h->type) h->type[i] = T[j];
        }
    return MMD_OK;
  }
};

void server_free(mmd_server *S) {
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stop = true;
  }
  S->cv_job.notify_all();
  for (auto &t : S->workers)
    if (t.joinable()) t.join();
  if (S->tab) {
    (void)hipSetDevice(S->o.device);
    if (S->cs) (void)hipStreamSynchronize(S->cs);
  }
  drop_graphs(S);
  free_lists(S->step);
  free_lists(S->prime);
  if (S->h_dom) (void)hipHostFree(S->h_dom);
  if (S->d_dom) (void)hipFree(S->d_dom);
  if (S->h_e) (void)hipHostFree(S->h_e);
  if (S->d_off_lds) (void)hipFree(S->d_off_lds);
  if (S->arena) (void)hipFree(S->arena);
  if (S->tab) (void)mmd_ctx_destroy(S->tab);
  delete S;
}

int check_opts(const mmd_server_opts *o) {
  if (!o) return fail(MMD_E_INVALID, "mmd_server_create: null options");
  if (o->precision != MMD_PREC_FP64 && o->precision != MMD_PREC_MIXED)
    return fail(MMD_E_INVALID, "mmd_server_create: precision %d", o->precision);
  if (!(o->lattice_const > 0.0) || !(o->cutoff > 0.0) || !(o->skin >= 0.0) || !(o->dt > 0.0) || o->ghost_cells < 1 ||
      !(o->sigma >= 0.0))
    return fail(MMD_E_INVALID, "mmd_server_create: lattice_const, cutoff, dt > 0; skin, sigma >= 0; ghost_cells >= 1");
  if (o->max_replicas < 1 || o->max_replicas > 4096 || o->chunk_steps < 1 || o->workers < 1 || o->workers > 64)
    return fail(MMD_E_INVALID, "mmd_server_create: max_replicas in [1, 4096], chunk_steps >= 1, workers in [1, 64]");
  if (!(o->hbm_fraction > 0.0) || o->hbm_fraction > 0.95)
    return fail(MMD_E_INVALID, "mmd_server_create: hbm_fraction in (0, 0.95]");
  return MMD_OK;
}

}  // namespace

// ================================================================== C ABI

extern "C" int mmd_server_replica_bytes(int32_t cells, int32_t ghost_cells, size_t *bytes) {
  GUARD({
    clear_error();
    if (!bytes) return fail(MMD_E_INVALID, "mmd_server_replica_bytes: null out");
    if (ghost_cells < 1 || cells < ghost_cells)
      return fail(MMD_E_INVALID, "mmd_server_replica_bytes: need 1 <= ghost_cells <= cells");
    RepLayout L;
    int rc = rep_layout(cells, ghost_cells, 1.0, L);
    if (rc) return rc;
    *bytes = L.bytes;
    return MMD_OK;
  })
}

extern "C" int mmd_server_simulate(const mmd_sim_params *p, const mmd_replica_spec *specs, int32_t n,
                                   mmd_replica_result *results, mmd_server_stats *stats) {
  GUARD({
    clear_error();
    if (!p || p->max_replicas < 1 || p->chunk_steps < 1 || p->ghost_cells < 1 || p->workers < 1 ||
        !(p->step_overhead_s >= 0.0) || !(p->step_atom_s >= 0.0) || !(p->load_atom_s >= 0.0) || (n > 0 && !results))
      return fail(MMD_E_INVALID, "mmd_server_simulate: bad parameters");
    int rc = check_specs(specs, n, p->ghost_cells);
    if (rc) return rc;
    std::vector<Rep> R;
    if ((rc = make_reps(specs, n, p->ghost_cells, 1.0, R))) return rc;
    SimExec X(*p, R);
    SchedStats st;
    rc = serve_loop(R, p->max_replicas, p->arena_bytes, p->chunk_steps, X, st, MMD_E_NOMEM);
    for (int32_t i = 0; i < n; i++) fill_result(R[(size_t)i], results[i]);
    fill_stats(R, st, ArenaAlloc(p->arena_bytes).size(), stats);
    if (rc < 0) return fail(MMD_E_STATE, "mmd_server_simulate: no progress possible");
    return rc;
  })
}

extern "C" int mmd_server_create(const mmd_server_opts *opts, const mmd_eam_tables *tab, mmd_server **out) {
  GUARD({
    clear_error();
    if (!out) return fail(MMD_E_INVALID, "mmd_server_create: null out");
    *out = nullptr;
    int rc = check_opts(opts);
    if (rc) return rc;
    mmd_server *S = new mmd_server();
    S->o = *opts;
    auto bail = [&](int code) {
      server_free(S);
      return code;
    };
    if ((rc = mmd_ctx_create(opts->device, &S->tab)) || (rc = mmd_set_precision(S->tab, opts->precision)) ||
        (rc = mmd_set_tables(S->tab, tab)))
      return bail(rc);
    S->cs = S->tab->stream;
    const int nt = S->tab->n_types;
    for (int t = 0; t + 1 < nt; t++)
      if (!(opts->type_cum[t] >= (t ? opts->type_cum[t - 1] : 0.0)) || opts->type_cum[t] > 1.0)
        return bail(fail(MMD_E_INVALID, "mmd_server_create: type_cum must be non-decreasing in [0, 1]"));
    if (!(opts->cutoff <= S->tab->r_max * (1.0 + 1e-12)))
      return bail(fail(MMD_E_INVALID, "mmd_server_create: cutoff %g beyond the tables (%g)", opts->cutoff,
                       S->tab->r_max));
    const int32_t cap = kMaxStencil - 8;
    for (int p = 0; p < 2; p++) {
      S->st3[p].assign(3 * (size_t)cap, 0);
      const int32_t k = mmd_stencil_build(opts->lattice_const, opts->cutoff + opts->skin, p, S->st3[p].data(), cap);
      if (k < 0) return bail(MMD_E_RANGE);
      S->nk[p] = k;
      for (int32_t j = 0; j < k; j++) {
        const int32_t *e = &S->st3[p][3 * (size_t)j];
        if (std::abs(e[0]) > 2 * opts->ghost_cells || std::abs(e[1]) > opts->ghost_cells ||
            std::abs(e[2]) > opts->ghost_cells)
          return bail(fail(MMD_E_RANGE, "mmd_server_create: the stencil (radius %g A) needs more than %d ghost cells",
                           opts->cutoff + opts->skin, opts->ghost_cells));
      }
    }
    {
      const int32_t *st3[2] = {S->st3[0].data(), S->st3[1].data()};
      S->halo = make_halo(st3, S->nk);
      if ((rc = check_tile_lds(opts->device, S->halo))) return bail(rc);
      const size_t need = tile_lds_bytes(S->halo, true);
      if (need > 65536 && batch_kernels_allow_lds(need) != hipSuccess)
        return bail(fail(MMD_E_RANGE, "mmd_server_create: LDS tile of %zu bytes refused", need));
      const std::vector<int32_t> off = stencil_image(st3, S->nk, S->halo);
      if (hipMalloc((void **)&S->d_off_lds, off.size() * sizeof(int32_t)) != hipSuccess)
        return bail(fail(MMD_E_NOMEM, "mmd_server_create: stencil offsets"));
      if (hipMemcpy(S->d_off_lds, off.data(), off.size() * sizeof(int32_t), hipMemcpyHostToDevice) != hipSuccess)
        return bail(fail(MMD_E_HIP, "mmd_server_create: stencil offsets upload"));
    }
    size_t free_b = 0, total_b = 0;
    HIP_TRY(hipMemGetInfo(&free_b, &total_b));
    size_t want = (size_t)((double)free_b * opts->hbm_fraction);
    if (opts->hbm_limit) want = std::min(want, opts->hbm_limit);
    want &= ~(size_t)255;
    if (want < 256) return bail(fail(MMD_E_NOMEM, "mmd_server_create: no device memory for the arena"));
    if (hipMalloc((void **)&S->arena, want) != hipSuccess) {
      (void)hipGetLastError();
      S->arena = nullptr;
      return bail(fail(MMD_E_NOMEM, "mmd_server_create: arena of %zu bytes refused", want));
    }
    S->arena_bytes = want;
    const size_t m = (size_t)opts->max_replicas;
    if (hipHostMalloc((void **)&S->h_dom, m * sizeof(BatchDom), hipHostMallocDefault) != hipSuccess ||
        hipMalloc((void **)&S->d_dom, m * sizeof(BatchDom)) != hipSuccess ||
        hipHostMalloc((void **)&S->h_e, 4 * m * sizeof(double), hipHostMallocDefault) != hipSuccess)
      return bail(fail(MMD_E_NOMEM, "mmd_server_create: slot tables"));
    std::memset(S->h_dom, 0, m * sizeof(BatchDom));
    for (ListBuf *b : {&S->step, &S->prime}) {
      if (hipHostMalloc((void **)&b->h_reps, m * sizeof(int), hipHostMallocDefault) != hipSuccess ||
          hipMalloc((void **)&b->d_reps, m * sizeof(int)) != hipSuccess ||
          hipHostMalloc((void **)&b->h_n, 4 * sizeof(int), hipHostMallocDefault) != hipSuccess ||
          hipMalloc((void **)&b->d_n, 4 * sizeof(int)) != hipSuccess)
        return bail(fail(MMD_E_NOMEM, "mmd_server_create: work lists"));
    }
    for (int w = 0; w < opts->workers; w++) S->workers.emplace_back(worker_main, S);
    *out = S;
    return MMD_OK;
  })
}

extern "C" int mmd_server_destroy(mmd_server *S) {
  GUARD({
    if (S) server_free(S);
    return MMD_OK;
  })
}

extern "C" int mmd_server_run(mmd_server *S, const mmd_replica_spec *specs, int32_t n, mmd_replica_result *results,
                              mmd_server_stats *stats) {
  GUARD({
    clear_error();
    if (!S || (n > 0 && !results)) return fail(MMD_E_INVALID, "mmd_server_run: bad arguments");
    HIP_TRY(hipSetDevice(S->o.device));
    int rc = check_specs(specs, n, S->o.ghost_cells);
    if (rc) return rc;
    std::vector<Rep> R;
    if ((rc = make_reps(specs, n, S->o.ghost_cells, S->o.lattice_const, R))) return rc;
    S->R = &R;
    S->specs = specs;
    S->step_set.clear();
    S->done_q.clear();
    S->t0 = std::chrono::steady_clock::now();
    const int64_t captures0 = S->graph_captures;
    GpuExec X(S);
    SchedStats st;
    rc = serve_loop(R, S->o.max_replicas, S->arena_bytes, S->o.chunk_steps, X, st, MMD_E_NOMEM);
    {
      // after an error, loads not yet started are dropped and running ones
      // are waited for: the workers must not touch R once it goes away
      std::unique_lock<std::mutex> lk(S->mu);
      S->in_flight -= (int)S->jobs.size();
      S->jobs.clear();
      S->cv_done.wait(lk, [&] { return S->in_flight == 0; });
      S->done_q.clear();
    }
    (void)hipStreamSynchronize(S->cs);
    for (int32_t i = 0; i < n; i++) fill_result(R[(size_t)i], results[i]);
    fill_stats(R, st, S->arena_bytes, stats);
    if (stats) stats->graph_captures = S->graph_captures - captures0;
    S->R = nullptr;
    S->specs = nullptr;
    S->step_set.clear();
    if (rc < 0) return fail(MMD_E_STATE, "mmd_server_run: no progress possible");
    return rc;
  })
}
