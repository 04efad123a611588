/*
 * mmd_eam.h -- C ABI of the MI355X-native EAM hot path (rho, dF/drho, force)
 * of MISA-MD, device-resident, for one GPU running one rank's sub-domain plus
 * its ghost shell.
 *
 * Every entry point returns MMD_OK (0) or a positive MMD_E_* status; no C++
 * exception crosses this boundary.  mmd_last_error() gives a message for the
 * last failing call on the calling thread.  A context owns one HIP stream and
 * all device memory of one sub-domain; different contexts are independent.
 * Calls on one context must not overlap.
 *
 * Host-visible arrays always use the reference plugin's layout: one array over
 * the ghost-extended lattice, x doubled, linear index
 *   ((z * ext_y) + y) * ext_x + x          (src/kernels/kernel_itl.hpp:21-26)
 * The device keeps its own layout (sub-lattice split rows); uploads and
 * downloads convert on the device.
 */
#ifndef MMD_EAM_H
#define MMD_EAM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------ status codes */
#define MMD_OK 0
#define MMD_E_INVALID 1   /* bad argument or inconsistent sizes             */
#define MMD_E_STATE 2     /* call order: domain, stencil or tables missing  */
#define MMD_E_HIP 3       /* HIP runtime error                              */
#define MMD_E_NOMEM 4     /* device or host allocation failed / refused     */
#define MMD_E_RANGE 5     /* size beyond a stated limit                     */
#define MMD_E_NONFINITE 6 /* non-finite value detected in state             */

/* field masks for upload / download / ghost refresh */
#define MMD_F_X 0x01u    /* positions x, y, z        */
#define MMD_F_V 0x02u    /* velocities               */
#define MMD_F_F 0x04u    /* forces                   */
#define MMD_F_RHO 0x08u  /* electron density         */
#define MMD_F_DF 0x10u   /* F'(rho)                  */
#define MMD_F_TYPE 0x20u /* atom type (int8, <0 = vacancy) */
#define MMD_F_ALL 0x3Fu

/* precision modes */
#define MMD_PREC_FP64 0  /* FP64 throughout, r-keyed splines (reference arithmetic) */
#define MMD_PREC_MIXED 1 /* FP64 displacements/accumulation, r^2-keyed splines with
                            FP64 leading term + FP32 correction polynomial          */

typedef struct mmd_ctx mmd_ctx;

/* Sub-domain geometry in lattice units with x doubled.  Mirrors
 * _hipDeviceDomain (src/kernel_types.h:51-64) as filled by hip_domain_init
 * (src/hip_arch_acc_imp.cpp:61-72): ghost = local_low - ghost_low, box = local
 * size, ext = ghost-extended size, start = global index of the first local
 * site. */
typedef struct {
  int32_t ghost[3];
  int32_t box[3];
  int32_t ext[3];
  int64_t start[3];
} mmd_domain;

/* Tabulated EAM potential (replaces the eam tables hip_pot_init copies,
 * src/hip_arch_acc_imp.cpp:146-150).  Knot values on uniform grids: rho_t(r)
 * and phi_ab(r) on r0 + i*dr (i < n_r), F_t(rho) on rho0 + i*drho (i < n_rho).
 * phi must be symmetric (phi_ab == phi_ba).  Masses in amu. */
typedef struct {
  int32_t n_types;
  int32_t n_r;
  double r0, dr;
  int32_t n_rho;
  double rho0, drho;
  const double *rho_r; /* [n_types][n_r]           */
  const double *emb;   /* [n_types][n_rho]         */
  const double *phi_r; /* [n_types][n_types][n_r]  */
  const double *mass;  /* [n_types]                */
} mmd_eam_tables;

/* Host SoA view in the reference layout (any pointer may be NULL when its
 * field is not in the mask).  Mirrors device_atoms::_type_atom_list_soa
 * (src/memory/device_atoms.h:21-28) with planar x/y/z. */
typedef struct {
  double *x, *y, *z;
  double *vx, *vy, *vz;
  double *fx, *fy, *fz;
  double *rho, *df;
  int8_t *type;
} mmd_host_soa;

/* Byte layout of one host AoS record (the reference's _cuAtomElement,
 * src/kernel_types.h:40-48: id@0 type@8 x@16 v@40 f@64 rho@88 df@96, 104 B). */
typedef struct {
  size_t stride;
  size_t off_type, off_x, off_v, off_f, off_rho, off_df;
  int32_t type_bytes; /* 1, 2 or 4 (signed) */
} mmd_aos_layout;

/* ------------------------------------------------------------- lifecycle */

/* hip_env_init (src/hip_arch_acc_imp.cpp:28-40): binds the context to
 * `device` (the reference always used device 0) and creates its stream. */
int mmd_ctx_create(int32_t device, mmd_ctx **out);
/* Number of visible GPUs (a rank picks its device from it). */
int mmd_device_count(int32_t *n);
/* hip_env_clean (src/hip_arch_acc_imp.cpp:42-45): frees everything the
 * context owns (the reference leaked its buffers). */
int mmd_ctx_destroy(mmd_ctx *ctx);
const char *mmd_last_error(void);

/* hip_domain_init (src/hip_arch_acc_imp.cpp:47-99).  Allocates the
 * device-resident ghost-extended SoA state, zero-filled, types = -1.
 * Requires even ext/ghost/start along x and ext = box + 2*ghost, and fewer
 * than 2^31 ext sites (MMD_E_RANGE; about 300 GB of state, more than one
 * MI355X holds).  A new domain needs mmd_md_setup again before MD calls. */
int mmd_set_domain(mmd_ctx *ctx, const mmd_domain *dom);
/* hip_nei_offset_init (src/hip_arch_acc_imp.cpp:101-144): linear offsets in
 * the reference layout for even / odd global x parity (NeiOffset values). */
int mmd_set_stencil(mmd_ctx *ctx, const int64_t *even, int32_t n_even, const int64_t *odd, int32_t n_odd);
/* hip_pot_init (src/hip_arch_acc_imp.cpp:146-150): builds the FP64 spline
 * tables (and the mixed-precision r^2-keyed tables) on the device.
 * MMD_E_RANGE when n_types^2 * n_r reaches 2^24 (pair-table record limit). */
int mmd_set_tables(mmd_ctx *ctx, const mmd_eam_tables *tab);
int mmd_set_precision(mmd_ctx *ctx, int32_t mode);
/* Where MMD_PREC_MIXED evaluates pairs from its r^2-keyed tables, as squared
 * distances: out[0..1] = [head, tail), out[2..5] = two interior holes
 * [lo, hi) (empty as [0, 0)), all on table knots (whole segments); pairs
 * elsewhere use the FP64 spline.  No
 * counterpart in the reference (FP64 only); for callers that check how much
 * of their workload the fast path covers. */
int mmd_mixed_ranges(mmd_ctx *ctx, double out[6]);

/* ------------------------------------------------------------ data movement */

/* Copies the masked fields of all ext sites host -> device (reference layout
 * on the host).  Synchronous.  Coordinates must lie within 1e6 A of the
 * origin: the pair passes stage vacant sites at a far point (1e7 A on every
 * axis) so that the cutoff test alone screens them. */
int mmd_upload(mmd_ctx *ctx, uint32_t fields, const mmd_host_soa *h);
/* Device -> host for all ext sites.  Synchronous. */
int mmd_download(mmd_ctx *ctx, uint32_t fields, const mmd_host_soa *h);

/* --------------------------------------------------------- EAM hot path */

/* hip_eam_rho_calc (src/hip_arch_acc_imp.cpp:163-181) on device-resident
 * state: rho and the fused df = F'(rho) for every local site with type >= 0
 * (overwritten; vacancies untouched), then interstitial contributions when an
 * interstitial list is set.  Asynchronous on the context stream. */
int mmd_eam_rho_df(mmd_ctx *ctx, double cutoff);
/* hip_eam_df_calc (src/hip_arch_acc_imp.cpp:183-205): standalone
 * df = F'(rho) from the device rho (the reference's unfused calDf). */
int mmd_eam_df(mmd_ctx *ctx);
/* hip_eam_force_calc (src/hip_arch_acc_imp.cpp:207-225): overwrites f of
 * local sites with type >= 0.  Ghost df must be current. */
int mmd_eam_force(mmd_ctx *ctx, double cutoff);
/* Blocks until the context stream is idle. */
int mmd_sync(mmd_ctx *ctx);

/* Drop-in host-array forms of the three hooks: upload x/type (and df for the
 * force pass) from a host AoS array of ext records, run the pass, and write
 * rho/df (or f) back into the local records, as the reference's streamed
 * passes do (src/double-buffer/rho_double_buffer_imp.cpp:81-121,
 * force_double_buffer_imp.cpp:78-126).  Synchronous.  A record whose type
 * has no table fails the call with MMD_E_INVALID; on pinned records
 * (mmd_host_alloc), which are processed in z-chunks with the outputs written
 * as each chunk's rows finish, some local records' output fields may already
 * hold this pass's values when it fails (other host memory is left
 * unchanged). */
int mmd_eam_rho_calc_aos(mmd_ctx *ctx, void *atoms, const mmd_aos_layout *lay, double cutoff);
int mmd_eam_df_calc_aos(mmd_ctx *ctx, void *atoms, const mmd_aos_layout *lay);
int mmd_eam_force_calc_aos(mmd_ctx *ctx, void *atoms, const mmd_aos_layout *lay, double cutoff);

/* hip_create_atoms_mem / hip_release_atoms_mem (src/mem.cpp:14-32): pinned
 * host allocation for the caller's lattice array. */
int mmd_host_alloc(void **out, size_t bytes);
int mmd_host_free(void *p);

/* ------------------------------------------------ device-resident MD step */

typedef struct {
  double box_len[3];    /* periodic box lengths in A (single-domain ghost refresh) */
  double lattice_const; /* A, for interstitial anchors                             */
  double dt;            /* fs                                                      */
  double cutoff;        /* A                                                       */
  int32_t energy_every; /* record energies every n steps (0 = never)               */
} mmd_md_params;

int mmd_md_setup(mmd_ctx *ctx, const mmd_md_params *p);
/* The reference's Newton's-third-law half-stencil mode (USE_NEWTONS_THIRD_LOW,
 * src/hip_arch_acc_imp.cpp:118-130, src/kernels/kernel_pairs.hpp:28-31,
 * :55-59), rebuilt deterministically for the periodic step (mmd_md_prime,
 * mmd_md_run, mmd_energy): each lattice pair's force is evaluated once and
 * its reaction pulled by the neighbour in a second pass, no atomics.  Needs a
 * symmetric stencil (MMD_E_INVALID otherwise); takes max-forward-entries x
 * local sites doubles of device memory; mmd_set_stencil and mmd_set_domain
 * turn it off.  mmd_eam_force and the drop-in hooks keep the full stencil
 * (their ghost shell belongs to other sub-domains). */
int mmd_set_half_stencil(mmd_ctx *ctx, int32_t on);
/* Periodic single-domain ghost refresh of MMD_F_X (positions + types) and/or
 * MMD_F_DF. */
int mmd_ghost_refresh(mmd_ctx *ctx, uint32_t fields);
/* Runs nsteps velocity-Verlet steps (kick+drift, ghost x, rho+df, ghost df,
 * force, kick) fully on the device.  Forces must be current on entry (call
 * mmd_md_prime once after uploading a new state).  energies, if not NULL,
 * receives [nsteps][2] {potential, kinetic} in eV for the recorded steps
 * (others are left untouched).  use_graph != 0 captures one step into a HIP
 * graph on first use and replays it. */
int mmd_md_prime(mmd_ctx *ctx);
int mmd_md_run(mmd_ctx *ctx, int32_t nsteps, double *energies, int32_t use_graph);
/* mmd_md_run (eager launches, no graph) that allocates nothing: no device or
 * host allocation, no graph instantiation, no event creation.  No counterpart
 * in the reference plugin; it is for callers that cannot allocate memory
 * while they step (a serving loop, a harness that forbids allocation).  The
 * per-block energy partials, the energy records and their counter live in a
 * library-owned device array of 2^20 doubles (8 MiB) per GPU, so
 *   2 * pair-pass tiles + site blocks + interstitial partials
 *     + 2 * max(nsteps / energy_every, 1) + 1
 * must fit in 2^20 doubles (MMD_E_RANGE otherwise, before any launch; a
 * 256^3-cell box with 0.5 % interstitials takes about 0.5 M).  That array
 * is shared by every call on one GPU: calls that overlap on one GPU race.
 * Interstitial bins must already be sized for the box (mmd_md_prime does it;
 * MMD_E_STATE otherwise).  Synchronous on return; not for stream capture.
 * Results equal mmd_md_run's bit for bit. */
int mmd_md_run_noalloc(mmd_ctx *ctx, int32_t nsteps, double *energies);
/* Potential and kinetic energy of the current state (forces current). */
int mmd_energy(mmd_ctx *ctx, double *epot, double *ekin);

/* ------------------------------------------------------ interstitial list */

/* Sets the off-lattice atom list (positions, velocities, types >= 0) and the
 * wide stencil W of 3-D doubled-x offsets for parity 0 (parity 1 uses the
 * negated entries) that covers every lattice site within reach of an
 * interstitial anchored at its nearest lattice site.  cap bounds the list. */
int mmd_itl_set(mmd_ctx *ctx, int32_t n, const double *x, const double *y, const double *z, const double *vx,
                const double *vy, const double *vz, const int8_t *type, const int32_t *wide3, int32_t n_wide,
                int32_t cap);
/* Wigner-Seitz re-assignment between the lattice and the interstitial list of
 * a single periodic domain (start = 0), after mmd_md_setup and mmd_itl_set.
 * No counterpart in the reference plugin, which skips off-lattice atoms
 * (src/kernels/kernel_itl.hpp:44-46, :52-54); it stands in for MISA-MD core's
 * lattice <-> interstitial bookkeeping between steps.  Rules:
 *   1. in increasing local site order (z, y, doubled x), a lattice atom whose
 *      nearest site (its Wigner-Seitz cell; ties to the corner sub-lattice) is
 *      not its own joins the end of the list with every field; its site
 *      becomes a vacancy (type -1; v, f, rho, df zeroed; position kept);
 *   2. a list entry whose nearest site, wrapped into the box, is vacant moves
 *      into it (position shifted by the periodic image); the lowest list index
 *      wins a contested site; the entries that moved leave the list, the rest
 *      keep their order.
 * Forces, rho and df travel with the atoms, so a run can continue without a
 * new force pass.  The ghost shell's positions, types and df are refreshed.
 * MMD_E_RANGE, with nothing changed, when the list would exceed its capacity.
 * *n_to_itl / *n_to_lattice receive the counts of each transition. */
int mmd_itl_migrate(mmd_ctx *ctx, int32_t *n_to_itl, int32_t *n_to_lattice);
/* Copies interstitial fields to the host (any pointer may be NULL). */
int mmd_itl_get(mmd_ctx *ctx, int32_t *n, double *x, double *y, double *z, double *vx, double *vy, double *vz,
                double *fx, double *fy, double *fz, double *rho, double *df, int8_t *type);

/* ----------------------------------------------------------- serving loop */

/* Continuous batching of independent MD replicas on one GPU (no counterpart
 * in the reference plugin, which runs one sub-domain per call).  Replicas
 * arrive over time; each is admitted when a slot and device memory (an
 * arena carved from free HBM at creation) are available, its lattice stays
 * resident, and all active replicas advance together through batched
 * launches over (replica, block) work items, replayed from a captured HIP
 * graph.  A replica's results
 * are identical whether it runs alone or in a mix. */
typedef struct mmd_server mmd_server;

typedef struct {
  int32_t device;
  int32_t precision;     /* MMD_PREC_FP64 or MMD_PREC_MIXED                       */
  double lattice_const;  /* A                                                     */
  double cutoff, skin;   /* A; stencil radius = cutoff + skin                      */
  double dt;             /* fs                                                    */
  int32_t ghost_cells;
  double type_cum[8];    /* cumulative type fractions of the random alloy         */
  double sigma;          /* initial Gaussian displacement, A                      */
  int32_t max_replicas;  /* concurrent replicas                                   */
  double hbm_fraction;   /* fraction of free HBM taken for the replica arena       */
  size_t hbm_limit;      /* 0 = none; else upper bound on the arena in bytes       */
  int32_t chunk_steps;   /* steps per scheduling round                            */
  int32_t workers;       /* host threads preparing initial conditions             */
  int32_t use_graph;     /* replay the batched step from a captured HIP graph     */
} mmd_server_opts;

typedef struct {
  uint64_t id;
  int32_t cells;           /* cubic box edge in BCC cells                          */
  double temperature;      /* K                                                    */
  int32_t steps;
  uint64_t seed;
  double arrival_s;        /* arrival time after the start of mmd_server_run       */
  const mmd_host_soa *keep; /* optional: receives the final ext state (reference layout) */
} mmd_replica_spec;

typedef struct {
  uint64_t id;
  int32_t status;          /* MMD_OK, or why the replica failed or was rejected    */
  int32_t n_atoms;
  double t_arrival;        /* seconds after the start of the run                   */
  double t_admit;          /* slot and memory granted, preparation starts          */
  double t_ready;          /* state resident, forces primed, joins the batch       */
  double t_done;
  double epot0, ekin0;     /* energies at admission (eV)                           */
  double epot, ekin;       /* energies after the last step (eV)                    */
  int32_t slot;            /* batch slot it occupied (-1: never admitted)          */
  uint64_t arena_offset, arena_bytes; /* device memory it occupied                 */
} mmd_replica_result;

typedef struct {
  double wall_s;
  double atom_steps;
  double atom_steps_per_s;
  double latency_p50_s, latency_p99_s, latency_mean_s;
  double wait_mean_s;      /* arrival -> admission                                  */
  int32_t completed, failed, max_active;
  int64_t batched_steps;
  double arena_bytes;
  int64_t graph_captures;  /* step graphs captured during the run (0 without graphs) */
} mmd_server_stats;

int mmd_server_create(const mmd_server_opts *opts, const mmd_eam_tables *tab, mmd_server **out);
int mmd_server_destroy(mmd_server *srv);
/* Runs a whole arrival trace in real time; returns when every replica has
 * completed or failed.  results[i] belongs to specs[i]. */
int mmd_server_run(mmd_server *srv, const mmd_replica_spec *specs, int32_t n, mmd_replica_result *results,
                   mmd_server_stats *stats);
/* Device bytes a replica of `cells` occupies (the admission size). */
int mmd_server_replica_bytes(int32_t cells, int32_t ghost_cells, size_t *bytes);

/* The same scheduling loop driven by a simulated executor: a virtual clock,
 * a fake device-memory budget and a linear cost model, no GPU.  Used to test
 * the admission and batching policy. */
typedef struct {
  size_t arena_bytes;
  int32_t max_replicas, chunk_steps, ghost_cells, workers;
  double step_overhead_s; /* per batched step                                   */
  double step_atom_s;     /* per atom-step                                      */
  double load_atom_s;     /* per atom of initial-condition preparation          */
} mmd_sim_params;
int mmd_server_simulate(const mmd_sim_params *p, const mmd_replica_spec *specs, int32_t n,
                        mmd_replica_result *results, mmd_server_stats *stats);

/* ---------------------------------------------------------------- timing */

#define MMD_K_RHO 0
#define MMD_K_FORCE 1
#define MMD_K_INTEGRATE 2
#define MMD_K_GHOST 3
#define MMD_K_ITL 4
#define MMD_K_COUNT 5
/* Enables HIP-event timing of each kernel class on the context stream. */
int mmd_timing_enable(mmd_ctx *ctx, int32_t on);
/* Total milliseconds and launch counts per kernel class since the last reset
 * (synchronises the stream). */
int mmd_timing_get(mmd_ctx *ctx, double *ms_total, int64_t *launches);
int mmd_timing_reset(mmd_ctx *ctx);

/* Streaming device-to-device copy of `bytes` (multiple of 16) on the context's
 * stream with a plain 16-byte-per-lane kernel; *ms receives its duration
 * from HIP events.  Used to measure achievable HBM bandwidth and to calibrate
 * the hardware traffic counters against a kernel of known traffic. */
int mmd_stream_copy(mmd_ctx *ctx, size_t bytes, int32_t repeats, double *ms);

/* ------------------------------------------------- host helpers (no GPU) */

/* Neighbour stencil of one parity in 3-D doubled-x offsets (dX,dY,dZ): all
 * sites with ideal BCC distance < radius, ordered by distance then dZ, dY, dX;
 * parity 1 = negated parity-0 entries.  Replaces MISA-MD's NeighbourIndex
 * generation.  Returns the count, or -1 if cap is too small. */
int32_t mmd_stencil_build(double lattice_const, double radius, int32_t parity, int32_t *out3, int32_t cap);
/* 3-D offsets -> linear offsets in the reference layout of `dom`. */
int mmd_stencil_linear(const mmd_domain *dom, const int32_t *in3, int32_t n, int64_t *out);

/* Synthetic EAM tables: phi(r)=A(rc-r)^2(1+B/r), rho(r)=(rc-r)^2 exp(-r),
 * F(rho)=-sqrt(rho)+C rho^2 with A,B per unordered type pair and C per type
 * jittered by +-jitter (relative) from seed; jitter 0 gives A,B,C exactly. */
typedef struct {
  int32_t n_types;
  double rc, A, B, C, jitter;
  uint64_t seed;
  int32_t n_r;
  double r0;
  int32_t n_rho;
  double rho_max;
} mmd_synth_params;
int mmd_synth_tables(const mmd_synth_params *p, double *rho_r, double *emb, double *phi_r, double *r0, double *dr,
                     double *rho0, double *drho);

/* Finnis-Sinclair tables for BCC Fe (Phil. Mag. A 50 (1984) 45): pair
 * V(r) = (r-c)^2 (c0 + c1 r + c2 r^2) below c = 3.40 A, density
 * (r-d)^2 + 1.8 (r-d)^3/d below d = 3.569745 A, F(rho) = -A sqrt(rho); zero
 * beyond, so every pair inside the passes' cutoff is evaluated while the
 * lattice stays mechanically stable (a = 2.8665 A, E_coh = 4.28 eV).  For
 * n_types > 1, A per type and the pair term per unordered type pair are scaled
 * by 1 +- jitter (seeded).  Knots on r0 + i*dr up to rc, rho = i*drho. */
typedef struct {
  int32_t n_types;
  double rc, jitter;
  uint64_t seed;
  int32_t n_r;
  double r0;
  int32_t n_rho;
  double rho_max;
} mmd_fs_params;
int mmd_fs_tables(const mmd_fs_params *p, double *rho_r, double *emb, double *phi_r, double *r0, double *dr,
                  double *rho0, double *drho);

/* EAM tables from a DYNAMO setfl file (LAMMPS eam/alloy format), the on-disk
 * source of the tables hip_pot_init copies (src/hip_arch_acc_imp.cpp:146-150).
 * mmd_setfl_info gives the sizes; mmd_setfl_read fills caller arrays shaped
 * as mmd_eam_tables expects (phi = (r phi) / r on the r grid, phi symmetric). */
int mmd_setfl_info(const char *path, int32_t *n_types, int32_t *n_r, int32_t *n_rho, double *cutoff);
int mmd_setfl_read(const char *path, double *rho_r, double *emb, double *phi_r, double *mass, double *r0, double *dr,
                   double *rho0, double *drho);
/* EAM tables from DYNAMO funcfl files (LAMMPS pair_style eam), one file per
 * element type, the other on-disk source of those tables: F(rho), rho(r) and
 * the effective charge Z(r), with phi_ij = 27.2 * 0.529 * Z_i(r) Z_j(r) / r
 * (eV, angstrom).  The files must share one grid (mmd_funcfl_info checks it
 * and gives the sizes); mmd_funcfl_read fills the arrays mmd_setfl_read does,
 * n_types = n_files. */
int mmd_funcfl_info(const char *const *paths, int32_t n_files, int32_t *n_r, int32_t *n_rho, double *cutoff);
int mmd_funcfl_read(const char *const *paths, int32_t n_files, double *rho_r, double *emb, double *phi_r,
                    double *mass, double *r0, double *dr, double *rho0, double *drho);

/* Initial conditions for a periodic single-domain BCC box of cells[3] unit
 * cells: types drawn from `type_frac` (cumulative, n_types entries), Gaussian
 * displacement sigma, Maxwell velocities at temperature (COM removed),
 * vacancy/interstitial fractions at random sites (interstitials at
 * tetrahedral sites), optionally clustered in n_blobs Gaussian blobs.  All
 * draws come from a counter-based generator keyed by (seed, global site id).
 * Fills the ext arrays in the reference layout (ghost shell refreshed) and the
 * interstitial arrays (capacity itl_cap; *n_itl receives the count). */
typedef struct {
  int32_t cells[3];
  int32_t ghost_cells;
  double lattice_const;
  int32_t n_types;
  double type_cum[8];
  double mass[8];
  double sigma;
  double temperature;
  double vacancy_frac;
  double interstitial_frac;
  int32_t n_blobs;
  double blob_sigma;
  uint64_t seed;
} mmd_lattice_spec;
int mmd_lattice_domain(const mmd_lattice_spec *s, mmd_domain *dom, double box_len[3]);
int mmd_lattice_init(const mmd_lattice_spec *s, const mmd_host_soa *h, int32_t itl_cap, int32_t *n_itl,
                     double *ix, double *iy, double *iz, double *ivx, double *ivy, double *ivz, int8_t *itype);

#ifdef __cplusplus
}
#endif
#endif
