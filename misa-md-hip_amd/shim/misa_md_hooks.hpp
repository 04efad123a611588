// Adapter from MISA-MD's arch-acceleration hooks (the hip_* functions the
// core calls, declared in MISA-MD's arch/arch_imp.h) to the C ABI in
// mmd_eam.h.  Header-only templates over MISA-MD's types, so the conversions
// are testable without the MISA-MD tree; misa_md_hooks.cpp instantiates them
// with the real types and defines the hooks themselves.
//
// State is process-global, as the reference plugin's is
// (src/hip_arch_acc_imp.cpp:24-26): one context per rank.  Errors print the
// library's message and abort, the reference's HIP_CHECK behaviour.
#ifndef MMD_MISA_MD_HOOKS_HPP
#define MMD_MISA_MD_HOOKS_HPP

#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <type_traits>
#include <vector>

#include "mmd_eam.h"

namespace mmd_shim {

struct HookState {
  mmd_ctx *ctx = nullptr;
  mmd_aos_layout layout{};
};

inline HookState &state() {
  static HookState s;
  return s;
}

inline void check(int rc, const char *what) {
  if (rc != MMD_OK) {
    std::fprintf(stderr, "mmd: %s failed (status %d): %s\n", what, rc, mmd_last_error());
    std::abort();
  }
}

// Device of this rank: MMD_DEVICE if set, else the node-local MPI rank modulo
// the device count (the reference always took device 0,
// src/hip_arch_acc_imp.cpp:31).
inline int32_t device_from_env(int32_t n_devices) {
  if (const char *d = std::getenv("MMD_DEVICE")) return (int32_t)std::atoi(d);
  for (const char *v : {"OMPI_COMM_WORLD_LOCAL_RANK", "MPI_LOCALRANKID", "SLURM_LOCALID"})
    if (const char *r = std::getenv(v)) return n_devices > 0 ? (int32_t)(std::atoi(r) % n_devices) : 0;
  return 0;
}

// hip_domain_init (src/hip_arch_acc_imp.cpp:47-99): ghost = local low - ghost
// low, box = local size, ext = ghost-extended size, start = local low.
template <class BccDomain>
mmd_domain domain_of(const BccDomain &d) {
  mmd_domain m{};
  const int64_t lo_local[3] = {d.dbx_sub_box_lattice_region.x_low, d.dbx_sub_box_lattice_region.y_low,
                               d.dbx_sub_box_lattice_region.z_low};
  const int64_t lo_ghost[3] = {d.dbx_ghost_ext_lattice_region.x_low, d.dbx_ghost_ext_lattice_region.y_low,
                               d.dbx_ghost_ext_lattice_region.z_low};
  for (int k = 0; k < 3; k++) {
    m.ghost[k] = (int32_t)(lo_local[k] - lo_ghost[k]);
    m.box[k] = (int32_t)d.dbx_sub_box_lattice_size[k];
    m.ext[k] = (int32_t)d.dbx_ghost_extended_lattice_size[k];
    m.start[k] = lo_local[k];
  }
  return m;
}

// hip_nei_offset_init (src/hip_arch_acc_imp.cpp:101-144), full stencil (the
// default build; the half-stencil variant is not used).
template <class NeighbourIndex>
void stencil_of(const NeighbourIndex &n, std::vector<int64_t> &even, std::vector<int64_t> &odd) {
  even.assign(n.nei_even_offsets.begin(), n.nei_even_offsets.end());
  odd.assign(n.nei_odd_offsets.begin(), n.nei_odd_offsets.end());
}

// Byte layout of the core's atom record (the reference mirrors it as
// _cuAtomElement, src/kernel_types.h:40-48: id, type, x[3], v[3], f[3], rho, df).
template <class Record>
mmd_aos_layout layout_of() {
  mmd_aos_layout l{};
  Record *r = nullptr;
  auto at = [&](const void *p) { return (size_t)((const char *)p - (const char *)r); };
  l.stride = sizeof(Record);
  l.off_type = at(&r->type);
  l.off_x = at(&r->x[0]);
  l.off_v = at(&r->v[0]);
  l.off_f = at(&r->f[0]);
  l.off_rho = at(&r->rho);
  l.off_df = at(&r->df);
  l.type_bytes = (int32_t)sizeof(r->type);
  return l;
}

// Host tables in the layout mmd_set_tables takes.
struct HostTables {
  int32_t n_types = 0, n_r = 0, n_rho = 0;
  double r0 = 0.0, dr = 0.0, rho0 = 0.0, drho = 0.0;
  std::vector<double> rho_r, emb, phi_r, mass;
  mmd_eam_tables view() const {
    mmd_eam_tables t{};
    t.n_types = n_types, t.n_r = n_r, t.r0 = r0, t.dr = dr, t.n_rho = n_rho, t.rho0 = rho0, t.drho = drho;
    t.rho_r = rho_r.data(), t.emb = emb.data(), t.phi_r = phi_r.data(), t.mass = mass.data();
    return t;
  }
  void load_setfl(const char *path) {
    double rc = 0.0;
    check(mmd_setfl_info(path, &n_types, &n_r, &n_rho, &rc), "mmd_setfl_info");
    rho_r.assign((size_t)n_types * n_r, 0.0);
    emb.assign((size_t)n_types * n_rho, 0.0);
    phi_r.assign((size_t)n_types * n_types * n_r, 0.0);
    mass.assign((size_t)n_types, 0.0);
    check(mmd_setfl_read(path, rho_r.data(), emb.data(), phi_r.data(), mass.data(), &r0, &dr, &rho0, &drho),
          "mmd_setfl_read");
  }
  // list = funcfl files separated by ':', one per element type in type order
  void load_funcfl(const char *list) {
    std::vector<std::string> files;
    std::string cur;
    for (const char *p = list;; p++) {
      if (*p == ':' || *p == '\0') {
        if (!cur.empty()) files.push_back(cur);
        cur.clear();
        if (*p == '\0') break;
      } else {
        cur.push_back(*p);
      }
    }
    std::vector<const char *> paths;
    for (const std::string &f : files) paths.push_back(f.c_str());
    double rc = 0.0;
    check(mmd_funcfl_info(paths.data(), (int32_t)paths.size(), &n_r, &n_rho, &rc), "mmd_funcfl_info");
    n_types = (int32_t)paths.size();
    rho_r.assign((size_t)n_types * n_r, 0.0);
    emb.assign((size_t)n_types * n_rho, 0.0);
    phi_r.assign((size_t)n_types * n_types * n_r, 0.0);
    mass.assign((size_t)n_types, 0.0);
    check(mmd_funcfl_read(paths.data(), n_types, rho_r.data(), emb.data(), phi_r.data(), mass.data(), &r0, &dr, &rho0,
                          &drho),
          "mmd_funcfl_read");
  }
};

// Knot tables of MISA-MD's eam object.  Its internals belong to the core,
// which is not part of the reference tree, so the extraction is a
// specialisation point: EamTablesOf<eam>::fill(const eam &, HostTables &).
// Without one, hip_pot_init reads the setfl file named by MMD_SETFL, or the
// funcfl files (one per element type, ':'-separated) named by MMD_FUNCFL.
template <class Eam, class = void>
struct EamTablesOf {
  static constexpr bool available = false;
};

template <class Eam>
void pot_tables(const Eam *pot, HostTables &t) {
  if (const char *path = std::getenv("MMD_SETFL")) {
    t.load_setfl(path);
    return;
  }
  if (const char *list = std::getenv("MMD_FUNCFL")) {
    t.load_funcfl(list);
    return;
  }
  if constexpr (EamTablesOf<Eam>::available) {
    EamTablesOf<Eam>::fill(*pot, t);
  } else {
    (void)pot;
    std::fprintf(stderr, "mmd: hip_pot_init needs MMD_SETFL=<setfl file>, MMD_FUNCFL=<funcfl files> or an "
                         "EamTablesOf<eam> specialisation\n");
    std::abort();
  }
}

// ---------------------------------------------------------------- hooks

inline void env_init() {
  int32_t n = 0;
  HookState &s = state();
  if (s.ctx) return;
  check(mmd_device_count(&n), "hip_env_init");
  check(mmd_ctx_create(device_from_env(n), &s.ctx), "hip_env_init");
}

inline void env_clean() {
  HookState &s = state();
  if (s.ctx) check(mmd_ctx_destroy(s.ctx), "hip_env_clean");
  s.ctx = nullptr;
}

template <class BccDomain>
void domain_init(const BccDomain *p) {
  const mmd_domain d = domain_of(*p);
  check(mmd_set_domain(state().ctx, &d), "hip_domain_init");
}

template <class NeighbourIndex>
void nei_offset_init(const NeighbourIndex *n) {
  std::vector<int64_t> even, odd;
  stencil_of(*n, even, odd);
  check(mmd_set_stencil(state().ctx, even.data(), (int32_t)even.size(), odd.data(), (int32_t)odd.size()),
        "hip_nei_offset_init");
}

template <class Eam>
void pot_init(const Eam *pot) {
  HostTables t;
  pot_tables(pot, t);
  const mmd_eam_tables v = t.view();
  check(mmd_set_tables(state().ctx, &v), "hip_pot_init");
}

// The atom collection in the AoS layout: .atoms points at the ext array of
// core records (src/memory/device_atoms.h:36-43).
template <class Collection>
void rho_calc(const Collection &atoms, double cutoff) {
  check(mmd_eam_rho_calc_aos(state().ctx, atoms.atoms, &state().layout, cutoff), "hip_eam_rho_calc");
}

// df is fused into the rho pass; the reference's df hook returns at once in
// the full-stencil build (src/hip_arch_acc_imp.cpp:183-186).
template <class Collection>
void df_calc(const Collection &, double) {}

template <class Collection>
void force_calc(const Collection &atoms, double cutoff) {
  check(mmd_eam_force_calc_aos(state().ctx, atoms.atoms, &state().layout, cutoff), "hip_eam_force_calc");
}

// hip_create_atoms_mem / hip_release_atoms_mem (src/mem.cpp:14-32): pinned host memory.
inline bool create_atoms_mem(void **atoms, size_t type_size, int64_t sx, int64_t sy, int64_t sz) {
  return mmd_host_alloc(atoms, type_size * (size_t)sx * (size_t)sy * (size_t)sz) == MMD_OK && *atoms;
}

inline bool release_atoms_mem(void *atoms) { return mmd_host_free(atoms) == MMD_OK; }

}  // namespace mmd_shim

#endif
