// The arch-acceleration hooks MISA-MD core calls, with the exact signatures of
// the reference plugin (src/hip_arch_acc_imp.cpp:28-225, src/cli.cpp:16-29,
// src/mem.cpp:14-32), forwarding to libmmd_eam.so.  Compiled inside the MISA-MD
// tree (its headers: arch/arch_imp.h, atom/atom_element.h, args.hpp) in place
// of the reference's libmd_arch_hip.a, with the core built for the AoS atom
// layout (the reference's SoA kernels are unimplemented:
// src/double-buffer/rho_double_buffer_imp.cpp:60-66,
// force_double_buffer_imp.cpp:61-66).
#include <args.hpp>

#include "arch/arch_imp.h"
#include "atom/atom_element.h"
#include "misa_md_hooks.hpp"

#ifndef MD_ATOM_HASH_ARRAY_MEMORY_LAYOUT_AOS
#error "build MISA-MD with the AoS atom layout: the reference plugin computes nothing in the SoA layout"
#endif

void hip_env_init() {
  mmd_shim::env_init();
  mmd_shim::state().layout = mmd_shim::layout_of<AtomElement>();
}

void hip_env_clean() { mmd_shim::env_clean(); }

void hip_domain_init(const comm::BccDomain *p_domain) { mmd_shim::domain_init(p_domain); }

void hip_nei_offset_init(const NeighbourIndex<_type_neighbour_index_ele> *nei_offset) {
  mmd_shim::nei_offset_init(nei_offset);
}

void hip_pot_init(eam *_pot) { mmd_shim::pot_init(_pot); }

void hip_eam_rho_calc(eam *, _type_atom_list_collection atoms, double cutoff_radius) {
  mmd_shim::rho_calc(atoms, cutoff_radius);
}

void hip_eam_df_calc(eam *, _type_atom_list_collection atoms, double cutoff_radius) {
  mmd_shim::df_calc(atoms, cutoff_radius);
}

void hip_eam_force_calc(eam *, _type_atom_list_collection atoms, double cutoff_radius) {
  mmd_shim::force_calc(atoms, cutoff_radius);
}

// The reference's -b (batches) flag sized its z-slab streaming; the lattice is
// device resident here, so the flag is accepted and ignored.
static args::ValueFlag<unsigned int> *flag_batches = nullptr;

void hip_cli_options(args::ArgumentParser &parser) {
  flag_batches = new args::ValueFlag<unsigned int>(parser, "batches", "batches (ignored)", {'b', "Batches number"});
}

bool hip_cli_options_parse(args::ArgumentParser &) { return true; }

bool hip_create_atoms_mem(void **atoms, std::size_t type_size, _type_atom_count size_x, _type_atom_count size_y,
                          _type_atom_count size_z) {
  return mmd_shim::create_atoms_mem(atoms, type_size, size_x, size_y, size_z);
}

bool hip_release_atoms_mem(void *atoms) { return mmd_shim::release_atoms_mem(atoms); }
