// Stand-in for MISA-MD core's arch/arch_imp.h: the types the hook signatures
// use (src/hip_arch_acc_imp.cpp:28-225, src/mem.cpp:14-32) with the member
// names the reference plugin reads (BccDomain :48-59, NeighbourIndex
// :108-118, the AoS atom collection's .atoms pointer), and the hook
// declarations.
#pragma once
#include <cstddef>
#include <cstdint>
#include <vector>

#include "args.hpp"
#include "atom/atom_element.h"

namespace comm {
struct Region {
  int64_t x_low, y_low, z_low;
};
struct BccDomain {
  Region dbx_sub_box_lattice_region, dbx_ghost_ext_lattice_region;
  int64_t dbx_sub_box_lattice_size[3], dbx_ghost_extended_lattice_size[3];
};
}  // namespace comm

typedef int64_t _type_neighbour_index_ele;
template <class T>
struct NeighbourIndex {
  std::vector<T> nei_even_offsets, nei_odd_offsets;
};

struct eam {};
typedef int64_t _type_atom_count;
struct _type_atom_list_collection {
  AtomElement *atoms;
};

void hip_env_init();
void hip_env_clean();
void hip_domain_init(const comm::BccDomain *p_domain);
void hip_nei_offset_init(const NeighbourIndex<_type_neighbour_index_ele> *nei_offset);
void hip_pot_init(eam *_pot);
void hip_eam_rho_calc(eam *pot, _type_atom_list_collection atoms, double cutoff_radius);
void hip_eam_df_calc(eam *pot, _type_atom_list_collection atoms, double cutoff_radius);
void hip_eam_force_calc(eam *pot, _type_atom_list_collection atoms, double cutoff_radius);
void hip_cli_options(args::ArgumentParser &parser);
bool hip_cli_options_parse(args::ArgumentParser &parser);
bool hip_create_atoms_mem(void **atoms, std::size_t type_size, _type_atom_count size_x, _type_atom_count size_y,
                          _type_atom_count size_z);
bool hip_release_atoms_mem(void *atoms);
