// Instantiates the MISA-MD hook adapter with stand-in types that carry the
// member names the reference plugin reads (src/hip_arch_acc_imp.cpp:48-59,
// :108-118; src/kernel_types.h:40-48) and prints the converted values.
#include <cstdio>
#include <vector>

#include "misa_md_hooks.hpp"

struct Region {
  long x_low, y_low, z_low;
};
struct BccDomain {
  Region dbx_sub_box_lattice_region, dbx_ghost_ext_lattice_region;
  long dbx_sub_box_lattice_size[3], dbx_ghost_extended_lattice_size[3];
};
struct NeighbourIndex {
  std::vector<long> nei_even_offsets, nei_odd_offsets;
};
struct AtomElement {
  long id;
  int type;
  double x[3], v[3], f[3];
  double rho, df;
};

int main(int argc, char **argv) {
  BccDomain d{{40, 10, 20}, {34, 7, 17}, {100, 50, 60}, {112, 56, 66}};
  const mmd_domain m = mmd_shim::domain_of(d);
  std::printf("domain %d %d %d | %d %d %d | %d %d %d | %lld %lld %lld\n", m.ghost[0], m.ghost[1], m.ghost[2],
              m.box[0], m.box[1], m.box[2], m.ext[0], m.ext[1], m.ext[2], (long long)m.start[0],
              (long long)m.start[1], (long long)m.start[2]);
  NeighbourIndex n{{1, -1, 112, -5000}, {2, -2}};
  std::vector<int64_t> even, odd;
  mmd_shim::stencil_of(n, even, odd);
  std::printf("stencil %zu %zu %lld %lld\n", even.size(), odd.size(), (long long)even[3], (long long)odd[1]);
  const mmd_aos_layout l = mmd_shim::layout_of<AtomElement>();
  std::printf("layout %zu %zu %zu %zu %zu %zu %zu %d\n", l.stride, l.off_type, l.off_x, l.off_v, l.off_f, l.off_rho,
              l.off_df, l.type_bytes);
  if (argc > 1) {
    mmd_shim::HostTables t;
    t.load_setfl(argv[1]);
    const mmd_eam_tables v = t.view();
    std::printf("tables %d %d %d %.17g %.17g %.17g\n", v.n_types, v.n_r, v.n_rho, v.dr, v.drho, v.phi_r[v.n_r + 7]);
  }
  if (argc > 2) {  // ':'-separated funcfl files, as MMD_FUNCFL names them
    mmd_shim::HostTables t;
    t.load_funcfl(argv[2]);
    const mmd_eam_tables v = t.view();
    std::printf("funcfl %d %d %d %.17g %.17g %.17g\n", v.n_types, v.n_r, v.n_rho, v.dr, v.phi_r[v.n_r + 7],
                v.mass[v.n_types - 1]);
  }
  return 0;
}
