// A minimal stand-in for MISA-MD core's EAM step on one rank, driving the
// arch-acceleration hooks exactly as the core does (src/hip_arch_acc_imp.cpp
// call order): hip_env_init, hip_domain_init, hip_nei_offset_init,
// hip_pot_init, then per step hip_eam_rho_calc, hip_eam_df_calc, the core's
// own ghost df exchange (here: copies from each ghost record's periodic
// source), hip_eam_force_calc; atom records from hip_create_atoms_mem.
//
// usage: mini_misa <input dir> <output file> [steps]
//   domain.txt  ghost[3] box[3] ext[3] start[3] cutoff n_even n_odd
//   even.bin, odd.bin   int64 stencil offsets (the core's NeiOffset values)
//   type.bin            int32 per ext record (-1 vacancy)
//   x.bin               3 doubles per ext record
//   ghost_src.bin       int64 per ext record: periodic source record, -1 if local
// output: 5 doubles per ext record (rho, df, fx, fy, fz) after the last step.
// The potential comes from the setfl file named by MMD_SETFL.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "arch/arch_imp.h"

template <class T>
static std::vector<T> read_bin(const std::string &path, size_t n) {
  std::vector<T> v(n);
  FILE *f = std::fopen(path.c_str(), "rb");
  if (!f || std::fread(v.data(), sizeof(T), n, f) != n) {
    std::fprintf(stderr, "mini_misa: cannot read %zu items from %s\n", n, path.c_str());
    std::exit(2);
  }
  std::fclose(f);
  return v;
}

int main(int argc, char **argv) {
  if (argc < 3) {
    std::fprintf(stderr, "usage: mini_misa <input dir> <output file> [steps]\n");
    return 2;
  }
  const std::string in = argv[1];
  const int steps = argc > 3 ? std::atoi(argv[3]) : 1;
  long long g[3], b[3], e[3], s[3];
  double cutoff;
  long long n_even, n_odd;
  FILE *f = std::fopen((in + "/domain.txt").c_str(), "r");
  if (!f || std::fscanf(f, "%lld %lld %lld %lld %lld %lld %lld %lld %lld %lld %lld %lld %lf %lld %lld", &g[0], &g[1],
                        &g[2], &b[0], &b[1], &b[2], &e[0], &e[1], &e[2], &s[0], &s[1], &s[2], &cutoff, &n_even,
                        &n_odd) != 15) {
    std::fprintf(stderr, "mini_misa: bad domain.txt\n");
    return 2;
  }
  std::fclose(f);
  const size_t n = (size_t)(e[0] * e[1] * e[2]);
  const auto types = read_bin<int32_t>(in + "/type.bin", n);
  const auto xs = read_bin<double>(in + "/x.bin", 3 * n);
  const auto src = read_bin<int64_t>(in + "/ghost_src.bin", n);

  comm::BccDomain dom{};
  dom.dbx_sub_box_lattice_region = {s[0], s[1], s[2]};
  dom.dbx_ghost_ext_lattice_region = {s[0] - g[0], s[1] - g[1], s[2] - g[2]};
  for (int k = 0; k < 3; k++) dom.dbx_sub_box_lattice_size[k] = b[k], dom.dbx_ghost_extended_lattice_size[k] = e[k];
  NeighbourIndex<_type_neighbour_index_ele> nei;
  nei.nei_even_offsets = read_bin<int64_t>(in + "/even.bin", (size_t)n_even);
  nei.nei_odd_offsets = read_bin<int64_t>(in + "/odd.bin", (size_t)n_odd);

  args::ArgumentParser parser;
  hip_cli_options(parser);
  if (!hip_cli_options_parse(parser)) return 3;
  void *mem = nullptr;
  if (!hip_create_atoms_mem(&mem, sizeof(AtomElement), e[0], e[1], e[2])) {
    std::fprintf(stderr, "mini_misa: hip_create_atoms_mem failed\n");
    return 3;
  }
  AtomElement *atoms = static_cast<AtomElement *>(mem);
  for (size_t i = 0; i < n; i++) {
    AtomElement &a = atoms[i];
    a.id = (int64_t)i;
    a.type = types[i];
    for (int k = 0; k < 3; k++) a.x[k] = xs[3 * i + k], a.v[k] = 0.0, a.f[k] = 0.0;
    a.rho = a.df = 0.0;
  }
  eam pot;
  hip_env_init();
  hip_domain_init(&dom);
  hip_nei_offset_init(&nei);
  hip_pot_init(&pot);
  _type_atom_list_collection coll{atoms};
  for (int step = 0; step < steps; step++) {
    hip_eam_rho_calc(&pot, coll, cutoff);
    hip_eam_df_calc(&pot, coll, cutoff);
    for (size_t i = 0; i < n; i++)
      if (src[i] >= 0) atoms[i].df = atoms[src[i]].df;
    hip_eam_force_calc(&pot, coll, cutoff);
  }
  std::vector<double> out(5 * n);
  for (size_t i = 0; i < n; i++) {
    out[5 * i] = atoms[i].rho, out[5 * i + 1] = atoms[i].df;
    for (int k = 0; k < 3; k++) out[5 * i + 2 + k] = atoms[i].f[k];
  }
  FILE *o = std::fopen(argv[2], "wb");
  if (!o || std::fwrite(out.data(), sizeof(double), out.size(), o) != out.size()) return 4;
  std::fclose(o);
  hip_release_atoms_mem(mem);
  hip_env_clean();
  std::printf("mini_misa: %zu records, %d step(s) done\n", n, steps);
  return 0;
}
