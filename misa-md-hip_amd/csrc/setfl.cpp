// Readers of EAM potential files in the DYNAMO "setfl" format (LAMMPS
// pair_style eam/alloy) and "funcfl" format (LAMMPS pair_style eam), the
// on-disk forms of the tables MISA-MD's eam object holds and hip_pot_init
// copies to the device (src/hip_arch_acc_imp.cpp:146-150).
//
// Format: three comment lines; "N el_1 .. el_N"; "Nrho drho Nr dr cutoff";
// per element a line "Z mass a lattice" followed by Nrho values of F(rho) on
// rho = k * drho and Nr values of rho(r) on r = k * dr; then, for i = 0..N-1
// and j = 0..i, Nr values of r * phi_ij(r).  Values may wrap across lines.
//
// funcfl, one file per element: a comment line; "Z mass a lattice"; "Nrho
// drho Nr dr cutoff"; Nrho values of F(rho), Nr values of the effective
// charge Z(r), Nr values of rho(r).  The pair term of elements i and j is
// r phi_ij = 27.2 * 0.529 * Z_i(r) Z_j(r) (Hartree and Bohr to eV and
// angstrom; the same product mixes two elements).  LAMMPS resamples files of
// different grids onto one; here the files must share their grid.
//
// The library's tables hold phi itself: phi_ij(r_k) = (r phi)_k / r_k for
// k >= 1 and a linear extrapolation from k = 1, 2 at r = 0 (a point no pair
// reaches).  LAMMPS instead interpolates r * phi and divides by r per pair; the
// two interpolants agree to the tables' O(dr^4) interpolation error.
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "host_helpers.h"
#include "mmd_eam.h"

namespace {

struct Reader {
  FILE *f = nullptr;
  std::string tok;
  int line = 1;
  ~Reader() {
    if (f) std::fclose(f);
  }
  // next line as text (for the comment lines and the element list)
  bool getline(std::string &out) {
    out.clear();
    int c;
    while ((c = std::fgetc(f)) != EOF && c != '\n') out.push_back((char)c);
    if (c == '\n') line++;
    return c != EOF || !out.empty();
  }
  bool next(std::string &t) {
    t.clear();
    int c;
    while ((c = std::fgetc(f)) != EOF && (c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == ','))
      if (c == '\n') line++;
    if (c == EOF) return false;
    do t.push_back((char)c);
    while ((c = std::fgetc(f)) != EOF && !(c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == ','));
    if (c == '\n') line++;
    return true;
  }
  bool number(double &v) {
    if (!next(tok)) return false;
    char *end = nullptr;
    errno = 0;
    v = std::strtod(tok.c_str(), &end);
    return end && *end == '\0' && errno == 0 && std::isfinite(v);
  }
  bool integer(long &v) {
    double d;
    if (!number(d) || d != std::floor(d)) return false;
    v = (long)d;
    return true;
  }
};

struct Header {
  int32_t n_types = 0, n_r = 0, n_rho = 0;
  double dr = 0.0, drho = 0.0, cutoff = 0.0;
};

int read_header(Reader &R, const char *path, Header &h) {
  R.f = std::fopen(path, "r");
  if (!R.f) return mmd::fail(MMD_E_INVALID, "setfl: cannot open %s", path);
  std::string s;
  for (int k = 0; k < 3; k++)
    if (!R.getline(s)) return mmd::fail(MMD_E_INVALID, "setfl %s: missing comment line %d", path, k + 1);
  long n;
  if (!R.integer(n) || n < 1 || n > 8) return mmd::fail(MMD_E_INVALID, "setfl %s: element count must be 1..8", path);
  h.n_types = (int32_t)n;
  if (!R.getline(s)) return mmd::fail(MMD_E_INVALID, "setfl %s: element names missing", path);
  long nrho, nr;
  if (!R.integer(nrho) || !R.number(h.drho) || !R.integer(nr) || !R.number(h.dr) || !R.number(h.cutoff) ||
      nrho < 5 || nr < 5 || !(h.drho > 0.0) || !(h.dr > 0.0) || !(h.cutoff > 0.0) || nrho > (1 << 24) ||
      nr > (1 << 24))
    return mmd::fail(MMD_E_INVALID, "setfl %s: bad grid line (Nrho drho Nr dr cutoff)", path);
  h.n_rho = (int32_t)nrho, h.n_r = (int32_t)nr;
  return MMD_OK;
}

// phi_ij = (r phi)_k / r_k on the r grid, extrapolated linearly to r = 0,
// into row (i, j) and its mirror (j, i)
void put_phi(double *phi_r, int nt, int nr, double dr, int i, int j, const double *z2r) {
  double *p = phi_r + ((size_t)i * nt + j) * nr;
  for (int k = 1; k < nr; k++) p[k] = z2r[k] / (k * dr);
  p[0] = 2.0 * p[1] - p[2];
  if (i != j) std::memcpy(phi_r + ((size_t)j * nt + i) * nr, p, sizeof(double) * (size_t)nr);
}

// funcfl header: comment line, element line, grid line
int read_funcfl_header(Reader &R, const char *path, Header &h, double &mass) {
  R.f = std::fopen(path, "r");
  if (!R.f) return mmd::fail(MMD_E_INVALID, "funcfl: cannot open %s", path);
  std::string s;
  if (!R.getline(s)) return mmd::fail(MMD_E_INVALID, "funcfl %s: missing comment line", path);
  long z;
  double a;
  if (!R.integer(z) || !R.number(mass) || !R.number(a) || !R.next(R.tok) || !(mass > 0.0))
    return mmd::fail(MMD_E_INVALID, "funcfl %s: bad element line (Z mass a lattice)", path);
  long nrho, nr;
  if (!R.integer(nrho) || !R.number(h.drho) || !R.integer(nr) || !R.number(h.dr) || !R.number(h.cutoff) ||
      nrho < 5 || nr < 5 || !(h.drho > 0.0) || !(h.dr > 0.0) || !(h.cutoff > 0.0) || nrho > (1 << 24) ||
      nr > (1 << 24))
    return mmd::fail(MMD_E_INVALID, "funcfl %s: bad grid line (Nrho drho Nr dr cutoff)", path);
  h.n_rho = (int32_t)nrho, h.n_r = (int32_t)nr;
  return MMD_OK;
}

// the common grid of n funcfl files
int funcfl_grid(const char *const *paths, int32_t n, Header &h) {
  if (!paths || n < 1 || n > 8) return mmd::fail(MMD_E_INVALID, "funcfl: need 1..8 files");
  for (int32_t t = 0; t < n; t++) {
    if (!paths[t]) return mmd::fail(MMD_E_INVALID, "funcfl: null path");
    Reader R;
    Header g;
    double mass;
    int rc = read_funcfl_header(R, paths[t], g, mass);
    if (rc) return rc;
    if (t == 0) {
      h = g;
    } else if (g.n_r != h.n_r || g.n_rho != h.n_rho || g.dr != h.dr || g.drho != h.drho || g.cutoff != h.cutoff) {
      return mmd::fail(MMD_E_INVALID, "funcfl %s: grid differs from %s's (the files must share one grid)", paths[t],
                       paths[0]);
    }
  }
  h.n_types = n;
  return MMD_OK;
}

}  // namespace

extern "C" int mmd_setfl_info(const char *path, int32_t *n_types, int32_t *n_r, int32_t *n_rho, double *cutoff) {
  mmd::clear_error();
  if (!path) return mmd::fail(MMD_E_INVALID, "mmd_setfl_info: null path");
  Reader R;
  Header h;
  int rc = read_header(R, path, h);
  if (rc) return rc;
  if (n_types) *n_types = h.n_types;
  if (n_r) *n_r = h.n_r;
  if (n_rho) *n_rho = h.n_rho;
  if (cutoff) *cutoff = h.cutoff;
  return MMD_OK;
}

extern "C" int mmd_setfl_read(const char *path, double *rho_r, double *emb, double *phi_r, double *mass,
                              double *r0, double *dr, double *rho0, double *drho) {
  mmd::clear_error();
  if (!path || !rho_r || !emb || !phi_r || !mass) return mmd::fail(MMD_E_INVALID, "mmd_setfl_read: null argument");
  Reader R;
  Header h;
  int rc = read_header(R, path, h);
  if (rc) return rc;
  const int nt = h.n_types, nr = h.n_r, nrho = h.n_rho;
  auto bad = [&](const char *what) {
    return mmd::fail(MMD_E_INVALID, "setfl %s: %s (near line %d)", path, what, R.line);
  };
  for (int t = 0; t < nt; t++) {
    long z;
    double a;
    if (!R.integer(z) || !R.number(mass[t]) || !R.number(a) || !R.next(R.tok) || !(mass[t] > 0.0))
      return bad("element line (Z mass a lattice)");
    for (int k = 0; k < nrho; k++)
      if (!R.number(emb[(size_t)t * nrho + k])) return bad("F(rho) value");
    for (int k = 0; k < nr; k++)
      if (!R.number(rho_r[(size_t)t * nr + k])) return bad("rho(r) value");
  }
  std::vector<double> z2r((size_t)nr);
  for (int i = 0; i < nt; i++)
    for (int j = 0; j <= i; j++) {
      for (int k = 0; k < nr; k++)
        if (!R.number(z2r[(size_t)k])) return bad("r*phi(r) value");
      put_phi(phi_r, nt, nr, h.dr, i, j, z2r.data());
    }
  *r0 = 0.0, *dr = h.dr, *rho0 = 0.0, *drho = h.drho;
  return MMD_OK;
}

extern "C" int mmd_funcfl_info(const char *const *paths, int32_t n_files, int32_t *n_r, int32_t *n_rho,
                               double *cutoff) {
  mmd::clear_error();
  Header h;
  int rc = funcfl_grid(paths, n_files, h);
  if (rc) return rc;
  if (n_r) *n_r = h.n_r;
  if (n_rho) *n_rho = h.n_rho;
  if (cutoff) *cutoff = h.cutoff;
  return MMD_OK;
}

extern "C" int mmd_funcfl_read(const char *const *paths, int32_t n_files, double *rho_r, double *emb, double *phi_r,
                               double *mass, double *r0, double *dr, double *rho0, double *drho) {
  mmd::clear_error();
  if (!rho_r || !emb || !phi_r || !mass || !r0 || !dr || !rho0 || !drho)
    return mmd::fail(MMD_E_INVALID, "mmd_funcfl_read: null argument");
  Header h;
  int rc = funcfl_grid(paths, n_files, h);
  if (rc) return rc;
  const int nt = n_files, nr = h.n_r, nrho = h.n_rho;
  std::vector<double> zr((size_t)nt * nr);  // Z_t(r) per element
  for (int t = 0; t < nt; t++) {
    Reader R;
    Header g;
    if ((rc = read_funcfl_header(R, paths[t], g, mass[t]))) return rc;
    auto bad = [&](const char *what) {
      return mmd::fail(MMD_E_INVALID, "funcfl %s: %s (near line %d)", paths[t], what, R.line);
    };
    for (int k = 0; k < nrho; k++)
      if (!R.number(emb[(size_t)t * nrho + k])) return bad("F(rho) value");
    for (int k = 0; k < nr; k++)
      if (!R.number(zr[(size_t)t * nr + k])) return bad("Z(r) value");
    for (int k = 0; k < nr; k++)
      if (!R.number(rho_r[(size_t)t * nr + k])) return bad("rho(r) value");
  }
  const double hartree_bohr = 27.2 * 0.529;
  std::vector<double> z2r((size_t)nr);
  for (int i = 0; i < nt; i++)
    for (int j = 0; j <= i; j++) {
      for (int k = 0; k < nr; k++) z2r[(size_t)k] = hartree_bohr * zr[(size_t)i * nr + k] * zr[(size_t)j * nr + k];
      put_phi(phi_r, nt, nr, h.dr, i, j, z2r.data());
    }
  *r0 = 0.0, *dr = h.dr, *rho0 = 0.0, *drho = h.drho;
  return MMD_OK;
}
