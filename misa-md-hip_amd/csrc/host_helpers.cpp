// Host-side helpers of the EAM library: status/error reporting, neighbour
// stencil generation, spline coefficient construction, synthetic EAM tables
// and the deterministic initial-condition generator.  Plain C++ (compiled by
// the host compiler with FP contraction off so the spline coefficients are the
// textbook FP64 values).
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <unordered_set>
#include <vector>

#include "mmd_eam.h"
#include "host_helpers.h"

namespace mmd {

static thread_local std::string g_last_error;

int fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

void clear_error() { g_last_error.clear(); }

const char *last_error_cstr() { return g_last_error.c_str(); }

// ---------------------------------------------------------------- splines

// Cubic Hermite spline with 4th-order finite-difference slope estimates at the
// knots (the LAMMPS pair_eam interpolation MISA-MD's eam tables use).  Layout
// per knot: [c0 c1 c2 | c3 c4 c5 c6] with value ((c3 p + c4) p + c5) p + c6
// and derivative (c0 p + c1) p + c2, p the fraction of the segment.
void spline7(const double *f, int n, double dx, double *c) {
  auto C = [&](int m, int j) -> double & { return c[(size_t)m * 7 + j]; };
  for (int m = 0; m < n; m++) C(m, 6) = f[m];
  if (n < 2) {
    for (int j = 0; j < 6; j++) C(0, j) = 0.0;
    return;
  }
  C(0, 5) = C(1, 6) - C(0, 6);
  if (n >= 3) {
    C(1, 5) = 0.5 * (C(2, 6) - C(0, 6));
    C(n - 2, 5) = 0.5 * (C(n - 1, 6) - C(n - 3, 6));
  }
  C(n - 1, 5) = C(n - 1, 6) - C(n - 2, 6);
  for (int m = 2; m <= n - 3; m++)
    C(m, 5) = ((C(m - 2, 6) - C(m + 2, 6)) + 8.0 * (C(m + 1, 6) - C(m - 1, 6))) / 12.0;
  for (int m = 0; m < n - 1; m++) {
    const double d = C(m + 1, 6) - C(m, 6);
    C(m, 4) = 3.0 * d - 2.0 * C(m, 5) - C(m + 1, 5);
    C(m, 3) = C(m, 5) + C(m + 1, 5) - 2.0 * d;
  }
  C(n - 1, 4) = 0.0;
  C(n - 1, 3) = 0.0;
  for (int m = 0; m < n; m++) {
    C(m, 2) = C(m, 5) / dx;
    C(m, 1) = 2.0 * C(m, 4) / dx;
    C(m, 0) = 3.0 * C(m, 3) / dx;
  }
}

double spline7_value(const double *c, int n, double x0, double inv_dx, double x) {
  double p = (x - x0) * inv_dx;
  if (p < 0.0) p = 0.0;
  int m = (int)p;
  m = std::min(m, n - 2);
  m = std::max(m, 0);
  p -= (double)m;
  p = std::min(p, 1.0);
  const double *k = c + (size_t)m * 7;
  return ((k[3] * p + k[4]) * p + k[5]) * p + k[6];
}

double spline7_deriv(const double *c, int n, double x0, double inv_dx, double x) {
  double p = (x - x0) * inv_dx;
  if (p < 0.0) p = 0.0;
  int m = (int)p;
  m = std::min(m, n - 2);
  m = std::max(m, 0);
  p -= (double)m;
  p = std::min(p, 1.0);
  const double *k = c + (size_t)m * 7;
  return (k[0] * p + k[1]) * p + k[2];
}

// ------------------------------------------------------------- RNG (ICs)

static inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

uint64_t rng_bits(uint64_t seed, uint64_t id, uint64_t stream) {
  return mix64(mix64(seed * 0x9E3779B97F4A7C15ull + stream) ^ (id * 0xD1B54A32D192ED03ull + 0x632BE59BD9B4E019ull));
}

double rng_u01(uint64_t seed, uint64_t id, uint64_t stream) {
  return ((double)(rng_bits(seed, id, stream) >> 11) + 0.5) * 0x1.0p-53;
}

double rng_gauss(uint64_t seed, uint64_t id, uint64_t stream) {
  const double u1 = rng_u01(seed, id, stream);
  const double u2 = rng_u01(seed, id, stream + 1);
  return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586476925 * u2);
}

}  // namespace mmd

using namespace mmd;

// ---------------------------------------------------------------- stencil

extern "C" int32_t mmd_stencil_build(double a, double radius, int32_t parity, int32_t *out3, int32_t cap) {
  if (!(a > 0.0) || !(radius > 0.0) || (parity != 0 && parity != 1) || (cap > 0 && !out3)) {
    fail(MMD_E_INVALID, "mmd_stencil_build: bad argument");
    return -1;
  }
  struct E {
    int64_t d2;
    int32_t dx, dy, dz;
  };
  std::vector<E> e;
  // half-lattice units h = a/2: site (X,Y,Z) of parity s sits at (X, 2Y+s, 2Z+s) h
  const double rh = radius / (0.5 * a);
  const int32_t reach = (int32_t)std::floor(rh) + 2;
  const double r2 = radius * radius, h2 = 0.25 * a * a;
  for (int32_t dz = -reach; dz <= reach; dz++)
    for (int32_t dy = -reach; dy <= reach; dy++)
      for (int32_t dx = -2 * reach; dx <= 2 * reach; dx++) {
        if (!dx && !dy && !dz) continue;
        const int64_t s1 = dx & 1;
        const int64_t hx = dx, hy = 2 * (int64_t)dy + s1, hz = 2 * (int64_t)dz + s1;
        const int64_t d2 = hx * hx + hy * hy + hz * hz;
        if ((double)d2 * h2 < r2) e.push_back({d2, dx, dy, dz});
      }
  std::sort(e.begin(), e.end(), [](const E &p, const E &q) {
    if (p.d2 != q.d2) return p.d2 < q.d2;
    if (p.dz != q.dz) return p.dz < q.dz;
    if (p.dy != q.dy) return p.dy < q.dy;
    return p.dx < q.dx;
  });
  if ((int64_t)e.size() > cap) {
    fail(MMD_E_RANGE, "mmd_stencil_build: %zu entries exceed cap %d", e.size(), cap);
    return -1;
  }
  const int32_t sg = parity ? -1 : 1;
  for (size_t k = 0; k < e.size(); k++) {
    out3[3 * k + 0] = sg * e[k].dx;
    out3[3 * k + 1] = sg * e[k].dy;
    out3[3 * k + 2] = sg * e[k].dz;
  }
  return (int32_t)e.size();
}

extern "C" int mmd_stencil_linear(const mmd_domain *d, const int32_t *in3, int32_t n, int64_t *out) {
  if (!d || (n > 0 && (!in3 || !out)) || n < 0) return fail(MMD_E_INVALID, "mmd_stencil_linear: bad argument");
  for (int32_t k = 0; k < n; k++) {
    const int64_t dx = in3[3 * k], dy = in3[3 * k + 1], dz = in3[3 * k + 2];
    if (std::llabs(dx) > d->ghost[0] || std::llabs(dy) > d->ghost[1] || std::llabs(dz) > d->ghost[2])
      return fail(MMD_E_RANGE, "mmd_stencil_linear: entry %d (%lld,%lld,%lld) reaches beyond the ghost shell", k,
                  (long long)dx, (long long)dy, (long long)dz);
    out[k] = (dz * d->ext[1] + dy) * d->ext[0] + dx;
  }
  return MMD_OK;
}

// --------------------------------------------------------- synthetic tables

extern "C" int mmd_synth_tables(const mmd_synth_params *p, double *rho_r, double *emb, double *phi_r, double *r0,
                                double *dr, double *rho0, double *drho) {
  if (!p || p->n_types < 1 || p->n_types > 8 || p->n_r < 5 || p->n_rho < 5 || !(p->rc > p->r0) ||
      !(p->r0 > 0.0) || !(p->rho_max > 0.0) || !rho_r || !emb || !phi_r)
    return fail(MMD_E_INVALID, "mmd_synth_tables: bad argument");
  const int nt = p->n_types, nr = p->n_r, nrho = p->n_rho;
  const double delta_r = (p->rc - p->r0) / (double)(nr - 1);
  const double delta_rho = p->rho_max / (double)(nrho - 1);
  auto jit = [&](uint64_t id, uint64_t stream) { return 1.0 + p->jitter * (2.0 * rng_u01(p->seed, id, stream) - 1.0); };
  for (int t = 0; t < nt; t++) {
    for (int i = 0; i < nr; i++) {
      const double r = p->r0 + (double)i * delta_r;
      const double q = p->rc - r;
      rho_r[(size_t)t * nr + i] = q * q * std::exp(-r);
    }
    const double Ct = p->C * jit((uint64_t)t, 102);
    for (int i = 0; i < nrho; i++) {
      const double rho = (double)i * delta_rho;
      emb[(size_t)t * nrho + i] = -std::sqrt(rho) + Ct * rho * rho;
    }
  }
  for (int a = 0; a < nt; a++)
    for (int b = a; b < nt; b++) {
      const uint64_t pid = (uint64_t)(a * nt + b);
      const double A = p->A * jit(pid, 100), B = p->B * jit(pid, 101);
      for (int i = 0; i < nr; i++) {
        const double r = p->r0 + (double)i * delta_r;
        const double q = p->rc - r;
        const double v = A * q * q * (1.0 + B / r);
        phi_r[((size_t)a * nt + b) * nr + i] = v;
        phi_r[((size_t)b * nt + a) * nr + i] = v;
      }
    }
  if (r0) *r0 = p->r0;
  if (dr) *dr = delta_r;
  if (rho0) *rho0 = 0.0;
  if (drho) *drho = delta_rho;
  return MMD_OK;
}

// Finnis-Sinclair EAM for BCC Fe (M. W. Finnis and J. E. Sinclair, Phil. Mag.
// A 50 (1984) 45, Table 1): pair V(r) = (r-c)^2 (c0 + c1 r + c2 r^2) for
// r < c, density phi(r) = (r-d)^2 + beta (r-d)^3 / d for r < d, embedding
// F(rho) = -A sqrt(rho); equilibrium a = 2.8665 A, cohesive energy 4.28 eV.
extern "C" int mmd_fs_tables(const mmd_fs_params *p, double *rho_r, double *emb, double *phi_r, double *r0,
                             double *dr, double *rho0, double *drho) {
  constexpr double d = 3.569745, A = 1.828905, beta = 1.8, c = 3.40;
  constexpr double c0 = 1.2371147, c1 = -0.3592185, c2 = -0.0385607;
  if (!p || p->n_types < 1 || p->n_types > 8 || p->n_r < 5 || p->n_rho < 5 || !(p->rc >= d) || !(p->r0 > 0.0) ||
      !(p->rc > p->r0) || !(p->rho_max > 0.0) || !(p->jitter >= 0.0 && p->jitter < 0.5) || !rho_r || !emb || !phi_r)
    return fail(MMD_E_INVALID, "mmd_fs_tables: bad argument (the table range must reach d = %g A)", d);
  const int nt = p->n_types, nr = p->n_r, nrho = p->n_rho;
  const double delta_r = (p->rc - p->r0) / (double)(nr - 1);
  const double delta_rho = p->rho_max / (double)(nrho - 1);
  auto jit = [&](uint64_t id, uint64_t stream) { return 1.0 + p->jitter * (2.0 * rng_u01(p->seed, id, stream) - 1.0); };
  for (int t = 0; t < nt; t++) {
    for (int i = 0; i < nr; i++) {
      const double q = p->r0 + (double)i * delta_r - d;
      rho_r[(size_t)t * nr + i] = q < 0.0 ? q * q * (1.0 + beta * q / d) : 0.0;
    }
    const double At = A * jit((uint64_t)t, 110);
    for (int i = 0; i < nrho; i++) emb[(size_t)t * nrho + i] = -At * std::sqrt((double)i * delta_rho);
  }
  for (int a = 0; a < nt; a++)
    for (int b = a; b < nt; b++) {
      const double s = jit((uint64_t)(a * nt + b), 111);
      for (int i = 0; i < nr; i++) {
        const double r = p->r0 + (double)i * delta_r;
        const double q = r - c;
        const double v = q < 0.0 ? s * q * q * (c0 + c1 * r + c2 * r * r) : 0.0;
        phi_r[((size_t)a * nt + b) * nr + i] = v;
        phi_r[((size_t)b * nt + a) * nr + i] = v;
      }
    }
  if (r0) *r0 = p->r0;
  if (dr) *dr = delta_r;
  if (rho0) *rho0 = 0.0;
  if (drho) *drho = delta_rho;
  return MMD_OK;
}

// ------------------------------------------------------ initial conditions

static constexpr double kBoltzmann = 8.617333262e-5;  // eV/K
static constexpr double kFtm2v = 9.64853321e-3;       // (eV/A/amu) -> A/fs^2

extern "C" int mmd_lattice_domain(const mmd_lattice_spec *s, mmd_domain *d, double box_len[3]) {
  if (!s || !d) return fail(MMD_E_INVALID, "mmd_lattice_domain: null argument");
  if (s->cells[0] < 1 || s->cells[1] < 1 || s->cells[2] < 1 || s->ghost_cells < 1 || !(s->lattice_const > 0.0))
    return fail(MMD_E_INVALID, "mmd_lattice_domain: bad sizes");
  if (s->ghost_cells > s->cells[0] || s->ghost_cells > s->cells[1] || s->ghost_cells > s->cells[2])
    return fail(MMD_E_INVALID, "mmd_lattice_domain: ghost shell wider than the box");
  const int g = s->ghost_cells;
  d->box[0] = 2 * s->cells[0];
  d->box[1] = s->cells[1];
  d->box[2] = s->cells[2];
  d->ghost[0] = 2 * g;
  d->ghost[1] = g;
  d->ghost[2] = g;
  for (int k = 0; k < 3; k++) {
    d->ext[k] = d->box[k] + 2 * d->ghost[k];
    d->start[k] = 0;
  }
  if (box_len)
    for (int k = 0; k < 3; k++) box_len[k] = s->cells[k] * s->lattice_const;
  return MMD_OK;
}

namespace {
struct Geo {
  mmd_domain d;
  int64_t lin(int64_t xe, int64_t ye, int64_t ze) const { return (ze * d.ext[1] + ye) * d.ext[0] + xe; }
};

inline int64_t wrapi(int64_t v, int64_t n) {
  int64_t r = v % n;
  return r < 0 ? r + n : r;
}

// tetrahedral site t (0..11) of a unit cell, fractional coordinates
void tet_site(int t, double f[3]) {
  const int zero_axis = t / 4;          // which coordinate is 0
  const int half_pick = (t / 2) & 1;    // which of the remaining two is 1/2
  const double quarter = (t & 1) ? 0.75 : 0.25;
  const int o1 = (zero_axis + 1) % 3, o2 = (zero_axis + 2) % 3;
  f[zero_axis] = 0.0;
  f[half_pick ? o2 : o1] = 0.5;
  f[half_pick ? o1 : o2] = quarter;
}
}  // namespace

extern "C" int mmd_lattice_init(const mmd_lattice_spec *s, const mmd_host_soa *h, int32_t itl_cap, int32_t *n_itl,
                                double *ix, double *iy, double *iz, double *ivx, double *ivy, double *ivz,
                                int8_t *itype) {
  clear_error();
  Geo G;
  double L[3];
  int rc = mmd_lattice_domain(s, &G.d, L);
  if (rc) return rc;
  if (!h || !h->x || !h->y || !h->z || !h->vx || !h->vy || !h->vz || !h->type)
    return fail(MMD_E_INVALID, "mmd_lattice_init: host arrays missing");
  if (s->n_types < 1 || s->n_types > 8) return fail(MMD_E_INVALID, "mmd_lattice_init: n_types out of range");
  if (s->vacancy_frac < 0 || s->vacancy_frac > 0.5 || s->interstitial_frac < 0 || s->interstitial_frac > 0.05)
    return fail(MMD_E_INVALID, "mmd_lattice_init: defect fraction out of range");
  const double a = s->lattice_const;
  const int64_t NX = G.d.box[0], NY = G.d.box[1], NZ = G.d.box[2];
  const int64_t nsite = NX * NY * NZ;
  const int64_t next = (int64_t)G.d.ext[0] * G.d.ext[1] * G.d.ext[2];
  const uint64_t seed = s->seed;

  for (int64_t i = 0; i < next; i++) {
    h->type[i] = -1;
    h->x[i] = h->y[i] = h->z[i] = 0.0;
    h->vx[i] = h->vy[i] = h->vz[i] = 0.0;
  }
  auto draw_type = [&](uint64_t id, uint64_t stream) -> int8_t {
    const double u = rng_u01(seed, id, stream);
    for (int t = 0; t < s->n_types - 1; t++)
      if (u < s->type_cum[t]) return (int8_t)t;
    return (int8_t)(s->n_types - 1);
  };
  auto vel_scale = [&](int t) { return std::sqrt(kBoltzmann * s->temperature / s->mass[t] * kFtm2v); };

  // lattice sites: type, displacement, velocity
  for (int64_t Z = 0; Z < NZ; Z++)
    for (int64_t Y = 0; Y < NY; Y++)
      for (int64_t X = 0; X < NX; X++) {
        const uint64_t id = (uint64_t)((Z * NY + Y) * NX + X);
        const int64_t e = G.lin(X + G.d.ghost[0], Y + G.d.ghost[1], Z + G.d.ghost[2]);
        const int8_t t = draw_type(id, 1);
        const double par = (double)(X & 1);
        h->type[e] = t;
        h->x[e] = 0.5 * a * (double)X + s->sigma * rng_gauss(seed, id, 3);
        h->y[e] = a * ((double)Y + 0.5 * par) + s->sigma * rng_gauss(seed, id, 5);
        h->z[e] = a * ((double)Z + 0.5 * par) + s->sigma * rng_gauss(seed, id, 7);
        const double vs = vel_scale(t);
        h->vx[e] = vs * rng_gauss(seed, id, 9);
        h->vy[e] = vs * rng_gauss(seed, id, 11);
        h->vz[e] = vs * rng_gauss(seed, id, 13);
      }

  // blob centres (fractional box coordinates)
  std::vector<double> blob(3 * std::max(s->n_blobs, 0));
  for (int b = 0; b < s->n_blobs; b++)
    for (int k = 0; k < 3; k++) blob[3 * b + k] = rng_u01(seed, (uint64_t)b, 20 + k) * L[k];

  // vacancies
  if (s->vacancy_frac > 0.0) {
    if (s->n_blobs <= 0) {
      for (int64_t Z = 0; Z < NZ; Z++)
        for (int64_t Y = 0; Y < NY; Y++)
          for (int64_t X = 0; X < NX; X++) {
            const uint64_t id = (uint64_t)((Z * NY + Y) * NX + X);
            if (rng_u01(seed, id, 2) < s->vacancy_frac)
              h->type[G.lin(X + G.d.ghost[0], Y + G.d.ghost[1], Z + G.d.ghost[2])] = -1;
          }
    } else {
      // a blob holds at most its lattice sites: every rejected draw widens that
      // blob by 1%, so dense blobs grow into larger clusters instead of stalling
      const int64_t nvac = (int64_t)std::llround(s->vacancy_frac * (double)nsite);
      std::vector<double> bsig(s->n_blobs, s->blob_sigma);
      uint64_t ctr = 0;
      for (int64_t v = 0; v < nvac; ctr++) {
        if (ctr > (uint64_t)nvac * 64 + 1024) return fail(MMD_E_INVALID, "mmd_lattice_init: vacancy placement stalled");
        const int b = (int)(rng_u01(seed, ctr, 30) * s->n_blobs);
        double p[3];
        for (int k = 0; k < 3; k++) p[k] = blob[3 * b + k] + bsig[b] * rng_gauss(seed, ctr, 31 + 2 * k);
        // nearest site (pick sub-lattice by distance)
        const double ux = p[0] / a, uy = p[1] / a, uz = p[2] / a;
        const double cx = std::floor(ux + 0.5), cy = std::floor(uy + 0.5), cz = std::floor(uz + 0.5);
        const double bx = std::floor(ux) + 0.5, by = std::floor(uy) + 0.5, bz = std::floor(uz) + 0.5;
        const double dc = (ux - cx) * (ux - cx) + (uy - cy) * (uy - cy) + (uz - cz) * (uz - cz);
        const double db = (ux - bx) * (ux - bx) + (uy - by) * (uy - by) + (uz - bz) * (uz - bz);
        int64_t X, Y, Z;
        if (dc <= db) {
          X = 2 * (int64_t)cx, Y = (int64_t)cy, Z = (int64_t)cz;
        } else {
          X = 2 * (int64_t)std::floor(ux) + 1, Y = (int64_t)std::floor(uy), Z = (int64_t)std::floor(uz);
        }
        X = wrapi(X, NX), Y = wrapi(Y, NY), Z = wrapi(Z, NZ);
        const int64_t e = G.lin(X + G.d.ghost[0], Y + G.d.ghost[1], Z + G.d.ghost[2]);
        if (h->type[e] < 0) {
          bsig[b] *= 1.01;
          continue;
        }
        h->type[e] = -1;
        v++;
      }
    }
  }

  // interstitials at tetrahedral sites, at most one per 3x3x3 block of cells
  int32_t nitl = 0;
  if (s->interstitial_frac > 0.0) {
    const int64_t want = (int64_t)std::llround(s->interstitial_frac * (double)nsite);
    if (want > itl_cap || !ix || !iy || !iz || !ivx || !ivy || !ivz || !itype)
      return fail(MMD_E_RANGE, "mmd_lattice_init: %lld interstitials exceed capacity %d", (long long)want, itl_cap);
    const int64_t cx = s->cells[0], cy = s->cells[1], cz = s->cells[2];
    std::unordered_set<int64_t> taken;
    std::vector<double> bsig(std::max(s->n_blobs, 0), s->blob_sigma);
    uint64_t ctr = 0;
    int blob_of_draw = -1;
    while (nitl < want) {
      if (ctr > (uint64_t)want * 256 + 4096)
        return fail(MMD_E_INVALID, "mmd_lattice_init: interstitial placement stalled (density too high)");
      int64_t c[3];
      if (s->n_blobs > 0) {
        const int b = (int)(rng_u01(seed, ctr, 40) * s->n_blobs);
        blob_of_draw = b;
        for (int k = 0; k < 3; k++) {
          const double p = blob[3 * b + k] + bsig[b] * rng_gauss(seed, ctr, 41 + 2 * k);
          c[k] = (int64_t)std::floor(p / a);
        }
      } else {
        c[0] = (int64_t)(rng_u01(seed, ctr, 50) * cx);
        c[1] = (int64_t)(rng_u01(seed, ctr, 51) * cy);
        c[2] = (int64_t)(rng_u01(seed, ctr, 52) * cz);
      }
      c[0] = wrapi(c[0], cx), c[1] = wrapi(c[1], cy), c[2] = wrapi(c[2], cz);
      const int tsite = (int)(rng_u01(seed, ctr, 53) * 12.0) % 12;
      const uint64_t id = ctr;
      ctr++;
      bool clash = false;
      for (int dz = -1; dz <= 1 && !clash; dz++)
        for (int dy = -1; dy <= 1 && !clash; dy++)
          for (int dx = -1; dx <= 1 && !clash; dx++) {
            const int64_t key = (wrapi(c[2] + dz, cz) * cy + wrapi(c[1] + dy, cy)) * cx + wrapi(c[0] + dx, cx);
            if (taken.count(key)) clash = true;
          }
      if (clash) {
        if (blob_of_draw >= 0) bsig[blob_of_draw] *= 1.01;
        continue;
      }
      taken.insert((c[2] * cy + c[1]) * cx + c[0]);
      double f[3];
      tet_site(tsite, f);
      const int8_t t = draw_type(id, 60);
      ix[nitl] = a * ((double)c[0] + f[0]) + s->sigma * rng_gauss(seed, id, 61);
      iy[nitl] = a * ((double)c[1] + f[1]) + s->sigma * rng_gauss(seed, id, 63);
      iz[nitl] = a * ((double)c[2] + f[2]) + s->sigma * rng_gauss(seed, id, 65);
      const double vs = vel_scale(t);
      ivx[nitl] = vs * rng_gauss(seed, id, 67);
      ivy[nitl] = vs * rng_gauss(seed, id, 69);
      ivz[nitl] = vs * rng_gauss(seed, id, 71);
      itype[nitl] = t;
      nitl++;
    }
  }
  if (n_itl) *n_itl = nitl;

  // remove the centre-of-mass velocity (lattice atoms + interstitials)
  double P[3] = {0, 0, 0}, M = 0.0;
  for (int64_t Z = 0; Z < NZ; Z++)
    for (int64_t Y = 0; Y < NY; Y++)
      for (int64_t X = 0; X < NX; X++) {
        const int64_t e = G.lin(X + G.d.ghost[0], Y + G.d.ghost[1], Z + G.d.ghost[2]);
        const int t = h->type[e];
        if (t < 0) continue;
        const double m = s->mass[t];
        P[0] += m * h->vx[e], P[1] += m * h->vy[e], P[2] += m * h->vz[e];
        M += m;
      }
  for (int32_t k = 0; k < nitl; k++) {
    const double m = s->mass[itype[k]];
    P[0] += m * ivx[k], P[1] += m * ivy[k], P[2] += m * ivz[k];
    M += m;
  }
  if (M > 0.0) {
    const double u[3] = {P[0] / M, P[1] / M, P[2] / M};
    for (int64_t Z = 0; Z < NZ; Z++)
      for (int64_t Y = 0; Y < NY; Y++)
        for (int64_t X = 0; X < NX; X++) {
          const int64_t e = G.lin(X + G.d.ghost[0], Y + G.d.ghost[1], Z + G.d.ghost[2]);
          if (h->type[e] < 0) {
            h->vx[e] = h->vy[e] = h->vz[e] = 0.0;
            continue;
          }
          h->vx[e] -= u[0], h->vy[e] -= u[1], h->vz[e] -= u[2];
        }
    for (int32_t k = 0; k < nitl; k++) ivx[k] -= u[0], ivy[k] -= u[1], ivz[k] -= u[2];
  }

  // periodic ghost shell (types, shifted positions, velocities)
  for (int64_t ze = 0; ze < G.d.ext[2]; ze++)
    for (int64_t ye = 0; ye < G.d.ext[1]; ye++)
      for (int64_t xe = 0; xe < G.d.ext[0]; xe++) {
        const int64_t xl = xe - G.d.ghost[0], yl = ye - G.d.ghost[1], zl = ze - G.d.ghost[2];
        if (xl >= 0 && xl < NX && yl >= 0 && yl < NY && zl >= 0 && zl < NZ) continue;
        const int64_t sx = wrapi(xl, NX), sy = wrapi(yl, NY), sz = wrapi(zl, NZ);
        const int64_t g = G.lin(xe, ye, ze), src = G.lin(sx + G.d.ghost[0], sy + G.d.ghost[1], sz + G.d.ghost[2]);
        h->type[g] = h->type[src];
        h->x[g] = h->x[src] + (double)((xl - sx) / NX) * L[0];
        h->y[g] = h->y[src] + (double)((yl - sy) / NY) * L[1];
        h->z[g] = h->z[src] + (double)((zl - sz) / NZ) * L[2];
        h->vx[g] = h->vx[src], h->vy[g] = h->vy[src], h->vz[g] = h->vz[src];
      }
  return MMD_OK;
}

extern "C" const char *mmd_last_error(void) { return mmd::last_error_cstr(); }
