/*
 * eam_oracle.c -- CPU FP64 restatement of MISA-MD-hip's EAM hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for tests/, __graft_entry__.smoke()
 * and the CPU baseline leg of bench.py.  The product library never links it.
 *
 * Reference behaviour restated here (file:line in misa-md-hip):
 *   - site loop, ghost-extended linear index, skip type < 0:
 *       src/kernels/kernel_itl.hpp:21-26, :32-46, :69-80
 *   - parity-selected stencil of linear offsets: kernel_itl.hpp:82-86
 *   - neighbour skip type < 0, del = x_i - x_j, dist2: kernel_itl.hpp:48-57
 *   - rho by neighbour type, strict dist2 < rc^2: kernel_pairs.hpp:21-32
 *   - fused df = F'_{t_i}(rho_i): kernel_itl.hpp:99-103; standalone calDf:
 *       src/kernels/hip_kernels.hip.cpp:21-40
 *   - force f_i = sum del * fpair, overwrite: kernel_pairs.hpp:39-60,
 *       kernel_itl.hpp:93-97
 *   - hipToForce(t_i, t_j, dist2, df_i, df_j) (hip-potential, absent):
 *       fpair = -(df_i*rho'_{t_j}(r) + df_j*rho'_{t_i}(r) + phi'_{t_i t_j}(r)) / r
 *     which is the EAM force of E = sum_i F(rho_i) + 1/2 sum_ij phi(r_ij).
 * Ghost refresh, velocity Verlet, energies and interstitials live in the
 * absent MISA-MD core; they are restated from the physics.
 */
#include "eam_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* 1 eV/(A*amu) in A/fs^2, and its inverse (amu*A^2/fs^2 -> eV). */
#define ORC_FTM2V 9.64853321e-3
#define ORC_MVV2E (1.0 / ORC_FTM2V)

/* ------------------------------------------------------------------ splines */

void orc_spline_build(const double *f, int32_t n, double dx, double *c) {
  /* c[m*7 + j]; value = ((c3 p + c4) p + c5) p + c6, derivative =
   * (c0 p + c1) p + c2 with p the fraction inside segment m. */
  for (int32_t m = 0; m < n; m++) c[m * 7 + 6] = f[m];
  if (n == 1) {
    for (int j = 0; j < 6; j++) c[j] = 0.0;
    return;
  }
  /* slope estimates per knot (in units of one segment) */
  c[0 * 7 + 5] = c[1 * 7 + 6] - c[0 * 7 + 6];
  if (n >= 3) {
    c[1 * 7 + 5] = 0.5 * (c[2 * 7 + 6] - c[0 * 7 + 6]);
    c[(n - 2) * 7 + 5] = 0.5 * (c[(n - 1) * 7 + 6] - c[(n - 3) * 7 + 6]);
  }
  c[(n - 1) * 7 + 5] = c[(n - 1) * 7 + 6] - c[(n - 2) * 7 + 6];
  for (int32_t m = 2; m <= n - 3; m++) {
    c[m * 7 + 5] = ((c[(m - 2) * 7 + 6] - c[(m + 2) * 7 + 6]) + 8.0 * (c[(m + 1) * 7 + 6] - c[(m - 1) * 7 + 6])) / 12.0;
  }
  for (int32_t m = 0; m < n - 1; m++) {
    const double d = c[(m + 1) * 7 + 6] - c[m * 7 + 6];
    c[m * 7 + 4] = 3.0 * d - 2.0 * c[m * 7 + 5] - c[(m + 1) * 7 + 5];
    c[m * 7 + 3] = c[m * 7 + 5] + c[(m + 1) * 7 + 5] - 2.0 * d;
  }
  c[(n - 1) * 7 + 4] = 0.0;
  c[(n - 1) * 7 + 3] = 0.0;
  for (int32_t m = 0; m < n; m++) {
    c[m * 7 + 2] = c[m * 7 + 5] / dx;
    c[m * 7 + 1] = 2.0 * c[m * 7 + 4] / dx;
    c[m * 7 + 0] = 3.0 * c[m * 7 + 3] / dx;
  }
}

static inline const double *orc_seg(const orc_spline *s, double x, double *frac) {
  double p = (x - s->x0) * s->inv_dx;
  if (p < 0.0) p = 0.0;
  int32_t m = (int32_t)p;
  if (m > s->n - 2) m = s->n - 2;
  if (m < 0) m = 0;
  p -= (double)m;
  if (p > 1.0) p = 1.0;
  *frac = p;
  return s->c + (size_t)m * 7;
}

double orc_spline_value(const orc_spline *s, double x) {
  double p;
  const double *c = orc_seg(s, x, &p);
  return ((c[3] * p + c[4]) * p + c[5]) * p + c[6];
}

double orc_spline_deriv(const orc_spline *s, double x) {
  double p;
  const double *c = orc_seg(s, x, &p);
  return (c[0] * p + c[1]) * p + c[2];
}

/* ------------------------------------------------------------------ stencil */

typedef struct {
  int64_t d2q; /* squared ideal distance in units of (a/2)^2 */
  int32_t dx, dy, dz;
} orc_st_ent;

static int orc_st_cmp(const void *pa, const void *pb) {
  const orc_st_ent *a = (const orc_st_ent *)pa, *b = (const orc_st_ent *)pb;
  if (a->d2q != b->d2q) return a->d2q < b->d2q ? -1 : 1;
  if (a->dz != b->dz) return a->dz < b->dz ? -1 : 1;
  if (a->dy != b->dy) return a->dy < b->dy ? -1 : 1;
  if (a->dx != b->dx) return a->dx < b->dx ? -1 : 1;
  return 0;
}

int32_t orc_stencil3(double a, double radius, int32_t parity, int32_t *out, int32_t cap) {
  /* Work in half-lattice units h = a/2: site (X,Y,Z) of parity s sits at
   * (X, 2Y + s, 2Z + s) * h.  Ideal distances squared are integers in h^2. */
  const double rh = radius / (0.5 * a);
  const int32_t reach = (int32_t)floor(rh) + 2;
  const size_t maxn = (size_t)(2 * reach + 1) * (2 * reach + 1) * (2 * reach + 1);
  orc_st_ent *e = (orc_st_ent *)malloc(maxn * sizeof(orc_st_ent));
  int32_t n = 0;
  for (int32_t dz = -reach; dz <= reach; dz++)
    for (int32_t dy = -reach; dy <= reach; dy++)
      for (int32_t dx = -2 * reach; dx <= 2 * reach; dx++) {
        if (dx == 0 && dy == 0 && dz == 0) continue;
        const int32_t s1 = (0 + dx) & 1; /* neighbour parity relative to a parity-0 site */
        const int64_t hx = dx, hy = 2 * dy + s1, hz = 2 * dz + s1;
        const int64_t d2q = hx * hx + hy * hy + hz * hz;
        if ((double)d2q * (0.25 * a * a) < radius * radius) {
          if ((size_t)n >= maxn) continue;
          e[n].d2q = d2q;
          e[n].dx = dx;
          e[n].dy = dy;
          e[n].dz = dz;
          n++;
        }
      }
  qsort(e, (size_t)n, sizeof(orc_st_ent), orc_st_cmp);
  if (n > cap) {
    free(e);
    return -1;
  }
  const int32_t sign = parity ? -1 : 1;
  for (int32_t k = 0; k < n; k++) {
    out[3 * k + 0] = sign * e[k].dx;
    out[3 * k + 1] = sign * e[k].dy;
    out[3 * k + 2] = sign * e[k].dz;
  }
  free(e);
  return n;
}

/* ------------------------------------------------------------------ passes */

static inline int64_t orc_lin(const orc_domain *d, int64_t xe, int64_t ye, int64_t ze) {
  return (ze * d->ext[1] + ye) * d->ext[0] + xe;
}

static inline int32_t orc_parity(const orc_domain *d, int64_t xl) {
  return (int32_t)(((xl + d->start[0]) % 2 + 2) % 2);
}

static void orc_set_threads(int32_t nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
}

static void orc_rho_lattice(const orc_domain *d, const int64_t *off_even, int32_t n_even, const int64_t *off_odd,
                            int32_t n_odd, const orc_tables *t, double rc, const double *x, const double *y,
                            const double *z, const int8_t *type, double *rho) {
  const double rc2 = rc * rc;
  const int64_t bz = d->box[2], by = d->box[1], bx = d->box[0];
#pragma omp parallel for collapse(2) schedule(static)
  for (int64_t zl = 0; zl < bz; zl++)
    for (int64_t yl = 0; yl < by; yl++) {
      for (int64_t xl = 0; xl < bx; xl++) {
        const int64_t i = orc_lin(d, xl + d->ghost[0], yl + d->ghost[1], zl + d->ghost[2]);
        const int t0 = type[i];
        if (t0 < 0) continue;
        const int32_t par = orc_parity(d, xl);
        const int64_t *off = par == 0 ? off_even : off_odd;
        const int32_t nk = par == 0 ? n_even : n_odd;
        double acc = 0.0;
        for (int32_t k = 0; k < nk; k++) {
          const int64_t j = i + off[k];
          const int tj = type[j];
          if (tj < 0) continue;
          const double dx = x[i] - x[j], dy = y[i] - y[j], dz = z[i] - z[j];
          const double d2 = dx * dx + dy * dy + dz * dz;
          if (d2 >= rc2) continue;
          acc += orc_spline_value(&t->rho[tj], sqrt(d2));
        }
        rho[i] = acc;
      }
    }
}

static double orc_df_and_embed(const orc_domain *d, const orc_tables *t, const int8_t *type, const double *rho,
                               double *df, int want_e) {
  const int64_t bz = d->box[2], by = d->box[1], bx = d->box[0];
  double e = 0.0;
  for (int64_t zl = 0; zl < bz; zl++)
    for (int64_t yl = 0; yl < by; yl++)
      for (int64_t xl = 0; xl < bx; xl++) {
        const int64_t i = orc_lin(d, xl + d->ghost[0], yl + d->ghost[1], zl + d->ghost[2]);
        const int t0 = type[i];
        if (t0 < 0) continue;
        df[i] = orc_spline_deriv(&t->emb[t0], rho[i]);
        if (want_e) e += orc_spline_value(&t->emb[t0], rho[i]);
      }
  return e;
}

void orc_rho_df(const orc_domain *d, const int64_t *off_even, int32_t n_even, const int64_t *off_odd,
                int32_t n_odd, const orc_tables *t, double rc, const double *x, const double *y,
                const double *z, const int8_t *type, double *rho, double *df, double *e_embed,
                int32_t nthreads) {
  orc_set_threads(nthreads);
  orc_rho_lattice(d, off_even, n_even, off_odd, n_odd, t, rc, x, y, z, type, rho);
  const double e = orc_df_and_embed(d, t, type, rho, df, e_embed != NULL);
  if (e_embed) *e_embed = e;
}

void orc_df(const orc_domain *d, const orc_tables *t, const int8_t *type, const double *rho, double *df) {
  (void)orc_df_and_embed(d, t, type, rho, df, 0);
}

/* fpair and (optionally) phi for one pair, types ti (centre) and tj. */
static inline double orc_fpair(const orc_tables *t, int ti, int tj, double r, double dfi, double dfj,
                               double *phi_out) {
  const orc_spline *ph = &t->phi[ti * t->n_types + tj];
  const double rhojp = orc_spline_deriv(&t->rho[tj], r);
  const double rhoip = orc_spline_deriv(&t->rho[ti], r);
  const double phip = orc_spline_deriv(ph, r);
  if (phi_out) *phi_out = orc_spline_value(ph, r);
  const double psip = dfi * rhojp + dfj * rhoip + phip;
  return -psip / r;
}

static double orc_force_lattice(const orc_domain *d, const int64_t *off_even, int32_t n_even,
                                const int64_t *off_odd, int32_t n_odd, const orc_tables *t, double rc,
                                const double *x, const double *y, const double *z, const int8_t *type,
                                const double *df, double *fx, double *fy, double *fz, int want_e) {
  const double rc2 = rc * rc;
  const int64_t bz = d->box[2], by = d->box[1], bx = d->box[0];
  double esum = 0.0;
  /* per-row energy partials keep the sum independent of the thread count */
  double *rowe = want_e ? (double *)calloc((size_t)(bz * by), sizeof(double)) : NULL;
#pragma omp parallel for collapse(2) schedule(static)
  for (int64_t zl = 0; zl < bz; zl++)
    for (int64_t yl = 0; yl < by; yl++) {
      double erow = 0.0;
      for (int64_t xl = 0; xl < bx; xl++) {
        const int64_t i = orc_lin(d, xl + d->ghost[0], yl + d->ghost[1], zl + d->ghost[2]);
        const int t0 = type[i];
        if (t0 < 0) continue;
        const int32_t par = orc_parity(d, xl);
        const int64_t *off = par == 0 ? off_even : off_odd;
        const int32_t nk = par == 0 ? n_even : n_odd;
        double ax = 0.0, ay = 0.0, az = 0.0, ep = 0.0;
        for (int32_t k = 0; k < nk; k++) {
          const int64_t j = i + off[k];
          const int tj = type[j];
          if (tj < 0) continue;
          const double dx = x[i] - x[j], dy = y[i] - y[j], dz = z[i] - z[j];
          const double d2 = dx * dx + dy * dy + dz * dz;
          if (d2 >= rc2) continue;
          double phi = 0.0;
          const double fp = orc_fpair(t, t0, tj, sqrt(d2), df[i], df[j], want_e ? &phi : NULL);
          ax += dx * fp;
          ay += dy * fp;
          az += dz * fp;
          ep += phi;
        }
        fx[i] = ax;
        fy[i] = ay;
        fz[i] = az;
        erow += 0.5 * ep;
      }
      if (rowe) rowe[zl * by + yl] = erow;
    }
  if (rowe) {
    for (int64_t r = 0; r < bz * by; r++) esum += rowe[r];
    free(rowe);
  }
  return esum;
}

void orc_force(const orc_domain *d, const int64_t *off_even, int32_t n_even, const int64_t *off_odd,
               int32_t n_odd, const orc_tables *t, double rc, const double *x, const double *y,
               const double *z, const int8_t *type, const double *df, double *fx, double *fy, double *fz,
               double *e_pair, int32_t nthreads) {
  orc_set_threads(nthreads);
  const double e =
      orc_force_lattice(d, off_even, n_even, off_odd, n_odd, t, rc, x, y, z, type, df, fx, fy, fz, e_pair != NULL);
  if (e_pair) *e_pair = e;
}

/* ------------------------------------------------------------- ghost shell */

static inline int64_t orc_wrap(int64_t v, int64_t n, int64_t *img) {
  int64_t q = v / n;
  int64_t r = v - q * n;
  if (r < 0) {
    r += n;
    q -= 1;
  }
  *img = q;
  return r;
}

static void orc_ghost_loop(const orc_domain *d, const double *box_len, double *x, double *y, double *z,
                           int8_t *type, double *df) {
  for (int64_t ze = 0; ze < d->ext[2]; ze++)
    for (int64_t ye = 0; ye < d->ext[1]; ye++)
      for (int64_t xe = 0; xe < d->ext[0]; xe++) {
        const int64_t xl = xe - d->ghost[0], yl = ye - d->ghost[1], zl = ze - d->ghost[2];
        if (xl >= 0 && xl < d->box[0] && yl >= 0 && yl < d->box[1] && zl >= 0 && zl < d->box[2]) continue;
        int64_t ix, iy, iz;
        const int64_t sx = orc_wrap(xl, d->box[0], &ix);
        const int64_t sy = orc_wrap(yl, d->box[1], &iy);
        const int64_t sz = orc_wrap(zl, d->box[2], &iz);
        const int64_t g = orc_lin(d, xe, ye, ze);
        const int64_t s = orc_lin(d, sx + d->ghost[0], sy + d->ghost[1], sz + d->ghost[2]);
        if (df) {
          df[g] = df[s];
        } else {
          type[g] = type[s];
          x[g] = x[s] + (double)ix * box_len[0];
          y[g] = y[s] + (double)iy * box_len[1];
          z[g] = z[s] + (double)iz * box_len[2];
        }
      }
}

void orc_ghost_periodic_xt(const orc_domain *d, const double box_len[3], double *x, double *y, double *z,
                           int8_t *type) {
  orc_ghost_loop(d, box_len, x, y, z, type, NULL);
}

void orc_ghost_periodic_df(const orc_domain *d, double *df) {
  orc_ghost_loop(d, NULL, NULL, NULL, NULL, NULL, df);
}

/* ---------------------------------------------------------------- integrate */

double orc_kinetic(const orc_domain *d, const orc_tables *t, const double *vx, const double *vy,
                   const double *vz, const int8_t *type) {
  double ke = 0.0;
  for (int64_t zl = 0; zl < d->box[2]; zl++)
    for (int64_t yl = 0; yl < d->box[1]; yl++)
      for (int64_t xl = 0; xl < d->box[0]; xl++) {
        const int64_t i = orc_lin(d, xl + d->ghost[0], yl + d->ghost[1], zl + d->ghost[2]);
        const int ti = type[i];
        if (ti < 0) continue;
        ke += 0.5 * t->mass[ti] * (vx[i] * vx[i] + vy[i] * vy[i] + vz[i] * vz[i]);
      }
  return ke * ORC_MVV2E;
}

static void orc_kick(const orc_domain *d, const orc_tables *t, double h, double *vx, double *vy, double *vz,
                     const int8_t *type, const double *fx, const double *fy, const double *fz, double *x,
                     double *y, double *z, double dt) {
  for (int64_t zl = 0; zl < d->box[2]; zl++)
    for (int64_t yl = 0; yl < d->box[1]; yl++)
      for (int64_t xl = 0; xl < d->box[0]; xl++) {
        const int64_t i = orc_lin(d, xl + d->ghost[0], yl + d->ghost[1], zl + d->ghost[2]);
        const int ti = type[i];
        if (ti < 0) continue;
        const double s = h * ORC_FTM2V / t->mass[ti];
        vx[i] += s * fx[i];
        vy[i] += s * fy[i];
        vz[i] += s * fz[i];
        if (x) {
          x[i] += dt * vx[i];
          y[i] += dt * vy[i];
          z[i] += dt * vz[i];
        }
      }
}

void orc_md_run(const orc_domain *d, const int64_t *off_even, int32_t n_even, const int64_t *off_odd,
                int32_t n_odd, const orc_tables *t, double rc, const double box_len[3], double dt,
                int32_t nsteps, double *x, double *y, double *z, double *vx, double *vy, double *vz,
                const int8_t *type_in, double *rho, double *df, double *fx, double *fy, double *fz,
                double *energies, int32_t nthreads) {
  int8_t *type = (int8_t *)type_in; /* ghost refresh rewrites ghost types with identical values */
  for (int32_t s = 0; s < nsteps; s++) {
    orc_kick(d, t, 0.5 * dt, vx, vy, vz, type, fx, fy, fz, x, y, z, dt);
    orc_ghost_periodic_xt(d, box_len, x, y, z, type);
    double ee = 0.0, ep = 0.0;
    orc_rho_df(d, off_even, n_even, off_odd, n_odd, t, rc, x, y, z, type, rho, df, energies ? &ee : NULL,
               nthreads);
    orc_ghost_periodic_df(d, df);
    orc_force(d, off_even, n_even, off_odd, n_odd, t, rc, x, y, z, type, df, fx, fy, fz, energies ? &ep : NULL,
              nthreads);
    orc_kick(d, t, 0.5 * dt, vx, vy, vz, type, fx, fy, fz, NULL, NULL, NULL, dt);
    if (energies) {
      energies[2 * s + 0] = ee + ep;
      energies[2 * s + 1] = orc_kinetic(d, t, vx, vy, vz, type);
    }
  }
}

/* ------------------------------------------------------------ interstitials */

void orc_anchor(double a, double px, double py, double pz, int64_t *X, int64_t *Y, int64_t *Z) {
  /* corner sub-lattice: (2i, j, k) at (i, j, k) a; body sub-lattice:
   * (2i+1, j, k) at (i+1/2, j+1/2, k+1/2) a.  Pick the closer candidate; a tie
   * goes to the corner.  The squared distances are written with fma() so the
   * product library's device code, which uses the same expression, reaches
   * the same decision bit for bit (the choice is integer bookkeeping). */
  const double ux = px / a, uy = py / a, uz = pz / a;
  const double cx = floor(ux + 0.5), cy = floor(uy + 0.5), cz = floor(uz + 0.5);
  const double fx = floor(ux), fy = floor(uy), fz = floor(uz);
  const double ex = ux - cx, ey = uy - cy, ez = uz - cz;
  const double gx = ux - (fx + 0.5), gy = uy - (fy + 0.5), gz = uz - (fz + 0.5);
  const double dc = fma(ex, ex, fma(ey, ey, ez * ez));
  const double db = fma(gx, gx, fma(gy, gy, gz * gz));
  if (dc <= db) {
    *X = 2 * (int64_t)cx;
    *Y = (int64_t)cy;
    *Z = (int64_t)cz;
  } else {
    *X = 2 * (int64_t)fx + 1;
    *Y = (int64_t)fy;
    *Z = (int64_t)fz;
  }
}

static inline double orc_minimg(double v, double L) { return v - L * nearbyint(v / L); }

/* Visits every lattice<->interstitial pair of interstitial k; the callback gets
 * the ext index of the site, the periodic source (local) index and the
 * displacement del = x_site - x_k. */
typedef void (*orc_li_visit)(void *ctx, int32_t k, int64_t ext_i, int64_t src_i, double dx, double dy, double dz,
                             double r);

static void orc_li_pairs(const orc_domain *d, const int32_t *wide3, int32_t n_wide, double a,
                         const double box_len[3], double rc, const double *x, const double *y, const double *z,
                         const int8_t *type, const orc_itl *itl, int32_t k, orc_li_visit fn, void *ctx) {
  /* anchor in global doubled-x coordinates, wrapped into the periodic box */
  const int64_t nX = d->box[0], nY = d->box[1], nZ = d->box[2];
  double px = itl->x[k], py = itl->y[k], pz = itl->z[k];
  int64_t X, Y, Z, img;
  orc_anchor(a, px, py, pz, &X, &Y, &Z);
  int64_t Xw = orc_wrap(X, nX, &img);
  px -= (double)img * box_len[0];
  int64_t Yw = orc_wrap(Y, nY, &img);
  py -= (double)img * box_len[1];
  int64_t Zw = orc_wrap(Z, nZ, &img);
  pz -= (double)img * box_len[2];
  const int32_t par = (int32_t)(Xw & 1);
  const int32_t sign = par ? -1 : 1; /* W stored for parity 0, parity 1 is the negation */
  const double rc2 = rc * rc;
  for (int32_t w = 0; w < n_wide; w++) {
    const int64_t xe = Xw + sign * wide3[3 * w + 0] + d->ghost[0];
    const int64_t ye = Yw + sign * wide3[3 * w + 1] + d->ghost[1];
    const int64_t ze = Zw + sign * wide3[3 * w + 2] + d->ghost[2];
    const int64_t e = orc_lin(d, xe, ye, ze);
    if (type[e] < 0) continue;
    const double dx = x[e] - px, dy = y[e] - py, dz = z[e] - pz;
    const double d2 = dx * dx + dy * dy + dz * dz;
    if (d2 >= rc2) continue;
    int64_t i0, i1, i2;
    const int64_t sx = orc_wrap(xe - d->ghost[0], nX, &i0);
    const int64_t sy = orc_wrap(ye - d->ghost[1], nY, &i1);
    const int64_t sz = orc_wrap(ze - d->ghost[2], nZ, &i2);
    const int64_t src = orc_lin(d, sx + d->ghost[0], sy + d->ghost[1], sz + d->ghost[2]);
    fn(ctx, k, e, src, dx, dy, dz, sqrt(d2));
  }
}

typedef struct {
  const orc_tables *t;
  const int8_t *type;
  double *rho;
  const double *df;
  double *fx, *fy, *fz;
  orc_itl *itl;
  double e;
} orc_li_ctx;

static void orc_li_rho(void *vc, int32_t k, int64_t e, int64_t src, double dx, double dy, double dz, double r) {
  orc_li_ctx *c = (orc_li_ctx *)vc;
  (void)dx;
  (void)dy;
  (void)dz;
  const int ti = c->type[e], tk = c->itl->type[k];
  c->rho[src] += orc_spline_value(&c->t->rho[tk], r);
  c->itl->rho[k] += orc_spline_value(&c->t->rho[ti], r);
}

static void orc_li_force(void *vc, int32_t k, int64_t e, int64_t src, double dx, double dy, double dz, double r) {
  orc_li_ctx *c = (orc_li_ctx *)vc;
  const int ti = c->type[e], tk = c->itl->type[k];
  double phi = 0.0;
  const double fp = orc_fpair(c->t, ti, tk, r, c->df[src], c->itl->df[k], &phi);
  c->fx[src] += dx * fp;
  c->fy[src] += dy * fp;
  c->fz[src] += dz * fp;
  c->itl->fx[k] -= dx * fp;
  c->itl->fy[k] -= dy * fp;
  c->itl->fz[k] -= dz * fp;
  c->e += phi; /* one full phi per lattice<->interstitial pair */
}

void orc_rho_df_itl(const orc_domain *d, const int64_t *off_even, int32_t n_even, const int64_t *off_odd,
                    int32_t n_odd, const int32_t *wide3, int32_t n_wide, double a, const double box_len[3],
                    const orc_tables *t, double rc, const double *x, const double *y, const double *z,
                    const int8_t *type, double *rho, double *df, orc_itl *itl, double *e_embed,
                    int32_t nthreads) {
  orc_set_threads(nthreads);
  orc_rho_lattice(d, off_even, n_even, off_odd, n_odd, t, rc, x, y, z, type, rho);
  orc_li_ctx c = {t, type, rho, NULL, NULL, NULL, NULL, itl, 0.0};
  for (int32_t k = 0; k < itl->n; k++) itl->rho[k] = 0.0;
  for (int32_t k = 0; k < itl->n; k++) orc_li_pairs(d, wide3, n_wide, a, box_len, rc, x, y, z, type, itl, k, orc_li_rho, &c);
  const double rc2 = rc * rc;
  for (int32_t k = 0; k < itl->n; k++)
    for (int32_t m = 0; m < itl->n; m++) {
      if (m == k) continue;
      const double dx = orc_minimg(itl->x[k] - itl->x[m], box_len[0]);
      const double dy = orc_minimg(itl->y[k] - itl->y[m], box_len[1]);
      const double dz = orc_minimg(itl->z[k] - itl->z[m], box_len[2]);
      const double d2 = dx * dx + dy * dy + dz * dz;
      if (d2 >= rc2) continue;
      itl->rho[k] += orc_spline_value(&t->rho[itl->type[m]], sqrt(d2));
    }
  double e = orc_df_and_embed(d, t, type, rho, df, e_embed != NULL);
  for (int32_t k = 0; k < itl->n; k++) {
    itl->df[k] = orc_spline_deriv(&t->emb[itl->type[k]], itl->rho[k]);
    if (e_embed) e += orc_spline_value(&t->emb[itl->type[k]], itl->rho[k]);
  }
  if (e_embed) *e_embed = e;
}

void orc_force_itl(const orc_domain *d, const int64_t *off_even, int32_t n_even, const int64_t *off_odd,
                   int32_t n_odd, const int32_t *wide3, int32_t n_wide, double a, const double box_len[3],
                   const orc_tables *t, double rc, const double *x, const double *y, const double *z,
                   const int8_t *type, const double *df, double *fx, double *fy, double *fz, orc_itl *itl,
                   double *e_pair, int32_t nthreads) {
  orc_set_threads(nthreads);
  double e = orc_force_lattice(d, off_even, n_even, off_odd, n_odd, t, rc, x, y, z, type, df, fx, fy, fz,
                               e_pair != NULL);
  orc_li_ctx c = {t, type, NULL, df, fx, fy, fz, itl, 0.0};
  for (int32_t k = 0; k < itl->n; k++) itl->fx[k] = itl->fy[k] = itl->fz[k] = 0.0;
  for (int32_t k = 0; k < itl->n; k++)
    orc_li_pairs(d, wide3, n_wide, a, box_len, rc, x, y, z, type, itl, k, orc_li_force, &c);
  const double rc2 = rc * rc;
  double eii = 0.0;
  for (int32_t k = 0; k < itl->n; k++)
    for (int32_t m = 0; m < itl->n; m++) {
      if (m == k) continue;
      const double dx = orc_minimg(itl->x[k] - itl->x[m], box_len[0]);
      const double dy = orc_minimg(itl->y[k] - itl->y[m], box_len[1]);
      const double dz = orc_minimg(itl->z[k] - itl->z[m], box_len[2]);
      const double d2 = dx * dx + dy * dy + dz * dz;
      if (d2 >= rc2) continue;
      double phi = 0.0;
      const double fp = orc_fpair(t, itl->type[k], itl->type[m], sqrt(d2), itl->df[k], itl->df[m], &phi);
      itl->fx[k] += dx * fp;
      itl->fy[k] += dy * fp;
      itl->fz[k] += dz * fp;
      eii += 0.5 * phi;
    }
  if (e_pair) *e_pair = e + c.e + eii;
}

void orc_md_run_itl(const orc_domain *d, const int64_t *off_even, int32_t n_even, const int64_t *off_odd,
                    int32_t n_odd, const int32_t *wide3, int32_t n_wide, double a, const orc_tables *t, double rc,
                    const double box_len[3], double dt, int32_t nsteps, double *x, double *y, double *z, double *vx,
                    double *vy, double *vz, const int8_t *type_in, double *rho, double *df, double *fx, double *fy,
                    double *fz, orc_itl *itl, double *energies, int32_t nthreads) {
  int8_t *type = (int8_t *)type_in;
  for (int32_t s = 0; s < nsteps; s++) {
    orc_kick(d, t, 0.5 * dt, vx, vy, vz, type, fx, fy, fz, x, y, z, dt);
    for (int32_t k = 0; k < itl->n; k++) {
      const double sc = 0.5 * dt * ORC_FTM2V / t->mass[itl->type[k]];
      itl->vx[k] += sc * itl->fx[k];
      itl->vy[k] += sc * itl->fy[k];
      itl->vz[k] += sc * itl->fz[k];
      itl->x[k] += dt * itl->vx[k];
      itl->y[k] += dt * itl->vy[k];
      itl->z[k] += dt * itl->vz[k];
    }
    orc_ghost_periodic_xt(d, box_len, x, y, z, type);
    double ee = 0.0, ep = 0.0;
    orc_rho_df_itl(d, off_even, n_even, off_odd, n_odd, wide3, n_wide, a, box_len, t, rc, x, y, z, type, rho, df, itl,
                   energies ? &ee : NULL, nthreads);
    orc_ghost_periodic_df(d, df);
    orc_force_itl(d, off_even, n_even, off_odd, n_odd, wide3, n_wide, a, box_len, t, rc, x, y, z, type, df, fx, fy, fz,
                  itl, energies ? &ep : NULL, nthreads);
    orc_kick(d, t, 0.5 * dt, vx, vy, vz, type, fx, fy, fz, NULL, NULL, NULL, dt);
    double ke_itl = 0.0;
    for (int32_t k = 0; k < itl->n; k++) {
      const double m = t->mass[itl->type[k]];
      const double sc = 0.5 * dt * ORC_FTM2V / m;
      itl->vx[k] += sc * itl->fx[k];
      itl->vy[k] += sc * itl->fy[k];
      itl->vz[k] += sc * itl->fz[k];
      ke_itl += 0.5 * m * (itl->vx[k] * itl->vx[k] + itl->vy[k] * itl->vy[k] + itl->vz[k] * itl->vz[k]);
    }
    if (energies) {
      energies[2 * s + 0] = ee + ep;
      energies[2 * s + 1] = orc_kinetic(d, t, vx, vy, vz, type) + ke_itl * ORC_MVV2E;
    }
  }
}

/* ------------------------------------------------- interstitial migration */

/* Wigner-Seitz re-assignment between the lattice and the interstitial list.
 * MISA-MD core keeps off-lattice atoms apart from the lattice array (the
 * plugin skips them, kernel_itl.hpp:44-46, :52-54); the core's own transition
 * rules are not in the reference tree, so the rules are stated here:
 *   1. in increasing local site order (z, then y, then doubled x), a lattice
 *      atom whose nearest site (orc_anchor) is not its own site is appended to
 *      the list with every field; its site becomes a vacancy (type -1; v, f,
 *      rho, df zeroed; position kept);
 *   2. in list order, an entry whose nearest site (wrapped into the periodic
 *      box) is vacant moves into it, its position shifted by the periodic
 *      image; visiting in increasing order means the lowest index wins a site
 *      several entries claim.  The entries that moved leave the list; the rest
 *      keep their order.
 * Then the ghost shell's positions, types and df are refreshed.  A single
 * periodic domain (start = 0) only.  Returns 0; -1 (nothing changed) when the
 * list would exceed cap; -2 for a domain with start != 0. */
int32_t orc_migrate(const orc_domain *d, double a, const double box_len[3], double *x, double *y, double *z,
                    double *vx, double *vy, double *vz, double *fx, double *fy, double *fz, double *rho, double *df,
                    int8_t *type, orc_itl *itl, int32_t cap, int32_t *n_to_itl, int32_t *n_to_lattice) {
  if (d->start[0] || d->start[1] || d->start[2]) return -2;
  const int64_t nX = d->box[0], nY = d->box[1], nZ = d->box[2];
  int64_t leave = 0;
  for (int pass = 0; pass < 2; pass++) {
    int32_t n = itl->n;
    for (int64_t zl = 0; zl < nZ; zl++)
      for (int64_t yl = 0; yl < nY; yl++)
        for (int64_t xl = 0; xl < nX; xl++) {
          const int64_t i = orc_lin(d, xl + d->ghost[0], yl + d->ghost[1], zl + d->ghost[2]);
          if (type[i] < 0) continue;
          int64_t X, Y, Z;
          orc_anchor(a, x[i], y[i], z[i], &X, &Y, &Z);
          if (X == xl && Y == yl && Z == zl) continue;
          if (pass == 0) {
            leave++;
            continue;
          }
          const int32_t k = n++;
          itl->x[k] = x[i], itl->y[k] = y[i], itl->z[k] = z[i];
          itl->vx[k] = vx[i], itl->vy[k] = vy[i], itl->vz[k] = vz[i];
          itl->fx[k] = fx[i], itl->fy[k] = fy[i], itl->fz[k] = fz[i];
          itl->rho[k] = rho[i], itl->df[k] = df[i];
          itl->type[k] = type[i];
          type[i] = -1;
          vx[i] = vy[i] = vz[i] = 0.0;
          fx[i] = fy[i] = fz[i] = 0.0;
          rho[i] = df[i] = 0.0;
        }
    if (pass == 0 && (int64_t)itl->n + leave > cap) return -1;
    if (pass == 1) itl->n = n;
  }
  int32_t kept = 0, filled = 0;
  for (int32_t k = 0; k < itl->n; k++) {
    int64_t X, Y, Z, ix, iy, iz;
    orc_anchor(a, itl->x[k], itl->y[k], itl->z[k], &X, &Y, &Z);
    X = orc_wrap(X, nX, &ix);
    Y = orc_wrap(Y, nY, &iy);
    Z = orc_wrap(Z, nZ, &iz);
    const int64_t i = orc_lin(d, X + d->ghost[0], Y + d->ghost[1], Z + d->ghost[2]);
    if (type[i] < 0) {
      x[i] = fma(-(double)ix, box_len[0], itl->x[k]);
      y[i] = fma(-(double)iy, box_len[1], itl->y[k]);
      z[i] = fma(-(double)iz, box_len[2], itl->z[k]);
      vx[i] = itl->vx[k], vy[i] = itl->vy[k], vz[i] = itl->vz[k];
      fx[i] = itl->fx[k], fy[i] = itl->fy[k], fz[i] = itl->fz[k];
      rho[i] = itl->rho[k], df[i] = itl->df[k];
      type[i] = itl->type[k];
      filled++;
      continue;
    }
    if (kept != k) {
      itl->x[kept] = itl->x[k], itl->y[kept] = itl->y[k], itl->z[kept] = itl->z[k];
      itl->vx[kept] = itl->vx[k], itl->vy[kept] = itl->vy[k], itl->vz[kept] = itl->vz[k];
      itl->fx[kept] = itl->fx[k], itl->fy[kept] = itl->fy[k], itl->fz[kept] = itl->fz[k];
      itl->rho[kept] = itl->rho[k], itl->df[kept] = itl->df[k];
      itl->type[kept] = itl->type[k];
    }
    kept++;
  }
  itl->n = kept;
  *n_to_itl = (int32_t)leave;
  *n_to_lattice = filled;
  orc_ghost_periodic_xt(d, box_len, x, y, z, type);
  orc_ghost_periodic_df(d, df);
  return 0;
}
