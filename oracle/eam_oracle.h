/*
 * eam_oracle.h -- CPU FP64 restatement of MISA-MD-hip's EAM hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this code.  It is used by tests/ (as the checker), by
 * __graft_entry__.smoke() (as the checker) and by bench.py's cpu_baseline leg
 * (timed as the "port" CPU baseline).
 *
 * Data layout follows the reference plugin exactly: one linear array over the
 * ghost-extended lattice, x dimension doubled (two BCC sub-lattice sites per
 * unit cell along x), linear index ((z * ext_y) + y) * ext_x + x
 * (src/kernels/kernel_itl.hpp:21-26), neighbour stencils given as signed linear
 * offsets selected by the parity of the global doubled-x index
 * (src/kernels/kernel_itl.hpp:82-86).
 *
 * Spline convention: the hip-potential v0.2.0 device tables are absent from the
 * reference tree; the restatement uses the published LAMMPS pair_eam
 * interpolation (7 coefficients per knot, derivative estimates from a 5-point
 * finite-difference stencil), which MISA-MD's eam tables feed.  Parity of this
 * spline with hip-potential itself is unpinned; the spline is pinned by closed
 * forms (exact knot values, exact reproduction of cubics in the interior,
 * O(dx^4) error on smooth functions).
 */
#ifndef MMD_EAM_ORACLE_H
#define MMD_EAM_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Mirrors _hipDeviceDomain (src/kernel_types.h:51-64). */
typedef struct {
  int32_t ghost[3];
  int32_t box[3];
  int32_t ext[3];
  int64_t start[3];
} orc_domain;

/* One tabulated function: n knots at x0 + i*dx, 7 coefficients per knot. */
typedef struct {
  int32_t n;
  double x0;
  double dx;
  double inv_dx;
  const double *c; /* [n][7] */
} orc_spline;

typedef struct {
  int32_t n_types;
  const orc_spline *rho; /* [n_types]            rho_t(r) by source type */
  const orc_spline *emb; /* [n_types]            F_t(rho)                */
  const orc_spline *phi; /* [n_types * n_types]  phi_ab(r), symmetric    */
  const double *mass;    /* [n_types] amu                                */
} orc_tables;

/* Builds the 7 coefficients per knot from n knot values f (LAMMPS pair_eam
 * interpolate): c[m][6]=f, c[m][5]=slope estimate, c[m][4], c[m][3] the cubic
 * Hermite terms, c[m][2..0] the derivative polynomial divided by dx. */
void orc_spline_build(const double *f, int32_t n, double dx, double *c7);
double orc_spline_value(const orc_spline *s, double x);
double orc_spline_deriv(const orc_spline *s, double x);

/* Neighbour stencil of one BCC sub-lattice parity (0 = cube corner, 1 = body
 * centre): every (dX,dY,dZ) != 0 in doubled-x lattice units whose ideal
 * distance is strictly below radius.  Ordered by ideal distance, then dZ, dY,
 * dX for parity 0; parity 1 lists the negated parity-0 entries in the same
 * order, so entry k has the same ideal distance for both parities.  Returns the
 * count (or -1 if cap is too small); out is [count][3]. */
int32_t orc_stencil3(double a, double radius, int32_t parity, int32_t *out, int32_t cap);

/* rho + fused df over local sites (src/kernels/kernel_itl.hpp:69-105 with
 * pair_rho src/kernels/kernel_pairs.hpp:21-32).  Sites with type < 0 are
 * skipped and left untouched.  Overwrites rho and df.  If e_embed is not NULL
 * it receives sum_i F_{t_i}(rho_i). */
void orc_rho_df(const orc_domain *d, const int64_t *off_even, int32_t n_even, const int64_t *off_odd,
                int32_t n_odd, const orc_tables *t, double rc, const double *x, const double *y,
                const double *z, const int8_t *type, double *rho, double *df, double *e_embed,
                int32_t nthreads);

/* Standalone df = F'_{t_i}(rho_i) (src/kernels/hip_kernels.hip.cpp:21-40). */
void orc_df(const orc_domain *d, const orc_tables *t, const int8_t *type, const double *rho, double *df);

/* Force pass (pair_force src/kernels/kernel_pairs.hpp:39-60): overwrites f on
 * local sites with type >= 0.  If e_pair is not NULL it receives
 * 0.5 * sum_i sum_j phi(r_ij) over local sites. */
void orc_force(const orc_domain *d, const int64_t *off_even, int32_t n_even, const int64_t *off_odd,
               int32_t n_odd, const orc_tables *t, double rc, const double *x, const double *y,
               const double *z, const int8_t *type, const double *df, double *fx, double *fy, double *fz,
               double *e_pair, int32_t nthreads);

/* Single-domain periodic ghost refresh: every ghost site copies its periodic
 * source site (type, and position shifted by the box length, or df). */
void orc_ghost_periodic_xt(const orc_domain *d, const double box_len[3], double *x, double *y, double *z,
                           int8_t *type);
void orc_ghost_periodic_df(const orc_domain *d, double *df);

/* Velocity Verlet over nsteps (kick-drift, ghost x, rho+df, ghost df, force,
 * kick); forces in fx/fy/fz must be current on entry.  energies (optional)
 * receives [nsteps][2] = {potential, kinetic} after each step, in eV. */
void orc_md_run(const orc_domain *d, const int64_t *off_even, int32_t n_even, const int64_t *off_odd,
                int32_t n_odd, const orc_tables *t, double rc, const double box_len[3], double dt,
                int32_t nsteps, double *x, double *y, double *z, double *vx, double *vy, double *vz,
                const int8_t *type, double *rho, double *df, double *fx, double *fy, double *fz,
                double *energies, int32_t nthreads);

/* Kinetic energy (eV) of local sites with type >= 0. */
double orc_kinetic(const orc_domain *d, const orc_tables *t, const double *vx, const double *vy,
                   const double *vz, const int8_t *type);

/* ---- interstitials (off-lattice atoms) ------------------------------------
 * Interstitial k interacts with the lattice sites of the wide stencil W around
 * its anchor (the nearest ideal lattice site of its position, wrapped into the
 * periodic box) and with every other interstitial within rc (minimum image).
 * Lattice<->interstitial pairs are the set {(i,k): site(i) in W(anchor(k)),
 * type_i >= 0, r_ik < rc}; a ghost site in that set stands for its periodic
 * source site.  Contributions are added on top of the lattice-only passes. */
typedef struct {
  int32_t n;
  double *x, *y, *z;
  double *vx, *vy, *vz;
  double *fx, *fy, *fz;
  double *rho, *df;
  int8_t *type;
} orc_itl;

/* Nearest ideal lattice site (global doubled-x coordinates) of a position. */
void orc_anchor(double a, double px, double py, double pz, int64_t *X, int64_t *Y, int64_t *Z);

/* rho: lattice part from the stencil pass plus interstitial contributions;
 * then df for lattice sites and interstitials.  wide3 is W as [n_wide][3]. */
void orc_rho_df_itl(const orc_domain *d, const int64_t *off_even, int32_t n_even, const int64_t *off_odd,
                    int32_t n_odd, const int32_t *wide3, int32_t n_wide, double a, const double box_len[3],
                    const orc_tables *t, double rc, const double *x, const double *y, const double *z,
                    const int8_t *type, double *rho, double *df, orc_itl *itl, double *e_embed,
                    int32_t nthreads);
void orc_force_itl(const orc_domain *d, const int64_t *off_even, int32_t n_even, const int64_t *off_odd,
                   int32_t n_odd, const int32_t *wide3, int32_t n_wide, double a, const double box_len[3],
                   const orc_tables *t, double rc, const double *x, const double *y, const double *z,
                   const int8_t *type, const double *df, double *fx, double *fy, double *fz, orc_itl *itl,
                   double *e_pair, int32_t nthreads);

/* Velocity Verlet with interstitials (kick-drift of lattice atoms and
 * interstitials, ghost x, rho+df with interstitials, ghost df, force with
 * interstitials, kick).  Forces of both lists must be current on entry. */
void orc_md_run_itl(const orc_domain *d, const int64_t *off_even, int32_t n_even, const int64_t *off_odd,
                    int32_t n_odd, const int32_t *wide3, int32_t n_wide, double a, const orc_tables *t, double rc,
                    const double box_len[3], double dt, int32_t nsteps, double *x, double *y, double *z, double *vx,
                    double *vy, double *vz, const int8_t *type, double *rho, double *df, double *fx, double *fy,
                    double *fz, orc_itl *itl, double *energies, int32_t nthreads);

/* Wigner-Seitz migration between the lattice and the interstitial list of a
 * single periodic domain (rules in eam_oracle.c); itl arrays hold cap
 * entries.  Returns 0, -1 over capacity (nothing changed), -2 for start != 0. */
int32_t orc_migrate(const orc_domain *d, double a, const double box_len[3], double *x, double *y, double *z,
                    double *vx, double *vy, double *vz, double *fx, double *fy, double *fz, double *rho, double *df,
                    int8_t *type, orc_itl *itl, int32_t cap, int32_t *n_to_itl, int32_t *n_to_lattice);

#ifdef __cplusplus
}
#endif
#endif
