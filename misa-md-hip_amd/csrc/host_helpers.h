// Internal host-side helpers shared by the library's translation units.
#ifndef MMD_HOST_HELPERS_H
#define MMD_HOST_HELPERS_H

#include <cstdint>

namespace mmd {

// Records a message for mmd_last_error() and returns `code`.
int fail(int code, const char *fmt, ...);
void clear_error();
const char *last_error_cstr();

// LAMMPS-style cubic Hermite spline: 7 coefficients per knot (see
// host_helpers.cpp for the layout).
void spline7(const double *f, int n, double dx, double *c7);
double spline7_value(const double *c7, int n, double x0, double inv_dx, double x);
double spline7_deriv(const double *c7, int n, double x0, double inv_dx, double x);

// Counter-based random numbers keyed by (seed, id, stream).
uint64_t rng_bits(uint64_t seed, uint64_t id, uint64_t stream);
double rng_u01(uint64_t seed, uint64_t id, uint64_t stream);
double rng_gauss(uint64_t seed, uint64_t id, uint64_t stream);

}  // namespace mmd

#endif
