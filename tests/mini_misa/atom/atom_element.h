// Stand-in for MISA-MD core's atom record (atom/atom_element.h): the fields the
// reference plugin mirrors as _cuAtomElement (src/kernel_types.h:40-48), 104 B.
#pragma once
#include <cstdint>

struct AtomElement {
  int64_t id;
  int32_t type;
  double x[3], v[3], f[3];
  double rho, df;
};
static_assert(sizeof(AtomElement) == 104, "record size of the reference's _cuAtomElement");
