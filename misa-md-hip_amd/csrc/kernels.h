// Host-callable launchers of the library's gfx950 kernels (internal).
#ifndef MMD_KERNELS_H
#define MMD_KERNELS_H

#include <hip/hip_runtime.h>

#include "device_common.h"

namespace mmd {

// local-site blocks (streaming passes), pair-pass tile blocks; both multiples of the 8 XCDs
int64_t site_blocks(const Geom &g);
int64_t tile_blocks(const Geom &g);

// EAM passes; off = the stencil's device image (stencil_image): LDS offsets within the tile, energy masks.
// lean (mixed, MD steps): rho and df from the quadratic records; with e_part the embedding energy
// from a second sum over the cubic value records.
// pair passes over tiles [lb0, lb1) (lb1 < 0: to the last tile)
hipError_t launch_rho_df(hipStream_t st, const Geom &g, const Halo &h, const int32_t *off, int32_t nk_even,
                         int32_t nk_odd, const double4 *P, double *RHO, double *DF, const TabF64 &tf,
                         const TabMixed &tm, double rc2, int prec, double *e_part, bool lean, int64_t lb0 = 0,
                         int64_t lb1 = -1);
hipError_t launch_df(hipStream_t st, const Geom &g, const int8_t *T, const double *RHO, double *DF,
                     const TabF64 &tf);
hipError_t launch_force(hipStream_t st, const Geom &g, const Halo &h, const int32_t *off, int32_t nk_even,
                        int32_t nk_odd, const double4 *P, const double *DF, double *FX, double *FY, double *FZ,
                        const TabF64 &tf, const TabMixed &tm, double rc2, int prec, double *e_part, int64_t lb0 = 0,
                        int64_t lb1 = -1);
hipError_t pair_kernels_allow_lds(size_t bytes);
// Newton's-third-law half stencil (periodic step only): forward pass (own
// force + pair-force slots SL[nf_max][n_local]) then backward pass (adds the
// reactions read from the slots); img = half_stencil_image, nf / nb = forward
// / backward entries per parity
hipError_t launch_force_half(hipStream_t st, const Geom &g, const Halo &h, const int32_t *img, const int32_t nf[2],
                             const int32_t nb[2], const double4 *P, const double *DF, double *FX, double *FY,
                             double *FZ, double *SL, const TabF64 &tf, const TabMixed &tm, double rc2, int prec,
                             double *e_part);
size_t half_back_lds_bytes(const Halo &h);
hipError_t half_kernels_allow_lds(size_t bytes);

// layout conversion between the reference layout (host view) and the device layout;
// positions: three staged reference-layout arrays <-> (x, y, z, type) records
hipError_t launch_ref2dev_pos(hipStream_t st, const Geom &g, const double *rx, const double *ry, const double *rz,
                              const int8_t *T, double4 *P, bool set_xyz);
hipError_t launch_dev2ref_pos(hipStream_t st, const Geom &g, const double4 *P, double *rx, double *ry, double *rz);
hipError_t launch_ref2dev_f64(hipStream_t st, const Geom &g, const double *ref, double *dev);
hipError_t launch_dev2ref_f64(hipStream_t st, const Geom &g, const double *dev, double *ref);
hipError_t launch_ref2dev_i8(hipStream_t st, const Geom &g, const int8_t *ref, int8_t *dev);
hipError_t launch_dev2ref_i8(hipStream_t st, const Geom &g, const int8_t *dev, int8_t *ref);

// host AoS records (reference _cuAtomElement or caller-described) <-> device SoA
struct AosDesc {
  int64_t stride, off_type, off_x, off_v, off_f, off_rho, off_df;
  int32_t type_bytes;
};
// largest type unpacked: the maximum over kMaxTypeSlots ints kMaxTypeStride apart
constexpr int kMaxTypeSlots = 64, kMaxTypeStride = 32;
// device-layout ext sites [i0, i1) (i1 < 0: to the last)
hipError_t launch_aos_unpack(hipStream_t st, const Geom &g, const char *aos, AosDesc d, uint32_t fields, double4 *P,
                             int8_t *T, double *DF, double *RHO, int32_t n_types, int32_t *max_type, int64_t i0 = 0,
                             int64_t i1 = -1);
// local sites [q0, q1) in device order (q1 < 0: to the last)
hipError_t launch_aos_pack(hipStream_t st, const Geom &g, char *aos, AosDesc d, uint32_t fields, const int8_t *T,
                           const double *RHO, const double *DF, const double *FX, const double *FY,
                           const double *FZ, int64_t q0 = 0, int64_t q1 = -1);

// periodic single-domain ghost refresh
hipError_t launch_ghost(hipStream_t st, const Geom &g, double Lx, double Ly, double Lz, double4 *P, int8_t *T,
                        double *DF, uint32_t fields);

// velocity Verlet: v += h a; optionally x += dt v.  ke_part (optional) gets the
// kinetic energy partials of the updated velocities (half-kick variant).
hipError_t launch_kick(hipStream_t st, const Geom &g, const Masses &M, double h, double dt, bool drift, double4 *P,
                       double *VX, double *VY, double *VZ, const double *FX, const double *FY, const double *FZ,
                       const int8_t *T, double *ke_part);

// the closing half kick of a step and the opening kick + drift of the next in
// one pass; ke_part (optional) gets the completed step's kinetic energy partials
hipError_t launch_kick_fused(hipStream_t st, const Geom &g, const Masses &M, double h, double dt, double4 *P,
                             double *VX, double *VY, double *VZ, const double *FX, const double *FY, const double *FZ,
                             const int8_t *T, double *ke_part);

// energies: out[slot] = {sum(ep_a) + sum(ep_b), sum(ke)}, slot = (*counter)++
hipError_t launch_energy_finalize(hipStream_t st, const double *ep_a, int64_t na, const double *ep_b, int64_t nb,
                                  const double *ep_c, int64_t nc, const double *ke, int64_t nk, const double *ke_b,
                                  int64_t nkb, double *out, int32_t *counter);
hipError_t launch_stream_copy(hipStream_t st, const void *src, void *dst, int64_t bytes);
hipError_t launch_fill_i32(hipStream_t st, int32_t *p, int64_t n, int32_t v);
hipError_t launch_fill_f64(hipStream_t st, double *p, int64_t n, double v);
hipError_t launch_fill_i8(hipStream_t st, int8_t *p, int64_t n, int8_t v);

}  // namespace mmd

#endif
