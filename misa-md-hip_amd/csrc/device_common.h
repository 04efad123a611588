// Device-side geometry, table descriptors and small helpers shared by the
// gfx950 kernels.
//
// Device layout ("sub-lattice split rows"): the ghost-extended lattice is
// stored as rows of one BCC sub-lattice each.  Site (xe, ye, ze) of the
// reference layout (x doubled, parity s = xe & 1, cell xh = xe >> 1) lives at
//     ((ze * ext_y + ye) * 2 + s) * exh + xh,      exh = ext_x / 2.
// A wavefront takes 64 consecutive cells of one row, so every lane of a wave
// has the same sub-lattice parity: the neighbour stencil is wave-uniform
// (scalar loads) and the k-th neighbours of the 64 lanes are 64 consecutive
// sites of one row (fully coalesced 512-byte loads per coordinate).
#ifndef MMD_DEVICE_COMMON_H
#define MMD_DEVICE_COMMON_H

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mmd {

constexpr int kWave = 64;
constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;
// Force-pass kernels: their LDS tile fits three blocks on a CU (12 waves,
// 3 per SIMD), so their registers are capped to keep 3 waves per SIMD
// (at most 168 VGPRs).
constexpr int kForceWavesPerSimd = 3;
constexpr int kMaxStencil = 512;  // entries per parity
// Device image of a stencil for the pair passes (stencil_image): [2][kMaxStencil]
// LDS offsets, rows zero-padded; then [2][kMaxStencil / 32] bit masks of the
// entries that count a pair's energy (each unordered pair on exactly one side
// when the stencil is symmetric); then one word, 1 when the masks hold that
// half (pair energy weight 1) and 0 when every entry counts with weight 1/2.
constexpr int kStencilBits = 2 * kMaxStencil;
constexpr int kStencilHalfFlag = kStencilBits + 2 * (kMaxStencil / 32);
constexpr int kStencilImage = kStencilHalfFlag + 1;
// Device image of the Newton's-third-law half stencil (half_stencil_image):
// [2][kMaxStencil] LDS offsets of the forward entries (ideal displacement
// lexicographically positive in (z, y, x)), then [2][kMaxStencil] of the
// backward entries, then [2][kMaxStencil] giving for each backward entry the
// number of the forward entry, at the neighbour's parity, that reaches back
// to the centre.  Rows are zero-padded.
constexpr int kHalfFwd = 0, kHalfBack = 2 * kMaxStencil, kHalfSlot = 4 * kMaxStencil;
constexpr int kHalfImage = 6 * kMaxStencil;
constexpr int kMaxTypes = 8;
constexpr int kXcds = 8;          // MI355X: 8 accelerator complex dies, blocks dispatched round-robin

// Pair-pass tiles: a block of 256 threads computes kTileX x kTileY x kTileZ
// cells (both sub-lattices) after staging the tile plus its stencil halo in LDS.
constexpr int kTileX = 8, kTileY = 4, kTileZ = 4;

struct Geom {
  int32_t ext_x, ext_y, ext_z;  // ext_x in doubled units
  int32_t exh;                  // ext_x / 2 (cells per row)
  int32_t gxh, gy, gz;          // ghost width (x in cells)
  int32_t bxh, by, bz;          // local box (x in cells)
  int64_t n_ext;                // ext sites
  int64_t n_local;              // local sites (2 bxh by bz)
  int64_t n_ghost;              // ghost sites (n_ext - n_local)
  int32_t ntx, nty, ntz;        // pair-pass tiles along x, y, z
  int64_t n_tiles;
};

// Stencil reach (cells along x, rows along y and z) and the LDS tile it
// implies: lx x ly x lz cells of both parities, n records.
struct Halo {
  int32_t rx, ry, rz;
  int32_t lx, ly, lz;
  int32_t n;
};

// Unlike-pair queue of the force pass, one per wave: kSlots record numbers
// (uint32) followed by kSlots doubles.
constexpr int kSlots = kWave;
constexpr size_t kSlotWaveBytes = (size_t)kSlots * 12;

// LDS offset of the force pass's queues: after the staged records, 16-byte aligned
__host__ __device__ inline size_t tile_queue_offset(const Halo &h) { return ((size_t)h.n * 33 + 15) / 16 * 16; }

// LDS bytes of a pair-pass block: (x, y), z or (z, df), type per staged
// record; the force pass adds the per-wave unlike-pair queues
__host__ __device__ inline size_t tile_lds_bytes(const Halo &h, bool with_df) {
  if (!with_df) return (size_t)h.n * 25;
  return tile_queue_offset(h) + (size_t)kWavesPerBlock * kSlotWaveBytes;
}

// Loads through a uniform base pointer with a 32-bit element index, so the
// address is an SGPR base + 32-bit VGPR offset (no 64-bit VALU arithmetic).
// The byte offset is 32-bit: only for arrays below 4 GiB (the spline tables,
// at most 2^24 records of 48 bytes: mmd_set_tables checks the record count).
// Lattice arrays (up to 2^31 sites) are indexed with 64-bit offsets.
template <typename T>
__device__ __forceinline__ T ld(const T *__restrict__ base, uint32_t idx) {
  return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + idx * (uint32_t)sizeof(T));
}

// FP64 cubic spline of one function family (all types or type pairs back to
// back, record = function * n + knot).  The seven coefficients per knot are
// split into a value record and derivative arrays, so a lane's gather touches
// as few cache lines as possible:
//   value = ((c3 p + c4) p + c5) p + c6   from val[2r] = (c3, c4), val[2r+1] = (c5, c6)
//   deriv = (c0 p + c1) p + c2            from d01 = (c0, c1), d2 = c2
// (the value record's two 16-byte halves share a cache line)
struct SplF64 {
  const double2 *val, *d01;
  const double *d2;
};

struct TabF64 {
  SplF64 rho;  // [nt][n_r]      rho_t(r) by source type
  SplF64 emb;  // [nt][n_rho]    F_t(rho)
  SplF64 phi;  // [nt*nt][n_r]   phi_ab(r)
  // force records per type pair (a, b) and knot: the derivative coefficients
  // of rho_b and phi_ab side by side, three 16-byte loads per pair:
  //   fc[3r] = (rho_b c0, c1), fc[3r+1] = (rho_b c2, phi_ab c0), fc[3r+2] = (phi_ab c1, c2)
  const double2 *fc;  // [nt*nt][n_r][3]
  int32_t nt, n_r, n_rho;
  double r0, inv_dr, rho0, inv_drho;
  double dr;  // knot spacing of the r tables
};

__device__ __forceinline__ double spl_value(const SplF64 &s, uint32_t rec, double p) {
  const double2 a = ld(s.val, 2 * rec), b = ld(s.val, 2 * rec + 1);
  return ((a.x * p + a.y) * p + b.x) * p + b.y;
}

__device__ __forceinline__ double spl_deriv(const SplF64 &s, uint32_t rec, double p) {
  const double2 a = ld(s.d01, rec);
  return (a.x * p + a.y) * p + ld(s.d2, rec);
}

// Mixed-precision spline keyed by s = r^2 (no square root or division per
// pair).  Per knot, with q the fraction of the segment:
//   value = base + q * (a1 + q * c),   c = a2 + q * a3   (c in FP32)
//   d/ds  = dbase + q * e,             e = b1 + q * b2   (e in FP32)
// The FP64 base (and the value's linear term) carry the leading digits; the
// FP32 terms are corrections of relative size ds * f''/f' (a few per cent)
// or smaller, so their rounding stays far below the table's interpolation
// error.  A derivative record is 16 bytes — one load: (dbase, bits of (b1, b2)).
struct SplMixed {
  const double2 *vb;  // (base, a1)
  const float2 *vc;   // (a2, a3)
  const double2 *dp;  // -2 x (dbase, (b1, b2) as two floats): the force pass's form
};

struct TabMixed {
  SplMixed rho;  // [nt][n_s]
  SplMixed phi;  // [nt*nt][n_s]
  // force records per type pair (a, b) and knot: fb[2r] = rho_b derivative
  // record, fb[2r+1] = phi_ab derivative record (two 16-byte loads per pair).
  // The force pass reads derivatives only from fb and rho.dp, both stored
  // times -2 (the pair force is -2 dpsi/ds along the displacement).
  const double2 *fb;  // [nt*nt][n_s][2]
  // rho(s) for the passes that compute no energy: one 16-byte record per
  // segment holding the quadratic through the cubic at q = 0, 1/2 and 1 (FP64
  // base, FP32 linear and quadratic terms, evaluated like a derivative
  // record), NaN-flagged where the cubic is and where the quadratic misses
  // its own bound.
  const double2 *rho_lean;  // [nt][n_s]
  int32_t nt, n_s;
  double s0, inv_ds;
  // Segments that miss the accuracy bounds evaluate the FP64 r-spline instead:
  // those below s_head or from s_tail on (the reference spline's low-order end
  // slopes leave kinks a uniform s grid cannot follow) and those in up to
  // kMixedHoles interior holes [hole_lo, hole_hi) (a tabulated function that
  // is only C1 somewhere, e.g. where a potential term reaches its own cutoff
  // inside the table).  Unused holes are empty ([0, 0)).  The first and last
  // segments are never served.  The kernels do not compare s against these
  // bounds: every record of a segment outside them carries a NaN FP32 term,
  // so its evaluation is NaN (host-side bookkeeping only).
  double s_head, s_tail;
  double hole_lo[2], hole_hi[2];
};
constexpr int kMixedHoles = 2;

// derivative from one 16-byte record; qf = (float)q
__device__ __forceinline__ double mix_deriv_rec(double2 d, double q, float qf) {
  const float2 b = __builtin_bit_cast(float2, d.y);
  return fma(q, (double)fmaf(qf, b.y, b.x), d.x);
}

struct Masses {
  double inv_m[kMaxTypes];  // 1 / mass (amu^-1)
  double m[kMaxTypes];
};

__host__ __device__ inline int64_t dev_index(const Geom &g, int32_t s, int32_t xh, int32_t ye, int32_t ze) {
  return ((int64_t)(ze * (int64_t)g.ext_y + ye) * 2 + s) * g.exh + xh;
}

// Blocks are dispatched to the 8 XCDs round-robin.  Remap so that XCD k takes
// a contiguous range of logical blocks (contiguous z-slabs of the lattice):
// neighbour halos then stay in that XCD's L2.  gridDim.x is a multiple of 8.
__device__ inline int64_t xcd_block() {
  const int64_t per = gridDim.x / kXcds;
  return (int64_t)(blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
}

template <typename T>
__device__ inline T wave_sum(T v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

// Deterministic block sum (fixed tree): every thread of the block must call.
__device__ inline double block_sum(double v, double *lds4) {
  v = wave_sum(v);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) lds4[w] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) r = (lds4[0] + lds4[1]) + (lds4[2] + lds4[3]);
  __syncthreads();
  return r;
}

// One thread per local site, in device order (z, y, parity, x): a wave takes
// 64 consecutive local sites -- one half-row of a wide box, several short
// rows of a small one -- so no lane idles on rows shorter than a wave.  Used
// by the streaming passes (integration, standalone df, host record pack).

struct Site {
  int64_t i;    // device index
  bool active;  // inside the local box
};

// lb = logical block (the single-domain kernels use xcd_block(); the batched
// serving kernels take it from their work list, so a replica's sites map to
// blocks identically in both)
// local site q (device order) -> its device index; false past the last
__device__ __forceinline__ bool map_site_q(const Geom &g, int64_t q, Site &st) {
  if (q >= g.n_local) return false;
  const uint32_t u = (uint32_t)q, bxh = (uint32_t)g.bxh;
  const uint32_t r = u / bxh, xh = u - r * bxh;  // r = (z * by + y) * 2 + parity
  const uint32_t yz = r >> 1, zl = yz / (uint32_t)g.by, yl = yz - zl * (uint32_t)g.by;
  st.active = true;
  st.i = dev_index(g, (int32_t)(r & 1u), (int32_t)xh + g.gxh, (int32_t)yl + g.gy, (int32_t)zl + g.gz);
  return true;
}

__device__ __forceinline__ bool map_site_lb(const Geom &g, int64_t lb, Site &st) {
  return map_site_q(g, lb * kBlock + threadIdx.x, st);
}

__device__ __forceinline__ bool map_site(const Geom &g, Site &st) { return map_site_lb(g, xcd_block(), st); }

// Pair-pass tile lb: wave w of the block takes parity w & 1 and z-half w >> 1,
// its 64 lanes 8 cells along x x 4 rows along y x 2 planes along z, so the
// stencil row (parity) is wave-uniform.
struct TileSite {
  int32_t x0, y0, z0;  // tile origin, local cells / rows / planes
  int32_t tx, ty, tz;  // this thread's cell in the tile
  int s;               // parity (wave-uniform)
  bool active;         // inside the local box
  int64_t i;           // device index
  int32_t c;           // LDS record of this site
};

__device__ __forceinline__ TileSite map_tile(const Geom &g, const Halo &h, int64_t lb) {
  TileSite t;
  const int32_t txi = (int32_t)(lb % g.ntx);
  const int64_t r = lb / g.ntx;
  const int32_t tyi = (int32_t)(r % g.nty), tzi = (int32_t)(r / g.nty);
  t.x0 = txi * kTileX, t.y0 = tyi * kTileY, t.z0 = tzi * kTileZ;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & (kWave - 1);
  t.s = __builtin_amdgcn_readfirstlane(w & 1);
  t.tx = lane & 7;
  t.ty = (lane >> 3) & 3;
  t.tz = (w >> 1) * 2 + (lane >> 5);
  t.active = lb < g.n_tiles && t.x0 + t.tx < g.bxh && t.y0 + t.ty < g.by && t.z0 + t.tz < g.bz;
  t.i = dev_index(g, t.s, t.x0 + t.tx + g.gxh, t.y0 + t.ty + g.gy, t.z0 + t.tz + g.gz);
  t.c = (((t.tz + h.rz) * h.ly + (t.ty + h.ry)) * 2 + t.s) * h.lx + (t.tx + h.rx);
  return t;
}

// device index of the site thread tid takes in tile t (map_tile's t.i)
__device__ __forceinline__ int64_t tile_site_index(const Geom &g, const TileSite &t, uint32_t tid) {
  const uint32_t lane = tid & (kWave - 1), w = tid >> 6;
  return dev_index(g, t.s, t.x0 + (int32_t)(lane & 7) + g.gxh, t.y0 + (int32_t)((lane >> 3) & 3) + g.gy,
                   t.z0 + (int32_t)((w >> 1) * 2 + (lane >> 5)) + g.gz);
}

// LDS record idx -> device index of the ext site it mirrors, or -1 outside
// the ghost-extended lattice.
__device__ __forceinline__ int64_t tile_record_src(const Geom &g, const Halo &h, const TileSite &t, int32_t idx) {
  const int32_t lxh = idx % h.lx;
  const int32_t r = idx / h.lx;
  const int32_t s = r & 1;
  const int32_t yz = r >> 1;
  const int32_t ly = yz % h.ly, lz = yz / h.ly;
  const int32_t xh = t.x0 + lxh - h.rx + g.gxh, ye = t.y0 + ly - h.ry + g.gy, ze = t.z0 + lz - h.rz + g.gz;
  if (xh < 0 || xh >= g.exh || ye < 0 || ye >= g.ext_y || ze < 0 || ze >= g.ext_z) return -1;
  return dev_index(g, s, xh, ye, ze);
}

}  // namespace mmd

#endif
