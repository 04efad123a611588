// Interstitial (off-lattice atom) passes on the device.
//
// The reference plugin skips every off-lattice atom (src/kernels/kernel_itl.hpp:44-46,
// :52-54); MISA-MD core keeps them in a separate list.  Here the list is
// device resident and the passes run on top of the lattice passes:
//   * interstitial k sees the lattice sites of the wide stencil W around its
//     anchor (nearest lattice site of its position, wrapped into the periodic
//     box) and every other interstitial within rc (minimum image);
//   * lattice<->interstitial pairs are {(i,k): site(i) in W(anchor k),
//     type_i >= 0, r_ik < rc}; a ghost site stands for its periodic source.
// Every step the interstitials are binned by the cell of their anchor with an
// exact counting sort (counts, prefix sum, placement, then a rank by id inside
// each bin), so the bins hold any density, in a fixed order.  Two passes then
// evaluate every pair from each side, with no atomics on physical quantities:
//   * one wave per interstitial: its lattice sites of W (loads of a chunk of W
//     in flight together, in-cutoff sites compacted through LDS) and the
//     interstitials of the bins around its anchor;
//   * one wave per bin of lattice sites: each lane a site, every interstitial
//     anchored within W's reach tested against it in bin order, the site's
//     rho and df (or force) updated by that lane alone.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "ctx.h"
#include "eam_body.h"
#include "host_helpers.h"
#include "kernels.h"

namespace mmd {

// Interstitial bins: the local box's cells cut into nb[0] x nb[1] x nb[2]
// bins of at least bc[d] cells along each axis; an interstitial belongs to
// the bin of its anchor's cell.
struct ItlDev {
  const double *x, *y, *z;  // positions as stored (unwrapped)
  double *fx, *fy, *fz, *rho, *df;
  const int8_t *type;
  int32_t *ax, *ay, *az;  // anchor site, local lattice coordinates (x doubled), wrapped into the box
  double *px, *py, *pz;   // positions shifted by the same periodic image as the anchor
  int32_t *bin;           // [cap] bin of each interstitial
  int32_t *bin_count;     // [nbins] interstitials per bin, then the placement counters
  int32_t *bin_start;     // [nbins + 1] exclusive prefix sum of the counts
  int32_t *scan_tot;      // [scan blocks] per-block totals of that sum
  int32_t *tmp;           // [cap] placement by bin before the rank by id
  double4 *spos;          // [cap] (px, py, pz, type) in (bin, id) order
  int4 *sanc;             // [cap] (ax, ay, az, id) in (bin, id) order
  int32_t *err;           // bit 0: an interstitial position beyond the lattice's coordinate range
  const int32_t *wide;    // [n_wide][3] wide stencil W, parity 0
  int32_t n, n_wide, max_d2q;
  int32_t nb[3], nbins;   // bins per axis
  int32_t bc[3];          // fewest cells per bin along each axis
  int32_t rii[3];         // interstitial <-> interstitial search reach, cells
  int32_t rw[3];          // reach of W, cells
  int32_t ii_staged;      // anchors within twice W's radius include every interstitial within rc
  double Lx, Ly, Lz, a, rc2;
};
namespace {

constexpr double kFtm2v = 9.64853321e-3;
constexpr double kMvv2e = 1.0 / kFtm2v;

__device__ __forceinline__ int seg(double x, double x0, double inv_dx, int n, double &p) {
  double q = (x - x0) * inv_dx;
  q = fmax(q, 0.0);
  int m = min((int)q, n - 2);
  p = fmin(q - (double)m, 1.0);
  return m;
}

__device__ __forceinline__ int wrap_n(int v, int n) {
  v %= n;
  return v < 0 ? v + n : v;
}

// centred periodic difference s - a of two coordinates in [0, n): in [-n/2, n/2)
// (n even; for odd n in [-(n-1)/2, (n-1)/2])
__device__ __forceinline__ int cdiff0(int s, int a, int n) {
  int v = s - a;
  v = v >= n - n / 2 ? v - n : v;
  return v < -(n / 2) ? v + n : v;
}

__device__ __forceinline__ double minimg(double d, double L) { return d - L * rint(d / L); }

// Is lattice site (sx, sy, sz) (local, x doubled) in W around anchor (aX, aY,
// aZ) and closer than rc to the interstitial at (mx, my, mz) (its position
// shifted by the anchor's periodic image)?  The site is read through its image
// seen from the anchor (a ghost record near the box faces), so the
// displacement site - interstitial is computed exactly as from the
// interstitial's side.
__device__ __forceinline__ bool li_test(const Geom &g, const ItlDev &I, const double4 *__restrict__ P, int aX, int aY,
                                        int aZ, double mx, double my, double mz, int sx, int sy, int sz, double &d2,
                                        double &dx, double &dy, double &dz) {
  const int dX = cdiff0(sx, aX, 2 * g.bxh), dY = cdiff0(sy, aY, g.by), dZ = cdiff0(sz, aZ, g.bz);
  const int sa = aX & 1, s1 = sx & 1;
  const int hy = 2 * dY + s1 - sa, hz = 2 * dZ + s1 - sa;
  if (dX * dX + hy * hy + hz * hz > I.max_d2q) return false;
  const int xe = aX + dX + 2 * g.gxh, ye = aY + dY + g.gy, ze = aZ + dZ + g.gz;
  const double4 q = P[dev_index(g, xe & 1, xe >> 1, ye, ze)];
  dx = q.x - mx, dy = q.y - my, dz = q.z - mz;
  d2 = dx * dx + dy * dy + dz * dz;
  return d2 < I.rc2;
}
// Nearest BCC site of a position (global lattice coordinates, x doubled): the
// closer of the nearest corner-sub-lattice and body-sub-lattice sites, a tie
// going to the corner, i.e. the site whose Wigner-Seitz cell holds the point.
// The squared distances are written with fma() so that no contraction choice
// of the compiler can change the decision (the oracle uses the same
// expression, oracle/eam_oracle.c orc_anchor).
__device__ __forceinline__ void ws_site(double a, double px, double py, double pz, long long &X, long long &Y,
                                        long long &Z) {
  const double ux = px / a, uy = py / a, uz = pz / a;
  const double cx = floor(ux + 0.5), cy = floor(uy + 0.5), cz = floor(uz + 0.5);
  const double fx = floor(ux), fy = floor(uy), fz = floor(uz);
  const double ex = ux - cx, ey = uy - cy, ez = uz - cz;
  const double gx = ux - (fx + 0.5), gy = uy - (fy + 0.5), gz = uz - (fz + 0.5);
  const double dc = fma(ex, ex, fma(ey, ey, ez * ez));
  const double db = fma(gx, gx, fma(gy, gy, gz * gz));
  if (dc <= db) {
    X = 2 * (long long)cx, Y = (long long)cy, Z = (long long)cz;
  } else {
    X = 2 * (long long)fx + 1, Y = (long long)fy, Z = (long long)fz;
  }
}

// floor division image of v in [0, n): v - q n, q returned
__device__ __forceinline__ long long wrap_img64(long long &v, long long n) {
  const long long q = v / n - (v % n < 0 ? 1 : 0);
  v -= q * n;
  return q;
}

// exclusive prefix sum of v over the block (kBlock threads); total in tot
__device__ __forceinline__ int32_t block_excl_scan(int32_t v, int32_t &tot, int32_t *lds) {
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  int32_t incl = v;
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    const int32_t t = __shfl_up(incl, d, kWave);
    if (lane >= d) incl += t;
  }
  if (lane == kWave - 1) lds[w] = incl;
  __syncthreads();
  int32_t off = 0, all = 0;
#pragma unroll
  for (int k = 0; k < kWavesPerBlock; k++) {
    const int32_t c = lds[k];
    off += k < w ? c : 0;
    all += c;
  }
  __syncthreads();
  tot = all;
  return off + incl - v;
}

// Positions this far from the origin (in lattice constants) are refused: the
// anchor's integer coordinates must stay exact.
constexpr double kItlMaxCoord = 1e9;

// Bin of cell c along an axis of n cells cut into nb bins of floor(n / nb) or
// ceil(n / nb) cells (no short last bin).
__host__ __device__ __forceinline__ int bin_of_cell(int c, int n, int nb) { return (int)((int64_t)c * nb / n); }

// anchor (nearest lattice site) of each interstitial, wrapped into the box, and its bin's count
__global__ __launch_bounds__(kBlock) void k_itl_anchor(Geom g, ItlDev I) {
  const int k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= I.n) return;
  double px = I.x[k], py = I.y[k], pz = I.z[k];
  long long X = 0, Y = 0, Z = 0;
  const double lim = kItlMaxCoord * I.a;
  if (fabs(px) < lim && fabs(py) < lim && fabs(pz) < lim) {  // false for NaN too
    ws_site(I.a, px, py, pz, X, Y, Z);
  } else {
    atomicOr(I.err, 1);
    px = py = pz = 0.0;
  }
  px = fma(-(double)wrap_img64(X, 2 * (long long)g.bxh), I.Lx, px);
  py = fma(-(double)wrap_img64(Y, (long long)g.by), I.Ly, py);
  pz = fma(-(double)wrap_img64(Z, (long long)g.bz), I.Lz, pz);
  I.ax[k] = (int)X, I.ay[k] = (int)Y, I.az[k] = (int)Z;
  I.px[k] = px, I.py[k] = py, I.pz[k] = pz;
  const int b = (bin_of_cell((int)Z, g.bz, I.nb[2]) * I.nb[1] + bin_of_cell((int)Y, g.by, I.nb[1])) * I.nb[0] +
                bin_of_cell((int)X >> 1, g.bxh, I.nb[0]);
  I.bin[k] = b;
  atomicAdd(&I.bin_count[b], 1);
}

constexpr int kScanPer = 8;                     // bin counts per thread in the prefix sum
constexpr int kScanChunk = kScanPer * kBlock;   // bins per block

// per block: exclusive prefix sum of its bins' counts, block total to scan_tot
__global__ __launch_bounds__(kBlock) void k_bin_scan_local(ItlDev I) {
  __shared__ int32_t lds[kWavesPerBlock];
  const int64_t base = (int64_t)blockIdx.x * kScanChunk + (int64_t)threadIdx.x * kScanPer;
  int32_t v[kScanPer], s = 0;
#pragma unroll
  for (int u = 0; u < kScanPer; u++) {
    v[u] = base + u < I.nbins ? I.bin_count[base + u] : 0;
    s += v[u];
  }
  int32_t tot;
  int32_t run = block_excl_scan(s, tot, lds);
#pragma unroll
  for (int u = 0; u < kScanPer; u++) {
    if (base + u < I.nbins) I.bin_start[base + u] = run;
    run += v[u];
  }
  if (threadIdx.x == 0) I.scan_tot[blockIdx.x] = tot;
}

// one block: exclusive prefix sum of the block totals; the grand total closes bin_start
__global__ __launch_bounds__(kBlock) void k_bin_scan_tot(ItlDev I, int32_t nblk) {
  __shared__ int32_t lds[kWavesPerBlock];
  int32_t carry = 0;
  for (int32_t c0 = 0; c0 < nblk; c0 += kBlock) {
    const int32_t i = c0 + threadIdx.x;
    const int32_t x = i < nblk ? I.scan_tot[i] : 0;
    int32_t tot;
    const int32_t ex = block_excl_scan(x, tot, lds);
    if (i < nblk) I.scan_tot[i] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0) I.bin_start[I.nbins] = carry;
}

// adds each block's offset; the counts restart at zero as placement counters
__global__ __launch_bounds__(kBlock) void k_bin_scan_add(ItlDev I) {
  const int64_t base = (int64_t)blockIdx.x * kScanChunk + (int64_t)threadIdx.x * kScanPer;
  const int32_t off = I.scan_tot[blockIdx.x];
#pragma unroll
  for (int u = 0; u < kScanPer; u++)
    if (base + u < I.nbins) {
      I.bin_start[base + u] += off;
      I.bin_count[base + u] = 0;
    }
}

// placement in the bin's range (in whatever order the atomics give)
__global__ __launch_bounds__(kBlock) void k_itl_place(ItlDev I) {
  const int k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= I.n) return;
  const int b = I.bin[k];
  I.tmp[I.bin_start[b] + atomicAdd(&I.bin_count[b], 1)] = k;
}

// final position = bin start + the number of lower ids in the bin: the (bin,
// id) order, whatever the placement's order was
__global__ __launch_bounds__(kBlock) void k_itl_rank(ItlDev I) {
  const int k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= I.n) return;
  const int b = I.bin[k];
  const int32_t s0 = I.bin_start[b], s1 = I.bin_start[b + 1];
  int32_t r = 0;
  for (int32_t j = s0; j < s1; j++) r += I.tmp[j] < k ? 1 : 0;
  I.spos[s0 + r] = make_double4(I.px[k], I.py[k], I.pz[k], (double)I.type[k]);
  I.sanc[s0 + r] = make_int4(I.ax[k], I.ay[k], I.az[k], k);
  I.bin_count[b] = 0;  // the next binning starts from zero counts (every interstitial of the bin stores the same 0)
}

// Bins along one periodic axis of n cells (nb bins, bin_of_cell) that hold
// cells lo..hi: the first bin and how many follow it cyclically, each bin at
// most once.
__host__ __device__ __forceinline__ void bin_span(int lo, int hi, int n, int nb, int &b0, int &cnt) {
  if (hi - lo + 1 >= n) {
    b0 = 0, cnt = nb;
    return;
  }
  const int wl = ((lo % n) + n) % n, wh = ((hi % n) + n) % n;
  const int bl = bin_of_cell(wl, n, nb), bh = bin_of_cell(wh, n, nb);
  b0 = bl;
  if (wl <= wh)
    cnt = bh - bl + 1;
  else
    cnt = bh >= bl ? nb : nb - bl + bh + 1;
}

__device__ __forceinline__ int bin_next(int b, int nb) { return b + 1 == nb ? 0 : b + 1; }

// exclusive prefix sum of v over the wave; total in tot
__device__ __forceinline__ int32_t wave_excl_scan(int32_t v, int32_t &tot) {
  const int lane = threadIdx.x & (kWave - 1);
  int32_t incl = v;
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    const int32_t t = __shfl_up(incl, d, kWave);
    if (lane >= d) incl += t;
  }
  tot = __shfl(incl, kWave - 1, kWave);
  return incl - v;
}

// The candidate interstitials of one search: the bins of the spans b0 / nb
// (at most 4 per axis, so at most 64: the set-up checks it), flattened in bin
// order.  Lane l < nbt holds bin l's start, count and exclusive prefix.
struct Cands {
  int32_t start, cnt, pre;  // this lane's bin
  int32_t nbt, total;
};

__device__ __forceinline__ void cands_setup(const ItlDev &I, const int b0[3], const int nb[3], Cands &C) {
  const int lane = threadIdx.x & (kWave - 1);
  C.nbt = nb[0] * nb[1] * nb[2];
  C.start = C.cnt = 0;
  if (lane < C.nbt) {
    int bx = b0[0] + lane % nb[0], by = b0[1] + (lane / nb[0]) % nb[1], bz = b0[2] + lane / (nb[0] * nb[1]);
    bx -= bx >= I.nb[0] ? I.nb[0] : 0;
    by -= by >= I.nb[1] ? I.nb[1] : 0;
    bz -= bz >= I.nb[2] ? I.nb[2] : 0;
    const int b = (bz * I.nb[1] + by) * I.nb[0] + bx;
    C.start = I.bin_start[b];
    C.cnt = I.bin_start[b + 1] - C.start;
  }
  int32_t tot;
  C.pre = wave_excl_scan(C.cnt, tot);
  C.total = tot;
}

// index in the sorted arrays of flattened candidate t (every lane calls; the
// result is meaningful for t < total)
__device__ __forceinline__ int32_t cand_index(const Cands &C, int32_t t) {
  int32_t j = 0;
  for (int l = 0; l < C.nbt; l++) {
    const int32_t p = __shfl(C.pre, l, kWave), c = __shfl(C.cnt, l, kWave), s = __shfl(C.start, l, kWave);
    j = t >= p && t < p + c ? s + (t - p) : j;
  }
  return j;
}

// ext device index of wide-stencil entry w around anchor (aX, aY, aZ)
__device__ __forceinline__ int64_t wide_site(const Geom &g, const ItlDev &I, int aX, int aY, int aZ, int sgn, int w) {
  const int xe = aX + sgn * I.wide[3 * w] + 2 * g.gxh, ye = aY + sgn * I.wide[3 * w + 1] + g.gy,
            ze = aZ + sgn * I.wide[3 * w + 2] + g.gz;
  return dev_index(g, xe & 1, xe >> 1, ye, ze);
}

// displacement site - interstitial and its square (one expression for both phases)
__device__ __forceinline__ double itl_d2(const double4 &q, double px, double py, double pz, double &dx, double &dy,
                                         double &dz) {
  dx = q.x - px, dy = q.y - py, dz = q.z - pz;
  return dx * dx + dy * dy + dz * dz;
}

// Radial derivatives of a pair (a, b) at segment m, fraction p: rho_b' and
// phi_ab' from the combined force record (three 16-byte loads, the same
// coefficients as the per-function tables), rho_a' read only when a != b.
__device__ __forceinline__ void itl_derivs(const TabF64 &tf, int a, int b, int m, double p, double &rho_b,
                                           double &phi_ab, double &rho_a) {
  const uint32_t rp = (uint32_t)((a * tf.nt + b) * tf.n_r + m);
  const double2 c0 = ld(tf.fc, 3 * rp), c1 = ld(tf.fc, 3 * rp + 1), c2 = ld(tf.fc, 3 * rp + 2);
  rho_b = (c0.x * p + c0.y) * p + c1.x;
  phi_ab = (c1.y * p + c2.x) * p + c2.y;
  rho_a = a == b ? rho_b : spl_deriv(tf.rho, (uint32_t)(a * tf.n_r + m), p);
}

// Pair lookups at squared distance s (segment m, fraction p of r = sqrt(s) in
// the FP64 tables).  PREC 1 reads the r^2-keyed mixed tables and falls back
// to the FP64 spline where they do not serve s (a NaN result), as the
// lattice passes do; the embedding F(rho) stays FP64 in both.
template <int PREC>
__device__ __forceinline__ double itl_rho(const TabF64 &tf, const TabMixed &tmx, int t, double s, int m, double p) {
  if (PREC == 1) {
    const SegS sg = seg_s(s, tmx);
    const uint32_t rec = __umul24((uint32_t)t, (uint32_t)tmx.n_s) + sg.m;
    const double2 b = ld(tmx.rho.vb, rec);
    const float2 c = ld(tmx.rho.vc, rec);
    const double v = fma(sg.q, fma(sg.q, (double)fmaf((float)sg.q, c.y, c.x), b.y), b.x);
    if (!__builtin_isnan(v)) return v;
  }
  return spl_value(tf.rho, (uint32_t)(t * tf.n_r + m), p);
}

template <int PREC>
__device__ __forceinline__ double itl_phi(const TabF64 &tf, const TabMixed &tmx, int a, int b, double s, int m,
                                          double p) {
  if (PREC == 1) {
    const SegS sg = seg_s(s, tmx);
    const uint32_t rec = __umul24((uint32_t)(a * tmx.nt + b), (uint32_t)tmx.n_s) + sg.m;
    const double2 vb = ld(tmx.phi.vb, rec);
    const float2 vc = ld(tmx.phi.vc, rec);
    const double v = fma(sg.q, fma(sg.q, (double)fmaf((float)sg.q, vc.y, vc.x), vb.y), vb.x);
    if (!__builtin_isnan(v)) return v;
  }
  return spl_value(tf.phi, (uint32_t)((a * tf.nt + b) * tf.n_r + m), p);
}

// -(dpsi/dr)/r of pair (a, b) = -(df_a rho_b' + df_b rho_a' + phi_ab') / r,
// the scalar along the displacement from b to a
template <int PREC>
__device__ __forceinline__ double itl_fpair(const TabF64 &tf, const TabMixed &tmx, int a, int b, double s, double r,
                                            int m, double p, double dfa, double dfb) {
  if (PREC == 1) {  // records hold -2 d/ds = -(d/dr)/r
    const SegS sg = seg_s(s, tmx);
    const uint32_t rp = __umul24((uint32_t)(a * tmx.nt + b), (uint32_t)tmx.n_s) + sg.m;
    const float qf = (float)sg.q;
    const double rho_b = mix_deriv_rec(ld(tmx.fb, 2 * rp), sg.q, qf);
    const double phi_ab = mix_deriv_rec(ld(tmx.fb, 2 * rp + 1), sg.q, qf);
    const double rho_a =
        a == b ? rho_b : mix_deriv_rec(ld(tmx.rho.dp, __umul24((uint32_t)a, (uint32_t)tmx.n_s) + sg.m), sg.q, qf);
    const double f = fma(dfa, rho_b, fma(dfb, rho_a, phi_ab));
    if (!__builtin_isnan(f)) return f;
  }
  double rho_b, phi_ab, rho_a;
  itl_derivs(tf, a, b, m, p, rho_b, phi_ab, rho_a);
  const double psip = dfa * rho_b + dfb * rho_a + phi_ab;
  return -psip / r;
}

// LDS written by some lanes of a wave and read by others: order the wave's
// LDS accesses (all lanes of the wave run the same path here)
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

constexpr int kItlPer = 4;                 // wide-stencil entries per lane in one chunk
constexpr int kItlChunk = kItlPer * kWave;  // entries tested per chunk

// local lattice coordinates (x doubled) of the periodic source of ext site (xe, ye, ze)
__device__ __forceinline__ void src_local(const Geom &g, int xe, int ye, int ze, int &sx, int &sy, int &sz) {
  sx = wrap_n(xe - 2 * g.gxh, 2 * g.bxh);
  sy = wrap_n(ye - g.gy, g.by);
  sz = wrap_n(ze - g.gz, g.bz);
}

// One wave per interstitial k.
//   * Its lattice sites of W: phase 1 loads a chunk of W's records together
//     and compacts the in-cutoff sites through LDS; then a round of those
//     pairs takes one per lane: k's own term (MODE 0: rho_k; MODE 1: the
//     force on k), and the site's side.  The site's side is computed by its
//     owner, the lowest-numbered interstitial interacting with it: the lane
//     scans the interstitials anchored within twice W's radius of k's anchor
//     and within 2 rc of k (every one that can share a site with k), from the
//     bins around it in bin order, staged in LDS (all at once when at most 64
//     remain, else 64 at a time), stops as soon as one with a lower number
//     interacts with the site, and otherwise adds every interacting one's
//     term in that scan order (MODE 0: to the site's lattice rho, then a new
//     df, the energy partial taking F(rho) - F(lattice rho); MODE 1: to the
//     site's lattice force, one full pair energy per pair).
//   * The interstitials within rc (minimum image), anchored in the bins
//     around k's anchor, a lane per candidate (MODE 1: half the pair energy
//     per side).
template <int MODE, bool WANT_E, int PREC>
__global__ __launch_bounds__(kBlock) void k_itl_side(Geom g, ItlDev I, const double4 *__restrict__ P,
                                                       double *__restrict__ RHO, double *__restrict__ DF,
                                                       double *__restrict__ FX, double *__restrict__ FY,
                                                       double *__restrict__ FZ, TabF64 tf, TabMixed tmx,
                                                       double *__restrict__ e_part) {
  __shared__ double lds[kWavesPerBlock];
  // in-cutoff entries of the current chunk, per wave: wide-stencil index | site type << 24, and the site's position
  __shared__ int32_t wlist[kWavesPerBlock][kItlChunk];
  __shared__ double3 wpos[kWavesPerBlock][kItlChunk];
  // staged candidates of the sites' side, per wave
  __shared__ double4 cpos[kWavesPerBlock][kWave];
  __shared__ int4 canc[kWavesPerBlock][kWave];
  __shared__ double cdf[kWavesPerBlock][kWave];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int k = blockIdx.x * kWavesPerBlock + wv;
  double acc = 0.0, fxk = 0.0, fyk = 0.0, fzk = 0.0, e_loc = 0.0;
  if (k < I.n) {
    const int tk = I.type[k];
    const int aX = I.ax[k], aY = I.ay[k], aZ = I.az[k];
    const int sgn = (aX & 1) ? -1 : 1;
    const double dfk = MODE == 1 ? I.df[k] : 0.0;
    const double pxk = I.px[k], pyk = I.py[k], pzk = I.pz[k];
    int32_t *cw = wlist[wv];
    double3 *cq = wpos[wv];
    double4 *sp = cpos[wv];
    int4 *sa = canc[wv];
    double *sd = cdf[wv];
    const int ac[3] = {aX >> 1, aY, aZ}, nc[3] = {g.bxh, g.by, g.bz};
    // candidates of the sites' side: anchored within twice W's reach of k's anchor
    Cands C2;
    {
      int b0[3], nb[3];
#pragma unroll
      for (int d = 0; d < 3; d++)
        bin_span(ac[d] - 2 * I.rw[d], ac[d] + 2 * I.rw[d], nc[d], I.nb[d], b0[d], nb[d]);
      cands_setup(I, b0, nb, C2);
    }
    // A candidate is kept, in its (bin, id) order, when it can interact with
    // one of k's in-cutoff sites: anchored within twice W's radius of k's
    // anchor, and within 2 rc of k's position (a site within rc of both k
    // and m puts m within 2 rc of k; the margin covers rounding).  Dropping
    // the others changes neither a site's owner nor its summation order.
    const double lim = 2.0 * sqrt(I.rc2) + 1e-6, lim2 = lim * lim;
    // stages the kept candidates of flattened range [t0, t0 + 64) from LDS
    // slot base on (slots past 63 are not written); returns how many it kept
    auto stage_chunk = [&](int32_t t0, int base) {
      const int32_t t = t0 + lane;
      const int32_t j = cand_index(C2, t);
      bool keep = false;
      int4 am = make_int4(0, 0, 0, 0);
      double4 q = make_double4(0.0, 0.0, 0.0, 0.0);
      if (t < C2.total) {
        am = I.sanc[j];
        const int dX = cdiff0(am.x, aX, 2 * g.bxh), dY = cdiff0(am.y, aY, g.by), dZ = cdiff0(am.z, aZ, g.bz);
        const int sd2 = (am.x & 1) - (aX & 1), hy = 2 * dY + sd2, hz = 2 * dZ + sd2;
        if (dX * dX + hy * hy + hz * hz <= 4 * I.max_d2q) {
          q = I.spos[j];
          const double ux = minimg(q.x - pxk, I.Lx), uy = minimg(q.y - pyk, I.Ly), uz = minimg(q.z - pzk, I.Lz);
          keep = ux * ux + uy * uy + uz * uz <= lim2;
        }
      }
      const uint64_t mk = __ballot(keep);
      const int slot = base + __popcll(mk & ((1ull << lane) - 1));
      if (keep && slot < kWave) {
        sp[slot] = q;
        sa[slot] = am;
        if (MODE == 1) sd[slot] = I.df[am.w];
      }
      return (int)__popcll(mk);
    };
    // every kept candidate in one staging when they fit (one_chunk) ...
    int n_kept = 0;
    for (int32_t t0 = 0; t0 < C2.total && n_kept <= kWave; t0 += kWave) n_kept += stage_chunk(t0, n_kept);
    wave_lds_sync();
    const bool one_chunk = n_kept <= kWave;
    const int n_one = one_chunk ? n_kept : 0;
    // ... else chunk by chunk of the flattened list, restaged from slot 0
    auto stage = [&](int32_t t0) {
      wave_lds_sync();  // the previous staging is no longer read
      const int n = stage_chunk(t0, 0);
      wave_lds_sync();
      return n;
    };
    const int32_t scan_total = one_chunk ? 1 : C2.total;
    for (int w0 = 0; w0 < I.n_wide; w0 += kItlChunk) {
      // phase 1: every lane tests its entries of the chunk, the record loads
      // (position and type) of all of them in flight together; the entries
      // inside the cutoff are compacted in stencil order with their records
      int32_t found = 0;
      {
        double4 qq[kItlPer];
#pragma unroll
        for (int j = 0; j < kItlPer; j++) {
          const int w = w0 + j * kWave + lane;
          qq[j] = P[w < I.n_wide ? wide_site(g, I, aX, aY, aZ, sgn, w) : 0];
        }
#pragma unroll
        for (int j = 0; j < kItlPer; j++) {
          const int w = w0 + j * kWave + lane;
          const int tj = w < I.n_wide ? (int)qq[j].w : -1;
          double dx, dy, dz;
          const bool in = tj >= 0 && itl_d2(qq[j], pxk, pyk, pzk, dx, dy, dz) < I.rc2;
          const uint64_t m = __ballot(in);
          if (in) {
            const int slot = found + __popcll(m & ((1ull << lane) - 1));
            cw[slot] = w | (tj << 24);
            cq[slot] = make_double3(qq[j].x, qq[j].y, qq[j].z);
          }
          found += __popcll(m);
        }
      }
      wave_lds_sync();
      // phase 2: rounds of one in-cutoff pair per lane
      for (int c0 = 0; c0 < found; c0 += kWave) {
        const int c = c0 + lane;
        bool own = false;
        int ti = 0, sx = 0, sy = 0, sz = 0;
        int64_t src = 0;
        if (c < found) {
          const int32_t wt = cw[c];
          const int w = wt & 0xffffff;
          ti = wt >> 24;
          const double3 qp = cq[c];
          double dx, dy, dz;
          const double d2 = itl_d2(make_double4(qp.x, qp.y, qp.z, 0.0), pxk, pyk, pzk, dx, dy, dz);
          const int xe = aX + sgn * I.wide[3 * w] + 2 * g.gxh, ye = aY + sgn * I.wide[3 * w + 1] + g.gy,
                    ze = aZ + sgn * I.wide[3 * w + 2] + g.gz;
          const int64_t e = dev_index(g, xe & 1, xe >> 1, ye, ze);
          const double r = sqrt(d2);
          double p;
          const int m = seg(r, tf.r0, tf.inv_dr, tf.n_r, p);
          if (MODE == 0) {
            acc += itl_rho<PREC>(tf, tmx, ti, d2, m, p);
          } else {
            // the pair force along del = x_site - x_k, with the site's df read at its ext record
            const double fp = itl_fpair<PREC>(tf, tmx, ti, tk, d2, r, m, p, DF[e], dfk);
            fxk -= dx * fp, fyk -= dy * fp, fzk -= dz * fp;
          }
          src_local(g, xe, ye, ze, sx, sy, sz);
          src = dev_index(g, sx & 1, (sx >> 1) + g.gxh, sy + g.gy, sz + g.gz);
          own = true;
        }
        double rho_i = 0.0, dfi = 0.0, gx = 0.0, gy = 0.0, gz = 0.0, ep = 0.0;
        if (own) {
          if (MODE == 0) rho_i = RHO[src];
          if (MODE == 1) dfi = DF[src];
        }
        const double rho_lat = rho_i;
        for (int32_t t0 = 0; t0 < scan_total; t0 += kWave) {
          const int n_chunk = one_chunk ? n_one : stage(t0);
          for (int cc = 0; own && cc < n_chunk; cc++) {
            const int4 am = sa[cc];
            const double4 pm = sp[cc];
            double d2b, ex, ey, ez;
            if (!li_test(g, I, P, am.x, am.y, am.z, pm.x, pm.y, pm.z, sx, sy, sz, d2b, ex, ey, ez)) continue;
            if (am.w < k) {
              own = false;  // a lower-numbered interstitial interacts with this site: it owns the site
              break;
            }
            const int tm = (int)pm.w;
            const double rb = sqrt(d2b);
            double pb;
            const int mb = seg(rb, tf.r0, tf.inv_dr, tf.n_r, pb);
            if (MODE == 0) {
              rho_i += itl_rho<PREC>(tf, tmx, tm, d2b, mb, pb);
            } else {
              const double fp = itl_fpair<PREC>(tf, tmx, ti, tm, d2b, rb, mb, pb, dfi, sd[cc]);
              gx += ex * fp, gy += ey * fp, gz += ez * fp;
              if (WANT_E) ep += itl_phi<PREC>(tf, tmx, ti, tm, d2b, mb, pb);
            }
          }
        }
        if (own) {
          if (MODE == 0) {
            double pe;
            const int me = seg(rho_i, tf.rho0, tf.inv_drho, tf.n_rho, pe);
            if (WANT_E) {
              double po;
              const int mo = seg(rho_lat, tf.rho0, tf.inv_drho, tf.n_rho, po);
              e_loc += spl_value(tf.emb, ti * tf.n_rho + me, pe) - spl_value(tf.emb, ti * tf.n_rho + mo, po);
            }
            RHO[src] = rho_i;
            DF[src] = spl_deriv(tf.emb, ti * tf.n_rho + me, pe);
          } else {
            FX[src] += gx, FY[src] += gy, FZ[src] += gz;
            if (WANT_E) e_loc += ep;
          }
        }
      }
      wave_lds_sync();  // the list is rewritten by the next chunk
    }
    // interstitials within rc (minimum image), a lane per candidate: the
    // staged candidates of the sites' side when they fit one staging and hold
    // every interstitial close enough (ii_staged), else those anchored in the
    // bins around k's anchor cell
    auto ii_pair = [&](const double4 &q, int mm, double dfm) {
      if (mm == k) return;
      const double dx = minimg(pxk - q.x, I.Lx), dy = minimg(pyk - q.y, I.Ly), dz = minimg(pzk - q.z, I.Lz);
      const double d2 = dx * dx + dy * dy + dz * dz;
      if (d2 >= I.rc2) return;
      const int tm = (int)q.w;
      const double r = sqrt(d2);
      double p;
      const int m = seg(r, tf.r0, tf.inv_dr, tf.n_r, p);
      if (MODE == 0) {
        acc += itl_rho<PREC>(tf, tmx, tm, d2, m, p);
      } else {
        const double fp = itl_fpair<PREC>(tf, tmx, tk, tm, d2, r, m, p, dfk, dfm);
        fxk += dx * fp, fyk += dy * fp, fzk += dz * fp;
        if (WANT_E) e_loc += 0.5 * itl_phi<PREC>(tf, tmx, tk, tm, d2, m, p);
      }
    };
    if (one_chunk && I.ii_staged) {
      if (lane < n_one) ii_pair(sp[lane], sa[lane].w, MODE == 1 ? sd[lane] : 0.0);
    } else {
      int b0[3], nb[3];
#pragma unroll
      for (int d = 0; d < 3; d++) bin_span(ac[d] - I.rii[d], ac[d] + I.rii[d], nc[d], I.nb[d], b0[d], nb[d]);
      Cands C;
      cands_setup(I, b0, nb, C);
      for (int32_t t0 = 0; t0 < C.total; t0 += kWave) {
        const int32_t t = t0 + lane;
        const int32_t j = cand_index(C, t);
        if (t >= C.total) continue;
        const int mm = I.sanc[j].w;
        ii_pair(I.spos[j], mm, MODE == 1 ? I.df[mm] : 0.0);
      }
    }
    if (MODE == 0) {
      acc = wave_sum(acc);
      if (lane == 0) {
        double pe;
        const int me = seg(acc, tf.rho0, tf.inv_drho, tf.n_rho, pe);
        I.rho[k] = acc;
        I.df[k] = spl_deriv(tf.emb, tk * tf.n_rho + me, pe);
        if (WANT_E) e_loc += spl_value(tf.emb, tk * tf.n_rho + me, pe);
      }
    } else {
      fxk = wave_sum(fxk), fyk = wave_sum(fyk), fzk = wave_sum(fzk);
      if (lane == 0) I.fx[k] = fxk, I.fy[k] = fyk, I.fz[k] = fzk;
    }
  }
  if (WANT_E) {
    const double b = block_sum(e_loc, lds);
    if (threadIdx.x == 0) e_part[blockIdx.x] = b;
  }
}

__global__ __launch_bounds__(kBlock) void k_itl_df(ItlDev I, TabF64 tf) {
  const int k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= I.n) return;
  double p;
  const int m = seg(I.rho[k], tf.rho0, tf.inv_drho, tf.n_rho, p);
  I.df[k] = spl_deriv(tf.emb, I.type[k] * tf.n_rho + m, p);
}

template <bool DRIFT, bool WANT_KE>
__global__ __launch_bounds__(kBlock) void k_itl_kick(ItlDev I, double *vx, double *vy, double *vz, double *x,
                                                       double *y, double *z, Masses M, double h, double dt,
                                                       double *ke_part) {
  __shared__ double lds[kWavesPerBlock];
  const int k = blockIdx.x * kBlock + threadIdx.x;
  double ke = 0.0;
  if (k < I.n) {
    const int t = I.type[k];
    const double sc = h * kFtm2v * M.inv_m[t];
    const double a = vx[k] + sc * I.fx[k], b = vy[k] + sc * I.fy[k], c = vz[k] + sc * I.fz[k];
    vx[k] = a, vy[k] = b, vz[k] = c;
    if (DRIFT) x[k] += dt * a, y[k] += dt * b, z[k] += dt * c;
    if (WANT_KE) ke = 0.5 * M.m[t] * (a * a + b * b + c * c) * kMvv2e;
  }
  if (WANT_KE) {
    const double s = block_sum(ke, lds);
    if (threadIdx.x == 0) ke_part[blockIdx.x] = s;
  }
}

// ------------------------------------------------------------- migration
//
// Wigner-Seitz re-assignment between the lattice and the interstitial list
// (rules: a lattice atom whose nearest site is not its own joins the list, in
// increasing local site order, and leaves a vacancy; a list entry whose
// nearest site is vacant moves into it, the lowest index winning a contested
// site, and the list is compacted in order).  Every decision is integer work
// on positions read as stored, so the result is bit-identical to the oracle's
// sequential restatement.

struct MigDev {
  Geom g;
  double4 *P;
  int8_t *T;
  double *VX, *VY, *VZ, *FX, *FY, *FZ, *RHO, *DF;   // lattice (ext, device layout)
  double *x, *y, *z, *vx, *vy, *vz, *fx, *fy, *fz;  // interstitial list
  double *rho, *df;
  int8_t *type;
  int32_t *blk;    // per-block leaver counts, then exclusive offsets
  int32_t *claim;  // [n_local] lowest list index claiming a vacant site
  int32_t *cand;   // [cap] claimed local site of each entry (-1: none), then keep flags
  int32_t *cnt;    // [0] leavers, [1] entries kept
  int64_t n_local;
  int32_t n, n_all;
  double a, Lx, Ly, Lz;
};

// device index of local site li (reference order: z, y, doubled x)
__device__ __forceinline__ int64_t local_dev(const Geom &g, int64_t li, int32_t &xl, int32_t &yl, int32_t &zl) {
  const int64_t nx = 2 * (int64_t)g.bxh;
  xl = (int32_t)(li % nx);
  const int64_t r = li / nx;
  yl = (int32_t)(r % g.by);
  zl = (int32_t)(r / g.by);
  return dev_index(g, xl & 1, (xl >> 1) + g.gxh, yl + g.gy, zl + g.gz);
}

// Does the atom on local site li belong to another site's cell?
__device__ __forceinline__ bool mig_leaves(const MigDev &M, int64_t li, int64_t &e) {
  e = -1;
  if (li >= M.n_local) return false;
  int32_t xl, yl, zl;
  e = local_dev(M.g, li, xl, yl, zl);
  if (M.T[e] < 0) return false;
  const double4 p = M.P[e];
  long long X, Y, Z;
  ws_site(M.a, p.x, p.y, p.z, X, Y, Z);
  return X != xl || Y != yl || Z != zl;
}

__global__ __launch_bounds__(kBlock) void k_mig_count(MigDev M) {
  __shared__ int32_t lds[kWavesPerBlock];
  int64_t e;
  const bool f = mig_leaves(M, (int64_t)blockIdx.x * kBlock + threadIdx.x, e);
  int32_t tot;
  (void)block_excl_scan(f ? 1 : 0, tot, lds);
  if (threadIdx.x == 0) M.blk[blockIdx.x] = tot;
}

// one block: exclusive scan of v[0..n) in place, total into *total
__global__ __launch_bounds__(kBlock) void k_mig_scan(int32_t *v, int64_t n, int32_t *total) {
  __shared__ int32_t lds[kWavesPerBlock];
  int32_t carry = 0;
  for (int64_t c0 = 0; c0 < n; c0 += kBlock) {
    const int64_t i = c0 + threadIdx.x;
    const int32_t x = i < n ? v[i] : 0;
    int32_t tot;
    const int32_t ex = block_excl_scan(x, tot, lds);
    if (i < n) v[i] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0) *total = carry;
}

// leavers join the list at n + (their rank in site order); their sites empty
__global__ __launch_bounds__(kBlock) void k_mig_append(MigDev M) {
  __shared__ int32_t lds[kWavesPerBlock];
  int64_t e;
  const bool f = mig_leaves(M, (int64_t)blockIdx.x * kBlock + threadIdx.x, e);
  int32_t tot;
  const int32_t r = block_excl_scan(f ? 1 : 0, tot, lds);
  if (!f) return;
  const int32_t k = M.n + M.blk[blockIdx.x] + r;
  double4 p = M.P[e];
  M.x[k] = p.x, M.y[k] = p.y, M.z[k] = p.z;
  M.vx[k] = M.VX[e], M.vy[k] = M.VY[e], M.vz[k] = M.VZ[e];
  M.fx[k] = M.FX[e], M.fy[k] = M.FY[e], M.fz[k] = M.FZ[e];
  M.rho[k] = M.RHO[e], M.df[k] = M.DF[e];
  M.type[k] = M.T[e];
  p.w = -1.0;
  M.P[e] = p;
  M.T[e] = -1;
  M.VX[e] = 0.0, M.VY[e] = 0.0, M.VZ[e] = 0.0;
  M.FX[e] = 0.0, M.FY[e] = 0.0, M.FZ[e] = 0.0;
  M.RHO[e] = 0.0, M.DF[e] = 0.0;
}

// nearest site of entry k wrapped into the box; returns the local index and
// the periodic image
__device__ __forceinline__ int64_t mig_target(const MigDev &M, int32_t k, long long &qx, long long &qy,
                                              long long &qz) {
  long long X, Y, Z;
  ws_site(M.a, M.x[k], M.y[k], M.z[k], X, Y, Z);
  qx = wrap_img64(X, 2 * (long long)M.g.bxh);
  qy = wrap_img64(Y, (long long)M.g.by);
  qz = wrap_img64(Z, (long long)M.g.bz);
  return (Z * M.g.by + Y) * (2 * (long long)M.g.bxh) + X;
}

__global__ __launch_bounds__(kBlock) void k_mig_claim(MigDev M) {
  const int32_t k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= M.n_all) return;
  long long qx, qy, qz;
  const int64_t li = mig_target(M, k, qx, qy, qz);
  int32_t xl, yl, zl;
  const int64_t e = local_dev(M.g, li, xl, yl, zl);
  if (M.T[e] < 0) {
    M.cand[k] = (int32_t)li;
    atomicMin(&M.claim[li], k);
  } else {
    M.cand[k] = -1;
  }
}

// winners move into their site; cand[k] becomes the keep flag
__global__ __launch_bounds__(kBlock) void k_mig_fill(MigDev M) {
  const int32_t k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= M.n_all) return;
  const int32_t li = M.cand[k];
  if (li < 0 || M.claim[li] != k) {
    M.cand[k] = 1;
    return;
  }
  long long qx, qy, qz;
  (void)mig_target(M, k, qx, qy, qz);
  int32_t xl, yl, zl;
  const int64_t e = local_dev(M.g, li, xl, yl, zl);
  const int8_t t = M.type[k];
  M.P[e] = make_double4(fma(-(double)qx, M.Lx, M.x[k]), fma(-(double)qy, M.Ly, M.y[k]),
                        fma(-(double)qz, M.Lz, M.z[k]), (double)t);
  M.T[e] = t;
  M.VX[e] = M.vx[k], M.VY[e] = M.vy[k], M.VZ[e] = M.vz[k];
  M.FX[e] = M.fx[k], M.FY[e] = M.fy[k], M.FZ[e] = M.fz[k];
  M.RHO[e] = M.rho[k], M.DF[e] = M.df[k];
  M.cand[k] = 0;
}

// one block: in-place, order-preserving compaction of the list by the keep
// flags.  Each chunk is staged through LDS, so all of its loads have landed
// before any store of the chunk (stores go to indices <= the chunk's own).
__global__ __launch_bounds__(kBlock) void k_mig_compact(MigDev M) {
  constexpr int kF = 11;
  __shared__ double vals[kF][kBlock];
  __shared__ int8_t typ[kBlock];
  __shared__ int32_t lds[kWavesPerBlock];
  double *const arr[kF] = {M.x, M.y, M.z, M.vx, M.vy, M.vz, M.fx, M.fy, M.fz, M.rho, M.df};
  int32_t base = 0;
  for (int32_t c0 = 0; c0 < M.n_all; c0 += kBlock) {
    const int32_t k = c0 + threadIdx.x;
    const int32_t keep = k < M.n_all ? M.cand[k] : 0;
    if (keep) {
#pragma unroll
      for (int f = 0; f < kF; f++) vals[f][threadIdx.x] = arr[f][k];
      typ[threadIdx.x] = M.type[k];
    }
    int32_t tot;
    const int32_t r = block_excl_scan(keep, tot, lds);  // its barriers order the loads above before the stores
    if (keep) {
      const int32_t d = base + r;
#pragma unroll
      for (int f = 0; f < kF; f++) arr[f][d] = vals[f][threadIdx.x];
      M.type[d] = typ[threadIdx.x];
    }
    base += tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) M.cnt[1] = base;
}

// closing half kick of a step + opening kick and drift of the next (the same
// operations, in the same order, as k_itl_kick<false, K> then k_itl_kick<true, false>)
template <bool WANT_KE>
__global__ __launch_bounds__(kBlock) void k_itl_kick_fused(ItlDev I, double *vx, double *vy, double *vz, double *x,
                                                             double *y, double *z, Masses M, double h, double dt,
                                                             double *ke_part) {
  __shared__ double lds[kWavesPerBlock];
  const int k = blockIdx.x * kBlock + threadIdx.x;
  double ke = 0.0;
  if (k < I.n) {
    const int t = I.type[k];
    const double sc = h * kFtm2v * M.inv_m[t];
    const double fx = I.fx[k], fy = I.fy[k], fz = I.fz[k];
    const double a = vx[k] + sc * fx, b = vy[k] + sc * fy, c = vz[k] + sc * fz;
    if (WANT_KE) ke = 0.5 * M.m[t] * (a * a + b * b + c * c) * kMvv2e;
    const double a2 = a + sc * fx, b2 = b + sc * fy, c2 = c + sc * fz;
    vx[k] = a2, vy[k] = b2, vz[k] = c2;
    x[k] += dt * a2, y[k] += dt * b2, z[k] += dt * c2;
  }
  if (WANT_KE) {
    const double s = block_sum(ke, lds);
    if (threadIdx.x == 0) ke_part[blockIdx.x] = s;
  }
}

}  // namespace

// ------------------------------------------------------------------- host

#define ITL_TRY(expr)                                                                                       \
  do {                                                                                                      \
    hipError_t e_ = (expr);                                                                                 \
    if (e_ != hipSuccess)                                                                                   \
      return fail(MMD_E_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); \
  } while (0)

// the bin grid of a box: cells per bin, bins, search reaches (cells per axis)
struct BinPlan {
  int32_t nb[3], bc[3], rii[3], rw[3];
  int32_t nbins;
  bool operator==(const BinPlan &o) const {
    for (int d = 0; d < 3; d++)
      if (nb[d] != o.nb[d] || bc[d] != o.bc[d] || rii[d] != o.rii[d] || rw[d] != o.rw[d]) return false;
    return nbins == o.nbins;
  }
};

struct ItlState {
  int32_t *ax = nullptr, *ay = nullptr, *az = nullptr, *bin = nullptr, *tmp = nullptr;
  double *px = nullptr, *py = nullptr, *pz = nullptr;
  double4 *spos = nullptr;
  int4 *sanc = nullptr;
  int32_t *err = nullptr;
  int32_t max_d2q = 0;
  int32_t wreach[3] = {0, 0, 0};  // reach of W in cells per axis
  // bins, sized by ensure_bins for the box, lattice constant, cutoff and W
  int32_t *bin_count = nullptr, *bin_start = nullptr, *scan_tot = nullptr;
  BinPlan plan{};
  int32_t nscan = 0;
  int64_t side_blocks = 0, kick_blocks = 0;  // energy partials of a pass, of a kick
  bool bins_ok = false;
  // migration work arrays
  int32_t *mig_blk = nullptr, *mig_claim = nullptr, *mig_cand = nullptr, *mig_cnt = nullptr;
  int64_t mig_nblk = 0, mig_nlocal = 0;
  int32_t mig_cap = 0;
};

static ItlState &state(mmd_ctx *c) {
  if (!c->itl_state) c->itl_state = new ItlState();
  return *static_cast<ItlState *>(c->itl_state);
}

template <typename T>
static void dfree_t(T *&p) {
  if (p) (void)hipFree((void *)p);
  p = nullptr;
}

static void free_bins(mmd_ctx *c, ItlState &s) {
  dfree_t(s.bin_count), dfree_t(s.bin_start), dfree_t(s.scan_tot), dfree_t(c->e_itl);
  s.bins_ok = false;
  c->itl_part_n = c->itl_kick_n = 0;
}

void itl_free(mmd_ctx *c) {
  dfree_t(c->ix), dfree_t(c->iy), dfree_t(c->iz), dfree_t(c->ivx), dfree_t(c->ivy), dfree_t(c->ivz);
  dfree_t(c->ifx), dfree_t(c->ify), dfree_t(c->ifz), dfree_t(c->irho), dfree_t(c->idf), dfree_t(c->itype);
  dfree_t(c->d_wide), dfree_t(c->e_itl);
  if (c->itl_state) {
    ItlState &s = state(c);
    free_bins(c, s);
    dfree_t(s.ax), dfree_t(s.ay), dfree_t(s.az), dfree_t(s.bin), dfree_t(s.tmp);
    dfree_t(s.px), dfree_t(s.py), dfree_t(s.pz), dfree_t(s.spos), dfree_t(s.sanc), dfree_t(s.err);
    dfree_t(s.mig_blk), dfree_t(s.mig_claim), dfree_t(s.mig_cand), dfree_t(s.mig_cnt);
    delete static_cast<ItlState *>(c->itl_state);
    c->itl_state = nullptr;
  }
  c->n_itl = c->itl_cap = c->n_wide = 0;
  c->itl_part_n = c->itl_kick_n = 0;
}

static ItlDev make_dev(mmd_ctx *c) {
  ItlState &s = state(c);
  ItlDev I{};
  I.x = c->ix, I.y = c->iy, I.z = c->iz;
  I.fx = c->ifx, I.fy = c->ify, I.fz = c->ifz, I.rho = c->irho, I.df = c->idf;
  I.type = c->itype;
  I.ax = s.ax, I.ay = s.ay, I.az = s.az, I.px = s.px, I.py = s.py, I.pz = s.pz;
  I.bin = s.bin, I.bin_count = s.bin_count, I.bin_start = s.bin_start, I.scan_tot = s.scan_tot;
  I.tmp = s.tmp, I.spos = s.spos, I.sanc = s.sanc;
  I.err = s.err;
  I.wide = c->d_wide;
  I.n = c->n_itl, I.n_wide = c->n_wide, I.max_d2q = s.max_d2q;
  for (int d = 0; d < 3; d++) I.nb[d] = s.plan.nb[d], I.bc[d] = s.plan.bc[d], I.rii[d] = s.plan.rii[d], I.rw[d] = s.plan.rw[d];
  I.nbins = s.plan.nbins;
  // two interstitials closer than rc have anchors closer than rc + 2 r_ws
  // (r_ws = a sqrt(5) / 4, the Wigner-Seitz cell's reach); the sites' side
  // stages the anchors within twice W's radius, sqrt(max_d2q) a / 2, from
  // bins searched 2 wreach cells wide
  const double a = c->md.lattice_const;
  bool staged = std::sqrt((double)s.max_d2q) * a >= c->md.cutoff + 0.5 * std::sqrt(5.0) * a + 1e-9 * a;
  for (int d = 0; d < 3; d++) staged = staged && 2 * s.plan.rw[d] >= s.plan.rii[d];
  I.ii_staged = staged ? 1 : 0;
  I.Lx = c->md.box_len[0], I.Ly = c->md.box_len[1], I.Lz = c->md.box_len[2];
  I.a = c->md.lattice_const;
  I.rc2 = c->md.cutoff * c->md.cutoff;
  return I;
}

// Bins of the current box.  Two interstitials closer than rc have anchor cells
// at most rc / a + 1.5 apart along an axis (a position lies within a / 2 of
// its anchor along each axis, the Wigner-Seitz cell's reach, and the anchor
// within a / 2 of its cell's corner): the interstitial pairs' search reaches
// rii cells.  W reaches wreach cells, so two interstitials that share a
// lattice site have anchors at most 2 wreach apart: the sites' side searches
// that far.  A bin spans at least a third of the wider search (and
// kBinCells), so a search covers at most four bins per axis (checked by
// ensure_bins).
constexpr int32_t kBinCells = 5;

static BinPlan bin_plan(const mmd_ctx *c, const ItlState &s) {
  BinPlan p{};
  const int32_t nc[3] = {c->g.bxh, c->g.by, c->g.bz};
  const int32_t rii = (int32_t)std::floor(c->md.cutoff / c->md.lattice_const + 1.62);
  p.nbins = 1;
  for (int d = 0; d < 3; d++) {
    p.rii[d] = rii;
    p.rw[d] = s.wreach[d];
    const int32_t reach = std::max(rii, 2 * s.wreach[d]);
    p.bc[d] = std::min(std::max((2 * reach + 1 + 2) / 3, kBinCells), nc[d]);  // fewest cells per bin
    p.nb[d] = std::max(nc[d] / p.bc[d], 1);
    p.nbins *= p.nb[d];
  }
  return p;
}

bool itl_bins_ready(mmd_ctx *c) {
  if (!c->itl_state || !c->md_ready) return false;
  ItlState &s = state(c);
  return s.bins_ok && bin_plan(c, s) == s.plan;
}

// The most bins a search of the given reach spans along one axis (the
// device's bin_span), over every cell it can start from.
static int max_span(int n, int nb, int reach) {
  int m = 0;
  for (int c = 0; c < n; c++) {
    int b0, cnt;
    bin_span(c - reach, c + reach, n, nb, b0, cnt);
    m = std::max(m, cnt);
  }
  return m;
}

// Sizes the bins and the energy partials for the current box (outside any
// capture: the passes themselves never allocate).
static int ensure_bins(mmd_ctx *c) {
  ItlState &s = state(c);
  const BinPlan p = bin_plan(c, s);
  if (s.bins_ok && p == s.plan) return MMD_OK;
  const int32_t nc[3] = {c->g.bxh, c->g.by, c->g.bz};
  int spans = 1;
  for (int d = 0; d < 3; d++) spans *= max_span(nc[d], p.nb[d], std::max(p.rii[d], 2 * p.rw[d]));
  if (spans > kWave)
    return fail(MMD_E_RANGE, "interstitial bins: a search would span %d bins (the passes take at most %d)", spans,
                kWave);
  ITL_TRY(hipStreamSynchronize(c->stream));
  ctx_drop_graphs(c);  // graphs hold the bin arrays and the grids
  free_bins(c, s);
  const int64_t nscan = ((int64_t)p.nbins + kScanChunk - 1) / kScanChunk;
  const int64_t cap = std::max(c->itl_cap, 1);
  s.side_blocks = (cap + kWavesPerBlock - 1) / kWavesPerBlock;
  s.kick_blocks = (cap + kBlock - 1) / kBlock;
  const int64_t n_e = 2 * s.side_blocks + s.kick_blocks;
  ITL_TRY(hipMalloc((void **)&s.bin_count, (size_t)p.nbins * sizeof(int32_t)));
  ITL_TRY(hipMalloc((void **)&s.bin_start, ((size_t)p.nbins + 1) * sizeof(int32_t)));
  ITL_TRY(hipMalloc((void **)&s.scan_tot, (size_t)nscan * sizeof(int32_t)));
  ITL_TRY(hipMalloc((void **)&c->e_itl, (size_t)n_e * sizeof(double)));
  ITL_TRY(launch_fill_f64(c->stream, c->e_itl, n_e, 0.0));
  ITL_TRY(launch_fill_i32(c->stream, s.bin_count, p.nbins, 0));
  ITL_TRY(hipStreamSynchronize(c->stream));
  s.plan = p;
  s.nscan = (int32_t)nscan;
  // energy partials: rho pass, force pass, kicks
  c->itl_part_n = 2 * s.side_blocks;
  c->itl_kick_n = s.kick_blocks;
  s.bins_ok = true;
  return MMD_OK;
}

int itl_setup(mmd_ctx *c) {
  if (!c->md_ready || !c->itl_state || c->itl_cap < 1) return MMD_OK;
  return ensure_bins(c);
}

// anchors, wrapped positions and the (bin, id) order of this step's positions
static int itl_prepare(mmd_ctx *c) {
  if (!itl_bins_ready(c))
    return fail(MMD_E_STATE, "interstitial bins are not sized for this box (mmd_md_setup and mmd_itl_set size them)");
  ItlState &s = state(c);
  ItlDev I = make_dev(c);
  const dim3 grid((c->n_itl + kBlock - 1) / kBlock);
  k_itl_anchor<<<grid, kBlock, 0, c->stream>>>(c->g, I);  // counts are zero: set at sizing, cleared by k_itl_rank
  ITL_TRY(hipGetLastError());
  k_bin_scan_local<<<dim3((unsigned)s.nscan), kBlock, 0, c->stream>>>(I);
  ITL_TRY(hipGetLastError());
  k_bin_scan_tot<<<1, kBlock, 0, c->stream>>>(I, s.nscan);
  ITL_TRY(hipGetLastError());
  k_bin_scan_add<<<dim3((unsigned)s.nscan), kBlock, 0, c->stream>>>(I);
  ITL_TRY(hipGetLastError());
  k_itl_place<<<grid, kBlock, 0, c->stream>>>(I);
  ITL_TRY(hipGetLastError());
  k_itl_rank<<<grid, kBlock, 0, c->stream>>>(I);
  ITL_TRY(hipGetLastError());
  c->itl_stale = false;
  return MMD_OK;
}

// a pass in the context's precision; ep = the pass's energy partials
template <int MODE, bool WANT_E, int PREC>
static void launch_itl_both(mmd_ctx *c, const ItlDev &I, double *ep) {
  const dim3 g_side((c->n_itl + kWavesPerBlock - 1) / kWavesPerBlock);
  k_itl_side<MODE, WANT_E, PREC><<<g_side, kBlock, 0, c->stream>>>(c->g, I, c->pos, c->rho, c->df, c->fx, c->fy, c->fz,
                                                                   c->tf, c->tm, ep);
}

template <int MODE>
static int launch_itl_pass(mmd_ctx *c, const ItlDev &I, double *ep, bool want_e) {
  if (c->prec == MMD_PREC_MIXED) {
    if (want_e)
      launch_itl_both<MODE, true, 1>(c, I, ep);
    else
      launch_itl_both<MODE, false, 1>(c, I, ep);
  } else {
    if (want_e)
      launch_itl_both<MODE, true, 0>(c, I, ep);
    else
      launch_itl_both<MODE, false, 0>(c, I, ep);
  }
  ITL_TRY(hipGetLastError());
  return MMD_OK;
}

int itl_rho_df(mmd_ctx *c, double cutoff, bool want_e) {
  if (!c->md_ready) return fail(MMD_E_STATE, "interstitial passes need mmd_md_setup (box, lattice constant, cutoff)");
  if (std::fabs(cutoff - c->md.cutoff) > 1e-12 * cutoff)
    return fail(MMD_E_INVALID, "interstitial passes use the cutoff given to mmd_md_setup (%g), got %g", c->md.cutoff,
                cutoff);
  hipEvent_t ev;
  int rc = timed_begin(c, MMD_K_ITL, &ev);
  if (rc || (rc = itl_prepare(c))) return rc;
  if ((rc = launch_itl_pass<0>(c, make_dev(c), c->e_itl, want_e))) return rc;
  return timed_end(c, MMD_K_ITL, ev);
}
int itl_df(mmd_ctx *c) {
  ItlDev I = make_dev(c);
  k_itl_df<<<dim3((c->n_itl + kBlock - 1) / kBlock), kBlock, 0, c->stream>>>(I, c->tf);
  ITL_TRY(hipGetLastError());
  return MMD_OK;
}

int itl_force(mmd_ctx *c, double cutoff, bool want_e) {
  if (!c->md_ready) return fail(MMD_E_STATE, "interstitial passes need mmd_md_setup (box, lattice constant, cutoff)");
  if (std::fabs(cutoff - c->md.cutoff) > 1e-12 * cutoff)
    return fail(MMD_E_INVALID, "interstitial passes use the cutoff given to mmd_md_setup (%g), got %g", c->md.cutoff,
                cutoff);
  hipEvent_t ev;
  int rc = timed_begin(c, MMD_K_ITL, &ev);
  if (rc) return rc;
  if (c->itl_stale && (rc = itl_prepare(c))) return rc;  // positions changed since the rho pass
  // the rho pass's partials fill the first half of e_itl
  if ((rc = launch_itl_pass<1>(c, make_dev(c), c->e_itl + c->itl_part_n / 2, want_e))) return rc;
  return timed_end(c, MMD_K_ITL, ev);
}
int itl_kick(mmd_ctx *c, double h, bool drift, bool want_ke) {
  ItlDev I = make_dev(c);
  const dim3 grid((c->n_itl + kBlock - 1) / kBlock);
  double *kp = c->e_itl + c->itl_part_n;  // kinetic partials follow the energy partials
  if (drift) {
    if (want_ke)
      k_itl_kick<true, true><<<grid, kBlock, 0, c->stream>>>(I, c->ivx, c->ivy, c->ivz, c->ix, c->iy, c->iz, c->masses,
                                                               h, c->md.dt, kp);
    else
      k_itl_kick<true, false><<<grid, kBlock, 0, c->stream>>>(I, c->ivx, c->ivy, c->ivz, c->ix, c->iy, c->iz,
                                                                c->masses, h, c->md.dt, kp);
  } else {
    if (want_ke)
      k_itl_kick<false, true><<<grid, kBlock, 0, c->stream>>>(I, c->ivx, c->ivy, c->ivz, c->ix, c->iy, c->iz,
                                                                c->masses, h, c->md.dt, kp);
    else
      k_itl_kick<false, false><<<grid, kBlock, 0, c->stream>>>(I, c->ivx, c->ivy, c->ivz, c->ix, c->iy, c->iz,
                                                                 c->masses, h, c->md.dt, kp);
  }
  ITL_TRY(hipGetLastError());
  if (drift) c->itl_stale = true;
  return MMD_OK;
}

int itl_kick_fused(mmd_ctx *c, double h, bool want_ke) {
  ItlDev I = make_dev(c);
  const dim3 grid((c->n_itl + kBlock - 1) / kBlock);
  double *kp = c->e_itl + c->itl_part_n;
  if (want_ke)
    k_itl_kick_fused<true><<<grid, kBlock, 0, c->stream>>>(I, c->ivx, c->ivy, c->ivz, c->ix, c->iy, c->iz, c->masses, h,
                                                           c->md.dt, kp);
  else
    k_itl_kick_fused<false><<<grid, kBlock, 0, c->stream>>>(I, c->ivx, c->ivy, c->ivz, c->ix, c->iy, c->iz, c->masses,
                                                            h, c->md.dt, kp);
  ITL_TRY(hipGetLastError());
  c->itl_stale = true;
  return MMD_OK;
}

int itl_check(mmd_ctx *c) {
  if (c->n_itl == 0 || !c->itl_state) return MMD_OK;
  ItlState &s = state(c);
  int32_t e = 0;
  ITL_TRY(hipMemcpyAsync(&e, s.err, sizeof(e), hipMemcpyDeviceToHost, c->stream));
  ITL_TRY(hipStreamSynchronize(c->stream));
  if (e & 1)
    return fail(MMD_E_NONFINITE, "an interstitial position is not finite or lies beyond %g lattice constants",
                kItlMaxCoord);
  return MMD_OK;
}

}  // namespace mmd

using namespace mmd;

extern "C" int mmd_itl_set(mmd_ctx *c, int32_t n, const double *x, const double *y, const double *z, const double *vx,
                           const double *vy, const double *vz, const int8_t *type, const int32_t *wide3,
                           int32_t n_wide, int32_t cap) {
  int rc = ctx_bind(c);
  if (rc) return rc;
  if (!c->has_domain) return fail(MMD_E_STATE, "mmd_itl_set: domain not set");
  if (n < 0 || cap < n || cap > (1 << 24)) return fail(MMD_E_RANGE, "mmd_itl_set: need 0 <= n <= cap <= 2^24");
  if (n > 0 && (!x || !y || !z || !vx || !vy || !vz || !type || !wide3 || n_wide < 1 || n_wide > 4096))
    return fail(MMD_E_INVALID, "mmd_itl_set: arrays or wide stencil missing");
  // W: within the ghost shell, a ball of sites around the anchor (every site
  // within its largest squared distance, each once), and narrower than half
  // the box, so a site's offset from an anchor is its centred periodic
  // difference
  int32_t max_d2q = 0, reach[3] = {0, 0, 0}, maxd[3] = {0, 0, 0};
  std::vector<int64_t> keys;
  for (int32_t w = 0; w < n_wide; w++) {
    const int dX = wide3[3 * w], dY = wide3[3 * w + 1], dZ = wide3[3 * w + 2];
    if (std::abs(dX) > c->dom.ghost[0] || std::abs(dY) > c->dom.ghost[1] || std::abs(dZ) > c->dom.ghost[2])
      return fail(MMD_E_RANGE, "mmd_itl_set: wide stencil entry %d leaves the ghost shell", w);
    const int s1 = dX & 1;
    max_d2q = std::max(max_d2q, dX * dX + (2 * dY + s1) * (2 * dY + s1) + (2 * dZ + s1) * (2 * dZ + s1));
    auto fd2 = [](int v) { return v >= 0 ? v / 2 : -((-v + 1) / 2); };
    reach[0] = std::max(reach[0], std::max(std::abs(fd2(dX)), std::abs(fd2(1 - dX))));
    reach[1] = std::max(reach[1], std::abs(dY));
    reach[2] = std::max(reach[2], std::abs(dZ));
    maxd[0] = std::max(maxd[0], std::abs(dX)), maxd[1] = std::max(maxd[1], std::abs(dY));
    maxd[2] = std::max(maxd[2], std::abs(dZ));
    keys.push_back(((int64_t)(dZ + 1024) * 2048 + (dY + 1024)) * 2048 + (dX + 1024));
  }
  if (n_wide > 0) {
    std::sort(keys.begin(), keys.end());
    int64_t ball = 0;
    const int rx = (int)std::ceil(std::sqrt((double)max_d2q)) + 1, ry = rx / 2 + 1;
    for (int dZ = -ry; dZ <= ry; dZ++)
      for (int dY = -ry; dY <= ry; dY++)
        for (int dX = -rx; dX <= rx; dX++) {
          const int s1 = dX & 1;
          if (dX * dX + (2 * dY + s1) * (2 * dY + s1) + (2 * dZ + s1) * (2 * dZ + s1) <= max_d2q) ball++;
        }
    if (std::adjacent_find(keys.begin(), keys.end()) != keys.end() || ball != n_wide)
      return fail(MMD_E_INVALID, "mmd_itl_set: the wide stencil must hold every site within its largest distance "
                                 "of the anchor, each once (%d entries, the ball has %lld)",
                  n_wide, (long long)ball);
    if (2 * maxd[0] >= c->dom.box[0] || 2 * maxd[1] >= c->dom.box[1] || 2 * maxd[2] >= c->dom.box[2])
      return fail(MMD_E_RANGE, "mmd_itl_set: the wide stencil reaches half the box or more (box %d x %d x %d)",
                  c->dom.box[0], c->dom.box[1], c->dom.box[2]);
  }
  int32_t max_type = -1;
  for (int32_t k = 0; k < n; k++) {
    if (type[k] < 0 || type[k] >= kMaxTypes || (c->has_tables && type[k] >= c->n_types))
      return fail(MMD_E_INVALID, "mmd_itl_set: interstitial %d has type %d", k, type[k]);
    max_type = std::max<int32_t>(max_type, type[k]);
  }
  ITL_TRY(hipStreamSynchronize(c->stream));
  ctx_drop_graphs(c);
  const bool realloc = cap > c->itl_cap || n_wide > c->n_wide || !c->d_wide;
  if (realloc) {
    itl_free(c);
    ItlState &s = state(c);
    const size_t nc = (size_t)std::max(cap, 1);
    double **d[] = {&c->ix, &c->iy, &c->iz, &c->ivx, &c->ivy, &c->ivz, &c->ifx, &c->ify, &c->ifz, &c->irho,
                    &c->idf, &s.px, &s.py, &s.pz};
    for (double **p : d) ITL_TRY(hipMalloc((void **)p, nc * sizeof(double)));
    int32_t **ip[] = {&s.ax, &s.ay, &s.az, &s.bin, &s.tmp};
    for (int32_t **p : ip) ITL_TRY(hipMalloc((void **)p, nc * sizeof(int32_t)));
    ITL_TRY(hipMalloc((void **)&s.spos, nc * sizeof(double4)));
    ITL_TRY(hipMalloc((void **)&s.sanc, nc * sizeof(int4)));
    ITL_TRY(hipMalloc((void **)&c->itype, nc));
    ITL_TRY(hipMalloc((void **)&s.err, sizeof(int32_t)));
    ITL_TRY(hipMalloc((void **)&c->d_wide, (size_t)std::max(n_wide, 1) * 3 * sizeof(int32_t)));
    c->itl_cap = cap;
  }
  ItlState &s = state(c);
  s.max_d2q = max_d2q;
  for (int d = 0; d < 3; d++) s.wreach[d] = reach[d];
  c->n_wide = n_wide;
  c->n_itl = n;
  c->itl_max_type = max_type;
  c->itl_stale = true;
  if (n_wide > 0) ITL_TRY(hipMemcpy(c->d_wide, wide3, (size_t)n_wide * 3 * sizeof(int32_t), hipMemcpyHostToDevice));
  ITL_TRY(launch_fill_i32(c->stream, s.err, 1, 0));
  if (n > 0) {
    const size_t b = (size_t)n * sizeof(double);
    ITL_TRY(hipMemcpy(c->ix, x, b, hipMemcpyHostToDevice));
    ITL_TRY(hipMemcpy(c->iy, y, b, hipMemcpyHostToDevice));
    ITL_TRY(hipMemcpy(c->iz, z, b, hipMemcpyHostToDevice));
    ITL_TRY(hipMemcpy(c->ivx, vx, b, hipMemcpyHostToDevice));
    ITL_TRY(hipMemcpy(c->ivy, vy, b, hipMemcpyHostToDevice));
    ITL_TRY(hipMemcpy(c->ivz, vz, b, hipMemcpyHostToDevice));
    ITL_TRY(hipMemcpy(c->itype, type, (size_t)n, hipMemcpyHostToDevice));
    ITL_TRY(launch_fill_f64(c->stream, c->ifx, n, 0.0));
    ITL_TRY(launch_fill_f64(c->stream, c->ify, n, 0.0));
    ITL_TRY(launch_fill_f64(c->stream, c->ifz, n, 0.0));
    ITL_TRY(launch_fill_f64(c->stream, c->irho, n, 0.0));
    ITL_TRY(launch_fill_f64(c->stream, c->idf, n, 0.0));
  }
  ITL_TRY(hipStreamSynchronize(c->stream));
  // a new W or capacity changes the bins and partials: sized now when the MD box is known (else by mmd_md_setup)
  s.bins_ok = false;
  return itl_setup(c);
}
extern "C" int mmd_itl_get(mmd_ctx *c, int32_t *n, double *x, double *y, double *z, double *vx, double *vy, double *vz,
                           double *fx, double *fy, double *fz, double *rho, double *df, int8_t *type) {
  int rc = ctx_bind(c);
  if (rc) return rc;
  if ((rc = itl_check(c))) return rc;
  if (n) *n = c->n_itl;
  const size_t b = (size_t)c->n_itl * sizeof(double);
  if (c->n_itl == 0) return MMD_OK;
  struct P {
    double *h;
    const double *d;
  } ps[] = {{x, c->ix}, {y, c->iy}, {z, c->iz}, {vx, c->ivx}, {vy, c->ivy}, {vz, c->ivz},
            {fx, c->ifx}, {fy, c->ify}, {fz, c->ifz}, {rho, c->irho}, {df, c->idf}};
  for (auto &p : ps)
    if (p.h) ITL_TRY(hipMemcpyAsync(p.h, p.d, b, hipMemcpyDeviceToHost, c->stream));
  if (type) ITL_TRY(hipMemcpyAsync(type, c->itype, (size_t)c->n_itl, hipMemcpyDeviceToHost, c->stream));
  ITL_TRY(hipStreamSynchronize(c->stream));
  return MMD_OK;
}

extern "C" int mmd_itl_migrate(mmd_ctx *c, int32_t *n_to_itl, int32_t *n_to_lattice) {
  int rc = ctx_bind(c);
  if (rc) return rc;
  if (!c->has_domain || !c->md_ready)
    return fail(MMD_E_STATE, "mmd_itl_migrate: needs the domain and mmd_md_setup (box, lattice constant)");
  if (c->itl_cap < 1 || !c->d_wide || !c->itl_state)
    return fail(MMD_E_STATE, "mmd_itl_migrate: call mmd_itl_set first (list capacity and wide stencil)");
  if (c->dom.start[0] || c->dom.start[1] || c->dom.start[2])
    return fail(MMD_E_STATE, "mmd_itl_migrate: a single periodic domain (start = 0) only");
  const Geom &g = c->g;
  const int64_t n_local = 2 * (int64_t)g.bxh * g.by * g.bz;
  if (n_local >= ((int64_t)1 << 31)) return fail(MMD_E_RANGE, "mmd_itl_migrate: more than 2^31 local sites");
  const int64_t nblk = (n_local + kBlock - 1) / kBlock;
  ItlState &s = state(c);
  if (s.mig_nlocal != n_local || s.mig_cap != c->itl_cap) {
    ITL_TRY(hipStreamSynchronize(c->stream));
    dfree_t(s.mig_blk), dfree_t(s.mig_claim), dfree_t(s.mig_cand), dfree_t(s.mig_cnt);
    s.mig_nlocal = s.mig_nblk = 0, s.mig_cap = 0;
    ITL_TRY(hipMalloc((void **)&s.mig_blk, (size_t)nblk * sizeof(int32_t)));
    ITL_TRY(hipMalloc((void **)&s.mig_claim, (size_t)n_local * sizeof(int32_t)));
    ITL_TRY(hipMalloc((void **)&s.mig_cand, (size_t)c->itl_cap * sizeof(int32_t)));
    ITL_TRY(hipMalloc((void **)&s.mig_cnt, 2 * sizeof(int32_t)));
    s.mig_nlocal = n_local, s.mig_nblk = nblk, s.mig_cap = c->itl_cap;
  }
  MigDev M{};
  M.g = g;
  M.P = c->pos, M.T = c->type;
  M.VX = c->vx, M.VY = c->vy, M.VZ = c->vz, M.FX = c->fx, M.FY = c->fy, M.FZ = c->fz;
  M.RHO = c->rho, M.DF = c->df;
  M.x = c->ix, M.y = c->iy, M.z = c->iz, M.vx = c->ivx, M.vy = c->ivy, M.vz = c->ivz;
  M.fx = c->ifx, M.fy = c->ify, M.fz = c->ifz, M.rho = c->irho, M.df = c->idf, M.type = c->itype;
  M.blk = s.mig_blk, M.claim = s.mig_claim, M.cand = s.mig_cand, M.cnt = s.mig_cnt;
  M.n_local = n_local;
  M.n = c->n_itl;
  M.a = c->md.lattice_const;
  M.Lx = c->md.box_len[0], M.Ly = c->md.box_len[1], M.Lz = c->md.box_len[2];
  // phase 1: count the leavers (site order), check the capacity before any change
  k_mig_count<<<dim3((unsigned)nblk), kBlock, 0, c->stream>>>(M);
  ITL_TRY(hipGetLastError());
  k_mig_scan<<<1, kBlock, 0, c->stream>>>(s.mig_blk, nblk, s.mig_cnt);
  ITL_TRY(hipGetLastError());
  int32_t leave = 0;
  ITL_TRY(hipMemcpyAsync(&leave, s.mig_cnt, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
  ITL_TRY(hipStreamSynchronize(c->stream));
  if ((int64_t)c->n_itl + leave > c->itl_cap)
    return fail(MMD_E_RANGE, "mmd_itl_migrate: %d interstitials plus %d leaving the lattice exceed the capacity %d",
                c->n_itl, leave, c->itl_cap);
  if (leave > 0) {
    k_mig_append<<<dim3((unsigned)nblk), kBlock, 0, c->stream>>>(M);
    ITL_TRY(hipGetLastError());
  }
  // phase 2: entries whose nearest site is vacant fill it, lowest index first
  M.n_all = c->n_itl + leave;
  int32_t kept = M.n_all;
  if (M.n_all > 0) {
    const dim3 grid((unsigned)((M.n_all + kBlock - 1) / kBlock));
    ITL_TRY(launch_fill_i32(c->stream, s.mig_claim, n_local, INT32_MAX));
    k_mig_claim<<<grid, kBlock, 0, c->stream>>>(M);
    ITL_TRY(hipGetLastError());
    k_mig_fill<<<grid, kBlock, 0, c->stream>>>(M);
    ITL_TRY(hipGetLastError());
    k_mig_compact<<<1, kBlock, 0, c->stream>>>(M);
    ITL_TRY(hipGetLastError());
    ITL_TRY(hipMemcpyAsync(&kept, s.mig_cnt + 1, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
  }
  ITL_TRY(launch_ghost(c->stream, g, M.Lx, M.Ly, M.Lz, c->pos, c->type, c->df, MMD_F_X | MMD_F_DF));
  ITL_TRY(hipStreamSynchronize(c->stream));
  const int32_t filled = M.n_all - kept;
  if (kept != c->n_itl) {
    ctx_drop_graphs(c);  // launch grids follow the list length
    // partials of blocks beyond the new grids would otherwise stay in the energy sums
    ITL_TRY(launch_fill_f64(c->stream, c->e_itl, c->itl_part_n + c->itl_kick_n, 0.0));
    ITL_TRY(hipStreamSynchronize(c->stream));
  }
  if (leave > 0 || filled > 0) {
    const int32_t m = std::max(c->max_type, c->itl_max_type);
    c->max_type = c->itl_max_type = m;
  }
  c->n_itl = kept;
  c->itl_stale = true;
  if (n_to_itl) *n_to_itl = leave;
  if (n_to_lattice) *n_to_lattice = filled;
  return MMD_OK;
}
