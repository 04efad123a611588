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
// Determinism without atomics on physical quantities: each affected lattice
// site has one owner, the lowest-numbered interstitial interacting with it,
// whose lane adds every interstitial contribution to that site in increasing
// interstitial order.  Interstitial neighbour lists (pairs closer than 2 rc,
// the only interstitials that can share a lattice site) come from periodic
// bins filled with atomics and are then sorted, so their order is fixed.
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

constexpr int kItlNbrCap = 32;  // interstitials within 2 rc of one interstitial
constexpr int kBinCap = 64;     // interstitials per bin

struct ItlDev {
  const double *x, *y, *z;  // positions as stored (unwrapped)
  double *fx, *fy, *fz, *rho, *df;
  const int8_t *type;
  int32_t *ax, *ay, *az;        // anchor site, local lattice coordinates (x doubled)
  double *px, *py, *pz;         // positions shifted by the same periodic image as the anchor
  int32_t *nbr, *nnb;           // [n][kItlNbrCap] sorted neighbour ids, counts
  int32_t *bin_count, *bin_list;
  int32_t *err;
  const int32_t *wide;          // [n_wide][3], parity 0
  int32_t n, n_wide, max_d2q;
  int32_t nbx, nby, nbz;
  double bsx, bsy, bsz;
  double Lx, Ly, Lz, a, rc2, rc2x4;
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

// centred periodic difference in [-n/2, n/2)
__device__ __forceinline__ int cdiff(int v, int n) {
  v = wrap_n(v, n);
  return v >= n - n / 2 ? v - n : v;
}

__device__ __forceinline__ double minimg(double d, double L) { return d - L * rint(d / L); }

// Is lattice site (sx, sy, sz) (local, x doubled) in W of interstitial m and
// closer than rc?  Returns the ext device index of the image seen from m.
__device__ __forceinline__ bool li_pair(const Geom &g, const ItlDev &I, const double4 *P, int m, int sx, int sy,
                                        int sz, int64_t &e, double &d2, double &dx, double &dy, double &dz) {
  const int aX = I.ax[m], aY = I.ay[m], aZ = I.az[m];
  const int dX = cdiff(sx - aX, 2 * g.bxh), dY = cdiff(sy - aY, g.by), dZ = cdiff(sz - aZ, g.bz);
  const int sa = aX & 1, s1 = (aX + dX) & 1;
  const int hy = 2 * dY + s1 - sa, hz = 2 * dZ + s1 - sa;
  if (dX * dX + hy * hy + hz * hz > I.max_d2q) return false;
  const int xe = aX + dX + 2 * g.gxh, ye = aY + dY + g.gy, ze = aZ + dZ + g.gz;
  e = dev_index(g, xe & 1, xe >> 1, ye, ze);
  const double4 q = P[e];
  dx = q.x - I.px[m];
  dy = q.y - I.py[m];
  dz = q.z - I.pz[m];
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

// anchor (nearest lattice site), wrapped into the box, and bins
__global__ __launch_bounds__(kBlock) void k_itl_prepare(Geom g, ItlDev I) {
  const int k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= I.n) return;
  double px = I.x[k], py = I.y[k], pz = I.z[k];
  long long X, Y, Z;
  ws_site(I.a, px, py, pz, X, Y, Z);
  px = fma(-(double)wrap_img64(X, 2 * (long long)g.bxh), I.Lx, px);
  py = fma(-(double)wrap_img64(Y, (long long)g.by), I.Ly, py);
  pz = fma(-(double)wrap_img64(Z, (long long)g.bz), I.Lz, pz);
  I.ax[k] = (int)X, I.ay[k] = (int)Y, I.az[k] = (int)Z;
  I.px[k] = px, I.py[k] = py, I.pz[k] = pz;
  // bin of the wrapped position (positions may sit a little outside [0, L))
  const double wx = px - I.Lx * floor(px / I.Lx), wy = py - I.Ly * floor(py / I.Ly), wz = pz - I.Lz * floor(pz / I.Lz);
  const int ibx = min((int)(wx / I.bsx), I.nbx - 1), iby = min((int)(wy / I.bsy), I.nby - 1),
            ibz = min((int)(wz / I.bsz), I.nbz - 1);
  const int b = (ibz * I.nby + iby) * I.nbx + ibx;
  const int slot = atomicAdd(&I.bin_count[b], 1);
  if (slot < kBinCap)
    I.bin_list[b * kBinCap + slot] = k;
  else
    atomicOr(I.err, 1);
}

__device__ __forceinline__ int bin_of(const ItlDev &I, double p, double L, double bs, int nb) {
  const double w = p - L * floor(p / L);
  return min((int)(w / bs), nb - 1);
}

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

// Sorted list of the interstitials closer than 2 rc (minimum image): one wave
// per interstitial, a lane per neighbouring bin; candidates are gathered in
// LDS and placed by rank (ids are distinct), so the list is in increasing id
// order.  The 27 bins are those of the offsets -1..1 per axis, fewer along an
// axis with fewer than 3 bins so that no bin is visited twice.
__global__ __launch_bounds__(kBlock) void k_itl_nbr(ItlDev I) {
  __shared__ int32_t cand[kWavesPerBlock][kItlNbrCap];
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  const int k = blockIdx.x * kWavesPerBlock + w;
  const bool active = k < I.n;
  int32_t found = 0, b = 0, nb = 0;
  double pkx = 0.0, pky = 0.0, pkz = 0.0;
  if (active && lane < 27) {
    pkx = I.px[k], pky = I.py[k], pkz = I.pz[k];
    const int ox = lane % 3 - 1, oy = (lane / 3) % 3 - 1, oz = lane / 9 - 1;
    const bool vx = I.nbx >= 3 || (ox >= 0 && ox < I.nbx), vy = I.nby >= 3 || (oy >= 0 && oy < I.nby),
               vz = I.nbz >= 3 || (oz >= 0 && oz < I.nbz);
    if (vx && vy && vz) {
      const int bx = bin_of(I, pkx, I.Lx, I.bsx, I.nbx), by = bin_of(I, pky, I.Ly, I.bsy, I.nby),
                bz = bin_of(I, pkz, I.Lz, I.bsz, I.nbz);
      b = (wrap_n(bz + oz, I.nbz) * I.nby + wrap_n(by + oy, I.nby)) * I.nbx + wrap_n(bx + ox, I.nbx);
      nb = min(I.bin_count[b], kBinCap);
    }
  }
  // pass 1: count this lane's candidates; pass 2: store them at the wave offset
  for (int pass = 0; pass < 2; pass++) {
    int32_t off = 0, tot = 0;
    if (pass == 1) {
      off = wave_excl_scan(found, tot);
      if (active && lane == 0 && tot > kItlNbrCap) atomicOr(I.err, 2);
    }
    int32_t c = 0;
    for (int q = 0; q < nb; q++) {
      const int m = I.bin_list[b * kBinCap + q];
      if (m == k) continue;
      const double dx = minimg(pkx - I.px[m], I.Lx), dy = minimg(pky - I.py[m], I.Ly), dz = minimg(pkz - I.pz[m], I.Lz);
      if (dx * dx + dy * dy + dz * dz >= I.rc2x4) continue;
      if (pass == 1 && off + c < kItlNbrCap) cand[w][off + c] = m;
      c++;
    }
    if (pass == 0) found = c;
    if (pass == 1) found = min(tot, kItlNbrCap);
  }
  __syncthreads();
  if (active && lane < found) {
    const int32_t v = cand[w][lane];
    int32_t rank = 0;
    for (int q = 0; q < found; q++) rank += cand[w][q] < v ? 1 : 0;
    I.nbr[k * kItlNbrCap + rank] = v;
  }
  if (active && lane == 0) I.nnb[k] = found;
}

// local lattice coordinates (x doubled) of the periodic source of ext site (xe, ye, ze)
__device__ __forceinline__ void src_local(const Geom &g, int xe, int ye, int ze, int &sx, int &sy, int &sz) {
  sx = wrap_n(xe - 2 * g.gxh, 2 * g.bxh);
  sy = wrap_n(ye - g.gy, g.by);
  sz = wrap_n(ze - g.gz, g.bz);
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

// One wave per interstitial.  MODE 0: rho (+ owner-side lattice rho/df update),
// MODE 1: force (+ owner-side lattice force update).
template <int MODE, bool WANT_E, int PREC>
__global__ __launch_bounds__(kBlock) void k_itl_pass(Geom g, ItlDev I, const double4 *__restrict__ P,
                                                       double *__restrict__ RHO,
                                                       double *__restrict__ DF, double *__restrict__ FX,
                                                       double *__restrict__ FY, double *__restrict__ FZ, TabF64 tf,
                                                       TabMixed tmx, double *__restrict__ e_part) {
  __shared__ double lds[kWavesPerBlock];
  // in-cutoff entries of the current chunk, per wave: wide-stencil index | site type << 24, and the site's position
  __shared__ int32_t wlist[kWavesPerBlock][kItlChunk];
  __shared__ double3 wpos[kWavesPerBlock][kItlChunk];
  const int lane = threadIdx.x & 63;
  const int k = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
  double acc = 0.0, fxk = 0.0, fyk = 0.0, fzk = 0.0, e_loc = 0.0;
  if (k < I.n) {
    const int tk = I.type[k];
    const int aX = I.ax[k], aY = I.ay[k], aZ = I.az[k];
    const int sgn = (aX & 1) ? -1 : 1;
    const int nn = I.nnb[k];
    const int *nb = I.nbr + k * kItlNbrCap;
    const double dfk = MODE == 1 ? I.df[k] : 0.0;
    const double pxk = I.px[k], pyk = I.py[k], pzk = I.pz[k];
    int32_t *cw = wlist[threadIdx.x >> 6];
    double3 *cq = wpos[threadIdx.x >> 6];
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
      // phase 2: one in-cutoff pair per lane (rarely two)
      for (int c = lane; c < found; c += kWave) {
        const int32_t wt = cw[c];
        const int w = wt & 0xffffff, ti = wt >> 24;
        const int xe = aX + sgn * I.wide[3 * w] + 2 * g.gxh, ye = aY + sgn * I.wide[3 * w + 1] + g.gy,
                  ze = aZ + sgn * I.wide[3 * w + 2] + g.gz;
        const int64_t e = dev_index(g, xe & 1, xe >> 1, ye, ze);
        const double3 qp = cq[c];
        double dx, dy, dz;
        const double d2 = itl_d2(make_double4(qp.x, qp.y, qp.z, 0.0), pxk, pyk, pzk, dx, dy, dz);
        if (d2 >= I.rc2) continue;  // the phase-1 test, repeated on the same values
        const double r = sqrt(d2);
        double p;
        const int m = seg(r, tf.r0, tf.inv_dr, tf.n_r, p);
        // this pair's term, reused when the owner adds interstitial k to the site
        double v_k = 0.0, fp_k = 0.0;
        if (MODE == 0) {
          v_k = itl_rho<PREC>(tf, tmx, ti, d2, m, p);
          acc += v_k;
        } else {
          // force on the lattice atom along del = x_i - x_k
          fp_k = itl_fpair<PREC>(tf, tmx, ti, tk, d2, r, m, p, DF[e], dfk);
          fxk -= dx * fp_k, fyk -= dy * fp_k, fzk -= dz * fp_k;
        }
        // owner: no lower-numbered neighbour interacts with this lattice site
        int sx, sy, sz;
        src_local(g, xe, ye, ze, sx, sy, sz);
        bool owner = true;
        for (int q = 0; q < nn && owner; q++) {
          const int mm = nb[q];
          if (mm > k) break;
          int64_t e2;
          double d2b, ex, ey, ez;
          if (li_pair(g, I, P, mm, sx, sy, sz, e2, d2b, ex, ey, ez)) owner = false;
        }
        if (!owner) continue;
        const int64_t src = dev_index(g, sx & 1, (sx >> 1) + g.gxh, sy + g.gy, sz + g.gz);
        double rho_i = 0.0, gx = 0.0, gy = 0.0, gz = 0.0, epair = 0.0;
        if (MODE == 0) rho_i = RHO[src];
        const double dfi = MODE == 1 ? DF[src] : 0.0;
        // every interacting interstitial in increasing order (k merged into its neighbours)
        int placed_k = 0;
        for (int q = 0; q < nn + 1; q++) {
          int mm;
          if (!placed_k && (q - placed_k >= nn || nb[q - placed_k] > k)) {
            mm = k;
            placed_k = 1;
          } else {
            mm = nb[q - placed_k];
          }
          if (mm == k) {  // the pair evaluated above: same image, same distance
            if (MODE == 0) {
              rho_i += tk == ti ? v_k : itl_rho<PREC>(tf, tmx, tk, d2, m, p);
            } else {
              gx += dx * fp_k, gy += dy * fp_k, gz += dz * fp_k;
              if (WANT_E) epair += itl_phi<PREC>(tf, tmx, ti, tk, d2, m, p);
            }
            continue;
          }
          int64_t e2;
          double d2b, ex, ey, ez;
          if (!li_pair(g, I, P, mm, sx, sy, sz, e2, d2b, ex, ey, ez)) continue;
          const int tm = I.type[mm];
          const double rb = sqrt(d2b);
          double pb;
          const int mb = seg(rb, tf.r0, tf.inv_dr, tf.n_r, pb);
          if (MODE == 0) {
            rho_i += itl_rho<PREC>(tf, tmx, tm, d2b, mb, pb);
          } else {
            const double fp = itl_fpair<PREC>(tf, tmx, ti, tm, d2b, rb, mb, pb, dfi, I.df[mm]);
            gx += ex * fp, gy += ey * fp, gz += ez * fp;
            if (WANT_E) epair += itl_phi<PREC>(tf, tmx, ti, tm, d2b, mb, pb);
          }
        }
        if (MODE == 0) {
          double pe;
          const int me = seg(rho_i, tf.rho0, tf.inv_drho, tf.n_rho, pe);
          if (WANT_E) {
            double po;
            const int mo = seg(RHO[src], tf.rho0, tf.inv_drho, tf.n_rho, po);
            e_loc += spl_value(tf.emb, ti * tf.n_rho + me, pe) - spl_value(tf.emb, ti * tf.n_rho + mo, po);
          }
          RHO[src] = rho_i;
          DF[src] = spl_deriv(tf.emb, ti * tf.n_rho + me, pe);
        } else {
          FX[src] += gx, FY[src] += gy, FZ[src] += gz;
          if (WANT_E) e_loc += epair;
        }
      }
      wave_lds_sync();  // the list is rewritten by the next chunk
    }
    // interstitial <-> interstitial
    for (int q = lane; q < nn; q += kWave) {
      const int mm = nb[q];
      const double dx = minimg(I.px[k] - I.px[mm], I.Lx), dy = minimg(I.py[k] - I.py[mm], I.Ly),
                   dz = minimg(I.pz[k] - I.pz[mm], I.Lz);
      const double d2 = dx * dx + dy * dy + dz * dz;
      if (d2 >= I.rc2) continue;
      const int tm = I.type[mm];
      const double r = sqrt(d2);
      double p;
      const int m = seg(r, tf.r0, tf.inv_dr, tf.n_r, p);
      if (MODE == 0) {
        acc += itl_rho<PREC>(tf, tmx, tm, d2, m, p);
      } else {
        const double fp = itl_fpair<PREC>(tf, tmx, tk, tm, d2, r, m, p, dfk, I.df[mm]);
        fxk += dx * fp, fyk += dy * fp, fzk += dz * fp;
        if (WANT_E) e_loc += 0.5 * itl_phi<PREC>(tf, tmx, tk, tm, d2, m, p);
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

struct ItlState {
  int32_t *ax = nullptr, *ay = nullptr, *az = nullptr, *nbr = nullptr, *nnb = nullptr;
  double *px = nullptr, *py = nullptr, *pz = nullptr;
  int32_t *bins = nullptr;  // counts followed by lists
  int32_t *err = nullptr;
  int32_t nbx = 0, nby = 0, nbz = 0, max_d2q = 0;
  // migration work arrays
  int32_t *mig_blk = nullptr, *mig_claim = nullptr, *mig_cand = nullptr, *mig_cnt = nullptr;
  int64_t mig_nblk = 0, mig_nlocal = 0;
  int32_t mig_cap = 0;
};

static ItlState &state(mmd_ctx *c) {
  static_assert(sizeof(ItlState) > 0, "");
  if (!c->itl_state) c->itl_state = new ItlState();
  return *static_cast<ItlState *>(c->itl_state);
}

template <typename T>
static void dfree_t(T *&p) {
  if (p) (void)hipFree((void *)p);
  p = nullptr;
}

void itl_free(mmd_ctx *c) {
  dfree_t(c->ix), dfree_t(c->iy), dfree_t(c->iz), dfree_t(c->ivx), dfree_t(c->ivy), dfree_t(c->ivz);
  dfree_t(c->ifx), dfree_t(c->ify), dfree_t(c->ifz), dfree_t(c->irho), dfree_t(c->idf), dfree_t(c->itype);
  dfree_t(c->d_wide), dfree_t(c->e_itl);
  if (c->itl_state) {
    ItlState &s = state(c);
    dfree_t(s.ax), dfree_t(s.ay), dfree_t(s.az), dfree_t(s.nbr), dfree_t(s.nnb);
    dfree_t(s.px), dfree_t(s.py), dfree_t(s.pz), dfree_t(s.bins), dfree_t(s.err);
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
  I.nbr = s.nbr, I.nnb = s.nnb;
  const int nbins = s.nbx * s.nby * s.nbz;
  I.bin_count = s.bins, I.bin_list = s.bins + nbins;
  I.err = s.err;
  I.wide = c->d_wide;
  I.n = c->n_itl, I.n_wide = c->n_wide, I.max_d2q = s.max_d2q;
  I.nbx = s.nbx, I.nby = s.nby, I.nbz = s.nbz;
  I.Lx = c->md.box_len[0], I.Ly = c->md.box_len[1], I.Lz = c->md.box_len[2];
  I.bsx = I.Lx / s.nbx, I.bsy = I.Ly / s.nby, I.bsz = I.Lz / s.nbz;
  I.a = c->md.lattice_const;
  I.rc2 = c->md.cutoff * c->md.cutoff;
  I.rc2x4 = 4.0 * I.rc2;
  return I;
}

static void bin_grid(const mmd_ctx *c, int &nx, int &ny, int &nz) {
  const double bs = 2.0 * c->md.cutoff;
  nx = std::max(1, (int)(c->md.box_len[0] / bs)), ny = std::max(1, (int)(c->md.box_len[1] / bs)),
  nz = std::max(1, (int)(c->md.box_len[2] / bs));
}

bool itl_bins_ready(mmd_ctx *c) {
  if (!c->itl_state || !c->md_ready) return false;
  const ItlState &s = state(c);
  int nx, ny, nz;
  bin_grid(c, nx, ny, nz);
  return s.bins && nx == s.nbx && ny == s.nby && nz == s.nbz;
}

// bins sized from the MD box; (re)allocated when the box changes
static int ensure_bins(mmd_ctx *c) {
  ItlState &s = state(c);
  int nx, ny, nz;
  bin_grid(c, nx, ny, nz);
  if (s.bins && nx == s.nbx && ny == s.nby && nz == s.nbz) return MMD_OK;
  ctx_drop_graphs(c);
  dfree_t(s.bins);
  const size_t n = (size_t)nx * ny * nz;
  ITL_TRY(hipMalloc((void **)&s.bins, n * (1 + kBinCap) * sizeof(int32_t)));
  s.nbx = nx, s.nby = ny, s.nbz = nz;
  return MMD_OK;
}

static int itl_prepare(mmd_ctx *c) {
  if (!c->md_ready) return fail(MMD_E_STATE, "interstitial passes need mmd_md_setup (box, lattice constant, cutoff)");
  int rc = ensure_bins(c);
  if (rc) return rc;
  ItlState &s = state(c);
  ItlDev I = make_dev(c);
  const int nbins = s.nbx * s.nby * s.nbz;
  ITL_TRY(launch_fill_i32(c->stream, s.bins, nbins, 0));
  const dim3 grid((c->n_itl + kBlock - 1) / kBlock);
  k_itl_prepare<<<grid, kBlock, 0, c->stream>>>(c->g, I);
  ITL_TRY(hipGetLastError());
  k_itl_nbr<<<dim3((c->n_itl + kWavesPerBlock - 1) / kWavesPerBlock), kBlock, 0, c->stream>>>(I);
  ITL_TRY(hipGetLastError());
  c->itl_stale = false;
  return MMD_OK;
}

// the interstitial pass in the context's precision
template <int MODE>
static void launch_itl_pass(mmd_ctx *c, const ItlDev &I, dim3 grid, double *ep, bool want_e) {
  if (c->prec == MMD_PREC_MIXED) {
    if (want_e)
      k_itl_pass<MODE, true, 1><<<grid, kBlock, 0, c->stream>>>(c->g, I, c->pos, c->rho, c->df, c->fx, c->fy, c->fz,
                                                                c->tf, c->tm, ep);
    else
      k_itl_pass<MODE, false, 1><<<grid, kBlock, 0, c->stream>>>(c->g, I, c->pos, c->rho, c->df, c->fx, c->fy, c->fz,
                                                                 c->tf, c->tm, ep);
  } else {
    if (want_e)
      k_itl_pass<MODE, true, 0><<<grid, kBlock, 0, c->stream>>>(c->g, I, c->pos, c->rho, c->df, c->fx, c->fy, c->fz,
                                                                c->tf, c->tm, ep);
    else
      k_itl_pass<MODE, false, 0><<<grid, kBlock, 0, c->stream>>>(c->g, I, c->pos, c->rho, c->df, c->fx, c->fy, c->fz,
                                                                 c->tf, c->tm, ep);
  }
}

int itl_rho_df(mmd_ctx *c, double cutoff, bool want_e) {
  if (!c->md_ready) return fail(MMD_E_STATE, "interstitial passes need mmd_md_setup (box, lattice constant, cutoff)");
  if (std::fabs(cutoff - c->md.cutoff) > 1e-12 * cutoff)
    return fail(MMD_E_INVALID, "interstitial passes use the cutoff given to mmd_md_setup (%g), got %g", c->md.cutoff,
                cutoff);
  hipEvent_t ev;
  int rc = timed_begin(c, MMD_K_ITL, &ev);
  if (rc || (rc = itl_prepare(c))) return rc;
  ItlDev I = make_dev(c);
  const dim3 grid((c->n_itl + kWavesPerBlock - 1) / kWavesPerBlock);
  launch_itl_pass<0>(c, I, grid, c->e_itl, want_e);
  ITL_TRY(hipGetLastError());
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
  ItlDev I = make_dev(c);
  const dim3 grid((c->n_itl + kWavesPerBlock - 1) / kWavesPerBlock);
  // the embedding-energy partials of the rho pass live in the first half of e_itl
  double *ep = c->e_itl + c->itl_part_n / 2;
  launch_itl_pass<1>(c, I, grid, ep, want_e);
  ITL_TRY(hipGetLastError());
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
  if (e & 1) return fail(MMD_E_RANGE, "interstitial bin overflow (more than %d per bin)", kBinCap);
  if (e & 2) return fail(MMD_E_RANGE, "more than %d interstitials within 2 rc of one interstitial", kItlNbrCap);
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
  int32_t max_d2q = 0;
  for (int32_t w = 0; w < n_wide; w++) {
    const int dX = wide3[3 * w], dY = wide3[3 * w + 1], dZ = wide3[3 * w + 2];
    if (std::abs(dX) > c->dom.ghost[0] || std::abs(dY) > c->dom.ghost[1] || std::abs(dZ) > c->dom.ghost[2])
      return fail(MMD_E_RANGE, "mmd_itl_set: wide stencil entry %d leaves the ghost shell", w);
    const int s1 = dX & 1;
    max_d2q = std::max(max_d2q, dX * dX + (2 * dY + s1) * (2 * dY + s1) + (2 * dZ + s1) * (2 * dZ + s1));
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
    int32_t **ip[] = {&s.ax, &s.ay, &s.az, &s.nnb};
    for (int32_t **p : ip) ITL_TRY(hipMalloc((void **)p, nc * sizeof(int32_t)));
    ITL_TRY(hipMalloc((void **)&s.nbr, nc * kItlNbrCap * sizeof(int32_t)));
    ITL_TRY(hipMalloc((void **)&c->itype, nc));
    ITL_TRY(hipMalloc((void **)&s.err, sizeof(int32_t)));
    ITL_TRY(hipMalloc((void **)&c->d_wide, (size_t)std::max(n_wide, 1) * 3 * sizeof(int32_t)));
    // energy partials: rho-pass blocks, force-pass blocks, kinetic blocks
    const int64_t nb_pass = (cap + kWavesPerBlock - 1) / kWavesPerBlock;
    const int64_t nb_kick = (cap + kBlock - 1) / kBlock;
    c->itl_part_n = 2 * nb_pass;
    c->itl_kick_n = nb_kick;
    ITL_TRY(hipMalloc((void **)&c->e_itl, (size_t)(2 * nb_pass + nb_kick) * sizeof(double)));
    c->itl_cap = cap;
  }
  ItlState &s = state(c);
  s.max_d2q = max_d2q;
  c->n_wide = n_wide;
  c->n_itl = n;
  c->itl_max_type = max_type;
  c->itl_stale = true;
  if (n_wide > 0) ITL_TRY(hipMemcpy(c->d_wide, wide3, (size_t)n_wide * 3 * sizeof(int32_t), hipMemcpyHostToDevice));
  ITL_TRY(launch_fill_i32(c->stream, s.err, 1, 0));
  const int64_t nb_pass = (c->itl_cap + kWavesPerBlock - 1) / kWavesPerBlock;
  const int64_t nb_kick = (c->itl_cap + kBlock - 1) / kBlock;
  ITL_TRY(launch_fill_f64(c->stream, c->e_itl, 2 * nb_pass + nb_kick, 0.0));
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
  return MMD_OK;
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
