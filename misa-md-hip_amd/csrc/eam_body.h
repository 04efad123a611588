// Device code of the EAM passes shared by the single-domain kernels
// (kernels_eam.hip) and the batched serving kernels (kernels_batch.hip): pair
// terms, spline evaluation and the per-block bodies of rho + df and force.
//
// Reference semantics (misa-md-hip): one result per local lattice site with
// type >= 0 (src/kernels/kernel_itl.hpp:69-105), neighbours from the parity
// stencil (:82-86), neighbours with type < 0 skipped (:48-57), pair included
// only for dist2 < rc^2 (src/kernels/kernel_pairs.hpp:21-32, :39-60), rho by
// the neighbour's type, del = x_i - x_j, results overwritten.
#ifndef MMD_EAM_BODY_H
#define MMD_EAM_BODY_H

#include "device_common.h"

namespace mmd {


#ifndef MMD_PAIR_UNROLL
#define MMD_PAIR_UNROLL 4
#endif
constexpr int kUnroll = MMD_PAIR_UNROLL;  // stencil entries whose loads are in flight together
// Stencil entries per trip of the rho pass.  Its occupancy is set by its LDS
// tile (four blocks per CU), so a pass that stays within 128 registers at
// eight entries keeps eight entries' loads in flight: the mixed MD pass (one
// 16-byte gather per pair, under 100 registers) and the FP64 pass (106; half
// the trips' waits and ballots, 0.3-1.5 % faster than at four).  The mixed
// passes that sum two rho records, or that keep the cubic records for
// energies, measured slower at eight and keep four.
template <int PREC, bool LEAN, bool WANT_E>
constexpr int rho_unroll() {
  return PREC == 0 || (PREC == 1 && LEAN && !WANT_E) ? 8 : 4;
}
static_assert(32 % kUnroll == 0, "a trip's entries share one word of the energy masks");
// stencil rows are read up to index nk + U - 1 for U entries per trip (the
// next trip's offsets are read ahead); rows are zero-padded to kMaxStencil and
// mmd_set_stencil accepts at most kMaxStencil - 8 entries
static_assert(kUnroll <= 8, "stencil padding too short for the unroll");

// 1/sqrt(d2) for d2 > 0: hardware reciprocal-square-root seed refined by two
// Newton steps to full FP64 accuracy (a few ulp), instead of the IEEE sqrt and
// divide sequences.
__device__ __forceinline__ double rsqrt64(double d2) {
  double y = __builtin_amdgcn_rsq(d2);
  double t = fma(-d2 * y, y, 1.0);
  y = fma(0.5 * y, t, y);
  t = fma(-d2 * y, y, 1.0);
  return fma(0.5 * y, t, y);
}

// sqrt(d2) for d2 > 0, where 1/sqrt(d2) is not needed (the rho passes): from the
// hardware reciprocal-square-root seed y = (1 + e) / sqrt(d2), g = d2 y and
// h = y / 2, two residual corrections g += (d2 - g^2) h leave a relative error
// of 1.5 e^3 (e below 2^-23: under 2^-67), so the result is within a few ulp,
// like d2 * rsqrt64(d2), in seven instructions instead of ten.
__device__ __forceinline__ double sqrt64(double d2) {
  const double y = __builtin_amdgcn_rsq(d2), h = 0.5 * y;
  double g = d2 * y;
  g = fma(fma(-g, g, d2), h, g);
  return fma(fma(-g, g, d2), h, g);
}

// ------------------------------------------------------------ spline eval

struct Seg {
  int m;
  double p;
};

// Segment and fraction of x in a table of n knots, clamped to the first and
// last segment (fraction 1 of the last one past the table's end).  p is
// capped at n - 1 before the conversion, which keeps any x -- the far point of
// a vacancy included -- in the converter's range and the fraction p - m at
// most 1, since m = n - 2 wherever p > n - 2.
__device__ __forceinline__ Seg seg_clamp(double p, int n) {
  p = fmin(fmax(p, 0.0), (double)(n - 1));
  int m = (int)p;
  m = min(m, n - 2);
  Seg s;
  s.p = p - (double)m;
  s.m = m;
  return s;
}
__device__ __forceinline__ Seg seg_f64(double x, double x0, double inv_dx, int n) {
  return seg_clamp((x - x0) * inv_dx, n);
}
// The rho passes' form: p = (x - x0) / dx from one FMA with the start folded
// into the addend, exactly the two-step rounding where x0 = 0 (the tables'
// usual start) and within an ulp of it otherwise; one instruction fewer per
// pair (the force pass keeps the two-step form: the folded constant's
// register made it spill).
__device__ __forceinline__ Seg seg_f64_fma(double x, double x0, double inv_dx, int n) {
  return seg_clamp(fma(x, inv_dx, -x0 * inv_dx), n);
}

struct SegS {
  uint32_t m;
  double q;
};

// Segment of s = r^2 >= 0 in the mixed tables.  p is capped below n_s - 1 so
// the conversion stays in range; p >= -s0 / ds > INT_MIN.  A p below 0 maps to
// segment 0 (p > -1) or, through the unsigned clamp, to segment n_s - 2, like
// every p past the table: neither segment is served by the mixed tables (their
// records are NaN-flagged), so q needs no clamping and is taken as the
// fractional part of p (exactly p - m wherever the segment is served).
__device__ __forceinline__ SegS seg_s(double s2, const TabMixed &t) {
  const double p = fmin(fma(s2, t.inv_ds, -t.s0 * t.inv_ds), (double)t.n_s - 1.5);
  SegS r;
  r.m = min((uint32_t)(int32_t)p, (uint32_t)(t.n_s - 2));
  r.q = __builtin_amdgcn_fract(p);
  return r;
}

// ------------------------------------------------------ lane-pair gathers

// v of the other lane of the pair (2m, 2m + 1): DPP quad permutation
// [1, 0, 3, 2].  Both lanes of the pair must be active.
__device__ __forceinline__ uint32_t pair_swap(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xF, 0xF, false);
}
__device__ __forceinline__ double pair_swap(double v) {
  const uint64_t b = __builtin_bit_cast(uint64_t, v);
  const uint64_t lo = pair_swap((uint32_t)b), hi = pair_swap((uint32_t)(b >> 32));
  return __builtin_bit_cast(double, (hi << 32) | lo);
}
__device__ __forceinline__ double2 pair_swap(double2 v) { return make_double2(pair_swap(v.x), pair_swap(v.y)); }
// v of the pair's even / odd lane, in both lanes: quad permutations
// [0, 0, 2, 2] / [1, 1, 3, 3] (a lane-dependent pick with no select)
__device__ __forceinline__ uint32_t pair_even(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xA0, 0xF, 0xF, false);
}
__device__ __forceinline__ uint32_t pair_odd(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xF5, 0xF, 0xF, false);
}

// A lane pair's two 32-byte records (two 16-byte halves each) gathered one
// record per instruction: the first load brings the even lane's record (its
// first half to the even lane, its second half to the odd lane), the second
// the odd lane's with the halves crossed (its second half to the even lane,
// its first half to the odd lane).  Each lane then holds its own first half
// in one of the two registers and the half its partner lacks in the other,
// so a select and one swap complete both records.  A gather touches the
// cache lines of 32 records instead of 64, and the two lanes of a pair read
// adjacent 16-byte halves (pairing lanes l and l + 32 instead, with
// v_permlane32_swap and no selects, measured 8-13 % slower: the partner
// lanes' reads no longer fall together).  Both lanes of every pair must be
// active.
struct PairRec32 {
  double2 e;  // the even lane's record: first half in the even lane, second half in the odd lane
  double2 o;  // the odd lane's record: second half in the even lane, first half in the odd lane
};
__device__ __forceinline__ PairRec32 pair_rec32_load(const double2 *__restrict__ base, uint32_t rec) {
  const uint32_t h = threadIdx.x & 1u;
  return {ld(base, 2 * pair_even(rec) + h), ld(base, 2 * pair_odd(rec) + 1 - h)};
}
// this lane's record: a = first half, b = second half
__device__ __forceinline__ void pair_rec32_halves(const PairRec32 &r, double2 &a, double2 &b) {
  const bool h = threadIdx.x & 1u;
  a = h ? r.o : r.e;
  b = pair_swap(h ? r.e : r.o);
}


// ------------------------------------------------------------- pair terms
//
// A trip's kUnroll pairs are evaluated in three phases — table indices, then
// every table load, then the arithmetic — with no control flow around a
// pair's loads, so all of a trip's loads are in flight together (a branch or
// an early use makes the compiler wait for each pair's loads before issuing
// the next pair's).  The mixed path evaluates every lane with the r^2-keyed
// tables (index clamped into the table); segments outside the range where
// they meet the accuracy bound hold NaN-flagged records, so a NaN result marks
// the few pairs that are redone with the FP64 r-spline in a separate,
// wave-uniformly guarded pass.

// FP64 r-keyed spline: rho_{tj}(r), r^2 = d2
__device__ __forceinline__ double rho_f64(const TabF64 &tf, int tj, double d2) {
  const double r = sqrt64(d2);
  const Seg sg = seg_f64_fma(r, tf.r0, tf.inv_dr, tf.n_r);
  return spl_value(tf.rho, (uint32_t)(tj * tf.n_r + sg.m), sg.p);
}

// rho'_t of FP64 derivative record rec at fraction p
__device__ __forceinline__ double rho_deriv_f64(const TabF64 &tf, uint32_t rec, double p) {
  const double2 a = ld(tf.rho.d01, rec);
  const double c = ld(tf.rho.d2, rec);
  return (a.x * p + a.y) * p + c;
}

// phi_ab at fraction p of a record from the derivative coefficients of its
// force record (e0 p + e1) p + e2 and the knot value c6: the spline's value
// coefficients are (e0 dr / 3, e1 dr / 2, e2 dr, phi(r_m)), so the value needs
// one 8-byte gather (the knot value, the last word of the value record)
// instead of the 32-byte value record.
__device__ __forceinline__ double phi_from_deriv(const TabF64 &tf, double c6, double p, double e0, double e1,
                                                 double e2) {
  const double t = fma(p, e0 * (1.0 / 3.0), 0.5 * e1);
  return fma(tf.dr * p, fma(p, t, e2), c6);
}
// the knot value phi(r_m) of record rp (the last word of its value record)
__device__ __forceinline__ double phi_knot(const TabF64 &tf, uint32_t rp) {
  return ld(reinterpret_cast<const double *>(tf.phi.val), 4 * rp + 3);
}

// fpair = -(df_i rho'_{tj} + df_j rho'_{ti} + phi'_{ti tj}) / r (the scalar that
// multiplies del = x_i - x_j) from the pair's force record rp at fraction p,
// rinv = 1/r; an unlike pair brings rho'_{ti} in rhoip_u (a like pair reuses
// rho'_{tj}).  phi receives phi_{ti tj}(r) when WANT_E and the (wave-uniform)
// entry counts the pair's energy.
// c0..c2: the record's three 16-byte pieces; c6 its knot value (phi_knot,
// read only when the entry counts the pair energy).
template <bool WANT_E>
__device__ __forceinline__ double fpair_f64_c(const TabF64 &tf, double2 c0, double2 c1, double2 c2, double c6,
                                              double p, double rinv, double dfi, double dfj, bool unlike,
                                              double rhoip_u, bool count_e, double &phi) {
  const double rhojp = (c0.x * p + c0.y) * p + c1.x;
  const double phip = (c1.y * p + c2.x) * p + c2.y;
  const double rhoip = unlike ? rhoip_u : rhojp;
  if (WANT_E && count_e) phi = phi_from_deriv(tf, c6, p, c1.y, c2.x, c2.y);
  return -(dfi * rhojp + dfj * rhoip + phip) * rinv;
}

template <bool WANT_E>
__device__ __forceinline__ double fpair_f64_rec(const TabF64 &tf, uint32_t rp, double p, double rinv, double dfi,
                                                double dfj, bool unlike, double rhoip_u, bool count_e, double &phi) {
  const double c6 = WANT_E && count_e ? phi_knot(tf, rp) : 0.0;
  return fpair_f64_c<WANT_E>(tf, ld(tf.fc, 3 * rp), ld(tf.fc, 3 * rp + 1), ld(tf.fc, 3 * rp + 2), c6, p, rinv, dfi,
                             dfj, unlike, rhoip_u, count_e, phi);
}

// One FP64 pair from its squared distance, rho'_{ti} loaded here (the mixed
// path's redo of a flagged pair)
template <bool WANT_E>
__device__ __forceinline__ double fpair_f64(const TabF64 &tf, int ti, int tj, double d2, double dfi, double dfj,
                                            bool count_e, double &phi) {
  const double rinv = rsqrt64(d2);
  const Seg sg = seg_f64(d2 * rinv, tf.r0, tf.inv_dr, tf.n_r);
  const uint32_t rp = __umul24((uint32_t)(ti * tf.nt + tj), (uint32_t)tf.n_r) + (uint32_t)sg.m;
  const bool unlike = ti != tj;
  double rhoip = 0.0;
  if (unlike) rhoip = rho_deriv_f64(tf, __umul24((uint32_t)ti, (uint32_t)tf.n_r) + (uint32_t)sg.m, sg.p);
  return fpair_f64_rec<WANT_E>(tf, rp, sg.p, rinv, dfi, dfj, unlike, rhoip, count_e, phi);
}

// ------------------------------------------------------ unlike-pair queue
//
// An unlike pair (ti != tj) also needs rho'_{ti}(r).  Gathering it on the
// unlike lanes of every stencil entry costs a full gather instruction per
// entry although few lanes use it: the load unit's time follows the
// instructions, not the active lanes.  In the FP64 force pass (two gathers
// per unlike pair, 24-byte records) a trip instead numbers its unlike
// in-cutoff pairs across its kUnroll entries (ballot + mbcnt) and writes
// them to the wave's LDS queue as (record, fraction); each active lane of
// the wave then evaluates one queued pair with one gather (more rounds only
// when more pairs are queued than lanes are active) and the owners read the
// values back.  A wave's LDS operations execute in order, so the phases
// need only compiler barriers.

__device__ __forceinline__ void lds_order() { asm volatile("" ::: "memory"); }

__device__ __forceinline__ uint32_t lanes_below(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// out[u] = ev(rec[u], x[u]) for every u with ul[u]; qr / qx: the wave's queue
template <int U, class Eval>
__device__ __forceinline__ void unlike_queue(uint32_t *qr, double *qx, const bool *ul, const uint32_t *rec,
                                             const double *x, double *out, Eval ev) {
  uint32_t slot[U], tot = 0;
#pragma unroll
  for (int u = 0; u < U; u++) {
    const uint64_t m = __ballot(ul[u]);
    slot[u] = tot + lanes_below(m);
    tot += (uint32_t)__popcll(m);
  }
  if (tot == 0) return;
  const uint64_t act = __ballot(true);
  const uint32_t nact = (uint32_t)__popcll(act), rank = lanes_below(act);
  for (uint32_t r0 = 0; r0 < tot; r0 += nact) {
#pragma unroll
    for (int u = 0; u < U; u++)
      if (ul[u] && slot[u] - r0 < nact) qr[slot[u] - r0] = rec[u], qx[slot[u] - r0] = x[u];
    lds_order();
    if (r0 + rank < tot) qx[rank] = ev(qr[rank], qx[rank]);
    lds_order();
#pragma unroll
    for (int u = 0; u < U; u++)
      if (ul[u] && slot[u] - r0 < nact) out[u] = qx[slot[u] - r0];
    lds_order();
  }
}

// Phased evaluation of a trip (mixed): d/ds tables, -(dpsi/dr)/r = -2 dpsi/ds
// (the force-side derivative records hold -2 dpsi/ds), no square root or
// divide; NaN where the pair needs the FP64 spline.  rho'_{ti} is gathered
// on the unlike lanes together with the trip's other loads: routed through
// the unlike-pair queue it measured slower here, its round trip adding to
// the phased loads' latency.
// row_i = ti * nt * n_s and ra_i = ti * n_s (the centre's rows).
template <bool WANT_E, int U>
__device__ __forceinline__ void trip_force_mixed(const TabMixed &tm, uint32_t ti, uint32_t row_i, uint32_t ra_i,
                                                 const uint32_t *tj, const double *d2, double dfi, const double *dfj,
                                                 const bool *count_e, double *fp, double *phi) {
  double q[U];
  uint32_t rp[U], ra[U];
#pragma unroll
  for (int u = 0; u < U; u++) {
    const SegS sg = seg_s(d2[u], tm);
    q[u] = sg.q;
    rp[u] = __umul24(tj[u], (uint32_t)tm.n_s) + row_i + sg.m;
    ra[u] = ra_i + sg.m;
  }
  // the 32-byte force records are gathered per lane pair
  PairRec32 fr[U];
  double2 ua[U], vb[U];
  float2 vc[U];
#pragma unroll
  for (int u = 0; u < U; u++) {
    fr[u] = pair_rec32_load(tm.fb, rp[u]);
    // rho'_{ti} only on unlike lanes (like pairs reuse rho'_{tj}), into
    // registers set first so that no copy, and no wait, is needed where the
    // lanes rejoin
    ua[u] = make_double2(0.0, 0.0);
    if (ti != tj[u]) ua[u] = ld(tm.rho.dp, ra[u]);
    if (WANT_E && count_e[u]) {
      vb[u] = ld(tm.phi.vb, rp[u]);
      vc[u] = ld(tm.phi.vc, rp[u]);
    }
  }
#pragma unroll
  for (int u = 0; u < U; u++) {
    const float qf = (float)q[u];
    double2 b0, b1;
    pair_rec32_halves(fr[u], b0, b1);
    const double rhojp = mix_deriv_rec(b0, q[u], qf);
    const double phip = mix_deriv_rec(b1, q[u], qf);
    const double rhoip = ti != tj[u] ? mix_deriv_rec(ua[u], q[u], qf) : rhojp;
    if (WANT_E && count_e[u]) {
      const float cc = fmaf(qf, vc[u].y, vc[u].x);
      phi[u] = fma(q[u], fma(q[u], (double)cc, vb[u].y), vb[u].x);
    }
    fp[u] = fma(dfi, rhojp, fma(dfj[u], rhoip, phip));
  }
}

// Phased rho of a trip (mixed): rho_{tj}(s), NaN where the pair needs the
// FP64 spline.  LEAN: the one-load quadratic records (passes whose rho feeds
// only df); else the cubic value records (two loads).
template <int U, bool LEAN>
__device__ __forceinline__ void trip_rho_mixed(const TabMixed &tm, const uint32_t *tj, const double *d2, double *v) {
  double q[U];
  uint32_t rec[U];
#pragma unroll
  for (int u = 0; u < U; u++) {
    const SegS sg = seg_s(d2[u], tm);
    q[u] = sg.q;
    rec[u] = __umul24(tj[u], (uint32_t)tm.n_s) + sg.m;
  }
  if constexpr (LEAN) {
    double2 lr[U];
#pragma unroll
    for (int u = 0; u < U; u++) lr[u] = ld(tm.rho_lean, rec[u]);
#pragma unroll
    for (int u = 0; u < U; u++) v[u] = mix_deriv_rec(lr[u], q[u], (float)q[u]);
    return;
  }
  double2 vb[U];
  float2 vc[U];
#pragma unroll
  for (int u = 0; u < U; u++) {
    vb[u] = ld(tm.rho.vb, rec[u]);
    vc[u] = ld(tm.rho.vc, rec[u]);
  }
#pragma unroll
  for (int u = 0; u < U; u++) {
    const float cc = fmaf((float)q[u], vc[u].y, vc[u].x);
    v[u] = fma(q[u], fma(q[u], (double)cc, vb[u].y), vb[u].x);
  }
}

// Phased rho of a trip (FP64)
template <int U>
__device__ __forceinline__ void trip_rho_f64(const TabF64 &tf, const uint32_t *tj, const double *d2, double *v) {
  double p[U];
  uint32_t rec[U];
#pragma unroll
  for (int u = 0; u < U; u++) {
    const double r = sqrt64(d2[u]);
    const Seg sg = seg_f64_fma(r, tf.r0, tf.inv_dr, tf.n_r);
    p[u] = sg.p;
    rec[u] = __umul24(tj[u], (uint32_t)tf.n_r) + (uint32_t)sg.m;
  }
  // the 32-byte value records are gathered per lane pair
  PairRec32 vr[U];
#pragma unroll
  for (int u = 0; u < U; u++) vr[u] = pair_rec32_load(tf.rho.val, rec[u]);
#pragma unroll
  for (int u = 0; u < U; u++) {
    double2 a, b;
    pair_rec32_halves(vr[u], a, b);
    v[u] = ((a.x * p[u] + a.y) * p[u] + b.x) * p[u] + b.y;
  }
}

// --------------------------------------------------------------- kernels
//
// One block per tile of kTileX x kTileY x kTileZ cells (both sub-lattices),
// one thread per site.  The block first stages the tile plus its stencil halo
// in LDS — (x, y) and (z, df) of every record as 16-byte arrays (z alone for
// the rho pass) plus the int8 types — with coalesced row loads, so a neighbour costs LDS
// reads instead of vector-memory instructions; only the spline tables are
// gathered from memory.  Stencil entries are LDS offsets (wave-uniform: a wave
// holds one parity), kUnroll per trip with the tables' loads of a trip issued
// together; pairs are masked without per-pair branches and a trip with no
// in-cutoff pair in the whole wave is skipped.  Records outside the
// ghost-extended lattice are staged as vacancies, and vacancies at a far
// point (kFarPos on every axis), so the mixed path screens them with the
// cutoff test alone.  The body starts with a barrier so a block that walks
// several work items can restage for its next tile.

// LDS tile layout (n records): lxy = (x, y) [n], then lzd = (z, df) [n] for
// the force pass or lz = z [n] for the rho pass, then the int8 types [n].
// A vacancy's staged coordinates: any occupied site lies beyond the cutoff
// of it (coordinates are within 1e6 A of the origin), so a staged record is
// occupied exactly when its x is not kFarPos.  Types are staged as table
// rows, clamped to [0, tmax] (a vacancy's -1 to 0), so the pair loops index
// the tables with them directly; the cutoff test screens the vacancies.
constexpr double kFarPos = 1e7;

template <bool WITH_DF>
__device__ __forceinline__ void stage_tile(const Geom &g, const Halo &h, const TileSite &t, int64_t lb,
                                           const double4 *__restrict__ P, const double *__restrict__ DF,
                                           double2 *lxy, void *lz_or_lzd, int8_t *lt, int tmax) {
  __syncthreads();
  if (lb < g.n_tiles) {
    for (int32_t idx = threadIdx.x; idx < h.n; idx += kBlock) {
      const int64_t src = tile_record_src(g, h, t, idx);
      double4 q = make_double4(0.0, 0.0, 0.0, -1.0);
      double f = 0.0;
      if (src >= 0) {
        q = P[src];
        if (WITH_DF) f = DF[src];
      }
      if (q.w < 0.0) q.x = q.y = q.z = kFarPos, f = 0.0;
      lxy[idx] = make_double2(q.x, q.y);
      if (WITH_DF)
        static_cast<double2 *>(lz_or_lzd)[idx] = make_double2(q.z, f);
      else
        static_cast<double *>(lz_or_lzd)[idx] = q.z;
      lt[idx] = (int8_t)min(max((int)q.w, 0), tmax);
    }
  }
  __syncthreads();
}

// LEAN (mixed, the MD steps): rho and df from the one-load quadratic records
// on every step, so a trajectory does not depend on which steps record
// energies; a step that records them also sums rho from the cubic value
// records, for the embedding energy alone.  The passes whose rho reaches the
// caller use the cubic value records.
template <int PREC, bool WANT_E, bool LEAN = false>
__device__ __forceinline__ void rho_df_block(const Geom &g, const Halo &h, int64_t lb,
                                             const int32_t *__restrict__ off, int32_t nk_even, int32_t nk_odd,
                                             const double4 *__restrict__ P, double *__restrict__ RHO,
                                             double *__restrict__ DF, const TabF64 &tf, const TabMixed &tm, double rc2,
                                             double *__restrict__ e_part, double2 *lds, double *red) {
  constexpr int kUnrollRho = rho_unroll<PREC, LEAN, WANT_E>();
  static_assert(kUnrollRho <= 8, "stencil padding too short for the unroll");
  const TileSite t = map_tile(g, h, lb);
  double2 *lxy = lds;
  double *lz = reinterpret_cast<double *>(lds + h.n);
  int8_t *lt = reinterpret_cast<int8_t *>(lz + h.n);
  stage_tile<false>(g, h, t, lb, P, nullptr, lxy, lz, lt, tf.nt - 1);
  double e_loc = 0.0;
  // every lane of a wave with an active site runs the stencil loop, pairs
  // masked on lanes without an occupied site (as in force_block)
  if (__ballot(t.active) != 0) {
    const double2 cxy = lxy[t.c];
    const double cz = lz[t.c];
    const bool live = t.active && cxy.x != kFarPos;
    const int t0 = lt[t.c];
    {
      const int32_t *o = off + t.s * kMaxStencil;
      const int nk = t.s ? nk_odd : nk_even;
      constexpr bool BOTH = PREC == 1 && LEAN && WANT_E;  // dynamics from lean records, energy from cubic ones
      double acc = 0.0, acc_e = 0.0;
      // vacancies sit at the far point, so the cutoff test screens them
      // (their staged type is row 0)
      const uint8_t *ltu = reinterpret_cast<const uint8_t *>(lt);
      // a trip's neighbour records, read from the LDS tile
      struct Nbr {
        double2 xy;
        double z;
        uint32_t t;
      };
      auto fetch = [&](const int32_t *oc, Nbr *nb) {
#pragma unroll
        for (int u = 0; u < kUnrollRho; u++) {
          const int32_t j = t.c + oc[u];
          nb[u].xy = lxy[j];
          nb[u].z = lz[j];
          nb[u].t = ltu[j];
        }
      };
      auto trip = [&](int k, bool tail, const Nbr *nb) {
        uint32_t tq[kUnrollRho];
        double d2[kUnrollRho];
        bool in[kUnrollRho], any = false;
#pragma unroll
        for (int u = 0; u < kUnrollRho; u++) {
          const double2 pxy = nb[u].xy;
          tq[u] = nb[u].t;
          const double dx = cxy.x - pxy.x, dy = cxy.y - pxy.y, dz = cz - nb[u].z;
          d2[u] = dx * dx + dy * dy + dz * dz;
          in[u] = live && d2[u] < rc2 && (!tail || k + u < nk);
          any |= in[u];
        }
        if (__ballot(any) == 0) return;
        double v[kUnrollRho], ve[kUnrollRho];
        if constexpr (PREC == 1) {
          trip_rho_mixed<kUnrollRho, LEAN>(tm, tq, d2, v);
          if constexpr (BOTH) trip_rho_mixed<kUnrollRho, false>(tm, tq, d2, ve);
          uint64_t redo = 0;  // lanes with a pair to redo (scalar masks)
#pragma unroll
          for (int u = 0; u < kUnrollRho; u++)
            redo |= __ballot(in[u] && (__builtin_isnan(v[u]) || (BOTH && __builtin_isnan(ve[u]))));
          if (redo != 0) {
#pragma unroll
            for (int u = 0; u < kUnrollRho; u++) {
              if (in[u] && __builtin_isnan(v[u])) v[u] = rho_f64(tf, (int)tq[u], d2[u]);
              if (BOTH && in[u] && __builtin_isnan(ve[u])) ve[u] = rho_f64(tf, (int)tq[u], d2[u]);
            }
          }
        } else {
          trip_rho_f64<kUnrollRho>(tf, tq, d2, v);
        }
        // FP64: every lane's value is finite (clamped segment of a finite
        // table), so the masked sum is one FMA with a 0 / 1 weight, equal to
        // the selected add bit for bit (acc is never -0); the mixed values
        // of lanes outside the cutoff may be NaN and are selected away
#pragma unroll
        for (int u = 0; u < kUnrollRho; u++) {
          if constexpr (PREC == 0)
            acc = fma(v[u], in[u] ? 1.0 : 0.0, acc);
          else
            acc += in[u] ? v[u] : 0.0;
          if (BOTH) acc_e += in[u] ? ve[u] : 0.0;
        }
      };
      // A trip's stencil offsets are read one trip ahead (rows are padded
      // far past nk + 2 kUnrollRho), so no trip waits on its scalar loads;
      // the next trip's neighbour records are read from LDS as a trip ends.
      // (Reading them before a trip's gathers, to overlap their latency,
      // measured 3 % slower in the FP64 pass.)
      const int nk4 = nk & ~(kUnrollRho - 1);
      int32_t oc[kUnrollRho];
#pragma unroll
      for (int u = 0; u < kUnrollRho; u++) oc[u] = o[u];
      Nbr cur[kUnrollRho];
      fetch(oc, cur);
      for (int k = 0; k < nk4; k += kUnrollRho) {
        int32_t on[kUnrollRho];
#pragma unroll
        for (int u = 0; u < kUnrollRho; u++) on[u] = o[k + kUnrollRho + u];
        trip(k, false, cur);
        fetch(on, cur);
      }
      if (nk4 < nk) trip(nk4, true, cur);
      if (live) {
        const uint32_t i = (uint32_t)t.i;
        RHO[i] = acc;
        const Seg se = seg_f64(acc, tf.rho0, tf.inv_drho, tf.n_rho);
        const uint32_t re = (uint32_t)(t0 * tf.n_rho + se.m);
        DF[i] = spl_deriv(tf.emb, re, se.p);
        if (BOTH) {
          const Seg sv = seg_f64(acc_e, tf.rho0, tf.inv_drho, tf.n_rho);
          e_loc = spl_value(tf.emb, (uint32_t)(t0 * tf.n_rho + sv.m), sv.p);
        } else if (WANT_E) {
          e_loc = spl_value(tf.emb, re, se.p);
        }
      }
    }
  }
  if (WANT_E) {
    const double b = block_sum(e_loc, red);
    if (threadIdx.x == 0) e_part[lb] = b;
  }
}

// local site number of local cell (xl, yl, zl), parity s: the device order of
// the local box alone (the half stencil's slot index)
__device__ __forceinline__ uint32_t local_q(const Geom &g, int32_t s, int32_t xl, int32_t yl, int32_t zl) {
  return ((uint32_t)(zl * g.by + yl) * 2u + (uint32_t)s) * (uint32_t)g.bxh + (uint32_t)xl;
}

// HALF: the forward pass of the Newton's-third-law half stencil (kernels_half.hip):
// off is the half-stencil image (forward rows first), nk_even / nk_odd the
// forward entry counts, every entry counts its pair energy once, and the pair
// force scalar of forward entry k is also stored in SL[k][q] (0 beyond the
// cutoff, for a vacant neighbour or a vacant centre).
template <int PREC, bool WANT_E, bool HALF = false>
__device__ __forceinline__ void force_block(const Geom &g, const Halo &h, int64_t lb, const int32_t *__restrict__ off,
                                            int32_t nk_even, int32_t nk_odd, const double4 *__restrict__ P,
                                            const double *__restrict__ DF, double *__restrict__ FX,
                                            double *__restrict__ FY, double *__restrict__ FZ, const TabF64 &tf,
                                            const TabMixed &tm, double rc2, double *__restrict__ e_part, double2 *lds,
                                            double *red, double *__restrict__ SL = nullptr) {
  const TileSite t = map_tile(g, h, lb);
  double2 *lxy = lds, *lzd = lds + h.n;
  int8_t *lt = reinterpret_cast<int8_t *>(lds + 2 * h.n);
  char *qw = reinterpret_cast<char *>(lds) + tile_queue_offset(h) + (threadIdx.x >> 6) * kSlotWaveBytes;
  uint32_t *qr = reinterpret_cast<uint32_t *>(qw);
  double *qx = reinterpret_cast<double *>(qw + kSlots * sizeof(uint32_t));
  stage_tile<true>(g, h, t, lb, P, DF, lxy, lzd, lt, tf.nt - 1);
  double e_loc = 0.0;
  // Every lane of a wave that holds an active site runs the stencil loop:
  // lanes without an occupied site (a vacancy, a cell past a ragged box edge)
  // take part with all their pairs masked, so the lane-pair record exchanges
  // always find both lanes of a pair.
  if (__ballot(t.active) != 0) {
    const double2 cxy = lxy[t.c], czd = lzd[t.c];
    const bool live = t.active && cxy.x != kFarPos;
    const int t0 = lt[t.c];
    const int nk = t.s ? nk_odd : nk_even;
    const size_t nl = (size_t)g.n_local;
    const uint32_t q = HALF ? local_q(g, t.s, t.x0 + t.tx, t.y0 + t.ty, t.z0 + t.tz) : 0u;
    {
      static_assert(!HALF || kHalfFwd == 0, "forward rows lead the half-stencil image");
      const int32_t *o = off + t.s * kMaxStencil;
      // entries that count the pair energy (each pair once for a symmetric stencil)
      const int32_t *ebits = off + kStencilBits + t.s * (kMaxStencil / 32);
      const double ew = HALF || off[kStencilHalfFlag] ? 1.0 : 0.5;
      const double dfi = czd.y;
      double ax = 0.0, ay = 0.0, az = 0.0, ep = 0.0;
      // vacancies sit at the far point (see rho_df_block)
      const uint8_t *ltu = reinterpret_cast<const uint8_t *>(lt);
      const uint32_t ti = (uint32_t)t0;
      const uint32_t ra_i = ti * (uint32_t)tm.n_s, row_i = ra_i * (uint32_t)tm.nt;
      auto trip = [&](int k, bool tail, const int32_t *oc) {
        uint32_t tq[kUnroll];
        double dx[kUnroll], dy[kUnroll], dz[kUnroll], d2[kUnroll], dfj[kUnroll];
        bool in[kUnroll], any = false;
        // entry k + u exists: tested on every FP64 trip (its loop's last trip
        // is the partial one) and on the mixed pass's tail trip, as a plain
        // scalar comparison (a select on the run-time tail flag was
        // materialised per lane)
        auto entry = [&](int u) { return PREC == 0 || tail ? k + u < nk : true; };
#pragma unroll
        for (int u = 0; u < kUnroll; u++) {
          const int32_t j = t.c + oc[u];
          const double2 pxy = lxy[j], pzd = lzd[j];
          tq[u] = ltu[j];
          dx[u] = cxy.x - pxy.x;
          dy[u] = cxy.y - pxy.y;
          dz[u] = czd.x - pzd.x;
          dfj[u] = pzd.y;
          d2[u] = dx[u] * dx[u] + dy[u] * dy[u] + dz[u] * dz[u];
          in[u] = live && d2[u] < rc2 && entry(u);
          any |= in[u];
        }
        if (__ballot(any) == 0) {
          if (HALF && t.active) {
#pragma unroll
            for (int u = 0; u < kUnroll; u++)
              if (entry(u)) SL[(size_t)(k + u) * nl + q] = 0.0;
          }
          return;
        }
        double fp[kUnroll], phi[kUnroll];
        bool ce[kUnroll];  // wave-uniform: this entry counts the pair energy
        const uint32_t hw = WANT_E && !HALF ? (uint32_t)ebits[k >> 5] : 0u;
#pragma unroll
        for (int u = 0; u < kUnroll; u++) {
          ce[u] = WANT_E && (HALF || ((hw >> ((k & 31) + u)) & 1u));
          phi[u] = 0.0;
        }
        if constexpr (PREC == 1) {
          trip_force_mixed<WANT_E, kUnroll>(tm, ti, row_i, ra_i, tq, d2, dfi, dfj, ce, fp, phi);
          uint64_t redo = 0;  // lanes with a pair to redo (scalar masks)
#pragma unroll
          for (int u = 0; u < kUnroll; u++) redo |= __ballot(in[u] && __builtin_isnan(fp[u]));
          if (redo != 0) {
#pragma unroll
            for (int u = 0; u < kUnroll; u++)
              if (in[u] && __builtin_isnan(fp[u]))
                fp[u] = fpair_f64<WANT_E>(tf, t0, (int)tq[u], d2[u], dfi, dfj[u], ce[u], phi[u]);
          }
        } else {
          // FP64: the trip's record numbers and fractions, rho'_{ti} of its
          // unlike pairs through the queue, then pair by pair (its large
          // tables are gathered from L2; this measured faster than phased loads)
          double rinv[kUnroll], pr[kUnroll], rhoq[kUnroll];
          uint32_t rp[kUnroll], ra[kUnroll];
          bool ul[kUnroll];
#pragma unroll
          for (int u = 0; u < kUnroll; u++) {
            rinv[u] = rsqrt64(d2[u]);
            const Seg sg = seg_f64(d2[u] * rinv[u], tf.r0, tf.inv_dr, tf.n_r);
            pr[u] = sg.p;
            rp[u] = __umul24(ti * (uint32_t)tf.nt + tq[u], (uint32_t)tf.n_r) + (uint32_t)sg.m;
            ra[u] = __umul24(ti, (uint32_t)tf.n_r) + (uint32_t)sg.m;
            ul[u] = in[u] && tq[u] != ti;
            rhoq[u] = 0.0;
          }
          // 48-byte force records (pieces 0-2) gathered per lane pair, each
          // gather touching the lines of 32 records in two of three loads.
          // Full stencil: the first load brings each lane its own piece 0,
          // the second pieces 1 and 2 of the even lane's record (to the even
          // and the odd lane), the third pieces 2 and 1 of the odd lane's; a
          // select and one swap complete pieces 1 and 2.  Half stencil: the
          // first load brings pieces 0 and 1 of the even lane's record, the
          // second its piece 2 and piece 0 of the odd lane's, the third the
          // odd lane's pieces 1 and 2, and a swap of piece 1 completes both
          // (one select more per record, but the forward pass, which holds
          // its slot stores, measured 4 % slower with the first form).  The
          // plain full-stencil pass issues them before the unlike queue, so
          // its round trips overlap theirs; the passes that hold more
          // registers (pair energies, the half stencil's slot stores) issue
          // them after (issued first they spill there).
          const uint32_t hl = threadIdx.x & 1u;
          double2 r0[kUnroll], r1[kUnroll], r2[kUnroll];
          double kv[kUnroll];  // knot values of the entries that count the pair energy
          auto gather_records = [&] {
#pragma unroll
            for (int u = 0; u < kUnroll; u++) {
              // issued with the records (read one by one where they are
              // used, each waited for in turn)
              if constexpr (WANT_E) {
                kv[u] = 0.0;
                if (ce[u]) kv[u] = phi_knot(tf, rp[u]);
              }
              const uint32_t re = pair_even(rp[u]), ro = pair_odd(rp[u]);
              if constexpr (!HALF) {
                // (the 24-bit multiply keeps 3 rp off the quarter-rate 32-bit
                // one; r1 and r2 stay in the form that keeps every load on a
                // scalar base plus a 32-bit lane offset)
                r0[u] = ld(tf.fc, __umul24(rp[u], 3u));
                r1[u] = ld(tf.fc, 3 * re + 1 + hl);
                r2[u] = ld(tf.fc, 3 * ro + 2 - hl);
              } else {
                r0[u] = ld(tf.fc, 3 * re + hl);
                r1[u] = ld(tf.fc, hl ? 3 * ro : 3 * re + 2);
                r2[u] = ld(tf.fc, 3 * ro + 1 + hl);
              }
            }
          };
          constexpr bool kEarly = !HALF && !WANT_E;
          if (kEarly) gather_records();
          unlike_queue<kUnroll>(qr, qx, ul, ra, pr, rhoq,
                                [&](uint32_t rec, double x) { return rho_deriv_f64(tf, rec, x); });
          if (!kEarly) gather_records();
#pragma unroll
          for (int u = 0; u < kUnroll; u++) {
            double2 c0, c1, c2;
            if constexpr (!HALF) {
              c0 = r0[u];
              c1 = hl ? r2[u] : r1[u];
              c2 = pair_swap(hl ? r1[u] : r2[u]);
            } else {
              c0 = hl ? r1[u] : r0[u];
              c1 = pair_swap(hl ? r0[u] : r2[u]);
              c2 = hl ? r2[u] : r1[u];
            }
            fp[u] = fpair_f64_c<WANT_E>(tf, c0, c1, c2, WANT_E ? kv[u] : 0.0, pr[u], rinv[u], dfi, dfj[u], tq[u] != ti,
                                        rhoq[u], ce[u], phi[u]);
          }
        }
#pragma unroll
        for (int u = 0; u < kUnroll; u++) {
          const double f = in[u] ? fp[u] : 0.0;
          ax = fma(dx[u], f, ax);
          ay = fma(dy[u], f, ay);
          az = fma(dz[u], f, az);
          if (WANT_E) ep += in[u] ? phi[u] : 0.0;
          if (HALF && t.active && entry(u)) SL[(size_t)(k + u) * nl + q] = f;
        }
      };
      // a trip's stencil offsets are read one trip ahead (rows are padded
      // far past nk + 2 kUnroll), so no trip waits on its scalar loads
      // FP64: one loop over every trip, the last one partial when nk is not
      // a multiple of kUnroll (its entry test is a scalar compare), and the
      // site's index recomputed from the thread id after the loop (the empty
      // asm keeps the compiler from reusing map_tile's): a separate copy of
      // the partial trip and the index held across the loop spilled to
      // scratch there.  The mixed pass keeps the separate partial trip
      // (folded it measured 2.5 % slower).
      int32_t oc[kUnroll];
#pragma unroll
      for (int u = 0; u < kUnroll; u++) oc[u] = o[u];
      const int nk_loop = PREC == 0 ? nk : nk & ~(kUnroll - 1);
      for (int k = 0; k < nk_loop; k += kUnroll) {
        int32_t on[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; u++) on[u] = o[k + kUnroll + u];
        trip(k, PREC == 0 && k + kUnroll > nk, oc);
#pragma unroll
        for (int u = 0; u < kUnroll; u++) oc[u] = on[u];
      }
      if (nk_loop < nk) trip(nk_loop, true, oc);
      if (live) {
        uint32_t i = (uint32_t)t.i;
        if constexpr (PREC == 0) {
          uint32_t tid = threadIdx.x;
          asm volatile("" : "+v"(tid));
          i = (uint32_t)tile_site_index(g, t, tid);
        }
        FX[i] = ax;
        FY[i] = ay;
        FZ[i] = az;
        if (WANT_E) e_loc = ew * ep;
      }
    }
  }
  if (WANT_E) {
    const double b = block_sum(e_loc, red);
    if (threadIdx.x == 0) e_part[lb] = b;
  }
}

}  // namespace mmd

#endif
