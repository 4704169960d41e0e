// Population scoring kernel for gfx950 (CDNA4, wave64).
//
// One workgroup (4 x wave64) scores one pose (or one receptor range of a
// pose): it rigidly transforms the ligand in FP64 (structure.rs:80-131
// operation order), keeps K ligand atoms per lane in registers as fixed-point
// coordinates, and walks the receptor.  Receptor atoms are wave-uniform, so
// they are read with scalar loads into SGPRs and every pair costs VALU work
// only; the receptor (16 B/atom) stays resident in the scalar cache / L2
// across the whole launch.
//
// Per pair (fitness.rs:57-105,122-184,284-302): soft-core LJ, Coulomb with
// eps = r, and the two desolvation burial counts.  Coordinates are integers in
// units of 2^-qexp A (qexp up to 24), so a pair's coordinate difference is
// exact and only its FP32 conversion rounds: the squared distance carries an
// error relative to the distance itself, not to the size of the frame.  Pair
// math is FP32.  Every cutoff decision (d > 1, d < 8, d < 12, d < 15) is taken
// in FP32 against a guard band; a pair inside a band is rare, the wave records
// it as an event in LDS and, once per ligand chunk, decides it again in FP64
// from the original coordinates with the reference's operation order and
// applies the difference, so every decision matches the FP64 reference.  (A
// wave whose event buffer overflows flags its pose, and the exact variant of
// the kernel, launched right after, rescores that pose deciding such pairs
// inline.)  Receptor charges are exact integers in units of 1e-4 e (OPLS
// charges), 1/d^2 is a Newton-refined reciprocal, Coulomb terms are summed
// per lane in FP32 over a group of 4 receptor atoms and then in FP64,
// soft-core LJ terms in FP32 per 64-slot tile and class and then in FP64;
// the cross-lane and cross-wave reductions are FP64.  Desolvation burial
// counts are integers; the per-atom desolvation and AIR restraint terms and
// the weighted fitness are computed in FP64 in the reference's order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>

#include "common.h"
#include "internal.h"

// The fast scoring kernels run at 4 waves per SIMD (at most 128 VGPRs): the
// pair loop is VALU-bound and needs the fourth wave to hide its latencies.

namespace gdock {

// Complex descriptors and receptor tables are read-only for a whole launch;
// reading them through the constant address space lets the compiler use
// scalar loads for every wave-uniform access.
#define GDOCK_AS4 __attribute__((address_space(4)))
using CxRef = const DevComplex GDOCK_AS4&;

// Receptor atom: fixed-point x, y, z (origin frame) and its charge in units
// of 1 / q_scale.
struct RecAtom { int32_t x, y, z; float w; };

__device__ __forceinline__ RecAtom rec_atom(const int32_t GDOCK_AS4* rf, int j) {
  return RecAtom{rf[4 * j], rf[4 * j + 1], rf[4 * j + 2], __int_as_float(rf[4 * j + 3])};
}

struct PoseCoef {
  double ca, sa, cb, sb, cg, sg;
  double m01, m02, m11, m12;  // the loop-invariant sums inside structure.rs:96-101
  double dx, dy, dz;
};

// Every field comes from the pose record (the host forms the products), so a
// workgroup reads its pose with scalar loads and does no FP64 work for it.
__device__ __forceinline__ PoseCoef load_pose(const double* __restrict__ p) {
  PoseCoef c;
  c.dx = p[3]; c.dy = p[4]; c.dz = p[5];
  c.ca = p[6]; c.sa = p[7]; c.cb = p[8]; c.sb = p[9]; c.cg = p[10]; c.sg = p[11];
  c.m01 = p[12]; c.m02 = p[13]; c.m11 = p[14]; c.m12 = p[15];
  return c;
}

// FP64 pose of ligand atom l: rotate about the COM, then displace
// (chromosome.rs:79-83 -> structure.rs:80-131).  This file is compiled with
// -ffp-contract=off, so every product and sum rounds as in the reference.
__device__ __forceinline__ void ligand_atom(CxRef cx, const PoseCoef& pc, const double* __restrict__ coords, int l,
                                            double& X, double& Y, double& Z) {
  if (coords) {  // this pose's coordinates, n_lig x 3 in input order
    const double* p = coords + static_cast<size_t>(cx.lig_orig[l]) * 3;
    X = p[0]; Y = p[1]; Z = p[2];
    return;
  }
  const double x = cx.lig_x[l] - cx.com[0];
  const double y = cx.lig_y[l] - cx.com[1];
  const double z = cx.lig_z[l] - cx.com[2];
  const double nx = ((x * pc.ca) * pc.cb + y * pc.m01) + z * pc.m02;
  const double ny = ((x * pc.sa) * pc.cb + y * pc.m11) + z * pc.m12;
  const double nz = ((-x) * pc.sb + (y * pc.cb) * pc.sg) + (z * pc.cb) * pc.cg;
  X = (nx + cx.com[0]) + pc.dx;
  Y = (ny + cx.com[1]) + pc.dy;
  Z = (nz + cx.com[2]) + pc.dz;
}

__device__ __forceinline__ float uniform(float v) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}

// Fixed-point value of a posed FP64 coordinate, the same operations as the
// host applies to the receptor: (p - origin) 2^qexp, clamped, rounded to
// nearest even.
__device__ __forceinline__ int32_t to_fixed(double p, double o, double scale) {
  double v = (p - o) * scale;
  v = fmin(fmax(v, -static_cast<double>(kFixedClamp)), static_cast<double>(kFixedClamp));
  return static_cast<int32_t>(rint(v));
}

// FP32 difference of two fixed-point coordinates (exact integer difference,
// one rounding).  Real pairs never overflow (|values| <= kFixedClamp); the
// unsigned form keeps the padding sentinel's wrap-around defined.
__device__ __forceinline__ float fixed_diff(int32_t a, int32_t b) {
  return static_cast<float>(static_cast<int32_t>(static_cast<uint32_t>(a) - static_cast<uint32_t>(b)));
}

// Squared distance in units^2.
__device__ __forceinline__ float pair_s(const RecAtom& r, int32_t x, int32_t y, int32_t z) {
  const float dx = fixed_diff(r.x, x), dy = fixed_diff(r.y, y), dz = fixed_diff(r.z, z);
  return __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
}

// The FP32 cutoff decisions on s (units^2) against the complex's guard bands,
// kept as integers: for a non-negative float, its bit pattern orders like its
// value.  c_T = bits(lo_T) + 1, so bits(s) - c_T borrows exactly when s <=
// lo_T (the FP32 decision), and the same difference, as an unsigned number
// below w, puts s inside the band of T.  The four bands share one width.
struct Bands {
  float lo1, lo64, lo144, lo225;
  uint32_t c1, c64, c144, c225, w;
};

__device__ __forceinline__ Bands make_bands(CxRef cx) {
  Bands B;
  B.c1 = cx.band_c[0]; B.c64 = cx.band_c[1]; B.c144 = cx.band_c[2]; B.c225 = cx.band_c[3];
  B.w = cx.band_w;
  B.lo1 = __uint_as_float(B.c1 - 1u); B.lo64 = __uint_as_float(B.c64 - 1u);
  B.lo144 = __uint_as_float(B.c144 - 1u); B.lo225 = __uint_as_float(B.c225 - 1u);
  return B;
}

// Ligand atom held in slot k of this lane for the chunk starting at `base`:
// wave w holds the K consecutive blocks of kLigBlock atoms at base + w * 64 K.
// Each block is spatially compact, and the complex stores its blocks in
// bit-reversed order, so any K consecutive blocks come from all over the
// ligand: every wave holds buried and exposed parts alike and the waves' pair
// work is even (spatially consecutive blocks per wave made the wave with the
// buried region the pose's critical path, 10% slower).  Consecutive slots
// keep the K addresses of a lane within load-offset range of one base.
template <int K>
__device__ __forceinline__ int lig_index(int base, int k) {
  return base + static_cast<int>(threadIdx.x >> 6) * (kLigBlock * K) + k * kLigBlock +
         static_cast<int>(threadIdx.x & 63);
}

// Axis-aligned boxes of a wave's K ligand blocks (FP32 A, origin frame),
// kept in LDS per block as the block's box and the boxes of its four
// kLigSub-lane quarters (8 floats each), and the test that a receptor atom is
// farther from a box than every cutoff.  A box with no atom is the single
// point (kFarBox, kFarBox, kFarBox): every distance to it overflows to +inf.
constexpr int kBoxFloats = 8 * (1 + kLigBlock / kLigSub);
constexpr float kFarBox = 1.0e30f;

// The distance from a point to any point of a box is at least the
// point-to-box distance, whose axis components are the point minus its clamp
// to the box (v_med3_f32; the same values, up to sign, as max(lo - x, x - hi,
// 0)); ten VALU operations.  FP32 rounding of it (and of the fixed-point to A
// conversions) is far below the 0.01 A margin of kCullReach over 15 A.
__device__ __forceinline__ bool beyond_reach(float rx, float ry, float rz, const float* __restrict__ b) {
  const float dx = rx - __builtin_amdgcn_fmed3f(rx, b[0], b[3]);
  const float dy = ry - __builtin_amdgcn_fmed3f(ry, b[1], b[4]);
  const float dz = rz - __builtin_amdgcn_fmed3f(rz, b[2], b[5]);
  return __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx)) > kCullReach * kCullReach;
}

// The same test for every point of the box [lo, hi]: for a point inside it
// each difference below is at most the point's magnitude in beyond_reach
// (FP32 subtraction, products and fma are monotone), so a box beyond reach
// implies every point in it is, as beyond_reach would find it.
__device__ __forceinline__ bool box_beyond_reach(float4 lo, float4 hi, const float* __restrict__ b) {
  const float dx = fmaxf(fmaxf(b[0] - hi.x, lo.x - b[3]), 0.0f);
  const float dy = fmaxf(fmaxf(b[1] - hi.y, lo.y - b[4]), 0.0f);
  const float dz = fmaxf(fmaxf(b[2] - hi.z, lo.z - b[5]), 0.0f);
  return __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx)) > kCullReach * kCullReach;
}

// Lane t's receptor slot g0 + t of the tile [g0, g1) (slot g0 past its end).
__device__ __forceinline__ int4 tile_slot(CxRef cx, int g0, int g1) {
  const int lane = threadIdx.x & 63;
  return cx.rec_fixed[g0 + lane < g1 ? g0 + lane : g0];
}

// Receptor atoms of the tile [g0, g1) against the wave's ligand blocks: bit
// t of keep[k] is set when slot g0 + t can reach block k (lane t tests its
// slot against every block; padding slots sit far from every block), and,
// for a block some slot reaches, one of the block's quarters.  Blocks whose
// bit in kmask is clear are beyond reach of every sub-tile box of the tile
// (so of every slot in it) and are not tested.  A culled pair would only have
// added exact zeros, so the tests decide work, not results.  r: the lane's
// slot of the tile, as tile_slot loads it.
template <int K>
__device__ __forceinline__ void atoms_in_reach(CxRef cx, float unit, int g0, int g1, const float* __restrict__ boxes,
                                               const int4 r, uint32_t kmask, uint64_t (&keep)[K]) {
  const int lane = threadIdx.x & 63;
  const bool valid = g0 + lane < g1;
  const float rx = static_cast<float>(r.x) * unit, ry = static_cast<float>(r.y) * unit,
              rz = static_cast<float>(r.z) * unit;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    keep[k] = 0ull;
    if (((kmask >> k) & 1u) == 0u) continue;
    const float* b = boxes + kBoxFloats * k;
    keep[k] = __builtin_amdgcn_ballot_w64(valid && !beyond_reach(rx, ry, rz, b));
    if (keep[k] != 0ull)
      keep[k] &= __builtin_amdgcn_ballot_w64(!beyond_reach(rx, ry, rz, b + 8) || !beyond_reach(rx, ry, rz, b + 16) ||
                                             !beyond_reach(rx, ry, rz, b + 24) || !beyond_reach(rx, ry, rz, b + 32));
  }
}

// Minimum and maximum over the 16-lane quarter of the wave holding this lane.
__device__ __forceinline__ int32_t quarter_min(int32_t v) {
#pragma unroll
  for (int off = 8; off > 0; off >>= 1) v = min(v, __shfl_xor(v, off));
  return v;
}

__device__ __forceinline__ int32_t quarter_max(int32_t v) {
#pragma unroll
  for (int off = 8; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off));
  return v;
}

template <int K>
struct LaneAtoms {
  // x, y, z: fixed point.  acv: sum of rho (rho - 2) over the current
  // receptor class within the tile (times eps when folded into FP64); ace:
  // sum of q_r / s over the current group of kRecGroup receptor slots.
  int32_t x[K], y[K], z[K];
  float eps[K], ir2[K], acv[K], ace[K];
  int32_t cnt[K];
  double kq[K];  // the atom's Coulomb factor (lig_kq), loaded once per chunk
};

// Soft-core LJ (fitness.rs:57-70) and the Coulomb factor from s, the squared
// distance in units^2.  The pair's vdw term is eps (rho^2 - 2 rho), rho = R^6
// / (d^6 + R^6/2) = 1 / (t^3 + 1/2), t = d^2/R^2 = s ir2 (ir2: the class
// factor q^2 / R^2 of the lane's ligand atom, q the fixed-point unit); u =
// rho - 2.  rho and 1/s come from two reciprocals.  On gfx950 a v_rcp_f32
// costs two to three FP32 VALU issue slots (measured), so this form times
// the same as deriving both from one reciprocal of (t^3 + 1/2) d^2 (three
// more products); it needs no rescaling to keep that product normal.  rho
// lies in (0, 2] for every pair and t^3 stays finite even for padding lanes
// (t < 1e10).  1/s is refined by one Newton step (e = 1 - s y,
// y + y e, two FMAs): v_rcp_f32 is within 1 ulp but biased (-0.07 ulp on
// average), and over ~10^6 Coulomb terms of one sign a bias adds up where
// rounding errors cancel; after the step the reciprocal is correctly rounded
// but for rare half-ulp cases.  s = 0 gives rho = 2 and u = 0, a zero vdw
// term (the reference's term is 0 there too), and a non-finite 1/s that the
// elec mask never selects (d <= 1 A).
__device__ __forceinline__ void pair_terms(float s, float ir2, float& rho, float& u, float& is) {
  const float t = s * ir2;
  rho = __builtin_amdgcn_rcpf(__builtin_fmaf(t * t, t, 0.5f));
  u = rho - 2.0f;
  const float y = __builtin_amdgcn_rcpf(s);
  is = __builtin_fmaf(y, __builtin_fmaf(-s, y, 1.0f), y);
}

// The vdw term of a pair as the pair loop counts it: eps (u rho), in FP64
// from the FP32 factors.
__device__ __forceinline__ double vdw_term(float eps, float u, float rho) {
  return static_cast<double>(eps) * (static_cast<double>(u) * static_cast<double>(rho));
}

// s (units^2) inside one of the four guard bands (the pair loop's test).
__device__ __forceinline__ bool in_band(float s, const Bands& B) {
  const uint32_t b = __float_as_uint(s);
  return min(min(b - B.c1, b - B.c64), min(b - B.c144, b - B.c225)) < B.w;
}

// Per-lane Lennard-Jones factors of ligand atom l against receptor class c:
// q^2 / R^2 2^-16 (q: the fixed-point unit, R = rmin2_r + rmin2_l) and
// sqrt(eps_r eps_l) from FP64 (fitness.rs:58-63), rounded once, tabulated
// per complex (one coalesced 8-byte load per lane).
__device__ __forceinline__ void class_factors(CxRef cx, int c, int l, float& ir2, float& eps) {
  const float2 f = cx.lig_class[static_cast<size_t>(c) * cx.n_lig + l];
  ir2 = f.x;
  eps = f.y;
}

// The reference's decisions for one pair, in FP64 on the pose's coordinates
// (px, py, pz: ligand atom l after the pose transform).
__device__ __forceinline__ void exact_flags_at(CxRef cx, double px, double py, double pz, int l, int j, int rser,
                                               bool& x64, bool& x144, bool& x225) {
  const double ex = cx.rec_x[j] - px, ey = cx.rec_y[j] - py, ez = cx.rec_z[j] - pz;
  const double e2 = ex * ex + ey * ey + ez * ez;
  x64 = e2 <= kD2DesolvIncl;
  x144 = e2 > 0.0 && e2 <= kD2VdwIncl && cx.lig_serial[l] != rser;
  x225 = e2 > kD2ElecMinExcl && e2 <= kD2ElecMaxIncl;
}

__device__ __forceinline__ void exact_flags(CxRef cx, const PoseCoef& pc, const double* __restrict__ coords, int l,
                                            int j, int rser, bool& x64, bool& x144, bool& x225) {
  double px, py, pz;
  ligand_atom(cx, pc, coords, l, px, py, pz);
  exact_flags_at(cx, px, py, pz, l, j, rser, x64, x144, x225);
}

struct SettleEvent { int32_t j, pad; uint64_t mask; };  // first receptor slot of a group, lanes
constexpr int kEventsPerWave = 64;

// Fast path of one receptor atom against the lane's ligand atom in slot k;
// returns this wave's burial count of the receptor atom from that slot.
// bmin keeps the least of the pair's four distances above the band floors
// (bits(s) - c_T, unsigned) over the lane's pairs of the receptor group: the
// group holds a band pair for the lane when it ends below the band width, one
// compare per group instead of one per pair.
template <int K>
__device__ __forceinline__ int32_t pair_fast(const RecAtom r, const Bands& B, LaneAtoms<K>& a, int k,
                                             uint32_t& bmin) {
  const float s = pair_s(r, a.x[k], a.y[k], a.z[k]);
  // FP32 decisions below each band, as the borrows of bits(s) - c_T (pairs
  // of equal serial are included in the vdw sum and taken out afterwards, see
  // collision_correction), and the band test on the same differences.
  const uint32_t bd = __float_as_uint(s);
  uint32_t q1, q64, q144, q225;
  const bool le1 = __builtin_sub_overflow(bd, B.c1, &q1);
  // d <= 8 A as a lane mask (the borrow): it is the carry-in that counts the
  // ligand atom's burial and, popcounted, the receptor atom's.  Written out
  // because the compiler would first turn the borrow into a 0/1 value to
  // take its ballot.  The s_nop gives the 2 wait states a VALU-written SGPR
  // needs before a VALU reads it.
  uint64_t md, carry;
  asm("v_sub_co_u32_e64 %0, %1, %4, %5\n\t"
      "s_nop 1\n\t"
      "v_addc_co_u32_e64 %2, %3, 0, %2, %1"
      : "=&v"(q64), "=&s"(md), "+v"(a.cnt[k]), "=&s"(carry)
      : "v"(bd), "s"(B.c64));
  const bool mv = __builtin_sub_overflow(bd, B.c144, &q144);
  const bool me0 = __builtin_sub_overflow(bd, B.c225, &q225);
  bmin = min(min(bmin, q1), q64);
  bmin = min(min(bmin, q144), q225);
  const bool me = me0 && !le1;
  float rho, u, id2;
  pair_terms(s, a.ir2[k], rho, u, id2);
  a.acv[k] = __builtin_fmaf(mv ? u : 0.0f, rho, a.acv[k]);
  a.ace[k] = __builtin_fmaf(me ? id2 : 0.0f, r.w, a.ace[k]);
  const int32_t cj = __popcll(md);
  // Pin this pair's results here (an empty asm, no instruction): otherwise
  // the compiler sinks the math below the group's rare band-pair branch and
  // keeps every pair's lane masks alive across it, spilling SGPRs.
  asm volatile("" : "+v"(a.acv[k]), "+v"(a.ace[k]), "+v"(a.cnt[k]), "+v"(bmin));
  return cj;
}

// The fast path's FP32 decisions of a pair from its s: d <= 8, within the
// vdw cutoff (d <= 12), within the elec cutoffs (1 < d <= 15).
__device__ __forceinline__ void fast_flags(float s, const Bands& B, bool& f64, bool& f144, bool& f225) {
  f64 = s <= B.lo64;
  f144 = s <= B.lo144;
  f225 = s <= B.lo225 && !(s <= B.lo1);
}

// Band pairs of one receptor atom: recorded as events (fast kernel) or
// settled on the spot in FP64 (exact kernel).
template <int K, bool EXACT>
__device__ __forceinline__ void band_pairs(CxRef cx, const PoseCoef& pc, const double* __restrict__ coords,
                                           int base, int j, const RecAtom r, const Bands& B,
                                           LaneAtoms<K>& a, double& sum_e, SettleEvent* __restrict__ ev, int& nev,
                                           int& overflow, int32_t* __restrict__ s_rcnt, int r0) {
  const int lane = threadIdx.x & 63;
  const int2 m = cx.rec_meta[j];
  const int rser = m.y ? m.x : INT_MIN;  // INT_MIN: no ligand atom shares this receptor atom's serial
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const float s = pair_s(r, a.x[k], a.y[k], a.z[k]);
    const int l = lig_index<K>(base, k);
    const uint64_t am = __builtin_amdgcn_ballot_w64(in_band(s, B) && l < cx.n_lig);
    if (am == 0ull) continue;
    if (!EXACT) {  // settled once per ligand chunk
      if (nev < kEventsPerWave) {
        if (lane == 0) ev[nev] = SettleEvent{j, k, am};
        ++nev;
      } else {
        overflow = 1;
      }
      continue;
    }
    if ((am >> lane) & 1ull) {
      bool x64, x144, x225, f64, f144, f225;
      exact_flags(cx, pc, coords, l, j, rser, x64, x144, x225);
      fast_flags(s, B, f64, f144, f225);
      float rho, u, id2;
      pair_terms(s, a.ir2[k], rho, u, id2);
      if (x144 != f144) a.acv[k] = __builtin_fmaf(x144 ? u : -u, rho, a.acv[k]);
      // the elec difference goes straight to FP64: the lane's FP32 partial
      // may already have been folded
      if (x225 != f225) sum_e += static_cast<double>(x225 ? id2 : -id2) * static_cast<double>(r.w) * cx.lig_kq[l];
      const int dc = static_cast<int>(x64) - static_cast<int>(f64);
      if (dc) {
        a.cnt[k] += dc;
        atomicAdd(&s_rcnt[j - r0], dc);
      }
    }
  }
}

// Settles the wave's recorded events.  An event names a group of kRecGroup
// receptor atoms and the lanes that saw a band pair in it; each band pair is
// decided in FP64 and the difference to the FP32 decisions is added to the
// FP64 sums, the lane's burial count and the receptor atom's count.  The
// ligand slot is a loop variable here, so its fixed-point coordinates are
// rebuilt from the FP64 transform (the same operations that filled the
// lane's registers) and its burial-count changes are carried as 16-bit fields
// of two packed integers: the lane's register arrays are only ever indexed by
// constants and stay in registers.  s_rcnt is indexed from the part's first
// receptor slot r0.
template <int K>
__device__ __forceinline__ void settle_events(CxRef cx, const double* __restrict__ pose_rec,
                                              const double* __restrict__ coords, int base, int r0, const Bands& B,
                                              const SettleEvent* __restrict__ ev, int nev, LaneAtoms<K>& a,
                                              double& sum_v, double& sum_e, int32_t* __restrict__ s_rcnt) {
  const int lane = threadIdx.x & 63;
  const int32_t GDOCK_AS4* cstart = (const int32_t GDOCK_AS4*)cx.class_start;
  const int32_t GDOCK_AS4* rf = (const int32_t GDOCK_AS4*)cx.rec_fixed;
  const double scale = ldexp(1.0, cx.qexp);
  PoseCoef pc{};
  if (!coords) pc = load_pose(pose_rec);
  int64_t dpack0 = 0, dpack1 = 0;  // slot k's burial-count change in bits [16 (k % 4), +16)
  for (int e = 0; e < nev; ++e) {
    const SettleEvent E = ev[e];
    if (!((E.mask >> lane) & 1ull)) continue;
    int c = 0;
    while (cstart[c + 1] <= E.j) ++c;  // groups never straddle classes
#pragma unroll 1
    for (int k = 0; k < K; ++k) {  // one copy of the FP64 path
      const int l = lig_index<K>(base, k);
      if (l >= cx.n_lig) break;
      double px, py, pz;
      ligand_atom(cx, pc, coords, l, px, py, pz);
      const int32_t lx = to_fixed(px, cx.origin[0], scale), ly = to_fixed(py, cx.origin[1], scale),
                    lz = to_fixed(pz, cx.origin[2], scale);
      bool have_factors = false;
      float ir2 = 0.0f, eps = 0.0f;
      for (int j = E.j; j < E.j + kRecGroup; ++j) {
        const RecAtom r = rec_atom(rf, j);
        if (r.x == kFixedPad) continue;  // padding slot: never evaluated
        const float s = pair_s(r, lx, ly, lz);
        if (!in_band(s, B)) continue;
        const int2 m = cx.rec_meta[j];
        const int rser = m.y ? m.x : INT_MIN;
        if (!have_factors) {
          class_factors(cx, c, l, ir2, eps);
          have_factors = true;
        }
        bool x64, x144, x225, f64, f144, f225;
        exact_flags_at(cx, px, py, pz, l, j, rser, x64, x144, x225);
        fast_flags(s, B, f64, f144, f225);
        float rho, u, id2;
        pair_terms(s, ir2, rho, u, id2);
        if (x144 != f144) sum_v += x144 ? vdw_term(eps, u, rho) : -vdw_term(eps, u, rho);
        if (x225 != f225) sum_e += static_cast<double>(x225 ? id2 : -id2) * static_cast<double>(r.w) * cx.lig_kq[l];
        const int dc = static_cast<int>(x64) - static_cast<int>(f64);
        if (dc) {
          const int64_t d = static_cast<int64_t>(dc) << (16 * (k & 3));
          if (k < 4) dpack0 += d; else dpack1 += d;
          atomicAdd(&s_rcnt[j - r0], dc);
        }
      }
    }
  }
#pragma unroll
  for (int k = 0; k < K; ++k) {  // unpack: fields are signed, so peel them from the bottom
    int64_t& p = k < 4 ? dpack0 : dpack1;
    const int32_t v = static_cast<int16_t>(p & 0xffff);
    p = (p - v) >> 16;
    a.cnt[k] += v;
  }
}

// The receptor slots a workgroup covers for one pose.  A pose of a large
// complex can be split across `parts` workgroups by receptor range: each part
// adds its ligand burial counts into the pose's shared count table and leaves
// its {vdw, elec, receptor desolvation} sums; the last part to finish
// combines them in part order, adds the ligand desolvation from the summed
// counts, finishes the pose and clears the shared state for the next launch.
struct PosePart {
  int r0, r1;          // receptor slots [r0, r1)
  int part, parts;     // parts == 1: the whole pose in this workgroup
  int32_t* lig_cnt;    // parts > 1: n_lig burial counts of the pose's ligand atoms
  double* sums;        // parts > 1: 3 doubles per part
  int32_t* sync;       // parts > 1: {parts finished, overflow seen}
  int32_t* redo_count;  // batched launches: counts the poses flagged for the exact variant
};

// AIR, restraint penalty and the weighted fitness (thread 0), then the outputs.
__device__ __forceinline__ void finish_pose(CxRef cx, const PoseCoef& pc, const double* __restrict__ coords,
                                            const Weights& w, double vdw, double elec, double desolv,
                                            double* __restrict__ o, int32_t* __restrict__ nsat_out) {
  // fitness.rs:205-267, restraints.rs:11-40: FP64 CA-CA flat-bottom term
  double air = 0.0;
  int32_t nsat = 0;
  for (int i = 0; i < cx.n_restr; ++i) {
    double px, py, pz;
    ligand_atom(cx, pc, coords, cx.restr_lig[i], px, py, pz);
    const double ex = cx.restr_rec_xyz[3 * i] - px, ey = cx.restr_rec_xyz[3 * i + 1] - py,
                 ez = cx.restr_rec_xyz[3 * i + 2] - pz;
    const double e2 = ex * ex + ey * ey + ez * ez;
    const double dist = sqrt(e2);
    const double viol = e2 > kD2RestraintIncl ? dist - kAirUpper : 0.0;
    air += kAirK * viol * viol;
    nsat += e2 <= kD2RestraintIncl ? 1 : 0;
  }
  const double n = static_cast<double>(cx.n_restr);
  const double ratio = cx.n_restr > 0 ? static_cast<double>(nsat) / n : 0.0;
  o[0] = vdw;
  o[1] = elec;
  o[2] = desolv;
  o[3] = air;
  o[4] = (1.0 - ratio) * n;  // chromosome.rs:117-118
  o[5] = w.vdw * vdw + w.elec * elec + w.desolv * desolv + w.air * air;  // chromosome.rs:121-124
  if (nsat_out) *nsat_out = nsat;
}

// fitness.rs:165-181: one ligand atom's desolvation term.
__device__ __forceinline__ double ligand_desolv(CxRef cx, int l, int32_t cnt) {
  if (cnt == 0) return 0.0;  // burial 0: the term is exactly zero
  const double asp = cx.lig_asp[l];
  if (!(fabs(asp) >= 0.001)) return 0.0;
  const double burial = fmin(static_cast<double>(cnt) / 10.0, 1.0);
  const double rad = cx.lig_radius[l];
  return asp * burial * rad * rad * 4.0 * kPI;
}

// Scores one pose (or the receptor range pp of one pose) with the whole
// workgroup.  pose_rec: the pose record (kPoseDoubles doubles) or null;
// coords: this pose's ligand coordinates (n_lig x 3) when scoring
// coordinates as given.  o: 6 outputs; *redo is set when the fast variant
// could not settle every band pair (the exact variant then rescores the
// pose).  The LDS receptor count table covers the part's slots [r0, r1).
// COUNT (a diagnostic instantiation): counts[0] += the pair steps the wave
// executes x 64 lanes, counts[1] += the pairs among them within 15 A (below
// the 15 A guard band).
template <int K, bool EXACT, bool COUNT = false>
__device__ __forceinline__ void score_pose(CxRef cx, const double* __restrict__ pose_rec,
                                           const double* __restrict__ coords, const Weights& w,
                                           double* __restrict__ o, int32_t* __restrict__ nsat_out,
                                           int32_t* __restrict__ redo, int32_t* __restrict__ s_rcnt,
                                           int32_t* s_overflow, const PosePart& pp,
                                           unsigned long long* __restrict__ counts = nullptr) {
  const int range = pp.r1 - pp.r0;
  double* s_red = reinterpret_cast<double*>(s_rcnt + ((range + 1) & ~1));
  SettleEvent* s_ev = reinterpret_cast<SettleEvent*>(s_red + 4 * (kScoreThreads / 64));
  float* s_box = reinterpret_cast<float*>(s_ev + (kScoreThreads / 64) * kEventsPerWave);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  SettleEvent* my_ev = s_ev + wave * kEventsPerWave;
  const int32_t GDOCK_AS4* rf = (const int32_t GDOCK_AS4*)cx.rec_fixed;
  const int32_t GDOCK_AS4* cstart = (const int32_t GDOCK_AS4*)cx.class_start;

  PoseCoef pc{};
  if (!coords) pc = load_pose(pose_rec);
  const Bands B = make_bands(cx);
  const double scale = ldexp(1.0, cx.qexp);
  const float unit = ldexpf(1.0f, -cx.qexp);

  for (int i = tid; i < range; i += kScoreThreads) s_rcnt[i] = 0;
  if (tid == 0) *s_overflow = 0;
  __syncthreads();

  double sum_v = 0.0, sum_e = 0.0, sum_d = 0.0;
  int overflow = 0;
  uint32_t n_steps = 0, n_within = 0;  // COUNT only
  for (int base = 0; base < cx.n_lig; base += kScoreThreads * K) {
    LaneAtoms<K> a;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int l = lig_index<K>(base, k);
      if (l < cx.n_lig) {
        double px, py, pz;
        ligand_atom(cx, pc, coords, l, px, py, pz);
        a.x[k] = to_fixed(px, cx.origin[0], scale);
        a.y[k] = to_fixed(py, cx.origin[1], scale);
        a.z[k] = to_fixed(pz, cx.origin[2], scale);
      } else {
        a.x[k] = a.y[k] = a.z[k] = kFixedClamp;
      }
      a.eps[k] = 0.0f; a.ir2[k] = 0.0f; a.acv[k] = 0.0f; a.ace[k] = 0.0f; a.cnt[k] = 0;
      a.kq[k] = cx.lig_kq[l];  // zero-padded past n_lig
    }
    float* my_box = s_box + wave * (kMaxAtomsPerLane * kBoxFloats);
    __syncthreads();  // the previous chunk's boxes are no longer read
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const bool in = lig_index<K>(base, k) < cx.n_lig;
      // quarter boxes first, then the block's box from them
      int32_t x0 = quarter_min(in ? a.x[k] : INT_MAX), y0 = quarter_min(in ? a.y[k] : INT_MAX),
              z0 = quarter_min(in ? a.z[k] : INT_MAX), x1 = quarter_max(in ? a.x[k] : INT_MIN),
              y1 = quarter_max(in ? a.y[k] : INT_MIN), z1 = quarter_max(in ? a.z[k] : INT_MIN);
      auto put_box = [unit](float* b, int32_t x0, int32_t y0, int32_t z0, int32_t x1, int32_t y1, int32_t z1) {
        const bool empty = x0 > x1;
        b[0] = empty ? kFarBox : static_cast<float>(x0) * unit;
        b[1] = empty ? kFarBox : static_cast<float>(y0) * unit;
        b[2] = empty ? kFarBox : static_cast<float>(z0) * unit;
        b[3] = empty ? kFarBox : static_cast<float>(x1) * unit;
        b[4] = empty ? kFarBox : static_cast<float>(y1) * unit;
        b[5] = empty ? kFarBox : static_cast<float>(z1) * unit;
      };
      if ((lane & (kLigSub - 1)) == 0)
        put_box(my_box + kBoxFloats * k + 8 * (1 + lane / kLigSub), x0, y0, z0, x1, y1, z1);
#pragma unroll
      for (int off = kLigSub; off < 64; off <<= 1) {
        x0 = min(x0, __shfl_xor(x0, off)); y0 = min(y0, __shfl_xor(y0, off)); z0 = min(z0, __shfl_xor(z0, off));
        x1 = max(x1, __shfl_xor(x1, off)); y1 = max(y1, __shfl_xor(y1, off)); z1 = max(z1, __shfl_xor(z1, off));
      }
      if (lane == 0) put_box(my_box + kBoxFloats * k, x0, y0, z0, x1, y1, z1);
    }
    __syncthreads();
    int nev = 0;
    int cls = -1, cend = 0;
    // Receptor sub-tiles (kRecSub slots) near any of the wave's blocks, 64
    // at a time (lane t tests sub-tile s0 + t); a 64-slot tile is walked when
    // one of its four sub-tiles is, and tests its slots against the blocks
    // that one of its sub-tiles reaches.  Parts start on whole tiles.
    const int s_end = (pp.r1 + kRecSub - 1) / kRecSub;
    for (int s0 = pp.r0 / kRecSub; s0 < s_end; s0 += 64) {
      const int si = s0 + lane;
      uint32_t near_k = 0;  // bit k: sub-tile si within reach of block k
      if (si < s_end) {
        const float4 lo = cx.rec_sub_box[2 * si], hi = cx.rec_sub_box[2 * si + 1];
#pragma unroll
        for (int k = 0; k < K; ++k) near_k |= box_beyond_reach(lo, hi, my_box + kBoxFloats * k) ? 0u : 1u << k;
      }
      uint64_t tiles = __builtin_amdgcn_ballot_w64(near_k != 0u);
      // each tile's blocks, in the four lanes of its sub-tiles
      near_k |= __shfl_xor(near_k, 1);
      near_k |= __shfl_xor(near_k, 2);
      // a tile's first sub-tile; the next near tile's slots load while this
      // one is tested and walked
      auto tile_first = [](uint64_t t) { return (__builtin_ctzll(t) / (64 / kRecSub)) * (64 / kRecSub); };
      int4 r_next = make_int4(0, 0, 0, 0);
      if (tiles != 0ull) {
        const int g = (s0 + tile_first(tiles)) * kRecSub;
        r_next = tile_slot(cx, g, min(g + 64, pp.r1));
      }
      while (tiles != 0ull) {
      const int first = tile_first(tiles);
      tiles &= ~(((1ull << (64 / kRecSub)) - 1ull) << first);
      const int g0 = (s0 + first) * kRecSub;
      const int g1 = min(g0 + 64, pp.r1);
      const int4 r_tile = r_next;
      if (tiles != 0ull) {
        const int g = (s0 + tile_first(tiles)) * kRecSub;
        r_next = tile_slot(cx, g, min(g + 64, pp.r1));
      }
      // A receptor atom beyond every cutoff from all atoms of a ligand block
      // adds nothing and holds no band pair with it: that block skips the
      // atom (a group when no block is near, the tile when nothing is).
      uint64_t keep[K];
      atoms_in_reach<K>(cx, unit, g0, g1, my_box, r_tile, __builtin_amdgcn_readlane(near_k, first), keep);
      uint64_t any_keep = 0;
#pragma unroll
      for (int k = 0; k < K; ++k) any_keep |= keep[k];
      if (any_keep == 0ull) continue;
      int32_t rc = 0;  // lane t: burial count of receptor slot g0 + t from this wave
      for (int j = g0; j < g1; j += kRecGroup) {
        const int sh = j - g0;
        if (((any_keep >> sh) & 15u) == 0u) continue;
        if (j >= cend) {  // next (rmin2, epsilon) class: fold the vdw sum, per-lane LJ factors
          do { ++cls; cend = cstart[cls + 1]; } while (cend <= j);
#pragma unroll
          for (int k = 0; k < K; ++k) {
            sum_v += static_cast<double>(a.eps[k]) * static_cast<double>(a.acv[k]);
            a.acv[k] = 0.0f;
            const int l = lig_index<K>(base, k);
            if (l < cx.n_lig) class_factors(cx, cls, l, a.ir2[k], a.eps[k]);
          }
        }
        const RecAtom r0 = rec_atom(rf, j), r1 = rec_atom(rf, j + 1), r2 = rec_atom(rf, j + 2),
                      r3 = rec_atom(rf, j + 3);
        uint32_t bmin = 0xffffffffu;  // see pair_fast
        int32_t c0 = 0, c1 = 0, c2 = 0, c3 = 0;
#pragma unroll
        for (int k = 0; k < K; ++k) {  // wave-uniform branches
          const uint32_t mk = static_cast<uint32_t>(keep[k] >> sh);
          if ((mk & 15u) == 0u) continue;
          if constexpr (COUNT) {
            n_steps += __builtin_popcount(mk & 15u);
            const RecAtom rr[4] = {r0, r1, r2, r3};
#pragma unroll
            for (int q = 0; q < 4; ++q)
              if ((mk >> q) & 1u) n_within += pair_s(rr[q], a.x[k], a.y[k], a.z[k]) <= B.lo225 ? 1u : 0u;
          }
          if (mk & 1u) c0 += pair_fast<K>(r0, B, a, k, bmin);
          if (mk & 2u) c1 += pair_fast<K>(r1, B, a, k, bmin);
          if (mk & 4u) c2 += pair_fast<K>(r2, B, a, k, bmin);
          if (mk & 8u) c3 += pair_fast<K>(r3, B, a, k, bmin);
          // Coulomb: the group's FP32 partial (at most 4 terms) goes to FP64
          // at once; a longer FP32 partial carries a close pair's large term
          // and rounds every later term at its scale.
          sum_e = __builtin_fma(static_cast<double>(a.ace[k]), a.kq[k], sum_e);
          a.ace[k] = 0.0f;
        }
        const int t = lane - sh;
        rc = t == 0 ? c0 : rc;
        rc = t == 1 ? c1 : rc;
        rc = t == 2 ? c2 : rc;
        rc = t == 3 ? c3 : rc;
        // lanes that saw a band pair in this group, as a lane mask straight
        // from the compare (predicate 36: unsigned less-than); a ballot of a
        // bool would first widen it to 0/1
        const uint64_t amask = __builtin_amdgcn_uicmp(bmin, B.w, 36);
        if (amask != 0ull) {  // rare: this group holds band pairs
          if (EXACT) {
            band_pairs<K, EXACT>(cx, pc, coords, base, j, r0, B, a, sum_e, my_ev, nev, overflow, s_rcnt, pp.r0);
            band_pairs<K, EXACT>(cx, pc, coords, base, j + 1, r1, B, a, sum_e, my_ev, nev, overflow, s_rcnt, pp.r0);
            band_pairs<K, EXACT>(cx, pc, coords, base, j + 2, r2, B, a, sum_e, my_ev, nev, overflow, s_rcnt, pp.r0);
            band_pairs<K, EXACT>(cx, pc, coords, base, j + 3, r3, B, a, sum_e, my_ev, nev, overflow, s_rcnt, pp.r0);
          } else if (nev < kEventsPerWave) {  // settled once per ligand chunk
            if (lane == 0) my_ev[nev] = SettleEvent{j, 0, amask};
            ++nev;
          } else {
            overflow = 1;
          }
        }
      }
      if (g0 + lane < g1 && rc) atomicAdd(&s_rcnt[g0 + lane - pp.r0], rc);
#pragma unroll
      for (int k = 0; k < K; ++k) {  // fold the tile's FP32 LJ sums into FP64
        sum_v += static_cast<double>(a.eps[k]) * static_cast<double>(a.acv[k]);
        a.acv[k] = 0.0f;
      }
      }
    }
    if (!EXACT && nev > 0)
      settle_events<K>(cx, pose_rec, coords, base, pp.r0, B, my_ev, min(nev, kEventsPerWave), a, sum_v, sum_e,
                       s_rcnt);
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int l = lig_index<K>(base, k);
      if (l < cx.n_lig) {
        if (pp.parts > 1) {  // the count over this part's receptor range only
          if (a.cnt[k]) atomicAdd(&pp.lig_cnt[l], a.cnt[k]);
        } else {
          sum_d -= ligand_desolv(cx, l, a.cnt[k]);
        }
      }
    }
  }
  if constexpr (COUNT) {
    uint32_t c = n_within;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off);
    if (lane == 0) {
      atomicAdd(&counts[0], 64ull * n_steps);
      atomicAdd(&counts[1], static_cast<unsigned long long>(c));
    }
  }
  if (overflow && lane == 0) *s_overflow = 1;
  __syncthreads();
  const bool ovf = !EXACT && *s_overflow != 0;
  if (ovf && pp.parts == 1) {  // the exact variant rescores this pose
    if (tid == 0) {
      *redo = 1;
      if (pp.redo_count) atomicAdd(pp.redo_count, 1);
    }
    __syncthreads();
    return;
  }
  // fitness.rs:92: a receptor and a ligand atom with the same serial number
  // add no vdw term.  The pair loop includes them; here each such pair that
  // was not re-decided as a band pair is taken out again with the very FP32
  // value the pair loop added (same coordinates, factors and operations).
  for (int i = tid; i < cx.n_coll; i += kScoreThreads) {
    const int4 P = cx.coll_pairs[i];  // receptor slot, ligand atom, receptor class
    if (P.x < pp.r0 || P.x >= pp.r1) continue;  // another part's receptor slot
    double px, py, pz;
    ligand_atom(cx, pc, coords, P.y, px, py, pz);
    const RecAtom r = rec_atom(rf, P.x);
    const float s = pair_s(r, to_fixed(px, cx.origin[0], scale), to_fixed(py, cx.origin[1], scale),
                           to_fixed(pz, cx.origin[2], scale));
    if (s <= B.lo144 && !in_band(s, B)) {
      float ir2, eps, rho, u, id2;
      class_factors(cx, P.z, P.y, ir2, eps);
      pair_terms(s, ir2, rho, u, id2);
      sum_v -= vdw_term(eps, u, rho);
    }
  }
  for (int j = pp.r0 + tid; j < pp.r1; j += kScoreThreads) {  // fitness.rs:141-162
    const int32_t cnt = s_rcnt[j - pp.r0];
    if (cnt == 0) continue;  // burial 0: the term is exactly zero
    const double asp = cx.rec_asp[j];
    if (fabs(asp) >= 0.001) {
      const double burial = fmin(static_cast<double>(cnt) / 10.0, 1.0);
      const double rad = cx.rec_radius[j];
      sum_d -= asp * burial * rad * rad * 4.0 * kPI;
    }
  }
  // FP64 reductions: wave (cross-lane shuffles), then the 4 waves through LDS
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    sum_v += __shfl_xor(sum_v, off);
    sum_e += __shfl_xor(sum_e, off);
    sum_d += __shfl_xor(sum_d, off);
  }
  if (lane == 0) {
    s_red[3 * wave + 0] = sum_v;
    s_red[3 * wave + 1] = sum_e;
    s_red[3 * wave + 2] = sum_d;
  }
  __syncthreads();
  double vdw = 0.0, elec = 0.0, desolv = 0.0;
  if (tid == 0)
    for (int i = 0; i < kScoreThreads / 64; ++i) {
      vdw += s_red[3 * i]; elec += s_red[3 * i + 1]; desolv += s_red[3 * i + 2];
    }
  if (pp.parts == 1) {
    if (tid == 0) {
      finish_pose(cx, pc, coords, w, vdw, elec, desolv, o, nsat_out);
      if (!EXACT) *redo = 0;
    }
  } else {
    int32_t* s_last = s_overflow + 1;
    if (tid == 0) {
      pp.sums[3 * pp.part + 0] = vdw;
      pp.sums[3 * pp.part + 1] = elec;
      pp.sums[3 * pp.part + 2] = desolv;
      if (ovf) atomicOr(&pp.sync[1], 1);
      __threadfence();
      *s_last = atomicAdd(&pp.sync[0], 1) == pp.parts - 1 ? 1 : 0;
    }
    __syncthreads();
    if (*s_last) {  // every part has finished: combine
      __threadfence();
      double sdl = 0.0;
      for (int l = tid; l < cx.n_lig; l += kScoreThreads) sdl -= ligand_desolv(cx, l, atomicExch(&pp.lig_cnt[l], 0));
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) sdl += __shfl_xor(sdl, off);
      __syncthreads();
      if (lane == 0) s_red[wave] = sdl;
      __syncthreads();
      if (tid == 0) {
        double lig_d = 0.0;
        for (int i = 0; i < kScoreThreads / 64; ++i) lig_d += s_red[i];
        double v = 0.0, e = 0.0, d = 0.0;
        for (int q = 0; q < pp.parts; ++q) {
          v += pp.sums[3 * q]; e += pp.sums[3 * q + 1]; d += pp.sums[3 * q + 2];
        }
        const int any_ovf = atomicExch(&pp.sync[1], 0);
        atomicExch(&pp.sync[0], 0);
        if (any_ovf) {
          *redo = 1;  // the exact variant rescores the whole pose
          if (pp.redo_count) atomicAdd(pp.redo_count, 1);
        } else {
          finish_pose(cx, pc, coords, w, v, e, d + lig_d, o, nsat_out);
          // the exact variant leaves the flag set: its other parts read it
          if (!EXACT) *redo = 0;
        }
      }
    }
  }
  __syncthreads();  // LDS is reused by the workgroup's next pose
}

// Receptor slots [r0, r1) of part `part` of a pose split into `parts`: an
// equal share of the 64-slot tiles.
__device__ __forceinline__ int2 part_range(int n_rec, int part, int parts) {
  const int tiles = (n_rec + 63) / 64;
  const int t0 = part * tiles / parts, t1 = (part + 1) * tiles / parts;
  return make_int2(64 * t0, min(64 * t1, n_rec));
}

// Plain path: pose order[blockIdx.x / parts] (blockIdx.x / parts without an
// order), receptor part blockIdx.x % parts (one part unless the receptor
// exceeds one workgroup's LDS table).  A split pose keeps its shared state at
// split_* (cleared before the launch; the last part clears it again).
template <int K, bool EXACT, bool COUNT = false>
__global__ __launch_bounds__(kScoreThreads) __attribute__((amdgpu_waves_per_eu(EXACT ? 1 : 4, EXACT ? 8 : 4))) void score_poses_kernel(
    const DevComplex* __restrict__ d_cx, const double* __restrict__ poses, const double* __restrict__ coords,
    const Weights w, double* __restrict__ out, int32_t* __restrict__ nsat_out, int32_t* __restrict__ redo,
    int parts, int32_t* __restrict__ split_lig, double* __restrict__ split_sums, int32_t* __restrict__ split_sync,
    const int32_t* __restrict__ order, unsigned long long* __restrict__ counts) {
  const int slot = blockIdx.x / parts, part = blockIdx.x % parts;
  const int pose = order ? order[slot] : slot;
  if (EXACT && redo[pose] == 0) return;  // only poses whose settle buffer overflowed
  extern __shared__ int32_t s_lds[];
  __shared__ int32_t s_flags[2];
  CxRef cx = *(const DevComplex GDOCK_AS4*)d_cx;
  const int2 rr = part_range(cx.n_rec, part, parts);
  PosePart pp{rr.x, rr.y, part, parts, nullptr, nullptr, nullptr, nullptr};
  if (parts > 1) {
    pp.lig_cnt = split_lig + static_cast<size_t>(pose) * cx.n_lig;
    pp.sums = split_sums + static_cast<size_t>(pose) * 3 * parts;
    pp.sync = split_sync + static_cast<size_t>(pose) * 2;
  }
  score_pose<K, EXACT, COUNT>(cx, coords ? nullptr : poses + static_cast<size_t>(kPoseDoubles) * pose,
                              coords ? coords + static_cast<size_t>(pose) * cx.n_lig * 3 : nullptr, w,
                              out + static_cast<size_t>(kOutDoubles) * pose, nsat_out ? nsat_out + pose : nullptr,
                              redo + pose, s_lds, s_flags, pp, counts);
}

int score_atoms_per_lane(int n_lig) {
  int k = (n_lig + kScoreThreads - 1) / kScoreThreads;
  return k < 1 ? 1 : (k > kMaxAtomsPerLane ? kMaxAtomsPerLane : k);
}

// Batched form over many complexes.  An item is {complex slot, global pose
// index}; each complex carries its own ligand atoms per lane (lanes_k), so the
// workgroup picks the matching instantiation of score_pose per item.
template <bool EXACT>
__device__ __forceinline__ void score_item(CxRef cx, const double* __restrict__ pose_rec, const Weights& w,
                                           double* __restrict__ o, int32_t* __restrict__ nsat,
                                           int32_t* __restrict__ redo, int32_t* __restrict__ s_lds,
                                           int32_t* s_flags, const PosePart& pp) {
  switch (cx.lanes_k) {
#define GDOCK_ITEM(KK)                                                                                   \
  case KK:                                                                                               \
    score_pose<KK, EXACT>(cx, pose_rec, nullptr, w, o, nsat, redo, s_lds, s_flags, pp);                 \
    break;
    GDOCK_ITEM(1) GDOCK_ITEM(2) GDOCK_ITEM(3) GDOCK_ITEM(4)
    GDOCK_ITEM(5) GDOCK_ITEM(6) GDOCK_ITEM(7) GDOCK_ITEM(8)
#undef GDOCK_ITEM
    default: break;  // lanes_k is validated on the host
  }
}

// Part `part` of a pose split into `parts` (1: the whole pose): an equal share
// of the receptor tiles; the split state has room for cx.parts parts.
__device__ __forceinline__ PosePart part_of(CxRef cx, int part, int parts, int pose_row) {
  if (parts <= 1) return PosePart{0, cx.n_rec, 0, 1, nullptr, nullptr, nullptr, nullptr};
  const int2 rr = part_range(cx.n_rec, part, parts);
  const size_t q = static_cast<size_t>(pose_row - cx.split_pose_base);
  return PosePart{rr.x, rr.y, part, parts, cx.split_lig + q * cx.n_lig, cx.split_sums + q * 3 * cx.parts,
                  cx.split_sync + q * 2, nullptr};
}

__device__ __forceinline__ int item_slot(int2 I) { return I.x & kItemSlotMask; }
__device__ __forceinline__ int item_part(int2 I) { return (I.x >> kItemPartShift) & 15; }
__device__ __forceinline__ int item_parts(int2 I) { return (I.x >> kItemPartsShift) & 15; }

// Persistent scoring of a batch: every workgroup takes items from a shared
// counter until the list (most expensive first) is exhausted.  An item is
// {complex slot | part << kItemPartShift | parts << kItemPartsShift, global
// pose row}.  The item count is read from device memory so that one captured
// launch serves any number of items up to the capacity.  The last workgroup
// to leave resets the counter and the exit count, so the next launch starts
// from zero with no memset.
__global__ __launch_bounds__(kScoreThreads) __attribute__((amdgpu_waves_per_eu(4, 4))) void score_batch_kernel(
    const DevComplex* __restrict__ table, const int2* __restrict__ items, const int32_t* __restrict__ n_items_p,
    int32_t* __restrict__ counter, int32_t* __restrict__ exit_count, const double* __restrict__ poses,
    const Weights w, double* __restrict__ out, int32_t* __restrict__ nsat_out, int32_t* __restrict__ redo,
    int32_t* __restrict__ redo_count) {
  extern __shared__ int32_t s_lds[];
  __shared__ int32_t s_flags[2], s_item;
  const int n_items = *n_items_p;
  for (;;) {
    if (threadIdx.x == 0) s_item = atomicAdd(counter, 1);
    __syncthreads();
    const int it = s_item;
    __syncthreads();
    if (it >= n_items) break;
    const int2 I = items[it];
    CxRef cx = ((const DevComplex GDOCK_AS4*)table)[item_slot(I)];
    PosePart pp = part_of(cx, item_part(I), item_parts(I), I.y);
    pp.redo_count = redo_count;
    score_item<false>(cx, poses + static_cast<size_t>(kPoseDoubles) * I.y, w,
                      out + static_cast<size_t>(kOutDoubles) * I.y, nsat_out + I.y, redo + I.y, s_lds, s_flags, pp);
  }
  if (threadIdx.x == 0) {
    __threadfence();
    if (atomicAdd(exit_count, 1) == static_cast<int>(gridDim.x) - 1) {
      atomicExch(counter, 0);
      atomicExch(exit_count, 0);
    }
  }
}

// Exact rescoring of the batch's poses whose settle buffer overflowed: every
// item (part) of a flagged pose runs again, deciding band pairs inline; the
// items of clean poses return at once.
__global__ __launch_bounds__(kScoreThreads) void score_batch_exact_kernel(
    const DevComplex* __restrict__ table, const int2* __restrict__ items, const int32_t* __restrict__ n_items_p,
    const double* __restrict__ poses, const Weights w, double* __restrict__ out, int32_t* __restrict__ nsat_out,
    int32_t* __restrict__ redo, const int32_t* __restrict__ redo_count) {
  extern __shared__ int32_t s_lds[];
  __shared__ int32_t s_flags[2];
  if (*redo_count == 0) return;  // no pose flagged in this launch
  const int n_items = *n_items_p;
  for (int it = blockIdx.x; it < n_items; it += gridDim.x) {
    const int2 I = items[it];
    if (redo[I.y] == 0) continue;  // uniform per workgroup
    CxRef cx = ((const DevComplex GDOCK_AS4*)table)[item_slot(I)];
    const PosePart pp = part_of(cx, item_part(I), item_parts(I), I.y);
    score_item<true>(cx, poses + static_cast<size_t>(kPoseDoubles) * I.y, w,
                     out + static_cast<size_t>(kOutDoubles) * I.y, nsat_out + I.y, redo + I.y, s_lds, s_flags, pp);
  }
}

// Dynamic LDS of a scoring workgroup covering at most `slots` receptor slots:
// the burial-count table, the cross-wave reduction, the settle events and the
// ligand block boxes.
size_t score_lds_bytes(int slots) {
  return static_cast<size_t>((slots + 1) & ~1) * sizeof(int32_t) + 4 * (kScoreThreads / 64) * sizeof(double) +
         (kScoreThreads / 64) * kEventsPerWave * sizeof(SettleEvent) +
         (kScoreThreads / 64) * kMaxAtomsPerLane * kBoxFloats * sizeof(float);
}

// Most receptor slots one part covers when a receptor of n_rec slots is
// split into `parts` (whole tiles).
int part_slots(int n_rec, int parts) {
  const int tiles = (n_rec + 63) / 64;
  return std::min(n_rec, 64 * ((tiles + parts - 1) / parts));
}

// Parts of the plain path: one workgroup per pose while the receptor fits one
// workgroup's LDS table, else the fewest equal receptor ranges that fit; with
// min_items > 0, at least enough ranges (whole tiles, each part non-empty)
// that the launch holds min_items workgroups.
int fill_parts(const DevComplex& cx, int n_poses, int min_items) {
  if (min_items <= 0 || n_poses <= 0) return 1;
  const int tiles = (cx.n_rec + 63) / 64;
  int want = std::max(1, std::min({(min_items + n_poses - 1) / n_poses, tiles, kMaxPoseParts}));
  while (want > 1 && (want - 1) * ((tiles + want - 1) / want) >= tiles) --want;  // no empty part
  return want;
}

int plain_parts(const DevComplex& cx, int n_poses, int min_items) {
  const int tiles = (cx.n_rec + 63) / 64, per = kMaxLdsReceptor / 64;
  return std::max({1, (tiles + per - 1) / per, fill_parts(cx, n_poses, min_items)});
}

// Plain-path scratch: the per-pose rescore flags, the launch order of the
// poses and their cost buckets, then the split state (each piece 256-byte
// aligned).
struct PlainScratch { size_t order, keys, split, total; };
static PlainScratch plain_scratch_layout(const DevComplex& cx, int n_poses, int min_items) {
  auto al = [](size_t b) { return (b + 255) / 256 * 256; };
  const size_t n = static_cast<size_t>(std::max(n_poses, 1));
  PlainScratch L;
  L.order = al(n * sizeof(int32_t));
  L.keys = L.order + al(n * sizeof(int32_t));
  L.split = L.keys + al(n);
  L.total = L.split + split_state_bytes(cx, plain_parts(cx, n_poses, min_items), n_poses);
  return L;
}

size_t plain_scratch_bytes(const DevComplex& cx, int n_poses, int min_items) {
  return plain_scratch_layout(cx, n_poses, min_items).total;
}

// Estimated work of scoring a pose (pose record p) with k ligand atoms per
// lane, by one whole wave: the receptor slots within the ligand's radius +
// kCullReach of the pose's ligand centre (the rotation keeps the centre; the
// genes move it; every fourth slot is tested), plus 1/16 of the slots for the
// per-tile tests, times the ligand lane-slots walked.
__device__ double pose_cost(CxRef cx, const double* __restrict__ p, int k) {
  const int lane = threadIdx.x & 63;
  const double ox = cx.com[0] - cx.origin[0], oy = cx.com[1] - cx.origin[1], oz = cx.com[2] - cx.origin[2];
  const double lig_r = cx.lig_extent - fmax(fabs(ox), fmax(fabs(oy), fabs(oz)));
  const float px = static_cast<float>(ox + p[3]), py = static_cast<float>(oy + p[4]),
              pz = static_cast<float>(oz + p[5]);
  const float reach = static_cast<float>(lig_r) + kCullReach;
  const float unit = ldexpf(1.0f, -cx.qexp);
  int32_t near = 0;
#pragma unroll 4
  for (int j = lane; j < cx.n_rec; j += 256) {  // every fourth slot: an estimate
    const int4 r = cx.rec_fixed[j];
    const float dx = static_cast<float>(r.x) * unit - px, dy = static_cast<float>(r.y) * unit - py,
                dz = static_cast<float>(r.z) * unit - pz;
    near += dx * dx + dy * dy + dz * dz <= reach * reach ? 4 : 0;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) near += __shfl_xor(near, off);
  const int per_chunk = kScoreThreads * k;
  const double slots = static_cast<double>((cx.n_lig + per_chunk - 1) / per_chunk) * per_chunk;
  return (near + 0.0625 * cx.n_rec) * slots;
}

// Estimated work of scoring a pose (pose record p) with k ligand atoms per
// lane, by one whole wave, for the plain path's launch order.  The pose moves
// each ligand block's bounding sphere, and every receptor sub-tile box within
// the sphere's radius + kCullReach counts its kRecSub slots (the pairs the
// kernel's culling keeps, nearly).  A workgroup ends with its slowest wave,
// and wave w of a chunk holds the chunk's blocks w k .. w k + k - 1, so the
// estimate sums, over chunks, the largest of the four waves' block counts;
// plus 1/16 of the slots per block for the per-tile tests.
__device__ double pose_cost_blocks(CxRef cx, const double* __restrict__ p, int k) {
  const int lane = threadIdx.x & 63;
  const float ca = p[6], sa = p[7], cb = p[8], sb = p[9], cg = p[10], sg = p[11];
  const float m01 = p[12], m02 = p[13], m11 = p[14], m12 = p[15];
  const float ox = static_cast<float>(cx.com[0] - cx.origin[0] + p[3]),
              oy = static_cast<float>(cx.com[1] - cx.origin[1] + p[4]),
              oz = static_cast<float>(cx.com[2] - cx.origin[2] + p[5]);
  const int n_blocks = (cx.n_lig + kLigBlock - 1) / kLigBlock, n_sub = (cx.n_rec + kRecSub - 1) / kRecSub;
  const int waves = kScoreThreads / 64;
  int32_t total = 0, wave_sum = 0, wave_max = 0;
  for (int b = 0; b < n_blocks; ++b) {
    const float4 S = cx.lig_block_sphere[b];
    const float x = ((S.x * ca) * cb + S.y * m01) + S.z * m02 + ox;
    const float y = ((S.x * sa) * cb + S.y * m11) + S.z * m12 + oy;
    const float z = ((-S.x) * sb + (S.y * cb) * sg) + (S.z * cb) * cg + oz;
    const float reach = S.w + kCullReach;
    int32_t near = 0;
    for (int t = lane; t < n_sub; t += 64) {
      const float4 lo = cx.rec_sub_box[2 * t], hi = cx.rec_sub_box[2 * t + 1];
      const float dx = fmaxf(fmaxf(lo.x - x, x - hi.x), 0.0f), dy = fmaxf(fmaxf(lo.y - y, y - hi.y), 0.0f),
                  dz = fmaxf(fmaxf(lo.z - z, z - hi.z), 0.0f);
      near += dx * dx + dy * dy + dz * dz <= reach * reach ? 1 : 0;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) near += __shfl_xor(near, off);
    wave_sum += near;
    const int in_chunk = b % (waves * k);
    if (in_chunk % k == k - 1 || b == n_blocks - 1) {  // a wave's last block
      wave_max = max(wave_max, wave_sum);
      wave_sum = 0;
    }
    if (in_chunk == waves * k - 1 || b == n_blocks - 1) {  // a chunk's last block
      total += wave_max;
      wave_max = 0;
    }
  }
  return static_cast<double>(total) * (kRecSub * kLigBlock) + 0.0625 * cx.n_rec * n_blocks * kLigBlock / waves;
}

// Half-octave bucket of an estimated cost.
__device__ __forceinline__ int cost_bucket(double cost) {
  return cost < 1.0 ? 0 : min(kCostBuckets - 1, static_cast<int>(2.0 * log2(cost)));
}

// Plain path, before the scoring launch: each pose's cost bucket, in eighth
// octaves (one wave per pose).
__global__ __launch_bounds__(256) void plain_cost_kernel(const DevComplex* __restrict__ d_cx,
                                                         const double* __restrict__ poses, int n, int k,
                                                         uint8_t* __restrict__ keys) {
  CxRef cx = *(const DevComplex GDOCK_AS4*)d_cx;
  const int i = __builtin_amdgcn_readfirstlane(static_cast<int>(blockIdx.x * 4 + (threadIdx.x >> 6)));
  if (i >= n) return;
  const double cost = pose_cost_blocks(cx, poses + static_cast<size_t>(kPoseDoubles) * i, k);
  if ((threadIdx.x & 63) == 0)
    keys[i] = static_cast<uint8_t>(cost < 16.0 ? 0 : min(kOrderBuckets - 1, static_cast<int>(8.0 * log2(cost)) - 32));
}

// Plain path: the launch order of the poses, most expensive bucket first, so
// that a launch of several rounds of workgroups does not end on expensive
// poses dispatched last (a pose costs up to twentyfold another of the same
// complex).  One workgroup: counts per bucket, their descending prefix sums,
// then every pose placed at its bucket's cursor.  Each pose is scored by one
// workgroup whatever its place, so the order changes no result.
__global__ __launch_bounds__(kOrderThreads) void plain_order_kernel(const uint8_t* __restrict__ keys, int n,
                                                                    int32_t* __restrict__ order) {
  __shared__ int32_t s_count[kOrderBuckets], s_cursor[kOrderBuckets];
  const int tid = threadIdx.x;
  if (tid < kOrderBuckets) s_count[tid] = 0;
  __syncthreads();
  for (int i = tid; i < n; i += kOrderThreads) atomicAdd(&s_count[keys[i]], 1);
  __syncthreads();
  if (tid < kOrderBuckets) {
    int32_t start = 0;
    for (int c = kOrderBuckets - 1; c > tid; --c) start += s_count[c];
    s_cursor[tid] = start;
  }
  __syncthreads();
  for (int i = tid; i < n; i += kOrderThreads) order[atomicAdd(&s_cursor[keys[i]], 1)] = i;
}

template <int K>
static void launch_pair(const DevComplex* d_cx, const double* d_poses, const double* d_coords, int n_poses,
                        const Weights& w, double* d_out, int32_t* d_nsat, int32_t* d_redo, int parts,
                        const DevComplex& split, const int32_t* d_order, size_t lds, hipStream_t stream,
                        unsigned long long* counts) {
  const dim3 grid(static_cast<unsigned>(n_poses) * parts);
  if (counts) {  // the counting instantiation of the fast kernel alone
    hipLaunchKernelGGL((score_poses_kernel<K, false, true>), grid, dim3(kScoreThreads), lds, stream, d_cx, d_poses,
                       d_coords, w, d_out, d_nsat, d_redo, parts, split.split_lig, split.split_sums, split.split_sync,
                       nullptr, counts);
    return;
  }
  hipLaunchKernelGGL((score_poses_kernel<K, false>), grid, dim3(kScoreThreads), lds, stream, d_cx, d_poses, d_coords,
                     w, d_out, d_nsat, d_redo, parts, split.split_lig, split.split_sums, split.split_sync, d_order,
                     nullptr);
  // the exact variant returns at once for every pose not flagged: no order
  hipLaunchKernelGGL((score_poses_kernel<K, true>), grid, dim3(kScoreThreads), lds, stream, d_cx, d_poses, d_coords,
                     w, d_out, d_nsat, d_redo, parts, split.split_lig, split.split_sums, split.split_sync, nullptr,
                     nullptr);
}

// Fast kernel over all poses, then the exact kernel (a no-op for every pose
// that did not overflow its settle buffer).  scratch: plain_scratch_bytes
// of device memory (per-pose flags, then the split state of a receptor
// larger than one workgroup's LDS table, cleared here).
hipError_t launch_score(const DevComplex& cx, const DevComplex* d_cx, const double* d_poses, const double* d_coords,
                        int n_poses, const Weights& w, double* d_out, int32_t* d_nsat, void* scratch,
                        hipStream_t stream, int lanes_k, unsigned long long* counts, int min_items) {
  if (n_poses <= 0) return hipSuccess;
  const int parts = plain_parts(cx, n_poses, min_items);
  const size_t lds = score_lds_bytes(part_slots(cx.n_rec, parts));
  const PlainScratch SL = plain_scratch_layout(cx, n_poses, min_items);
  char* m = static_cast<char*>(scratch);
  int32_t* d_redo = reinterpret_cast<int32_t*>(m);
  DevComplex split{};
  split.n_lig = cx.n_lig;
  if (parts > 1) {
    bind_split_state(&split, parts, n_poses, 0, m + SL.split);
    const hipError_t e = hipMemsetAsync(m + SL.split, 0, split_state_bytes(cx, parts, n_poses), stream);
    if (e != hipSuccess) return e;
  }
  const int k = lanes_k > 0 ? lanes_k : score_atoms_per_lane(cx.n_lig);
  // Poses in order of estimated cost, most expensive first, when the launch
  // takes more than one round of the device's workgroup slots (pose records
  // only: coordinates as given carry no pose centre to estimate from).
  int32_t* d_order = nullptr;
  if (d_poses && !d_coords && !counts && static_cast<int64_t>(n_poses) * parts > kOrderMinItems &&
      n_poses <= kOrderMaxPoses) {
    d_order = reinterpret_cast<int32_t*>(m + SL.order);
    uint8_t* d_keys = reinterpret_cast<uint8_t*>(m + SL.keys);
    hipLaunchKernelGGL(plain_cost_kernel, dim3((n_poses + 3) / 4), dim3(256), 0, stream, d_cx, d_poses, n_poses, k,
                       d_keys);
    hipLaunchKernelGGL(plain_order_kernel, dim3(1), dim3(kOrderThreads), 0, stream, d_keys, n_poses, d_order);
  }
  switch (k) {
#define GDOCK_CASE(KK)                                                                                  \
  case KK:                                                                                              \
    launch_pair<KK>(d_cx, d_poses, d_coords, n_poses, w, d_out, d_nsat, d_redo, parts, split, d_order, lds, stream, \
                    counts);                                                                            \
    break;
    GDOCK_CASE(1) GDOCK_CASE(2) GDOCK_CASE(3) GDOCK_CASE(4)
    GDOCK_CASE(5) GDOCK_CASE(6) GDOCK_CASE(7) GDOCK_CASE(8)
#undef GDOCK_CASE
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

size_t batch_lds_bytes(int max_slots) { return score_lds_bytes(max_slots); }

int batch_grid(size_t lds, int device) {
  int per_cu = 0, cus = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, score_batch_kernel, kScoreThreads, lds) != hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cus < 1) cus = 256;
  return per_cu * cus;
}

// Per-launch plan of the work items, most expensive first, so that the
// persistent workgroups start the long items early and the launch does not
// end on a long item taken late.  The cost of a pose varies twentyfold
// within one complex (a pose inside the receptor against one far from it),
// so it is estimated from the pose itself: the receptor slots within the
// ligand's radius + kCullReach of the pose's ligand centre (the rotation keeps
// the centre; the genes move it; every fourth slot is tested), plus 1/16 of
// the slots for the per-tile tests, times the ligand lane-slots walked.  A pose
// estimated above kSplitCost is split into ceil(cost / kSplitCost) items over
// receptor ranges (at most the complex's part count), so that no single item
// holds the launch; cheap poses stay whole, as splitting has a cost of its
// own.  Items are counting-sorted into half-octave buckets of their share of
// the cost.  The plan is a function of the pose and the complex only, so a
// pose gives the same bits in any launch and any mix.
__global__ __launch_bounds__(256) void pose_plan_kernel(const DevComplex* __restrict__ table,
                                                        const int2* __restrict__ pose_list,
                                                        const int32_t* __restrict__ n_list_p,
                                                        const double* __restrict__ poses,
                                                        uint16_t* __restrict__ plan, int32_t* __restrict__ hist) {
  __shared__ int32_t s_hist[kCostBuckets];
  if (threadIdx.x < kCostBuckets) s_hist[threadIdx.x] = 0;
  __syncthreads();
  const int n_list = *n_list_p;
  const int lane = threadIdx.x & 63;
  for (int i = __builtin_amdgcn_readfirstlane(static_cast<int>(blockIdx.x * 4 + (threadIdx.x >> 6))); i < n_list;
       i += gridDim.x * 4) {  // one wave per pose
    const int2 P = pose_list[i];
    CxRef cx = ((const DevComplex GDOCK_AS4*)table)[P.x];
    const double* p = poses + static_cast<size_t>(kPoseDoubles) * P.y;
    const double cost = pose_cost(cx, p, cx.lanes_k > 0 ? cx.lanes_k : 1);
    const int parts = min(max(cx.parts, 1), max(cx.min_parts, static_cast<int>(ceil(cost / kSplitCost))));
    const double share = cost / parts;
    const int b = cost_bucket(share);
    if (lane == 0) {
      plan[i] = static_cast<uint16_t>(b | (parts << 8));
      atomicAdd(&s_hist[b], parts);
    }
  }
  __syncthreads();
  if (threadIdx.x < kCostBuckets && s_hist[threadIdx.x] != 0) atomicAdd(&hist[threadIdx.x], s_hist[threadIdx.x]);
}

// One workgroup of kCostBuckets threads: bucket b starts after every item of
// the costlier buckets; the item count of the launch is the sum; the counts
// are cleared for the next launch.
__global__ __launch_bounds__(kCostBuckets) void item_order_scan_kernel(int32_t* __restrict__ hist,
                                                                       int32_t* __restrict__ n_items) {
  __shared__ int32_t s_n[kCostBuckets];
  const int b = threadIdx.x;
  s_n[b] = hist[b];
  hist[b] = 0;
  if (b == 0) hist[2 * kCostBuckets] = 0;  // poses flagged for the exact variant
  __syncthreads();
  int32_t start = 0;
  for (int c = kCostBuckets - 1; c > b; --c) start += s_n[c];
  hist[kCostBuckets + b] = start;  // cursor
  if (b == 0) *n_items = start + s_n[0];
}

// 256 poses per workgroup: each writes its items at places within its bucket
// from LDS atomics, with one global reservation per bucket present in the
// workgroup.
__global__ __launch_bounds__(256) void item_scatter_kernel(const int2* __restrict__ pose_list,
                                                           const int32_t* __restrict__ n_list_p,
                                                           const uint16_t* __restrict__ plan,
                                                           int32_t* __restrict__ hist, int2* __restrict__ items) {
  __shared__ int32_t s_cnt[kCostBuckets], s_base[kCostBuckets];
  const int n_list = *n_list_p;
  if (static_cast<int>(blockIdx.x) * 256 >= n_list) return;  // whole workgroup
  if (threadIdx.x < kCostBuckets) s_cnt[threadIdx.x] = 0;
  __syncthreads();
  const int i = blockIdx.x * 256 + threadIdx.x;
  const bool valid = i < n_list;
  const int pl = valid ? plan[i] : 0;
  const int b = pl & 0xFF, parts = pl >> 8;
  const int local = valid ? atomicAdd(&s_cnt[b], parts) : 0;
  __syncthreads();
  if (threadIdx.x < kCostBuckets && s_cnt[threadIdx.x] != 0)
    s_base[threadIdx.x] = atomicAdd(&hist[kCostBuckets + threadIdx.x], s_cnt[threadIdx.x]);
  __syncthreads();
  if (valid) {
    const int2 P = pose_list[i];
    for (int q = 0; q < parts; ++q)
      items[s_base[b] + local + q] = make_int2(P.x | (q << kItemPartShift) | (parts << kItemPartsShift), P.y);
  }
}

size_t order_bytes(size_t capacity) {
  const size_t c = std::max<size_t>(capacity, 1);
  return (c * sizeof(uint16_t) + 255) / 256 * 256 + (2 * kCostBuckets + 1) * sizeof(int32_t);
}

void bind_order(BatchLaunch* b, char* mem, size_t capacity) {
  const size_t c = std::max<size_t>(capacity, 1);
  b->plan = reinterpret_cast<uint16_t*>(mem);
  b->hist = reinterpret_cast<int32_t*>(mem + (c * sizeof(uint16_t) + 255) / 256 * 256);
  b->redo_count = b->hist + 2 * kCostBuckets;
}

hipError_t launch_batch(const BatchLaunch& b, const Weights& w, hipStream_t stream) {
  const int cap = std::max(b.capacity, 1);
  hipLaunchKernelGGL(pose_plan_kernel, dim3(std::min((cap + 3) / 4, 4096)), dim3(256), 0, stream, b.table,
                     b.pose_list, b.n_pose_list, b.poses, b.plan, b.hist);
  hipLaunchKernelGGL(item_order_scan_kernel, dim3(1), dim3(kCostBuckets), 0, stream, b.hist, b.n_items);
  hipLaunchKernelGGL(item_scatter_kernel, dim3((cap + 255) / 256), dim3(256), 0, stream, b.pose_list, b.n_pose_list,
                     b.plan, b.hist, b.items);
  hipLaunchKernelGGL(score_batch_kernel, dim3(b.grid), dim3(kScoreThreads), b.lds, stream, b.table, b.items,
                     b.n_items, b.counter, b.exit_count, b.poses, w, b.out, b.nsat, b.redo, b.redo_count);
  hipLaunchKernelGGL(score_batch_exact_kernel, dim3(std::min(b.item_capacity, b.grid)), dim3(kScoreThreads), b.lds,
                     stream, b.table, b.items, b.n_items, b.poses, w, b.out, b.nsat, b.redo, b.redo_count);
  return hipGetLastError();
}

}  // namespace gdock
