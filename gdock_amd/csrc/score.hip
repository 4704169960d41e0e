// Population scoring kernel for gfx950 (CDNA4, wave64).
//
// One workgroup (4 x wave64) scores one pose: it rigidly transforms the
// ligand in FP64 (structure.rs:80-131 operation order), keeps K ligand atoms
// per lane in registers as FP32 coordinates, and walks the receptor.  Receptor
// atoms are wave-uniform, so they are read with scalar loads into SGPRs and
// every pair costs VALU work only; the receptor (16 B/atom) stays resident in
// the scalar cache / L2 across the whole launch.
//
// Per pair (fitness.rs:57-105,122-184,284-302): soft-core LJ, Coulomb with
// eps = r, and the two desolvation burial counts.  Pair math is FP32.  Every
// cutoff decision (d > 1, d < 8, d < 12, d < 15) is taken in FP32 against a
// guard band derived from the coordinate magnitudes.  A pair inside a band is
// rare; the wave records it as an event in LDS and, once per ligand chunk,
// decides it again in FP64 from the original coordinates with the reference's
// operation order and applies the difference, so every decision matches the
// FP64 reference.  (A wave whose event buffer overflows flags its pose, and
// the exact variant of the kernel, launched right after, rescores that pose
// deciding such pairs inline.)  Energies are accumulated per lane in FP32 over
// 64-atom receptor tiles, then in FP64; the cross-lane and cross-wave
// reductions are FP64.  Desolvation burial counts are integers; the per-atom
// desolvation and AIR restraint terms and the weighted fitness are computed
// in FP64 in the reference's order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>

#include "common.h"

// The fast scoring kernels run at 4 waves per SIMD (at most 128 VGPRs): the
// pair loop is VALU-bound and needs the fourth wave to hide its latencies.

namespace gdock {

// Complex descriptors and receptor tables are read-only for a whole launch;
// reading them through the constant address space lets the compiler use
// scalar loads for every wave-uniform access.
#define GDOCK_AS4 __attribute__((address_space(4)))
using CxRef = const DevComplex GDOCK_AS4&;

struct RecAtom { float x, y, z, w; };  // receptor x, y, z (FP32, origin frame), charge

__device__ __forceinline__ RecAtom rec_atom(const float GDOCK_AS4* rf, int j) {
  return RecAtom{rf[4 * j], rf[4 * j + 1], rf[4 * j + 2], rf[4 * j + 3]};
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

// Guard band around a squared cutoff T for FP32 d2 computed from coordinates
// rounded to FP32 in a frame where every component is at most X in magnitude.
// Bound: |d2_f32 - d2| <= 2 sqrt(3) d e + 3 e^2 + 3.01 u T with e = 2uX + u d,
// u = 2^-24; the band is twice that.  Returned as FP32 bounds rounded outward.
__device__ __forceinline__ void guard_band(double T, double X, float& lo, float& hi) {
  const double u = 5.9604644775390625e-08;
  const double s = sqrt(T) * 1.01 + 0.01;
  const double e = 2.0 * u * X + u * s;
  const double b = 2.0 * (3.4641016151377544 * s * e * (1.0 + 2.0 * u) + 3.0 * e * e + 3.01 * u * T * 1.01) + 1e-9;
  float l = static_cast<float>(T - b), h = static_cast<float>(T + b);
  if (static_cast<double>(l) > T - b) l = nextafterf(l, -INFINITY);
  if (static_cast<double>(h) < T + b) h = nextafterf(h, INFINITY);
  lo = uniform(l);
  hi = uniform(h);
}

// One receptor atom r (FP32, wave-uniform) against the K ligand atoms of the
// lane.  The fast path decides every cutoff in FP32 (d2 <= lo_T: below the
// guard band of T; d2 > hi_1: beyond the elec lower cutoff).  A pair whose d2
// lies inside a band, or at d <= ~1 A, flags the receptor atom; settle_pairs()
// then re-decides those pairs in FP64 exactly as the reference does and adds
// the difference to what the fast path accumulated.
//
// Bands are kept as integers: for a non-negative float, its bit pattern
// orders like its value.  c_T = bits(lo_T) + 1, so bits(d2) - c_T borrows
// exactly when d2 <= lo_T (the FP32 decision), and the same difference, as
// an unsigned number below a width in units in the last place, puts d2
// inside the band of T.  The 1 A band has a width of its own (w1): near
// d2 = 1 the coordinate rounding error spans several hundred ulps, ten times
// more than at the 8, 12 and 15 A cutoffs, which share the widest of their
// three widths (w); a narrower band widened to w only sends a few more pairs
// to the FP64 decision.
struct Bands {
  float lo1, lo64, lo144, lo225;
  uint32_t c1, c64, c144, c225, w1, w;
};

__device__ __forceinline__ Bands make_bands(double X) {
  Bands B;
  float h1, h64, h144, h225;
  guard_band(1.0, X, B.lo1, h1);
  guard_band(64.0, X, B.lo64, h64);
  guard_band(144.0, X, B.lo144, h144);
  guard_band(225.0, X, B.lo225, h225);
  B.c1 = __float_as_uint(B.lo1) + 1u;
  B.c64 = __float_as_uint(B.lo64) + 1u;
  B.c144 = __float_as_uint(B.lo144) + 1u;
  B.c225 = __float_as_uint(B.lo225) + 1u;
  B.w1 = __float_as_uint(h1) - B.c1;
  B.w = max(__float_as_uint(h64) - B.c64, max(__float_as_uint(h144) - B.c144, __float_as_uint(h225) - B.c225));
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

// Axis-aligned boxes of a wave's K ligand blocks (FP32, origin frame; empty
// when a block holds no atom), kept in LDS as 8 floats per block, and the
// test that a receptor group's box is farther from a block's box than every
// cutoff.
constexpr int kBoxFloats = 8;

__device__ __forceinline__ float wave_min(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fminf(v, __shfl_xor(v, off));
  return uniform(v);
}

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off));
  return uniform(v);
}

// The distance from a point to any point of a box is at least the
// point-to-box distance; FP32 rounding of it is far below the 0.01 A margin
// of kCullReach over the 15 A cutoff.
__device__ __forceinline__ bool beyond_reach(const float4 r, const float* __restrict__ b) {
  const float dx = fmaxf(fmaxf(b[0] - r.x, r.x - b[3]), 0.0f);
  const float dy = fmaxf(fmaxf(b[1] - r.y, r.y - b[4]), 0.0f);
  const float dz = fmaxf(fmaxf(b[2] - r.z, r.z - b[5]), 0.0f);
  return dx * dx + dy * dy + dz * dz > kCullReach * kCullReach;
}

// Receptor atoms of the tile [g0, g1) against the wave's ligand blocks: bit
// t of keep[k] is set when slot g0 + t can reach block k (lane t tests its
// slot against every block; padding slots sit far from every block).
template <int K>
__device__ __forceinline__ void atoms_in_reach(CxRef cx, int g0, int g1, const float* __restrict__ boxes,
                                               uint64_t (&keep)[K]) {
  const int lane = threadIdx.x & 63;
  const bool valid = g0 + lane < g1;
  const float4 r = cx.rec_xyzq[valid ? g0 + lane : g0];
#pragma unroll
  for (int k = 0; k < K; ++k) keep[k] = __builtin_amdgcn_ballot_w64(valid && !beyond_reach(r, boxes + kBoxFloats * k));
}

template <int K>
struct LaneAtoms {
  // acv: sum of rho (rho - 2) over the current receptor class within the
  // tile (times eps when folded into FP64); ace: sum of q_r / d^2.
  float x[K], y[K], z[K], eps[K], ir6[K], acv[K], ace[K];
  int32_t cnt[K];
};

__device__ __forceinline__ float pair_d2(float rx, float ry, float rz, float lx, float ly, float lz) {
  const float dx = rx - lx, dy = ry - ly, dz = rz - lz;
  return __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
}

// Soft-core LJ (fitness.rs:57-70): the pair's vdw term is eps (rho^2 - 2 rho)
// = eps u rho with u = rho - 2, rho = R^6 / (d^6 + R^6/2); rho and 1/d^2
// come from one reciprocal, 1 / (d^2 (d^6/R^6 + 1/2)).  The pair loop sums
// u rho over the receptor atoms of one (rmin2, eps) class and multiplies by
// eps once per class and tile.
__device__ __forceinline__ void pair_terms(float d2, float ir6, float& rho, float& u, float& id2) {
  const float d6 = (d2 * d2) * d2;
  const float den = __builtin_fmaf(d6, ir6, 0.5f);
  const float rp = __builtin_amdgcn_rcpf(__builtin_fmaf(den, d2, 1.0e-30f));
  rho = rp * d2;
  id2 = rp * den;
  u = rho - 2.0f;
}

// The vdw term of a pair as the pair loop counts it: eps (u rho), in FP64
// from the FP32 factors.
__device__ __forceinline__ double vdw_term(float eps, float u, float rho) {
  return static_cast<double>(eps) * (static_cast<double>(u) * static_cast<double>(rho));
}

// d2 inside one of the four guard bands (the same test as the pair loop's).
__device__ __forceinline__ bool in_band(float d2, const Bands& B) {
  const uint32_t b = __float_as_uint(d2);
  return b - B.c1 < B.w1 || min(b - B.c64, min(b - B.c144, b - B.c225)) < B.w;
}

// Per-lane Lennard-Jones factors of ligand atom l against receptor class c:
// 1/R^6 and sqrt(eps_r eps_l) from FP64 (fitness.rs:58-63), rounded once,
// tabulated per complex (one coalesced 8-byte load per lane).
__device__ __forceinline__ void class_factors(CxRef cx, int c, int l, float& ir6, float& eps) {
  const float2 f = cx.lig_class[static_cast<size_t>(c) * cx.n_lig + l];
  ir6 = f.x;
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
constexpr int kEventsPerWave = 128;

// Fast path of one receptor atom against the lane's ligand atom in slot k;
// returns this wave's burial count of the receptor atom from that slot and
// flags band pairs.
template <int K>
__device__ __forceinline__ int32_t pair_fast(const RecAtom r, const Bands& B, LaneAtoms<K>& a, int k,
                                             uint64_t& ambm) {
  const float d2 = pair_d2(r.x, r.y, r.z, a.x[k], a.y[k], a.z[k]);
  // FP32 decisions below each band, as the borrows of bits(d2) - c_T (pairs
  // of equal serial are included in the vdw sum and taken out afterwards, see
  // collision_correction), and the band test on the same differences.
  const uint32_t bd = __float_as_uint(d2);
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
  // The two band tests as lane masks straight from the compares (predicate
  // 36: unsigned less-than), joined by a scalar OR; a ballot of the combined
  // bool would first widen it to a 0/1 value.
  ambm |= __builtin_amdgcn_uicmp(q1, B.w1, 36) | __builtin_amdgcn_uicmp(min(q64, min(q144, q225)), B.w, 36);
  const bool me = me0 && !le1;
  float rho, u, id2;
  pair_terms(d2, a.ir6[k], rho, u, id2);
  a.acv[k] = __builtin_fmaf(mv ? u : 0.0f, rho, a.acv[k]);
  a.ace[k] = __builtin_fmaf(me ? id2 : 0.0f, r.w, a.ace[k]);
  const int32_t cj = __popcll(md);
  // Pin this pair's results here (an empty asm, no instruction): otherwise
  // the compiler sinks the math below the group's rare band-pair branch and
  // keeps every pair's lane masks alive across it, spilling SGPRs.
  asm volatile("" : "+v"(a.acv[k]), "+v"(a.ace[k]), "+v"(a.cnt[k]), "+s"(ambm));
  return cj;
}

// Band pairs of one receptor atom: recorded as events (fast kernel) or
// settled on the spot in FP64 (exact kernel).
template <int K, bool EXACT>
__device__ __forceinline__ void band_pairs(CxRef cx, const PoseCoef& pc, const double* __restrict__ coords,
                                           int base, int j, const RecAtom r, const Bands& B,
                                           LaneAtoms<K>& a, SettleEvent* __restrict__ ev, int& nev, int& overflow,
                                           int32_t* __restrict__ s_rcnt) {
  const int lane = threadIdx.x & 63;
  const int2 m = cx.rec_meta[j];
  const int rser = m.y ? m.x : INT_MIN;  // INT_MIN: no ligand atom shares this receptor atom's serial
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const float d2 = pair_d2(r.x, r.y, r.z, a.x[k], a.y[k], a.z[k]);
    const int l = lig_index<K>(base, k);
    const uint64_t am = __builtin_amdgcn_ballot_w64(in_band(d2, B) && l < cx.n_lig);
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
      bool x64, x144, x225;
      exact_flags(cx, pc, coords, l, j, rser, x64, x144, x225);
      const bool f144 = d2 <= B.lo144;
      const bool f225 = d2 <= B.lo225 && !(d2 <= B.lo1), f64 = d2 <= B.lo64;
      float rho, u, id2;
      pair_terms(d2, a.ir6[k], rho, u, id2);
      if (x144 != f144) a.acv[k] = __builtin_fmaf(x144 ? u : -u, rho, a.acv[k]);
      if (x225 != f225) a.ace[k] = __builtin_fmaf(x225 ? id2 : -id2, r.w, a.ace[k]);
      const int dc = static_cast<int>(x64) - static_cast<int>(f64);
      if (dc) {
        a.cnt[k] += dc;
        atomicAdd(&s_rcnt[j], dc);
      }
    }
  }
}

// Settles the wave's recorded events.  An event names a group of kRecGroup
// receptor atoms and the lanes that saw a band pair in it; each band pair is
// decided in FP64 and the difference to the FP32 decisions is added to the
// FP64 sums, the lane's burial count and the receptor atom's count.  The
// ligand slot is a loop variable here, so its FP32 coordinates are rebuilt
// from the FP64 transform (the same operations that filled the lane's
// registers) and its burial-count changes are carried as 16-bit fields of
// two packed integers: the lane's register arrays are only ever indexed by
// constants and stay in registers.
template <int K>
__device__ __forceinline__ void settle_events(CxRef cx, const double* __restrict__ pose_rec,
                                              const double* __restrict__ coords, int base, const Bands& B,
                                              const SettleEvent* __restrict__ ev, int nev, LaneAtoms<K>& a,
                                              double& sum_v, double& sum_e, int32_t* __restrict__ s_rcnt) {
  const int lane = threadIdx.x & 63;
  const int32_t GDOCK_AS4* cstart = (const int32_t GDOCK_AS4*)cx.class_start;
  const float GDOCK_AS4* rq = (const float GDOCK_AS4*)cx.rec_xyzq;
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
      const float lx = static_cast<float>(px - cx.origin[0]);
      const float ly = static_cast<float>(py - cx.origin[1]);
      const float lz = static_cast<float>(pz - cx.origin[2]);
      bool have_factors = false;
      float ir6 = 0.0f, eps = 0.0f;
      for (int j = E.j; j < E.j + kRecGroup; ++j) {
        const RecAtom r = rec_atom(rq, j);
        const float d2 = pair_d2(r.x, r.y, r.z, lx, ly, lz);
        if (!in_band(d2, B)) continue;
        const int2 m = cx.rec_meta[j];
        const int rser = m.y ? m.x : INT_MIN;
        if (!have_factors) {
          class_factors(cx, c, l, ir6, eps);
          have_factors = true;
        }
        bool x64, x144, x225;
        exact_flags_at(cx, px, py, pz, l, j, rser, x64, x144, x225);
        const bool f144 = d2 <= B.lo144;
        const bool f225 = d2 <= B.lo225 && !(d2 <= B.lo1), f64 = d2 <= B.lo64;
        float rho, u, id2;
        pair_terms(d2, ir6, rho, u, id2);
        if (x144 != f144) sum_v += x144 ? vdw_term(eps, u, rho) : -vdw_term(eps, u, rho);
        if (x225 != f225)
          sum_e += static_cast<double>(x225 ? id2 : -id2) * static_cast<double>(r.w) * (kCoulomb * cx.lig_q[l]);
        const int dc = static_cast<int>(x64) - static_cast<int>(f64);
        if (dc) {
          const int64_t d = static_cast<int64_t>(dc) << (16 * (k & 3));
          if (k < 4) dpack0 += d; else dpack1 += d;
          atomicAdd(&s_rcnt[j], dc);
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

// Scores one pose with the whole workgroup.  pose_rec: the pose record
// (kPoseDoubles doubles) or null; coords: this pose's ligand coordinates (n_lig x 3) when
// scoring coordinates as given, with `extent` = their largest |x - origin|.
// o: 6 outputs; *redo is set when the fast variant could not settle every
// band pair (the exact variant then rescores the pose).
template <int K, bool EXACT>
__device__ __forceinline__ void score_pose(CxRef cx, const double* __restrict__ pose_rec,
                                           const double* __restrict__ coords, double extent, const Weights& w,
                                           double* __restrict__ o, int32_t* __restrict__ nsat_out,
                                           int32_t* __restrict__ redo, int32_t* __restrict__ s_rcnt,
                                           int32_t* s_overflow, const PosePart& pp) {
  double* s_red = reinterpret_cast<double*>(s_rcnt + ((cx.n_rec + 1) & ~1));
  SettleEvent* s_ev = reinterpret_cast<SettleEvent*>(s_red + 4 * (kScoreThreads / 64));
  float* s_box = reinterpret_cast<float*>(s_ev + (kScoreThreads / 64) * kEventsPerWave);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  SettleEvent* my_ev = s_ev + wave * kEventsPerWave;
  const float GDOCK_AS4* rq = (const float GDOCK_AS4*)cx.rec_xyzq;
  const int32_t GDOCK_AS4* cstart = (const int32_t GDOCK_AS4*)cx.class_start;

  PoseCoef pc{};
  double X;
  if (coords) {
    X = fmax(cx.rec_extent, extent);
  } else {
    pc = load_pose(pose_rec);
    X = fmax(cx.rec_extent, cx.lig_extent + fmax(fabs(pc.dx), fmax(fabs(pc.dy), fabs(pc.dz))));
  }
  const Bands B = make_bands(X);

  for (int i = pp.r0 + tid; i < pp.r1; i += kScoreThreads) s_rcnt[i] = 0;
  if (tid == 0) *s_overflow = 0;
  __syncthreads();

  double sum_v = 0.0, sum_e = 0.0, sum_d = 0.0;
  int overflow = 0;
  for (int base = 0; base < cx.n_lig; base += kScoreThreads * K) {
    LaneAtoms<K> a;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int l = lig_index<K>(base, k);
      if (l < cx.n_lig) {
        double px, py, pz;
        ligand_atom(cx, pc, coords, l, px, py, pz);
        a.x[k] = static_cast<float>(px - cx.origin[0]);
        a.y[k] = static_cast<float>(py - cx.origin[1]);
        a.z[k] = static_cast<float>(pz - cx.origin[2]);
      } else {
        a.x[k] = a.y[k] = a.z[k] = kPadCoord;
      }
      a.eps[k] = 0.0f; a.ir6[k] = 0.0f; a.acv[k] = 0.0f; a.ace[k] = 0.0f; a.cnt[k] = 0;
    }
    float* my_box = s_box + wave * (kMaxAtomsPerLane * kBoxFloats);
    __syncthreads();  // the previous chunk's boxes are no longer read
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const bool in = lig_index<K>(base, k) < cx.n_lig;
      const float x0 = wave_min(in ? a.x[k] : INFINITY), y0 = wave_min(in ? a.y[k] : INFINITY),
                  z0 = wave_min(in ? a.z[k] : INFINITY), x1 = wave_max(in ? a.x[k] : -INFINITY),
                  y1 = wave_max(in ? a.y[k] : -INFINITY), z1 = wave_max(in ? a.z[k] : -INFINITY);
      if (lane == 0) {
        float* b = my_box + kBoxFloats * k;
        b[0] = x0; b[1] = y0; b[2] = z0; b[3] = x1; b[4] = y1; b[5] = z1;
      }
    }
    __syncthreads();
    int nev = 0;
    int cls = -1, cend = 0;
    for (int g0 = pp.r0; g0 < pp.r1; g0 += 64) {
      const int g1 = min(g0 + 64, pp.r1);
      // A receptor atom beyond every cutoff from all atoms of a ligand block
      // adds nothing and holds no band pair with it: that block skips the
      // atom (a group when no block is near, the tile when nothing is).
      uint64_t keep[K];
      atoms_in_reach<K>(cx, g0, g1, my_box, keep);
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
            if (l < cx.n_lig) class_factors(cx, cls, l, a.ir6[k], a.eps[k]);
          }
        }
        const RecAtom r0 = rec_atom(rq, j), r1 = rec_atom(rq, j + 1), r2 = rec_atom(rq, j + 2), r3 = rec_atom(rq, j + 3);
        uint64_t ambm = 0;  // lanes that saw a band pair in this group
        int32_t c0 = 0, c1 = 0, c2 = 0, c3 = 0;
#pragma unroll
        for (int k = 0; k < K; ++k) {  // wave-uniform branches
          const uint32_t mk = static_cast<uint32_t>(keep[k] >> sh);
          if (mk & 1u) c0 += pair_fast<K>(r0, B, a, k, ambm);
          if (mk & 2u) c1 += pair_fast<K>(r1, B, a, k, ambm);
          if (mk & 4u) c2 += pair_fast<K>(r2, B, a, k, ambm);
          if (mk & 8u) c3 += pair_fast<K>(r3, B, a, k, ambm);
        }
        const int t = lane - (j - g0);
        rc = t == 0 ? c0 : rc;
        rc = t == 1 ? c1 : rc;
        rc = t == 2 ? c2 : rc;
        rc = t == 3 ? c3 : rc;
        const uint64_t amask = ambm;
        if (amask != 0ull) {  // rare: this group holds band pairs
          if (EXACT) {
            band_pairs<K, EXACT>(cx, pc, coords, base, j, r0, B, a, my_ev, nev, overflow, s_rcnt);
            band_pairs<K, EXACT>(cx, pc, coords, base, j + 1, r1, B, a, my_ev, nev, overflow, s_rcnt);
            band_pairs<K, EXACT>(cx, pc, coords, base, j + 2, r2, B, a, my_ev, nev, overflow, s_rcnt);
            band_pairs<K, EXACT>(cx, pc, coords, base, j + 3, r3, B, a, my_ev, nev, overflow, s_rcnt);
          } else if (nev < kEventsPerWave) {  // settled once per ligand chunk
            if (lane == 0) my_ev[nev] = SettleEvent{j, 0, amask};
            ++nev;
          } else {
            overflow = 1;
          }
        }
      }
      if (g0 + lane < g1 && rc) atomicAdd(&s_rcnt[g0 + lane], rc);
#pragma unroll
      for (int k = 0; k < K; ++k) {  // fold the tile's FP32 partial sums into FP64
        const int l = lig_index<K>(base, k);
        if (l < cx.n_lig) {
          sum_v += static_cast<double>(a.eps[k]) * static_cast<double>(a.acv[k]);
          sum_e += static_cast<double>(a.ace[k]) * (kCoulomb * cx.lig_q[l]);
        }
        a.acv[k] = 0.0f; a.ace[k] = 0.0f;
      }
    }
    if (!EXACT && nev > 0)
      settle_events<K>(cx, pose_rec, coords, base, B, my_ev, min(nev, kEventsPerWave), a, sum_v, sum_e, s_rcnt);
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
    const float lx = static_cast<float>(px - cx.origin[0]);
    const float ly = static_cast<float>(py - cx.origin[1]);
    const float lz = static_cast<float>(pz - cx.origin[2]);
    const float4 r = cx.rec_xyzq[P.x];
    const float d2 = pair_d2(r.x, r.y, r.z, lx, ly, lz);
    if (d2 <= B.lo144 && !in_band(d2, B)) {
      float ir6, eps, rho, u, id2;
      class_factors(cx, P.z, P.y, ir6, eps);
      pair_terms(d2, ir6, rho, u, id2);
      sum_v -= vdw_term(eps, u, rho);
    }
  }
  for (int j = pp.r0 + tid; j < pp.r1; j += kScoreThreads) {  // fitness.rs:141-162
    const int32_t cnt = s_rcnt[j];
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
          *redo = 0;
        }
      }
    }
  }
  __syncthreads();  // LDS is reused by the workgroup's next pose
}

template <int K, bool EXACT>
__global__ __launch_bounds__(kScoreThreads) __attribute__((amdgpu_waves_per_eu(EXACT ? 1 : 4, EXACT ? 8 : 4))) void score_poses_kernel(
    const DevComplex* __restrict__ d_cx, const double* __restrict__ poses, const double* __restrict__ coords,
    const double* __restrict__ pose_extent, const Weights w, double* __restrict__ out,
    int32_t* __restrict__ nsat_out, int32_t* __restrict__ redo) {
  const int pose = blockIdx.x;
  if (EXACT && redo[pose] == 0) return;  // only poses whose settle buffer overflowed
  extern __shared__ int32_t s_lds[];
  __shared__ int32_t s_flags[2];
  CxRef cx = *(const DevComplex GDOCK_AS4*)d_cx;
  const PosePart whole{0, cx.n_rec, 0, 1, nullptr, nullptr, nullptr};
  score_pose<K, EXACT>(cx, coords ? nullptr : poses + static_cast<size_t>(kPoseDoubles) * pose,
                             coords ? coords + static_cast<size_t>(pose) * cx.n_lig * 3 : nullptr,
                             coords ? pose_extent[pose] : 0.0, w, out + static_cast<size_t>(kOutDoubles) * pose,
                             nsat_out ? nsat_out + pose : nullptr, redo + pose, s_lds, s_flags, whole);
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
    score_pose<KK, EXACT>(cx, pose_rec, nullptr, 0.0, w, o, nsat, redo, s_lds, s_flags, pp);            \
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
  if (parts <= 1) return PosePart{0, cx.n_rec, 0, 1, nullptr, nullptr, nullptr};
  const int tiles = (cx.n_rec + 63) / 64;
  const int t0 = part * tiles / parts, t1 = (part + 1) * tiles / parts;
  const size_t q = static_cast<size_t>(pose_row - cx.split_pose_base);
  return PosePart{64 * t0, min(64 * t1, cx.n_rec), part, parts, cx.split_lig + q * cx.n_lig,
                  cx.split_sums + q * 3 * cx.parts, cx.split_sync + q * 2};
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

// Exact rescoring of the batch's poses whose settle buffer overflowed: the
// whole pose in one workgroup (items of part 0; the other parts' items and
// every clean pose return at once).
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
    if (item_part(I) != 0 || redo[I.y] == 0) continue;  // uniform per workgroup
    CxRef cx = ((const DevComplex GDOCK_AS4*)table)[item_slot(I)];
    const PosePart whole{0, cx.n_rec, 0, 1, nullptr, nullptr, nullptr};
    score_item<true>(cx, poses + static_cast<size_t>(kPoseDoubles) * I.y, w,
                     out + static_cast<size_t>(kOutDoubles) * I.y, nsat_out + I.y, redo + I.y, s_lds, s_flags, whole);
  }
}

size_t score_lds_bytes(const DevComplex& cx) {
  return static_cast<size_t>((cx.n_rec + 1) & ~1) * sizeof(int32_t) + 4 * (kScoreThreads / 64) * sizeof(double) +
         (kScoreThreads / 64) * kEventsPerWave * sizeof(SettleEvent) +
         (kScoreThreads / 64) * kMaxAtomsPerLane * kBoxFloats * sizeof(float);
}

template <int K>
static void launch_pair(const DevComplex* d_cx, const double* d_poses, const double* d_coords, const double* d_extent,
                        int n_poses, const Weights& w, double* d_out, int32_t* d_nsat, int32_t* d_redo, size_t lds,
                        hipStream_t stream) {
  hipLaunchKernelGGL((score_poses_kernel<K, false>), dim3(n_poses), dim3(kScoreThreads), lds, stream, d_cx,
                     d_poses, d_coords, d_extent, w, d_out, d_nsat, d_redo);
  hipLaunchKernelGGL((score_poses_kernel<K, true>), dim3(n_poses), dim3(kScoreThreads), lds, stream, d_cx,
                     d_poses, d_coords, d_extent, w, d_out, d_nsat, d_redo);
}

// Fast kernel over all poses, then the exact kernel (a no-op for every pose
// that did not overflow its settle buffer).  d_redo: n_poses int32 scratch.
hipError_t launch_score(const DevComplex& cx, const DevComplex* d_cx, const double* d_poses, const double* d_coords,
                        const double* d_extent, int n_poses, const Weights& w, double* d_out, int32_t* d_nsat,
                        int32_t* d_redo, hipStream_t stream, int lanes_k) {
  if (n_poses <= 0) return hipSuccess;
  const size_t lds = score_lds_bytes(cx);
  switch (lanes_k > 0 ? lanes_k : score_atoms_per_lane(cx.n_lig)) {
#define GDOCK_CASE(KK)                                                                                            \
  case KK:                                                                                                        \
    launch_pair<KK>(d_cx, d_poses, d_coords, d_extent, n_poses, w, d_out, d_nsat, d_redo, lds, stream);         \
    break;
    GDOCK_CASE(1) GDOCK_CASE(2) GDOCK_CASE(3) GDOCK_CASE(4)
    GDOCK_CASE(5) GDOCK_CASE(6) GDOCK_CASE(7) GDOCK_CASE(8)
#undef GDOCK_CASE
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

size_t batch_lds_bytes(int max_n_rec) {
  DevComplex cx{};
  cx.n_rec = max_n_rec;
  return score_lds_bytes(cx);
}

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
    const double ox = cx.com[0] - cx.origin[0], oy = cx.com[1] - cx.origin[1], oz = cx.com[2] - cx.origin[2];
    const double lig_r = cx.lig_extent - fmax(fabs(ox), fmax(fabs(oy), fabs(oz)));
    const float px = static_cast<float>(ox + p[3]), py = static_cast<float>(oy + p[4]),
                pz = static_cast<float>(oz + p[5]);
    const float reach = static_cast<float>(lig_r) + kCullReach;
    int32_t near = 0;
#pragma unroll 4
    for (int j = lane; j < cx.n_rec; j += 256) {  // every fourth slot: an estimate
      const float4 r = cx.rec_xyzq[j];
      const float dx = r.x - px, dy = r.y - py, dz = r.z - pz;
      near += dx * dx + dy * dy + dz * dz <= reach * reach ? 4 : 0;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) near += __shfl_xor(near, off);
    const int k = cx.lanes_k > 0 ? cx.lanes_k : 1;
    const int per_chunk = kScoreThreads * k;
    const double slots = static_cast<double>((cx.n_lig + per_chunk - 1) / per_chunk) * per_chunk;
    const double cost = (near + 0.0625 * cx.n_rec) * slots;
    const int parts = min(max(cx.parts, 1), max(1, static_cast<int>(ceil(cost / kSplitCost))));
    const double share = cost / parts;
    const int b = share < 1.0 ? 0 : min(kCostBuckets - 1, static_cast<int>(2.0 * log2(share)));
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
