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
// guard band derived from the coordinate magnitudes; a pair inside a band is
// decided again in FP64 from the original coordinates with the reference's
// operation order, so decisions match the FP64 reference exactly.  Energies
// are accumulated per lane in FP32 over 64-atom receptor tiles, then in FP64;
// the cross-lane and cross-wave reductions are FP64.  Desolvation burial
// counts are integers; the per-atom desolvation and AIR restraint terms and
// the weighted fitness are computed in FP64 in the reference's order.
#include <hip/hip_runtime.h>

#include "common.h"

namespace gdock {

struct PoseCoef {
  double ca, sa, cb, sb, cg, sg;
  double m01, m02, m11, m12;  // the loop-invariant sums inside structure.rs:96-101
  double dx, dy, dz;
};

__device__ __forceinline__ PoseCoef load_pose(const double* __restrict__ p) {
  PoseCoef c;
  c.dx = p[3]; c.dy = p[4]; c.dz = p[5];
  c.ca = p[6]; c.sa = p[7]; c.cb = p[8]; c.sb = p[9]; c.cg = p[10]; c.sg = p[11];
  c.m01 = (-c.sa * c.cg) + ((c.ca * c.sb) * c.sg);
  c.m02 = (c.sa * c.sg) + ((c.ca * c.sb) * c.cg);
  c.m11 = (c.ca * c.cg) + ((c.sa * c.sb) * c.sg);
  c.m12 = (-c.ca * c.sg) + ((c.sa * c.sb) * c.cg);
  return c;
}

// FP64 pose of ligand atom l: rotate about the COM, then displace
// (chromosome.rs:79-83 -> structure.rs:80-131).  This file is compiled with
// -ffp-contract=off, so every product and sum rounds as in the reference.
__device__ __forceinline__ void ligand_atom(const DevComplex& cx, const PoseCoef& pc,
                                            const double* __restrict__ coords, int pose, int l,
                                            double& X, double& Y, double& Z) {
  if (coords) {
    const double* p = coords + (static_cast<size_t>(pose) * cx.n_lig + l) * 3;
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

template <int K>
__global__ __launch_bounds__(kScoreThreads) void score_poses_kernel(
    const DevComplex cx, const double* __restrict__ poses, const double* __restrict__ coords,
    const double* __restrict__ pose_extent, const Weights w, double* __restrict__ out,
    int32_t* __restrict__ nsat_out) {
  extern __shared__ int32_t s_rcnt[];  // n_rec burial counters, then reduction scratch
  double* s_red = reinterpret_cast<double*>(s_rcnt + ((cx.n_rec + 1) & ~1));
  const int pose = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;

  PoseCoef pc{};
  double X;
  if (coords) {
    X = fmax(cx.rec_extent, pose_extent[pose]);
  } else {
    pc = load_pose(poses + static_cast<size_t>(kPoseDoubles) * pose);
    X = fmax(cx.rec_extent, cx.lig_extent + fmax(fabs(pc.dx), fmax(fabs(pc.dy), fabs(pc.dz))));
  }
  float lo1, hi1, lo64, hi64, lo144, hi144, lo225, hi225;
  guard_band(1.0, X, lo1, hi1);
  guard_band(64.0, X, lo64, hi64);
  guard_band(144.0, X, lo144, hi144);
  guard_band(225.0, X, lo225, hi225);

  for (int i = tid; i < cx.n_rec; i += kScoreThreads) s_rcnt[i] = 0;
  __syncthreads();

  double sum_v = 0.0, sum_e = 0.0, sum_d = 0.0;

  for (int base = 0; base < cx.n_lig; base += kScoreThreads * K) {
    float lx[K], ly[K], lz[K], epsij[K], invr6[K], acv[K], ace[K];
    double adv[K], ade[K];
    int32_t lcnt[K], lser[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int l = base + tid + kScoreThreads * k;
      if (l < cx.n_lig) {
        double px, py, pz;
        ligand_atom(cx, pc, coords, pose, l, px, py, pz);
        lx[k] = static_cast<float>(px - cx.origin[0]);
        ly[k] = static_cast<float>(py - cx.origin[1]);
        lz[k] = static_cast<float>(pz - cx.origin[2]);
        lser[k] = cx.lig_serial[l];
      } else {
        lx[k] = ly[k] = lz[k] = kPadCoord;
        lser[k] = 0;
      }
      epsij[k] = 0.0f; invr6[k] = 0.0f; acv[k] = 0.0f; ace[k] = 0.0f;
      adv[k] = 0.0; ade[k] = 0.0; lcnt[k] = 0;
    }

    int cls = -1, cend = 0;
    for (int t0 = 0; t0 < cx.n_rec; t0 += 64) {
      const int t1 = min(t0 + 64, cx.n_rec);
      int32_t rc = 0;  // lane t: burial count of receptor atom t0 + t from this wave
      for (int j = t0; j < t1; ++j) {
        if (j == cend) {  // next (rmin2, epsilon) class: per-lane LJ factors
          do { ++cls; cend = cx.class_start[cls + 1]; } while (cend == j);
          const double rmin_c = cx.class_rmin2[cls], eps_c = cx.class_eps[cls];
#pragma unroll
          for (int k = 0; k < K; ++k) {
            const int l = base + tid + kScoreThreads * k;
            if (l < cx.n_lig) {
              const double R = rmin_c + cx.lig_rmin2[l];
              const double R2 = R * R, R6 = R2 * (R2 * R2);
              invr6[k] = R6 > 1e-10 ? static_cast<float>(1.0 / R6) : 1.0e10f;
              epsij[k] = static_cast<float>(sqrt(eps_c * cx.lig_eps[l]));
            }
          }
        }
        const float4 r = cx.rec_xyzq[j];
        int coll = 0, rser = 0;
        if (cx.has_serial_collision) {
          const int2 m = cx.rec_meta[j];
          rser = m.x; coll = m.y;
        }
        int32_t cj = 0;
#pragma unroll
        for (int k = 0; k < K; ++k) {
          const float dx = r.x - lx[k], dy = r.y - ly[k], dz = r.z - lz[k];
          const float d2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
          const bool a1 = d2 < hi1, a64 = d2 < hi64, a144 = d2 < hi144, a225 = d2 < hi225;
          const bool b1 = d2 <= lo1, b64 = d2 <= lo64, b144 = d2 <= lo144, b225 = d2 <= lo225;
          bool m1 = !a1, m64 = a64, m144 = a144, m225 = a225;
          const bool amb = (a1 != b1) | (a64 != b64) | (a144 != b144) | (a225 != b225);
          if (__any(amb)) {
            if (amb && base + tid + kScoreThreads * k < cx.n_lig) {  // FP64 decision, structure.rs:312-317 order
              double px, py, pz;
              ligand_atom(cx, pc, coords, pose, base + tid + kScoreThreads * k, px, py, pz);
              const double ex = cx.rec_x[j] - px, ey = cx.rec_y[j] - py, ez = cx.rec_z[j] - pz;
              const double e2 = ex * ex + ey * ey + ez * ez;
              m1 = e2 > kD2ElecMinExcl;
              m64 = e2 <= kD2DesolvIncl;
              m144 = e2 > 0.0 && e2 <= kD2VdwIncl;
              m225 = e2 <= kD2ElecMaxIncl;
            }
          }
          if (coll) m144 = m144 && (lser[k] != rser);  // fitness.rs:92 serial skip
          // soft-core LJ (rho = R^6 / (d^6 + R^6/2)) and 1/d^2 from one reciprocal
          const float d6 = (d2 * d2) * d2;
          const float den = __builtin_fmaf(d6, invr6[k], 0.5f);
          const float p = __builtin_fmaf(den, d2, 1.0e-30f);
          const float rp = __builtin_amdgcn_rcpf(p);
          const float rho = rp * d2;
          const float id2 = rp * den;
          const float tv = (rho - 2.0f) * rho;
          acv[k] = __builtin_fmaf(m144 ? tv : 0.0f, epsij[k], acv[k]);
          ace[k] = __builtin_fmaf((m1 && m225) ? id2 : 0.0f, r.w, ace[k]);
          lcnt[k] += m64 ? 1 : 0;
          cj += __popcll(__ballot(m64));
        }
        rc = lane == j - t0 ? cj : rc;
      }
      if (t0 + lane < t1 && rc) atomicAdd(&s_rcnt[t0 + lane], rc);
#pragma unroll
      for (int k = 0; k < K; ++k) {
        adv[k] += static_cast<double>(acv[k]);
        ade[k] += static_cast<double>(ace[k]);
        acv[k] = 0.0f; ace[k] = 0.0f;
      }
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int l = base + tid + kScoreThreads * k;
      if (l < cx.n_lig) {
        sum_v += adv[k];
        sum_e += ade[k] * (kCoulomb * cx.lig_q[l]);
        const double asp = cx.lig_asp[l];
        if (fabs(asp) >= 0.001) {  // fitness.rs:165-181
          const double burial = fmin(static_cast<double>(lcnt[k]) / 10.0, 1.0);
          const double rad = cx.lig_radius[l];
          sum_d -= asp * burial * rad * rad * 4.0 * kPI;
        }
      }
    }
  }
  __syncthreads();
  for (int j = tid; j < cx.n_rec; j += kScoreThreads) {  // fitness.rs:141-162
    const double asp = cx.rec_asp[j];
    if (fabs(asp) >= 0.001) {
      const double burial = fmin(static_cast<double>(s_rcnt[j]) / 10.0, 1.0);
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
  if (tid == 0) {
    double vdw = 0.0, elec = 0.0, desolv = 0.0;
    for (int i = 0; i < kScoreThreads / 64; ++i) {
      vdw += s_red[3 * i]; elec += s_red[3 * i + 1]; desolv += s_red[3 * i + 2];
    }
    // fitness.rs:205-267, restraints.rs:11-40: FP64 CA-CA flat-bottom term
    double air = 0.0;
    int32_t nsat = 0;
    for (int i = 0; i < cx.n_restr; ++i) {
      double px, py, pz;
      ligand_atom(cx, pc, coords, pose, cx.restr_lig[i], px, py, pz);
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
    double* o = out + static_cast<size_t>(kOutDoubles) * pose;
    o[0] = vdw;
    o[1] = elec;
    o[2] = desolv;
    o[3] = air;
    o[4] = (1.0 - ratio) * n;  // chromosome.rs:117-118
    o[5] = w.vdw * vdw + w.elec * elec + w.desolv * desolv + w.air * air;  // chromosome.rs:121-124
    if (nsat_out) nsat_out[pose] = nsat;
  }
}

size_t score_lds_bytes(const DevComplex& cx) {
  return static_cast<size_t>((cx.n_rec + 1) & ~1) * sizeof(int32_t) + 3 * (kScoreThreads / 64) * sizeof(double);
}

int score_atoms_per_lane(int n_lig) {
  int k = (n_lig + kScoreThreads - 1) / kScoreThreads;
  return k < 1 ? 1 : (k > kMaxAtomsPerLane ? kMaxAtomsPerLane : k);
}

hipError_t launch_score(const DevComplex& cx, const double* d_poses, const double* d_coords,
                        const double* d_extent, int n_poses, const Weights& w, double* d_out,
                        int32_t* d_nsat, hipStream_t stream) {
  if (n_poses <= 0) return hipSuccess;
  const size_t lds = score_lds_bytes(cx);
  const dim3 grid(n_poses), block(kScoreThreads);
  switch (score_atoms_per_lane(cx.n_lig)) {
#define GDOCK_CASE(KK) \
  case KK: hipLaunchKernelGGL(score_poses_kernel<KK>, grid, block, lds, stream, cx, d_poses, d_coords, d_extent, w, d_out, d_nsat); break;
    GDOCK_CASE(1) GDOCK_CASE(2) GDOCK_CASE(3) GDOCK_CASE(4)
    GDOCK_CASE(5) GDOCK_CASE(6) GDOCK_CASE(7) GDOCK_CASE(8)
#undef GDOCK_CASE
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace gdock
