// The GA generation on the device: one workgroup per job takes its
// population's scores, offers the best to the Hall of Fame, tests
// convergence and breeds the next population (run.rs:381-531,
// population.rs:74-158, ga.rs:7-39, chromosome.rs:56-77), then writes the
// next generation's pose records, so consecutive generations need nothing
// from the host.  Every decision and every draw of the project stream
// (stream.h) follows the host driver (ga.cpp) step by step: a job run here is
// bit-identical to the same job run by the host driver.
//
// The parts that are sequential by nature - tournaments, the Fisher-Yates
// shuffle, crossover and mutation, whose draws come one after another from
// one stream - run on one lane; ranking, the Hall of Fame's similarity tests
// and sorts, trace copies and pose records run on the whole workgroup.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/gdock_hip.h"
#include "common.h"
#include "ga.h"
#include "ga_device.h"
#include "internal.h"
#include "stream.h"

namespace gdock {

namespace {
constexpr double kTwoPi = 2.0 * kPI;
constexpr int kHofPrepass = 256;  // best candidates whose Hall of Fame matches are counted up front
}  // namespace

// structure.rs:96-103 through the pose record of the scoring kernels: genes,
// then cos and sin of the three angles and the four mixed products (the
// bracketed sums, no contraction).  The device's sin and cos.
__device__ __forceinline__ void write_pose_record(const double* __restrict__ g, double* __restrict__ p) {
  for (int k = 0; k < 6; ++k) p[k] = g[k];
  double ca, sa, cb, sb, cg, sg;
  sincos(g[0], &sa, &ca);
  sincos(g[1], &sb, &cb);
  sincos(g[2], &sg, &cg);
  p[6] = ca; p[7] = sa; p[8] = cb; p[9] = sb; p[10] = cg; p[11] = sg;
  p[12] = (-sa * cg) + ((ca * sb) * sg);
  p[13] = (sa * sg) + ((ca * sb) * cg);
  p[14] = (ca * cg) + ((sa * sb) * sg);
  p[15] = (-ca * sg) + ((sa * sb) * cg);
}

__global__ void ga_pose_records_kernel(const double* __restrict__ genes, int n, double* __restrict__ poses) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) write_pose_record(genes + 6 * i, poses + static_cast<size_t>(kPoseDoubles) * i);
}

hipError_t launch_ga_pose_records(const double* d_genes, int n, double* d_poses, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(ga_pose_records_kernel, dim3((n + 255) / 256), dim3(256), 0, s, d_genes, n, d_poses);
  return hipGetLastError();
}

// run.rs:160-169 and 139-158: a pose within 0.2 rad per angle and 2 A per
// shift of a kept one is the same pose.
__device__ __forceinline__ bool hof_similar(const double* __restrict__ g, const double* __restrict__ h) {
  for (int i = 0; i < 3; ++i) {
    double d = fabs(g[i] - h[i]);
    d = d > kPI ? kTwoPi - d : d;
    if (!(d <= 0.2)) return false;
  }
  for (int i = 3; i < 6; ++i)
    if (!(fabs(g[i] - h[i]) <= 2.0)) return false;
  return true;
}

// Rust's f64::rem_euclid(a, 2 pi) as the host computes it (fmod, then + 2 pi
// when negative); fmod is exact, and for |a| < 4 pi it is a, a - 2 pi or
// a + 2 pi, each exact by Sterbenz's lemma.
__device__ __forceinline__ double rem_euclid_2pi(double a) {
  double r;
  if (fabs(a) < kTwoPi) r = a;
  else if (a >= kTwoPi && a < 2.0 * kTwoPi) r = a - kTwoPi;
  else if (a <= -kTwoPi && a > -2.0 * kTwoPi) r = a + kTwoPi;
  else r = fmod(a, kTwoPi);
  return r < 0.0 ? r + kTwoPi : r;
}

__device__ __forceinline__ void copy_entry(const DevGaJob& J, int from, double* tg, double* tf, int32_t* tn, int to) {
  for (int k = 0; k < 6; ++k) tg[6 * to + k] = J.hof_genes[6 * from + k];
  tf[to] = J.hof_fit[from];
  tn[to] = J.hof_gen[from];
}

// The Hall of Fame holds kGaHofCapacity + 1 entries: the one just appended
// (run.rs:171-177) makes it sort by fitness, stably, and drop the last.  The
// first time the entries are in insertion order (a full stable sort through
// the spare arrays); from then on they stay sorted, and the stable sort of a
// sorted run plus one entry puts that entry after every entry of lower or
// equal fitness.
// The genes of the entry that leaves go to dropped (LDS).
__device__ void hof_trim(const DevGaJob& J, bool sorted, double* dropped) {
  const int tid = threadIdx.x, n = kGaHofCapacity + 1;
  if (!sorted) {
    for (int e = tid; e < n; e += kGaThreads) {
      const double fe = J.hof_fit[e];
      int r = 0;
      for (int j = 0; j < n; ++j) {
        const double fj = J.hof_fit[j];
        r += (fj < fe || (fj == fe && j < e)) ? 1 : 0;
      }
      copy_entry(J, e, J.hof_tmp_genes, J.hof_tmp_fit, J.hof_tmp_gen, r);
      if (r == n - 1)
        for (int k = 0; k < 6; ++k) dropped[k] = J.hof_genes[6 * e + k];
    }
    __syncthreads();
    for (int e = tid; e < kGaHofCapacity; e += kGaThreads) {
      for (int k = 0; k < 6; ++k) J.hof_genes[6 * e + k] = J.hof_tmp_genes[6 * e + k];
      J.hof_fit[e] = J.hof_tmp_fit[e];
      J.hof_gen[e] = J.hof_tmp_gen[e];
    }
    __syncthreads();
    return;
  }
  const double fn = J.hof_fit[kGaHofCapacity];
  int cnt = 0;
  for (int e = tid; e < kGaHofCapacity; e += kGaThreads) cnt += J.hof_fit[e] <= fn ? 1 : 0;
  __shared__ int s_cnt;
  if (tid == 0) s_cnt = 0;
  __syncthreads();
  if (cnt) atomicAdd(&s_cnt, cnt);
  __syncthreads();
  const int p = s_cnt;
  if (tid == 0) {  // the last entry leaves, or the new one when it sorts last
    const int gone = p < kGaHofCapacity ? kGaHofCapacity - 1 : kGaHofCapacity;
    for (int k = 0; k < 6; ++k) dropped[k] = J.hof_genes[6 * gone + k];
  }
  if (p < kGaHofCapacity) {  // shift [p, capacity - 1) up by one, then place the new entry at p
    double g[2][6], f[2];
    int32_t gen[2];
    int m = 0;
    for (int e = p + tid; e < kGaHofCapacity - 1; e += kGaThreads, ++m) {
      for (int k = 0; k < 6; ++k) g[m][k] = J.hof_genes[6 * e + k];
      f[m] = J.hof_fit[e];
      gen[m] = J.hof_gen[e];
    }
    __syncthreads();
    m = 0;
    for (int e = p + tid; e < kGaHofCapacity - 1; e += kGaThreads, ++m) {
      for (int k = 0; k < 6; ++k) J.hof_genes[6 * (e + 1) + k] = g[m][k];
      J.hof_fit[e + 1] = f[m];
      J.hof_gen[e + 1] = gen[m];
    }
    __syncthreads();
    if (tid == 0) copy_entry(J, kGaHofCapacity, J.hof_genes, J.hof_fit, J.hof_gen, p);
  }
  __syncthreads();
}

// One generation of every listed job (one workgroup each); see the file
// comment.  A job that has ended does nothing.
// x^(kJumpStride k) mod P per k (stream.h), loaded by ga_device_init.
__device__ uint64_t g_stream_jump[4 * stream::kJumpEntries];

// LDS of one phase at a time: the Hall of Fame's offers, then the draws.
union GaPhaseLds {
  struct {
    double hof[6][kGaHofCapacity + 1];       // the entries at the generation's start, gene-major
    double cand[kHofPrepass * 6];             // the best candidates, in order
    double kept[kGaHofPerGeneration * 6];     // entries kept this generation
    double dropped[6];                        // the entry the last trim removed
    uint64_t unmatched[kHofPrepass / 64];     // bit r: candidate r matches no start entry still present
    int32_t cnt[kHofPrepass];                 // per candidate: start entries similar to it, still present
    int32_t gone[kGaHofPerGeneration];        // kept this generation, then removed by a trim
  } h;
  struct {
    uint64_t base[kGaBaseStates][4];          // T^i s for the stream state s at the generation's start
    uint64_t nib[kGaBaseStates / 4][16][4];   // [c][v]: XOR of base[4 c + b] over the set bits b of v
    uint16_t idx[kGaMaxIndexDraws];           // tournament and shuffle draws as drawn indices
    uint32_t flag[kGaThreads * stream::kJumpStride * (stream::kJumpEntries / kGaThreads) / 32];  // unit() <= rate
    uint32_t fn[kGaThreads / 64];             // per wave: its lanes' mutation-draw automaton, composed
    int32_t cnt[kGaThreads / 64];             // per wave: mutation decisions
    int32_t reject, end;                      // a bounded draw Lemire would repeat; the position after the last draw
  } d;
};

// Mutation draws: a child gene's draw is followed by its normal's two
// uniforms when it mutates.  State z before a draw: 0 when it decides a gene,
// else the uniforms still to pass; one draw maps z = 0 to 2 (mutates) or 0,
// 1 to 0 and 2 to 1.  A map of {0, 1, 2} is held as three 2-bit fields, so
// runs of draws compose (in order) and scan across lanes.
constexpr uint32_t kZIdentity = 0u | (1u << 2) | (2u << 4);
__device__ __forceinline__ uint32_t z_draw(bool mutates) { return (mutates ? 2u : 0u) | (0u << 2) | (1u << 4); }
__device__ __forceinline__ uint32_t z_then(uint32_t first, uint32_t second) {
  uint32_t r = 0;
#pragma unroll
  for (int z = 0; z < 3; ++z) r |= ((second >> (2 * ((first >> (2 * z)) & 3u))) & 3u) << (2 * z);
  return r;
}

__device__ __forceinline__ double unit_of(uint64_t x) { return static_cast<double>(x >> 11) * 0x1.0p-53; }

__device__ __forceinline__ bool same_genes(const double* a, const double* b) {
  bool eq = true;
  for (int k = 0; k < 6; ++k) eq = eq && stream::bits_of(a[k]) == stream::bits_of(b[k]);
  return eq;
}

__global__ __launch_bounds__(kGaThreads) void ga_generation_kernel(const DevGaJob* __restrict__ jobs) {
  __shared__ GaPhaseLds sh;
  __shared__ double s_fit[kGaMaxPopulation];    // fitness of the population; then the children's inherited fitness
  __shared__ double s_cfit[kGaMaxPopulation];
  __shared__ int32_t s_order[kGaMaxPopulation];  // ascending stable order of the population by fitness
  __shared__ int32_t s_chosen[kGaMaxPopulation];
  __shared__ int32_t s_pos[kGaMaxPopulation];    // each child's place in the descending stable order
  const DevGaJob J = jobs[blockIdx.x];
  DevGaState* st = J.st;
  const int tid = threadIdx.x;
  if (st->done) return;
  const int n = st->n, cur = st->cur, gen = st->generation;
  const double* __restrict__ genes = J.genes[cur];
  const DevGaParams P = J.p;
  for (int i = tid; i < n; i += kGaThreads) s_fit[i] = J.out[static_cast<size_t>(kOutDoubles) * i + 5];
  __syncthreads();
  for (int i = tid; i < n; i += kGaThreads) {
    const double fi = s_fit[i];
    int r = 0;
#pragma unroll 8
    for (int j = 0; j < n; ++j) {
      const double fj = s_fit[j];
      r += (fj < fi || (fj == fi && j < i)) ? 1 : 0;
    }
    s_order[r] = i;
  }
  __syncthreads();
  const int trace_at = st->trace_at;
  if (P.keep_trace) {
    for (int i = tid; i < 6 * n; i += kGaThreads) J.trace_genes[6 * static_cast<size_t>(trace_at) + i] = genes[i];
    for (int i = tid; i < n; i += kGaThreads) J.trace_fit[trace_at + i] = s_fit[i];
    if (tid == 0) J.trace_best[gen] = s_order[0];
  }
  // run.rs:194-216: the population's best, in order, until 10 are kept.  A
  // candidate is kept unless it is similar to an entry present when it is
  // offered: one of the generation's start (matches counted up front for the
  // best kHofPrepass candidates, every pair at once, less the matches of
  // start entries a trim removes later) or one kept earlier in the generation.
  int hof_n = st->hof_n;
  bool sorted = st->hof_sorted != 0;
  int added = 0, n_kept = 0;
  const int npre = min(n, kHofPrepass);
  for (int i = tid; i < 6 * hof_n; i += kGaThreads) sh.h.hof[i % 6][i / 6] = J.hof_genes[i];
  for (int i = tid; i < 6 * npre; i += kGaThreads) sh.h.cand[i] = genes[6 * s_order[i / 6] + i % 6];
  if (tid < kHofPrepass / 64) sh.h.unmatched[tid] = 0;
  __syncthreads();
  {  // a wave takes 4 candidates at a time, its lanes over the start entries
    const int lane = tid & 63;
    for (int r0 = 4 * (tid >> 6); r0 < npre; r0 += 4 * (kGaThreads / 64)) {
      double g[4][6];
#pragma unroll
      for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int k = 0; k < 6; ++k) g[c][k] = r0 + c < npre ? sh.h.cand[6 * (r0 + c) + k] : 0.0;
      int cnt[4] = {0, 0, 0, 0};
      for (int e0 = 0; e0 < hof_n; e0 += 64) {
        const int e = e0 + lane;
        const bool valid = e < hof_n;
        double h[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) h[k] = valid ? sh.h.hof[k][e] : 0.0;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          bool sim = valid;
#pragma unroll
          for (int k = 0; k < 3; ++k) {
            double d = fabs(g[c][k] - h[k]);
            d = d > kPI ? kTwoPi - d : d;
            sim &= d <= 0.2;
          }
#pragma unroll
          for (int k = 3; k < 6; ++k) sim &= fabs(g[c][k] - h[k]) <= 2.0;
          cnt[c] += __popcll(__builtin_amdgcn_ballot_w64(sim));
        }
      }
      if (lane < 4 && r0 + lane < npre) {
        const int c = lane == 0 ? cnt[0] : (lane == 1 ? cnt[1] : (lane == 2 ? cnt[2] : cnt[3]));
        const int r = r0 + lane;
        sh.h.cnt[r] = c;
        if (c == 0) atomicOr(reinterpret_cast<unsigned long long*>(&sh.h.unmatched[r >> 6]), 1ull << (r & 63));
      }
    }
  }
  __syncthreads();
  int r = 0;
  while (r < n && added < kGaHofPerGeneration) {
    double g[6];
    if (r < npre) {  // the next candidate no start entry matches (LDS alone: every thread decides alike)
      int nx = npre;
      for (int w = r >> 6; w < kHofPrepass / 64; ++w) {
        const uint64_t m = sh.h.unmatched[w] & (w == (r >> 6) ? ~0ull << (r & 63) : ~0ull);
        if (m) {
          nx = 64 * w + __builtin_ctzll(m);
          break;
        }
      }
      r = nx;
      if (r >= npre) continue;
      for (int k = 0; k < 6; ++k) g[k] = sh.h.cand[6 * r + k];
      bool sim = false;
      for (int a = 0; a < n_kept && !sim; ++a) sim = !sh.h.gone[a] && hof_similar(g, sh.h.kept + 6 * a);
      if (sim) {
        ++r;
        continue;
      }
    } else {
      for (int k = 0; k < 6; ++k) g[k] = genes[6 * s_order[r] + k];
      bool sim = false;
      for (int e = tid; e < hof_n && !sim; e += kGaThreads) sim = hof_similar(g, J.hof_genes + 6 * e);
      if (__syncthreads_or(sim ? 1 : 0)) {
        ++r;
        continue;
      }
    }
    const int i = s_order[r];
    if (tid == 0) {
      for (int k = 0; k < 6; ++k) J.hof_genes[6 * hof_n + k] = g[k];
      J.hof_fit[hof_n] = s_fit[i];
      J.hof_gen[hof_n] = gen;
      for (int k = 0; k < 6; ++k) sh.h.kept[6 * n_kept + k] = g[k];
      sh.h.gone[n_kept] = 0;
    }
    ++hof_n;
    ++added;
    ++n_kept;
    __syncthreads();
    if (hof_n > kGaHofCapacity) {
      hof_trim(J, sorted, sh.h.dropped);
      sorted = true;
      hof_n = kGaHofCapacity;
      // Entries are pairwise dissimilar, so equal genes name the entry.
      int which = -1;
      for (int a = 0; a < n_kept; ++a)
        if (!sh.h.gone[a] && same_genes(sh.h.kept + 6 * a, sh.h.dropped)) which = a;
      __syncthreads();  // every thread has read the list
      if (which >= 0) {
        if (tid == 0) sh.h.gone[which] = 1;
      } else {  // a start entry left: the later candidates it matched lose a match
        for (int q = r + 1 + tid; q < npre; q += kGaThreads)
          if (hof_similar(sh.h.cand + 6 * q, sh.h.dropped) && --sh.h.cnt[q] == 0)
            atomicOr(reinterpret_cast<unsigned long long*>(&sh.h.unmatched[q >> 6]), 1ull << (q & 63));
      }
      __syncthreads();
    }
    ++r;
  }
  // run.rs:480-531: the first individual of least fitness; the relative
  // improvement of the best against the threshold; early stop or the last
  // generation
  const int best = s_order[0];
  const double f = s_fit[best];
  int stall = st->stall, converged_at = st->converged_at, done = 0;
  if (gen > 0) {
    const double last = st->last_best;
    const double improvement = (last - f) / fabs(last);
    stall = improvement < P.threshold ? stall + 1 : 0;
  }
  int next_gen = gen;
  bool final_breed = false;
  if (P.early && stall >= P.window) {
    converged_at = gen;
    done = 1;
  } else {
    ++next_gen;
    if (next_gen >= P.max_gen) {
      done = 1;
      final_breed = true;  // the loop evolves once more after its last generation (run.rs:530)
    }
  }
  __syncthreads();  // every thread has read the state
  if (tid == 0) {
    st->hof_n = hof_n;
    st->hof_sorted = sorted ? 1 : 0;
    st->stall = stall;
    st->last_best = f;
    st->generations = gen + 1;
    st->generation = next_gen;
    st->converged_at = converged_at;
    st->trace_at = trace_at + n;
    st->done = done;
    if (J.done_flag) *J.done_flag = done;
  }
  if (done && !final_breed) return;

  // population.rs:74-158: breed the next population.  The generation's draws
  // come from one stream in the host driver's order (see ga_draw_span); the
  // workgroup takes them as kGaThreads runs of `span` consecutive outputs,
  // each lane jumping to its first through the stream's jump table, and keeps
  // only decisions: tournament winners, the shuffle, a swap bit per crossed
  // gene, a mutation bit per child gene with the two uniforms of its normal.
  // Where a draw's use depends on earlier draws (a mutation takes two more),
  // a scan over the lanes finds it.  Should a bounded draw land where
  // Lemire's method draws again (probability below n / 2^64 per draw), or the
  // population need more draws than the table reaches, one lane draws them
  // all in order instead.  The workgroup then builds the children from the
  // decisions: the same operations on the same values as the host driver, so
  // the same bits.
  const int n_child = n & ~1;
  const int elite = min(P.elitism, n_child);
  const int npairs = n / 2;
  __shared__ uint8_t s_xbits[kGaMaxPopulation / 2];  // per pair: genes swapped by the crossover
  __shared__ uint32_t s_mbits[kGaMaxPopulation];     // per child: genes mutated
  const int span = J.draws ? ga_draw_span(n, P.tsize) : 0;
  bool serial = !(span > 0 && span <= J.draw_span);
  if (!serial) {
    const int t_end = n * P.tsize, f_end = t_end + n - 1, x_end = f_end + 6 * npairs;
    const int R = kGaThreads * span, p0 = span * tid;
    if (st->base_ok) {
      for (int i = tid; i < 4 * kGaBaseStates; i += kGaThreads) sh.d.base[i / 4][i % 4] = J.rng_base[i];
    } else if (tid == 0) {
      stream::Xoshiro256 w;
      for (int k = 0; k < 4; ++k) w.s[k] = st->rng[k];
      for (int i = 0; i < kGaBaseStates; ++i) {
        for (int k = 0; k < 4; ++k) sh.d.base[i][k] = w.s[k];
        w.next();
      }
    }
    for (int i = tid; i < R / 32; i += kGaThreads) sh.d.flag[i] = 0;
    for (int c = tid; c < n_child; c += kGaThreads) s_mbits[c] = 0;
    if (tid == 0) {
      sh.d.reject = 0;
      sh.d.end = x_end;
    }
    const uint64_t* jq = g_stream_jump + 4 * (span / stream::kJumpStride) * tid;
    const uint64_t q0 = jq[0], q1 = jq[1], q2 = jq[2], q3 = jq[3];
    __syncthreads();
    for (int i = tid; i < kGaBaseStates * 4; i += kGaThreads) {  // the XOR of every subset of 4 base states
      const int c = i >> 4, v = i & 15;
      uint64_t a[4] = {0, 0, 0, 0};
      for (int b = 0; b < 4; ++b)
        if ((v >> b) & 1)
          for (int k = 0; k < 4; ++k) a[k] ^= sh.d.base[4 * c + b][k];
      for (int k = 0; k < 4; ++k) sh.d.nib[c][v][k] = a[k];
    }
    __syncthreads();
    // The lane's first state (its jump polynomial, 4 bits at a time), then
    // its outputs [p0, p0 + span): indices of the bounded draws, bits of the
    // uniform ones, and the raw mutation-region outputs.
    uint64_t first[4] = {0, 0, 0, 0};
#pragma unroll
    for (int c = 0; c < kGaBaseStates / 4; ++c) {
      const uint64_t qw = c < 16 ? q0 : (c < 32 ? q1 : (c < 48 ? q2 : q3));
      const int v = static_cast<int>((qw >> (4 * (c & 15))) & 15u);
      for (int k = 0; k < 4; ++k) first[k] ^= sh.d.nib[c][v][k];
    }
    {
      stream::Xoshiro256 w;
      for (int k = 0; k < 4; ++k) w.s[k] = first[k];
      for (int j = 0; j < span; ++j) {
        const int p = p0 + j;
        const uint64_t x = w.next();
        if (p < f_end) {  // the tournaments' n, then the shuffle's n, n - 1, ..., 2
          const uint64_t bound = static_cast<uint64_t>(p < t_end ? n : n - (p - t_end));
          const unsigned __int128 m = static_cast<unsigned __int128>(x) * bound;
          if (static_cast<uint64_t>(m) < bound) sh.d.reject = 1;
          sh.d.idx[p] = static_cast<uint16_t>(m >> 64);
        } else {
          if (unit_of(x) <= (p < x_end ? P.cross_rate : P.mut_rate)) atomicOr(&sh.d.flag[p >> 5], 1u << (p & 31));
          if (p >= x_end) J.draws[p] = x;
        }
      }
    }
    __syncthreads();
    serial = sh.d.reject != 0;
    if (!serial) {
      auto flag = [&](int p) { return ((sh.d.flag[p >> 5] >> (p & 31)) & 1u) != 0; };
      for (int k = tid; k < n; k += kGaThreads) {  // tournaments: the first of least fitness wins
        int b = -1;
        for (int t = 0; t < P.tsize; ++t) {
          const int c = sh.d.idx[k * P.tsize + t];
          if (b < 0 || s_fit[c] < s_fit[b]) b = c;
        }
        s_chosen[k] = b;
      }
      for (int m = tid; m < npairs; m += kGaThreads) {  // uniform crossover of neighbours
        uint32_t bits = 0;
        for (int g = 0; g < 6; ++g) bits |= (flag(f_end + 6 * m + g) ? 1u : 0u) << g;
        s_xbits[m] = static_cast<uint8_t>(bits);
      }
      // the lane's mutation-region draws as one map of z, scanned over the lanes
      const int lo = max(p0, x_end), hi = p0 + span, lane = tid & 63, wave = tid >> 6;
      uint32_t fn = kZIdentity;
      for (int p = lo; p < hi; ++p) fn = z_then(fn, z_draw(flag(p)));
      uint32_t inc = fn;
      for (int off = 1; off < 64; off <<= 1) {
        const uint32_t u = static_cast<uint32_t>(__shfl_up(static_cast<int>(inc), off));
        if (lane >= off) inc = z_then(u, inc);
      }
      uint32_t pre = static_cast<uint32_t>(__shfl_up(static_cast<int>(inc), 1));
      if (lane == 0) pre = kZIdentity;
      if (lane == 63) sh.d.fn[wave] = inc;
      __syncthreads();
      // Fisher-Yates: swapping places i - 1 and j_i for i = n down to 2
      // leaves at place x the winner from place tau_n(...tau_2(x)...), tau_i
      // the transposition (i - 1, j_i); each place follows its own.
      for (int x = tid; x < n; x += kGaThreads) {
        int y = x;
#pragma unroll 8
        for (int i = 2; i <= n; ++i) {
          const int j = sh.d.idx[t_end + (n - i)];
          y = y == i - 1 ? j : (y == j ? i - 1 : y);
        }
        s_pos[x] = s_chosen[y];
      }
      uint32_t before = kZIdentity;
      for (int v = 0; v < wave; ++v) before = z_then(before, sh.d.fn[v]);
      const uint32_t z0 = z_then(before, pre) & 3u;  // z at the lane's first mutation-region draw
      int cnt = 0;
      uint32_t z = z0;
      for (int p = lo; p < hi; ++p) {
        cnt += z == 0u ? 1 : 0;
        z = (z_draw(flag(p)) >> (2 * z)) & 3u;
      }
      int csum = cnt;
      for (int off = 1; off < 64; off <<= 1) {
        const int u = __shfl_up(csum, off);
        if (lane >= off) csum += u;
      }
      if (lane == 63) sh.d.cnt[wave] = csum;
      __syncthreads();
      int k = csum - cnt;  // decisions before the lane's first draw
      for (int v = 0; v < wave; ++v) k += sh.d.cnt[v];
      const int n_dec = 6 * n_child;
      z = z0;
      for (int p = lo; p < hi && k < n_dec; ++p) {  // mutation: Box-Muller takes 1 - u1 and u2
        const bool f = flag(p);
        if (z == 0u) {
          if (f) {
            atomicOr(&s_mbits[k / 6], 1u << (k % 6));
            J.mut_u[2 * k] = 1.0 - unit_of(J.draws[p + 1]);
            J.mut_u[2 * k + 1] = unit_of(J.draws[p + 2]);
          }
          if (k == n_dec - 1) sh.d.end = p + (f ? 3 : 1);
          ++k;
        }
        z = (z_draw(f) >> (2 * z)) & 3u;
      }
      __syncthreads();
      for (int c = tid; c < n; c += kGaThreads) s_chosen[c] = s_pos[c];
      for (int c = tid; c < n_child; c += kGaThreads) s_cfit[c] = s_fit[s_pos[c]];
      // the stream after the generation's last draw, and the next base states
      const int D = sh.d.end, qlo = max(D, p0), qhi = min(D + kGaBaseStates, p0 + span);
      if (qlo < qhi) {
        stream::Xoshiro256 w;
        for (int kk = 0; kk < 4; ++kk) w.s[kk] = first[kk];
        for (int q = p0; q < qlo; ++q) w.next();
        for (int q = qlo; q < qhi; ++q) {
          for (int kk = 0; kk < 4; ++kk) J.rng_base[4 * (q - D) + kk] = w.s[kk];
          if (q == D)
            for (int kk = 0; kk < 4; ++kk) st->rng[kk] = w.s[kk];
          w.next();
        }
      }
    }
  }
  if (serial && tid == 0) {
    stream::Xoshiro256 rng;
    for (int k = 0; k < 4; ++k) rng.s[k] = st->rng[k];
    for (int k = 0; k < n; ++k) {  // tournaments: the first of least fitness wins
      int b = -1;
      for (int t = 0; t < P.tsize; ++t) {
        const int c = static_cast<int>(rng.below(static_cast<uint64_t>(n)));
        if (b < 0 || s_fit[c] < s_fit[b]) b = c;
      }
      s_chosen[k] = b;
    }
    for (int i = n; i > 1; --i) {  // Fisher-Yates
      const int j = static_cast<int>(rng.below(static_cast<uint64_t>(i)));
      const int t = s_chosen[i - 1];
      s_chosen[i - 1] = s_chosen[j];
      s_chosen[j] = t;
    }
    for (int m = 0; 2 * m + 1 < n; ++m) {  // uniform crossover of neighbours
      uint32_t bits = 0;
      for (int g = 0; g < 6; ++g) bits |= (rng.unit() <= P.cross_rate ? 1u : 0u) << g;
      s_xbits[m] = static_cast<uint8_t>(bits);
      s_cfit[2 * m] = s_fit[s_chosen[2 * m]];
      s_cfit[2 * m + 1] = s_fit[s_chosen[2 * m + 1]];
    }
    for (int c = 0; c < n_child; ++c) {  // mutation: Box-Muller takes 1 - u1 and u2
      uint32_t bits = 0;
      for (int g = 0; g < 6; ++g) {
        if (!(rng.unit() <= P.mut_rate)) continue;
        bits |= 1u << g;
        const double u1 = 1.0 - rng.unit(), u2 = rng.unit();
        J.mut_u[2 * (6 * c + g)] = u1;
        J.mut_u[2 * (6 * c + g) + 1] = u2;
      }
      s_mbits[c] = bits;
    }
    for (int k = 0; k < 4; ++k) st->rng[k] = rng.s[k];
  }
  __syncthreads();
  for (int i = tid; i < 6 * n_child; i += kGaThreads) {
    const int c = i / 6, g = i - 6 * c, m = c >> 1;
    const int a = s_chosen[2 * m], b = s_chosen[2 * m + 1];
    const bool swap = (s_xbits[m] >> g) & 1u;
    const double x = genes[6 * a + g], y = genes[6 * b + g];
    double v = (c & 1) == 0 ? (swap ? y : x) : (swap ? x : y);
    if ((s_mbits[c] >> g) & 1u) {
      const double u1 = J.mut_u[2 * i], u2 = J.mut_u[2 * i + 1];
      const double sd = g < 3 ? 0.174 : 1.0;
      const double z = 0.0 + sd * (stream::sqrt_cr(-2.0 * stream::ln(u1)) * stream::cos2pi(u2));
      v = g < 3 ? rem_euclid_2pi(v + z) : fmin(fmax(v + z, -P.max_disp), P.max_disp);
    }
    J.child[i] = v;
  }
  __syncthreads();
  // the elite replace the children that inherited the worst fitness: stable
  // descending order of the children, then the elite over its first places
  for (int c = tid; c < n_child; c += kGaThreads) {
    const double fc = s_cfit[c];
    int r = 0;
#pragma unroll 8
    for (int j = 0; j < n_child; ++j) {
      const double fj = s_cfit[j];
      r += (fj > fc || (fj == fc && j < c)) ? 1 : 0;
    }
    s_pos[c] = r;
  }
  __syncthreads();
  double* __restrict__ out_genes = J.genes[cur ^ 1];
  for (int c = tid; c < n_child; c += kGaThreads) {
    const int q = s_pos[c];
    if (q >= elite)
      for (int g = 0; g < 6; ++g) out_genes[6 * q + g] = J.child[6 * c + g];
  }
  for (int q = tid; q < elite; q += kGaThreads)
    for (int g = 0; g < 6; ++g) out_genes[6 * q + g] = genes[6 * s_order[q] + g];
  __syncthreads();
  if (!done)
    for (int q = tid; q < n_child; q += kGaThreads)
      write_pose_record(out_genes + 6 * q, J.poses + static_cast<size_t>(kPoseDoubles) * q);
  if (tid == 0) {
    st->n = n_child;
    st->cur = cur ^ 1;
    st->base_ok = serial ? 0 : 1;
  }
}

hipError_t launch_ga_generation(const DevGaJob* d_jobs, int n_jobs, hipStream_t s) {
  if (n_jobs <= 0) return hipSuccess;
  hipLaunchKernelGGL(ga_generation_kernel, dim3(n_jobs), dim3(kGaThreads), 0, s, d_jobs);
  return hipGetLastError();
}

// ---------------------------------------------------------------- host side
namespace {
std::mutex g_jump_mu;
std::vector<uint64_t> g_jump_host;  // empty until built; stays empty if P fails its check
bool g_jump_built = false;
uint64_t g_jump_devices = 0;        // bit d: loaded on device d

// The table is on the current device (then the kernel may take the lanes).
bool jump_table_ready() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return false;
  std::lock_guard<std::mutex> g(g_jump_mu);
  return (g_jump_devices >> dev) & 1u;
}
}  // namespace

hipError_t ga_device_init(hipStream_t s) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  std::lock_guard<std::mutex> g(g_jump_mu);
  if (!g_jump_built) {
    std::vector<uint64_t> t(4 * static_cast<size_t>(stream::kJumpEntries));
    if (stream::jump_polys(t.data())) g_jump_host.swap(t);
    g_jump_built = true;
  }
  if (g_jump_host.empty() || dev < 0 || dev >= 64 || ((g_jump_devices >> dev) & 1u)) return hipSuccess;
  e = hipMemcpyToSymbolAsync(HIP_SYMBOL(g_stream_jump), g_jump_host.data(), g_jump_host.size() * sizeof(uint64_t), 0,
                             hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e == hipSuccess) g_jump_devices |= 1ull << dev;
  return e;
}

size_t ga_job_bytes(int max_pop, const gdock_ga_params& p) {
  Layout L;
  ga_job_layout(&L, max_pop, p, nullptr);
  return L.size;
}

void ga_job_layout(Layout* L, int max_pop, const gdock_ga_params& p, GaJobOffsets* o) {
  const int max_gen = p.max_generations;
  const bool trace = p.keep_trace != 0;
  GaJobOffsets t{};
  t.state = L->add<DevGaState>(1);
  t.genes0 = L->add<double>(static_cast<size_t>(max_pop) * 6);
  t.genes1 = L->add<double>(static_cast<size_t>(max_pop) * 6);
  t.child = L->add<double>(static_cast<size_t>(max_pop) * 6);
  t.mut_u = L->add<double>(static_cast<size_t>(max_pop) * 12);
  t.hof_genes = L->add<double>((kGaHofCapacity + 1) * 6);
  t.hof_fit = L->add<double>(kGaHofCapacity + 1);
  t.hof_gen = L->add<int32_t>(kGaHofCapacity + 1);
  t.tmp_genes = L->add<double>((kGaHofCapacity + 1) * 6);
  t.tmp_fit = L->add<double>(kGaHofCapacity + 1);
  t.tmp_gen = L->add<int32_t>(kGaHofCapacity + 1);
  t.draw_span = ga_draw_span(max_pop, p.tournament_size);  // populations only shrink
  if (t.draw_span > 0) {
    t.rng_base = L->add<uint64_t>(4 * kGaBaseStates);
    t.draws = L->add<uint64_t>(static_cast<size_t>(kGaThreads) * t.draw_span);
  }
  if (trace) {
    t.trace_genes = L->add<double>(static_cast<size_t>(max_gen) * max_pop * 6);
    t.trace_fit = L->add<double>(static_cast<size_t>(max_gen) * max_pop);
    t.trace_best = L->add<int32_t>(max_gen);
  }
  if (o) *o = t;
}

DevGaJob ga_job_bind(char* base, const GaJobOffsets& o, const gdock_ga_params& p, double* d_poses,
                     const double* d_out) {
  DevGaJob J{};
  J.st = reinterpret_cast<DevGaState*>(base + o.state);
  J.genes[0] = reinterpret_cast<double*>(base + o.genes0);
  J.genes[1] = reinterpret_cast<double*>(base + o.genes1);
  J.child = reinterpret_cast<double*>(base + o.child);
  J.mut_u = reinterpret_cast<double*>(base + o.mut_u);
  J.hof_genes = reinterpret_cast<double*>(base + o.hof_genes);
  J.hof_fit = reinterpret_cast<double*>(base + o.hof_fit);
  J.hof_gen = reinterpret_cast<int32_t*>(base + o.hof_gen);
  J.hof_tmp_genes = reinterpret_cast<double*>(base + o.tmp_genes);
  J.hof_tmp_fit = reinterpret_cast<double*>(base + o.tmp_fit);
  J.hof_tmp_gen = reinterpret_cast<int32_t*>(base + o.tmp_gen);
  if (p.keep_trace) {
    J.trace_genes = reinterpret_cast<double*>(base + o.trace_genes);
    J.trace_fit = reinterpret_cast<double*>(base + o.trace_fit);
    J.trace_best = reinterpret_cast<int32_t*>(base + o.trace_best);
  }
  if (o.draw_span > 0 && jump_table_ready()) {
    J.rng_base = reinterpret_cast<uint64_t*>(base + o.rng_base);
    J.draws = reinterpret_cast<uint64_t*>(base + o.draws);
    J.draw_span = o.draw_span;
  }
  J.poses = d_poses;
  J.out = d_out;
  J.p.pop = p.population_size;
  J.p.max_gen = p.max_generations;
  J.p.elitism = p.elitism;
  J.p.tsize = p.tournament_size;
  J.p.window = p.convergence_window;
  J.p.early = p.early_stopping;
  J.p.keep_trace = p.keep_trace;
  J.p.mut_rate = p.mutation_rate;
  J.p.cross_rate = p.crossover_rate;
  J.p.threshold = p.convergence_threshold;
  J.p.max_disp = p.max_displacement;
  return J;
}

DevGaState ga_initial_state(const GaJob& job) {
  DevGaState s{};
  for (int k = 0; k < 4; ++k) s.rng[k] = job.rng_state()[k];
  s.n = job.population();
  s.cur = 0;
  s.generation = 0;
  s.generations = 0;
  s.converged_at = -1;
  s.stall = 0;
  s.done = 0;
  s.hof_n = 0;
  s.hof_sorted = 0;
  s.trace_at = 0;
  s.base_ok = 0;
  s.last_best = 1.7976931348623157e308;
  return s;
}

hipError_t ga_job_download(const DevGaJob& J, const gdock_ga_params& p, hipStream_t s, GaJob* job) {
  DevGaState st{};
  hipError_t e = hipMemcpyAsync(&st, J.st, sizeof st, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return e;
  const int hn = st.hof_n, n = st.n;
  std::vector<double> hg(static_cast<size_t>(hn) * 6), hf(hn), fin(static_cast<size_t>(n) * 6);
  std::vector<int32_t> hgen(hn);
  std::vector<double> tg, tf;
  std::vector<int32_t> tb;
  if (hn) {
    e = hipMemcpyAsync(hg.data(), J.hof_genes, hg.size() * 8, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(hf.data(), J.hof_fit, hf.size() * 8, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(hgen.data(), J.hof_gen, hgen.size() * 4, hipMemcpyDeviceToHost, s);
  }
  if (e == hipSuccess && n) e = hipMemcpyAsync(fin.data(), J.genes[st.cur], fin.size() * 8, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess && p.keep_trace) {
    tg.resize(static_cast<size_t>(st.trace_at) * 6);
    tf.resize(st.trace_at);
    tb.resize(st.generations);
    if (!tg.empty()) e = hipMemcpyAsync(tg.data(), J.trace_genes, tg.size() * 8, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess && !tf.empty()) e = hipMemcpyAsync(tf.data(), J.trace_fit, tf.size() * 8, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess && !tb.empty()) e = hipMemcpyAsync(tb.data(), J.trace_best, tb.size() * 4, hipMemcpyDeviceToHost, s);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return e;
  std::vector<HofEntry> hof(hn);
  for (int i = 0; i < hn; ++i) {
    std::memcpy(hof[i].genes, &hg[6 * i], 6 * sizeof(double));
    hof[i].fitness = hf[i];
    hof[i].generation = static_cast<uint32_t>(hgen[i]);
  }
  job->load_device_run(st.generations, st.converged_at, std::move(hof), std::move(fin), std::move(tg), std::move(tf),
                       std::move(tb));
  return hipSuccess;
}

}  // namespace gdock

using namespace gdock;

extern "C" {

gdock_status gdock_ga_run_device(gdock_ctx* ctx, gdock_complex* cx, const gdock_weights* w, uint64_t seed,
                                 const gdock_ga_params* params, gdock_ga_result** out) {
  if (!ctx || !cx || !w || !out) {
    if (ctx) set_error(ctx, "ga_run_device: invalid arguments");
    return GDOCK_ERR_INVALID_ARG;
  }
  gdock_ga_params p;
  if (params) p = *params; else gdock_ga_default_params(&p);
  gdock_status st = validate_ga_params(ctx, &p);
  if (st != GDOCK_OK) return st;
  if (p.population_size > kGaMaxPopulation) {
    set_error(ctx, "ga_run_device: population %d exceeds the device GA's %d", p.population_size, kGaMaxPopulation);
    return GDOCK_ERR_INVALID_ARG;
  }
  hipError_t e = hipSetDevice(ctx->device);
  if (e != hipSuccess) return hip_fail(ctx, e, "ga_run_device: set device");
  auto* r = new gdock_ga_result(seed, p);  // draws the initial population, as the host driver does
  e = ga_device_init(ctx->stream);
  if (e != hipSuccess) { delete r; return hip_fail(ctx, e, "ga_run_device: jump table"); }
  const int pop = p.population_size;
  Layout L;
  GaJobOffsets o;
  ga_job_layout(&L, pop, p, &o);
  const size_t o_job = L.add<DevGaJob>(1), o_poses = L.add<double>(static_cast<size_t>(pop) * kPoseDoubles),
               o_out = L.add<double>(static_cast<size_t>(pop) * kOutDoubles),
               o_scr = L.add<uint8_t>(std::max(plain_scratch_bytes(cx->dev, pop, kGaFillItems),
                                               plain_scratch_bytes(cx->dev, pop & ~1, kGaFillItems)));
  char* m = nullptr;
  e = hipMalloc(reinterpret_cast<void**>(&m), L.size);
  if (e != hipSuccess) { delete r; return hip_fail(ctx, e, "ga_run_device: memory"); }
  const DevGaJob J = ga_job_bind(m, o, p, reinterpret_cast<double*>(m + o_poses), reinterpret_cast<double*>(m + o_out));
  const DevGaState s0 = ga_initial_state(r->job);
  const Weights ww{w->vdw, w->elec, w->desolv, w->air};
  e = hipMemcpyAsync(J.st, &s0, sizeof s0, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(J.genes[0], r->job.genes().data(), static_cast<size_t>(pop) * 6 * 8, hipMemcpyHostToDevice,
                       ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(m + o_job, &J, sizeof J, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = launch_ga_pose_records(J.genes[0], pop, J.poses, ctx->stream);
  // One plain scoring launch (filled as the host driver fills it) and one GA
  // launch per generation, kGaAhead generations between looks at the done
  // flag (the first alone: an odd population breeds one child fewer).
  // Launches past the end change nothing the result reads: the GA launch
  // returns at once and the scoring launch rescores the last pose records.
  constexpr int kGaAhead = 4;
  int n = pop, ahead = 1;
  while (e == hipSuccess) {
    for (int g = 0; g < ahead && e == hipSuccess; ++g) {
      e = launch_score(cx->dev, cx->d_dev, J.poses, nullptr, n, ww, const_cast<double*>(J.out), nullptr, m + o_scr,
                       ctx->stream, 0, nullptr, kGaFillItems);
      if (e == hipSuccess) e = launch_ga_generation(reinterpret_cast<const DevGaJob*>(m + o_job), 1, ctx->stream);
    }
    ahead = kGaAhead;
    int32_t dn[2] = {0, 0};
    if (e == hipSuccess) e = hipMemcpyAsync(&dn[0], &J.st->done, 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(&dn[1], &J.st->n, 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess || dn[0]) break;
    n = dn[1];
  }
  if (e == hipSuccess) e = ga_job_download(J, p, ctx->stream, &r->job);
  (void)hipFree(m);
  if (e != hipSuccess) { delete r; return hip_fail(ctx, e, "ga_run_device"); }
  if (p.cluster) {
    st = finish_job(ctx, cx, &r->job, &r->clusters);
    if (st != GDOCK_OK) { delete r; return st; }
  }
  *out = r;
  return GDOCK_OK;
}

}  // extern "C"
