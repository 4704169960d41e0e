// Device GA (ga_device.hip): per-job state in device memory and the kernels
// that advance it, shared by the single-job entry point and the server.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/gdock_hip.h"
#include "ga.h"
#include "internal.h"
#include "stream.h"

namespace gdock {

constexpr int kGaThreads = 256;          // one workgroup per job
constexpr int kGaMaxPopulation = 1024;   // a job's population in one workgroup's LDS
constexpr int kGaHofCapacity = 500;      // run.rs:84-99
constexpr int kGaHofPerGeneration = 10;  // run.rs:194-216

struct DevGaParams {
  int32_t pop, max_gen, elitism, tsize, window, early, keep_trace, pad;
  double mut_rate, cross_rate, threshold, max_disp;
};

// A job's GA state between generations (device memory).
struct DevGaState {
  uint64_t rng[4];      // the project stream
  double last_best;     // best fitness of the previous generation
  int32_t n;            // population (the children of an odd population are one fewer)
  int32_t cur;          // which of the two gene buffers holds the population
  int32_t generation;   // next generation to score
  int32_t generations;  // generations scored
  int32_t converged_at; // -1 unless stopped early
  int32_t stall;        // consecutive generations below the improvement threshold
  int32_t done;         // 1 once the GA has ended (and, after its last generation, bred once more)
  int32_t hof_n, hof_sorted;
  int32_t trace_at;     // individuals traced so far
  int32_t base_ok;      // 1 when the job's rng_base holds T^i rng for i < kGaBaseStates
  int32_t pad;
};

// Kernel view of one job: its state, buffers, parameters, and where its
// pose records are read from and its scores are written to.
struct DevGaJob {
  DevGaState* st;
  double* genes[2];     // population genes, double-buffered (max_pop x 6)
  double* child;        // children before the elite replacement (max_pop x 6)
  double* mut_u;        // per child gene: the two uniforms of its mutation draw (2 x max_pop x 6)
  double* hof_genes;    // kGaHofCapacity + 1 entries
  double* hof_fit;
  int32_t* hof_gen;
  double* hof_tmp_genes;  // the first full sort's output
  double* hof_tmp_fit;
  int32_t* hof_tmp_gen;
  double* trace_genes;  // with keep_trace: per generation genes and fitness, and the best index
  double* trace_fit;
  int32_t* trace_best;
  double* poses;        // the generation's pose records (scored), then the next generation's
  const double* out;    // the generation's scores (kOutDoubles per pose)
  int32_t* done_flag;   // optional: set to the state's done flag after every generation
  uint64_t* rng_base;   // kGaBaseStates x 4: the stream's next kGaBaseStates states
  uint64_t* draws;      // kGaThreads x draw_span raw outputs of the generation's stream
  int32_t draw_span;    // outputs per lane the job's buffers allow (0: one lane draws them all)
  int32_t pad;
  DevGaParams p;
};

struct GaJobOffsets {
  size_t state, genes0, genes1, child, mut_u, hof_genes, hof_fit, hof_gen, tmp_genes, tmp_fit, tmp_gen;
  size_t rng_base, draws;
  size_t trace_genes, trace_fit, trace_best;
  int32_t draw_span;
};

// The generation's draws in the host driver's order, for a population of n:
// n x tsize tournament draws, n - 1 for the shuffle, 6 per crossed pair, then
// per child gene one mutation draw and, when it mutates, two more.  The
// workgroup takes them as kGaThreads lanes of draw_span consecutive outputs
// each, every lane jumping to its first (kGaBaseStates states past the last
// possible draw are also covered: the next generation's base states).  0
// when a population needs more than the jump table or the index buffer holds.
constexpr int kGaBaseStates = 256;
constexpr int kGaMaxIndexDraws = 8192;  // tournament and shuffle draws held as indices in LDS
__host__ __device__ inline int ga_draw_span(int n, int tsize) {
  if (n < 1 || tsize < 1) return 0;
  const long long idx = static_cast<long long>(n) * tsize + (n - 1);
  const long long worst = idx + 6LL * (n / 2) + 18LL * (n & ~1);
  const long long per = (worst + kGaBaseStates + kGaThreads - 1) / kGaThreads;
  const long long span = (per + stream::kJumpStride - 1) / stream::kJumpStride * stream::kJumpStride;
  if (idx > kGaMaxIndexDraws || (span / stream::kJumpStride) * (kGaThreads - 1) >= stream::kJumpEntries) return 0;
  return static_cast<int>(span);
}

// Device bytes and layout of one job's GA buffers (appended to L).
size_t ga_job_bytes(int max_pop, const gdock_ga_params& p);
void ga_job_layout(Layout* L, int max_pop, const gdock_ga_params& p, GaJobOffsets* o);
// The stream's jump table on the current device (once per device; call it
// before the first GA launch, outside graph capture).
hipError_t ga_device_init(hipStream_t s);
DevGaJob ga_job_bind(char* base, const GaJobOffsets& o, const gdock_ga_params& p, double* d_poses,
                     const double* d_out);
// The state after the host drew the initial population (job.genes()).
DevGaState ga_initial_state(const GaJob& job);
// Pose records of n gene vectors on the device (the device's sin and cos).
hipError_t launch_ga_pose_records(const double* d_genes, int n, double* d_poses, hipStream_t s);
// One generation of n_jobs jobs (d_jobs: device array), one workgroup each.
hipError_t launch_ga_generation(const DevGaJob* d_jobs, int n_jobs, hipStream_t s);
// Reads an ended job's results into the host driver's state.
hipError_t ga_job_download(const DevGaJob& J, const gdock_ga_params& p, hipStream_t s, GaJob* job);

}  // namespace gdock
