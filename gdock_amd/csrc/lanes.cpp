// Ligand atoms per lane (K) of the scoring kernels, chosen from the complex's
// shape alone.
//
// A ligand of n atoms is scored in ceil(n / (256 K)) chunks of 256 K lane
// slots; each chunk walks the whole receptor.  K fixes the FP32 summation
// order of a pose, so it must not depend on anything but the complex: the
// choice is a pure function of (ligand atoms, receptor slots), read from the
// committed table lanes_table.inc.  The autotuner (gdock_amd/tune.py) times
// the candidates of each tabulated shape with gdock_time_lanes and rewrites
// that table; it is never consulted at run time.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/gdock_hip.h"
#include "common.h"
#include "internal.h"

namespace gdock {

namespace {

// One row per measured shape: ligand atoms (the midpoint of a 128-atom
// interval, within which the candidate K values do not change: they change
// at multiples of 256, 512 and 768), receptor size class (half octaves of
// receptor atoms), chosen chunks.
struct LanesRow { int32_t n_lig, rec_class, chunks; };
#include "lanes_table.inc"

constexpr int kLigInterval = 128;

int rec_size_class(int n_rec_atoms) { return static_cast<int>(std::lround(2.0 * std::log2(std::max(n_rec_atoms, 1)))); }

int lanes_for_chunks(int n_lig, int chunks) {
  const int k = (n_lig + kScoreThreads * chunks - 1) / (kScoreThreads * chunks);
  return std::min(std::max(k, 1), kMaxAtomsPerLane);
}

}  // namespace

// The K of each chunk count 1..16 (the smallest K that covers the ligand in
// that many chunks), at most three distinct values: the autotuner's
// candidates.
std::vector<int> lane_candidates(int n_lig) {
  std::vector<int> ks;
  for (int chunks = 1; chunks <= 16; ++chunks) {
    const int k = (n_lig + kScoreThreads * chunks - 1) / (kScoreThreads * chunks);
    if (k < 1) break;
    if (k > kMaxAtomsPerLane) continue;
    if (std::find(ks.begin(), ks.end(), k) == ks.end()) ks.push_back(k);
    if (ks.size() >= 3) break;
  }
  if (ks.empty()) ks.push_back(1);
  return ks;
}

// The chunk count of the tabulated shape in the same 128-atom ligand
// interval with the nearest receptor size class (ties: the smaller class);
// without a row for the interval, the fewest chunks with K <= 6 (K = 7 and 8
// hold their ligand atoms in more registers than the 4-waves-per-SIMD
// budget leaves).
int choose_lanes(const DevComplex& d) {
  const int n_lig = std::max(d.n_lig, 1), cls = rec_size_class(d.n_rec_atoms);
  const int interval = (n_lig - 1) / kLigInterval;
  const LanesRow* best = nullptr;
  for (const LanesRow& r : kLanesTable) {
    if (r.rec_class < 0 || (r.n_lig - 1) / kLigInterval != interval) continue;
    if (!best || std::abs(r.rec_class - cls) < std::abs(best->rec_class - cls)) best = &r;
  }
  if (best) return lanes_for_chunks(n_lig, best->chunks);
  const int chunks = (n_lig + kScoreThreads * 6 - 1) / (kScoreThreads * 6);
  return lanes_for_chunks(n_lig, std::max(chunks, 1));
}

}  // namespace gdock

using namespace gdock;

extern "C" {

gdock_status gdock_time_lanes(gdock_ctx* ctx, gdock_complex* cx, int32_t lanes_k, int32_t n_poses, int32_t reps,
                              float* ms) {
  if (!ctx || !cx || !ms || lanes_k < 1 || lanes_k > kMaxAtomsPerLane || n_poses < 1 || n_poses > 65536 ||
      reps < 1) {
    if (ctx) set_error(ctx, "time_lanes: invalid arguments");
    return GDOCK_ERR_INVALID_ARG;
  }
  // Probe poses: fixed rotations, the ligand kept on the receptor (shifts of
  // up to 5 A), from a fixed LCG: the same poses for every candidate.
  std::vector<double> genes(static_cast<size_t>(n_poses) * 6), poses(static_cast<size_t>(n_poses) * kPoseDoubles);
  uint64_t s = 0x9e3779b97f4a7c15ull;
  auto u01 = [&s]() {
    s = s * 6364136223846793005ull + 1442695040888963407ull;
    return static_cast<double>(s >> 11) * (1.0 / 9007199254740992.0);
  };
  for (int32_t i = 0; i < n_poses; ++i) {
    for (int a = 0; a < 3; ++a) genes[6 * i + a] = 6.283185307179586 * u01();
    for (int a = 3; a < 6; ++a) genes[6 * i + a] = 10.0 * (u01() - 0.5);
  }
  gdock_prepare_poses(genes.data(), n_poses, poses.data());
  hipError_t e = hipSetDevice(ctx->device);
  const size_t scratch = plain_scratch_bytes(cx->dev, n_poses);
  char* d = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (e == hipSuccess)
    e = hipMalloc(reinterpret_cast<void**>(&d), poses.size() * 8 + static_cast<size_t>(n_poses) * kOutDoubles * 8 + scratch);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e == hipSuccess) e = hipMemcpyAsync(d, poses.data(), poses.size() * 8, hipMemcpyHostToDevice, ctx->stream);
  double* d_poses = reinterpret_cast<double*>(d);
  double* d_out = d_poses + poses.size();
  void* d_scratch = d_out + static_cast<size_t>(n_poses) * kOutDoubles;
  const Weights w{0.4, 0.05, 3.4, 1.0};
  if (e == hipSuccess)  // warm-up launch
    e = launch_score(cx->dev, cx->d_dev, d_poses, nullptr, n_poses, w, d_out, nullptr, d_scratch, ctx->stream, lanes_k);
  if (e == hipSuccess) e = hipEventRecord(e0, ctx->stream);
  for (int32_t r = 0; r < reps && e == hipSuccess; ++r)
    e = launch_score(cx->dev, cx->d_dev, d_poses, nullptr, n_poses, w, d_out, nullptr, d_scratch, ctx->stream, lanes_k);
  if (e == hipSuccess) e = hipEventRecord(e1, ctx->stream);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float t = 0.0f;
  if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (d) (void)hipFree(d);
  if (e != hipSuccess) return hip_fail(ctx, e, "time_lanes");
  *ms = t / static_cast<float>(reps);
  return GDOCK_OK;
}

int32_t gdock_choose_lanes(const gdock_complex* cx) { return cx ? choose_lanes(cx->dev) : -1; }

}  // extern "C"
