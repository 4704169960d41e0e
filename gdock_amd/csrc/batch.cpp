// Batched scoring of many complexes in one persistent launch, replayed from a
// captured HIP graph, with a measured choice of ligand atoms per lane.
//
// A batch is a fixed set of complexes with a fixed number of poses each (one
// GA generation of every complex).  Each launch plans its work items from the
// poses (a pose split over receptor ranges when its estimated cost is high,
// items ordered most expensive first); one persistent kernel drains them from
// a device counter, and a second kernel rescores the rare poses whose settle
// buffer overflowed.  The launches are captured once into a graph; every
// generation replays it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <numeric>
#include <vector>

#include "../../include/gdock_hip.h"
#include "common.h"
#include "internal.h"

using namespace gdock;

struct gdock_batch {
  gdock_ctx* ctx = nullptr;
  std::vector<gdock_complex*> cx;
  std::vector<int32_t> pose_off;  // n + 1
  std::vector<int32_t> lanes_k;
  std::vector<int32_t> parts;
  std::vector<int2> pose_list;  // {complex, pose row} per pose
  int64_t item_capacity = 0;
  int32_t n_poses = 0;
  int32_t max_n_rec = 0;
  BatchLaunch L{};
  void* mem = nullptr;  // device: table, pose list, counts, poses, out, nsat, redo, items, plan
  double* h_poses = nullptr;  // pinned host staging
  double* h_out = nullptr;
  int32_t* h_nsat = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  Weights gw{};
  bool have_graph = false;
};

namespace gdock {

// Largest number of workgroups one pose of a complex may be split over in
// batched launches: enough parts of kSplitCost for a pose with every
// receptor slot within reach, at least 512 receptor slots each, at most
// kMaxPoseParts, and never fewer than min_batch_parts.  Each launch picks
// every pose's own count between the two (pose_plan_kernel).  A function of
// the complex's shape only.
int min_batch_parts(const DevComplex& d) {
  const int tiles = (d.n_rec + 63) / 64, per = kPartSlots / 64;
  return std::max(1, (tiles + per - 1) / per);
}

int pose_parts(const DevComplex& d, int lanes_k) {
  const int64_t per_chunk = static_cast<int64_t>(kScoreThreads) * std::max(lanes_k, 1);
  const double slots = static_cast<double>((d.n_lig + per_chunk - 1) / per_chunk * per_chunk);
  const double worst = 1.0625 * d.n_rec * slots;
  const int64_t p = std::min<int64_t>(static_cast<int64_t>(std::ceil(worst / kSplitCost)), d.n_rec / 512);
  const int64_t q = std::max<int64_t>(p, min_batch_parts(d));
  return static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(q, kMaxPoseParts)));
}

size_t split_state_bytes(const DevComplex& d, int parts, int poses) {
  if (parts <= 1 || poses <= 0) return 0;
  return (static_cast<size_t>(poses) * d.n_lig * 4 + 255) / 256 * 256 +
         (static_cast<size_t>(poses) * parts * 3 * 8 + 255) / 256 * 256 + (static_cast<size_t>(poses) * 2 * 4 + 255) / 256 * 256;
}

void bind_split_state(DevComplex* d, int parts, int poses, int pose_base, char* mem) {
  d->parts = parts;
  d->split_pose_base = pose_base;
  d->split_lig = nullptr;
  d->split_sums = nullptr;
  d->split_sync = nullptr;
  if (parts <= 1 || !mem) { d->parts = 1; return; }
  const size_t a = (static_cast<size_t>(poses) * d->n_lig * 4 + 255) / 256 * 256;
  const size_t b = (static_cast<size_t>(poses) * parts * 3 * 8 + 255) / 256 * 256;
  d->split_lig = reinterpret_cast<int32_t*>(mem);
  d->split_sums = reinterpret_cast<double*>(mem + a);
  d->split_sync = reinterpret_cast<int32_t*>(mem + a + b);
}

// Most receptor slots one item of a complex covers in batched launches (the
// LDS burial table of the batch kernel must hold them).
int batch_part_slots(const DevComplex& d) { return part_slots(d.n_rec, min_batch_parts(d)); }

}  // namespace gdock

extern "C" {

gdock_status gdock_batch_create(gdock_ctx* ctx, gdock_complex* const* complexes, int32_t n,
                                const int32_t* poses_per_complex, gdock_batch** out) {
  if (!ctx || !out || n < 0 || n > kItemSlotMask + 1 || (n > 0 && (!complexes || !poses_per_complex))) {
    if (ctx) set_error(ctx, "batch_create: invalid arguments");
    return GDOCK_ERR_INVALID_ARG;
  }
  *out = nullptr;
  int64_t total = 0;
  int max_rec = 1;
  for (int32_t i = 0; i < n; ++i) {
    if (!complexes[i] || poses_per_complex[i] < 0) {
      set_error(ctx, "batch_create: complex %d: null complex or negative pose count", i);
      return GDOCK_ERR_INVALID_ARG;
    }
    total += poses_per_complex[i];
    max_rec = std::max(max_rec, batch_part_slots(complexes[i]->dev));
  }
  if (total > (1 << 30)) {
    set_error(ctx, "batch_create: %lld poses is too many", static_cast<long long>(total));
    return GDOCK_ERR_INVALID_ARG;
  }
  hipError_t e = hipSetDevice(ctx->device);
  if (e != hipSuccess) return hip_fail(ctx, e, "batch_create: hipSetDevice");

  auto* b = new gdock_batch();
  b->ctx = ctx;
  b->cx.assign(complexes, complexes + n);
  b->pose_off.assign(n + 1, 0);
  for (int32_t i = 0; i < n; ++i) b->pose_off[i + 1] = b->pose_off[i] + poses_per_complex[i];
  b->n_poses = static_cast<int32_t>(total);
  b->max_n_rec = max_rec;
  b->lanes_k.resize(n);
  for (int32_t i = 0; i < n; ++i) b->lanes_k[i] = choose_lanes(b->cx[i]->dev);

  b->parts.resize(n);
  for (int32_t i = 0; i < n; ++i) {
    b->parts[i] = pose_parts(b->cx[i]->dev, b->lanes_k[i]);
    b->item_capacity += static_cast<int64_t>(poses_per_complex[i]) * b->parts[i];
  }
  if (b->item_capacity > (1 << 30)) {
    set_error(ctx, "batch_create: %lld work items is too many", static_cast<long long>(b->item_capacity));
    delete b;
    return GDOCK_ERR_INVALID_ARG;
  }
  // Every launch plans its own work items from this list (pose_plan_kernel).
  for (int32_t i = 0; i < n; ++i)
    for (int32_t p = b->pose_off[i]; p < b->pose_off[i + 1]; ++p) b->pose_list.push_back(make_int2(i, p));

  // Device memory.
  std::vector<DevComplex> table(std::max(n, 1));
  for (int32_t i = 0; i < n; ++i) {
    table[i] = b->cx[i]->dev;
    table[i].lanes_k = b->lanes_k[i];
    table[i].min_parts = min_batch_parts(table[i]);
  }
  const size_t np = std::max<int64_t>(total, 1);
  const size_t ni = std::max<int64_t>(b->item_capacity, 1);
  Layout Ly;
  const size_t o_tab = Ly.add<DevComplex>(table.size()), o_list = Ly.add<int2>(np), o_n = Ly.add<int32_t>(4),
               o_poses = Ly.add<double>(np * kPoseDoubles), o_out = Ly.add<double>(np * kOutDoubles),
               o_ns = Ly.add<int32_t>(np), o_redo = Ly.add<int32_t>(np), o_items = Ly.add<int2>(ni),
               o_order = Ly.add<uint8_t>(order_bytes(np));
  std::vector<size_t> o_split(n, 0);
  const size_t split_begin = Ly.size;
  for (int32_t i = 0; i < n; ++i) {
    const size_t bytes = split_state_bytes(b->cx[i]->dev, b->parts[i], poses_per_complex[i]);
    if (bytes) o_split[i] = Ly.add<uint8_t>(bytes);
  }
  const size_t split_end = Ly.size;
  e = hipMalloc(&b->mem, Ly.size);
  if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&b->h_poses), np * kPoseDoubles * sizeof(double));
  if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&b->h_out), np * kOutDoubles * sizeof(double));
  if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&b->h_nsat), np * sizeof(int32_t));
  if (e != hipSuccess) {
    gdock_batch_destroy(ctx, b);
    return hip_fail(ctx, e, "batch_create: allocation");
  }
  char* m = static_cast<char*>(b->mem);
  for (int32_t i = 0; i < n; ++i)
    bind_split_state(&table[i], b->parts[i], poses_per_complex[i], b->pose_off[i],
                     b->parts[i] > 1 ? m + o_split[i] : nullptr);
  // n_items (written by each launch's plan), counter, exit count, pose count
  int32_t head[4] = {0, 0, 0, static_cast<int32_t>(b->pose_list.size())};
  std::vector<uint8_t> blob(o_poses, 0);
  std::memcpy(blob.data() + o_tab, table.data(), table.size() * sizeof(DevComplex));
  if (!b->pose_list.empty())
    std::memcpy(blob.data() + o_list, b->pose_list.data(), b->pose_list.size() * sizeof(int2));
  std::memcpy(blob.data() + o_n, head, sizeof head);
  e = hipMemcpyAsync(m, blob.data(), blob.size(), hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipMemsetAsync(m + o_ns, 0, (o_redo - o_ns) + np * sizeof(int32_t), ctx->stream);
  if (e == hipSuccess) e = hipMemsetAsync(m + o_order, 0, order_bytes(np), ctx->stream);
  if (e == hipSuccess && split_end > split_begin)  // split state starts cleared; the kernel clears it after use
    e = hipMemsetAsync(m + split_begin, 0, split_end - split_begin, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    gdock_batch_destroy(ctx, b);
    return hip_fail(ctx, e, "batch_create: upload");
  }
  BatchLaunch& L = b->L;
  L.table = reinterpret_cast<const DevComplex*>(m + o_tab);
  L.pose_list = reinterpret_cast<const int2*>(m + o_list);
  L.n_pose_list = reinterpret_cast<const int32_t*>(m + o_n) + 3;
  L.items = reinterpret_cast<int2*>(m + o_items);
  L.n_items = reinterpret_cast<int32_t*>(m + o_n);
  L.counter = reinterpret_cast<int32_t*>(m + o_n) + 1;
  L.exit_count = reinterpret_cast<int32_t*>(m + o_n) + 2;
  L.poses = reinterpret_cast<const double*>(m + o_poses);
  L.out = reinterpret_cast<double*>(m + o_out);
  L.nsat = reinterpret_cast<int32_t*>(m + o_ns);
  L.redo = reinterpret_cast<int32_t*>(m + o_redo);
  bind_order(&L, m + o_order, np);
  L.lds = batch_lds_bytes(max_rec);
  L.capacity = static_cast<int>(np);
  L.item_capacity = static_cast<int>(ni);
  L.grid = std::min(batch_grid(L.lds, ctx->device), L.item_capacity);
  *out = b;
  return GDOCK_OK;
}

void gdock_batch_destroy(gdock_ctx* ctx, gdock_batch* b) {
  if (!b) return;
  if (ctx) { (void)hipSetDevice(ctx->device); (void)hipStreamSynchronize(ctx->stream); }
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->mem) (void)hipFree(b->mem);
  if (b->h_poses) (void)hipHostFree(b->h_poses);
  if (b->h_out) (void)hipHostFree(b->h_out);
  if (b->h_nsat) (void)hipHostFree(b->h_nsat);
  delete b;
}

int32_t gdock_batch_n_poses(const gdock_batch* b) { return b ? b->n_poses : -1; }

gdock_status gdock_batch_plan(const gdock_batch* b, int32_t* lanes_k, int32_t* parts, int32_t* grid,
                              int32_t* lds_bytes) {
  if (!b) return GDOCK_ERR_INVALID_ARG;
  if (lanes_k) std::copy(b->lanes_k.begin(), b->lanes_k.end(), lanes_k);
  if (parts) std::copy(b->parts.begin(), b->parts.end(), parts);
  if (grid) *grid = b->L.grid;
  if (lds_bytes) *lds_bytes = static_cast<int32_t>(b->L.lds);
  return GDOCK_OK;
}

gdock_status gdock_batch_last_items(const gdock_batch* b, int32_t* n_items) {
  if (!b || !n_items) return GDOCK_ERR_INVALID_ARG;
  *n_items = 0;
  if (b->pose_list.empty()) return GDOCK_OK;
  hipError_t e = hipSetDevice(b->ctx->device);
  if (e == hipSuccess) e = hipStreamSynchronize(b->ctx->stream);
  if (e == hipSuccess) e = hipMemcpy(n_items, b->L.n_items, sizeof(int32_t), hipMemcpyDeviceToHost);
  return e == hipSuccess ? GDOCK_OK : hip_fail(b->ctx, e, "batch_last_items");
}

gdock_status gdock_batch_device_buffers(gdock_batch* b, double** d_poses, double** d_out, int32_t** d_nsat) {
  if (!b) return GDOCK_ERR_INVALID_ARG;
  if (d_poses) *d_poses = const_cast<double*>(b->L.poses);
  if (d_out) *d_out = b->L.out;
  if (d_nsat) *d_nsat = b->L.nsat;
  return GDOCK_OK;
}

// Captures (once per set of weights) and replays the generation graph.
gdock_status gdock_batch_launch(gdock_ctx* ctx, gdock_batch* b, const gdock_weights* w, void* stream) {
  if (!ctx || !b || !w) {
    if (ctx) set_error(ctx, "batch_launch: invalid arguments");
    return GDOCK_ERR_INVALID_ARG;
  }
  if (b->pose_list.empty()) return GDOCK_OK;
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
  const Weights ww{w->vdw, w->elec, w->desolv, w->air};
  hipError_t e = hipSetDevice(ctx->device);
  if (e != hipSuccess) return hip_fail(ctx, e, "batch_launch: hipSetDevice");
  if (!b->have_graph || std::memcmp(&ww, &b->gw, sizeof ww) != 0) {
    if (b->exec) { (void)hipGraphExecDestroy(b->exec); b->exec = nullptr; }
    if (b->graph) { (void)hipGraphDestroy(b->graph); b->graph = nullptr; }
    b->have_graph = false;
    hipStream_t cap = nullptr;
    e = hipStreamCreateWithFlags(&cap, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
      const hipError_t el = launch_batch(b->L, ww, cap);
      e = hipStreamEndCapture(cap, &b->graph);
      if (e == hipSuccess) e = el;
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
    if (cap) (void)hipStreamDestroy(cap);
    if (e != hipSuccess) return hip_fail(ctx, e, "batch_launch: graph capture");
    b->gw = ww;
    b->have_graph = true;
  }
  e = hipGraphLaunch(b->exec, s);
  if (e != hipSuccess) return hip_fail(ctx, e, "batch_launch: hipGraphLaunch");
  return GDOCK_OK;
}

gdock_status gdock_batch_score(gdock_ctx* ctx, gdock_batch* b, const double* genes, const gdock_weights* w,
                               double* out, int32_t* n_satisfied) {
  if (!ctx || !b || !w || (b->n_poses > 0 && (!genes || !out))) {
    if (ctx) set_error(ctx, "batch_score: invalid arguments");
    return GDOCK_ERR_INVALID_ARG;
  }
  const int32_t n = b->n_poses;
  if (n == 0) return GDOCK_OK;
  for (int64_t i = 0; i < static_cast<int64_t>(n) * 6; ++i)
    if (!std::isfinite(genes[i])) {
      set_error(ctx, "batch_score: gene %lld is not finite", static_cast<long long>(i));
      return GDOCK_ERR_INVALID_ARG;
    }
  gdock_prepare_poses(genes, n, b->h_poses);
  hipError_t e = hipSetDevice(ctx->device);
  if (e == hipSuccess)
    e = hipMemcpyAsync(const_cast<double*>(b->L.poses), b->h_poses, static_cast<size_t>(n) * kPoseDoubles * 8,
                       hipMemcpyHostToDevice, ctx->stream);
  if (e != hipSuccess) return hip_fail(ctx, e, "batch_score: upload");
  gdock_status st = gdock_batch_launch(ctx, b, w, nullptr);
  if (st != GDOCK_OK) return st;
  e = hipMemcpyAsync(b->h_out, b->L.out, static_cast<size_t>(n) * kOutDoubles * 8, hipMemcpyDeviceToHost,
                     ctx->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(b->h_nsat, b->L.nsat, static_cast<size_t>(n) * 4, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return hip_fail(ctx, e, "batch_score: download");
  std::memcpy(out, b->h_out, static_cast<size_t>(n) * kOutDoubles * 8);
  if (n_satisfied) std::memcpy(n_satisfied, b->h_nsat, static_cast<size_t>(n) * 4);
  return GDOCK_OK;
}

}  // extern "C"
