// DockQ evaluation of poses against a reference ligand (evaluator.rs) on the
// device: one workgroup per pose.
//
// Host side (once per evaluator): the interface residues (receptor and
// reference atoms within 5 A, evaluator.rs:170-187), the receptor and
// reference interface atoms (structure.rs:319-327 filter), and the native
// contacts (evaluator.rs:147-168) as a bitset over (receptor key, ligand key),
// a key being (chain, residue, atom name) as in evaluator.rs:20-27.
// Device side (per pose): the ligand moved by the pose in FP64 (the
// reference's operation order), L-RMSD and interface RMSD over the zipped atom
// lists (sums of distances, not squared distances: evaluator.rs:88-132), fnat
// from the docked interface contacts decided in FP64 (evaluator.rs:134-145),
// and DockQ (evaluator.rs:68-87).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <set>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/gdock_hip.h"
#include "common.h"
#include "internal.h"

namespace gdock {

struct DevEvaluator {
  const double* rint_xyz;    // receptor interface atoms, FP64 x, y, z
  const int32_t* rint_key;   // receptor key of each
  const int32_t* lint_atom;  // model ligand interface atoms (device ligand index), input order
  const int32_t* lint_key;   // ligand key of each (-1: no native contact can match)
  const int32_t* rmsd_atom;  // model ligand atom i (device index), i < n_rmsd
  const double* rmsd_ref;    // reference atom i, FP64 x, y, z
  const int32_t* irmsd_atom; // model ligand interface atom k (device index), k < n_irmsd
  const double* irmsd_ref;   // reference interface atom k, FP64 x, y, z
  const uint32_t* native;    // bit rkey * n_lkeys + lkey
  int32_t n_rint, n_lint, n_rmsd, n_irmsd, n_lkeys;
  int32_t n_ref, n_ref_interface, n_native;  // denominators (evaluator.rs:91,113,139-144)
};

namespace {

struct PoseMat {
  double ca, sa, cb, sb, cg, sg, m01, m02, m11, m12, dx, dy, dz;
};

__device__ __forceinline__ PoseMat pose_mat(const double* __restrict__ p) {
  PoseMat c;
  c.dx = p[3]; c.dy = p[4]; c.dz = p[5];
  c.ca = p[6]; c.sa = p[7]; c.cb = p[8]; c.sb = p[9]; c.cg = p[10]; c.sg = p[11];
  c.m01 = p[12]; c.m02 = p[13]; c.m11 = p[14]; c.m12 = p[15];
  return c;
}

// structure.rs:80-131 (rotate about the COM, then displace); -ffp-contract=off
__device__ __forceinline__ void moved(const DevComplex& cx, const PoseMat& pc, int l, double& X, double& Y, double& Z) {
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

__device__ __forceinline__ double dist_to(double X, double Y, double Z, const double* __restrict__ r) {
  const double dx = X - r[0], dy = Y - r[1], dz = Z - r[2];  // structure.rs:178-183
  return sqrt(dx * dx + dy * dy + dz * dz);
}

constexpr int kEvalThreads = 256;

template <typename T>
__device__ T block_sum(T v, T* s_red) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) s_red[wave] = v;
  __syncthreads();
  T t = 0;
  for (int w = 0; w < kEvalThreads / 64; ++w) t += s_red[w];
  return t;
}

}  // namespace

__global__ __launch_bounds__(kEvalThreads) void dockq_kernel(const DevComplex cx, const DevEvaluator ev,
                                                             const double* __restrict__ poses,
                                                             double* __restrict__ metrics) {
  __shared__ double s_red[kEvalThreads / 64];
  __shared__ long long s_cnt[kEvalThreads / 64];
  const int pose = blockIdx.x, tid = threadIdx.x;
  const PoseMat pc = pose_mat(poses + static_cast<size_t>(kPoseDoubles) * pose);
  double sr = 0.0, si = 0.0;
  for (int i = tid; i < ev.n_rmsd; i += kEvalThreads) {
    double X, Y, Z;
    moved(cx, pc, ev.rmsd_atom[i], X, Y, Z);
    sr += dist_to(X, Y, Z, ev.rmsd_ref + 3 * i);
  }
  for (int k = tid; k < ev.n_irmsd; k += kEvalThreads) {
    double X, Y, Z;
    moved(cx, pc, ev.irmsd_atom[k], X, Y, Z);
    si += dist_to(X, Y, Z, ev.irmsd_ref + 3 * k);
  }
  // docked interface contacts that are native (FP64 distance <= 5 A)
  long long common = 0;
  for (int b = tid; b < ev.n_lint; b += kEvalThreads) {
    const int lk = ev.lint_key[b];
    if (lk < 0) continue;
    double X, Y, Z;
    moved(cx, pc, ev.lint_atom[b], X, Y, Z);
    for (int a = 0; a < ev.n_rint; ++a) {
      const double* r = ev.rint_xyz + 3 * a;
      const double ex = r[0] - X, ey = r[1] - Y, ez = r[2] - Z;
      if (ex * ex + ey * ey + ez * ez <= kD2ContactIncl) {
        const uint32_t bit = static_cast<uint32_t>(ev.rint_key[a]) * static_cast<uint32_t>(ev.n_lkeys) +
                             static_cast<uint32_t>(lk);
        common += (ev.native[bit >> 5] >> (bit & 31)) & 1u;
      }
    }
  }
  sr = block_sum(sr, s_red);
  si = block_sum(si, s_red);
  common = block_sum(common, s_cnt);
  if (tid == 0) {
    const double rmsd = ev.n_ref > 0 ? sqrt(sr / static_cast<double>(ev.n_ref)) : 0.0;
    const double irmsd = ev.n_ref_interface > 0 ? sqrt(si / static_cast<double>(ev.n_ref_interface)) : 0.0;
    const double fnat = ev.n_native > 0 ? static_cast<double>(common) / static_cast<double>(ev.n_native) : 0.0;
    const double qi = irmsd / 1.5, ql = rmsd / 8.5;
    const double dockq = (fnat + 1.0 / (1.0 + qi * qi) + 1.0 / (1.0 + ql * ql)) / 3.0;
    double* o = metrics + 4 * static_cast<size_t>(pose);
    o[0] = rmsd; o[1] = irmsd; o[2] = fnat; o[3] = dockq;
  }
}

}  // namespace gdock

using namespace gdock;

struct gdock_evaluator {
  gdock_complex* cx = nullptr;
  DevEvaluator dev{};
  void* mem = nullptr;
  int32_t n_irec_res = 0, n_iref_res = 0;
};

namespace {

double host_distance(const gdock_atom& a, const gdock_atom& b) {  // structure.rs:312-317
  const double dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
  return std::sqrt(dx * dx + dy * dy + dz * dz);
}

using AtomKey = std::tuple<char, int16_t, std::string>;
AtomKey key_of(const gdock_atom& a) { return AtomKey{a.chainid, a.resseq, std::string(field(a.name, sizeof a.name))}; }

}  // namespace

extern "C" {

gdock_status gdock_evaluator_create(gdock_ctx* ctx, gdock_complex* cx, const gdock_atom* reference, int32_t n_ref,
                                    gdock_evaluator** out) {
  if (!ctx || !cx || !out || n_ref < 0 || (n_ref > 0 && !reference)) {
    if (ctx) set_error(ctx, "evaluator_create: invalid arguments");
    return GDOCK_ERR_INVALID_ARG;
  }
  *out = nullptr;
  const std::vector<gdock_atom>& rec = cx->rec;
  const std::vector<gdock_atom>& lig = cx->lig;  // the model topology (positioned ligand)
  // evaluator.rs:170-187 interface residues
  std::set<int16_t> irec, iref;
  for (const gdock_atom& a : rec)
    for (int32_t j = 0; j < n_ref; ++j)
      if (host_distance(a, reference[j]) <= 5.0) { irec.insert(a.resseq); iref.insert(reference[j].resseq); }
  std::vector<int32_t> rint, refint, lint;  // filtered atom lists, input order
  for (int32_t i = 0; i < static_cast<int32_t>(rec.size()); ++i) if (irec.count(rec[i].resseq)) rint.push_back(i);
  for (int32_t j = 0; j < n_ref; ++j) if (iref.count(reference[j].resseq)) refint.push_back(j);
  for (int32_t i = 0; i < static_cast<int32_t>(lig.size()); ++i) if (iref.count(lig[i].resseq)) lint.push_back(i);
  // keys and native contacts (evaluator.rs:147-168 over the interfaces)
  std::map<AtomKey, int32_t> rkeys, lkeys;
  std::vector<int32_t> rint_key(rint.size());
  for (size_t a = 0; a < rint.size(); ++a) rint_key[a] = rkeys.emplace(key_of(rec[rint[a]]), static_cast<int32_t>(rkeys.size())).first->second;
  std::vector<int32_t> refint_key(refint.size());
  for (size_t b = 0; b < refint.size(); ++b)
    refint_key[b] = lkeys.emplace(key_of(reference[refint[b]]), static_cast<int32_t>(lkeys.size())).first->second;
  const int64_t nbits = static_cast<int64_t>(std::max<size_t>(rkeys.size(), 1)) * std::max<size_t>(lkeys.size(), 1);
  if (nbits > (int64_t{1} << 31)) {
    set_error(ctx, "evaluator_create: interface key table too large");
    return GDOCK_ERR_INVALID_ARG;
  }
  std::vector<uint32_t> native(static_cast<size_t>((nbits + 31) / 32), 0u);
  int32_t n_native = 0;
  const int32_t n_lkeys = static_cast<int32_t>(std::max<size_t>(lkeys.size(), 1));
  for (size_t a = 0; a < rint.size(); ++a)
    for (size_t b = 0; b < refint.size(); ++b)
      if (host_distance(rec[rint[a]], reference[refint[b]]) <= 5.0) {
        ++n_native;  // every pair is a Vec entry, duplicates included
        const int64_t bit = static_cast<int64_t>(rint_key[a]) * n_lkeys + refint_key[b];
        native[bit >> 5] |= 1u << (bit & 31);
      }
  std::vector<int32_t> lint_key(lint.size()), lint_atom(lint.size());
  for (size_t b = 0; b < lint.size(); ++b) {
    const auto it = lkeys.find(key_of(lig[lint[b]]));
    lint_key[b] = it == lkeys.end() ? -1 : it->second;
    lint_atom[b] = cx->lig_slot[lint[b]];
  }
  std::vector<double> rint_xyz;
  for (int32_t i : rint) rint_xyz.insert(rint_xyz.end(), {rec[i].x, rec[i].y, rec[i].z});
  // zipped lists (evaluator.rs:95-101,119-125)
  const int32_t n_rmsd = std::min<int32_t>(static_cast<int32_t>(lig.size()), n_ref);
  std::vector<int32_t> rmsd_atom(n_rmsd);
  std::vector<double> rmsd_ref;
  for (int32_t i = 0; i < n_rmsd; ++i) {
    rmsd_atom[i] = cx->lig_slot[i];
    rmsd_ref.insert(rmsd_ref.end(), {reference[i].x, reference[i].y, reference[i].z});
  }
  const int32_t n_irmsd = static_cast<int32_t>(std::min(lint.size(), refint.size()));
  std::vector<int32_t> irmsd_atom(n_irmsd);
  std::vector<double> irmsd_ref;
  for (int32_t k = 0; k < n_irmsd; ++k) {
    irmsd_atom[k] = cx->lig_slot[lint[k]];
    const gdock_atom& r = reference[refint[k]];
    irmsd_ref.insert(irmsd_ref.end(), {r.x, r.y, r.z});
  }

  Layout L;
  const size_t o1 = L.add<double>(rint_xyz.size()), o2 = L.add<int32_t>(rint_key.size()),
               o3 = L.add<int32_t>(lint_atom.size()), o4 = L.add<int32_t>(lint_key.size()),
               o5 = L.add<int32_t>(rmsd_atom.size()), o6 = L.add<double>(rmsd_ref.size()),
               o7 = L.add<int32_t>(irmsd_atom.size()), o8 = L.add<double>(irmsd_ref.size()),
               o9 = L.add<uint32_t>(native.size());
  std::vector<uint8_t> blob(std::max<size_t>(L.size, 256), 0);
  auto put = [&](size_t off, const void* src, size_t bytes) { if (bytes) std::memcpy(blob.data() + off, src, bytes); };
  put(o1, rint_xyz.data(), rint_xyz.size() * 8); put(o2, rint_key.data(), rint_key.size() * 4);
  put(o3, lint_atom.data(), lint_atom.size() * 4); put(o4, lint_key.data(), lint_key.size() * 4);
  put(o5, rmsd_atom.data(), rmsd_atom.size() * 4); put(o6, rmsd_ref.data(), rmsd_ref.size() * 8);
  put(o7, irmsd_atom.data(), irmsd_atom.size() * 4); put(o8, irmsd_ref.data(), irmsd_ref.size() * 8);
  put(o9, native.data(), native.size() * 4);
  auto* ev = new gdock_evaluator();
  ev->cx = cx;
  hipError_t e = hipSetDevice(ctx->device);
  if (e == hipSuccess) e = hipMalloc(&ev->mem, blob.size());
  if (e == hipSuccess) e = hipMemcpy(ev->mem, blob.data(), blob.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    if (ev->mem) (void)hipFree(ev->mem);
    delete ev;
    return hip_fail(ctx, e, "evaluator_create: upload");
  }
  char* m = static_cast<char*>(ev->mem);
  DevEvaluator& d = ev->dev;
  d.rint_xyz = reinterpret_cast<const double*>(m + o1);
  d.rint_key = reinterpret_cast<const int32_t*>(m + o2);
  d.lint_atom = reinterpret_cast<const int32_t*>(m + o3);
  d.lint_key = reinterpret_cast<const int32_t*>(m + o4);
  d.rmsd_atom = reinterpret_cast<const int32_t*>(m + o5);
  d.rmsd_ref = reinterpret_cast<const double*>(m + o6);
  d.irmsd_atom = reinterpret_cast<const int32_t*>(m + o7);
  d.irmsd_ref = reinterpret_cast<const double*>(m + o8);
  d.native = reinterpret_cast<const uint32_t*>(m + o9);
  d.n_rint = static_cast<int32_t>(rint.size());
  d.n_lint = static_cast<int32_t>(lint.size());
  d.n_rmsd = n_rmsd;
  d.n_irmsd = n_irmsd;
  d.n_lkeys = n_lkeys;
  d.n_ref = n_ref;
  d.n_ref_interface = static_cast<int32_t>(refint.size());
  d.n_native = n_native;
  ev->n_irec_res = static_cast<int32_t>(irec.size());
  ev->n_iref_res = static_cast<int32_t>(iref.size());
  *out = ev;
  return GDOCK_OK;
}

void gdock_evaluator_destroy(gdock_ctx* ctx, gdock_evaluator* ev) {
  if (!ev) return;
  if (ctx) { (void)hipSetDevice(ctx->device); (void)hipStreamSynchronize(ctx->stream); }
  if (ev->mem) (void)hipFree(ev->mem);
  delete ev;
}

gdock_status gdock_evaluator_info(const gdock_evaluator* ev, int32_t* n_receptor_interface_res,
                                  int32_t* n_reference_interface_res, int32_t* n_native_contacts) {
  if (!ev) return GDOCK_ERR_INVALID_ARG;
  if (n_receptor_interface_res) *n_receptor_interface_res = ev->n_irec_res;
  if (n_reference_interface_res) *n_reference_interface_res = ev->n_iref_res;
  if (n_native_contacts) *n_native_contacts = ev->dev.n_native;
  return GDOCK_OK;
}

gdock_status gdock_calc_metrics(gdock_ctx* ctx, gdock_evaluator* ev, const double* genes, int32_t n, double* metrics) {
  if (!ctx || !ev || n < 0 || (n > 0 && (!genes || !metrics))) {
    if (ctx) set_error(ctx, "calc_metrics: invalid arguments");
    return GDOCK_ERR_INVALID_ARG;
  }
  if (n == 0) return GDOCK_OK;
  std::vector<double> poses(static_cast<size_t>(n) * kPoseDoubles);
  if (gdock_prepare_poses(genes, n, poses.data()) != GDOCK_OK) {
    set_error(ctx, "calc_metrics: non-finite gene");
    return GDOCK_ERR_INVALID_ARG;
  }
  hipError_t e = hipSetDevice(ctx->device);
  Layout L;
  const size_t o_p = L.add<double>(poses.size()), o_m = L.add<double>(static_cast<size_t>(n) * 4);
  if (e == hipSuccess && !ctx->staging.reserve(L.size, &e)) return hip_fail(ctx, e, "calc_metrics: scratch");
  char* m = static_cast<char*>(ctx->staging.ptr);
  if (e == hipSuccess)
    e = hipMemcpyAsync(m + o_p, poses.data(), poses.size() * 8, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(dockq_kernel, dim3(n), dim3(kEvalThreads), 0, ctx->stream, ev->cx->dev, ev->dev,
                       reinterpret_cast<const double*>(m + o_p), reinterpret_cast<double*>(m + o_m));
    e = hipGetLastError();
  }
  if (e == hipSuccess)
    e = hipMemcpyAsync(metrics, m + o_m, static_cast<size_t>(n) * 4 * 8, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  return e == hipSuccess ? GDOCK_OK : hip_fail(ctx, e, "calc_metrics");
}

}  // extern "C"
