// C ABI of the gfx950 scoring / FCC library: contexts, complexes (device
// resident receptor + positioned ligand), population scoring entry points.
#include <hip/hip_runtime.h>

#include <pthread.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <condition_variable>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <string>
#include <string_view>
#include <thread>
#include <unordered_map>
#include <utility>
#include <vector>

#include "../../include/gdock_hip.h"
#include "common.h"
#include "internal.h"

using namespace gdock;

namespace gdock {

void set_error(gdock_ctx* ctx, const char* fmt, ...) {
  if (!ctx) return;
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  ctx->err = buf;
}

gdock_status hip_fail(gdock_ctx* ctx, hipError_t e, const char* what) {
  set_error(ctx, "%s: %s", what, hipGetErrorString(e));
  return e == hipErrorOutOfMemory ? GDOCK_ERR_OOM : GDOCK_ERR_DEVICE;
}

bool DeviceArena::reserve(size_t bytes, hipError_t* err) {
  if (bytes <= cap) return true;
  if (ptr) (void)hipFree(ptr);
  ptr = nullptr;
  cap = 0;
  const hipError_t e = hipMalloc(&ptr, bytes);
  if (e != hipSuccess) { *err = e; ptr = nullptr; return false; }
  cap = bytes;
  return true;
}

void DeviceArena::release() {
  if (ptr) (void)hipFree(ptr);
  ptr = nullptr;
  cap = 0;
}

// Trimmed string of a fixed-size record field.
std::string_view field(const char* f, size_t n) {
  std::string_view s(f, strnlen(f, n));
  while (!s.empty() && s.front() == ' ') s.remove_prefix(1);
  while (!s.empty() && s.back() == ' ') s.remove_suffix(1);
  return s;
}

// fitness.rs:129-138 atomic solvation parameter by element.
double asp_of(const gdock_atom& a) {
  const std::string_view e = field(a.element, sizeof a.element);
  if (e == "C" || e == "S") return 0.012;
  if (e == "N" || e == "O") return -0.160;
  return 0.0;
}

// structure.rs:63-69: sequential sums divided by the count.
void com_of(const gdock_atom* a, int32_t n, double* c) {
  double sx = 0.0, sy = 0.0, sz = 0.0;
  for (int32_t i = 0; i < n; ++i) sx += a[i].x;
  for (int32_t i = 0; i < n; ++i) sy += a[i].y;
  for (int32_t i = 0; i < n; ++i) sz += a[i].z;
  const double dn = static_cast<double>(n);
  c[0] = sx / dn; c[1] = sy / dn; c[2] = sz / dn;
}

bool finite_atoms(const gdock_atom* a, int32_t n) {
  for (int32_t i = 0; i < n; ++i)
    if (!std::isfinite(a[i].x) || !std::isfinite(a[i].y) || !std::isfinite(a[i].z) ||
        !std::isfinite(a[i].charge) || !std::isfinite(a[i].epsilon) || !std::isfinite(a[i].rmin2) ||
        !std::isfinite(a[i].vdw_radius))
      return false;
  return true;
}

// Spatial order by recursive bisection: idx[lo, hi) is split across its widest
// extent at a multiple of `block`, so every aligned run of `block` entries of
// the result is a compact cluster of atoms.  Deterministic (stable sorts).
void bisect_order(const gdock_atom* atoms, std::vector<int32_t>& idx, size_t lo, size_t hi, size_t block) {
  if (hi - lo <= block) return;
  double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
  for (size_t i = lo; i < hi; ++i) {
    const gdock_atom& a = atoms[idx[i]];
    const double p[3] = {a.x, a.y, a.z};
    for (int d = 0; d < 3; ++d) { mn[d] = std::min(mn[d], p[d]); mx[d] = std::max(mx[d], p[d]); }
  }
  int ax = 0;
  for (int d = 1; d < 3; ++d)
    if (mx[d] - mn[d] > mx[ax] - mn[ax]) ax = d;
  auto coord = [&](int32_t i) { return ax == 0 ? atoms[i].x : (ax == 1 ? atoms[i].y : atoms[i].z); };
  std::stable_sort(idx.begin() + lo, idx.begin() + hi, [&](int32_t a, int32_t b) { return coord(a) < coord(b); });
  const size_t n = hi - lo;
  const size_t k = std::max(block, (n / block / 2) * block);
  bisect_order(atoms, idx, lo, lo + k, block);
  bisect_order(atoms, idx, lo + k, hi, block);
}


// Reorders the full blocks of `block` entries of idx (the partial block, if
// any, stays last) by bit-reversed block index, so that any run of
// consecutive blocks samples the whole order evenly.  Applied after
// bisect_order: each block stays spatially compact, while consecutive blocks
// come from all over the molecule.
void spread_blocks(std::vector<int32_t>& idx, size_t block) {
  const size_t nb = idx.size() / block;
  if (nb < 3) return;
  int bits = 0;
  while ((size_t{1} << bits) < nb) ++bits;
  auto rev = [bits](size_t b) {
    size_t r = 0;
    for (int i = 0; i < bits; ++i) r |= ((b >> i) & 1u) << (bits - 1 - i);
    return r;
  };
  std::vector<size_t> order(nb);
  for (size_t b = 0; b < nb; ++b) order[b] = b;
  std::sort(order.begin(), order.end(), [&](size_t a, size_t b) { return rev(a) < rev(b); });
  std::vector<int32_t> out(idx.size());
  for (size_t p = 0; p < nb; ++p)
    std::copy(idx.begin() + order[p] * block, idx.begin() + (order[p] + 1) * block, out.begin() + p * block);
  std::copy(idx.begin() + nb * block, idx.end(), out.begin() + nb * block);
  idx.swap(out);
}

// Fixed-point value of an FP64 coordinate: (x - o) in units of 2^-qexp A,
// clamped to +-kFixedClamp, rounded to nearest even.  The scoring kernel
// applies the same operations to the posed ligand atoms.
int32_t fixed_of(double x, double o, double scale) {
  double v = (x - o) * scale;
  v = std::min(std::max(v, -static_cast<double>(kFixedClamp)), static_cast<double>(kFixedClamp));
  return static_cast<int32_t>(std::nearbyint(v));
}

// Finest fixed-point unit that holds the receptor within kFixedClamp units
// minus 16 A (so a clamped ligand atom is beyond every cutoff); -1 if none.
int fixed_exponent(double rec_extent) {
  for (int e = kFixedMaxExp; e >= kFixedMinExp; --e)
    if (rec_extent + 16.0 <= static_cast<double>(kFixedClamp) * std::ldexp(1.0, -e)) return e;
  return -1;
}

// Guard bands of the FP32 cutoff decisions on s, the FP32 squared distance in
// units^2 (u = 2^-qexp A).  For a pair with reference FP64 coordinates r, p:
//   n_i = round(r_i 2^e) - round(p_i 2^e) (after subtracting the origin in
//   FP64) is within a = 1 + 1e-6 units of the exact difference D_i;
//   f_i = fl32(n_i) is within u32 |n_i| of n_i (u32 = 2^-24), so
//   |f_i - D_i| <= eps = a + u32 (|D| + a);
//   |sum f_i^2 - |D|^2| <= 2 sqrt(3) eps |D| + 3 eps^2;
//   the three FP32 roundings of s = fma(z, z, fma(y, y, x x)) add at most
//   3.0001 u32 s, and the reference's own FP64 rounding of d^2 6e-16 d^2.
// Near a cutoff T (|D|^2 ~ T 2^2e) the band is twice the sum; a pair whose s
// lies outside it gets the reference's decision from the FP32 comparison, a
// pair inside it is decided again in FP64.  The four cutoffs share the
// widest band width (in ulps of s).
void fixed_bands(int qexp, uint32_t* c, uint32_t* w) {
  const double T[4] = {1.0, 64.0, 144.0, 225.0};
  const double u32 = 5.9604644775390625e-08, a = 1.000001;
  uint32_t wmax = 0;
  for (int i = 0; i < 4; ++i) {
    const double Tu = std::ldexp(T[i], 2 * qexp), Du = std::sqrt(Tu) * 1.001;
    const double eps = a + u32 * (Du + a);
    const double b = 2.0 * (2.0 * 1.7320508075688772 * eps * Du + 3.0 * eps * eps + 3.0001 * u32 * Tu * 1.001 +
                            6e-16 * Tu) + 1.0;
    float lo = static_cast<float>(Tu - b), hi = static_cast<float>(Tu + b);
    if (static_cast<double>(lo) > Tu - b) lo = std::nextafter(lo, -INFINITY);
    if (static_cast<double>(hi) < Tu + b) hi = std::nextafter(hi, INFINITY);
    uint32_t lb, hb;
    std::memcpy(&lb, &lo, 4);
    std::memcpy(&hb, &hi, 4);
    c[i] = lb + 1u;
    wmax = std::max(wmax, hb - c[i]);
  }
  *w = wmax;
}

// Charge unit of the receptor's FP32 charges: 1e4 when every charge is an
// exact multiple of 1e-4 e (as every OPLS charge of the toppar table is), so
// the FP32 values q * 1e4 are exact integers; otherwise 1 (charges rounded
// to FP32).
double charge_scale(const gdock_atom* a, int32_t n) {
  constexpr double kScale = 1e4;
  for (int32_t i = 0; i < n; ++i) {
    const double m = std::nearbyint(a[i].charge * kScale);
    if (std::fabs(m) >= 16777216.0 || m / kScale != a[i].charge) return 1.0;
  }
  return kScale;
}

namespace {

// Pose records [lo, hi) from genes (genes already checked finite).
void pose_records(const double* genes, double* poses, int64_t lo, int64_t hi) {
  for (int64_t i = lo; i < hi; ++i) {
    const double* g = genes + 6 * i;
    double* p = poses + kPoseDoubles * i;
    for (int k = 0; k < 6; ++k) p[k] = g[k];
    // structure.rs:96-103 takes f64::cos and f64::sin of each angle; a sin and
    // a cos of one operand compile to one glibc sincos call (LLVM on GNU
    // targets, as g++ does too), which can differ from separate cos and sin
    // calls in the last bit.  So: sincos, explicitly.
    ::sincos(g[0], &p[7], &p[6]);
    ::sincos(g[1], &p[9], &p[8]);
    ::sincos(g[2], &p[11], &p[10]);
    // the bracketed sums of structure.rs:96-101 (no FMA contraction: -ffp-contract=off)
    const double ca = p[6], sa = p[7], sb = p[9], cg = p[10], sg = p[11];
    p[12] = (-sa * cg) + ((ca * sb) * sg);
    p[13] = (sa * sg) + ((ca * sb) * cg);
    p[14] = (ca * cg) + ((sa * sb) * sg);
    p[15] = (-ca * sg) + ((sa * sb) * cg);
  }
}

// Fork-join over [0, n) for large host loops: a few worker threads, created
// on first use and kept for the life of the process, each take one slice and
// the caller takes the first.  One call at a time uses the workers; a call
// that finds them busy runs its loop alone, and so does every call in a
// child process forked after the pool started (the workers are not there).
std::atomic<bool> g_pool_forked{false};

class HostPool {
 public:
  static HostPool& get() {
    static HostPool* pool = new HostPool();  // never destroyed: workers live as long as the process
    return *pool;
  }
  template <class F>
  void run(int64_t n, F&& f) {
    if (g_pool_forked.load(std::memory_order_relaxed)) {
      f(0, n);
      return;
    }
    std::unique_lock<std::mutex> busy(call_mu_, std::try_to_lock);
    const int parts = static_cast<int>(std::min<int64_t>(n, workers_ + 1));
    if (!busy.owns_lock() || parts <= 1) {
      f(0, n);
      return;
    }
    auto slice = [n, parts](int q, int64_t* lo, int64_t* hi) {
      *lo = n * q / parts;
      *hi = n * (q + 1) / parts;
    };
    std::function<void(int)> job = [&](int q) {
      int64_t lo, hi;
      slice(q, &lo, &hi);
      f(lo, hi);
    };
    {
      std::lock_guard<std::mutex> g(mu_);
      job_ = &job;
      parts_ = parts;
      pending_ = parts - 1;
      ++gen_;
    }
    cv_.notify_all();
    job(0);
    std::unique_lock<std::mutex> g(mu_);
    done_cv_.wait(g, [this] { return pending_ == 0; });
    job_ = nullptr;
  }

 private:
  HostPool() {
    pthread_atfork(nullptr, nullptr, [] { g_pool_forked.store(true, std::memory_order_relaxed); });
    const unsigned hw = std::thread::hardware_concurrency();
    workers_ = static_cast<int>(std::min(7u, hw > 1 ? hw - 1 : 0u));
    for (int w = 0; w < workers_; ++w) std::thread([this, w] { loop(w + 1); }).detach();
  }
  void loop(int q) {
    uint64_t seen = 0;
    for (;;) {
      std::unique_lock<std::mutex> g(mu_);
      cv_.wait(g, [&] { return gen_ != seen; });
      seen = gen_;
      if (q >= parts_) continue;  // this call uses fewer slices
      const std::function<void(int)>* job = job_;
      g.unlock();
      (*job)(q);
      g.lock();
      if (--pending_ == 0) done_cv_.notify_one();
    }
  }
  int workers_ = 0;
  std::mutex call_mu_, mu_;
  std::condition_variable cv_, done_cv_;
  const std::function<void(int)>* job_ = nullptr;
  int parts_ = 0, pending_ = 0;
  uint64_t gen_ = 0;
};

template <class F>
void host_pool_for(int64_t n, F&& f) {
  HostPool::get().run(n, std::forward<F>(f));
}

}  // namespace

void host_parallel_for(int64_t n, const std::function<void(int64_t, int64_t)>& f) { host_pool_for(n, f); }

}  // namespace gdock

// Device scratch of gdock_score_poses_device_noalloc: per pose an
// exact-rescore flag, a place in the launch order and a cost bucket (576 KiB
// on every device the library runs on, held until the process ends).  Host code never uses the name `scratch` as an address: its
// device address is resolved once per device (scratch_resolve) and that
// pointer is what kernels receive.
namespace {
constexpr int32_t kNoAllocMaxPoses = 1 << 16;
static __device__ __align__(256) int32_t scratch[2 * kNoAllocMaxPoses + kNoAllocMaxPoses / 4];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};
hipError_t g_scratch_err[kMaxDevices] = {};
bool g_scratch_done[kMaxDevices] = {};

// Resolves the scratch address on `device`, which must be current; once per
// device, at context creation.
void scratch_resolve(int device) {
  if (device < 0 || device >= kMaxDevices) return;
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (g_scratch_done[device]) return;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  g_scratch_ptr[device] = e == hipSuccess ? p : nullptr;
  g_scratch_err[device] = e;
  g_scratch_done[device] = true;
}

int32_t* scratch_of(int device, hipError_t* err) {
  if (device < 0 || device >= kMaxDevices) { *err = hipErrorInvalidDevice; return nullptr; }
  std::lock_guard<std::mutex> g(g_scratch_mu);
  *err = g_scratch_done[device] ? g_scratch_err[device] : hipErrorNotInitialized;
  return static_cast<int32_t*>(g_scratch_ptr[device]);
}
}  // namespace

extern "C" {

const char* gdock_version(void) { return "gdock-mi355x 0.1 (gfx950)"; }

gdock_status gdock_ctx_create(int device, gdock_ctx** out) {
  if (!out) return GDOCK_ERR_INVALID_ARG;
  *out = nullptr;
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || device < 0 || device >= n) return GDOCK_ERR_DEVICE;
  e = hipSetDevice(device);
  if (e != hipSuccess) return GDOCK_ERR_DEVICE;
  auto* ctx = new gdock_ctx();
  ctx->device = device;
  e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
  if (e != hipSuccess) { delete ctx; return GDOCK_ERR_DEVICE; }
  scratch_resolve(device);  // the no-allocation entry point's device scratch
  *out = ctx;
  return GDOCK_OK;
}

void gdock_ctx_destroy(gdock_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  ctx->staging.release();
  ctx->redo.release();
  release_fcc_scratch(ctx);
  if (ctx->pinned) (void)hipHostFree(ctx->pinned);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

const char* gdock_last_error(const gdock_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

gdock_status gdock_read_pdb(const char* text, size_t len, gdock_atom** atoms, int32_t** model_offsets,
                            int32_t* n_models) {
  if (!text || !atoms || !model_offsets || !n_models) return GDOCK_ERR_INVALID_ARG;
  std::vector<gdock_atom> a;
  std::vector<int32_t> off;
  read_pdb_text(std::string_view(text, len), &a, &off);
  *atoms = static_cast<gdock_atom*>(std::malloc(std::max<size_t>(1, a.size()) * sizeof(gdock_atom)));
  *model_offsets = static_cast<int32_t*>(std::malloc(off.size() * sizeof(int32_t)));
  if (!*atoms || !*model_offsets) return GDOCK_ERR_OOM;
  if (!a.empty()) std::memcpy(*atoms, a.data(), a.size() * sizeof(gdock_atom));
  std::memcpy(*model_offsets, off.data(), off.size() * sizeof(int32_t));
  *n_models = static_cast<int32_t>(off.size()) - 1;
  return GDOCK_OK;
}

void gdock_free(void* p) { std::free(p); }

gdock_status gdock_position_ligand(const gdock_atom* rec, int32_t n_rec, gdock_atom* lig, int32_t n_lig) {
  if ((!rec && n_rec) || (!lig && n_lig) || n_rec < 0 || n_lig < 0) return GDOCK_ERR_INVALID_ARG;
  double r[3], l[3];
  com_of(rec, n_rec, r);
  com_of(lig, n_lig, l);
  for (int32_t i = 0; i < n_lig; ++i) { lig[i].x -= l[0]; lig[i].y -= l[1]; lig[i].z -= l[2]; }
  for (int32_t i = 0; i < n_lig; ++i) { lig[i].x += r[0]; lig[i].y += r[1]; lig[i].z += r[2]; }
  return GDOCK_OK;
}

gdock_status gdock_complex_create(gdock_ctx* ctx, const gdock_atom* rec, int32_t n_rec, const gdock_atom* lig,
                                  int32_t n_lig, const int32_t* pairs, int32_t n_pairs, gdock_complex** out) {
  if (!ctx || !out) return GDOCK_ERR_INVALID_ARG;
  *out = nullptr;
  if (n_rec < 0 || n_lig < 0 || n_pairs < 0 || (n_rec && !rec) || (n_lig && !lig) || (n_pairs && !pairs)) {
    set_error(ctx, "complex_create: bad sizes or null arrays");
    return GDOCK_ERR_INVALID_ARG;
  }
  if (n_rec > kMaxReceptorAtoms) {
    set_error(ctx, "complex_create: %d receptor atoms exceeds the supported %d", n_rec, kMaxReceptorAtoms);
    return GDOCK_ERR_INVALID_ARG;
  }
  if (!finite_atoms(rec, n_rec) || !finite_atoms(lig, n_lig)) {
    set_error(ctx, "complex_create: non-finite coordinate or parameter");
    return GDOCK_ERR_INVALID_ARG;
  }
  auto* cx = new gdock_complex();
  cx->rec.assign(rec, rec + n_rec);
  cx->lig.assign(lig, lig + n_lig);

  // Receptor permutation: group atoms by (rmin2, epsilon) class so that the
  // kernel updates its per-lane Lennard-Jones factors once per class.  Each
  // class segment is padded to a multiple of 4 with inert atoms (far away, no
  // charge, no solvation parameter) so the kernel reads receptor atoms 4 at a
  // time (one 64-byte scalar load) and class boundaries fall on those groups.
  std::map<std::pair<double, double>, int> cls_of;
  for (auto& a : cx->rec) cls_of.emplace(std::make_pair(a.rmin2, a.epsilon), 0);
  int nc = 0;
  for (auto& kv : cls_of) kv.second = nc++;
  std::vector<std::vector<int32_t>> members(nc);
  for (int32_t i = 0; i < n_rec; ++i) members[cls_of[{cx->rec[i].rmin2, cx->rec[i].epsilon}]].push_back(i);
  // within a class, groups of 4 consecutive atoms are spatially compact (their
  // bounding spheres let a wave skip groups beyond every cutoff)
  for (auto& mem : members) bisect_order(cx->rec.data(), mem, 0, mem.size(), kRecGroup);
  std::vector<int32_t> class_start(nc + 1, 0), perm;  // perm[j] = receptor atom or -1 (padding)
  std::vector<double> class_rmin(nc), class_eps(nc);
  for (auto& kv : cls_of) { class_rmin[kv.second] = kv.first.first; class_eps[kv.second] = kv.first.second; }
  for (int c = 0; c < nc; ++c) {
    perm.insert(perm.end(), members[c].begin(), members[c].end());
    while (perm.size() % kRecGroup) perm.push_back(-1);
    class_start[c + 1] = static_cast<int32_t>(perm.size());
  }
  const int32_t n_pad = static_cast<int32_t>(perm.size());

  double origin[3] = {0, 0, 0}, lcom[3] = {0, 0, 0};
  if (n_rec) com_of(rec, n_rec, origin);
  if (n_lig) com_of(lig, n_lig, lcom);
  // Ligand device order: blocks of kLigBlock compact atoms (a wave's slots).
  std::vector<int32_t> lig_perm(n_lig);  // device index -> input index
  for (int32_t i = 0; i < n_lig; ++i) lig_perm[i] = i;
  bisect_order(cx->lig.data(), lig_perm, 0, lig_perm.size(), kLigBlock);
  spread_blocks(lig_perm, kLigBlock);
  // within a block, aligned quarters of kLigSub atoms are compact too (the
  // kernel culls by the boxes of both)
  for (size_t b = 0; b < lig_perm.size(); b += kLigBlock)
    bisect_order(cx->lig.data(), lig_perm, b, std::min(lig_perm.size(), b + kLigBlock), kLigSub);
  cx->lig_slot.assign(n_lig, 0);
  for (int32_t q = 0; q < n_lig; ++q) cx->lig_slot[lig_perm[q]] = q;
  std::unordered_map<int32_t, std::vector<int32_t>> lig_by_serial;  // serial -> device indices
  for (int32_t q = 0; q < n_lig; ++q) lig_by_serial[cx->lig[lig_perm[q]].serial].push_back(q);

  // Fixed-point frame: the receptor COM as origin, the finest unit that holds
  // the receptor (with room to clamp far ligand atoms).
  double rec_ext = 0.0;
  for (int32_t i = 0; i < n_rec; ++i)
    rec_ext = std::max({rec_ext, std::fabs(rec[i].x - origin[0]), std::fabs(rec[i].y - origin[1]),
                        std::fabs(rec[i].z - origin[2])});
  const int qexp = fixed_exponent(rec_ext);
  if (qexp < 0) {
    delete cx;
    set_error(ctx, "complex_create: receptor extent %.1f A is too large for the fixed-point frame", rec_ext);
    return GDOCK_ERR_INVALID_ARG;
  }
  const double qscale = std::ldexp(1.0, qexp);
  const double charge_unit = charge_scale(rec, n_rec);
  std::vector<int4> fixed(n_pad);
  std::vector<double> rx(n_pad), ry(n_pad), rz(n_pad), rasp(n_pad, 0.0), rrad(n_pad, 0.0);
  std::vector<int2> meta(n_pad, make_int2(0, 0));
  std::vector<int4> coll;  // fitness.rs:92 pairs: same serial in receptor and ligand
  int cls = 0;
  for (int32_t j = 0; j < n_pad; ++j) {
    while (class_start[cls + 1] <= j) ++cls;
    if (perm[j] < 0) {  // padding: far from every (clamped) ligand atom, never evaluated
      fixed[j] = make_int4(kFixedPad, kFixedPad, kFixedPad, 0);
      rx[j] = ry[j] = rz[j] = -1.0e9;
      continue;
    }
    const gdock_atom& a = cx->rec[perm[j]];
    float q = static_cast<float>(a.charge * charge_unit);  // an exact integer when charge_unit > 1
    int32_t qbits;
    std::memcpy(&qbits, &q, 4);
    fixed[j] = make_int4(fixed_of(a.x, origin[0], qscale), fixed_of(a.y, origin[1], qscale),
                         fixed_of(a.z, origin[2], qscale), qbits);
    rx[j] = a.x; ry[j] = a.y; rz[j] = a.z;
    rasp[j] = asp_of(a);
    rrad[j] = a.vdw_radius;
    const auto hit = lig_by_serial.find(a.serial);
    meta[j] = make_int2(a.serial, hit != lig_by_serial.end() ? 1 : 0);
    if (hit != lig_by_serial.end())
      for (int32_t l : hit->second) coll.push_back(make_int4(j, l, cls, 0));
  }
  // Sub-tile boxes: the kernel's own FP32 A coordinates of each receptor slot
  // (FP32 of the fixed-point integer times the power-of-two unit), so a box
  // holds every slot's coordinates exactly and a box beyond reach of a
  // ligand block implies every slot in it is.
  const int32_t n_sub = (n_pad + kRecSub - 1) / kRecSub;
  std::vector<float4> sub_box(2 * static_cast<size_t>(std::max(n_sub, 1)));
  const float unit_f = std::ldexp(1.0f, -qexp);
  for (int32_t t = 0; t < n_sub; ++t) {
    float4 lo = make_float4(INFINITY, INFINITY, INFINITY, 0.0f), hi = make_float4(-INFINITY, -INFINITY, -INFINITY, 0.0f);
    for (int32_t j = t * kRecSub; j < std::min(n_pad, (t + 1) * kRecSub); ++j) {
      if (perm[j] < 0) continue;
      const float x = static_cast<float>(fixed[j].x) * unit_f, y = static_cast<float>(fixed[j].y) * unit_f,
                  z = static_cast<float>(fixed[j].z) * unit_f;
      lo.x = std::min(lo.x, x); lo.y = std::min(lo.y, y); lo.z = std::min(lo.z, z);
      hi.x = std::max(hi.x, x); hi.y = std::max(hi.y, y); hi.z = std::max(hi.z, z);
    }
    sub_box[2 * t] = lo;
    sub_box[2 * t + 1] = hi;
  }
  std::vector<double> lx(n_lig), ly(n_lig), lz(n_lig), lrm(n_lig), leps(n_lig), lq(n_lig), lasp(n_lig), lrad(n_lig);
  std::vector<int32_t> lser(n_lig);
  // padded with zeros to whole chunks of the largest lane count: the kernel
  // reads a lane's factor without a bounds test
  const size_t n_lig_pad = (static_cast<size_t>(n_lig) + kScoreThreads * kMaxAtomsPerLane - 1) /
                           (kScoreThreads * kMaxAtomsPerLane) * (kScoreThreads * kMaxAtomsPerLane);
  std::vector<double> lkq(std::max<size_t>(n_lig_pad, 1), 0.0);
  double lig_r = 0.0;
  for (int32_t i = 0; i < n_lig; ++i) {  // device index i holds input atom lig_perm[i]
    const gdock_atom& a = cx->lig[lig_perm[i]];
    lx[i] = a.x; ly[i] = a.y; lz[i] = a.z;
    lrm[i] = a.rmin2; leps[i] = a.epsilon; lq[i] = a.charge;
    lasp[i] = asp_of(a); lrad[i] = a.vdw_radius; lser[i] = a.serial;
    lkq[i] = kCoulomb * a.charge / charge_unit * std::ldexp(1.0, 2 * qexp);  // the sums hold q_r / s, s = d^2 / q^2
    const double dx = a.x - lcom[0], dy = a.y - lcom[1], dz = a.z - lcom[2];
    lig_r = std::max(lig_r, std::sqrt(dx * dx + dy * dy + dz * dz));
  }
  const double com_off = std::max({std::fabs(lcom[0] - origin[0]), std::fabs(lcom[1] - origin[1]),
                                   std::fabs(lcom[2] - origin[2])});
  // Bounding sphere of each ligand block (centre relative to the COM): the
  // per-pose cost estimates of the launch order rotate these, not the atoms.
  const int32_t n_blocks = (n_lig + kLigBlock - 1) / kLigBlock;
  std::vector<float4> lsph(static_cast<size_t>(std::max(n_blocks, 1)), make_float4(0.0f, 0.0f, 0.0f, 0.0f));
  for (int32_t b = 0; b < n_blocks; ++b) {
    const int32_t i0 = b * kLigBlock, i1 = std::min(n_lig, i0 + kLigBlock);
    double c[3] = {0.0, 0.0, 0.0};
    for (int32_t i = i0; i < i1; ++i) { c[0] += lx[i] - lcom[0]; c[1] += ly[i] - lcom[1]; c[2] += lz[i] - lcom[2]; }
    for (double& v : c) v /= (i1 - i0);
    double r = 0.0;
    for (int32_t i = i0; i < i1; ++i) {
      const double dx = lx[i] - lcom[0] - c[0], dy = ly[i] - lcom[1] - c[1], dz = lz[i] - lcom[2] - c[2];
      r = std::max(r, std::sqrt(dx * dx + dy * dy + dz * dz));
    }
    lsph[b] = make_float4(static_cast<float>(c[0]), static_cast<float>(c[1]), static_cast<float>(c[2]),
                          static_cast<float>(r + 1e-3));
  }

  // restraints.rs:44-77: first "CA" with the residue number, receptor and ligand.
  std::vector<double> rr_xyz;
  std::vector<int32_t> rl_idx;
  for (int32_t p = 0; p < n_pairs; ++p) {
    int32_t ri = -1, li = -1;
    for (int32_t i = 0; i < n_rec && ri < 0; ++i)
      if (field(rec[i].name, sizeof rec[i].name) == "CA" && rec[i].resseq == pairs[2 * p]) ri = i;
    for (int32_t i = 0; i < n_lig && li < 0; ++i)
      if (field(lig[i].name, sizeof lig[i].name) == "CA" && lig[i].resseq == pairs[2 * p + 1]) li = i;
    if (ri < 0 || li < 0) continue;  // dropped with a warning in the reference
    rr_xyz.insert(rr_xyz.end(), {rec[ri].x, rec[ri].y, rec[ri].z});
    rl_idx.push_back(cx->lig_slot[li]);
  }
  cx->n_restr = static_cast<int32_t>(rl_idx.size());

  Layout L;
  const size_t o_fixed = L.add<int4>(n_pad), o_rx = L.add<double>(n_pad), o_ry = L.add<double>(n_pad),
               o_rz = L.add<double>(n_pad), o_rasp = L.add<double>(n_pad), o_rrad = L.add<double>(n_pad),
               o_meta = L.add<int2>(n_pad), o_coll = L.add<int4>(coll.size()),
               o_lorig = L.add<int32_t>(n_lig), o_cs = L.add<int32_t>(nc + 1), o_crm = L.add<double>(nc),
               o_ceps = L.add<double>(nc), o_lx = L.add<double>(n_lig), o_ly = L.add<double>(n_lig),
               o_lz = L.add<double>(n_lig), o_lrm = L.add<double>(n_lig), o_leps = L.add<double>(n_lig),
               o_lq = L.add<double>(n_lig), o_lasp = L.add<double>(n_lig), o_lrad = L.add<double>(n_lig),
               o_lser = L.add<int32_t>(n_lig), o_lkq = L.add<double>(lkq.size()), o_rrx = L.add<double>(rr_xyz.size()),
               o_rli = L.add<int32_t>(rl_idx.size()), o_dev = L.add<DevComplex>(1),
               o_lcls = L.add<float2>(static_cast<size_t>(nc) * n_lig), o_sbox = L.add<float4>(sub_box.size()),
               o_lsph = L.add<float4>(lsph.size());
  // Lennard-Jones factors of every (receptor class, ligand atom) pair, with
  // the FP64 operations of fitness.rs:58-63 (R = rmin2_r + rmin2_l;
  // sqrt(eps_r eps_l)), as q^2 / R^2 in the kernel's squared units (q =
  // 2^-qexp A), rounded once to FP32.  1 / R^2 is capped at 5e4 (R >= 4.5e-3 A): only
  // unparameterized pairs (R = 0, eps = 0, a zero term) and nearly
  // coincident atoms notice the cap, and with it every product of the pair
  // step stays finite.
  std::vector<float2> lcls(static_cast<size_t>(nc) * n_lig);
  const double q2 = std::ldexp(1.0, -2 * qexp);
  for (int c = 0; c < nc; ++c)
    for (int32_t i = 0; i < n_lig; ++i) {
      const double R = class_rmin[c] + lrm[i];
      const double inv = R * R > 2e-5 ? 1.0 / (R * R) : 5e4;
      float2& f = lcls[static_cast<size_t>(c) * n_lig + i];
      f.x = static_cast<float>(std::min(inv, 5e4) * q2);
      f.y = static_cast<float>(std::sqrt(class_eps[c] * leps[i]));
    }
  cx->contacts_layout_begin = L.size;
  std::vector<uint8_t> host;
  build_contact_tables(cx, &L, &host);  // FCC residue tables appended to the same arena
  std::vector<uint8_t> blob(L.size, 0);
  auto put = [&](size_t off, const void* src, size_t bytes) { if (bytes) std::memcpy(blob.data() + off, src, bytes); };
  put(o_fixed, fixed.data(), fixed.size() * sizeof(int4));
  put(o_rx, rx.data(), n_pad * 8); put(o_ry, ry.data(), n_pad * 8); put(o_rz, rz.data(), n_pad * 8);
  put(o_rasp, rasp.data(), n_pad * 8); put(o_rrad, rrad.data(), n_pad * 8);
  put(o_meta, meta.data(), n_pad * sizeof(int2));
  put(o_coll, coll.data(), coll.size() * sizeof(int4));
  put(o_lorig, lig_perm.data(), lig_perm.size() * 4);
  put(o_cs, class_start.data(), (nc + 1) * 4); put(o_crm, class_rmin.data(), nc * 8); put(o_ceps, class_eps.data(), nc * 8);
  put(o_lx, lx.data(), n_lig * 8); put(o_ly, ly.data(), n_lig * 8); put(o_lz, lz.data(), n_lig * 8);
  put(o_lrm, lrm.data(), n_lig * 8); put(o_leps, leps.data(), n_lig * 8); put(o_lq, lq.data(), n_lig * 8);
  put(o_lasp, lasp.data(), n_lig * 8); put(o_lrad, lrad.data(), n_lig * 8); put(o_lser, lser.data(), n_lig * 4);
  put(o_lkq, lkq.data(), lkq.size() * 8);
  put(o_rrx, rr_xyz.data(), rr_xyz.size() * 8); put(o_rli, rl_idx.data(), rl_idx.size() * 4);
  put(o_lcls, lcls.data(), lcls.size() * sizeof(float2));
  put(o_sbox, sub_box.data(), sub_box.size() * sizeof(float4));
  put(o_lsph, lsph.data(), lsph.size() * sizeof(float4));
  if (!host.empty()) std::memcpy(blob.data() + cx->contacts_layout_begin, host.data(), host.size());

  hipError_t e = hipSetDevice(ctx->device);
  if (e == hipSuccess) e = hipMalloc(&cx->mem, std::max<size_t>(L.size, 256));
  if (e != hipSuccess) { delete cx; return hip_fail(ctx, e, "complex_create: hipMalloc"); }
  cx->mem_bytes = std::max<size_t>(L.size, 256);
  char* m = static_cast<char*>(cx->mem);
  DevComplex& d = cx->dev;
  d.rec_fixed = reinterpret_cast<const int4*>(m + o_fixed);
  d.rec_x = reinterpret_cast<const double*>(m + o_rx);
  d.rec_y = reinterpret_cast<const double*>(m + o_ry);
  d.rec_z = reinterpret_cast<const double*>(m + o_rz);
  d.rec_asp = reinterpret_cast<const double*>(m + o_rasp);
  d.rec_radius = reinterpret_cast<const double*>(m + o_rrad);
  d.rec_meta = reinterpret_cast<const int2*>(m + o_meta);
  d.coll_pairs = reinterpret_cast<const int4*>(m + o_coll);
  d.lig_orig = reinterpret_cast<const int32_t*>(m + o_lorig);
  d.class_start = reinterpret_cast<const int32_t*>(m + o_cs);
  d.class_rmin2 = reinterpret_cast<const double*>(m + o_crm);
  d.class_eps = reinterpret_cast<const double*>(m + o_ceps);
  d.lig_x = reinterpret_cast<const double*>(m + o_lx);
  d.lig_y = reinterpret_cast<const double*>(m + o_ly);
  d.lig_z = reinterpret_cast<const double*>(m + o_lz);
  d.lig_rmin2 = reinterpret_cast<const double*>(m + o_lrm);
  d.lig_eps = reinterpret_cast<const double*>(m + o_leps);
  d.lig_q = reinterpret_cast<const double*>(m + o_lq);
  d.lig_asp = reinterpret_cast<const double*>(m + o_lasp);
  d.lig_radius = reinterpret_cast<const double*>(m + o_lrad);
  d.lig_serial = reinterpret_cast<const int32_t*>(m + o_lser);
  d.lig_kq = reinterpret_cast<const double*>(m + o_lkq);
  d.restr_rec_xyz = reinterpret_cast<const double*>(m + o_rrx);
  d.restr_lig = reinterpret_cast<const int32_t*>(m + o_rli);
  d.lig_class = reinterpret_cast<const float2*>(m + o_lcls);
  d.rec_sub_box = reinterpret_cast<const float4*>(m + o_sbox);
  d.lig_block_sphere = reinterpret_cast<const float4*>(m + o_lsph);
  d.n_rec = n_pad; d.n_rec_atoms = n_rec; d.n_lig = n_lig; d.n_class = nc; d.n_restr = cx->n_restr;
  d.n_coll = static_cast<int32_t>(coll.size());
  for (int i = 0; i < 3; ++i) { d.com[i] = lcom[i]; d.origin[i] = origin[i]; }
  d.rec_extent = rec_ext;
  d.qexp = qexp;
  fixed_bands(qexp, d.band_c, &d.band_w);
  d.q_scale = charge_unit;
  d.parts = 1;
  d.min_parts = 1;
  d.lig_extent = lig_r + com_off;
  bind_contact_tables(cx);
  cx->d_dev = reinterpret_cast<const DevComplex*>(m + o_dev);
  std::memcpy(blob.data() + o_dev, &cx->dev, sizeof(DevComplex));
  e = hipMemcpyAsync(cx->mem, blob.data(), L.size, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) { (void)hipFree(cx->mem); delete cx; return hip_fail(ctx, e, "complex_create: upload"); }
  *out = cx;
  return GDOCK_OK;
}

void gdock_complex_destroy(gdock_ctx* ctx, gdock_complex* cx) {
  if (!cx) return;
  if (ctx) { (void)hipSetDevice(ctx->device); (void)hipStreamSynchronize(ctx->stream); }
  if (cx->mem) (void)hipFree(cx->mem);
  delete cx;
}

int32_t gdock_complex_n_restraints(const gdock_complex* cx) { return cx ? cx->n_restr : -1; }

gdock_status gdock_prepare_poses(const double* genes, int32_t n, double* poses) {
  if (n < 0 || (n && (!genes || !poses))) return GDOCK_ERR_INVALID_ARG;
  for (int64_t i = 0; i < static_cast<int64_t>(n) * 6; ++i)
    if (!std::isfinite(genes[i])) return GDOCK_ERR_INVALID_ARG;
  // A generation of a large batch is thousands of poses: split it over the
  // host pool (each pose is independent, so the split changes no value).
  constexpr int32_t kSplitAt = 2048;
  if (n >= kSplitAt) {
    host_pool_for(n, [&](int64_t lo, int64_t hi) { pose_records(genes, poses, lo, hi); });
  } else {
    pose_records(genes, poses, 0, n);
  }
  return GDOCK_OK;
}

gdock_status gdock_score_poses_device(gdock_ctx* ctx, gdock_complex* cx, const double* d_poses, int32_t n,
                                      const gdock_weights* w, double* d_out, int32_t* d_nsat, void* stream) {
  if (!ctx || !cx || !w || n < 0 || (n && (!d_poses || !d_out))) return GDOCK_ERR_INVALID_ARG;
  const Weights ww{w->vdw, w->elec, w->desolv, w->air};
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
  hipError_t e = hipSuccess;
  if (!ctx->redo.reserve(plain_scratch_bytes(cx->dev, n), &e)) return hip_fail(ctx, e, "score_poses_device: scratch");
  e = launch_score(cx->dev, cx->d_dev, d_poses, nullptr, n, ww, d_out, d_nsat, ctx->redo.ptr, s);
  return e == hipSuccess ? GDOCK_OK : hip_fail(ctx, e, "score_poses_device: launch");
}

// The same kernels as gdock_score_poses_device, with the per-pose rescore
// flags in the device's static scratch instead of the context's arena: no
// allocation of any kind.  The fast kernel writes every pose's flag before
// the exact kernel reads it, so the scratch needs no clearing; two calls that
// overlap on one device share the scratch and race.
gdock_status gdock_score_poses_device_noalloc(gdock_ctx* ctx, gdock_complex* cx, const double* d_poses, int32_t n,
                                              const gdock_weights* w, double* d_out, int32_t* d_nsat, void* stream) {
  if (!ctx || !cx || !w || n < 0 || (n && (!d_poses || !d_out))) {
    if (ctx) set_error(ctx, "score_poses_device_noalloc: bad arguments");
    return GDOCK_ERR_INVALID_ARG;
  }
  if (n > kNoAllocMaxPoses) {
    set_error(ctx, "score_poses_device_noalloc: %d poses exceed the limit of %d per call", n, kNoAllocMaxPoses);
    return GDOCK_ERR_INVALID_ARG;
  }
  if (plain_parts(cx->dev) > 1) {
    set_error(ctx, "score_poses_device_noalloc: a receptor of %d slots needs split-pose state; use "
              "gdock_score_poses_device", cx->dev.n_rec);
    return GDOCK_ERR_INVALID_ARG;
  }
  if (plain_scratch_bytes(cx->dev, n) > sizeof(scratch)) {  // the layout for n <= kNoAllocMaxPoses fits by design
    set_error(ctx, "score_poses_device_noalloc: %d poses need more than the %zu bytes of device scratch", n,
              sizeof(scratch));
    return GDOCK_ERR_INVALID_ARG;
  }
  hipError_t e = hipSuccess;
  int32_t* flags = scratch_of(ctx->device, &e);
  if (!flags) return hip_fail(ctx, e, "score_poses_device_noalloc: device scratch");
  const Weights ww{w->vdw, w->elec, w->desolv, w->air};
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
  e = launch_score(cx->dev, cx->d_dev, d_poses, nullptr, n, ww, d_out, d_nsat, flags, s);
  return e == hipSuccess ? GDOCK_OK : hip_fail(ctx, e, "score_poses_device_noalloc: launch");
}

gdock_status gdock_score_poses(gdock_ctx* ctx, gdock_complex* cx, const double* genes, int32_t n,
                               const gdock_weights* w, double* out, int32_t* n_satisfied) {
  if (!ctx || !cx || !w || n < 0 || (n && (!genes || !out))) {
    if (ctx) set_error(ctx, "score_poses: bad arguments");
    return GDOCK_ERR_INVALID_ARG;
  }
  if (n == 0) return GDOCK_OK;
  std::vector<double> poses(static_cast<size_t>(n) * kPoseDoubles);
  if (gdock_prepare_poses(genes, n, poses.data()) != GDOCK_OK) {
    set_error(ctx, "score_poses: non-finite gene");
    return GDOCK_ERR_INVALID_ARG;
  }
  hipError_t e = hipSetDevice(ctx->device);
  if (e != hipSuccess) return hip_fail(ctx, e, "score_poses: set device");
  const size_t pb = poses.size() * 8, ob = static_cast<size_t>(n) * kOutDoubles * 8, nb = static_cast<size_t>(n) * 4;
  const size_t o_out = (pb + 255) & ~static_cast<size_t>(255), o_ns = (o_out + ob + 255) & ~static_cast<size_t>(255),
               o_rd = (o_ns + nb + 255) & ~static_cast<size_t>(255);
  if (!ctx->staging.reserve(o_rd + plain_scratch_bytes(cx->dev, n), &e)) return hip_fail(ctx, e, "score_poses: staging");
  char* base = static_cast<char*>(ctx->staging.ptr);
  e = hipMemcpyAsync(base, poses.data(), pb, hipMemcpyHostToDevice, ctx->stream);
  if (e != hipSuccess) return hip_fail(ctx, e, "score_poses: upload");
  const Weights ww{w->vdw, w->elec, w->desolv, w->air};
  e = launch_score(cx->dev, cx->d_dev, reinterpret_cast<double*>(base), nullptr, n, ww,
                   reinterpret_cast<double*>(base + o_out), reinterpret_cast<int32_t*>(base + o_ns), base + o_rd,
                   ctx->stream);
  if (e != hipSuccess) return hip_fail(ctx, e, "score_poses: launch");
  e = hipMemcpyAsync(out, base + o_out, ob, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess && n_satisfied) e = hipMemcpyAsync(n_satisfied, base + o_ns, nb, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  return e == hipSuccess ? GDOCK_OK : hip_fail(ctx, e, "score_poses: download");
}

gdock_status gdock_count_pairs(gdock_ctx* ctx, gdock_complex* cx, const double* genes, int32_t n, int64_t* counts) {
  if (!ctx || !cx || !counts || n < 0 || (n && !genes)) {
    if (ctx) set_error(ctx, "count_pairs: bad arguments");
    return GDOCK_ERR_INVALID_ARG;
  }
  counts[0] = counts[1] = 0;
  if (n == 0) return GDOCK_OK;
  std::vector<double> poses(static_cast<size_t>(n) * kPoseDoubles);
  if (gdock_prepare_poses(genes, n, poses.data()) != GDOCK_OK) {
    set_error(ctx, "count_pairs: non-finite gene");
    return GDOCK_ERR_INVALID_ARG;
  }
  hipError_t e = hipSetDevice(ctx->device);
  if (e != hipSuccess) return hip_fail(ctx, e, "count_pairs: set device");
  const size_t pb = poses.size() * 8, ob = static_cast<size_t>(n) * kOutDoubles * 8;
  const size_t o_out = (pb + 255) & ~static_cast<size_t>(255), o_cnt = (o_out + ob + 255) & ~static_cast<size_t>(255),
               o_rd = o_cnt + 256;
  if (!ctx->staging.reserve(o_rd + plain_scratch_bytes(cx->dev, n), &e)) return hip_fail(ctx, e, "count_pairs: staging");
  char* base = static_cast<char*>(ctx->staging.ptr);
  e = hipMemcpyAsync(base, poses.data(), pb, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipMemsetAsync(base + o_cnt, 0, 16, ctx->stream);
  if (e != hipSuccess) return hip_fail(ctx, e, "count_pairs: upload");
  const Weights ww{0.4, 0.05, 3.4, 1.0};
  e = launch_score(cx->dev, cx->d_dev, reinterpret_cast<double*>(base), nullptr, n, ww,
                   reinterpret_cast<double*>(base + o_out), nullptr, base + o_rd, ctx->stream, 0,
                   reinterpret_cast<unsigned long long*>(base + o_cnt));
  if (e != hipSuccess) return hip_fail(ctx, e, "count_pairs: launch");
  e = hipMemcpyAsync(counts, base + o_cnt, 16, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  return e == hipSuccess ? GDOCK_OK : hip_fail(ctx, e, "count_pairs: download");
}

gdock_status gdock_score_coords(gdock_ctx* ctx, gdock_complex* cx, const double* lig_xyz, int32_t n,
                                const gdock_weights* w, double* out, int32_t* n_satisfied) {
  if (!ctx || !cx || !w || n < 0 || (n && (!lig_xyz || !out))) return GDOCK_ERR_INVALID_ARG;
  if (n == 0) return GDOCK_OK;
  const int32_t nl = cx->dev.n_lig;
  const size_t ncoord = static_cast<size_t>(n) * nl * 3;
  for (size_t i = 0; i < ncoord; ++i)
    if (!std::isfinite(lig_xyz[i])) { set_error(ctx, "score_coords: non-finite coordinate"); return GDOCK_ERR_INVALID_ARG; }
  hipError_t e = hipSetDevice(ctx->device);
  if (e != hipSuccess) return hip_fail(ctx, e, "score_coords: set device");
  const size_t cb = ncoord * 8, ob = static_cast<size_t>(n) * kOutDoubles * 8, nb = static_cast<size_t>(n) * 4;
  auto al = [](size_t v) { return (v + 255) & ~static_cast<size_t>(255); };
  const size_t o_out = al(cb), o_ns = al(o_out + ob), o_rd = al(o_ns + nb);
  if (!ctx->staging.reserve(o_rd + plain_scratch_bytes(cx->dev, n), &e)) return hip_fail(ctx, e, "score_coords: staging");
  char* base = static_cast<char*>(ctx->staging.ptr);
  e = hipMemcpyAsync(base, lig_xyz, cb, hipMemcpyHostToDevice, ctx->stream);
  if (e != hipSuccess) return hip_fail(ctx, e, "score_coords: upload");
  const Weights ww{w->vdw, w->elec, w->desolv, w->air};
  e = launch_score(cx->dev, cx->d_dev, nullptr, reinterpret_cast<double*>(base), n, ww,
                   reinterpret_cast<double*>(base + o_out), reinterpret_cast<int32_t*>(base + o_ns), base + o_rd,
                   ctx->stream);
  if (e != hipSuccess) return hip_fail(ctx, e, "score_coords: launch");
  e = hipMemcpyAsync(out, base + o_out, ob, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess && n_satisfied) e = hipMemcpyAsync(n_satisfied, base + o_ns, nb, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  return e == hipSuccess ? GDOCK_OK : hip_fail(ctx, e, "score_coords: download");
}

gdock_status gdock_eval_fitness(gdock_ctx* ctx, gdock_complex* cx, gdock_chromosome* ch, int32_t n,
                                const gdock_weights* w) {
  if (!ctx || !cx || !w || n < 0 || (n && !ch)) return GDOCK_ERR_INVALID_ARG;
  std::vector<double> genes(static_cast<size_t>(n) * 6), out(static_cast<size_t>(n) * kOutDoubles);
  for (int32_t i = 0; i < n; ++i) std::memcpy(&genes[6 * i], ch[i].genes, 6 * sizeof(double));
  const gdock_status s = gdock_score_poses(ctx, cx, genes.data(), n, w, out.data(), nullptr);
  if (s != GDOCK_OK) return s;
  for (int32_t i = 0; i < n; ++i) {  // chromosome.rs:111-127
    const double* o = &out[kOutDoubles * i];
    ch[i].vdw = o[0]; ch[i].elec = o[1]; ch[i].desolv = o[2]; ch[i].air = o[3];
    ch[i].restraint_penalty = o[4]; ch[i].fitness = o[5];
  }
  return GDOCK_OK;
}

}  // extern "C"
