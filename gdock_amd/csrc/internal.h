// Host-side internals shared by the library's translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <mutex>
#include <string>
#include <string_view>
#include <vector>

#include "../../include/gdock_hip.h"
#include "common.h"

namespace gdock {

struct DeviceArena {
  void* ptr = nullptr;
  size_t cap = 0;
  bool reserve(size_t bytes, hipError_t* err);
  void release();
};

// Device tables of the inter-chain residue contacts of a complex.
//
// Residue keys are (chain, resseq) over the heavy atoms of receptor and
// ligand together, so a key may hold receptor atoms (a receptor group),
// ligand atoms (a ligand group) or both.  The contacts of a pose are the key
// pairs of different chains with an atom pair within the cutoff: receptor x
// ligand atom pairs move with the pose (the kernel), receptor x receptor
// pairs never do, ligand x ligand pairs only by the rounding of the pose
// transform (decided on the host, except pairs within a rounding window of
// the cutoff, which the kernel re-decides per pose).
struct DevContacts {
  const float4* rec_c4;       // receptor heavy atoms grouped by receptor group: {-2x, -2y, -2z, x^2+y^2+z^2} of the FP32 origin-frame coordinates
  const double* rec_x;        // FP64 coordinates of the same atoms
  const double* rec_y;
  const double* rec_z;
  const int32_t* rec_res_start;  // n_rec_res + 1
  const int32_t* lig_atom;    // ligand heavy atom -> ligand atom index
  const int32_t* lig_res;     // ligand heavy atom -> ligand group index
  const float4* rec_res_sphere;  // per receptor group: bounding sphere of its heavy atoms (FP32, origin frame)
  const int32_t* pair_bit;    // (receptor group * n_lig_res + ligand group) -> bit, -1: same chain; null: the identity
  const uint64_t* static_mask;  // per call: bits set in every pose (words64), or null
  const int4* window;         // per call: {ligand atom a, ligand atom b, bit, 0} pairs re-decided per pose
  int32_t n_window;
  int32_t n_rec_heavy, n_lig_heavy, n_rec_res, n_lig_res;
  int32_t n_dyn;              // bits of receptor x ligand group pairs (0: no such candidate)
  int32_t words64;            // per call: uint64 words per pose bitset
  int32_t supported;          // 0: tables not built
};

// Contact bit layout of one cutoff: bits [0, n_dyn) are the receptor x
// ligand key pairs (cutoff independent), then key pairs that only receptor x
// receptor or ligand x ligand atom pairs can bring into contact at this
// cutoff.
struct ContactCut {
  bool valid = false;
  double d2 = -1.0;                                   // FP64 squared-distance threshold it was built for
  int32_t n_bits = 0, words64 = 0;
  std::vector<std::pair<int32_t, int32_t>> extra_keys;  // key pairs of bits n_dyn..
  std::vector<uint64_t> static_mask;                  // bits set in every pose
  std::vector<int4> window;                           // ligand atom pairs re-decided per pose
  double max_translation = 0.0;                       // largest pose translation for which the ligand pairs' decisions hold
  std::string err;                                    // non-empty: this cutoff is not supported
};

}  // namespace gdock

struct gdock_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  gdock::DeviceArena staging;   // per-call poses / outputs
  gdock::DeviceArena fcc;       // FCC bitsets, overlap and cluster scratch
  gdock::DeviceArena redo;      // per-pose exact-rescore flags of the device-resident entry
  void* pinned = nullptr;       // pinned host staging
  size_t pinned_cap = 0;
};

struct gdock_complex {
  gdock::DevComplex dev{};                  // host copy (sizes, pointers)
  const gdock::DevComplex* d_dev = nullptr;  // the same struct in device memory
  gdock::DevContacts con{};
  void* mem = nullptr;
  size_t mem_bytes = 0;
  std::vector<gdock_atom> rec, lig;
  std::vector<int32_t> lig_slot;  // original ligand atom -> its index in the device arrays
  int32_t n_restr = 0;
  // contact tables (host)
  size_t contacts_layout_begin = 0;
  std::vector<size_t> con_offsets;
  std::vector<std::pair<char, int16_t>> res_key;        // residue keys, ascending
  std::vector<int32_t> rec_group_key, lig_group_key;    // group -> key
  std::vector<int32_t> rg_of_key, lg_of_key;            // key -> group or -1
  std::vector<int32_t> pair_bit;                        // host copy (empty: the identity)
  std::vector<std::pair<int32_t, int32_t>> dyn_keys;    // bit < n_dyn -> key pair
  std::vector<double> rec_heavy_xyz, lig_heavy_xyz;     // FP64 coordinates (x, y, z) of heavy atoms
  std::vector<int32_t> rec_heavy_key, lig_heavy_key;    // their keys
  std::vector<int32_t> lig_heavy_slot;                  // ligand heavy atom -> device ligand index
  int32_t rec_chains = 0, lig_chains = 0;               // distinct chains among receptor / ligand heavy atoms
  std::string con_unsupported;
  std::mutex con_mu;                                    // guards `cut`
  gdock::ContactCut cut;                                // layout of the last cutoff used
};

namespace gdock {

void set_error(gdock_ctx* ctx, const char* fmt, ...);
gdock_status hip_fail(gdock_ctx* ctx, hipError_t e, const char* what);
std::string_view field(const char* f, size_t n);

void read_pdb_text(std::string_view text, std::vector<gdock_atom>* atoms, std::vector<int32_t>* offsets);

// Scores n_poses poses: the fast kernel, then the exact kernel for any pose
// whose settle buffer overflowed.  scratch: plain_scratch_bytes(cx, n_poses)
// of device memory.  lanes_k: ligand atoms per lane (1-8); 0 picks
// ceil(n_lig / 256) capped at 8.  counts (diagnostic): when set, only the
// counting instantiation of the fast kernel runs and adds {pair steps x 64,
// pairs within 15 A} to counts[0..1] (the outputs are then not valid).
// min_items > 0 splits the poses of a small launch over receptor ranges until
// the launch holds that many workgroups: a pose's sums then associate by
// range, so its last bits depend on the launch's pose count (the GA drivers,
// whose launches are one population each, pass kGaFillItems; the scoring
// entry points pass 0, and a pose scores the same alone and in any launch).
hipError_t launch_score(const DevComplex& cx, const DevComplex* d_cx, const double* d_poses, const double* d_coords,
                        int n_poses, const Weights& w, double* d_out, int32_t* d_nsat, void* scratch,
                        hipStream_t stream, int lanes_k = 0, unsigned long long* counts = nullptr,
                        int min_items = 0);
int score_atoms_per_lane(int n_lig);
// Receptor-range parts of the plain path (1 unless the receptor exceeds one
// workgroup's LDS table; see launch_score for min_items) and its scratch
// bytes (flags, split state).
int plain_parts(const DevComplex& cx, int n_poses = 0, int min_items = 0);
// Equal whole-tile receptor ranges (none empty, at most kMaxPoseParts) that
// bring n_poses poses to min_items workgroups; 1 when min_items is 0.
int fill_parts(const DevComplex& cx, int n_poses, int min_items);
size_t plain_scratch_bytes(const DevComplex& cx, int n_poses, int min_items = 0);
// Most receptor slots one of `parts` equal tile ranges covers.
int part_slots(int n_rec, int parts);
size_t score_lds_bytes(int slots);

size_t batch_lds_bytes(int max_slots);
int batch_grid(size_t lds, int device);
hipError_t launch_batch(const BatchLaunch& b, const Weights& w, hipStream_t stream);
// Device bytes of the per-launch item order for `capacity` items, and the
// binding of BatchLaunch::sorted / bucket / hist into such a block (which
// must start zeroed).
size_t order_bytes(size_t capacity);
void bind_order(BatchLaunch* b, char* mem, size_t capacity);

// Fork-join of f(lo, hi) over slices of [0, n) on the library's host threads
// (the caller takes one slice; a call made while another runs, or in a child
// process forked after the threads started, runs alone).
void host_parallel_for(int64_t n, const std::function<void(int64_t, int64_t)>& f);

// Host-side bump layout of one device allocation (256-byte aligned pieces).
struct Layout {
  size_t size = 0;
  template <class T>
  size_t add(size_t count) {
    const size_t off = (size + 255) & ~static_cast<size_t>(255);
    size = off + count * sizeof(T);
    return off;
  }
};
// Orders idx[lo, hi) (atom indices) so that aligned runs of `block` entries are
// spatially compact (recursive bisection across the widest extent).
void bisect_order(const gdock_atom* atoms, std::vector<int32_t>& idx, size_t lo, size_t hi, size_t block);

// Ligand atoms per lane of a complex shape (a pure function of the shape),
// and the autotuner's candidates for a ligand size.
int choose_lanes(const DevComplex& d);
std::vector<int> lane_candidates(int n_lig);
// Parts of a pose in batched launches: at least min_batch_parts (receptor
// ranges of at most kPartSlots slots), at most pose_parts (also 1 = never
// split, at most kMaxPoseParts); each launch picks every pose's count.
int min_batch_parts(const DevComplex& d);
int pose_parts(const DevComplex& d, int lanes_k);
int batch_part_slots(const DevComplex& d);
// Device bytes of the split state of `poses` poses, and its binding into a
// table entry (pose rows [pose_base, pose_base + poses), state cleared).
size_t split_state_bytes(const DevComplex& d, int parts, int poses);
void bind_split_state(DevComplex* d, int parts, int poses, int pose_base, char* mem);

void build_contact_tables(gdock_complex* cx, Layout* L, std::vector<uint8_t>* host);
void bind_contact_tables(gdock_complex* cx);
void release_fcc_scratch(gdock_ctx* ctx);

}  // namespace gdock
