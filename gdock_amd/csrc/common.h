// Shared host/device definitions of the gfx950 scoring and FCC library.
#pragma once
#include <cstdint>
#include <climits>

namespace gdock {

// constants.rs:22-72, fitness.rs:129-138,286
constexpr double kPI = 3.14159265358979323846;
constexpr double kCoulomb = 332.0636;
constexpr double kAirK = 10.0;
constexpr double kAirUpper = 7.0;

// Exact FP64 squared-distance forms of the reference's sqrt-based predicates
// (a correctly rounded sqrt(d2) compared with the cutoff).  Each equality is
// checked by brute force around the boundary in the test suite.
constexpr double kD2ElecMinExcl = 1.0000000000000002;    // dist > 1.0  <=> d2 >  this
constexpr double kD2DesolvIncl = 63.99999999999999;      // dist < 8.0  <=> d2 <= this
constexpr double kD2VdwIncl = 143.99999999999997;        // dist < 12.0 <=> d2 <= this
constexpr double kD2ElecMaxIncl = 224.99999999999997;    // dist < 15.0 <=> d2 <= this
constexpr double kD2RestraintIncl = 49.0;                // dist <= 7.0 <=> d2 <= this
constexpr double kD2ContactIncl = 25.000000000000004;    // dist <= 5.0 <=> d2 <= this

// Pose record: genes[6], cos/sin of alpha, beta, gamma, then the four
// trigonometric products of the rotation that do not involve an atom
// coordinate (m01, m02, m11, m12, structure.rs:96-101), all FP64 from the host.
constexpr int kPoseDoubles = 16;
constexpr int kOutDoubles = 6;    // vdw, elec, desolv, air, restraint_penalty, fitness

constexpr int kScoreThreads = 256;  // 4 wave64 per pose workgroup
constexpr int kMaxAtomsPerLane = 8;
constexpr int kMaxPoseParts = 15;  // workgroups one pose can be split over (a 4-bit field of a work item)
// The GA drivers' plain launches hold at least this many workgroups (two per
// CU of a 256-CU device), splitting a small population's poses over receptor
// ranges; a fixed count, so a run gives the same bits on any device.
constexpr int kGaFillItems = 512;
// Receptor slots one workgroup covers: the whole receptor in the plain
// one-workgroup-per-pose path up to kMaxLdsReceptor slots (larger receptors
// are split over receptor ranges), at most kPartSlots slots per workgroup in
// batched launches (the burial-count table is 4 B per slot of LDS: 32 KiB
// leaves room for 4 workgroups per CU).
constexpr int kMaxLdsReceptor = 16384;
constexpr int kPartSlots = 8192;
constexpr int kMaxReceptorAtoms = kMaxPoseParts * kPartSlots - 4 * 64;
constexpr float kPadCoord = 1.0e4f;      // padding ligand lanes of the contact kernel sit far away
constexpr int kRecGroup = 4;             // receptor atoms per scalar load
constexpr int kLigBlock = 64;            // ligand atoms per spatial block (one wave's slot)
constexpr int kLigSub = 16;              // ligand atoms per spatial sub-block (a quarter of a block)
constexpr int kRecSub = 16;              // receptor slots per sub-tile box
// A receptor group and a wave's ligand atoms farther apart than this have no
// pair within any cutoff (15 A) or its guard band.
constexpr float kCullReach = 15.01f;

// Fixed-point coordinates of the scoring kernels: integers in units of
// 2^-qexp A in the receptor-COM frame, so a pair's difference is exact and
// only its FP32 conversion rounds (relative to the distance, not the frame).
// Ligand coordinates are clamped to +-kFixedClamp units; receptor atoms lie
// within kFixedClamp units - 16 A, so a clamped atom is beyond every cutoff of
// every receptor atom.  Receptor padding slots hold kFixedPad (far from the
// clamped cube; never evaluated) and ligand padding lanes kFixedClamp.
constexpr int32_t kFixedClamp = (1 << 30) - 1;
constexpr int32_t kFixedPad = INT32_MIN;
constexpr int kFixedMaxExp = 24;
constexpr int kFixedMinExp = 8;

// Device view of one complex (kernel argument, plain pointers).
struct DevComplex {
  // receptor, permuted so that atoms of one (rmin2, epsilon) class are contiguous
  const int4* rec_fixed;       // x, y, z fixed point (units of 2^-qexp A, origin frame); w: FP32 bits of the charge in units of 1 / q_scale
  const double* rec_x;         // FP64 original coordinates (permuted)
  const double* rec_y;
  const double* rec_z;
  const double* rec_asp;       // atomic solvation parameter (0 = skipped)
  const double* rec_radius;
  const int2* rec_meta;        // {serial, 1 if the serial also occurs in the ligand}
  const int4* coll_pairs;      // {receptor slot, ligand atom, receptor class, 0}: equal serials
  const int32_t* class_start;  // n_class + 1
  const double* class_rmin2;
  const double* class_eps;
  const float2* lig_class;     // n_class x n_lig: {q^2 / R^2 (q: the fixed-point unit), sqrt(eps_r eps_l)} per receptor class and ligand atom
  const float4* lig_block_sphere;  // per kLigBlock ligand atoms (device order): {centre x, y, z relative to com, radius} in FP32 A
  const float4* rec_sub_box;   // per kRecSub receptor slots: {lo x, y, z, 0}, {hi x, y, z, 0} of the slots' FP32 A coordinates (padding excluded; empty: lo +inf, hi -inf)
  // ligand (positioned), in spatial block order (lig_orig: index in the input)
  const double* lig_x;
  const double* lig_y;
  const double* lig_z;
  const double* lig_rmin2;
  const double* lig_eps;
  const double* lig_q;
  const double* lig_asp;
  const double* lig_radius;
  const int32_t* lig_serial;
  const int32_t* lig_orig;
  const double* lig_kq;        // kCoulomb * charge / q_scale / q^2 per ligand atom (the elec fold factor; sums hold q_r / s), zero-padded to whole chunks
  // restraints: receptor CA coordinates (FP64) and ligand CA index
  const double* restr_rec_xyz;  // n_restr x 3
  const int32_t* restr_lig;     // n_restr
  int32_t n_rec;        // receptor slots including class padding (multiple of kRecGroup)
  int32_t n_rec_atoms;  // real receptor atoms
  int32_t n_lig, n_class, n_restr;
  int32_t n_coll;       // pairs in coll_pairs
  int32_t qexp;         // fixed-point exponent: coordinates in units of 2^-qexp A
  // FP32 decisions on the squared distance in units^2, as integers of its bit
  // pattern: bits(s) - band_c[T] borrows exactly when s is below the guard band
  // of cutoff T (1, 8, 12, 15 A), and lies within band_w of it inside the band.
  uint32_t band_c[4];
  uint32_t band_w;
  double q_scale;       // receptor charges are stored as charge * q_scale (exact integers when q_scale > 1)
  int32_t lanes_k;      // ligand atoms per lane chosen for batched launches (0: ceil(n_lig / 256), at most 8)
  // batched launches: a pose may be split across up to `parts` workgroups by
  // receptor range (1: never split; each launch picks every pose's count),
  // with the split state of the complex's poses, indexed by global pose row -
  // split_pose_base
  int32_t parts;
  int32_t min_parts;    // parts every pose takes at least (receptor slots over the per-workgroup LDS table)
  int32_t split_pose_base;
  int32_t* split_lig;   // n_lig ints per pose
  double* split_sums;   // 3 doubles per part (up to `parts`) per pose
  int32_t* split_sync;  // 2 ints per pose
  double com[3];     // ligand centre of mass (rotation centre, structure.rs:82)
  double origin[3];  // fixed-point frame origin (receptor COM)
  double rec_extent;   // max |rec - origin| component
  double lig_extent;   // max |lig - com| (2-norm) + max |com - origin| component
};

struct Weights { double vdw, elec, desolv, air; };

// Cost buckets of the per-launch item order (half-octave steps of the
// estimated pair work).
constexpr int kCostBuckets = 64;
// The plain path orders the poses of a launch by cost when its workgroups
// exceed one round of an MI355X's slots (256 CUs x 4 scoring workgroups) and
// it has at most kOrderMaxPoses poses (one workgroup of kOrderThreads places
// them).
constexpr int kOrderMinItems = 1024;
constexpr int kOrderMaxPoses = 1 << 16;
constexpr int kOrderThreads = 1024;
constexpr int kOrderBuckets = 256;  // eighth-octave steps of the estimated cost
// Estimated cost of one work item (receptor slots within reach of the
// ligand, plus 1/16 of the slots walked for the per-tile tests, times the
// ligand lane-slots): a pose estimated above it is split by receptor range
// into ceil(cost / kSplitCost) parts, up to the complex's largest part count.
constexpr double kSplitCost = 1.5e6;
constexpr int kItemPartShift = 20, kItemPartsShift = 24;
constexpr int kItemSlotMask = (1 << kItemPartShift) - 1;  // complex slots per launch table

// Device buffers of one batched (persistent) scoring launch.
struct BatchLaunch {
  const DevComplex* table;     // complex slots (device)
  const int2* pose_list;       // {complex slot, pose row} per pose of the launch
  const int32_t* n_pose_list;  // device-resident pose count
  // Work items, written per launch by the planning kernels: {slot | part <<
  // kItemPartShift | parts << kItemPartsShift, pose row}, most expensive first.
  int2* items;
  int32_t* n_items;
  int32_t* counter;         // work counter and exit count: zero between launches
  int32_t* exit_count;
  const double* poses;      // pose records, kPoseDoubles each
  double* out;              // kOutDoubles per pose
  int32_t* nsat;
  int32_t* redo;
  uint16_t* plan;           // per listed pose: cost bucket | parts << 8
  int32_t* hist;            // kCostBuckets counts, then kCostBuckets cursors: zero between launches
  int32_t* redo_count;      // poses of the launch flagged for the exact variant
  int grid;                 // persistent workgroups
  int capacity;             // pose capacity of the list
  int item_capacity;        // item capacity (capacity x the largest part count)
  size_t lds;
};


}  // namespace gdock
