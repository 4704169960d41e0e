// FCC clustering on gfx950: residue contact bitsets, pairwise overlap with
// AND + popcount, threshold graph, greedy clustering (clustering.rs:62-293).
//
// Contacts: one workgroup per pose; ligand heavy atoms (FP64 pose transform,
// then FP32) sit in lanes, receptor heavy atoms are wave-uniform scalar loads
// walked residue by residue.  The 5 A test is FP32 against a guard band with
// an FP64 re-decision inside the band, so every contact bit equals the FP64
// reference's decision.  Residues are (chain, resseq) keys over both
// molecules; receptor x ligand key pairs of different chains take the first
// bits (bit = receptor group * n_lig_res + ligand group for one receptor
// chain against one other ligand chain), and key pairs that only receptor x
// receptor or ligand x ligand atoms can bring into contact follow, decided
// once per cutoff on the host (a ligand x ligand atom pair within 1e-6 A^2 of
// the threshold is re-decided per pose in FP64).  The residue-pair string of
// a bit is reported by gdock_contact_pair.
//
// Overlap: bitsets are compacted to the 32-bit words that are set in any
// pose and stored word-major, so a wave reads 16 row words with one scalar
// load and its 64 column words with one coalesced vector load; each (pair,
// word) costs v_and_b32 + v_bcnt_u32_b32.  The neighbour decisions
// fcc >= cutoff and fcc_v >= cutoff * strictness are exact integer compares
// against per-set-size minima that the host derives with the reference's own
// FP64 division, and are written as two bit matrices: A (row i = nbr(i)) and
// its transpose.  Greedy clustering runs in one workgroup with incremental
// unused-neighbour counts.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <map>
#include <set>
#include <tuple>
#include <vector>

#include "../../include/gdock_hip.h"
#include "common.h"
#include "internal.h"

namespace gdock {

size_t contacts_lds_bytes(const DevContacts& c);

// ---------------------------------------------------------------- tables
// Residue keys (clustering.rs:66-76): (chain, resseq) over the heavy atoms of
// receptor and ligand; a key's receptor atoms form its receptor group, its
// ligand atoms its ligand group.  Receptor x ligand group pairs of different
// chains get the bits [0, n_dyn), in row-major order of (receptor group,
// ligand group); a key pair reached from both sides (both keys hold receptor
// and ligand atoms) shares one bit.  With one receptor chain against one
// other ligand chain the map is the identity and is not stored.
void build_contact_tables(gdock_complex* cx, Layout* L, std::vector<uint8_t>* host) {
  cx->con = DevContacts{};
  cx->con_offsets.clear();
  cx->res_key.clear();
  cx->rec_group_key.clear();
  cx->lig_group_key.clear();
  cx->pair_bit.clear();
  cx->dyn_keys.clear();
  cx->rec_heavy_xyz.clear();
  cx->lig_heavy_xyz.clear();
  cx->rec_heavy_key.clear();
  cx->lig_heavy_key.clear();
  cx->lig_heavy_slot.clear();
  cx->con_unsupported.clear();
  cx->cut = ContactCut{};
  auto heavy = [](const gdock_atom& a) { return field(a.element, sizeof a.element) != "H"; };
  std::map<std::pair<char, int16_t>, int32_t> key_of;
  for (auto& a : cx->rec) if (heavy(a)) key_of[{a.chainid, a.resseq}] = 0;
  for (auto& a : cx->lig) if (heavy(a)) key_of[{a.chainid, a.resseq}] = 0;
  for (auto& kv : key_of) { kv.second = static_cast<int32_t>(cx->res_key.size()); cx->res_key.push_back(kv.first); }
  const int32_t nkeys = static_cast<int32_t>(cx->res_key.size());
  std::vector<std::vector<int32_t>> rres(nkeys), lres(nkeys);
  std::set<char> rch, lch;
  for (int32_t i = 0; i < static_cast<int32_t>(cx->rec.size()); ++i)
    if (heavy(cx->rec[i])) { rres[key_of[{cx->rec[i].chainid, cx->rec[i].resseq}]].push_back(i); rch.insert(cx->rec[i].chainid); }
  for (int32_t i = 0; i < static_cast<int32_t>(cx->lig.size()); ++i)
    if (heavy(cx->lig[i])) { lres[key_of[{cx->lig[i].chainid, cx->lig[i].resseq}]].push_back(i); lch.insert(cx->lig[i].chainid); }
  cx->rec_chains = static_cast<int32_t>(rch.size());
  cx->lig_chains = static_cast<int32_t>(lch.size());
  cx->rg_of_key.assign(nkeys, -1);
  cx->lg_of_key.assign(nkeys, -1);
  for (int32_t k = 0; k < nkeys; ++k) {
    if (!rres[k].empty()) { cx->rg_of_key[k] = static_cast<int32_t>(cx->rec_group_key.size()); cx->rec_group_key.push_back(k); }
    if (!lres[k].empty()) { cx->lg_of_key[k] = static_cast<int32_t>(cx->lig_group_key.size()); cx->lig_group_key.push_back(k); }
  }
  const int32_t n_rg = static_cast<int32_t>(cx->rec_group_key.size()), n_lg = static_cast<int32_t>(cx->lig_group_key.size());
  // bits of the receptor x ligand group pairs
  std::vector<int32_t> pb(static_cast<size_t>(n_rg) * n_lg, -1);
  bool identity = true;
  for (int32_t rg = 0; rg < n_rg; ++rg)
    for (int32_t lg = 0; lg < n_lg; ++lg) {
      const int32_t k1 = cx->rec_group_key[rg], k2 = cx->lig_group_key[lg];
      const size_t at = static_cast<size_t>(rg) * n_lg + lg;
      if (cx->res_key[k1].first == cx->res_key[k2].first) { identity = false; continue; }  // same chain (or key)
      int32_t bit = -1;
      if (cx->rg_of_key[k2] >= 0 && cx->lg_of_key[k1] >= 0)  // the mirror pair, if already numbered
        bit = pb[static_cast<size_t>(cx->rg_of_key[k2]) * n_lg + cx->lg_of_key[k1]];
      if (bit < 0) {
        bit = static_cast<int32_t>(cx->dyn_keys.size());
        cx->dyn_keys.push_back({std::min(k1, k2), std::max(k1, k2)});
      }
      pb[at] = bit;
      identity = identity && bit == static_cast<int32_t>(at);
    }
  if (!identity) cx->pair_bit = pb;
  // FP64 heavy atoms of both sides for the static key pairs (contact_cut)
  for (int32_t k = 0; k < nkeys; ++k)
    for (int32_t i : rres[k]) {
      cx->rec_heavy_xyz.insert(cx->rec_heavy_xyz.end(), {cx->rec[i].x, cx->rec[i].y, cx->rec[i].z});
      cx->rec_heavy_key.push_back(k);
    }
  for (int32_t k = 0; k < nkeys; ++k)
    for (int32_t i : lres[k]) {
      cx->lig_heavy_xyz.insert(cx->lig_heavy_xyz.end(), {cx->lig[i].x, cx->lig[i].y, cx->lig[i].z});
      cx->lig_heavy_key.push_back(k);
      cx->lig_heavy_slot.push_back(cx->lig_slot[i]);
    }

  std::vector<float4> rxyz;
  std::vector<double> rx, ry, rz;
  std::vector<int32_t> rstart{0}, latom, lresi;
  double origin[3] = {0, 0, 0};
  if (!cx->rec.empty()) {  // same origin as the scoring tables (receptor COM)
    double sx = 0, sy = 0, sz = 0;
    for (auto& a : cx->rec) sx += a.x;
    for (auto& a : cx->rec) sy += a.y;
    for (auto& a : cx->rec) sz += a.z;
    const double n = static_cast<double>(cx->rec.size());
    origin[0] = sx / n; origin[1] = sy / n; origin[2] = sz / n;
  }
  std::vector<float4> rsph;
  for (int32_t k : cx->rec_group_key) {
    const size_t first = rxyz.size();
    for (int32_t i : rres[k]) {
      const gdock_atom& a = cx->rec[i];
      rxyz.push_back(make_float4(static_cast<float>(a.x - origin[0]), static_cast<float>(a.y - origin[1]),
                                 static_cast<float>(a.z - origin[2]), 0.0f));
      rx.push_back(a.x); ry.push_back(a.y); rz.push_back(a.z);
    }
    rstart.push_back(static_cast<int32_t>(rxyz.size()));
    double c[3] = {0, 0, 0};
    for (size_t j = first; j < rxyz.size(); ++j) { c[0] += rxyz[j].x; c[1] += rxyz[j].y; c[2] += rxyz[j].z; }
    const double nres = static_cast<double>(rxyz.size() - first);
    const float cf[3] = {static_cast<float>(c[0] / nres), static_cast<float>(c[1] / nres), static_cast<float>(c[2] / nres)};
    double r = 0.0;
    for (size_t j = first; j < rxyz.size(); ++j) {
      const double dx = static_cast<double>(rxyz[j].x) - cf[0], dy = static_cast<double>(rxyz[j].y) - cf[1],
                   dz = static_cast<double>(rxyz[j].z) - cf[2];
      r = std::max(r, std::sqrt(dx * dx + dy * dy + dz * dz));
    }
    rsph.push_back(make_float4(cf[0], cf[1], cf[2], std::nextafter(static_cast<float>(r * (1.0 + 1e-6) + 1e-4), 1e30f)));
  }
  // the kernel's form of each receptor atom: -2 r and |r|^2 (FP64 sum of the FP32 squares, rounded once)
  for (float4& v : rxyz) {
    const double x = v.x, y = v.y, z = v.z;
    v = make_float4(-2.0f * v.x, -2.0f * v.y, -2.0f * v.z, static_cast<float>(x * x + y * y + z * z));
  }
  // ligand heavy atoms in blocks of kLigBlock spatially compact atoms (a wave's slots)
  std::vector<int32_t> lheavy, res_of(cx->lig.size(), -1);
  for (int32_t lg = 0; lg < n_lg; ++lg)
    for (int32_t i : lres[cx->lig_group_key[lg]]) { lheavy.push_back(i); res_of[i] = lg; }
  bisect_order(cx->lig.data(), lheavy, 0, lheavy.size(), kLigBlock);
  for (int32_t i : lheavy) { latom.push_back(cx->lig_slot[i]); lresi.push_back(res_of[i]); }
  DevContacts& c = cx->con;
  c.n_rec_heavy = static_cast<int32_t>(rxyz.size());
  c.n_lig_heavy = static_cast<int32_t>(latom.size());
  c.n_rec_res = n_rg;
  c.n_lig_res = n_lg;
  c.n_dyn = static_cast<int32_t>(cx->dyn_keys.size());
  c.supported = 1;
  const size_t begin = cx->contacts_layout_begin;
  const size_t o1 = L->add<float4>(rxyz.size()), o2 = L->add<double>(rx.size()), o3 = L->add<double>(ry.size()),
               o4 = L->add<double>(rz.size()), o5 = L->add<int32_t>(rstart.size()),
               o6 = L->add<int32_t>(latom.size()), o7 = L->add<int32_t>(lresi.size()),
               o8 = L->add<float4>(rsph.size()), o9 = L->add<int32_t>(cx->pair_bit.size());
  host->assign(L->size - begin, 0);
  auto put = [&](size_t off, const void* src, size_t bytes) {
    if (bytes) std::memcpy(host->data() + (off - begin), src, bytes);
  };
  put(o1, rxyz.data(), rxyz.size() * sizeof(float4));
  put(o2, rx.data(), rx.size() * 8); put(o3, ry.data(), ry.size() * 8); put(o4, rz.data(), rz.size() * 8);
  put(o5, rstart.data(), rstart.size() * 4); put(o6, latom.data(), latom.size() * 4); put(o7, lresi.data(), lresi.size() * 4);
  put(o8, rsph.data(), rsph.size() * sizeof(float4));
  put(o9, cx->pair_bit.data(), cx->pair_bit.size() * 4);
  cx->con_offsets = {o1, o2, o3, o4, o5, o6, o7, o8, o9};
}

void bind_contact_tables(gdock_complex* cx) {
  if (!cx->con.supported) return;
  char* m = static_cast<char*>(cx->mem);
  const auto& o = cx->con_offsets;
  DevContacts& c = cx->con;
  c.rec_c4 = reinterpret_cast<const float4*>(m + o[0]);
  c.rec_x = reinterpret_cast<const double*>(m + o[1]);
  c.rec_y = reinterpret_cast<const double*>(m + o[2]);
  c.rec_z = reinterpret_cast<const double*>(m + o[3]);
  c.rec_res_start = reinterpret_cast<const int32_t*>(m + o[4]);
  c.lig_atom = reinterpret_cast<const int32_t*>(m + o[5]);
  c.lig_res = reinterpret_cast<const int32_t*>(m + o[6]);
  c.rec_res_sphere = reinterpret_cast<const float4*>(m + o[7]);
  c.pair_bit = cx->pair_bit.empty() ? nullptr : reinterpret_cast<const int32_t*>(m + o[8]);
}

// Pairs (i, j) of points (FP64 x, y, z triples) closer than `reach` in
// each coordinate's cell neighbourhood, i < j, whose keys have different
// chains: f(i, j, d2) with d2 = ((xi - xj)^2 + (yi - yj)^2) + (zi - zj)^2,
// the operation order of clustering.rs:97.  Cells of side `reach` on a
// sorted grid, so every pair within `reach` is visited exactly once.
template <class F>
void near_pairs(const std::vector<double>& xyz, const std::vector<int32_t>& key,
                const std::vector<std::pair<char, int16_t>>& res_key, double reach, F&& f) {
  const size_t n = key.size();
  if (n < 2) return;
  double lo[3] = {INFINITY, INFINITY, INFINITY};
  for (size_t i = 0; i < n; ++i)
    for (int a = 0; a < 3; ++a) lo[a] = std::min(lo[a], xyz[3 * i + a]);
  const double cell = std::max(reach, 0.5) * (1.0 + 1e-9) + 1e-9;  // pairs within `reach` sit in adjacent cells
  struct Cell { int64_t x, y, z; };
  std::vector<Cell> cc(n);
  std::vector<int32_t> order(n);
  for (size_t i = 0; i < n; ++i) {
    cc[i] = {static_cast<int64_t>(std::floor((xyz[3 * i] - lo[0]) / cell)),
             static_cast<int64_t>(std::floor((xyz[3 * i + 1] - lo[1]) / cell)),
             static_cast<int64_t>(std::floor((xyz[3 * i + 2] - lo[2]) / cell))};
    order[i] = static_cast<int32_t>(i);
  }
  auto less = [&](const Cell& a, const Cell& b) {
    return a.x != b.x ? a.x < b.x : (a.y != b.y ? a.y < b.y : a.z < b.z);
  };
  std::sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
    return less(cc[a], cc[b]) || (!less(cc[b], cc[a]) && a < b);
  });
  std::map<std::tuple<int64_t, int64_t, int64_t>, std::pair<size_t, size_t>> span;
  for (size_t s = 0; s < n;) {
    size_t e = s + 1;
    const Cell& c0 = cc[order[s]];
    while (e < n && !less(c0, cc[order[e]])) ++e;
    span[{c0.x, c0.y, c0.z}] = {s, e};
    s = e;
  }
  for (auto& kv : span) {
    const auto [x, y, z] = kv.first;
    for (int64_t dx = -1; dx <= 1; ++dx)
      for (int64_t dy = -1; dy <= 1; ++dy)
        for (int64_t dz = -1; dz <= 1; ++dz) {
          const std::tuple<int64_t, int64_t, int64_t> nb{x + dx, y + dy, z + dz};
          if (nb < kv.first) continue;  // each unordered cell pair once
          auto it = span.find(nb);
          if (it == span.end()) continue;
          for (size_t p = kv.second.first; p < kv.second.second; ++p)
            for (size_t q = (nb == kv.first ? p + 1 : it->second.first); q < it->second.second; ++q) {
              const int32_t i = std::min(order[p], order[q]), j = std::max(order[p], order[q]);
              if (res_key[key[i]].first == res_key[key[j]].first) continue;
              const double ex = xyz[3 * i] - xyz[3 * j], ey = xyz[3 * i + 1] - xyz[3 * j + 1], ez = xyz[3 * i + 2] - xyz[3 * j + 2];
              f(i, j, (ex * ex + ey * ey) + ez * ez);
            }
        }
  }
}

// Window of the ligand x ligand decisions: an atom pair whose unmoved FP64
// squared distance is within kLigPairWindow of the threshold is re-decided
// per pose; the others are decided once, which is exact while the rounding
// of the pose transform moves squared distances by less than the window.
constexpr double kLigPairWindow = 1e-6;

// Bound on |d2(posed, FP64) - d2(unmoved)| for a ligand atom pair at distance
// at most d, when every posed coordinate is within e of the transform of the
// unmoved atom under the (FP-valued, nearly orthogonal) rotation matrix.
double lig_pair_error(double d, double e) {
  const double u = 1.1102230246251565e-16;
  const double dl = 2.0 * e + u * (d + 2.0 * e);  // error of one coordinate difference
  return 2.0 * 1.7320508075688772 * d * dl + 3.0 * dl * dl + 4.0 * u * (d + 1.7320508075688772 * dl) * (d + 1.7320508075688772 * dl) +
         1e-14 * d * d;
}

// The layout of the contact bits at squared-distance threshold d2 (cached on
// the complex; callers hold cx->con_mu).
const ContactCut& contact_cut(gdock_complex* cx, double d2) {
  ContactCut& cut = cx->cut;
  if (cut.valid && cut.d2 == d2) return cut;
  cut = ContactCut{};
  cut.d2 = d2;
  cut.valid = true;
  const int32_t n_dyn = cx->con.n_dyn, nkeys = static_cast<int32_t>(cx->res_key.size());
  auto dyn_bit = [&](int32_t k1, int32_t k2) -> int32_t {  // bit of key pair {k1, k2} among [0, n_dyn), or -1
    const int32_t n_lg = cx->con.n_lig_res;
    for (int side = 0; side < 2; ++side) {
      const int32_t a = side ? k2 : k1, b = side ? k1 : k2;
      const int32_t rg = cx->rg_of_key[a], lg = cx->lg_of_key[b];
      if (rg < 0 || lg < 0) continue;
      const size_t at = static_cast<size_t>(rg) * n_lg + lg;
      return cx->pair_bit.empty() ? static_cast<int32_t>(at) : cx->pair_bit[at];
    }
    return -1;
  };
  std::set<std::pair<int32_t, int32_t>> fixed_in;            // key pairs in contact in every pose
  std::map<std::pair<int32_t, int32_t>, std::vector<std::pair<int32_t, int32_t>>> win;  // key pair -> ligand atom pairs
  if (cx->rec_chains > 1)  // receptor x receptor: the receptor never moves, decided exactly
    near_pairs(cx->rec_heavy_xyz, cx->rec_heavy_key, cx->res_key, std::sqrt(d2), [&](int32_t i, int32_t j, double e2) {
      if (e2 <= d2) fixed_in.insert({std::min(cx->rec_heavy_key[i], cx->rec_heavy_key[j]), std::max(cx->rec_heavy_key[i], cx->rec_heavy_key[j])});
    });
  if (cx->lig_chains > 1) {
    near_pairs(cx->lig_heavy_xyz, cx->lig_heavy_key, cx->res_key, std::sqrt(d2 + kLigPairWindow),
               [&](int32_t i, int32_t j, double e2) {
                 const std::pair<int32_t, int32_t> kp{std::min(cx->lig_heavy_key[i], cx->lig_heavy_key[j]),
                                                      std::max(cx->lig_heavy_key[i], cx->lig_heavy_key[j])};
                 if (e2 <= d2 - kLigPairWindow) fixed_in.insert(kp);
                 else if (e2 <= d2 + kLigPairWindow) win[kp].push_back({i, j});
               });
    // the largest translation for which every ligand pair decision above holds
    double A = 0.0, C = 0.0;
    for (int a = 0; a < 3; ++a) C = std::max(C, std::fabs(cx->dev.com[a]));
    for (size_t i = 0; i < cx->lig_heavy_key.size(); ++i)
      for (int a = 0; a < 3; ++a) A = std::max(A, std::fabs(cx->lig_heavy_xyz[3 * i + a] - cx->dev.com[a]));
    const double u = 1.1102230246251565e-16, dmax = std::sqrt(d2 + kLigPairWindow);
    double D = 0.0;
    if (lig_pair_error(dmax, 32.0 * u * (A + C)) < kLigPairWindow) {
      double hi = 1e12;
      for (int it = 0; it < 200; ++it) {
        const double mid = 0.5 * (D + hi);
        (lig_pair_error(dmax, 32.0 * u * (A + C + mid)) < kLigPairWindow ? D : hi) = mid;
      }
    } else {
      D = -1.0;  // no translation is safe: the ligand itself is too large for the window
    }
    cut.max_translation = D;
  } else {
    cut.max_translation = INFINITY;
  }
  // bits beyond the receptor x ligand pairs, in key-pair order
  std::set<std::pair<int32_t, int32_t>> extra;
  for (auto& kp : fixed_in) if (dyn_bit(kp.first, kp.second) < 0) extra.insert(kp);
  for (auto& kv : win) if (!fixed_in.count(kv.first) && dyn_bit(kv.first.first, kv.first.second) < 0) extra.insert(kv.first);
  cut.extra_keys.assign(extra.begin(), extra.end());
  std::map<std::pair<int32_t, int32_t>, int32_t> extra_bit;
  for (size_t i = 0; i < cut.extra_keys.size(); ++i) extra_bit[cut.extra_keys[i]] = n_dyn + static_cast<int32_t>(i);
  auto bit_of = [&](const std::pair<int32_t, int32_t>& kp) {
    const int32_t b = dyn_bit(kp.first, kp.second);
    return b >= 0 ? b : extra_bit.at(kp);
  };
  cut.n_bits = n_dyn + static_cast<int32_t>(cut.extra_keys.size());
  cut.words64 = (cut.n_bits + 63) / 64;
  cut.static_mask.assign(cut.words64, 0ull);
  for (auto& kp : fixed_in) { const int32_t b = bit_of(kp); cut.static_mask[b >> 6] |= 1ull << (b & 63); }
  for (auto& kv : win) {
    if (fixed_in.count(kv.first)) continue;
    const int32_t b = bit_of(kv.first);
    for (auto& ij : kv.second)
      cut.window.push_back(make_int4(cx->lig_heavy_slot[ij.first], cx->lig_heavy_slot[ij.second], b, 0));
  }
  (void)nkeys;
  DevContacts probe = cx->con;
  probe.words64 = cut.words64;
  if (contacts_lds_bytes(probe) > 64 * 1024)
    cut.err = "contacts: residue-pair bitset and residue masks exceed the 64 KiB workgroup budget";
  return cut;
}

void release_fcc_scratch(gdock_ctx* ctx) { ctx->fcc.release(); }

// ---------------------------------------------------------------- device helpers
struct PoseT {
  double ca, sa, cb, sb, m01, m02, m11, m12, cg, sg, dx, dy, dz;
};

__device__ __forceinline__ PoseT pose_of(const double* __restrict__ p) {
  PoseT c;
  c.dx = p[3]; c.dy = p[4]; c.dz = p[5];
  c.ca = p[6]; c.sa = p[7]; c.cb = p[8]; c.sb = p[9]; c.cg = p[10]; c.sg = p[11];
  c.m01 = p[12]; c.m02 = p[13]; c.m11 = p[14]; c.m12 = p[15];
  return c;
}

// structure.rs:80-131 operation order (file built with -ffp-contract=off)
__device__ __forceinline__ void pose_atom(const DevComplex& cx, const PoseT& pc, int l, double& X, double& Y, double& Z) {
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

__device__ __forceinline__ float uniform_f(float v) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}

__device__ __forceinline__ void band(double T, double X, float& lo, float& hi) {
  const double u = 5.9604644775390625e-08;
  const double s = sqrt(T) * 1.01 + 0.01;
  const double e = 2.0 * u * X + u * s;
  const double b = 2.0 * (3.4641016151377544 * s * e * (1.0 + 2.0 * u) + 3.0 * e * e + 3.01 * u * T * 1.01) + 1e-9;
  float l = static_cast<float>(T - b), h = static_cast<float>(T + b);
  if (static_cast<double>(l) > T - b) l = nextafterf(l, -INFINITY);
  if (static_cast<double>(h) < T + b) h = nextafterf(h, INFINITY);
  lo = uniform_f(l);
  hi = uniform_f(h);
}

// Guard band for d2 evaluated as |r|^2 + |l|^2 - 2 r.l in FP32 (one add, three
// FMAs) from coordinates rounded to FP32 in a frame where every component is
// at most X: |r|^2, |l|^2 <= 3 X^2 and every intermediate is at most 12 X^2, so
// the roundings of |r|^2 (3 u X^2), |l|^2 (9 u X^2), the add (6 u X^2) and the
// three FMAs (36 u X^2) stay below 60 u X^2; the coordinate rounding adds at
// most 4 sqrt(3) u X d.  The band is twice that, rounded outward.
__device__ __forceinline__ void band_expanded(double T, double X, float& lo, float& hi) {
  const double u = 5.9604644775390625e-08;
  const double e = 60.0 * u * X * X + 6.9282032302755092 * u * X * (sqrt(T) + 0.01) + 1e-7;
  float l = static_cast<float>(T - 2.0 * e), h = static_cast<float>(T + 2.0 * e);
  if (static_cast<double>(l) > T - 2.0 * e) l = nextafterf(l, -INFINITY);
  if (static_cast<double>(h) < T + 2.0 * e) h = nextafterf(h, INFINITY);
  lo = uniform_f(l);
  hi = uniform_f(h);
}

// ---------------------------------------------------------------- K6 contacts
// Wave w holds K blocks of kLigBlock spatially compact ligand heavy atoms
// (contact_slot).  A receptor residue whose bounding sphere lies farther than
// `reach` (the cutoff plus a margin well above the FP32 error) from the box
// of a block has no contact with its atoms; each residue is tested against
// each block's box.
constexpr int kContactBoxFloats = 8;

// Dynamic LDS of contacts_kernel: the pose's bitset (32-bit words), then per
// wave kMaxAtomsPerLane block boxes and one mask byte per receptor residue.
size_t contacts_lds_bytes(const DevContacts& c) {
  return static_cast<size_t>(c.words64) * 8 + 4 * kContactBoxFloats * kMaxAtomsPerLane * sizeof(float) +
         4 * static_cast<size_t>((c.n_rec_res + 3) & ~3);
}

// Ligand heavy atom in slot k of a lane for the chunk at `base`: the chunk's
// blocks are dealt to the waves in turn (wave w: blocks w, w + 4, ...), so
// every wave holds parts of the whole ligand and the waves' work is even.
__device__ __forceinline__ int contact_slot(int base, int k, int wave, int lane) {
  return base + (k * 4 + wave) * kLigBlock + lane;
}

template <int K>
__global__ __launch_bounds__(256) void contacts_kernel(const DevComplex cx, const DevContacts ct,
                                                       const double* __restrict__ poses, double d2_incl, float reach,
                                                       unsigned long long* __restrict__ bits_out,
                                                       int32_t* __restrict__ sizes) {
  extern __shared__ uint32_t s_bits[];  // words32 bitset words, then the per-wave block boxes and residue masks
  __shared__ int32_t s_cnt[4];
  const int pose = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int words32 = 2 * ct.words64;
  float* my_box = reinterpret_cast<float*>(s_bits + words32) + wave * (kContactBoxFloats * kMaxAtomsPerLane);
  uint8_t* my_mask = reinterpret_cast<uint8_t*>(reinterpret_cast<float*>(s_bits + words32) +
                                                4 * kContactBoxFloats * kMaxAtomsPerLane) +
                     wave * ((ct.n_rec_res + 3) & ~3);
  const uint32_t* fixed_in = reinterpret_cast<const uint32_t*>(ct.static_mask);
  for (int i = tid; i < words32; i += 256) s_bits[i] = fixed_in ? fixed_in[i] : 0u;
  const PoseT pc = pose_of(poses + static_cast<size_t>(kPoseDoubles) * pose);
  const double X = fmax(cx.rec_extent, cx.lig_extent + fmax(fabs(pc.dx), fmax(fabs(pc.dy), fabs(pc.dz))));
  float lo, hi;
  band_expanded(d2_incl, X, lo, hi);
  __syncthreads();
  // Ligand x ligand atom pairs near the cutoff: the FP64 distance of the
  // posed atoms, as clustering.rs:97-99 computes it.
  for (int p = tid; p < ct.n_window; p += 256) {
    const int4 wp = ct.window[p];
    double ax, ay, az, bx, by, bz;
    pose_atom(cx, pc, wp.x, ax, ay, az);
    pose_atom(cx, pc, wp.y, bx, by, bz);
    const double ex = ax - bx, ey = ay - by, ez = az - bz;
    if ((ex * ex + ey * ey) + ez * ez <= d2_incl) atomicOr(&s_bits[wp.z >> 5], 1u << (wp.z & 31));
  }

  for (int base = 0; ct.n_dyn > 0 && base < ct.n_lig_heavy; base += 256 * K) {
    float lx[K], ly[K], lz[K], ll[K];
    int32_t lres[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int h = contact_slot(base, k, wave, lane);
      if (h < ct.n_lig_heavy) {
        double px, py, pz;
        pose_atom(cx, pc, ct.lig_atom[h], px, py, pz);
        lx[k] = static_cast<float>(px - cx.origin[0]);
        ly[k] = static_cast<float>(py - cx.origin[1]);
        lz[k] = static_cast<float>(pz - cx.origin[2]);
        lres[k] = ct.lig_res[h];
      } else {
        lx[k] = ly[k] = lz[k] = kPadCoord;
        lres[k] = 0;
      }
      ll[k] = __builtin_fmaf(lz[k], lz[k], __builtin_fmaf(ly[k], ly[k], lx[k] * lx[k]));
    }
    // Box of each of the wave's ligand blocks (empty: +inf / -inf, never in reach).
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const bool in = contact_slot(base, k, wave, lane) < ct.n_lig_heavy;
      float x0 = in ? lx[k] : INFINITY, y0 = in ? ly[k] : INFINITY, z0 = in ? lz[k] : INFINITY;
      float x1 = in ? lx[k] : -INFINITY, y1 = in ? ly[k] : -INFINITY, z1 = in ? lz[k] : -INFINITY;
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) {
        x0 = fminf(x0, __shfl_xor(x0, off)); y0 = fminf(y0, __shfl_xor(y0, off)); z0 = fminf(z0, __shfl_xor(z0, off));
        x1 = fmaxf(x1, __shfl_xor(x1, off)); y1 = fmaxf(y1, __shfl_xor(y1, off)); z1 = fmaxf(z1, __shfl_xor(z1, off));
      }
      if (lane == 0) {
        float* bx = my_box + kContactBoxFloats * k;
        bx[0] = x0; bx[1] = y0; bx[2] = z0; bx[3] = x1; bx[4] = y1; bx[5] = z1;
      }
    }
    // Bit k of a residue's mask: its bounding sphere reaches block k.  A
    // residue beyond reach of a block has no atom pair within the cutoff or
    // its guard band with it, so skipping the pair is exact.
    for (int r0 = 0; r0 < ct.n_rec_res; r0 += 64) {
      const int ri = r0 + lane;
      if (ri >= ct.n_rec_res) break;
      const float4 S = ct.rec_res_sphere[ri];
      const float rr = S.w + reach;
      uint32_t m = 0;
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const float* bx = my_box + kContactBoxFloats * k;
        const float ex = fmaxf(fmaxf(bx[0] - S.x, S.x - bx[3]), 0.0f), ey = fmaxf(fmaxf(bx[1] - S.y, S.y - bx[4]), 0.0f),
                    ez = fmaxf(fmaxf(bx[2] - S.z, S.z - bx[5]), 0.0f);
        m |= (ex * ex + ey * ey + ez * ez <= rr * rr ? 1u : 0u) << k;
      }
      my_mask[ri] = static_cast<uint8_t>(m);
    }
    for (int ri = 0; ri < ct.n_rec_res; ++ri) {
      const uint32_t mk = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(my_mask[ri]));
      if (mk == 0u) continue;
      // closest receptor atom of the residue to each of the lane's atoms
      float mn[K];
#pragma unroll
      for (int k = 0; k < K; ++k) mn[k] = INFINITY;
      const int j0 = ct.rec_res_start[ri], j1 = ct.rec_res_start[ri + 1];
      auto expanded = [&](const float4& r, int k) {
        return __builtin_fmaf(r.x, lx[k], __builtin_fmaf(r.y, ly[k], __builtin_fmaf(r.z, lz[k], r.w + ll[k])));
      };
      int j = j0;
      // four atoms per step: their loads issue together and each block's
      // mask is tested once (the minimum of exact values in any order)
      for (; j + 4 <= j1; j += 4) {
        const float4 r0 = ct.rec_c4[j], r1 = ct.rec_c4[j + 1], r2 = ct.rec_c4[j + 2], r3 = ct.rec_c4[j + 3];
#pragma unroll
        for (int k = 0; k < K; ++k)
          if ((mk >> k) & 1u)
            mn[k] = fminf(fminf(mn[k], fminf(expanded(r0, k), expanded(r1, k))), fminf(expanded(r2, k), expanded(r3, k)));
      }
      for (; j < j1; ++j) {
        const float4 r = ct.rec_c4[j];
#pragma unroll
        for (int k = 0; k < K; ++k)
          if ((mk >> k) & 1u) mn[k] = fminf(mn[k], expanded(r, k));
      }
      bool hit[K], any_amb = false;
#pragma unroll
      for (int k = 0; k < K; ++k) {
        hit[k] = mn[k] <= lo;
        any_amb = any_amb || (!hit[k] && mn[k] < hi);
      }
      if (__any(any_amb)) {  // rare: decide this residue in FP64 (clustering.rs:98-99)
#pragma unroll
        for (int k = 0; k < K; ++k) {
          const bool amb = !hit[k] && mn[k] < hi;
          const int h = contact_slot(base, k, wave, lane);
          if (amb && h < ct.n_lig_heavy) {
            double px, py, pz;
            pose_atom(cx, pc, ct.lig_atom[h], px, py, pz);
            for (int j = j0; j < j1 && !hit[k]; ++j) {
              const double ex = ct.rec_x[j] - px, ey = ct.rec_y[j] - py, ez = ct.rec_z[j] - pz;
              hit[k] = ex * ex + ey * ey + ez * ez <= d2_incl;
            }
          }
        }
      }
      bool any_hit = false;
#pragma unroll
      for (int k = 0; k < K; ++k) any_hit = any_hit || hit[k];
      if (__any(any_hit)) {
#pragma unroll
        for (int k = 0; k < K; ++k)
          if (hit[k]) {
            const int32_t at = ri * ct.n_lig_res + lres[k];
            const int32_t bit = ct.pair_bit ? ct.pair_bit[at] : at;
            if (bit >= 0) atomicOr(&s_bits[bit >> 5], 1u << (bit & 31));
          }
      }
    }
  }
  __syncthreads();
  int cnt = 0;
  unsigned long long* dst = bits_out + static_cast<size_t>(pose) * ct.words64;
  for (int w = tid; w < ct.words64; w += 256) {
    const unsigned long long v = static_cast<unsigned long long>(s_bits[2 * w]) |
                                 (static_cast<unsigned long long>(s_bits[2 * w + 1]) << 32);
    dst[w] = v;
    cnt += __popcll(v);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off);
  if (lane == 0) s_cnt[wave] = cnt;
  __syncthreads();
  if (tid == 0) sizes[pose] = s_cnt[0] + s_cnt[1] + s_cnt[2] + s_cnt[3];
}

// ---------------------------------------------------------------- K7 overlap
// OR of all poses' words (32-bit view) into uni[w] (zeroed beforehand); each
// workgroup row folds a slice of kUnionPoses poses, coalesced across words.
constexpr int kUnionPoses = 64;
__global__ void union_kernel(const uint32_t* __restrict__ bits, int n, int words32, uint32_t* __restrict__ uni) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= words32) return;
  const int p0 = blockIdx.y * kUnionPoses, p1 = min(n, p0 + kUnionPoses);
  uint32_t u = 0;
  for (int p = p0; p < p1; ++p) u |= bits[static_cast<size_t>(p) * words32 + w];
  if (u) atomicOr(&uni[w], u);
}

// Word-major compacted copy: T[k * n_pad + p] = bits[p][col[k]].
__global__ void compact_kernel(const uint32_t* __restrict__ bits, int n, int n_pad, int words32,
                               const int32_t* __restrict__ col, int wc, uint32_t* __restrict__ T) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = blockIdx.y;
  if (p >= n_pad || k >= wc) return;
  T[static_cast<size_t>(k) * n_pad + p] = p < n ? bits[static_cast<size_t>(p) * words32 + col[k]] : 0u;
}

constexpr int kRowsPerWave = 16;

// Tile: 64 rows (4 waves x 16 rows) x 64 columns (one per lane).  The common
// count c is symmetric, so only tiles with jblk >= iblk are computed.  A tile
// writes the words A[i][jblk] (bits j with j in nbr(i)) and AT[i][jblk] (bits j
// with i in nbr(j)) of its rows and, off the diagonal, the mirrored words
// A[j][iblk] (bits i with i in nbr(j)) and AT[j][iblk] (bits i with j in
// nbr(i)), each lane collecting its column's bits over the tile's rows.
__global__ __launch_bounds__(256) void overlap_kernel(const uint32_t* __restrict__ T, int n, int n_pad, int wc,
                                                      const int32_t* __restrict__ sizes,
                                                      const int32_t* __restrict__ cmin_a,
                                                      const int32_t* __restrict__ cmin_b, int zero_ok,
                                                      unsigned long long* __restrict__ A,
                                                      unsigned long long* __restrict__ AT, int nwords) {
  __shared__ unsigned long long s_ta[4][64], s_tat[4][64];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int jblk = blockIdx.x, iblk = blockIdx.y;
  if (jblk < iblk) return;  // the mirror of tile (jblk, iblk)
  const int j = jblk * 64 + lane;
  const int i0 = iblk * 64 + wave * kRowsPerWave;
  unsigned long long ta = 0ull, tat = 0ull;  // mirrored words of column j over this wave's rows
  if (i0 < n) {
    uint32_t c[kRowsPerWave];
#pragma unroll
    for (int r = 0; r < kRowsPerWave; ++r) c[r] = 0u;
    const uint32_t* colp = T + j;
    const uint32_t* rowp = T + i0;
    for (int k = 0; k < wc; ++k) {
      const uint32_t xj = colp[static_cast<size_t>(k) * n_pad];
      const uint32_t* rk = rowp + static_cast<size_t>(k) * n_pad;
#pragma unroll
      for (int r = 0; r < kRowsPerWave; ++r) c[r] += __popc(rk[r] & xj);
    }
    const int sj = j < n ? sizes[j] : 0;
    const int cbj = j < n && sj > 0 ? cmin_b[sj] : INT_MAX;
    const int caj = j < n && sj > 0 ? cmin_a[sj] : INT_MAX;
#pragma unroll
    for (int r = 0; r < kRowsPerWave; ++r) {
      const int i = i0 + r;
      if (i >= n) break;
      const int si = sizes[i];
      bool fwd, bwd;  // j in nbr(i); i in nbr(j)   (clustering.rs:169-177)
      if (si == 0 || sj == 0) {
        fwd = bwd = zero_ok && j < n;
      } else {
        const int c_ = static_cast<int>(c[r]);
        fwd = c_ >= cmin_a[si] && c_ >= cbj;
        bwd = c_ >= caj && c_ >= cmin_b[si];
      }
      fwd = fwd && j < n;
      bwd = bwd && j < n;
      const unsigned long long wf = __ballot(fwd);
      const unsigned long long wb = __ballot(bwd);
      if (lane == 0) {
        A[static_cast<size_t>(i) * nwords + jblk] = wf;
        AT[static_cast<size_t>(i) * nwords + jblk] = wb;
      }
      const unsigned long long bit = 1ull << (wave * kRowsPerWave + r);
      ta |= bwd ? bit : 0ull;
      tat |= fwd ? bit : 0ull;
    }
  }
  if (jblk == iblk) return;  // the diagonal tile is its own mirror
  s_ta[wave][lane] = ta;
  s_tat[wave][lane] = tat;
  __syncthreads();
  if (wave == 0 && j < n) {
    A[static_cast<size_t>(j) * nwords + iblk] = s_ta[0][lane] | s_ta[1][lane] | s_ta[2][lane] | s_ta[3][lane];
    AT[static_cast<size_t>(j) * nwords + iblk] = s_tat[0][lane] | s_tat[1][lane] | s_tat[2][lane] | s_tat[3][lane];
  }
}

// ---------------------------------------------------------------- K8 greedy
// clustering.rs:183-254 in one workgroup.  count[a] = unused neighbours of a
// (self included).  Each step of the reference takes the unused element with
// the largest count (lowest index on ties), forms {centre} + its unused
// neighbours (in ascending order), marks them used and decrements the counts
// of every element that has a new member as neighbour (rows of AT).
//
// Counts only decrease, so the steps are taken in rounds: a round finds the
// largest count m among unused elements and lists the unused elements of
// count m in ascending order; walking that list, each element still unused
// and still of count m is the reference's next centre (any unused element of
// count m now was of count m at the start of the round, and every lower one
// of the list was taken or fell below m), the others are passed over for a
// later round.  This gives the reference's sequence of clusters with one
// block-wide reduction per round instead of one per cluster.  Cluster sizes
// are non-increasing along the sequence, so the loop stops as soon as no
// element can reach min_size (count + 1 < min_size).
// Neighbour count of every row of A (self included), one wave per row.
__global__ __launch_bounds__(256) void row_count_kernel(const unsigned long long* __restrict__ A, int n, int nwords,
                                                        int32_t* __restrict__ counts) {
  const int row = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (row >= n) return;
  int c = 0;
  for (int w = lane; w < nwords; w += 64) c += __popcll(A[static_cast<size_t>(row) * nwords + w]);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off);
  if (lane == 0) counts[row] = c;
}

constexpr int kGreedyThreads = 1024;
constexpr int kGreedyMaxN = 15000;  // LDS: 4 B count, 2 B member slot, 2 B candidate slot, 1 bit used flag per element

__global__ __launch_bounds__(kGreedyThreads) void greedy_kernel(const unsigned long long* __restrict__ A,
                                                                const unsigned long long* __restrict__ AT,
                                                                const int32_t* __restrict__ counts, int n, int nwords,
                                                                long long min_size,
                                                                int64_t* __restrict__ out_center,
                                                                int64_t* __restrict__ out_size,
                                                                int64_t* __restrict__ out_offset,
                                                                int64_t* __restrict__ out_members,
                                                                int32_t* __restrict__ out_n) {
  // LDS: used bits (nwords), count[n], then the current members[n] and the
  // round's candidates[n] as 16-bit indices
  extern __shared__ unsigned long long s_used[];
  int32_t* s_count = reinterpret_cast<int32_t*>(s_used + nwords);
  uint16_t* s_members = reinterpret_cast<uint16_t*>(s_count + n);
  uint16_t* s_cand = s_members + n;
  __shared__ int32_t s_red[kGreedyThreads / 64];
  __shared__ int32_t s_scan[kGreedyThreads / 64];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int a = tid; a < n; a += kGreedyThreads) s_count[a] = counts[a];
  for (int w = tid; w < nwords; w += kGreedyThreads) s_used[w] = 0ull;
  __syncthreads();
  auto used = [&](int a) { return ((s_used[a >> 6] >> (a & 63)) & 1ull) != 0ull; };
  // block-wide exclusive prefix of v (all threads); *total gets the sum
  auto prefix = [&](int v, int* total) {
    int incl = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const int o = __shfl_up(incl, off);
      if (lane >= off) incl += o;
    }
    if (lane == 63) s_scan[wave] = incl;
    __syncthreads();
    int base = 0, sum = 0;
    for (int i = 0; i < kGreedyThreads / 64; ++i) {
      const int x = s_scan[i];
      base += i < wave ? x : 0;
      sum += x;
    }
    *total = sum;
    return base + incl - v;
  };
  const int per = (n + kGreedyThreads - 1) / kGreedyThreads;  // thread t owns [t per, t per + per)
  const int lo = min(tid * per, n), hi = min(lo + per, n);
  int32_t nclu = 0;
  int64_t written = 0;
  for (;;) {
    // m = the largest count among unused elements (-1: every element used)
    int mx = -1;
    for (int a = tid; a < n; a += kGreedyThreads)
      if (!used(a)) mx = max(mx, s_count[a]);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) mx = max(mx, __shfl_xor(mx, off));
    if (lane == 0) s_red[wave] = mx;
    __syncthreads();
    int m = -1;
    for (int i = 0; i < kGreedyThreads / 64; ++i) m = max(m, s_red[i]);
    if (m < 0) break;  // every element used
    if (min_size > 1 && static_cast<long long>(m) + 1 < min_size) break;
    // the round's candidates: unused elements of count m, ascending
    int mine_n = 0;
    for (int a = lo; a < hi; ++a) mine_n += (!used(a) && s_count[a] == m) ? 1 : 0;
    int ncand = 0;
    int pos = prefix(mine_n, &ncand);
    for (int a = lo; a < hi; ++a)
      if (!used(a) && s_count[a] == m) s_cand[pos++] = static_cast<uint16_t>(a);
    __syncthreads();
    for (int t = 0; t < ncand; ++t) {
      const int center = s_cand[t];
      if (used(center) || s_count[center] != m) continue;  // taken, or below m now: a later round
      // members = centre + its unused neighbours, ascending (one word per thread)
      unsigned long long mine = 0ull;
      if (tid < nwords) {
        mine = A[static_cast<size_t>(center) * nwords + tid] & ~s_used[tid];
        if (tid == (center >> 6)) mine |= 1ull << (center & 63);
      }
      const int local = __popcll(mine);
      int size32 = 0;
      int mpos = prefix(local, &size32);
      for (unsigned long long mm = mine; mm; mm &= mm - 1) {
        const int e = tid * 64 + (__ffsll(static_cast<long long>(mm)) - 1);
        s_members[mpos] = static_cast<uint16_t>(e);
        out_members[written + mpos++] = e;
      }
      if (tid < nwords) s_used[tid] |= mine;
      __syncthreads();
      const int64_t size = size32;
      // decrement the counts of elements that have a new member as neighbour
      for (int64_t q = tid; q < size * nwords; q += kGreedyThreads) {
        const int64_t u = s_members[q / nwords];
        const int w = static_cast<int>(q % nwords);
        for (unsigned long long mm = AT[static_cast<size_t>(u) * nwords + w]; mm; mm &= mm - 1)
          atomicSub(&s_count[w * 64 + (__ffsll(static_cast<long long>(mm)) - 1)], 1);
      }
      if (tid == 0) {
        out_center[nclu] = center;
        out_size[nclu] = size;
        out_offset[nclu] = written;
      }
      ++nclu;
      written += size;
      __syncthreads();
    }
  }
  if (tid == 0) *out_n = nclu;
}

template <int K>
hipError_t launch_contacts_k(const gdock_complex* cx, const DevContacts& ct, const double* d_poses, int n,
                             double d2_incl, unsigned long long* d_bits, int32_t* d_sizes, hipStream_t s) {
  const float reach = static_cast<float>(std::sqrt(std::max(d2_incl, 0.0))) + 0.01f;
  hipLaunchKernelGGL(contacts_kernel<K>, dim3(n), dim3(256), contacts_lds_bytes(ct), s, cx->dev, ct, d_poses, d2_incl,
                     reach, d_bits, d_sizes);
  return hipGetLastError();
}

// Contact bitsets of n poses under the layout `cut` of threshold d2_incl;
// d_static / d_window hold cut.static_mask / cut.window on the device.
// Without any bit the sets are empty: sizes are cleared, no kernel runs.
hipError_t launch_contacts(const gdock_complex* cx, const ContactCut& cut, const uint64_t* d_static,
                           const int4* d_window, const double* d_poses, int n, double d2_incl,
                           unsigned long long* d_bits, int32_t* d_sizes, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  if (cut.words64 == 0) return hipMemsetAsync(d_sizes, 0, static_cast<size_t>(n) * 4, s);
  DevContacts ct = cx->con;
  ct.words64 = cut.words64;
  ct.static_mask = d_static;
  ct.window = d_window;
  ct.n_window = static_cast<int32_t>(cut.window.size());
  int k = (ct.n_lig_heavy + 255) / 256;
  k = k < 1 ? 1 : (k > 8 ? 8 : k);
  switch (k) {
    case 1: return launch_contacts_k<1>(cx, ct, d_poses, n, d2_incl, d_bits, d_sizes, s);
    case 2: return launch_contacts_k<2>(cx, ct, d_poses, n, d2_incl, d_bits, d_sizes, s);
    case 3: return launch_contacts_k<3>(cx, ct, d_poses, n, d2_incl, d_bits, d_sizes, s);
    case 4: return launch_contacts_k<4>(cx, ct, d_poses, n, d2_incl, d_bits, d_sizes, s);
    case 5: return launch_contacts_k<5>(cx, ct, d_poses, n, d2_incl, d_bits, d_sizes, s);
    case 6: return launch_contacts_k<6>(cx, ct, d_poses, n, d2_incl, d_bits, d_sizes, s);
    case 7: return launch_contacts_k<7>(cx, ct, d_poses, n, d2_incl, d_bits, d_sizes, s);
    default: return launch_contacts_k<8>(cx, ct, d_poses, n, d2_incl, d_bits, d_sizes, s);
  }
}

// Largest d2 with sqrt(d2) <= c (correctly rounded sqrt), i.e. the exact
// FP64 form of clustering.rs:99 `dist <= contact_distance`.
double contact_d2_threshold(double c) {
  double t = c * c;
  while (std::sqrt(t) > c) t = std::nextafter(t, -INFINITY);
  while (std::sqrt(std::nextafter(t, INFINITY)) <= c) t = std::nextafter(t, INFINITY);
  return t;
}

// Smallest c with (double)c / (double)s >= cut, for s = 1..max (clustering.rs:127-129,171-174).
void fcc_minima(double cut, int max_size, std::vector<int32_t>* out) {
  out->assign(static_cast<size_t>(max_size) + 1, INT_MAX);
  for (int s = 1; s <= max_size; ++s) {
    int c = static_cast<int>(std::floor(cut * s)) - 2;
    if (c < 0) c = 0;
    while (c <= s && !(static_cast<double>(c) / static_cast<double>(s) >= cut)) ++c;
    (*out)[s] = c <= s ? c : INT_MAX;
  }
}

}  // namespace gdock

using namespace gdock;

extern "C" {

// Layout of cutoff `contact_distance` (callers hold cx->con_mu), or null
// with the reason in *why.
static const ContactCut* cut_of(gdock_complex* cx, double contact_distance, const char** why) {
  if (!cx->con.supported) { *why = cx->con_unsupported.c_str(); return nullptr; }
  if (!(contact_distance >= 0.0) || !std::isfinite(contact_distance)) { *why = "contacts: bad cutoff"; return nullptr; }
  const ContactCut& cut = contact_cut(cx, contact_d2_threshold(contact_distance));
  if (!cut.err.empty()) { *why = cut.err.c_str(); return nullptr; }
  return &cut;
}

int32_t gdock_contact_words(gdock_complex* cx, double contact_distance) {
  if (!cx) return -1;
  std::lock_guard<std::mutex> lk(cx->con_mu);
  const char* why = nullptr;
  const ContactCut* cut = cut_of(cx, contact_distance, &why);
  return cut ? cut->words64 : -1;
}

int32_t gdock_contact_candidates(gdock_complex* cx, double contact_distance) {
  if (!cx) return -1;
  std::lock_guard<std::mutex> lk(cx->con_mu);
  const char* why = nullptr;
  const ContactCut* cut = cut_of(cx, contact_distance, &why);
  return cut ? cut->n_bits : -1;
}

gdock_status gdock_contact_pair(gdock_complex* cx, double contact_distance, int32_t bit, char* chain_a, int16_t* res_a,
                                char* chain_b, int16_t* res_b) {
  if (!cx || !chain_a || !res_a || !chain_b || !res_b) return GDOCK_ERR_INVALID_ARG;
  std::lock_guard<std::mutex> lk(cx->con_mu);
  const char* why = nullptr;
  const ContactCut* cut = cut_of(cx, contact_distance, &why);
  if (!cut || bit < 0 || bit >= cut->n_bits) return GDOCK_ERR_INVALID_ARG;
  const int32_t n_dyn = cx->con.n_dyn;
  const auto kp = bit < n_dyn ? cx->dyn_keys[bit] : cut->extra_keys[bit - n_dyn];
  const auto& r = cx->res_key[kp.first];
  const auto& l = cx->res_key[kp.second];
  const bool rf = r.first < l.first;  // smaller chain first (clustering.rs:101-105)
  *chain_a = rf ? r.first : l.first; *res_a = rf ? r.second : l.second;
  *chain_b = rf ? l.first : r.first; *res_b = rf ? l.second : r.second;
  return GDOCK_OK;
}

// Uploads the poses of n gene vectors; *max_shift gets their largest
// translation component.
static gdock_status upload_poses(gdock_ctx* ctx, const double* genes, int32_t n, double* d_poses, double* max_shift) {
  std::vector<double> poses(static_cast<size_t>(n) * kPoseDoubles);
  if (gdock_prepare_poses(genes, n, poses.data()) != GDOCK_OK) {
    set_error(ctx, "non-finite gene");
    return GDOCK_ERR_INVALID_ARG;
  }
  double m = 0.0;
  for (int32_t i = 0; i < n; ++i)
    for (int a = 3; a < 6; ++a) m = std::max(m, std::fabs(poses[static_cast<size_t>(i) * kPoseDoubles + a]));
  *max_shift = m;
  hipError_t e = hipMemcpyAsync(d_poses, poses.data(), poses.size() * 8, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);  // `poses` is freed on return
  return e == hipSuccess ? GDOCK_OK : hip_fail(ctx, e, "upload poses");
}

// Poses, the layout's static words and window pairs on the device.
static gdock_status stage_contacts(gdock_ctx* ctx, const ContactCut& cut, const double* genes, int32_t n, char* m,
                                   size_t o_p, size_t o_st, size_t o_w, const char* what) {
  double shift = 0.0;
  gdock_status st = upload_poses(ctx, genes, n, reinterpret_cast<double*>(m + o_p), &shift);
  if (st != GDOCK_OK) return st;
  if (!(shift <= cut.max_translation)) {
    set_error(ctx, "%s: a pose translation of %g A exceeds the %g A within which the ligand's inter-chain contacts are exact",
              what, shift, cut.max_translation);
    return GDOCK_ERR_INVALID_ARG;
  }
  hipError_t e = hipSuccess;
  if (!cut.static_mask.empty())
    e = hipMemcpyAsync(m + o_st, cut.static_mask.data(), cut.static_mask.size() * 8, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess && !cut.window.empty())
    e = hipMemcpyAsync(m + o_w, cut.window.data(), cut.window.size() * sizeof(int4), hipMemcpyHostToDevice, ctx->stream);
  return e == hipSuccess ? GDOCK_OK : hip_fail(ctx, e, what);
}

gdock_status gdock_contacts(gdock_ctx* ctx, gdock_complex* cx, const double* genes, int32_t n, double contact_distance,
                            uint64_t* bitsets, int32_t* set_sizes) {
  if (!ctx || !cx || n < 0 || (n && !genes)) return GDOCK_ERR_INVALID_ARG;
  std::lock_guard<std::mutex> lk(cx->con_mu);
  const char* why = nullptr;
  const ContactCut* cut = cut_of(cx, contact_distance, &why);
  if (!cut) { set_error(ctx, "%s", why); return GDOCK_ERR_INVALID_ARG; }
  if (n && cut->words64 && !bitsets) return GDOCK_ERR_INVALID_ARG;
  if (n == 0) return GDOCK_OK;
  hipError_t e = hipSetDevice(ctx->device);
  if (e != hipSuccess) return hip_fail(ctx, e, "contacts: set device");
  const int W64 = cut->words64;
  Layout L;
  const size_t o_p = L.add<double>(static_cast<size_t>(n) * kPoseDoubles),
               o_b = L.add<uint64_t>(static_cast<size_t>(n) * W64), o_s = L.add<int32_t>(n),
               o_st = L.add<uint64_t>(W64), o_w = L.add<int4>(cut->window.size());
  if (!ctx->fcc.reserve(L.size, &e)) return hip_fail(ctx, e, "contacts: scratch");
  char* m = static_cast<char*>(ctx->fcc.ptr);
  gdock_status st = stage_contacts(ctx, *cut, genes, n, m, o_p, o_st, o_w, "contacts");
  if (st != GDOCK_OK) return st;
  e = launch_contacts(cx, *cut, cut->static_mask.empty() ? nullptr : reinterpret_cast<const uint64_t*>(m + o_st),
                      reinterpret_cast<const int4*>(m + o_w), reinterpret_cast<double*>(m + o_p), n, cut->d2,
                      reinterpret_cast<unsigned long long*>(m + o_b), reinterpret_cast<int32_t*>(m + o_s), ctx->stream);
  if (e != hipSuccess) return hip_fail(ctx, e, "contacts: launch");
  if (W64 > 0)
    e = hipMemcpyAsync(bitsets, m + o_b, static_cast<size_t>(n) * W64 * 8, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess && set_sizes) e = hipMemcpyAsync(set_sizes, m + o_s, static_cast<size_t>(n) * 4, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  return e == hipSuccess ? GDOCK_OK : hip_fail(ctx, e, "contacts: download");
}

gdock_status gdock_cluster_structures(gdock_ctx* ctx, gdock_complex* cx, const double* genes, int32_t n,
                                      const gdock_clustering_config* cfg, gdock_cluster_result* out, int64_t* members,
                                      int32_t* n_clusters) {
  if (!ctx || !cx || !cfg || !n_clusters || n < 0 || (n && (!genes || !out || !members))) return GDOCK_ERR_INVALID_ARG;
  *n_clusters = 0;
  if (n == 0) return GDOCK_OK;  // clustering.rs:269-271
  if (n > kGreedyMaxN) { set_error(ctx, "cluster_structures: %d structures exceeds the supported %d", n, kGreedyMaxN); return GDOCK_ERR_INVALID_ARG; }
  if (!std::isfinite(cfg->contact_distance) || cfg->contact_distance < 0.0 || !std::isfinite(cfg->fcc_cutoff) ||
      !std::isfinite(cfg->strictness)) {
    set_error(ctx, "cluster_structures: bad config");
    return GDOCK_ERR_INVALID_ARG;
  }
  std::lock_guard<std::mutex> lk(cx->con_mu);
  const char* why = nullptr;
  const ContactCut* cut = cut_of(cx, cfg->contact_distance, &why);
  if (!cut) { set_error(ctx, "%s", why); return GDOCK_ERR_INVALID_ARG; }
  hipError_t e = hipSetDevice(ctx->device);
  if (e != hipSuccess) return hip_fail(ctx, e, "cluster: set device");
  const int W64 = cut->words64, W32 = 2 * W64;
  const int n_pad = (n + 63) & ~63;
  const int nwords = (n + 63) / 64;
  const int max_set = cut->n_bits;
  Layout L;
  const size_t o_st = L.add<uint64_t>(W64), o_w = L.add<int4>(cut->window.size());
  const size_t o_p = L.add<double>(static_cast<size_t>(n) * kPoseDoubles),
               o_b = L.add<uint64_t>(static_cast<size_t>(n) * W64), o_s = L.add<int32_t>(n_pad),
               o_u = L.add<uint32_t>(W32), o_col = L.add<int32_t>(W32),
               o_t = L.add<uint32_t>(static_cast<size_t>(W32) * n_pad), o_ca = L.add<int32_t>(max_set + 1),
               o_cb = L.add<int32_t>(max_set + 1), o_A = L.add<uint64_t>(static_cast<size_t>(n) * nwords),
               o_AT = L.add<uint64_t>(static_cast<size_t>(n) * nwords), o_cc = L.add<int64_t>(n), o_cs = L.add<int64_t>(n),
               o_co = L.add<int64_t>(n), o_cm = L.add<int64_t>(n), o_cn = L.add<int32_t>(1),
               o_rc = L.add<int32_t>(n);
  if (!ctx->fcc.reserve(L.size, &e)) return hip_fail(ctx, e, "cluster: scratch");
  char* m = static_cast<char*>(ctx->fcc.ptr);
  auto P = [&](size_t o) { return static_cast<void*>(m + o); };
  gdock_status st = stage_contacts(ctx, *cut, genes, n, m, o_p, o_st, o_w, "cluster_structures");
  if (st != GDOCK_OK) return st;
  e = launch_contacts(cx, *cut, cut->static_mask.empty() ? nullptr : static_cast<const uint64_t*>(P(o_st)),
                      static_cast<const int4*>(P(o_w)), static_cast<double*>(P(o_p)), n, cut->d2,
                      static_cast<unsigned long long*>(P(o_b)), static_cast<int32_t*>(P(o_s)), ctx->stream);
  if (e != hipSuccess) return hip_fail(ctx, e, "cluster: contacts");
  // Integer overlap minima per set size (host), computed while the device
  // finds the contacts.
  std::vector<int32_t> ca, cb;
  fcc_minima(cfg->fcc_cutoff, max_set, &ca);
  fcc_minima(cfg->fcc_cutoff * cfg->strictness, max_set, &cb);
  // Words set in any pose (none without candidate bits: every set is empty).
  std::vector<uint32_t> uni(W32);
  if (W32 > 0) {
    e = hipMemsetAsync(P(o_u), 0, W32 * 4, ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "cluster: union");
    hipLaunchKernelGGL(union_kernel, dim3((W32 + 255) / 256, (n + kUnionPoses - 1) / kUnionPoses), dim3(256), 0,
                       ctx->stream, static_cast<const uint32_t*>(P(o_b)), n, W32, static_cast<uint32_t*>(P(o_u)));
    e = hipMemcpyAsync(uni.data(), P(o_u), W32 * 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "cluster: union");
  }
  std::vector<int32_t> col;
  for (int w = 0; w < W32; ++w) if (uni[w]) col.push_back(w);
  const int wc = static_cast<int>(col.size());
  const int zero_ok = (0.0 >= cfg->fcc_cutoff && 0.0 >= cfg->fcc_cutoff * cfg->strictness) ? 1 : 0;
  if (wc > 0) e = hipMemcpyAsync(P(o_col), col.data(), wc * 4, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(P(o_ca), ca.data(), ca.size() * 4, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(P(o_cb), cb.data(), cb.size() * 4, hipMemcpyHostToDevice, ctx->stream);
  if (e != hipSuccess) return hip_fail(ctx, e, "cluster: tables");
  if (wc > 0)
    hipLaunchKernelGGL(compact_kernel, dim3(n_pad / 256 + 1, wc), dim3(256), 0, ctx->stream,
                       static_cast<const uint32_t*>(P(o_b)), n, n_pad, W32, static_cast<const int32_t*>(P(o_col)), wc,
                       static_cast<uint32_t*>(P(o_t)));
  hipLaunchKernelGGL(overlap_kernel, dim3(nwords, nwords), dim3(256), 0, ctx->stream,
                     static_cast<const uint32_t*>(P(o_t)), n, n_pad, wc, static_cast<const int32_t*>(P(o_s)),
                     static_cast<const int32_t*>(P(o_ca)), static_cast<const int32_t*>(P(o_cb)), zero_ok,
                     static_cast<unsigned long long*>(P(o_A)), static_cast<unsigned long long*>(P(o_AT)), nwords);
  hipLaunchKernelGGL(row_count_kernel, dim3((n + 3) / 4), dim3(256), 0, ctx->stream,
                     static_cast<const unsigned long long*>(P(o_A)), n, nwords, static_cast<int32_t*>(P(o_rc)));
  hipLaunchKernelGGL(greedy_kernel, dim3(1), dim3(kGreedyThreads), static_cast<size_t>(nwords) * 8 + static_cast<size_t>(n) * 8,
                     ctx->stream, static_cast<const unsigned long long*>(P(o_A)),
                     static_cast<const unsigned long long*>(P(o_AT)), static_cast<const int32_t*>(P(o_rc)), n, nwords,
                     static_cast<long long>(cfg->min_cluster_size), static_cast<int64_t*>(P(o_cc)),
                     static_cast<int64_t*>(P(o_cs)), static_cast<int64_t*>(P(o_co)), static_cast<int64_t*>(P(o_cm)),
                     static_cast<int32_t*>(P(o_cn)));
  e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(ctx, e, "cluster: launch");
  int32_t nclu = 0;
  e = hipMemcpyAsync(&nclu, P(o_cn), 4, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return hip_fail(ctx, e, "cluster: result");
  std::vector<int64_t> cc(nclu), cs(nclu), co(nclu);
  if (nclu > 0) {
    e = hipMemcpyAsync(cc.data(), P(o_cc), nclu * 8, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(cs.data(), P(o_cs), nclu * 8, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(co.data(), P(o_co), nclu * 8, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "cluster: result");
  }
  // clustering.rs:250-251 stable sort by size (descending), then 288-292 filter
  std::vector<int> order(nclu);
  for (int i = 0; i < nclu; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cs[a] > cs[b]; });
  std::vector<int64_t> all_members(static_cast<size_t>(co.empty() ? 0 : co.back() + cs.back()));
  if (!all_members.empty()) {
    e = hipMemcpy(all_members.data(), P(o_cm), all_members.size() * 8, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(ctx, e, "cluster: members");
  }
  int32_t k = 0;
  int64_t off = 0;
  for (int i : order) {
    if (cs[i] < cfg->min_cluster_size) continue;
    out[k].center_idx = cc[i];
    out[k].size = cs[i];
    out[k].members_offset = off;
    std::memcpy(members + off, all_members.data() + co[i], cs[i] * 8);
    off += cs[i];
    ++k;
  }
  *n_clusters = k;
  return GDOCK_OK;
}

}  // extern "C"
