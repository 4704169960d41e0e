// gdock FP64 oracle — TEST INFRASTRUCTURE ONLY.
//
// A CPU restatement, operation for operation, of the reference's population
// scoring and FCC clustering path (Rust crate gdock 2.0.0-rc.2).  Only the
// tests, the smoke check and the bench's cpu_baseline leg load this library,
// and only as the checker; the product (gdock_amd/) never links or calls it.
//
// Build: g++ -O3 -ffp-contract=off (no fast-math) so that every FP64 operation
// rounds exactly as the reference's (Rust never contracts a*b+c into an FMA).
// glibc sqrt/sin/cos/fmod are the ones the reference links on Linux.
//
// Parity pins (tests/test_oracle_*.py): the reference's own known-answer tests
// (fitness.rs:357-607, clustering.rs:299-363, restraints.rs:173-203,
// structure.rs:433-479, chromosome.rs:256-279, population.rs:492-530,
// run.rs:914-1033, toppar.rs:665-880), closed-form single-pair energies, the
// exact squared-distance thresholds checked against correctly rounded sqrt,
// invariants (rigid-motion invariance, COM preservation, monotone contacts) and
// brute force on tiny inputs.  The GA random stream is project-defined
// (xoshiro256** seeded by SplitMix64) because rand 0.8.5 / rand_distr 0.4.3
// are not available offline: that stream is "parity unpinned" against the
// reference's StdRng; every decision the GA takes on a given stream is pinned.
#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <functional>
#include <set>
#include <sstream>
#include <string>
#include <thread>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "toppar_data.inc"

namespace {

// ---------------------------------------------------------------- constants
// constants.rs:2-72, fitness.rs:129-138,286, run.rs:56-62, clustering.rs:26-35
constexpr double kPI = 3.14159265358979323846;
constexpr int kPopulationSize = 150;
constexpr double kMutationRate = 0.1;
constexpr double kCrossoverRate = 0.6;
constexpr int kTournamentSize = 3;
constexpr int kElitism = 5;
constexpr int kMaxGenerations = 250;
constexpr double kMaxDisplacement = 50.0;
constexpr double kConvergenceThreshold = 0.001;
constexpr int kConvergenceWindow = 10;
constexpr double kVdwCutoff = 12.0, kSoftcoreAlpha = 0.5, kDesolvCutoff = 8.0;
constexpr double kElecCutoff = 15.0, kElecMin = 1.0, kAirK = 10.0;
constexpr double kCoulomb = 332.0636;
constexpr int kHofMax = 500, kHofTopK = 10, kNumOutputModels = 5;
constexpr double kHofRot = 0.2, kHofTrans = 2.0;

// ---------------------------------------------------------------- data model
// structure.rs:28-49
struct Atom {
  int32_t serial = 0;
  std::string name;
  char altloc = ' ';
  std::string resname;
  char chainid = ' ';
  int16_t resseq = 0;
  char icode = ' ';
  double x = 0, y = 0, z = 0;
  float occupancy = 0, tempfactor = 0;
  std::string element;
  double charge = 0, vdw_radius = 0, epsilon = 0, rmin2 = 0;
};
using Molecule = std::vector<Atom>;

// Interchange record, same layout as the C-ABI's gdock_atom.
struct RecordAtom {
  double x, y, z, charge, vdw_radius, epsilon, rmin2;
  float occupancy, tempfactor;
  int32_t serial;
  int16_t resseq;
  char chainid, altloc, icode;
  char name[5], resname[5], element[5];
};
static_assert(sizeof(RecordAtom) == 88, "record layout");

// ---------------------------------------------------------------- toppar
struct ToppType { bool has_eps, has_rmin, has_q; double eps, rmin, q; };
struct Toppar {
  std::map<std::string, ToppType> types;
  std::map<std::string, std::map<std::string, std::string>> residues;
  Toppar() {
    std::istringstream in(kToppar);
    std::string line;
    while (std::getline(in, line)) {
      std::istringstream ls(line);
      std::string tag, key;
      ls >> tag >> key;
      if (tag == "T") {
        std::string e, r, q;
        ls >> e >> r >> q;
        ToppType t{e != "-", r != "-", q != "-", 0, 0, 0};
        if (t.has_eps) t.eps = std::strtod(e.c_str(), nullptr);
        if (t.has_rmin) t.rmin = std::strtod(r.c_str(), nullptr);
        if (t.has_q) t.q = std::strtod(q.c_str(), nullptr);
        types[key] = t;
      } else if (tag == "R") {
        std::string pair;
        auto& m = residues[key];
        while (ls >> pair) {
          auto c = pair.find(':');
          m[pair.substr(0, c)] = pair.substr(c + 1);
        }
      }
    }
  }
};
const Toppar& toppar() { static Toppar t; return t; }

// toppar.rs:152 get_atom
const std::string* get_atom(const std::string& res, const std::string& name) {
  auto& R = toppar().residues;
  auto it = R.find(res);
  if (it == R.end()) return nullptr;
  auto jt = it->second.find(name);
  return jt == it->second.end() ? nullptr : &jt->second;
}

// ---------------------------------------------------------------- PDB parsing
std::string trim(const std::string& s) {
  size_t a = 0, b = s.size();
  while (a < b && std::isspace(static_cast<unsigned char>(s[a]))) ++a;
  while (b > a && std::isspace(static_cast<unsigned char>(s[b - 1]))) --b;
  return s.substr(a, b - a);
}

// Rust's str::parse::<iN>: optional sign then decimal digits only.
bool parse_int(const std::string& s, long lo, long hi, long* out) {
  if (s.empty()) return false;
  size_t i = 0;
  bool neg = false;
  if (s[0] == '+' || s[0] == '-') { neg = s[0] == '-'; i = 1; }
  if (i == s.size()) return false;
  long v = 0;
  for (; i < s.size(); ++i) {
    if (s[i] < '0' || s[i] > '9') return false;
    v = v * 10 + (s[i] - '0');
    if (v > 100000000L) return false;
  }
  v = neg ? -v : v;
  if (v < lo || v > hi) return false;
  *out = v;
  return true;
}

// Rust's str::parse::<f64> for the decimal forms found in PDB columns.
bool parse_f64(const std::string& s, double* out) {
  if (s.empty()) return false;
  for (char c : s)
    if (!(std::isdigit(static_cast<unsigned char>(c)) || c == '.' || c == '-' ||
          c == '+' || c == 'e' || c == 'E'))
      return false;
  char* end = nullptr;
  double v = std::strtod(s.c_str(), &end);
  if (end != s.c_str() + s.size()) return false;
  *out = v;
  return true;
}

// Rust's f32 parse of the same forms: correctly rounded straight to f32
// (not through f64, which could round twice).
bool parse_f32(const std::string& s, float* out) {
  double probe;
  if (!parse_f64(s, &probe)) return false;
  *out = std::strtof(s.c_str(), nullptr);
  return true;
}

// structure.rs:187-270 process_atom_line
bool process_atom_line(const std::string& line, Atom* atom) {
  if (line.compare(0, 4, "ATOM") != 0 || line.size() < 54) return false;
  auto get = [&](size_t a, size_t b) -> std::string {
    return line.size() >= b ? line.substr(a, b - a) : std::string();
  };
  Atom t;
  long iv;
  if (!parse_int(trim(get(6, 11)), INT32_MIN, INT32_MAX, &iv)) return false;
  t.serial = static_cast<int32_t>(iv);
  t.name = trim(get(12, 16));
  { auto s = get(16, 17); t.altloc = s.empty() ? ' ' : s[0]; }
  t.resname = trim(get(17, 20));
  { auto s = get(21, 22); t.chainid = s.empty() ? ' ' : s[0]; }
  if (!parse_int(trim(get(22, 26)), -32768, 32767, &iv)) return false;
  t.resseq = static_cast<int16_t>(iv);
  { auto s = get(26, 27); t.icode = s.empty() ? ' ' : s[0]; }
  if (!parse_f64(trim(get(30, 38)), &t.x)) return false;
  if (!parse_f64(trim(get(38, 46)), &t.y)) return false;
  if (!parse_f64(trim(get(46, 54)), &t.z)) return false;
  if (!parse_f32(trim(get(54, 60)), &t.occupancy)) t.occupancy = 1.0f;
  if (!parse_f32(trim(get(60, 66)), &t.tempfactor)) t.tempfactor = 0.0f;
  t.element = trim(get(76, 78));
  if (t.element.empty()) t.element = std::string(1, t.name.empty() ? ' ' : t.name[0]);
  const std::string el = trim(t.element);
  t.vdw_radius = el == "H" ? 1.2 : el == "C" ? 1.7 : el == "N" ? 1.55 : el == "O" ? 1.52 : 1.0;
  if (const std::string* ty = get_atom(t.resname, t.name)) {
    auto it = toppar().types.find(*ty);
    if (it != toppar().types.end()) {
      t.epsilon = it->second.has_eps ? it->second.eps : 0.0;
      t.rmin2 = it->second.has_rmin ? it->second.rmin : 0.0;
      t.charge = it->second.has_q ? it->second.q : 0.0;
    }
  }
  *atom = t;
  return true;
}

// structure.rs:273-300 read_pdb (from text instead of a file)
std::vector<Molecule> read_pdb_text(const char* text) {
  std::vector<Molecule> model;
  Molecule mol;
  bool has_models = false;
  std::istringstream in(text);
  std::string line;
  while (std::getline(in, line)) {
    if (!line.empty() && line.back() == '\r') line.pop_back();
    Atom a;
    if (line.compare(0, 5, "MODEL") == 0) {
      has_models = true;
      mol.clear();
    } else if (line.compare(0, 6, "ENDMDL") == 0) {
      model.push_back(mol);
      mol.clear();
    } else if (process_atom_line(line, &a)) {
      mol.push_back(a);
    }
  }
  if (!has_models) model.push_back(mol);
  return model;
}

// ---------------------------------------------------------------- geometry
// structure.rs:63-69
void center_of_mass(const Molecule& m, double* cx, double* cy, double* cz) {
  double n = static_cast<double>(m.size()), sx = 0, sy = 0, sz = 0;
  for (auto& a : m) sx += a.x;
  for (auto& a : m) sy += a.y;
  for (auto& a : m) sz += a.z;
  *cx = sx / n; *cy = sy / n; *cz = sz / n;
}

// structure.rs:80-117 rotate (Euler, about the COM), op order kept exactly
Molecule rotate(Molecule m, double alpha, double beta, double gamma) {
  double cx, cy, cz;
  center_of_mass(m, &cx, &cy, &cz);
  for (auto& a : m) { a.x -= cx; a.y -= cy; a.z -= cz; }
  // f64::cos and f64::sin of one angle compile to one glibc sincos call
  // (LLVM on GNU targets; g++ -O3 does the same): called explicitly here.
  double ca, sa, cb, sb, cg, sg;
  ::sincos(alpha, &sa, &ca);
  ::sincos(beta, &sb, &cb);
  ::sincos(gamma, &sg, &cg);
  for (auto& a : m) {
    const double x = a.x, y = a.y, z = a.z;
    const double nx = x * ca * cb + y * (-sa * cg + ca * sb * sg) + z * (sa * sg + ca * sb * cg);
    const double ny = x * sa * cb + y * (ca * cg + sa * sb * sg) + z * (-ca * sg + sa * sb * cg);
    const double nz = -x * sb + y * cb * sg + z * cb * cg;
    a.x = nx; a.y = ny; a.z = nz;
  }
  for (auto& a : m) { a.x += cx; a.y += cy; a.z += cz; }
  return m;
}

// structure.rs:119-131
Molecule displace(Molecule m, double dx, double dy, double dz) {
  for (auto& a : m) { a.x += dx; a.y += dy; a.z += dz; }
  return m;
}

// chromosome.rs:79-83
Molecule apply_genes(const double* g, const Molecule& lig) {
  return displace(rotate(lig, g[0], g[1], g[2]), g[3], g[4], g[5]);
}

// utils.rs:77-95
Molecule position_ligand(const Molecule& rec, Molecule lig) {
  double rx, ry, rz, lx, ly, lz;
  center_of_mass(rec, &rx, &ry, &rz);
  center_of_mass(lig, &lx, &ly, &lz);
  for (auto& a : lig) { a.x -= lx; a.y -= ly; a.z -= lz; }
  for (auto& a : lig) { a.x += rx; a.y += ry; a.z += rz; }  // Molecule::translate 71-77
  return lig;
}

// structure.rs:312-317
inline double distance(const Atom& a, const Atom& b) {
  const double dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
  return std::sqrt(dx * dx + dy * dy + dz * dz);
}

// Rust f64::powi with a constant exponent (LLVM ExpandPowI / compiler-rt
// __powidf2 both evaluate by repeated squaring).
inline double powi6(double x) { const double x2 = x * x; return x2 * (x2 * x2); }
inline double powi2(double x) { return x * x; }

// ---------------------------------------------------------------- energies
// fitness.rs:57-70
inline double softcore_lj(const Atom& a1, const Atom& a2, double d, double alpha) {
  const double eps_ij = std::sqrt(a1.epsilon * a2.epsilon);
  const double rmin_ij = a1.rmin2 + a2.rmin2;
  const double rmin6 = powi6(rmin_ij);
  const double r6 = powi6(d);
  const double r_eff6 = r6 + alpha * rmin6;
  const double ratio = rmin6 / r_eff6;
  return eps_ij * (powi2(ratio) - 2.0 * ratio);
}

// fitness.rs:86-105
double vdw_energy(const Molecule& rec, const Molecule& lig) {
  double e = 0.0;
  for (auto& a1 : rec)
    for (auto& a2 : lig)
      if (a1.serial != a2.serial) {
        const double d = distance(a1, a2);
        if (d > 0.0 && d < kVdwCutoff) {
          const double v = softcore_lj(a1, a2, d, kSoftcoreAlpha);
          e += std::min(std::max(v, -50.0), 500.0);
        }
      }
  return e;
}

// fitness.rs:284-302
double elec_energy(const Molecule& rec, const Molecule& lig) {
  const double k = kCoulomb;
  double e = 0.0;
  for (auto& a1 : rec)
    for (auto& a2 : lig) {
      const double d = distance(a1, a2);
      if (d > kElecMin && d < kElecCutoff) e += k * (a1.charge * a2.charge) / (d * d);
    }
  return e;
}

// fitness.rs:129-138
inline double get_asp(const std::string& element) {
  const std::string e = trim(element);
  if (e == "C") return 0.012;
  if (e == "N") return -0.160;
  if (e == "O") return -0.160;
  if (e == "S") return 0.012;
  return 0.0;
}

// fitness.rs:122-184
double desolv_energy(const Molecule& rec, const Molecule& lig) {
  double e = 0.0;
  for (auto& a1 : rec) {
    const double asp = get_asp(a1.element);
    if (std::fabs(asp) < 0.001) continue;
    int cnt = 0;
    for (auto& a2 : lig)
      if (distance(a1, a2) < kDesolvCutoff) ++cnt;
    const double burial = std::min(static_cast<double>(cnt) / 10.0, 1.0);
    e -= asp * burial * a1.vdw_radius * a1.vdw_radius * 4.0 * kPI;
  }
  for (auto& a2 : lig) {
    const double asp = get_asp(a2.element);
    if (std::fabs(asp) < 0.001) continue;
    int cnt = 0;
    for (auto& a1 : rec)
      if (distance(a1, a2) < kDesolvCutoff) ++cnt;
    const double burial = std::min(static_cast<double>(cnt) / 10.0, 1.0);
    e -= asp * burial * a2.vdw_radius * a2.vdw_radius * 4.0 * kPI;
  }
  return e;
}

// restraints.rs:4 Restraint(receptor atom, ligand atom); only resseq is used later.
struct Restraint { int16_t rec_resseq, lig_resseq; };

const Atom* find_ca(const Molecule& m, int16_t resseq) {
  for (auto& a : m)
    if (a.resseq == resseq && trim(a.name) == "CA") return &a;
  return nullptr;
}

// restraints.rs:44-77
std::vector<Restraint> restraints_from_pairs(const Molecule& m1, const Molecule& m2,
                                             const int32_t* pairs, int n) {
  std::vector<Restraint> out;
  for (int i = 0; i < n; ++i) {
    const Atom *a1 = nullptr, *a2 = nullptr;
    for (auto& a : m1) if (trim(a.name) == "CA" && a.resseq == pairs[2 * i]) { a1 = &a; break; }
    for (auto& a : m2) if (trim(a.name) == "CA" && a.resseq == pairs[2 * i + 1]) { a2 = &a; break; }
    if (a1 && a2) out.push_back({a1->resseq, a2->resseq});
  }
  return out;
}

// restraints.rs:11-40
bool is_satisfied(const Restraint& r, const Molecule& rec, const Molecule& lig) {
  const Atom* c1 = find_ca(rec, r.rec_resseq);
  const Atom* c2 = find_ca(lig, r.lig_resseq);
  if (c1 && c2) return distance(*c1, *c2) <= 7.0;
  return false;
}

// fitness.rs:205-249
double air_energy(const std::vector<Restraint>& rs, const Molecule& rec, const Molecule& lig) {
  double e = 0.0;
  const double lower = 0.0, upper = 7.0;
  for (auto& r : rs) {
    const Atom* c1 = find_ca(rec, r.rec_resseq);
    const Atom* c2 = find_ca(lig, r.lig_resseq);
    if (c1 && c2) {
      const double d = distance(*c1, *c2);
      const double v = d < lower ? lower - d : d > upper ? d - upper : 0.0;
      e += kAirK * v * v;
    }
  }
  return e;
}

// fitness.rs:251-267
double satisfaction_ratio(const std::vector<Restraint>& rs, const Molecule& rec, const Molecule& lig) {
  size_t s = 0;
  for (auto& r : rs) s += is_satisfied(r, rec, lig) ? 1 : 0;
  return rs.empty() ? 0.0 : static_cast<double>(s) / static_cast<double>(rs.size());
}

// chromosome.rs:85-128 (normal mode).  out = {vdw, elec, desolv, air,
// restraint_penalty, fitness}; *n_sat = satisfied restraint count.
void chromosome_fitness(const Molecule& rec, const Molecule& lig, const std::vector<Restraint>& rs,
                        const double* w, const double* genes, bool transform, double* out,
                        int32_t* n_sat) {
  const Molecule t = transform ? apply_genes(genes, lig) : lig;
  out[0] = vdw_energy(rec, t);
  out[1] = elec_energy(rec, t);
  out[2] = desolv_energy(rec, t);
  out[3] = air_energy(rs, rec, t);
  const double ratio = satisfaction_ratio(rs, rec, t);
  out[4] = (1.0 - ratio) * static_cast<double>(rs.size());
  out[5] = w[0] * out[0] + w[1] * out[1] + w[2] * out[2] + w[3] * out[3];
  if (n_sat) {
    int32_t s = 0;
    for (auto& r : rs) s += is_satisfied(r, rec, t) ? 1 : 0;
    *n_sat = s;
  }
}

// ---------------------------------------------------------------- clustering
struct ClusteringConfig { double contact_distance, fcc_cutoff, strictness; int64_t min_cluster_size; };
struct ClusterResult { int64_t center; std::vector<int64_t> members; };

// clustering.rs:62-115; a contact is kept as (chainA, resA, chainB, resB)
// with the smaller chain first, which is the string's content.
using Contact = std::tuple<char, int16_t, char, int16_t>;
std::set<Contact> calculate_contacts(const Molecule& m, double cutoff) {
  std::map<std::pair<char, int16_t>, std::vector<std::array<double, 3>>> residues;
  for (auto& a : m) {
    if (trim(a.element) == "H") continue;
    residues[{a.chainid, a.resseq}].push_back({a.x, a.y, a.z});
  }
  std::vector<std::pair<char, int16_t>> keys;
  for (auto& kv : residues) keys.push_back(kv.first);
  std::set<Contact> contacts;
  for (size_t i = 0; i < keys.size(); ++i)
    for (size_t j = i + 1; j < keys.size(); ++j) {
      auto [ca, ra] = keys[i];
      auto [cb, rb] = keys[j];
      if (ca == cb) continue;
      bool hit = false;
      for (auto& pa : residues[keys[i]]) {
        for (auto& pb : residues[keys[j]]) {
          const double d = std::sqrt(powi2(pa[0] - pb[0]) + powi2(pa[1] - pb[1]) + powi2(pa[2] - pb[2]));
          if (d <= cutoff) { hit = true; break; }
        }
        if (hit) break;
      }
      if (hit) contacts.insert(ca < cb ? Contact{ca, ra, cb, rb} : Contact{cb, rb, ca, ra});
    }
  return contacts;
}

// clustering.rs:122-132
std::pair<double, double> calculate_fcc(const std::set<Contact>& x, const std::set<Contact>& y) {
  if (x.empty() || y.empty()) return {0.0, 0.0};
  size_t common = 0;
  for (auto& c : x) common += y.count(c);
  return {static_cast<double>(common) / static_cast<double>(x.size()),
          static_cast<double>(common) / static_cast<double>(y.size())};
}

// clustering.rs:151-180 (create_elements) on per-pair intersection counts.
std::vector<std::vector<uint8_t>> neighbour_matrix(const std::vector<size_t>& sizes,
                                                   const std::vector<uint32_t>& common,
                                                   const ClusteringConfig& cfg) {
  const size_t n = sizes.size();
  std::vector<std::vector<uint8_t>> nb(n, std::vector<uint8_t>(n, 0));
  for (size_t i = 0; i < n; ++i)
    for (size_t j = 0; j < n; ++j) {
      double fcc = 0, fccv = 0;  // calculate_fcc for (i, j)
      if (sizes[i] && sizes[j]) {
        const double c = static_cast<double>(common[i * n + j]);
        fcc = c / static_cast<double>(sizes[i]);
        fccv = c / static_cast<double>(sizes[j]);
      }
      if (fcc >= cfg.fcc_cutoff && fccv >= cfg.fcc_cutoff * cfg.strictness) nb[i][j] = 1;
      if (fccv >= cfg.fcc_cutoff && fcc >= cfg.fcc_cutoff * cfg.strictness) nb[j][i] = 1;
    }
  return nb;
}

// clustering.rs:183-254 cluster_elements, followed by the min-size filter of
// cluster_structures 288-292.  Straight transcription: O(n^2) per step.
std::vector<ClusterResult> cluster_elements(const std::vector<std::vector<uint8_t>>& nb,
                                            size_t min_size) {
  const size_t n = nb.size();
  std::vector<uint8_t> used(n, 0);
  std::vector<ClusterResult> clusters;
  for (;;) {
    int64_t center = -1;
    size_t best = 0;
    for (size_t a = 0; a < n; ++a) {
      if (used[a]) continue;
      size_t cnt = 0;
      for (size_t b = 0; b < n; ++b) cnt += nb[a][b] && !used[b];
      if (center < 0 || cnt > best) { center = static_cast<int64_t>(a); best = cnt; }
    }
    if (center < 0) break;
    ClusterResult c;
    c.center = center;
    c.members.push_back(center);
    for (size_t b = 0; b < n; ++b)
      if (nb[center][b] && !used[b] && static_cast<int64_t>(b) != center) {
        c.members.push_back(static_cast<int64_t>(b));
        used[b] = 1;
      }
    used[center] = 1;
    std::sort(c.members.begin(), c.members.end());
    clusters.push_back(std::move(c));
  }
  std::stable_sort(clusters.begin(), clusters.end(),
                   [](const ClusterResult& a, const ClusterResult& b) {
                     return a.members.size() > b.members.size();
                   });
  std::vector<ClusterResult> out;
  for (auto& c : clusters)
    if (c.members.size() >= min_size) out.push_back(c);
  return out;
}

// clustering.rs:265-293
std::vector<ClusterResult> cluster_structures(const std::vector<Molecule>& structs,
                                              const ClusteringConfig& cfg,
                                              std::vector<std::set<Contact>>* sets_out) {
  if (structs.empty()) return {};
  std::vector<std::set<Contact>> sets(structs.size());
  for (size_t i = 0; i < structs.size(); ++i) sets[i] = calculate_contacts(structs[i], cfg.contact_distance);
  const size_t n = sets.size();
  std::vector<size_t> sizes(n);
  std::vector<uint32_t> common(n * n);
  for (size_t i = 0; i < n; ++i) sizes[i] = sets[i].size();
  for (size_t i = 0; i < n; ++i)
    for (size_t j = 0; j < n; ++j) {
      size_t c = 0;
      for (auto& k : sets[i]) c += sets[j].count(k);
      common[i * n + j] = static_cast<uint32_t>(c);
    }
  auto nb = neighbour_matrix(sizes, common, cfg);
  if (sets_out) *sets_out = std::move(sets);
  return cluster_elements(nb, static_cast<size_t>(cfg.min_cluster_size));
}

// ---------------------------------------------------------------- GA
// Test builds of the oracle define ORACLE_MUTANT to plant one plausible
// mistake in the GA decisions (see the switches below); the known-answer
// tests must fail on every such build.  0: the restatement itself.
#ifndef ORACLE_MUTANT
#define ORACLE_MUTANT 0
#endif

// The project stream's logarithm and cosine (the library's stream.h defines
// them): IEEE basic operations only, so the library's host and device code
// and this restatement compute the same bits.  ln: x = m 2^e, m in
// [sqrt(1/2), sqrt(2)); ln m = 2 atanh(s), s = (m - 1) / (m + 1), as the odd
// series to s^25 in Horner form on z = s^2; plus e ln 2 with ln 2 split in a
// high part exact under multiplication by e and a low part.
inline double stream_ln(double x) {
  uint64_t b;
  std::memcpy(&b, &x, 8);
  int e = static_cast<int>((b >> 52) & 0x7ff) - 1023;
  const uint64_t mb = (b & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL;
  double m;
  std::memcpy(&m, &mb, 8);
  if (m > 1.4142135623730951) { m = m * 0.5; e += 1; }
  const double f = m - 1.0, s = f / (2.0 + f), z = s * s;
  double p = 1.0 / 25.0;
  for (int k = 11; k >= 1; --k) p = 1.0 / static_cast<double>(2 * k + 1) + z * p;  // 1/23, ..., 1/3
  const double s2 = 2.0 * s;
  const double de = static_cast<double>(e);
  return de * 6.93147180369123816490e-01 + (de * 1.90821492927058770002e-10 + (s2 + s2 * (z * p)));
}

// cos(2 pi u), 0 <= u < 1: nearest quarter k (u - k/4 exact), x = 2 pi r,
// Taylor series of cos (to x^18) or sin (to x^19) chosen by k mod 4.
inline double stream_cos2pi(double u) {
  const double k = static_cast<double>(static_cast<int64_t>(u * 4.0 + 0.5));
  const double x = (u - k * 0.25) * 6.283185307179586, z = x * x;
  double fact[20];
  fact[0] = 1.0;
  for (int i = 1; i < 20; ++i) fact[i] = fact[i - 1] * static_cast<double>(i);  // exact to 18!; 19! rounds once
  double c = -1.0 / fact[18];
  for (int i = 16; i >= 4; i -= 2) c = ((i / 2) % 2 ? -1.0 : 1.0) / fact[i] + z * c;
  c = -0.5 + z * c;
  const double cosx = 1.0 + z * c;
  double sn = -1.0 / fact[19];
  for (int i = 17; i >= 3; i -= 2) sn = ((i / 2) % 2 ? -1.0 : 1.0) / fact[i] + z * sn;
  const double sinx = x + x * (z * sn);
  switch (static_cast<int>(k) & 3) {
    case 0: return cosx;
    case 1: return -sinx;
    case 2: return -cosx;
    default: return sinx;
  }
}

// Project-defined random stream: SplitMix64 seeding, xoshiro256**.  Test
// hook: with a script, every draw replays the next scripted value instead
// (unit and unit_incl: the value itself; below(n): the index; normal: the
// standard deviate z, returned as mean + sd z) and is logged as (kind, n).
struct Rng {
  uint64_t s[4];
  const double* script = nullptr;
  int script_len = 0, script_pos = 0;
  bool script_error = false;  // ran out, or a scripted index out of range
  std::vector<std::pair<int, double>>* log = nullptr;
  double scripted(int kind, double arg) {
    if (log) log->emplace_back(kind, arg);
    if (script_pos >= script_len) { script_error = true; return 0.0; }
    return script[script_pos++];
  }
  explicit Rng(uint64_t seed) {
    uint64_t x = seed;
    for (auto& v : s) {
      uint64_t z = (x += 0x9e3779b97f4a7c15ULL);
      z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
      z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
      v = z ^ (z >> 31);
    }
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    const uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
    s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
    s[2] ^= t; s[3] = rotl(s[3], 45);
    return r;
  }
  double unit() {  // [0,1)
    if (script) return scripted(0, 0.0);
    return static_cast<double>(next() >> 11) * 0x1.0p-53;
  }
  double unit_incl() {  // [0,1]
    if (script) return scripted(1, 0.0);
    return static_cast<double>(next() >> 11) / 9007199254740991.0;
  }
  double range_incl(double lo, double hi) { double v = lo + (hi - lo) * unit_incl(); return v > hi ? hi : v; }
  uint64_t below(uint64_t n) {  // Lemire
    if (script) {
      const double v = scripted(2, static_cast<double>(n));
      if (!(v >= 0.0 && v < static_cast<double>(n))) { script_error = true; return 0; }
      return static_cast<uint64_t>(v);
    }
    __uint128_t m = static_cast<__uint128_t>(next()) * n;
    uint64_t l = static_cast<uint64_t>(m);
    if (l < n) {
      const uint64_t t = (0 - n) % n;
      while (l < t) { m = static_cast<__uint128_t>(next()) * n; l = static_cast<uint64_t>(m); }
    }
    return static_cast<uint64_t>(m >> 64);
  }
  double normal(double mean, double sd) {  // Box-Muller, one draw per call
    if (script) return mean + sd * scripted(3, 0.0);
    const double u1 = 1.0 - unit(), u2 = unit();
    return mean + sd * (std::sqrt(-2.0 * stream_ln(u1)) * stream_cos2pi(u2));
  }
};

struct Chromo {
  double genes[6];
  double fitness = 0, vdw = 0, elec = 0, desolv = 0, air = 0, penalty = 0;
};

inline double rem_euclid(double a, double b) {  // Rust f64::rem_euclid
  const double r = std::fmod(a, b);
  return r < 0.0 ? r + std::fabs(b) : r;
}

// chromosome.rs:27-54
Chromo new_chromosome(Rng& rng) {
  Chromo c;
  for (int i = 0; i < 3; ++i) c.genes[i] = rng.range_incl(0.0, 2.0 * kPI);
  for (int i = 3; i < 6; ++i) c.genes[i] = rng.range_incl(-kMaxDisplacement, kMaxDisplacement);
  return c;
}

// chromosome.rs:56-77
void mutate(Chromo& c, Rng& rng, double rate) {
  for (int i = 0; i < 6; ++i)
#if ORACLE_MUTANT == 4  // planted: mutation at u < rate
    if (rng.unit() < rate) {
#else
    if (rng.unit() <= rate) {
#endif
#if ORACLE_MUTANT == 8  // planted: angles wrapped with fmod alone (negative angles stay negative)
      if (i < 3) c.genes[i] = std::fmod(c.genes[i] + rng.normal(0.0, 0.174), 2.0 * kPI);
#else
      if (i < 3) c.genes[i] = rem_euclid(c.genes[i] + rng.normal(0.0, 0.174), 2.0 * kPI);
#endif
      else c.genes[i] = std::min(std::max(c.genes[i] + rng.normal(0.0, 1.0), -kMaxDisplacement), kMaxDisplacement);
    }
}

// ga.rs:7-39
void crossover(Rng& rng, const Chromo& p1, const Chromo& p2, Chromo* o1, Chromo* o2) {
  *o1 = p1; *o2 = p2;
  for (int i = 0; i < 6; ++i)
#if ORACLE_MUTANT == 3  // planted: crossover at u < rate
    if (rng.unit() < kCrossoverRate) { o1->genes[i] = p2.genes[i]; o2->genes[i] = p1.genes[i]; }
#else
    if (rng.unit() <= kCrossoverRate) { o1->genes[i] = p2.genes[i]; o2->genes[i] = p1.genes[i]; }
#endif
}

// population.rs:124-158
std::vector<Chromo> tournament(const std::vector<Chromo>& pop, Rng& rng) {
  std::vector<Chromo> out;
  for (size_t k = 0; k < pop.size(); ++k) {
    const Chromo* best = nullptr;
    for (int t = 0; t < kTournamentSize; ++t) {
      const Chromo* c = &pop[rng.below(pop.size())];
#if ORACLE_MUTANT == 5  // planted: the last minimum wins
      if (!best || c->fitness <= best->fitness) best = c;
#else
      if (!best || c->fitness < best->fitness) best = c;  // min_by keeps the first minimum
#endif
    }
    out.push_back(*best);
  }
  return out;
}

// population.rs:74-121
std::vector<Chromo> evolve(const std::vector<Chromo>& pop, Rng& rng) {
  std::vector<Chromo> elite = pop;
#if ORACLE_MUTANT == 2  // planted: an unstable sort (equal fitness in reverse order)
  std::reverse(elite.begin(), elite.end());
#endif
  std::stable_sort(elite.begin(), elite.end(), [](const Chromo& a, const Chromo& b) { return a.fitness < b.fitness; });
  elite.resize(std::min<size_t>(kElitism, elite.size()));
  std::vector<Chromo> np = tournament(pop, rng);
#if ORACLE_MUTANT == 10  // planted: a forward shuffle (swap i with below(n - i) + i)
  for (size_t i = 0; i + 1 < np.size(); ++i) std::swap(np[i], np[i + rng.below(np.size() - i)]);
#else
  for (size_t i = np.size(); i > 1; --i) std::swap(np[i - 1], np[rng.below(i)]);  // shuffle
#endif
  std::vector<Chromo> kids;
  for (size_t i = 0; i + 1 < np.size(); i += 2) {
    Chromo a, b;
    crossover(rng, np[i], np[i + 1], &a, &b);
    kids.push_back(a); kids.push_back(b);
  }
  for (auto& c : kids) mutate(c, rng, kMutationRate);
#if ORACLE_MUTANT == 1  // planted: the elites overwrite the best children by stale fitness
  std::stable_sort(kids.begin(), kids.end(), [](const Chromo& a, const Chromo& b) { return a.fitness < b.fitness; });
#else
  std::stable_sort(kids.begin(), kids.end(), [](const Chromo& a, const Chromo& b) { return b.fitness < a.fitness; });
#endif
  for (size_t i = 0; i < elite.size() && i < kids.size(); ++i) kids[i] = elite[i];
  return kids;
}

// run.rs:77-216 HallOfFame
struct HofEntry { double genes[6]; double fitness; uint32_t generation; };
struct HallOfFame {
  std::vector<HofEntry> entries;
  size_t max_size = kHofMax;
  static double angdiff(double a, double b) { const double d = std::fabs(a - b); return d > kPI ? 2.0 * kPI - d : d; }
  static bool similar(const double* a, const double* b) {
    for (int i = 0; i < 3; ++i) if (angdiff(a[i], b[i]) > kHofRot) return false;
    for (int i = 3; i < 6; ++i) if (std::fabs(a[i] - b[i]) > kHofTrans) return false;
    return true;
  }
  bool try_add(const double* g, double f, uint32_t gen) {
    for (auto& e : entries) if (similar(g, e.genes)) return false;
    HofEntry e; std::memcpy(e.genes, g, sizeof e.genes); e.fitness = f; e.generation = gen;
    entries.push_back(e);
    if (entries.size() > max_size) {
      std::stable_sort(entries.begin(), entries.end(), [](const HofEntry& a, const HofEntry& b) { return a.fitness < b.fitness; });
      entries.resize(max_size);
    }
    return true;
  }
  void add_from_population(const std::vector<Chromo>& pop, uint32_t gen) {
    std::vector<size_t> idx(pop.size());
    for (size_t i = 0; i < idx.size(); ++i) idx[i] = i;
    std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) { return pop[a].fitness < pop[b].fitness; });
    int added = 0;
    for (size_t i : idx) {
      if (added >= kHofTopK) break;
      if (try_add(pop[i].genes, pop[i].fitness, gen)) ++added;
    }
  }
};

// run.rs:694-733: cluster centres stable-sorted by fitness, the first five
// taken; fewer than five are filled from the best-scored Hall of Fame
// entries not chosen yet (stable by fitness), each with cluster size 1.
std::vector<std::pair<int64_t, int64_t>> select_models(const std::vector<HofEntry>& hof,
                                                      const std::vector<ClusterResult>& clusters) {
  std::vector<std::pair<int64_t, int64_t>> selected;
  std::vector<std::tuple<int64_t, double, int64_t>> centres;
  for (auto& c : clusters)
    centres.emplace_back(c.center, hof[c.center].fitness, static_cast<int64_t>(c.members.size()));
  std::stable_sort(centres.begin(), centres.end(), [](auto& a, auto& b) { return std::get<1>(a) < std::get<1>(b); });
  for (size_t i = 0; i < centres.size() && i < static_cast<size_t>(kNumOutputModels); ++i)
    selected.emplace_back(std::get<0>(centres[i]), std::get<2>(centres[i]));
  if (selected.size() < static_cast<size_t>(kNumOutputModels)) {
    std::vector<std::pair<int64_t, double>> all;
    for (size_t i = 0; i < hof.size(); ++i) all.emplace_back(static_cast<int64_t>(i), hof[i].fitness);
#if ORACLE_MUTANT != 7  // planted (7): filled in Hall of Fame order instead of by fitness
    std::stable_sort(all.begin(), all.end(), [](auto& a, auto& b) { return a.second < b.second; });
#endif
    std::vector<int64_t> chosen;
    for (auto& s : selected) chosen.push_back(s.first);
    for (auto& [idx, f] : all) {
      if (selected.size() >= static_cast<size_t>(kNumOutputModels)) break;
      if (std::find(chosen.begin(), chosen.end(), idx) == chosen.end()) selected.emplace_back(idx, 1);
    }
  }
  return selected;
}

// Per-job state the parity test reads back.
struct GaTrace {
  std::vector<double> genes;    // per generation: pop x 6, as evaluated
  std::vector<double> fitness;  // per generation: pop
  std::vector<int32_t> best_idx;
  int32_t generations = 0;      // number of evaluated generations
  int32_t converged_at = -1;
  HallOfFame hof;
  std::vector<ClusterResult> clusters;
  std::vector<std::pair<int64_t, int64_t>> selected;  // (hof index, cluster size)
  std::vector<double> final_genes;  // the population at loop exit (run.rs:553-554 scores it for --no-clust)
};

void score_population(const Molecule& rec, const Molecule& lig, const std::vector<Restraint>& rs,
                      const double* w, std::vector<Chromo>& pop, int nthreads) {
  std::atomic<size_t> next{0};
  auto worker = [&]() {
    for (size_t i; (i = next.fetch_add(1)) < pop.size();) {
      double out[6];
      chromosome_fitness(rec, lig, rs, w, pop[i].genes, true, out, nullptr);
      pop[i].vdw = out[0]; pop[i].elec = out[1]; pop[i].desolv = out[2];
      pop[i].air = out[3]; pop[i].penalty = out[4]; pop[i].fitness = out[5];
    }
  };
  if (nthreads <= 1) { worker(); return; }
  std::vector<std::thread> th;
  for (int t = 0; t < nthreads; ++t) th.emplace_back(worker);
  for (auto& t : th) t.join();
}

// run.rs:361-531 (generation loop) and 665-733 (HoF -> FCC -> model selection)
// `score` fills every chromosome's fitness (and components); null scores the
// energies (Population::eval_fitness, normal mode).
void run_ga(const Molecule& rec, const Molecule& lig_pos, const std::vector<Restraint>& rs,
            const double* w, uint64_t seed, int pop_size, int max_gen, bool cluster, int nthreads,
            GaTrace* tr, const std::function<void(std::vector<Chromo>&)>& score = nullptr) {
  Rng rng(seed);
  std::vector<Chromo> pop;
  for (int i = 0; i < pop_size; ++i) pop.push_back(new_chromosome(rng));
  int gen = 0, stall = 0;
  double last = 1.7976931348623157e308;  // f64::MAX
  while (gen < max_gen) {
    for (auto& c : pop) tr->genes.insert(tr->genes.end(), c.genes, c.genes + 6);
    if (score) score(pop); else score_population(rec, lig_pos, rs, w, pop, nthreads);
    for (auto& c : pop) tr->fitness.push_back(c.fitness);
    tr->hof.add_from_population(pop, static_cast<uint32_t>(gen));
    size_t bi = 0;
    for (size_t i = 1; i < pop.size(); ++i) if (pop[i].fitness < pop[bi].fitness) bi = i;
    tr->best_idx.push_back(static_cast<int32_t>(bi));
    const double best = pop[bi].fitness;
    if (gen > 0) {
#if ORACLE_MUTANT == 9  // planted: the improvement divided by the signed previous best
      const double imp = (last - best) / last;
#else
      const double imp = (last - best) / std::fabs(last);
#endif
#if ORACLE_MUTANT == 6  // planted: a stall counted at imp <= threshold
      stall = imp <= kConvergenceThreshold ? stall + 1 : 0;
#else
      stall = imp < kConvergenceThreshold ? stall + 1 : 0;
#endif
    }
    last = best;
    tr->generations = gen + 1;
    if (stall >= kConvergenceWindow) { tr->converged_at = gen; break; }
    ++gen;
    pop = evolve(pop, rng);
  }
  for (auto& c : pop) tr->final_genes.insert(tr->final_genes.end(), c.genes, c.genes + 6);
  if (!cluster) return;
  std::vector<Molecule> structs;
  for (auto& e : tr->hof.entries) {
    Molecule c = rec;
    Molecule l = apply_genes(e.genes, lig_pos);
    c.insert(c.end(), l.begin(), l.end());  // run.rs:35-49 combine_molecules
    structs.push_back(std::move(c));
  }
  ClusteringConfig cfg{5.0, 0.60, 0.75, 4};
  tr->clusters = cluster_structures(structs, cfg, nullptr);
  tr->selected = select_models(tr->hof.entries, tr->clusters);
}

Molecule from_records(const RecordAtom* r, int n) {
  Molecule m(n);
  for (int i = 0; i < n; ++i) {
    Atom& a = m[i];
    a.x = r[i].x; a.y = r[i].y; a.z = r[i].z;
    a.charge = r[i].charge; a.vdw_radius = r[i].vdw_radius; a.epsilon = r[i].epsilon; a.rmin2 = r[i].rmin2;
    a.occupancy = r[i].occupancy; a.tempfactor = r[i].tempfactor;
    a.serial = r[i].serial; a.resseq = r[i].resseq; a.chainid = r[i].chainid;
    a.altloc = r[i].altloc; a.icode = r[i].icode;
    a.name = std::string(r[i].name, strnlen(r[i].name, 5));
    a.resname = std::string(r[i].resname, strnlen(r[i].resname, 5));
    a.element = std::string(r[i].element, strnlen(r[i].element, 5));
  }
  return m;
}

void to_records(const Molecule& m, RecordAtom* r) {
  for (size_t i = 0; i < m.size(); ++i) {
    const Atom& a = m[i];
    std::memset(&r[i], 0, sizeof(RecordAtom));
    r[i].x = a.x; r[i].y = a.y; r[i].z = a.z;
    r[i].charge = a.charge; r[i].vdw_radius = a.vdw_radius; r[i].epsilon = a.epsilon; r[i].rmin2 = a.rmin2;
    r[i].occupancy = a.occupancy; r[i].tempfactor = a.tempfactor;
    r[i].serial = a.serial; r[i].resseq = a.resseq; r[i].chainid = a.chainid;
    r[i].altloc = a.altloc; r[i].icode = a.icode;
    std::strncpy(r[i].name, a.name.c_str(), 4);
    std::strncpy(r[i].resname, a.resname.c_str(), 4);
    std::strncpy(r[i].element, a.element.c_str(), 4);
  }
}

}  // namespace

// =================================================================== C ABI
// ---------------------------------------------------------------- evaluator
// evaluator.rs: DockQ against a reference ligand.  Contacts are atom pairs
// keyed by (chain, resseq, atom name) on both sides, kept as a Vec in
// receptor-major order; RMSDs sum distances (not squared distances), as the
// reference does (evaluator.rs:88-132).
struct EvContact {
  char ci; int16_t ri; std::string ai; char cj; int16_t rj; std::string aj;
  bool operator==(const EvContact& o) const {
    return ci == o.ci && ri == o.ri && ai == o.ai && cj == o.cj && rj == o.rj && aj == o.aj;
  }
};

// evaluator.rs:147-168
std::vector<EvContact> ev_contacts(const Molecule& r, const Molecule& l) {
  std::vector<EvContact> out;
  for (const Atom& a : r)
    for (const Atom& b : l)
      if (distance(a, b) <= 5.0) out.push_back({a.chainid, a.resseq, a.name, b.chainid, b.resseq, b.name});
  return out;
}

// evaluator.rs:170-187
void ev_interface(const Molecule& r, const Molecule& ref, std::set<int16_t>* ir, std::set<int16_t>* il) {
  for (const Atom& a : r)
    for (const Atom& b : ref)
      if (distance(a, b) <= 5.0) { ir->insert(a.resseq); il->insert(b.resseq); }
}

// structure.rs:319-327
Molecule filter_by_resseq(const Molecule& m, const std::set<int16_t>& keep) {
  Molecule out;
  for (const Atom& a : m) if (keep.count(a.resseq)) out.push_back(a);
  return out;
}

struct OracleEvaluator {  // evaluator.rs:50-66
  Molecule reference, receptor_interface, reference_interface;
  std::set<int16_t> irec, iref;
  std::vector<EvContact> native;
  OracleEvaluator(const Molecule& receptor, const Molecule& ref) : reference(ref) {
    ev_interface(receptor, ref, &irec, &iref);
    receptor_interface = filter_by_resseq(receptor, irec);
    reference_interface = filter_by_resseq(ref, iref);
    native = ev_contacts(receptor_interface, reference_interface);
  }
  static double zip_rmsd(const Molecule& a, const Molecule& b, size_t n) {  // evaluator.rs:88-132
    if (n == 0) return 0.0;
    double sum = 0.0;
    for (size_t i = 0; i < a.size() && i < b.size(); ++i) sum += distance(a[i], b[i]);
    return std::sqrt(sum / static_cast<double>(n));
  }
  double rmsd(const Molecule& lig) const { return zip_rmsd(lig, reference, reference.size()); }
  double irmsd(const Molecule& lig) const {
    return zip_rmsd(filter_by_resseq(lig, iref), reference_interface, reference_interface.size());
  }
  double fnat(const Molecule& lig) const {  // evaluator.rs:134-145
    const auto docked = ev_contacts(receptor_interface, filter_by_resseq(lig, iref));
    size_t common = 0;
    for (const auto& c : docked)
      if (std::find(native.begin(), native.end(), c) != native.end()) ++common;
    if (native.empty()) return 0.0;
    return static_cast<double>(common) / static_cast<double>(native.size());
  }
  void metrics(const Molecule& lig, double* out) const {  // evaluator.rs:68-87
    const double f = fnat(lig), ir = irmsd(lig), r = rmsd(lig);
    const double irs = 1.0 / (1.0 + powi2(ir / 1.5));
    const double lrs = 1.0 / (1.0 + powi2(r / 8.5));
    out[0] = r; out[1] = ir; out[2] = f; out[3] = (f + irs + lrs) / 3.0;
  }
};

extern "C" {

// -- molecules ------------------------------------------------------------
int or_pdb_count_models(const char* text) { return static_cast<int>(read_pdb_text(text).size()); }

// Writes model `m` of `text` as records; returns its atom count (or needed
// size when `out` is null / cap too small).
int or_pdb_model(const char* text, int m, RecordAtom* out, int cap) {
  auto model = read_pdb_text(text);
  if (m < 0 || m >= static_cast<int>(model.size())) return -1;
  const int n = static_cast<int>(model[m].size());
  if (out && cap >= n) to_records(model[m], out);
  return n;
}

void or_position_ligand(const RecordAtom* rec, int nr, const RecordAtom* lig, int nl, RecordAtom* out) {
  to_records(position_ligand(from_records(rec, nr), from_records(lig, nl)), out);
}

void or_apply_genes(const RecordAtom* lig, int nl, const double* genes, RecordAtom* out) {
  to_records(apply_genes(genes, from_records(lig, nl)), out);
}

void or_center_of_mass(const RecordAtom* m, int n, double* out3) {
  center_of_mass(from_records(m, n), &out3[0], &out3[1], &out3[2]);
}

// -- energies on given coordinates (fitness.rs signatures) ------------------
double or_vdw_energy(const RecordAtom* r, int nr, const RecordAtom* l, int nl) { return vdw_energy(from_records(r, nr), from_records(l, nl)); }
double or_elec_energy(const RecordAtom* r, int nr, const RecordAtom* l, int nl) { return elec_energy(from_records(r, nr), from_records(l, nl)); }
double or_desolv_energy(const RecordAtom* r, int nr, const RecordAtom* l, int nl) { return desolv_energy(from_records(r, nr), from_records(l, nl)); }

// Restraints are given as residue pairs (restraints.rs:44-77), resolved on (r, l).
double or_air_energy(const RecordAtom* r, int nr, const RecordAtom* l, int nl, const int32_t* pairs, int np) {
  auto R = from_records(r, nr), L = from_records(l, nl);
  return air_energy(restraints_from_pairs(R, L, pairs, np), R, L);
}
double or_satisfaction_ratio(const RecordAtom* r, int nr, const RecordAtom* l, int nl, const int32_t* pairs, int np) {
  auto R = from_records(r, nr), L = from_records(l, nl);
  return satisfaction_ratio(restraints_from_pairs(R, L, pairs, np), R, L);
}
int or_restraints_resolved(const RecordAtom* r, int nr, const RecordAtom* l, int nl, const int32_t* pairs, int np) {
  return static_cast<int>(restraints_from_pairs(from_records(r, nr), from_records(l, nl), pairs, np).size());
}

// Population::eval_fitness over n poses.  The restraints are resolved once on
// (rec, lig) as run.rs:310 does; `transform` = 0 scores lig as given for
// every entry (score.rs:35-46 scores each model untransformed).
// out: n x {vdw, elec, desolv, air, restraint_penalty, fitness}.
void or_score_poses(const RecordAtom* r, int nr, const RecordAtom* l, int nl, const int32_t* pairs,
                    int np, const double* genes, int n, const double* w, int transform, int nthreads,
                    double* out, int32_t* n_sat) {
  const Molecule R = from_records(r, nr), L = from_records(l, nl);
  const auto rs = restraints_from_pairs(R, L, pairs, np);
  std::atomic<int> next{0};
  auto worker = [&]() {
    for (int i; (i = next.fetch_add(1)) < n;)
      chromosome_fitness(R, L, rs, w, genes + 6 * i, transform != 0, out + 6 * i, n_sat ? n_sat + i : nullptr);
  };
  if (nthreads <= 1) { worker(); return; }
  std::vector<std::thread> th;
  for (int t = 0; t < nthreads; ++t) th.emplace_back(worker);
  for (auto& t : th) t.join();
}

// Desolvation burial counts (fitness.rs:147-153,171-177) for one transformed
// pose: rec_cnt[nr], lig_cnt[nl] (counted for every atom, ASP ignored).
void or_burial_counts(const RecordAtom* r, int nr, const RecordAtom* l, int nl, const double* genes,
                      int32_t* rec_cnt, int32_t* lig_cnt) {
  const Molecule R = from_records(r, nr);
  const Molecule L = genes ? apply_genes(genes, from_records(l, nl)) : from_records(l, nl);
  for (int i = 0; i < nr; ++i) { int c = 0; for (int j = 0; j < nl; ++j) c += distance(R[i], L[j]) < kDesolvCutoff; rec_cnt[i] = c; }
  for (int j = 0; j < nl; ++j) { int c = 0; for (int i = 0; i < nr; ++i) c += distance(R[i], L[j]) < kDesolvCutoff; lig_cnt[j] = c; }
}

// -- clustering --------------------------------------------------------------
// Contacts of one complex (combine_molecules of r and l, run.rs:35-49).
// Returns the contact count; out = 4 int32 per contact (chainA, resA, chainB,
// resB) in sorted order.
int or_contacts(const RecordAtom* m, int n, double cutoff, int32_t* out, int cap) {
  auto s = calculate_contacts(from_records(m, n), cutoff);
  int k = 0;
  for (auto& c : s) {
    if (out && k < cap) {
      out[4 * k] = std::get<0>(c); out[4 * k + 1] = std::get<1>(c);
      out[4 * k + 2] = std::get<2>(c); out[4 * k + 3] = std::get<3>(c);
    }
    ++k;
  }
  return k;
}

// (fcc, fcc_v) of two contact sets given as 4-int32 tuples.
void or_calculate_fcc(const int32_t* a, int na, const int32_t* b, int nb, double* out2) {
  std::set<Contact> x, y;
  for (int i = 0; i < na; ++i) x.insert({static_cast<char>(a[4 * i]), static_cast<int16_t>(a[4 * i + 1]), static_cast<char>(a[4 * i + 2]), static_cast<int16_t>(a[4 * i + 3])});
  for (int i = 0; i < nb; ++i) y.insert({static_cast<char>(b[4 * i]), static_cast<int16_t>(b[4 * i + 1]), static_cast<char>(b[4 * i + 2]), static_cast<int16_t>(b[4 * i + 3])});
  auto f = calculate_fcc(x, y);
  out2[0] = f.first; out2[1] = f.second;
}

// Greedy clustering of a given neighbour matrix (n x n bytes, row i = nbr(i)).
// out_center[k], out_offsets[k..k+1] into out_members; returns cluster count.
int or_cluster_elements(const uint8_t* nb, int n, int min_size, int64_t* out_center,
                        int64_t* out_offsets, int64_t* out_members) {
  std::vector<std::vector<uint8_t>> m(n, std::vector<uint8_t>(n));
  for (int i = 0; i < n; ++i) for (int j = 0; j < n; ++j) m[i][j] = nb[static_cast<size_t>(i) * n + j];
  auto cl = cluster_elements(m, static_cast<size_t>(min_size));
  int64_t off = 0;
  for (size_t k = 0; k < cl.size(); ++k) {
    out_center[k] = cl[k].center; out_offsets[k] = off;
    for (auto v : cl[k].members) out_members[off++] = v;
  }
  out_offsets[cl.size()] = off;
  return static_cast<int>(cl.size());
}

// cluster_structures over poses of one complex: structure i = rec + (lig
// transformed by genes i) (or lig as given when genes is null and n == 1...).
// Also returns every pose's contact count and the ordered contact list.
int or_cluster_poses(const RecordAtom* r, int nr, const RecordAtom* l, int nl, const double* genes,
                     int n, const double* cfg4, int nthreads, int64_t* out_center, int64_t* out_offsets,
                     int64_t* out_members, int32_t* out_set_sizes) {
  const Molecule R = from_records(r, nr), L = from_records(l, nl);
  ClusteringConfig cfg{cfg4[0], cfg4[1], cfg4[2], static_cast<int64_t>(cfg4[3])};
  if (n == 0) return 0;
  std::vector<std::set<Contact>> sets(n);
  std::atomic<int> next{0};
  auto worker = [&]() {
    for (int i; (i = next.fetch_add(1)) < n;) {
      Molecule c = R;
      Molecule t = apply_genes(genes + 6 * i, L);
      c.insert(c.end(), t.begin(), t.end());
      sets[i] = calculate_contacts(c, cfg.contact_distance);
    }
  };
  std::vector<std::thread> th;
  for (int t = 0; t < std::max(1, nthreads); ++t) th.emplace_back(worker);
  for (auto& t : th) t.join();
  std::vector<size_t> sizes(n);
  for (int i = 0; i < n; ++i) { sizes[i] = sets[i].size(); if (out_set_sizes) out_set_sizes[i] = static_cast<int32_t>(sizes[i]); }
  std::vector<uint32_t> common(static_cast<size_t>(n) * n);
  next = 0;
  auto w2 = [&]() {
    for (int i; (i = next.fetch_add(1)) < n;)
      for (int j = 0; j < n; ++j) {
        size_t c = 0;
        if (sets[i].size() <= sets[j].size()) { for (auto& k : sets[i]) c += sets[j].count(k); }
        else { for (auto& k : sets[j]) c += sets[i].count(k); }
        common[static_cast<size_t>(i) * n + j] = static_cast<uint32_t>(c);
      }
  };
  th.clear();
  for (int t = 0; t < std::max(1, nthreads); ++t) th.emplace_back(w2);
  for (auto& t : th) t.join();
  auto nbm = neighbour_matrix(sizes, common, cfg);
  auto cl = cluster_elements(nbm, static_cast<size_t>(cfg.min_cluster_size));
  int64_t off = 0;
  for (size_t k = 0; k < cl.size(); ++k) {
    out_center[k] = cl[k].center; out_offsets[k] = off;
    for (auto v : cl[k].members) out_members[off++] = v;
  }
  out_offsets[cl.size()] = off;
  return static_cast<int>(cl.size());
}

// -- GA --------------------------------------------------------------------
void* or_ga_run(const RecordAtom* r, int nr, const RecordAtom* l_positioned, int nl, const int32_t* pairs,
                int np, const double* w, uint64_t seed, int pop_size, int max_gen, int cluster, int nthreads) {
  const Molecule R = from_records(r, nr), L = from_records(l_positioned, nl);
  const auto rs = restraints_from_pairs(R, L, pairs, np);
  auto* tr = new GaTrace();
  run_ga(R, L, rs, w, seed, pop_size > 0 ? pop_size : kPopulationSize,
         max_gen > 0 ? max_gen : kMaxGenerations, cluster != 0, nthreads, tr);
  return tr;
}
// Debug mode (chromosome.rs:97-108): fitness = -DockQ against a reference
// ligand, energy components zero.
void* or_ga_run_dockq(const RecordAtom* r, int nr, const RecordAtom* l_positioned, int nl, const RecordAtom* ref,
                      int nref, uint64_t seed, int pop_size, int max_gen, int cluster) {
  const Molecule R = from_records(r, nr), L = from_records(l_positioned, nl);
  const OracleEvaluator ev(R, from_records(ref, nref));
  auto* tr = new GaTrace();
  const double w[4] = {0.4, 0.05, 3.4, 1.0};
  run_ga(R, L, {}, w, seed, pop_size > 0 ? pop_size : kPopulationSize, max_gen > 0 ? max_gen : kMaxGenerations,
         cluster != 0, 1, tr, [&](std::vector<Chromo>& pop) {
           for (auto& c : pop) {
             double m[4];
             ev.metrics(apply_genes(c.genes, L), m);
             c.vdw = c.elec = c.desolv = c.air = c.penalty = 0.0;
             c.fitness = -m[3];
           }
         });
  return tr;
}
void or_ga_free(void* h) { delete static_cast<GaTrace*>(h); }
int or_ga_generations(void* h) { return static_cast<GaTrace*>(h)->generations; }
int or_ga_converged_at(void* h) { return static_cast<GaTrace*>(h)->converged_at; }
void or_ga_trace(void* h, double* genes, double* fitness, int32_t* best_idx) {
  auto* t = static_cast<GaTrace*>(h);
  std::memcpy(genes, t->genes.data(), t->genes.size() * sizeof(double));
  std::memcpy(fitness, t->fitness.data(), t->fitness.size() * sizeof(double));
  std::memcpy(best_idx, t->best_idx.data(), t->best_idx.size() * sizeof(int32_t));
}
int or_ga_hof(void* h, double* genes, double* fitness, int32_t* generation) {
  auto* t = static_cast<GaTrace*>(h);
  for (size_t i = 0; i < t->hof.entries.size(); ++i) {
    if (genes) std::memcpy(genes + 6 * i, t->hof.entries[i].genes, 6 * sizeof(double));
    if (fitness) fitness[i] = t->hof.entries[i].fitness;
    if (generation) generation[i] = static_cast<int32_t>(t->hof.entries[i].generation);
  }
  return static_cast<int>(t->hof.entries.size());
}
int or_ga_clusters(void* h, int64_t* center, int64_t* size) {
  auto* t = static_cast<GaTrace*>(h);
  for (size_t i = 0; i < t->clusters.size(); ++i) {
    if (center) center[i] = t->clusters[i].center;
    if (size) size[i] = static_cast<int64_t>(t->clusters[i].members.size());
  }
  return static_cast<int>(t->clusters.size());
}
int or_ga_final_population(void* h, double* genes) {
  const auto* t = static_cast<const GaTrace*>(h);
  if (genes) std::memcpy(genes, t->final_genes.data(), t->final_genes.size() * sizeof(double));
  return static_cast<int>(t->final_genes.size() / 6);
}

int or_ga_selected(void* h, int64_t* idx, int64_t* size) {
  auto* t = static_cast<GaTrace*>(h);
  for (size_t i = 0; i < t->selected.size(); ++i) { idx[i] = t->selected[i].first; size[i] = t->selected[i].second; }
  return static_cast<int>(t->selected.size());
}

// HallOfFame with a given capacity fed explicit entries (run.rs:101-176).
void* or_hof_new(int max_size) { auto* h = new HallOfFame(); h->max_size = static_cast<size_t>(max_size); return h; }
int or_hof_try_add(void* h, const double* genes, double fitness, int gen) { return static_cast<HallOfFame*>(h)->try_add(genes, fitness, static_cast<uint32_t>(gen)); }
int or_hof_entries(void* h, double* fitness) {
  auto* t = static_cast<HallOfFame*>(h);
  for (size_t i = 0; i < t->entries.size(); ++i) if (fitness) fitness[i] = t->entries[i].fitness;
  return static_cast<int>(t->entries.size());
}
double or_hof_angular_difference(double a, double b) { return HallOfFame::angdiff(a, b); }
void or_hof_free(void* h) { delete static_cast<HallOfFame*>(h); }

// Project random stream, for the product GA's stream test.
void or_rng_stream(uint64_t seed, int n, uint64_t* out) { Rng r(seed); for (int i = 0; i < n; ++i) out[i] = r.next(); }
void or_rng_normals(uint64_t seed, int n, double mean, double sd, double* out) {
  Rng r(seed);
  for (int i = 0; i < n; ++i) out[i] = r.normal(mean, sd);
}
// which: 0 the stream's ln, 1 its cos(2 pi u)
void or_stream_math(const double* x, int n, int which, double* out) {
  for (int i = 0; i < n; ++i) out[i] = which == 0 ? stream_ln(x[i]) : stream_cos2pi(x[i]);
}

// One evolve step on given (genes, fitness), for GA operator tests.
void or_evolve(const double* genes, const double* fitness, int n, uint64_t seed, double* out_genes, double* out_fitness) {
  std::vector<Chromo> pop(n);
  for (int i = 0; i < n; ++i) { std::memcpy(pop[i].genes, genes + 6 * i, 6 * sizeof(double)); pop[i].fitness = fitness[i]; }
  Rng rng(seed);
  auto np = evolve(pop, rng);
  for (size_t i = 0; i < np.size(); ++i) { std::memcpy(out_genes + 6 * i, np[i].genes, 6 * sizeof(double)); out_fitness[i] = np[i].fitness; }
}

// One evolve step on a scripted random stream (Rng's test hook): returns the
// number of draws taken, or -1 when the script ran out or a scripted index
// was out of range.  log_kind / log_arg (capacity log_cap) receive each
// draw's kind (0 unit, 1 unit_incl, 2 below, 3 normal) and below's n.
int or_evolve_scripted(const double* genes, const double* fitness, int n, const double* script, int script_len,
                       double* out_genes, double* out_fitness, int32_t* log_kind, double* log_arg, int log_cap) {
  std::vector<Chromo> pop(n);
  for (int i = 0; i < n; ++i) { std::memcpy(pop[i].genes, genes + 6 * i, 6 * sizeof(double)); pop[i].fitness = fitness[i]; }
  Rng rng(0);
  std::vector<std::pair<int, double>> log;
  rng.script = script; rng.script_len = script_len; rng.log = &log;
  auto np = evolve(pop, rng);
  for (size_t i = 0; i < np.size(); ++i) { std::memcpy(out_genes + 6 * i, np[i].genes, 6 * sizeof(double)); out_fitness[i] = np[i].fitness; }
  for (size_t i = 0; i < log.size() && static_cast<int>(i) < log_cap; ++i) { log_kind[i] = log[i].first; log_arg[i] = log[i].second; }
  if (rng.script_error) return -1;
  return static_cast<int>(log.size());
}

// The generation loop of run.rs:361-531 with every individual of generation
// g given fitness fseq[min(g, nseq - 1)] (the random stream of `seed` breeds
// the populations, which do not matter here): returns the number of
// evaluated generations; *converged_at receives the early-stop generation or
// -1.
int or_ga_fitness_sequence(int pop_size, int max_gen, const double* fseq, int nseq, uint64_t seed, int32_t* converged_at) {
  GaTrace tr;
  const Molecule none;
  const std::vector<Restraint> rs;
  int g = 0;
  auto score = [&](std::vector<Chromo>& pop) {
    const double f = fseq[std::min(g, nseq - 1)];
    for (auto& c : pop) c.fitness = f;
    ++g;
  };
  run_ga(none, none, rs, nullptr, seed, pop_size, max_gen, false, 1, &tr, score);
  if (converged_at) *converged_at = tr.converged_at;
  return tr.generations;
}

// Model selection (run.rs:694-733) from Hall of Fame fitness values and a
// cluster list (centre, size) in cluster_structures' order; returns the
// number of models.
int or_select_models(const double* hof_fitness, int n_hof, const int64_t* centers, const int64_t* sizes,
                     int n_clusters, int64_t* out_idx, int64_t* out_size) {
  std::vector<HofEntry> hof(n_hof);
  for (int i = 0; i < n_hof; ++i) { std::memset(hof[i].genes, 0, sizeof hof[i].genes); hof[i].fitness = hof_fitness[i]; hof[i].generation = 0; }
  std::vector<ClusterResult> cl(n_clusters);
  for (int i = 0; i < n_clusters; ++i) { cl[i].center = centers[i]; cl[i].members.assign(static_cast<size_t>(sizes[i]), 0); }
  auto sel = select_models(hof, cl);
  for (size_t i = 0; i < sel.size(); ++i) { out_idx[i] = sel[i].first; out_size[i] = sel[i].second; }
  return static_cast<int>(sel.size());
}

// HallOfFame::add_from_population (run.rs:194-216) on a given population.
void or_hof_add_population(void* h, const double* genes, const double* fitness, int n, int gen) {
  std::vector<Chromo> pop(n);
  for (int i = 0; i < n; ++i) { std::memcpy(pop[i].genes, genes + 6 * i, 6 * sizeof(double)); pop[i].fitness = fitness[i]; }
  static_cast<HallOfFame*>(h)->add_from_population(pop, static_cast<uint32_t>(gen));
}
int or_hof_genes(void* h, double* genes) {
  auto* t = static_cast<HallOfFame*>(h);
  for (size_t i = 0; i < t->entries.size(); ++i) std::memcpy(genes + 6 * i, t->entries[i].genes, 6 * sizeof(double));
  return static_cast<int>(t->entries.size());
}

// Evaluator (evaluator.rs): receptor and reference ligand as given.
void* or_evaluator_new(const RecordAtom* r, int nr, const RecordAtom* ref, int nref) {
  return new OracleEvaluator(from_records(r, nr), from_records(ref, nref));
}
void or_evaluator_free(void* h) { delete static_cast<OracleEvaluator*>(h); }
void or_evaluator_info(void* h, int* n_irec, int* n_iref, int* n_native, int16_t* irec, int16_t* iref) {
  auto* e = static_cast<OracleEvaluator*>(h);
  *n_irec = static_cast<int>(e->irec.size());
  *n_iref = static_cast<int>(e->iref.size());
  *n_native = static_cast<int>(e->native.size());
  if (irec) { int k = 0; for (int16_t v : e->irec) irec[k++] = v; }
  if (iref) { int k = 0; for (int16_t v : e->iref) iref[k++] = v; }
}
// out: rmsd, irmsd, fnat, dockq for each of n models (model i: atoms [i*nl, (i+1)*nl))
void or_calc_metrics(void* h, const RecordAtom* models, int nl, int n, double* out) {
  auto* e = static_cast<OracleEvaluator*>(h);
  for (int i = 0; i < n; ++i) e->metrics(from_records(models + static_cast<size_t>(i) * nl, nl), out + 4 * i);
}
// evaluator.rs:147-168 contact count between two molecules
int or_eval_contacts(const RecordAtom* a, int na, const RecordAtom* b, int nb) {
  return static_cast<int>(ev_contacts(from_records(a, na), from_records(b, nb)).size());
}

}  // extern "C"
