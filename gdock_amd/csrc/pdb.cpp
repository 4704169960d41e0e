// PDB text -> parameterized gdock_atom records (the input format feeding the
// scoring path).  Semantics follow structure.rs:187-300 (process_atom_line,
// read_pdb) and toppar.rs:12-662 (OPLS-AA lookups); see the table include.
#include <cctype>
#include <cstdlib>
#include <cstring>
#include <string>
#include <string_view>
#include <unordered_map>
#include <vector>

#include "../../include/gdock_hip.h"
#include "toppar_table.inc"

namespace gdock {
namespace {

struct ToppIndex {
  std::unordered_map<std::string, const ToppType*> type;
  std::unordered_map<std::string, const char*> atom;  // "RES NAME" -> type
  ToppIndex() {
    for (const auto& t : kToppTypes) type.emplace(t.type, &t);
    for (const auto& a : kToppResAtoms) atom.emplace(std::string(a.residue) + " " + a.atom, a.type);
  }
};
const ToppIndex& index() {
  static const ToppIndex idx;
  return idx;
}

std::string_view trim(std::string_view s) {
  while (!s.empty() && std::isspace(static_cast<unsigned char>(s.front()))) s.remove_prefix(1);
  while (!s.empty() && std::isspace(static_cast<unsigned char>(s.back()))) s.remove_suffix(1);
  return s;
}

// Integer field as Rust's str::parse: [+-]digits, range checked.
bool to_int(std::string_view s, long lo, long hi, long* v) {
  if (s.empty()) return false;
  bool neg = false;
  if (s[0] == '+' || s[0] == '-') { neg = s[0] == '-'; s.remove_prefix(1); }
  if (s.empty() || s.size() > 10) return false;
  long x = 0;
  for (char c : s) {
    if (c < '0' || c > '9') return false;
    x = x * 10 + (c - '0');
  }
  x = neg ? -x : x;
  if (x < lo || x > hi) return false;
  *v = x;
  return true;
}

// Decimal floating-point field (Rust's f64 parse on the forms PDB files use).
bool to_f64(std::string_view s, double* v) {
  if (s.empty() || s.size() > 63) return false;
  for (char c : s)
    if (!(std::isdigit(static_cast<unsigned char>(c)) || c == '.' || c == '+' || c == '-' || c == 'e' || c == 'E'))
      return false;
  char buf[64];
  std::memcpy(buf, s.data(), s.size());
  buf[s.size()] = 0;
  char* end = nullptr;
  *v = std::strtod(buf, &end);
  return end == buf + s.size();
}

// Rust's f32 parse of the same forms: correctly rounded straight to f32.
bool to_f32(std::string_view s, float* v) {
  double probe;
  if (!to_f64(s, &probe)) return false;
  char buf[64];
  std::memcpy(buf, s.data(), s.size());
  buf[s.size()] = 0;
  *v = std::strtof(buf, nullptr);
  return true;
}

void copy_field(char* dst, std::string_view s) {
  const size_t n = s.size() < 4 ? s.size() : 4;
  std::memcpy(dst, s.data(), n);
  dst[n] = 0;
}

bool parse_atom_line(std::string_view line, gdock_atom* a) {
  if (line.size() < 54 || line.substr(0, 4) != "ATOM") return false;
  auto col = [&](size_t b, size_t e) -> std::string_view {
    return line.size() >= e ? line.substr(b, e - b) : std::string_view();
  };
  std::memset(a, 0, sizeof(*a));
  long v;
  if (!to_int(trim(col(6, 11)), -2147483648L, 2147483647L, &v)) return false;
  a->serial = static_cast<int32_t>(v);
  const std::string_view name = trim(col(12, 16));
  copy_field(a->name, name);
  a->altloc = col(16, 17).empty() ? ' ' : col(16, 17)[0];
  const std::string_view resname = trim(col(17, 20));
  copy_field(a->resname, resname);
  a->chainid = col(21, 22).empty() ? ' ' : col(21, 22)[0];
  if (!to_int(trim(col(22, 26)), -32768, 32767, &v)) return false;
  a->resseq = static_cast<int16_t>(v);
  a->icode = col(26, 27).empty() ? ' ' : col(26, 27)[0];
  if (!to_f64(trim(col(30, 38)), &a->x) || !to_f64(trim(col(38, 46)), &a->y) ||
      !to_f64(trim(col(46, 54)), &a->z))
    return false;
  if (!to_f32(trim(col(54, 60)), &a->occupancy)) a->occupancy = 1.0f;  // structure.rs:236-240
  if (!to_f32(trim(col(60, 66)), &a->tempfactor)) a->tempfactor = 0.0f;
  std::string_view el = trim(col(76, 78));
  if (el.empty()) {
    a->element[0] = name.empty() ? ' ' : name[0];  // structure.rs:245-247
    a->element[1] = 0;
  } else {
    copy_field(a->element, el);
  }
  const std::string_view e = trim(std::string_view(a->element));
  a->vdw_radius = e == "H" ? 1.2 : e == "C" ? 1.7 : e == "N" ? 1.55 : e == "O" ? 1.52 : 1.0;
  const auto& idx = index();
  auto it = idx.atom.find(std::string(resname) + " " + std::string(name));
  if (it != idx.atom.end()) {
    auto jt = idx.type.find(it->second);
    if (jt != idx.type.end()) {
      const ToppType* t = jt->second;
      a->epsilon = (t->has & 1) ? t->epsilon : 0.0;
      a->rmin2 = (t->has & 2) ? t->rmin2 : 0.0;
      a->charge = (t->has & 4) ? t->charge : 0.0;
    }
  }
  return true;
}

}  // namespace

// read_pdb over text: MODEL starts a fresh molecule (dropping unterminated
// atoms), ENDMDL pushes it; without MODEL records the single molecule is the
// only model (structure.rs:273-300).
void read_pdb_text(std::string_view text, std::vector<gdock_atom>* atoms, std::vector<int32_t>* offsets) {
  atoms->clear();
  offsets->assign(1, 0);
  std::vector<gdock_atom> cur;
  bool has_models = false;
  size_t pos = 0;
  while (pos < text.size()) {
    size_t nl = text.find('\n', pos);
    if (nl == std::string_view::npos) nl = text.size();
    std::string_view line = text.substr(pos, nl - pos);
    pos = nl + 1;
    if (!line.empty() && line.back() == '\r') line.remove_suffix(1);
    gdock_atom a;
    if (line.substr(0, 5) == "MODEL") {
      has_models = true;
      cur.clear();
    } else if (line.substr(0, 6) == "ENDMDL") {
      atoms->insert(atoms->end(), cur.begin(), cur.end());
      offsets->push_back(static_cast<int32_t>(atoms->size()));
      cur.clear();
    } else if (parse_atom_line(line, &a)) {
      cur.push_back(a);
    }
  }
  if (!has_models) {
    atoms->insert(atoms->end(), cur.begin(), cur.end());
    offsets->push_back(static_cast<int32_t>(atoms->size()));
  }
}

}  // namespace gdock
