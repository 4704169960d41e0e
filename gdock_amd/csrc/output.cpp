// Output stage of a docking run: PDB records of posed complexes
// (structure.rs:159-176 to_pdb_string, structure.rs:303-310 write_pdb),
// Chromosome::apply_genes on the host (chromosome.rs:79-83), and the files
// commands::run writes after clustering (run.rs:738-880): model_<i>.pdb,
// ranked_<i>.pdb and metrics.tsv.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/gdock_hip.h"
#include "common.h"
#include "internal.h"

using namespace gdock;

namespace {

std::string field(const char* s, size_t cap) { return std::string(s, strnlen(s, cap)); }

// Rust's {:>w} / {:<w} / {:^w} padding of a string (the odd pad of a
// centred field goes to the right); longer values are not cut.
void pad(std::string& out, const std::string& v, size_t w, char align) {
  const size_t n = v.size() < w ? w - v.size() : 0;
  const size_t left = align == '>' ? n : align == '^' ? n / 2 : 0;
  out.append(left, ' ');
  out += v;
  out.append(n - left, ' ');
}

// Rust's {:.p} of a float: the exact binary value rounded to p decimals,
// ties to even, as glibc's %.*f in the default rounding mode; NaN and the
// infinities are spelled as Rust spells them.
std::string fixed(double v, int p) {
  if (std::isnan(v)) return "NaN";
  if (std::isinf(v)) return v > 0 ? "inf" : "-inf";
  char buf[512];
  std::snprintf(buf, sizeof buf, "%.*f", p, v);
  return buf;
}

char blank(char c) { return c ? c : ' '; }

// structure.rs:159-176:
// "ATOM  {:>5} {:^4}{:>1}{:3} {:>1}{:>4}{:>1}   {:>8.3}{:>8.3}{:>8.3}{:>6.2}{:>6.2}          {:>2}\n"
void pdb_line(std::string& out, const gdock_atom& a) {
  out += "ATOM  ";
  pad(out, std::to_string(a.serial), 5, '>');
  out += ' ';
  pad(out, field(a.name, sizeof a.name), 4, '^');
  out += blank(a.altloc);
  pad(out, field(a.resname, sizeof a.resname), 3, '<');
  out += ' ';
  out += blank(a.chainid);
  pad(out, std::to_string(a.resseq), 4, '>');
  out += blank(a.icode);
  out += "   ";
  pad(out, fixed(a.x, 3), 8, '>');
  pad(out, fixed(a.y, 3), 8, '>');
  pad(out, fixed(a.z, 3), 8, '>');
  pad(out, fixed(static_cast<double>(a.occupancy), 2), 6, '>');
  pad(out, fixed(static_cast<double>(a.tempfactor), 2), 6, '>');
  out += "          ";
  pad(out, field(a.element, sizeof a.element), 2, '>');
  out += '\n';
}

// Molecule::rotate about its centre of mass, then displace (structure.rs:
// 63-131): the same FP64 operation order as the device transform, cos and
// sin of each angle from one glibc sincos call (as gdock_prepare_poses), no
// contraction (-ffp-contract=off).
void apply_genes(const gdock_atom* lig, int32_t n, const double* g, gdock_atom* out) {
  double sx = 0.0, sy = 0.0, sz = 0.0;
  for (int32_t i = 0; i < n; ++i) { sx += lig[i].x; sy += lig[i].y; sz += lig[i].z; }
  const double nn = static_cast<double>(n);
  const double cx = sx / nn, cy = sy / nn, cz = sz / nn;
  double ca, sa, cb, sb, cg, sg;
  ::sincos(g[0], &sa, &ca);
  ::sincos(g[1], &sb, &cb);
  ::sincos(g[2], &sg, &cg);
  const double m01 = (-sa * cg) + ((ca * sb) * sg), m02 = (sa * sg) + ((ca * sb) * cg);
  const double m11 = (ca * cg) + ((sa * sb) * sg), m12 = (-ca * sg) + ((sa * sb) * cg);
  for (int32_t i = 0; i < n; ++i) {
    gdock_atom a = lig[i];
    const double x = a.x - cx, y = a.y - cy, z = a.z - cz;
    const double nx = ((x * ca) * cb + y * m01) + z * m02;
    const double ny = ((x * sa) * cb + y * m11) + z * m12;
    const double nz = ((-x) * sb + (y * cb) * sg) + (z * cb) * cg;
    a.x = (nx + cx) + g[3];
    a.y = (ny + cy) + g[4];
    a.z = (nz + cz) + g[5];
    out[i] = a;
  }
}

bool write_text(const std::filesystem::path& p, const std::string& text) {
  std::ofstream f(p, std::ios::binary | std::ios::trunc);
  if (!f) return false;
  f.write(text.data(), static_cast<std::streamsize>(text.size()));
  return static_cast<bool>(f);
}

// <dir>/<name>.pdb: the complex's receptor followed by its positioned ligand
// moved by genes (run.rs:35-49 combine_molecules, structure.rs:303-310).
bool write_complex(gdock_ctx* ctx, const gdock_complex* cx, const std::filesystem::path& dir, const std::string& name,
                   const double* genes, std::vector<gdock_atom>* scratch) {
  const std::vector<gdock_atom>& rec = cx->rec;
  const std::vector<gdock_atom>& lig = cx->lig;
  scratch->resize(rec.size() + lig.size());
  std::copy(rec.begin(), rec.end(), scratch->begin());
  apply_genes(lig.data(), static_cast<int32_t>(lig.size()), genes, scratch->data() + rec.size());
  std::string text;
  text.reserve(scratch->size() * 80);
  for (const gdock_atom& a : *scratch) pdb_line(text, a);
  if (!write_text(dir / (name + ".pdb"), text)) {
    set_error(ctx, "write outputs: cannot write %s.pdb", name.c_str());
    return false;
  }
  return true;
}

bool make_dir(gdock_ctx* ctx, const char* out_dir) {
  std::error_code ec;
  std::filesystem::create_directories(std::filesystem::path(out_dir), ec);  // run.rs:540-543
  if (ec) set_error(ctx, "write outputs: cannot create %s: %s", out_dir, ec.message().c_str());
  return !ec;
}

}  // namespace

extern "C" {

gdock_status gdock_format_pdb(const gdock_atom* atoms, int32_t n, char* out, size_t cap, size_t* len) {
  if (n < 0 || (n && !atoms) || !len) return GDOCK_ERR_INVALID_ARG;
  std::string s;
  s.reserve(static_cast<size_t>(n) * 80);
  for (int32_t i = 0; i < n; ++i) pdb_line(s, atoms[i]);
  *len = s.size();
  if (!out) return GDOCK_OK;
  if (cap <= s.size()) return GDOCK_ERR_INVALID_ARG;
  std::memcpy(out, s.data(), s.size());
  out[s.size()] = 0;
  return GDOCK_OK;
}

gdock_status gdock_apply_genes(const gdock_atom* lig, int32_t n, const double* genes, gdock_atom* out) {
  if (n < 0 || (n && (!lig || !out)) || !genes) return GDOCK_ERR_INVALID_ARG;
  for (int k = 0; k < 6; ++k)
    if (!std::isfinite(genes[k])) return GDOCK_ERR_INVALID_ARG;
  if (n) apply_genes(lig, n, genes, out);
  return GDOCK_OK;
}

gdock_status gdock_write_run_outputs(gdock_ctx* ctx, gdock_complex* cx, gdock_evaluator* ev, const double* hof_genes,
                                     const double* hof_fitness, int32_t n_hof, const int64_t* selected_hof_index,
                                     const int64_t* selected_cluster_size, int32_t n_selected, const char* out_dir) {
  if (!ctx || !cx || !out_dir || n_hof < 0 || n_selected < 0 || (n_hof && (!hof_genes || !hof_fitness)) ||
      (n_selected && (!selected_hof_index || !selected_cluster_size))) {
    if (ctx) set_error(ctx, "write_run_outputs: bad arguments");
    return GDOCK_ERR_INVALID_ARG;
  }
  for (int32_t i = 0; i < n_selected; ++i)
    if (selected_hof_index[i] < 0 || selected_hof_index[i] >= n_hof) {
      set_error(ctx, "write_run_outputs: selected model %d names Hall of Fame entry %lld of %d", i,
                static_cast<long long>(selected_hof_index[i]), n_hof);
      return GDOCK_ERR_INVALID_ARG;
    }
  const std::filesystem::path dir(out_dir);
  if (!make_dir(ctx, out_dir)) return GDOCK_ERR_INVALID_ARG;
  // run.rs:826-831: Hall of Fame indices by fitness, best first (stable)
  std::vector<int32_t> ranked(n_hof);
  std::iota(ranked.begin(), ranked.end(), 0);
  std::stable_sort(ranked.begin(), ranked.end(),
                   [&](int32_t a, int32_t b) { return hof_fitness[a] < hof_fitness[b]; });
  const int32_t n_ranked = std::min<int32_t>(n_hof, 5);  // NUM_OUTPUT_MODELS, run.rs:62

  // DockQ metrics of every model written, one evaluator call.
  std::vector<double> metrics;
  if (ev) {
    std::vector<double> genes;
    for (int32_t i = 0; i < n_selected; ++i)
      genes.insert(genes.end(), hof_genes + 6 * selected_hof_index[i], hof_genes + 6 * selected_hof_index[i] + 6);
    for (int32_t r = 0; r < n_ranked; ++r)
      genes.insert(genes.end(), hof_genes + 6 * ranked[r], hof_genes + 6 * ranked[r] + 6);
    const int32_t n = n_selected + n_ranked;
    metrics.resize(static_cast<size_t>(n) * 4);
    if (n) {
      const gdock_status st = gdock_calc_metrics(ctx, ev, genes.data(), n, metrics.data());
      if (st != GDOCK_OK) return st;
    }
  }

  std::vector<gdock_atom> complex;
  auto write_model = [&](const std::string& name, const double* g) { return write_complex(ctx, cx, dir, name, g, &complex); };
  // metrics row: model, cluster size or '-', score, then DockQ, L-RMSD, iRMSD, fnat with an evaluator
  auto row = [&](const std::string& name, const std::string& size, double fitness, int32_t m) {
    std::string r = name + "\t" + size + "\t" + fixed(fitness, 4);
    if (ev) {
      const double* q = &metrics[4 * static_cast<size_t>(m)];  // {rmsd, irmsd, fnat, dockq}
      r += "\t" + fixed(q[3], 4) + "\t" + fixed(q[0], 4) + "\t" + fixed(q[1], 4) + "\t" + fixed(q[2], 4);
    }
    return r + "\n";
  };
  std::string tsv = ev ? "model\tcluster_size\tscore\tdockq\trmsd\tirmsd\tfnat\n" : "model\tcluster_size\tscore\n";
  for (int32_t i = 0; i < n_selected; ++i) {  // run.rs:754-817
    const int64_t h = selected_hof_index[i];
    const std::string name = "model_" + std::to_string(i + 1);
    if (!write_model(name, hof_genes + 6 * h)) return GDOCK_ERR_INVALID_ARG;
    tsv += row(name, std::to_string(selected_cluster_size[i]), hof_fitness[h], i);
  }
  for (int32_t r = 0; r < n_ranked; ++r) {  // run.rs:833-882
    const std::string name = "ranked_" + std::to_string(r + 1);
    if (!write_model(name, hof_genes + 6 * ranked[r])) return GDOCK_ERR_INVALID_ARG;
    tsv += row(name, "-", hof_fitness[ranked[r]], n_selected + r);
  }
  if (!write_text(dir / "metrics.tsv", tsv)) {
    set_error(ctx, "write_run_outputs: cannot write metrics.tsv");
    return GDOCK_ERR_INVALID_ARG;
  }
  return GDOCK_OK;
}

namespace {
// run.rs:549-637 from the final population's fitness (kOutDoubles per pose,
// the fitness at 5) and, with an evaluator, its metrics: best_by_score.pdb is
// the first pose of least fitness (Iterator::min_by), best_by_dockq.pdb the
// last pose of greatest DockQ (Iterator::max_by), and metrics.tsv both rows.
gdock_status write_best(gdock_ctx* ctx, gdock_complex* cx, gdock_evaluator* ev, const double* genes, int32_t n,
                        const std::vector<double>& out, const std::vector<double>* metrics, const char* out_dir) {
  const std::filesystem::path dir(out_dir);
  if (!make_dir(ctx, out_dir)) return GDOCK_ERR_INVALID_ARG;
  int32_t best = 0;
  for (int32_t i = 1; i < n; ++i)
    if (out[kOutDoubles * i + 5] < out[kOutDoubles * best + 5]) best = i;
  std::vector<gdock_atom> complex;
  if (!write_complex(ctx, cx, dir, "best_by_score", genes + 6 * best, &complex)) return GDOCK_ERR_INVALID_ARG;
  if (!ev) return GDOCK_OK;
  std::vector<double> m(static_cast<size_t>(n) * 4);  // {rmsd, irmsd, fnat, dockq}
  if (metrics) {
    m = *metrics;
  } else {
    const gdock_status st = gdock_calc_metrics(ctx, ev, genes, n, m.data());
    if (st != GDOCK_OK) return st;
  }
  int32_t top = 0;
  for (int32_t i = 1; i < n; ++i)
    if (m[4 * i + 3] >= m[4 * top + 3]) top = i;
  if (!write_complex(ctx, cx, dir, "best_by_dockq", genes + 6 * top, &complex)) return GDOCK_ERR_INVALID_ARG;
  auto row = [&](const char* name, int32_t i) {
    const double* q = &m[4 * static_cast<size_t>(i)];
    return std::string(name) + "\t" + fixed(q[3], 4) + "\t" + fixed(q[0], 4) + "\t" + fixed(q[1], 4) + "\t" +
           fixed(q[2], 4) + "\t" + fixed(out[kOutDoubles * i + 5], 4) + "\n";
  };
  const std::string tsv = "model\tdockq\trmsd\tirmsd\tfnat\tscore\n" + row("best_by_score", best) + row("best_by_dockq", top);
  if (!write_text(dir / "metrics.tsv", tsv)) {
    set_error(ctx, "write_best_outputs: cannot write metrics.tsv");
    return GDOCK_ERR_INVALID_ARG;
  }
  return GDOCK_OK;
}
}  // namespace

gdock_status gdock_write_best_outputs(gdock_ctx* ctx, gdock_complex* cx, gdock_evaluator* ev, const gdock_weights* w,
                                      const double* genes, int32_t n, const char* out_dir) {
  if (!ctx || !cx || !w || !out_dir || n <= 0 || !genes) {
    if (ctx) set_error(ctx, "write_best_outputs: bad arguments");
    return GDOCK_ERR_INVALID_ARG;
  }
  // run.rs:553-565: the final population scored with the energy weights
  std::vector<double> out(static_cast<size_t>(n) * kOutDoubles);
  const gdock_status st = gdock_score_poses(ctx, cx, genes, n, w, out.data(), nullptr);
  if (st != GDOCK_OK) return st;
  return write_best(ctx, cx, ev, genes, n, out, nullptr, out_dir);
}

gdock_status gdock_write_best_outputs_dockq(gdock_ctx* ctx, gdock_complex* cx, gdock_evaluator* ev,
                                            const double* genes, int32_t n, const char* out_dir) {
  if (!ctx || !cx || !ev || !out_dir || n <= 0 || !genes) {
    if (ctx) set_error(ctx, "write_best_outputs_dockq: bad arguments");
    return GDOCK_ERR_INVALID_ARG;
  }
  // debug mode (chromosome.rs:97-108): the fitness of a pose is -DockQ and
  // its energies are zero
  std::vector<double> m(static_cast<size_t>(n) * 4);
  const gdock_status st = gdock_calc_metrics(ctx, ev, genes, n, m.data());
  if (st != GDOCK_OK) return st;
  std::vector<double> out(static_cast<size_t>(n) * kOutDoubles, 0.0);
  for (int32_t i = 0; i < n; ++i) out[kOutDoubles * static_cast<size_t>(i) + 5] = -m[4 * static_cast<size_t>(i) + 3];
  return write_best(ctx, cx, ev, genes, n, out, &m, out_dir);
}

}  // extern "C"
