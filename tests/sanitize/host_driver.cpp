// Host-code workout for the AddressSanitizer / UndefinedBehaviorSanitizer
// build of libgdock_hip.so (tests/test_host_sanitizers.py builds both with
// the sanitizers on the host side only and runs this program).  It touches
// no GPU: PDB parsing (every fixture, plus damaged and edge-case text),
// record formatting, ligand positioning, the host pose transform, pose
// records over the host thread pool, the project stream and its jump table,
// and the GA's host logic (tournaments, crossover, mutation, the Hall of Fame
// past its capacity, early stop, the final breed, model selection) over whole
// runs with synthetic fitness.  Any sanitizer report aborts the program; it prints
// "host driver ok" at the end.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../gdock_amd/csrc/ga.h"
#include "../../include/gdock_hip.h"

#define CHECK(c)                                                          \
  do {                                                                    \
    if (!(c)) {                                                           \
      std::fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
      std::exit(2);                                                       \
    }                                                                     \
  } while (0)

static std::string read_file(const char* path) {
  std::ifstream f(path, std::ios::binary);
  std::stringstream ss;
  ss << f.rdbuf();
  return ss.str();
}

// Parse, write back, parse again: the atoms must survive the round trip.
static void pdb_round_trip(const std::string& text, std::vector<gdock_atom>* first_model) {
  gdock_atom* atoms = nullptr;
  int32_t* offs = nullptr;
  int32_t n_models = 0;
  CHECK(gdock_read_pdb(text.data(), text.size(), &atoms, &offs, &n_models) == GDOCK_OK);
  if (n_models > 0) {
    const int32_t n = offs[1] - offs[0];
    size_t len = 0;
    CHECK(gdock_format_pdb(atoms, n, nullptr, 0, &len) == GDOCK_OK);
    std::vector<char> out(len + 1);
    CHECK(gdock_format_pdb(atoms, n, out.data(), out.size(), &len) == GDOCK_OK);
    CHECK(gdock_format_pdb(atoms, n, out.data(), len, &len) == GDOCK_ERR_INVALID_ARG);  // no room for the NUL
    gdock_atom* back = nullptr;
    int32_t* boffs = nullptr;
    int32_t bm = 0;
    CHECK(gdock_read_pdb(out.data(), len, &back, &boffs, &bm) == GDOCK_OK);
    CHECK(bm == 1 && boffs[1] == n);
    for (int32_t i = 0; i < n; ++i) CHECK(back[i].serial == atoms[i].serial && back[i].resseq == atoms[i].resseq);
    if (first_model) first_model->assign(atoms + offs[0], atoms + offs[1]);
    gdock_free(back);
    gdock_free(boffs);
  }
  gdock_free(atoms);
  gdock_free(offs);
}

static void pdb_edge_cases() {
  const char* cases[] = {
      "",
      "ATOM",
      "ATOM      1  N   ALA A   1",                                        // truncated record
      "ATOM      1  N   ALA A   1      11.104   6.134  -6.504  1.00  0.00           N\n"
      "ATOM      2  CA  ALA A   1      11.639   6.071  -5.147\n"             // short line
      "MODEL        2\nATOM      3  C   ALA A   1       1.000   2.000   3.000  1.00  0.00\nENDMDL\n",
      "HETATM    1  O   HOH W   1       0.000   0.000   0.000  1.00  0.00           O\n",
      "ATOM  99999  CB  XYZ Z9999    -999.999-999.999-999.999  1.00  0.00           C\n",  // unknown residue
      "ATOM      1  N   ALA A   1        nan     inf    -inf  1.00  0.00           N\n",
      "\r\n\r\nEND\n",
  };
  for (const char* t : cases) {
    gdock_atom* atoms = nullptr;
    int32_t* offs = nullptr;
    int32_t nm = 0;
    const gdock_status st = gdock_read_pdb(t, std::strlen(t), &atoms, &offs, &nm);
    CHECK(st == GDOCK_OK || st == GDOCK_ERR_INVALID_ARG);
    gdock_free(atoms);
    gdock_free(offs);
  }
  CHECK(gdock_read_pdb(nullptr, 10, nullptr, nullptr, nullptr) == GDOCK_ERR_INVALID_ARG);
}

// Synthetic fitness: a smooth bowl around one pose with a ripple, so that
// the GA improves, stalls and stops early, and offers many distinct poses.
static double fitness_of(const double* g, double ripple) {
  const double t[6] = {1.0, 2.0, 3.0, 4.0, -5.0, 6.0};
  double f = 0.0;
  for (int k = 0; k < 6; ++k) {
    const double d = g[k] - t[k];
    f += d * d;
  }
  return f + ripple * std::sin(7.0 * g[0] + 3.0 * g[4]);
}

static void ga_runs() {
  int overflowed = 0;  // runs whose Hall of Fame reached its capacity
  for (int seed = 1; seed <= 6; ++seed) {
    gdock_ga_params p;
    CHECK(gdock_ga_default_params(&p) == GDOCK_OK);
    p.population_size = seed % 2 ? 150 : 151;  // odd populations breed one child fewer
    p.max_generations = seed <= 2 ? 250 : 60;
    p.early_stopping = seed == 3 ? 0 : 1;
    p.keep_trace = 1;
    p.elitism = seed == 4 ? 0 : 5;
    p.tournament_size = seed == 5 ? 1 : 3;
    gdock::GaJob job(static_cast<uint64_t>(seed), p);
    std::vector<double> out;
    int gens = 0;
    while (!job.done()) {
      const std::vector<double>& g = job.genes();
      const int n = job.population();
      CHECK(static_cast<int>(g.size()) == 6 * n);
      out.assign(static_cast<size_t>(n) * 6, 0.0);
      for (int i = 0; i < n; ++i) {
        const double f = fitness_of(&g[6 * static_cast<size_t>(i)], seed == 2 ? 0.0 : 0.5);
        out[6 * i + 0] = f;      // vdw
        out[6 * i + 4] = 1.0;    // restraint penalty
        out[6 * i + 5] = f;      // fitness
      }
      job.absorb(out.data());
      ++gens;
      CHECK(gens <= p.max_generations);
    }
    CHECK(job.generations() == gens);
    const std::vector<double>& fin = job.final_genes();
    CHECK(!fin.empty() && fin.size() % 6 == 0);
    const auto& hof = job.hof().entries();
    CHECK(!hof.empty() && hof.size() <= 500);
    if (hof.size() == 500) {  // kept sorted by fitness once it has overflowed (run.rs:171-177)
      ++overflowed;
      for (size_t i = 1; i < hof.size(); ++i) CHECK(hof[i - 1].fitness <= hof[i].fitness);
    }
    // model selection: more clusters than models, one cluster, none
    const int64_t h = static_cast<int64_t>(hof.size());
    std::vector<gdock_cluster_result> cl;
    for (int64_t c = 0; c < 8 && c < h; ++c) cl.push_back(gdock_cluster_result{(c * 7) % h, 8 - c, 0});
    job.select_models(cl);
    CHECK(job.selected().size() <= 5);
    job.select_models(std::vector<gdock_cluster_result>(cl.begin(), cl.begin() + 1));
    job.select_models({});
    CHECK(job.selected().size() == static_cast<size_t>(std::min<int64_t>(5, h)));
    CHECK(job.trace_best.size() == static_cast<size_t>(job.generations()));
  }
  CHECK(overflowed >= 1);
}

// Hall of Fame offers decided on first-angle bins equal offers tested
// against every entry: populations crowded at the 0 / 2 pi wrap and at the
// bin edges (multiples of 0.2 rad), offered to a Hall of Fame grown past its
// capacity, against the same offers made one by one through offer().
static void hof_bins() {
  int kept_total = 0;
  for (int trial = 0; trial < 40; ++trial) {
    gdock::HallOfFame binned, plain;
    for (int gen = 0; gen < 80; ++gen) {
      std::vector<gdock::Individual> pop(150);
      gdock::Xoshiro256 rng(static_cast<uint64_t>(7919 * trial + gen));
      for (auto& c : pop) {
        const double u = rng.unit();
        double a;
        if (u < 0.35) a = 6.283185307179586 - 0.25 * rng.unit();   // just below 2 pi
        else if (u < 0.7) a = 0.25 * rng.unit();                    // just above 0
        else a = 0.2 * static_cast<double>(rng.below(32)) + (rng.unit() - 0.5) * 1e-9;  // bin edges
        if (a < 0.0) a = 0.0;
        c.genes[0] = a;
        c.genes[1] = rng.unit() * 0.6;
        c.genes[2] = rng.unit() * 0.6;
        for (int k = 3; k < 6; ++k) c.genes[k] = (rng.unit() - 0.5) * 6.0;
        c.fitness = rng.unit();
      }
      binned.offer_population(pop, static_cast<uint32_t>(gen));
      std::vector<size_t> order(pop.size());
      for (size_t i = 0; i < order.size(); ++i) order[i] = i;
      std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return pop[x].fitness < pop[y].fitness; });
      int added = 0;
      for (size_t i : order) {
        if (added >= 10) break;
        added += plain.offer(pop[i].genes, pop[i].fitness, static_cast<uint32_t>(gen)) ? 1 : 0;
      }
      const auto& a = binned.entries();
      const auto& b = plain.entries();
      CHECK(a.size() == b.size());
      for (size_t i = 0; i < a.size(); ++i) {
        CHECK(std::memcmp(a[i].genes, b[i].genes, sizeof a[i].genes) == 0);
        CHECK(a[i].fitness == b[i].fitness && a[i].generation == b[i].generation);
      }
    }
    kept_total += static_cast<int>(binned.entries().size());
  }
  CHECK(kept_total > 40 * 100);  // the offers did fill the Hall of Fame
}

// The stream's jump table (x^(16 k) mod P, P the characteristic polynomial
// of its transition, found by Berlekamp-Massey): the state it gives for 16 k
// outputs ahead continues the stream exactly as stepping one output at a time.
static void jump_table() {
  using namespace gdock::stream;
  std::vector<uint64_t> tab(4 * static_cast<size_t>(kJumpEntries));
  CHECK(jump_polys(tab.data()));
  for (uint64_t seed : {1ull, 42ull, 0xfeedull}) {
    const Xoshiro256 r(seed);
    Xoshiro256 w = r;
    uint64_t base[256][4];
    for (int i = 0; i < 256; ++i) {
      for (int k = 0; k < 4; ++k) base[i][k] = w.s[k];
      w.next();
    }
    std::vector<uint64_t> seq(16 * static_cast<size_t>(kJumpEntries) + 16);
    w = r;
    for (uint64_t& v : seq) v = w.next();
    for (int k : {0, 1, 2, 15, 16, 17, 255, 256, 1000, 1785, 2047}) {
      uint64_t st[4];
      jump_combine(base, &tab[4 * static_cast<size_t>(k)], st);
      Xoshiro256 j;
      for (int q = 0; q < 4; ++q) j.s[q] = st[q];
      for (int t = 0; t < 16; ++t) CHECK(j.next() == seq[16 * static_cast<size_t>(k) + t]);
    }
  }
}

int main(int argc, char** argv) {
  std::vector<gdock_atom> rec, lig;
  for (int i = 1; i < argc; ++i) {  // the fixtures: receptor first, then the ligand
    std::vector<gdock_atom> m;
    pdb_round_trip(read_file(argv[i]), &m);
    if (i == 1) rec = m;
    if (i == 2) lig = m;
  }
  pdb_edge_cases();
  CHECK(!rec.empty() && !lig.empty());
  CHECK(gdock_position_ligand(rec.data(), static_cast<int32_t>(rec.size()), lig.data(),
                              static_cast<int32_t>(lig.size())) == GDOCK_OK);

  // pose records over the host thread pool, and the host transform
  const int32_t n = 100000;
  std::vector<double> genes(static_cast<size_t>(n) * 6), poses(static_cast<size_t>(n) * 16);
  std::vector<uint64_t> r(static_cast<size_t>(n) * 6);
  gdock_rng_stream(7, static_cast<int32_t>(r.size()), r.data());
  for (size_t i = 0; i < genes.size(); ++i) {
    const double u = static_cast<double>(r[i] >> 11) * 0x1.0p-53;
    genes[i] = (i % 6) < 3 ? u * 6.283185307179586 : (u - 0.5) * 40.0;
  }
  CHECK(gdock_prepare_poses(genes.data(), n, poses.data()) == GDOCK_OK);
  genes[6 * 123 + 2] = NAN;
  CHECK(gdock_prepare_poses(genes.data(), n, poses.data()) == GDOCK_ERR_INVALID_ARG);
  std::vector<gdock_atom> moved(lig.size());
  for (int k = 0; k < 50; ++k)
    CHECK(gdock_apply_genes(lig.data(), static_cast<int32_t>(lig.size()), &genes[6 * k], moved.data()) == GDOCK_OK);

  // the project stream and its own functions
  std::vector<double> nv(4096), xs(4096), ys(4096);
  gdock_rng_normals(11, 4096, 0.0, 1.0, nv.data());
  for (int which = 0; which < 3; ++which) {  // ln and sqrt of positive normals, cos(2 pi x) on [0, 1)
    for (int i = 0; i < 4096; ++i) xs[i] = which == 1 ? std::fmod(i * 0.37, 1.0) : 1e-300 + i * 0.37;
    gdock_stream_math(xs.data(), 4096, which, ys.data());
  }

  ga_runs();
  hof_bins();
  jump_table();
  std::printf("host driver ok (%zu receptor and %zu ligand atoms)\n", rec.size(), lig.size());
  return 0;
}
