// GA driver of a docking job: the reference's generation loop (run.rs), its
// operators (population.rs, ga.rs, chromosome.rs) and Hall of Fame, on the
// host, with every population scored on the GPU.
#include "ga.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "common.h"
#include "ga_device.h"
#include "internal.h"

namespace gdock {

namespace {
constexpr double kTwoPi = 2.0 * kPI;
constexpr double kHofRotation = 0.2;     // run.rs:139-158
constexpr double kHofTranslation = 2.0;
constexpr size_t kHofCapacity = 500;     // run.rs:84-99
constexpr int kHofPerGeneration = 10;    // run.rs:194-216
constexpr size_t kOutputModels = 5;      // run.rs:61

double rem_euclid(double a, double b) {  // Rust f64::rem_euclid
  const double r = std::fmod(a, b);
  return r < 0.0 ? r + std::fabs(b) : r;
}

double angular_difference(double a, double b) {  // run.rs:160-169
  const double d = std::fabs(a - b);
  return d > kPI ? kTwoPi - d : d;
}

bool by_fitness(const Individual& a, const Individual& b) { return a.fitness < b.fitness; }

// run.rs:129-158: a pose within 0.2 rad per angle and 2 A per shift of a kept one.
bool similar(const double* g, const HofEntry& e) {
  bool sim = true;
  for (int i = 0; i < 3 && sim; ++i) sim = angular_difference(g[i], e.genes[i]) <= kHofRotation;
  for (int i = 3; i < 6 && sim; ++i) sim = std::fabs(g[i] - e.genes[i]) <= kHofTranslation;
  return sim;
}

// Entries by their first angle, in bins of 0.2 rad over [0, 2 pi] (the last
// one shorter): an entry within 0.2 rad of a pose's first angle, wrapping at
// 2 pi, lies at most two bins from the pose's, so "similar to some entry" is
// decided on five bins.  Angles outside [0, 2 pi] (the GA never makes them)
// are kept aside and always tested.
constexpr int kHofBins = 32;
int hof_bin(double a) {
  if (!(a >= 0.0 && a <= kTwoPi)) return -1;
  return std::min(static_cast<int>(a / kHofRotation), kHofBins - 1);
}

struct HofIndex {
  std::vector<int> bin[kHofBins], aside;
  void add(const HofEntry& e, int i) {
    const int b = hof_bin(e.genes[0]);
    (b < 0 ? aside : bin[b]).push_back(i);
  }
  void build(const std::vector<HofEntry>& entries) {
    for (auto& v : bin) v.clear();
    aside.clear();
    for (size_t i = 0; i < entries.size(); ++i) add(entries[i], static_cast<int>(i));
  }
  bool any_similar(const double* g, const std::vector<HofEntry>& entries) const {
    const int b = hof_bin(g[0]);
    if (b < 0) {
      for (const HofEntry& e : entries)
        if (similar(g, e)) return true;
      return false;
    }
    for (int d = -2; d <= 2; ++d)
      for (int i : bin[(b + d + kHofBins) % kHofBins])
        if (similar(g, entries[i])) return true;
    for (int i : aside)
      if (similar(g, entries[i])) return true;
    return false;
  }
};
}  // namespace

// ---- Hall of Fame -------------------------------------------------------------
bool HallOfFame::offer(const double* g, double fitness, uint32_t generation) {
  for (const HofEntry& e : entries_)  // run.rs:129-158: similar to a kept pose?
    if (similar(g, e)) return false;
  return keep(g, fitness, generation);
}

bool HallOfFame::keep(const double* g, double fitness, uint32_t generation) {
  HofEntry e;
  std::memcpy(e.genes, g, sizeof e.genes);
  e.fitness = fitness;
  e.generation = generation;
  entries_.push_back(e);
  if (entries_.size() > kHofCapacity) {  // run.rs:171-177: keep the best
    std::stable_sort(entries_.begin(), entries_.end(),
                     [](const HofEntry& a, const HofEntry& b) { return a.fitness < b.fitness; });
    entries_.resize(kHofCapacity);
  }
  return true;
}

void HallOfFame::offer_population(const std::vector<Individual>& pop, uint32_t generation) {
  std::vector<size_t> order(pop.size());
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return pop[a].fitness < pop[b].fitness; });
  HofIndex index;
  index.build(entries_);
  int added = 0;
  for (size_t i : order) {
    if (added >= kHofPerGeneration) break;
    if (index.any_similar(pop[i].genes, entries_)) continue;
    const size_t before = entries_.size();
    keep(pop[i].genes, pop[i].fitness, generation);
    if (entries_.size() > before) index.add(entries_.back(), static_cast<int>(before));
    else index.build(entries_);  // trimmed: sorted anew
    ++added;
  }
}

// ---- GA job ---------------------------------------------------------------------
GaJob::GaJob(uint64_t seed, const gdock_ga_params& p) : p_(p), rng_(seed) {
  pop_.resize(static_cast<size_t>(p.population_size));
  for (Individual& c : pop_) {  // chromosome.rs:27-54
    for (int i = 0; i < 3; ++i) c.genes[i] = rng_.range_closed(0.0, kTwoPi);
    for (int i = 3; i < 6; ++i) c.genes[i] = rng_.range_closed(-p.max_displacement, p.max_displacement);
  }
  refresh_genes();
}

void GaJob::refresh_genes() {
  genes_.resize(pop_.size() * 6);
  for (size_t i = 0; i < pop_.size(); ++i) std::memcpy(&genes_[6 * i], pop_[i].genes, 6 * sizeof(double));
}

void GaJob::absorb(const double* out) {
  if (done_) return;
  for (size_t i = 0; i < pop_.size(); ++i) {
    Individual& c = pop_[i];
    c.vdw = out[6 * i];
    c.elec = out[6 * i + 1];
    c.desolv = out[6 * i + 2];
    c.air = out[6 * i + 3];
    c.penalty = out[6 * i + 4];
    c.fitness = out[6 * i + 5];
  }
  if (p_.keep_trace) {
    trace_genes.insert(trace_genes.end(), genes_.begin(), genes_.end());
    for (const Individual& c : pop_) trace_fitness.push_back(c.fitness);
  }
  hof_.offer_population(pop_, static_cast<uint32_t>(generation_));
  size_t best = 0;  // the first individual of least fitness
  for (size_t i = 1; i < pop_.size(); ++i)
    if (pop_[i].fitness < pop_[best].fitness) best = i;
  if (p_.keep_trace) trace_best.push_back(static_cast<int32_t>(best));
  const double f = pop_.empty() ? 0.0 : pop_[best].fitness;
  if (generation_ > 0) {  // run.rs:480-531: relative improvement of the best
    const double improvement = (last_best_ - f) / std::fabs(last_best_);
    stall_ = improvement < p_.convergence_threshold ? stall_ + 1 : 0;
  }
  last_best_ = f;
  generations_ = generation_ + 1;
  if (p_.early_stopping && stall_ >= p_.convergence_window) {
    converged_at_ = generation_;
    done_ = true;
    return;
  }
  ++generation_;
  if (generation_ >= p_.max_generations) {
    done_ = true;
    return;
  }
  breed();
  refresh_genes();
}

// population.rs:74-158 with ga.rs:7-39 and chromosome.rs:56-77.
const std::vector<double>& GaJob::final_genes() {
  if (done_ && !final_bred_ && converged_at_ < 0 && !pop_.empty()) {
    breed();
    refresh_genes();
  }
  final_bred_ = true;
  return genes_;
}

void GaJob::load_device_run(int generations, int converged_at, std::vector<HofEntry> hof,
                            std::vector<double> final_genes, std::vector<double> trace_g, std::vector<double> trace_f,
                            std::vector<int32_t> trace_b) {
  generations_ = generations;
  generation_ = generations;
  converged_at_ = converged_at;
  hof_.assign(std::move(hof));
  genes_ = std::move(final_genes);
  pop_.resize(genes_.size() / 6);
  for (size_t i = 0; i < pop_.size(); ++i) std::memcpy(pop_[i].genes, &genes_[6 * i], 6 * sizeof(double));
  trace_genes = std::move(trace_g);
  trace_fitness = std::move(trace_f);
  trace_best = std::move(trace_b);
  done_ = true;
  final_bred_ = true;  // the device bred the final population already
}

void GaJob::breed() {
  const size_t n = pop_.size();
  std::vector<Individual> elite = pop_;
  std::stable_sort(elite.begin(), elite.end(), by_fitness);
  elite.resize(std::min<size_t>(static_cast<size_t>(p_.elitism), elite.size()));

  std::vector<Individual> chosen;  // tournament selection, first minimum wins
  chosen.reserve(n);
  for (size_t k = 0; k < n; ++k) {
    const Individual* best = nullptr;
    for (int t = 0; t < p_.tournament_size; ++t) {
      const Individual* c = &pop_[rng_.below(n)];
      if (!best || c->fitness < best->fitness) best = c;
    }
    chosen.push_back(*best);
  }
  for (size_t i = chosen.size(); i > 1; --i) std::swap(chosen[i - 1], chosen[rng_.below(i)]);  // Fisher-Yates

  std::vector<Individual> next;
  next.reserve(n);
  for (size_t i = 0; i + 1 < chosen.size(); i += 2) {  // uniform crossover of neighbours
    Individual a = chosen[i], b = chosen[i + 1];
    for (int g = 0; g < 6; ++g)
      if (rng_.unit() <= p_.crossover_rate) std::swap(a.genes[g], b.genes[g]);
    next.push_back(a);
    next.push_back(b);
  }
  for (Individual& c : next)
    for (int g = 0; g < 6; ++g) {
      if (!(rng_.unit() <= p_.mutation_rate)) continue;
      if (g < 3)
        c.genes[g] = rem_euclid(c.genes[g] + rng_.normal(0.0, 0.174), kTwoPi);
      else
        c.genes[g] = std::min(std::max(c.genes[g] + rng_.normal(0.0, 1.0), -p_.max_displacement),
                              p_.max_displacement);
    }
  // the elite replace the children that inherited the worst fitness
  std::stable_sort(next.begin(), next.end(), [](const Individual& a, const Individual& b) { return b.fitness < a.fitness; });
  for (size_t i = 0; i < elite.size() && i < next.size(); ++i) next[i] = elite[i];
  pop_ = std::move(next);
}

void GaJob::select_models(const std::vector<gdock_cluster_result>& clusters) {
  selected_.clear();
  const auto& h = hof_.entries();
  struct Centre { int64_t idx; double fitness; int64_t size; };
  std::vector<Centre> centres;
  for (const gdock_cluster_result& c : clusters) centres.push_back({c.center_idx, h[c.center_idx].fitness, c.size});
  std::stable_sort(centres.begin(), centres.end(), [](const Centre& a, const Centre& b) { return a.fitness < b.fitness; });
  for (size_t i = 0; i < centres.size() && selected_.size() < kOutputModels; ++i)
    selected_.push_back({centres[i].idx, centres[i].size});
  if (selected_.size() < kOutputModels) {
    std::vector<int64_t> order(h.size());
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return h[a].fitness < h[b].fitness; });
    for (int64_t idx : order) {
      if (selected_.size() >= kOutputModels) break;
      bool taken = false;
      for (const GaSelected& s : selected_) taken = taken || s.hof_index == idx;
      if (!taken) selected_.push_back({idx, 1});
    }
  }
}

}  // namespace gdock

using namespace gdock;

namespace gdock {

gdock_status validate_ga_params(gdock_ctx* ctx, const gdock_ga_params* p) {
  if (!p || p->population_size < 2 || p->population_size > (1 << 20) || p->max_generations < 1 || p->elitism < 0 ||
      p->tournament_size < 1 || p->convergence_window < 1 || !(p->max_displacement >= 0.0) ||
      !std::isfinite(p->max_displacement) || !std::isfinite(p->mutation_rate) || !std::isfinite(p->crossover_rate)) {
    if (ctx) set_error(ctx, "ga: invalid parameters");
    return GDOCK_ERR_INVALID_ARG;
  }
  return GDOCK_OK;
}

gdock_status finish_job(gdock_ctx* ctx, gdock_complex* cx, GaJob* job, std::vector<gdock_cluster_result>* clusters) {
  const auto& h = job->hof().entries();
  const int32_t n = static_cast<int32_t>(h.size());
  std::vector<double> genes(static_cast<size_t>(n) * 6);
  for (int32_t i = 0; i < n; ++i) std::memcpy(&genes[6 * i], h[i].genes, 6 * sizeof(double));
  const gdock_clustering_config cfg{5.0, 0.60, 0.75, 4};  // ClusteringConfig::default()
  std::vector<gdock_cluster_result> res(std::max(n, 1));
  std::vector<int64_t> members(std::max(n, 1));
  int32_t nc = 0;
  if (n > 0) {
    const gdock_status st = gdock_cluster_structures(ctx, cx, genes.data(), n, &cfg, res.data(), members.data(), &nc);
    if (st != GDOCK_OK) return st;
  }
  clusters->assign(res.begin(), res.begin() + nc);
  job->select_models(*clusters);
  return GDOCK_OK;
}

}  // namespace gdock

extern "C" {

gdock_status gdock_ga_default_params(gdock_ga_params* p) {
  if (!p) return GDOCK_ERR_INVALID_ARG;
  *p = gdock_ga_params{};
  p->population_size = 150;      // constants.rs:2-16
  p->max_generations = 250;
  p->elitism = 5;
  p->tournament_size = 3;
  p->convergence_window = 10;
  p->early_stopping = 1;
  p->cluster = 1;
  p->keep_trace = 0;
  p->mutation_rate = 0.1;
  p->crossover_rate = 0.6;
  p->convergence_threshold = 0.001;
  p->max_displacement = 50.0;
  return GDOCK_OK;
}

namespace {
// Scores one generation of the host driver: genes up, pose records made on
// the device (as the device GA makes them, so both drivers score the same
// records), the plain scoring launch filled to kGaFillItems workgroups (as
// the device GA launches it), results down.
gdock_status score_generation(gdock_ctx* ctx, gdock_complex* cx, const double* genes, int32_t n,
                              const gdock_weights* w, double* out) {
  hipError_t e = hipSetDevice(ctx->device);
  if (e != hipSuccess) return hip_fail(ctx, e, "ga: set device");
  Layout L;
  const size_t o_g = L.add<double>(static_cast<size_t>(n) * 6), o_p = L.add<double>(static_cast<size_t>(n) * kPoseDoubles),
               o_o = L.add<double>(static_cast<size_t>(n) * kOutDoubles),
               o_s = L.add<uint8_t>(plain_scratch_bytes(cx->dev, n, kGaFillItems));
  if (!ctx->staging.reserve(L.size, &e)) return hip_fail(ctx, e, "ga: staging");
  char* m = static_cast<char*>(ctx->staging.ptr);
  const Weights ww{w->vdw, w->elec, w->desolv, w->air};
  e = hipMemcpyAsync(m + o_g, genes, static_cast<size_t>(n) * 6 * 8, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess)
    e = launch_ga_pose_records(reinterpret_cast<double*>(m + o_g), n, reinterpret_cast<double*>(m + o_p), ctx->stream);
  if (e == hipSuccess)
    e = launch_score(cx->dev, cx->d_dev, reinterpret_cast<double*>(m + o_p), nullptr, n, ww,
                     reinterpret_cast<double*>(m + o_o), nullptr, m + o_s, ctx->stream, 0, nullptr, kGaFillItems);
  if (e == hipSuccess)
    e = hipMemcpyAsync(out, m + o_o, static_cast<size_t>(n) * kOutDoubles * 8, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  return e == hipSuccess ? GDOCK_OK : hip_fail(ctx, e, "ga: generation");
}
}  // namespace

gdock_status gdock_ga_run(gdock_ctx* ctx, gdock_complex* cx, const gdock_weights* w, uint64_t seed,
                          const gdock_ga_params* params, gdock_ga_result** out) {
  if (!ctx || !cx || !w || !out) {
    if (ctx) set_error(ctx, "ga_run: invalid arguments");
    return GDOCK_ERR_INVALID_ARG;
  }
  gdock_ga_params p;
  if (params) p = *params; else gdock_ga_default_params(&p);
  gdock_status st = validate_ga_params(ctx, &p);
  if (st != GDOCK_OK) return st;
  auto* r = new gdock_ga_result(seed, p);
  GaJob& job = r->job;
  std::vector<double> res(static_cast<size_t>(job.population()) * 6);
  while (!job.done()) {  // run.rs:361-531, one plain scoring launch per generation
    for (double g : job.genes())
      if (!std::isfinite(g)) { delete r; set_error(ctx, "ga_run: non-finite gene"); return GDOCK_ERR_INVALID_ARG; }
    st = score_generation(ctx, cx, job.genes().data(), job.population(), w, res.data());
    if (st != GDOCK_OK) { delete r; return st; }
    job.absorb(res.data());
  }
  if (p.cluster) {
    st = finish_job(ctx, cx, &job, &r->clusters);
    if (st != GDOCK_OK) { delete r; return st; }
  }
  *out = r;
  return GDOCK_OK;
}

gdock_status gdock_ga_run_dockq(gdock_ctx* ctx, gdock_complex* cx, gdock_evaluator* ev, uint64_t seed,
                                const gdock_ga_params* params, gdock_ga_result** out) {
  if (!ctx || !cx || !ev || !out) {
    if (ctx) set_error(ctx, "ga_run_dockq: invalid arguments");
    return GDOCK_ERR_INVALID_ARG;
  }
  gdock_ga_params p;
  if (params) p = *params; else gdock_ga_default_params(&p);
  gdock_status st = validate_ga_params(ctx, &p);
  if (st != GDOCK_OK) return st;
  auto* r = new gdock_ga_result(seed, p);
  GaJob& job = r->job;
  const size_t pop = static_cast<size_t>(job.population());
  std::vector<double> m(pop * 4), res(pop * 6, 0.0);
  while (!job.done()) {  // chromosome.rs:97-108: fitness = -DockQ, energy terms zero
    st = gdock_calc_metrics(ctx, ev, job.genes().data(), job.population(), m.data());
    if (st != GDOCK_OK) { delete r; return st; }
    for (size_t i = 0; i < pop; ++i) res[6 * i + 5] = -m[4 * i + 3];
    job.absorb(res.data());
  }
  if (p.cluster) {
    st = finish_job(ctx, cx, &job, &r->clusters);
    if (st != GDOCK_OK) { delete r; return st; }
  }
  *out = r;
  return GDOCK_OK;
}

void gdock_ga_free(gdock_ga_result* r) { delete r; }

gdock_status gdock_ga_info(const gdock_ga_result* r, int32_t* generations, int32_t* converged_at,
                           int32_t* hof_size, int32_t* n_clusters, int32_t* n_selected) {
  if (!r) return GDOCK_ERR_INVALID_ARG;
  if (generations) *generations = r->job.generations();
  if (converged_at) *converged_at = r->job.converged_at();
  if (hof_size) *hof_size = static_cast<int32_t>(r->job.hof().entries().size());
  if (n_clusters) *n_clusters = static_cast<int32_t>(r->clusters.size());
  if (n_selected) *n_selected = static_cast<int32_t>(r->job.selected().size());
  return GDOCK_OK;
}

int32_t gdock_ga_final_population(gdock_ga_result* r, double* genes) {
  if (!r || !r->job.done()) return -1;
  const std::vector<double>& g = r->job.final_genes();
  if (genes) std::memcpy(genes, g.data(), g.size() * sizeof(double));
  return static_cast<int32_t>(g.size() / 6);
}

gdock_status gdock_ga_trace(const gdock_ga_result* r, double* genes, double* fitness, int32_t* best_idx) {
  if (!r) return GDOCK_ERR_INVALID_ARG;
  const GaJob& j = r->job;
  if (genes) std::copy(j.trace_genes.begin(), j.trace_genes.end(), genes);
  if (fitness) std::copy(j.trace_fitness.begin(), j.trace_fitness.end(), fitness);
  if (best_idx) std::copy(j.trace_best.begin(), j.trace_best.end(), best_idx);
  return GDOCK_OK;
}

int32_t gdock_ga_hof(const gdock_ga_result* r, double* genes, double* fitness, int32_t* generation) {
  if (!r) return -1;
  const auto& h = r->job.hof().entries();
  for (size_t i = 0; i < h.size(); ++i) {
    if (genes) std::memcpy(genes + 6 * i, h[i].genes, 6 * sizeof(double));
    if (fitness) fitness[i] = h[i].fitness;
    if (generation) generation[i] = static_cast<int32_t>(h[i].generation);
  }
  return static_cast<int32_t>(h.size());
}

int32_t gdock_ga_clusters(const gdock_ga_result* r, int64_t* center, int64_t* size) {
  if (!r) return -1;
  for (size_t i = 0; i < r->clusters.size(); ++i) {
    if (center) center[i] = r->clusters[i].center_idx;
    if (size) size[i] = r->clusters[i].size;
  }
  return static_cast<int32_t>(r->clusters.size());
}

int32_t gdock_ga_selected(const gdock_ga_result* r, int64_t* hof_index, int64_t* cluster_size) {
  if (!r) return -1;
  const auto& s = r->job.selected();
  for (size_t i = 0; i < s.size(); ++i) {
    if (hof_index) hof_index[i] = s[i].hof_index;
    if (cluster_size) cluster_size[i] = s[i].cluster_size;
  }
  return static_cast<int32_t>(s.size());
}

void gdock_rng_stream(uint64_t seed, int32_t n, uint64_t* out) {
  Xoshiro256 r(seed);
  for (int32_t i = 0; i < n; ++i) out[i] = r.next();
}

void gdock_rng_normals(uint64_t seed, int32_t n, double mean, double sd, double* out) {
  Xoshiro256 r(seed);
  for (int32_t i = 0; i < n; ++i) out[i] = r.normal(mean, sd);
}

void gdock_stream_math(const double* x, int32_t n, int32_t which, double* out) {
  for (int32_t i = 0; i < n; ++i)
    out[i] = which == 0 ? stream::ln(x[i]) : (which == 1 ? stream::cos2pi(x[i]) : stream::sqrt_cr(x[i]));
}

}  // extern "C"
