// Genetic-algorithm driver (host side) of a docking job.  Population scoring
// runs on the GPU; selection, crossover, mutation, the Hall of Fame and model
// selection are cheap sequential host work.
#pragma once
#include <cstdint>
#include <vector>

#include "../../include/gdock_hip.h"
#include "stream.h"

namespace gdock {

using Xoshiro256 = stream::Xoshiro256;  // the project stream (stream.h)

struct Individual {
  double genes[6];
  double fitness = 0.0, vdw = 0.0, elec = 0.0, desolv = 0.0, air = 0.0, penalty = 0.0;
};

struct HofEntry {
  double genes[6];
  double fitness;
  uint32_t generation;
};

// run.rs:77-216: unique poses (angles within 0.2 rad, shifts within 2 A count
// as the same pose), the 10 best of each generation offered, at most 500 kept.
class HallOfFame {
 public:
  bool offer(const double* genes, double fitness, uint32_t generation);
  void offer_population(const std::vector<Individual>& pop, uint32_t generation);
  // Appends the pose (then trims to the capacity), without the similarity test.
  bool keep(const double* genes, double fitness, uint32_t generation);
  const std::vector<HofEntry>& entries() const { return entries_; }
  void assign(std::vector<HofEntry> e) { entries_ = std::move(e); }

 private:
  std::vector<HofEntry> entries_;
};

struct GaSelected { int64_t hof_index, cluster_size; };

// One docking job's GA, advanced one generation at a time: genes() are the
// poses to score; absorb() takes their six outputs per pose (vdw, elec,
// desolv, air, restraint_penalty, fitness), updates the Hall of Fame and the
// convergence test, and breeds the next generation unless the job is done.
class GaJob {
 public:
  GaJob(uint64_t seed, const gdock_ga_params& p);
  const std::vector<double>& genes() const { return genes_; }
  int population() const { return static_cast<int>(pop_.size()); }
  void absorb(const double* out);
  bool done() const { return done_; }
  int generations() const { return generations_; }
  int converged_at() const { return converged_at_; }
  const HallOfFame& hof() const { return hof_; }
  // run.rs:694-733: cluster centres sorted by fitness, then the best-scored
  // remaining entries, up to five models.
  void select_models(const std::vector<gdock_cluster_result>& clusters);
  // The population the generation loop leaves behind (run.rs:381-531): the
  // last one scored after an early stop; after the last generation the loop
  // has evolved once more (run.rs:530), so that unscored population.  Breeds
  // it on first call once the job is done.
  const std::vector<double>& final_genes();
  const std::vector<GaSelected>& selected() const { return selected_; }
  // The stream's state after the initial population was drawn (a device GA
  // continues from it), and the end state of a job the device GA ran.
  const uint64_t* rng_state() const { return rng_.s; }
  void load_device_run(int generations, int converged_at, std::vector<HofEntry> hof, std::vector<double> final_genes,
                       std::vector<double> trace_g, std::vector<double> trace_f, std::vector<int32_t> trace_b);
  // per evaluated generation: genes (pop x 6), fitness (pop), best index
  std::vector<double> trace_genes, trace_fitness;
  std::vector<int32_t> trace_best;

 private:
  void breed();
  void refresh_genes();
  gdock_ga_params p_;
  Xoshiro256 rng_;
  std::vector<Individual> pop_;
  std::vector<double> genes_;
  HallOfFame hof_;
  std::vector<GaSelected> selected_;
  int generation_ = 0, generations_ = 0, converged_at_ = -1, stall_ = 0;
  double last_best_ = 1.7976931348623157e308;
  bool done_ = false, final_bred_ = false;
};

gdock_status validate_ga_params(gdock_ctx* ctx, const gdock_ga_params* p);

// Clusters the job's Hall of Fame on the device (run.rs:665-693) and selects
// its output models.
gdock_status finish_job(gdock_ctx* ctx, gdock_complex* cx, GaJob* job, std::vector<gdock_cluster_result>* clusters);

}  // namespace gdock

// A finished job: its GA state and the FCC clusters of its Hall of Fame.
struct gdock_ga_result {
  gdock::GaJob job;
  std::vector<gdock_cluster_result> clusters;
  gdock_ga_result(uint64_t seed, const gdock_ga_params& p) : job(seed, p) {}
};
