// Docking server: continuous batching of GA jobs on one GPU.
//
// Jobs are queued by gdock_server_submit and admitted by a worker thread as
// slots and device memory free up.  Each admitted job keeps its complex
// (receptor, positioned ligand, restraints, contact tables) and its GA state
// (population, Hall of Fame, random stream) resident in HBM until it
// finishes.  Every scheduling step scores the current generation of every
// active job in one persistent launch (replayed from a graph captured once at
// fixed capacity), then advances every job's GA: on the host threads (the
// default: a generation's draws are a sequential chain per job, which a host
// core runs faster than one GPU lane), or, with ga_on_device, in one launch of
// the device GA (one workgroup per job), which keeps the population resident
// and writes the next generation's pose records in place, so the host only
// reads one done flag per slot between steps.  Finished jobs are retired
// after clustering their Hall of Fame.  A job's results do
// not depend on which other jobs share its launches: every pose is scored by
// the same kernel code with the same per-shape choice of lanes, and every
// job's GA has its own stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/gdock_hip.h"
#include "common.h"
#include "ga.h"
#include "ga_device.h"
#include "internal.h"

using namespace gdock;

namespace {

enum JobState : int32_t { kQueued = 0, kActive = 1, kFinished = 2, kFailed = 3 };

struct Job {
  int64_t id = 0;
  std::vector<gdock_atom> rec, lig;
  std::vector<int32_t> pairs;
  uint64_t seed = 0;
  gdock_ga_params params{};
  JobState state = kQueued;
  gdock_status error = GDOCK_OK;
  std::string message;
  double t_submit = 0, t_admit = 0, t_finish = 0;
  gdock_complex* cx = nullptr;
  size_t bytes = 0;   // device bytes held while active
  int slot = -1;
  int lanes = 0;
  int parts = 1;
  void* split_mem = nullptr;  // split-pose state (parts > 1)
  std::unique_ptr<gdock_ga_result> result;
  DevComplex dev{};      // the job's launch-table entry, once prepared
  void* ga_mem = nullptr;  // the job's device GA state
  DevGaJob ga{};           // its kernel view (pose records and scores at the slot's rows)
  int n_cur = 0;           // poses of the generation to score next
  bool ready = false;    // prepared: the worker scores it from its next step
  bool in_table = false; // the worker has uploaded its table entry
};

double now_s(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

struct gdock_server {
  gdock_server_config cfg{};
  gdock_ctx* ctx = nullptr;
  std::chrono::steady_clock::time_point t0;
  // device state of the launches
  BatchLaunch L{};
  void* mem = nullptr;
  int2* h_list = nullptr;  // {slot, pose row} per pose of a step (pinned)
  DevComplex* h_table = nullptr;
  double* h_poses = nullptr;  // host GA: the step's pose records (pinned)
  double* h_out = nullptr;    // host GA: the step's scores (pinned)
  DevGaJob* d_ga = nullptr;   // the active jobs' GA views, in step order
  DevGaJob* h_ga = nullptr;   // (pinned staging)
  int32_t* d_done = nullptr;  // per slot: the GA's done flag after the last step
  int32_t* h_done = nullptr;  // (pinned)
  std::vector<int64_t> step_key;  // the {job, slot, poses} list last uploaded
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;
  int max_pop = 0;
  int lds_slots = 0;  // receptor slots the launch's LDS table holds per item
  // jobs
  std::mutex mu;
  std::condition_variable cv;      // worker: new work or stop
  std::condition_variable done_cv; // waiters: a job finished
  std::condition_variable idle_cv; // destroy: the last waiter left
  int waiters = 0;                 // threads inside gdock_server_wait
  bool closing = false;            // destroy has begun: no new waits
  std::deque<Job*> queue;
  std::map<int64_t, std::unique_ptr<Job>> jobs;
  std::vector<Job*> slots;
  size_t bytes_in_use = 0;
  int64_t next_id = 1;
  bool stop = false;
  std::thread worker;
  // Admitted jobs are prepared (ligand positioning, complex tables, lanes,
  // split state) on a thread and stream of their own while scoring goes on.
  gdock_ctx* pctx = nullptr;
  std::deque<Job*> to_prepare;
  std::condition_variable prep_cv;  // preparer: a job to prepare or stop
  bool prep_stop = false;
  std::thread preparer;
  // Jobs whose GA has ended are finished (FCC of the Hall of Fame, model
  // selection, release of their complex) on a thread and stream of their own,
  // while the worker goes on scoring the other jobs.
  gdock_ctx* fctx = nullptr;
  std::deque<Job*> to_finish;
  int n_finishing = 0;             // handed to the finisher and not yet released
  std::condition_variable fin_cv;  // finisher: a job to finish or stop
  bool fin_stop = false;
  std::thread finisher;
  // statistics
  int64_t n_steps = 0, n_poses = 0, n_finished = 0;
  int32_t peak_active = 0;
  double busy_s = 0, kernel_s = 0, admit_s = 0, ga_s = 0, finish_s = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr;

  void run();
  void admit_locked(std::vector<Job*>* fresh);
  bool step();
  // One step's launches and GA advance (jobs prepared and in the table);
  // returns the launch status.
  hipError_t step_host_ga(const std::vector<Job*>& active);
  hipError_t step_device_ga(const std::vector<Job*>& active);
  void finish_loop();
  void prepare_loop();
  // Frees the job's device state, then marks it ended (failed with `st` unless
  // GDOCK_OK) and wakes its waiters; called holding no lock.
  void release(Job* j, gdock_status st = GDOCK_OK, const std::string& message = std::string());
};

// Moves queued jobs into free slots while their device memory fits the
// budget (first come, first served; the head of the queue waits for memory).
void gdock_server::admit_locked(std::vector<Job*>* fresh) {
  while (!queue.empty()) {
    Job* j = queue.front();
    const auto free_slot = std::find(slots.begin(), slots.end(), nullptr);
    if (free_slot == slots.end()) return;
    // Device bytes of a job: admission uses an estimate from the atom counts
    // (256 B per atom and 1 MiB for contact tables and launch state); once
    // the complex exists the larger of the estimate and its arena is charged.
    const size_t est = 256 * (j->rec.size() + j->lig.size()) + (size_t{1} << 20);
    const bool nothing_active =
        n_finishing == 0 && std::all_of(slots.begin(), slots.end(), [](Job* s) { return s == nullptr; });
    if (bytes_in_use + est > static_cast<size_t>(cfg.memory_budget) && !nothing_active) return;
    queue.pop_front();
    j->slot = static_cast<int>(free_slot - slots.begin());
    *free_slot = j;
    j->state = kActive;
    j->t_admit = now_s(t0);
    j->bytes = est;
    bytes_in_use += est;
    fresh->push_back(j);
  }
}

// One scheduling step; returns false when there was nothing to run.
bool gdock_server::step() {
  std::vector<Job*> fresh, active;
  {
    std::unique_lock<std::mutex> lk(mu);
    cv.wait(lk, [&] { return stop || !queue.empty() || std::any_of(slots.begin(), slots.end(), [](Job* s) { return s; }); });
    if (stop) return false;
    admit_locked(&fresh);
  }
  if (!fresh.empty()) {
    {
      std::lock_guard<std::mutex> g(mu);
      to_prepare.insert(to_prepare.end(), fresh.begin(), fresh.end());
    }
    prep_cv.notify_one();
  }
  // The prepared jobs take part from this step: their table entries are
  // uploaded on the worker's stream, ahead of the step's launch.
  {
    std::lock_guard<std::mutex> g(mu);
    for (Job* s : slots)
      if (s && s->state == kActive && s->ready) active.push_back(s);
    peak_active = std::max<int32_t>(peak_active, static_cast<int32_t>(active.size()));
  }
  hipError_t te = hipSuccess;
  for (Job* j : active)
    if (!j->in_table && te == hipSuccess) {
      h_table[j->slot] = j->dev;
      te = hipMemcpyAsync(const_cast<DevComplex*>(L.table) + j->slot, &h_table[j->slot], sizeof(DevComplex),
                          hipMemcpyHostToDevice, ctx->stream);
      j->in_table = true;
    }
  if (active.empty()) {  // queued jobs wait for memory held by finishing ones
    std::unique_lock<std::mutex> lk(mu);
    cv.wait_for(lk, std::chrono::milliseconds(2));
    return true;
  }
  hipError_t e = te;
  if (e == hipSuccess) e = cfg.ga_on_device ? step_device_ga(active) : step_host_ga(active);

  std::vector<Job*> failed, ended;
  const gdock_status launch_st = e != hipSuccess ? hip_fail(ctx, e, "server: generation launch") : GDOCK_OK;
  for (Job* j : active) {
    if (e != hipSuccess) {
      failed.push_back(j);
    } else if (cfg.ga_on_device ? h_done[j->slot] != 0 : j->result->job.done()) {
      ended.push_back(j);
    } else {
      j->n_cur &= ~1;  // the children of an odd population are one fewer (population.rs:74-158)
    }
  }
  {
    // Ended jobs leave their slots (their rows are no longer scored) and go
    // to the finisher, which holds their memory until it has finished them.
    std::lock_guard<std::mutex> g(mu);
    for (Job* j : failed) slots[j->slot] = nullptr;
    for (Job* j : ended) {
      slots[j->slot] = nullptr;
      to_finish.push_back(j);
      ++n_finishing;
    }
  }
  if (!ended.empty()) fin_cv.notify_one();
  for (Job* j : failed) release(j, launch_st, ctx->err);
  return true;
}

// The host GA: pose records prepared on the host threads, one graph
// replay, the scores down, then every job's GA advanced on the host threads
// (the jobs are independent).
hipError_t gdock_server::step_host_ga(const std::vector<Job*>& active) {
  const auto t_step = std::chrono::steady_clock::now();
  std::vector<int> list_at(active.size() + 1, 0);
  int max_row = 0;
  for (size_t i = 0; i < active.size(); ++i) {
    const int pop = active[i]->result->job.population();
    list_at[i + 1] = list_at[i] + pop;
    max_row = std::max(max_row, active[i]->slot * max_pop + pop);
  }
  const int n_list = list_at.back();
  host_parallel_for(static_cast<int64_t>(active.size()), [&](int64_t lo, int64_t hi) {
    for (int64_t i = lo; i < hi; ++i) {
      const Job* j = active[i];
      const GaJob& ga = j->result->job;
      const int pop = ga.population(), row0 = j->slot * max_pop;
      gdock_prepare_poses(ga.genes().data(), pop, h_poses + static_cast<size_t>(row0) * kPoseDoubles);
      for (int p = 0; p < pop; ++p) h_list[list_at[i] + p] = make_int2(j->slot, row0 + p);
    }
  });
  const int32_t head = n_list;
  hipError_t e = hipMemcpyAsync(const_cast<double*>(L.poses), h_poses,
                                static_cast<size_t>(max_row) * kPoseDoubles * sizeof(double), hipMemcpyHostToDevice,
                                ctx->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(const_cast<int2*>(L.pose_list), h_list, static_cast<size_t>(n_list) * sizeof(int2),
                       hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(const_cast<int32_t*>(L.n_pose_list), &head, sizeof head, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipEventRecord(ev0, ctx->stream);
  if (e == hipSuccess) e = hipGraphLaunch(exec, ctx->stream);
  if (e == hipSuccess) e = hipEventRecord(ev1, ctx->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(h_out, L.out, static_cast<size_t>(max_row) * kOutDoubles * sizeof(double),
                       hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  busy_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - t_step).count();
  float kms = 0.0f;
  if (e == hipSuccess && hipEventElapsedTime(&kms, ev0, ev1) == hipSuccess) kernel_s += kms * 1e-3;
  ++n_steps;
  n_poses += n_list;
  if (e == hipSuccess) {
    const auto t_ga = std::chrono::steady_clock::now();
    host_parallel_for(static_cast<int64_t>(active.size()), [&](int64_t lo, int64_t hi) {
      for (int64_t i = lo; i < hi; ++i)
        active[i]->result->job.absorb(h_out + static_cast<size_t>(active[i]->slot) * max_pop * kOutDoubles);
    });
    ga_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - t_ga).count();
  }
  return e;
}

// The device GA: scoring graph, then one GA launch for every job, which
// writes the next pose records in place; the host reads one done flag per slot.
hipError_t gdock_server::step_device_ga(const std::vector<Job*>& active) {
  const auto t_step = std::chrono::steady_clock::now();

  // The poses of the step (rows slot * max_pop + p) and the jobs' GA views,
  // uploaded when the set of jobs or a population size changed; the launch
  // plans its work items from the list.
  std::vector<int64_t> key;  // a job that takes over a slot is another key
  key.reserve(3 * active.size());
  for (const Job* j : active) key.insert(key.end(), {j->id, j->slot, j->n_cur});
  int n_list = 0;
  for (const Job* j : active) n_list += j->n_cur;
  hipError_t e = hipSuccess;
  if (key != step_key) {
    int at = 0;
    for (size_t i = 0; i < active.size(); ++i) {
      const Job* j = active[i];
      const int row0 = j->slot * max_pop;
      for (int q = 0; q < j->n_cur; ++q) h_list[at++] = make_int2(j->slot, row0 + q);
      h_ga[i] = j->ga;
    }
    const int32_t head = n_list;
    if (e == hipSuccess)
      e = hipMemcpyAsync(const_cast<int2*>(L.pose_list), h_list, static_cast<size_t>(n_list) * sizeof(int2),
                         hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(const_cast<int32_t*>(L.n_pose_list), &head, sizeof head, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(d_ga, h_ga, active.size() * sizeof(DevGaJob), hipMemcpyHostToDevice, ctx->stream);
    // the pageable count must be read before it leaves scope
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    step_key = e == hipSuccess ? key : std::vector<int64_t>();
  }
  if (e == hipSuccess) e = hipEventRecord(ev0, ctx->stream);
  if (e == hipSuccess) e = hipGraphLaunch(exec, ctx->stream);
  if (e == hipSuccess) e = hipEventRecord(ev1, ctx->stream);
  if (e == hipSuccess) e = launch_ga_generation(d_ga, static_cast<int>(active.size()), ctx->stream);
  if (e == hipSuccess) e = hipEventRecord(ev2, ctx->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(h_done, d_done, slots.size() * sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  busy_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - t_step).count();
  float kms = 0.0f, gms = 0.0f;
  if (e == hipSuccess && hipEventElapsedTime(&kms, ev0, ev1) == hipSuccess) kernel_s += kms * 1e-3;
  if (e == hipSuccess && hipEventElapsedTime(&gms, ev1, ev2) == hipSuccess) ga_s += gms * 1e-3;
  ++n_steps;
  n_poses += n_list;
  return e;
}

void gdock_server::release(Job* j, gdock_status st, const std::string& message) {
  if (j->cx) gdock_complex_destroy(fctx ? fctx : ctx, j->cx);
  j->cx = nullptr;
  if (j->split_mem) { (void)hipFree(j->split_mem); j->split_mem = nullptr; }
  if (j->ga_mem) { (void)hipFree(j->ga_mem); j->ga_mem = nullptr; }
  {
    std::lock_guard<std::mutex> g(mu);
    if (st != GDOCK_OK) {
      j->state = kFailed;
      j->error = st;
      j->message = message;
    } else if (j->state == kActive) {
      j->state = kFinished;
    }
    j->t_finish = now_s(t0);
    bytes_in_use -= j->bytes;
    ++n_finished;
  }
  done_cv.notify_all();
  cv.notify_all();  // memory freed: the worker may admit queued jobs
}

void gdock_server::prepare_loop() {
  (void)hipSetDevice(pctx->device);
  for (;;) {
    Job* j = nullptr;
    {
      std::unique_lock<std::mutex> lk(mu);
      prep_cv.wait(lk, [&] { return prep_stop || !to_prepare.empty(); });
      if (prep_stop) return;
      j = to_prepare.front();
      to_prepare.pop_front();
    }
    const auto t_prep = std::chrono::steady_clock::now();
    gdock_status st = GDOCK_OK;
    std::vector<gdock_atom> lig = j->lig;
    if (!lig.empty() && !j->rec.empty())
      st = gdock_position_ligand(j->rec.data(), static_cast<int32_t>(j->rec.size()), lig.data(),
                                 static_cast<int32_t>(lig.size()));
    if (st == GDOCK_OK)
      st = gdock_complex_create(pctx, j->rec.data(), static_cast<int32_t>(j->rec.size()), lig.data(),
                                static_cast<int32_t>(lig.size()), j->pairs.data(),
                                static_cast<int32_t>(j->pairs.size() / 2), &j->cx);
    if (st == GDOCK_OK && (j->cx->dev.n_rec_atoms > cfg.max_receptor_atoms || batch_part_slots(j->cx->dev) > lds_slots)) {
      set_error(pctx, "receptor of %d atoms (%d slots) exceeds the server's %d", j->cx->dev.n_rec_atoms,
                j->cx->dev.n_rec, cfg.max_receptor_atoms);
      st = GDOCK_ERR_INVALID_ARG;
    }
    hipError_t he = hipSuccess;
    if (st == GDOCK_OK) j->lanes = choose_lanes(j->cx->dev);
    size_t split_bytes = 0;
    // A job's poses take at least the parts that would fill the device with
    // its population alone (as the GA drivers fill their plain launches): a
    // function of the job, so its results stay the same in any mix.
    int fill = 1;
    if (st == GDOCK_OK) {
      fill = fill_parts(j->cx->dev, j->params.population_size, kGaFillItems);
      j->parts = std::max(pose_parts(j->cx->dev, j->lanes), fill);
      split_bytes = split_state_bytes(j->cx->dev, j->parts, max_pop);
      if (split_bytes) {
        he = hipMalloc(&j->split_mem, split_bytes);
        if (he == hipSuccess) he = hipMemsetAsync(j->split_mem, 0, split_bytes, pctx->stream);
        if (he != hipSuccess) st = hip_fail(pctx, he, "server: split state");
      }
    }
    // The job's GA on the device: the initial population (drawn on the host,
    // as every driver draws it), the stream state after it, and its pose
    // records written straight into the slot's rows of the launch.
    size_t ga_bytes = 0;
    DevGaState s0{};
    if (st == GDOCK_OK) j->result = std::make_unique<gdock_ga_result>(j->seed, j->params);
    if (st == GDOCK_OK && cfg.ga_on_device) {
      const int pop = j->result->job.population();
      Layout GL;
      GaJobOffsets go;
      ga_job_layout(&GL, pop, j->params, &go);
      ga_bytes = GL.size;
      he = hipMalloc(&j->ga_mem, ga_bytes);
      if (he == hipSuccess) {
        const size_t row0 = static_cast<size_t>(j->slot) * max_pop;
        j->ga = ga_job_bind(static_cast<char*>(j->ga_mem), go, j->params,
                            const_cast<double*>(L.poses) + row0 * kPoseDoubles, L.out + row0 * kOutDoubles);
        j->ga.done_flag = d_done + j->slot;
        j->n_cur = pop;
        s0 = ga_initial_state(j->result->job);
        he = hipMemcpyAsync(j->ga.st, &s0, sizeof s0, hipMemcpyHostToDevice, pctx->stream);
      }
      if (he == hipSuccess)
        he = hipMemcpyAsync(j->ga.genes[0], j->result->job.genes().data(), static_cast<size_t>(pop) * 6 * sizeof(double),
                            hipMemcpyHostToDevice, pctx->stream);
      if (he == hipSuccess) he = hipMemsetAsync(j->ga.done_flag, 0, sizeof(int32_t), pctx->stream);
      if (he == hipSuccess) he = launch_ga_pose_records(j->ga.genes[0], pop, j->ga.poses, pctx->stream);
      if (he != hipSuccess) st = hip_fail(pctx, he, "server: GA state");
    }
    if (st == GDOCK_OK) {  // everything on the preparer's stream is done before the worker uses it
      he = hipStreamSynchronize(pctx->stream);
      if (he != hipSuccess) st = hip_fail(pctx, he, "server: job preparation");
    }
    if (st == GDOCK_OK) {
      j->dev = j->cx->dev;
      j->dev.lanes_k = j->lanes;
      j->dev.min_parts = std::max(min_batch_parts(j->dev), fill);
      bind_split_state(&j->dev, j->parts, max_pop, j->slot * max_pop, static_cast<char*>(j->split_mem));
    }
    const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_prep).count();
    if (st == GDOCK_OK) {
      std::lock_guard<std::mutex> g(mu);
      const size_t charged = std::max(j->bytes, j->cx->mem_bytes + split_bytes + ga_bytes);  // the larger of estimate and arena
      bytes_in_use += charged - j->bytes;
      j->bytes = charged;
      j->ready = true;
      admit_s += dt;
    } else {
      {
        std::lock_guard<std::mutex> g(mu);
        slots[j->slot] = nullptr;
        admit_s += dt;
      }
      release(j, st, pctx->err);
      continue;
    }
    cv.notify_all();  // the worker scores the job from its next step
  }
}

void gdock_server::finish_loop() {
  (void)hipSetDevice(fctx->device);
  for (;;) {
    Job* j = nullptr;
    {
      std::unique_lock<std::mutex> lk(mu);
      fin_cv.wait(lk, [&] { return fin_stop || !to_finish.empty(); });
      if (to_finish.empty()) return;  // stopping with nothing left
      j = to_finish.front();
      to_finish.pop_front();
    }
    gdock_status st = GDOCK_OK;
    if (j->ga_mem) {  // the device GA's results into the job's host state, then its memory is free
      const hipError_t he = ga_job_download(j->ga, j->params, fctx->stream, &j->result->job);
      if (he != hipSuccess) st = hip_fail(fctx, he, "server: GA results");
      (void)hipFree(j->ga_mem);
      j->ga_mem = nullptr;
    }
    if (st == GDOCK_OK && j->params.cluster) {
      const auto t_fcc = std::chrono::steady_clock::now();
      st = finish_job(fctx, j->cx, &j->result->job, &j->result->clusters);
      const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_fcc).count();
      std::lock_guard<std::mutex> g(mu);
      finish_s += dt;
    }
    release(j, st, st != GDOCK_OK ? std::string(fctx->err) : std::string());
    std::lock_guard<std::mutex> g(mu);
    --n_finishing;
  }
}

void gdock_server::run() {
  (void)hipSetDevice(ctx->device);
  while (step()) {
  }
}

extern "C" {

gdock_status gdock_server_default_config(gdock_server_config* c) {
  if (!c) return GDOCK_ERR_INVALID_ARG;
  *c = gdock_server_config{};
  // A job's launches already fill the device (kGaFillItems), so more active
  // jobs only share the steps: 160 jobs at once ran at 8,241 jobs/min with
  // 16 slots and 7,615 with 64 (lower median latency too).
  c->max_jobs = 16;
  c->max_population = 150;
  c->max_receptor_atoms = 8192;
  c->ga_on_device = 0;
  c->memory_budget = 0;
  c->weights = gdock_weights{0.4, 0.05, 3.4, 1.0};
  return GDOCK_OK;
}

gdock_status gdock_server_create(int device, const gdock_server_config* config, gdock_server** out) {
  if (!out) return GDOCK_ERR_INVALID_ARG;
  *out = nullptr;
  gdock_server_config c;
  if (config) c = *config; else gdock_server_default_config(&c);
  if (c.max_jobs < 1 || c.max_jobs > 4096 || c.max_population < 2 || c.max_population > 65536 ||
      (c.ga_on_device && c.max_population > kGaMaxPopulation) || c.max_receptor_atoms < 1 ||
      c.max_receptor_atoms > kMaxReceptorAtoms || c.memory_budget < 0 ||
      static_cast<int64_t>(c.max_jobs) * c.max_population * kMaxPoseParts > (int64_t{1} << 30))
    return GDOCK_ERR_INVALID_ARG;
  auto* s = new gdock_server();
  s->cfg = c;
  s->t0 = std::chrono::steady_clock::now();
  gdock_status st = gdock_ctx_create(device, &s->ctx);
  if (st != GDOCK_OK) { delete s; return st; }
  if (s->cfg.memory_budget == 0) {
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) == hipSuccess) s->cfg.memory_budget = static_cast<int64_t>(fr / 2);
    else s->cfg.memory_budget = int64_t{8} << 30;
  }
  s->max_pop = c.max_population;
  s->slots.assign(c.max_jobs, nullptr);
  const size_t cap = static_cast<size_t>(c.max_jobs) * c.max_population;
  const size_t cap_items = cap * kMaxPoseParts;  // a pose is at most kMaxPoseParts items
  Layout Ly;
  const size_t o_tab = Ly.add<DevComplex>(c.max_jobs), o_list = Ly.add<int2>(cap),
               o_items = Ly.add<int2>(cap_items), o_n = Ly.add<int32_t>(4),
               o_poses = Ly.add<double>(cap * kPoseDoubles), o_out = Ly.add<double>(cap * kOutDoubles),
               o_ns = Ly.add<int32_t>(cap), o_redo = Ly.add<int32_t>(cap),
               o_order = Ly.add<uint8_t>(order_bytes(cap)), o_ga = Ly.add<DevGaJob>(c.max_jobs),
               o_done = Ly.add<int32_t>(c.max_jobs);
  hipError_t e = hipMalloc(&s->mem, Ly.size);
  if (e == hipSuccess) e = hipMemsetAsync(s->mem, 0, Ly.size, s->ctx->stream);
  if (e == hipSuccess && !c.ga_on_device)
    e = hipHostMalloc(reinterpret_cast<void**>(&s->h_poses), cap * kPoseDoubles * sizeof(double));
  if (e == hipSuccess && !c.ga_on_device)
    e = hipHostMalloc(reinterpret_cast<void**>(&s->h_out), cap * kOutDoubles * sizeof(double));
  if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&s->h_list), cap * sizeof(int2));
  if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&s->h_table), c.max_jobs * sizeof(DevComplex));
  if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&s->h_ga), c.max_jobs * sizeof(DevGaJob));
  if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&s->h_done), c.max_jobs * sizeof(int32_t));
  if (e == hipSuccess) e = hipStreamSynchronize(s->ctx->stream);
  if (e == hipSuccess && c.ga_on_device) e = ga_device_init(s->ctx->stream);
  if (e == hipSuccess) e = hipEventCreate(&s->ev0);
  if (e == hipSuccess) e = hipEventCreate(&s->ev1);
  if (e == hipSuccess) e = hipEventCreate(&s->ev2);
  char* m = static_cast<char*>(s->mem);
  BatchLaunch& L = s->L;
  if (e == hipSuccess) {
    L.table = reinterpret_cast<const DevComplex*>(m + o_tab);
    L.pose_list = reinterpret_cast<const int2*>(m + o_list);
    L.n_pose_list = reinterpret_cast<const int32_t*>(m + o_n) + 3;
    L.items = reinterpret_cast<int2*>(m + o_items);
    L.n_items = reinterpret_cast<int32_t*>(m + o_n);
    L.counter = reinterpret_cast<int32_t*>(m + o_n) + 1;
    L.exit_count = reinterpret_cast<int32_t*>(m + o_n) + 2;
    L.poses = reinterpret_cast<const double*>(m + o_poses);
    L.out = reinterpret_cast<double*>(m + o_out);
    L.nsat = reinterpret_cast<int32_t*>(m + o_ns);
    L.redo = reinterpret_cast<int32_t*>(m + o_redo);
    bind_order(&L, m + o_order, cap);
    s->d_ga = reinterpret_cast<DevGaJob*>(m + o_ga);
    s->d_done = reinterpret_cast<int32_t*>(m + o_done);
    // a job's items cover at most min(its receptor slots, kPartSlots) slots;
    // class padding adds at most 3 slots per (rmin2, eps) class
    s->lds_slots = std::min(kPartSlots, (c.max_receptor_atoms + 256 + 63) / 64 * 64);
    L.lds = batch_lds_bytes(s->lds_slots);
    L.capacity = static_cast<int>(cap);
    L.item_capacity = static_cast<int>(cap_items);
    L.grid = std::min(batch_grid(L.lds, device), L.item_capacity);
    // One graph for every step: the kernels read the item count from device memory.
    hipStream_t cs = nullptr;
    const Weights w{c.weights.vdw, c.weights.elec, c.weights.desolv, c.weights.air};
    e = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
      const hipError_t el = launch_batch(L, w, cs);
      e = hipStreamEndCapture(cs, &s->graph);
      if (e == hipSuccess) e = el;
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&s->exec, s->graph, nullptr, nullptr, 0);
    if (cs) (void)hipStreamDestroy(cs);
  }
  if (e != hipSuccess) {
    gdock_server_destroy(s);
    return GDOCK_ERR_DEVICE;
  }
  if (gdock_ctx_create(device, &s->fctx) != GDOCK_OK || gdock_ctx_create(device, &s->pctx) != GDOCK_OK) {
    gdock_server_destroy(s);
    return GDOCK_ERR_DEVICE;
  }
  s->worker = std::thread([s] { s->run(); });
  s->preparer = std::thread([s] { s->prepare_loop(); });
  s->finisher = std::thread([s] { s->finish_loop(); });
  *out = s;
  return GDOCK_OK;
}

void gdock_server_destroy(gdock_server* s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> g(s->mu);
    s->stop = true;
  }
  s->cv.notify_all();
  if (s->worker.joinable()) s->worker.join();
  {  // jobs waiting for preparation are dropped with the rest
    std::lock_guard<std::mutex> g(s->mu);
    s->prep_stop = true;
  }
  s->prep_cv.notify_all();
  if (s->preparer.joinable()) s->preparer.join();
  {  // the finisher completes the jobs already handed to it, then stops
    std::lock_guard<std::mutex> g(s->mu);
    s->fin_stop = true;
  }
  s->fin_cv.notify_all();
  if (s->finisher.joinable()) s->finisher.join();
  {  // every job still queued or running ends as failed; waiters wake, read it and leave
    std::unique_lock<std::mutex> lk(s->mu);
    s->closing = true;
    for (auto& kv : s->jobs) {
      Job* j = kv.second.get();
      if (j->cx) gdock_complex_destroy(s->ctx, j->cx);
      if (j->split_mem) (void)hipFree(j->split_mem);
      if (j->ga_mem) (void)hipFree(j->ga_mem);
      j->cx = nullptr;
      j->split_mem = nullptr;
      j->ga_mem = nullptr;
      if (j->state != kFinished && j->state != kFailed) {
        j->state = kFailed;
        j->error = GDOCK_ERR_NOT_FOUND;
        j->message = "server destroyed before the job ended";
        j->result.reset();
      }
    }
    s->done_cv.notify_all();
    s->idle_cv.wait(lk, [s] { return s->waiters == 0; });
  }
  if (s->ctx) (void)hipSetDevice(s->ctx->device);
  if (s->ev0) (void)hipEventDestroy(s->ev0);
  if (s->ev1) (void)hipEventDestroy(s->ev1);
  if (s->ev2) (void)hipEventDestroy(s->ev2);
  if (s->exec) (void)hipGraphExecDestroy(s->exec);
  if (s->graph) (void)hipGraphDestroy(s->graph);
  if (s->mem) (void)hipFree(s->mem);
  if (s->h_poses) (void)hipHostFree(s->h_poses);
  if (s->h_out) (void)hipHostFree(s->h_out);
  if (s->h_list) (void)hipHostFree(s->h_list);
  if (s->h_ga) (void)hipHostFree(s->h_ga);
  if (s->h_done) (void)hipHostFree(s->h_done);
  if (s->h_table) (void)hipHostFree(s->h_table);
  if (s->ctx) gdock_ctx_destroy(s->ctx);
  if (s->fctx) gdock_ctx_destroy(s->fctx);
  if (s->pctx) gdock_ctx_destroy(s->pctx);
  delete s;
}

gdock_status gdock_server_submit(gdock_server* s, const gdock_atom* rec, int32_t n_rec, const gdock_atom* lig,
                                 int32_t n_lig, const int32_t* restraint_pairs, int32_t n_pairs, uint64_t seed,
                                 const gdock_ga_params* params, int64_t* job_id) {
  if (!s || !job_id || n_rec < 0 || n_lig < 0 || n_pairs < 0 || (n_rec && !rec) || (n_lig && !lig) ||
      (n_pairs && !restraint_pairs))
    return GDOCK_ERR_INVALID_ARG;
  gdock_ga_params p;
  if (params) p = *params; else gdock_ga_default_params(&p);
  if (validate_ga_params(nullptr, &p) != GDOCK_OK || p.population_size > s->max_pop) return GDOCK_ERR_INVALID_ARG;
  auto j = std::make_unique<Job>();
  j->rec.assign(rec, rec + n_rec);
  j->lig.assign(lig, lig + n_lig);
  j->pairs.assign(restraint_pairs, restraint_pairs + 2 * n_pairs);
  j->seed = seed;
  j->params = p;
  {
    std::lock_guard<std::mutex> g(s->mu);
    j->id = s->next_id++;
    j->t_submit = now_s(s->t0);
    *job_id = j->id;
    s->queue.push_back(j.get());
    s->jobs.emplace(j->id, std::move(j));
  }
  s->cv.notify_all();
  return GDOCK_OK;
}

gdock_status gdock_server_wait(gdock_server* s, int64_t job_id, double timeout_s, gdock_ga_result** result,
                               gdock_job_stats* stats) {
  if (!s) return GDOCK_ERR_INVALID_ARG;
  std::unique_lock<std::mutex> lk(s->mu);
  if (s->closing) return GDOCK_ERR_NOT_FOUND;
  if (s->jobs.find(job_id) == s->jobs.end()) return GDOCK_ERR_NOT_FOUND;
  auto finished = [s, job_id] {  // ended, or taken by another waiter of the same id
    auto f = s->jobs.find(job_id);
    return f == s->jobs.end() || f->second->state == kFinished || f->second->state == kFailed;
  };
  ++s->waiters;
  const bool ended = timeout_s < 0 ? (s->done_cv.wait(lk, finished), true)
                                   : s->done_cv.wait_for(lk, std::chrono::duration<double>(timeout_s), finished);
  struct Leave {  // the last waiter out lets a pending destroy go on
    gdock_server* s;
    ~Leave() { if (--s->waiters == 0 && s->closing) s->idle_cv.notify_all(); }
  } leave{s};
  if (!ended) return GDOCK_ERR_BUSY;
  auto it = s->jobs.find(job_id);
  if (it == s->jobs.end()) return GDOCK_ERR_NOT_FOUND;
  Job* j = it->second.get();
  if (stats) {
    stats->submit_s = j->t_submit;
    stats->admit_s = j->t_admit;
    stats->finish_s = j->t_finish;
    stats->generations = j->result ? j->result->job.generations() : 0;
    stats->slot = j->slot;
  }
  const gdock_status err = j->state == kFailed ? j->error : GDOCK_OK;
  if (result) *result = err == GDOCK_OK ? j->result.release() : nullptr;
  s->jobs.erase(it);
  return err;
}

gdock_status gdock_server_stats(gdock_server* s, gdock_server_stats_t* out) {
  if (!s || !out) return GDOCK_ERR_INVALID_ARG;
  std::lock_guard<std::mutex> g(s->mu);
  out->steps = s->n_steps;
  out->poses_scored = s->n_poses;
  out->jobs_finished = s->n_finished;
  out->jobs_queued = static_cast<int64_t>(s->queue.size());
  out->peak_active = s->peak_active;
  out->busy_s = s->busy_s;
  out->kernel_s = s->kernel_s;
  out->admit_s = s->admit_s;
  out->ga_s = s->ga_s;
  out->finish_s = s->finish_s;
  out->bytes_in_use = static_cast<int64_t>(s->bytes_in_use);
  out->memory_budget = s->cfg.memory_budget;
  return GDOCK_OK;
}

}  // extern "C"
