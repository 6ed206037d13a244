// Serving loop: continuous batching of simulation jobs over one device batch.
//
// Jobs arrive over time (arrival_s on the server clock).  Each pass of the
// loop waits for the persistent launch in flight ("tick") and launches the
// next, which advances every resident box by up to sweeps_per_tick moves;
// while it runs, the loop retires the boxes the finished tick completed
// (freeing their slot and memory) and admits every arrived job whose device
// footprint fits the HBM budget and a free slot (FIFO by arrival); they run
// from the following tick on.  Box state and RNG stay on the device from
// admission to retirement.  Latency of a job = finish time - arrival time.
//
// The loop runs on a host thread of the server (mcgpu_server_start), which
// alone drives the device batch.  Other threads submit jobs, poll them, read
// live statistics and wait for completions at any time; one mutex guards
// the job table and the statistics, and a condition variable wakes the loop
// on a submission and waiters on a completion.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <deque>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "hostbox.h"
#include "mcgpu_host.h"

namespace mcgpu {
HostBox& hostbox_ref(mcgpu_hostbox* h);
}

namespace {

struct Job {
  mcgpu_hostbox* box;  // the caller's, read at admission
  mcgpu_run_params params;
  mcgpu_job_result res;
  int32_t slot = -1;
};

using Clock = std::chrono::steady_clock;

}  // namespace

struct mcgpu_server {
  mcgpu_server_config cfg{};
  mcgpu_batch* batch = nullptr;  // driven by the loop thread only
  std::mutex mu;                 // guards everything below
  std::condition_variable cv_loop;  // submissions and stop wake the loop
  std::condition_variable cv_done;  // completions wake waiters
  // a deque: a submission (push_back) leaves references to earlier jobs
  // valid, which the loop holds across its unlocked device calls
  std::deque<Job> jobs;
  std::vector<int64_t> pending;  // queued job ids, in arrival order
  mcgpu_server_stats stats{};
  int64_t completed = 0;
  int32_t resident_now = 0;
  bool started = false, stopping = false, finished = false;
  int loop_rc = MCGPU_OK;
  std::string loop_error;
  Clock::time_point t0;
  std::thread loop;

  double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
};

namespace {

// Queues job id k in arrival order (stable for equal arrivals).
void enqueue_pending(mcgpu_server* s, int64_t k) {
  const double a = s->jobs[k].res.arrival_s;
  auto it = std::upper_bound(s->pending.begin(), s->pending.end(), a,
                             [&](double t, int64_t j) { return t < s->jobs[j].res.arrival_s; });
  s->pending.insert(it, k);
}

// The continuous-batching loop (runs on the server's thread).
int serve_loop(mcgpu_server* s) {
  std::vector<int64_t> resident;  // job ids on the device
  mcgpu_box_stats st;
  // MCGPU_SERVER_LOG=<seconds>: a progress line on stderr at that period
  const char* log_env = std::getenv("MCGPU_SERVER_LOG");
  const double log_every = log_env ? std::atof(log_env) : 0.0;
  double next_log = log_every;
  // No-progress watchdog: with jobs resident or already arrived, a tick or an
  // admission must complete within this many seconds.
  const double kStallLimit = 60.0;
  double last_progress = 0.0;
  double t_admit = 0.0, t_run = 0.0, t_retire = 0.0;  // host time per phase
  double kernel_ms = 0.0;  // device time of the ticks (HIP events), summed
  int64_t peak = 0;

  // Admits arrived jobs while they fit (FIFO; the head waits for memory
  // when boxes are resident).  Returns MCGPU_OK or a loop-ending error.
  auto admit = [&](bool allow_empty_retry) {
    for (;;) {
      int64_t k;
      mcgpu_box_desc d;
      mcgpu_run_params p;
      {
        std::lock_guard<std::mutex> lock(s->mu);
        if (s->pending.empty() || s->jobs[s->pending.front()].res.arrival_s > s->now()) break;
        k = s->pending.front();
        d = mcgpu::hostbox_ref(s->jobs[k].box).desc();
        p = s->jobs[k].params;
      }
      int32_t slot = -1;
      const double ta = s->now();
      const int rc = mcgpu_batch_add(s->batch, &d, &p, &slot);
      if (log_every > 0 && s->now() - ta > log_every) {
        std::fprintf(stderr, "[server] slow admission: job %lld (%d molecules) took %.2fs rc=%d\n",
                     (long long)k, d.n_molecules, s->now() - ta, rc);
        std::fflush(stderr);
      }
      if (rc == MCGPU_ERR_NOMEM && (!resident.empty() || !allow_empty_retry)) break;
      std::lock_guard<std::mutex> lock(s->mu);
      s->pending.erase(s->pending.begin());
      Job& j = s->jobs[k];
      j.res.admit_s = s->now();
      if (rc != MCGPU_OK) {  // failed job: isolated, reported, not retried
        j.res.status = rc;
        j.res.finish_s = j.res.admit_s;
        j.res.state = MCGPU_JOB_DONE;
        s->completed++;
        s->cv_done.notify_all();
        continue;
      }
      j.slot = slot;
      j.res.state = MCGPU_JOB_RUNNING;
      last_progress = j.res.admit_s;
      resident.push_back(k);
      s->resident_now = (int32_t)resident.size();
    }
  };

  for (;;) {
    const double t = s->now();
    admit(true);
    t_admit += s->now() - t;
    peak = std::max<int64_t>(peak, (int64_t)resident.size());
    bool arrived, stop_now;
    double next_arrival = -1.0;
    {
      std::lock_guard<std::mutex> lock(s->mu);
      arrived = !s->pending.empty() && s->jobs[s->pending.front()].res.arrival_s <= s->now();
      if (!s->pending.empty()) next_arrival = s->jobs[s->pending.front()].res.arrival_s;
      stop_now = s->stopping && s->pending.empty() && resident.empty();
      s->stats.peak_resident = peak;
    }
    if (stop_now) break;
    if ((arrived || !resident.empty()) && s->now() - last_progress > kStallLimit) {
      mcgpu::set_error("mcgpu_server: no progress for 60 s (" + std::to_string(resident.size()) +
                       " resident)");
      return MCGPU_ERR_STATE;
    }
    if (resident.empty()) {
      if (!arrived) last_progress = s->now();
      // idle: sleep until the next arrival, a submission or stop
      std::unique_lock<std::mutex> lock(s->mu);
      const double wait = next_arrival < 0 ? 0.05 : std::min(0.05, next_arrival - s->now());
      if (wait > 0 && !s->stopping)
        s->cv_loop.wait_for(lock, std::chrono::duration<double>(wait));
      continue;
    }
    // Pipelined ticks: the run call consumes the previous tick (waits for
    // it) and launches the next one.  While that one runs, the jobs the
    // consumed tick finished are retired (peek reads the statistics as of
    // the consumed tick, and removing a box that the running tick does not
    // run waits for nothing) and arrived jobs are admitted (host set-up and
    // queued uploads), so the device idles only for the consume-and-launch
    // step between ticks.
    const double tt = s->now();
    int rc = mcgpu_batch_run(s->batch, s->cfg.sweeps_per_tick, 0);
    if (rc != MCGPU_OK) return rc;
    t_run += s->now() - tt;
    last_progress = s->now();
    if (log_every > 0 && s->now() - tt > log_every) {
      std::fprintf(stderr, "[server] slow tick: %.2fs with %zu resident\n", s->now() - tt,
                   resident.size());
      std::fflush(stderr);
    }
    if (log_every > 0 && s->now() >= next_log) {
      double kms = 0.0;
      int64_t nl = 0;
      if (mcgpu_batch_kernel_ms(s->batch, &kms, &nl) == MCGPU_OK) kernel_ms += kms * nl;
      double tw = 0, tv = 0, tl = 0;
      (void)mcgpu_batch_host_seconds(s->batch, &tw, &tv, &tl);
      std::lock_guard<std::mutex> lock(s->mu);
      std::fprintf(stderr, "[server] t=%.2fs ticks=%lld resident=%zu pending=%zu done=%lld moves=%lld "
                   "host s: admit %.2f run %.2f (kernels %.2f; wait %.2f volume %.2f launch %.2f) "
                   "retire %.2f\n",
                   s->now(), (long long)s->stats.ticks + 1, resident.size(), s->pending.size(),
                   (long long)s->completed, (long long)s->stats.total_moves, t_admit, t_run,
                   kernel_ms / 1e3, tw, tv, tl, t_retire);
      std::fflush(stderr);
      next_log = s->now() + log_every;
    }
    // live progress of every resident job, and retirement of finished ones
    const double tr = s->now();
    int64_t moved = 0;
    for (size_t k = 0; k < resident.size();) {
      const int64_t id = resident[k];
      int32_t slot;
      {
        std::lock_guard<std::mutex> lock(s->mu);
        slot = s->jobs[id].slot;
      }
      rc = mcgpu_batch_peek(s->batch, slot, &st);
      if (rc != MCGPU_OK) return rc;
      std::unique_lock<std::mutex> lock(s->mu);
      Job& j = s->jobs[id];
      moved += st.steps_done - j.res.steps_done;
      j.res.steps_done = st.steps_done;
      j.res.accepted = st.accepted;
      j.res.final_energy = st.energy;
      j.res.start_energy = st.start_energy;
      if (!st.done) {
        k++;
        continue;
      }
      j.res.finish_s = s->now();
      j.res.steps = st.steps_done;
      j.res.status = st.fault ? MCGPU_ERR_STATE : MCGPU_OK;
      s->stats.work_candidates += st.work_candidates;
      s->stats.work_sites += st.work_sites;
      lock.unlock();
      if (st.fault && log_every > 0) {
        std::fprintf(stderr, "[server] job %lld stopped by device fault %d after %lld moves\n",
                     (long long)id, st.fault, (long long)st.steps_done);
        std::fflush(stderr);
      }
      rc = mcgpu_batch_remove(s->batch, slot);
      if (rc != MCGPU_OK) return rc;
      lock.lock();
      j.slot = -1;
      j.res.state = MCGPU_JOB_DONE;
      s->completed++;
      s->stats.wall_s = j.res.finish_s;
      resident[k] = resident.back();
      resident.pop_back();
      s->resident_now = (int32_t)resident.size();
      s->cv_done.notify_all();
    }
    t_retire += s->now() - tr;
    {
      const double t2 = s->now();
      admit(false);
      t_admit += s->now() - t2;
    }
    std::lock_guard<std::mutex> lock(s->mu);
    s->stats.ticks++;
    s->stats.total_moves += moved;
  }
  size_t used = 0, pk = 0, capb = 0;
  double avg_ms = 0;
  int64_t launches = 0;
  if (mcgpu_batch_kernel_ms(s->batch, &avg_ms, &launches) == MCGPU_OK)
    kernel_ms += avg_ms * launches;
  std::lock_guard<std::mutex> lock(s->mu);
  if (mcgpu_batch_memory(s->batch, &used, &pk, &capb) == MCGPU_OK) s->stats.peak_bytes = pk;
  s->stats.kernel_ms_total = kernel_ms;
  return MCGPU_OK;
}

void loop_main(mcgpu_server* s) {
  int rc = mcgpu_batch_bind_thread(s->batch);
  if (rc == MCGPU_OK) rc = serve_loop(s);
  std::lock_guard<std::mutex> lock(s->mu);
  s->loop_rc = rc;
  if (rc != MCGPU_OK) s->loop_error = mcgpu_last_error();
  s->finished = true;
  s->cv_done.notify_all();
}

}  // namespace

extern "C" int mcgpu_server_create(const mcgpu_server_config* cfg,
                                   mcgpu_server** out) {
  if (!cfg || !out || cfg->max_boxes < 1 || cfg->sweeps_per_tick < 1) {
    mcgpu::set_error("mcgpu_server_create: bad configuration");
    return MCGPU_ERR_INVALID;
  }
  auto* s = new (std::nothrow) mcgpu_server();
  if (!s) return MCGPU_ERR_NOMEM;
  s->cfg = *cfg;
  int rc = mcgpu_batch_create(cfg->max_boxes, cfg->hbm_budget_bytes, &s->batch);
  if (rc == MCGPU_OK && cfg->tick_budget_us > 0)
    rc = mcgpu_batch_set_tick_budget(s->batch, cfg->tick_budget_us);
  if (rc != MCGPU_OK) {
    if (s->batch) mcgpu_batch_destroy(s->batch);
    delete s;
    return rc;
  }
  *out = s;
  return MCGPU_OK;
}

extern "C" int mcgpu_server_stop(mcgpu_server* s) {
  if (!s) return MCGPU_ERR_INVALID;
  {
    std::lock_guard<std::mutex> lock(s->mu);
    if (!s->started) return MCGPU_OK;
    s->stopping = true;
  }
  s->cv_loop.notify_all();
  if (s->loop.joinable()) s->loop.join();
  std::lock_guard<std::mutex> lock(s->mu);
  if (s->loop_rc != MCGPU_OK) mcgpu::set_error(s->loop_error);
  return s->loop_rc;
}

extern "C" int mcgpu_server_destroy(mcgpu_server* s) {
  if (!s) return MCGPU_OK;
  (void)mcgpu_server_stop(s);
  if (s->batch) mcgpu_batch_destroy(s->batch);
  delete s;
  return MCGPU_OK;
}

extern "C" int mcgpu_server_submit(mcgpu_server* s, mcgpu_hostbox* box,
                                   const mcgpu_run_params* params,
                                   double arrival_s, int64_t* job_id) {
  if (!s || !box || !params) {
    mcgpu::set_error("mcgpu_server_submit: bad argument");
    return MCGPU_ERR_INVALID;
  }
  {
    std::lock_guard<std::mutex> lock(s->mu);
    if (s->stopping) {
      mcgpu::set_error("mcgpu_server_submit: the server is stopping");
      return MCGPU_ERR_STATE;
    }
    Job j;
    j.box = box;
    j.params = *params;
    std::memset(&j.res, 0, sizeof(j.res));
    j.res.job_id = (int64_t)s->jobs.size();
    j.res.arrival_s = arrival_s >= 0 ? arrival_s : (s->started ? s->now() : 0.0);
    j.res.steps = params->steps;
    j.res.state = MCGPU_JOB_QUEUED;
    s->jobs.push_back(j);
    enqueue_pending(s, j.res.job_id);
    if (job_id) *job_id = j.res.job_id;
  }
  s->cv_loop.notify_all();
  return MCGPU_OK;
}

extern "C" int mcgpu_server_start(mcgpu_server* s) {
  if (!s) return MCGPU_ERR_INVALID;
  std::lock_guard<std::mutex> lock(s->mu);
  if (s->started) {
    mcgpu::set_error("mcgpu_server_start: already started");
    return MCGPU_ERR_STATE;
  }
  s->started = true;
  s->t0 = Clock::now();
  try {
    s->loop = std::thread(loop_main, s);
  } catch (...) {
    s->started = false;
    mcgpu::set_error("mcgpu_server_start: cannot start the loop thread");
    return MCGPU_ERR_NOMEM;
  }
  return MCGPU_OK;
}

extern "C" int mcgpu_server_wait(mcgpu_server* s, int64_t job_id, double timeout_s) {
  if (!s) return MCGPU_ERR_INVALID;
  std::unique_lock<std::mutex> lock(s->mu);
  if (job_id >= (int64_t)s->jobs.size()) {
    mcgpu::set_error("mcgpu_server_wait: unknown job");
    return MCGPU_ERR_RANGE;
  }
  if (!s->started) {
    mcgpu::set_error("mcgpu_server_wait: the server has not started");
    return MCGPU_ERR_STATE;
  }
  auto done = [&]() {
    if (s->finished) return true;
    return job_id >= 0 ? s->jobs[job_id].res.state == MCGPU_JOB_DONE
                       : s->completed == (int64_t)s->jobs.size();
  };
  if (timeout_s < 0) {
    s->cv_done.wait(lock, done);
  } else if (!s->cv_done.wait_for(lock, std::chrono::duration<double>(timeout_s), done)) {
    mcgpu::set_error("mcgpu_server_wait: timed out");
    return MCGPU_ERR_STATE;
  }
  if (s->finished && s->loop_rc != MCGPU_OK) {
    mcgpu::set_error(s->loop_error);
    return s->loop_rc;
  }
  return MCGPU_OK;
}

extern "C" int mcgpu_server_run(mcgpu_server* s) {
  if (!s) return MCGPU_ERR_INVALID;
  int rc = mcgpu_server_start(s);
  if (rc == MCGPU_OK) rc = mcgpu_server_wait(s, -1, -1.0);
  const int rs = mcgpu_server_stop(s);
  return rc != MCGPU_OK ? rc : rs;
}

extern "C" int mcgpu_server_poll(mcgpu_server* s, int64_t job_id, mcgpu_job_result* out) {
  if (!s || !out) return MCGPU_ERR_INVALID;
  std::lock_guard<std::mutex> lock(s->mu);
  if (job_id < 0 || job_id >= (int64_t)s->jobs.size()) {
    mcgpu::set_error("mcgpu_server_poll: unknown job");
    return MCGPU_ERR_RANGE;
  }
  *out = s->jobs[job_id].res;
  return MCGPU_OK;
}

extern "C" int mcgpu_server_results(mcgpu_server* s, mcgpu_job_result* out,
                                    int64_t cap, int64_t* n) {
  if (!s || !n || (cap > 0 && !out)) return MCGPU_ERR_INVALID;
  std::lock_guard<std::mutex> lock(s->mu);
  const int64_t k = std::max<int64_t>(0, std::min<int64_t>(cap, (int64_t)s->jobs.size()));
  for (int64_t i = 0; i < k; i++) out[i] = s->jobs[i].res;
  *n = k;
  return MCGPU_OK;
}

extern "C" int mcgpu_server_get_stats(mcgpu_server* s, mcgpu_server_stats* out) {
  if (!s || !out) return MCGPU_ERR_INVALID;
  std::lock_guard<std::mutex> lock(s->mu);
  *out = s->stats;
  out->resident = s->resident_now;
  out->queued = (int64_t)s->pending.size();
  out->completed = s->completed;
  out->submitted = (int64_t)s->jobs.size();
  out->clock_s = s->started ? s->now() : 0.0;
  return MCGPU_OK;
}
