// Serving loop: continuous batching of simulation jobs over one device batch.
//
// Jobs arrive over time (arrival_s on the server clock).  Between persistent
// launches ("ticks") the loop admits every arrived job whose device
// footprint fits the HBM budget and a free slot (FIFO), then launches one
// tick that advances every resident box by up to sweeps_per_tick moves, then
// retires finished boxes, freeing their slot and memory for the next
// admissions.  Box state and RNG stay on the device from admission to
// retirement.  Latency of a job = finish time - arrival time.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <string>
#include <thread>
#include <vector>

#include "hostbox.h"
#include "mcgpu_host.h"

namespace mcgpu {
HostBox& hostbox_ref(mcgpu_hostbox* h);
}

using mcgpu::HostBox;

namespace {

struct Job {
  mcgpu_hostbox* box;  // owned by the caller until mcgpu_server_run returns
  mcgpu_run_params params;
  mcgpu_job_result res;
  int32_t slot = -1;
};

}  // namespace

struct mcgpu_server {
  mcgpu_server_config cfg{};
  mcgpu_batch* batch = nullptr;
  std::vector<Job> jobs;
  mcgpu_server_stats stats{};
  bool ran = false;
};

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

extern "C" int mcgpu_server_destroy(mcgpu_server* s) {
  if (!s) return MCGPU_OK;
  if (s->batch) mcgpu_batch_destroy(s->batch);
  delete s;
  return MCGPU_OK;
}

extern "C" int mcgpu_server_submit(mcgpu_server* s, mcgpu_hostbox* box,
                                   const mcgpu_run_params* params,
                                   double arrival_s, int64_t* job_id) {
  if (!s || !box || !params || s->ran) {
    mcgpu::set_error("mcgpu_server_submit: bad argument or server already ran");
    return MCGPU_ERR_INVALID;
  }
  Job j;
  j.box = box;
  j.params = *params;
  std::memset(&j.res, 0, sizeof(j.res));
  j.res.job_id = (int64_t)s->jobs.size();
  j.res.arrival_s = arrival_s;
  j.res.steps = params->steps;
  s->jobs.push_back(j);
  if (job_id) *job_id = j.res.job_id;
  return MCGPU_OK;
}

extern "C" int mcgpu_server_run(mcgpu_server* s) {
  if (!s || s->ran) return MCGPU_ERR_INVALID;
  s->ran = true;
  std::vector<size_t> order(s->jobs.size());
  for (size_t i = 0; i < order.size(); i++) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) {
    return s->jobs[a].res.arrival_s < s->jobs[b].res.arrival_s;
  });
  using clk = std::chrono::steady_clock;
  const auto t0 = clk::now();
  auto now = [&]() { return std::chrono::duration<double>(clk::now() - t0).count(); };
  std::vector<int32_t> resident;  // job indices on the device
  size_t next = 0, completed = 0;
  int64_t peak = 0;
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
  while (completed < s->jobs.size()) {
    // admit every arrived job that fits (FIFO; the head waits for memory)
    const double t = now();
    while (next < order.size() && s->jobs[order[next]].res.arrival_s <= t) {
      Job& j = s->jobs[order[next]];
      const mcgpu_box_desc d = mcgpu::hostbox_ref(j.box).desc();
      int32_t slot = -1;
      const double ta = now();
      const int rc = mcgpu_batch_add(s->batch, &d, &j.params, &slot);
      if (log_every > 0 && now() - ta > log_every) {
        std::fprintf(stderr, "[server] slow admission: job %lld (%d molecules) took %.2fs rc=%d\n",
                     (long long)j.res.job_id, d.n_molecules, now() - ta, rc);
        std::fflush(stderr);
      }
      if (rc == MCGPU_ERR_NOMEM && !resident.empty()) break;  // wait for room
      next++;
      if (rc != MCGPU_OK) {  // failed job: isolated, reported, not retried
        j.res.status = rc;
        j.res.admit_s = j.res.finish_s = now();
        completed++;
        continue;
      }
      j.slot = slot;
      j.res.admit_s = now();
      last_progress = j.res.admit_s;
      resident.push_back((int32_t)order[next - 1]);
    }
    t_admit += now() - t;
    peak = std::max<int64_t>(peak, (int64_t)resident.size());
    const bool arrived = next < order.size() && s->jobs[order[next]].res.arrival_s <= now();
    if ((arrived || !resident.empty()) && now() - last_progress > kStallLimit) {
      mcgpu::set_error("mcgpu_server_run: no progress for 60 s (" +
                       std::to_string(resident.size()) + " resident, " +
                       std::to_string(completed) + " done)");
      return MCGPU_ERR_STATE;
    }
    if (resident.empty()) {
      if (!arrived) last_progress = now();
      if (next < order.size()) {
        const double wait = s->jobs[order[next]].res.arrival_s - now();
        if (wait > 0)
          std::this_thread::sleep_for(std::chrono::duration<double>(std::min(wait, 0.001)));
      }
      continue;
    }
    // The tick runs asynchronously: jobs that arrive meanwhile are admitted
    // (host set-up and queued uploads) while it runs; retirement below
    // waits for it.
    const double tt = now();
    int rc = mcgpu_batch_run(s->batch, s->cfg.sweeps_per_tick, 0);
    if (rc != MCGPU_OK) return rc;
    t_run += now() - tt;
    {
      const double t2 = now();
      while (next < order.size() && s->jobs[order[next]].res.arrival_s <= t2) {
        Job& j = s->jobs[order[next]];
        const mcgpu_box_desc d = mcgpu::hostbox_ref(j.box).desc();
        int32_t slot = -1;
        rc = mcgpu_batch_add(s->batch, &d, &j.params, &slot);
        if (rc == MCGPU_ERR_NOMEM) break;  // wait for room after retirement
        next++;
        if (rc != MCGPU_OK) {
          j.res.status = rc;
          j.res.admit_s = j.res.finish_s = now();
          completed++;
          continue;
        }
        j.slot = slot;
        j.res.admit_s = now();
        resident.push_back((int32_t)order[next - 1]);
      }
      t_admit += now() - t2;
    }
    s->stats.ticks++;
    last_progress = now();
    if (log_every > 0 && now() - tt > log_every) {
      std::fprintf(stderr, "[server] slow tick %lld: %.2fs with %zu resident\n",
                   (long long)s->stats.ticks, now() - tt, resident.size());
      std::fflush(stderr);
    }
    if (log_every > 0 && now() >= next_log) {
      double kms = 0.0;
      int64_t nl = 0;
      if (mcgpu_batch_kernel_ms(s->batch, &kms, &nl) == MCGPU_OK) kernel_ms += kms * nl;
      double tw = 0, tv = 0, tl = 0;
      (void)mcgpu_batch_host_seconds(s->batch, &tw, &tv, &tl);
      std::fprintf(stderr, "[server] t=%.2fs ticks=%lld resident=%zu admitted=%zu done=%zu moves=%lld "
                   "host s: admit %.2f run %.2f (kernels %.2f; wait %.2f volume %.2f launch %.2f) "
                   "retire %.2f\n",
                   now(), (long long)s->stats.ticks, resident.size(), next, completed,
                   (long long)s->stats.total_moves, t_admit, t_run, kernel_ms / 1e3, tw, tv, tl,
                   t_retire);
      std::fflush(stderr);
      next_log = now() + log_every;
    }
    // retire finished jobs
    const double tr = now();
    for (size_t k = 0; k < resident.size();) {
      Job& j = s->jobs[resident[k]];
      rc = mcgpu_batch_stats(s->batch, j.slot, &st);
      if (rc != MCGPU_OK) return rc;
      if (st.done) {
        j.res.finish_s = now();
        j.res.final_energy = st.energy;
        j.res.start_energy = st.start_energy;
        j.res.accepted = st.accepted;
        j.res.steps = st.steps_done;
        j.res.status = st.fault ? MCGPU_ERR_STATE : MCGPU_OK;
        if (st.fault && log_every > 0) {
          std::fprintf(stderr, "[server] job %lld stopped by device fault %d after %lld moves\n",
                       (long long)j.res.job_id, st.fault, (long long)st.steps_done);
          std::fflush(stderr);
        }
        s->stats.total_moves += st.steps_done;
        s->stats.work_candidates += st.work_candidates;
        s->stats.work_sites += st.work_sites;
        rc = mcgpu_batch_remove(s->batch, j.slot);
        if (rc != MCGPU_OK) return rc;
        j.slot = -1;
        completed++;
        resident[k] = resident.back();
        resident.pop_back();
      } else {
        k++;
      }
    }
    t_retire += now() - tr;
  }
  s->stats.wall_s = now();
  s->stats.peak_resident = peak;
  size_t used = 0, pk = 0, capb = 0;
  if (mcgpu_batch_memory(s->batch, &used, &pk, &capb) == MCGPU_OK) s->stats.peak_bytes = pk;
  double avg_ms = 0;
  int64_t launches = 0;
  if (mcgpu_batch_kernel_ms(s->batch, &avg_ms, &launches) == MCGPU_OK)
    kernel_ms += avg_ms * launches;
  s->stats.kernel_ms_total = kernel_ms;
  return MCGPU_OK;
}

extern "C" int mcgpu_server_results(mcgpu_server* s, mcgpu_job_result* out,
                                    int64_t cap, int64_t* n) {
  if (!s || !n) return MCGPU_ERR_INVALID;
  const int64_t k = std::min<int64_t>(cap, (int64_t)s->jobs.size());
  for (int64_t i = 0; i < k; i++) out[i] = s->jobs[i].res;
  *n = k;
  return MCGPU_OK;
}

extern "C" int mcgpu_server_get_stats(mcgpu_server* s, mcgpu_server_stats* out) {
  if (!s || !out) return MCGPU_ERR_INVALID;
  *out = s->stats;
  return MCGPU_OK;
}
