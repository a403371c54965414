// pool.hip -- lanes, HBM-budget admission, and the batch / serve entry points.
//
// A lane is one context (HIP stream + grow-only work buffers) driven by one
// host worker thread.  Every lane runs the same single-field correction as
// extremz_correct, so a job's result does not depend on what runs beside it;
// the lanes' kernels overlap on the device (each stream is its own queue of
// dependent launches, and the host waits of one lane hide behind the kernels
// of the others).
//
// Admission: a job may start on an idle lane when the pool's device
// footprint, with that lane's charge raised to the job's working-set
// estimate, stays within the budget.  After each job the lane's charge is
// reset to the bytes its buffers actually hold.  An idle lane that cannot
// admit anything while jobs wait frees its buffers, so memory held by idle
// lanes never blocks a job.  Jobs behind a job that does not fit may pass it,
// until it has been passed `lanes` times; then it waits at the head.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "engine.h"

namespace {

using Clock = std::chrono::steady_clock;

// Device bytes of one correction of N voxels, in two parts.  fixed: the
// per-voxel work arrays every correction sizes by N (232 B, 31 of them the
// change lists, pointers and flags of the incremental CP pass) plus the
// small per-extremum arrays.  var: what depends on the data or the caller's
// buffers -- the saddle pair records, side and slot lists, stamps, indices
// and keys (92 B/voxel covers the 0.43 saddle sides per voxel of white noise;
// smooth fields have fewer) -- and the staging of host buffers / float32
// originals.
struct Need {
  int64_t fixed, var;
};
Need working_set(const extremz_job& j) {
  const int64_t N = (int64_t)j.W * j.H * j.D;
  Need n{232 * N + (1 << 20), 92 * N};
  if (j.on_host || j.orig_dtype == EXTREMZ_DTYPE_F32) n.var += 8 * N;  // orig as float64
  if (j.on_host) n.var += 9 * N;                                       // decp + counts
  return n;
}

struct PJob {
  extremz_job* job = nullptr;
  int64_t ticket = 0;
  Need need{0, 0};    // working-set estimate, bytes
  int passed = 0;     // admissions that passed this job over
  Clock::time_point t_submit;
  bool done = false;
  int status = 0;
  std::string err;
};

struct Lane {
  extremz_ctx* ctx = nullptr;
  int64_t charged = 0;  // bytes charged to this lane's buffers
  std::thread th;
};

double ms_between(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

}  // namespace

struct extremz_pool {
  int device = 0;
  int64_t budget = 0;
  extremz_opts opts{};
  std::mutex mu;
  std::condition_variable cv_work;  // lanes: a job or memory may be available
  std::condition_variable cv_done;  // waiters: a job finished
  std::deque<PJob*> queue;
  std::unordered_map<int64_t, PJob*> live;  // submitted, not yet waited for
  std::deque<int64_t> finished;             // finished, not yet waited for
  std::vector<Lane> lanes;
  int64_t footprint = 0, peak = 0;
  int64_t next_ticket = 1, submitted = 0, completed = 0, running = 0;
  bool stop = false;

  // Picks the job lane l may start now (caller holds mu), or nullptr.
  PJob* admit(Lane& l) {
    for (size_t i = 0; i < queue.size(); i++) {
      PJob* j = queue[i];
      // buffers grow one by one to the larger of what they hold and what the
      // job needs: the data-dependent part is charged on top
      const int64_t charge = std::max(l.charged, j->need.fixed) + j->need.var;
      if (footprint - l.charged + charge <= budget) {
        for (size_t k = 0; k < i; k++) queue[k]->passed++;
        queue.erase(queue.begin() + (long)i);
        footprint += charge - l.charged;
        l.charged = charge;
        peak = std::max(peak, footprint);
        return j;
      }
      if (j->passed >= (int)lanes.size()) break;  // it waits at the head now
    }
    return nullptr;
  }

  void worker(int li) {
    Lane& l = lanes[li];
    (void)hipSetDevice(device);
    std::unique_lock<std::mutex> lk(mu);
    while (true) {
      if (stop && queue.empty()) break;
      PJob* j = admit(l);
      if (!j) {
        if (!queue.empty() && l.charged > 0) {
          // idle buffers stand in the way of a waiting job: give them back
          lk.unlock();
          ez::ctx_release(l.ctx);
          lk.lock();
          footprint -= l.charged;
          l.charged = 0;
          cv_work.notify_all();
          continue;
        }
        cv_work.wait(lk);
        continue;
      }
      running++;
      lk.unlock();
      const Clock::time_point t0 = Clock::now();
      extremz_job& job = *j->job;
      job.wait_ms = ms_between(j->t_submit, t0);
      job.lane = li;
      int rc = ez::guarded([&] {
        return ez::correct_staged(l.ctx, job.orig, job.orig_dtype, job.on_host != 0, job.decp,
                                  job.counts, job.W, job.H, job.D, job.xi, &opts, &job.stats);
      });
      std::string err = ez::g_err;
      if (rc == EXTREMZ_EHIP) (void)hipGetLastError();
      const int64_t held = (int64_t)ez::ctx_bytes(l.ctx);
      job.run_ms = ms_between(t0, Clock::now());
      job.status = rc;
      lk.lock();
      footprint += held - l.charged;
      l.charged = held;
      peak = std::max(peak, footprint);
      j->status = rc;
      j->err = err;
      j->done = true;
      running--;
      completed++;
      finished.push_back(j->ticket);
      cv_done.notify_all();
      cv_work.notify_all();
    }
  }

  // Takes a finished job out of the books (caller holds mu).
  int retire(PJob* j) {
    live.erase(j->ticket);
    auto it = std::find(finished.begin(), finished.end(), j->ticket);
    if (it != finished.end()) finished.erase(it);
    int rc = j->status;
    ez::g_err = j->err;
    delete j;
    return rc;
  }
};

namespace {

bool job_valid(const extremz_job* j) {
  return j && j->orig && j->decp && ez::valid_shape(j->W, j->H, j->D) && j->xi > 0 &&
         std::isfinite(j->xi) &&
         (j->orig_dtype == EXTREMZ_DTYPE_F32 || j->orig_dtype == EXTREMZ_DTYPE_F64);
}

int submit(extremz_pool* p, extremz_job* job, int64_t* ticket) {
  if (!job_valid(job)) {
    if (job) job->status = EXTREMZ_EINVAL;
    return EXTREMZ_EINVAL;
  }
  PJob* j = new PJob();
  j->job = job;
  j->need = working_set(*job);
  j->t_submit = Clock::now();
  job->status = EXTREMZ_EAGAIN;
  job->lane = -1;
  job->wait_ms = job->run_ms = 0.0;
  std::lock_guard<std::mutex> lk(p->mu);
  j->ticket = p->next_ticket++;
  *ticket = j->ticket;
  p->live[j->ticket] = j;
  p->submitted++;
  if (j->need.fixed + j->need.var > p->budget) {  // can never be admitted
    j->status = job->status = EXTREMZ_ENOMEM;
    j->err = "job working set exceeds the pool's HBM budget";
    j->done = true;
    p->completed++;
    p->finished.push_back(j->ticket);
    p->cv_done.notify_all();
    return EXTREMZ_OK;
  }
  p->queue.push_back(j);
  p->cv_work.notify_all();
  return EXTREMZ_OK;
}

// Another thread may wait for the same ticket and retire (free) its job
// first, so the job is looked up again after every wake, never held.
int wait_ticket(extremz_pool* p, int64_t ticket, int timeout_ms) {
  std::unique_lock<std::mutex> lk(p->mu);
  if (p->live.find(ticket) == p->live.end()) {
    ez::g_err = "unknown ticket";
    return EXTREMZ_EINVAL;
  }
  auto settled = [&] {
    auto it = p->live.find(ticket);
    return it == p->live.end() || it->second->done;
  };
  if (timeout_ms < 0) {
    p->cv_done.wait(lk, settled);
  } else if (!p->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), settled)) {
    return EXTREMZ_EAGAIN;
  }
  auto it = p->live.find(ticket);
  if (it == p->live.end()) {
    ez::g_err = "ticket already waited for by another caller";
    return EXTREMZ_EINVAL;
  }
  return p->retire(it->second);
}

}  // namespace

extern "C" {

int extremz_pool_create(int device, int lanes, int64_t hbm_budget, const extremz_opts* opts,
                        extremz_pool** out) {
  if (!out || lanes < 1 || lanes > 64 || hbm_budget < 0 || !ez::valid_opts(opts))
    return EXTREMZ_EINVAL;
  *out = nullptr;
  extremz_pool* p = new extremz_pool();
  int rc = ez::guarded([&] {
    p->device = device;
    if (opts) p->opts = *opts;
    p->lanes.resize(lanes);
    for (Lane& l : p->lanes) {
      l.ctx = ez::new_ctx(device);
      if (!l.ctx) {
        ez::g_err = "no gfx950 device with that ordinal";
        return EXTREMZ_ENODEV;
      }
    }
    size_t free_b = 0, total_b = 0;
    ez::hip_check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
    p->budget = hbm_budget > 0 ? hbm_budget : (int64_t)(0.8 * (double)total_b);
    return EXTREMZ_OK;
  });
  if (rc != EXTREMZ_OK) {
    for (Lane& l : p->lanes) ez::free_ctx(l.ctx);
    delete p;
    return rc;
  }
  for (int i = 0; i < lanes; i++) p->lanes[i].th = std::thread([p, i] { p->worker(i); });
  *out = p;
  return EXTREMZ_OK;
}

int extremz_pool_destroy(extremz_pool* p) {
  if (!p) return EXTREMZ_OK;
  {
    std::lock_guard<std::mutex> lk(p->mu);
    p->stop = true;
    p->cv_work.notify_all();
  }
  for (Lane& l : p->lanes)
    if (l.th.joinable()) l.th.join();
  for (Lane& l : p->lanes) ez::free_ctx(l.ctx);
  for (auto& kv : p->live) delete kv.second;
  delete p;
  return EXTREMZ_OK;
}

int extremz_pool_info_get(extremz_pool* p, extremz_pool_info* info) {
  if (!p || !info) return EXTREMZ_EINVAL;
  std::lock_guard<std::mutex> lk(p->mu);
  info->budget_bytes = p->budget;
  info->footprint_bytes = p->footprint;
  info->peak_footprint_bytes = p->peak;
  info->submitted = p->submitted;
  info->completed = p->completed;
  info->running = p->running;
  info->queued = (int64_t)p->queue.size();
  info->lanes = (int)p->lanes.size();
  info->reserved = 0;
  return EXTREMZ_OK;
}

int extremz_serve_submit(extremz_pool* p, extremz_job* job, int64_t* ticket) {
  if (!p || !ticket) return EXTREMZ_EINVAL;
  return submit(p, job, ticket);
}

int extremz_serve_wait(extremz_pool* p, int64_t ticket, int timeout_ms) {
  if (!p) return EXTREMZ_EINVAL;
  return wait_ticket(p, ticket, timeout_ms);
}

int extremz_serve_wait_any(extremz_pool* p, int timeout_ms, int64_t* ticket) {
  if (!p || !ticket) return EXTREMZ_EINVAL;
  std::unique_lock<std::mutex> lk(p->mu);
  auto ready = [&] { return !p->finished.empty(); };
  if (timeout_ms < 0) {
    p->cv_done.wait(lk, ready);
  } else if (!p->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready)) {
    return EXTREMZ_EAGAIN;
  }
  *ticket = p->finished.front();
  return p->retire(p->live.at(*ticket));
}

int extremz_batch_correct(extremz_pool* p, extremz_job* jobs, int n) {
  if (!p || n < 0 || (n > 0 && !jobs)) return EXTREMZ_EINVAL;
  // largest first: the long jobs start early and the short ones fill in
  std::vector<int> order(n);
  for (int i = 0; i < n; i++) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    return (int64_t)jobs[a].W * jobs[a].H * jobs[a].D > (int64_t)jobs[b].W * jobs[b].H * jobs[b].D;
  });
  std::vector<int64_t> tickets(n, 0);
  for (int i : order)
    if (submit(p, &jobs[i], &tickets[i]) != EXTREMZ_OK) tickets[i] = 0;
  int first = EXTREMZ_OK;
  std::string first_err;
  for (int i = 0; i < n; i++) {
    int rc = tickets[i] ? wait_ticket(p, tickets[i], -1) : jobs[i].status;
    if (rc != EXTREMZ_OK && first == EXTREMZ_OK) {
      first = rc;
      first_err = ez::g_err;
    }
  }
  ez::g_err = first_err;
  return first;
}

}  // extern "C"
