// Serving loop (lpmdx_server_* in include/lpmdx.h): a worker thread owns one
// lpmdx_batch whose arenas are sized from an HBM budget.  Between rounds of
// graph-replayed steps it retires finished jobs (results copied to host,
// arena slices freed) and admits queued jobs that now fit, so jobs that
// arrive while others run join them on the device.  Every batch call is made
// by the worker; submitters only touch the queue under the mutex.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "host_util.hpp"

using namespace lpmdx;

namespace {

using Clock = std::chrono::steady_clock;

// Steps replayed between two looks at the queue (an arrival waits at most
// this many steps before it is admitted).
constexpr long kChunkSteps = 16;
// Admission passes a blocked head job lets later, smaller jobs go first
// before the queue stops for it.
constexpr int kMaxBypass = 4;

struct Job {
  long id = 0;
  long n = 0;
  double box[3] = {0, 0, 0};
  std::vector<int> z;
  std::vector<double> mass, x, v;
  std::string cm;
  std::vector<std::string> names, args;
  std::vector<int> s1, s2;
  double dt = 0.0;
  long steps = 0;
  int state = 0;  // 0 queued, 1 resident, 2 finished
  int status = 0; // LPMDX_OK or the code the job failed with
  std::string err;
  long sid = -1;
  int bypassed = 0;
  Clock::time_point t_submit, t_admit, t_done;
  std::vector<double> xo, vo, ao;
  double e[3] = {0, 0, 0};
  long builds = 0, steps_done = 0;
};

double seconds(Clock::time_point a, Clock::time_point b) { return std::chrono::duration<double>(b - a).count(); }

// numpy's default ("linear") percentile of sorted values
double percentile(const std::vector<double>& s, double q) {
  if (s.empty())
    return 0.0;
  const double pos = q * (double)(s.size() - 1);
  const size_t lo = (size_t)std::floor(pos);
  const size_t hi = std::min(lo + 1, s.size() - 1);
  return s[lo] + (s[hi] - s[lo]) * (pos - (double)lo);
}

} // namespace

struct lpmdx_server {
  lpmdx_batch* b = nullptr;
  int device = 0;
  long max_round = 256;
  std::thread worker;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  bool stop = false;
  long next_id = 0;
  long arrivals = 0; // submissions so far (the worker compares it between chunks)
  std::map<long, std::unique_ptr<Job>> jobs;
  std::deque<Job*> queue;
  std::map<long, Job*> resident; // batch slot -> job (worker-owned; changed under mu)
  int lanes[LPMDX_NKINDS + 1] = {0};
  bool lanes_dirty = false;
  // statistics (under mu)
  std::vector<double> latency;
  long n_done = 0, n_failed = 0, rounds = 0, max_resident = 0, graph_steps = 0;
  double atom_steps = 0.0;
  // worker seconds: admitting jobs, running steps, retiring jobs
  double t_admit = 0.0, t_run = 0.0, t_retire = 0.0;
  bool any = false;
  Clock::time_point t_first, t_last;

  ~lpmdx_server() {
    {
      std::lock_guard<std::mutex> lk(mu);
      stop = true;
    }
    cv_work.notify_all();
    if (worker.joinable())
      worker.join();
    if (b)
      (void)lpmdx_batch_destroy(b);
  }

  // ------------------------------------------------------------------ worker
  void finish(Job* j, int status, const std::string& err) {
    // caller holds mu
    j->state = 2;
    j->status = status;
    j->err = err;
    j->t_done = Clock::now();
    t_last = j->t_done;
    if (status == LPMDX_OK) {
      ++n_done;
      latency.push_back(seconds(j->t_submit, j->t_done));
      atom_steps += (double)j->n * (double)j->steps_done;
    } else {
      ++n_failed;
    }
  }

  void admit(long arrivals_seen, long& admitted_through) {
    std::vector<Job*> cand;
    {
      std::lock_guard<std::mutex> lk(mu);
      cand.assign(queue.begin(), queue.end());
      admitted_through = arrivals_seen;
    }
    bool blocked = false;
    for (Job* j : cand) {
      if (blocked)
        break;
      std::vector<const char*> nm, ar;
      for (size_t q = 0; q < j->names.size(); ++q) {
        nm.push_back(j->names[q].c_str());
        ar.push_back(j->args[q].c_str());
      }
      long sid = -1;
      const int rc = lpmdx_batch_add(b, j->n, j->box, j->z.data(), j->mass.empty() ? nullptr : j->mass.data(),
                                     j->x.data(), j->v.empty() ? nullptr : j->v.data(), j->cm.c_str(),
                                     (int)nm.size(), nm.data(), ar.data(), j->s1.data(), j->s2.data(), j->dt,
                                     j->steps, &sid);
      std::lock_guard<std::mutex> lk(mu);
      if (rc == LPMDX_OK) {
        j->state = 1;
        j->sid = sid;
        j->t_admit = Clock::now();
        resident[sid] = j;
        queue.erase(std::find(queue.begin(), queue.end(), j));
        max_resident = std::max<long>(max_resident, (long)resident.size());
        // inputs are on the device now
        std::vector<double>().swap(j->x);
        std::vector<double>().swap(j->v);
      } else if (rc == LPMDX_ENOMEM && !resident.empty()) {
        // no room yet: later jobs may go first a few times, then the queue waits
        if (++j->bypassed > kMaxBypass)
          blocked = true;
      } else {
        // invalid, or larger than the whole budget
        queue.erase(std::find(queue.begin(), queue.end(), j));
        finish(j, rc, lpmdx_last_error());
        cv_done.notify_all();
      }
    }
  }

  // Retires finished (or overflowed) jobs; returns false on a device error.
  bool retire(std::vector<long>& left, std::vector<int>& flags) {
    long nslots = resident.empty() ? 0 : resident.rbegin()->first + 1;
    left.assign(nslots, 0);
    flags.assign(nslots, 0);
    if (lpmdx_batch_progress(b, nslots, left.data(), flags.data()) != LPMDX_OK)
      return false;
    std::vector<std::pair<long, Job*>> res(resident.begin(), resident.end());
    for (auto& kv : res) {
      const long sid = kv.first;
      Job* j = kv.second;
      const bool overflow = flags[sid] & 2;
      if (!overflow && left[sid] > 0)
        continue;
      int status = LPMDX_OK;
      std::string err;
      long info[8] = {0};
      if (overflow) {
        status = LPMDX_ECAPACITY;
        err = "neighbour list capacity exceeded (raise maxneighbors)";
      } else {
        j->xo.resize(3 * j->n);
        j->vo.resize(3 * j->n);
        j->ao.resize(3 * j->n);
        int rc = lpmdx_batch_get(b, sid, j->xo.data(), j->vo.data(), j->ao.data(), info);
        if (rc == LPMDX_OK)
          rc = lpmdx_batch_energies(b, sid, j->e);
        if (rc != LPMDX_OK) {
          status = rc;
          err = lpmdx_last_error();
        }
        j->steps_done = info[0];
        j->builds = info[2];
      }
      if (lpmdx_batch_remove(b, sid) != LPMDX_OK)
        return false;
      std::lock_guard<std::mutex> lk(mu);
      resident.erase(sid);
      finish(j, status, err);
    }
    cv_done.notify_all();
    return true;
  }

  void fail_all(int code, const std::string& err) {
    std::lock_guard<std::mutex> lk(mu);
    for (auto& kv : resident)
      finish(kv.second, code, err);
    resident.clear();
    for (Job* j : queue)
      finish(j, code, err);
    queue.clear();
    stop = true;
    cv_done.notify_all();
  }

  void loop() {
    (void)hipSetDevice(device);
    std::vector<long> left;
    std::vector<int> flags;
    long seen = 0;
    for (;;) {
      long arr = 0;
      {
        std::unique_lock<std::mutex> lk(mu);
        cv_work.wait(lk, [&] { return stop || !queue.empty() || !resident.empty(); });
        if (stop)
          return;
        arr = arrivals;
        if (lanes_dirty) {
          for (int k = 1; k <= LPMDX_NKINDS; ++k)
            if (lanes[k])
              (void)lpmdx_batch_set_lanes(b, k, lanes[k]);
          lanes_dirty = false;
        }
      }
      auto t0 = std::chrono::steady_clock::now();
      admit(arr, seen);
      auto t1 = std::chrono::steady_clock::now();
      {
        std::lock_guard<std::mutex> lk(mu);
        t_admit += std::chrono::duration<double>(t1 - t0).count();
      }
      if (resident.empty())
        continue; // everything queued failed admission
      if (!retire(left, flags)) { // zero-step jobs finish at once; also sizes the round
        fail_all(LPMDX_EDEVICE, lpmdx_last_error());
        return;
      }
      auto t2 = std::chrono::steady_clock::now();
      if (resident.empty())
        continue;
      long round = max_round;
      for (auto& kv : resident)
        round = std::min(round, std::max<long>(1, left[kv.first]));
      long ran = 0;
      while (ran < round) {
        const long c = std::min(kChunkSteps, round - ran);
        const int rc = lpmdx_batch_run(b, c);
        if (rc != LPMDX_OK && rc != LPMDX_ECAPACITY) {
          fail_all(rc, lpmdx_last_error());
          return;
        }
        ran += c;
        std::lock_guard<std::mutex> lk(mu);
        graph_steps += c;
        if (rc == LPMDX_ECAPACITY || stop || arrivals != seen)
          break; // overflowed job to retire, or new arrivals to admit
      }
      auto t3 = std::chrono::steady_clock::now();
      if (!retire(left, flags)) {
        fail_all(LPMDX_EDEVICE, lpmdx_last_error());
        return;
      }
      auto t4 = std::chrono::steady_clock::now();
      {
        std::lock_guard<std::mutex> lk(mu);
        ++rounds;
        t_retire += std::chrono::duration<double>((t2 - t1) + (t4 - t3)).count();
        t_run += std::chrono::duration<double>(t3 - t2).count();
      }
    }
  }

  Job* find(long id) {
    auto it = jobs.find(id);
    if (it == jobs.end())
      fail(LPMDX_EINVAL, "no such job");
    return it->second.get();
  }
};

extern "C" {

int lpmdx_server_create(lpmdx_server** out, long hbm_budget, long max_round_steps) {
  return guard([&] {
    if (!out)
      fail(LPMDX_EINVAL, "out is NULL");
    *out = nullptr;
    if (hbm_budget < (1L << 20) || max_round_steps <= 0)
      fail(LPMDX_EINVAL, "hbm_budget must be at least 1 MiB and max_round_steps positive");
    auto S = std::make_unique<lpmdx_server>();
    hipchk(hipGetDevice(&S->device), "hipGetDevice");
    S->max_round = max_round_steps;
    // Lists take most of a job's memory (about 100-300 words per atom for the
    // usual cutoffs and skins), so a quarter of the budget goes to the
    // per-atom arenas, 72 % to list words and the rest to cell ranges.
    const double bud = (double)hbm_budget;
    const long atoms = std::max<long>(64, (long)(0.25 * bud / (double)lpmdx_batch_bytes_per_atom()));
    const long words = std::max<long>(64, (long)(0.72 * bud / 4.0));
    const long cells = std::max<long>(64, (long)(0.03 * bud / 8.0));
    const int rc = lpmdx_batch_create(&S->b, atoms, words, cells);
    if (rc != LPMDX_OK)
      fail(rc, lpmdx_last_error());
    lpmdx_server* p = S.get();
    S->worker = std::thread([p] { p->loop(); });
    *out = S.release();
  });
}

int lpmdx_server_destroy(lpmdx_server* s) {
  return guard([&] { delete s; });
}

int lpmdx_server_submit(lpmdx_server* s, long n, const double box[3], const int* z, const double* mass,
                        const double* x, const double* v, const char* cm_args, int npot, const char** names,
                        const char** args, const int* s1, const int* s2, double dt, long steps, long* job) {
  return guard([&] {
    if (!s || !job || !box || !z || !x || !names)
      fail(LPMDX_EINVAL, "server, job, box, z, x or names is NULL");
    if (n <= 0 || npot < 1 || npot > 4 || steps < 0)
      fail(LPMDX_EINVAL, "a job needs atoms, 1 to 4 potentials and a non-negative step count");
    auto j = std::make_unique<Job>();
    j->n = n;
    std::copy(box, box + 3, j->box);
    j->z.assign(z, z + n);
    if (mass)
      j->mass.assign(mass, mass + n);
    j->x.assign(x, x + 3 * n);
    if (v)
      j->v.assign(v, v + 3 * n);
    j->cm = cm_args ? cm_args : "";
    for (int q = 0; q < npot; ++q) {
      if (!names[q])
        fail(LPMDX_EINVAL, "potential name is NULL");
      j->names.push_back(names[q]);
      j->args.push_back(args && args[q] ? args[q] : "");
      j->s1.push_back(s1 ? s1[q] : -1);
      j->s2.push_back(s2 ? s2[q] : -1);
    }
    j->dt = dt;
    j->steps = steps;
    {
      std::lock_guard<std::mutex> lk(s->mu);
      if (s->stop)
        fail(LPMDX_ESTATE, "the server has stopped");
      j->id = s->next_id++;
      j->t_submit = Clock::now();
      if (!s->any) {
        s->any = true;
        s->t_first = j->t_submit;
      }
      *job = j->id;
      s->queue.push_back(j.get());
      s->jobs[j->id] = std::move(j);
      ++s->arrivals;
    }
    s->cv_work.notify_all();
  });
}

int lpmdx_server_wait(lpmdx_server* s, long job, double timeout_s, int* done) {
  return guard([&] {
    if (!s || !done)
      fail(LPMDX_EINVAL, "server or done is NULL");
    std::unique_lock<std::mutex> lk(s->mu);
    Job* j = job >= 0 ? s->find(job) : nullptr;
    auto ready = [&] {
      if (j)
        return j->state == 2;
      return s->queue.empty() && s->resident.empty();
    };
    if (timeout_s < 0)
      s->cv_done.wait(lk, ready);
    else
      s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), ready);
    *done = ready() ? 1 : 0;
  });
}

int lpmdx_server_result(lpmdx_server* s, long job, double* x, double* v, double* a, double out[8]) {
  return guard([&] {
    if (!s)
      fail(LPMDX_EINVAL, "server is NULL");
    std::lock_guard<std::mutex> lk(s->mu);
    Job* j = s->find(job);
    if (j->state != 2)
      fail(LPMDX_ESTATE, "the job has not finished");
    if (j->status == LPMDX_OK) {
      const size_t m = 3 * (size_t)j->n;
      if (x)
        std::copy(j->xo.begin(), j->xo.begin() + m, x);
      if (v)
        std::copy(j->vo.begin(), j->vo.begin() + m, v);
      if (a)
        std::copy(j->ao.begin(), j->ao.begin() + m, a);
    }
    if (out) {
      out[0] = j->e[0];
      out[1] = j->e[1];
      out[2] = j->e[2];
      out[3] = j->sid >= 0 ? seconds(j->t_submit, j->t_admit) : seconds(j->t_submit, j->t_done);
      out[4] = seconds(j->t_submit, j->t_done);
      out[5] = (double)j->builds;
      out[6] = (double)j->steps_done;
      out[7] = (double)j->status;
    }
    if (j->status != LPMDX_OK)
      g_err = j->err;
  });
}

int lpmdx_server_release(lpmdx_server* s, long job) {
  return guard([&] {
    if (!s)
      fail(LPMDX_EINVAL, "server is NULL");
    std::lock_guard<std::mutex> lk(s->mu);
    Job* j = s->find(job);
    if (j->state != 2)
      fail(LPMDX_ESTATE, "the job has not finished");
    s->jobs.erase(job);
  });
}

int lpmdx_server_stats(lpmdx_server* s, double out[LPMDX_NSERVER_STATS]) {
  return guard([&] {
    if (!s || !out)
      fail(LPMDX_EINVAL, "server or out is NULL");
    std::lock_guard<std::mutex> lk(s->mu);
    std::vector<double> l = s->latency;
    std::sort(l.begin(), l.end());
    double mean = 0.0;
    for (double t : l)
      mean += t;
    out[0] = (double)s->n_done;
    out[1] = (double)s->n_failed;
    out[2] = percentile(l, 0.50);
    out[3] = percentile(l, 0.99);
    out[4] = l.empty() ? 0.0 : mean / (double)l.size();
    out[5] = s->atom_steps;
    out[6] = s->any && (s->n_done + s->n_failed) ? seconds(s->t_first, s->t_last) : 0.0;
    out[7] = (double)s->rounds;
    out[8] = (double)s->max_resident;
    out[9] = (double)s->queue.size();
    out[10] = (double)s->resident.size();
    out[11] = (double)s->graph_steps;
    out[12] = s->t_admit;
    out[13] = s->t_run;
    out[14] = s->t_retire;
    out[15] = 0.0;
  });
}

int lpmdx_server_set_lanes(lpmdx_server* s, int kind, int lanes) {
  return guard([&] {
    if (!s)
      fail(LPMDX_EINVAL, "server is NULL");
    if (kind < 1 || kind > LPMDX_NKINDS || !(lanes == 1 || lanes == 2 || lanes == 4))
      fail(LPMDX_EINVAL, "kind 1.." + std::to_string(LPMDX_NKINDS) + " and lanes 1, 2 or 4");
    std::lock_guard<std::mutex> lk(s->mu);
    s->lanes[kind] = lanes;
    s->lanes_dirty = true;
  });
}

} // extern "C"
