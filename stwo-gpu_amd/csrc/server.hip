// Proof serving with continuous batching.
//
// Requests (wide-Fibonacci statements + trace seeds) queue in arrival order.  A worker
// thread admits the head of the queue as soon as a slot stream is free and the job's
// device footprint fits the HBM budget next to the jobs in flight, allocates its buffers
// and enqueues the whole proof on the slot's stream (no host round trip inside a proof,
// see prover.hip), so jobs that arrive while others run overlap with them on the
// device.  Finished jobs (their completion event has fired) release memory and slot at
// once, which admits the next waiting job.  The twiddle trees (one per log size) stay
// cached on the device for every later job.
//
// A work limit (trace elements in flight, n_cols x 2^log_n per job) keeps large,
// device-filling jobs from crowding each other while small ones still overlap.  A job
// larger than the whole budget (or the work limit) is admitted alone.  Results are returned in
// completion order with their submit / start / end times.
//
// Same-phase batching across jobs: with a batch limit set, admitting a small job (at most
// max_batch_work trace elements) also takes up to max_batch - 1 queued jobs of the same
// statement shape, and the group is proved as one batch (prover::Job with n_jobs > 1):
// every phase of the proof is one launch covering all of its jobs, so a burst of small
// proofs costs one proof's chain of launches instead of one chain per proof.
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "prover.hpp"

namespace sh {
namespace serve {

using Clock = std::chrono::steady_clock;

struct Request {
  uint64_t id;
  prover::Params p;
  uint64_t seed;
  double t_submit;
};
struct Running {
  std::vector<Request> rs;  // the batch's requests (job k of the batch proves rs[k])
  std::unique_ptr<prover::Job> job;
  int slot;
  uint64_t bytes, work;  // device footprint, trace elements
  double t_start, host_s;  // admission time, host time spent building and enqueuing the job
};
struct Result {
  uint64_t id;
  int status;
  std::vector<uint32_t> proof;
  double t_submit, t_start, t_end, host_s;
  std::string error;  // the worker's error message when status != OK
};

// trace elements of a job (n_cols x 2^log_n): its share of the device's compute
inline uint64_t job_work(const prover::Params& p) { return (uint64_t)p.n_cols << p.log_n; }

struct Server {
  uint64_t budget;
  uint32_t max_inflight;
  std::atomic<uint64_t> work_limit{0};  // 0: none
  std::atomic<uint32_t> max_batch{1};   // same-shape jobs proved together (1: no batching)
  std::atomic<uint64_t> max_batch_work{0};  // only jobs of at most this many trace elements batch
  Clock::time_point t0 = Clock::now();
  std::vector<hipStream_t> streams;
  std::vector<bool> slot_busy;

  std::mutex m;
  std::condition_variable cv_work, cv_done;
  std::deque<Request> queue;
  std::deque<Result> results;
  uint64_t next_id = 1;
  bool stopping = false;

  std::vector<Running> running;  // worker thread only
  uint64_t inflight_bytes = 0, inflight_work = 0;
  std::thread worker;

  double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }

  void finish(Running& r, int status) {
    const double t_end = now();
    std::vector<Result> out;
    for (size_t k = 0; k < r.rs.size(); k++) {
      Result res{r.rs[k].id, status, {}, r.rs[k].t_submit, r.t_start, t_end, r.host_s,
                 status == STWO_HIP_OK ? std::string() : std::string("job failed on the device")};
      if (status == STWO_HIP_OK) {
        const uint32_t* pr = r.job->result((uint32_t)k);
        res.proof.assign(pr, pr + r.job->L.total);
        // no proof-of-work nonce found within the search bound
        if (pr[r.job->L.nonce] == 0xffffffffu && pr[r.job->L.nonce + 1] == 0xffffffffu) {
          res.status = STWO_HIP_ERR_ARG;
          res.error = "no proof-of-work nonce within the search bound";
        }
      }
      out.push_back(std::move(res));
    }
    {
      auto g = lock();
      r.job.reset();
    }
    slot_busy[r.slot] = false;
    inflight_bytes -= r.bytes;
    inflight_work -= r.work;
    std::lock_guard<std::mutex> l(m);
    for (auto& res : out) results.push_back(std::move(res));
    cv_done.notify_all();
  }

  void loop() {
    for (;;) {
      bool progressed = false;
      // completions
      for (size_t i = 0; i < running.size();) {
        const hipError_t q = hipEventQuery(running[i].job->done);
        if (q == hipErrorNotReady) {
          i++;
          continue;
        }
        finish(running[i], q == hipSuccess ? STWO_HIP_OK : STWO_HIP_ERR_HIP);
        running.erase(running.begin() + i);
        progressed = true;
      }
      // admissions, in arrival order
      for (;;) {
        std::vector<Request> rs;
        uint64_t need = 0, w = 0;
        {
          std::lock_guard<std::mutex> l(m);
          if (queue.empty() || running.size() >= max_inflight) break;
          const prover::Params p = queue.front().p;
          // the head job, and (small jobs, batching on) queued jobs of the same shape with it
          uint32_t k = 1;
          const uint32_t mb = max_batch.load();
          if (mb > 1 && job_work(p) <= max_batch_work.load())
            for (size_t i = 1; i < queue.size() && k < mb; i++) k += queue[i].p == p;
          const uint64_t wl = work_limit.load();
          // a batch that does not fit next to the jobs in flight shrinks before it waits
          for (; k >= 1; k--) {
            need = prover::footprint_bytes(p, k);
            w = job_work(p) * k;
            if (running.empty() || (inflight_bytes + need <= budget && (!wl || inflight_work + w <= wl))) break;
          }
          if (k == 0) break;
          rs.push_back(queue.front());
          queue.pop_front();
          for (size_t i = 0; i < queue.size() && rs.size() < k;) {
            if (queue[i].p == p) {
              rs.push_back(queue[i]);
              queue.erase(queue.begin() + i);
            } else {
              i++;
            }
          }
        }
        int slot = 0;
        while (slot_busy[slot]) slot++;
        Running run{rs, nullptr, slot, need, w, now(), 0.0};
        int st = STWO_HIP_OK;
        {
          auto g = lock();
          try {
            std::vector<uint64_t> seeds;
            for (const Request& q : rs) seeds.push_back(q.seed);
            run.job = std::make_unique<prover::Job>(rs[0].p, streams[slot], (uint32_t)rs.size());
            run.job->enqueue(seeds.data());
            run.host_s = now() - run.t_start;
          } catch (const Failure& f) {
            st = f.status;
            run.job.reset();
          }
        }
        if (st != STWO_HIP_OK) {
          std::lock_guard<std::mutex> l(m);
          for (const Request& q : rs)
            results.push_back(Result{q.id, st, {}, q.t_submit, run.t_start, now(), 0.0, stwo_hip_last_error()});
          cv_done.notify_all();
          continue;
        }
        slot_busy[slot] = true;
        inflight_bytes += run.bytes;
        inflight_work += run.work;
        running.push_back(std::move(run));
        progressed = true;
      }
      if (progressed) continue;
      std::unique_lock<std::mutex> l(m);
      if (stopping && queue.empty() && running.empty()) return;
      if (running.empty())
        cv_work.wait(l, [&] { return stopping || !queue.empty(); });
      else
        cv_work.wait_for(l, std::chrono::microseconds(20));
    }
  }
};

}  // namespace serve
}  // namespace sh

using namespace sh;

struct StwoProofServer {
  serve::Server s;
};

static prover::Params params_of(const StwoWideFibParams* p) {
  require(p != nullptr, "wide fibonacci: null parameters");
  prover::Params q{p->n_cols, p->log_n, p->n_queries, p->pow_bits, p->last_log};
  prover::validate(q);
  return q;
}

extern "C" int stwo_hip_wide_fib_proof_size(const StwoWideFibParams* params, uint64_t* proof_words,
                                            uint64_t* device_bytes) {
  return guarded([&] {
    prover::Params p = params_of(params);
    if (proof_words) *proof_words = prover::Layout(p).total;
    if (device_bytes) *device_bytes = prover::footprint_bytes(p);
  });
}

extern "C" int stwo_hip_wide_fib_prove(const StwoWideFibParams* params, uint64_t seed, uint32_t* proof,
                                       uint64_t proof_capacity_words) {
  return guarded([&] {
    prover::Params p = params_of(params);
    require(proof != nullptr && proof_capacity_words >= prover::Layout(p).total,
            "wide fibonacci prove: proof buffer too small");
    prover::Job job(p, ctx().stream);
    job.enqueue(seed);
    hip_check(hipEventSynchronize(job.done), "hipEventSynchronize");
    sync();
    const uint32_t* r = job.result();
    require(!(r[job.L.nonce] == 0xffffffffu && r[job.L.nonce + 1] == 0xffffffffu),
            "wide fibonacci prove: no proof-of-work nonce within the search bound");
    std::copy(r, r + job.L.total, proof);
  });
}

extern "C" int stwo_hip_wide_fib_prove_batch(const StwoWideFibParams* params, const uint64_t* seeds, uint32_t n_jobs,
                                             uint32_t* proofs, uint64_t proof_capacity_words) {
  return guarded([&] {
    prover::Params p = params_of(params);
    require(seeds != nullptr && n_jobs >= 1 && n_jobs <= prover::MAX_BATCH,
            "wide fibonacci prove batch: 1 .. 16 jobs with their seeds");
    const uint64_t words = prover::Layout(p).total;
    require(proofs != nullptr && proof_capacity_words >= words * n_jobs,
            "wide fibonacci prove batch: proof buffer too small");
    prover::Job job(p, ctx().stream, n_jobs);
    job.enqueue(seeds);
    hip_check(hipEventSynchronize(job.done), "hipEventSynchronize");
    sync();
    for (uint32_t k = 0; k < n_jobs; k++) {
      const uint32_t* r = job.result(k);
      require(!(r[job.L.nonce] == 0xffffffffu && r[job.L.nonce + 1] == 0xffffffffu),
              "wide fibonacci prove batch: no proof-of-work nonce within the search bound");
      std::copy(r, r + words, proofs + k * words);
    }
  });
}

extern "C" int stwo_hip_server_create(StwoProofServer** out, uint64_t hbm_budget_bytes, uint32_t max_inflight) {
  return guarded([&] {
    require(out != nullptr && max_inflight >= 1 && max_inflight <= 16, "server: max_inflight must be in [1, 16]");
    auto h = std::make_unique<StwoProofServer>();
    serve::Server& s = h->s;
    s.budget = hbm_budget_bytes;
    s.max_inflight = max_inflight;
    s.streams.resize(max_inflight);
    s.slot_busy.assign(max_inflight, false);
    for (uint32_t i = 0; i < max_inflight; i++) s.streams[i] = slot_stream(i);
    const int dev = ctx().device;
    s.worker = std::thread([&s, dev] {
      (void)hipSetDevice(dev);
      s.loop();
    });
    *out = h.release();
  });
}

extern "C" int stwo_hip_server_limit_work(StwoProofServer* h, uint64_t max_trace_elements) {
  if (!h) return STWO_HIP_ERR_ARG;
  h->s.work_limit.store(max_trace_elements);
  return STWO_HIP_OK;
}

extern "C" int stwo_hip_server_batch(StwoProofServer* h, uint32_t max_batch, uint64_t max_batch_work) {
  if (!h || max_batch < 1 || max_batch > prover::MAX_BATCH) return STWO_HIP_ERR_ARG;
  h->s.max_batch.store(max_batch);
  h->s.max_batch_work.store(max_batch_work);
  return STWO_HIP_OK;
}

extern "C" int stwo_hip_server_submit(StwoProofServer* h, const StwoWideFibParams* params, uint64_t seed,
                                      uint64_t* job_id) {
  try {
    require(h != nullptr, "server: null handle");
    prover::Params p = params_of(params);
    serve::Server& s = h->s;
    std::lock_guard<std::mutex> l(s.m);
    require(!s.stopping, "server: stopped");
    const uint64_t id = s.next_id++;
    s.queue.push_back(serve::Request{id, p, seed, s.now()});
    if (job_id) *job_id = id;
    s.cv_work.notify_one();
    return STWO_HIP_OK;
  } catch (const Failure& f) {
    return f.status;
  }
}

extern "C" int stwo_hip_server_wait(StwoProofServer* h, uint32_t timeout_us, uint64_t* job_id, int* job_status,
                                    uint32_t* proof, uint64_t proof_capacity_words, uint64_t* proof_words,
                                    double* times) {
  try {
    require(h != nullptr, "server: null handle");
    serve::Server& s = h->s;
    std::unique_lock<std::mutex> l(s.m);
    if (!s.cv_done.wait_for(l, std::chrono::microseconds(timeout_us), [&] { return !s.results.empty(); }))
      return STWO_HIP_NOT_READY;
    serve::Result r = std::move(s.results.front());
    s.results.pop_front();
    l.unlock();
    if (job_id) *job_id = r.id;
    if (job_status) *job_status = r.status;
    if (r.status != STWO_HIP_OK) set_error(r.status, "server job %llu: %s", (unsigned long long)r.id, r.error.c_str());
    if (proof_words) *proof_words = r.proof.size();
    if (proof && proof_capacity_words >= r.proof.size()) std::copy(r.proof.begin(), r.proof.end(), proof);
    if (times) {
      times[0] = r.t_submit;
      times[1] = r.t_start;
      times[2] = r.t_end;
      times[3] = r.host_s;
    }
    return STWO_HIP_OK;
  } catch (const Failure& f) {
    return f.status;
  }
}

extern "C" int stwo_hip_server_destroy(StwoProofServer* h) {
  if (!h) return STWO_HIP_OK;
  {
    std::lock_guard<std::mutex> l(h->s.m);
    h->s.stopping = true;
    h->s.cv_work.notify_all();
  }
  if (h->s.worker.joinable()) h->s.worker.join();
  // the slot streams belong to the library context and stay alive
  return guarded([&] { delete h; });
}
