// Batched commitment of columns of many log sizes: one plan object per column set
// (the Rust trait path builds the same commitment with one synchronous call per
// operation, per column batch of equal size: stwo_gpu_backend/src/poly.rs:57-168 and
// blake2s.rs:19-97, driven layer by layer by stwo's MerkleProver::commit).
//
// The plan uploads every pointer table once, then each run enqueues:
//   * the LDE of each size class (interpolate into coefficients, evaluate on the
//     blowup domain), the classes fanned out over the auxiliary streams so small
//     classes run side by side, joined back into the library stream;
//   * the whole Merkle tree over the LDE columns with mixed-degree injection.
// A run can be captured once as a hipGraph and replayed (no per-launch CPU cost).
// Nothing is allocated while enqueuing, and no memset is recorded.
//
// Autotuner: for each size class of log >= 14 the candidate pass schedules
// (contiguous passes of 12 or 13 layers, strided passes fused or not) are timed on
// the plan's own buffers and the fastest is kept, process-wide, for that
// (log size, blowup, column-count bucket); then graph replay and stream launches of the
// whole plan are timed and run(use_graph) uses the graph only where it is faster.
#include <map>
#include <memory>
#include <tuple>
#include <vector>

#include "fft.hpp"
#include "merkle_plan.hpp"
#include "runtime.hpp"

namespace sh {
namespace fft {

namespace {
std::map<std::tuple<uint32_t, uint32_t, uint32_t>, LdeVariant> g_variants;
uint32_t bucket(uint64_t n_cols) { return log2_exact(n_cols ? n_cols : 1); }
}  // namespace

LdeVariant lde_variant_for(uint32_t n, uint32_t log_blowup, uint64_t n_cols) {
  auto it = g_variants.find({n, log_blowup, bucket(n_cols)});
  if (it != g_variants.end()) return it->second;
  // untuned default: the fused middle pass where it runs on 2^13-element tiles (at most 8
  // strided layers, log sizes 14..20), measured faster than separate strided passes there
  LdeVariant v;
  v.fused = n >= 14 && n <= 20;
  return v;
}
void set_lde_variant(uint32_t n, uint32_t log_blowup, uint64_t n_cols, LdeVariant v) {
  g_variants[{n, log_blowup, bucket(n_cols)}] = v;
}

}  // namespace fft

struct CommitPlan {
  struct Class {
    uint32_t log, n;
    DevBuf tables;  // [src | coeffs | lde] device pointer tables
    fft::LdeVariant var;
  };
  std::vector<Class> classes;
  uint32_t log_blowup;
  const uint32_t *tw, *itw;
  uint64_t tree_len;
  std::unique_ptr<merkle::MerklePlan> merkle;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool graph_pays = true;  // autotuned: replaying the graph beats launching from streams
  std::vector<hipEvent_t> class_done;  // one per class

  void enqueue_class(const Class& c) const {
    const void* const* base = c.tables.as<const void*>();
    fft::lde_same_size((const uint32_t* const*)base, (uint32_t* const*)(base + c.n),
                       (uint32_t* const*)(base + 2 * c.n), c.n, c.log, log_blowup, tw, itw, tree_len, c.var);
  }

  // Classes run on the auxiliary streams, largest first, round robin; each records an
  // event when its LDE columns are written, and every Merkle launch waits only for the
  // classes whose columns it hashes, so the top of the tree is built while smaller
  // classes are still being transformed.  Merkle launches are submitted as soon as the
  // classes they read are submitted (not after every class), so the device sees them
  // early.  Every auxiliary stream's last event is waited
  // for by some Merkle launch, which joins the streams back (as graph capture needs).
  void enqueue() const {
    Context& x = ctx();
    const int n_aux = (int)std::min<size_t>(classes.size(), N_AUX_STREAMS);
    if (n_aux > 1) {
      struct Restore {
        Context& x;
        ~Restore() { x.cur = x.stream; }
      } restore{x};
      hip_check(hipEventRecord(x.fork, x.stream), "hipEventRecord");
      for (int i = 0; i < n_aux; i++) hip_check(hipStreamWaitEvent(x.aux[i], x.fork, 0), "hipStreamWaitEvent");
      std::vector<hipEvent_t> ready(merkle->max_log + 1, nullptr);
      size_t next_step = 0;
      for (size_t i = 0; i < classes.size(); i++) {
        x.cur = x.aux[i % n_aux];
        enqueue_class(classes[i]);
        hip_check(hipEventRecord(class_done[i], x.cur), "hipEventRecord");
        ready[classes[i].log + log_blowup] = class_done[i];
        // the Merkle steps above the next (smaller) class read only classes enqueued so far
        x.cur = x.stream;
        const uint32_t below = i + 1 < classes.size() ? classes[i + 1].log + log_blowup + 1 : 0;
        merkle->enqueue_from(&next_step, below, &ready);
      }
      x.cur = x.stream;
      merkle->enqueue_from(&next_step, 0, &ready);
    } else {
      for (const Class& c : classes) enqueue_class(c);
      merkle->enqueue();
    }
  }

  void run(bool use_graph) {
    Context& x = ctx();
    if (!use_graph || !graph_pays) {
      enqueue();
      return;
    }
    if (!exec) {
      const bool timing = x.kernel_timing;
      x.kernel_timing = false;  // no event bookkeeping inside a captured graph
      hip_check(hipStreamBeginCapture(x.stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
      try {
        enqueue();
      } catch (...) {
        hipGraph_t g;
        (void)hipStreamEndCapture(x.stream, &g);
        x.cur = x.stream;
        x.kernel_timing = timing;
        throw;
      }
      hip_check(hipStreamEndCapture(x.stream, &graph), "hipStreamEndCapture");
      x.kernel_timing = timing;
      hip_check(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "hipGraphInstantiate");
    }
    hip_check(hipGraphLaunch(exec, x.stream), "hipGraphLaunch");
  }

  // time one class's LDE under each candidate schedule; keep the fastest
  void autotune() {
    Context& x = ctx();
    for (Class& c : classes) {
      if (c.log < 14) continue;
      std::vector<fft::LdeVariant> cand = {{12, false}};
      if (c.log - 13 <= 9 && c.log + log_blowup - 13 <= 9) cand.push_back({13, false});
      if (c.log <= 21) cand.push_back({12, true});
      if (cand.size() == 1) continue;
      float best = 1e30f;
      fft::LdeVariant best_v = cand[0];
      for (const auto& v : cand) {
        c.var = v;
        enqueue_class(c);  // warm-up
        float ms_min = 1e30f;
        for (int rep = 0; rep < 3; rep++) {
          hip_check(hipEventRecord(x.timer_start, x.stream), "hipEventRecord");
          enqueue_class(c);
          hip_check(hipEventRecord(x.timer_stop, x.stream), "hipEventRecord");
          hip_check(hipEventSynchronize(x.timer_stop), "hipEventSynchronize");
          float ms = 0;
          hip_check(hipEventElapsedTime(&ms, x.timer_start, x.timer_stop), "hipEventElapsedTime");
          ms_min = std::min(ms_min, ms);
        }
        if (ms_min < best) {
          best = ms_min;
          best_v = v;
        }
      }
      c.var = best_v;
      fft::set_lde_variant(c.log, log_blowup, c.n, best_v);
    }
    // graph replay against stream launches of the whole plan (a plan of many size classes
    // can keep more of them in flight from the streams than its replayed graph does)
    auto time_runs = [&](bool g) {
      run(g);  // warm-up (captures the graph on first use)
      float ms_min = 1e30f;
      for (int rep = 0; rep < 3; rep++) {
        hip_check(hipEventRecord(x.timer_start, x.stream), "hipEventRecord");
        run(g);
        hip_check(hipEventRecord(x.timer_stop, x.stream), "hipEventRecord");
        hip_check(hipEventSynchronize(x.timer_stop), "hipEventSynchronize");
        float ms = 0;
        hip_check(hipEventElapsedTime(&ms, x.timer_start, x.timer_stop), "hipEventElapsedTime");
        ms_min = std::min(ms_min, ms);
      }
      return ms_min;
    };
    graph_pays = true;
    const float ms_graph = time_runs(true);
    const float ms_stream = time_runs(false);
    graph_pays = ms_graph <= ms_stream;
  }

  ~CommitPlan() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (hipEvent_t e : class_done) (void)hipEventDestroy(e);
  }
};

}  // namespace sh

using namespace sh;

struct StwoCommitPlan {
  CommitPlan plan;
};

extern "C" int stwo_hip_commit_plan_create(StwoCommitPlan** out, uint32_t n_cols, const uint32_t* const* evals,
                                           uint32_t* const* coeffs, uint32_t* const* ldes, const uint32_t* log_sizes,
                                           uint32_t log_blowup, const uint32_t* twiddles, const uint32_t* itwiddles,
                                           uint64_t tree_len, StwoBlake2sHash* const* layers, uint32_t max_log,
                                           int autotune) {
  return guarded([&] {
    require(out != nullptr && n_cols > 0, "commit plan: no columns");
    auto holder = std::make_unique<StwoCommitPlan>();
    CommitPlan& p = holder->plan;
    p.log_blowup = log_blowup;
    p.tw = twiddles;
    p.itw = itwiddles;
    p.tree_len = tree_len;
    std::map<uint32_t, std::vector<uint32_t>, std::greater<uint32_t>> by;
    uint32_t top = 0;
    for (uint32_t i = 0; i < n_cols; i++) {
      require(log_sizes[i] >= 1 && log_sizes[i] + log_blowup <= 30, "commit plan: log size out of range");
      require(tree_len >= (1ull << (log_sizes[i] + log_blowup - 1)), "commit plan: twiddle tree too small");
      by[log_sizes[i]].push_back(i);
      top = std::max(top, log_sizes[i] + log_blowup);
    }
    require(max_log == top, "commit plan: max_log must be the largest LDE log size");
    for (auto& kv : by) {
      const uint32_t n = (uint32_t)kv.second.size();
      std::vector<const void*> t(3 * n);
      for (uint32_t i = 0; i < n; i++) {
        t[i] = evals[kv.second[i]];
        t[n + i] = coeffs[kv.second[i]];
        t[2 * n + i] = ldes[kv.second[i]];
      }
      CommitPlan::Class c{kv.first, n, upload(t.data(), t.size() * sizeof(void*)),
                          fft::lde_variant_for(kv.first, log_blowup, n)};
      p.classes.push_back(std::move(c));
    }
    p.class_done.resize(p.classes.size());
    for (auto& e : p.class_done) hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
    std::vector<uint32_t> lde_logs(n_cols);
    for (uint32_t i = 0; i < n_cols; i++) lde_logs[i] = log_sizes[i] + log_blowup;
    p.merkle = std::make_unique<merkle::MerklePlan>(n_cols, (const uint32_t* const*)ldes, lde_logs.data(),
                                                    (merkle::Hash* const*)layers, max_log);
    if (autotune) {
      // the autotuner runs the plan several times on the caller's buffers; a column whose
      // coefficients alias its evaluations is interpolated in place by every run, so its
      // evaluations are kept aside and put back before create returns
      struct Saved {
        void* col;
        size_t bytes;
        DevBuf copy;
      };
      std::vector<Saved> saved;
      for (uint32_t i = 0; i < n_cols; i++) {
        if ((const void*)coeffs[i] != (const void*)evals[i]) continue;
        const size_t bytes = sizeof(uint32_t) << log_sizes[i];
        Saved sv{coeffs[i], bytes, DevBuf(bytes)};
        hip_check(hipMemcpyAsync(sv.copy.p, sv.col, bytes, hipMemcpyDeviceToDevice, ctx().cur), "hipMemcpyAsync D2D");
        saved.push_back(std::move(sv));
      }
      p.autotune();
      for (Saved& sv : saved)
        hip_check(hipMemcpyAsync(sv.col, sv.copy.p, sv.bytes, hipMemcpyDeviceToDevice, ctx().cur),
                  "hipMemcpyAsync D2D");
    }
    sync();
    *out = holder.release();
  });
}

extern "C" int stwo_hip_commit_plan_run(StwoCommitPlan* plan, int use_graph) {
  return guarded([&] {
    require(plan != nullptr, "commit plan: null plan");
    plan->plan.run(use_graph != 0);
  });
}

extern "C" int stwo_hip_lde_schedule(uint32_t log_size, uint32_t log_blowup, uint64_t n_cols, uint32_t* contiguous_layers,
                                     int* fused) {
  return guarded([&] {
    const fft::LdeVariant v = fft::lde_variant_for(log_size, log_blowup, n_cols);
    if (contiguous_layers) *contiguous_layers = v.kc;
    if (fused) *fused = v.fused ? 1 : 0;
  });
}

extern "C" int stwo_hip_commit_plan_uses_graph(const StwoCommitPlan* plan, int* uses_graph) {
  return guarded([&] {
    require(plan != nullptr && uses_graph != nullptr, "commit plan: null argument");
    *uses_graph = plan->plan.graph_pays ? 1 : 0;
  });
}

extern "C" int stwo_hip_commit_plan_destroy(StwoCommitPlan* plan) {
  return guarded([&] {
    if (!plan) return;
    sync();
    delete plan;
  });
}
