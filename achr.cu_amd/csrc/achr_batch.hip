// Batches: several samplers (one per model, any mix of reaction counts) run
// as one job.
//
// Models whose chains use the same register schedule (threads per chain,
// slots per thread, direction-bank or centre form) form a bucket; one launch
// per bucket runs a segment of every chain of every model in it
// (step_batch_kernel over an item list of {model, first chain}: one workgroup
// per item for chains of 128-1,024 lanes, a persistent grid of resident
// workgroups walking the list for groups of 16-64-lane chains). A segment of the whole batch is every bucket's launch
// followed by every model's projection round; it is captured once in a HIP
// graph and replayed for each segment, so a run of S steps costs
// ceil(S / check_every) + 1 graph launches whatever the number of models.
//
// The autotuner times the candidate threads-per-chain of every size class
// (models with n in [2^k, 2^(k+1))) on a short probe of the real chain
// counts and keeps the fastest. Results depend on the threads per chain (the
// per-lane partial sums of the norm), never on the slot count, the grid or
// the bucket a model runs in, so a tuned batch reproduces bit for bit once its
// lanes are pinned (achr_batch_lanes / achr_batch_set_lanes).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <string>
#include <vector>

#include "../../include/achr.h"
#include "achr_host.h"
#include "achr_internal.h"

using achr_host::DevBuf;
using achr_host::fail;

namespace {

constexpr int kTpcs[] = {16, 32, 64, 128, 256, 512, 1024};
constexpr int kNumTpc = 7;

struct Bucket {
  achr::Schedule sched{};
  bool dir = false;
  std::vector<int> models;  // indices into the batch, largest n first
  std::vector<int2> items;
  DevBuf<int2> items_d;
  int resident = 0;
  int grid = 0;
};

}  // namespace

struct achr_batch {
  std::vector<achr_sampler*> s;
  std::vector<achr::Schedule> sched;  // per model
  hipStream_t stream = nullptr;
  bool own_stream = false;
  std::vector<achr::StepArgs> tab;
  DevBuf<achr::StepArgs> tab_d;
  std::vector<Bucket> buckets;
  hipEvent_t ev_begin = nullptr, ev_end = nullptr;
  // branches of a segment: bucket b and its models' projection rounds run
  // on aux[b % aux.size()], forked from and joined back into `stream`
  std::vector<hipStream_t> aux;
  std::vector<hipEvent_t> ev_join;
  hipEvent_t ev_fork = nullptr;
  int segments = 0;
  int graphs = 0;
  size_t graph_nodes = 0;
  bool have_run = false;
};

namespace {

achr::Schedule schedule_for(int n, int tpc) {
  achr::Schedule s{};
  s.tpc = tpc;
  s.e = achr::elems_for(n, tpc);
  s.cpb = tpc <= 64 ? 256 / tpc : 1;
  return s;
}

// Buckets of the models listed in `which` under the batch's schedules.
int build_buckets(achr_batch* B, const std::vector<int>& which, const std::vector<int>& P,
                  std::vector<Bucket>& out) {
  out.clear();
  std::vector<int> order = which;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return B->s[a]->n > B->s[b]->n; });
  for (int k : order) {
    const achr::Schedule& sc = B->sched[k];
    const bool dir = B->s[k]->use_dirs != 0;
    Bucket* bk = nullptr;
    for (auto& b : out)
      if (b.sched.tpc == sc.tpc && b.sched.e == sc.e && b.dir == dir) bk = &b;
    if (!bk) {
      out.emplace_back();
      bk = &out.back();
      bk->sched = sc;
      bk->dir = dir;
    }
    bk->models.push_back(k);
    for (int c = 0; c < P[k]; c += sc.cpb) bk->items.push_back(make_int2(k, c));
  }
  for (auto& b : out) {
    hipError_t e = achr::prepare_step_batch(b.sched, b.dir, &b.resident);
    if (e != hipSuccess) return fail(ACHR_E_DEVICE, "step kernel (%d lanes x %d): %s", b.sched.tpc, b.sched.e,
                                     hipGetErrorString(e));
    const int nit = static_cast<int>(b.items.size());
    b.grid = b.sched.cpb == 1 ? nit : std::min(nit, b.resident);  // one workgroup per item, or persistent
    e = b.items_d.upload(b.items.data(), b.items.size(), B->stream);
    if (e != hipSuccess) return fail(ACHR_E_DEVICE, "item upload: %s", hipGetErrorString(e));
  }
  return ACHR_OK;
}

int launch_buckets(achr_batch* B, const std::vector<Bucket>& bks) {
  for (const auto& b : bks) {
    const hipError_t e = achr::launch_step_batch(B->tab_d.p, b.items_d.p, static_cast<int>(b.items.size()), b.grid,
                                                 b.sched, b.dir, B->stream);
    if (e != hipSuccess) return fail(ACHR_E_DEVICE, "batched step launch: %s", hipGetErrorString(e));
  }
  return ACHR_OK;
}

// Resets the chains of the listed models and uploads their kernel arguments.
int prepare_models(achr_batch* B, const std::vector<int>& which, const std::vector<int>& P, int Ssteps,
                   uint64_t seed, uint32_t file) {
  for (int k : which) {
    int rc = achr_host::run_limits(B->s[k], B->sched[k], P[k], Ssteps);
    if (rc == ACHR_OK) rc = achr_host::prepare_run(B->s[k], P[k], Ssteps, seed, file, 0, B->stream, &B->tab[k]);
    if (rc != ACHR_OK) return rc;
  }
  // The previous run has finished (callers synchronise), so the host table
  // may be reused; wait for this copy before it can change again.
  const hipError_t e = B->tab_d.upload(B->tab.data(), B->tab.size(), B->stream);
  if (e != hipSuccess) return fail(ACHR_E_DEVICE, "argument table upload: %s", hipGetErrorString(e));
  return ACHR_OK;
}

}  // namespace

extern "C" {

int achr_batch_create(achr_sampler* const* samplers, int count, achr_batch** out) {
  if (!out) return fail(ACHR_E_ARG, "out is NULL");
  *out = nullptr;
  if (!samplers || count <= 0) return fail(ACHR_E_ARG, "no samplers");
  for (int k = 0; k < count; ++k) {
    if (!samplers[k]) return fail(ACHR_E_ARG, "sampler %d is NULL", k);
    for (int j = 0; j < k; ++j)
      if (samplers[j] == samplers[k]) return fail(ACHR_E_ARG, "sampler %d appears twice", k);
  }
  auto* B = new (std::nothrow) achr_batch();
  if (!B) return fail(ACHR_E_NOMEM, "batch allocation");
  B->s.assign(samplers, samplers + count);
  for (auto* S : B->s) B->sched.push_back(S->sched);
  B->tab.resize(count);
  hipError_t e = hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking);
  if (e == hipSuccess) B->own_stream = true;
  if (e == hipSuccess) e = hipEventCreate(&B->ev_begin);
  if (e == hipSuccess) e = hipEventCreate(&B->ev_end);
  if (e != hipSuccess) {
    achr_batch_destroy(B);
    return fail(ACHR_E_DEVICE, "batch stream: %s", hipGetErrorString(e));
  }
  *out = B;
  return ACHR_OK;
}

void achr_batch_destroy(achr_batch* B) {
  if (!B) return;
  if (B->stream) (void)hipStreamSynchronize(B->stream);
  for (hipStream_t a : B->aux) (void)hipStreamSynchronize(a);
  for (hipStream_t a : B->aux) (void)hipStreamDestroy(a);
  for (hipEvent_t e : B->ev_join) (void)hipEventDestroy(e);
  if (B->ev_fork) (void)hipEventDestroy(B->ev_fork);
  if (B->ev_begin) (void)hipEventDestroy(B->ev_begin);
  if (B->ev_end) (void)hipEventDestroy(B->ev_end);
  if (B->own_stream && B->stream) (void)hipStreamDestroy(B->stream);
  delete B;
}

int achr_batch_set_stream(achr_batch* B, void* stream) {
  if (!B) return fail(ACHR_E_ARG, "batch is NULL");
  if (!stream) return fail(ACHR_E_ARG, "a batch captures its work in a graph and needs a non-null stream");
  HIP_TRY(hipStreamSynchronize(B->stream));
  if (B->own_stream) (void)hipStreamDestroy(B->stream);
  B->own_stream = false;
  B->stream = static_cast<hipStream_t>(stream);
  return ACHR_OK;
}

int achr_batch_lanes(const achr_batch* B, int* lanes) {
  if (!B || !lanes) return fail(ACHR_E_ARG, "NULL argument");
  for (size_t k = 0; k < B->s.size(); ++k) lanes[k] = B->sched[k].tpc;
  return ACHR_OK;
}

int achr_batch_set_lanes(achr_batch* B, const int* lanes) {
  if (!B || !lanes) return fail(ACHR_E_ARG, "NULL argument");
  std::vector<achr::Schedule> next = B->sched;
  for (size_t k = 0; k < B->s.size(); ++k) {
    const achr::Schedule sc = lanes[k] == 0 ? B->s[k]->sched : schedule_for(B->s[k]->n, lanes[k]);
    if (!achr::schedule_supported(sc.tpc, sc.e))
      return fail(ACHR_E_ARG, "model %zu: lanes=%d cannot hold n=%d", k, lanes[k], B->s[k]->n);
    next[k] = sc;
  }
  B->sched = next;
  return ACHR_OK;
}

int achr_batch_autotune(achr_batch* B, const int* P, int probe_steps, uint64_t seed, double* ms) {
  if (!B || !P) return fail(ACHR_E_ARG, "NULL argument");
  if (probe_steps <= 0) return fail(ACHR_E_ARG, "probe_steps must be > 0");
  const int count = static_cast<int>(B->s.size());
  std::vector<int> Pv(P, P + count);
  for (int k = 0; k < count; ++k)
    if (Pv[k] <= 0) return fail(ACHR_E_ARG, "model %d: P=%d", k, Pv[k]);
  if (ms)
    for (int i = 0; i < count * kNumTpc; ++i) ms[i] = -1.0;
  HIP_TRY(hipStreamSynchronize(B->stream));
  // size classes: floor(log2 n)
  std::map<int, std::vector<int>> classes;
  for (int k = 0; k < count; ++k) classes[static_cast<int>(std::floor(std::log2(std::max(B->s[k]->n, 1))))].push_back(k);
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  int rc = ACHR_OK;
  const std::vector<achr::Schedule> saved = B->sched;
  for (auto& cl : classes) {
    const std::vector<int>& which = cl.second;
    double best = INFINITY;
    int best_tpc = 0;
    for (int ti = 0; ti < kNumTpc && rc == ACHR_OK; ++ti) {
      const int tpc = kTpcs[ti];
      bool ok = true;
      for (int k : which) {
        const achr::Schedule sc = schedule_for(B->s[k]->n, tpc);
        if (!achr::schedule_supported(sc.tpc, sc.e)) ok = false;
        // a candidate must not leave lanes idle for a whole slot row
        if (ok && static_cast<long>(sc.tpc) * sc.e >= 2L * std::max(B->s[k]->n, 64) + sc.tpc) ok = false;
        if (ok) {
          B->sched[k] = sc;
          if (achr_host::run_limits(B->s[k], sc, Pv[k], probe_steps) != ACHR_OK) ok = false;
        }
      }
      if (!ok) continue;
      int segs = 1;
      for (int k : which) segs = std::max(segs, achr_host::segments_for(B->s[k], probe_steps));
      float t = 0.f;
      for (int rep = 0; rep < 2 && rc == ACHR_OK; ++rep) {  // the first pass loads the code object
        rc = prepare_models(B, which, Pv, probe_steps, seed, 0);
        std::vector<Bucket> bks;
        if (rc == ACHR_OK) rc = build_buckets(B, which, Pv, bks);
        if (rc != ACHR_OK) break;
        HIP_TRY(hipEventRecord(e0, B->stream));
        for (int seg = 0; seg < segs && rc == ACHR_OK; ++seg) rc = launch_buckets(B, bks);
        HIP_TRY(hipEventRecord(e1, B->stream));
        HIP_TRY(hipEventSynchronize(e1));
        HIP_TRY(hipEventElapsedTime(&t, e0, e1));
      }
      if (rc != ACHR_OK) break;
      if (ms)
        for (int k : which) ms[k * kNumTpc + ti] = t;
      if (t < best) {
        best = t;
        best_tpc = tpc;
      }
    }
    if (rc != ACHR_OK) break;
    for (int k : which) B->sched[k] = best_tpc ? schedule_for(B->s[k]->n, best_tpc) : saved[k];
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc != ACHR_OK) B->sched = saved;
  for (auto* S : B->s) S->have_run = false;  // probe results are not samples
  return rc;
}

int achr_batch_sample(achr_batch* B, const int* P, int Ssteps, uint64_t seed, uint32_t file, int use_graph) {
  if (!B || !P) return fail(ACHR_E_ARG, "NULL argument");
  const int count = static_cast<int>(B->s.size());
  std::vector<int> Pv(P, P + count);
  std::vector<int> all(count);
  for (int k = 0; k < count; ++k) {
    if (Pv[k] <= 0) return fail(ACHR_E_ARG, "model %d: P=%d", k, Pv[k]);
    all[k] = k;
  }
  HIP_TRY(hipStreamSynchronize(B->stream));
  B->have_run = false;
  int rc = prepare_models(B, all, Pv, Ssteps, seed, file);
  if (rc == ACHR_OK) rc = build_buckets(B, all, Pv, B->buckets);
  if (rc != ACHR_OK) return rc;
  std::vector<int> segs(count);
  B->segments = 1;
  for (int k = 0; k < count; ++k) {
    segs[k] = achr_host::segments_for(B->s[k], Ssteps);
    B->segments = std::max(B->segments, segs[k]);
  }
  // Projection rounds after segment seg; segments with the same rounds share
  // one graph.
  auto signature = [&](int seg) {
    std::string sig(count, '\0');
    for (int k = 0; k < count; ++k)
      sig[k] = static_cast<char>(seg + 1 < segs[k] ? achr_host::projection_kind(B->s[k], seg) : achr_host::kProjNone);
    return sig;
  };
  // Branch streams (at most the four hardware queues a process gets),
  // created once per batch.
  const int nbr = std::min<int>(4, static_cast<int>(B->buckets.size()));
  while (static_cast<int>(B->aux.size()) < nbr) {
    hipStream_t a = nullptr;
    hipEvent_t e = nullptr;
    hipError_t he = hipStreamCreateWithFlags(&a, hipStreamNonBlocking);
    if (he == hipSuccess) he = hipEventCreateWithFlags(&e, hipEventDisableTiming);
    if (he != hipSuccess) return fail(ACHR_E_DEVICE, "batch branch stream: %s", hipGetErrorString(he));
    B->aux.push_back(a);
    B->ev_join.push_back(e);
  }
  if (!B->ev_fork) HIP_TRY(hipEventCreateWithFlags(&B->ev_fork, hipEventDisableTiming));
  // A segment: every bucket's launch and then its models' projection
  // rounds, one branch per bucket (a model's round depends only on its own
  // chains), the branches forked from and joined back into the batch stream.
  auto enqueue_segment = [&](const std::string& sig) {
    hipError_t he = hipEventRecord(B->ev_fork, B->stream);
    for (int a = 0; a < nbr && he == hipSuccess; ++a) he = hipStreamWaitEvent(B->aux[a], B->ev_fork, 0);
    if (he != hipSuccess) return fail(ACHR_E_DEVICE, "batch fork: %s", hipGetErrorString(he));
    int r = ACHR_OK;
    for (size_t b = 0; b < B->buckets.size() && r == ACHR_OK; ++b) {
      const Bucket& bk = B->buckets[b];
      hipStream_t st = B->aux[b % nbr];
      he = achr::launch_step_batch(B->tab_d.p, bk.items_d.p, static_cast<int>(bk.items.size()), bk.grid, bk.sched,
                                   bk.dir, st);
      if (he != hipSuccess) r = fail(ACHR_E_DEVICE, "batched step launch: %s", hipGetErrorString(he));
      for (size_t j = 0; j < bk.models.size() && r == ACHR_OK; ++j)
        r = achr_host::enqueue_projection(B->s[bk.models[j]], sig[bk.models[j]], st);
    }
    for (int a = 0; a < nbr; ++a) {
      he = hipEventRecord(B->ev_join[a], B->aux[a]);
      if (he == hipSuccess) he = hipStreamWaitEvent(B->stream, B->ev_join[a], 0);
      if (he != hipSuccess && r == ACHR_OK) r = fail(ACHR_E_DEVICE, "batch join: %s", hipGetErrorString(he));
    }
    return r;
  };
  std::map<std::string, std::pair<hipGraph_t, hipGraphExec_t>> graphs;
  auto release = [&]() {
    for (auto& g : graphs) {
      if (g.second.second) (void)hipGraphExecDestroy(g.second.second);
      if (g.second.first) (void)hipGraphDestroy(g.second.first);
    }
  };
  HIP_TRY(hipEventRecord(B->ev_begin, B->stream));
  for (auto* S : B->s) HIP_TRY(hipEventRecord(S->ev_begin, B->stream));
  B->graph_nodes = 0;
  for (int seg = 0; seg < B->segments && rc == ACHR_OK; ++seg) {
    const std::string sig = signature(seg);
    for (int k = 0; k < count; ++k) {
      if (sig[k] == achr_host::kProjFP64) ++B->s[k]->rounds_fp64;
      if (sig[k] == achr_host::kProjLow) ++B->s[k]->rounds_low;
    }
    if (!use_graph) {
      rc = enqueue_segment(sig);
      continue;
    }
    auto it = graphs.find(sig);
    if (it == graphs.end()) {
      hipGraph_t g = nullptr;
      hipGraphExec_t ge = nullptr;
      hipError_t e = hipStreamBeginCapture(B->stream, hipStreamCaptureModeThreadLocal);
      if (e != hipSuccess) {
        release();
        return fail(ACHR_E_DEVICE, "graph capture: %s", hipGetErrorString(e));
      }
      rc = enqueue_segment(sig);
      e = hipStreamEndCapture(B->stream, &g);
      if (rc == ACHR_OK && e == hipSuccess) e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
      size_t nodes = 0;
      if (rc == ACHR_OK && e == hipSuccess) e = hipGraphGetNodes(g, nullptr, &nodes);
      it = graphs.emplace(sig, std::make_pair(g, ge)).first;
      if (rc != ACHR_OK) break;
      if (e != hipSuccess) {
        release();
        return fail(ACHR_E_DEVICE, "graph build: %s", hipGetErrorString(e));
      }
      B->graph_nodes += nodes;
    }
    const hipError_t e = hipGraphLaunch(it->second.second, B->stream);
    if (e != hipSuccess) {
      release();
      return fail(ACHR_E_DEVICE, "graph launch: %s", hipGetErrorString(e));
    }
  }
  B->graphs = static_cast<int>(graphs.size());
  if (rc == ACHR_OK) {
    for (int k = 0; k < count; ++k) {
      achr_sampler* S = B->s[k];
      hipError_t e = hipEventRecord(S->ev_end, B->stream);
      if (e != hipSuccess) rc = fail(ACHR_E_DEVICE, "event: %s", hipGetErrorString(e));
      achr_host::note_run(S, B->sched[k]);
      S->have_run = true;
    }
  }
  if (rc == ACHR_OK) {
    const hipError_t e = hipEventRecord(B->ev_end, B->stream);
    if (e != hipSuccess) rc = fail(ACHR_E_DEVICE, "event: %s", hipGetErrorString(e));
  }
  // Executable graphs may be destroyed once their launches are queued only
  // after they complete: wait here (a batch run is one job).
  const hipError_t se = hipStreamSynchronize(B->stream);
  release();
  if (rc == ACHR_OK && se != hipSuccess) rc = fail(ACHR_E_DEVICE, "batch run: %s", hipGetErrorString(se));
  B->have_run = rc == ACHR_OK;
  return rc;
}

int achr_batch_timing(const achr_batch* B, double* ms_total, int* segments, int* buckets, int* graphs,
                      int* graph_nodes) {
  if (!B) return fail(ACHR_E_ARG, "batch is NULL");
  if (!B->have_run) return fail(ACHR_E_STATE, "no run");
  float ms = 0.f;
  HIP_TRY(hipEventSynchronize(B->ev_end));
  HIP_TRY(hipEventElapsedTime(&ms, B->ev_begin, B->ev_end));
  if (ms_total) *ms_total = ms;
  if (segments) *segments = B->segments;
  if (buckets) *buckets = static_cast<int>(B->buckets.size());
  if (graphs) *graphs = B->graphs;
  if (graph_nodes) *graph_nodes = static_cast<int>(B->graph_nodes);
  return ACHR_OK;
}

}  // extern "C"
