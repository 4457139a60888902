// Serving: circuit-sampling requests arrive at any time; a worker thread
// admits them (first-come first-served, or smallest estimated work first with
// an age limit) as soon as a slot (stream) is free and the request's device
// bytes fit the HBM budget, runs each as a persistent
// program on its slot stream (so requests overlap on the GPU), detects
// completion with events, copies the result planes to the host and releases
// the device memory stream-ordered (no device-wide synchronisation).
//
// A request's results depend only on (n, B, seed, ops): coins and noise are
// keyed by its own seed, sequence numbers and shot indices, never by launch
// position, so they are identical alone and in any mix.
#include <hip/hip_runtime.h>

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <set>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "plaquette_amd.h"
#include "pq_capi_internal.h"
#include "pq_common.h"
#include "pq_kernels.h"
#include "pq_program_internal.h"

namespace {

constexpr size_t kMaxCachedPlans = 64;  // frame plans kept between requests

double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

struct Req {
  uint64_t ticket = 0;
  uint64_t n = 0, B = 0;
  int64_t seed = 0;
  std::vector<pq_op> ops;
  uint64_t nrec = 0, W = 0, bytes = 0;
  double cost = 0;  // estimated work: word groups x rows x ops
  // frame form on a state that keeps only its R column: the server returns
  // records, never the tableau, so X/Z need not exist per shot
  bool lite = false;
  std::string plan_key;    // lite requests: frame_plan_key of the ops
  bool own_build = false;  // this request's submit queued its plan's build
  double t_submit = 0, t_start = 0, t_done = 0;
  int status = 0;  // 0 queued, 1 running, 2 done, < 0 failed (PQ_ERR_*)
  int slot = -1;
  pq_state st = nullptr;
  pq_program prog = nullptr;
  hipEvent_t ev = nullptr;
  std::vector<uint64_t> planes;  // [nrec][2][W] once done
};

}  // namespace

struct pq_server_s {
  uint64_t budget = 0, used = 0, peak = 0;
  std::vector<hipStream_t> streams;
  std::vector<char> busy;
  int device = 0;
  std::mutex mu;
  std::condition_variable cv;
  std::deque<Req *> pending;
  std::vector<Req *> running;
  std::unordered_map<uint64_t, std::unique_ptr<Req>> reqs;
  uint64_t next_ticket = 1, completed = 0;
  uint32_t max_running = 0;
  int policy = PQ_SERVE_FCFS;
  double age_limit = 0;
  int form = PQ_RUN_PERSISTENT;
  // frame plans by structure (frame_plan_key), built once and shared;
  // plans_mu guards them and the counters (a build can take a while, so it
  // is not `mu`, which submit and status take)
  std::mutex plans_mu;
  std::unordered_map<std::string, std::shared_ptr<FramePlan>> plans;
  uint64_t plan_hits = 0, plan_builds = 0, frame_runs = 0;
  // plans are built ahead on the builder thread, from submit on, so a
  // request's plan is usually ready when a slot frees up; `building` holds
  // the keys queued or in progress (plans_cv signals each finish)
  struct BuildJob {
    std::string key;
    uint64_t n = 0;
    std::vector<pq_op> ops;
  };
  std::deque<BuildJob> build_q;
  std::set<std::string> building;
  std::condition_variable plans_cv;
  bool builder_stop = false;
  std::thread builder;
  std::atomic<bool> stop{false};
  std::thread worker;
  void loop();
  void build_loop();
  Req *next_fitting(double now);
  int admit(Req *r);
  void finish(Req *r);
};

// Whether X/Z stay the same in every shot (no flagged H/S/CX/CZ), so the
// frame form can run the request.
static bool frame_eligible(const pq_op *ops, uint64_t nops) {
  for (uint64_t i = 0; i < nops; ++i) {
    const uint32_t c = ops[i].code;
    if (c == PQ_OP_H_FLAG || c == PQ_OP_S_FLAG || c == PQ_OP_CX_FLAG || c == PQ_OP_CZ_FLAG)
      return false;
  }
  return true;
}

// Device bytes a request reserves: tableau (its R column alone for a lite
// request), records, noise planes, program.
static uint64_t request_bytes(uint64_t n, uint64_t B, const pq_op *ops, uint64_t nops,
                              uint64_t *nrec_out, bool lite) {
  const uint64_t W = (B + 63) / 64, C = 2 * n + 1;
  uint64_t nrec = 0, nnoise = 0;
  for (uint64_t i = 0; i < nops; ++i) {
    if (ops[i].code == PQ_OP_M && ops[i].slot + 1 > nrec) nrec = ops[i].slot + 1;
    if (ops[i].code == PQ_OP_NOISE || (ops[i].code != PQ_OP_M && ops[i].code >= PQ_OP_H_FLAG &&
                                       ops[i].slot == PQ_FLAG_RANDOM))
      ++nnoise;
  }
  *nrec_out = nrec;
  return W * (lite ? 1 : C) * C * 8 + nrec * 2 * W * 8 +
         nnoise * (2 * W * 8 + sizeof(pq::NoiseDesc)) +
         nops * sizeof(pq::DInstr) + 3 * W * 8 + 4096;
}

// The queued request to start next under the policy, removed from `pending`,
// or nullptr if the policy's choice does not fit the free budget.
Req *pq_server_s::next_fitting(double now) {
  if (pending.empty()) return nullptr;
  auto pick = pending.begin();
  if (policy == PQ_SERVE_SMALLEST_FIRST) {
    const bool aged = age_limit > 0 && now - pending.front()->t_submit > age_limit;
    if (!aged) {
      pick = pending.end();
      for (auto it = pending.begin(); it != pending.end(); ++it)
        if (used + (*it)->bytes <= budget && (pick == pending.end() || (*it)->cost < (*pick)->cost))
          pick = it;
      if (pick == pending.end()) return nullptr;
    }
  }
  if (used + (*pick)->bytes > budget) return nullptr;
  Req *r = *pick;
  pending.erase(pick);
  return r;
}

int pq_server_s::admit(Req *r) {
  const hipStream_t s = streams[r->slot];
  int rc = create_state_on(r->n, r->B, r->seed, s, &r->st, r->lite);
  if (rc == PQ_OK) rc = create_program_on(r->st, r->ops.data(), r->ops.size(), &r->prog);
  int run_form = PQ_RUN_PERSISTENT;
  if (rc == PQ_OK && (r->lite || form == PQ_RUN_FRAME)) {
    // the frame form where the request allows it (no flagged H/S/CX/CZ),
    // with plans shared between requests of the same structure
    const std::string key =
        r->plan_key.empty() ? frame_plan_key(r->n, r->ops.data(), r->ops.size()) : r->plan_key;
    std::unique_lock<std::mutex> g(plans_mu);
    // a build of this structure under way on the builder thread: wait for it
    plans_cv.wait(g, [&] { return building.count(key) == 0; });
    auto it = plans.find(key);
    if (it != plans.end()) {
      r->prog->frame = it->second;
      if (!r->own_build) ++plan_hits;  // the request that queued the build is its first use
      run_form = PQ_RUN_FRAME;
    } else if ((rc = build_frame_plan(r->prog, UINT64_MAX)) == PQ_OK) {
      // bounded cache: plans in use stay alive through their programs
      if (plans.size() >= kMaxCachedPlans) plans.clear();
      plans.emplace(key, r->prog->frame);
      ++plan_builds;
      run_form = PQ_RUN_FRAME;
    } else if (!r->lite) {
      rc = PQ_OK;  // a full state runs persistent instead
    }
    if (run_form == PQ_RUN_FRAME) ++frame_runs;
  }
  if (rc == PQ_OK) rc = pq_program_run(r->st, r->prog, run_form);
  if (rc == PQ_OK && hipEventRecord(r->ev, s) != hipSuccess) rc = PQ_ERR_HIP;
  return rc;
}

void pq_server_s::finish(Req *r) {
  const hipStream_t s = streams[r->slot];
  if (r->status == 1 && r->st && r->st->recs && r->nrec) {
    r->planes.resize(r->nrec * 2 * r->W);
    if (hipMemcpyAsync(r->planes.data(), r->st->recs, r->planes.size() * 8,
                       hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      r->status = PQ_ERR_HIP;
  }
  if (r->prog) pq_program_destroy(r->prog);
  if (r->st) pq_destroy(r->st);
  r->prog = nullptr;
  r->st = nullptr;
}

// Builds queued plans ahead of their requests' admission (host only: the
// plan is compiled from the ops, as pq_frame_plan does).  A build that fails
// leaves no plan: admission then builds again and reports or falls back.
void pq_server_s::build_loop() {
  std::unique_lock<std::mutex> g(plans_mu);
  while (true) {
    plans_cv.wait(g, [&] { return builder_stop || !build_q.empty(); });
    if (builder_stop) return;
    BuildJob job = std::move(build_q.front());
    build_q.pop_front();
    g.unlock();
    pq_program_s prog;
    prog.n = job.n;
    prog.R = 2 * job.n + 1;
    prog.ops = std::move(job.ops);
    const bool ok =
        compile_frame_only(&prog) == PQ_OK && build_frame_plan(&prog, UINT64_MAX) == PQ_OK;
    g.lock();
    if (ok) {
      if (plans.size() >= kMaxCachedPlans) plans.clear();
      plans.emplace(job.key, prog.frame);
      ++plan_builds;
    }
    building.erase(job.key);
    plans_cv.notify_all();
  }
}

void pq_server_s::loop() {
  (void)hipSetDevice(device);
  while (!stop.load()) {
    std::vector<Req *> done, start;
    {
      std::unique_lock<std::mutex> lk(mu);
      for (size_t i = 0; i < running.size();) {
        hipError_t q = hipEventQuery(running[i]->ev);
        if (q == hipErrorNotReady) {
          ++i;
          continue;
        }
        Req *r = running[i];
        if (q != hipSuccess) r->status = PQ_ERR_HIP;
        done.push_back(r);
        running[i] = running.back();
        running.pop_back();
      }
      const double t = now_s();
      for (size_t sl = 0; sl < busy.size() && !pending.empty(); ++sl) {
        if (busy[sl]) continue;
        Req *r = next_fitting(t);
        if (!r) break;
        busy[sl] = 1;
        r->slot = (int)sl;
        used += r->bytes;
        peak = std::max(peak, used);
        start.push_back(r);
      }
      if (done.empty() && start.empty()) {
        cv.wait_for(lk, std::chrono::microseconds(200));
        continue;
      }
    }
    for (Req *r : start) {
      r->t_start = now_s();
      r->status = 1;
      const int rc = admit(r);
      if (rc != PQ_OK) {
        r->status = rc;
        done.push_back(r);
      } else {
        std::lock_guard<std::mutex> g(mu);
        running.push_back(r);
        max_running = std::max<uint32_t>(max_running, (uint32_t)running.size());
      }
    }
    for (Req *r : done) {
      finish(r);
      std::lock_guard<std::mutex> g(mu);
      if (r->status == 1) r->status = 2;
      r->t_done = now_s();
      busy[r->slot] = 0;
      used -= r->bytes;
      ++completed;
      cv.notify_all();
    }
  }
}

extern "C" int pq_server_create(uint64_t budget_bytes, uint32_t slots, pq_server *out) {
  if (!out || !budget_bytes || !slots || slots > 64)
    return pq_set_error(PQ_ERR_ARG, "server needs an output handle, a budget and 1..64 slots");
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess || c < 1)
    return pq_set_error(PQ_ERR_NODEV, "no HIP device available");
  auto sv = std::make_unique<pq_server_s>();
  sv->budget = budget_bytes;
  if (hipGetDevice(&sv->device) != hipSuccess) return pq_set_error(PQ_ERR_HIP, "hipGetDevice");
  // keep freed request memory in the stream-ordered pool for reuse
  hipMemPool_t pool = nullptr;
  if (hipDeviceGetDefaultMemPool(&pool, sv->device) == hipSuccess && pool) {
    uint64_t keep = UINT64_MAX;
    (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
  }
  sv->streams.resize(slots);
  sv->busy.assign(slots, 0);
  for (auto &s : sv->streams)
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess)
      return pq_set_error(PQ_ERR_HIP, "stream creation failed");
  if (pq::prepare_kernels(pq::measure_max_qubits()) != hipSuccess)
    return pq_set_error(PQ_ERR_HIP, "kernel preparation failed");
  pq_server_s *raw = sv.get();
  raw->worker = std::thread([raw] { raw->loop(); });
  raw->builder = std::thread([raw] { raw->build_loop(); });
  *out = sv.release();
  return PQ_OK;
}

extern "C" int pq_server_submit(pq_server sv, uint64_t n, uint64_t B, int64_t seed,
                                const pq_op *ops, uint64_t nops, uint64_t *ticket) {
  const double t_arrive = now_s();  // latency counts this call's own work too
  if (!sv || !ticket || (!ops && nops)) return pq_set_error(PQ_ERR_ARG, "null argument");
  if (n == 0 || B == 0) return pq_set_error(PQ_ERR_ARG, "request needs qubits and shots");
  if (n > pq::measure_max_qubits())
    return pq_set_error(PQ_ERR_UNSUPPORTED, "requests support at most 4095 qubits");
  for (uint64_t i = 0; i < nops; ++i) {
    const pq_op &o = ops[i];
    const bool two = o.code == PQ_OP_CX || o.code == PQ_OP_CZ || o.code == PQ_OP_CX_FLAG ||
                     o.code == PQ_OP_CZ_FLAG;
    if (o.code < PQ_OP_H || o.code > PQ_OP_CZ_FLAG || o.q0 >= n ||
        (two && (o.q1 >= n || o.q1 == o.q0)))
      return pq_set_error(PQ_ERR_ARG, ("request op " + std::to_string(i) + " is invalid").c_str());
  }
  auto r = std::make_unique<Req>();
  r->n = n;
  r->B = B;
  r->W = (B + 63) / 64;
  r->seed = seed;
  r->ops.assign(ops, ops + nops);
  {
    std::lock_guard<std::mutex> g(sv->mu);
    r->lite = sv->form == PQ_RUN_FRAME && frame_eligible(ops, nops);
  }
  r->bytes = request_bytes(n, B, ops, nops, &r->nrec, r->lite);
  r->cost = (double)r->W * (double)(2 * n + 1) * (double)(nops + 1);
  if (r->bytes > sv->budget)
    return pq_set_error(PQ_ERR_OOM, "request can never fit the server's memory budget");
  if (hipEventCreateWithFlags(&r->ev, hipEventDisableTiming) != hipSuccess)
    return pq_set_error(PQ_ERR_HIP, "event creation failed");
  if (r->lite) {
    // the plan of a structure not seen yet is built ahead, on the builder
    r->plan_key = frame_plan_key(n, ops, nops);
    std::lock_guard<std::mutex> g(sv->plans_mu);
    if (!sv->plans.count(r->plan_key) && !sv->building.count(r->plan_key)) {
      sv->building.insert(r->plan_key);
      sv->build_q.push_back(pq_server_s::BuildJob{r->plan_key, n, r->ops});
      r->own_build = true;
      sv->plans_cv.notify_all();
    }
  }
  std::lock_guard<std::mutex> g(sv->mu);
  r->ticket = sv->next_ticket++;
  r->t_submit = t_arrive;
  *ticket = r->ticket;
  sv->pending.push_back(r.get());
  sv->reqs[r->ticket] = std::move(r);
  sv->cv.notify_all();
  return PQ_OK;
}

extern "C" int pq_server_status(pq_server sv, uint64_t ticket, int *status, double *latency_s,
                                double *queue_s) {
  if (!sv || !status) return pq_set_error(PQ_ERR_ARG, "null argument");
  std::lock_guard<std::mutex> g(sv->mu);
  auto it = sv->reqs.find(ticket);
  if (it == sv->reqs.end()) return pq_set_error(PQ_ERR_RANGE, "unknown ticket");
  const Req &r = *it->second;
  *status = r.status;
  const bool fin = r.status == 2 || r.status < 0;
  if (latency_s) *latency_s = fin ? r.t_done - r.t_submit : -1.0;
  if (queue_s) *queue_s = r.t_start > 0 ? r.t_start - r.t_submit : -1.0;
  return PQ_OK;
}

extern "C" int pq_server_result(pq_server sv, uint64_t ticket, uint64_t *words,
                                uint64_t capacity, uint64_t *nrec, uint64_t *W) {
  if (!sv) return pq_set_error(PQ_ERR_ARG, "null server");
  std::lock_guard<std::mutex> g(sv->mu);
  auto it = sv->reqs.find(ticket);
  if (it == sv->reqs.end()) return pq_set_error(PQ_ERR_RANGE, "unknown ticket");
  const Req &r = *it->second;
  if (r.status != 2) return pq_set_error(PQ_ERR_BUSY, "request not finished");
  if (nrec) *nrec = r.nrec;
  if (W) *W = r.W;
  if (words) {
    if (capacity < r.planes.size()) return pq_set_error(PQ_ERR_ARG, "result buffer too small");
    std::copy(r.planes.begin(), r.planes.end(), words);
  }
  return PQ_OK;
}

extern "C" int pq_server_release(pq_server sv, uint64_t ticket) {
  if (!sv) return pq_set_error(PQ_ERR_ARG, "null server");
  std::lock_guard<std::mutex> g(sv->mu);
  auto it = sv->reqs.find(ticket);
  if (it == sv->reqs.end()) return pq_set_error(PQ_ERR_RANGE, "unknown ticket");
  if (it->second->status == 0 || it->second->status == 1)
    return pq_set_error(PQ_ERR_BUSY, "request not finished");
  (void)hipEventDestroy(it->second->ev);
  sv->reqs.erase(it);
  return PQ_OK;
}

extern "C" int pq_server_drain(pq_server sv) {
  if (!sv) return pq_set_error(PQ_ERR_ARG, "null server");
  std::unique_lock<std::mutex> lk(sv->mu);
  sv->cv.wait(lk, [sv] { return sv->pending.empty() && sv->running.empty() &&
                                std::none_of(sv->busy.begin(), sv->busy.end(),
                                             [](char b) { return b != 0; }); });
  return PQ_OK;
}

extern "C" int pq_server_stats(pq_server sv, uint64_t *completed, uint64_t *peak_bytes,
                               uint32_t *max_running, uint64_t *queued) {
  if (!sv) return pq_set_error(PQ_ERR_ARG, "null server");
  std::lock_guard<std::mutex> g(sv->mu);
  if (completed) *completed = sv->completed;
  if (peak_bytes) *peak_bytes = sv->peak;
  if (max_running) *max_running = sv->max_running;
  if (queued) *queued = sv->pending.size();
  return PQ_OK;
}

extern "C" int pq_server_set_policy(pq_server sv, int policy, double age_limit_s) {
  if (!sv) return pq_set_error(PQ_ERR_ARG, "null server");
  if (policy != PQ_SERVE_FCFS && policy != PQ_SERVE_SMALLEST_FIRST)
    return pq_set_error(PQ_ERR_ARG, "unknown serving policy");
  if (!(age_limit_s >= 0)) return pq_set_error(PQ_ERR_ARG, "age limit must be >= 0");
  std::lock_guard<std::mutex> g(sv->mu);
  sv->policy = policy;
  sv->age_limit = age_limit_s;
  return PQ_OK;
}

extern "C" int pq_server_set_form(pq_server sv, int form) {
  if (!sv) return pq_set_error(PQ_ERR_ARG, "null server");
  if (form != PQ_RUN_PERSISTENT && form != PQ_RUN_FRAME)
    return pq_set_error(PQ_ERR_ARG, "server form must be PQ_RUN_PERSISTENT or PQ_RUN_FRAME");
  std::lock_guard<std::mutex> g(sv->mu);
  sv->form = form;
  return PQ_OK;
}

extern "C" int pq_server_request_bytes(pq_server sv, uint64_t n, uint64_t B, const pq_op *ops,
                                       uint64_t nops, uint64_t *bytes) {
  if (!sv || !bytes || (!ops && nops)) return pq_set_error(PQ_ERR_ARG, "null argument");
  bool lite;
  {
    std::lock_guard<std::mutex> g(sv->mu);
    lite = sv->form == PQ_RUN_FRAME && frame_eligible(ops, nops);
  }
  uint64_t nrec = 0;
  *bytes = request_bytes(n, B, ops, nops, &nrec, lite);
  return PQ_OK;
}

extern "C" int pq_server_frame_stats(pq_server sv, uint64_t *frame_runs, uint64_t *plan_builds,
                                     uint64_t *plan_hits) {
  if (!sv) return pq_set_error(PQ_ERR_ARG, "null server");
  std::lock_guard<std::mutex> g(sv->plans_mu);
  if (frame_runs) *frame_runs = sv->frame_runs;
  if (plan_builds) *plan_builds = sv->plan_builds;
  if (plan_hits) *plan_hits = sv->plan_hits;
  return PQ_OK;
}

extern "C" int pq_server_destroy(pq_server sv) {
  if (!sv) return PQ_OK;
  (void)pq_server_drain(sv);
  sv->stop.store(true);
  sv->cv.notify_all();
  if (sv->worker.joinable()) sv->worker.join();
  {
    std::lock_guard<std::mutex> g(sv->plans_mu);
    sv->builder_stop = true;
  }
  sv->plans_cv.notify_all();
  if (sv->builder.joinable()) sv->builder.join();
  for (auto &kv : sv->reqs)
    if (kv.second->ev) (void)hipEventDestroy(kv.second->ev);
  for (hipStream_t s : sv->streams) {
    (void)hipStreamSynchronize(s);
    (void)hipStreamDestroy(s);
  }
  delete sv;
  return PQ_OK;
}
