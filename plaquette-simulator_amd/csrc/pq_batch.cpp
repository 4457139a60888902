// Mixed batches: many states (different qubit and shot counts), each with its
// own compiled program, run by one persistent kernel over a packed work list
// of (state, 64-shot word group) items, optionally as a captured hipGraph.
//
// Per run: every member's record slots are prepared and its coin/noise
// sequence numbers are taken and advanced exactly as a standalone run would,
// so each member's results are identical alone and in the batch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "plaquette_amd.h"
#include "pq_capi_internal.h"
#include "pq_common.h"
#include "pq_kernels.h"
#include "pq_program_internal.h"

using namespace pq;

#define HIPCHK(expr)                                                              \
  do {                                                                            \
    hipError_t e_ = (expr);                                                       \
    if (e_ != hipSuccess)                                                         \
      return pq_set_error(PQ_ERR_HIP,                                             \
                          (std::string(#expr) + ": " + hipGetErrorString(e_)).c_str()); \
  } while (0)

struct pq_batch_s {
  std::vector<std::pair<pq_state, pq_program>> members;
  hipStream_t stream = nullptr;
  BatchCircuit *d_circ = nullptr;
  BatchCircuit *h_circ = nullptr;  // pinned
  FrameArgs *d_fargs = nullptr;    // frame form: per-member kernel arguments
  FrameArgs *h_fargs = nullptr;    // pinned
  uint64_t *d_items = nullptr;
  unsigned long long *d_counter = nullptr;
  uint64_t nitems = 0;
  std::vector<uint64_t> h_items;   // the work list, host copy
  // persistent form: word groups per work item; with more than one, the
  // items are every wtile-th entry of h_items (d_titems, ntitems)
  uint32_t wtile = 1, titems_tile = 0;
  uint64_t *d_titems = nullptr;
  uint64_t ntitems = 0;
  // frame form: the work list split by LDS class (members whose plans let 4
  // groups share a CU, then the rest), one launch each
  uint64_t *d_fitems = nullptr;
  std::vector<hipEvent_t> member_ev;  // frame form: each member's queued work
  uint64_t fitems_n[2] = {0, 0};
  uint32_t flds[2] = {0, 0};
  bool fsplit = false;
  bool built = false;
  uint32_t threads = kProgramThreadsMax;
  uint32_t blocks = 0;             // 0: CUs x blocks per CU
  uint64_t lds_budget = kProgramLdsBytes;
  uint64_t noise_tasks = 0;  // the run's noise tasks (word group x 32 ops), all members
  hipEvent_t ev_upload = nullptr, ev_a = nullptr, ev_b = nullptr;
  bool timed = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  std::vector<uint64_t> graph_gens;  // arena generations at capture
  ~pq_batch_s();
  void drop_graph() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    gexec = nullptr;
    graph = nullptr;
  }
};

pq_batch_s::~pq_batch_s() {
  if (stream) (void)hipStreamSynchronize(stream);
  drop_graph();
  if (d_circ) (void)hipFree(d_circ);
  if (h_circ) (void)hipHostFree(h_circ);
  if (d_fargs) (void)hipFree(d_fargs);
  if (h_fargs) (void)hipHostFree(h_fargs);
  if (d_items) (void)hipFree(d_items);
  if (d_titems) (void)hipFree(d_titems);
  if (d_fitems) (void)hipFree(d_fitems);
  for (hipEvent_t e : member_ev) (void)hipEventDestroy(e);
  if (d_counter) (void)hipFree(d_counter);
  for (hipEvent_t e : {ev_upload, ev_a, ev_b})
    if (e) (void)hipEventDestroy(e);
  if (stream) (void)hipStreamDestroy(stream);
}

static int cu_count() {
  static int cus = 0;
  static std::once_flag once;
  std::call_once(once, [] {
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
      cus = prop.multiProcessorCount;
    if (cus <= 0) cus = 1;
  });
  return cus;
}

extern "C" int pq_batch_create(pq_batch *out) {
  if (!out) return pq_set_error(PQ_ERR_ARG, "null argument");
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess || c < 1)
    return pq_set_error(PQ_ERR_NODEV, "no HIP device available");
  auto b = std::make_unique<pq_batch_s>();
  HIPCHK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
  HIPCHK(hipMalloc(&b->d_counter, sizeof(unsigned long long)));
  HIPCHK(hipEventCreate(&b->ev_upload));
  HIPCHK(hipEventCreate(&b->ev_a));
  HIPCHK(hipEventCreate(&b->ev_b));
  HIPCHK(prepare_batch_kernel());
  *out = b.release();
  return PQ_OK;
}

extern "C" int pq_batch_destroy(pq_batch b) {
  delete b;
  return PQ_OK;
}

extern "C" int pq_batch_add(pq_batch b, pq_state s, pq_program p) {
  if (!b || !s || !p) return pq_set_error(PQ_ERR_ARG, "null argument");
  if (p->n != s->n || p->W != s->W)
    return pq_set_error(PQ_ERR_ARG, "program was compiled for a state of another shape");
  if (s->n > measure_max_qubits())
    return pq_set_error(PQ_ERR_UNSUPPORTED, "batched form supports at most 4095 qubits");
  for (auto &m : b->members) {
    if (m.first == s) return pq_set_error(PQ_ERR_ARG, "state already in the batch");
    // a program's noise planes and side buffer serve one run at a time
    if (m.second == p)
      return pq_set_error(PQ_ERR_ARG,
                          "program already in the batch: give each member a program of its own");
  }
  if (b->members.size() >= (1u << 23)) return pq_set_error(PQ_ERR_UNSUPPORTED, "too many members");
  b->members.emplace_back(s, p);
  b->built = false;
  b->drop_graph();
  return PQ_OK;
}

extern "C" int pq_batch_set_launch(pq_batch b, uint32_t threads, uint32_t blocks,
                                   uint64_t lds_budget) {
  if (!b) return pq_set_error(PQ_ERR_ARG, "null batch");
  if (threads) {
    if (threads % 64 || threads > kProgramThreadsMax)
      return pq_set_error(PQ_ERR_ARG,
                          "threads must be a multiple of 64, at most 768 (a block of more "
                          "than 512 keeps the chains of members of 193..512 qubits in LDS)");
    b->threads = threads;
  }
  b->blocks = blocks;
  if (lds_budget) {
    if (lds_budget < 40 * 1024 || lds_budget > kProgramLdsBytes)
      return pq_set_error(PQ_ERR_ARG, "LDS budget must be in [40 KB, 158 KB]");
    b->lds_budget = lds_budget;
  }
  b->drop_graph();
  return PQ_OK;
}

// Packed work list, heaviest circuits first (longest-processing-time order),
// word groups of one circuit consecutive.
static int build(pq_batch b) {
  const uint32_t nc = (uint32_t)b->members.size();
  std::vector<std::pair<double, uint32_t>> order;
  uint64_t items = 0;
  for (uint32_t c = 0; c < nc; ++c) {
    pq_state s = b->members[c].first;
    pq_program p = b->members[c].second;
    const double cost = (double)s->R * ((double)p->instrs.size() + 32.0 * p->nmeas);
    order.emplace_back(-cost, c);
    items += s->W;
  }
  std::stable_sort(order.begin(), order.end());
  std::vector<uint64_t> list;
  list.reserve(items);
  for (auto &o : order) {
    const uint32_t c = o.second;
    for (uint64_t w = 0; w < b->members[c].first->W; ++w)
      list.push_back(((uint64_t)c << 40) | w);
  }
  if (b->d_items) HIPCHK(hipFree(b->d_items));
  b->d_items = nullptr;
  if (b->d_titems) HIPCHK(hipFree(b->d_titems));
  b->d_titems = nullptr;
  b->titems_tile = 0;
  if (b->d_circ) HIPCHK(hipFree(b->d_circ));
  b->d_circ = nullptr;
  if (b->h_circ) HIPCHK(hipHostFree(b->h_circ));
  b->h_circ = nullptr;
  if (b->d_fargs) HIPCHK(hipFree(b->d_fargs));
  b->d_fargs = nullptr;
  if (b->h_fargs) HIPCHK(hipHostFree(b->h_fargs));
  b->h_fargs = nullptr;
  b->nitems = list.size();
  b->h_items = list;
  if (b->d_fitems) HIPCHK(hipFree(b->d_fitems));
  b->d_fitems = nullptr;
  b->fsplit = false;
  if (!list.empty()) {
    HIPCHK(hipMalloc(&b->d_items, list.size() * 8));
    HIPCHK(hipMemcpy(b->d_items, list.data(), list.size() * 8, hipMemcpyHostToDevice));
  }
  if (nc) {
    HIPCHK(hipMalloc(&b->d_circ, nc * sizeof(BatchCircuit)));
    HIPCHK(hipHostMalloc(&b->h_circ, nc * sizeof(BatchCircuit), hipHostMallocDefault));
    HIPCHK(hipMalloc(&b->d_fargs, nc * sizeof(FrameArgs)));
    HIPCHK(hipHostMalloc(&b->h_fargs, nc * sizeof(FrameArgs), hipHostMallocDefault));
  }
  b->built = true;
  return PQ_OK;
}

static uint32_t launch_blocks(pq_batch b) {
  if (b->blocks) return (uint32_t)std::min<uint64_t>(b->blocks, std::max<uint64_t>(1, b->nitems));
  const uint64_t per_cu = std::max<uint64_t>(1, (160ull * 1024) / (b->lds_budget + 2048));
  return (uint32_t)std::max<uint64_t>(
      1, std::min<uint64_t>((uint64_t)cu_count() * per_cu, b->nitems));
}

// Dynamic LDS of the launch: the largest any member asks for (members'
// arguments carry it, so row-tiled gate segments size their tiles to it).
static size_t launch_lds(pq_batch b) {
  size_t lds = 0;
  for (uint32_t c = 0; c < b->members.size(); ++c)
    lds = std::max<size_t>(lds, b->h_circ[c].a.lds_words * 8);
  return lds;
}

// The persistent form's work list for word tiles of b->wtile groups: the
// items of h_items whose word group is a multiple of the tile (each
// circuit's groups are consecutive there, from 0).
static int tiled_items(pq_batch b) {
  if (b->wtile <= 1 || b->titems_tile == b->wtile) return PQ_OK;
  std::vector<uint64_t> list;
  for (uint64_t it : b->h_items)
    if ((it & ((1ull << 40) - 1)) % b->wtile == 0) list.push_back(it);
  if (b->d_titems) HIPCHK(hipFree(b->d_titems));
  b->d_titems = nullptr;
  if (!list.empty()) {
    HIPCHK(hipMalloc(&b->d_titems, list.size() * 8));
    HIPCHK(hipMemcpy(b->d_titems, list.data(), list.size() * 8, hipMemcpyHostToDevice));
  }
  b->ntitems = list.size();
  b->titems_tile = b->wtile;
  return PQ_OK;
}

static int enqueue(pq_batch b) {
  const uint32_t nc = (uint32_t)b->members.size();
  // members whose determined chains the 512-thread build keeps in registers
  uint32_t chain_n = 0;
  for (uint32_t c = 0; c < nc; ++c)
    if (program_max_threads(b->h_circ[c].a.n, b->h_circ[c].a.wide) < kProgramThreadsMax)
      chain_n = std::max<uint32_t>(chain_n, b->h_circ[c].a.n);
  const bool tiled = b->wtile > 1;
  return launch_batch(b->stream, b->d_circ, nc, tiled ? b->d_titems : b->d_items,
                      tiled ? b->ntitems : b->nitems, b->d_counter, b->threads,
                      launch_blocks(b), launch_lds(b), b->noise_tasks, chain_n,
                      b->wtile) == hipSuccess
             ? PQ_OK
             : pq_set_error(PQ_ERR_HIP, "batch launch failed");
}

// The frame form of a batch: every member must allow it (no flagged
// H/S/CX/CZ) and start from the |0...0> tableau; members with equal
// structure share one plan.  One noise prologue for all members, frame
// kernels over the packed work list (by LDS class); each member's shared X/Z
// are written when something reads or updates its tableau.
static int run_frame_batch(pq_batch b) {
  const uint32_t nc = (uint32_t)b->members.size();
  // everything that can refuse the run, before any member changes
  std::unordered_map<std::string, std::shared_ptr<FramePlan>> plans;
  for (uint32_t c = 0; c < nc; ++c) {
    pq_state s = b->members[c].first;
    pq_program p = b->members[c].second;
    if (s->tab && !s->xz_identity)
      return pq_set_error(PQ_ERR_UNSUPPORTED,
                          "frame form runs from the |0...0> tableau: reset every member first");
    if (p->frame) continue;
    std::string key = frame_plan_key(p->n, p->ops.data(), p->ops.size());
    auto it = plans.find(key);
    if (it != plans.end()) {
      p->frame = it->second;
    } else {
      int rc = build_frame_plan(p, UINT64_MAX);
      if (rc) return rc;
      plans.emplace(std::move(key), p->frame);
    }
  }
  HIPCHK(hipEventSynchronize(b->ev_upload));
  for (uint32_t c = 0; c < nc; ++c) {
    pq_state s = b->members[c].first;
    pq_program p = b->members[c].second;
    int rc = prepare_records(s, p);
    if (rc) return rc;
    // the member's own queued work (its reset, a new record arena) first,
    // ordered on the device rather than waited for here
    while (b->member_ev.size() <= c) {
      hipEvent_t e;
      HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
      b->member_ev.push_back(e);
    }
    HIPCHK(hipEventRecord(b->member_ev[c], s->stream));
    HIPCHK(hipStreamWaitEvent(b->stream, b->member_ev[c], 0));
    // and the program's last standalone run, wherever it ran
    if (p->done_ev) HIPCHK(hipStreamWaitEvent(b->stream, p->done_ev, 0));
    if ((rc = upload_frame_plan(p, b->stream))) return rc;
    BatchCircuit &bc = b->h_circ[c];
    bc = BatchCircuit{};
    bc.a.planes = p->d_planes;
    bc.a.W = s->tab ? s->W : 0;
    bc.seed = (uint64_t)s->seed;
    bc.mseq0 = s->meas_seq;
    bc.nseq0 = s->noise_seq;
    bc.noise = p->d_noise;
    bc.nnoise = (uint32_t)p->nnoise;
    b->h_fargs[c] = frame_args(s, p);
    if (!s->tab) b->h_fargs[c].W = 0;
    s->meas_seq += p->nmeas;
    s->noise_seq += p->nnoise;
    if (s->tab) s->xz_identity = p->frame->final_identity;
  }
  if (!b->fsplit) {
    // small members: LDS for at least 4 groups per CU
    std::vector<uint8_t> small(nc);
    for (uint32_t c = 0; c < nc; ++c) {
      const uint32_t l = b->members[c].second->frame->lds_bytes;
      small[c] = l * 4ull <= 160ull * 1024;
    }
    std::vector<uint64_t> split;
    split.reserve(b->h_items.size());
    b->fitems_n[0] = b->fitems_n[1] = 0;
    b->flds[0] = b->flds[1] = 0;
    for (int k = 0; k < 2; ++k)
      for (uint64_t it : b->h_items) {
        const uint32_t c = (uint32_t)(it >> 40);
        if (small[c] != k) continue;
        split.push_back(it);
        ++b->fitems_n[k];
        b->flds[k] = std::max(b->flds[k], b->members[c].second->frame->lds_bytes);
      }
    if (!split.empty()) {
      HIPCHK(hipMalloc(&b->d_fitems, split.size() * 8));
      HIPCHK(hipMemcpy(b->d_fitems, split.data(), split.size() * 8, hipMemcpyHostToDevice));
    }
    b->fsplit = true;
  }
  HIPCHK(hipMemcpyAsync(b->d_fargs, b->h_fargs, nc * sizeof(FrameArgs), hipMemcpyHostToDevice,
                        b->stream));
  // the noise tasks (word group x 32 ops) of all members in one numbering
  uint64_t tasks = 0;
  for (uint32_t c = 0; c < nc; ++c) {
    BatchCircuit &bc = b->h_circ[c];
    bc.task0 = tasks;
    tasks += bc.a.W * ((bc.nnoise + 31) / 32);
  }
  HIPCHK(hipMemcpyAsync(b->d_circ, b->h_circ, nc * sizeof(BatchCircuit), hipMemcpyHostToDevice,
                        b->stream));
  HIPCHK(hipEventRecord(b->ev_upload, b->stream));
  if (b->timed) HIPCHK(hipEventRecord(b->ev_a, b->stream));
  HIPCHK(launch_batch_frame_noise(b->stream, b->d_circ, nc, tasks));
  // two launches: the members whose plans fit 4 groups per CU would run at
  // the largest member's occupancy in one launch sized for it
  for (int k = 0, off = 0; k < 2; off += (int)b->fitems_n[k], ++k)
    if (b->fitems_n[k])
      HIPCHK(launch_frame_batch(b->stream, b->d_fargs, b->d_fitems + off, b->fitems_n[k],
                                b->flds[k]));
  for (uint32_t c = 0; c < nc; ++c) {
    pq_state s = b->members[c].first;
    if (s->tab && !s->r_only) {
      // X/Z are the plan's shared X/Z, written when needed (materialize_tableau)
      // or now if the raw tableau pointer was handed out
      s->xz_pending = false;
      s->xz_plan = b->members[c].second->frame;
      if (s->exposed) {
        int rc = materialize_tableau(s);
        if (rc) return rc;
      }
    }
  }
  if (b->timed) HIPCHK(hipEventRecord(b->ev_b, b->stream));
  return PQ_OK;
}

extern "C" int pq_batch_run(pq_batch b, int form) {
  if (!b) return pq_set_error(PQ_ERR_ARG, "null batch");
  if (form != PQ_RUN_PERSISTENT && form != PQ_RUN_GRAPH && form != PQ_RUN_FRAME)
    return pq_set_error(PQ_ERR_ARG, "a batch runs in the persistent, graph or frame form");
  if (b->members.empty()) return PQ_OK;
  if (form != PQ_RUN_FRAME)
    for (auto &m : b->members)
      if (m.first->r_only)
        return pq_set_error(PQ_ERR_UNSUPPORTED,
                            "a sampling state in the batch: the batch runs in the frame form only");
  if (form == PQ_RUN_FRAME) {
    if (!b->built) {
      int rc = build(b);
      if (rc) return rc;
    }
    return run_frame_batch(b);
  }
  if (b->members.empty()) return PQ_OK;
  if (!b->built) {
    int rc = build(b);
    if (rc) return rc;
  }
  // the pinned descriptor buffer is reused: the previous upload must be done
  HIPCHK(hipEventSynchronize(b->ev_upload));
  const uint32_t nc = (uint32_t)b->members.size();
  std::vector<uint64_t> gens(nc);
  for (uint32_t c = 0; c < nc; ++c) {
    pq_state s = b->members[c].first;
    pq_program p = b->members[c].second;
    int rc = prepare_records(s, p);
    if (rc) return rc;
    if ((rc = materialize_tableau(s))) return rc;
    HIPCHK(hipStreamSynchronize(s->stream));  // the member's own queued work first
    if (p->done_ev) HIPCHK(hipEventSynchronize(p->done_ev));  // the program's last run
    BatchCircuit &bc = b->h_circ[c];
    bc.a = program_args(s, p, b->lds_budget);
    bc.a.census = nullptr;
    bc.seed = (uint64_t)s->seed;
    bc.mseq0 = s->meas_seq;
    bc.nseq0 = s->noise_seq;
    bc.noise = p->d_noise;
    bc.nnoise = (uint32_t)p->nnoise;

    s->meas_seq += p->nmeas;
    s->xz_identity = false;
    s->noise_seq += p->nnoise;
    gens[c] = s->arena_gen;
  }
  if (form != PQ_RUN_FRAME) {
    int rc = tiled_items(b);  // allocates: before any capture
    if (rc) return rc;
  }
  const uint64_t lds_words = launch_lds(b) / 8;
  // the noise tasks (word group x 32 ops) of all members in one numbering
  b->noise_tasks = 0;
  for (uint32_t c = 0; c < nc; ++c) {
    BatchCircuit &bc = b->h_circ[c];
    bc.a.lds_words = lds_words;
    bc.task0 = b->noise_tasks;
    b->noise_tasks += bc.a.W * ((bc.nnoise + 31) / 32);
  }
  HIPCHK(hipMemcpyAsync(b->d_circ, b->h_circ, nc * sizeof(BatchCircuit),
                        hipMemcpyHostToDevice, b->stream));
  HIPCHK(hipEventRecord(b->ev_upload, b->stream));
  if (b->timed) HIPCHK(hipEventRecord(b->ev_a, b->stream));
  if (form == PQ_RUN_PERSISTENT) {
    int rc = enqueue(b);
    if (rc) return rc;
  } else {
    if (b->gexec && b->graph_gens != gens) b->drop_graph();
    if (!b->gexec) {
      HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
      int rc = enqueue(b);
      hipGraph_t g = nullptr;
      hipError_t e = hipStreamEndCapture(b->stream, &g);
      if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
      }
      HIPCHK(e);
      b->graph = g;
      HIPCHK(hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0));
      b->graph_gens = gens;
    }
    HIPCHK(hipGraphLaunch(b->gexec, b->stream));
  }
  if (b->timed) HIPCHK(hipEventRecord(b->ev_b, b->stream));
  return PQ_OK;
}

extern "C" int pq_batch_wait(pq_batch b) {
  if (!b) return pq_set_error(PQ_ERR_ARG, "null batch");
  HIPCHK(hipStreamSynchronize(b->stream));
  return PQ_OK;
}

extern "C" int pq_batch_set_timing(pq_batch b, int enable) {
  if (!b) return pq_set_error(PQ_ERR_ARG, "null batch");
  b->timed = enable != 0;
  return PQ_OK;
}

extern "C" int pq_batch_last_ms(pq_batch b, float *ms) {
  if (!b || !ms) return pq_set_error(PQ_ERR_ARG, "null argument");
  HIPCHK(hipEventSynchronize(b->ev_b));
  HIPCHK(hipEventElapsedTime(ms, b->ev_a, b->ev_b));
  return PQ_OK;
}

extern "C" int pq_batch_set_word_tile(pq_batch b, uint32_t words) {
  if (!b) return pq_set_error(PQ_ERR_ARG, "null batch");
  if (words < 1 || words > 64) return pq_set_error(PQ_ERR_ARG, "word tile must be 1..64 groups");
  if (words != b->wtile) b->drop_graph();
  b->wtile = words;
  return PQ_OK;
}

extern "C" int pq_batch_word_tile(pq_batch b, uint32_t *words) {
  if (!b || !words) return pq_set_error(PQ_ERR_ARG, "null argument");
  *words = b->wtile;
  return PQ_OK;
}

extern "C" int pq_batch_info(pq_batch b, uint64_t *members, uint64_t *items,
                             uint32_t *threads, uint32_t *blocks, uint64_t *lds_budget) {
  if (!b) return pq_set_error(PQ_ERR_ARG, "null batch");
  if (!b->built && !b->members.empty()) {
    int rc = build(b);
    if (rc) return rc;
  }
  if (members) *members = b->members.size();
  if (items) *items = b->nitems;
  if (threads) *threads = b->threads;
  if (blocks) *blocks = launch_blocks(b);
  if (lds_budget) *lds_budget = b->lds_budget;
  return PQ_OK;
}

// Autotuner: times each launch shape (block size x LDS budget, the latter
// setting how many workgroups share a CU and how many waves run fully
// determined measurements in parallel) on the batch itself and keeps the
// fastest, then the word tile (1, 2 or 4 word groups per work item) at that
// shape.  Trial runs are real runs: afterwards every member is reset to
// |0...0> and its coin/noise sequence numbers are restored.  Choices are
// cached per batch shape (member qubit counts, word-group counts and program
// sizes) for the life of the process.
extern "C" int pq_batch_autotune(pq_batch b, int reps, uint32_t *threads_out,
                                 uint64_t *lds_out, float *ms_out) {
  if (!b) return pq_set_error(PQ_ERR_ARG, "null batch");
  if (b->members.empty()) return PQ_OK;
  static std::mutex mu;
  struct Choice {
    uint32_t threads;
    uint64_t lds;
    uint32_t wtile;
  };
  static std::map<std::vector<uint64_t>, Choice> cache;
  std::vector<uint64_t> key;
  for (auto &m : b->members) {
    key.push_back(m.first->n);
    key.push_back(m.first->W);
    key.push_back(m.second->instrs.size());
  }
  {
    std::lock_guard<std::mutex> g(mu);
    auto it = cache.find(key);
    if (it != cache.end()) {
      int rc = pq_batch_set_launch(b, it->second.threads, 0, it->second.lds);
      if (rc == PQ_OK) rc = pq_batch_set_word_tile(b, it->second.wtile);
      if (threads_out) *threads_out = it->second.threads;
      if (lds_out) *lds_out = it->second.lds;
      if (ms_out) *ms_out = -1.0f;
      return rc;
    }
  }
  std::vector<std::pair<uint64_t, uint64_t>> saved;
  for (auto &m : b->members) saved.emplace_back(m.first->meas_seq, m.first->noise_seq);
  const uint32_t cand_threads[] = {256, 512, 768};
  const uint64_t cand_lds[] = {kProgramLdsBytes, 72 * 1024, 48 * 1024};
  float best = 1e30f;
  uint32_t bt = b->threads, bw = 1;
  uint64_t bl = b->lds_budget;
  const bool was_timed = b->timed;
  b->timed = true;
  // average HIP-event time of `reps` runs after one warm-up run
  auto time_runs = [&](float *avg) -> int {
    float tot = 0;
    for (int r = 0; r < std::max(1, reps) + 1; ++r) {
      int rc = pq_batch_run(b, PQ_RUN_PERSISTENT);
      if (rc) return rc;
      float ms = 0;
      rc = pq_batch_last_ms(b, &ms);
      if (rc) return rc;
      if (r > 0) tot += ms;  // first run warms up
    }
    *avg = tot / std::max(1, reps);
    return PQ_OK;
  };
  int rc = pq_batch_set_word_tile(b, 1);
  if (rc) return rc;
  for (uint32_t t : cand_threads) {
    for (uint64_t l : cand_lds) {
      // every member must fit: base area + one determined-wave slab
      bool fits = true;
      for (auto &m : b->members)
        if (program_lds_bytes((uint32_t)m.first->n, 1) > l) fits = false;
      if (!fits) continue;
      rc = pq_batch_set_launch(b, t, 0, l);
      if (rc) return rc;
      float avg = 0;
      if ((rc = time_runs(&avg))) return rc;
      if (avg < best) {
        best = avg;
        bt = t;
        bl = l;
      }
    }
  }
  // then the word tile (word groups per work item) at the best shape
  if ((rc = pq_batch_set_launch(b, bt, 0, bl))) return rc;
  for (uint32_t wt : {2u, 4u}) {
    if ((rc = pq_batch_set_word_tile(b, wt))) return rc;
    float avg = 0;
    if ((rc = time_runs(&avg))) return rc;
    if (avg < best) {
      best = avg;
      bw = wt;
    }
  }
  b->timed = was_timed;
  for (size_t c = 0; c < b->members.size(); ++c) {
    pq_state s = b->members[c].first;
    int rc = pq_reset(s);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(s->stream));
    s->meas_seq = saved[c].first;
    s->noise_seq = saved[c].second;
  }
  rc = pq_batch_set_launch(b, bt, 0, bl);
  if (rc == PQ_OK) rc = pq_batch_set_word_tile(b, bw);
  if (rc) return rc;
  {
    std::lock_guard<std::mutex> g(mu);
    cache[key] = {bt, bl, bw};
  }
  if (threads_out) *threads_out = bt;
  if (lds_out) *lds_out = bl;
  if (ms_out) *ms_out = best;
  return PQ_OK;
}
