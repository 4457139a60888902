// Op-list programs: validation, compilation for the persistent kernel, and the
// three execution forms (plain per-op launches, one persistent kernel, a
// captured hipGraph of the persistent form).
//
// Compilation of a gate segment (a maximal run of non-measurement ops):
//  * ASAP layering on qubit dependencies.  Two ops on disjoint qubits commute
//    bit-exactly on the tableau: each rewrites only its own columns and XORs
//    into r a function of those columns (GateFunctors.hpp rules), so any
//    order that keeps per-qubit order gives identical words.
//  * A NOISE op on qubit q is fused into the op that last touched q in the
//    segment (applied to the columns that op produced), else kept alone.
//  * Each layer is cut into batches of at most 8 qubit-disjoint ops; the
//    kernel issues a whole batch's loads before its stores.
// Measurement segments keep program order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <memory>
#include <queue>
#include <string>
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

namespace {

bool is_two(uint32_t c) {
  return c == PQ_OP_CX || c == PQ_OP_CZ || c == PQ_OP_CX_FLAG || c == PQ_OP_CZ_FLAG;
}
bool is_flag(uint32_t c) {
  return c == PQ_OP_H_FLAG || c == PQ_OP_S_FLAG || c == PQ_OP_X_FLAG ||
         c == PQ_OP_Y_FLAG || c == PQ_OP_Z_FLAG || c == PQ_OP_CX_FLAG || c == PQ_OP_CZ_FLAG;
}
// A flagged op whose flag is a Bernoulli draw (a noise stream of its own).
bool is_random_flag(const pq_op &o) { return is_flag(o.code) && o.slot == PQ_FLAG_RANDOM; }

int validate(uint64_t n, const pq_op *ops, uint64_t nops) {
  for (uint64_t i = 0; i < nops; ++i) {
    const pq_op &o = ops[i];
    if (o.code < PQ_OP_H || o.code > PQ_OP_CZ_FLAG)
      return pq_set_error(PQ_ERR_ARG, ("op " + std::to_string(i) + ": unknown code").c_str());
    if (o.q0 >= n)
      return pq_set_error(PQ_ERR_ARG, ("op " + std::to_string(i) + ": qubit out of range").c_str());
    if (is_two(o.code) && (o.q1 >= n || o.q1 == o.q0))
      return pq_set_error(PQ_ERR_ARG,
                          ("op " + std::to_string(i) + ": two-qubit op needs distinct qubits in range").c_str());
    if (o.code == PQ_OP_NOISE && !(o.t0 <= o.t1 && o.t1 <= o.t2))
      return pq_set_error(PQ_ERR_ARG, ("op " + std::to_string(i) + ": noise thresholds not cumulative").c_str());
    if (is_random_flag(o) && o.t0 > (1u << 24))
      return pq_set_error(PQ_ERR_ARG, ("op " + std::to_string(i) + ": flag threshold above 2^24").c_str());
    if (is_flag(o.code) && !is_random_flag(o) && (o.slot & kSlotPlane))
      return pq_set_error(PQ_ERR_RANGE, ("op " + std::to_string(i) + ": flag record out of range").c_str());
  }
  return PQ_OK;
}

}  // namespace

int validate_ops(uint64_t n, const pq_op *ops, uint64_t nops) { return validate(n, ops, nops); }

// Compile `ops` into segments/instructions (host side only).
// Run lengths of a gate segment: a run starts at every layer start, at
// every staging-chunk start and where the op code changes (a layer's ops are
// sorted by code), and ends before the next one; its length goes into bits
// 8-19 of its first instruction's pad.  Every run then has one op code, so
// the row-tiled kernel dispatches on it once per wave share.
static void mark_runs(pq_program_s *p, const DSegment &sg) {
  const uint32_t end = (uint32_t)p->instrs.size();
  uint32_t start = sg.begin;
  for (uint32_t k = sg.begin + 1; k <= end; ++k) {
    const bool brk = k == end || (p->instrs[k].pad & kLayerStart) ||
                     (k - sg.begin) % kProgramStage == 0 ||
                     p->instrs[k].code != p->instrs[k - 1].code;
    if (brk) {
      p->instrs[start].pad |= (k - start) << kRunShift;
      start = k;
    }
  }
}

// Row tiling of a gate segment.  Gates act on each tableau row on its own,
// so a tile of rows restricted to the columns the segment touches can be
// loaded into LDS once, carry every op of the segment there, and be written
// back once.  Per row that moves 4 words per touched qubit (x and z, in and
// out); issuing the ops one by one against global memory moves the columns
// each op reads and writes (H 4, S 3, CX/CZ 6, Paulis and noise 2).  The
// segment is marked for tiling when that is at least 1.5x the tiled traffic;
// the kernel still falls back to the direct form when too few rows fit.
static void plan_tiling(pq_program_s *p, DSegment &sg) {
  const uint32_t end = (uint32_t)p->instrs.size();
  std::vector<int32_t> local(p->n, -1);
  std::vector<uint32_t> qs;
  uint64_t direct = 0;
  auto loc = [&](uint32_t q) -> uint32_t {
    if (local[q] < 0) {
      local[q] = (int32_t)qs.size();
      qs.push_back(q);
    }
    return (uint32_t)local[q];
  };
  for (uint32_t k = sg.begin; k < end; ++k) {
    DInstr &I = p->instrs[k];
    // one-qubit ops name their own column twice (the kernel's branch-free
    // update then rewrites it with the final value last)
    const uint32_t la = loc(I.qa);
    const uint32_t lb = is_two(I.code) ? loc(I.qb) : la;
    I.t0 = la | (lb << 16);
    switch (I.code) {
      case PQ_OP_H: case PQ_OP_H_FLAG: direct += 4; break;
      case PQ_OP_S: case PQ_OP_S_FLAG: direct += 3; break;
      case PQ_OP_CX: case PQ_OP_CZ: case PQ_OP_CX_FLAG: case PQ_OP_CZ_FLAG: direct += 6; break;
      default: direct += 2; break;
    }
  }
  const uint64_t tiled = 4ull * qs.size() + 2;
  if (qs.size() <= 0xFFFFu && 2 * direct >= 3 * tiled) {
    sg.nq = (uint32_t)qs.size();
    sg.qoff = (uint32_t)p->segq.size();
    p->segq.insert(p->segq.end(), qs.begin(), qs.end());
    p->tiled = true;
  }
}

// One gate segment from ops in an order that respects every qubit's program
// order: ASAP layers on qubit dependencies (ops of a layer act on distinct
// qubits, so they commute bit-exactly: each rewrites only its own columns and
// XORs into r a function of them), each layer's ops ordered by code and cut
// into batches, runs marked, row tiling planned.
static void emit_gate_segment(pq_program_s *p, const std::vector<DInstr> &seg) {
  if (seg.empty()) return;
  DSegment sg{0u, (uint32_t)p->instrs.size(), 0, 0, 0, 0, 0, 0};
  std::vector<int32_t> qlayer(p->n, -1);
  std::vector<uint32_t> layer(seg.size());
  for (size_t k = 0; k < seg.size(); ++k) {
    const DInstr &I = seg[k];
    int32_t L = qlayer[I.qa] + 1;
    if (is_two(I.code)) L = std::max(L, qlayer[I.qb] + 1);
    qlayer[I.qa] = L;
    if (is_two(I.code)) qlayer[I.qb] = L;
    layer[k] = (uint32_t)L;
  }
  std::vector<uint32_t> order(seg.size());
  for (uint32_t k = 0; k < order.size(); ++k) order[k] = k;
  // by layer, and inside a layer by op code
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
    return layer[a] != layer[b] ? layer[a] < layer[b] : seg[a].code < seg[b].code;
  });
  uint32_t k = 0;
  while (k < order.size()) {
    uint32_t e = k;
    while (e < order.size() && layer[order[e]] == layer[order[k]]) ++e;
    for (uint32_t b = k; b < e;) {
      // batches never straddle a staging chunk of the kernel
      const uint32_t pos = (uint32_t)(p->instrs.size() - sg.begin) % kProgramStage;
      const uint32_t len = std::min<uint32_t>({kProgramBatch, e - b, kProgramStage - pos});
      for (uint32_t u = 0; u < len; ++u) {
        DInstr I = seg[order[b + u]];
        I.pad = (u == 0) ? (len | (b == k ? kLayerStart : 0u)) : 0u;
        p->instrs.push_back(I);
      }
      b += len;
    }
    k = e;
  }
  mark_runs(p, sg);
  plan_tiling(p, sg);
  sg.count = (uint32_t)(p->instrs.size() - sg.begin);
  p->segs.push_back(sg);
}

static uint32_t count_qubits(const std::vector<DInstr> &seg, uint64_t n) {
  std::vector<char> seen(n, 0);
  uint32_t c = 0;
  for (const DInstr &I : seg) {
    if (!seen[I.qa]) { seen[I.qa] = 1; ++c; }
    if (is_two(I.code) && !seen[I.qb]) { seen[I.qb] = 1; ++c; }
  }
  return c;
}

// Most qubits whose row tile keeps `rows` rows in the default LDS budget.
static uint32_t tile_qubits(uint64_t n, uint32_t rows) {
  const uint64_t words = kProgramLdsBytes / 8;
  uint32_t lo = 0, hi = (uint32_t)std::min<uint64_t>(n, 0xFFFFu);
  while (lo < hi) {  // largest nq with program_tile_rows(words, nq) >= rows
    const uint32_t mid = (lo + hi + 1) / 2;
    if (pq::program_tile_rows(words, mid) >= rows) lo = mid; else hi = mid - 1;
  }
  return lo;
}

static std::atomic<uint32_t> g_window_qubits{0};

// Window size for gate segments wider than *cut qubits (0: windows off).
static uint32_t window_qubits(const pq_program_s *p, uint32_t *cut) {
  const uint32_t set = g_window_qubits.load();
  if (set) {
    *cut = set == ~0u ? 0u : set;
    return *cut;
  }
  *cut = tile_qubits(p->n, kWindowCutRows);
  return tile_qubits(p->n, kWindowRows);
}

// Cuts a gate segment (ops in program order) into windows of at most `cap`
// distinct qubits, each a list of op indices whose concatenation keeps every
// qubit's ops in program order (an op is taken only when it is the next one
// on each of its qubits).  A window takes, among the ready ops, one that needs
// the fewest qubits it does not hold yet (the earliest of those), while it
// fits: it grows over neighbouring qubits and runs the layers on them as far
// as the ops outside let it before it closes.  A ready op's qubits can join
// the window only through the op itself, so its count of new qubits holds
// until the window closes.
static std::vector<std::vector<uint32_t>> window_schedule(const std::vector<DInstr> &seg,
                                                          uint64_t n, uint32_t cap) {
  const uint32_t nops = (uint32_t)seg.size();
  auto nq_of = [&](uint32_t i) { return is_two(seg[i].code) ? 2u : 1u; };
  auto q_of = [&](uint32_t i, uint32_t t) { return t ? seg[i].qb : seg[i].qa; };
  // per-qubit op lists (CSR)
  std::vector<uint32_t> off(n + 1, 0), pos(n, 0);
  for (uint32_t i = 0; i < nops; ++i)
    for (uint32_t t = 0; t < nq_of(i); ++t) ++off[q_of(i, t) + 1];
  for (uint64_t q = 0; q < n; ++q) off[q + 1] += off[q];
  std::vector<uint32_t> list(off[n]);
  {
    std::vector<uint32_t> fill(off.begin(), off.end() - 1);
    for (uint32_t i = 0; i < nops; ++i)
      for (uint32_t t = 0; t < nq_of(i); ++t) list[fill[q_of(i, t)]++] = i;
  }
  auto next_on = [&](uint64_t q) -> int64_t {
    return off[q] + pos[q] < off[q + 1] ? (int64_t)list[off[q] + pos[q]] : -1;
  };
  auto ready = [&](uint32_t i) {
    for (uint32_t t = 0; t < nq_of(i); ++t)
      if (next_on(q_of(i, t)) != (int64_t)i) return false;
    return true;
  };
  std::vector<char> done(nops, 0), in_win(n, 0);
  std::vector<uint32_t> win_q;
  auto fresh = [&](uint32_t i) {
    uint32_t f = 0;
    for (uint32_t t = 0; t < nq_of(i); ++t) f += in_win[q_of(i, t)] ? 0u : 1u;
    return f;
  };
  // ready ops by the qubits they need beyond the window: 0, 1, 2
  typedef std::priority_queue<uint32_t, std::vector<uint32_t>, std::greater<uint32_t>> MinHeap;
  MinHeap heap[3];
  for (uint32_t i = 0; i < nops; ++i)
    if (ready(i)) heap[nq_of(i)].push(i);
  std::vector<std::vector<uint32_t>> parts;
  uint32_t left = nops;
  while (left) {
    for (uint32_t q : win_q) in_win[q] = 0;
    win_q.clear();
    // an empty window: every ready op needs all its qubits
    std::vector<uint32_t> keep;
    for (int h = 0; h < 2; ++h)
      for (; !heap[h].empty(); heap[h].pop())
        if (!done[heap[h].top()]) keep.push_back(heap[h].top());
    for (uint32_t i : keep) heap[nq_of(i)].push(i);
    std::vector<uint32_t> part;
    for (;;) {
      int64_t pick = -1;
      for (int h = 0; h < 3 && pick < 0; ++h) {
        while (!heap[h].empty() && done[heap[h].top()]) heap[h].pop();
        if (heap[h].empty()) continue;
        const uint32_t i = heap[h].top();
        if (!part.empty() && win_q.size() + fresh(i) > cap) continue;
        heap[h].pop();
        pick = i;
      }
      if (pick < 0) break;
      const uint32_t i = (uint32_t)pick;
      done[i] = 1;
      --left;
      part.push_back(i);
      for (uint32_t t = 0; t < nq_of(i); ++t) {
        const uint32_t q = q_of(i, t);
        if (!in_win[q]) {
          in_win[q] = 1;
          win_q.push_back(q);
        }
        ++pos[q];
      }
      for (uint32_t t = 0; t < nq_of(i); ++t) {
        const int64_t jn = next_on(q_of(i, t));
        if (jn < 0) continue;
        const uint32_t j = (uint32_t)jn;
        if (!done[j] && ready(j)) heap[fresh(j)].push(j);
      }
    }
    parts.push_back(std::move(part));
  }
  return parts;
}

int compile_program(pq_program_s *p) {
  const pq_op *ops = p->ops.data();
  const uint64_t nops = p->ops.size();
  uint32_t nnoise = 0, nmeas = 0;
  uint64_t i = 0;
  while (i < nops) {
    const bool meas = ops[i].code == PQ_OP_M;
    uint64_t j = i;
    while (j < nops && (ops[j].code == PQ_OP_M) == meas) ++j;
    DSegment sg{meas ? 1u : 0u, (uint32_t)p->instrs.size(), 0, 0, 0, 0, 0, 0};
    if (meas) {
      // a segment's fully determined measurements run in parallel, one wave
      // each; of several writes to one record only the last in program order
      // may happen (MeasureQubit with an existing id overwrites it,
      // CliffordState.hpp:608-611)
      for (uint64_t k = i; k < j; ++k) {
        uint32_t pad = 0;
        for (uint64_t l = k + 1; l < j; ++l)
          if (ops[l].slot == ops[k].slot) {
            pad = kMeasSuperseded;
            break;
          }
        DInstr I{PQ_OP_M, ops[k].q0, 0, ops[k].slot, nmeas, 0, ops[k].t0, pad};
        ++nmeas;
        p->instrs.push_back(I);
      }
    } else {
      // noise fusion: each NOISE op joins the op that last touched its qubit
      std::vector<DInstr> seg;
      std::vector<int64_t> last(p->n, -1);
      std::vector<char> noise_only;
      for (uint64_t k = i; k < j; ++k) {
        const pq_op &o = ops[k];
        if (o.code == PQ_OP_NOISE) {
          const uint32_t kn = nnoise++;
          p->noise.push_back(NoiseDesc{0, o.t0, o.t1, o.t2, kNoisePauli});
          const int64_t l = last[o.q0];
          if (l >= 0 && !noise_only[l]) {
            DInstr &I = seg[l];
            if (I.qa == o.q0 && I.na == kNone) { I.na = kn; continue; }
            // noise on qb fuses into unflagged CX/CZ only (a flagged op's
            // second mask is its flag)
            if ((I.code == PQ_OP_CX || I.code == PQ_OP_CZ) && I.qb == o.q0 && I.nb == kNone) {
              I.nb = kn;
              continue;
            }
          }
          last[o.q0] = (int64_t)seg.size();
          seg.push_back(DInstr{PQ_OP_NOISE, o.q0, 0, 0, kn, kNone, 0, 0});
          noise_only.push_back(1);
          continue;
        }
        last[o.q0] = (int64_t)seg.size();
        if (is_two(o.code)) last[o.q1] = (int64_t)seg.size();
        uint32_t slot = o.slot;
        if (is_random_flag(o)) {  // the flag is the X part of a noise plane
          const uint32_t kn = nnoise++;
          p->noise.push_back(NoiseDesc{0, 0, o.t0, 0, kNoiseFlag});
          slot = kSlotPlane | kn;
        }
        seg.push_back(DInstr{o.code, o.q0, is_two(o.code) ? o.q1 : 0, slot, kNone, kNone, 0, 0});
        noise_only.push_back(0);
        if (o.code >= PQ_OP_H && o.code <= PQ_OP_CZ) ++p->ngates;
      }
      // more qubits than a row tile of kWindowCutRows rows holds: windows of
      // at most as many as a kWindowRows-row tile holds, each its own
      // row-tiled segment
      uint32_t cut = 0;
      const uint32_t cap = window_qubits(p, &cut);
      if (cap && cut && count_qubits(seg, p->n) > cut) {
        for (const std::vector<uint32_t> &part : window_schedule(seg, p->n, cap)) {
          std::vector<DInstr> sub;
          sub.reserve(part.size());
          for (uint32_t k : part) sub.push_back(seg[k]);
          emit_gate_segment(p, sub);
        }
      } else {
        emit_gate_segment(p, seg);
      }
    }
    if (meas) {
      sg.count = (uint32_t)(p->instrs.size() - sg.begin);
      if (sg.count) p->segs.push_back(sg);
    }
    i = j;
  }
  p->nmeas = nmeas;
  p->nnoise = nnoise;
  return PQ_OK;
}

int compile_frame_only(pq_program_s *p) {
  for (const pq_op &o : p->ops) {
    if (o.code == PQ_OP_M) {
      ++p->nmeas;
    } else if (o.code == PQ_OP_NOISE) {
      p->noise.push_back(NoiseDesc{0, o.t0, o.t1, o.t2, kNoisePauli});
    } else {
      if (is_random_flag(o)) p->noise.push_back(NoiseDesc{0, 0, o.t0, 0, kNoiseFlag});
      if (o.code >= PQ_OP_H && o.code <= PQ_OP_CZ) ++p->ngates;
    }
  }
  p->nnoise = p->noise.size();
  return PQ_OK;
}

void pq_program_s::drop_graph() {
  if (gexec) (void)hipGraphExecDestroy(gexec);
  if (graph) (void)hipGraphDestroy(graph);
  gexec = nullptr;
  graph = nullptr;
}

pq_program_s::~pq_program_s() {
  drop_graph();
  for (hipEvent_t e : ev) (void)hipEventDestroy(e);
  if (done_ev) (void)hipEventDestroy(done_ev);
  for (void *q : {(void *)d_census, (void *)d_segs, (void *)d_segq, (void *)d_instrs,
                  (void *)d_noise, (void *)d_planes, (void *)d_side, (void *)d_sync})
    dev_free(q, alloc_stream);
}

template <class T>
static int upload(T **dst, const std::vector<T> &src, hipStream_t as) {
  if (src.empty()) return PQ_OK;
  const size_t bytes = src.size() * sizeof(T);
  HIPCHK(dev_alloc((void **)dst, bytes, as));
  if (as)
    HIPCHK(hipMemcpyAsync(*dst, src.data(), bytes, hipMemcpyHostToDevice, as));
  else
    HIPCHK(hipMemcpy(*dst, src.data(), bytes, hipMemcpyHostToDevice));
  return PQ_OK;
}

static int create_program(pq_state s, const pq_op *ops, uint64_t nops, hipStream_t as,
                          pq_program *out) {
  if (!s || !out || (!ops && nops)) return pq_set_error(PQ_ERR_ARG, "null argument");
  int rc = validate(s->n, ops, nops);
  if (rc) return rc;
  auto p = std::make_unique<pq_program_s>();
  p->alloc_stream = as;
  p->n = s->n;
  p->W = s->W;
  p->R = s->R;
  p->ops.assign(ops, ops + nops);
  // a state that keeps only its R column runs the frame form alone
  rc = s->r_only ? compile_frame_only(p.get()) : compile_program(p.get());
  if (rc) return rc;
  if ((rc = upload(&p->d_segs, p->segs, as))) return rc;
  if ((rc = upload(&p->d_instrs, p->instrs, as))) return rc;
  if ((rc = upload(&p->d_segq, p->segq, as))) return rc;
  if (p->nnoise) {
    if ((rc = upload(&p->d_noise, p->noise, as))) return rc;
    const uint64_t bytes = (uint64_t)p->nnoise * 2 * std::max<uint64_t>(p->W, 1) * 8;
    if (dev_alloc((void **)&p->d_planes, bytes, as) != hipSuccess) {
      (void)hipGetLastError();
      return pq_set_error(PQ_ERR_OOM, "noise plane allocation failed");
    }
  }
  // the measurement phases use every wave, so the default is the largest block
  p->threads = program_max_threads((uint32_t)s->n, p->tiled);
  p->blocks = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(s->W, 1u << 20));
  if (!as) HIPCHK(prepare_kernels((uint32_t)s->n));
  *out = p.release();
  return PQ_OK;
}

int create_program_on(pq_state s, const pq_op *ops, uint64_t nops, pq_program *out) {
  return create_program(s, ops, nops, s->stream, out);
}

extern "C" int pq_program_create(pq_state s, const pq_op *ops, uint64_t nops,
                                 pq_program *out) {
  return create_program(s, ops, nops, nullptr, out);
}

extern "C" int pq_program_destroy(pq_program p) {
  delete p;
  return PQ_OK;
}

extern "C" int pq_program_info(pq_program p, uint64_t *gates, uint64_t *measurements,
                               uint64_t *noise_ops, uint64_t *instructions) {
  if (!p) return pq_set_error(PQ_ERR_ARG, "null program");
  if (gates) *gates = p->ngates;
  if (measurements) *measurements = p->nmeas;
  if (noise_ops) *noise_ops = p->nnoise;
  if (instructions) *instructions = p->instrs.size();
  return PQ_OK;
}

extern "C" int pq_program_layout(pq_program p, uint64_t *segments, uint64_t *tiled_segments) {
  if (!p) return pq_set_error(PQ_ERR_ARG, "null program");
  if (segments) *segments = p->segs.size();
  if (tiled_segments) {
    uint64_t t = 0;
    for (const DSegment &sg : p->segs) t += (sg.kind == 0 && sg.nq > 0);
    *tiled_segments = t;
  }
  return PQ_OK;
}

extern "C" int pq_set_window_qubits(uint32_t qubits) {
  g_window_qubits.store(qubits);
  return PQ_OK;
}

extern "C" int pq_program_set_launch(pq_program p, uint32_t threads, uint32_t blocks) {
  if (!p) return pq_set_error(PQ_ERR_ARG, "null program");
  if (threads) {
    if (threads % 64 || threads > program_max_threads((uint32_t)p->n, p->tiled))
      return pq_set_error(PQ_ERR_ARG, ("threads must be a multiple of 64, at most " +
                                       std::to_string(program_max_threads((uint32_t)p->n, p->tiled)) +
                                       " for this state size").c_str());
    p->threads = threads;
  }
  if (blocks) p->blocks = blocks;
  p->drop_graph();
  return PQ_OK;
}

// Record slots: M ops write slot <= current count (== count appends), flagged
// ops read slot < current count, checked in program order.  Sets
// p->prep_end to the record count after the run; allocates nothing.
int check_records(pq_state s, pq_program p) {
  // the check depends only on the ops and the state's record count, so a
  // program rerun on a state with the same count reuses it
  if (!p->prep_valid || p->prep_start != s->nrec) {
    uint64_t cur = s->nrec;
    for (uint64_t i = 0; i < p->ops.size(); ++i) {
      const pq_op &o = p->ops[i];
      if (o.code == PQ_OP_M) {
        if (o.slot > cur)
          return pq_set_error(PQ_ERR_RANGE, ("op " + std::to_string(i) + ": record id out of range").c_str());
        if (o.slot == cur) ++cur;
      } else if (is_flag(o.code) && !is_random_flag(o) && o.slot >= cur) {
        return pq_set_error(PQ_ERR_RANGE, ("op " + std::to_string(i) + ": flag record out of range").c_str());
      }
    }
    if (!p->prep_valid) {
      for (const pq_op &o : p->ops)
        if (o.code == PQ_OP_M) p->m_slots.push_back(o.slot);
      std::sort(p->m_slots.begin(), p->m_slots.end());
      p->m_slots.erase(std::unique(p->m_slots.begin(), p->m_slots.end()), p->m_slots.end());
    }
    p->prep_valid = true;
    p->prep_start = s->nrec;
    p->prep_end = cur;
  }
  return PQ_OK;
}

int prepare_records(pq_state s, pq_program p) {
  int rc = check_records(s, p);
  if (rc) return rc;
  rc = ensure_records(s, p->prep_end);
  if (rc) return rc;
  for (uint32_t slot : p->m_slots) s->rec_valid[slot] = 0;
  return PQ_OK;
}

static int run_plain(pq_state s, pq_program p) {
  for (const pq_op &o : p->ops) {
    int rc = PQ_OK;
    switch (o.code) {
      case PQ_OP_H: case PQ_OP_S: case PQ_OP_X: case PQ_OP_Y: case PQ_OP_Z:
      case PQ_OP_CX: case PQ_OP_CZ:
        rc = plain_gate(s, o.code, o.q0, o.q1);
        break;
      case PQ_OP_NOISE:
        rc = plain_noise(s, o.q0, o.t0, o.t1, o.t2);
        break;
      case PQ_OP_M:
        rc = plain_measure(s, o.q0, o.t0, o.slot, nullptr);
        break;
      default:
        rc = is_random_flag(o) ? plain_random_flag(s, o.code, o.q0, o.q1, o.t0)
                               : plain_flag_plane(s, o.code, o.q0, o.q1, record_ptr(s, o.slot));
        break;
    }
    if (rc) return rc;
  }
  return PQ_OK;
}

ProgramArgs program_args(pq_state s, pq_program p, uint64_t lds_budget) {
  ProgramArgs a{};
  a.tab = s->tab;
  a.recs = s->recs;
  a.planes = p->d_planes;
  a.wide = p->tiled ? 1u : 0u;
  a.nk = (uint32_t)p->nnoise;
  a.segs = p->d_segs;
  a.instrs = p->d_instrs;
  a.coin_planes = nullptr;
  a.nseg = (uint32_t)p->segs.size();
  a.n = (uint32_t)s->n;
  a.R = (uint32_t)s->R;
  a.C = (uint32_t)s->C;
  a.W = s->W;
  a.word0 = 0;
  a.nwords = s->W;
  a.census = p->census_on ? p->d_census : nullptr;
  // waves that run fully-determined measurements in parallel, each with a
  // 2n-word LDS slab, within the kernel's dynamic-LDS budget
  const uint64_t waves = p->threads / 64;
  const uint64_t budget = lds_budget;
  const uint64_t base = program_lds_bytes(a.n, 0);
  const uint64_t slab = 16ull * a.n;
  uint64_t ndw = (budget > base && slab) ? (budget - base) / slab : 0;
  a.det_waves = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(waves, ndw));
  a.segq = p->d_segq;
  // row-tiled gate segments use whatever LDS the budget allows
  const uint64_t need = program_lds_bytes(a.n, a.det_waves);
  a.lds_words = (p->tiled ? std::max<uint64_t>(need, budget) : need) / 8;
  return a;
}

// Workgroups per 64-shot word group.  A state of 512+ qubits with at most
// half as many word groups as the device has compute units splits each
// group's rows between two workgroups.  The slices meet at a device barrier
// after every gate segment and measurement step; more than two did not pay
// (n = 1024, 64 groups: 1.69 s per step with one workgroup per group, 1.61 s
// with two, 1.96 s with four; the random measurements' scattered row
// updates are bound by the memory system's request rate, not by the compute
// units in use).  Runs by a server (stream-ordered allocation) and graph
// captures use one workgroup.
constexpr uint32_t kSliceMinQubits = 512;
constexpr uint32_t kAutoSlices = 2;

static int choose_slices(pq_state s, pq_program p, const ProgramArgs &a, bool allow,
                         uint32_t *out) {
  *out = 1;
  if (!allow || p->alloc_stream || p->slices_req == 1) return PQ_OK;
  uint32_t S = p->slices_req;
  if (!S) {
    if (a.n < kSliceMinQubits) return PQ_OK;
    int dev = 0, cus = 0;
    HIPCHK(hipGetDevice(&dev));
    HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    S = a.nwords * kAutoSlices <= (uint64_t)cus ? kAutoSlices : 1;
    while (S > 1 && !program_slices_fit(a.n, a.nwords, S, p->threads, a.lds_words * 8, a.wide))
      S /= 2;
    (void)hipGetLastError();
    if (S <= 1) return PQ_OK;
  } else if (!program_slices_fit(a.n, a.nwords, S, p->threads, a.lds_words * 8, a.wide)) {
    (void)hipGetLastError();
    return pq_set_error(PQ_ERR_ARG, ("slices " + std::to_string(S) + " x " +
                                     std::to_string(a.nwords) +
                                     " word groups do not fit the device at once")
                                        .c_str());
  }
  if (p->sync_words < a.nwords + 1) {
    if (p->d_sync) {
      HIPCHK(hipStreamSynchronize(s->stream));
      dev_free(p->d_sync, nullptr);
      p->d_sync = nullptr;
      p->sync_words = 0;
    }
    HIPCHK(dev_alloc((void **)&p->d_sync, (a.nwords + 1) * sizeof(unsigned int), nullptr));
    HIPCHK(hipMemset(p->d_sync, 0, (a.nwords + 1) * sizeof(unsigned int)));
    p->sync_words = a.nwords + 1;
  }
  *out = S;
  return PQ_OK;
}

static int enqueue_persistent(pq_state s, pq_program p, bool timed, bool allow_slices,
                              bool fresh = false) {
  ProgramArgs a = program_args(s, p, kProgramLdsBytes);
  a.fresh = fresh ? 1u : 0u;
  uint32_t S = 1;
  int rc = choose_slices(s, p, a, allow_slices, &S);
  if (rc) return rc;
  p->slices_used = S;
  if (S > 1) {
    a.slices = S;
    a.sync = p->d_sync;
    a.sync_err = p->d_sync + a.nwords;
    a.sync_spin = p->slice_wait;
    // barrier counters start at zero; the give-up flag keeps its value
    // until read
    HIPCHK(launch_zero_u32(s->stream, a.sync, (uint32_t)a.nwords));
  }
  if (p->nnoise)
    HIPCHK(launch_program_noise(s->stream, p->d_planes, p->d_noise,
                                (uint32_t)p->nnoise, s->W, s->params_dev));
  if (timed) {
    if (p->ev_used + 2 > p->ev.size()) {
      hipEvent_t a, b;
      HIPCHK(hipEventCreate(&a));
      HIPCHK(hipEventCreate(&b));
      p->ev.push_back(a);
      p->ev.push_back(b);
    }
    HIPCHK(hipEventRecord(p->ev[p->ev_used], s->stream));
  }
  HIPCHK(launch_program(s->stream, a, p->threads, p->blocks, s->params_dev));
  if (timed) {
    HIPCHK(hipEventRecord(p->ev[p->ev_used + 1], s->stream));
    p->ev_used += 2;
  }
  return PQ_OK;
}

static int program_run(pq_state s, pq_program p, int form);

// A program's device buffers (noise planes, the frame form's side buffer)
// serve every run, so a run on a stream other than the last run's waits for
// that run's device work first.
extern "C" int pq_program_run(pq_state s, pq_program p, int form) {
  if (!s || !p) return pq_set_error(PQ_ERR_ARG, "null argument");
  const bool device_run = s->tab && form != PQ_RUN_PLAIN;
  if (device_run && p->done_ev && p->last_stream != s->stream)
    HIPCHK(hipStreamWaitEvent(s->stream, p->done_ev, 0));
  const int rc = program_run(s, p, form);
  if (rc == PQ_OK && device_run) {
    if (!p->done_ev) HIPCHK(hipEventCreateWithFlags(&p->done_ev, hipEventDisableTiming));
    HIPCHK(hipEventRecord(p->done_ev, s->stream));
    p->last_stream = s->stream;
  }
  return rc;
}

static int program_run(pq_state s, pq_program p, int form) {
  if (!s || !p) return pq_set_error(PQ_ERR_ARG, "null argument");
  if (p->n != s->n || p->W != s->W)
    return pq_set_error(PQ_ERR_ARG, "program was compiled for a state of another shape");
  if (s->r_only && form != PQ_RUN_FRAME)
    return pq_set_error(PQ_ERR_UNSUPPORTED, "this state keeps only its phase column: frame form only");
  if (form == PQ_RUN_FRAME && s->tab) {
    // eligibility first, so a refused run leaves the state untouched
    if (!s->xz_identity)
      return pq_set_error(PQ_ERR_UNSUPPORTED,
                          "frame form runs from the |0...0> tableau: reset the state first");
    if (!p->frame) {
      int rc0 = build_frame_plan(p, s->W);
      if (rc0) return rc0;
    }
  }
  int rc = prepare_records(s, p);
  if (rc) return rc;
  if (!s->tab) {
    s->meas_seq += p->nmeas;
    s->noise_seq += p->nnoise;
    return PQ_OK;
  }
  if (form == PQ_RUN_FRAME && s->n > measure_max_qubits())
    return pq_set_error(PQ_ERR_UNSUPPORTED, "frame form supports at most 4095 qubits");
  if (form == PQ_RUN_FRAME) return run_frame(s, p);
  // X/Z still the identity after a reset, and the program opens with a
  // row-tiled gate segment over every qubit: the persistent kernel writes
  // that segment's tiles itself, so the identity fill is not needed
  const bool fresh = form == PQ_RUN_PERSISTENT && s->xz_pending && !s->xz_plan &&
                     !p->segs.empty() && p->segs[0].kind == 0 && p->segs[0].nq == s->n &&
                     program_tile_rows(program_args(s, p, kProgramLdsBytes).lds_words,
                                       (uint32_t)s->n) > 0;
  if (fresh)
    s->xz_pending = false;  // written by the run
  else if ((rc = materialize_tableau(s)))
    return rc;
  s->xz_identity = false;
  if (form == PQ_RUN_PLAIN) return run_plain(s, p);
  if (s->n > measure_max_qubits())
    return pq_set_error(PQ_ERR_UNSUPPORTED, "persistent form supports at most 4095 qubits");
  HIPCHK(launch_set_params(s->stream, s->params_dev, (uint64_t)s->seed, s->meas_seq,
                           s->noise_seq));
  s->meas_seq += p->nmeas;
  s->noise_seq += p->nnoise;
  if (form == PQ_RUN_PERSISTENT) return enqueue_persistent(s, p, p->timing_on, true, fresh);
  if (form != PQ_RUN_GRAPH) return pq_set_error(PQ_ERR_ARG, "unknown run form");
  if (p->gexec && (p->graph_state != s || p->graph_arena_gen != s->arena_gen)) {
    (void)hipGraphExecDestroy(p->gexec);
    (void)hipGraphDestroy(p->graph);
    p->gexec = nullptr;
    p->graph = nullptr;
  }
  if (!p->gexec) {
    HIPCHK(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
    rc = enqueue_persistent(s, p, false, false);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(s->stream, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    HIPCHK(e);
    p->graph = g;
    HIPCHK(hipGraphInstantiate(&p->gexec, p->graph, nullptr, nullptr, 0));
    p->graph_state = s;
    p->graph_arena_gen = s->arena_gen;
  }
  HIPCHK(hipGraphLaunch(p->gexec, s->stream));
  return PQ_OK;
}

extern "C" int pq_program_set_slices(pq_program p, uint32_t slices) {
  if (!p) return pq_set_error(PQ_ERR_ARG, "null program");
  if (slices > 1024) return pq_set_error(PQ_ERR_ARG, "slices must be at most 1024");
  p->slices_req = slices;
  return PQ_OK;
}

extern "C" int pq_program_set_slice_wait(pq_program p, uint32_t polls) {
  if (!p) return pq_set_error(PQ_ERR_ARG, "null program");
  p->slice_wait = polls;
  return PQ_OK;
}

extern "C" int pq_program_slice_status(pq_state s, pq_program p, uint32_t *slices,
                                       uint32_t *failed) {
  if (!s || !p) return pq_set_error(PQ_ERR_ARG, "null argument");
  if (slices) *slices = p->slices_used;
  if (failed) {
    *failed = 0;
    if (p->d_sync) {
      HIPCHK(hipStreamSynchronize(s->stream));
      unsigned int h = 0;
      HIPCHK(hipMemcpy(&h, p->d_sync + (p->sync_words - 1), sizeof(h), hipMemcpyDeviceToHost));
      *failed = h;
    }
  }
  return PQ_OK;
}

extern "C" int pq_program_set_census(pq_program p, int enable) {
  if (!p) return pq_set_error(PQ_ERR_ARG, "null program");
  if (enable && !p->d_census)
    HIPCHK(dev_alloc((void **)&p->d_census, 8 * sizeof(unsigned long long), p->alloc_stream));
  if (enable) HIPCHK(hipMemset(p->d_census, 0, 8 * sizeof(unsigned long long)));
  if ((bool)enable != p->census_on) p->drop_graph();
  p->census_on = enable != 0;
  return PQ_OK;
}

extern "C" int pq_program_read_census(pq_state s, pq_program p, uint64_t out[8]) {
  if (!s || !p || !out) return pq_set_error(PQ_ERR_ARG, "null argument");
  for (int i = 0; i < 8; ++i) out[i] = 0;
  if (!p->d_census) return PQ_OK;
  HIPCHK(hipStreamSynchronize(s->stream));
  unsigned long long h[8];
  HIPCHK(hipMemcpy(h, p->d_census, sizeof(h), hipMemcpyDeviceToHost));
  for (int i = 0; i < 8; ++i) out[i] = h[i];
  return PQ_OK;
}

extern "C" int pq_program_set_timing(pq_program p, int enable) {
  if (!p) return pq_set_error(PQ_ERR_ARG, "null program");
  p->timing_on = enable != 0;
  p->ev_used = 0;
  return PQ_OK;
}

extern "C" int pq_program_kernel_ms(pq_program p, float *total_ms, uint64_t *launches) {
  if (!p || !total_ms) return pq_set_error(PQ_ERR_ARG, "null argument");
  float tot = 0;
  for (size_t i = 0; i + 1 < p->ev_used; i += 2) {
    float ms = 0;
    HIPCHK(hipEventSynchronize(p->ev[i + 1]));
    HIPCHK(hipEventElapsedTime(&ms, p->ev[i], p->ev[i + 1]));
    tot += ms;
  }
  *total_ms = tot;
  if (launches) *launches = p->ev_used / 2;
  p->ev_used = 0;
  return PQ_OK;
}

extern "C" int pq_run_program(pq_state s, const pq_op *ops, uint64_t nops, int form) {
  pq_program p = nullptr;
  int rc = pq_program_create(s, ops, nops, &p);
  if (rc) return rc;
  rc = pq_program_run(s, p, form);
  if (rc == PQ_OK && form != PQ_RUN_PLAIN) {
    // the program's device buffers are freed below: wait for the kernels
    if (hipStreamSynchronize(s->stream) != hipSuccess)
      rc = pq_set_error(PQ_ERR_HIP, "program execution failed");
  }
  pq_program_destroy(p);
  return rc;
}
