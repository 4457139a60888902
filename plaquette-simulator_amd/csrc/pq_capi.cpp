// Host implementation of the C ABI declared in include/plaquette_amd.h.
//
// A state owns: the bit-plane tableau (W word groups of (2n+1) x (2n+1)
// uint64 words), a record arena ([rec][2][W] words: results, determined),
// one HIP stream (default priority) and small device scratch buffers.
// Every entry point validates its arguments on the host before any launch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "plaquette_amd.h"
#include "pq_capi_internal.h"
#include "pq_frame.h"
#include "pq_common.h"
#include "pq_kernels.h"

using namespace pq;

namespace {
thread_local std::string g_err;

int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}
}  // namespace

int pq_set_error(int code, const char *msg) { return fail(code, msg); }

#define HIPCHK(expr)                                                         \
  do {                                                                       \
    hipError_t e_ = (expr);                                                  \
    if (e_ != hipSuccess)                                                    \
      return fail(PQ_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define ARGCHK(cond, msg) \
  do {                    \
    if (!(cond)) return fail(PQ_ERR_ARG, msg); \
  } while (0)

static constexpr int64_t kDefaultSeed = 213434232223LL;  // CliffordState.hpp:135

// ---------------------------------------------------------------------------
// library
// ---------------------------------------------------------------------------
extern "C" const char *pq_last_error(void) { return g_err.c_str(); }

extern "C" const char *pq_backend_name(void) {
  static std::string name;
  static std::once_flag once;
  std::call_once(once, [] {
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipGetDeviceProperties(&prop, dev) == hipSuccess) {
      std::string arch = prop.gcnArchName;
      arch = arch.substr(0, arch.find(':'));
      name = "HIP(" + arch + ")";
    } else {
      name = "HIP(unavailable)";
    }
  });
  return name.c_str();
}

extern "C" int pq_device_count(int *count) {
  ARGCHK(count, "count is null");
  int c = 0;
  hipError_t e = hipGetDeviceCount(&c);
  if (e != hipSuccess) c = 0;
  *count = c;
  return PQ_OK;
}

extern "C" uint32_t pq_coin_threshold(float bias) {
  const double t = (double)bias * 16777216.0;
  if (!(t > 0.0)) return 0;
  if (t >= 16777216.0) return 16777216u;
  return (uint32_t)std::ceil(t);
}

extern "C" uint64_t pq_draw(uint64_t seed, uint64_t stream, uint64_t idx) {
  return draw(seed, stream, idx);
}

// ---------------------------------------------------------------------------
// state
// ---------------------------------------------------------------------------
pq_state_s::~pq_state_s() {
  if (aux) (void)hipStreamSynchronize(aux);
  if (stream && own_stream) (void)hipStreamSynchronize(stream);
  for (void *p : {(void *)tab, (void *)recs, (void *)scratch, params_dev, csum, (void *)stage})
    dev_free(p, alloc_stream);
  for (auto &e : ev)
    if (e) (void)hipEventDestroy(e);
  if (order_ev) (void)hipEventDestroy(order_ev);
  for (hipEvent_t e : aux_ev) (void)hipEventDestroy(e);
  if (aux) (void)hipStreamDestroy(aux);
  if (stream && own_stream) (void)hipStreamDestroy(stream);
}

// Entry points that work on the full tableau refuse states that keep only
// their phase column (sampling states: frame-form programs and records), and
// fill X/Z a reset left pending.
#define FULL_ONLY(st)                                                                   \
  do {                                                                                  \
    if ((st) && (st)->r_only)                                                           \
      return fail(PQ_ERR_UNSUPPORTED,                                                   \
                  "a sampling state keeps only its phase column: frame-form programs, " \
                  "records and reset only");                                            \
    if (st) {                                                                           \
      const int mrc_ = materialize_tableau(st);                                        \
      if (mrc_) return mrc_;                                                            \
    }                                                                                   \
  } while (0)

static int check_device() {
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess || c < 1)
    return fail(PQ_ERR_NODEV, "no HIP device available");
  return PQ_OK;
}

static int alloc_state(uint64_t n, uint64_t B, int64_t seed, hipStream_t borrowed,
                       pq_state *out, bool r_only = false) {
  ARGCHK(out, "output handle is null");
  int rc = check_device();
  if (rc) return rc;
  if (n > 0xFFFFFFF0ull / 2) return fail(PQ_ERR_UNSUPPORTED, "too many qubits");
  auto s = std::make_unique<pq_state_s>();
  s->n = n;
  s->B = B;
  s->R = 2 * n + 1;
  s->C = 2 * n + 1;
  s->W = (B + 63) / 64;
  s->seed = seed >= 0 ? seed : kDefaultSeed;
  HIPCHK(hipGetDevice(&s->device));
  if (borrowed) {
    s->stream = borrowed;
    s->own_stream = false;
    s->alloc_stream = borrowed;
  } else {
    HIPCHK(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
  }
  s->r_only = r_only;
  if (n > 0 && B > 0) {
    const uint64_t words = s->W * (r_only ? 1 : s->C) * s->R;
    if (dev_alloc((void **)&s->tab, words * 8, s->alloc_stream) != hipSuccess) {
      (void)hipGetLastError();
      return fail(PQ_ERR_OOM, "tableau allocation failed (" +
                                  std::to_string(words * 8) + " bytes)");
    }
  }
  HIPCHK(dev_alloc((void **)&s->scratch, std::max<uint64_t>(1, 3 * s->W) * 8, s->alloc_stream));
  HIPCHK(dev_alloc(&s->params_dev, 64, s->alloc_stream));
  HIPCHK(dev_alloc(&s->csum, 8, s->alloc_stream));
  if (!borrowed)
    HIPCHK(prepare_kernels((uint32_t)std::min<uint64_t>(n, measure_max_qubits())));
  *out = s.release();
  return PQ_OK;
}

Geom geom_of(const pq_state_s *s) {
  return Geom{(uint32_t)s->n, (uint32_t)s->R, (uint32_t)s->C, s->W, s->B};
}

int create_state_on(uint64_t n, uint64_t B, int64_t seed, hipStream_t stream,
                    pq_state *out, bool r_only) {
  int rc = alloc_state(n, B, seed, stream, out, r_only);
  if (rc) return rc;
  pq_state s = *out;
  if (s->tab) {
    if (r_only)
      HIPCHK(launch_fill_r_identity(s->stream, s->tab, s->W, (uint32_t)s->R));
    else
      HIPCHK(launch_fill_identity(s->stream, s->tab, geom_of(s)));
  }
  s->xz_identity = true;
  return PQ_OK;
}

extern "C" int pq_create(uint64_t n, uint64_t B, int64_t seed, int init_zero,
                         pq_state *out) {
  int rc = alloc_state(n, B, seed, nullptr, out);
  if (rc) return rc;
  pq_state s = *out;
  if (s->tab) {
    if (init_zero) {
      HIPCHK(launch_fill_identity(s->stream, s->tab, geom_of(s)));
      s->xz_identity = true;
    } else {
      HIPCHK(hipMemsetAsync(s->tab, 0, s->W * s->C * s->R * 8, s->stream));
    }
  }
  return PQ_OK;
}

extern "C" int pq_create_sampling(uint64_t n, uint64_t B, int64_t seed, pq_state *out) {
  int rc = alloc_state(n, B, seed, nullptr, out, true);
  if (rc) return rc;
  pq_state s = *out;
  if (s->tab) HIPCHK(launch_fill_r_identity(s->stream, s->tab, s->W, (uint32_t)s->R));
  s->xz_identity = true;
  return PQ_OK;
}

extern "C" int pq_create_from_host(const int32_t *x, const int32_t *z,
                                   const int32_t *r, uint64_t n, uint64_t B,
                                   int64_t seed, pq_state *out) {
  int rc = pq_create(n, B, seed, 0, out);
  if (rc) return rc;
  rc = pq_host_to_device(*out, x, z, r);
  if (rc) {
    pq_destroy(*out);
    *out = nullptr;
  }
  return rc;
}

extern "C" int pq_destroy(pq_state s) {
  delete s;
  return PQ_OK;
}

extern "C" int pq_info(pq_state s, uint64_t *n, uint64_t *B, int64_t *seed,
                       uint64_t *nrec) {
  ARGCHK(s, "null state");
  if (n) *n = s->n;
  if (B) *B = s->B;
  if (seed) *seed = s->seed;
  if (nrec) *nrec = s->nrec;
  return PQ_OK;
}

int materialize_tableau(pq_state s) {
  if (!s || !s->tab) return PQ_OK;
  if (s->xz_pending) {
    HIPCHK(launch_fill_identity(s->stream, s->tab, geom_of(s)));
    s->xz_pending = false;
  } else if (s->xz_plan) {
    const FramePlan &F = *s->xz_plan;
    HIPCHK(launch_frame_xz(s->stream, s->tab, geom_of(s), F.d_X, F.d_Z, F.nw));
    s->xz_plan.reset();
  }
  return PQ_OK;
}

// |0...0> again (the reference's ResetCliffordState, with its intended
// extent).  Only the R column is written now; X/Z are filled when something
// needs them (materialize_tableau), so a frame-form run, which writes X/Z
// itself, never pays for the fill.  A state whose raw pointer was handed out
// is filled at once.
extern "C" int pq_reset(pq_state s) {
  ARGCHK(s, "null state");
  if (s->tab && s->r_only) {
    HIPCHK(launch_fill_r_identity(s->stream, s->tab, s->W, (uint32_t)s->R));
  } else if (s->tab && !s->exposed) {
    HIPCHK(launch_fill_r_column(s->stream, s->tab, geom_of(s)));
    s->xz_pending = true;
  } else if (s->tab) {
    HIPCHK(launch_fill_identity(s->stream, s->tab, geom_of(s)));
    s->xz_pending = false;
  }
  s->xz_plan.reset();
  s->xz_identity = true;
  return PQ_OK;
}

extern "C" int pq_clone(pq_state src, pq_state *out) {
  FULL_ONLY(src);
  ARGCHK(src, "null state");
  int rc = pq_create(src->n, src->B, src->seed, 0, out);
  if (rc) return rc;
  rc = pq_copy_from(*out, src);
  if (rc) {
    pq_destroy(*out);
    *out = nullptr;
  }
  return rc;
}

extern "C" int pq_copy_from(pq_state dst, pq_state src) {
  FULL_ONLY(dst);
  FULL_ONLY(src);
  ARGCHK(dst && src, "null state");
  ARGCHK(dst->n == src->n && dst->B == src->B, "DeviceToDevice needs equal n and B");
  if (!dst->tab || dst == src) return PQ_OK;
  dst->xz_identity = src->xz_identity;
  // Kokkos::deep_copy has finished before either state changes again
  // (CliffordState.hpp:268-274).  Here the copy runs on the source's stream
  // after the destination's queued work, and the destination's stream waits
  // for it: later work on either state is ordered after the copy, and the
  // host does not block.
  for (pq_state st : {dst, src})
    if (!st->order_ev) HIPCHK(hipEventCreateWithFlags(&st->order_ev, hipEventDisableTiming));
  HIPCHK(hipEventRecord(dst->order_ev, dst->stream));
  HIPCHK(hipStreamWaitEvent(src->stream, dst->order_ev, 0));
  HIPCHK(hipMemcpyAsync(dst->tab, src->tab, dst->W * dst->C * dst->R * 8,
                        hipMemcpyDeviceToDevice, src->stream));
  HIPCHK(hipEventRecord(src->order_ev, src->stream));
  HIPCHK(hipStreamWaitEvent(dst->stream, src->order_ev, 0));
  return PQ_OK;
}

// ---------------------------------------------------------------------------
// records
// ---------------------------------------------------------------------------
int ensure_records(pq_state s, uint64_t count) {
  if (count <= s->rec_cap) {
    if (count > s->nrec) {
      s->nrec = count;
      s->rec_host.resize(count);
      s->rec_valid.resize(count, 0);
    }
    return PQ_OK;
  }
  // server-owned states size the arena exactly (they know their program)
  uint64_t cap = s->alloc_stream ? count : std::max<uint64_t>({count, 2 * s->rec_cap, 64});
  uint64_t *nr = nullptr;
  const uint64_t bytes = cap * 2 * std::max<uint64_t>(s->W, 1) * 8;
  if (dev_alloc((void **)&nr, bytes, s->alloc_stream) != hipSuccess) {
    (void)hipGetLastError();
    return fail(PQ_ERR_OOM, "record arena allocation failed");
  }
  HIPCHK(hipMemsetAsync(nr, 0, bytes, s->stream));
  if (s->recs) {
    HIPCHK(hipMemcpyAsync(nr, s->recs, s->rec_cap * 2 * s->W * 8,
                          hipMemcpyDeviceToDevice, s->stream));
    if (!s->alloc_stream) HIPCHK(hipStreamSynchronize(s->stream));
    dev_free(s->recs, s->alloc_stream);
  }
  s->recs = nr;
  s->rec_cap = cap;
  s->arena_gen++;
  s->nrec = count;
  s->rec_host.resize(count);
  s->rec_valid.resize(count, 0);
  return PQ_OK;
}

uint64_t *record_ptr(pq_state s, uint64_t id) { return s->recs + id * 2 * s->W; }

static int fetch_record(pq_state s, uint64_t id) {
  if (s->rec_valid[id]) return PQ_OK;
  auto &h = s->rec_host[id];
  h.resize(2 * s->W);
  HIPCHK(hipStreamSynchronize(s->stream));
  if (s->W)
    HIPCHK(hipMemcpy(h.data(), record_ptr(s, id), 2 * s->W * 8, hipMemcpyDeviceToHost));
  s->rec_valid[id] = 1;
  return PQ_OK;
}

void invalidate_records(pq_state s) {
  std::fill(s->rec_valid.begin(), s->rec_valid.end(), 0);
}

extern "C" int pq_get_measurement(pq_state s, uint64_t rec, uint64_t b,
                                  int32_t *result, int32_t *determined) {
  ARGCHK(s, "null state");
  if (rec >= s->nrec || b >= s->B)
    return fail(PQ_ERR_RANGE, "Measurement id or batch index is out of range.");
  int rc = fetch_record(s, rec);
  if (rc) return rc;
  const auto &h = s->rec_host[rec];
  if (result) *result = (int32_t)((h[b >> 6] >> (b & 63)) & 1);
  if (determined) *determined = (int32_t)((h[s->W + (b >> 6)] >> (b & 63)) & 1);
  return PQ_OK;
}

extern "C" int pq_get_record(pq_state s, uint64_t rec, int32_t *results,
                             int32_t *determined) {
  ARGCHK(s, "null state");
  if (rec >= s->nrec) return fail(PQ_ERR_RANGE, "Measurement id is out of range.");
  int rc = fetch_record(s, rec);
  if (rc) return rc;
  const auto &h = s->rec_host[rec];
  for (uint64_t b = 0; b < s->B; ++b) {
    if (results) results[b] = (int32_t)((h[b >> 6] >> (b & 63)) & 1);
    if (determined) determined[b] = (int32_t)((h[s->W + (b >> 6)] >> (b & 63)) & 1);
  }
  return PQ_OK;
}

// ---------------------------------------------------------------------------
// plain form: one launch per op
// ---------------------------------------------------------------------------
static bool is_gate(uint32_t c) { return c >= PQ_OP_H && c <= PQ_OP_CZ; }
static bool is_two_qubit(uint32_t c) { return c == PQ_OP_CX || c == PQ_OP_CZ; }
static bool is_flag_op(uint32_t c) {
  return c == PQ_OP_H_FLAG || c == PQ_OP_S_FLAG || c == PQ_OP_X_FLAG ||
         c == PQ_OP_Y_FLAG || c == PQ_OP_Z_FLAG;
}
static bool is_flag2_op(uint32_t c) { return c == PQ_OP_CX_FLAG || c == PQ_OP_CZ_FLAG; }
// The flagged op code of gate `c` (H..CZ).
static uint32_t flagged_code(uint32_t c) {
  switch (c) {
    case PQ_OP_H: return PQ_OP_H_FLAG;
    case PQ_OP_S: return PQ_OP_S_FLAG;
    case PQ_OP_X: return PQ_OP_X_FLAG;
    case PQ_OP_Y: return PQ_OP_Y_FLAG;
    case PQ_OP_Z: return PQ_OP_Z_FLAG;
    case PQ_OP_CX: return PQ_OP_CX_FLAG;
    default: return PQ_OP_CZ_FLAG;
  }
}

int plain_gate(pq_state s, uint32_t code, uint64_t q0, uint64_t q1) {
  ARGCHK(s, "null state");
  ARGCHK(is_gate(code), "not a gate op code");
  if (q0 >= s->n) return fail(PQ_ERR_ARG, "qubit index out of range");
  if (is_two_qubit(code) && (q1 >= s->n || q1 == q0))
    return fail(PQ_ERR_ARG, "two-qubit gate needs distinct qubits in range");
  if (!s->tab) return PQ_OK;
  if (code == PQ_OP_H || code == PQ_OP_S || code == PQ_OP_CX || code == PQ_OP_CZ)
    s->xz_identity = false;
  HIPCHK(launch_gate(s->stream, s->tab, geom_of(s), code, (uint32_t)q0,
                     (uint32_t)q1, nullptr, nullptr));
  return PQ_OK;
}

extern "C" int pq_gate(pq_state s, uint32_t code, uint64_t q0, uint64_t q1) {
  FULL_ONLY(s);
  return plain_gate(s, code, q0, q1);
}

// Flagged op with a device flag plane (W words).
static int zero_scratch_tail(pq_state s);

int plain_flag_plane(pq_state s, uint32_t code, uint64_t q, uint64_t q1,
                     const uint64_t *plane) {
  if (!s->tab) return PQ_OK;
  int rc = zero_scratch_tail(s);
  if (rc) return rc;
  const Geom g = geom_of(s);
  if (code == PQ_OP_H_FLAG || code == PQ_OP_S_FLAG || code == PQ_OP_CX_FLAG ||
      code == PQ_OP_CZ_FLAG)
    s->xz_identity = false;
  switch (code) {
    case PQ_OP_H_FLAG:
    case PQ_OP_S_FLAG:
      HIPCHK(launch_gate(s->stream, s->tab, g, code, (uint32_t)q, 0, plane, nullptr));
      break;
    case PQ_OP_CX_FLAG:
    case PQ_OP_CZ_FLAG:
      HIPCHK(launch_gate(s->stream, s->tab, g, code, (uint32_t)q, (uint32_t)q1, plane,
                         nullptr));
      break;
    case PQ_OP_X_FLAG:
      HIPCHK(launch_gate(s->stream, s->tab, g, PQ_OP_NOISE, (uint32_t)q, 0, plane,
                         s->scratch + 2 * s->W));
      break;
    case PQ_OP_Z_FLAG:
      HIPCHK(launch_gate(s->stream, s->tab, g, PQ_OP_NOISE, (uint32_t)q, 0,
                         s->scratch + 2 * s->W, plane));
      break;
    case PQ_OP_Y_FLAG:
      HIPCHK(launch_gate(s->stream, s->tab, g, PQ_OP_NOISE, (uint32_t)q, 0, plane, plane));
      break;
    default:
      return fail(PQ_ERR_ARG, "not a flagged op code");
  }
  return PQ_OK;
}

static int zero_scratch_tail(pq_state s) {
  // scratch[2W .. 3W) is kept all-zero for the "no Z part" / "no X part" mask
  if (!s->scratch_zeroed) {
    HIPCHK(hipMemsetAsync(s->scratch + 2 * s->W, 0, std::max<uint64_t>(s->W, 1) * 8, s->stream));
    s->scratch_zeroed = true;
  }
  return PQ_OK;
}

// Copy a host int32[B] vector to the device and turn it into a shot mask
// plane (nonzero -> bit set).  Synchronous: the staging buffer is reused.
static int stage_host_ints(pq_state s, const int32_t *v, uint64_t *plane) {
  if (!s->stage)
    HIPCHK(dev_alloc((void **)&s->stage, std::max<uint64_t>(s->B, 1) * 4, s->alloc_stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  HIPCHK(hipMemcpy(s->stage, v, s->B * 4, hipMemcpyHostToDevice));
  HIPCHK(launch_flags_to_plane(s->stream, s->stage, s->B, plane));
  return PQ_OK;
}

extern "C" int pq_gate_flag_record(pq_state s, uint32_t code, uint64_t q, uint64_t rec) {
  FULL_ONLY(s);
  ARGCHK(s, "null state");
  ARGCHK(is_flag_op(code), "not a flagged op code");
  if (q >= s->n) return fail(PQ_ERR_ARG, "qubit index out of range");
  if (rec >= s->nrec) return fail(PQ_ERR_RANGE, "Measurement id is out of range.");
  int rc = zero_scratch_tail(s);
  if (rc) return rc;
  return plain_flag_plane(s, code, q, 0, record_ptr(s, rec));
}

extern "C" int pq_gate2_flag_record(pq_state s, uint32_t code, uint64_t q0, uint64_t q1,
                                    uint64_t rec) {
  FULL_ONLY(s);
  ARGCHK(s, "null state");
  ARGCHK(is_flag2_op(code), "not a flagged two-qubit op code");
  if (q0 >= s->n || q1 >= s->n || q0 == q1)
    return fail(PQ_ERR_ARG, "two-qubit gate needs distinct qubits in range");
  if (rec >= s->nrec) return fail(PQ_ERR_RANGE, "Measurement id is out of range.");
  return plain_flag_plane(s, code, q0, q1, record_ptr(s, rec));
}

extern "C" int pq_gate_flag_host(pq_state s, uint32_t code, uint64_t q,
                                 const int32_t *flags) {
  FULL_ONLY(s);
  ARGCHK(s, "null state");
  ARGCHK(is_flag_op(code), "not a flagged op code");
  ARGCHK(flags || s->B == 0, "flags is null");
  if (q >= s->n) return fail(PQ_ERR_ARG, "qubit index out of range");
  if (!s->tab) return PQ_OK;
  int rc = zero_scratch_tail(s);
  if (rc) return rc;
  rc = stage_host_ints(s, flags, s->scratch);
  if (rc) return rc;
  return plain_flag_plane(s, code, q, 0, s->scratch);
}

extern "C" int pq_gate2_flag_host(pq_state s, uint32_t code, uint64_t q0, uint64_t q1,
                                  const int32_t *flags) {
  FULL_ONLY(s);
  ARGCHK(s, "null state");
  ARGCHK(is_flag2_op(code), "not a flagged two-qubit op code");
  ARGCHK(flags || s->B == 0, "flags is null");
  if (q0 >= s->n || q1 >= s->n || q0 == q1)
    return fail(PQ_ERR_ARG, "two-qubit gate needs distinct qubits in range");
  if (!s->tab) return PQ_OK;
  int rc = stage_host_ints(s, flags, s->scratch);
  if (rc) return rc;
  return plain_flag_plane(s, code, q0, q1, s->scratch);
}

// Flagged op `code` whose flag is a Bernoulli draw with coin threshold T
// (PQ_FLAG_RANDOM): the flag plane is drawn into scratch[0, W) from the
// state's next noise stream, then the flagged op runs on it.
int plain_random_flag(pq_state s, uint32_t code, uint64_t q0, uint64_t q1, uint32_t T) {
  const uint64_t k = s->noise_seq++;
  if (!s->tab) return PQ_OK;
  NoiseDesc d{stream_key((uint64_t)s->seed, 2 * k + 1), 0, T, 0, kNoiseFlag};
  HIPCHK(launch_noise_planes(s->stream, s->scratch, d, nullptr, 1, s->W));
  return plain_flag_plane(s, code, q0, q1, s->scratch);
}

extern "C" int pq_gate_prob(pq_state s, uint32_t code, uint64_t q0, uint64_t q1, float prob) {
  FULL_ONLY(s);
  ARGCHK(s, "null state");
  ARGCHK(is_gate(code), "not a gate op code");
  if (q0 >= s->n) return fail(PQ_ERR_ARG, "qubit index out of range");
  if (is_two_qubit(code) && (q1 >= s->n || q1 == q0))
    return fail(PQ_ERR_ARG, "two-qubit gate needs distinct qubits in range");
  return plain_random_flag(s, flagged_code(code), q0, is_two_qubit(code) ? q1 : 0,
                           pq_coin_threshold(prob));
}

int plain_noise(pq_state s, uint64_t q, uint32_t t0, uint32_t t1, uint32_t t2) {
  ARGCHK(s, "null state");
  if (q >= s->n) return fail(PQ_ERR_ARG, "qubit index out of range");
  ARGCHK(t0 <= t1 && t1 <= t2, "noise thresholds must be cumulative");
  const uint64_t k = s->noise_seq++;
  if (!s->tab) return PQ_OK;
  NoiseDesc d{stream_key((uint64_t)s->seed, 2 * k + 1), t0, t1, t2, kNoisePauli};
  HIPCHK(launch_noise_planes(s->stream, s->scratch, d, nullptr, 1, s->W));
  HIPCHK(launch_gate(s->stream, s->tab, geom_of(s), PQ_OP_NOISE, (uint32_t)q, 0,
                     s->scratch, s->scratch + s->W));
  return PQ_OK;
}

extern "C" int pq_noise(pq_state s, uint64_t q, uint32_t t0, uint32_t t1, uint32_t t2) {
  FULL_ONLY(s);
  return plain_noise(s, q, t0, t1, t2);
}

int plain_measure(pq_state s, uint64_t q, uint32_t T, uint64_t rec,
                  const uint64_t *coin_plane_dev) {
  const uint64_t m = s->meas_seq++;
  s->rec_valid[rec] = 0;
  if (!s->tab) return PQ_OK;
  s->xz_identity = false;
  if (s->n > measure_max_qubits())
    return fail(PQ_ERR_UNSUPPORTED, "measurement supports at most 4095 qubits");
  MeasParams mp{(uint32_t)q, T, stream_key((uint64_t)s->seed, 2 * m), coin_plane_dev,
                record_ptr(s, rec), false};
  HIPCHK(launch_measure(s->stream, s->tab, geom_of(s), mp));
  return PQ_OK;
}

extern "C" int pq_measure(pq_state s, uint64_t q, float bias, int64_t rec,
                          const int32_t *coins, uint64_t *out_rec) {
  FULL_ONLY(s);
  ARGCHK(s, "null state");
  if (q >= s->n) return fail(PQ_ERR_ARG, "qubit index out of range");
  uint64_t id;
  if (rec < 0) {
    id = s->nrec;
    int rc = ensure_records(s, s->nrec + 1);
    if (rc) return rc;
  } else {
    if ((uint64_t)rec >= s->nrec) return fail(PQ_ERR_RANGE, "Measurement id is out of range.");
    id = (uint64_t)rec;
  }
  const uint64_t *cp = nullptr;
  if (coins && s->B) {
    int rc = stage_host_ints(s, coins, s->scratch + s->W);
    if (rc) return rc;
    cp = s->scratch + s->W;
  }
  int rc = plain_measure(s, q, pq_coin_threshold(bias), id, cp);
  if (rc) return rc;
  if (out_rec) *out_rec = id;
  return PQ_OK;
}

// ---------------------------------------------------------------------------
// state I/O
// ---------------------------------------------------------------------------
static uint64_t io_chunk_shots(const pq_state_s *s) {
  const uint64_t per_shot = s->R * s->n * 8 + s->R * 4;  // x+z+r int32
  uint64_t c = (512ull << 20) / std::max<uint64_t>(per_shot, 1);
  c = std::max<uint64_t>(64, c - c % 64);
  return c;
}

extern "C" int pq_device_to_host(pq_state s, int32_t *x, int32_t *z, int32_t *r) {
  FULL_ONLY(s);
  ARGCHK(s, "null state");
  if (!s->tab) return PQ_OK;
  ARGCHK(x && z && r, "output buffer is null");
  const uint64_t chunk = std::min<uint64_t>(io_chunk_shots(s), s->B);
  const uint64_t per = s->R * s->n;
  int32_t *dx = nullptr, *dz = nullptr, *dr = nullptr;
  HIPCHK(hipMalloc(&dx, chunk * per * 4));
  HIPCHK(hipMalloc(&dz, chunk * per * 4));
  HIPCHK(hipMalloc(&dr, chunk * s->R * 4));
  int rc = PQ_OK;
  for (uint64_t b0 = 0; b0 < s->B && rc == PQ_OK; b0 += chunk) {
    const uint64_t nb = std::min(chunk, s->B - b0);
    hipError_t e = launch_unpack(s->stream, s->tab, geom_of(s), b0, nb, dx, dz, dr);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e == hipSuccess) e = hipMemcpy(x + b0 * per, dx, nb * per * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(z + b0 * per, dz, nb * per * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(r + b0 * s->R, dr, nb * s->R * 4, hipMemcpyDeviceToHost);
    if (e != hipSuccess) rc = fail(PQ_ERR_HIP, std::string("device_to_host: ") + hipGetErrorString(e));
  }
  (void)hipFree(dx);
  (void)hipFree(dz);
  (void)hipFree(dr);
  return rc;
}

extern "C" int pq_host_to_device(pq_state s, const int32_t *x, const int32_t *z,
                                 const int32_t *r) {
  FULL_ONLY(s);
  ARGCHK(s, "null state");
  if (!s->tab) return PQ_OK;
  ARGCHK(x && z && r, "input buffer is null");
  s->xz_identity = false;
  const uint64_t chunk = std::min<uint64_t>(io_chunk_shots(s), 64 * s->W);
  const uint64_t per = s->R * s->n;
  int32_t *dx = nullptr, *dz = nullptr, *dr = nullptr;
  HIPCHK(hipMalloc(&dx, chunk * per * 4));
  HIPCHK(hipMalloc(&dz, chunk * per * 4));
  HIPCHK(hipMalloc(&dr, chunk * s->R * 4));
  int rc = PQ_OK;
  HIPCHK(hipStreamSynchronize(s->stream));
  for (uint64_t b0 = 0; b0 < s->B && rc == PQ_OK; b0 += chunk) {
    const uint64_t nb = std::min(chunk, s->B - b0);
    hipError_t e = hipMemcpy(dx, x + b0 * per, nb * per * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dz, z + b0 * per, nb * per * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dr, r + b0 * s->R, nb * s->R * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess)
      e = launch_pack(s->stream, s->tab, geom_of(s), b0 / 64, (nb + 63) / 64, dx, dz, dr);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) rc = fail(PQ_ERR_HIP, std::string("host_to_device: ") + hipGetErrorString(e));
  }
  (void)hipFree(dx);
  (void)hipFree(dz);
  (void)hipFree(dr);
  return rc;
}

extern "C" int pq_checksum(pq_state s, int64_t b0, int64_t b1, int64_t *out) {
  FULL_ONLY(s);
  ARGCHK(s && out, "null argument");
  if (b1 < 0) b1 = (int64_t)s->B;
  ARGCHK(b0 >= 0 && b0 <= b1 && (uint64_t)b1 <= s->B, "shot range out of bounds");
  *out = 0;
  if (!s->tab || b0 == b1) return PQ_OK;
  HIPCHK(hipMemsetAsync(s->csum, 0, 8, s->stream));
  HIPCHK(launch_checksum(s->stream, s->tab, geom_of(s), (uint64_t)b0, (uint64_t)b1,
                         (unsigned long long *)s->csum));
  unsigned long long v = 0;
  HIPCHK(hipMemcpyAsync(&v, s->csum, 8, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  *out = (int64_t)v;
  return PQ_OK;
}

extern "C" int pq_compare(pq_state a, pq_state b, uint64_t *planes_differ,
                          uint64_t *records_differ) {
  FULL_ONLY(a);
  FULL_ONLY(b);
  ARGCHK(a && b, "null state");
  ARGCHK(a->n == b->n && a->B == b->B, "pq_compare needs equal n and B");
  // on a's stream, after b's queued work
  if (a != b) {
    for (pq_state st : {a, b})
      if (!st->order_ev) HIPCHK(hipEventCreateWithFlags(&st->order_ev, hipEventDisableTiming));
    HIPCHK(hipEventRecord(b->order_ev, b->stream));
    HIPCHK(hipStreamWaitEvent(a->stream, b->order_ev, 0));
  }
  const uint64_t nr = std::min(a->nrec, b->nrec);
  HIPCHK(launch_zero_word(a->stream, (unsigned long long *)a->csum));
  HIPCHK(launch_count_diff(a->stream, a->tab, b->tab, a->W * a->C * a->R,
                           (unsigned long long *)a->csum));
  unsigned long long v[2] = {0, 0};
  HIPCHK(hipMemcpyAsync(&v[0], a->csum, 8, hipMemcpyDeviceToHost, a->stream));
  HIPCHK(launch_zero_word(a->stream, (unsigned long long *)a->csum));
  if (nr && a->recs && b->recs)
    HIPCHK(launch_count_diff(a->stream, a->recs, b->recs, nr * 2 * a->W,
                             (unsigned long long *)a->csum));
  HIPCHK(hipMemcpyAsync(&v[1], a->csum, 8, hipMemcpyDeviceToHost, a->stream));
  HIPCHK(hipStreamSynchronize(a->stream));
  if (planes_differ) *planes_differ = v[0];
  if (records_differ) *records_differ = v[1] + (a->nrec != b->nrec ? 1 : 0);
  return PQ_OK;
}

extern "C" int pq_read_planes(pq_state s, uint64_t *host_words) {
  FULL_ONLY(s);
  ARGCHK(s && host_words, "null argument");
  if (!s->tab) return PQ_OK;
  HIPCHK(hipStreamSynchronize(s->stream));
  HIPCHK(hipMemcpy(host_words, s->tab, s->W * s->C * s->R * 8, hipMemcpyDeviceToHost));
  return PQ_OK;
}

extern "C" int pq_device_handles(pq_state s, void **tableau, void **records,
                                 uint64_t *words, uint64_t *rows, void **stream) {
  FULL_ONLY(s);
  ARGCHK(s, "null state");
  if (tableau) s->exposed = true;
  if (tableau) *tableau = s->tab;
  if (records) *records = s->recs;
  if (words) *words = s->W;
  if (rows) *rows = s->R;
  if (stream) *stream = (void *)s->stream;
  return PQ_OK;
}

extern "C" int pq_synchronize(pq_state s) {
  ARGCHK(s, "null state");
  HIPCHK(hipStreamSynchronize(s->stream));
  return PQ_OK;
}

extern "C" int pq_event_record(pq_state s, int slot) {
  ARGCHK(s && slot >= 0 && slot < 8, "bad event slot");
  if (!s->ev[slot]) HIPCHK(hipEventCreate(&s->ev[slot]));
  HIPCHK(hipEventRecord(s->ev[slot], s->stream));
  return PQ_OK;
}

extern "C" int pq_event_elapsed_ms(pq_state s, int a, int b, float *ms) {
  ARGCHK(s && ms && a >= 0 && a < 8 && b >= 0 && b < 8, "bad event slot");
  ARGCHK(s->ev[a] && s->ev[b], "event slot not recorded");
  HIPCHK(hipEventSynchronize(s->ev[b]));
  HIPCHK(hipEventElapsedTime(ms, s->ev[a], s->ev[b]));
  return PQ_OK;
}

// ---------------------------------------------------------------------------
// diagnostics mirroring the reference's test-only functors
// ---------------------------------------------------------------------------
extern "C" int pq_row_multiply(pq_state s, uint64_t i, uint64_t k) {
  FULL_ONLY(s);
  ARGCHK(s, "null state");
  ARGCHK(i < s->R && k < s->R, "row out of range");
  if (!s->tab) return PQ_OK;
  s->xz_identity = false;
  HIPCHK(launch_rowsum(s->stream, s->tab, geom_of(s), (uint32_t)i, (uint32_t)k, nullptr));
  return PQ_OK;
}

extern "C" int pq_row_product_sign(pq_state s, uint64_t i, uint64_t k, int32_t *out) {
  FULL_ONLY(s);
  ARGCHK(s && out, "null argument");
  ARGCHK(i < s->R && k < s->R, "row out of range");
  if (!s->tab) return PQ_OK;
  HIPCHK(launch_rowsum(s->stream, s->tab, geom_of(s), (uint32_t)i, (uint32_t)k, s->scratch));
  std::vector<uint64_t> h(s->W);
  HIPCHK(hipStreamSynchronize(s->stream));
  HIPCHK(hipMemcpy(h.data(), s->scratch, s->W * 8, hipMemcpyDeviceToHost));
  for (uint64_t b = 0; b < s->B; ++b) out[b] = (int32_t)((h[b >> 6] >> (b & 63)) & 1);
  return PQ_OK;
}

extern "C" int pq_measure_determined(pq_state s, uint64_t q, int32_t *out) {
  FULL_ONLY(s);
  ARGCHK(s && out, "null argument");
  if (q >= s->n) return fail(PQ_ERR_ARG, "qubit index out of range");
  if (!s->tab) return PQ_OK;
  uint64_t *rec = s->scratch;  // 2W words: results, determined (scratch is 3W)
  s->xz_identity = false;
  MeasParams mp{(uint32_t)q, 0, 0, nullptr, rec, true};
  HIPCHK(launch_measure(s->stream, s->tab, geom_of(s), mp));
  std::vector<uint64_t> h(s->W);
  HIPCHK(hipStreamSynchronize(s->stream));
  HIPCHK(hipMemcpy(h.data(), rec, s->W * 8, hipMemcpyDeviceToHost));
  s->scratch_zeroed = false;
  for (uint64_t b = 0; b < s->B; ++b) out[b] = (int32_t)((h[b >> 6] >> (b & 63)) & 1);
  return PQ_OK;
}
