// pq_run_program_static: the persistent form of pq_run_program for callers
// that cannot allocate device memory.  Everything the run needs beyond the
// state's own tableau and records -- the compiled segments, instructions,
// qubit lists and noise descriptors, the noise planes and the run's
// parameter block -- goes into one file-scope device array, `scratch`, of
// PQ_STATIC_SCRATCH_BYTES.  Each GPU holds its own copy of the array from
// the moment the library's code is loaded on it until the process ends.
//
// In host code the array's name is only a stand-in for the symbol; its device
// address is fetched once per device with hipGetSymbolAddress on the first
// call made on that device, and only that address is handed to kernels and
// copies.  Every call on a GPU shares the array and leaves the last call's
// contents behind; each call overwrites all of it that it reads (the program
// is uploaded, the noise planes are rewritten by the noise kernel, the
// parameter block by its kernel) and returns only after its kernels have
// finished.  Calls that overlap on one GPU (from several host threads) race
// on the array and must not be made.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "plaquette_amd.h"
#include "pq_capi_internal.h"
#include "pq_common.h"
#include "pq_kernels.h"
#include "pq_program_internal.h"

using namespace pq;

namespace {

constexpr uint64_t kScratchBytes = PQ_STATIC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) uint64_t scratch[kScratchBytes / 8];

std::mutex g_mu;
void *g_ptr[kMaxDevices] = {};

int fail(int code, const std::string &msg) { return pq_set_error(code, msg.c_str()); }

// The device address of `scratch` on the current device (looked up once).
int scratch_ptr(int dev, uint8_t **out) {
  if (dev < 0 || dev >= kMaxDevices) return fail(PQ_ERR_UNSUPPORTED, "device index too large");
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(PQ_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_ptr[dev] = p;
  }
  *out = static_cast<uint8_t *>(g_ptr[dev]);
  return PQ_OK;
}

uint64_t align256(uint64_t v) { return (v + 255) & ~uint64_t(255); }

// Host vector -> scratch at *off (advanced past it, 256-byte aligned).
template <class T>
hipError_t put(uint8_t *base, uint64_t &off, const std::vector<T> &v, hipStream_t st, T **dst) {
  *dst = nullptr;
  if (v.empty()) return hipSuccess;
  *dst = reinterpret_cast<T *>(base + off);
  const uint64_t bytes = v.size() * sizeof(T);
  off = align256(off + bytes);
  return hipMemcpyAsync(*dst, v.data(), bytes, hipMemcpyHostToDevice, st);
}

}  // namespace

extern "C" int pq_run_program_static(pq_state s, const pq_op *ops, uint64_t nops) {
  if (!s || (!ops && nops)) return fail(PQ_ERR_ARG, "null argument");
  if (s->r_only)
    return fail(PQ_ERR_UNSUPPORTED, "a sampling state runs frame-form programs only");
  if (s->n > measure_max_qubits())
    return fail(PQ_ERR_UNSUPPORTED, "persistent form supports at most 4095 qubits");
  int rc = validate_ops(s->n, ops, nops);
  if (rc) return rc;
  // compile on the host: the same compiler as pq_program_create
  pq_program_s p;
  p.n = s->n;
  p.W = s->W;
  p.R = s->R;
  p.ops.assign(ops, ops + nops);
  if ((rc = compile_program(&p))) return rc;
  // sizes first, so a refused call leaves the state untouched
  const uint64_t planes_bytes = (uint64_t)p.nnoise * 2 * s->W * 8;
  const uint64_t need = 256 + align256(p.segs.size() * sizeof(DSegment)) +
                        align256(p.instrs.size() * sizeof(DInstr)) +
                        align256(p.segq.size() * sizeof(uint32_t)) +
                        align256(p.noise.size() * sizeof(NoiseDesc)) + align256(planes_bytes);
  if (need > kScratchBytes)
    return fail(PQ_ERR_UNSUPPORTED,
                "the program needs " + std::to_string(need) +
                    " bytes of device scratch (compiled ops and noise planes); "
                    "pq_run_program_static holds at most " +
                    std::to_string(kScratchBytes));
  if ((rc = check_records(s, &p))) return rc;
  if (p.prep_end > s->rec_cap)
    return fail(PQ_ERR_OOM, "the state's record arena holds " + std::to_string(s->rec_cap) +
                                " records and the ops write up to " +
                                std::to_string(p.prep_end) +
                                ": size it with an allocating call first (pq_measure, "
                                "pq_run_program)");
  int dev = 0, cur = 0;
  if (hipGetDevice(&cur) != hipSuccess) return fail(PQ_ERR_HIP, "hipGetDevice failed");
  dev = s->device;
  if (cur != dev && hipSetDevice(dev) != hipSuccess) return fail(PQ_ERR_HIP, "hipSetDevice failed");
  uint8_t *base = nullptr;
  rc = scratch_ptr(dev, &base);
  hipError_t e = hipSuccess;
  if (rc == PQ_OK) {
    // the arena is large enough: this sizes nothing, it marks the measured
    // records' host copies stale
    rc = prepare_records(s, &p);
  }
  if (rc == PQ_OK) rc = materialize_tableau(s);
  if (rc == PQ_OK && s->tab && s->W) {
    s->xz_identity = false;
    uint64_t off = 0;
    void *params = base;
    off = 256;
    if (e == hipSuccess) e = put(base, off, p.segs, s->stream, &p.d_segs);
    if (e == hipSuccess) e = put(base, off, p.instrs, s->stream, &p.d_instrs);
    if (e == hipSuccess) e = put(base, off, p.segq, s->stream, &p.d_segq);
    if (e == hipSuccess) e = put(base, off, p.noise, s->stream, &p.d_noise);
    p.d_planes = p.nnoise ? reinterpret_cast<uint64_t *>(base + off) : nullptr;
    p.threads = program_max_threads((uint32_t)s->n, p.tiled);
    p.blocks = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(s->W, 1u << 20));
    if (e == hipSuccess) e = prepare_kernels((uint32_t)s->n);
    if (e == hipSuccess)
      e = launch_set_params(s->stream, params, (uint64_t)s->seed, s->meas_seq, s->noise_seq);
    if (e == hipSuccess && p.nnoise)
      e = launch_program_noise(s->stream, p.d_planes, p.d_noise, (uint32_t)p.nnoise, s->W,
                               params);
    if (e == hipSuccess)
      e = launch_program(s->stream, program_args(s, &p, kProgramLdsBytes), p.threads, p.blocks,
                         params);
    // the next call reuses the array: this one's kernels end first
    const hipError_t es = hipStreamSynchronize(s->stream);
    if (e == hipSuccess) e = es;
    if (e == hipSuccess) {
      s->meas_seq += p.nmeas;
      s->noise_seq += p.nnoise;
    }
  } else if (rc == PQ_OK) {
    s->meas_seq += p.nmeas;
    s->noise_seq += p.nnoise;
  }
  // the program's pointers lead into the array: nothing of it is freed
  p.d_segs = nullptr;
  p.d_instrs = nullptr;
  p.d_segq = nullptr;
  p.d_noise = nullptr;
  p.d_planes = nullptr;
  if (cur != dev) (void)hipSetDevice(cur);
  if (rc) return rc;
  if (e != hipSuccess)
    return fail(PQ_ERR_HIP, std::string("pq_run_program_static: ") + hipGetErrorString(e));
  return PQ_OK;
}
