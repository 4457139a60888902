// Launch wrappers of the gfx950 kernels (host C++ linkage, used by the C ABI).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pq_common.h"

namespace pq {

struct Geom {
  uint32_t n, R, C;  // qubits, rows (2n+1), columns (2n+1)
  uint64_t W;        // 64-shot word groups
  uint64_t B;        // shots
};

// Measurement parameters of one MeasureQubit call.
struct MeasParams {
  uint32_t q;
  uint32_t coin_T;           // coin threshold (24-bit scale)
  uint64_t coin_key;         // stream_key(seed, 2m)
  const uint64_t *coin_plane;  // optional replay plane (W words) or nullptr
  uint64_t *rec;             // record base: results [W], determined [W]
  bool force_determined = false;  // MeasureDetermined on every shot (KAT use)
};

hipError_t launch_fill_identity(hipStream_t s, uint64_t *tab, const Geom &g);
hipError_t launch_gate(hipStream_t s, uint64_t *tab, const Geom &g, uint32_t code,
                       uint32_t qa, uint32_t qb, const uint64_t *fa,
                       const uint64_t *fb);
// Noise planes for nk descriptors (desc on the device), or for the single
// descriptor `one` passed by value when desc is nullptr.
hipError_t launch_noise_planes(hipStream_t s, uint64_t *planes, NoiseDesc one,
                               const NoiseDesc *desc, uint32_t nk, uint64_t W);
hipError_t launch_measure(hipStream_t s, uint64_t *tab, const Geom &g,
                          const MeasParams &m);
size_t measure_lds_bytes(uint32_t n);
uint32_t measure_max_qubits();
hipError_t launch_unpack(hipStream_t s, const uint64_t *tab, const Geom &g,
                         uint64_t b0, uint64_t nb, int32_t *x, int32_t *z,
                         int32_t *r);
hipError_t launch_pack(hipStream_t s, uint64_t *tab, const Geom &g, uint64_t w0,
                       uint64_t nw, const int32_t *x, const int32_t *z,
                       const int32_t *r);
hipError_t launch_checksum(hipStream_t s, const uint64_t *tab, const Geom &g,
                           uint64_t b0, uint64_t b1,
                           unsigned long long *out_dev);
hipError_t launch_flags_to_plane(hipStream_t s, const int32_t *flags,
                                 uint64_t B, uint64_t *plane);
hipError_t launch_record_to_int(hipStream_t s, const uint64_t *rec, uint64_t B,
                                uint64_t W, int32_t *res, int32_t *det);

// Persistent program kernel: one workgroup per 64-shot word group.
struct ProgramArgs {
  uint64_t *tab;
  uint64_t *recs;         // record arena [rec][2][W]
  const uint64_t *planes;  // noise planes [k][2][W]
  const DSegment *segs;
  const DInstr *instrs;
  const uint64_t *coin_planes;  // optional replay planes [m][W] or nullptr
  uint32_t nseg;
  uint32_t n, R, C;
  uint64_t W;
  uint64_t word0;  // first word group handled by this launch
  uint64_t nwords; // word groups handled by this launch
  // optional work census (4 counters, accumulated over word groups):
  // random measurements, determined measurements, random rowsum rows,
  // determined rowsum steps
  unsigned long long *census;
  uint32_t det_waves;  // waves that run fully-determined measurements in parallel
};
// Dynamic LDS bytes k_program needs for n qubits and det_waves slabs.
size_t program_lds_bytes(uint32_t n, uint32_t det_waves);
// RowMultiply(i <- i*k) on every shot, or only the sign into sign_out.
hipError_t launch_rowsum(hipStream_t s, uint64_t *tab, const Geom &g, uint32_t i,
                         uint32_t k, uint64_t *sign_out);
// Writes {seed, mseq0, nseq0} into the device parameter block `params`
// (a kernel, so it is stream-ordered and safe to issue before a graph launch).
hipError_t launch_set_params(hipStream_t s, void *params, uint64_t seed,
                             uint64_t mseq0, uint64_t nseq0);
// Raise the dynamic-LDS limit of the measurement kernels for n qubits.
hipError_t prepare_kernels(uint32_t n);
// `params` points at a device ProgParams {seed, mseq0, nseq0} read by the
// kernels at run time, so a captured graph can be replayed with new values.
hipError_t launch_program(hipStream_t s, const ProgramArgs &a, uint32_t threads,
                          uint32_t blocks, const void *params);
hipError_t launch_program_noise(hipStream_t s, uint64_t *planes,
                                const NoiseDesc *desc, uint32_t nk, uint64_t W,
                                const void *params);

}  // namespace pq
