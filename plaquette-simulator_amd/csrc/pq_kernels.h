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
// The R column alone ([W][R] words) of the |0...0> tableau: r = 1 on the
// scratch row R-1 only.
hipError_t launch_fill_r_identity(hipStream_t s, uint64_t *rcol, uint64_t W, uint32_t R);
// The R column (column 2n) of every word group of a full tableau, as the
// |0...0> tableau has it; X/Z untouched.
hipError_t launch_fill_r_column(hipStream_t s, uint64_t *tab, const Geom &g);
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
// Zeroes one device word (a kernel, never hipMemsetAsync: graph-capturable).
hipError_t launch_zero_word(hipStream_t s, unsigned long long *p);
// Counts the words of a and b (n each) that differ into *out_dev.
hipError_t launch_count_diff(hipStream_t s, const uint64_t *a, const uint64_t *b, uint64_t n,
                             unsigned long long *out_dev);
hipError_t launch_flags_to_plane(hipStream_t s, const int32_t *flags,
                                 uint64_t B, uint64_t *plane);
hipError_t launch_record_to_int(hipStream_t s, const uint64_t *rec, uint64_t B,
                                uint64_t W, int32_t *res, int32_t *det);

// Persistent program kernel: one workgroup per 64-shot word group.
struct ProgramArgs {
  uint64_t *tab;
  uint64_t *recs;         // record arena [rec][2][W]
  const uint64_t *planes;  // noise planes, group-major [W][nk][2] (X part, Z part)
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
  const uint32_t *segq;  // qubit lists of row-tiled gate segments
  uint64_t lds_words;    // dynamic LDS of the launch, in 8-byte words
  // Workgroups sharing one word group (1: one workgroup per group).  With
  // slices > 1 the launch is cooperative (every workgroup resident), slice s
  // of group wi is workgroup wi * slices + s, the slices split the rows of
  // gate segments, the rowsums of random measurements and the determined
  // measurements, and meet at barriers on sync[wi] (zero at launch); a wait
  // that gives up sets *sync_err.
  uint32_t slices;
  unsigned int *sync;
  unsigned int *sync_err;
  uint32_t sync_spin;  // polls before a barrier gives up (0: the kernel's default)
  uint32_t nk;         // noise planes per word group
  // 1: X/Z are the identity and not in memory; segment 0 (a row-tiled gate
  // segment over all n qubits) writes its tiles instead of loading them
  uint32_t fresh;
  // 1: the program has row-tiled gate segments; above 512 qubits it runs in
  // the 768-thread build (program_max_threads)
  uint32_t wide;
};
// One state of a mixed batch: its program arguments, run parameters and
// noise descriptors (planes are generated by the batch prologue).
struct BatchCircuit {
  ProgramArgs a;
  uint64_t seed, mseq0, nseq0;
  const NoiseDesc *noise;
  uint32_t nnoise, pad;
  uint64_t task0;  // first noise task (word group x 32 ops) of this circuit in the batch
};
// Work items are (circuit << 40 | word group), in execution order; with
// wtile > 1 an item stands for wtile consecutive word groups of its circuit
// from the one it names.  noise_tasks > 0: the noise planes of every circuit
// first (launch_batch_frame_noise over circ[c].task0 numbering).
// chain_qubits: the largest member of 193..512 qubits (0: none); such members
// keep their determined chains in registers when the block has at most 512
// threads (the large build), in LDS otherwise.
hipError_t launch_batch(hipStream_t s, const BatchCircuit *circ, uint32_t ncirc,
                        const uint64_t *items, uint64_t nitems,
                        unsigned long long *counter, uint32_t threads,
                        uint32_t blocks, size_t lds_bytes, uint64_t noise_tasks,
                        uint32_t chain_qubits, uint32_t wtile = 1);
hipError_t prepare_batch_kernel();
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
// Zeroes n 32-bit device words (a kernel: graph-capturable).
hipError_t launch_zero_u32(hipStream_t s, unsigned int *p, uint32_t n);
// Whether `slices` workgroups per word group for `nwords` groups of n qubits
// fit the device at once (the cooperative launch's condition).
bool program_slices_fit(uint32_t n, uint64_t nwords, uint32_t slices, uint32_t threads,
                        uint64_t lds_bytes, bool wide = false);
// Largest block of the persistent kernel build used for n qubits: 768 up to
// 192 qubits, 512 for 193..512, and above 512 768 for programs with row-tiled
// gate segments (`wide`), else 512.  Launches clamp their block size to it.
uint32_t program_max_threads(uint32_t n, bool wide = false);
// Rows per tile of a row-tiled gate segment over nq qubits with `lds_words`
// of dynamic LDS (0: the segment runs in the direct form).
uint32_t program_tile_rows(uint64_t lds_words, uint32_t nq);
// The noise planes of a program run, group-major [W][nk][2] as the
// persistent kernel reads them (the frame form's layout and kernel).
hipError_t launch_program_noise(hipStream_t s, uint64_t *planes,
                                const NoiseDesc *desc, uint32_t nk, uint64_t W,
                                const void *params);

// Frame form: one workgroup per 64-shot word group runs the compiled stream
// on the group's R column and Pauli frame (see pq_frame.cpp for the stream
// layout).
#ifndef PQ_FRAME_THREADS
#define PQ_FRAME_THREADS 256
#endif
constexpr uint32_t kFrameThreads = PQ_FRAME_THREADS;  // threads per word group
constexpr uint32_t kFrameXor = 1, kFrameRandom = 2, kFrameDet = 3,
                   kFrameLayer = 4;  // stream op kinds
struct FrameArgs {
  uint64_t *tab;
  uint64_t *recs;          // record arena [rec][2][W]
  const uint64_t *planes;  // noise planes, group-major [W][nnoise][2]
  const uint32_t *stream;
  const uint32_t *chunks;  // {begin, words, k0, nk} per chunk
  const uint32_t *preload; // {slot, rec_l index} pairs
  uint32_t nchunk, npreload;
  uint32_t n, R, C;
  uint64_t W;
  uint64_t w0, wn;  // k_frame runs word groups [w0, w0 + wn)
  uint32_t c0, c1;  // ... and stream chunks [c0, c1)
  // a run in two launches: the first ends by saving each group's frame and
  // record words (LDS words R .. R + 2n + nref) to side[w][2n + nref], the
  // second starts from them (the R column goes through the tableau)
  uint32_t suspend, resume;
  uint64_t *side;
  uint64_t nnoise;
  uint32_t lds_rec, lds_const, lds_noise, nb_words;
  uint32_t sb_words;   // stream buffer (uint32), then 64 partial words per wave
  uint32_t nbufs;      // staging buffers: noise windows, then stream buffers
  uint32_t lds_bytes;
  uint64_t seed, mseq0;  // coin streams: measurement m of the run uses 2(mseq0 + m)
  // group w's R column: tab + w * col_stride + r_off (C*R and 2n*R for a
  // full tableau, R and 0 for an R-column-only state)
  uint64_t col_stride, r_off;
};
hipError_t launch_frame(hipStream_t s, const FrameArgs &a);
// A mixed batch in the frame form: items are (circuit << 40 | word group);
// lds_bytes the largest any circuit needs.
hipError_t launch_frame_batch(hipStream_t s, const FrameArgs *circ, const uint64_t *items,
                              uint64_t nitems, uint32_t lds_bytes);
// Group-major noise planes [W][nk][2] of the frame form (the same masks as
// launch_program_noise writes by plane), word groups [w0, w0 + wn), noise ops
// [ka, min(kb, nk)).
hipError_t launch_frame_noise(hipStream_t s, uint64_t *planes, const NoiseDesc *desc,
                              uint32_t nk, uint64_t w0, uint64_t wn, const void *params,
                              uint32_t ka = 0, uint32_t kb = ~0u);
// The group-major noise planes of every circuit of a frame batch: `tasks`
// (word group x 32 ops) over all circuits, circuit c's from circ[c].task0 on,
// split evenly over the waves.
hipError_t launch_batch_frame_noise(hipStream_t s, const BatchCircuit *circ, uint32_t ncirc,
                                    uint64_t tasks);
// Writes the shared X/Z ([R][nw] bit rows) into every word group's columns.
hipError_t launch_frame_xz(hipStream_t s, uint64_t *tab, const Geom &g, const uint64_t *X,
                           const uint64_t *Z, uint32_t nw);

}  // namespace pq
