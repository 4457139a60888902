// Shared host/device primitives of the bit-plane tableau engine.
//
// Layout: a state with n qubits and B shots keeps W = ceil(B/64) word groups;
// group w owns a contiguous block of C = 2n+1 columns x R = 2n+1 rows of
// uint64 words (column c < n: x_c, n <= c < 2n: z_{c-n}, c = 2n: r), and bit
// s of every word of group w belongs to shot 64w+s.  Word (w, c, row) lives at
// tab[(w*C + c)*R + row].
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define PQ_HD __host__ __device__ __forceinline__

namespace pq {

constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ull;

// SplitMix64 output function (Steele, Lea, Flood, "Fast splittable
// pseudorandom number generators", OOPSLA 2014).
PQ_HD uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Output idx+1 of a SplitMix64 generator whose state is output stream+1 of
// a SplitMix64 generator seeded with `seed`.  Coins of the m-th measurement
// use stream 2m, the k-th noise op uses stream 2k+1; idx is the shot.
PQ_HD uint64_t stream_key(uint64_t seed, uint64_t stream) {
  return mix64(seed + (stream + 1) * kGolden);
}
PQ_HD uint64_t draw_keyed(uint64_t key, uint64_t idx) {
  return mix64(key + (idx + 1) * kGolden);
}
PQ_HD uint64_t draw(uint64_t seed, uint64_t stream, uint64_t idx) {
  return draw_keyed(stream_key(seed, stream), idx);
}

// Bit-sliced mod-4 accumulation of the Pauli-product phase exponent
// g(dest, src) (MeasurementFunctors.hpp:10-72; destination first, as
// RowProductSign calls it at :112-113).  Per shot bit: the Paulis
// anticommute iff anti = x1 z2 ^ z1 x2; then g = +1 for (X,Y),(Y,Z),(Z,X)
// and -1 otherwise, and -1 == ~(z1 ^ x2) & (x1 | z2) on those inputs.
// (c0, c1) is the running sum mod 4; adding +1 carries c0 into c1, adding
// -1 (== +3) carries ~c0.
PQ_HD void phase_acc(uint64_t x1, uint64_t z1, uint64_t x2, uint64_t z2,
                     uint64_t &c0, uint64_t &c1) {
  const uint64_t anti = (x1 & z2) ^ (z1 & x2);
  const uint64_t minus = ~(z1 ^ x2) & (x1 | z2);
  c1 ^= anti & (c0 ^ minus);
  c0 ^= anti;
}

// (a0, a1) += (b0, b1) mod 4, bitwise per shot.
PQ_HD void add_mod4(uint64_t &a0, uint64_t &a1, uint64_t b0, uint64_t b1) {
  a1 ^= b1 ^ (a0 & b0);
  a0 ^= b0;
}

// Device-side instruction of a compiled program (32 bytes).
//  gates: code, qa, qb; flagged ops: slot = record id, or kSlotPlane | k
//  when the flag is the X part of noise plane k (a random flag);
//  na / nb: noise plane fused after the op on qa / qb (kNone if none);
//  t0 (gate segments): local column indices of qa (low 16 bits) and qb
//  (high 16 bits) in the segment's qubit list;
//  pad (gate segments): on the first op of a batch, the batch length (bits
//  0-7) and kLayerStart if the batch opens a new layer (all ops of one layer
//  act on pairwise distinct qubits); on the first op of a run -- a layer,
//  or the part of a layer inside one staging chunk -- also the run length
//  (bits 8-19);
//  PQ_OP_NOISE alone: qa, na;  M: qa, slot = record id, t0 = coin
//  threshold, (na, nb) = coin stream (low, high 32 bits), pad = kMeasSuperseded
//  when a later M of the same measurement segment writes the same record.
struct DInstr {
  uint32_t code, qa, qb, slot, na, nb, t0, pad;
};
constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr uint32_t kSlotPlane = 0x80000000u;
constexpr uint32_t kLayerStart = 0x80000000u;
constexpr uint32_t kBatchLenMask = 0xFFu;
constexpr uint32_t kRunShift = 8;
constexpr uint32_t kRunMask = 0xFFFu;
// An M whose record a later M of its segment overwrites: the measurement still
// runs (it may leave the scratch row), but its record write is skipped, so the
// last write in program order is the one that stays even when both run at
// once in a batch of fully determined measurements.
constexpr uint32_t kMeasSuperseded = 1u;

// Segment of a compiled program: a run of gate batches or a run of
// measurements.  Gate segments hold `count` instructions in batches.  A gate
// segment with nq > 0 may run row-tiled out of LDS: its nq distinct qubits
// are listed at segq[qoff .. qoff+nq) in first-use order (the local column
// indices of its instructions refer to this list).
struct DSegment {
  uint32_t kind;  // 0 = gates, 1 = measurements
  uint32_t begin, count, nq, qoff, pad0, pad1, pad2;
};

// Noise plane descriptor: stream, thresholds and kind.  A plane pair holds,
// per 64-shot word, the shots whose draw v = draw_keyed(key, shot) fires:
//  kNoisePauli: u = v >> 32; X part u < t1 (X or Y), Z part t0 <= u < t2
//               (Y or Z) for cumulative thresholds t0 <= t1 <= t2;
//  kNoiseFlag:  X part (v >> 40) < t1 (a coin threshold, 2^24 = always),
//               Z part empty: the flag of a randomly flagged op.
constexpr uint32_t kNoisePauli = 0, kNoiseFlag = 1;
struct NoiseDesc {
  uint64_t key;  // stream_key(seed, 2k+1)
  uint32_t t0, t1, t2, kind;
};
PQ_HD void noise_bits(const NoiseDesc &d, uint64_t v, bool &bx, bool &bz) {
  if (d.kind == kNoiseFlag) {
    bx = (uint32_t)(v >> 40) < d.t1;
    bz = false;
  } else {
    const uint32_t u = (uint32_t)(v >> 32);
    bx = u < d.t1;
    bz = (u >= d.t0) & (u < d.t2);
  }
}

}  // namespace pq
