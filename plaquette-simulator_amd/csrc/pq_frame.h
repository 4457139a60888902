// Frame form of a program (shared-X/Z circuits): the compiled plan.
#pragma once

#include <stdint.h>

#include <mutex>
#include <string>
#include <vector>

#include "pq_common.h"
#include "pq_kernels.h"

struct pq_program_s;
struct pq_state_s;

// LDS of one word group: its words (Rv, frame, records) plus staging
// buffers for the noise-mask window (8-byte words) and the stream (uint32),
// sized per plan (build_frame_plan) for the most groups per CU.
constexpr uint64_t kFrameLdsMax = 160 * 1024;

struct FramePlan {
  uint32_t R = 0, nw = 0;
  uint32_t nref = 0;                 // records read by flag ops (rec_l words)
  uint32_t lds_rec = 0, lds_const = 0, lds_noise = 0;  // LDS word offsets
  uint32_t nb_words = 0, sb_words = 0, lds_bytes = 0;
  uint32_t nbufs = 1;                // staging buffers (2: load ahead)
  std::vector<uint32_t> stream;      // ops
  std::vector<uint32_t> chunks;      // {begin, words, k0, nk} per chunk
  std::vector<uint32_t> preload;     // {slot, rec_l index} loaded at start
  std::vector<uint64_t> X, Z;        // final shared X/Z, [R][nw]
  bool final_identity = false;
  // a two-part run: chunks [0, split_chunk) read only noise ops below
  // split_k (0 when the plan does not split)
  uint32_t split_chunk = 0, split_k = 0;
  uint64_t nrandom = 0, ndet = 0;
  uint32_t *d_stream = nullptr, *d_chunks = nullptr, *d_preload = nullptr;
  uint64_t *d_X = nullptr, *d_Z = nullptr;
  std::mutex upload_mu;
  ~FramePlan();
};

// `groups`: the word groups one launch of the plan runs (UINT64_MAX: many,
// as in batches and the server); it sizes the LDS staging, not the results.
int build_frame_plan(pq_program_s *p, uint64_t groups);
int upload_frame_plan(pq_program_s *p, hipStream_t stream);
// Kernel arguments of program p's frame plan on state s (at s's current
// measurement sequence number).
pq::FrameArgs frame_args(pq_state_s *s, pq_program_s *p);
// What a frame plan depends on, as bytes: n and the op list with the noise
// thresholds left out (they live in the noise descriptors, not the plan).
// Programs with equal keys can share one plan.
std::string frame_plan_key(uint64_t n, const pq_op *ops, uint64_t nops);
int run_frame(pq_state_s *s, pq_program_s *p);
