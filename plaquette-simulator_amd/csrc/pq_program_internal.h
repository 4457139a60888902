// Internal representation of a compiled program.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <memory>
#include <vector>

#include "plaquette_amd.h"
#include "pq_common.h"
#include "pq_frame.h"

constexpr uint32_t kProgramBatch = 8;  // must equal kBatch in pq_kernels.hip
constexpr uint32_t kProgramThreadsMax = 768;  // largest block of the persistent kernel builds
constexpr uint32_t kProgramStage = 448;       // must equal kStage there
// Gate segments with more qubits than a row tile of kWindowCutRows rows can
// hold (in the default LDS budget) are cut into qubit windows, each small
// enough for a tile of kWindowRows rows.
constexpr uint32_t kWindowCutRows = 32;
constexpr uint32_t kWindowRows = 48;

struct pq_program_s {
  uint64_t n = 0, W = 0, R = 0;
  std::vector<pq_op> ops;
  uint64_t ngates = 0, nmeas = 0, nnoise = 0;
  std::vector<pq::DSegment> segs;
  std::vector<pq::DInstr> instrs;
  std::vector<pq::NoiseDesc> noise;
  std::vector<uint32_t> segq;  // qubit lists of row-tiled gate segments
  bool tiled = false;          // some gate segment runs row-tiled
  pq::DSegment *d_segs = nullptr;
  uint32_t *d_segq = nullptr;
  pq::DInstr *d_instrs = nullptr;
  pq::NoiseDesc *d_noise = nullptr;
  uint64_t *d_planes = nullptr;  // noise planes [nnoise][2][W]
  uint64_t *d_side = nullptr;    // frame form, two-part runs: [W][2n + nref] words
  uint32_t threads = 0, blocks = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  pq_state graph_state = nullptr;
  uint64_t graph_arena_gen = 0;
  hipStream_t alloc_stream = nullptr;  // stream-ordered allocation (server use)
  unsigned long long *d_census = nullptr;  // 8 counters when census is on
  bool census_on = false;
  bool timing_on = false;
  std::vector<hipEvent_t> ev;  // start/end pairs of timed kernel launches
  size_t ev_used = 0;
  std::shared_ptr<FramePlan> frame;  // frame form plan, built on first use (may be shared)
  // prepare_records' check of the record ids, for a state holding
  // prep_start records (prep_end after the run), and the measured slots
  bool prep_valid = false;
  uint64_t prep_start = 0, prep_end = 0;
  std::vector<uint32_t> m_slots;
  // end of the last run's device work (on last_stream): a run on another
  // stream waits for it, since the noise planes and side buffer are shared
  hipEvent_t done_ev = nullptr;
  hipStream_t last_stream = nullptr;
  // workgroups per word group in persistent runs: requested (0: chosen from
  // the state size and the device), last used, and the slices' barrier
  // counters ([W] words) and give-up flag (one word after them)
  uint32_t slices_req = 0, slices_used = 1;
  uint32_t slice_wait = 0;  // polls before a barrier gives up (0: default)
  unsigned int *d_sync = nullptr;
  uint64_t sync_words = 0;
  void drop_graph();
  ~pq_program_s();
};

int compile_program(pq_program_s *p);
// What the frame form alone needs: the noise descriptors (noise ops and
// random flags, numbered in program order as compile_program numbers them)
// and the op counts, without the persistent form's segments.
int compile_frame_only(pq_program_s *p);
int validate_ops(uint64_t n, const pq_op *ops, uint64_t nops);

constexpr uint64_t kProgramLdsBytes = 158 * 1024;  // dynamic-LDS cap (kProgramLdsBudget there)

// Shared with the batch runner.  check_records validates the record ids
// and sets p->prep_end without allocating; prepare_records also sizes the
// state's record arena and marks the measured records' host copies stale.
int check_records(pq_state s, pq_program p);
int prepare_records(pq_state s, pq_program p);
pq::ProgramArgs program_args(pq_state s, pq_program p, uint64_t lds_budget);
// pq_program_create for a server-owned state: stream-ordered allocation and
// copies on the state's stream.
int create_program_on(pq_state s, const pq_op *ops, uint64_t nops, pq_program *out);
