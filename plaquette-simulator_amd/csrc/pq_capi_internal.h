// Internal state of the C ABI (shared by pq_capi.cpp and pq_program.cpp).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <memory>
#include <vector>

#include "plaquette_amd.h"
#include "pq_kernels.h"

struct FramePlan;

struct pq_state_s {
  uint64_t n = 0, B = 0, R = 0, C = 0, W = 0;
  int64_t seed = 0;
  int device = 0;
  uint64_t *tab = nullptr;     // W * C * R words
  uint64_t *recs = nullptr;    // rec_cap * 2 * W words
  uint64_t nrec = 0, rec_cap = 0;
  uint64_t arena_gen = 0;      // bumped whenever `recs` moves
  uint64_t meas_seq = 0;       // measurements issued so far (coin streams)
  uint64_t noise_seq = 0;      // noise ops issued so far (noise streams)
  hipStream_t stream = nullptr;
  uint64_t *scratch = nullptr;  // 3W words: flag/noise planes for plain ops
  bool scratch_zeroed = false;
  int32_t *stage = nullptr;     // int32[B] host-vector staging
  void *params_dev = nullptr;   // ProgParams of the running program
  void *csum = nullptr;
  // States the server creates borrow a slot stream and allocate with the
  // stream-ordered allocator on it (alloc_stream != nullptr), so freeing
  // them never waits for work on other streams.
  hipStream_t alloc_stream = nullptr;
  bool own_stream = true;
  hipEvent_t ev[8] = {};
  // orders this state's stream against another state's (DeviceToDevice)
  hipEvent_t order_ev = nullptr;
  // X/Z hold the |0...0> tableau in every shot (frame programs start there)
  bool xz_identity = false;
  // reset left X/Z unwritten (the R column is reset): they hold |0...0>
  // once materialize_tableau fills them; frame runs overwrite them instead
  bool xz_pending = false;
  // the raw tableau pointer was handed out: resets write X/Z at once
  bool exposed = false;
  // a frame run left X/Z unwritten: every group's X/Z is this plan's shared
  // X/Z, written by materialize_tableau (never set with xz_pending)
  std::shared_ptr<FramePlan> xz_plan;
  // frame runs of many word groups: the noise of the next slice of groups is
  // drawn on this stream while k_frame runs the current one (created on use)
  hipStream_t aux = nullptr;
  std::vector<hipEvent_t> aux_ev;
  // only the R column is kept (tab is [W][R]): a server-internal state for
  // frame-form requests, whose X/Z nobody reads; only frame runs take it
  bool r_only = false;
  std::vector<std::vector<uint64_t>> rec_host;  // host copies of records
  std::vector<char> rec_valid;
  ~pq_state_s();
};

int pq_set_error(int code, const char *msg);

// Synchronous (hipMalloc/hipFree) or, with a stream, stream-ordered
// (hipMallocAsync/hipFreeAsync) device allocation.
inline hipError_t dev_alloc(void **p, size_t bytes, hipStream_t async_stream) {
  return async_stream ? hipMallocAsync(p, bytes, async_stream) : hipMalloc(p, bytes);
}
inline void dev_free(void *p, hipStream_t async_stream) {
  if (!p) return;
  if (async_stream)
    (void)hipFreeAsync(p, async_stream);
  else
    (void)hipFree(p);
}
// pq_create on a borrowed stream with stream-ordered allocation (server use).
int create_state_on(uint64_t n, uint64_t B, int64_t seed, hipStream_t stream, pq_state *out,
                    bool r_only = false);
pq::Geom geom_of(const pq_state_s *s);
// Fills a reset state's pending X/Z (the |0...0> tableau) before anything
// reads or updates the full tableau.
int materialize_tableau(pq_state s);
int ensure_records(pq_state s, uint64_t count);
uint64_t *record_ptr(pq_state s, uint64_t id);
int plain_gate(pq_state s, uint32_t code, uint64_t q0, uint64_t q1);
int plain_flag_plane(pq_state s, uint32_t code, uint64_t q, uint64_t q1,
                     const uint64_t *plane);
int plain_random_flag(pq_state s, uint32_t code, uint64_t q0, uint64_t q1, uint32_t T);
int plain_noise(pq_state s, uint64_t q, uint32_t t0, uint32_t t1, uint32_t t2);
int plain_measure(pq_state s, uint64_t q, uint32_t T, uint64_t rec,
                  const uint64_t *coin_plane_dev);
