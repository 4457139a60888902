/*
 * plaquette_amd.h -- C ABI of the MI355X (gfx950) batched stabilizer-tableau
 * engine.  Plain pointers and sizes only; every entry point returns an int
 * status (PQ_OK == 0) and never throws.  pq_last_error() gives the message
 * of the calling thread's last failure.
 *
 * The engine replaces the device side of the reference's
 * Plaquette::CliffordState<int> (plaquette_simulator/src/CliffordState.hpp).
 * Each entry point names the reference member it stands in for.
 *
 * Data model.  A state holds n qubits and B shots.  Its tableau has
 * R = 2n+1 rows (destabilizers 0..n-1, stabilizers n..2n-1, scratch row 2n)
 * and 2n+1 columns per row (x_0..x_{n-1}, z_0..z_{n-1}, r).  It is stored as
 * shot-sliced bit-planes: one uint64 word holds one tableau entry for 64
 * consecutive shots (bit s of word w is shot 64w+s).  With W = ceil(B/64)
 * the device layout is word[w][col][row], i.e. each 64-shot group owns a
 * contiguous (2n+1) x R block and every column of it is R contiguous words.
 * Host-side I/O uses the reference's flat int32 order
 * x/z[b][row][col], r[b][row] (CliffordState.hpp:238-259).
 *
 * Random measurement outcomes use coins drawn from a counter-based
 * SplitMix64 stream keyed by (seed, measurement sequence number, shot), or
 * coins the caller supplies (replay).  A state's results therefore do not
 * depend on launch shape, batching or the other work on the GPU.
 */
#ifndef PLAQUETTE_AMD_H
#define PLAQUETTE_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
#define PQ_OK 0
#define PQ_ERR_ARG (-1)         /* invalid argument (qubit out of range, bad op) */
#define PQ_ERR_RANGE (-2)       /* record id / shot index out of range */
#define PQ_ERR_HIP (-3)         /* HIP runtime failure */
#define PQ_ERR_OOM (-4)         /* device allocation failed or budget exceeded */
#define PQ_ERR_NODEV (-5)       /* no usable GPU */
#define PQ_ERR_UNSUPPORTED (-6) /* size beyond what a kernel supports */
#define PQ_ERR_BUSY (-7)        /* serving: no capacity right now */

/* ---- op codes (one pq_op per circuit operation) ---- */
#define PQ_OP_H 1       /* ApplyHadamardGate          CliffordState.hpp:389 */
#define PQ_OP_S 2       /* ApplyPhaseGate             CliffordState.hpp:467 */
#define PQ_OP_X 3       /* ApplyPauliXGate            CliffordState.hpp:488 */
#define PQ_OP_Y 4       /* ApplyPauliYGate            CliffordState.hpp:508 */
#define PQ_OP_Z 5       /* ApplyPauliZGate            CliffordState.hpp:529 */
#define PQ_OP_CX 6      /* ApplyControlNotGate(q0=control, q1=target) :412 */
#define PQ_OP_CZ 7      /* ApplyControlPhaseGate(q0, q1)              :442 */
#define PQ_OP_H_FLAG 8  /* ApplyHadamardGateWithApplyFlag, flag = record slot :348 */
#define PQ_OP_S_FLAG 9  /* ApplyPhaseGateWithApplyFlag, flag = record slot    :365 */
#define PQ_OP_NOISE 10  /* per-shot Pauli channel on q0 (no reference op; see below) */
#define PQ_OP_X_FLAG 11 /* X on shots whose record `slot` result is 1 (reset) */
#define PQ_OP_Y_FLAG 12 /* Y on those shots */
#define PQ_OP_Z_FLAG 13 /* Z on those shots */
#define PQ_OP_M 14      /* MeasureQubit(q0) into record `slot`       :587 */
#define PQ_OP_CX_FLAG 15 /* CX(q0 control, q1 target) on flagged shots (no reference op) */
#define PQ_OP_CZ_FLAG 16 /* CZ(q0, q1) on flagged shots (no reference op) */

/* `slot` of a flagged op whose flag is a per-shot Bernoulli draw instead of
 * a record: the op takes the state's next noise number k, and shot b is
 * flagged iff (draw(seed, 2k+1, b) >> 40) < t0, t0 = pq_coin_threshold(p).
 * With H..CZ this is the per-shot form of the reference's WithProb functors
 * (GateFunctors.hpp:38-75, 96-124, 172-202, 256-288, 330-357, 379-406),
 * which draw once per (shot, row) and so do not apply a gate per shot. */
#define PQ_FLAG_RANDOM 0xFFFFFFFFu

/*
 * One circuit operation (32 bytes).
 *  H..CZ:      q0 (control of CX/CZ), q1 (target of CX/CZ)
 *  *_FLAG:     q0 (q1: target of CX_FLAG/CZ_FLAG), slot = record whose
 *              result bit is the per-shot flag, or PQ_FLAG_RANDOM with
 *              t0 = flag probability threshold (see above)
 *  NOISE:      q0; each shot draws u = draw(seed, 2k+1, shot) >> 32 for the
 *              state's k-th noise op and gets X if u < t0, Y if u < t1,
 *              Z if u < t2 (t0 <= t1 <= t2 are cumulative 32-bit thresholds)
 *  M:          q0, slot = record id to write (existing or the next new id),
 *              t0 = coin threshold: the state's m-th measurement gives a
 *              random shot outcome 1 iff (draw(seed, 2m, shot) >> 40) < t0
 *              (pq_coin_threshold(bias) reproduces frand(0,1) < bias)
 */
typedef struct pq_op {
  uint32_t code, q0, q1, slot, t0, t1, t2, reserved;
} pq_op;

/* ---- circuit text front end ----
 * The line format of CliffordCircuit (CliffordCircuit.hpp:32-66): each line
 * is a name (H S X Y Z CX CZ M) and qubit indices; H..Z give one op per
 * qubit, CX/CZ one op per (control, target) pair, M one measurement per
 * qubit (bias 1/2, MeasureQubits :543-549) into records first_record,
 * first_record+1, ...  Any other name (a blank line's included) fails with
 * "Unsupported operation: <name>" as the reference throws.  Host only.
 * ops may be NULL (or capacity too small) to count: *nops = ops produced,
 * *nlines = lines read (GetNumOperations, :127-129), *qubits_used = 1 + the
 * largest index named.  Indices are checked against a state when the ops
 * are run. */
int pq_circuit_parse(const char *text, uint64_t len, uint32_t first_record, pq_op *ops,
                     uint64_t capacity, uint64_t *nops, uint64_t *nlines,
                     uint64_t *qubits_used);

/* Execution forms of pq_run_program. */
#define PQ_RUN_PLAIN 0      /* one kernel launch per op (reference style) */
#define PQ_RUN_PERSISTENT 1 /* one persistent kernel for the whole op list */
#define PQ_RUN_GRAPH 2      /* the persistent form captured once into a hipGraph and replayed */
/* X/Z simulated once on the host, only the per-shot phase column on the GPU.
 * For programs without flagged H/S/CX/CZ (so X/Z stay the same in every
 * shot), run from the |0...0> tableau (a fresh or reset state); results are
 * bit-identical to the other forms.  PQ_ERR_UNSUPPORTED otherwise. */
#define PQ_RUN_FRAME 3

typedef struct pq_state_s *pq_state;

/* ---- library / device ---- */
const char *pq_last_error(void);
/* GetKokkosBackend (CliffordState.hpp:280-282): e.g. "HIP(gfx950)". */
const char *pq_backend_name(void);
int pq_device_count(int *count);
/* Coin threshold for a bias: ceil(bias * 2^24) clamped to [0, 2^24]. */
uint32_t pq_coin_threshold(float bias);
/* SplitMix64 counter draw shared by coins and noise (for tests and tools). */
uint64_t pq_draw(uint64_t seed, uint64_t stream, uint64_t idx);

/* ---- state lifetime ---- */
/* CliffordState(num_qubits, batch_size, seed, init_to_zero)
 * (CliffordState.hpp:97-139).  seed < 0 selects 213434232223 (:135). */
int pq_create(uint64_t num_qubits, uint64_t batch_size, int64_t seed,
              int init_to_zero, pq_state *out);
/* A sampling state: no reference counterpart.  It keeps only the phase
 * column r per 64-shot group (W*(2n+1) words instead of W*(2n+1)^2) and runs
 * frame-form programs (PQ_RUN_FRAME) from |0...0>, in single runs and in
 * batches; records, pq_info, pq_reset and pq_destroy work as for any state,
 * and every other entry point returns PQ_ERR_UNSUPPORTED.  For sampling many
 * shots of circuits without flagged H/S/CX/CZ: d=25 surface-code memory with
 * 1M shots needs 0.33 GB here against 820 GB as a full tableau. */
int pq_create_sampling(uint64_t num_qubits, uint64_t batch_size, int64_t seed, pq_state *out);
/* CliffordState(x, z, r, n, B, seed) from flat host vectors (:153-159). */
int pq_create_from_host(const int32_t *x, const int32_t *z, const int32_t *r,
                        uint64_t num_qubits, uint64_t batch_size, int64_t seed,
                        pq_state *out);
/* Copy constructor (:168-174): copies the tableau and seed. */
int pq_clone(pq_state src, pq_state *out);
/* ~CliffordState (:665-688). */
int pq_destroy(pq_state s);
/* GetNumQubits / GetBatchSize / GetSeed (:289-303) plus the record count. */
int pq_info(pq_state s, uint64_t *num_qubits, uint64_t *batch_size,
            int64_t *seed, uint64_t *num_records);
/* ResetCliffordState (:309-316): re-initialise to |0...0> (the intended
 * semantics; the reference's {B,R,R} extent writes out of bounds).  The R
 * column is written at once; X/Z are written by the next call that reads or
 * updates the tableau (a frame-form run's own X/Z likewise).  After
 * pq_device_handles has handed out the tableau pointer, resets write the
 * whole tableau at once. */
int pq_reset(pq_state s);

/* ---- gates (plain form: one launch each) ---- */
/* Apply{Hadamard,Phase,PauliX,PauliY,PauliZ,ControlNot,ControlPhase}Gate
 * (:379-531; BatchXxxGateFunctor in GateFunctors.hpp). */
int pq_gate(pq_state s, uint32_t code, uint64_t q0, uint64_t q1);
/* Apply{Hadamard,Phase}GateWithApplyFlag (:348-372) and conditional
 * Paulis, with the flag taken from record `rec`'s result bits. */
int pq_gate_flag_record(pq_state s, uint32_t code, uint64_t q, uint64_t rec);
/* Same with a host flag vector int32[B] (nonzero = apply). */
int pq_gate_flag_host(pq_state s, uint32_t code, uint64_t q,
                      const int32_t *flags);
/* Flagged CX/CZ (PQ_OP_CX_FLAG, PQ_OP_CZ_FLAG: q0 control, q1 target) with
 * the flag from record `rec` or from a host int32[B] vector.  No reference
 * counterpart: the reference flags H and S only (CliffordState.hpp:348-372);
 * these extend that API to the two-qubit gates. */
int pq_gate2_flag_record(pq_state s, uint32_t code, uint64_t q0, uint64_t q1, uint64_t rec);
int pq_gate2_flag_host(pq_state s, uint32_t code, uint64_t q0, uint64_t q1,
                       const int32_t *flags);
/* Gate `code` (H..CZ) on each shot independently with probability `prob`:
 * the per-shot form of the WithProb functors (GateFunctors.hpp:38-406,
 * e.g. BatchHadamardGateWithProbFunctor :96-124), which are not reachable
 * from CliffordState and draw per (shot, row).  Takes the state's next
 * noise number k; shot b applies the gate iff (draw(seed, 2k+1, b) >> 40) <
 * pq_coin_threshold(prob), the reference's frand(0,1) < prob. */
int pq_gate_prob(pq_state s, uint32_t code, uint64_t q0, uint64_t q1, float prob);
/* Per-shot Pauli channel (PQ_OP_NOISE semantics); no reference counterpart:
 * the reference reaches per-shot Paulis only through the flagged H/S API. */
int pq_noise(pq_state s, uint64_t q, uint32_t t0, uint32_t t1, uint32_t t2);

/* ---- measurement ---- */
/* MeasureQubit(q, bias, measurement_id) (:587-623; BatchMeasureFunctor,
 * MeasurementFunctors.hpp:267-318).  rec < 0 appends a new record, else
 * overwrites record rec (PQ_ERR_RANGE if rec >= count, as :609-610).
 * coins: optional int32[B] coin per shot (replay); NULL draws them.
 * *out_rec receives the record id. */
int pq_measure(pq_state s, uint64_t q, float bias, int64_t rec,
               const int32_t *coins, uint64_t *out_rec);
/* GetMeasurement(id, b) (:635-648): PQ_ERR_RANGE when id or b is out of
 * range (the reference's nullopt). */
int pq_get_measurement(pq_state s, uint64_t rec, uint64_t b, int32_t *result,
                       int32_t *determined);
/* Whole record as int32[B] vectors (Measurement::CopyToHost :39-57). */
int pq_get_record(pq_state s, uint64_t rec, int32_t *results,
                  int32_t *determined);

/* ---- state I/O ---- */
/* DeviceToHost (:238-259): int32 x[B*R*n], z[B*R*n], r[B*R]. */
int pq_device_to_host(pq_state s, int32_t *x, int32_t *z, int32_t *r);
/* HostToDevice (:207-231).  Entries are read as (v & 1). */
int pq_host_to_device(pq_state s, const int32_t *x, const int32_t *z,
                      const int32_t *r);
/* DeviceToDevice (:268-274): dst's tableau := src's (same n, B). */
int pq_copy_from(pq_state dst, pq_state src);
/* No reference counterpart (tests, tools): the number of tableau words and of
 * record words (over the records both states hold, plus one if the record
 * counts differ) in which two states of equal n and B differ, counted on the
 * device after both states' queued work. */
int pq_compare(pq_state a, pq_state b, uint64_t *planes_differ, uint64_t *records_differ);
/* CheckSum(b0, b1) (:327-338) with a 64-bit total; b1 < 0 means B. */
int pq_checksum(pq_state s, int64_t b0, int64_t b1, int64_t *out);
/* Raw bit-plane tableau, W*(2n+1)*R words in the device layout. */
int pq_read_planes(pq_state s, uint64_t *host_words);
/* Device handles (GetX/GetZ/GetR :180-198 return Kokkos Views; this gives
 * the bit-plane base, record arena base, W, R and the state's hipStream_t).
 * Asking for the tableau pointer makes later resets write it in full. */
int pq_device_handles(pq_state s, void **tableau, void **records,
                      uint64_t *words, uint64_t *rows, void **stream);

/* ---- programs (whole op lists) ----
 * No single reference counterpart: the reference runs op lists through
 * CliffordCircuit::Execute (CliffordCircuit.hpp:131-135), one kernel launch
 * per op.  A program is validated and compiled once for a state's shape and
 * can be run many times; each run continues the state's coin and noise
 * sequence numbers exactly as the same ops issued one by one would.
 * M ops write record `slot` <= current record count (== count appends a new
 * record); flagged ops read records that exist at that point.  A program's
 * device buffers (its noise planes, the frame form's side buffer) serve all
 * of its runs: runs of one program on two states of the same shape must not
 * overlap (pq_synchronize the first state before running it on the other);
 * runs on one state are ordered by that state's stream. */
typedef struct pq_program_s *pq_program;
int pq_program_create(pq_state s, const pq_op *ops, uint64_t nops,
                      pq_program *out);
int pq_program_run(pq_state s, pq_program p, int form);
int pq_program_destroy(pq_program p);
/* Logical gates (H..CZ), measurements, noise ops, device instructions. */
int pq_program_info(pq_program p, uint64_t *gates, uint64_t *measurements,
                    uint64_t *noise_ops, uint64_t *instructions);
/* Compiled layout: segments (runs of gates or of measurements) and the gate
 * segments that run row-tiled out of LDS when the launch's LDS allows. */
int pq_program_layout(pq_program p, uint64_t *segments, uint64_t *tiled_segments);
/* Most distinct qubits per gate segment for programs created after the call
 * (process-wide): a segment over more qubits is cut into windows of at most
 * this many, each run row-tiled out of LDS.  0 (the default): segments wider
 * than a 32-row tile holds in the default LDS budget are cut into windows a
 * 48-row tile holds; UINT32_MAX: never cut.  No reference counterpart: the
 * compiler's window size, for tuning and tests. */
int pq_set_window_qubits(uint32_t qubits);
/* Frame form plan of the op list `ops` on n qubits, as PQ_RUN_FRAME compiles
 * it (host only; PQ_ERR_UNSUPPORTED if the ops hold flagged H/S/CX/CZ).
 * info[16] = {stream words, chunks,
 * preload pairs, words per X/Z bit row, flag records, LDS word offsets of
 * the records, the all-ones word (the zero word follows it) and the noise
 * window, noise-window words,
 * stream-buffer words, LDS bytes per word group, random measurements,
 * determined measurements, final X/Z is |0...0>, staging buffers, the chunk
 * where a run of one round of word groups splits in two (0: no split; the
 * chunks before it read only noise ops below the largest k0 + nk among
 * them)}; non-NULL buffers
 * receive the op stream, the chunk table ({begin, words, k0, nk} each), the
 * {slot, index} preload pairs and the final shared X/Z bit rows
 * ([2n+1][words per row] each).  The staging sizes are those of a launch
 * of many word groups.  No reference counterpart (tests, tools). */
int pq_frame_plan(uint64_t n, const pq_op *ops, uint64_t nops, uint64_t info[16],
                  uint32_t *stream, uint32_t *chunks, uint32_t *preload, uint64_t *x,
                  uint64_t *z);
/* Launch shape of the persistent kernel (0 keeps the current value; threads
 * a multiple of 64, at most 768, or 512 for states of 193..512 qubits and for
 * states above 512 qubits whose program has no row-tiled gate segment). */
int pq_program_set_launch(pq_program p, uint32_t threads, uint32_t blocks);
/* Workgroups per 64-shot word group in persistent runs: 0 chooses (states of
 * 512+ qubits with at most half as many word groups as the device has
 * compute units split each group's rows between two workgroups of one
 * cooperative launch),
 * 1 keeps one workgroup per group, k > 1 asks for k (PQ_ERR_ARG at the run
 * if they do not fit the device at once).  Graph-form runs and server runs
 * use one.  No reference counterpart (launch tuning, tests). */
int pq_program_set_slices(pq_program p, uint32_t slices);
/* How many times a barrier between slices polls before it gives up (0: the
 * default, 2^22).  A small value makes a slice that waits for a slower
 * partner give up: the run still finishes, every later barrier of it passes
 * at once and pq_program_slice_status reports the failure.  For tests of
 * that path; no reference counterpart. */
int pq_program_set_slice_wait(pq_program p, uint32_t polls);
/* Slices the last persistent run used, and whether any of the slices'
 * barriers gave up waiting (nonzero: the results of that run are not to be
 * trusted; it stays set).  Waits for the state's stream.  No reference
 * counterpart. */
int pq_program_slice_status(pq_state s, pq_program p, uint32_t *slices, uint32_t *failed);
/* Work census of persistent/graph runs (enable zeroes the counters):
 * out[0..3] = {random measurements, determined measurements, rows rowsummed
 * by random measurements, rowsum steps of determined measurements}, each
 * summed over 64-shot word groups (the algorithmic-byte accounting);
 * out[4..7] = wall-clock ticks (100 MHz) spent per workgroup, summed over
 * workgroups, in {gate segments, pivot checks, determined batches, other
 * measurements}. */
int pq_program_set_census(pq_program p, int enable);
int pq_program_read_census(pq_state s, pq_program p, uint64_t out[8]);
/* HIP-event timing of the persistent kernel launches of PQ_RUN_PERSISTENT
 * runs on the state's stream; pq_program_kernel_ms sums and clears them. */
int pq_program_set_timing(pq_program p, int enable);
int pq_program_kernel_ms(pq_program p, float *total_ms, uint64_t *launches);
/* Create, run once (and wait), destroy. */
int pq_run_program(pq_state s, const pq_op *ops, uint64_t nops, int form);
/* No reference counterpart.  For callers that cannot allocate device memory
 * (tight serving loops, harnesses that forbid hipMalloc): pq_run_program(s,
 * ops, nops, PQ_RUN_PERSISTENT) with the same results, but no hipMalloc,
 * hipFree or other allocation.  The compiled program, its noise planes
 * (noise ops x 2 x W words) and the run's parameters live in one library-wide
 * device array of PQ_STATIC_SCRATCH_BYTES per GPU, held until the process
 * ends; a program needing more fails with PQ_ERR_UNSUPPORTED (d=9 surface
 * code at 16384 shots needs 25 MB).  The state's record arena must already
 * hold every record the ops write (records made by earlier measurements or
 * runs), else PQ_ERR_OOM.  Returns after the run has finished.  Every call on
 * one GPU shares the array: calls that overlap on one GPU (from several host
 * threads) race and must not be made. */
#define PQ_STATIC_SCRATCH_BYTES (64ull << 20)
int pq_run_program_static(pq_state s, const pq_op *ops, uint64_t nops);
/* Wait for all work queued on the state. */
int pq_synchronize(pq_state s);

/* ---- mixed batches ----
 * No reference counterpart (the reference runs one state at a time).  A batch
 * holds (state, program) members of any sizes and runs all of them in one
 * persistent kernel over a packed work list of (member, 64-shot word group)
 * items, heaviest first, pulled from a global counter; PQ_RUN_GRAPH captures
 * the whole launch once and replays it.  Each run takes and advances every
 * member's coin/noise sequence numbers as a standalone pq_program_run would,
 * so every member's results are bit-identical alone and in the batch.
 * Members must not be used until pq_batch_wait returns. */
typedef struct pq_batch_s *pq_batch;
int pq_batch_create(pq_batch *out);
int pq_batch_destroy(pq_batch b);
int pq_batch_add(pq_batch b, pq_state s, pq_program p);
int pq_batch_run(pq_batch b, int form);
int pq_batch_wait(pq_batch b);
/* Launch shape: threads per workgroup (multiple of 64, <= 768; above 512 the
 * determined chains of members of 193..512 qubits keep their scratch row in
 * LDS instead of registers), workgroups (0 = CUs x workgroups per CU) and the
 * dynamic-LDS budget per workgroup (40 KB..158 KB; it sets workgroups per CU
 * and parallel determined waves). */
int pq_batch_set_launch(pq_batch b, uint32_t threads, uint32_t blocks,
                        uint64_t lds_budget);
/* Times candidate launch shapes on the batch and keeps the fastest, then the
 * word tile at that shape (cached per batch shape).  Members are reset to |0...0> afterwards with their
 * sequence numbers restored. */
int pq_batch_autotune(pq_batch b, int reps, uint32_t *threads, uint64_t *lds_budget,
                      float *best_ms);
int pq_batch_info(pq_batch b, uint64_t *members, uint64_t *items, uint32_t *threads,
                  uint32_t *blocks, uint64_t *lds_budget);
/* Word groups per work item of persistent/graph batch runs (1..64; an item
 * then covers that many consecutive 64-shot groups of one member, fewer at
 * its end), and its current value.  The autotuner also picks it.  Launch
 * tuning; no reference counterpart. */
int pq_batch_set_word_tile(pq_batch b, uint32_t words);
int pq_batch_word_tile(pq_batch b, uint32_t *words);
/* HIP-event time of the last run (enable first). */
int pq_batch_set_timing(pq_batch b, int enable);
int pq_batch_last_ms(pq_batch b, float *ms);

/* ---- serving ----
 * No reference counterpart.  A server keeps requests' tableaux resident in
 * HBM and runs them concurrently: a worker thread admits queued requests
 * first-come first-served whenever one of `slots` streams is free and the
 * request's device bytes (tableau + records + noise planes + program) fit the
 * budget, runs each as a persistent program on its stream, and on completion
 * copies the result planes to the host and releases the device memory
 * stream-ordered.  A request's results depend only on (n, B, seed, ops). */
typedef struct pq_server_s *pq_server;
int pq_server_create(uint64_t budget_bytes, uint32_t slots, pq_server *out);
/* Queue a request; PQ_ERR_OOM if it can never fit the budget. */
int pq_server_submit(pq_server sv, uint64_t num_qubits, uint64_t batch_size,
                     int64_t seed, const pq_op *ops, uint64_t nops, uint64_t *ticket);
/* status: 0 queued, 1 running, 2 done, < 0 failed (PQ_ERR_*); latency_s =
 * submit -> results on host (-1 until finished); queue_s = submit -> start. */
int pq_server_status(pq_server sv, uint64_t ticket, int *status, double *latency_s,
                     double *queue_s);
/* Result planes of a finished request: words[nrec][2][W] (result bits,
 * determined bits; bit s of word w = shot 64w+s).  words may be NULL to query
 * nrec and W.  PQ_ERR_BUSY until finished. */
int pq_server_result(pq_server sv, uint64_t ticket, uint64_t *words, uint64_t capacity,
                     uint64_t *nrec, uint64_t *words_per_plane);
int pq_server_release(pq_server sv, uint64_t ticket);
/* Block until every submitted request has finished. */
int pq_server_drain(pq_server sv);
int pq_server_stats(pq_server sv, uint64_t *completed, uint64_t *peak_bytes,
                    uint32_t *max_running, uint64_t *queued);
int pq_server_destroy(pq_server sv);
/* Admission order.  PQ_SERVE_FCFS (the default): arrival order, a request
 * that does not fit the free budget holds back the ones behind it.
 * PQ_SERVE_SMALLEST_FIRST: the queued request with the least estimated work
 * (word groups x rows x ops) that fits goes first, except that a request
 * queued longer than age_limit_s (> 0) goes before all younger ones, so large
 * requests are not starved under sustained load. */
#define PQ_SERVE_FCFS 0
#define PQ_SERVE_SMALLEST_FIRST 1
int pq_server_set_policy(pq_server sv, int policy, double age_limit_s);
/* Execution form of admitted requests: PQ_RUN_PERSISTENT (default) or
 * PQ_RUN_FRAME, which runs each request whose ops allow it (no flagged
 * H/S/CX/CZ) in the frame form and the others persistent; frame plans are
 * shared by requests of the same structure (n and ops, noise thresholds
 * aside).  Results are the same either way. */
int pq_server_set_form(pq_server sv, int form);
/* Device bytes a request of these ops would reserve against the budget under
 * the server's current form (host only). */
int pq_server_request_bytes(pq_server sv, uint64_t num_qubits, uint64_t batch_size,
                            const pq_op *ops, uint64_t nops, uint64_t *bytes);
/* Requests run in the frame form, frame plans built, plans reused. */
int pq_server_frame_stats(pq_server sv, uint64_t *frame_runs, uint64_t *plan_builds,
                          uint64_t *plan_hits);

/* ---- diagnostics: the reference's test-only functors ---- */
/* RowMultiply(i <- i*k) on every shot (BatchRowMultiplyFunctor,
 * MeasurementFunctors.hpp:168-187). */
int pq_row_multiply(pq_state s, uint64_t i, uint64_t k);
/* RowProductSign(i, k) per shot into out[B], tableau unchanged
 * (BatchRowProductSignFunctor, MeasurementFunctors.hpp:131-154). */
int pq_row_product_sign(pq_state s, uint64_t i, uint64_t k, int32_t *out);
/* MeasureDetermined(q) on every shot into out[B], regardless of pivots
 * (BatchMeasureDeterminedFunctor, MeasurementFunctors.hpp:344-364). */
int pq_measure_determined(pq_state s, uint64_t q, int32_t *out);

/* ---- timing on the state's stream (HIP events) ---- */
int pq_event_record(pq_state s, int slot); /* slot in [0, 8) */
int pq_event_elapsed_ms(pq_state s, int slot_a, int slot_b, float *ms);

#ifdef __cplusplus
}
#endif
#endif
