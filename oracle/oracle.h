/*
 * CPU oracle for the batched stabilizer-tableau engine.
 *
 * TEST INFRASTRUCTURE ONLY.  This file and oracle.c are a from-scratch,
 * plain-C restatement of the reference's CliffordState algorithm, kept in
 * the reference's own data layout (one int32 per tableau entry,
 * x/z: [B][2n+1][n], r: [B][2n+1], row-major).  Only tests, the smoke check
 * and the CPU-baseline leg of the benchmark may load it, and only as the
 * checker (or as the timed CPU baseline) -- never as the product path.
 *
 * Reference files restated (semantics only, no code copied):
 *   plaquette_simulator/src/MiscFunctors.hpp:8-50        (init, checksum)
 *   plaquette_simulator/src/GateFunctors.hpp:13-36,76-94,126-170,204-254,
 *                                           290-328,359-377 (gate rules)
 *   plaquette_simulator/src/MeasurementFunctors.hpp:10-318 (g, rowsum,
 *                                           determined/random measurement)
 *   plaquette_simulator/src/CliffordState.hpp:97-139,327-338,587-623
 *
 * Where the reference draws a Kokkos XorShift64 coin (MeasurementFunctors.hpp
 * :245-251) the oracle takes the coin bit from the caller, so both this
 * oracle and the GPU engine consume the same recorded coins.
 */
#ifndef PQ_ORACLE_H
#define PQ_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Pauli-product phase exponent g(x1,z1,x2,z2) in {-1,0,+1};
 * MeasurementFunctors.hpp:10-72 (destination Pauli first, source second). */
int32_t orc_g(int32_t x1, int32_t z1, int32_t x2, int32_t z2);

/* Identity tableau with the scratch row's r set to 1; MiscFunctors.hpp:8-28. */
void orc_init_zero(int64_t n, int64_t B, int32_t *x, int32_t *z, int32_t *r);

/* Sum of every x, z and r entry for shots [b0, b1); MiscFunctors.hpp:30-50,
 * CliffordState.hpp:327-338 (64-bit accumulator, no overflow). */
int64_t orc_checksum(int64_t n, int64_t B, const int32_t *x, const int32_t *z,
                     const int32_t *r, int64_t b0, int64_t b1);

/* Gate codes shared with the program format (see orc_run_program). */
enum {
  ORC_H = 1, ORC_S = 2, ORC_X = 3, ORC_Y = 4, ORC_Z = 5, ORC_CX = 6, ORC_CZ = 7,
  ORC_H_FLAG = 8, ORC_S_FLAG = 9, ORC_NOISE = 10, ORC_X_FLAG = 11,
  ORC_Y_FLAG = 12, ORC_Z_FLAG = 13, ORC_M = 14, ORC_CX_FLAG = 15,
  ORC_CZ_FLAG = 16
};
/* Slot value of a flagged op whose flag is a per-shot Bernoulli draw. */
#define ORC_FLAG_RANDOM (-1)

/* One unconditional gate over every shot and all 2n+1 rows
 * (GateFunctors.hpp: CZ :29-35, H :89-93, S :151-170, CX :247-253,
 *  X :303-307, Y :323-327, Z :372-376).  q1 is the target of CX/CZ. */
void orc_gate(int64_t n, int64_t B, int32_t *x, int32_t *z, int32_t *r,
              int code, int64_t q0, int64_t q1);

/* Shot-conditional gates: the H/S rule (GateFunctors.hpp:142-148, 220-228)
 * where flag[b] != 0; code is ORC_H_FLAG or ORC_S_FLAG. */
void orc_gate_flag(int64_t n, int64_t B, int32_t *x, int32_t *z, int32_t *r,
                   int code, int64_t q, const int32_t *flag);

/* Shot-conditional two-qubit gates: the CX (code ORC_CX_FLAG) or CZ
 * (ORC_CZ_FLAG) rule on q0 (control), q1 (target) where flag[b] != 0.  The
 * reference has no flagged CX/CZ; its per-draw variants are the WithProb
 * functors (GateFunctors.hpp:38-75, 256-288).  Each flagged shot is the
 * unflagged rule of orc_gate applied to that shot alone. */
void orc_gate2_flag(int64_t n, int64_t B, int32_t *x, int32_t *z, int32_t *r,
                    int code, int64_t q0, int64_t q1, const int32_t *flag);

/* Per-shot Pauli: shot b gets X^fx[b] Z^fz[b] on qubit q, i.e. the
 * reference's X rule (r ^= z_q) where fx[b] and Z rule (r ^= x_q) where
 * fz[b].  fz may be NULL (no Z part) and fx may be NULL (no X part). */
void orc_pauli_flag(int64_t n, int64_t B, int32_t *x, int32_t *z, int32_t *r,
                    int64_t q, const int32_t *fx, const int32_t *fz);

/* RowProductSign / RowMultiply for one shot; MeasurementFunctors.hpp:97-129,
 * 156-166.  Row i is the destination, row k the source. */
int32_t orc_row_product_sign(int64_t n, int64_t B, const int32_t *x,
                             const int32_t *z, const int32_t *r, int64_t b,
                             int64_t i, int64_t k);
void orc_row_multiply(int64_t n, int64_t B, int32_t *x, int32_t *z, int32_t *r,
                      int64_t b, int64_t i, int64_t k);

/* MeasureDetermined for one shot; MeasurementFunctors.hpp:189-201.
 * Adds the number of rowsums it performed to *kdet when kdet != NULL. */
int32_t orc_measure_determined(int64_t n, int64_t B, int32_t *x, int32_t *z,
                               int32_t *r, int64_t b, int64_t q, int64_t *kdet);

/* MeasureRandom for one shot with pivot p and the given coin bit;
 * MeasurementFunctors.hpp:225-265.  Adds its rowsum count to *krand. */
int32_t orc_measure_random(int64_t n, int64_t B, int32_t *x, int32_t *z,
                           int32_t *r, int64_t b, int64_t q, int64_t p,
                           int32_t coin, int64_t *krand);

/* BatchMeasureFunctor over every shot; MeasurementFunctors.hpp:291-317.
 * coin[b] is the coin the shot uses if its outcome is random.
 * stats (may be NULL) receives {#random shots, #determined shots,
 * total random rowsums, total determined rowsums}. */
void orc_measure(int64_t n, int64_t B, int32_t *x, int32_t *z, int32_t *r,
                 int64_t q, const int32_t *coin, int32_t *result,
                 int32_t *determined, int64_t *stats);

/* ---- counter-based draws (the coin / noise source both sides restate) ---- */

/* SplitMix64 output function (Steele, Lea, Flood, OOPSLA 2014). */
uint64_t orc_mix64(uint64_t z);
/* draw(seed, stream, idx) = mix64(mix64(seed + (stream+1)G) + (idx+1)G),
 * G = 0x9E3779B97F4A7C15: output idx+1 of a SplitMix64 generator whose
 * state is output stream+1 of a SplitMix64 generator seeded with seed. */
uint64_t orc_draw(uint64_t seed, uint64_t stream, uint64_t idx);
/* Coin threshold T = ceil(bias * 2^24) clamped to [0, 2^24]; a coin is 1
 * when (draw >> 40) < T, i.e. u < bias for u = (draw >> 40) * 2^-24 in
 * [0,1), which is the reference's frand(0,1) < bias. */
uint32_t orc_coin_threshold(float bias);

/* ---- program interpreter (the op list both engines execute) ---- */

/* One op = 8 int32 fields: {code, q0, q1, slot, t0, t1, t2, reserved}.
 *  gates H..CZ: q0 (control for CX/CZ), q1 (target)
 *  *_FLAG: q0 (and q1 for CX/CZ_FLAG), slot = record whose result bits
 *         are the per-shot flag, or ORC_FLAG_RANDOM: then the op takes the
 *         state's next noise number k and shot b is flagged iff
 *         (draw(seed, 2k+1, shot) >> 40) < t0 (t0 a coin threshold)
 *  NOISE: q0, t0 < t1 < t2 cumulative 32-bit thresholds for X, Y, Z;
 *         a shot draws u = draw(seed, 2k+1, shot) >> 32 for the k-th noise
 *         op of the state and gets X if u<t0, Y if u<t1, Z if u<t2
 *  M: q0, slot = record id to write, t0 = coin threshold; the m-th
 *         measurement of the state uses coin (draw(seed, 2m, shot)>>40) < t0
 * Records are int32 [nrec][2][B] (result, determined).
 * *meas_seq / *noise_seq are the state's running counters (updated).
 * stats (may be NULL): int64[4] accumulated as in orc_measure. */
int orc_run_program(int64_t n, int64_t B, int32_t *x, int32_t *z, int32_t *r,
                    const int32_t *ops, int64_t nops, uint64_t seed,
                    int64_t shot0, int32_t *records, int64_t nrec,
                    int64_t *meas_seq, int64_t *noise_seq, int64_t *stats);

/* Set the OpenMP thread count used by the oracle (0 = leave as is). */
void orc_set_threads(int nthreads);
int orc_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
