/*
 * CPU oracle for the batched stabilizer-tableau engine -- TEST
 * INFRASTRUCTURE ONLY (the checker and the timed CPU baseline; never the
 * product path).  Plain-C restatement of the reference algorithm in the
 * reference layout: x/z int32 [B][2n+1][n], r int32 [B][2n+1].
 * Every function cites the reference lines it restates.  OpenMP mirrors the
 * reference's Kokkos index spaces: (shot,row) for gates, shot for measurement.
 */
#include "oracle.h"

#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define IDX3(b, i, j) (((b) * R + (i)) * n + (j))
#define IDX2(b, i) ((b) * R + (i))

/* MeasurementFunctors.hpp:10-72: phase exponent of P(x1,z1) * P(x2,z2). */
int32_t orc_g(int32_t x1, int32_t z1, int32_t x2, int32_t z2) {
  if (x1 == 1 && z1 == 1) {            /* Y times I,X,Y,Z */
    if (x2 == 0 && z2 == 0) return 0;
    if (x2 == 1 && z2 == 0) return -1;
    if (x2 == 1 && z2 == 1) return 0;
    return 1;
  }
  if (x1 == 1 && z1 == 0) {            /* X times I,X,Y,Z */
    if (x2 == 0 && z2 == 0) return 0;
    if (x2 == 1 && z2 == 0) return 0;
    if (x2 == 1 && z2 == 1) return 1;
    return -1;
  }
  if (z1 == 1 && x1 == 0) {            /* Z times I,X,Y,Z */
    if (x2 == 0 && z2 == 0) return 0;
    if (x2 == 1 && z2 == 0) return 1;
    if (x2 == 1 && z2 == 1) return -1;
    return 0;
  }
  return 0;                            /* I */
}

/* MiscFunctors.hpp:22-27 over the {B, 2n+1, n} index space
 * (CliffordState.hpp:122-127). */
void orc_init_zero(int64_t n, int64_t B, int32_t *x, int32_t *z, int32_t *r) {
  const int64_t R = 2 * n + 1;
#pragma omp parallel for collapse(2) schedule(static)
  for (int64_t b = 0; b < B; ++b)
    for (int64_t i = 0; i < R; ++i) {
      for (int64_t j = 0; j < n; ++j) {
        x[IDX3(b, i, j)] = (i == j) ? 1 : 0;
        z[IDX3(b, i, j)] = (i - n == j) ? 1 : 0;
      }
      r[IDX2(b, i)] = (i == 2 * n) ? 1 : 0;
    }
}

/* MiscFunctors.hpp:43-49: x + z + r (r counted once per row, at j == 0). */
int64_t orc_checksum(int64_t n, int64_t B, const int32_t *x, const int32_t *z,
                     const int32_t *r, int64_t b0, int64_t b1) {
  const int64_t R = 2 * n + 1;
  if (b1 < 0) b1 = B;
  int64_t s = 0;
#pragma omp parallel for reduction(+ : s) schedule(static)
  for (int64_t b = b0; b < b1; ++b)
    for (int64_t i = 0; i < R; ++i)
      for (int64_t j = 0; j < n; ++j) {
        s += x[IDX3(b, i, j)];
        s += z[IDX3(b, i, j)];
        if (j == 0) s += r[IDX2(b, i)];
      }
  return s;
}

/* The per-(shot,row) gate rules of GateFunctors.hpp. */
void orc_gate(int64_t n, int64_t B, int32_t *x, int32_t *z, int32_t *r,
              int code, int64_t q0, int64_t q1) {
  const int64_t R = 2 * n + 1;
#pragma omp parallel for collapse(2) schedule(static)
  for (int64_t b = 0; b < B; ++b)
    for (int64_t i = 0; i < R; ++i) {
      int32_t *xr = x + IDX3(b, i, 0);
      int32_t *zr = z + IDX3(b, i, 0);
      int32_t *rr = r + IDX2(b, i);
      switch (code) {
        case ORC_H: { /* GateFunctors.hpp:90-93 */
          *rr ^= xr[q0] & zr[q0];
          int32_t t = xr[q0]; xr[q0] = zr[q0]; zr[q0] = t;
          break;
        }
        case ORC_S: /* GateFunctors.hpp:151-170 (operator body) */
          *rr ^= xr[q0] & zr[q0];
          zr[q0] ^= xr[q0];
          break;
        case ORC_CX: /* GateFunctors.hpp:248-252; q0 control, q1 target */
          *rr ^= xr[q0] & zr[q1] & (xr[q1] ^ zr[q0] ^ 1);
          xr[q1] ^= xr[q0];
          zr[q0] ^= zr[q1];
          break;
        case ORC_CZ: /* GateFunctors.hpp:30-35 */
          *rr ^= xr[q0] * xr[q1] * (zr[q1] ^ zr[q0]);
          zr[q1] ^= xr[q0];
          zr[q0] ^= xr[q1];
          break;
        case ORC_X: *rr ^= zr[q0]; break;            /* :304-307 */
        case ORC_Y: *rr ^= zr[q0] ^ xr[q0]; break;   /* :324-327 */
        case ORC_Z: *rr ^= xr[q0]; break;            /* :373-376 */
        default: break;
      }
    }
}

/* GateFunctors.hpp:142-148 (H) and :220-228 (S), gated by flag[b]. */
void orc_gate_flag(int64_t n, int64_t B, int32_t *x, int32_t *z, int32_t *r,
                   int code, int64_t q, const int32_t *flag) {
  const int64_t R = 2 * n + 1;
#pragma omp parallel for collapse(2) schedule(static)
  for (int64_t b = 0; b < B; ++b)
    for (int64_t i = 0; i < R; ++i) {
      if (!flag[b]) continue;
      int32_t *xr = x + IDX3(b, i, 0);
      int32_t *zr = z + IDX3(b, i, 0);
      int32_t *rr = r + IDX2(b, i);
      if (code == ORC_H_FLAG) {
        *rr ^= xr[q] & zr[q];
        int32_t t = xr[q]; xr[q] = zr[q]; zr[q] = t;
      } else {
        *rr ^= xr[q] & zr[q];
        zr[q] ^= xr[q];
      }
    }
}

/* The unflagged CX/CZ rule on each flagged shot's own [R][n] slice. */
void orc_gate2_flag(int64_t n, int64_t B, int32_t *x, int32_t *z, int32_t *r,
                    int code, int64_t q0, int64_t q1, const int32_t *flag) {
  const int64_t R = 2 * n + 1;
  const int gate = code == ORC_CX_FLAG ? ORC_CX : ORC_CZ;
  for (int64_t b = 0; b < B; ++b)
    if (flag[b])
      orc_gate(n, 1, x + IDX3(b, 0, 0), z + IDX3(b, 0, 0), r + IDX2(b, 0), gate, q0, q1);
}

/* X rule (GateFunctors.hpp:306) where fx[b], Z rule (:375) where fz[b]. */
void orc_pauli_flag(int64_t n, int64_t B, int32_t *x, int32_t *z, int32_t *r,
                    int64_t q, const int32_t *fx, const int32_t *fz) {
  const int64_t R = 2 * n + 1;
#pragma omp parallel for collapse(2) schedule(static)
  for (int64_t b = 0; b < B; ++b)
    for (int64_t i = 0; i < R; ++i) {
      int32_t *rr = r + IDX2(b, i);
      if (fx && fx[b]) *rr ^= z[IDX3(b, i, q)];
      if (fz && fz[b]) *rr ^= x[IDX3(b, i, q)];
    }
}

/* MeasurementFunctors.hpp:98-129: r_i ^ r_k ^ bit 1 of (sum g mod 4). */
int32_t orc_row_product_sign(int64_t n, int64_t B, const int32_t *x,
                             const int32_t *z, const int32_t *r, int64_t b,
                             int64_t i, int64_t k) {
  const int64_t R = 2 * n + 1;
  (void)B;
  int32_t s = 0;
  for (int64_t j = 0; j < n; ++j)
    s += orc_g(x[IDX3(b, i, j)], z[IDX3(b, i, j)], x[IDX3(b, k, j)],
               z[IDX3(b, k, j)]);
  /* (s >> 1) & 1 with an arithmetic shift == bit 1 of s mod 4 */
  int32_t m4 = ((s % 4) + 4) % 4;
  int32_t p = (m4 >> 1) & 1;
  return r[IDX2(b, i)] ^ r[IDX2(b, k)] ^ p;
}

/* MeasurementFunctors.hpp:157-166. */
void orc_row_multiply(int64_t n, int64_t B, int32_t *x, int32_t *z, int32_t *r,
                      int64_t b, int64_t i, int64_t k) {
  const int64_t R = 2 * n + 1;
  r[IDX2(b, i)] = orc_row_product_sign(n, B, x, z, r, b, i, k);
  for (int64_t j = 0; j < n; ++j) {
    x[IDX3(b, i, j)] ^= x[IDX3(b, k, j)];
    z[IDX3(b, i, j)] ^= z[IDX3(b, k, j)];
  }
}

/* MeasurementFunctors.hpp:86-95 */
static void zero_row(int64_t n, int32_t *x, int32_t *z, int32_t *r, int64_t b,
                     int64_t row) {
  const int64_t R = 2 * n + 1;
  for (int64_t j = 0; j < n; ++j) {
    x[IDX3(b, row, j)] = 0;
    z[IDX3(b, row, j)] = 0;
  }
  r[IDX2(b, row)] = 0;
}

/* MeasurementFunctors.hpp:190-201 */
int32_t orc_measure_determined(int64_t n, int64_t B, int32_t *x, int32_t *z,
                               int32_t *r, int64_t b, int64_t q, int64_t *kdet) {
  const int64_t R = 2 * n + 1;
  zero_row(n, x, z, r, b, 2 * n);
  for (int64_t i = 0; i < n; ++i)
    if (x[IDX3(b, i, q)]) {
      orc_row_multiply(n, B, x, z, r, b, 2 * n, i + n);
      if (kdet) ++*kdet;
    }
  return r[IDX2(b, 2 * n)];
}

/* MeasurementFunctors.hpp:226-265 (the coin replaces the frand draw at
 * :245-251). */
int32_t orc_measure_random(int64_t n, int64_t B, int32_t *x, int32_t *z,
                           int32_t *r, int64_t b, int64_t q, int64_t p,
                           int32_t coin, int64_t *krand) {
  const int64_t R = 2 * n + 1;
  /* copy row n+p -> row p (:235, CopyFromRowAtoRowB :74-84) */
  for (int64_t j = 0; j < n; ++j) {
    x[IDX3(b, p, j)] = x[IDX3(b, p + n, j)];
    z[IDX3(b, p, j)] = z[IDX3(b, p + n, j)];
  }
  r[IDX2(b, p)] = r[IDX2(b, p + n)];
  zero_row(n, x, z, r, b, p + n);      /* :238 */
  z[IDX3(b, p + n, q)] = 1;            /* :240 */
  r[IDX2(b, p + n)] = coin ? 1 : 0;    /* :247 */
  for (int64_t i = 0; i < 2 * n; ++i) /* :256-260 */
    if (x[IDX3(b, i, q)] && i != p && i != p + n) {
      orc_row_multiply(n, B, x, z, r, b, i, p);
      if (krand) ++*krand;
    }
  return r[IDX2(b, p + n)];
}

/* MeasurementFunctors.hpp:292-317 over RangePolicy{B}
 * (CliffordState.hpp:613-620). */
void orc_measure(int64_t n, int64_t B, int32_t *x, int32_t *z, int32_t *r,
                 int64_t q, const int32_t *coin, int32_t *result,
                 int32_t *determined, int64_t *stats) {
  const int64_t R = 2 * n + 1;
  int64_t nrand = 0, ndet = 0, krand = 0, kdet = 0;
#pragma omp parallel for reduction(+ : nrand, ndet, krand, kdet) schedule(dynamic, 4)
  for (int64_t b = 0; b < B; ++b) {
    int64_t piv = -1;
    for (int64_t p = 0; p < n; ++p)
      if (x[IDX3(b, p + n, q)]) { piv = p; break; }
    if (piv >= 0) {
      result[b] = orc_measure_random(n, B, x, z, r, b, q, piv, coin[b], &krand);
      determined[b] = 0;
      ++nrand;
    } else {
      result[b] = orc_measure_determined(n, B, x, z, r, b, q, &kdet);
      determined[b] = 1;
      ++ndet;
    }
  }
  if (stats) {
    stats[0] += nrand;
    stats[1] += ndet;
    stats[2] += krand;
    stats[3] += kdet;
  }
}

/* ---- SplitMix64-based counter draws ---- */
#define GOLDEN 0x9E3779B97F4A7C15ULL

uint64_t orc_mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

uint64_t orc_draw(uint64_t seed, uint64_t stream, uint64_t idx) {
  uint64_t key = orc_mix64(seed + (stream + 1) * GOLDEN);
  return orc_mix64(key + (idx + 1) * GOLDEN);
}

uint32_t orc_coin_threshold(float bias) {
  double t = (double)bias * 16777216.0;
  if (!(t > 0.0)) return 0;            /* also catches NaN */
  if (t >= 16777216.0) return 16777216u;
  double c = (double)(int64_t)t;
  if (c < t) c += 1.0;                 /* ceil */
  return (uint32_t)c;
}

int orc_run_program(int64_t n, int64_t B, int32_t *x, int32_t *z, int32_t *r,
                    const int32_t *ops, int64_t nops, uint64_t seed,
                    int64_t shot0, int32_t *records, int64_t nrec,
                    int64_t *meas_seq, int64_t *noise_seq, int64_t *stats) {
  int32_t *a = (int32_t *)malloc(sizeof(int32_t) * (size_t)(B > 0 ? B : 1));
  int32_t *c = (int32_t *)malloc(sizeof(int32_t) * (size_t)(B > 0 ? B : 1));
  if (!a || !c) { free(a); free(c); return -1; }
  int rc = 0;
  for (int64_t k = 0; k < nops && rc == 0; ++k) {
    const int32_t *op = ops + 8 * k;
    int code = op[0];
    int64_t q0 = op[1], q1 = op[2], slot = op[3];
    uint32_t t0 = (uint32_t)op[4], t1 = (uint32_t)op[5], t2 = (uint32_t)op[6];
    const int two = code == ORC_CX || code == ORC_CZ || code == ORC_CX_FLAG ||
                    code == ORC_CZ_FLAG;
    if (q0 < 0 || q0 >= n || (two && (q1 < 0 || q1 >= n || q1 == q0))) {
      rc = -2;
      break;
    }
    switch (code) {
      case ORC_H: case ORC_S: case ORC_X: case ORC_Y: case ORC_Z:
      case ORC_CX: case ORC_CZ:
        orc_gate(n, B, x, z, r, code, q0, q1);
        break;
      case ORC_H_FLAG: case ORC_S_FLAG: case ORC_X_FLAG: case ORC_Y_FLAG:
      case ORC_Z_FLAG: case ORC_CX_FLAG: case ORC_CZ_FLAG: {
        const int32_t *flag;
        if (slot == ORC_FLAG_RANDOM) {
          uint64_t kk = (uint64_t)(*noise_seq)++;
          for (int64_t b = 0; b < B; ++b)
            c[b] = (uint32_t)(orc_draw(seed, 2 * kk + 1, (uint64_t)(shot0 + b)) >> 40) < t0;
          flag = c;
        } else if (slot < 0 || slot >= nrec) {
          rc = -3;
          break;
        } else {
          flag = records + (size_t)slot * 2 * B;
        }
        if (code == ORC_CX_FLAG || code == ORC_CZ_FLAG)
          orc_gate2_flag(n, B, x, z, r, code, q0, q1, flag);
        else if (code == ORC_H_FLAG || code == ORC_S_FLAG)
          orc_gate_flag(n, B, x, z, r, code, q0, flag);
        else
          orc_pauli_flag(n, B, x, z, r, q0,
                         code != ORC_Z_FLAG ? flag : NULL,
                         code != ORC_X_FLAG ? flag : NULL);
        break;
      }
      case ORC_NOISE: {
        uint64_t kk = (uint64_t)(*noise_seq)++;
        for (int64_t b = 0; b < B; ++b) {
          uint32_t u = (uint32_t)(orc_draw(seed, 2 * kk + 1, (uint64_t)(shot0 + b)) >> 32);
          a[b] = u < t1;                  /* X or Y */
          c[b] = (u >= t0) && (u < t2);   /* Y or Z */
        }
        orc_pauli_flag(n, B, x, z, r, q0, a, c);
        break;
      }
      case ORC_M: {
        if (slot < 0 || slot >= nrec) { rc = -3; break; }
        uint64_t m = (uint64_t)(*meas_seq)++;
        for (int64_t b = 0; b < B; ++b)
          a[b] = (uint32_t)(orc_draw(seed, 2 * m, (uint64_t)(shot0 + b)) >> 40) < t0;
        int32_t *res = records + (size_t)slot * 2 * B;
        orc_measure(n, B, x, z, r, q0, a, res, res + B, stats);
        break;
      }
      default:
        rc = -4;
        break;
    }
  }
  free(a);
  free(c);
  return rc;
}

void orc_set_threads(int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
}

int orc_get_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
