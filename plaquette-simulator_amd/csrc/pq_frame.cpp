// Frame form of a program: circuits whose X/Z tableau is the same in every
// shot run with X/Z simulated once, here on the host, and only the per-shot
// phase column R on the GPU.
//
// X and Z change only through H, S, CX, CZ and measurements, and those act on
// every shot alike (GateFunctors.hpp; MeasurementFunctors.hpp:225-265 picks
// the pivot from X alone), so without flagged H/S/CX/CZ a state that starts
// with the same X/Z in every shot keeps it.  Every update of R is then an XOR
// of words whose positions X/Z fix:
//  * a gate, Pauli or noise op XORs into R[i] a constant, a noise mask or a
//    record's result bits, on the rows i its rule selects (the gate rules of
//    GateFunctors.hpp read x_q, z_q of row i);
//  * a random measurement (MeasureRandom, :225-265) sets R[p] = R[n+p],
//    R[n+p] = coin and R[i] ^= R[p] ^ phase_i on the rows i it rowsums, with
//    phase_i a function of X/Z only (RowProductSign, :97-129);
//  * a determined measurement (MeasureDetermined, :189-201) gives
//    R[2n] = c ^ XOR of R[n+i] over its rowsum rows.
// The compiler below runs the shared X/Z once from the |0...0> tableau,
// records those rows, constants and phases, and packs them into a stream the
// GPU kernel runs per 64-shot word group (k_frame in pq_kernels.hip).
//
// Stream layout (uint32 words), one op after another, each with an 8-word
// header {kind, a, b, c, d, e, f, g}:
//  kXor:    a = T rows, b = entries; then rows[T], offsets[T+1] (relative to
//           the entry array), entries: LDS word index of the mask XORed in;
//           ops of one XOR block commute, so a block splits freely.
//  kRandom: a = p, b = record slot, c = measurement index in the program,
//           d = coin threshold, e = rec_l index or kNone, f = entries; then
//           entries row | phase << 31 (rows other than p and n+p).
//  kDet:    a = G measurements (<= 64); then per measurement {slot, rec_l
//           index or kNone, c, row offset}, the end offset, and the row lists
//           (rows n+i).  The last one's result is R[2n].
// The stream is cut into chunks that fit the kernel's LDS staging buffer;
// each chunk also names the noise numbers [k0, k0+nk) whose masks it reads.
// With two staging buffers (when they fit the LDS) chunk c uses buffer c % 2
// and the kernel loads chunk c+1 while it runs chunk c; a chunk's noise
// entries name the mask words of its own buffer.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "plaquette_amd.h"
#include "pq_capi_internal.h"
#include "pq_common.h"
#include "pq_frame.h"
#include "pq_kernels.h"
#include "pq_program_internal.h"

using namespace pq;
using pq::kFrameXor;

#define HIPCHK(expr)                                                              \
  do {                                                                            \
    hipError_t e_ = (expr);                                                       \
    if (e_ != hipSuccess)                                                         \
      return pq_set_error(PQ_ERR_HIP,                                             \
                          (std::string(#expr) + ": " + hipGetErrorString(e_)).c_str()); \
  } while (0)

namespace {

// Shared X/Z of a state: rows 0..2n, bit-packed columns.
struct XZ {
  uint32_t n = 0, R = 0, nw = 0;
  std::vector<uint64_t> X, Z;
  void init_identity(uint32_t nq) {
    n = nq;
    R = 2 * nq + 1;
    nw = (nq + 63) / 64;
    X.assign((size_t)R * nw, 0);
    Z.assign((size_t)R * nw, 0);
    for (uint32_t i = 0; i < nq; ++i) {
      set(X, i, i, true);       // destabilizer i = X_i
      set(Z, nq + i, i, true);  // stabilizer i = Z_i
    }
  }
  bool get(const std::vector<uint64_t> &M, uint32_t row, uint32_t q) const {
    return (M[(size_t)row * nw + q / 64] >> (q % 64)) & 1;
  }
  void set(std::vector<uint64_t> &M, uint32_t row, uint32_t q, bool v) {
    uint64_t &w = M[(size_t)row * nw + q / 64];
    const uint64_t b = 1ull << (q % 64);
    w = v ? (w | b) : (w & ~b);
  }
  bool x(uint32_t row, uint32_t q) const { return get(X, row, q); }
  bool z(uint32_t row, uint32_t q) const { return get(Z, row, q); }
  // Bit 1 of sum_j g(row i, row k) mod 4 (RowProductSign, :97-129; the
  // destination row first), with g's sign as in phase_acc.
  bool phase(uint32_t i, uint32_t k) const {
    int64_t s = 0;
    for (uint32_t j = 0; j < nw; ++j) {
      const uint64_t x1 = X[(size_t)i * nw + j], z1 = Z[(size_t)i * nw + j];
      const uint64_t x2 = X[(size_t)k * nw + j], z2 = Z[(size_t)k * nw + j];
      const uint64_t anti = (x1 & z2) ^ (z1 & x2);
      const uint64_t minus = anti & ~(z1 ^ x2) & (x1 | z2);
      s += (int64_t)__builtin_popcountll(anti & ~minus) - (int64_t)__builtin_popcountll(minus);
    }
    return ((((s % 4) + 4) % 4) >> 1) & 1;
  }
  void row_xor(uint32_t i, uint32_t k) {
    for (uint32_t j = 0; j < nw; ++j) {
      X[(size_t)i * nw + j] ^= X[(size_t)k * nw + j];
      Z[(size_t)i * nw + j] ^= Z[(size_t)k * nw + j];
    }
  }
  void row_copy(uint32_t i, uint32_t k) {
    for (uint32_t j = 0; j < nw; ++j) {
      X[(size_t)i * nw + j] = X[(size_t)k * nw + j];
      Z[(size_t)i * nw + j] = Z[(size_t)k * nw + j];
    }
  }
  void row_zero(uint32_t i) {
    for (uint32_t j = 0; j < nw; ++j) X[(size_t)i * nw + j] = Z[(size_t)i * nw + j] = 0;
  }
  bool is_identity() const {
    XZ id;
    id.init_identity(n);
    return id.X == X && id.Z == Z;
  }
};

// LSB-first 64x64 bit-matrix transpose in place: bit b of a[k] <-> bit k of
// a[b] (swapping ever smaller off-diagonal blocks).
void transpose64(uint64_t a[64]) {
  static const uint64_t kMask[6] = {0x00000000FFFFFFFFull, 0x0000FFFF0000FFFFull,
                                    0x00FF00FF00FF00FFull, 0x0F0F0F0F0F0F0F0Full,
                                    0x3333333333333333ull, 0x5555555555555555ull};
  for (int s = 0, j = 32; j; j >>= 1, ++s)
    for (int k = 0; k < 64; ++k)
      if (!(k & j)) {
        const uint64_t t = ((a[k] >> j) ^ a[k | j]) & kMask[s];
        a[k] ^= t << j;
        a[k | j] ^= t;
      }
}

// The same X/Z by columns: X[q] is the set of rows with x_q (bit i of word
// i / 64), so a gate is a few word operations over rows.
struct Cols {
  uint32_t n = 0, R = 0, rw = 0;
  std::vector<uint64_t> X, Z;  // [n][rw]
  uint64_t *x(uint32_t q) { return &X[(size_t)q * rw]; }
  uint64_t *z(uint32_t q) { return &Z[(size_t)q * rw]; }
};

// rows [R][nw] <-> columns [n][rw] by 64x64 blocks
void rows_to_cols(const std::vector<uint64_t> &rows, uint32_t R, uint32_t n, uint32_t nw,
                  std::vector<uint64_t> &cols, uint32_t rw) {
  cols.assign((size_t)n * rw, 0);
  uint64_t a[64];
  for (uint32_t bi = 0; bi < rw; ++bi)
    for (uint32_t bj = 0; bj < nw; ++bj) {
      for (uint32_t k = 0; k < 64; ++k) {
        const uint32_t row = bi * 64 + k;
        a[k] = row < R ? rows[(size_t)row * nw + bj] : 0;
      }
      transpose64(a);
      for (uint32_t k = 0; k < 64; ++k) {
        const uint32_t q = bj * 64 + k;
        if (q < n) cols[(size_t)q * rw + bi] = a[k];
      }
    }
}
void cols_to_rows(const std::vector<uint64_t> &cols, uint32_t R, uint32_t n, uint32_t rw,
                  std::vector<uint64_t> &rows, uint32_t nw) {
  uint64_t a[64];
  for (uint32_t bj = 0; bj < nw; ++bj)
    for (uint32_t bi = 0; bi < rw; ++bi) {
      for (uint32_t k = 0; k < 64; ++k) {
        const uint32_t q = bj * 64 + k;
        a[k] = q < n ? cols[(size_t)q * rw + bi] : 0;
      }
      transpose64(a);
      for (uint32_t k = 0; k < 64; ++k) {
        const uint32_t row = bi * 64 + k;
        if (row < R) rows[(size_t)row * nw + bj] = a[k];
      }
    }
}

// Mask source of an XOR entry: the constant all-ones word, a record's
// result word (by rec_l index) or a noise mask part (2k + part: part 0 is
// the X-or-Y mask, also a random flag; part 1 the Y-or-Z mask).
enum : uint32_t { kSrcConst = 0, kSrcRec = 1, kSrcNoise = 2 };
struct Entry {
  uint32_t row, src, arg;
};

struct Logical {
  uint32_t kind = 0;  // kFrameXor / kFrameRandom / kFrameDet
  std::vector<Entry> ents;      // xor
  uint32_t p = 0, slot = 0, mi = 0, T = 0, ridx = kNone, c = 0;
  std::vector<uint32_t> rows;   // random: row | phase << 31; det: rows
};

bool frame_flag_breaks_xz(uint32_t c) {
  return c == PQ_OP_H_FLAG || c == PQ_OP_S_FLAG || c == PQ_OP_CX_FLAG || c == PQ_OP_CZ_FLAG;
}

// The shared-X/Z run of `ops` from the identity tableau.
int simulate(pq_program_s *p, FramePlan &F, std::vector<Logical> &out) {
  const uint32_t n = (uint32_t)p->n, R = 2 * n + 1;
  XZ t;
  t.init_identity(n);
  // LDS words for record values that flag ops read.  A value lives from the
  // M that writes it (or the program start) to its last reader; words are
  // reused once a value is dead, so the count is the most values live at
  // once, not the number of records.  op_ridx[i]: the word op i writes (M)
  // or reads (flag), kNone if none.
  const size_t nops = p->ops.size();
  std::vector<uint32_t> op_ridx(nops, kNone);
  {
    auto is_rec_flag = [](const pq_op &o) {
      return o.code >= PQ_OP_X_FLAG && o.code <= PQ_OP_Z_FLAG && o.slot != PQ_FLAG_RANDOM;
    };
    // last reader of each value: producer -1 = the value from before the program
    std::map<uint32_t, int64_t> producer;           // slot -> op index of its current value
    std::map<std::pair<uint32_t, int64_t>, int64_t> last_read;  // (slot, producer) -> reader
    for (size_t i = 0; i < nops; ++i) {
      const pq_op &o = p->ops[i];
      if (o.code == PQ_OP_M) {
        producer[o.slot] = (int64_t)i;
      } else if (is_rec_flag(o)) {
        auto it = producer.find(o.slot);
        last_read[{o.slot, it == producer.end() ? -1 : it->second}] = (int64_t)i;
      }
    }
    std::vector<uint32_t> free_words;
    uint32_t high = 0;
    auto take = [&]() {
      if (!free_words.empty()) {
        const uint32_t w = free_words.back();
        free_words.pop_back();
        return w;
      }
      return high++;
    };
    std::map<uint32_t, uint32_t> word_of;  // slot -> word of its live value
    for (const auto &kv : last_read)
      if (kv.first.second < 0) {  // read before any M of this program writes it
        const uint32_t w = take();
        word_of[kv.first.first] = w;
        F.preload.push_back(kv.first.first);
        F.preload.push_back(w);
      }
    producer.clear();
    for (size_t i = 0; i < nops; ++i) {
      const pq_op &o = p->ops[i];
      if (o.code == PQ_OP_M) {
        producer[o.slot] = (int64_t)i;
        word_of.erase(o.slot);
        if (last_read.count({o.slot, (int64_t)i})) {
          op_ridx[i] = take();
          word_of[o.slot] = op_ridx[i];
        }
      } else if (is_rec_flag(o)) {
        auto it = producer.find(o.slot);
        const int64_t prod = it == producer.end() ? -1 : it->second;
        op_ridx[i] = word_of.at(o.slot);
        if (last_read.at({o.slot, prod}) == (int64_t)i) {  // the value's last reader
          free_words.push_back(op_ridx[i]);
          word_of.erase(o.slot);
        }
      }
    }
    F.nref = high;
  }
  // Between measurements X/Z are held by columns (gates are word operations
  // over rows); measurements work on rows.
  Cols cl;
  cl.n = n;
  cl.R = R;
  cl.rw = (R + 63) / 64;
  const uint32_t rw = cl.rw;
  bool by_cols = false;
  auto need_cols = [&]() {
    if (by_cols) return;
    rows_to_cols(t.X, R, n, t.nw, cl.X, rw);
    rows_to_cols(t.Z, R, n, t.nw, cl.Z, rw);
    by_cols = true;
  };
  auto need_rows = [&]() {
    if (!by_cols) return;
    cols_to_rows(cl.X, R, n, rw, t.X, t.nw);
    cols_to_rows(cl.Z, R, n, rw, t.Z, t.nw);
    by_cols = false;
  };
  std::vector<uint64_t> flip(rw, 0);  // rows whose r the gates so far flip
  const uint64_t last_mask = (R % 64) ? ((1ull << (R % 64)) - 1) : ~0ull;
  Logical cur;
  cur.kind = kFrameXor;
  // rows set in `bits` (rw words) each get mask source (src, arg)
  auto add_rows = [&](const uint64_t *bits, uint32_t src, uint32_t arg) {
    for (uint32_t j = 0; j < rw; ++j) {
      uint64_t v = bits[j] & (j + 1 == rw ? last_mask : ~0ull);
      while (v) {
        const uint32_t b = (uint32_t)__builtin_ctzll(v);
        v &= v - 1;
        cur.ents.push_back(Entry{j * 64 + b, src, arg});
      }
    }
  };
  auto close_xor = [&]() {
    add_rows(flip.data(), kSrcConst, 0);
    std::fill(flip.begin(), flip.end(), 0);
    if (!cur.ents.empty()) out.push_back(std::move(cur));
    cur = Logical();
    cur.kind = kFrameXor;
  };
  std::vector<uint64_t> tmp(rw), tmp2(rw), sel(rw);
  bool seg_rows = false;  // the current gate segment runs row by row
  // Paulis, noise and flagged Paulis on qubit q, given the rows with z_q
  // (zr) and with x_q (xr): X flips r where z_q, Z where x_q, Y where one of
  // them (GateFunctors.hpp:304-307, 324-327, 373-376)
  uint32_t kk = 0, mi = 0;
  auto apply_pauli_like = [&](const pq_op &o, size_t oi, const uint64_t *zr, const uint64_t *xr) {
    switch (o.code) {
      case PQ_OP_X:
        for (uint32_t j = 0; j < rw; ++j) flip[j] ^= zr[j];
        break;
      case PQ_OP_Y:
        for (uint32_t j = 0; j < rw; ++j) flip[j] ^= zr[j] ^ xr[j];
        break;
      case PQ_OP_Z:
        for (uint32_t j = 0; j < rw; ++j) flip[j] ^= xr[j];
        break;
      case PQ_OP_NOISE: {
        const uint32_t k = kk++;
        add_rows(zr, kSrcNoise, 2 * k);      // X or Y part
        add_rows(xr, kSrcNoise, 2 * k + 1);  // Y or Z part
        break;
      }
      default: {  // X/Y/Z_FLAG
        uint32_t src = kSrcRec, arg = op_ridx[oi];
        if (o.slot == PQ_FLAG_RANDOM) {
          src = kSrcNoise;
          arg = 2 * kk++;
        }
        for (uint32_t j = 0; j < rw; ++j)
          sel[j] = o.code == PQ_OP_X_FLAG ? zr[j] : o.code == PQ_OP_Z_FLAG ? xr[j] : (zr[j] ^ xr[j]);
        add_rows(sel.data(), src, arg);
        break;
      }
    }
  };
  for (size_t oi = 0; oi < p->ops.size(); ++oi) {
    const pq_op &o = p->ops[oi];
    const uint32_t q = o.q0, q1 = o.q1;
    if (frame_flag_breaks_xz(o.code))
      return pq_set_error(PQ_ERR_UNSUPPORTED,
                          ("frame form: op " + std::to_string(oi) +
                           " is a flagged H/S/CX/CZ (X/Z would differ between shots)").c_str());
    if (o.code != PQ_OP_M && !by_cols && !seg_rows) {
      // a new gate segment: by columns when its gates outweigh the two
      // transposes (64x64 blocks of X and Z, there and back)
      uint64_t g = 0;
      for (size_t k = oi; k < nops && p->ops[k].code != PQ_OP_M; ++k) ++g;
      const uint64_t by_row = g * R * 8, by_col = g * rw * 8 + 4ull * rw * t.nw * 600;
      if (by_col < by_row)
        need_cols();
      else
        seg_rows = true;
    }
    if (o.code == PQ_OP_M) seg_rows = false;
    if (seg_rows) {
      // the same updates row by row
      auto bit = [&](uint32_t i) { flip[i / 64] ^= 1ull << (i % 64); };
      switch (o.code) {
        case PQ_OP_H:
          for (uint32_t i = 0; i < R; ++i) {
            const bool x = t.x(i, q), z = t.z(i, q);
            if (x & z) bit(i);
            t.set(t.X, i, q, z);
            t.set(t.Z, i, q, x);
          }
          break;
        case PQ_OP_S:
          for (uint32_t i = 0; i < R; ++i) {
            const bool x = t.x(i, q), z = t.z(i, q);
            if (x & z) bit(i);
            t.set(t.Z, i, q, z ^ x);
          }
          break;
        case PQ_OP_CX:
          for (uint32_t i = 0; i < R; ++i) {
            const bool xc = t.x(i, q), zc = t.z(i, q), xt = t.x(i, q1), zt = t.z(i, q1);
            if (xc & zt & !(xt ^ zc)) bit(i);
            t.set(t.X, i, q1, xt ^ xc);
            t.set(t.Z, i, q, zc ^ zt);
          }
          break;
        case PQ_OP_CZ:
          for (uint32_t i = 0; i < R; ++i) {
            const bool xc = t.x(i, q), zc = t.z(i, q), xt = t.x(i, q1), zt = t.z(i, q1);
            if (xc & xt & (zt ^ zc)) bit(i);
            t.set(t.Z, i, q1, zt ^ xc);
            t.set(t.Z, i, q, zc ^ xt);
          }
          break;
        case PQ_OP_X: case PQ_OP_Y: case PQ_OP_Z: case PQ_OP_NOISE:
        case PQ_OP_X_FLAG: case PQ_OP_Y_FLAG: case PQ_OP_Z_FLAG: {
          // rows by z_q (X part) and x_q (Z part) as column bitsets
          std::fill(tmp.begin(), tmp.end(), 0);
          std::fill(tmp2.begin(), tmp2.end(), 0);
          for (uint32_t i = 0; i < R; ++i) {
            if (t.z(i, q)) tmp[i / 64] |= 1ull << (i % 64);
            if (t.x(i, q)) tmp2[i / 64] |= 1ull << (i % 64);
          }
          apply_pauli_like(o, oi, tmp.data(), tmp2.data());
          break;
        }
        default:
          return pq_set_error(PQ_ERR_ARG, "frame form: unknown op code");
      }
      continue;
    }
    switch (o.code) {
      case PQ_OP_H: {  // GateFunctors.hpp:90-93
        uint64_t *x = cl.x(q), *z = cl.z(q);
        for (uint32_t j = 0; j < rw; ++j) {
          flip[j] ^= x[j] & z[j];
          std::swap(x[j], z[j]);
        }
        break;
      }
      case PQ_OP_S: {  // :151-170
        uint64_t *x = cl.x(q), *z = cl.z(q);
        for (uint32_t j = 0; j < rw; ++j) {
          flip[j] ^= x[j] & z[j];
          z[j] ^= x[j];
        }
        break;
      }
      case PQ_OP_CX: {  // :248-252
        uint64_t *xc = cl.x(q), *zc = cl.z(q), *xt = cl.x(q1), *zt = cl.z(q1);
        for (uint32_t j = 0; j < rw; ++j) {
          flip[j] ^= xc[j] & zt[j] & ~(xt[j] ^ zc[j]);
          xt[j] ^= xc[j];
          zc[j] ^= zt[j];
        }
        break;
      }
      case PQ_OP_CZ: {  // :30-35
        uint64_t *xc = cl.x(q), *zc = cl.z(q), *xt = cl.x(q1), *zt = cl.z(q1);
        for (uint32_t j = 0; j < rw; ++j) {
          flip[j] ^= xc[j] & xt[j] & (zt[j] ^ zc[j]);
          zt[j] ^= xc[j];
          zc[j] ^= xt[j];
        }
        break;
      }
      case PQ_OP_X: case PQ_OP_Y: case PQ_OP_Z: case PQ_OP_NOISE:
      case PQ_OP_X_FLAG: case PQ_OP_Y_FLAG: case PQ_OP_Z_FLAG:
        apply_pauli_like(o, oi, cl.z(q), cl.x(q));
        break;
      case PQ_OP_M: {
        close_xor();
        need_rows();
        Logical m;
        m.slot = o.slot;
        m.mi = mi++;
        m.T = o.t0;
        m.ridx = op_ridx[oi];
        uint32_t piv = kNone;
        for (uint32_t i = 0; i < n; ++i)
          if (t.x(n + i, q)) {
            piv = i;
            break;
          }
        if (piv != kNone) {  // MeasureRandom, :225-265
          m.kind = kFrameRandom;
          m.p = piv;
          t.row_copy(piv, n + piv);
          t.row_zero(n + piv);
          t.set(t.Z, n + piv, q, true);
          for (uint32_t i = 0; i < 2 * n; ++i) {
            if (i == piv || i == n + piv || !t.x(i, q)) continue;
            const bool ph = t.phase(i, piv);
            t.row_xor(i, piv);
            m.rows.push_back(i | (ph ? 0x80000000u : 0u));
          }
        } else {  // MeasureDetermined, :189-201
          m.kind = kFrameDet;
          t.row_zero(2 * n);
          bool c = false;
          for (uint32_t i = 0; i < n; ++i) {
            if (!t.x(i, q)) continue;
            c ^= t.phase(2 * n, n + i);
            t.row_xor(2 * n, n + i);
            m.rows.push_back(n + i);
          }
          m.c = c ? 1u : 0u;
        }
        out.push_back(std::move(m));
        break;
      }
      default:
        return pq_set_error(PQ_ERR_ARG, "frame form: unknown op code");
    }
  }
  close_xor();
  need_rows();
  if (kk != p->nnoise || mi != p->nmeas)
    return pq_set_error(PQ_ERR_HIP, "frame form: noise/measurement numbering mismatch");
  F.final_identity = t.is_identity();
  F.X = t.X;
  F.Z = t.Z;
  F.nw = t.nw;
  return PQ_OK;
}

// Packs logical ops into chunks of at most F.sb_words stream words and
// F.nb_words noise-mask words.
struct Packer {
  FramePlan &F;
  uint32_t chunk_begin = 0;
  int64_t k0 = -1, k1 = -1;  // noise numbers of the open chunk
  explicit Packer(FramePlan &f) : F(f) {}
  uint32_t used() const { return (uint32_t)F.stream.size() - chunk_begin; }
  uint32_t room() const { return F.sb_words - used(); }
  void close_chunk() {
    if (used() == 0) {
      k0 = k1 = -1;
      return;
    }
    F.chunks.push_back(chunk_begin);
    F.chunks.push_back(used());
    F.chunks.push_back(k0 < 0 ? 0u : (uint32_t)k0);
    F.chunks.push_back(k0 < 0 ? 0u : (uint32_t)(k1 - k0));
    chunk_begin = (uint32_t)F.stream.size();
    k0 = k1 = -1;
  }
  void ensure(uint32_t words) {
    if (room() < words) close_chunk();
  }
  uint32_t lds_index(const Entry &e) const {
    if (e.src == kSrcConst) return F.lds_const;
    if (e.src == kSrcRec) return F.lds_rec + e.arg;
    const uint32_t buf = (uint32_t)(F.chunks.size() / 4) % F.nbufs;  // the open chunk's
    return F.lds_noise + buf * F.nb_words + (e.arg - 2 * (uint32_t)k0);
  }
  // One XOR op of `es` (all noise numbers inside the open chunk's window).
  void emit_xor(std::vector<Entry> &es) {
    std::stable_sort(es.begin(), es.end(),
                     [](const Entry &a, const Entry &b) { return a.row < b.row; });
    std::vector<uint32_t> rows, offs;
    for (size_t i = 0; i < es.size(); ++i)
      if (i == 0 || es[i].row != es[i - 1].row) {
        rows.push_back(es[i].row);
        offs.push_back((uint32_t)i);
      }
    offs.push_back((uint32_t)es.size());
    const uint32_t T = (uint32_t)rows.size();
    const uint32_t hdr[8] = {kFrameXor, T, (uint32_t)es.size(), 0, 0, 0, 0, 0};
    F.stream.insert(F.stream.end(), hdr, hdr + 8);
    F.stream.insert(F.stream.end(), rows.begin(), rows.end());
    F.stream.insert(F.stream.end(), offs.begin(), offs.end());
    for (const Entry &e : es) F.stream.push_back(lds_index(e));
    es.clear();
  }
  void add_xor(std::vector<Entry> ents) {
    // noise entries in noise-number order, so a block splits at window ends
    std::stable_sort(ents.begin(), ents.end(), [](const Entry &a, const Entry &b) {
      const int64_t ka = a.src == kSrcNoise ? (int64_t)(a.arg / 2) : -1;
      const int64_t kb = b.src == kSrcNoise ? (int64_t)(b.arg / 2) : -1;
      return ka < kb;
    });
    const int64_t nbk = F.nb_words / 2;  // noise numbers per window
    std::vector<Entry> part;
    std::vector<uint32_t> seen(F.R, 0);
    uint32_t gen = 1, T = 0;
    auto flush = [&]() {
      if (!part.empty()) emit_xor(part);
      ++gen;
      T = 0;
    };
    ensure(64);
    for (const Entry &e : ents) {
      if (e.src == kSrcNoise) {
        const int64_t k = e.arg / 2;
        if (k0 >= 0 && k >= k0 + nbk) {  // past the window: new chunk
          flush();
          close_chunk();
        }
        if (k0 < 0) {
          k0 = k;
          k1 = k + 1;
        } else if (k + 1 > k1) {
          k1 = k + 1;
        }
      }
      // header, rows, offsets and entries if this entry joins
      const uint32_t newrow = seen[e.row] != gen ? 1u : 0u;
      const uint32_t need = 8 + 2 * (T + newrow) + 1 + (uint32_t)part.size() + 1;
      if (need > room()) {
        flush();
        close_chunk();
        if (e.src == kSrcNoise) {
          k0 = e.arg / 2;
          k1 = k0 + 1;
        }
      }
      if (seen[e.row] != gen) {
        seen[e.row] = gen;
        ++T;
      }
      part.push_back(e);
    }
    flush();
  }
  void add_random(const Logical &m) {
    ensure(8 + (uint32_t)m.rows.size());
    const uint32_t hdr[8] = {kFrameRandom, m.p, m.slot, m.mi, m.T, m.ridx,
                             (uint32_t)m.rows.size(), 0};
    F.stream.insert(F.stream.end(), hdr, hdr + 8);
    F.stream.insert(F.stream.end(), m.rows.begin(), m.rows.end());
  }
  void add_det_batch(const std::vector<const Logical *> &ms) {
    uint32_t total = 0;
    for (const Logical *m : ms) total += (uint32_t)m->rows.size();
    ensure(8 + 4 * (uint32_t)ms.size() + 1 + total);
    const uint32_t hdr[8] = {kFrameDet, (uint32_t)ms.size(), 0, 0, 0, 0, 0, 0};
    F.stream.insert(F.stream.end(), hdr, hdr + 8);
    uint32_t off = 0;
    for (const Logical *m : ms) {
      F.stream.push_back(m->slot);
      F.stream.push_back(m->ridx);
      F.stream.push_back(m->c);
      F.stream.push_back(off);
      off += (uint32_t)m->rows.size();
    }
    F.stream.push_back(off);
    for (const Logical *m : ms) F.stream.insert(F.stream.end(), m->rows.begin(), m->rows.end());
  }
};

}  // namespace

int build_frame_plan(pq_program_s *p) {
  auto F = std::make_shared<FramePlan>();
  const uint32_t n = (uint32_t)p->n, R = 2 * n + 1;
  F->R = R;
  std::vector<Logical> ops;
  int rc = simulate(p, *F, ops);
  if (rc) return rc;
  // LDS of one wave, in 8-byte words: R, records read by flags, the
  // all-ones word, the noise-mask window, then the stream buffer (uint32)
  F->lds_rec = R;
  F->lds_const = R + F->nref;
  F->lds_noise = F->lds_const + 1;
  // a window of up to kFrameNoiseWords / 2 noise ops (none without noise)
  F->nb_words = (uint32_t)std::min<uint64_t>(kFrameNoiseWords, (2 * p->nnoise + 63) / 64 * 64);
  // the stream buffer holds the largest single measurement op
  uint32_t need = 64;
  for (const Logical &m : ops)
    if (m.kind != kFrameXor) need = std::max<uint32_t>(need, 8 + 4 + 1 + (uint32_t)m.rows.size());
  F->sb_words = std::max<uint32_t>(kFrameStreamWords, (need + 63) / 64 * 64);
  // ... and 64 words of partial results per wave (up to 16 waves); two
  // staging buffers (noise window and stream) when they fit
  auto lds_for = [&](uint64_t nbufs) {
    return 8ull * (F->lds_noise + nbufs * F->nb_words) + 4ull * nbufs * F->sb_words +
           8ull * 64 * 16;
  };
  F->nbufs = lds_for(2) <= kFrameLdsMax ? 2 : 1;
  const uint64_t lds = lds_for(F->nbufs);
  if (lds > kFrameLdsMax)
    return pq_set_error(PQ_ERR_UNSUPPORTED,
                        ("frame form needs " + std::to_string(lds) +
                         " bytes of LDS per word group (too many qubits or flag records)").c_str());
  F->lds_bytes = (uint32_t)lds;
  Packer pk(*F);
  for (size_t i = 0; i < ops.size();) {
    const Logical &m = ops[i];
    if (m.kind == kFrameXor) {
      pk.add_xor(m.ents);
      ++i;
    } else if (m.kind == kFrameRandom) {
      pk.add_random(m);
      ++i;
    } else {
      // consecutive determined measurements, at most 64, distinct slots,
      // fitting one chunk
      std::vector<const Logical *> batch;
      uint32_t words = 8 + 1;
      std::vector<uint32_t> slots;
      while (i < ops.size() && ops[i].kind == kFrameDet && batch.size() < 64) {
        const Logical &d = ops[i];
        const uint32_t w = 4 + (uint32_t)d.rows.size();
        if (words + w > F->sb_words) break;
        if (std::find(slots.begin(), slots.end(), d.slot) != slots.end()) break;
        slots.push_back(d.slot);
        batch.push_back(&d);
        words += w;
        ++i;
      }
      pk.add_det_batch(batch);
    }
  }
  pk.close_chunk();
  for (const Logical &m : ops) {
    if (m.kind == kFrameRandom) ++F->nrandom;
    if (m.kind == kFrameDet) ++F->ndet;
  }
  p->frame = std::move(F);
  return PQ_OK;
}

std::string frame_plan_key(uint64_t n, const pq_op *ops, uint64_t nops) {
  std::string key(reinterpret_cast<const char *>(&n), sizeof(n));
  key.reserve(sizeof(n) + nops * sizeof(pq_op));
  for (uint64_t i = 0; i < nops; ++i) {
    pq_op o = ops[i];
    if (o.code != PQ_OP_M) o.t0 = o.t1 = o.t2 = 0;  // noise and random-flag thresholds
    o.reserved = 0;
    key.append(reinterpret_cast<const char *>(&o), sizeof(o));
  }
  return key;
}

int upload_frame_plan(pq_program_s *p, hipStream_t stream) {
  FramePlan &F = *p->frame;
  std::lock_guard<std::mutex> g(F.upload_mu);
  if (F.d_stream) return PQ_OK;
  // plain allocations: a plan may be shared by programs on several streams
  // and outlive the program that built it
  auto up = [&](void **dst, const void *src, size_t bytes) -> hipError_t {
    hipError_t e = hipMalloc(dst, std::max<size_t>(bytes, 8));
    if (e != hipSuccess) return e;
    if (!bytes) return hipSuccess;
    return hipMemcpyAsync(*dst, src, bytes, hipMemcpyHostToDevice, stream);
  };
  HIPCHK(up((void **)&F.d_stream, F.stream.data(), F.stream.size() * 4));
  HIPCHK(up((void **)&F.d_chunks, F.chunks.data(), F.chunks.size() * 4));
  HIPCHK(up((void **)&F.d_preload, F.preload.data(), F.preload.size() * 4));
  HIPCHK(up((void **)&F.d_X, F.X.data(), F.X.size() * 8));
  HIPCHK(up((void **)&F.d_Z, F.Z.data(), F.Z.size() * 8));
  HIPCHK(hipStreamSynchronize(stream));
  return PQ_OK;
}

FramePlan::~FramePlan() {
  for (void *q : {(void *)d_stream, (void *)d_chunks, (void *)d_preload, (void *)d_X,
                  (void *)d_Z})
    if (q) (void)hipFree(q);
}

FrameArgs frame_args(pq_state s, pq_program_s *p) {
  const FramePlan &F = *p->frame;
  FrameArgs a{};
  a.tab = s->tab;
  a.recs = s->recs;
  a.planes = p->d_planes;
  a.stream = F.d_stream;
  a.chunks = F.d_chunks;
  a.nchunk = (uint32_t)(F.chunks.size() / 4);
  a.preload = F.d_preload;
  a.npreload = (uint32_t)(F.preload.size() / 2);
  a.n = (uint32_t)s->n;
  a.R = (uint32_t)s->R;
  a.C = (uint32_t)s->C;
  a.W = s->W;
  a.lds_rec = F.lds_rec;
  a.lds_const = F.lds_const;
  a.lds_noise = F.lds_noise;
  a.nb_words = F.nb_words;
  a.sb_words = F.sb_words;
  a.nbufs = F.nbufs;
  a.lds_bytes = F.lds_bytes;
  a.seed = (uint64_t)s->seed;
  a.mseq0 = s->meas_seq;
  a.col_stride = s->r_only ? s->R : s->C * s->R;
  a.r_off = s->r_only ? 0 : 2 * s->n * s->R;
  return a;
}

int run_frame(pq_state s, pq_program_s *p) {
  if (!s->xz_identity)
    return pq_set_error(PQ_ERR_UNSUPPORTED,
                        "frame form runs from the |0...0> tableau: reset the state first");
  if (!p->frame) {
    int rc = build_frame_plan(p);
    if (rc) return rc;
  }
  int rc = upload_frame_plan(p, s->stream);
  if (rc) return rc;
  const FramePlan &F = *p->frame;
  HIPCHK(launch_set_params(s->stream, s->params_dev, (uint64_t)s->seed, s->meas_seq,
                           s->noise_seq));
  FrameArgs a = frame_args(s, p);
  s->meas_seq += p->nmeas;
  s->noise_seq += p->nnoise;
  if (p->nnoise)
    HIPCHK(launch_program_noise(s->stream, p->d_planes, p->d_noise, (uint32_t)p->nnoise, s->W,
                                s->params_dev));
  if (p->timing_on) {
    if (p->ev_used + 2 > p->ev.size()) {
      hipEvent_t e0, e1;
      HIPCHK(hipEventCreate(&e0));
      HIPCHK(hipEventCreate(&e1));
      p->ev.push_back(e0);
      p->ev.push_back(e1);
    }
    HIPCHK(hipEventRecord(p->ev[p->ev_used], s->stream));
  }
  HIPCHK(launch_frame(s->stream, a));
  if (p->timing_on) {
    HIPCHK(hipEventRecord(p->ev[p->ev_used + 1], s->stream));
    p->ev_used += 2;
  }
  if (!s->r_only)  // an R-column-only state keeps no X/Z
    HIPCHK(launch_frame_xz(s->stream, s->tab, geom_of(s), F.d_X, F.d_Z, F.nw));
  s->xz_identity = F.final_identity;
  return PQ_OK;
}

extern "C" int pq_frame_plan(uint64_t n, const pq_op *ops, uint64_t nops, uint64_t info[16],
                             uint32_t *stream, uint32_t *chunks, uint32_t *preload, uint64_t *x,
                             uint64_t *z) {
  if (!info || (!ops && nops)) return pq_set_error(PQ_ERR_ARG, "null argument");
  if (n == 0 || n > 4095) return pq_set_error(PQ_ERR_UNSUPPORTED, "frame plans need 1..4095 qubits");
  // host only: a program object with no device buffers
  pq_program_s prog;
  prog.n = n;
  prog.R = 2 * n + 1;
  prog.ops.assign(ops, ops + nops);
  int rc = validate_ops(n, ops, nops);
  if (rc) return rc;
  rc = compile_program(&prog);
  if (rc) return rc;
  rc = build_frame_plan(&prog);
  if (rc) return rc;
  const FramePlan &F = *prog.frame;
  const uint64_t v[16] = {F.stream.size(), F.chunks.size() / 4, F.preload.size() / 2, F.nw,
                          F.nref, F.lds_rec, F.lds_const, F.lds_noise, F.nb_words, F.sb_words,
                          F.lds_bytes, F.nrandom, F.ndet, F.final_identity ? 1u : 0u, F.nbufs,
                          0};
  for (int i = 0; i < 16; ++i) info[i] = v[i];
  if (stream) std::copy(F.stream.begin(), F.stream.end(), stream);
  if (chunks) std::copy(F.chunks.begin(), F.chunks.end(), chunks);
  if (preload) std::copy(F.preload.begin(), F.preload.end(), preload);
  if (x) std::copy(F.X.begin(), F.X.end(), x);
  if (z) std::copy(F.Z.begin(), F.Z.end(), z);
  return PQ_OK;
}
