#include <cstdlib>
// Frame form of a program: circuits whose X/Z tableau is the same in every
// shot run with X/Z simulated once, here on the host, and only per-shot words
// on the GPU.
//
// X and Z change only through H, S, CX, CZ and measurements, and those act on
// every shot alike (GateFunctors.hpp; MeasurementFunctors.hpp:225-265 picks
// the pivot from X alone), so without flagged H/S/CX/CZ a state that starts
// with the same X/Z in every shot keeps it.  The phase column R then differs
// between shots only through noise, random flags, record flags and
// measurement coins; simulate() below keeps it as R = Rv ^ cst ^ <E, P> (a
// per-shot part Rv, a compile-time constant cst and a per-shot Pauli frame
// E) and compiles the program into a stream the GPU kernel runs per 64-shot
// word group (k_frame in pq_kernels.hip), with Rv, E and the record words
// held in LDS.
//
// Stream layout (uint32 words), one op after another, each with an 8-word
// header {kind, a, b, c, d, e, f, g}:
//  kLayer:  a = T tasks, b = body words; then offsets[T+1] (in updates,
//           relative to the body), a pad word if the body would start at an
//           odd word of the chunk, and the bodies.  A body is a list of
//           updates, two words each: 16-bit LDS word indices {dst | s0 << 16,
//           s1 | s2 << 16}, LDS word dst ^= s0 ^ s1 ^ s2 (unused sources name
//           the zero word), in order.  Tasks touch disjoint frame words, one
//           thread runs each.
//  kRandom: a = p, b = record slot, c = measurement index in the program,
//           d = coin threshold, e = rec_l index or kNone, f = entries,
//           g = LDS word of E.x[q]; then the rowsum rows (other than p, n+p).
//  kDet:    a = G measurements (<= 64); then per measurement {slot, rec_l
//           index or kNone, LDS word of E.x[q] | c << 31 | last << 30, row
//           offset}, the end offset, and the row lists (rows n+i).  The
//           program's last determined measurement (last = 1) puts its Rv part
//           into Rv[2n].
//  kXor:    a = T rows, b = entries; then rows[T], offsets[T+1] (relative to
//           the entry array), entries: LDS words XORed into Rv[row] (the
//           closing R = Rv ^ cst ^ <E, P> update).
// The stream is cut into chunks that fit the kernel's LDS staging buffer;
// each chunk also names the noise numbers [k0, k0+nk) whose masks it reads.
// With two staging buffers (when they fit the LDS) chunk c uses buffer c % 2
// and the kernel loads chunk c+1 while it runs chunk c; a chunk's noise
// sources name the mask words of its own buffer.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
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

// Sources of the words a frame update XORs in: LDS words by absolute index
// (frame words, record words, the all-ones word), or noise mask parts
// kNoiseSrc | (2k + part) (part 0 the X-or-Y mask, also a random flag; part 1
// the Y-or-Z mask), placed in the noise window when the packer picks it.
constexpr uint32_t kNoiseSrc = 0x80000000u;
constexpr uint32_t kMaxSrc = 3;       // sources per update
constexpr uint32_t kTaskWords = 256;  // stream words of one task at most
constexpr uint32_t kRecPool = 128;    // record words handed out before reuse

// dst ^= XOR of s[0..ns) (LDS words)
struct TEnt {
  uint32_t dst = 0, ns = 0;
  uint32_t s[kMaxSrc] = {};
};
// Updates one thread runs in order; the tasks of one layer touch disjoint
// frame words, so they run in parallel.
struct Task {
  std::vector<TEnt> ents;
  int64_t kmin = -1, kmax = -1;  // noise numbers read
  uint32_t words = 0;            // stream words of the body
};
struct Meas {
  uint32_t exq = 0;  // LDS word of the frame's x part on the measured qubit
  uint32_t p = 0, slot = 0, mi = 0, T = 0, ridx = kNone, c = 0;
  bool scratch = false;  // determined: the program's last, writes Rv[2n]
  std::vector<uint32_t> rows;
};
// One barrier-separated step of the kernel: a layer of frame tasks, one
// random measurement or a batch of determined measurements.
struct Step {
  uint32_t kind = kFrameLayer;
  std::vector<Task> tasks;
  std::vector<Meas> ms;
  uint32_t words = 0;  // determined batch: stream words
};

bool frame_flag_breaks_xz(uint32_t c) {
  return c == PQ_OP_H_FLAG || c == PQ_OP_S_FLAG || c == PQ_OP_CX_FLAG || c == PQ_OP_CZ_FLAG;
}

// The shared-X/Z run of `ops` from the identity tableau, scheduled into
// steps; `fin` receives the closing R update ({row, LDS word} pairs).
//
// Per shot the R column is kept split as R[i] = Rv[i] ^ cst[i] ^ <E, P_i>:
// Rv the part the GPU holds, cst a constant the compiler tracks (the same in
// every shot), E a Pauli frame (x and z words per qubit, 64 shots a word) and
// <E, P_i> its symplectic product with row i's Pauli (bit 1 where they
// anticommute).  This holds because every R update is one of:
//  * a gate's phase rule (GateFunctors.hpp): a function of X/Z only, so a
//    constant, into cst;
//  * a Pauli, noise or flagged Pauli P (:304-307, 324-327, 373-376): R[i]
//    flips where P anticommutes with row i, i.e. E ^= P (constant Paulis go
//    into cst instead); a gate U then maps every row P_i to U P_i U^dag,
//    which keeps <E, P_i> when E is mapped to U E U^dag alike (H swaps the
//    x and z words, S: z ^= x, CX: x_t ^= x_c, z_c ^= z_t, CZ: z_b ^= x_a,
//    z_a ^= x_b);
//  * a random measurement (MeasureRandom, MeasurementFunctors.hpp:225-265)
//    with pivot p: rowsums multiply row i by row n+p, so <E, .> adds up and
//    Rv, cst follow the reference's sign rule; the new row n+p = Z_q gets
//    R = coin, i.e. Rv[n+p] = coin ^ E.x[q] with cst 0;
//  * a determined measurement (MeasureDetermined, :189-201): the product of
//    its stabilizer rows is Z_q, so result = XOR Rv[rows] ^ (c ^ XOR
//    cst[rows]) ^ E.x[q].
// At the end R[i] = Rv[i] ^ cst[i] ^ <E, P_i> is written out.  The frame
// words of a noisy circuit change with each gate and noise op, O(1) words
// each, instead of the R words of every row the op anticommutes with.
int simulate(pq_program_s *p, FramePlan &F, std::vector<Step> &steps,
             std::vector<std::pair<uint32_t, uint32_t>> &fin) {
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
    // Up to kRecPool words are handed out fresh, then the longest-dead word
    // is reused first: a word reused at once would order a measurement
    // after the flag that read the word's previous value, and so keep
    // consecutive measurements out of one batch.
    std::deque<uint32_t> free_words;
    uint32_t high = 0;
    auto take = [&]() {
      if (!free_words.empty() && high >= kRecPool) {
        const uint32_t w = free_words.front();
        free_words.pop_front();
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
  // LDS words: Rv [R], the frame [2n], records [nref], the all-ones and the
  // zero word, then the staging buffers
  const uint32_t E0 = R;
  F.lds_rec = R + 2 * n;
  F.lds_const = F.lds_rec + F.nref;  // the all-ones word, then the zero word
  F.lds_noise = F.lds_const + 2;

  // ---- scheduling of frame updates and measurements into steps ----
  // Frame slot s lives at LDS word E0 + s; sx[q] / sz[q] are the slots of
  // qubit q's x / z word (H swaps them, which costs no update).
  std::vector<uint32_t> sx(n), sz(n);
  for (uint32_t q = 0; q < n; ++q) {
    sx[q] = q;
    sz[q] = n + q;
  }
  std::vector<int64_t> last(2 * n, -1);  // last step touching each slot
  std::vector<int64_t> own(2 * n, -1);   // its task there (-1: a measurement)
  std::vector<uint8_t> nz(2 * n, 0);     // slot may be nonzero
  std::vector<int64_t> rec_w(F.nref, -1), rec_r(F.nref, -1);
  const int64_t krange = std::max<int64_t>(1, F.nb_words / 4);  // noise span of one task
  auto is_frame = [&](uint32_t v) { return !(v & kNoiseSrc) && v >= E0 && v < E0 + 2 * n; };
  auto is_rec = [&](uint32_t v) {
    return !(v & kNoiseSrc) && v >= F.lds_rec && v < F.lds_rec + F.nref;
  };
  auto ent_noise = [&](const TEnt &e, int64_t &lo, int64_t &hi) {
    for (uint32_t i = 0; i < e.ns; ++i)
      if (e.s[i] & kNoiseSrc) {
        const int64_t k = (e.s[i] & ~kNoiseSrc) / 2;
        lo = lo < 0 ? k : std::min(lo, k);
        hi = std::max(hi, k);
      }
  };
  // e joins the latest update of the same word when no update after it
  // reads that word or writes one of e's sources
  auto append = [&](Task &T, const TEnt &e) {
    int64_t lim = 8;
    for (int64_t j = (int64_t)T.ents.size() - 1; j >= 0 && lim > 0; --j, --lim) {
      TEnt &f = T.ents[(size_t)j];
      if (f.dst == e.dst) {
        if (f.ns + e.ns <= kMaxSrc) {
          for (uint32_t i = 0; i < e.ns; ++i) f.s[f.ns++] = e.s[i];
          return;
        }
        break;
      }
      bool clash = false;
      for (uint32_t i = 0; i < f.ns; ++i) clash |= f.s[i] == e.dst;
      for (uint32_t i = 0; i < e.ns; ++i) clash |= e.s[i] == f.dst;
      if (clash) break;
    }
    T.ents.push_back(e);
    T.words += 2;
  };
  // the frame updates of one op, in order
  auto add_frame = [&](std::initializer_list<TEnt> list) {
    TEnt keep[4];
    uint32_t nk = 0;
    for (TEnt e : list) {
      uint32_t k = 0;
      for (uint32_t i = 0; i < e.ns; ++i)
        if (!is_frame(e.s[i]) || nz[e.s[i] - E0]) e.s[k++] = e.s[i];  // zero words add nothing
      e.ns = k;
      if (!k) continue;
      nz[e.dst - E0] = 1;
      keep[nk++] = e;
    }
    if (!nk) return;
    int64_t smax = -1, rmax = -1, lo = -1, hi = -1;
    uint32_t words = 0;
    for (uint32_t j = 0; j < nk; ++j) {
      const TEnt &e = keep[j];
      smax = std::max(smax, last[e.dst - E0]);
      for (uint32_t i = 0; i < e.ns; ++i) {
        if (is_frame(e.s[i])) smax = std::max(smax, last[e.s[i] - E0]);
        if (is_rec(e.s[i])) rmax = std::max(rmax, rec_w[e.s[i] - F.lds_rec]);
      }
      ent_noise(e, lo, hi);
      words += 2;
    }
    int64_t step = -1, task = -1;
    // fuse into the task that last touched these words when every word it
    // touches was last touched there or in an earlier step
    if (smax >= 0 && steps[(size_t)smax].kind == kFrameLayer && rmax < smax) {
      int64_t tk = -1;
      bool ok = true;
      auto see = [&](uint32_t slot) {
        if (last[slot] != smax) return;
        if (own[slot] < 0 || (tk >= 0 && own[slot] != tk)) ok = false;
        tk = own[slot];
      };
      for (uint32_t j = 0; j < nk; ++j) {
        see(keep[j].dst - E0);
        for (uint32_t i = 0; i < keep[j].ns; ++i)
          if (is_frame(keep[j].s[i])) see(keep[j].s[i] - E0);
      }
      if (ok && tk >= 0) {
        const Task &T = steps[(size_t)smax].tasks[(size_t)tk];
        const int64_t nlo = T.kmin < 0 ? lo : (lo < 0 ? T.kmin : std::min(T.kmin, lo));
        const int64_t nhi = std::max(T.kmax, hi);
        if (T.words + words <= kTaskWords && (nlo < 0 || nhi - nlo < krange)) {
          step = smax;
          task = tk;
        }
      }
    }
    if (step < 0) {
      // the first layer after every step these words and records were in
      const int64_t lb = std::max(smax, rmax) + 1;
      for (int64_t i = lb; i < (int64_t)steps.size(); ++i)
        if (steps[(size_t)i].kind == kFrameLayer) {
          step = i;
          break;
        }
      if (step < 0) {
        steps.emplace_back();
        step = (int64_t)steps.size() - 1;
      }
      steps[(size_t)step].tasks.emplace_back();
      task = (int64_t)steps[(size_t)step].tasks.size() - 1;
    }
    Task &T = steps[(size_t)step].tasks[(size_t)task];
    for (uint32_t j = 0; j < nk; ++j) {
      const TEnt &e = keep[j];
      append(T, e);
      last[e.dst - E0] = step;
      own[e.dst - E0] = task;
      for (uint32_t i = 0; i < e.ns; ++i) {
        if (is_frame(e.s[i])) {
          last[e.s[i] - E0] = step;
          own[e.s[i] - E0] = task;
        }
        if (is_rec(e.s[i])) {
          int64_t &r = rec_r[e.s[i] - F.lds_rec];
          r = std::max(r, step);
        }
      }
    }
    if (lo >= 0) {
      T.kmin = T.kmin < 0 ? lo : std::min(T.kmin, lo);
      T.kmax = std::max(T.kmax, hi);
    }
  };
  auto ent1 = [](uint32_t dst, uint32_t src) {
    TEnt e;
    e.dst = dst;
    e.ns = 1;
    e.s[0] = src;
    return e;
  };
  // Measurement steps.  Random measurements rewrite Rv rows and take a new
  // step at the end.  A determined one reads its rows and writes only the
  // scratch row (whose value matters only after the last of them, which is
  // flagged to write it), so it joins the first determined batch after the
  // steps that wrote its rows, its frame word, its record word and its slot.
  std::vector<int64_t> rv_w(R, -1);  // last step writing each Rv row
  std::unordered_map<uint32_t, int64_t> slot_step;
  int64_t sc_step = -1;
  size_t sc_idx = 0;
  auto place_meas = [&](Meas &&m, bool random, uint32_t exq_slot, uint32_t stream_words) {
    int64_t lb = last[exq_slot];
    if (m.ridx != kNone) lb = std::max(lb, std::max(rec_w[m.ridx], rec_r[m.ridx]));
    auto it = slot_step.find(m.slot);
    if (it != slot_step.end()) lb = std::max(lb, it->second);
    int64_t step = -1;
    if (!random) {
      for (uint32_t r : m.rows) lb = std::max(lb, rv_w[r]);
      for (int64_t i = lb + 1; i < (int64_t)steps.size(); ++i) {
        Step &s = steps[(size_t)i];
        if (s.kind != kFrameDet || s.ms.size() >= 64 || s.words + stream_words > F.sb_words)
          continue;
        bool dup = false;
        for (const Meas &o : s.ms) dup |= o.slot == m.slot;
        if (!dup) {
          step = i;
          break;
        }
      }
    }
    if (step < 0) {
      steps.emplace_back();
      step = (int64_t)steps.size() - 1;
      steps[(size_t)step].kind = random ? kFrameRandom : kFrameDet;
      steps[(size_t)step].words = 8 + 1;
    }
    Step &s = steps[(size_t)step];
    s.words += stream_words;
    if (m.ridx != kNone) rec_w[m.ridx] = step;
    last[exq_slot] = std::max(last[exq_slot], step);
    own[exq_slot] = -1;
    slot_step[m.slot] = step;
    if (random) {
      for (uint32_t r : m.rows) rv_w[r] = step;
      rv_w[m.p] = rv_w[n + m.p] = step;
    } else {
      sc_step = step;
      sc_idx = s.ms.size();
    }
    s.ms.push_back(std::move(m));
  };

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
  std::vector<uint64_t> flip(rw, 0);  // rows whose cst the gates since the last fold flip
  std::vector<uint8_t> cst(R, 0);
  auto fold_flip = [&]() {  // word by word: most words have no flips
    for (uint32_t j = 0; j < rw; ++j) {
      for (uint64_t v = flip[j]; v; v &= v - 1) cst[j * 64 + (uint32_t)__builtin_ctzll(v)] ^= 1;
      flip[j] = 0;
    }
  };
  std::vector<uint64_t> tmp(rw), tmp2(rw);
  bool seg_rows = false;  // the current gate segment runs row by row
  uint32_t kk = 0, mi = 0;
  // constant Paulis on qubit q, given the rows with z_q (zr) and with x_q
  // (xr): X flips r where z_q, Z where x_q, Y where one of them
  auto const_pauli = [&](uint32_t code, const uint64_t *zr, const uint64_t *xr) {
    for (uint32_t j = 0; j < rw; ++j)
      flip[j] ^= code == PQ_OP_X ? zr[j] : code == PQ_OP_Z ? xr[j] : (zr[j] ^ xr[j]);
  };
  // noise and flagged Paulis: frame updates (X part into x, Z part into z)
  auto frame_pauli = [&](const pq_op &o, size_t oi) {
    const uint32_t q = o.q0;
    if (o.code == PQ_OP_NOISE) {
      const uint32_t k = kk++;
      add_frame({ent1(E0 + sx[q], kNoiseSrc | (2 * k)), ent1(E0 + sz[q], kNoiseSrc | (2 * k + 1))});
      return;
    }
    uint32_t src = F.lds_rec + op_ridx[oi];
    if (o.slot == PQ_FLAG_RANDOM) src = kNoiseSrc | (2 * kk++);
    if (o.code == PQ_OP_X_FLAG)
      add_frame({ent1(E0 + sx[q], src)});
    else if (o.code == PQ_OP_Z_FLAG)
      add_frame({ent1(E0 + sz[q], src)});
    else
      add_frame({ent1(E0 + sx[q], src), ent1(E0 + sz[q], src)});
  };
  // frame map of the gates (conjugation, phases dropped)
  auto frame_gate = [&](const pq_op &o) {
    const uint32_t q = o.q0, q1 = o.q1;
    switch (o.code) {
      case PQ_OP_H:
        std::swap(sx[q], sz[q]);
        break;
      case PQ_OP_S:
        add_frame({ent1(E0 + sz[q], E0 + sx[q])});
        break;
      case PQ_OP_CX:
        add_frame({ent1(E0 + sx[q1], E0 + sx[q]), ent1(E0 + sz[q], E0 + sz[q1])});
        break;
      case PQ_OP_CZ:
        add_frame({ent1(E0 + sz[q1], E0 + sx[q]), ent1(E0 + sz[q], E0 + sx[q1])});
        break;
      default:
        break;
    }
  };
  for (size_t oi = 0; oi < p->ops.size(); ++oi) {
    const pq_op &o = p->ops[oi];
    const uint32_t q = o.q0, q1 = o.q1;
    if (frame_flag_breaks_xz(o.code))
      return pq_set_error(PQ_ERR_UNSUPPORTED,
                          ("frame form: op " + std::to_string(oi) +
                           " is a flagged H/S/CX/CZ (X/Z would differ between shots)").c_str());
    switch (o.code) {
      case PQ_OP_H: case PQ_OP_S: case PQ_OP_CX: case PQ_OP_CZ:
        frame_gate(o);
        break;
      case PQ_OP_NOISE: case PQ_OP_X_FLAG: case PQ_OP_Y_FLAG: case PQ_OP_Z_FLAG:
        frame_pauli(o, oi);
        continue;  // X/Z and cst unchanged
      case PQ_OP_X: case PQ_OP_Y: case PQ_OP_Z: case PQ_OP_M:
        break;
      default:
        return pq_set_error(PQ_ERR_ARG, "frame form: unknown op code");
    }
    if (o.code != PQ_OP_M && !by_cols && !seg_rows) {
      // a new gate segment: by columns when its gates outweigh the two
      // transposes (64x64 blocks of X and Z, there and back)
      uint64_t g = 0;
      for (size_t k = oi; k < nops && p->ops[k].code != PQ_OP_M; ++k)
        if (p->ops[k].code != PQ_OP_NOISE && !(p->ops[k].code >= PQ_OP_X_FLAG &&
                                                p->ops[k].code <= PQ_OP_Z_FLAG))
          ++g;
      const uint64_t by_row = g * R * 8, by_col = g * rw * 8 + 4ull * rw * t.nw * 600;
      if (by_col < by_row)
        need_cols();
      else
        seg_rows = true;
    }
    if (o.code == PQ_OP_M) seg_rows = false;
    if (seg_rows) {
      // X/Z and the phase flips row by row
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
        default: {  // X, Y, Z
          std::fill(tmp.begin(), tmp.end(), 0);
          std::fill(tmp2.begin(), tmp2.end(), 0);
          for (uint32_t i = 0; i < R; ++i) {
            if (t.z(i, q)) tmp[i / 64] |= 1ull << (i % 64);
            if (t.x(i, q)) tmp2[i / 64] |= 1ull << (i % 64);
          }
          const_pauli(o.code, tmp.data(), tmp2.data());
          break;
        }
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
      case PQ_OP_X: case PQ_OP_Y: case PQ_OP_Z:
        const_pauli(o.code, cl.z(q), cl.x(q));
        break;
      case PQ_OP_M: {
        need_rows();
        fold_flip();
        Meas m;
        m.exq = E0 + sx[q];
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
          m.p = piv;
          const uint8_t cs = cst[n + piv];
          t.row_copy(piv, n + piv);
          t.row_zero(n + piv);
          t.set(t.Z, n + piv, q, true);
          for (uint32_t i = 0; i < 2 * n; ++i) {
            if (i == piv || i == n + piv || !t.x(i, q)) continue;
            cst[i] ^= cs ^ (t.phase(i, piv) ? 1 : 0);
            t.row_xor(i, piv);
            m.rows.push_back(i);
          }
          cst[piv] = cs;
          cst[n + piv] = 0;
          const uint32_t words = 8 + (uint32_t)m.rows.size();
          place_meas(std::move(m), true, sx[q], words);
        } else {  // MeasureDetermined, :189-201
          t.row_zero(2 * n);
          uint8_t c = 0;
          for (uint32_t i = 0; i < n; ++i) {
            if (!t.x(i, q)) continue;
            c ^= (t.phase(2 * n, n + i) ? 1 : 0) ^ cst[n + i];
            t.row_xor(2 * n, n + i);
            m.rows.push_back(n + i);
          }
          // the product is Z_q, whose frame part is E.x[q]
          for (uint32_t j = 0; j < t.nw; ++j) {
            const uint64_t zq = (q / 64 == j) ? (1ull << (q % 64)) : 0ull;
            if (t.X[(size_t)(2 * n) * t.nw + j] != 0 || t.Z[(size_t)(2 * n) * t.nw + j] != zq)
              return pq_set_error(PQ_ERR_HIP, "frame form: determined product is not Z_q");
          }
          cst[2 * n] = c;
          m.c = c;
          const uint32_t words = 4 + (uint32_t)m.rows.size();
          place_meas(std::move(m), false, sx[q], words);
        }
        break;
      }
      default:
        break;
    }
  }
  need_rows();
  fold_flip();
  if (sc_step >= 0) steps[(size_t)sc_step].ms[sc_idx].scratch = true;
  if (kk != p->nnoise || mi != p->nmeas)
    return pq_set_error(PQ_ERR_HIP, "frame form: noise/measurement numbering mismatch");
  // closing update: R[i] = Rv[i] ^ cst[i] ^ <E, P_i>
  for (uint32_t i = 0; i < R; ++i) {
    if (cst[i]) fin.emplace_back(i, F.lds_const);
    for (uint32_t j = 0; j < n; ++j) {
      if (t.z(i, j) && nz[sx[j]]) fin.emplace_back(i, E0 + sx[j]);
      if (t.x(i, j) && nz[sz[j]]) fin.emplace_back(i, E0 + sz[j]);
    }
  }
  F.final_identity = t.is_identity();
  F.X = t.X;
  F.Z = t.Z;
  F.nw = t.nw;
  return PQ_OK;
}

// Packs steps into chunks of at most F.sb_words stream words, each with a
// noise window of at most F.nb_words mask words.
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
  uint32_t src_index(uint32_t v) const {
    if (!(v & kNoiseSrc)) return v;
    const uint32_t buf = (uint32_t)(F.chunks.size() / 4) % F.nbufs;  // the open chunk's
    return F.lds_noise + buf * F.nb_words + ((v & ~kNoiseSrc) - 2 * (uint32_t)k0);
  }
  void emit_layer(const std::vector<const Task *> &ts) {
    uint32_t body = 0;
    for (const Task *t : ts) body += t->words;
    const uint32_t T = (uint32_t)ts.size();
    const bool pad = (used() + 8 + T + 1) & 1;  // updates start at an even word of the chunk
    const uint32_t hdr[8] = {kFrameLayer, T, body, 0, 0, 0, 0, 0};
    F.stream.insert(F.stream.end(), hdr, hdr + 8);
    uint32_t off = 0;
    for (const Task *t : ts) {
      F.stream.push_back(off);
      off += t->words / 2;
    }
    F.stream.push_back(off);
    if (pad) F.stream.push_back(0);
    const uint32_t zero = F.lds_const + 1;
    for (const Task *t : ts)
      for (const TEnt &e : t->ents) {
        uint32_t v[kMaxSrc];
        for (uint32_t i = 0; i < kMaxSrc; ++i) v[i] = i < e.ns ? src_index(e.s[i]) : zero;
        F.stream.push_back(e.dst | (v[0] << 16));
        F.stream.push_back(v[1] | (v[2] << 16));
      }
  }
  // A layer in as few pieces as room and noise windows allow (its tasks are
  // independent, so any split runs them all before the next step).
  int add_layer(const std::vector<Task> &tasks) {
    std::vector<const Task *> rest;
    for (const Task &t : tasks) rest.push_back(&t);
    std::stable_sort(rest.begin(), rest.end(),
                     [](const Task *a, const Task *b) { return a->kmin < b->kmin; });
    const int64_t nbk = F.nb_words / 2;  // noise numbers per window
    while (!rest.empty()) {
      std::vector<const Task *> take, keep;
      uint32_t words = 8 + 1 + 1;  // header, end offset, alignment pad
      bool full = false;
      for (const Task *t : rest) {
        const bool fit_w = t->kmin < 0 || k0 < 0 || (t->kmin >= k0 && t->kmax < k0 + nbk);
        if (full || !fit_w || words + 1 + t->words > room()) {
          full |= fit_w && words + 1 + t->words > room();
          keep.push_back(t);
          continue;
        }
        take.push_back(t);
        words += 1 + t->words;
        if (t->kmin >= 0) {
          if (k0 < 0) k0 = t->kmin;
          k1 = std::max<int64_t>(k1, t->kmax + 1);
        }
      }
      if (take.empty()) {
        if (used() == 0)
          return pq_set_error(PQ_ERR_HIP, "frame form: a task exceeds the stream buffer");
        close_chunk();
        continue;
      }
      emit_layer(take);
      rest.swap(keep);
      if (!rest.empty()) close_chunk();
    }
    return PQ_OK;
  }
  void add_random(const Meas &m) {
    ensure(8 + (uint32_t)m.rows.size());
    const uint32_t hdr[8] = {kFrameRandom, m.p, m.slot, m.mi, m.T, m.ridx,
                             (uint32_t)m.rows.size(), m.exq};
    F.stream.insert(F.stream.end(), hdr, hdr + 8);
    F.stream.insert(F.stream.end(), m.rows.begin(), m.rows.end());
  }
  void add_det_batch(const std::vector<Meas> &ms) {
    uint32_t total = 0;
    for (const Meas &m : ms) total += (uint32_t)m.rows.size();
    ensure(8 + 4 * (uint32_t)ms.size() + 1 + total);
    const uint32_t hdr[8] = {kFrameDet, (uint32_t)ms.size(), 0, 0, 0, 0, 0, 0};
    F.stream.insert(F.stream.end(), hdr, hdr + 8);
    uint32_t off = 0;
    for (const Meas &m : ms) {
      F.stream.push_back(m.slot);
      F.stream.push_back(m.ridx);
      F.stream.push_back(m.exq | (m.c << 31) | (m.scratch ? 1u << 30 : 0u));
      F.stream.push_back(off);
      off += (uint32_t)m.rows.size();
    }
    F.stream.push_back(off);
    for (const Meas &m : ms) F.stream.insert(F.stream.end(), m.rows.begin(), m.rows.end());
  }
  // The closing update ({row, LDS word} pairs, by row) as XOR ops; a row
  // whose entries outgrow the room continues in the next op.
  void add_xor(const std::vector<std::pair<uint32_t, uint32_t>> &es) {
    std::vector<uint32_t> rows, offs, ents;
    auto flush = [&]() {
      if (rows.empty()) return;
      offs.push_back((uint32_t)ents.size());
      const uint32_t hdr[8] = {kFrameXor, (uint32_t)rows.size(), (uint32_t)ents.size(),
                               0, 0, 0, 0, 0};
      F.stream.insert(F.stream.end(), hdr, hdr + 8);
      F.stream.insert(F.stream.end(), rows.begin(), rows.end());
      F.stream.insert(F.stream.end(), offs.begin(), offs.end());
      F.stream.insert(F.stream.end(), ents.begin(), ents.end());
      rows.clear();
      offs.clear();
      ents.clear();
    };
    ensure(64);
    size_t i = 0;
    while (i < es.size()) {
      const bool newrow = rows.empty() || rows.back() != es[i].first;
      const uint32_t T = (uint32_t)rows.size() + (newrow ? 1u : 0u);
      if (8 + 2 * T + 1 + (uint32_t)ents.size() + 1 > room()) {
        flush();
        close_chunk();
        continue;
      }
      if (newrow) {
        rows.push_back(es[i].first);
        offs.push_back((uint32_t)ents.size());
      }
      ents.push_back(es[i].second);
      ++i;
    }
    flush();
  }
};

}  // namespace

// Staging sizes, smallest first: {noise window words, stream words, buffers}.
struct Staging {
  uint32_t nb, sb, bufs;
};
constexpr Staging kStaging[] = {{512, 2048, 1}, {1024, 4096, 1}, {2048, 4096, 1},
                                {2048, 8192, 1}, {2048, 8192, 2}};
constexpr uint32_t kFrameGroupsPerCu = 4;  // resident groups that hide a group's latencies
constexpr uint64_t kCus = 256;             // compute units of an MI355X

int build_frame_plan(pq_program_s *p, uint64_t groups) {
  auto F = std::make_shared<FramePlan>();
  const uint32_t n = (uint32_t)p->n, R = 2 * n + 1;
  F->R = R;
  // the compiler sizes tasks and determined batches for the smallest staging
  F->nb_words = kStaging[0].nb;
  F->sb_words = kStaging[0].sb;
  std::vector<Step> steps;
  std::vector<std::pair<uint32_t, uint32_t>> fin;
  int rc = simulate(p, *F, steps, fin);
  if (rc) return rc;
  // the stream buffer holds the largest single measurement op
  uint32_t need = 64;
  for (const Step &s : steps)
    for (const Meas &m : s.ms) need = std::max<uint32_t>(need, 8 + 4 + 1 + (uint32_t)m.rows.size());
  need = (need + 63) / 64 * 64;
  // LDS in bytes: Rv, frame, records, the all-ones word, the noise windows
  // (none without noise), the stream buffers and 64 words of partial results
  // per wave.  Each group runs a chain of short dependent steps, so several
  // groups per CU hide each other's latencies: the largest staging that
  // keeps as many groups per CU as the launch has (up to kFrameGroupsPerCu)
  // is taken; a launch of at most one group per CU gets the largest staging,
  // double-buffered.
  const uint64_t nb_cap = (2 * p->nnoise + 63) / 64 * 64;
  auto lds_of = [&](const Staging &o) {
    const uint64_t nb = std::min<uint64_t>(o.nb, nb_cap), sb = std::max(o.sb, need);
    return 8ull * (F->lds_noise + o.bufs * nb) + 4ull * o.bufs * sb + 8ull * kFrameThreads;
  };
  auto per_cu = [&](const Staging &o) {
    const uint64_t l = lds_of(o);
    return l <= kFrameLdsMax ? kFrameLdsMax / l : 0;
  };
  const uint64_t want = std::max<uint64_t>(1, groups / kCus + (groups % kCus != 0));
  const uint64_t target = std::min<uint64_t>(std::min<uint64_t>(kFrameGroupsPerCu, want),
                                             per_cu(kStaging[0]));
  if (!target)
    return pq_set_error(PQ_ERR_UNSUPPORTED,
                        ("frame form needs " + std::to_string(lds_of(kStaging[0])) +
                         " bytes of LDS per word group (too many qubits or flag records)").c_str());
  const Staging *pick = &kStaging[0];
  for (const Staging &o : kStaging)
    if (per_cu(o) >= target) pick = &o;
  F->nb_words = (uint32_t)std::min<uint64_t>(pick->nb, nb_cap);
  F->sb_words = std::max(pick->sb, need);
  F->nbufs = pick->bufs;
  F->lds_bytes = (uint32_t)lds_of(*pick);
  Packer pk(*F);
  for (const Step &s : steps) {
    if (s.kind == kFrameLayer) {
      if ((rc = pk.add_layer(s.tasks))) return rc;
    } else if (s.kind == kFrameRandom) {
      pk.add_random(s.ms[0]);
      ++F->nrandom;
    } else {
      pk.add_det_batch(s.ms);
      F->ndet += s.ms.size();
    }
  }
  pk.add_xor(fin);
  pk.close_chunk();
  // where a two-part run splits: the first chunk whose predecessors' noise
  // reaches half the ops (they then read only ops below split_k)
  // (PQ_FRAME_SPLIT_PCT: that share of the ops in percent, default 50)
  {
    static const uint64_t pct = [] {
      const char *e = getenv("PQ_FRAME_SPLIT_PCT");
      const long v = e ? atol(e) : 50;
      return (uint64_t)(v >= 1 && v <= 99 ? v : 50);
    }();
    const uint32_t nchunk = (uint32_t)(F->chunks.size() / 4);
    uint32_t kmax = 0;
    for (uint32_t c = 1; c < nchunk; ++c) {
      const uint32_t nk = F->chunks[4 * (c - 1) + 3];
      if (nk) kmax = std::max(kmax, F->chunks[4 * (c - 1) + 2] + nk);
      if (100 * (uint64_t)kmax >= pct * p->nnoise && kmax < p->nnoise) {
        F->split_chunk = c;
        F->split_k = kmax;
        break;
      }
    }
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
  a.w0 = 0;
  a.wn = s->W;
  a.c0 = 0;
  a.c1 = (uint32_t)(F.chunks.size() / 4);
  a.suspend = a.resume = 0;
  a.side = nullptr;
  a.nnoise = p->nnoise;
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
    int rc = build_frame_plan(p, s->W);
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
  // A launch of more word groups than the GPU holds at once (CUs x groups
  // per CU) runs in slices of about that many: the noise kernel, bound by
  // its draws, fills slice i+1 on the auxiliary stream while k_frame, bound
  // by LDS latency, runs slice i on the state's stream.
  const uint64_t per_cu =
      std::max<uint64_t>(1, std::min<uint64_t>(kFrameGroupsPerCu, kFrameLdsMax / F.lds_bytes));
  const uint64_t held = kCus * per_cu;
  const uint64_t slices =
      (p->nnoise && s->own_stream && s->W >= 2 * held) ? std::min<uint64_t>(64, s->W / held) : 1;
  if (slices > 1 && !s->aux) HIPCHK(hipStreamCreateWithFlags(&s->aux, hipStreamNonBlocking));
  while (s->aux_ev.size() < slices + 1) {
    hipEvent_t e;
    HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    s->aux_ev.push_back(e);
  }
  // A launch of at most that many groups runs the plan in two parts when
  // it splits: the noise of the second part's ops is drawn on the auxiliary
  // stream while k_frame runs the first part, and each group's frame and
  // record words wait in p->d_side between the two k_frame launches.
  const bool two_part = slices == 1 && p->nnoise && s->own_stream && F.split_chunk > 0;
  if (two_part && !s->aux) HIPCHK(hipStreamCreateWithFlags(&s->aux, hipStreamNonBlocking));
  if (two_part && !p->d_side)
    HIPCHK(hipMalloc((void **)&p->d_side,
                     std::max<uint64_t>(8, s->W * (uint64_t)(F.lds_const - F.R) * 8)));
  if (slices == 1 && p->nnoise)
    HIPCHK(launch_frame_noise(s->stream, p->d_planes, p->d_noise, (uint32_t)p->nnoise, 0, s->W,
                              s->params_dev, 0, two_part ? F.split_k : ~0u));
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
  if (slices == 1 && two_part) {
    // the second half of the noise starts once the first is drawn
    HIPCHK(hipEventRecord(s->aux_ev[0], s->stream));
    HIPCHK(hipStreamWaitEvent(s->aux, s->aux_ev[0], 0));
    HIPCHK(launch_frame_noise(s->aux, p->d_planes, p->d_noise, (uint32_t)p->nnoise, 0, s->W,
                              s->params_dev, F.split_k, ~0u));
    HIPCHK(hipEventRecord(s->aux_ev[1], s->aux));
    const uint32_t nchunk = a.c1;
    a.side = p->d_side;
    a.c1 = F.split_chunk;
    a.suspend = 1;
    HIPCHK(launch_frame(s->stream, a));
    HIPCHK(hipStreamWaitEvent(s->stream, s->aux_ev[1], 0));
    a.c0 = F.split_chunk;
    a.c1 = nchunk;
    a.suspend = 0;
    a.resume = 1;
    HIPCHK(launch_frame(s->stream, a));
  } else if (slices == 1) {
    HIPCHK(launch_frame(s->stream, a));
  } else {
    // the auxiliary stream starts after everything queued on the state's
    // stream (the parameters, the reset, the last run's k_frame reading
    // the planes); each slice's k_frame waits for that slice's noise
    HIPCHK(hipEventRecord(s->aux_ev[0], s->stream));
    HIPCHK(hipStreamWaitEvent(s->aux, s->aux_ev[0], 0));
    for (uint64_t i = 0; i < slices; ++i) {
      const uint64_t w0 = i * s->W / slices, w1 = (i + 1) * s->W / slices;
      HIPCHK(launch_frame_noise(s->aux, p->d_planes, p->d_noise, (uint32_t)p->nnoise, w0,
                                w1 - w0, s->params_dev));
      HIPCHK(hipEventRecord(s->aux_ev[i + 1], s->aux));
      HIPCHK(hipStreamWaitEvent(s->stream, s->aux_ev[i + 1], 0));
      a.w0 = w0;
      a.wn = w1 - w0;
      HIPCHK(launch_frame(s->stream, a));
    }
  }
  if (p->timing_on) {
    HIPCHK(hipEventRecord(p->ev[p->ev_used + 1], s->stream));
    p->ev_used += 2;
  }
  if (!s->r_only) {  // an R-column-only state keeps no X/Z
    // every group's X/Z is the plan's shared X/Z: written when something
    // reads or updates the tableau (materialize_tableau), or now if the raw
    // tableau pointer was handed out
    s->xz_pending = false;
    s->xz_plan = p->frame;
    if (s->exposed) {
      int rc2 = materialize_tableau(s);
      if (rc2) return rc2;
    }
  }
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
  rc = compile_frame_only(&prog);
  if (rc) return rc;
  rc = build_frame_plan(&prog, UINT64_MAX);
  if (rc) return rc;
  const FramePlan &F = *prog.frame;
  const uint64_t v[16] = {F.stream.size(), F.chunks.size() / 4, F.preload.size() / 2, F.nw,
                          F.nref, F.lds_rec, F.lds_const, F.lds_noise, F.nb_words, F.sb_words,
                          F.lds_bytes, F.nrandom, F.ndet, F.final_identity ? 1u : 0u, F.nbufs,
                          F.split_chunk};
  for (int i = 0; i < 16; ++i) info[i] = v[i];
  if (stream) std::copy(F.stream.begin(), F.stream.end(), stream);
  if (chunks) std::copy(F.chunks.begin(), F.chunks.end(), chunks);
  if (preload) std::copy(F.preload.begin(), F.preload.end(), preload);
  if (x) std::copy(F.X.begin(), F.X.end(), x);
  if (z) std::copy(F.Z.begin(), F.Z.end(), z);
  return PQ_OK;
}
