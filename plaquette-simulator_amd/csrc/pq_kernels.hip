// gfx950 kernels of the bit-plane stabilizer-tableau engine.
//
// Word (w, c, row) of the tableau is tab[(w*C + c)*R + row]; every column of
// a 64-shot word group is R contiguous words, so a gate sweep over column q
// is a fully coalesced stream and a rowsum over rows i of column j is too.
#include "pq_kernels.h"

#include <algorithm>

#include "plaquette_amd.h"

namespace pq {

namespace {

constexpr uint32_t kMeasThreads = 256;
constexpr uint32_t kMaxWaves = 16;  // up to 1024 threads per block
constexpr uint32_t kMeasMaxQubits = 4095;

__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & 63u; }

// draw_keyed(key, 64 w + lane) with the product split as
// (64 w + lane + 1) G = 64 w G + (lane + 1) G: the first term is wave-uniform
// (scalar) and lane_g = (lane + 1) G is computed once per thread, so a draw
// costs one 64-bit add and the SplitMix64 output function.
__device__ __forceinline__ uint64_t lane_golden() { return (uint64_t)(lane_id() + 1) * kGolden; }
__device__ __forceinline__ uint64_t draw_word_lane(uint64_t key, uint64_t w, uint64_t lane_g) {
  return mix64(key + (64 * w) * kGolden + lane_g);
}

// ---------------------------------------------------------------------------
// Init: word (c, row) of every group is all-ones iff row == c, which is the
// identity tableau with r = 1 on the scratch row (MiscFunctors.hpp:22-27).
// ---------------------------------------------------------------------------
// R-column-only states: word (w, row) = all ones on the scratch row R-1.
__global__ __launch_bounds__(256) void k_fill_r_identity(uint64_t *__restrict__ rcol, uint32_t R,
                                                         uint64_t W) {
  for (uint64_t w = blockIdx.x; w < W; w += gridDim.x)
    for (uint32_t row = threadIdx.x; row < R; row += 256)
      rcol[w * R + row] = row == R - 1 ? ~0ull : 0ull;
}

// Column 2n (the R column) of every group of a full [W][C][R] tableau.
__global__ __launch_bounds__(256) void k_fill_r_column(uint64_t *__restrict__ tab, uint32_t R,
                                                       uint32_t C, uint64_t W) {
  for (uint64_t w = blockIdx.x; w < W; w += gridDim.x) {
    uint64_t *col = tab + (w * C + (C - 1)) * R;
    for (uint32_t row = threadIdx.x; row < R; row += 256) col[row] = row == R - 1 ? ~0ull : 0ull;
  }
}

// One block per column (w, c) at a time, its threads over the column's R
// contiguous rows: one index split per column, coalesced stores.
__global__ __launch_bounds__(256) void k_fill_identity(uint64_t *__restrict__ tab, uint32_t R,
                                                       uint32_t C, uint64_t ncols) {
  for (uint64_t col = blockIdx.x; col < ncols; col += gridDim.x) {
    const uint32_t c = (uint32_t)(col % C);
    uint64_t *dst = tab + col * R;
    for (uint32_t row = threadIdx.x; row < R; row += 256) dst[row] = (row == c) ? ~0ull : 0ull;
  }
}

// ---------------------------------------------------------------------------
// Gate sweep: one thread per (word group, row).  The rules are the
// reference's per-(shot,row) functors (GateFunctors.hpp) applied to 64 shots
// at once; fa/fb are optional per-word shot masks (flag, or noise X/Z parts).
// ---------------------------------------------------------------------------
struct GateArgs {
  uint32_t code, n, R, C, qa, qb;
  uint64_t W;
  const uint64_t *fa;
  const uint64_t *fb;
};

__global__ __launch_bounds__(256) void k_gate(uint64_t *__restrict__ tab,
                                              GateArgs a) {
  const uint64_t total = a.W * a.R;
  const uint64_t R = a.R;
  for (uint64_t idx = blockIdx.x * 256ull + threadIdx.x; idx < total;
       idx += (uint64_t)gridDim.x * 256ull) {
    const uint64_t w = idx / R;
    const uint64_t row = idx - w * R;
    uint64_t *T = tab + w * (uint64_t)a.C * R + row;
    uint64_t *pxa = T + (uint64_t)a.qa * R;
    uint64_t *pza = T + (uint64_t)(a.n + a.qa) * R;
    uint64_t *pr = T + (uint64_t)(2 * a.n) * R;
    switch (a.code) {
      case PQ_OP_H: {  // GateFunctors.hpp:90-93
        const uint64_t x = *pxa, z = *pza;
        *pr ^= x & z;
        *pxa = z;
        *pza = x;
        break;
      }
      case PQ_OP_S: {  // GateFunctors.hpp:151-170
        const uint64_t x = *pxa, z = *pza;
        *pr ^= x & z;
        *pza = z ^ x;
        break;
      }
      case PQ_OP_CX: {  // GateFunctors.hpp:248-252, qa control, qb target
        uint64_t *pxb = T + (uint64_t)a.qb * R;
        uint64_t *pzb = T + (uint64_t)(a.n + a.qb) * R;
        const uint64_t xc = *pxa, zc = *pza, xt = *pxb, zt = *pzb;
        *pr ^= xc & zt & ~(xt ^ zc);
        *pxb = xt ^ xc;
        *pza = zc ^ zt;
        break;
      }
      case PQ_OP_CZ: {  // GateFunctors.hpp:30-35
        uint64_t *pxb = T + (uint64_t)a.qb * R;
        uint64_t *pzb = T + (uint64_t)(a.n + a.qb) * R;
        const uint64_t xc = *pxa, zc = *pza, xt = *pxb, zt = *pzb;
        *pr ^= xc & xt & (zt ^ zc);
        *pzb = zt ^ xc;
        *pza = zc ^ xt;
        break;
      }
      case PQ_OP_X: *pr ^= *pza; break;            // :304-307
      case PQ_OP_Y: *pr ^= *pza ^ *pxa; break;     // :324-327
      case PQ_OP_Z: *pr ^= *pxa; break;            // :373-376
      case PQ_OP_H_FLAG: {                          // :142-148
        const uint64_t f = a.fa[w];
        const uint64_t x = *pxa, z = *pza;
        const uint64_t d = f & (x ^ z);
        *pr ^= f & x & z;
        *pxa = x ^ d;
        *pza = z ^ d;
        break;
      }
      case PQ_OP_S_FLAG: {                          // :220-228
        const uint64_t f = a.fa[w];
        const uint64_t x = *pxa, z = *pza;
        *pr ^= f & x & z;
        *pza = z ^ (f & x);
        break;
      }
      case PQ_OP_NOISE: {  // X part where fa, Z part where fb
        const uint64_t fx = a.fa[w], fz = a.fb[w];
        *pr ^= (fx & *pza) ^ (fz & *pxa);
        break;
      }
      case PQ_OP_CX_FLAG: {  // the CX rule (:248-252) on flagged shots
        const uint64_t f = a.fa[w];
        uint64_t *pxb = T + (uint64_t)a.qb * R;
        uint64_t *pzb = T + (uint64_t)(a.n + a.qb) * R;
        const uint64_t xc = *pxa, zc = *pza, xt = *pxb, zt = *pzb;
        *pr ^= f & xc & zt & ~(xt ^ zc);
        *pxb = xt ^ (f & xc);
        *pza = zc ^ (f & zt);
        break;
      }
      case PQ_OP_CZ_FLAG: {  // the CZ rule (:30-35) on flagged shots
        const uint64_t f = a.fa[w];
        uint64_t *pxb = T + (uint64_t)a.qb * R;
        uint64_t *pzb = T + (uint64_t)(a.n + a.qb) * R;
        const uint64_t xc = *pxa, zc = *pza, xt = *pxb, zt = *pzb;
        *pr ^= f & xc & xt & (zt ^ zc);
        *pzb = zt ^ (f & xc);
        *pza = zc ^ (f & xt);
        break;
      }
      default: break;
    }
  }
}

// ---------------------------------------------------------------------------
// Noise planes: plane k holds, per word group, the X-part and Z-part shot
// masks of the k-th noise op.  One wave per (k, w): each lane draws its shot.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_noise_planes(
    uint64_t *__restrict__ planes, NoiseDesc one,
    const NoiseDesc *__restrict__ desc, uint32_t nk, uint64_t W) {
  const uint64_t waves = (uint64_t)gridDim.x * (blockDim.x / 64);
  const uint64_t wave = blockIdx.x * (uint64_t)(blockDim.x / 64) + (threadIdx.x >> 6);
  const uint32_t lane = lane_id();
  for (uint64_t p = wave; p < (uint64_t)nk * W; p += waves) {
    const uint64_t k = p / W;
    const uint64_t w = p - k * W;
    const NoiseDesc d = desc ? desc[k] : one;
    bool bx, bz;
    noise_bits(d, draw_word_lane(d.key, w, lane_golden()), bx, bz);
    const uint64_t mx = __ballot(bx);  // X or Y (or the random flag)
    const uint64_t mz = __ballot(bz);  // Y or Z
    if (lane == 0) {
      planes[(2 * k) * W + w] = mx;
      planes[(2 * k + 1) * W + w] = mz;
    }
  }
}

// ---------------------------------------------------------------------------
// Block-wide scans (blockDim a multiple of 64, at most 1024 threads).
// ---------------------------------------------------------------------------
struct ScanShared {
  uint64_t w64[kMaxWaves];
  uint32_t w32[kMaxWaves];
  uint64_t tot64;
  uint32_t tot32;
};

// Exclusive OR-scan of v over threads in thread order; *total = OR of all.
__device__ uint64_t block_excl_or(uint64_t v, ScanShared &sh, uint64_t *total) {
  const uint32_t lane = lane_id(), wid = threadIdx.x >> 6;
  const uint32_t nw = blockDim.x >> 6;
  uint64_t inc = v;
#pragma unroll
  for (uint32_t d = 1; d < 64; d <<= 1) {
    const uint64_t t = __shfl_up(inc, d, 64);
    if (lane >= d) inc |= t;
  }
  uint64_t exc = __shfl_up(inc, 1, 64);
  if (lane == 0) exc = 0;
  if (lane == 63) sh.w64[wid] = inc;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t run = 0;
    for (uint32_t i = 0; i < nw; ++i) {
      const uint64_t t = sh.w64[i];
      sh.w64[i] = run;
      run |= t;
    }
    sh.tot64 = run;
  }
  __syncthreads();
  exc |= sh.w64[wid];
  *total = sh.tot64;
  __syncthreads();
  return exc;
}

// Exclusive sum-scan of v over threads in thread order; *total = sum.
__device__ uint32_t block_excl_sum(uint32_t v, ScanShared &sh, uint32_t *total) {
  const uint32_t lane = lane_id(), wid = threadIdx.x >> 6;
  const uint32_t nw = blockDim.x >> 6;
  uint32_t inc = v;
#pragma unroll
  for (uint32_t d = 1; d < 64; d <<= 1) {
    const uint32_t t = __shfl_up(inc, d, 64);
    if (lane >= d) inc += t;
  }
  if (lane == 63) sh.w32[wid] = inc;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t run = 0;
    for (uint32_t i = 0; i < nw; ++i) {
      const uint32_t t = sh.w32[i];
      sh.w32[i] = run;
      run += t;
    }
    sh.tot32 = run;
  }
  __syncthreads();
  const uint32_t exc = inc - v + sh.w32[wid];
  *total = sh.tot32;
  __syncthreads();
  return exc;
}

// ---------------------------------------------------------------------------
// MeasureQubit for one 64-shot word group (BatchMeasureFunctor,
// MeasurementFunctors.hpp:291-317), executed by a whole workgroup.
//
//  1. Pivot per shot = first stabilizer row with x_q set (:297-301): an
//     exclusive OR-scan over the stabilizer rows turns column q into
//     disjoint "first pivot" masks, one group (p, M) per distinct pivot.
//  2. Random shots, per group (MeasureRandom :226-265): copy row n+p to row
//     p, clear row n+p, set its z_q and its r to the coin, all masked by M;
//     then every row i (not p, n+p) with x_q set in M gets rowsum(i <- p),
//     masked per shot, with the bit-sliced mod-4 phase (only columns where
//     the source row is non-identity can contribute).
//  3. Determined shots D (MeasureDetermined :190-201): one wave clears the
//     scratch row in D and folds in stabilizer n+i for ascending i with
//     destabilizer x_q set, reducing each step's phase across lanes.
//  4. Record: result = coin (random) | scratch r (determined), determined = D.
// ---------------------------------------------------------------------------
// Workgroups that share one 64-shot word group (`slices` of them) meet at
// barriers on a device counter of that group: each arrival adds 1, and the
// k-th barrier waits until the counter reaches k * slices.  The launch is
// cooperative, so all of them are resident; a wait still gives up after
// kSliceSpinLimit polls (or the launch's own limit), sets *err and lets
// every later barrier pass at once,
// so no kernel can hang on it (the host reports the error).  Thread 0 releases
// the workgroup's writes at device scope before arriving and acquires after
// the wait; the workgroup barriers around it order the other threads.
constexpr uint32_t kSliceSpinLimit = 1u << 22;
struct SliceSync {
  uint32_t slice = 0, slices = 1;
  unsigned int *ctr = nullptr;
  unsigned int *err = nullptr;
  uint32_t epoch = 0;
  uint32_t spin_limit = kSliceSpinLimit;
};
__device__ __forceinline__ void slice_barrier(SliceSync &ss) {
  if (ss.slices <= 1) return;
  __syncthreads();
  ++ss.epoch;
  if (threadIdx.x == 0) {
    __threadfence();
    atomicAdd(ss.ctr, 1u);
    const unsigned int target = ss.epoch * ss.slices;
    uint32_t spins = 0;
    while (__hip_atomic_load(ss.ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
      if (__hip_atomic_load(ss.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
      __builtin_amdgcn_s_sleep(2);
      if (++spins >= ss.spin_limit) {
        atomicOr(ss.err, 1u);
        break;
      }
    }
    __threadfence();
  }
  __syncthreads();
}

struct MeasShared {
  ScanShared scan;
  uint64_t gm[64];
  uint32_t gp[64];
  uint32_t ngroups;
  uint32_t nlist;
  uint32_t krows;  // census: rows rowsummed by random groups
  uint64_t rm;    // random shots
  uint64_t coin;  // coin word
};

__device__ void measure_word(uint64_t *__restrict__ T, uint64_t w, uint32_t n,
                             uint32_t R, uint32_t q, uint32_t coin_T,
                             uint64_t coin_key, const uint64_t *coin_plane,
                             uint64_t *rec_res, uint64_t *rec_det,
                             uint64_t *s_src, uint32_t *s_list,
                             MeasShared &sh, bool force_determined = false,
                             unsigned long long *census = nullptr,
                             SliceSync *ss = nullptr) {
  const uint32_t tid = threadIdx.x, NT = blockDim.x;
  // with slices, every slice finds the same pivots and coins; slice 0 alone
  // rewrites the pivot rows and runs the determined shots, and the rowsum
  // rows are split between the slices
  const uint32_t slice = ss ? ss->slice : 0, S = ss ? ss->slices : 1;
  const uint64_t Ru = R;
  const uint64_t *xq = T + (uint64_t)q * Ru;

  // ---- 1. pivot groups ----
  // When the block has a thread per row (one workgroup per group, 2n <= NT),
  // thread t reads x_q of row t once: the pivot scan takes the stabilizer
  // rows from it and the rowsums their masks.  The random shots' rewrites
  // change x_q only in the shots of their own group, and every later use
  // masks it by another group's shots (or by D), so the value stays valid.
  const bool own_row = S == 1 && 2 * n <= NT;
  const uint64_t xv = (own_row && tid < 2 * n) ? xq[tid] : 0ull;
  const uint32_t cp = (n + NT - 1) / NT;
  const uint32_t p0 = min(n, tid * cp), p1 = min(n, p0 + cp);
  uint64_t acc = 0;
  if (own_row) {
    acc = tid >= n ? xv : 0ull;
  } else {
    for (uint32_t p = p0; p < p1; ++p) acc |= xq[n + p];
  }
  if (tid == 0) {
    sh.ngroups = 0;
    sh.krows = 0;
  }
  uint64_t rm;
  uint64_t before = block_excl_or(acc, sh.scan, &rm);
  if (force_determined) rm = 0;
  if (own_row) {
    const uint64_t first = acc & ~before;
    if (first && !force_determined) {
      const uint32_t g = atomicAdd(&sh.ngroups, 1u);
      sh.gp[g] = tid - n;
      sh.gm[g] = first;
    }
  } else if (S == 1) {
    for (uint32_t p = p0; p < p1 && !force_determined; ++p) {
      const uint64_t v = xq[n + p];
      const uint64_t first = v & ~before;
      before |= v;
      if (first) {
        const uint32_t g = atomicAdd(&sh.ngroups, 1u);
        sh.gp[g] = p;
        sh.gm[g] = first;
      }
    }
  } else {
    // slices rewrite shared words group by group, so they must all walk the
    // groups in one order: ascending pivot row, placed by a prefix count
    uint32_t cnt = 0;
    uint64_t seen = before;
    for (uint32_t p = p0; p < p1 && !force_determined; ++p) {
      const uint64_t v = xq[n + p];
      cnt += (v & ~seen) ? 1u : 0u;
      seen |= v;
    }
    uint32_t total;
    uint32_t g = block_excl_sum(cnt, sh.scan, &total);
    for (uint32_t p = p0; p < p1 && !force_determined; ++p) {
      const uint64_t v = xq[n + p];
      const uint64_t first = v & ~before;
      before |= v;
      if (first) {
        sh.gp[g] = p;
        sh.gm[g] = first;
        ++g;
      }
    }
    if (tid == 0) sh.ngroups = total;
  }
  if (tid < 64 && rm) {
    uint64_t coin;
    if (coin_plane) {
      coin = coin_plane[w];
    } else {
      const uint32_t u = (uint32_t)(draw_keyed(coin_key, 64 * w + tid) >> 40);
      coin = __ballot(u < coin_T);
    }
    if (tid == 0) sh.coin = coin;
  }
  __syncthreads();
  const uint64_t coin = rm ? sh.coin : 0;
  const uint32_t ngroups = sh.ngroups;

  // ---- 2. random shots, one pivot group at a time ----
  const uint32_t C = 2 * n + 1;
  for (uint32_t g = 0; g < ngroups; ++g) {
    const uint32_t p = sh.gp[g];
    const uint64_t M = sh.gm[g];
    if (S == 1) {
      for (uint32_t c = tid; c < C; c += NT) {
        uint64_t *col = T + (uint64_t)c * Ru;
        const uint64_t a = col[n + p];
        const uint64_t b = col[p];
        col[p] = (b & ~M) | (a & M);
        uint64_t na = a & ~M;
        if (c == n + q) na |= M;
        if (c == 2 * n) na |= coin & M;
        col[n + p] = na;
        s_src[c] = a;
      }
    } else {
      // every slice takes the source row before slice 0 rewrites it
      for (uint32_t c = tid; c < C; c += NT) s_src[c] = T[(uint64_t)c * Ru + n + p];
      slice_barrier(*ss);
      if (slice == 0) {
        for (uint32_t c = tid; c < C; c += NT) {
          uint64_t *col = T + (uint64_t)c * Ru;
          const uint64_t a = s_src[c];
          const uint64_t b = col[p];
          col[p] = (b & ~M) | (a & M);
          uint64_t na = a & ~M;
          if (c == n + q) na |= M;
          if (c == 2 * n) na |= coin & M;
          col[n + p] = na;
        }
      }
    }
    if (tid == 0) sh.nlist = 0;
    __syncthreads();
    for (uint32_t j = tid; j < n; j += NT)
      if ((s_src[j] | s_src[n + j]) & M) s_list[atomicAdd(&sh.nlist, 1u)] = j;
    __syncthreads();
    const uint32_t nl = sh.nlist;
    const uint64_t rsrc = s_src[2 * n];
    uint32_t my_rows = 0;
    for (uint32_t i = slice * NT + tid; i < 2 * n; i += NT * S) {
      if (i == p || i == n + p) continue;
      const uint64_t mi = M & (own_row ? xv : xq[i]);
      if (!mi) continue;
      ++my_rows;
      // the phase word is read with the first source columns, not after
      // the row's stores
      const uint64_t ri = T[(uint64_t)(2 * n) * Ru + i];
      uint64_t c0 = 0, c1 = 0;
      // eight source columns at a time: all loads issued before any store
      for (uint32_t k0 = 0; k0 < nl; k0 += 8) {
        uint64_t xs[8], zs[8];
#pragma unroll
        for (uint32_t u = 0; u < 8; ++u) {
          if (k0 + u < nl) {
            const uint32_t j = s_list[k0 + u];
            xs[u] = T[(uint64_t)j * Ru + i];
            zs[u] = T[(uint64_t)(n + j) * Ru + i];
          }
        }
#pragma unroll
        for (uint32_t u = 0; u < 8; ++u) {
          if (k0 + u < nl) {
            const uint32_t j = s_list[k0 + u];
            const uint64_t x2 = s_src[j], z2 = s_src[n + j];
            phase_acc(xs[u], zs[u], x2, z2, c0, c1);
            T[(uint64_t)j * Ru + i] = xs[u] ^ (mi & x2);
            T[(uint64_t)(n + j) * Ru + i] = zs[u] ^ (mi & z2);
          }
        }
      }
      T[(uint64_t)(2 * n) * Ru + i] = ri ^ (mi & (rsrc ^ c1));
    }
    if (census && my_rows) atomicAdd(&sh.krows, my_rows);
    __syncthreads();
    // all slices' rowsums are in before the next pivot group or phase
    if (S > 1) slice_barrier(*ss);
  }

  // ---- 3. determined shots ----
  const uint64_t D = ~rm;
  uint64_t rdet = 0;
  uint32_t nchain = 0;
  if (D && slice == 0) {
    // ordered list of destabilizers i with x_q set in some determined shot
    uint32_t cnt = 0;
    for (uint32_t i = p0; i < p1; ++i) cnt += (xq[i] & D) ? 1u : 0u;
    uint32_t pos = block_excl_sum(cnt, sh.scan, &nchain);
    for (uint32_t i = p0; i < p1; ++i)
      if (xq[i] & D) s_list[pos++] = i;
    __syncthreads();
    if (tid < 64) {
      const uint32_t lane = tid;
      const uint64_t scratch = 2 * n;
      for (uint32_t j = lane; j < 2 * n; j += 64)
        s_src[j] = T[(uint64_t)j * Ru + scratch] & ~D;
      uint64_t racc = T[scratch * Ru + scratch] & ~D;
      for (uint32_t k = 0; k < nchain; ++k) {
        const uint32_t s = n + s_list[k];
        const uint64_t m = xq[s_list[k]] & D;
        uint64_t c0 = 0, c1 = 0;
        for (uint32_t j = lane; j < n; j += 64) {
          const uint64_t x2 = T[(uint64_t)j * Ru + s];
          const uint64_t z2 = T[(uint64_t)(n + j) * Ru + s];
          const uint64_t x1 = s_src[j], z1 = s_src[n + j];
          phase_acc(x1, z1, x2, z2, c0, c1);
          s_src[j] = x1 ^ (m & x2);
          s_src[n + j] = z1 ^ (m & z2);
        }
#pragma unroll
        for (uint32_t off = 32; off > 0; off >>= 1) {
          const uint64_t t0 = __shfl_xor(c0, off, 64);
          const uint64_t t1 = __shfl_xor(c1, off, 64);
          add_mod4(c0, c1, t0, t1);
        }
        racc ^= m & (T[scratch * Ru + s] ^ c1);
      }
      for (uint32_t j = lane; j < 2 * n; j += 64)
        T[(uint64_t)j * Ru + scratch] = s_src[j];
      if (lane == 0) T[scratch * Ru + scratch] = racc;
      rdet = racc;
    }
  }
  if (tid == 0 && slice == 0) {
    rec_res[w] = (coin & rm) | (rdet & D);
    rec_det[w] = D;
  }
  if (tid == 0 && census) {
    if (slice == 0) {
      atomicAdd(census + 0, rm ? 1ull : 0ull);
      atomicAdd(census + 1, D ? 1ull : 0ull);
      atomicAdd(census + 3, (unsigned long long)nchain);
    }
    atomicAdd(census + 2, (unsigned long long)sh.krows);
  }
  __syncthreads();
}

__global__ __launch_bounds__(kMeasThreads) void k_measure(
    uint64_t *__restrict__ tab, uint32_t n, uint32_t R, uint64_t W, uint32_t q,
    uint32_t coin_T, uint64_t coin_key, const uint64_t *coin_plane,
    uint64_t *rec, uint32_t force_determined) {
  extern __shared__ uint64_t dyn[];
  __shared__ MeasShared sh;
  uint64_t *s_src = dyn;
  uint32_t *s_list = reinterpret_cast<uint32_t *>(dyn + (2 * n + 1));
  const uint64_t C = 2ull * n + 1;
  for (uint64_t w = blockIdx.x; w < W; w += gridDim.x)
    measure_word(tab + w * C * R, w, n, R, q, coin_T, coin_key, coin_plane,
                 rec, rec + W, s_src, s_list, sh, force_determined != 0);
}

// ---------------------------------------------------------------------------
// Host I/O: flat int32 [b][row][col] <-> bit-planes.
// ---------------------------------------------------------------------------
__global__ void k_unpack(const uint64_t *__restrict__ tab, uint32_t n,
                         uint32_t R, uint32_t C, uint64_t b0, uint64_t nb,
                         int32_t *__restrict__ x, int32_t *__restrict__ z,
                         int32_t *__restrict__ r) {
  const uint64_t per = (uint64_t)R * n;
  const uint64_t total = nb * per;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < total;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t b = i / per;
    const uint64_t rem = i - b * per;
    const uint32_t row = (uint32_t)(rem / n);
    const uint32_t col = (uint32_t)(rem - (uint64_t)row * n);
    const uint64_t sb = b0 + b;
    const uint64_t *G = tab + (sb >> 6) * (uint64_t)C * R;
    const uint32_t bit = (uint32_t)(sb & 63);
    x[i] = (int32_t)((G[(uint64_t)col * R + row] >> bit) & 1);
    z[i] = (int32_t)((G[(uint64_t)(n + col) * R + row] >> bit) & 1);
    if (col == 0)
      r[b * R + row] = (int32_t)((G[(uint64_t)(2 * n) * R + row] >> bit) & 1);
  }
}

__global__ void k_pack(uint64_t *__restrict__ tab, uint32_t n, uint32_t R,
                       uint32_t C, uint64_t B, uint64_t w0, uint64_t nw,
                       const int32_t *__restrict__ x,
                       const int32_t *__restrict__ z,
                       const int32_t *__restrict__ r) {
  // x/z/r hold shots [64*w0, 64*w0 + 64*nw) (clipped to B)
  const uint64_t blk = (uint64_t)C * R;
  const uint64_t total = nw * blk;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < total;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t wl = i / blk;
    const uint64_t rem = i - wl * blk;
    const uint32_t c = (uint32_t)(rem / R);
    const uint32_t row = (uint32_t)(rem - (uint64_t)c * R);
    uint64_t word = 0;
    for (uint32_t s = 0; s < 64; ++s) {
      const uint64_t gb = 64 * (w0 + wl) + s;
      if (gb >= B) break;
      const uint64_t lb = gb - 64 * w0;
      int32_t v;
      if (c < n) v = x[(lb * R + row) * n + c];
      else if (c < 2 * n) v = z[(lb * R + row) * n + (c - n)];
      else v = r[lb * R + row];
      word |= (uint64_t)(v & 1) << s;
    }
    tab[(w0 + wl) * blk + rem] = word;
  }
}

// CheckSum (MiscFunctors.hpp:43-49): popcount of x and z over all rows and
// of r once per row (the reference adds r at column 0, so only when n > 0).
__global__ __launch_bounds__(256) void k_checksum(
    const uint64_t *__restrict__ tab, uint32_t n, uint32_t R, uint32_t C,
    uint64_t W, uint64_t b0, uint64_t b1, unsigned long long *out) {
  const uint64_t blk = (uint64_t)C * R;
  const uint64_t total = W * blk;
  unsigned long long acc = 0;
  for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < total;
       i += (uint64_t)gridDim.x * 256ull) {
    const uint64_t w = i / blk;
    const uint64_t rem = i - w * blk;
    const uint32_t c = (uint32_t)(rem / R);
    if (c == 2 * n && n == 0) continue;
    const uint64_t lo = 64 * w, hi = lo + 64;
    const uint64_t a = b0 > lo ? b0 : lo, e = b1 < hi ? b1 : hi;
    if (a >= e) continue;
    const uint32_t sh = (uint32_t)(a - lo), len = (uint32_t)(e - a);
    const uint64_t mask = (len == 64 ? ~0ull : ((1ull << len) - 1)) << sh;
    acc += __popcll(tab[i] & mask);
  }
  __shared__ unsigned long long red[256];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (uint32_t s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0 && red[0]) atomicAdd(out, red[0]);
}

__global__ void k_flags_to_plane(const int32_t *__restrict__ flags, uint64_t B,
                                 uint64_t *__restrict__ plane) {
  const uint64_t W = (B + 63) / 64;
  const uint32_t lane = lane_id();
  const uint64_t waves = (uint64_t)gridDim.x * (blockDim.x / 64);
  for (uint64_t w = blockIdx.x * (uint64_t)(blockDim.x / 64) + (threadIdx.x >> 6); w < W;
       w += waves) {
    const uint64_t b = 64 * w + lane;
    const uint64_t m = __ballot(b < B && flags[b] != 0);
    if (lane == 0) plane[w] = m;
  }
}

__global__ void k_record_to_int(const uint64_t *__restrict__ rec, uint64_t B,
                                uint64_t W, int32_t *__restrict__ res,
                                int32_t *__restrict__ det) {
  for (uint64_t b = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; b < B;
       b += (uint64_t)gridDim.x * blockDim.x) {
    res[b] = (int32_t)((rec[b >> 6] >> (b & 63)) & 1);
    det[b] = (int32_t)((rec[W + (b >> 6)] >> (b & 63)) & 1);
  }
}

// ---------------------------------------------------------------------------
// Persistent program kernel: one workgroup runs a whole compiled op list on
// one 64-shot word group at a time.  Shots never interact, so no grid-wide
// synchronisation is needed, measurements included.
//
// Gate segments: each thread owns rows; r stays in a register for the whole
// segment; instructions come in batches of pairwise qubit-disjoint ops, so a
// batch's operand loads are all issued before any of its stores (memory-level
// parallelism), and noise planes fused into an op are applied to the columns
// that op just produced.  Measurement segments call measure_word.
// ---------------------------------------------------------------------------
__global__ void k_set_params(uint64_t *p, uint64_t seed, uint64_t mseq0,
                             uint64_t nseq0) {
  if (threadIdx.x == 0) {
    p[0] = seed;
    p[1] = mseq0;
    p[2] = nseq0;
  }
}

struct ProgParams {
  uint64_t seed;
  uint64_t mseq0;  // measurement sequence number of the program's first M
  uint64_t nseq0;  // noise sequence number of the program's first NOISE
};

constexpr int kBatch = 8;

// The compiled program and the noise planes are read-only while k_program
// runs; reading them through the constant address space lets uniform loads
// become scalar (SMEM) loads, so a batch's column loads are not serialised
// behind per-op vector loads of its instructions.
#define PQ_CONST __attribute__((address_space(4)))
typedef const PQ_CONST DInstr *cinstr_ptr;
typedef const PQ_CONST uint64_t *cword_ptr;

__device__ __forceinline__ DInstr ld_instr(cinstr_ptr p, uint32_t i) {
  DInstr I;
  I.code = p[i].code;
  I.qa = p[i].qa;
  I.qb = p[i].qb;
  I.slot = p[i].slot;
  I.na = p[i].na;
  I.nb = p[i].nb;
  I.t0 = p[i].t0;
  I.pad = p[i].pad;
  return I;
}

__device__ __forceinline__ DSegment ld_segment(const DSegment *g, uint32_t i) {
  const PQ_CONST DSegment *p = (const PQ_CONST DSegment *)g;
  DSegment sg;
  sg.kind = p[i].kind;
  sg.begin = p[i].begin;
  sg.count = p[i].count;
  sg.nq = p[i].nq;
  sg.qoff = p[i].qoff;
  sg.pad0 = sg.pad1 = sg.pad2 = 0;
  return sg;
}
constexpr uint32_t kSmallQubits = 192;         // the small build's state size bound
constexpr int kSmallRegs = 3, kSmallThreads = 768;
constexpr int kLargeRegs = 8, kLargeThreads = 512;
constexpr uint32_t kMeasBatch = 64;   // most measurements pivot-checked per pass
constexpr uint32_t kMeasWindow0 = 8;  // first window of a run

// Gate segment, processed in chunks of up to kStage instructions.  Each
// chunk's instructions and the per-word shot masks they need (fused noise
// planes, flags) are staged in LDS by the whole workgroup in one round trip;
// after that every batch costs exactly one global round trip (its column
// loads), and the stores are fire-and-forget.
constexpr uint32_t kStage = 448;
constexpr uint32_t kStageWords = kStage * (sizeof(DInstr) / 8 + 4);
// Row-tiled segments stage each op in one word (TileOp) beside its masks.
constexpr uint32_t kTileStageWords = kStage * 5;

// Stages instructions [cb, cb+cn) of a gate segment and the per-word masks
// they need (fused noise planes of qa and qb, or the flag of a flagged op)
// in LDS; a staged instruction's `slot` (read only here) is replaced by the
// kOp* flags below.  No barrier.
constexpr uint32_t kOpNoisy = 1, kOpIdentity = 2;
// the same flags in a row-tiled op's broadcast word of tile offsets
constexpr uint32_t kTileNoisyBit = 1u << 15, kTileIdentityBit = 1u << 31;
// A staged op of a row-tiled segment: `code` is the op code (bits 0-7), its
// kOp* flags (bits 8-9), the length of the run it starts (bits 10-21) and
// kLayerStart (bit 31, the op opens a layer); `t0` its local columns.
struct TileOp {
  uint32_t code, t0;
};
constexpr uint32_t kTileRunShift = 10;

template <bool kCompact>
__device__ __forceinline__ void stage_chunk_t(const pq::ProgramArgs &a, uint64_t w,
                                              uint32_t cb, uint32_t cn, void *s_ops,
                                              uint64_t *s_msk) {
  const uint64_t W = a.W;
  cinstr_ptr ins = (cinstr_ptr)a.instrs;
  // this word group's planes, group-major: plane k's X and Z words at 2k, 2k+1
  cword_ptr gp = (cword_ptr)a.planes + 2ull * a.nk * w;
  const uint64_t *__restrict__ recs = a.recs;
  for (uint32_t t = threadIdx.x; t < cn; t += blockDim.x) {
    const DInstr I = ld_instr(ins, cb + t);
    uint64_t m0 = 0, m1 = 0, m2 = 0, m3 = 0;
    if (I.na != kNone) {
      m0 = gp[2ull * I.na];
      m1 = gp[2ull * I.na + 1];
    }
    if (I.nb != kNone) {
      m2 = gp[2ull * I.nb];
      m3 = gp[2ull * I.nb + 1];
    } else if (I.code >= PQ_OP_H_FLAG && I.code != PQ_OP_NOISE) {
      // flag of a flagged op: a record's results or a random-flag plane
      m2 = (I.slot & kSlotPlane) ? gp[2ull * (I.slot & ~kSlotPlane)]
                                 : recs[(2ull * I.slot) * W + w];
    }
    // for the row-tiled path: whether the op's noise masks are non-zero on
    // this word group, and whether the op is the identity here (a noise op
    // drawing nothing, a flagged op whose flag is clear in all 64 shots);
    // only unflagged CX/CZ carry noise on qb (m2 is a flagged op's flag)
    const bool two = I.code == PQ_OP_CX || I.code == PQ_OP_CZ;
    const bool flagged = I.code >= PQ_OP_H_FLAG && I.code != PQ_OP_NOISE;
    const bool noisy = (m0 | m1) != 0 || (two && (m2 | m3) != 0);
    const bool identity = !noisy && (I.code == PQ_OP_NOISE || (flagged && m2 == 0));
    const uint32_t flags = (noisy ? kOpNoisy : 0u) | (identity ? kOpIdentity : 0u);
    if constexpr (kCompact) {
      TileOp J;
      J.code = I.code | (flags << 8) | (((I.pad >> kRunShift) & kRunMask) << kTileRunShift) |
               (I.pad & kLayerStart);
      J.t0 = I.t0;
      reinterpret_cast<TileOp *>(s_ops)[t] = J;
    } else {
      DInstr J = I;
      J.slot = flags;
      reinterpret_cast<DInstr *>(s_ops)[t] = J;
    }
    s_msk[4 * t + 0] = m0;
    s_msk[4 * t + 1] = m1;
    s_msk[4 * t + 2] = m2;
    s_msk[4 * t + 3] = m3;
  }
}

__device__ __forceinline__ void stage_chunk(const pq::ProgramArgs &a, uint64_t w,
                                            uint32_t cb, uint32_t cn, DInstr *s_ins,
                                            uint64_t *s_msk) {
  stage_chunk_t<false>(a, w, cb, cn, s_ins, s_msk);
}

__device__ void run_gate_segment(uint64_t *__restrict__ T, uint64_t w,
                                 const pq::ProgramArgs &a, const DSegment &sg,
                                 DInstr *s_ins, uint64_t *s_msk, uint32_t slice = 0,
                                 uint32_t slices = 1) {
  const uint64_t Ru = a.R;
  const uint32_t n = a.n;
  const uint32_t seg_end = sg.begin + sg.count;
  for (uint32_t cb = sg.begin; cb < seg_end; cb += kStage) {
    const uint32_t cn = min(kStage, seg_end - cb);
    stage_chunk(a, w, cb, cn, s_ins, s_msk);
    __syncthreads();
    for (uint32_t row = slice * blockDim.x + threadIdx.x; row < a.R;
         row += blockDim.x * slices) {
      uint64_t *Tr = T + row;
      uint64_t r = Tr[(uint64_t)(2 * n) * Ru];
      uint32_t k = 0;
      while (k < cn) {
        const uint32_t len = s_ins[k].pad & kBatchLenMask;
        uint64_t xa[kBatch], za[kBatch], xb[kBatch], zb[kBatch];
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
          if (u < (int)len) {
            const DInstr &I = s_ins[k + u];
            xa[u] = Tr[(uint64_t)I.qa * Ru];
            za[u] = Tr[(uint64_t)(n + I.qa) * Ru];
            if (I.code == PQ_OP_CX || I.code == PQ_OP_CZ || I.code == PQ_OP_CX_FLAG ||
                I.code == PQ_OP_CZ_FLAG) {
              xb[u] = Tr[(uint64_t)I.qb * Ru];
              zb[u] = Tr[(uint64_t)(n + I.qb) * Ru];
            }
          }
        }
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
          if (u < (int)len) {
            const DInstr &I = s_ins[k + u];
            const uint64_t *M = s_msk + 4 * (k + u);
            uint64_t *pxa = Tr + (uint64_t)I.qa * Ru;
            uint64_t *pza = Tr + (uint64_t)(n + I.qa) * Ru;
            uint64_t x = xa[u], z = za[u];
            const uint64_t f = M[2];
            switch (I.code) {
              case PQ_OP_H:
                r ^= x & z;
                *pxa = z; *pza = x;
                { const uint64_t t = x; x = z; z = t; }
                break;
              case PQ_OP_S:
                r ^= x & z;
                z ^= x;
                *pza = z;
                break;
              case PQ_OP_X: r ^= z; break;
              case PQ_OP_Y: r ^= z ^ x; break;
              case PQ_OP_Z: r ^= x; break;
              case PQ_OP_CX:
              case PQ_OP_CX_FLAG: {
                // g: the flag of CX_FLAG (whose m2 is that flag, not noise)
                const bool fl = I.code == PQ_OP_CX_FLAG;
                const uint64_t g = fl ? f : ~0ull, n2 = fl ? 0ull : M[2];
                uint64_t xt = xb[u], zt = zb[u];
                r ^= g & x & zt & ~(xt ^ z);
                xt ^= g & x;
                z ^= g & zt;
                Tr[(uint64_t)I.qb * Ru] = xt;
                *pza = z;
                r ^= (n2 & zt) ^ (M[3] & xt);
                break;
              }
              case PQ_OP_CZ:
              case PQ_OP_CZ_FLAG: {
                const bool fl = I.code == PQ_OP_CZ_FLAG;
                const uint64_t g = fl ? f : ~0ull, n2 = fl ? 0ull : M[2];
                uint64_t xt = xb[u], zt = zb[u];
                r ^= g & x & xt & (zt ^ z);
                zt ^= g & x;
                z ^= g & xt;
                Tr[(uint64_t)(n + I.qb) * Ru] = zt;
                *pza = z;
                r ^= (n2 & zt) ^ (M[3] & xt);
                break;
              }
              case PQ_OP_H_FLAG: {
                const uint64_t d = f & (x ^ z);
                r ^= f & x & z;
                x ^= d; z ^= d;
                *pxa = x; *pza = z;
                break;
              }
              case PQ_OP_S_FLAG:
                r ^= f & x & z;
                z ^= f & x;
                *pza = z;
                break;
              case PQ_OP_X_FLAG: r ^= f & z; break;
              case PQ_OP_Y_FLAG: r ^= f & (z ^ x); break;
              case PQ_OP_Z_FLAG: r ^= f & x; break;
              default: break;  // PQ_OP_NOISE: only the fused plane below
            }
            r ^= (M[0] & z) ^ (M[1] & x);
          }
        }
        k += len;
      }
      Tr[(uint64_t)(2 * n) * Ru] = r;
    }
    __syncthreads();
  }
}

#ifndef PQ_TILE_OPS
#define PQ_TILE_OPS 2
#endif


// Rows of a row-tiled gate segment: the most, up to kTileMaxRows, whose
// tile fits in the launch's LDS next to the staging area, the qubit list and
// the tile's r words (0 if fewer than kTileMinRows fit: the direct form
// runs).  The tile is [2 nq][TR] words, rounded up to whole 32-word DMA
// chunks; each tile column has TR + 1 words (a spare word per column).
constexpr uint32_t kTileMinRows = 8;
constexpr uint32_t kTileMaxRows = 64;
constexpr uint32_t kDmaWords = 32;  // words one wave-wide 4-byte LDS DMA moves
constexpr int kTileOps = PQ_TILE_OPS;

__host__ __device__ __forceinline__ uint64_t tile_words(uint32_t nq, uint32_t col_words) {
  const uint64_t w = 2ull * nq * col_words;
  return (w + kDmaWords - 1) / kDmaWords * kDmaWords;
}
__host__ __device__ __forceinline__ uint32_t tile_rows(uint64_t lds_words, uint32_t nq) {
  if (!nq) return 0;
  const uint64_t fixed = kTileStageWords + (nq + 1) / 2;
  if (lds_words <= fixed + kDmaWords + 2ull * nq) return 0;
  // TR r words + 2 nq (TR + 1) tile words + up to a chunk of rounding
  const uint64_t fit = (lds_words - fixed - kDmaWords - 2ull * nq) / (2ull * nq + 1);
  const uint64_t tr = fit < kTileMaxRows ? fit : kTileMaxRows;
  return tr >= kTileMinRows ? (uint32_t)tr : 0;
}

// Branch-free form of every gate-segment op, for lanes that may hold
// different ops.  Per op code, 8 coefficient bits: A, Bs, Cxz, Cz, Cx (the
// one-qubit update below), F (the op is flagged), CX, CZ.
//   x' = x ^ (A & (x ^ z))          z' = z ^ (A & (x ^ z)) ^ (Bs & x)
//   r ^= (Cxz & x & z) ^ (Cz & z) ^ (Cx & x)
// where each coefficient is the shot flag for a flagged op and all ones
// otherwise: H = A Cxz, S = Bs Cxz, X = Cz, Y = Cz Cx, Z = Cx; NOISE has
// none (only its fused planes).  CX and CZ are masked by the flag too.
enum : uint32_t {
  kCoefA = 1, kCoefBs = 2, kCoefCxz = 4, kCoefCz = 8, kCoefCx = 16,
  kCoefF = 32, kCoefCX = 64, kCoefCZ = 128
};
constexpr uint64_t coef_byte(uint32_t code) {
  return code == PQ_OP_H ? (kCoefA | kCoefCxz)
       : code == PQ_OP_S ? (kCoefBs | kCoefCxz)
       : code == PQ_OP_X ? kCoefCz
       : code == PQ_OP_Y ? (kCoefCz | kCoefCx)
       : code == PQ_OP_Z ? kCoefCx
       : code == PQ_OP_CX ? kCoefCX
       : code == PQ_OP_CZ ? kCoefCZ
       : code == PQ_OP_H_FLAG ? (kCoefA | kCoefCxz | kCoefF)
       : code == PQ_OP_S_FLAG ? (kCoefBs | kCoefCxz | kCoefF)
       : code == PQ_OP_X_FLAG ? (kCoefCz | kCoefF)
       : code == PQ_OP_Y_FLAG ? (kCoefCz | kCoefCx | kCoefF)
       : code == PQ_OP_Z_FLAG ? (kCoefCx | kCoefF)
       : code == PQ_OP_CX_FLAG ? (kCoefCX | kCoefF)
       : code == PQ_OP_CZ_FLAG ? (kCoefCZ | kCoefF)
       : 0;
}
constexpr uint64_t coef_table(uint32_t c0) {
  uint64_t t = 0;
  for (uint32_t c = 0; c < 8; ++c) t |= coef_byte(c0 + c) << (8 * c);
  return t;
}
constexpr uint64_t kCoefLo = coef_table(0), kCoefHi = coef_table(8), kCoefTop = coef_table(16);
__device__ __forceinline__ uint64_t bitmask(uint32_t bits, uint32_t bit) {
  return 0ull - (uint64_t)((bits >> bit) & 1u);
}

// Branch-free op on the row a lane holds: xa/za and xb/zb point at the
// row's entries of the op's two columns in LDS (a one-qubit op names its own
// column as b too) and x, z, xt, zt are their current values; m0..m3 are the
// op's per-word masks (noise planes of a, noise planes of b or the flag).
// Returns the phase contribution.
template <class Ptr>
__device__ __forceinline__ uint64_t tile_op_any(uint32_t code, Ptr pxa, Ptr pza,
                                                Ptr pxb, Ptr pzb, uint64_t x,
                                                uint64_t z, uint64_t xt, uint64_t zt,
                                                uint64_t m0, uint64_t m1, uint64_t m2,
                                                uint64_t m3) {
  const uint64_t coefs = code < 8 ? kCoefLo : (code < 16 ? kCoefHi : kCoefTop);
  const uint32_t cf = (uint32_t)((coefs >> (8 * (code & 7))) & 0xFFu);
  const uint64_t F = m2 | ~bitmask(cf, 5);  // the flag, or all ones
  const uint64_t CX = bitmask(cf, 6) & F, CZ = bitmask(cf, 7) & F;
  // noise planes on qb: unflagged CX/CZ only
  const uint64_t two = (bitmask(cf, 6) | bitmask(cf, 7)) & ~bitmask(cf, 5);
  const uint64_t d = bitmask(cf, 0) & F & (x ^ z);
  const uint64_t x2 = x ^ d;
  const uint64_t z2 = z ^ d ^ (bitmask(cf, 1) & F & x) ^ (CX & zt) ^ (CZ & xt);
  const uint64_t xt2 = xt ^ (CX & x), zt2 = zt ^ (CZ & x);
  uint64_t rp = F & ((bitmask(cf, 2) & x & z) ^ (bitmask(cf, 3) & z) ^ (bitmask(cf, 4) & x));
  rp ^= (CX & x & zt & ~(xt ^ z)) ^ (CZ & x & xt & (zt ^ z));
  rp ^= (m0 & z2) ^ (m1 & x2) ^ (two & ((m2 & zt2) ^ (m3 & xt2)));
  *pxb = xt2;
  *pzb = zt2;
  *pxa = x2;
  *pza = z2;
  return rp;
}

// The same op when every lane of the wave holds it: `code` is wave-uniform,
// so the switch compiles to scalar branches.  x, z (and xt, zt) are the
// row's current values of the op's columns, already loaded.
template <bool kNoise, class Ptr>
__device__ __forceinline__ uint64_t tile_op_uniform(uint32_t code, Ptr pxa, Ptr pza,
                                                    Ptr pxb, Ptr pzb, uint64_t x,
                                                    uint64_t z, uint64_t xt, uint64_t zt,
                                                    uint64_t m0, uint64_t m1, uint64_t m2,
                                                    uint64_t m3) {
  uint64_t rp = 0;
  const uint64_t f = m2;
  switch (code) {
    case PQ_OP_H:
      rp = x & z;
      *pxa = z; *pza = x;
      { const uint64_t t = x; x = z; z = t; }
      break;
    case PQ_OP_S:
      rp = x & z;
      z ^= x;
      *pza = z;
      break;
    case PQ_OP_X: rp = z; break;
    case PQ_OP_Y: rp = z ^ x; break;
    case PQ_OP_Z: rp = x; break;
    case PQ_OP_CX:
      rp = x & zt & ~(xt ^ z);
      xt ^= x;
      z ^= zt;
      *pxb = xt;
      *pza = z;
      if (kNoise) rp ^= (m2 & zt) ^ (m3 & xt);
      break;
    case PQ_OP_CZ:
      rp = x & xt & (zt ^ z);
      zt ^= x;
      z ^= xt;
      *pzb = zt;
      *pza = z;
      if (kNoise) rp ^= (m2 & zt) ^ (m3 & xt);
      break;
    case PQ_OP_H_FLAG: {
      const uint64_t d = f & (x ^ z);
      rp = f & x & z;
      x ^= d; z ^= d;
      *pxa = x; *pza = z;
      break;
    }
    case PQ_OP_S_FLAG:
      rp = f & x & z;
      z ^= f & x;
      *pza = z;
      break;
    case PQ_OP_X_FLAG: rp = f & z; break;
    case PQ_OP_Y_FLAG: rp = f & (z ^ x); break;
    case PQ_OP_Z_FLAG: rp = f & x; break;
    case PQ_OP_CX_FLAG:
      rp = f & x & zt & ~(xt ^ z);
      xt ^= f & x;
      z ^= f & zt;
      *pxb = xt;
      *pza = z;
      break;
    case PQ_OP_CZ_FLAG:
      rp = f & x & xt & (zt ^ z);
      zt ^= f & x;
      z ^= f & xt;
      *pzb = zt;
      *pza = z;
      break;
    default: break;  // PQ_OP_NOISE: only the fused plane below
  }
  return kNoise ? rp ^ (m0 & z) ^ (m1 & x) : rp;
}

typedef __attribute__((address_space(3))) uint64_t lds_u64;

// A wave's share of one run, one op at a time on every lane's tile row (the
// spare word for lanes without a row).  Op o of the share has its code in
// lane o of my_code and, in lane o of my_t0, the tile word offsets of its qa
// and qb columns (bits 0-14 and 16-30; x column, z follows TRp words on) and
// its flags (kTileNoisyBit, kTileIdentityBit); its masks are at
// s_msk[4 * (i0 + o * S)].  CODE != 0 when every op of the share has that
// code: the dispatch is then resolved at compile time, and my_code is not read.
template <uint32_t CODE>
__device__ __forceinline__ uint64_t tile_run_ops(lds_u64 *tl, uint32_t TRp, uint32_t lr,
                                                 uint32_t my_code, uint32_t my_t0, uint32_t cnt,
                                                 const uint64_t *s_msk, uint32_t i0,
                                                 uint32_t S) {
  uint64_t rp = 0;
  for (uint32_t o = 0; o < cnt; ++o) {
    const uint32_t t0 = __builtin_amdgcn_readlane(my_t0, o);
    // only noise-only and flagged ops can be the identity on a word group
    constexpr bool kMayBeIdentity = CODE == 0 || CODE >= PQ_OP_H_FLAG;
    if (kMayBeIdentity && (t0 & kTileIdentityBit)) continue;  // nothing to do here
    const uint32_t code = CODE ? CODE : __builtin_amdgcn_readlane(my_code, o);
    lds_u64 *pa = tl + ((t0 & 0x7FFFu) + lr);
    lds_u64 *pb = tl + (((t0 >> 16) & 0x7FFFu) + lr);
    const uint64_t vx = pa[0], vz = pa[TRp];
    uint64_t vxt = 0, vzt = 0;
    if (code == PQ_OP_CX || code == PQ_OP_CZ || code == PQ_OP_CX_FLAG ||
        code == PQ_OP_CZ_FLAG) {
      vxt = pb[0];
      vzt = pb[TRp];
    }
    const uint64_t *M = s_msk + 4 * (i0 + o * S);
    if (t0 & kTileNoisyBit) {
      rp ^= tile_op_uniform<true>(code, pa, pa + TRp, pb, pb + TRp, vx, vz, vxt, vzt, M[0], M[1],
                                  M[2], M[3]);
    } else {
      // flagged ops still need their flag
      const uint64_t f = (code >= PQ_OP_H_FLAG && code != PQ_OP_NOISE) ? M[2] : 0ull;
      rp ^= tile_op_uniform<false>(code, pa, pa + TRp, pb, pb + TRp, vx, vz, vxt, vzt, 0ull,
                                   0ull, f, 0ull);
    }
  }
  return rp;
}

// Row-tiled gate segment.  Gates act on every row independently, so the
// segment runs tile by tile over rows [r0, r0+TR): the x and z columns of
// the segment's nq qubits (local column 2j = x, 2j+1 = z of qubit segq[j])
// are loaded into LDS once, every op of the segment is applied there, and
// the columns are written back once.  Each column holds TR + 1 words: TR
// rows and a spare word that lanes without a row of this tile work on, so
// the op loop needs no lane masking.  A wave holds P ops' worth of rows (one
// op for TR >= 32, else 64 / TR): lane l works on row l % TR of op slot
// wave * P + l / TR, and the S = waves * P slots take ops slot, slot + S,
// ... of each run (a layer, or its part in one staging chunk); ops of one
// layer touch distinct qubits, so a run needs one barrier.  Phase
// contributions never feed back into x or z, so each lane accumulates its
// own and they are XOR-combined into r at the end of the tile.
__device__ void run_gate_segment_tiled(uint64_t *__restrict__ T, uint64_t w,
                                       const pq::ProgramArgs &a, const DSegment &sg,
                                       uint64_t *dyn, uint32_t TR, uint32_t slice = 0,
                                       uint32_t slices = 1, bool fresh = false) {
  const uint64_t Ru = a.R;
  const uint32_t n = a.n;
  const uint32_t nq = sg.nq;
  const uint32_t TRp = TR + 1;  // words per tile column
  TileOp *s_ins = reinterpret_cast<TileOp *>(dyn);
  uint64_t *s_msk = dyn + kStage;
  uint32_t *s_q = reinterpret_cast<uint32_t *>(dyn + kTileStageWords);
  uint64_t *s_r = dyn + kTileStageWords + (nq + 1) / 2;
  uint64_t *tile = s_r + TR;
  lds_u64 *tl = (lds_u64 *)tile;
  const uint32_t seg_end = sg.begin + sg.count;
  const bool one_chunk = sg.count <= kStage;
  for (uint32_t j = threadIdx.x; j < nq; j += blockDim.x) s_q[j] = a.segq[sg.qoff + j];
  if (one_chunk) stage_chunk_t<true>(a, w, sg.begin, sg.count, s_ins, s_msk);
  const uint32_t lane = lane_id();
  // tiles of 32+ rows give each wave one op (wave-uniform control flow)
  const uint32_t P = TR >= 32 ? 1 : 64 / TR;
  const uint32_t wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), nwv = blockDim.x >> 6;
  const uint32_t sub = lane / TR, ti = lane - sub * TR;
  const uint32_t slot = wv * P + sub;
  const uint32_t S = nwv * P;
  const uint32_t ncol = 2 * nq;
  const uint32_t nwords = ncol * TRp;  // tile words
  __syncthreads();
  for (uint32_t r0 = slice * TR; r0 < a.R; r0 += TR * slices) {
    const uint32_t TRv = min(TR, a.R - r0);
    const bool active = sub < P && ti < TRv;
    const uint32_t lr = active ? ti : TR;  // this lane's word in a column
    // The tile arrives by LDS DMA: each wave-wide 4-byte global_load_lds
    // fills one 32-word chunk of the tile (lane l carries half l&1 of word
    // l/2), with no register round trip, so a wave keeps dozens of chunk
    // loads in flight.  Spare words and rows past the tableau's last row
    // load that last row (they are never stored back).
    if (fresh) {
      // X/Z are still the identity (word (c, row) all ones iff row == c) and
      // not in memory: the tile is written here, not loaded
      // (a wave per column; the rounding words past the last column are
      // never read)
      const uint32_t last = a.R - 1 - r0;
      for (uint32_t c = wv; c < ncol; c += nwv) {
        const uint32_t q = s_q[c >> 1];
        const uint32_t col = (c & 1) ? n + q : q;
        for (uint32_t j = lane; j < TRp; j += 64)
          tile[c * TRp + j] = (r0 + min(j, last) == col) ? ~0ull : 0ull;
      }
      if (slot == 0 && active) s_r[ti] = T[(uint64_t)(2 * n) * Ru + r0 + ti];
    } else {
      const uint32_t nchunk = (uint32_t)(tile_words(nq, TRp) / kDmaWords);
      const uint32_t last = a.R - 1 - r0;
      const uint32_t step = kDmaWords * nwv;  // tile words between a lane's chunks
      const uint32_t dcc = step / TRp, dj = step - dcc * TRp;
      uint32_t t = wv * kDmaWords + (lane >> 1);
      uint32_t cc = t / TRp, j = t - cc * TRp;
      for (uint32_t ch = wv; ch < nchunk; ch += nwv) {
        const uint32_t c = min(cc, ncol - 1);
        const uint32_t q = s_q[c >> 1];
        const uint32_t off = ((c & 1) ? n + q : q) * (uint32_t)Ru + r0 + min(j, last);
        const uint32_t *src = reinterpret_cast<const uint32_t *>(T + off) + (lane & 1);
        __builtin_amdgcn_global_load_lds(
            src, (__attribute__((address_space(3))) void *)(tile + (uint64_t)ch * kDmaWords), 4,
            0, 0);
        cc += dcc;
        j += dj;
        if (j >= TRp) {
          j -= TRp;
          ++cc;
        }
      }
      if (slot == 0 && active) s_r[ti] = T[(uint64_t)(2 * n) * Ru + r0 + ti];
      asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    }
    __syncthreads();
    uint64_t rp = 0;
    for (uint32_t cb = sg.begin; cb < seg_end; cb += kStage) {
      const uint32_t cn = min(kStage, seg_end - cb);
      if (!one_chunk) {
        stage_chunk_t<true>(a, w, cb, cn, s_ins, s_msk);
        __syncthreads();
      }
      uint32_t k = 0;
      while (k < cn) {
        const uint32_t e = k + ((s_ins[k].code >> kTileRunShift) & kRunMask);
        if (P == 1) {
          // One op per wave at a time, wave-uniform.  The wave's ops of the
          // run (k + wv, k + wv + S, ...) are fetched in one LDS round trip,
          // one per lane, then broadcast with readlane; ops that are the
          // identity on this word group are skipped and ops whose noise drew
          // nothing here skip the noise terms.  Every lane runs every op
          // (lanes without a row on the spare word).
          const uint32_t nmine = (e - k > wv) ? (e - k - wv + S - 1) / S : 0;
          for (uint32_t base = 0; base < nmine; base += 64) {
            uint32_t my_code = 0, my_t0 = 0;
            if (base + lane < nmine) {
              const uint32_t idx = k + wv + (base + lane) * S;
              const uint32_t cw = s_ins[idx].code;
              my_code = cw & 0xFFu;
              // local columns -> tile word offsets (a tile holds fewer
              // than 2^15 words: LDS is 160 KB), flags in bits 15 and 31
              const uint32_t t = s_ins[idx].t0, f = (cw >> 8) & 3u;
              my_t0 = (2u * (t & 0xFFFFu) * TRp) | ((2u * (t >> 16) * TRp) << 16) |
                      ((f & kOpNoisy) ? kTileNoisyBit : 0u) |
                      ((f & kOpIdentity) ? kTileIdentityBit : 0u);
            }
            const uint32_t cnt = min(64u, nmine - base);
            const uint32_t i0 = k + wv + base * S;
            // one code for the whole share (the usual case: a layer of one
            // gate kind): dispatch once, outside the op loop
            const uint32_t c0 = __builtin_amdgcn_readlane(my_code, 0);
            const bool uni = __ballot(lane < cnt && my_code != c0) == 0;
            if (uni && c0 == PQ_OP_CX)
              rp ^= tile_run_ops<PQ_OP_CX>(tl, TRp, lr, my_code, my_t0, cnt, s_msk, i0, S);
            else if (uni && c0 == PQ_OP_H)
              rp ^= tile_run_ops<PQ_OP_H>(tl, TRp, lr, my_code, my_t0, cnt, s_msk, i0, S);
            else if (uni && c0 == PQ_OP_CZ)
              rp ^= tile_run_ops<PQ_OP_CZ>(tl, TRp, lr, my_code, my_t0, cnt, s_msk, i0, S);
            else if (uni && c0 == PQ_OP_S)
              rp ^= tile_run_ops<PQ_OP_S>(tl, TRp, lr, my_code, my_t0, cnt, s_msk, i0, S);
            else if (uni && c0 == PQ_OP_X_FLAG)
              rp ^= tile_run_ops<PQ_OP_X_FLAG>(tl, TRp, lr, my_code, my_t0, cnt, s_msk, i0, S);
            else if (uni && c0 == PQ_OP_NOISE)
              rp ^= tile_run_ops<PQ_OP_NOISE>(tl, TRp, lr, my_code, my_t0, cnt, s_msk, i0, S);
            else
              rp ^= tile_run_ops<0>(tl, TRp, lr, my_code, my_t0, cnt, s_msk, i0, S);
          }
        } else if (active) {
          lds_u64 *Tl = tl + ti;
          for (uint32_t i0 = k + slot; i0 < e; i0 += kTileOps * S) {
            uint32_t code[kTileOps], idxs[kTileOps];
            lds_u64 *pa[kTileOps], *pb[kTileOps];
            uint64_t vx[kTileOps], vz[kTileOps], vxt[kTileOps], vzt[kTileOps];
#pragma unroll
            for (int u = 0; u < kTileOps; ++u) {
              idxs[u] = i0 + u * S;
              code[u] = 0;
              if (idxs[u] < e) {
                code[u] = s_ins[idxs[u]].code & 0xFFu;
                const uint32_t t0 = s_ins[idxs[u]].t0;
                pa[u] = Tl + 2 * (t0 & 0xFFFFu) * TRp;
                pb[u] = Tl + 2 * (t0 >> 16) * TRp;
                vx[u] = pa[u][0];
                vz[u] = pa[u][TRp];
                vxt[u] = pb[u][0];
                vzt[u] = pb[u][TRp];
              }
            }
#pragma unroll
            for (int u = 0; u < kTileOps; ++u) {
              if (idxs[u] < e) {
                const uint64_t *M = s_msk + 4 * idxs[u];
                rp ^= tile_op_any(code[u], pa[u], pa[u] + TRp, pb[u], pb[u] + TRp, vx[u], vz[u],
                                  vxt[u], vzt[u], M[0], M[1], M[2], M[3]);
              }
            }
          }
        }
        // runs of one layer act on distinct qubits: a barrier only before
        // a new layer and at the end of the staged chunk
        if (e == cn || (s_ins[e].code & kLayerStart)) __syncthreads();
        k = e;
      }
    }
    if (active && rp) atomicXor(reinterpret_cast<unsigned long long *>(s_r + ti), rp);
    __syncthreads();
    {
      // write back: consecutive threads take consecutive tile words
      const uint32_t step = blockDim.x;
      const uint32_t dcc = step / TRp, dj = step - dcc * TRp;
      uint32_t cc = threadIdx.x / TRp, j = threadIdx.x - cc * TRp;
      for (uint32_t t = threadIdx.x; t < nwords; t += step) {
        if (j < TRv) {
          const uint32_t q = s_q[cc >> 1];
          T[((cc & 1) ? n + q : q) * (uint32_t)Ru + r0 + j] = tl[t];
        }
        cc += dcc;
        j += dj;
        if (j >= TRp) {
          j -= TRp;
          ++cc;
        }
      }
      if (slot == 0 && active) T[(uint64_t)(2 * n) * Ru + r0 + ti] = s_r[ti];
    }
    __syncthreads();
  }
}

__device__ __forceinline__ uint64_t wave_or(uint64_t v) {
#pragma unroll
  for (uint32_t off = 32; off > 0; off >>= 1) v |= __shfl_xor(v, off, 64);
  return v;
}

// MeasureDetermined (MeasurementFunctors.hpp:190-201) of a measurement that
// is determined in all 64 shots of the group, run by one wave.  Chain
// members (destabilizers i with x_q set, ascending) are compacted with
// ballots; each step folds stabilizer n+i into the scratch row and reduces
// that step's mod-4 phase across the wave.  Only the last measurement of a
// batch writes its scratch row back to the tableau.
//
// n <= 64*NREG: each lane owns at most NREG columns, the scratch row lives
// in registers and the next step's source words are loaded while the
// current step is folded in.  Larger n: scratch row in the wave's LDS slab.
// NREG sets the kernel's register footprint (see k_program_t).
template <int NREG>
__device__ void determined_wave(uint64_t *__restrict__ T, uint64_t w, uint32_t n,
                                uint32_t R, uint32_t q, uint64_t *slab,
                                uint64_t *rec, uint64_t W, bool write_scratch,
                                unsigned long long *census) {
  const uint32_t lane = lane_id();
  const uint64_t Ru = R;
  const uint64_t *xq = T + (uint64_t)q * Ru;
  const uint64_t scratch = 2ull * n;
  const uint64_t *rq = T + scratch * Ru;
  uint64_t racc = 0;
  uint32_t steps = 0;
  if (n <= 64u * NREG) {
    // chain list into the wave's slab: rows in slab[0..], masks in slab[n..]
    uint32_t count = 0;
    for (uint32_t base = 0; base < n; base += 4 * 64) {
      uint64_t xi[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const uint32_t i = base + 64 * u + lane;
        xi[u] = i < n ? xq[i] : 0ull;
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const uint64_t bal = __ballot(xi[u] != 0);
        const uint32_t pos = count + (uint32_t)__popcll(bal & ((1ull << lane) - 1));
        if (xi[u]) {
          slab[pos] = base + 64 * u + lane;
          slab[n + pos] = xi[u];
        }
        count += (uint32_t)__popcll(bal);
      }
    }
    steps = count;
    uint64_t sx[NREG], sz[NREG], nx[NREG], nz[NREG];
#pragma unroll
    for (int u = 0; u < NREG; ++u) sx[u] = sz[u] = nx[u] = nz[u] = 0;
    uint64_t nm = 0, nr = 0;
    if (count) {
      const uint64_t s = n + slab[0];
      nm = slab[n];
#pragma unroll
      for (int u = 0; u < NREG; ++u) {
        const uint32_t j = lane + 64 * u;
        if (j < n) {
          nx[u] = T[(uint64_t)j * Ru + s];
          nz[u] = T[(uint64_t)(n + j) * Ru + s];
        }
      }
      nr = rq[s];
    }
    for (uint32_t t = 0; t < count; ++t) {
      uint64_t cx[NREG], cz[NREG];
#pragma unroll
      for (int u = 0; u < NREG; ++u) {
        cx[u] = nx[u];
        cz[u] = nz[u];
      }
      const uint64_t m = nm, rs = nr;
      if (t + 1 < count) {
        const uint64_t s = n + slab[t + 1];
        nm = slab[n + t + 1];
#pragma unroll
        for (int u = 0; u < NREG; ++u) {
          const uint32_t j = lane + 64 * u;
          if (j < n) {
            nx[u] = T[(uint64_t)j * Ru + s];
            nz[u] = T[(uint64_t)(n + j) * Ru + s];
          }
        }
        nr = rq[s];
      }
      uint64_t c0 = 0, c1 = 0;
#pragma unroll
      for (int u = 0; u < NREG; ++u) {
        phase_acc(sx[u], sz[u], cx[u], cz[u], c0, c1);
        sx[u] ^= m & cx[u];
        sz[u] ^= m & cz[u];
      }
#pragma unroll
      for (uint32_t off = 32; off > 0; off >>= 1) {
        const uint64_t t0 = __shfl_xor(c0, off, 64);
        const uint64_t t1 = __shfl_xor(c1, off, 64);
        add_mod4(c0, c1, t0, t1);
      }
      racc ^= m & (rs ^ c1);
    }
    if (write_scratch) {
#pragma unroll
      for (int u = 0; u < NREG; ++u) {
        const uint32_t j = lane + 64 * u;
        if (j < n) {
          T[(uint64_t)j * Ru + scratch] = sx[u];
          T[(uint64_t)(n + j) * Ru + scratch] = sz[u];
        }
      }
    }
  } else {
    for (uint32_t j = lane; j < n; j += 64) {
      slab[j] = 0;
      slab[n + j] = 0;
    }
    for (uint32_t base = 0; base < n; base += 64) {
      const uint32_t i = base + lane;
      const uint64_t xi = i < n ? xq[i] : 0ull;
      uint64_t bal = __ballot(xi != 0);
      while (bal) {
        const uint32_t t = (uint32_t)__ffsll((long long)bal) - 1;
        bal &= bal - 1;
        const uint64_t m = __shfl(xi, t, 64);
        const uint64_t s = n + base + t;
        uint64_t c0 = 0, c1 = 0;
        for (uint32_t j = lane; j < n; j += 64) {
          const uint64_t x2 = T[(uint64_t)j * Ru + s];
          const uint64_t z2 = T[(uint64_t)(n + j) * Ru + s];
          const uint64_t x1 = slab[j], z1 = slab[n + j];
          phase_acc(x1, z1, x2, z2, c0, c1);
          slab[j] = x1 ^ (m & x2);
          slab[n + j] = z1 ^ (m & z2);
        }
#pragma unroll
        for (uint32_t off = 32; off > 0; off >>= 1) {
          const uint64_t t0 = __shfl_xor(c0, off, 64);
          const uint64_t t1 = __shfl_xor(c1, off, 64);
          add_mod4(c0, c1, t0, t1);
        }
        racc ^= m & (rq[s] ^ c1);
        ++steps;
      }
    }
    if (write_scratch) {
      for (uint32_t j = lane; j < n; j += 64) {
        T[(uint64_t)j * Ru + scratch] = slab[j];
        T[(uint64_t)(n + j) * Ru + scratch] = slab[n + j];
      }
    }
  }
  if (lane == 0) {
    if (rec) {
      rec[w] = racc;
      rec[W + w] = ~0ull;
    }
    if (census) {
      atomicAdd(census + 1, 1ull);
      atomicAdd(census + 3, (unsigned long long)steps);
    }
    if (write_scratch) T[scratch * Ru + scratch] = racc;
  }
}

// Dynamic LDS of k_program: s_src[2n+1] u64 | s_list[n+1] u32 (padded) |
// s_rm[kMeasBatch] u64 | ndw slabs of 2n u64.
__host__ __device__ inline size_t program_lds_words(uint32_t n, uint32_t ndw) {
  return (2ull * n + 1) + ((uint64_t)n + 2) / 2 + kMeasBatch + (uint64_t)ndw * 2 * n;
}

// Wall-clock ticks per phase (census mode, thread 0).
struct PhaseTimes {
  uint64_t gate = 0, check = 0, det = 0, gen = 0;
};

// The whole compiled program on 64-shot word group w of one state.
template <int NREG>
__device__ void program_word(const pq::ProgramArgs &a, uint64_t seed, uint64_t mseq0,
                             uint64_t w, uint64_t *dyn, MeasShared &sh,
                             PhaseTimes *tm, SliceSync *ss = nullptr) {
  const uint32_t slice = ss ? ss->slice : 0, S = ss ? ss->slices : 1;
  const uint32_t n = a.n;
  uint64_t *s_src = dyn;
  uint32_t *s_list = reinterpret_cast<uint32_t *>(dyn + (2 * n + 1));
  uint64_t *s_rm = dyn + (2 * n + 1) + (n + 2) / 2;
  uint64_t *s_slab = s_rm + kMeasBatch;
  const uint64_t C = 2ull * n + 1;
  const uint64_t Ru = a.R;
  const uint32_t lane = lane_id(), wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
  // waves running determined measurements: as many as the arguments size
  // LDS slabs for, but no more than the launched block has
  const uint32_t ndw = min(a.det_waves, nw);
  const bool timing = tm != nullptr;
  uint64_t t0 = 0;
  uint64_t *T = a.tab + w * C * a.R;
  for (uint32_t si = 0; si < a.nseg; ++si) {
    const DSegment sg = ld_segment(a.segs, si);
    if (sg.kind == 0) {
      if (timing) t0 = wall_clock64();
      const uint32_t TR = tile_rows(a.lds_words, sg.nq);
      if (TR)
        run_gate_segment_tiled(T, w, a, sg, dyn, TR, slice, S, a.fresh && si == 0);
      else
        run_gate_segment(T, w, a, sg, reinterpret_cast<DInstr *>(dyn),
                         dyn + kStage * (sizeof(DInstr) / 8), slice, S);
      // every slice's rows are in before a measurement reads whole columns
      if (S > 1) slice_barrier(*ss);
      if (timing) tm->gate += wall_clock64() - t0;
      continue;
    }
    uint32_t k = sg.begin;
    const uint32_t end = sg.begin + sg.count;
    // Windows are whole multiples of the waves that run determined
    // measurements (every slice's), at least kMeasWindow0, so a batch of
    // determined measurements fills its waves.
    const uint32_t dw = ndw * S;
    const uint32_t wunit = min(kMeasBatch, dw * ((kMeasWindow0 + dw - 1) / dw));
    const uint32_t wmax = (kMeasBatch / wunit) * wunit;
    uint32_t win = wunit;
    while (k < end) {
      if (timing) t0 = wall_clock64();
      // which of the next measurements have a pivot in some shot?  The window
      // grows while every checked measurement is determined and restarts
      // small after a random one (the state then changes).  One wave per
      // measurement.
      const uint32_t nm = min(end - k, win);
      for (uint32_t idx = wv; idx < nm; idx += nw) {
        const uint64_t *xq = T + (uint64_t)((cinstr_ptr)a.instrs)[k + idx].qa * Ru;
        uint64_t acc = 0;
        for (uint32_t p0 = 0; p0 < n; p0 += 4 * 64) {
          uint64_t v[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const uint32_t p = p0 + 64 * u + lane;
            v[u] = p < n ? xq[n + p] : 0ull;
          }
          acc |= v[0] | v[1] | v[2] | v[3];
        }
        acc = wave_or(acc);
        if (lane == 0) s_rm[idx] = acc;
      }
      __syncthreads();
      // first measurement with a pivot: one LDS read per lane and a ballot
      // (nm <= kMeasBatch = 64 = lanes)
      const uint64_t piv = __ballot(lane < nm && s_rm[lane] != 0);
      const uint32_t jr = piv ? (uint32_t)__ffsll((long long)piv) - 1 : nm;
      win = (jr == nm) ? min(2 * win, wmax) : wunit;
      if (timing) {
        const uint64_t t1 = wall_clock64();
        tm->check += t1 - t0;
        t0 = t1;
      }
      // measurements [k, k+jr) are determined in every shot: one wave each
      if (jr > 0) {
        if (wv < ndw) {
          uint64_t *slab = s_slab + (uint64_t)wv * 2 * n;
          // the slices' determined waves share the batch
          for (uint32_t idx = slice * ndw + wv; idx < jr; idx += ndw * S) {
            const DInstr I = ld_instr((cinstr_ptr)a.instrs, k + idx);
            // a record that a later measurement of the segment overwrites is
            // not written: waves of one batch are not ordered
            uint64_t *rec =
                (I.pad & kMeasSuperseded) ? nullptr : a.recs + (2ull * I.slot) * a.W;
            determined_wave<NREG>(T, w, n, a.R, I.qa, slab, rec, a.W, idx == jr - 1,
                                  a.census);
          }
        }
        __syncthreads();
        // the scratch row and records of the batch are in before what follows
        if (S > 1) slice_barrier(*ss);
      }
      if (timing) {
        const uint64_t t1 = wall_clock64();
        tm->det += t1 - t0;
        t0 = t1;
      }
      if (jr < nm) {
        const DInstr I = ld_instr((cinstr_ptr)a.instrs, k + jr);
        const uint64_t m = (uint64_t)I.na | ((uint64_t)I.nb << 32);
        const uint64_t key = stream_key(seed, 2 * (mseq0 + m));
        const uint64_t *cp = a.coin_planes ? a.coin_planes + m * a.W : nullptr;
        uint64_t *rec = a.recs + (2ull * I.slot) * a.W;
        measure_word(T, w, n, a.R, I.qa, I.t0, key, cp, rec, rec + a.W, s_src,
                     s_list, sh, false, a.census, ss);
        // slice 0's pivot rows, record and scratch row are in for every slice
        if (S > 1) slice_barrier(*ss);
        k += jr + 1;
        if (timing) tm->gen += wall_clock64() - t0;
      } else {
        k += nm;
      }
    }
  }
}

// Two builds of the persistent kernels: up to 192 qubits, determined chains
// keep 3 columns per lane in registers, which leaves room for 768-thread
// workgroups (12 waves); 193..512 qubits keep 8 per lane in 512-thread
// workgroups; above 512 the chains' scratch row is in LDS and the 768-thread
// build runs.
template <int NREG, int MAXT, bool SLICED>
__global__ __launch_bounds__(MAXT) void k_program_t(pq::ProgramArgs a,
                                                   const ProgParams *__restrict__ pp) {
  extern __shared__ uint64_t dyn[];
  __shared__ MeasShared sh;
  __shared__ unsigned long long s_census[4];
  const uint64_t seed = pp->seed, mseq0 = pp->mseq0;
  // census mode: work counts gather in LDS (one global add per workgroup at
  // the end, so counting does not contend) and thread 0 times the phases
  // (wall-clock ticks)
  unsigned long long *census = a.census;
  pq::ProgramArgs la = a;
  if (census) {
    if (threadIdx.x < 4) s_census[threadIdx.x] = 0;
    __syncthreads();
    la.census = s_census;
  }
  PhaseTimes times;
  PhaseTimes *tm = (census != nullptr && threadIdx.x == 0) ? &times : nullptr;
  if constexpr (SLICED) {
    // cooperative launch: one workgroup per (word group, slice), all resident
    const uint64_t wi = blockIdx.x / a.slices;
    if (wi < a.nwords) {
      SliceSync ss;
      ss.slice = blockIdx.x % a.slices;
      ss.slices = a.slices;
      ss.ctr = a.sync + wi;
      ss.err = a.sync_err;
      if (a.sync_spin) ss.spin_limit = a.sync_spin;
      program_word<NREG>(la, seed, mseq0, a.word0 + wi, dyn, sh, tm, &ss);
    }
  } else {
    for (uint64_t wi = blockIdx.x; wi < a.nwords; wi += gridDim.x)
      program_word<NREG>(la, seed, mseq0, a.word0 + wi, dyn, sh, tm);
  }
  if (census) {
    __syncthreads();
    if (threadIdx.x < 4) atomicAdd(census + threadIdx.x, s_census[threadIdx.x]);
  }
  if (tm) {
    atomicAdd(census + 4, (unsigned long long)times.gate);
    atomicAdd(census + 5, (unsigned long long)times.check);
    atomicAdd(census + 6, (unsigned long long)times.det);
    atomicAdd(census + 7, (unsigned long long)times.gen);
  }
}

// Mixed batch: many states, each with its own compiled program, in one
// persistent launch.  Work items (state, word group), heaviest first, are
// taken from a global counter until the list is exhausted; every workgroup
// leaves the loop once the counter passes the item count.
template <int NREG, int MAXT>
__global__ __launch_bounds__(MAXT) void k_batch_t(
    const pq::BatchCircuit *__restrict__ circ, const uint64_t *__restrict__ items,
    uint64_t nitems, unsigned long long *counter, uint32_t wtile) {
  extern __shared__ uint64_t dyn[];
  __shared__ MeasShared sh;
  __shared__ uint64_t s_item;
  for (;;) {
    if (threadIdx.x == 0) s_item = atomicAdd(counter, 1ull);
    __syncthreads();
    const uint64_t it = s_item;
    __syncthreads();
    if (it >= nitems) break;
    const uint64_t code = items[it];
    const uint32_t c = (uint32_t)(code >> 40);
    const uint64_t w = code & ((1ull << 40) - 1);
    const pq::BatchCircuit &bc = circ[c];
    // an item is `wtile` consecutive word groups of one circuit (fewer at
    // the circuit's end)
    const uint64_t wend = min(w + wtile, bc.a.nwords);
    for (uint64_t wi = w; wi < wend; ++wi)
      program_word<NREG>(bc.a, bc.seed, bc.mseq0, wi, dyn, sh, nullptr);
  }
}

// Noise planes of every circuit of a batch (blockIdx.y = circuit).  The
// block row's waves are laid out as (noise-op lane) x (word-group lane): one
// stream key per op and wave, one draw per lane per (op, word group).
// Lane `lane` of v takes the wave-uniform value val (v_writelane_b32; the
// lane select goes through M0, so the instruction reads one SGPR).
// The kernels that call it keep nothing in M0 (no LDS DMA), so the clobber
// only documents the write.
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Winline-asm"
__device__ __forceinline__ uint32_t write_lane(uint32_t v, uint32_t val, uint32_t lane) {
  asm volatile("s_mov_b32 m0, %2\n\ts_nop 0\n\tv_writelane_b32 %0, %1, m0"
               : "+v"(v)
               : "s"(val), "s"(lane)
               : "m0");
  return v;
}
#pragma clang diagnostic pop

// Group-major noise planes [W][nk][2] (the frame form).  One wave task
// covers word group w and noise ops [k0, k0 + 32): lanes 2i and 2i+1 end up
// holding op k0 + i's X and Z mask words, so a task writes one contiguous run
// of up to 512 bytes.  Op k's draws use stream 2 (nseq0 + k) + 1, as by plane.
__device__ __forceinline__ void frame_noise_task(uint64_t *__restrict__ planes,
                                                 const NoiseDesc *__restrict__ desc,
                                                 uint32_t nk, uint64_t w, uint32_t k0,
                                                 uint64_t seed, uint64_t nseq0,
                                                 uint32_t kend) {
  const uint32_t lane = lane_id();
  const uint64_t lane_g = lane_golden();
  const uint32_t cnt = min(32u, kend - k0);  // ops k0 .. min(k0 + 32, kend) - 1
  // lane i < cnt derives op k0 + i's stream key
  const uint64_t key_l = stream_key(seed, 2 * (nseq0 + k0 + min(lane, cnt - 1)) + 1);
  uint32_t out_lo = 0, out_hi = 0;
  for (uint32_t i = 0; i < cnt; ++i) {
    const NoiseDesc d = desc[k0 + i];
    const uint32_t klo = __builtin_amdgcn_readlane((uint32_t)key_l, i);
    const uint32_t khi = __builtin_amdgcn_readlane((uint32_t)(key_l >> 32), i);
    const uint64_t key = ((uint64_t)khi << 32) | klo;
    bool bx, bz;
    noise_bits(d, draw_word_lane(key, w, lane_g), bx, bz);
    const uint64_t mx = __ballot(bx);
    const uint64_t mz = __ballot(bz);
    out_lo = write_lane(out_lo, (uint32_t)mx, 2 * i);
    out_hi = write_lane(out_hi, (uint32_t)(mx >> 32), 2 * i);
    out_lo = write_lane(out_lo, (uint32_t)mz, 2 * i + 1);
    out_hi = write_lane(out_hi, (uint32_t)(mz >> 32), 2 * i + 1);
  }
  if (lane < 2 * cnt)
    planes[(w * nk + k0) * 2 + lane] = ((uint64_t)out_hi << 32) | out_lo;
}

// Wave task t: word group t / nkb, ops from 32 (t % nkb), nkb = ceil(nk / 32).
// Word groups [w0, w0 + wn) and noise ops [ka, kb) of the planes.
__global__ __launch_bounds__(256) void k_frame_noise(uint64_t *__restrict__ planes,
                                                     const NoiseDesc *__restrict__ desc,
                                                     uint32_t nk, uint64_t w0, uint64_t wn,
                                                     uint32_t ka, uint32_t kb,
                                                     const ProgParams *__restrict__ pp) {
  const uint64_t nwv = (uint64_t)gridDim.x * (blockDim.x / 64);
  const uint64_t wave =
      blockIdx.x * (uint64_t)(blockDim.x / 64) + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint64_t nkb = (kb - ka + 31) / 32;
  for (uint64_t t = wave; t < wn * nkb; t += nwv)
    frame_noise_task(planes, desc, nk, w0 + t / nkb, ka + (uint32_t)(t % nkb) * 32, pp->seed,
                     pp->nseq0, kb);
}

// Frame batch: every circuit's tasks in one numbering (circuit c's from
// circ[c].task0, circ[ncirc] standing for the end), each wave a contiguous
// share; the wave finds its first circuit by binary search, then walks on.
__global__ __launch_bounds__(256) void k_batch_frame_noise(const pq::BatchCircuit *__restrict__ circ,
                                                           uint32_t ncirc, uint64_t tasks) {
  const uint64_t nwv = (uint64_t)gridDim.x * (blockDim.x / 64);
  const uint64_t wave =
      blockIdx.x * (uint64_t)(blockDim.x / 64) + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint64_t per = (tasks + nwv - 1) / nwv;
  uint64_t t = wave * per;
  const uint64_t t1 = min(t + per, tasks);
  if (t >= t1) return;
  uint32_t lo = 0, hi = ncirc;  // the last circuit whose first task is <= t
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) / 2;
    if (circ[mid].task0 <= t)
      lo = mid;
    else
      hi = mid;
  }
  uint32_t c = lo;
  for (; t < t1; ++t) {
    while (c + 1 < ncirc && circ[c + 1].task0 <= t) ++c;
    const pq::BatchCircuit &bc = circ[c];
    const uint64_t nkb = (bc.nnoise + 31) / 32;
    const uint64_t lt = t - bc.task0;
    frame_noise_task(const_cast<uint64_t *>(bc.a.planes), bc.noise, bc.nnoise, lt / nkb,
                     (uint32_t)(lt % nkb) * 32, bc.seed, bc.nseq0, bc.nnoise);
  }
}

__global__ void k_zero_counter(unsigned long long *counter) {
  if (threadIdx.x == 0) *counter = 0;
}


// RowMultiply / RowProductSign (MeasurementFunctors.hpp:98-166) on every
// shot: one thread per word group; sign_out (if set) receives the sign plane
// without modifying the tableau.
__global__ void k_rowsum(uint64_t *__restrict__ tab, uint32_t n, uint32_t R,
                         uint64_t W, uint32_t i, uint32_t k, uint64_t *sign_out) {
  const uint64_t C = 2ull * n + 1;
  for (uint64_t w = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; w < W;
       w += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t *T = tab + w * C * R;
    uint64_t c0 = 0, c1 = 0;
    for (uint32_t j = 0; j < n; ++j) {
      uint64_t *px = T + (uint64_t)j * R, *pz = T + (uint64_t)(n + j) * R;
      phase_acc(px[i], pz[i], px[k], pz[k], c0, c1);
    }
    uint64_t *pr = T + (uint64_t)(2 * n) * R;
    const uint64_t sign = pr[i] ^ pr[k] ^ c1;
    if (sign_out) {
      sign_out[w] = sign;
    } else {
      pr[i] = sign;
      for (uint32_t j = 0; j < n; ++j) {
        uint64_t *px = T + (uint64_t)j * R, *pz = T + (uint64_t)(n + j) * R;
        px[i] ^= px[k];
        pz[i] ^= pz[k];
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Frame form (X/Z shared by all shots): one workgroup per 64-shot word group
// runs the compiled op stream on the group's R column and Pauli frame, held
// in LDS.  The stream, its layout and the LDS map are described in
// pq_frame.cpp; the shared X/Z themselves are written by k_frame_xz.
// ---------------------------------------------------------------------------
typedef __attribute__((address_space(3))) void lds_void;

__device__ __forceinline__ uint32_t rfl(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

// The kFrameThreads threads of a block share one word group: its R column
// (the per-shot part Rv), Pauli frame, record words, the noise-mask window
// and the stream chunk live in LDS, the threads split each op's rows or
// tasks (within an op no two threads write one word) and a barrier separates
// ops.
constexpr uint32_t kFrameWaves = kFrameThreads / 64;

__device__ __forceinline__ void frame_word_group(const FrameArgs &a, uint64_t w) {
  extern __shared__ uint64_t dyn[];
  const uint32_t tid = threadIdx.x, lane = lane_id();
  const uint32_t wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  const uint32_t n = a.n, R = a.R;
  uint64_t *Rl = dyn;
  uint64_t *nbuf0 = dyn + a.lds_noise;
  uint32_t *sbuf0 = reinterpret_cast<uint32_t *>(dyn + a.lds_noise + a.nbufs * a.nb_words);
  uint64_t *part = reinterpret_cast<uint64_t *>(sbuf0 + a.nbufs * a.sb_words);  // [waves][64]
  uint64_t *Tr = a.tab + w * a.col_stride + a.r_off;  // the group's R column
  // chunk c's stream words and noise masks (ops k0 .. k0+nk-1 of this group,
  // contiguous) into staging buffer c % nbufs by LDS DMA, one wave-wide 4-byte
  // load per 64 dwords, the waves taking turns; lanes past the end reload
  // the last dword into the buffer's padding
  auto stage = [&](uint32_t c) {
    const uint32_t cb = a.chunks[4 * c], cw = a.chunks[4 * c + 1];
    const uint32_t k0 = a.chunks[4 * c + 2], nk = a.chunks[4 * c + 3];
    const uint32_t buf = c % a.nbufs;
    uint32_t *sb = sbuf0 + buf * a.sb_words;
    uint32_t *nb = reinterpret_cast<uint32_t *>(nbuf0 + buf * a.nb_words);
    for (uint32_t i = wv * 64; i < cw; i += kFrameThreads) {
      const uint32_t j = min(i + lane, cw - 1);
      __builtin_amdgcn_global_load_lds(a.stream + cb + j, (lds_void *)(sb + i), 4, 0, 0);
    }
    const uint32_t dw = 4 * nk;
    const uint32_t *nsrc = reinterpret_cast<const uint32_t *>(a.planes + (w * a.nnoise + k0) * 2);
    for (uint32_t i = wv * 64; i < dw; i += kFrameThreads) {
      const uint32_t t = min(i + lane, dw - 1);
      __builtin_amdgcn_global_load_lds(nsrc + t, (lds_void *)(nb + i), 4, 0, 0);
    }
  };
  const uint32_t nside = a.lds_const - R;  // the frame and the record words
  for (uint32_t i = tid; i < R; i += kFrameThreads) Rl[i] = Tr[i];
  if (a.resume) {
    for (uint32_t i = tid; i < nside; i += kFrameThreads) Rl[R + i] = a.side[w * nside + i];
  } else {
    for (uint32_t i = tid; i < 2 * n; i += kFrameThreads) Rl[R + i] = 0;  // the frame E
    for (uint32_t i = tid; i < a.npreload; i += kFrameThreads)
      dyn[a.lds_rec + a.preload[2 * i + 1]] = a.recs[(2ull * a.preload[2 * i]) * a.W + w];
  }
  if (tid == 0) {
    dyn[a.lds_const] = ~0ull;  // the all-ones word
    dyn[a.lds_const + 1] = 0;  // the zero word
  }
  const uint64_t seed = a.seed, mseq0 = a.mseq0;
  if (a.c0 < a.c1) stage(a.c0);
  for (uint32_t c = a.c0; c < a.c1; ++c) {
    const uint32_t cw = a.chunks[4 * c + 1];
    // chunk c has landed, and every thread is done with the buffer that
    // chunk c+1 goes to (chunk c-1's, or chunk c's own with one buffer)
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    if (a.nbufs > 1 && c + 1 < a.c1) stage(c + 1);
    const uint32_t *sbuf = sbuf0 + (c % a.nbufs) * a.sb_words;
    uint32_t pos = 0;
    while (pos < cw) {
      const uint32_t kind = rfl(sbuf[pos]);
      const uint32_t h1 = rfl(sbuf[pos + 1]);
      if (kind == kFrameLayer) {
        // independent tasks, one per thread, each running its updates in
        // order: 16-bit LDS word indices {dst, s0, s1, s2}, dyn[dst] ^=
        // dyn[s0] ^ dyn[s1] ^ dyn[s2] (unused sources name the zero word);
        // the next update is read before this one writes
        const uint32_t T = h1, nbody = rfl(sbuf[pos + 2]);
        const uint32_t *offs = sbuf + pos + 8;
        const uint32_t bstart = (pos + 8 + T + 2) & ~1u;
        const uint2 *body = reinterpret_cast<const uint2 *>(sbuf + bstart);
        for (uint32_t t = tid; t < T; t += kFrameThreads) {
          const uint32_t e1 = offs[t + 1];
          uint32_t e = offs[t];
          if (e >= e1) continue;
          uint2 q = body[e];
          for (; e < e1; ++e) {
            const uint2 qn = body[min(e + 1, e1 - 1)];
            const uint64_t v = dyn[q.x >> 16] ^ dyn[q.y & 0xFFFFu] ^ dyn[q.y >> 16];
            dyn[q.x & 0xFFFFu] ^= v;
            q = qn;
          }
        }
        pos = bstart + nbody;
        __syncthreads();
      } else if (kind == kFrameXor) {
        // Rv[row] ^= XOR of the LDS words its entries name
        const uint32_t T = h1, ne = rfl(sbuf[pos + 2]);
        const uint32_t *rows = sbuf + pos + 8;
        const uint32_t *offs = rows + T;
        const uint32_t *ents = offs + T + 1;
        for (uint32_t t = tid; t < T; t += kFrameThreads) {
          const uint32_t row = rows[t], e1 = offs[t + 1];
          uint64_t acc = 0;
          uint32_t e = offs[t];
          for (; e + 4 <= e1; e += 4) {
            const uint32_t i0 = ents[e], i1 = ents[e + 1], i2 = ents[e + 2], i3 = ents[e + 3];
            acc ^= (dyn[i0] ^ dyn[i1]) ^ (dyn[i2] ^ dyn[i3]);
          }
          for (; e < e1; ++e) acc ^= dyn[ents[e]];
          Rl[row] ^= acc;
        }
        pos += 8 + 2 * T + 1 + ne;
        __syncthreads();
      } else if (kind == kFrameRandom) {
        // MeasureRandom (MeasurementFunctors.hpp:225-265) with pivot p: the
        // rowsum rows first (all threads read Rv[n+p] before it changes),
        // then Rv[p], Rv[n+p] = coin ^ E.x[q] and the record
        const uint32_t p = h1, slot = rfl(sbuf[pos + 2]), mi = rfl(sbuf[pos + 3]);
        const uint32_t Tc = rfl(sbuf[pos + 4]), ridx = rfl(sbuf[pos + 5]);
        const uint32_t ne = rfl(sbuf[pos + 6]), exq = rfl(sbuf[pos + 7]);
        const uint64_t Rs = Rl[n + p];
        const uint32_t *ents = sbuf + pos + 8;
        for (uint32_t e = tid; e < ne; e += kFrameThreads) Rl[ents[e]] ^= Rs;
        __syncthreads();
        if (wv == 0) {
          const uint64_t key = stream_key(seed, 2 * (mseq0 + mi));
          const uint32_t u = (uint32_t)(draw_keyed(key, 64 * w + lane) >> 40);
          const uint64_t coin = __ballot(u < Tc);
          if (lane == 0) {
            Rl[p] = Rs;
            Rl[n + p] = coin ^ dyn[exq];
            a.recs[(2ull * slot) * a.W + w] = coin;
            a.recs[(2ull * slot + 1) * a.W + w] = 0ull;
            if (ridx != kNone) dyn[a.lds_rec + ridx] = coin;
          }
        }
        pos += 8 + ne;
        __syncthreads();
      } else {
        // G determined measurements (MeasureDetermined, :189-201): lane j of
        // every wave XORs every kFrameWaves-th row of measurement j, wave 0
        // combines and adds the constant and E.x[q]; their rows exclude the
        // scratch row, which takes the Rv part of the one flagged last
        const uint32_t G = h1;
        const uint32_t *desc = sbuf + pos + 8;
        const uint32_t *rows = desc + 4 * G + 1;
        if (lane < G) {
          const uint32_t o1 = lane + 1 < G ? desc[4 * lane + 7] : desc[4 * G];
          uint64_t acc = 0;
          for (uint32_t e = desc[4 * lane + 3] + wv; e < o1; e += kFrameWaves)
            acc ^= Rl[rows[e]];
          part[wv * 64 + lane] = acc;
        }
        __syncthreads();
        if (wv == 0 && lane < G) {
          uint64_t acc = 0;
#pragma unroll
          for (uint32_t v = 0; v < kFrameWaves; ++v) acc ^= part[v * 64 + lane];
          const uint32_t slot = desc[4 * lane], ridx = desc[4 * lane + 1];
          const uint32_t xq = desc[4 * lane + 2];
          const uint64_t res = acc ^ ((xq >> 31) ? ~0ull : 0ull) ^ dyn[xq & 0xFFFFu];
          a.recs[(2ull * slot) * a.W + w] = res;
          a.recs[(2ull * slot + 1) * a.W + w] = ~0ull;
          if (ridx != kNone) dyn[a.lds_rec + ridx] = res;
          if ((xq >> 30) & 1) Rl[2 * n] = acc;
        }
        pos += 8 + 4 * G + 1 + rfl(desc[4 * G]);
        __syncthreads();
      }
    }
    if (a.nbufs == 1 && c + 1 < a.c1) stage(c + 1);  // after the ops' last barrier
  }
  __syncthreads();
  for (uint32_t i = tid; i < R; i += kFrameThreads) Tr[i] = Rl[i];
  if (a.suspend)
    for (uint32_t i = tid; i < nside; i += kFrameThreads) a.side[w * nside + i] = Rl[R + i];
}

__global__ __launch_bounds__(kFrameThreads) void k_frame(FrameArgs a) {
  frame_word_group(a, a.w0 + blockIdx.x);
}

// Frame form of a mixed batch: block b runs work item items[b] = (circuit
// << 40 | word group) with that circuit's arguments.
__global__ __launch_bounds__(kFrameThreads) void k_frame_batch(const FrameArgs *__restrict__ circ,
                                                               const uint64_t *__restrict__ items) {
  const uint64_t it = items[blockIdx.x];
  frame_word_group(circ[it >> 40], it & ((1ull << 40) - 1));
}

// Shared X/Z (bit rows [R][nw]) into columns 0..2n-1 of every word group.
__global__ __launch_bounds__(256) void k_frame_xz(uint64_t *__restrict__ tab,
                                                  const uint64_t *__restrict__ X,
                                                  const uint64_t *__restrict__ Z, uint32_t n,
                                                  uint32_t R, uint32_t C, uint64_t W,
                                                  uint32_t nw) {
  const uint32_t row = blockIdx.x * 256 + threadIdx.x;
  const uint32_t c = blockIdx.y;
  if (row >= R) return;
  const uint32_t q = c < n ? c : c - n;
  const uint64_t *M = c < n ? X : Z;
  const uint64_t v = ((M[(uint64_t)row * nw + q / 64] >> (q % 64)) & 1) ? ~0ull : 0ull;
  // every word group of this (column, row range): 2 KB per block and group
  for (uint64_t w = blockIdx.z; w < W; w += gridDim.z) tab[(w * C + c) * R + row] = v;
}

inline uint32_t grid_for(uint64_t items, uint32_t threads, uint32_t cap = 1u << 20) {
  uint64_t g = (items + threads - 1) / threads;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (uint32_t)g;
}

}  // namespace

// ---------------------------------------------------------------------------
// Persistent program kernel (defined in pq_program.hip, shares measure_word).
// ---------------------------------------------------------------------------

hipError_t launch_fill_r_identity(hipStream_t s, uint64_t *rcol, uint64_t W, uint32_t R) {
  if (!W || !R) return hipSuccess;
  hipLaunchKernelGGL(k_fill_r_identity, dim3((uint32_t)std::min<uint64_t>(W, 1u << 20)), dim3(256),
                     0, s, rcol, R, W);
  return hipGetLastError();
}

hipError_t launch_fill_r_column(hipStream_t s, uint64_t *tab, const Geom &g) {
  if (!g.W || !g.R) return hipSuccess;
  hipLaunchKernelGGL(k_fill_r_column, dim3((uint32_t)std::min<uint64_t>(g.W, 1u << 20)), dim3(256),
                     0, s, tab, g.R, g.C, g.W);
  return hipGetLastError();
}

hipError_t launch_fill_identity(hipStream_t s, uint64_t *tab, const Geom &g) {
  const uint64_t ncols = g.W * (uint64_t)g.C;
  if (!ncols || !g.R) return hipSuccess;
  // at most 16384 blocks, each filling columns in turn (many small blocks
  // would be bound by their dispatch, not by HBM)
  hipLaunchKernelGGL(k_fill_identity, dim3((uint32_t)std::min<uint64_t>(ncols, 16384)),
                     dim3(256), 0, s, tab, g.R, g.C, ncols);
  return hipGetLastError();
}

hipError_t launch_gate(hipStream_t s, uint64_t *tab, const Geom &g, uint32_t code,
                       uint32_t qa, uint32_t qb, const uint64_t *fa,
                       const uint64_t *fb) {
  const uint64_t total = g.W * g.R;
  if (!total) return hipSuccess;
  GateArgs a{code, g.n, g.R, g.C, qa, qb, g.W, fa, fb};
  hipLaunchKernelGGL(k_gate, dim3(grid_for(total, 256)), dim3(256), 0, s, tab, a);
  return hipGetLastError();
}

hipError_t launch_noise_planes(hipStream_t s, uint64_t *planes, NoiseDesc one,
                               const NoiseDesc *desc, uint32_t nk, uint64_t W) {
  const uint64_t items = (uint64_t)nk * W;
  if (!items) return hipSuccess;
  hipLaunchKernelGGL(k_noise_planes, dim3(grid_for(items * 64, 256, 16384)),
                     dim3(256), 0, s, planes, one, desc, nk, W);
  return hipGetLastError();
}

size_t measure_lds_bytes(uint32_t n) {
  return sizeof(uint64_t) * (2ull * n + 1) + sizeof(uint32_t) * (uint64_t)(n + 1);
}

uint32_t measure_max_qubits() { return kMeasMaxQubits; }

constexpr size_t kProgramLdsBudget = 158 * 1024;  // dynamic LDS cap of k_program (160 KB less static LDS)

static hipError_t ensure_lds_attr(const void *fn, size_t bytes) {
  if (bytes <= 65536) return hipSuccess;
  return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)bytes);
}

hipError_t launch_measure(hipStream_t s, uint64_t *tab, const Geom &g,
                          const MeasParams &m) {
  if (!g.W) return hipSuccess;
  const size_t lds = measure_lds_bytes(g.n);
  const uint32_t blocks = (uint32_t)std::min<uint64_t>(g.W, 1u << 20);
  hipLaunchKernelGGL(k_measure, dim3(blocks), dim3(kMeasThreads), lds, s, tab,
                     g.n, g.R, g.W, m.q, m.coin_T, m.coin_key, m.coin_plane,
                     m.rec, m.force_determined ? 1u : 0u);
  return hipGetLastError();
}

hipError_t launch_unpack(hipStream_t s, const uint64_t *tab, const Geom &g,
                         uint64_t b0, uint64_t nb, int32_t *x, int32_t *z,
                         int32_t *r) {
  const uint64_t total = nb * (uint64_t)g.R * g.n;
  if (!total) return hipSuccess;
  hipLaunchKernelGGL(k_unpack, dim3(grid_for(total, 256, 65536)), dim3(256), 0, s,
                     tab, g.n, g.R, g.C, b0, nb, x, z, r);
  return hipGetLastError();
}

hipError_t launch_pack(hipStream_t s, uint64_t *tab, const Geom &g, uint64_t w0,
                       uint64_t nw, const int32_t *x, const int32_t *z,
                       const int32_t *r) {
  const uint64_t total = nw * (uint64_t)g.C * g.R;
  if (!total) return hipSuccess;
  hipLaunchKernelGGL(k_pack, dim3(grid_for(total, 256, 65536)), dim3(256), 0, s,
                     tab, g.n, g.R, g.C, g.B, w0, nw, x, z, r);
  return hipGetLastError();
}

// Words of a and b that differ (n words each), accumulated into *out.
__global__ __launch_bounds__(256) void k_count_diff(const uint64_t *__restrict__ a,
                                                    const uint64_t *__restrict__ b, uint64_t n,
                                                    unsigned long long *out) {
  unsigned long long c = 0;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    c += a[i] != b[i];
#pragma unroll
  for (uint32_t off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off, 64);
  if (lane_id() == 0 && c) atomicAdd(out, c);
}

__global__ __launch_bounds__(256) void k_zero_u32(unsigned int *p, uint32_t n) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    p[i] = 0u;
}

hipError_t launch_zero_u32(hipStream_t s, unsigned int *p, uint32_t n) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(k_zero_u32, dim3(std::min<uint32_t>((n + 255) / 256, 64)), dim3(256), 0,
                     s, p, n);
  return hipGetLastError();
}

hipError_t launch_zero_word(hipStream_t s, unsigned long long *p) {
  hipLaunchKernelGGL(k_zero_counter, dim3(1), dim3(64), 0, s, p);
  return hipGetLastError();
}

hipError_t launch_count_diff(hipStream_t s, const uint64_t *a, const uint64_t *b, uint64_t n,
                             unsigned long long *out_dev) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(k_count_diff, dim3(grid_for(n, 256, 8192)), dim3(256), 0, s, a, b, n,
                     out_dev);
  return hipGetLastError();
}

hipError_t launch_checksum(hipStream_t s, const uint64_t *tab, const Geom &g,
                           uint64_t b0, uint64_t b1,
                           unsigned long long *out_dev) {
  const uint64_t total = g.W * (uint64_t)g.C * g.R;
  if (!total) return hipSuccess;
  hipLaunchKernelGGL(k_checksum, dim3(grid_for(total, 256, 8192)), dim3(256), 0, s,
                     tab, g.n, g.R, g.C, g.W, b0, b1, out_dev);
  return hipGetLastError();
}

hipError_t launch_flags_to_plane(hipStream_t s, const int32_t *flags,
                                 uint64_t B, uint64_t *plane) {
  const uint64_t W = (B + 63) / 64;
  if (!W) return hipSuccess;
  hipLaunchKernelGGL(k_flags_to_plane, dim3(grid_for(W * 64, 256, 4096)), dim3(256),
                     0, s, flags, B, plane);
  return hipGetLastError();
}

hipError_t launch_record_to_int(hipStream_t s, const uint64_t *rec, uint64_t B,
                                uint64_t W, int32_t *res, int32_t *det) {
  if (!B) return hipSuccess;
  hipLaunchKernelGGL(k_record_to_int, dim3(grid_for(B, 256, 4096)), dim3(256), 0,
                     s, rec, B, W, res, det);
  return hipGetLastError();
}

// The large build (8 scratch-row columns per lane, 512 threads) serves
// states of 193..512 qubits, whose determined chains it keeps in registers;
// the small one (768 threads) serves states up to 192 qubits (3 columns per
// lane).  Above 512 the chains keep their scratch row in LDS in either build:
// programs with row-tiled gate segments (`wide`) take the wider block for
// them, the others (bound by random measurements' row updates, which a wider
// block slows down) the 512-thread one.
static bool large_build(uint32_t n, bool wide) {
  return n > kSmallQubits && (n <= 64u * kLargeRegs || !wide);
}

// The persistent kernel build for n qubits, one workgroup per word group or
// sliced word groups.
static const void *program_kernel(uint32_t n, bool sliced, bool wide) {
  if (!large_build(n, wide))
    return sliced ? (const void *)k_program_t<kSmallRegs, kSmallThreads, true>
                  : (const void *)k_program_t<kSmallRegs, kSmallThreads, false>;
  return sliced ? (const void *)k_program_t<kLargeRegs, kLargeThreads, true>
                : (const void *)k_program_t<kLargeRegs, kLargeThreads, false>;
}

hipError_t launch_program(hipStream_t s, const ProgramArgs &a, uint32_t threads,
                          uint32_t blocks, const void *params) {
  if (!a.nwords) return hipSuccess;
  const size_t lds = a.lds_words * 8;
  if (a.slices > 1) {
    // one workgroup per (word group, slice), all resident at once: the
    // slices wait for one another
    const void *fn = program_kernel(a.n, true, a.wide);
    const uint32_t tb = std::min<uint32_t>(threads, program_max_threads(a.n, a.wide));
    ProgramArgs la = a;
    const ProgParams *pp = (const ProgParams *)params;
    void *args[] = {&la, &pp};
    return hipLaunchCooperativeKernel(fn, dim3((uint32_t)(a.nwords * a.slices)), dim3(tb), args,
                                      (unsigned int)lds, s);
  }
  if (!large_build(a.n, a.wide))
    hipLaunchKernelGGL((k_program_t<kSmallRegs, kSmallThreads, false>), dim3(blocks),
                       dim3(std::min<uint32_t>(threads, kSmallThreads)), lds, s, a,
                       (const ProgParams *)params);
  else
    hipLaunchKernelGGL((k_program_t<kLargeRegs, kLargeThreads, false>), dim3(blocks),
                       dim3(std::min<uint32_t>(threads, kLargeThreads)), lds, s, a,
                       (const ProgParams *)params);
  return hipGetLastError();
}

bool program_slices_fit(uint32_t n, uint64_t nwords, uint32_t slices, uint32_t threads,
                        uint64_t lds_bytes, bool wide) {
  const void *fn = program_kernel(n, true, wide);
  int dev = 0, cus = 0, per_cu = 0, coop = 0;
  if (hipGetDevice(&dev) != hipSuccess) return false;
  if (hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev) != hipSuccess ||
      !coop)
    return false;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return false;
  const uint32_t tb = std::min<uint32_t>(threads, program_max_threads(n, wide));
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, (int)tb, (size_t)lds_bytes) !=
      hipSuccess)
    return false;
  return (uint64_t)per_cu * (uint64_t)cus >= nwords * slices;
}

uint32_t program_tile_rows(uint64_t lds_words, uint32_t nq) { return tile_rows(lds_words, nq); }

uint32_t program_max_threads(uint32_t n, bool wide) {
  return large_build(n, wide) ? kLargeThreads : kSmallThreads;
}

size_t program_lds_bytes(uint32_t n, uint32_t det_waves) {
  // the gate-segment staging area and the measurement area share the LDS
  return 8 * std::max<size_t>(program_lds_words(n, det_waves), kStageWords);
}

hipError_t launch_program_noise(hipStream_t s, uint64_t *planes,
                                const NoiseDesc *desc, uint32_t nk, uint64_t W,
                                const void *params) {
  // plane k uses stream 2 (nseq0 + k) + 1, as the frame form's planes
  return launch_frame_noise(s, planes, desc, nk, 0, W, params, 0, nk);
}

hipError_t launch_frame(hipStream_t s, const FrameArgs &a) {
  if (!a.wn) return hipSuccess;
  hipError_t e = ensure_lds_attr((const void *)k_frame, a.lds_bytes);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_frame, dim3((uint32_t)a.wn), dim3(kFrameThreads), a.lds_bytes, s, a);
  return hipGetLastError();
}

hipError_t launch_frame_batch(hipStream_t s, const FrameArgs *circ, const uint64_t *items,
                              uint64_t nitems, uint32_t lds_bytes) {
  if (!nitems) return hipSuccess;
  hipError_t e = ensure_lds_attr((const void *)k_frame_batch, lds_bytes);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_frame_batch, dim3((uint32_t)nitems), dim3(kFrameThreads), lds_bytes, s,
                     circ, items);
  return hipGetLastError();
}

hipError_t launch_frame_noise(hipStream_t s, uint64_t *planes, const NoiseDesc *desc,
                              uint32_t nk, uint64_t w0, uint64_t wn, const void *params,
                              uint32_t ka, uint32_t kb) {
  kb = std::min(kb, nk);
  if (ka >= kb || !wn) return hipSuccess;
  const uint64_t tasks = wn * ((kb - ka + 31) / 32);
  const uint32_t blocks = (uint32_t)std::min<uint64_t>((tasks + 3) / 4, 1u << 16);
  hipLaunchKernelGGL(k_frame_noise, dim3(blocks), dim3(256), 0, s, planes, desc, nk, w0, wn, ka,
                     kb, (const ProgParams *)params);
  return hipGetLastError();
}

hipError_t launch_batch_frame_noise(hipStream_t s, const BatchCircuit *circ, uint32_t ncirc,
                                    uint64_t tasks) {
  if (!ncirc || !tasks) return hipSuccess;
  // about 8 tasks per wave, at most 65536 blocks
  const uint32_t blocks = (uint32_t)std::min<uint64_t>((tasks + 31) / 32, 65536);
  hipLaunchKernelGGL(k_batch_frame_noise, dim3(blocks), dim3(256), 0, s, circ, ncirc, tasks);
  return hipGetLastError();
}


hipError_t launch_frame_xz(hipStream_t s, uint64_t *tab, const Geom &g, const uint64_t *X,
                           const uint64_t *Z, uint32_t nw) {
  if (!g.W || !g.n) return hipSuccess;
  // about 65536 blocks, each writing its column slice in several word
  // groups: one 2 KB block per (column slice, group) would be bound by
  // dispatch, not by HBM, on a large state
  const uint64_t blocks = (uint64_t)((g.R + 255) / 256) * 2 * g.n;
  const uint64_t z = std::max<uint64_t>(1, std::min<uint64_t>(g.W, (65536 + blocks - 1) / blocks));
  const dim3 grid((g.R + 255) / 256, 2 * g.n, (uint32_t)std::min<uint64_t>(z, 64));
  hipLaunchKernelGGL(k_frame_xz, grid, dim3(256), 0, s, tab, X, Z, g.n, g.R, g.C, g.W, nw);
  return hipGetLastError();
}

hipError_t launch_rowsum(hipStream_t s, uint64_t *tab, const Geom &g, uint32_t i,
                         uint32_t k, uint64_t *sign_out) {
  if (!g.W) return hipSuccess;
  hipLaunchKernelGGL(k_rowsum, dim3(grid_for(g.W, 64, 4096)), dim3(64), 0, s, tab,
                     g.n, g.R, g.W, i, k, sign_out);
  return hipGetLastError();
}

hipError_t launch_set_params(hipStream_t s, void *params, uint64_t seed,
                             uint64_t mseq0, uint64_t nseq0) {
  hipLaunchKernelGGL(k_set_params, dim3(1), dim3(64), 0, s, (uint64_t *)params,
                     seed, mseq0, nseq0);
  return hipGetLastError();
}

hipError_t prepare_kernels(uint32_t n) {
  hipError_t e = ensure_lds_attr((const void *)k_measure, measure_lds_bytes(n));
  for (uint32_t nn : {1u, kSmallQubits + 1})
    for (bool sliced : {false, true})
      if (e == hipSuccess) e = ensure_lds_attr(program_kernel(nn, sliced, false), kProgramLdsBudget);
  return e;
}

hipError_t launch_batch(hipStream_t s, const BatchCircuit *circ, uint32_t ncirc,
                        const uint64_t *items, uint64_t nitems,
                        unsigned long long *counter, uint32_t threads,
                        uint32_t blocks, size_t lds_bytes, uint64_t noise_tasks,
                        uint32_t chain_qubits, uint32_t wtile) {
  if (!nitems) return hipSuccess;
  hipLaunchKernelGGL(k_zero_counter, dim3(1), dim3(64), 0, s, counter);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (noise_tasks) {
    e = launch_batch_frame_noise(s, circ, ncirc, noise_tasks);
    if (e != hipSuccess) return e;
  }
  // the large build only where a member's chains use it and the block fits it
  if (!large_build(chain_qubits, false) || threads > kLargeThreads)
    hipLaunchKernelGGL((k_batch_t<kSmallRegs, kSmallThreads>), dim3(blocks),
                       dim3(std::min<uint32_t>(threads, kSmallThreads)), lds_bytes, s, circ,
                       items, nitems, counter, std::max<uint32_t>(wtile, 1));
  else
    hipLaunchKernelGGL((k_batch_t<kLargeRegs, kLargeThreads>), dim3(blocks),
                       dim3(std::min<uint32_t>(threads, kLargeThreads)), lds_bytes, s, circ,
                       items, nitems, counter, std::max<uint32_t>(wtile, 1));
  return hipGetLastError();
}

hipError_t prepare_batch_kernel() {
  hipError_t e =
      ensure_lds_attr((const void *)k_batch_t<kSmallRegs, kSmallThreads>, kProgramLdsBudget);
  if (e == hipSuccess)
    e = ensure_lds_attr((const void *)k_batch_t<kLargeRegs, kLargeThreads>, kProgramLdsBudget);
  return e;
}

}  // namespace pq
