// tbg_kernels.hip -- CDNA4 (gfx950) kernels of the king-slice retrograde
// engine.  Wave64 throughout; one work-group = 4 waves.
//
// Frontier kernels (pred, uncap, verify) give each wave a chunk of 64
// bitmap words (4096 positions, one word per lane).  The wave scans its
// words, compacts the set bits into an LDS queue by a wave prefix sum, and
// then works through the queue 64 entries at a time, so every lane holds
// one frontier position regardless of how the bits are spread.  Dense
// kernels (init, resolve, exits) give each lane one position and assemble
// result words with a wave ballot.  Sweeps handle 2 words per lane with
// 16-byte loads.
#include "tbg_dev.h"
#include "tbg_kern.h"

namespace tbg {

__device__ Tabs g_tabs;

static_assert(sizeof(Tabs) % 4 == 0, "Tabs copy granularity");

__device__ __forceinline__ void load_tabs(Tabs &S)
{
  const uint32_t *src = reinterpret_cast<const uint32_t *>(&g_tabs);
  uint32_t *dst = reinterpret_cast<uint32_t *>(&S);
  for (int i = threadIdx.x; i < (int)(sizeof(Tabs) / 4); i += blockDim.x) dst[i] = src[i];
  __syncthreads();
}

__device__ __forceinline__ void atomic_or64(uint64_t *a, uint64_t v)
{
  __hip_atomic_fetch_or(reinterpret_cast<unsigned long long *>(a), (unsigned long long)v,
                        __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ bool bm_test(const uint64_t *p, uint64_t i)
{
  return (p[i >> 6] >> (i & 63)) & 1;
}

__device__ __forceinline__ int wave_excl_scan(int v, int lane, int *total)
{
  int x = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    int t = __shfl_up(x, d, 64);
    if (lane >= d) x += t;
  }
  *total = __shfl(x, 63, 64);
  return x - v;
}

__device__ __forceinline__ uint64_t wave_sum64(uint64_t v)
{
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  return v;
}

__device__ __forceinline__ uint64_t kings_of(const Tabs &T, int s)
{
  return (uint64_t)T.kksq[s][0] | ((uint64_t)T.kksq[s][1] << 8);
}

__device__ __forceinline__ uint64_t phys(const Mat &M, int s) { return M.synth ? 0 : (uint64_t)s; }

// ------------------------------------------------------------- unmoves
// mark_king_unmoves (src/tb8gen.c:246-262)
// Targets already in wins(stm) (illegal or won faster) can never become new
// W_n or L_n positions (W_n = C & ~wins; a won position fails verification),
// so they are tested with a cached read before the memory-side atomic.
__device__ void mark_king_unmoves(const Tabs &T, const Mat &M, int stm, uint64_t occ, uint64_t p,
                                  int s, uint64_t *cand, const uint64_t *__restrict__ won)
{
  uint64_t b = T.king[sq_get(p, stm)] & ~(occ | T.king[sq_get(p, stm ^ 1)]);
  while (b) {
    int t = ctz64(b);
    b &= b - 1;
    uint64_t q;
    int s2 = normalize(T, sq_set(p, stm, t), &q);
    if (M.synth && s2 != M.synth_slice) continue;  // leaves a single-slice job
    uint64_t *base = cand + phys(M, s2) * M.words;
    const uint64_t *wb = won + phys(M, s2) * M.words;
    uint64_t i2 = sq_to_idx(T, M.ii, q);
    if (!bm_test(wb, i2)) atomic_or64(base + (i2 >> 6), bit(i2 & 63));
    if (s < 441 && s2 >= 441) {
      i2 = sq_to_idx(T, M.ii, mirror_nonkings(q));
      if (!bm_test(wb, i2)) atomic_or64(base + (i2 >> 6), bit(i2 & 63));
    }
  }
}

// mark_unmoves (src/tb8gen.c:264-276) of the non-king piece in slot j; the
// unmoved position is q with slot j replaced.
__device__ __forceinline__ void mark_unmoves(const Tabs &T, const Mat &M, int j, uint64_t occ,
                                             uint64_t q, uint64_t *base,
                                             const uint64_t *__restrict__ wb)
{
  uint64_t b = piece_att(T, M.pt[j], sq_get(q, j), occ) & ~occ;
  while (b) {
    int t = ctz64(b);
    b &= b - 1;
    uint64_t i2 = sq_to_idx(T, M.ii, sq_set(q, j, t));
    if (!bm_test(wb, i2)) atomic_or64(base + (i2 >> 6), bit(i2 & 63));
  }
}

// ----------------------------------------------------------- frontier
struct Queue {
  uint16_t e[4][4096];
  uint64_t res[4][64];
};

// Fills the wave's queue from its chunk; returns the entry count.
__device__ __forceinline__ int fill_queue(uint16_t *q, uint64_t x, int lane)
{
  int total;
  int off = wave_excl_scan(popc64(x), lane, &total);
  while (x) {
    int b = ctz64(x);
    x &= x - 1;
    q[off++] = (uint16_t)((lane << 6) | b);
  }
  return total;
}

// predecessors_worker (src/tb8gen.c:484-516): every set bit of the stm^1
// frontier `src` -> stm unmoves OR-ed into `cand`.
__global__ __launch_bounds__(256) void k_pred(const Mat *__restrict__ Mp, int stm,
                                               const uint64_t *__restrict__ src, uint64_t *cand,
                                               const uint64_t *__restrict__ won, int s0, int ns,
                                               uint64_t cps)
{
  __shared__ Tabs T;
  __shared__ Queue Q;
  load_tabs(T);
  const Mat &M = *Mp;
  int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  uint64_t nchunks = (uint64_t)ns * cps;
  for (uint64_t g0 = (uint64_t)blockIdx.x * 4; g0 < nchunks; g0 += (uint64_t)gridDim.x * 4) {
    uint64_t g = g0 + wave;
    int s = s0 + (int)(g / cps);
    uint64_t w = (g % cps) * 64 + lane;
    uint64_t x = 0;
    if (g < nchunks && w < M.words) x = src[phys(M, s) * M.words + w];
    int total = fill_queue(Q.e[wave], x, lane);
    __syncthreads();
    uint64_t *own = cand + phys(M, s) * M.words;
    const uint64_t *ownw = won + phys(M, s) * M.words;
    uint64_t wbase = (g % cps) * 64;
    for (int e = lane; e < total; e += 64) {
      uint16_t v = Q.e[wave][e];
      uint64_t idx = (wbase + (v >> 6)) * 64 + (v & 63);
      if (idx >= M.ii.size) continue;
      uint64_t p = kings_of(T, s);
      uint64_t occ = idx_to_sq(T, M.ii, idx, &p);
      mark_king_unmoves(T, M, stm, occ, p, s, cand, won);
      for (int t = 0; t < MAXP - 1; t++) {
        if (t >= M.nk[stm]) break;
        mark_unmoves(T, M, M.pcs[stm][t], occ, p, own, ownw);
      }
    }
    __syncthreads();
  }
}

// Block-privatised form of k_pred.  One work-group owns one block of the
// index (positions [lo, hi) of slice s, a run of whole low-order digits, so
// every unmove of a piece in the low sets stays inside it): those targets
// are OR-ed into an LDS bitmap with LDS atomics, and only king unmoves and
// unmoves of high-set pieces go to the global bitmap.  The LDS bitmap is
// merged with one global atomic per nonzero word.  Device-scope atomics on
// gfx950 travel to the memory side, so this cuts the scatter's HBM write
// traffic by the ratio of same-block to all unmoves.
constexpr int PB_WAVES = 8;       // waves per work-group
constexpr int PB_CAP = 1024;      // queue entries per wave per round

__global__ __launch_bounds__(512) void k_pred_blk(const Mat *__restrict__ Mp, int stm,
                                                   const uint64_t *__restrict__ src, uint64_t *cand,
                                                   const uint64_t *__restrict__ won, int s0, int ns,
                                                   uint64_t blk_pos, uint64_t nblk,
                                                   uint64_t blk_words)
{
  __shared__ Tabs T;
  __shared__ int rmax[PB_WAVES];
  extern __shared__ uint64_t dyn[];
  load_tabs(T);
  const Mat &M = *Mp;
  uint64_t *B = dyn;
  uint16_t *Qall = reinterpret_cast<uint16_t *>(dyn + blk_words);
  int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  uint16_t *q = Qall + wave * PB_CAP;
  uint64_t nb = (uint64_t)ns * nblk;
  for (uint64_t b = blockIdx.x; b < nb; b += gridDim.x) {
    int s = s0 + (int)(b / nblk);
    uint64_t lo = (b % nblk) * blk_pos;
    uint64_t hi = lo + blk_pos < M.ii.size ? lo + blk_pos : M.ii.size;
    uint64_t wlo = lo >> 6, nbw = ((hi + 63) >> 6) - wlo;
    for (uint64_t w = threadIdx.x; w < nbw; w += blockDim.x) B[w] = 0;
    __syncthreads();
    const uint64_t *srcs = src + phys(M, s) * M.words;
    uint64_t *own = cand + phys(M, s) * M.words;
    const uint64_t *ownw = won + phys(M, s) * M.words;
    for (uint64_t c0 = 0; c0 < nbw; c0 += (uint64_t)PB_WAVES * 64) {
      uint64_t wi = c0 + (uint64_t)wave * 64 + lane;
      uint64_t x = wi < nbw ? srcs[wlo + wi] : 0;
      if (wi == 0) x &= ~0ull << (lo & 63);                      // not below the block
      if (wi + 1 == nbw && (hi & 63)) x &= (1ull << (hi & 63)) - 1;  // nor above it
      int total;
      int off = wave_excl_scan(popc64(x), lane, &total);
      // every wave runs the same number of rounds (they share barriers)
      if (lane == 0) rmax[wave] = (total + PB_CAP - 1) / PB_CAP;
      __syncthreads();
      int allr = 0;
      for (int v = 0; v < PB_WAVES; v++) allr = rmax[v] > allr ? rmax[v] : allr;
      __syncthreads();
      for (int r = 0; r < allr; r++) {
        int base = r * PB_CAP;
        uint64_t y = x;
        int k = off;
        while (y && k < base + PB_CAP) {
          int bt = ctz64(y);
          y &= y - 1;
          if (k >= base) q[k - base] = (uint16_t)((lane << 6) | bt);
          k++;
        }
        __syncthreads();
        int cnt = total - base;
        cnt = cnt < 0 ? 0 : cnt > PB_CAP ? PB_CAP : cnt;
        for (int e = lane; e < cnt; e += 64) {
          uint16_t v = q[e];
          uint64_t idx = (wlo + c0 + (uint64_t)wave * 64 + (v >> 6)) * 64 + (v & 63);
          uint64_t p = kings_of(T, s);
          uint64_t occ = idx_to_sq(T, M.ii, idx, &p);
          mark_king_unmoves(T, M, stm, occ, p, s, cand, won);
          for (int t = 0; t < MAXP - 1; t++) {
            if (t >= M.nk[stm]) break;
            int j = M.pcs[stm][t];
            uint64_t bb = piece_att(T, M.pt[j], sq_get(p, j), occ) & ~occ;
            while (bb) {
              int tq = ctz64(bb);
              bb &= bb - 1;
              uint64_t i2 = sq_to_idx(T, M.ii, sq_set(p, j, tq));
              if (i2 >= lo && i2 < hi)
                __hip_atomic_fetch_or(reinterpret_cast<unsigned long long *>(&B[(i2 >> 6) - wlo]),
                                      (unsigned long long)bit(i2 & 63), __ATOMIC_RELAXED,
                                      __HIP_MEMORY_SCOPE_WORKGROUP);
              else if (!bm_test(ownw, i2))
                atomic_or64(own + (i2 >> 6), bit(i2 & 63));
            }
          }
        }
        __syncthreads();
      }
    }
    for (uint64_t w = threadIdx.x; w < nbw; w += blockDim.x) {
      uint64_t v = B[w];
      if (v) atomic_or64(own + wlo + w, v);
    }
    __syncthreads();
  }
}

// predecessors_sub_worker (src/tb8gen.c:346-394): uncapture from the
// stm^1 capture-exit bitmap of set k (src0 & ~src1 when src1 is given).
__global__ __launch_bounds__(256) void k_uncap(const Mat *__restrict__ Mp, int stm, int k,
                                                const uint64_t *__restrict__ src0,
                                                const uint64_t *__restrict__ src1, int legality,
                                                uint64_t *cand, const uint64_t *__restrict__ won,
                                                int s0, int ns, uint64_t cps)
{
  __shared__ Tabs T;
  __shared__ Queue Q;
  load_tabs(T);
  const Mat &M = *Mp;
  const Idx &ci = M.capt[k];
  int o = stm ^ 1, m = M.ii.last[k];
  int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  uint64_t nchunks = (uint64_t)ns * cps;
  for (uint64_t g0 = (uint64_t)blockIdx.x * 4; g0 < nchunks; g0 += (uint64_t)gridDim.x * 4) {
    uint64_t g = g0 + wave;
    int s = s0 + (int)(g / cps);
    uint64_t w = (g % cps) * 64 + lane;
    uint64_t x = 0;
    if (g < nchunks && w < M.cwords[k]) {
      uint64_t a = phys(M, s) * M.ctotal[o] + M.coff[k] + w;
      x = src0[a];
      if (src1) x &= ~src1[a];
    }
    int total = fill_queue(Q.e[wave], x, lane);
    __syncthreads();
    uint64_t *own = cand + phys(M, s) * M.words;
    const uint64_t *ownw = won + phys(M, s) * M.words;
    uint64_t wbase = (g % cps) * 64;
    for (int e = lane; e < total; e += 64) {
      uint16_t v = Q.e[wave][e];
      uint64_t idx = (wbase + (v >> 6)) * 64 + (v & 63);
      if (idx >= ci.size) continue;
      uint64_t p = kings_of(T, s);
      uint64_t occ = idx_to_sq(T, ci, idx, &p);
      if (legality && attacked_by(T, M, p, occ, sq_get(p, stm), o, m)) continue;
      // uncapture by the king
      mark_king_unmoves(T, M, stm, occ, sq_set(p, m, sq_get(p, stm)), s, cand, won);
      // uncapture by the non-king pieces
      for (int t = 0; t < MAXP - 1; t++) {
        if (t >= M.nk[stm]) break;
        int j = M.pcs[stm][t];
        mark_unmoves(T, M, j, occ, sq_set(p, m, sq_get(p, j)), own, ownw);
      }
    }
    __syncthreads();
  }
}

// Synthetic king-move exit value for o after stm's king moved from `from`
// (position q0 not normalised, q normalised into slice s2); captured slot
// jj of set l, or jj < 0.  3 = illegal move.
__device__ int synth_exit(const Tabs &T, const Mat &M, int stm, uint64_t q0, uint64_t q, uint64_t occ,
                          int from, int jj, int l, int s2)
{
  int o = stm ^ 1, to = sq_get(q0, stm);
  if (jj >= 0) {
    if (attacked_by(T, M, q0, occ ^ bit(from), to, o, jj)) return 3;
    uint64_t c = sq_set(q, jj, sq_get(q, M.ii.last[l]));
    return synth_value(synth_hash(M.seed, o, 2, s2 * 8 + l, sq_to_idx(T, M.capt[l], c)));
  }
  if (attacked_by(T, M, q0, occ ^ bit(from) ^ bit(to), to, o, -1)) return 3;
  return synth_value(synth_hash(M.seed, o, 1, s2, sq_to_idx(T, M.ii, q)));
}

// check_successors_worker + check_moves + check_king_moves
// (src/tb8gen.c:529-650): is candidate idx of slice s a loss for stm?
__device__ bool verify_pos(const Tabs &T, const Mat &M, int stm, int s, uint64_t idx,
                           const uint64_t *__restrict__ wins_o, const uint64_t *__restrict__ sub_o)
{
  int o = stm ^ 1;
  uint64_t p = kings_of(T, s);
  uint64_t occ = idx_to_sq(T, M.ii, idx, &p);
  int ks = sq_get(p, stm), oks = sq_get(p, o);
  uint64_t att[MAXS];
  uint64_t all = 0;
#pragma unroll
  for (int t = 0; t < MAXS; t++) {
    att[t] = 0;
    if (t < M.nk[stm]) {
      int j = M.pcs[stm][t];
      att[t] = piece_att(T, M.pt[j], sq_get(p, j), occ);
    }
    all |= att[t];
  }
  if (all & bit(oks)) return false;  // illegal: stm attacks the enemy king
  uint64_t so[MAXS];
  set_occ(M.ii, p, so);
  const uint64_t *wins_s = wins_o + phys(M, s) * M.words;
  const uint64_t *sub_s = sub_o + phys(M, s) * M.ctotal[o];
#pragma unroll
  for (int t = 0; t < MAXS; t++) {
    if (t >= M.nk[stm]) break;
    int j = M.pcs[stm][t];
    uint64_t a = att[t];
    for (int l = 0; l < MAXS; l++) {
      if (l >= M.numsets) break;
      if (M.setcol[l] != o) continue;
      uint64_t c = a & so[l];
      while (c) {
        int to = ctz64(c);
        c &= c - 1;
        int jj = slot_of(p, to, M.setmask80[l]);
        uint64_t q = sq_set(p, j, to);
        q = sq_set(q, jj, sq_get(q, M.ii.last[l]));
        if (!bm_test(sub_s + M.coff[l], sq_to_idx(T, M.capt[l], q))) return false;
      }
    }
    uint64_t qm = a & ~occ;
    while (qm) {
      int to = ctz64(qm);
      qm &= qm - 1;
      if (!bm_test(wins_s, sq_to_idx(T, M.ii, sq_set(p, j, to)))) return false;
    }
  }
  uint64_t kb = T.king[ks] & ~T.king[oks];
  for (int l = 0; l < MAXS; l++) {
    if (l >= M.numsets) break;
    if (M.setcol[l] != o) continue;
    uint64_t c = kb & so[l];
    while (c) {
      int to = ctz64(c);
      c &= c - 1;
      int jj = slot_of(p, to, M.setmask80[l]);
      uint64_t q0 = sq_set(p, stm, to), q;
      int s2 = normalize(T, q0, &q);
      if (M.synth && s2 != M.synth_slice) {
        int v = synth_exit(T, M, stm, q0, q, occ, ks, jj, l, s2);
        if (v == 3 || v == 2) continue;
        return false;
      }
      q = sq_set(q, jj, sq_get(q, M.ii.last[l]));
      const uint64_t *sb = sub_o + phys(M, s2) * M.ctotal[o] + M.coff[l];
      if (!bm_test(sb, sq_to_idx(T, M.capt[l], q))) return false;
    }
  }
  uint64_t kq = kb & ~occ;
  while (kq) {
    int to = ctz64(kq);
    kq &= kq - 1;
    uint64_t q0 = sq_set(p, stm, to), q;
    int s2 = normalize(T, q0, &q);
    if (M.synth && s2 != M.synth_slice) {
      int v = synth_exit(T, M, stm, q0, q, occ, ks, -1, 0, s2);
      if (v == 3 || v == 2) continue;
      return false;
    }
    if (!bm_test(wins_o + phys(M, s2) * M.words, sq_to_idx(T, M.ii, q))) return false;
  }
  return true;
}

// check_successors (src/tb8gen.c:607-661): prune the candidates of stm in
// `cand` to L_n, written to newL (cand is cleared for the next half-step),
// with the count and, if codes != null, the L_n codes.
__global__ __launch_bounds__(256) void k_verify(const Mat *__restrict__ Mp, int stm, int n,
                                                 uint64_t *cand, const uint64_t *__restrict__ wins_o,
                                                 const uint64_t *__restrict__ sub_o,
                                                 const uint64_t *__restrict__ won, uint64_t *newL,
                                                 uint16_t *codes, int s0, int ns, uint64_t cps,
                                                 unsigned long long *cnt)
{
  __shared__ Tabs T;
  __shared__ Queue Q;
  load_tabs(T);
  const Mat &M = *Mp;
  int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  uint64_t nchunks = (uint64_t)ns * cps;
  uint64_t mycnt = 0;
  for (uint64_t g0 = (uint64_t)blockIdx.x * 4; g0 < nchunks; g0 += (uint64_t)gridDim.x * 4) {
    uint64_t g = g0 + wave;
    int s = s0 + (int)(g / cps);
    uint64_t w = (g % cps) * 64 + lane;
    bool live = g < nchunks && w < M.words;
    uint64_t a = phys(M, s) * M.words + w;
    uint64_t x0 = live ? cand[a] : 0;
    // a position already in wins(stm) can not be lost (and illegal ones are
    // not positions): only the rest is verified
    uint64_t x = x0 ? x0 & ~won[a] : 0;
    Q.res[wave][lane] = x;
    int total = fill_queue(Q.e[wave], x, lane);
    __syncthreads();
    uint64_t wbase = (g % cps) * 64;
    for (int e = lane; e < total; e += 64) {
      uint16_t v = Q.e[wave][e];
      uint64_t idx = (wbase + (v >> 6)) * 64 + (v & 63);
      bool ok = idx < M.ii.size && verify_pos(T, M, stm, s, idx, wins_o, sub_o);
      if (!ok)
        __hip_atomic_fetch_and(reinterpret_cast<unsigned long long *>(&Q.res[wave][v >> 6]),
                               ~(1ull << (v & 63)), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    __syncthreads();
    if (live) {
      uint64_t r = Q.res[wave][lane];
      newL[a] = r;
      if (x0) cand[a] = 0;
      mycnt += popc64(r);
      if (codes && r) {
        uint16_t *c = codes + phys(M, s) * M.ii.size + w * 64;
        uint16_t code = (uint16_t)((n << 2) | 2);
        while (r) {
          int b = ctz64(r);
          r &= r - 1;
          c[b] = code;
        }
      }
    }
    __syncthreads();
  }
  mycnt = wave_sum64(mycnt);
  if (lane == 0 && mycnt) atomicAdd(cnt, (unsigned long long)mycnt);
}

// calc_W_n tail (src/tb8gen.c:1072-1077) fused: W = C & ~wins; wins |= W;
// count; codes; C cleared.  One plane region of nw words (slices s0..).
__global__ __launch_bounds__(256) void k_wsweep(uint64_t *cand, uint64_t *wins, uint64_t *newW,
                                                 uint16_t *codes, uint64_t nw, uint64_t words,
                                                 uint64_t size, uint64_t w0, int n,
                                                 unsigned long long *cnt)
{
  uint64_t mycnt = 0;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x * 2;
  for (uint64_t i = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) * 2; i < nw; i += stride) {
    uint64_t a = w0 + i;
    ulonglong2 c = *reinterpret_cast<const ulonglong2 *>(cand + a);
    ulonglong2 wi = *reinterpret_cast<const ulonglong2 *>(wins + a);
    ulonglong2 wn;
    wn.x = c.x & ~wi.x;
    wn.y = c.y & ~wi.y;
    *reinterpret_cast<ulonglong2 *>(newW + a) = wn;
    if (wn.x | wn.y) {
      ulonglong2 o2;
      o2.x = wi.x | wn.x;
      o2.y = wi.y | wn.y;
      *reinterpret_cast<ulonglong2 *>(wins + a) = o2;
    }
    if (c.x | c.y) *reinterpret_cast<ulonglong2 *>(cand + a) = make_ulonglong2(0, 0);
    mycnt += popc64(wn.x) + popc64(wn.y);
    if (codes && (wn.x | wn.y)) {
      uint16_t code = (uint16_t)((n << 2) | 1);
      for (int h = 0; h < 2; h++) {
        uint64_t r = h ? wn.y : wn.x;
        uint64_t wa = a + h;
        uint16_t *cp = codes + (wa / words) * size + (wa % words) * 64;
        while (r) {
          int b = ctz64(r);
          r &= r - 1;
          cp[b] = code;
        }
      }
    }
  }
  int lane = threadIdx.x & 63;
  mycnt = wave_sum64(mycnt);
  if (lane == 0 && mycnt) atomicAdd(cnt, (unsigned long long)mycnt);
}

// Synthetic job, ply 1: does stm's king have a legal move or capture into
// an exit (a position of another slice, valued by the seeded hash) that is
// a loss (*w: W_1 candidate) or a win (*l: L_1 candidate) for the opponent?
// The opponent's attack sets are computed once with stm's king lifted, so
// each king move's legality is a bit test.
__device__ void synth_exits_pos(const Tabs &T, const Mat &M, uint64_t p, uint64_t occ, int stm,
                                bool *w, bool *l)
{
  int o = stm ^ 1, s = M.synth_slice, from = sq_get(p, stm), oks = sq_get(p, o);
  uint64_t occk = occ ^ bit(from);
  uint64_t att[MAXS];
  uint64_t own = bit(from);
#pragma unroll
  for (int t = 0; t < MAXS; t++) {
    att[t] = 0;
    if (t < M.nk[o]) {
      int j = M.pcs[o][t];
      att[t] = piece_att(T, M.pt[j], sq_get(p, j), occk);
    }
    if (t < M.nk[stm]) own |= bit(sq_get(p, M.pcs[stm][t]));
  }
  uint64_t so[MAXS];
  set_occ(M.ii, p, so);
  uint64_t b = T.king[from] & ~T.king[oks] & ~own;
  bool ww = false, ll = false;
  while (b) {
    int to = ctz64(b);
    b &= b - 1;
    int jj = -1, l = 0;
    if (occ & bit(to)) {
      for (int x = 0; x < MAXS; x++)
        if (x < M.numsets && (so[x] & bit(to))) l = x;
      jj = slot_of(p, to, M.setmask80[l]);
    }
    uint64_t q0 = sq_set(p, stm, to), q;
    int s2 = normalize(T, q0, &q);
    if (s2 == s) continue;
    uint64_t a = T.king[oks];
#pragma unroll
    for (int t = 0; t < MAXS; t++)
      if (t < M.nk[o] && M.pcs[o][t] != jj) a |= att[t];
    if (a & bit(to)) continue;  // the king would stand in check
    int v;
    if (jj >= 0) {
      uint64_t c = sq_set(q, jj, sq_get(q, M.ii.last[l]));
      v = synth_value(synth_hash(M.seed, o, 2, s2 * 8 + l, sq_to_idx(T, M.capt[l], c)));
    } else {
      v = synth_value(synth_hash(M.seed, o, 1, s2, sq_to_idx(T, M.ii, q)));
    }
    ww |= v == -2;
    ll |= v == 2;
  }
  *w = ww;
  *l = ll;
}

// has_legal_moves || has_legal_caps (src/movegen.c:75-111) for colour c.
__device__ bool has_legal(const Tabs &T, const Mat &M, uint64_t p, uint64_t occ, int c)
{
  int o = c ^ 1;
  uint64_t own = bit(sq_get(p, c));
  for (int t = 0; t < MAXP - 1; t++) {
    if (t >= M.nk[c]) break;
    own |= bit(sq_get(p, M.pcs[c][t]));
  }
  uint64_t enemy80 = 0;
  for (int t = 0; t < MAXP - 1; t++) {
    if (t >= M.nk[o]) break;
    enemy80 |= 0x80ull << (8 * M.pcs[o][t]);
  }
  for (int t = -1; t < MAXP - 1; t++) {
    if (t >= M.nk[c]) break;
    int j = t < 0 ? c : M.pcs[c][t];
    int from = sq_get(p, j);
    uint64_t b = piece_att(T, M.pt[j], from, occ) & ~own;
    while (b) {
      int to = ctz64(b);
      b &= b - 1;
      int cap = (occ & bit(to)) ? slot_of(p, to, enemy80) : -1;
      if ((occ & bit(to)) && cap < 0) continue;  // the enemy king: never legal to reach
      uint64_t q = sq_set(p, j, to);
      uint64_t occ2 = (occ ^ bit(from)) | bit(to);
      int ksq = j == c ? to : sq_get(p, c);
      if (!attacked_by(T, M, q, occ2, ksq, o, cap)) return true;
    }
  }
  return false;
}

// calc_illegal_worker + calc_mate_worker (src/tb8gen.c:664-727), dense: one
// position per lane.  Writes illegal(c) into wins[c], mates into mate[c],
// and every code.
__global__ __launch_bounds__(256) void k_init(const Mat *__restrict__ Mp, uint64_t *winsW,
                                               uint64_t *winsB, uint64_t *mateW, uint64_t *mateB,
                                               uint16_t *codesW, uint16_t *codesB,
                                               uint64_t *kexW0, uint64_t *kexW1, uint64_t *kexL0,
                                               uint64_t *kexL1, int s0, int ns,
                                               unsigned long long *cnt)
{
  __shared__ Tabs T;
  load_tabs(T);
  const Mat &M = *Mp;
  int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  uint64_t nw = (uint64_t)ns * M.words;
  uint64_t c0 = 0, c1 = 0, c2 = 0, c3 = 0;
  for (uint64_t g = (uint64_t)blockIdx.x * 4 + wave; g < nw; g += (uint64_t)gridDim.x * 4) {
    int s = s0 + (int)(g / M.words);
    uint64_t w = g % M.words;
    uint64_t idx = w * 64 + lane;
    bool wi = false, bi = false, wm = false, bm = false;
    bool xw0 = false, xw1 = false, xl0 = false, xl1 = false;
    if (idx < M.ii.size) {
      uint64_t p = kings_of(T, s);
      uint64_t occ = idx_to_sq(T, M.ii, idx, &p);
      if (M.synth) {
        synth_exits_pos(T, M, p, occ, WHITE, &xw0, &xl0);
        synth_exits_pos(T, M, p, occ, BLACK, &xw1, &xl1);
      }
      wi = attacked_by(T, M, p, occ, sq_get(p, BLACK), WHITE, -1);
      bi = attacked_by(T, M, p, occ, sq_get(p, WHITE), BLACK, -1);
      if (wi != bi) {
        int c = bi ? WHITE : BLACK;
        bool mate = !has_legal(T, M, p, occ, c);
        wm = mate && c == WHITE;
        bm = mate && c == BLACK;
      }
      uint64_t ca = phys(M, s) * M.ii.size + idx;
      if (codesW) codesW[ca] = wi ? 3 : wm ? 2 : 0;
      if (codesB) codesB[ca] = bi ? 3 : bm ? 2 : 0;
    }
    uint64_t a = phys(M, s) * M.words + w;
    uint64_t bwi = __ballot(wi), bbi = __ballot(bi), bwm = __ballot(wm), bbm = __ballot(bm);
    if (M.synth) {
      uint64_t e0 = __ballot(xw0), e1 = __ballot(xw1), e2 = __ballot(xl0), e3 = __ballot(xl1);
      if (lane == 0) {
        kexW0[a] = e0;
        kexW1[a] = e1;
        kexL0[a] = e2;
        kexL1[a] = e3;
      }
    }
    if (lane == 0) {
      winsW[a] = bwi;
      winsB[a] = bbi;
      mateW[a] = bwm;
      mateB[a] = bbm;
      c0 += popc64(bwi);
      c1 += popc64(bbi);
      c2 += popc64(bwm);
      c3 += popc64(bbm);
    }
  }
  if (lane == 0) {
    if (c0) atomicAdd(cnt + 0, (unsigned long long)c0);
    if (c1) atomicAdd(cnt + 1, (unsigned long long)c1);
    if (c2) atomicAdd(cnt + 2, (unsigned long long)c2);
    if (c3) atomicAdd(cnt + 3, (unsigned long long)c3);
  }
}

// calc_sub_worker (src/tb8gen.c:280-314) for set k of colour c: classify
// every position with one piece of set k removed, c to move, by the
// smaller material's finished table (probe_wdl, src/probe.c:1677-1684).
__global__ __launch_bounds__(256) void k_resolve(const Mat *__restrict__ Mp,
                                                  const SubRef *__restrict__ Rp, int c, int k,
                                                  uint64_t *loss, uint64_t *bloss, uint64_t *cwin,
                                                  uint64_t *win, int s0, int ns)
{
  __shared__ Tabs T;
  load_tabs(T);
  const Mat &M = *Mp;
  const SubRef &R = *Rp;
  const Idx &ci = M.capt[k];
  int m = M.ii.last[k];
  int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  uint64_t cw = M.cwords[k];
  uint64_t nw = (uint64_t)ns * cw;
  for (uint64_t g = (uint64_t)blockIdx.x * 4 + wave; g < nw; g += (uint64_t)gridDim.x * 4) {
    int s = s0 + (int)(g / cw);
    uint64_t w = g % cw;
    uint64_t idx = w * 64 + lane;
    int v = 0;
    bool ill = false;
    if (idx < ci.size) {
      uint64_t p = kings_of(T, s);
      uint64_t occ = idx_to_sq(T, ci, idx, &p);
      ill = attacked_by(T, M, p, occ, sq_get(p, c ^ 1), c, m);
      if (!ill) {
        if (M.synth) {
          v = synth_value(synth_hash(M.seed, c, 0, k, idx));
        } else if (R.present) {
          uint64_t tsq = 0;
          for (int j = 0; j < MAXP; j++) {
            if (j >= R.num) break;
            tsq = sq_set(tsq, j, sq_get(p, sq_get(R.map, j)) ^ (R.flip ? 0x38 : 0));
          }
          uint64_t q;
          int s2 = normalize(T, tsq, &q);
          uint64_t i2 = sq_to_idx(T, R.ii, q);
          v = code_wdl(R.codes[c ^ R.flip][(uint64_t)s2 * R.ii.size + i2]);
        }
      }
    }
    uint64_t b0 = __ballot(!ill && v == -2), b1 = __ballot(!ill && v == -1);
    uint64_t b2 = __ballot(ill || v >= 1), b3 = __ballot(ill || v == 2);
    if (lane == 0) {
      uint64_t a = phys(M, s) * M.ctotal[c] + M.coff[k] + w;
      loss[a] = b0;
      bloss[a] = b1;
      cwin[a] = b2;
      win[a] = b3;
    }
  }
}

// dst |= src over nw words (2 per lane)
__global__ __launch_bounds__(256) void k_or(uint64_t *dst, const uint64_t *__restrict__ src, uint64_t nw)
{
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x * 2;
  for (uint64_t i = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) * 2; i < nw; i += stride) {
    ulonglong2 a = *reinterpret_cast<const ulonglong2 *>(src + i);
    if (a.x | a.y) {
      ulonglong2 d = *reinterpret_cast<const ulonglong2 *>(dst + i);
      d.x |= a.x;
      d.y |= a.y;
      *reinterpret_cast<ulonglong2 *>(dst + i) = d;
    }
  }
}

}  // namespace tbg

namespace tbg {
hipError_t upload_tabs(const Tabs *h) { return hipMemcpyToSymbol(HIP_SYMBOL(g_tabs), h, sizeof(Tabs)); }
}  // namespace tbg
