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

// Bitmaps and codes live in device global memory.  Their pointers come
// from job records, so the compiler sees generic (flat) pointers; the hot
// accesses cast them to the global address space (global loads and
// atomics, whose waits do not also cover LDS traffic).
typedef __attribute__((address_space(1))) uint64_t gu64;
typedef __attribute__((address_space(1))) const uint64_t cgu64;
typedef __attribute__((address_space(1))) unsigned long long gull;

__device__ __forceinline__ uint64_t gload(const uint64_t *p, uint64_t i)
{
  return ((cgu64 *)p)[i];
}

// a value that is the same in every lane, moved to a scalar register
__device__ __forceinline__ int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ uint64_t uni64(uint64_t x)
{
  uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)x);
  uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(x >> 32));
  return (uint64_t)lo | ((uint64_t)hi << 32);
}
// later_sets of a piece with its factors in scalar registers (the same for
// every lane): the loads of the factors complete before the caller's loop
// over targets, not inside it
__device__ __forceinline__ Later later_sets_u(const Idx &ii, uint64_t p, int kj)
{
  Later L = later_sets(ii, p, kj);
  L.fj = (uint32_t)uni((int)L.fj);
  L.n = uni(L.n);
#pragma unroll
  for (int k = 0; k < MAXS - 1; k++) L.f[k] = (uint32_t)uni((int)L.f[k]);
  return L;
}
// moved_idx_l as a 32-bit offset from the position's own index, for a
// piece whose set's factor is below 2^23 (a set inside a pred block, whose
// factor is below the block size of at most 327,552 positions; any set of a
// slice of up to 2^22 positions): the rank change is below 62, so the
// product fits a 24-bit multiply and the offset 32 bits.
__device__ __forceinline__ int moved_delta(const Later &L, int a, int b, uint64_t occb,
                                           uint32_t rank_a)
{
  int rank_b = b - popc64(occb & (bit(b) - 1));
  int lo = a < b ? a : b, hi = a < b ? b : a;
  int acc = 0;
#pragma unroll
  for (int k = 0; k < MAXS - 1; k++) {
    if (k >= L.n) break;
    acc += (L.s[k] > lo && L.s[k] < hi) ? (int)L.f[k] : 0;
  }
  int d = __mul24(rank_b - (int)rank_a, (int)L.fj);
  return a < b ? d + acc : d - acc;
}
template <class X> __device__ __forceinline__ X *uniptr(X *p)
{
  return reinterpret_cast<X *>(uni64(reinterpret_cast<uint64_t>(p)));
}

__device__ __forceinline__ void atomic_or64(uint64_t *a, uint64_t v)
{
  __hip_atomic_fetch_or((gull *)a, (unsigned long long)v, __ATOMIC_RELAXED,
                        __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ bool bm_test(const uint64_t *p, uint64_t i)
{
  return (gload(p, i >> 6) >> (i & 63)) & 1;
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
// Up to 4 scatter targets (global bit indices into a plane; ~0: none):
// their words of `won` are loaded together, then each target not already
// won gets its 64-bit atomic OR.  One memory round trip per group instead
// of one per target.
// sm (nullable): the plane's chunk summaries (a one-slice plane), set for
// every chunk marked
__device__ __forceinline__ void mark4(uint64_t *plane, const uint64_t *__restrict__ won,
                                      const uint64_t (&g)[4], uint8_t *sm = nullptr)
{
  uint64_t w[4];
#pragma unroll
  for (int k = 0; k < 4; k++) w[k] = g[k] != ~0ull ? gload(won, g[k] >> 6) : ~0ull;
#pragma unroll
  for (int k = 0; k < 4; k++)
    if (!((w[k] >> (g[k] & 63)) & 1)) {
      atomic_or64(plane + (g[k] >> 6), bit(g[k] & 63));
      if (sm) sm[g[k] >> 12] = 1;
    }
}

// mark_king_unmoves (src/tb8gen.c:246-262)
// Targets already in wins(stm) (illegal or won faster) can never become new
// W_n or L_n positions (W_n = C & ~wins; a won position fails verification),
// so they are tested with a cached read before the memory-side atomic.
// The targets of a position go out in groups of 4 (mark4).
// idx: the position's index in slice s, or ~0 when unknown (an uncapture's
// constructed position); with it, a king move that needs no symmetry is
// indexed incrementally (king_moved_idx).
__device__ __forceinline__ void mark_king_unmoves(const Tabs &T, const Mat &M, int stm, uint64_t occ, uint64_t p,
                                  int s, uint64_t *cand, const uint64_t *__restrict__ won,
                                  uint64_t idx, uint8_t *sm = nullptr)
{
  int from = sq_get(p, stm);
  bool inc = M.all_single && idx != ~0ull;
  uint64_t b = T.king[from] & ~(occ | T.king[sq_get(p, stm ^ 1)]);
  bool dual = false;  // entering a diagonal slice from another: also the mirror
  uint64_t dq = 0;
  int ds = 0;
  while (b || dual) {
    uint64_t g[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      g[k] = ~0ull;
      if (dual) {
        uint64_t qd = mirror_nonkings(dq);
        g[k] = phys(M, ds) * M.words * 64 +
               (M.all_single ? sq_to_idx_1(M.ii, qd) : sq_to_idx(T, M.ii, qd));
        dual = false;
      } else if (b) {
        int t = ctz64(b);
        b &= b - 1;
        uint64_t q0 = sq_set(p, stm, t), q;
        int s2 = normalize(T, q0, &q);
        if (M.synth && s2 != M.synth_slice) continue;  // leaves a single-slice job
        g[k] = phys(M, s2) * M.words * 64 +
               (inc && q == q0  ? king_moved_idx(M.ii, idx, q0, from, t)
                : M.all_single ? sq_to_idx_1(M.ii, q)
                               : sq_to_idx(T, M.ii, q));
        if (s < 441 && s2 >= 441) {
          dual = true;
          dq = q;
          ds = s2;
        }
      }
    }
    mark4(cand, won, g, sm);
  }
}

// A 32-bit value moved between lanes by DPP (gfx9 controls: 0x138 wave_shr:1,
// 0x110 + n row_shr:n within rows of 16 lanes, 0x142 row_bcast:15, 0x143
// row_bcast:31); lanes with no source lane, or in rows outside ROWS, read
// `old`.  No LDS round trip, unlike a shuffle.
template <int CTRL, int ROWS>
__device__ __forceinline__ uint32_t dpp_move(uint32_t old, uint32_t v)
{
  return (uint32_t)__builtin_amdgcn_update_dpp((int)old, (int)v, CTRL, ROWS, 0xf, false);
}

// Exclusive prefix sum over the wave by DPP lane moves (shifts 1, 2, 4, 8
// inside each row of 16 lanes, then the row totals broadcast to the rows
// above); *total: the whole wave's sum.  Call with the whole wave (a lane
// left out would pass on a stale value).
__device__ __forceinline__ int wave_excl_scan(int v, int lane, int *total)
{
  (void)lane;
  uint32_t x = (uint32_t)v;
  x += dpp_move<0x111, 0xf>(0u, x);
  x += dpp_move<0x112, 0xf>(0u, x);
  x += dpp_move<0x114, 0xf>(0u, x);
  x += dpp_move<0x118, 0xf>(0u, x);
  x += dpp_move<0x142, 0xa>(0u, x);
  x += dpp_move<0x143, 0xc>(0u, x);
  *total = __builtin_amdgcn_readlane((int)x, 63);
  return (int)x - v;
}

// one step of the segmented OR-scan: take the moved lane's bits if it has
// the same key
template <int CTRL, int ROWS>
__device__ __forceinline__ void seg_or_step(uint32_t key, uint32_t ku, uint64_t &val)
{
  uint32_t lo = dpp_move<CTRL, ROWS>(0u, (uint32_t)val);
  uint32_t hi = dpp_move<CTRL, ROWS>(0u, (uint32_t)(val >> 32));
  if (ku == key) val |= (uint64_t)hi << 32 | lo;
}
template <int CTRL, int ROWS>
__device__ __forceinline__ void seg_or_step(uint32_t key, uint64_t &val)
{
  seg_or_step<CTRL, ROWS>(key, dpp_move<CTRL, ROWS>(~0u, key), val);
}

// Scatter with combining ("sort-then-OR" within a wave): every lane brings
// one target (bit index into the slice bitmap `plane`, ~0: none); a
// segmented OR-scan over runs of lanes whose targets share a 64-bit word
// (lanes hold consecutive frontier positions under the same unmove, so
// their targets cluster) gathers each run's bits into its last lane, which
// alone issues one atomic OR.  A word that recurs in separate runs just
// gets one atomic per run (lanes of another run with the same word may add
// their bits too: the same word, so harmless).  The scan is the usual
// wave64 DPP scan: lane shift 1, then shifts 2, 4, 8 inside each row of 16
// lanes, then the row totals broadcast to the rows above.  A slice has at
// most 62*61*60*59*58*57 < 2^36 positions (six non-king pieces), so its
// word indices fit 32 bits below the ~0u of "no target".  Targets already
// in wins(stm) are not filtered here: the verification and the W sweep mask
// them, and a read of `won` before each combined atomic costs more than the
// atomics it saves.  Call with the whole wave.
// sm (nullable): the plane's chunk summaries, set for every chunk marked
__device__ __forceinline__ void wave_mark(uint64_t *plane, uint64_t g, int lane,
                                          uint8_t *sm = nullptr)
{
  uint32_t key = g == ~0ull ? ~0u : (uint32_t)(g >> 6);
  uint64_t val = g == ~0ull ? 0 : bit(g & 63);
  uint32_t kp = dpp_move<0x138, 0xf>(~0u, key);  // the key of lane - 1
  uint64_t heads = __ballot(kp != key);         // lanes that start a run
  // every target alone in its run (scattered targets, as a piece's moves
  // across unrelated positions are): no scan
  if (__ballot(key != ~0u && kp == key)) {
    seg_or_step<0x138, 0xf>(key, kp, val);
    seg_or_step<0x112, 0xf>(key, val);
    seg_or_step<0x114, 0xf>(key, val);
    seg_or_step<0x118, 0xf>(key, val);
    seg_or_step<0x142, 0xa>(key, val);  // lane 15 of rows 0, 2 into rows 1, 3
    seg_or_step<0x143, 0xc>(key, val);  // lane 31 into rows 2, 3
  }
  bool last = ((heads >> 1 | 1ull << 63) >> lane) & 1;
  if (key != ~0u && last) {
    atomic_or64(plane + key, val);
    if (sm) sm[key >> 6] = 1;
  }
}

// mark_king_unmoves for the whole wave with combining.  Every lane of a
// pred unit has the same kings, so the loop runs over the king's target
// squares in step across the wave: each target's slice and symmetry are
// computed once (scalar), and for a move that needs no symmetry the lanes'
// targets keep the order of their positions, so the segmented OR folds
// neighbouring lanes into one atomic per word.  ok: the lane holds a
// position.
__device__ void king_unmoves_wave(const Tabs &T, const Mat &M, int stm, bool ok, uint64_t occ,
                                  uint64_t p, int s, uint64_t *cand, int lane,
                                  uint64_t idx, uint8_t *sm = nullptr)
{
  int from = uni(sq_get(p, stm)), oks = uni(sq_get(p, stm ^ 1));
  uint64_t targets = uni64(T.king[from] & ~T.king[oks]);
  // the sets' squares and (scalar) factors, gathered once for every target
  AllSets A = all_sets(M.ii, p);
#pragma unroll
  for (int k = 0; k < MAXS; k++) A.f[k] = (uint32_t)uni((int)A.f[k]);
  while (targets) {
    int t = ctz64(targets);
    targets &= targets - 1;
    // normalize (src/tb8gen.c:230-238) of the moved king pair
    int wk = stm == WHITE ? t : oks, bk = stm == WHITE ? oks : t;
    int m = mirror_mask(wk);
    bool fl = flip_test(wk, bk);
    int w2 = wk ^ m, b2 = bk ^ m;
    if (fl) {
      w2 = ((w2 & 7) << 3) | (w2 >> 3);
      b2 = ((b2 & 7) << 3) | (b2 >> 3);
    }
    int s2 = uni(T.kkidx[T.tri[w2] * 64 + b2]);
    if (M.synth && s2 != M.synth_slice) continue;  // it leaves a single-slice job
    bool dual = s < 441 && s2 >= 441;              // mark_king_unmoves' mirror image
    uint64_t g = ~0ull, gd = ~0ull;
    if (ok && !(occ & bit(t))) {
      uint64_t q0 = sq_set(p, stm, t);
      uint64_t q = q0 ^ (0x0101010101010101ull * (uint64_t)m);
      if (fl) q = flip_diag_bytes(q);
      g = !M.all_single ? sq_to_idx(T, M.ii, q)
          : m == 0 && !fl ? king_moved_idx_a(A, idx, from, t)
                          : sq_to_idx_a(A, q);
      if (dual) {
        uint64_t qd = mirror_nonkings(q);
        gd = M.all_single ? sq_to_idx_a(A, qd) : sq_to_idx(T, M.ii, qd);
      }
    }
    uint64_t *plane = cand + phys(M, s2) * M.words;
    wave_mark(plane, g, lane, sm);
    if (dual) wave_mark(plane, gd, lane, sm);
  }
}

// mark_unmoves (src/tb8gen.c:264-276) of the non-king piece in slot j; the
// unmoved position is q with slot j replaced.
// xb (nullable): a second plane of targets to leave out, tested only for
// those not in wb
__device__ __forceinline__ void mark_unmoves(const Tabs &T, const Mat &M, int j, uint64_t occ,
                                             uint64_t q, uint64_t *base,
                                             const uint64_t *__restrict__ wb,
                                             const uint64_t *__restrict__ xb = nullptr)
{
  uint64_t b = piece_att(T, M.pt[j], sq_get(q, j), occ) & ~occ;
  while (b) {
    int t = ctz64(b);
    b &= b - 1;
    uint64_t q2 = sq_set(q, j, t);
    uint64_t i2 = M.all_single ? sq_to_idx_1(M.ii, q2) : sq_to_idx(T, M.ii, q2);
    bool skip = bm_test(wb, i2);
    if (!skip && xb) skip = bm_test(xb, i2);
    if (!skip) atomic_or64(base + (i2 >> 6), bit(i2 & 63));
  }
}

// ----------------------------------------------------------- frontier
constexpr int QCAP = 1024;  // queue entries per wave per round
constexpr int VERIFY_QUICK = 4;    // bitmap reads of verification's first pass
constexpr int VERIFY_SPLIT = 256;  // candidates per round worth a quick first pass
// a ply is sparse when its frontier times this is below the chunks of a plane
constexpr unsigned long long SPARSE_UNMOVES = 64;
constexpr int LQCAP = 256;         // deferred candidates per wave (two-pass verification)
// an L step runs in the pull form when PULL_RATIO times its frontier is at
// least the positions it would verify (JobDev.nopen)
constexpr unsigned long long PULL_RATIO = 2;
// ... in slices of more positions than VERIFY_SPLIT_SIZE (tbg_kern.h): in
// smaller tables the bitmaps stay in the caches, their reads are cheap, and
// the second decode of the open candidates costs more than the idle lanes
// it saves

struct Queue {
  uint16_t e[4][QCAP];
  uint64_t res[4][64];
  int n[4];  // verification: candidates left open by the first pass
};

// LDS written by some lanes of a wave becomes visible to the other lanes
__device__ __forceinline__ void wave_sync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Writes the wave's set bits of rank [base, base + QCAP) (rank = position
// in the wave's bit order: lane, then bit) to its queue.
__device__ __forceinline__ void fill_round(uint16_t *q, uint64_t x, int off, int base, int lane)
{
  int k = off;
  while (x && k < base + QCAP) {
    int bt = ctz64(x);
    x &= x - 1;
    if (k >= base) q[k - base] = (uint16_t)((lane << 6) | bt);
    k++;
  }
}

__device__ __forceinline__ int find_pair(const uint64_t *__restrict__ P, int np, uint64_t u)
{
  int lo = 0, hi = np;
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (P[mid] <= u) lo = mid; else hi = mid;
  }
  return lo;
}

// Copies the material of job j into LDS when the work-group's unit moves to
// another job (j is uniform across the work-group).  Kernels then read the
// material from LDS: no reloads after their global atomics and stores.
__device__ __forceinline__ void cache_mat(Mat &Ms, const Mat *src, int j, int *cached)
{
  if (j == *cached) return;
  __syncthreads();
  const uint32_t *a = reinterpret_cast<const uint32_t *>(src);
  uint32_t *b = reinterpret_cast<uint32_t *>(&Ms);
  for (int i = threadIdx.x; i < (int)(sizeof(Mat) / 4); i += blockDim.x) b[i] = a[i];
  __syncthreads();
  *cached = j;
}

// Slice of a job-local work unit: a synthetic job has its one slice, a
// material covers slices sw0, sw0+1, ... of the launch window.
__device__ __forceinline__ int unit_slice(const Mat &M, int sw0, uint64_t local, uint64_t ups,
                                          uint64_t *within)
{
  if (M.synth) {  // one slice: no division
    *within = local;
    return M.synth_slice;
  }
  if (local < (1ull << 32)) {  // a 32-bit division is a fraction of a 64-bit one
    uint32_t q = (uint32_t)local / (uint32_t)ups;
    *within = local - (uint64_t)q * ups;
    return sw0 + (int)q;
  }
  *within = local % ups;
  return sw0 + (int)(local / ups);
}

// Word wi of a pred block's frontier: bits below lo and from hi on (the
// neighbouring blocks' positions sharing its first and last words) masked.
__device__ __forceinline__ uint64_t block_mask(uint64_t x, uint64_t wi, uint64_t nbw, uint64_t lo,
                                               uint64_t hi)
{
  if (wi == 0) x &= ~0ull << (lo & 63);
  if (wi + 1 == nbw && (hi & 63)) x &= (1ull << (hi & 63)) - 1;
  return x;
}
__device__ __forceinline__ uint64_t block_word(const uint64_t *srcs, uint64_t wlo, uint64_t wi,
                                               uint64_t nbw, uint64_t lo, uint64_t hi)
{
  return block_mask(gload(srcs, wlo + wi), wi, nbw, lo, hi);
}

// OR of words [w0, w1) of p (w0 < w1) over the work-group (thread t of nt
// reads w0 + t, w0 + t + nt, ...), eight loads per thread in flight at a
// time: the addresses past w1 are clamped to w1 - 1 and their words
// dropped, so no branch separates the loads.  edge(x, w) masks word w.
template <class Edge>
__device__ __forceinline__ uint64_t or_words(const uint64_t *p, uint64_t w0, uint64_t w1, int t,
                                             int nt, Edge edge)
{
  uint64_t acc = 0;
  for (uint64_t b = w0 + (uint64_t)t; b < w1; b += (uint64_t)nt * 8) {
    uint64_t v[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
      uint64_t w = b + (uint64_t)k * nt;
      v[k] = gload(p, w < w1 ? w : w1 - 1);
    }
#pragma unroll
    for (int k = 0; k < 8; k++) {
      uint64_t w = b + (uint64_t)k * nt;
      acc |= w < w1 ? edge(v[k], w) : 0;
    }
  }
  return acc;
}

// predecessors_worker (src/tb8gen.c:484-516) over every (job, side) of a
// launch: each set bit of the stm^1 frontier (W_{n-1} for an L step, kind
// 0; L_{n-1} for a W step, kind 1) -> stm unmoves OR-ed into cand[stm].
//
// One work-group owns one block of the index (positions [lo, hi) of slice
// s, a run of whole low-order digits, so every unmove of a piece in the low
// sets stays inside it): with the block scatter those targets are OR-ed
// into an LDS bitmap with LDS atomics and merged with one global atomic per
// nonzero word; king unmoves and unmoves of high-set pieces (all targets,
// with the global scatter) use 64-bit device atomics.  Device-scope atomics
// on gfx950 travel to the memory side, so privatising cuts the scatter's HBM
// write traffic by the share of same-block unmoves.
//
// Two register budgets: k_pred_j (no spills, 4 waves per SIMD) when every
// unmove but the king's lands in LDS; k_pred_wide_j (80 VGPRs, a few
// spilled, 6 waves per SIMD) when pieces of high sets scatter to device
// memory, whose atomics' latency more resident waves hide.  The _sum forms
// add the chunk summaries of sparse plies (a launch with a job that keeps
// them); the others leave that code out.  They are four plain kernels, not
// instances of a template: the template's instances compiled to slower code
// (8-piece pred 474 -> 490 ms) from the same body.
constexpr int PB_WAVES = 8;   // waves per work-group
constexpr int PB_CAP = PB_QUEUE_BYTES / 4 / PB_WAVES;  // 32-bit queue entries per wave
constexpr int PB_ROOM = PB_CAP - 64;  // a chunk's entries per fill (64 may be waiting)

// Writes the wave's set bits of rank [base, base + PB_ROOM) (rank = position
// in the wave's bit order: lane, then bit) to q[rank - base], each as the
// word's offset in the unit << 6 | the bit.
__device__ __forceinline__ void fill_pred(uint32_t *q, uint64_t x, int off, int base,
                                          uint32_t word)
{
  int k = off;
  while (x && k < base + PB_ROOM) {
    int bt = ctz64(x);
    x &= x - 1;
    if (k >= base) q[k - base] = word << 6 | (uint32_t)bt;
    k++;
  }
}

__global__ __launch_bounds__(512) void k_pred_j(JobDev *const *__restrict__ jobs,
                                                const Pair *__restrict__ pairs,
                                                const uint64_t *__restrict__ P, int np, int sw0,
                                                int kind)
{
  __shared__ Tabs T;
  __shared__ Mat Ms;
  extern __shared__ uint64_t dyn[];
  constexpr bool kDecode1 = true;  // registers to spare at 4 waves/SIMD
  constexpr bool SUM = false;
#include "tbg_pred.inc"
}

__global__ __launch_bounds__(512) void k_pred_sum_j(JobDev *const *__restrict__ jobs,
                                                    const Pair *__restrict__ pairs,
                                                    const uint64_t *__restrict__ P, int np, int sw0,
                                                    int kind)
{
  __shared__ Tabs T;
  __shared__ Mat Ms;
  extern __shared__ uint64_t dyn[];
  constexpr bool kDecode1 = true;  // registers to spare at 4 waves/SIMD
  constexpr bool SUM = true;
#include "tbg_pred.inc"
}

__global__ __launch_bounds__(512, 6) void k_pred_wide_j(JobDev *const *__restrict__ jobs,
                                                        const Pair *__restrict__ pairs,
                                                        const uint64_t *__restrict__ P, int np, int sw0,
                                                        int kind)
{
  __shared__ Tabs T;
  __shared__ Mat Ms;
  extern __shared__ uint64_t dyn[];
  constexpr bool kDecode1 = false;  // stepped decodes instead
  constexpr bool SUM = false;
#include "tbg_pred.inc"
}

__global__ __launch_bounds__(512, 6) void k_pred_wide_sum_j(JobDev *const *__restrict__ jobs,
                                                            const Pair *__restrict__ pairs,
                                                            const uint64_t *__restrict__ P, int np, int sw0,
                                                            int kind)
{
  __shared__ Tabs T;
  __shared__ Mat Ms;
  extern __shared__ uint64_t dyn[];
  constexpr bool kDecode1 = false;  // stepped decodes instead
  constexpr bool SUM = true;
#include "tbg_pred.inc"
}

// predecessors_sub_worker (src/tb8gen.c:346-414) over every (job, side):
// uncapture from the stm^1 capture-exit bitmaps, at n == 1 (L: sub/win with
// the legality filter; W: sub/loss) and n == 101 (L: sub/cwin & ~sub/win;
// W: sub/bloss).
// With the block scatter, a job whose whole slice bitmap fits in lds_words
// (the dynamic LDS of the launch) OR-s the non-king uncaptures of a unit
// into an LDS copy of the slice and merges it with one device atomic per
// nonzero word, as pred's block scatter does; king uncaptures, and every
// uncapture of larger slices, use device atomics.
__global__ __launch_bounds__(256, 5) void k_uncap_j(JobDev *const *__restrict__ jobs,
                                                  const Pair *__restrict__ pairs,
                                                  const uint64_t *__restrict__ P, int np, int sw0,
                                                  int kind, const unsigned *__restrict__ flags,
                                                  uint64_t lds_words)
{
  __shared__ Tabs T;
  __shared__ Queue Q;
  __shared__ Mat Ms;
  extern __shared__ uint64_t dyn[];
  if (flags && !flags[1]) return;  // no job is at ply 1 or 101 (uniform)
  load_tabs(T);
  int cached = -1;
  int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  uint64_t total = P[np];
  // the (job, side) pair's fields, re-read only when a unit leaves its range
  // (the job record does not change during a launch)
  uint64_t plo = 1, phi = 0;
  bool en = false;
  int stm = 0, o = 1, n = 0, scatter = 0;
  const JobDev *Jp = nullptr;
  for (uint64_t u = blockIdx.x; u < total; u += gridDim.x) {
    if (u < plo || u >= phi) {  // uniform across the work-group
      int pi = find_pair(P, np, u);
      plo = P[pi];
      phi = P[pi + 1];
      Pair pr = pairs[pi];
      Jp = uniptr(jobs[pr.job]);
      stm = pr.side;
      o = stm ^ 1;
      n = uni(Jp->n);
      en = !uni(Jp->done) && ((uni(Jp->mask) >> (kind * 2 + stm)) & 1) &&
           (n == 1 || n == DRAW_RULE + 1) && !(kind == 1 && n == 1 && uni(Jp->w1cap));
      if (en) {
        cache_mat(Ms, Jp->M, pr.job, &cached);
        scatter = uni(Jp->scatter);
      }
    }
    if (!en) continue;
    const JobDev &J = *Jp;
    const Mat &M = Ms;
    uint64_t within;
    int s = unit_slice(M, sw0, u - plo, uncap_units(M, stm), &within);
    int k;
    for (k = 0; k < M.numsets; k++) {
      if (M.setcol[k] == stm) continue;
      uint64_t cpb = (M.cwords[k] + UNIT_WORDS - 1) / UNIT_WORDS;
      if (within < cpb) break;
      within -= cpb;
    }
    // at n == 1 the legal wins only: the legality filter of the reference's
    // uncapture (src/tb8gen.c:964-970) was applied by k_resolve
    const uint64_t *src0, *src1 = nullptr;
    if (kind == 0) {
      src0 = J.sub[o][n == 1 ? SUB_LWIN : TBG_SUB_CWIN];
      if (n != 1) src1 = J.sub[o][TBG_SUB_WIN];
    } else {
      src0 = J.sub[o][n == 1 ? TBG_SUB_LOSS : TBG_SUB_BLOSS];
    }
    // a wave per 64 words; a capture space of fewer than 256 words has its
    // words interleaved over the 4 waves instead (word = base + 4 * lane +
    // wave), so that every wave has work
    uint64_t ubase = within * UNIT_WORDS;
    bool il = M.cwords[k] < UNIT_WORDS;
    uint64_t wstep = il ? 4 : 1, woff = il ? (uint64_t)wave : (uint64_t)wave * 64;
    uint64_t w = ubase + (uint64_t)lane * wstep + woff;
    uint64_t x = 0;
    if (w < M.cwords[k]) {
      uint64_t a = phys(M, s) * M.ctotal[o] + M.coff[k] + w;
      x = src0[a];
      if (src1) x &= ~src1[a];
    }
    int tot;
    int off = wave_excl_scan(popc64(x), lane, &tot);
    const Idx &ci = M.capt[k];
    int m = M.ii.last[k];
    uint64_t *cand = kind == 0 ? J.cand[stm] : J.candW[stm];
    const uint64_t *won = J.wins[stm];
    uint64_t *own = cand + phys(M, s) * M.words;
    const uint64_t *ownw = won + phys(M, s) * M.words;
    // the L step's candidates up to ply 100 skip the statically refuted
    // positions too (the verification would drop them): at ply 1 nearly all
    // legal ones, so most device atomics are saved for one more read
    const uint64_t *oxr =
        kind == 0 && n <= DRAW_RULE && J.xr[stm] ? J.xr[stm] + phys(M, s) * M.words : nullptr;
    uint64_t nbw = (M.ii.size + 63) >> 6;
    bool lds = (scatter & TBG_SCATTER_BLOCK) && nbw <= lds_words;  // uniform
    uint64_t *B = dyn;
    if (lds) {
      if (!__syncthreads_or(tot != 0)) continue;  // nothing to uncapture in this unit
      for (uint64_t i = threadIdx.x; i < nbw; i += blockDim.x) B[i] = 0;
      __syncthreads();
    }
    for (int base = 0; base < tot; base += QCAP) {
      fill_round(Q.e[wave], x, off, base, lane);
      wave_sync();
      int cnt = tot - base < QCAP ? tot - base : QCAP;
      for (int e = lane; e < cnt; e += 64) {
        uint16_t v = Q.e[wave][e];
        uint64_t idx = (ubase + (uint64_t)(v >> 6) * wstep + woff) * 64 + (v & 63);
        if (idx >= ci.size) continue;
        uint64_t p = kings_of(T, s);
        uint64_t occ = idx_to_sq(T, ci, idx, &p);
        // uncapture by the king
        mark_king_unmoves(T, M, stm, occ, sq_set(p, m, sq_get(p, stm)), s, cand, won, ~0ull);
        // uncapture by the non-king pieces
        for (int t = 0; t < MAXP - 1; t++) {
          if (t >= M.nk[stm]) break;
          int j = M.pcs[stm][t];
          uint64_t q = sq_set(p, m, sq_get(p, j));
          if (lds) {
            uint64_t b = piece_att(T, M.pt[j], sq_get(q, j), occ) & ~occ;
            while (b) {
              int tq = ctz64(b);
              b &= b - 1;
              uint64_t q2 = sq_set(q, j, tq);
              uint64_t i2 = M.all_single ? sq_to_idx_1(M.ii, q2) : sq_to_idx(T, M.ii, q2);
              // into the 32-bit half-word holding the bit: one LDS bank
              __hip_atomic_fetch_or(reinterpret_cast<unsigned *>(B) + (i2 >> 5),
                                    1u << (i2 & 31), __ATOMIC_RELAXED,
                                    __HIP_MEMORY_SCOPE_WORKGROUP);
            }
          } else {
            mark_unmoves(T, M, j, occ, q, own, ownw, oxr);
          }
        }
      }
      wave_sync();
    }
    if (lds) {
      __syncthreads();
      for (uint64_t i = threadIdx.x; i < nbw; i += blockDim.x) {
        uint64_t v = B[i];
        if (v) atomic_or64(own + i, v);
      }
      __syncthreads();
    }
  }
}

// Attacks of colour c's non-king pieces on occupancy `o` (the caller lifts
// the enemy king from it, so rays pass through the square that king
// leaves), without the piece in slot `skip` (-1: none).
__device__ __forceinline__ uint64_t side_attacks(const Tabs &T, const Mat &M, uint64_t p,
                                                 uint64_t o, int c, int skip)
{
  uint64_t a = 0;
  for (int t = 0; t < MAXP - 1; t++) {
    if (t >= M.nk[c]) break;
    int j = M.pcs[c][t];
    if (j == skip) continue;
    a |= piece_att(T, M.pt[j], sq_get(p, j), o);
  }
  return a;
}

// Synthetic king-move exit value for o after stm's king moved from `from`
// (position q0 not normalised, q normalised into slice s2); captured slot
// jj of set l, or jj < 0.  3 = illegal move: the target is in `unsafe`, the
// opponent's attacks with stm's king lifted plus its king's squares (a
// captured piece neither attacks its own square nor blocks a ray that ends
// there).
__device__ int synth_exit(const Tabs &T, const Mat &M, int stm, uint64_t q0, uint64_t q,
                          uint64_t unsafe, int from, int jj, int l, int s2, uint64_t idx)
{
  int o = stm ^ 1, to = sq_get(q0, stm);
  if ((unsafe >> to) & 1) return 3;
  if (jj >= 0) {
    uint64_t c = sq_set(q, jj, sq_get(q, M.ii.last[l]));
    uint64_t ci = M.all_single ? capt_idx_1(M.capt[l], c, l) : sq_to_idx(T, M.capt[l], c);
    return synth_value(synth_hash(M.seedmix, o, 2, s2 * 8 + l, ci));
  }
  uint64_t i2 = !M.all_single ? sq_to_idx(T, M.ii, q)
                : q == q0     ? king_moved_idx(M.ii, idx, q0, from, to)
                              : sq_to_idx_1(M.ii, q);
  return synth_value(synth_hash(M.seedmix, o, 1, s2, i2));
}

// The king's moves of a verification unit: every candidate of a slice has
// the same kings, so where stm's king can step and how each step is
// normalised (slice, mirror, diagonal flip: src/tb8gen.c:230-238, a function
// of the kings alone) is worked out once per (slice, side) and kept in LDS,
// one entry per target square: s2 | m << 9 | fl << 15.  Lane i fills
// square i; call with the whole wave.
__device__ __forceinline__ void king_table(const Tabs &T, int s, int stm, uint32_t *kt, int lane)
{
  int ks = T.kksq[s][stm], oks = T.kksq[s][stm ^ 1];
  uint32_t e = 0;
  if (((T.king[ks] & ~T.king[oks]) >> lane) & 1) {
    int wk = stm == WHITE ? lane : oks, bk = stm == WHITE ? oks : lane;
    uint64_t q;
    int s2 = normalize(T, (uint64_t)wk | ((uint64_t)bk << 8), &q);
    e = (uint32_t)s2 | (uint32_t)mirror_mask(wk) << 9 | (uint32_t)flip_test(wk, bk) << 15;
  }
  kt[lane] = e;
}

// q0 (stm's king moved to a target of kt) normalised by its table entry
__device__ __forceinline__ uint64_t king_norm(uint32_t e, uint64_t q0)
{
  uint64_t q = q0 ^ (0x0101010101010101ull * (uint64_t)((e >> 9) & 63));
  return (e >> 15) & 1 ? flip_diag_bytes(q) : q;
}

// squares strictly between a and b on a shared rank, file or diagonal (0
// when a and b share none)
__device__ __forceinline__ uint64_t between_sq(int a, int b)
{
  int dr = (b >> 3) - (a >> 3), df = (b & 7) - (a & 7);
  if (dr == 0 || df == 0) return rook_att(a, bit(b)) & rook_att(b, bit(a));
  if (dr == df || dr == -df) return bishop_att(a, bit(b)) & bishop_att(b, bit(a));
  return 0;
}

// check_successors_worker + check_moves + check_king_moves
// (src/tb8gen.c:529-650): is candidate idx of slice s a loss for stm?
// Candidates come from cand & ~wins(stm), and wins(stm) holds every position
// that is illegal with stm to move from init on, so the legality test of
// the reference is never needed here; each piece's attack set is computed
// only when its moves are reached (most candidates fail at their first
// move).  Captures look up the captured piece's set from its slot.
// Returns 0 (not lost), 1 (lost), or 2 when `budget` bitmap reads were
// spent without an answer.
// SMALL: the slice has at most VERIFY_SPLIT_SIZE positions, so quiet-move
// targets are found with 32-bit offsets (moved_delta).
// xr: the candidate passed the static refutations (k_xr_pl), so every
// capture by a non-king piece and every synthetic king exit loses; only the
// quiet moves and the king's captures within the table are read.
template <bool SMALL>
__device__ __forceinline__ int verify_pos(const Tabs &T, const Mat &M, int stm, int s, uint64_t idx,
                          const uint64_t *__restrict__ wins_o, const uint64_t *__restrict__ sub_o,
                          const uint32_t *kt, int budget, uint64_t &hidx, uint64_t &hD,
                          uint64_t &hp, bool xr)
{
  int used = 0;  // bitmap reads so far; at budget the answer is left open (2)
  int o = stm ^ 1;
  // the lane's previous candidate of the slice (hidx, its digits hD and
  // squares hp) is stepped forward when it lies shortly before this one
  uint64_t p, occ;
  if (M.all_single && hidx != ~0ull && idx > hidx && idx - hidx < (1u << 24)) {
    p = hp;
    occ = step_by(M.ii.numsets, &hD, &p, (uint32_t)(idx - hidx));
  } else {
    p = kings_of(T, s);
    occ = idx_to_sq(T, M.ii, idx, &p);
    if (M.all_single) hD = digits_of(p, M.ii.numsets);
  }
  hidx = idx;
  hp = p;
  int ks = sq_get(p, stm), oks = sq_get(p, o);
  uint64_t oppocc = 0, orq = 0, obq = 0, on = 0;  // o's pieces; rooks and queens, ...
  for (int t = 0; t < MAXP - 1; t++) {
    if (t >= M.nk[o]) break;
    int jo = M.pcs[o][t], pt = M.pt[jo] & 7;
    uint64_t b = bit(sq_get(p, jo));
    oppocc |= b;
    orq |= pt == ROOK || pt == QUEEN ? b : 0;
    obq |= pt == BISHOP || pt == QUEEN ? b : 0;
    on |= pt == KNIGHT ? b : 0;
  }
  // stm in check: a quiet move of a piece other than the king is legal only
  // when it blocks a single slider's check; every other one leads to a
  // position illegal for o to move, which wins(o) holds from init on, so its
  // bit is known to be set and is not read.  (Large slices only: on small
  // ones the test's registers cost the one-pass form more than it saves.)
  uint64_t quiet_ok = ~0ull;
  if (!SMALL) {
    uint64_t chk = (rook_att(ks, occ) & orq) | (bishop_att(ks, occ) & obq) | (T.knight[ks] & on);
    quiet_ok = !chk ? ~0ull : chk & (chk - 1) ? 0 : between_sq(ctz64(chk), ks);
  }
  const uint64_t *wins_s = wins_o + phys(M, s) * M.words;
  const uint64_t *sub_s = sub_o + phys(M, s) * M.ctotal[o];
  for (int t = 0; t < MAXP - 1; t++) {
    if (t >= M.nk[stm]) break;
    int j = M.pcs[stm][t];
    int fr = sq_get(p, j);
    uint64_t a = piece_att(T, M.pt[j], fr, occ);
    uint64_t c = xr ? 0 : a & oppocc;
    while (c) {
      int to = ctz64(c);
      c &= c - 1;
      int jj = slot_of(p, to, M.col80[o]);
      int l = sq_get(M.set_of, jj);
      uint64_t q = sq_set(p, j, to);
      q = sq_set(q, jj, sq_get(q, M.ii.last[l]));
      if (used++ >= budget) return 2;
      uint64_t ci = M.all_single ? capt_idx_1(M.capt[l], q, l) : sq_to_idx(T, M.capt[l], q);
      if (!bm_test(sub_s + M.coff[l], ci)) return 0;
    }
    uint64_t qm = a & ~occ & quiet_ok;
    // quiet moves four at a time: the four bitmap reads are in flight
    // together instead of one round trip after another
    uint64_t ob = 0;
    uint32_t ra = 0;
    Later LS = {};
    if (M.all_single) {
      ob = occ_before(p, j - 2);
      ra = (uint32_t)(fr - popc64(ob & (bit(fr) - 1)));
      LS = later_sets_u(M.ii, p, j - 2);
    }
    while (qm) {
      if (used >= budget) return 2;
      used += 4;
      uint64_t i4[4], w4[4];
#pragma unroll
      for (int k = 0; k < 4; k++) {
        i4[k] = ~0ull;
        if (qm) {
          int to = ctz64(qm);
          qm &= qm - 1;
          if (SMALL && M.all_single)
            i4[k] = (uint32_t)((int)idx + moved_delta(LS, fr, to, ob, ra));
          else
            i4[k] = M.all_single ? moved_idx_l(LS, idx, fr, to, ob, ra)
                                 : sq_to_idx(T, M.ii, sq_set(p, j, to));
        }
      }
#pragma unroll
      for (int k = 0; k < 4; k++) w4[k] = i4[k] != ~0ull ? gload(wins_s, i4[k] >> 6) : ~0ull;
#pragma unroll
      for (int k = 0; k < 4; k++)
        if (!((w4[k] >> (i4[k] & 63)) & 1)) return 0;
    }
  }
  uint64_t kb = T.king[ks] & ~T.king[oks];
  // a synthetic job's king moves mostly leave the slice, and each such exit
  // needs a legality test: the opponent's attacks once, with the king lifted
  uint64_t unsafe =
      M.synth && !xr ? T.king[oks] | side_attacks(T, M, p, occ ^ bit(ks), o, -1) : 0;
  uint64_t c = kb & oppocc;
  while (c) {
    int to = ctz64(c);
    c &= c - 1;
    int jj = slot_of(p, to, M.col80[o]);
    int l = sq_get(M.set_of, jj);
    uint32_t e = kt[to];
    uint64_t q0 = sq_set(p, stm, to), q = king_norm(e, q0);
    int s2 = (int)(e & 511);
    if (M.synth && s2 != M.synth_slice && xr) continue;
    if (used++ >= budget) return 2;
    if (M.synth && s2 != M.synth_slice) {
      int v = synth_exit(T, M, stm, q0, q, unsafe, ks, jj, l, s2, idx);
      if (v == 3 || v == 2) continue;
      return 0;
    }
    q = sq_set(q, jj, sq_get(q, M.ii.last[l]));
    const uint64_t *sb = sub_o + phys(M, s2) * M.ctotal[o] + M.coff[l];
    uint64_t ci = M.all_single ? capt_idx_1(M.capt[l], q, l) : sq_to_idx(T, M.capt[l], q);
    if (!bm_test(sb, ci)) return 0;
  }
  // quiet king moves four at a time, like the pieces' (a lost candidate
  // reads every one); a synthetic job's exits need no read
  uint64_t kq = kb & ~occ;
  while (kq) {
    if (used >= budget) return 2;
    used += 4;
    uint64_t i4[4], w4[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      i4[k] = ~0ull;
      if (kq) {
        int to = ctz64(kq);
        kq &= kq - 1;
        uint32_t e = kt[to];
        uint64_t q0 = sq_set(p, stm, to), q = king_norm(e, q0);
        int s2 = (int)(e & 511);
        if (M.synth && s2 != M.synth_slice) {
          if (!xr) {
            int v = synth_exit(T, M, stm, q0, q, unsafe, ks, -1, 0, s2, idx);
            if (v != 3 && v != 2) return 0;
          }
        } else {
          uint64_t i2 = !M.all_single ? sq_to_idx(T, M.ii, q)
                        : q == q0     ? king_moved_idx(M.ii, idx, q0, ks, to)
                                      : sq_to_idx_1(M.ii, q);
          // SMALL: at most 462 slices of 2^22 positions, below 2^32 bits
          i4[k] = SMALL ? (uint64_t)((uint32_t)phys(M, s2) * (uint32_t)M.words * 64u + (uint32_t)i2)
                        : phys(M, s2) * M.words * 64 + i2;
        }
      }
    }
#pragma unroll
    for (int k = 0; k < 4; k++) w4[k] = i4[k] != ~0ull ? gload(wins_o, i4[k] >> 6) : ~0ull;
#pragma unroll
    for (int k = 0; k < 4; k++)
      if (!((w4[k] >> (i4[k] & 63)) & 1)) return 0;
  }
  return 1;
}

// The pull form's second condition for L_n: does some quiet move of
// candidate idx (stm to move, slice s) reach the frontier W_{n-1}(opp)
// (prev_o), i.e. is idx a predecessor of it (src/tb8gen.c:484-527 run
// backwards)?  Only candidates that passed the verification ask, so this
// is not on the hot path.  Captures and synthetic exits lead out of the
// table and never reach it.
__device__ bool has_hit(const Tabs &T, const Mat &M, int stm, int s, uint64_t idx,
                        const uint64_t *__restrict__ prev_o, const uint32_t *kt)
{
  int o = stm ^ 1;
  uint64_t p = kings_of(T, s);
  uint64_t occ = idx_to_sq(T, M.ii, idx, &p);
  int ks = sq_get(p, stm), oks = sq_get(p, o);
  const uint64_t *prev_s = prev_o + phys(M, s) * M.words;
  for (int t = 0; t < MAXP - 1; t++) {
    if (t >= M.nk[stm]) break;
    int j = M.pcs[stm][t];
    int fr = sq_get(p, j);
    uint64_t qm = piece_att(T, M.pt[j], fr, occ) & ~occ;
    uint64_t ob = 0;
    uint32_t ra = 0;
    Later LS = {};
    if (M.all_single) {
      ob = occ_before(p, j - 2);
      ra = (uint32_t)(fr - popc64(ob & (bit(fr) - 1)));
      LS = later_sets(M.ii, p, j - 2);
    }
    while (qm) {
      int to = ctz64(qm);
      qm &= qm - 1;
      uint64_t i2 = M.all_single ? moved_idx_l(LS, idx, fr, to, ob, ra)
                                 : sq_to_idx(T, M.ii, sq_set(p, j, to));
      if (bm_test(prev_s, i2)) return true;
    }
  }
  uint64_t kq = T.king[ks] & ~T.king[oks] & ~occ;
  while (kq) {
    int to = ctz64(kq);
    kq &= kq - 1;
    uint32_t e = kt[to];
    uint64_t q0 = sq_set(p, stm, to), q = king_norm(e, q0);
    int s2 = (int)(e & 511);
    if (M.synth && s2 != M.synth_slice) continue;
    uint64_t i2 = !M.all_single ? sq_to_idx(T, M.ii, q)
                  : q == q0     ? king_moved_idx(M.ii, idx, q0, ks, to)
                                : sq_to_idx_1(M.ii, q);
    if (bm_test(prev_o + phys(M, s2) * M.words, i2)) return true;
  }
  return false;
}

// check_successors (src/tb8gen.c:607-661) over every (job, side): prune
// the candidates in cand[stm] to L_n, written to newL[stm] (cand is cleared
// for the next half-step), counting into the job and writing L_n codes.  A
// candidate already in wins(stm) can not be lost, so it is dropped without
// a decode.  A step that does not run this ply writes an empty newL.  For a
// synthetic job at n == 1 the king-move exit candidates (computed by k_init
// into newW[stm]) join the candidates.
// Verification waits on chains of random bitmap reads.  The one-pass form
// runs at 5 waves per SIMD (at most 96 VGPRs, a few spilled), which beats 4
// waves without spills and 6 with more spills on small slices.  The
// two-pass form runs at 4 without spills: its deferred queue keeps every
// lane at work on the dense pull-form plies of large slices, where the
// spill traffic of 5 waves went to HBM (8-piece verify 490 -> 382 ms).
// SPLIT: the two-pass form for slices larger than VERIFY_SPLIT_SIZE (one
// pass otherwise, compiled without the second pass's bookkeeping).
// Pull form (the job's pull bit of stm, set by k_ctrl for a dense frontier
// when the static refutations are in place): no predecessor pass ran; the
// candidates are every position not won, not statically refuted and not in
// L_{n-1} or L_{n-2}, and one that passes the verification must also reach
// the frontier (has_hit).  The same L_n: a position is a predecessor of the
// frontier exactly when one of its quiet moves reaches it.
// A job's counter is one address for the whole grid: each wave adds its
// count once per (job, side) pair it worked on, not once per unit.
__device__ __forceinline__ void flush_count(unsigned long long *ctr, uint64_t &v, int lane)
{
  uint64_t t = __ballot(v != 0) ? wave_sum64(v) : 0;
  if (lane == 0 && t) atomicAdd(ctr, (unsigned long long)t);
  v = 0;
}

template <bool SPLIT, bool SUM>
__global__ __launch_bounds__(256, SPLIT ? 4 : 5) void k_verify_j(JobDev *const *__restrict__ jobs,
                                                   const Pair *__restrict__ pairs,
                                                   const uint64_t *__restrict__ P, int np, int sw0,
                                                   int uw)
{
  __shared__ Tabs T;
  __shared__ Queue Q;
  __shared__ uint32_t KT[4][64];  // king_table per wave
  __shared__ Mat Ms;
  // SPLIT: per wave, the candidates its first passes left open, deferred
  // across the unit's chunks (word offset in the unit << 6 | bit)
  __shared__ uint32_t LQ[4][SPLIT ? LQCAP : 1];
  __shared__ int LQn[4];
  load_tabs(T);
  int cached = -1;
  int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  uint64_t total = P[np];
  // The (job, side) pair of the current unit and what is read from its job
  // record, refreshed only when a unit leaves the pair's range: the record
  // does not change during a launch, and a sparse ply then costs one
  // candidate load per unit instead of a chain of dependent loads.  The
  // plane pointers are uniform and kept in scalar registers.
  uint64_t plo = 1, phi = 0;
  bool skip = true, run = false, exits = false, pull = false;
  int stm = 0, n = 0, kt_key = -1;
  JobDev *Jp = nullptr;
  const uint64_t *candr = nullptr, *extra = nullptr, *wins_s = nullptr, *wins_o = nullptr,
                 *sub_o = nullptr, *xrp = nullptr, *prevW_o = nullptr, *prevL_s = nullptr;
  uint64_t *candw = nullptr, *newL = nullptr;
  uint16_t *codes = nullptr;
  // chunk summaries of the candidates and of the L plane written (one-slice
  // jobs; null otherwise); sp: a sparse ply, only chunks summarised as
  // holding candidates or old L bits are read and written
  uint8_t *csm = nullptr, *lsm = nullptr;
  bool sp = false;
  uint64_t mysum = 0;  // L positions found for the current pair
  unsigned long long *ctr = nullptr;
  for (uint64_t u = blockIdx.x; u < total; u += gridDim.x) {
    if (u < plo || u >= phi) {  // uniform across the work-group
      if (ctr) flush_count(ctr, mysum, lane);
      int pi = find_pair(P, np, u);
      plo = P[pi];
      phi = P[pi + 1];
      Pair pr = pairs[pi];
      Jp = uniptr(jobs[pr.job]);
      skip = uni(Jp->done) != 0;
      if (!skip) {
        cache_mat(Ms, Jp->M, pr.job, &cached);
        stm = pr.side;
        n = uni(Jp->n);
        run = (uni(Jp->mask) >> stm) & 1;
        exits = run && Ms.synth && n == 1;
        int o = stm ^ 1;
        candw = uniptr(Jp->cand[stm]);
        candr = candw;
        extra = uniptr(Jp->newW[stm]);
        wins_s = uniptr(Jp->wins[stm]);
        wins_o = uniptr(Jp->wins[o]);
        sub_o = uniptr(Jp->sub[o][n <= DRAW_RULE ? TBG_SUB_WIN : TBG_SUB_CWIN]);
        // the static refutations hold for the 50-move rule's plies only
        xrp = n <= DRAW_RULE ? uniptr(Jp->xr[stm]) : nullptr;
        pull = run && xrp && ((uni(Jp->pull) >> stm) & 1);
        prevW_o = uniptr(Jp->prevW[o]);
        prevL_s = uniptr(Jp->prevL[stm]);
        newL = uniptr(Jp->newL[stm]);
        codes = uniptr(Jp->codes[stm]);
        bool has = SUM && uniptr(Jp->sum) != nullptr;
        csm = has ? uniptr(plane_sum(*Jp, candw)) : nullptr;
        lsm = has ? uniptr(plane_sum(*Jp, newL)) : nullptr;
        sp = has && uni(Jp->sparse) != 0;
      }
      ctr = skip ? nullptr : &Jp->cnt[stm];
    }
    if (skip) continue;
    const Mat &M = Ms;
    uint64_t within;
    uint32_t uwe = verify_unit(M, (uint32_t)uw);
    int s = unit_slice(M, sw0, u - plo, (M.words + uwe - 1) / uwe, &within);
    uint64_t u0 = within * (uint64_t)uwe;
    uint64_t uend = u0 + uwe < M.words ? u0 + uwe : M.words;
    uint64_t hidx = ~0ull, hD = 0, hp = 0;  // the lane's last decoded candidate of the unit
    // A large unit without candidates (and exits) only has its L words
    // cleared.  Every thread's candidate loads go out together; a unit with
    // some is then verified as below, its words found in cache.  (Large
    // units come with large slices, which only the SPLIT form runs.)
    if (SPLIT && uwe >= VERIFY_BIG_UNIT && !pull) {
      uint64_t sb = phys(M, s) * M.words;
      uint64_t any = 0;
      if (sp) {  // (a sparse ply runs no exits and no pull form)
        for (uint64_t c = (u0 >> 6) + threadIdx.x; c < ((uend + 63) >> 6); c += 256)
          any |= (uint64_t)(csm[c] | lsm[c]);
      } else {
        auto none = [](uint64_t x, uint64_t) { return x; };
        any = or_words(candr + sb, u0, uend, threadIdx.x, 256, none);
        if (exits) any |= or_words(extra + sb, u0, uend, threadIdx.x, 256, none);
      }
      if (!__syncthreads_or(any != 0)) {
        if (!sp) {  // (sparse: the unit's L words and summaries are zero already)
          for (uint64_t w = u0 + threadIdx.x; w < uend; w += 256) newL[sb + w] = 0;
          if (lsm)
            for (uint64_t c = (u0 >> 6) + threadIdx.x; c < ((uend + 63) >> 6); c += 256)
              lsm[c] = csm[c] = 0;
        }
        continue;
      }
    }
    if (s * 2 + stm != kt_key) {  // uniform; the wave's earlier reads of KT are done
      kt_key = s * 2 + stm;
      king_table(T, s, stm, KT[wave], lane);
      wave_sync();
    }
    if (SPLIT && lane == 0) LQn[wave] = 0;
    uint16_t code = (uint16_t)((n << 2) | 2);
    // one 64-word chunk per wave per pass (the 4 waves cover 256 words); a
    // chunk without candidates costs one load and one store per lane
    for (uint64_t wbase = u0 + (uint64_t)wave * 64; wbase < uend; wbase += 256) {
      uint64_t w = wbase + lane;
      // sparse: a chunk without candidates or old L bits is left alone
      bool need = !sp || (csm[wbase >> 6] | lsm[wbase >> 6]) != 0;
      bool live = w < uend && need;
      uint64_t a = phys(M, s) * M.words + w;
      uint64_t x0 = 0, x = 0;
      if (live) {
        x0 = gload(candr, a);
        uint64_t c = x0;
        if (pull) {  // every position of the word
          uint64_t b0 = w * 64;
          c = b0 >= M.ii.size ? 0 : M.ii.size - b0 >= 64 ? ~0ull : bit(M.ii.size - b0) - 1;
        }
        if (exits) c |= gload(extra, a);
        x = run && c ? c & ~gload(wins_s, a) : 0;
        if (x && xrp) x &= ~gload(xrp, a);
        // the pull form also drops L_{n-1} and, still in the plane this
        // step overwrites, L_{n-2} (at n = 2 the mates): an earlier loss
        // has no move into W_{n-1}(opp), so it never reaches the frontier,
        // and every move of a mate leads to wins(opp)
        if (x && pull) x &= ~gload(prevL_s, a) & ~gload(newL, a);
      }
      // SPLIT: the deferred candidates are finished once half a queue has
      // gathered, and at the wave's last chunk of the unit
      bool last = wbase + 256 >= uend;
      uint64_t r = 0;
      wave_sync();
      bool drain = SPLIT && (last || LQn[wave] >= LQCAP / 2);
      if (__ballot(x != 0) || drain) {
        Q.res[wave][lane] = x;
        int tot;
        int off = wave_excl_scan(popc64(x), lane, &tot);
        for (int base = 0; base < tot || base == 0; base += QCAP) {
          fill_round(Q.e[wave], x, off, base, lane);
          wave_sync();
          int cnt = tot - base < QCAP ? tot - base : QCAP;
          bool last_round = base + QCAP >= tot;
          // Most candidates are refuted by one of their first moves, a lost
          // one reads every move.  SPLIT: in a round of at least
          // VERIFY_SPLIT candidates a first pass spends at most
          // VERIFY_QUICK reads per candidate and defers the open ones to
          // the wave's queue LQ; a smaller round defers all of its
          // candidates there unread (in place, pass 1, when the queue is
          // full); pass 2 finishes the deferred ones of all chunks so far
          // with every lane busy.  Otherwise one pass.
          int todo = cnt;
          bool quick = cnt >= VERIFY_SPLIT;
          for (int pass = SPLIT ? 0 : 1; pass < (SPLIT ? 3 : 2); pass++) {
            if (pass == 2) {
              if (!(drain && last_round)) break;
              todo = LQn[wave] < LQCAP ? LQn[wave] : LQCAP;
              // the earlier chunks' L words are stored: complete before the
              // results are OR-ed into them
              asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
            }
            if (lane == 0) Q.n[wave] = 0;
            wave_sync();
            int budget = pass == 0 ? VERIFY_QUICK : 1 << 30;
            for (int e = lane; e < todo; e += 64) {
              uint64_t ww;  // the candidate's word
              int bt;
              if (pass < 2) {
                uint16_t v = Q.e[wave][e];
                ww = wbase + (v >> 6);
                bt = v & 63;
              } else {
                uint32_t v = LQ[wave][e];
                ww = u0 + (v >> 6);
                bt = (int)(v & 63);
              }
              uint64_t idx = ww * 64 + (uint64_t)bt;
              int r = idx >= M.ii.size ? 0
                      : pass == 0 && !quick
                          ? 2
                          : verify_pos<!SPLIT>(T, M, stm, s, idx, wins_o, sub_o, KT[wave], budget,
                                               hidx, hD, hp, xrp != nullptr);
              if (r == 1 && pull && !has_hit(T, M, stm, s, idx, prevW_o, KT[wave])) r = 0;
              bool here = ww >= wbase && ww < wbase + 64;  // a word of this chunk
              unsigned *res32 = reinterpret_cast<unsigned *>(Q.res[wave]);
              int hb = (int)(ww - wbase) * 64 + bt;  // its bit in Q.res when here
              if (pass < 2 && r != 1) {
                bool keep = false;  // open: stays a candidate of pass 1
                if (r == 2 && pass == 0) {
                  int k = __hip_atomic_fetch_add(&LQn[wave], 1, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_WORKGROUP);
                  if (k < LQCAP) {
                    LQ[wave][k] = (uint32_t)((ww - u0) << 6) | (uint32_t)bt;
                  } else {
                    Q.e[wave][__hip_atomic_fetch_add(&Q.n[wave], 1, __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_WORKGROUP)] =
                        (uint16_t)((ww - wbase) << 6 | bt);
                    keep = true;
                  }
                }
                if (!keep)
                  __hip_atomic_fetch_and(res32 + (hb >> 5), ~(1u << (hb & 31)), __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_WORKGROUP);
              } else if (pass == 2 && r == 1) {
                if (here) {  // this chunk's word is stored below, with its result
                  __hip_atomic_fetch_or(res32 + (hb >> 5), 1u << (hb & 31), __ATOMIC_RELAXED,
                                        __HIP_MEMORY_SCOPE_WORKGROUP);
                } else {  // an earlier chunk's word, stored already
                  uint64_t a2 = phys(M, s) * M.words + ww;
                  atomicOr(reinterpret_cast<unsigned long long *>(newL + a2),
                           (unsigned long long)bit(bt));
                  if (lsm) lsm[ww >> 6] = 1;
                  if (codes) codes[phys(M, s) * M.ii.size + idx] = code;
                  mysum++;
                }
              }
            }
            wave_sync();
            if (pass == 0) todo = Q.n[wave];
            if (pass == 2 && lane == 0) LQn[wave] = 0;
          }
        }
        r = Q.res[wave][lane];
        wave_sync();  // the next chunk overwrites Q.res
      }
      if (live) {
        newL[a] = r;
        if (x0) candw[a] = 0;
        if (codes && r) {
          uint16_t *c = codes + phys(M, s) * M.ii.size + w * 64;
          uint64_t rr = r;
          while (rr) {
            int b = ctz64(rr);
            rr &= rr - 1;
            c[b] = code;
          }
        }
      }
      if (lsm && need) {  // the chunk's summaries: L bits or none; candidates consumed
        bool nz = __ballot(r != 0) != 0;
        if (lane == 0) {
          lsm[wbase >> 6] = nz;
          csm[wbase >> 6] = 0;
        }
      }
      mysum += popc64(r);
    }
  }
  if (ctr) flush_count(ctr, mysum, lane);
}

template __global__ void k_verify_j<false, false>(JobDev *const *, const Pair *, const uint64_t *,
                                                  int, int, int);
template __global__ void k_verify_j<true, false>(JobDev *const *, const Pair *, const uint64_t *,
                                                 int, int, int);
template __global__ void k_verify_j<false, true>(JobDev *const *, const Pair *, const uint64_t *,
                                                 int, int, int);
template __global__ void k_verify_j<true, true>(JobDev *const *, const Pair *, const uint64_t *,
                                                int, int, int);

// calc_W_n tail (src/tb8gen.c:1072-1077) fused, over every (job, side):
// W = C & ~wins; wins |= W; count; codes; C cleared.  uw words per unit
// (a multiple of 512), two per lane per pass with 16-byte accesses.  A step that does not run
// writes an empty newW.  A synthetic job at n == 1 adds the king-move exit
// candidates computed by k_init into prevW[stm].
template <bool SUM>
__global__ __launch_bounds__(256) void k_sweep_j(JobDev *const *__restrict__ jobs,
                                                  const Pair *__restrict__ pairs,
                                                  const uint64_t *__restrict__ P, int np, int sw0,
                                                  int uw)
{
  __shared__ Mat Ms;
  int cached = -1;
  int lane = threadIdx.x & 63;
  uint64_t total = P[np];
  // the (job, side) pair's fields, re-read only when a unit leaves its range
  // (the job record does not change during a launch), in scalar registers
  uint64_t plo = 1, phi = 0;
  bool skip = true, run = false, exits = false;
  int stm = 0, n = 0;
  JobDev *Jp = nullptr;
  uint64_t *cand = nullptr, *wins = nullptr, *newW = nullptr;
  const uint64_t *kex = nullptr;
  uint16_t *codes = nullptr;
  // chunk summaries of the candidates and of the W plane written (one-slice
  // jobs); sp: a sparse ply, only chunks summarised as holding candidates
  // or old W bits are read and written
  uint8_t *csm = nullptr, *wsm = nullptr;
  bool sp = false;
  uint64_t mycnt = 0;  // W positions found for the current pair
  unsigned long long *ctr = nullptr;
  for (uint64_t u = blockIdx.x; u < total; u += gridDim.x) {
    if (u < plo || u >= phi) {  // uniform across the work-group
      if (ctr) flush_count(ctr, mycnt, lane);
      int pi = find_pair(P, np, u);
      plo = P[pi];
      phi = P[pi + 1];
      Pair pr = pairs[pi];
      Jp = uniptr(jobs[pr.job]);
      skip = uni(Jp->done) != 0;
      if (!skip) {
        cache_mat(Ms, Jp->M, pr.job, &cached);
        stm = pr.side;
        n = uni(Jp->n);
        run = (uni(Jp->mask) >> (2 + stm)) & 1;
        exits = run && Ms.synth && n == 1;
        cand = uniptr(Jp->candW[stm]);
        wins = uniptr(Jp->wins[stm]);
        newW = uniptr(Jp->newW[stm]);
        kex = uniptr(Jp->prevW[stm]);
        codes = uniptr(Jp->codes[stm]);
        bool has = SUM && uniptr(Jp->sum) != nullptr;
        csm = has ? uniptr(plane_sum(*Jp, cand)) : nullptr;
        wsm = has ? uniptr(plane_sum(*Jp, newW)) : nullptr;
        sp = has && uni(Jp->sparse) != 0;
      }
      ctr = skip ? nullptr : &Jp->cnt[2 + stm];
    }
    if (skip) continue;
    const Mat &M = Ms;
    uint64_t within;
    int s = unit_slice(M, sw0, u - plo, (M.words + uw - 1) / uw, &within);
    uint64_t base = phys(M, s) * M.words;
    uint64_t wend = (within + 1) * uw < M.words ? (within + 1) * uw : M.words;
    if (sp) {  // a unit whose chunks hold no candidates and no old W bits is left alone
      bool any = false;
      for (uint64_t ch = within * uw / 64 + threadIdx.x; ch < (wend + 63) / 64; ch += 256)
        any |= (csm[ch] | wsm[ch]) != 0;
      if (!__syncthreads_or(any)) continue;
    }
    // each pass of a wave covers two 64-word chunks: lanes 0-31 the first;
    // two passes at a time, so that both passes' loads are in flight together
    for (uint64_t w0 = within * uw + threadIdx.x * 2; w0 < wend; w0 += 1024) {
      ulonglong2 c[2], cc[2], wi[2];
      bool in[2], need[2];
#pragma unroll
      for (int k = 0; k < 2; k++) {
        uint64_t w = w0 + k * 512;
        in[k] = w < wend;
        // sparse: a chunk without candidates or old W bits is left alone
        need[k] = in[k] && (!sp || (csm[w >> 6] | wsm[w >> 6]) != 0);
        c[k] = make_ulonglong2(0, 0);
        if (need[k]) c[k] = *reinterpret_cast<const ulonglong2 *>(cand + base + w);
        cc[k] = c[k];
        if (exits && in[k]) {
          ulonglong2 e = *reinterpret_cast<const ulonglong2 *>(kex + base + w);
          cc[k].x |= e.x;
          cc[k].y |= e.y;
        }
        if (!run) cc[k] = make_ulonglong2(0, 0);
      }
#pragma unroll
      for (int k = 0; k < 2; k++) {
        wi[k] = make_ulonglong2(0, 0);
        if (cc[k].x | cc[k].y) wi[k] = *reinterpret_cast<const ulonglong2 *>(wins + base + w0 + k * 512);
      }
#pragma unroll
      for (int k = 0; k < 2; k++) {
        uint64_t w = w0 + k * 512;
        if (sp && !__ballot(need[k])) continue;
        uint64_t a = base + w;
        uint64_t ch = w >> 6;
        ulonglong2 wn = make_ulonglong2(0, 0);
        if (cc[k].x | cc[k].y) {
          wn.x = cc[k].x & ~wi[k].x;
          wn.y = cc[k].y & ~wi[k].y;
          if (wn.x | wn.y) {
            ulonglong2 o2;
            o2.x = wi[k].x | wn.x;
            o2.y = wi[k].y | wn.y;
            *reinterpret_cast<ulonglong2 *>(wins + a) = o2;
          }
        }
        if (need[k]) *reinterpret_cast<ulonglong2 *>(newW + a) = wn;
        if (c[k].x | c[k].y) *reinterpret_cast<ulonglong2 *>(cand + a) = make_ulonglong2(0, 0);
        if (wsm) {  // each half-wave's chunk: W bits or none; candidates consumed
          uint64_t nzb = __ballot((wn.x | wn.y) != 0);
          if ((lane & 31) == 0 && need[k]) {
            wsm[ch] = ((nzb >> (lane & 32)) & 0xffffffffull) != 0;
            csm[ch] = 0;
          }
        }
        mycnt += popc64(wn.x) + popc64(wn.y);
        if (codes && (wn.x | wn.y)) {
          uint16_t code = (uint16_t)((n << 2) | 1);
          for (int h = 0; h < 2; h++) {
            uint64_t r = h ? wn.y : wn.x;
            uint16_t *cp = codes + phys(M, s) * M.ii.size + (w + h) * 64;
            while (r) {
              int b = ctz64(r);
              r &= r - 1;
              cp[b] = code;
            }
          }
        }
      }
    }
  }
  if (ctr) flush_count(ctr, mycnt, lane);
}

template __global__ void k_sweep_j<false>(JobDev *const *, const Pair *, const uint64_t *, int, int,
                                          int);
template __global__ void k_sweep_j<true>(JobDev *const *, const Pair *, const uint64_t *, int, int,
                                         int);

// End of a ply for every job (one lane per job): record the counts, derive
// the next ply's flags exactly as main() does (src/tb8gen.c:1416-1423: an
// L(stm) result enables W(stm^1), a W(stm) result enables L(stm^1)), swap
// the W/L planes, and advance n.  Plies 2..100 in which nothing can run are
// skipped to ply 101, which always runs (src/tb8gen.c:1426-1433); after it
// the job ends once every flag is false (src/tb8gen.c:1435).
__device__ void ctrl_one(JobDev &J, unsigned *flags, unsigned *any_edge);

__global__ __launch_bounds__(1024) void k_ctrl(JobDev *const *jobs, int nj, unsigned *flags,
                                                unsigned *done_out)
{
  __shared__ unsigned any_edge;
  if (threadIdx.x == 0) any_edge = 0;
  __syncthreads();
  for (int j = threadIdx.x; j < nj; j += blockDim.x) {
    if (!jobs[j]->done) ctrl_one(*jobs[j], flags, &any_edge);
    if (done_out) done_out[j] = (unsigned)jobs[j]->done | ((unsigned)jobs[j]->maxn << 2);
  }
  __syncthreads();
  if (threadIdx.x == 0) flags[1] = any_edge;
}

__device__ void ctrl_one(JobDev &J, unsigned *flags, unsigned *any_edge)
{
  int n = J.n;
  if (J.hist)
    for (int i = 0; i < 4; i++)
      J.hist[(uint64_t)n * 4 + i] = ((J.mask >> i) & 1) ? (long long)J.cnt[i] : -1;
  bool lw = (J.mask & 1) && J.cnt[0], lb = (J.mask & 2) && J.cnt[1];
  bool ww = (J.mask & 4) && J.cnt[2], wb = (J.mask & 8) && J.cnt[3];
  // more_ww <- L_n(B), more_wb <- L_n(W), more_lw <- W_n(B), more_lb <- W_n(W)
  unsigned more = (lb ? 1u : 0u) | (lw ? 2u : 0u) | (wb ? 4u : 0u) | (ww ? 8u : 0u);
  if (more) J.maxn = n;
  for (int c = 0; c < 2; c++) {
    uint64_t *x = J.prevW[c];
    J.prevW[c] = J.newW[c];
    J.newW[c] = x;
    x = J.prevL[c];
    J.prevL[c] = J.newL[c];
    J.newL[c] = x;
  }
  // W_n of each side: the next ply's L frontier of the other
  unsigned long long wcnt[2] = { (J.mask & 4) ? J.cnt[2] : 0ull, (J.mask & 8) ? J.cnt[3] : 0ull };
  unsigned long long front = 0;  // the next ply's frontiers, all four
  for (int i = 0; i < 4; i++) front += ((J.mask >> i) & 1) ? J.cnt[i] : 0ull;
  for (int i = 0; i < 4; i++) J.cnt[i] = 0;
  // this ply wrote the summaries of its planes
  J.sum_plies = J.sum_plies < 2 ? J.sum_plies + 1 : 2;
  n++;
  J.more = more;
  if ((n > DRAW_RULE + 1 && !more) || n >= J.max_ply) {
    J.n = n;
    J.mask = 0;
    J.sparse = 0;
    J.done = 1;
    atomicAdd(&flags[0], 1u);
    return;
  }
  if (n < DRAW_RULE + 1 && !more) n = DRAW_RULE + 1;
  bool forced = n == DRAW_RULE + 1;
  J.mask = ((forced || (more & 4)) ? 1u : 0u) | ((forced || (more & 8)) ? 2u : 0u) |
           ((forced || (more & 1)) ? 4u : 0u) | ((forced || (more & 2)) ? 8u : 0u);
  J.n = n;
  // the pull form for an L step whose frontier is large beside the
  // positions it would verify: each frontier position costs the predecessor
  // pass tens of unmoves, each open position the pull form about one
  // refuted verification
  unsigned pull = 0;
  if (n >= 2 && n <= DRAW_RULE) {
    for (int c = 0; c < 2; c++)
      if (((J.mask >> c) & 1) && J.xr[c] && J.nopen && wcnt[c ^ 1] * PULL_RATIO >= J.nopen[c])
        pull |= 1u << c;
  }
  J.pull = pull;
  // a sparse ply: a frontier whose unmoves (about SPARSE_UNMOVES each) can
  // mark few of the planes' chunks, the summaries of both ping-pong planes
  // known, no uncapture and no pull-form step
  J.sparse = J.sum && !forced && !pull && J.sum_plies >= 2 && front * SPARSE_UNMOVES < J.sum_pl;
  if (forced) atomicOr(any_edge, 1u);
  if (J.stop_at > 0 && n >= J.stop_at) {  // a ply boundary the host asked to stop at
    J.done = 2;
    atomicAdd(&flags[0], 1u);
  }
}

__global__ __launch_bounds__(256) void k_code_plane(const uint16_t *codes, uint64_t size,
                                                    uint64_t words, int nslices, uint16_t code,
                                                    uint64_t *out)
{
  int lane = threadIdx.x & 63;
  uint64_t nw = (uint64_t)nslices * words;
  for (uint64_t g = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; g < nw;
       g += ((uint64_t)gridDim.x * blockDim.x) >> 6) {
    uint64_t s = g / words, w = g % words;
    uint64_t idx = w * 64 + lane;
    bool hit = idx < size && codes[s * size + idx] == code;
    uint64_t b = __ballot(hit);
    if (lane == 0) out[g] = b;
  }
}

// Finished codes narrowed to one byte each (a table whose plies are all
// below 64 keeps n << 2 | tag below 256): the served form's copy of a
// finished slice to host memory then moves half the bytes.  Eight codes per
// thread: a 16-byte load and an 8-byte store.
__global__ __launch_bounds__(256) void k_codes8(const uint16_t *__restrict__ codes, uint64_t n,
                                                uint8_t *__restrict__ out)
{
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x * 8;
  for (uint64_t i = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) * 8; i < n; i += stride) {
    if (i + 8 <= n) {
      uint4 v = *reinterpret_cast<const uint4 *>(codes + i);
      uint32_t lo = (v.x & 0xff) | ((v.x >> 16) & 0xff) << 8 | (v.y & 0xff) << 16 |
                    ((v.y >> 16) & 0xff) << 24;
      uint32_t hi = (v.z & 0xff) | ((v.z >> 16) & 0xff) << 8 | (v.w & 0xff) << 16 |
                    ((v.w >> 16) & 0xff) << 24;
      *reinterpret_cast<uint2 *>(out + i) = make_uint2(lo, hi);
    } else {
      for (uint64_t k = i; k < n; k++) out[k] = (uint8_t)codes[k];
    }
  }
}

// Synthetic job, ply 1: does stm's king have a legal move or capture into
// an exit (a position of another slice, valued by the seeded hash) that is
// a loss (*w: W_1 candidate) or a win (*l: L_1 candidate) for the opponent?
// uo: the opponent's piece attacks with stm's king lifted, so a king
// move's legality, capture or not, is a bit test.
__device__ void synth_exits_pos(const Tabs &T, const Mat &M, uint64_t p, uint64_t occ, int stm,
                                uint64_t idx, uint64_t uo, bool *w, bool *l, bool *r)
{
  int o = stm ^ 1, from = sq_get(p, stm), oks = sq_get(p, o);
  uint64_t own = 0;
  for (int t = 0; t < MAXP - 1; t++) {
    if (t >= M.nk[stm]) break;
    own |= bit(sq_get(p, M.pcs[stm][t]));
  }
  uint64_t blocked = own | T.king[oks] | uo;
  bool ww = false, ll = false, rr = false;
  AllSets A = all_sets(M.ii, p);
#pragma unroll
  for (int k = 0; k < MAXS; k++) A.f[k] = (uint32_t)uni((int)A.f[k]);
  // the king's moves out of the slice, their slice and symmetry prepared on
  // the host
  for (int e = 0; e < 8; e++) {
    if (e >= M.nexit[stm]) break;
    int to = M.exit_to[stm][e];
    if (blocked & bit(to)) continue;  // own piece, or the king would stand in check
    int m = M.exit_m[stm][e], s2 = M.exit_s2[stm][e];
    bool fl = M.exit_fl[stm][e] != 0;
    uint64_t q0 = sq_set(p, stm, to);
    uint64_t q = q0 ^ (0x0101010101010101ull * (uint64_t)m);
    if (fl) q = flip_diag_bytes(q);
    int v;
    if (occ & bit(to)) {
      int jj = slot_of(p, to, M.col80[o]);
      int k = sq_get(M.set_of, jj);
      uint64_t c = sq_set(q, jj, sq_get(q, M.ii.last[k]));
      uint64_t ci = M.all_single ? capt_idx_1(M.capt[k], c, k) : sq_to_idx(T, M.capt[k], c);
      v = synth_value(synth_hash(M.seedmix, o, 2, s2 * 8 + k, ci));
    } else {
      uint64_t i2 = !M.all_single ? sq_to_idx(T, M.ii, q)
                    : m == 0 && !fl ? king_moved_idx_a(A, idx, from, to)
                                    : sq_to_idx_a(A, q);
      v = synth_value(synth_hash(M.seedmix, o, 1, s2, i2));
    }
    ww |= v == -2;
    ll |= v == 2;
    rr |= v != 2;  // a legal exit not into a win of the opponent
  }
  *w = ww;
  *l = ll;
  *r = rr;
}

// has_legal_moves || has_legal_caps (src/movegen.c:75-111) for colour c:
// the same answer, found with less work.  uo: the opponent's piece attacks
// with c's king lifted (a slider's ray passes through the square the king
// leaves), so a king move's legality, capture or not, is a bit test.  Only
// if no king move is legal are the other pieces' moves tried: under a
// double check none can be legal, under a single check exactly the moves of
// unpinned pieces onto the checker or between it and the king are, and
// without a check each move gets a full attack test (pins).
__device__ bool has_legal(const Tabs &T, const Mat &M, uint64_t p, uint64_t occ, int c, uint64_t uo)
{
  int o = c ^ 1;
  int k = sq_get(p, c), ok = sq_get(p, o);
  uint64_t own = bit(k);
  for (int t = 0; t < MAXP - 1; t++) {
    if (t >= M.nk[c]) break;
    own |= bit(sq_get(p, M.pcs[c][t]));
  }
  uint64_t kb = T.king[k] & ~own & ~T.king[ok];
  // a king move or capture to a safe square (a captured piece neither
  // attacks its own square nor blocks a ray that ends there)
  if (kb & ~uo) return true;
  // checkers (the lifted king does not block rays onto its own square)
  uint64_t occk = occ ^ bit(k);
  uint64_t block = ~0ull;
  int nchk = 0;
  for (int t = 0; t < MAXP - 1; t++) {
    if (t >= M.nk[o]) break;
    int j = M.pcs[o][t];
    int s = sq_get(p, j);
    int pt = M.pt[j] & 7;
    if (!(piece_att(T, M.pt[j], s, occk) & bit(k))) continue;
    uint64_t b = bit(s);
    if (pt == ROOK || pt == QUEEN) {
      uint64_t r = rook_att(s, occ) & rook_att(k, occ);
      if ((s >> 3) == (k >> 3) || (s & 7) == (k & 7)) b |= r;
    }
    if (pt == BISHOP || pt == QUEEN) {
      uint64_t r = bishop_att(s, occ) & bishop_att(k, occ);
      if ((s >> 3) - (k >> 3) == (s & 7) - (k & 7) || (s >> 3) + (s & 7) == (k >> 3) + (k & 7))
        b |= r;
    }
    block = nchk ? 0 : b;
    nchk++;
  }
  if (nchk > 1) return false;
  block &= ~bit(ok);  // the enemy king is never captured
  if (nchk == 1) {
    // A pinned piece can not answer the check (its pin line and the check's
    // meet only at the king), and any other piece's move onto the checker
    // or between it and the king does (it opens no line onto the king): one
    // attack test per piece instead of a full attack test per move.
    uint64_t pinned = 0;
    for (int t = 0; t < MAXP - 1; t++) {
      if (t >= M.nk[o]) break;
      int j = M.pcs[o][t];
      int s = sq_get(p, j), pt = M.pt[j] & 7;
      bool line = (s >> 3) == (k >> 3) || (s & 7) == (k & 7);
      if (!(line ? pt == ROOK || pt == QUEEN : pt == BISHOP || pt == QUEEN)) continue;
      uint64_t on = between_sq(s, k) & occ;  // empty when not aligned
      if (on && !(on & (on - 1)) && (on & own)) pinned |= on;
    }
    for (int t = 0; t < MAXP - 1; t++) {
      if (t >= M.nk[c]) break;
      int j = M.pcs[c][t];
      int from = sq_get(p, j);
      if (pinned & bit(from)) continue;
      if (piece_att(T, M.pt[j], from, occ) & ~own & block) return true;
    }
    return false;
  }
  for (int t = 0; t < MAXP - 1; t++) {
    if (t >= M.nk[c]) break;
    int j = M.pcs[c][t];
    int from = sq_get(p, j);
    uint64_t b = piece_att(T, M.pt[j], from, occ) & ~own & block;
    while (b) {
      int to = ctz64(b);
      b &= b - 1;
      int cap = (occ & bit(to)) ? slot_of(p, to, M.col80[o]) : -1;
      uint64_t q = sq_set(p, j, to);
      uint64_t occ2 = (occ ^ bit(from)) | bit(to);
      if (!attacked_by(T, M, q, occ2, k, o, cap)) return true;
    }
  }
  return false;
}

// calc_illegal_worker + calc_mate_worker (src/tb8gen.c:664-727), dense: one
// position per lane.  Writes illegal(c) into wins[c], mates into mate[c],
// and every code.
__global__ __launch_bounds__(256, 6) void k_init(const Mat *__restrict__ Mp, uint64_t *winsW,
                                               uint64_t *winsB, uint64_t *mateW, uint64_t *mateB,
                                               uint16_t *codesW, uint16_t *codesB,
                                               uint64_t *kexW0, uint64_t *kexW1, uint64_t *kexL0,
                                               uint64_t *kexL1, uint64_t *xrW, uint64_t *xrB,
                                               int s0, int ns, unsigned long long *cnt)
{
  __shared__ Tabs T;
  load_tabs(T);
  const Mat &M = *Mp;
  int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  uint64_t nw = (uint64_t)ns * M.words;
  uint64_t c0 = 0, c1 = 0, c2 = 0, c3 = 0;
  // each wave takes INIT_CHUNK consecutive words; for an all-single material
  // a lane's position then steps by 64 from word to word (step64) instead of
  // a full decode
  bool stepping = M.all_single;
  int nsets = M.ii.numsets;
  for (uint64_t g0 = ((uint64_t)blockIdx.x * 4 + wave) * INIT_CHUNK; g0 < nw;
       g0 += (uint64_t)gridDim.x * 4 * INIT_CHUNK)
  for (uint64_t g = g0, D = 0, pk = 0; g < g0 + INIT_CHUNK && g < nw; g++) {
    int s = s0 + (int)(g / M.words);
    uint64_t w = g % M.words;
    uint64_t idx = w * 64 + lane;
    bool wi = false, bi = false, wm = false, bm = false;
    bool xw0 = false, xw1 = false, xl0 = false, xl1 = false, xr0 = false, xr1 = false;
    if (idx < M.ii.size) {
      // the previous word's position of this lane is index - 64 of the same
      // slice unless the chunk or the slice starts here
      uint64_t p, occ;
      if (stepping && g != g0 && w != 0) {
        p = pk;
        occ = step64(nsets, &D, &p);
      } else {
        p = kings_of(T, s);
        occ = idx_to_sq(T, M.ii, idx, &p);
        if (stepping) D = digits_of(p, nsets);
      }
      pk = p;
      int wk = sq_get(p, WHITE), bk = sq_get(p, BLACK);
      // each side's piece attacks with the enemy king lifted: the legality
      // tests, the exits' king moves and the mate search all read them
      uint64_t uw = side_attacks(T, M, p, occ ^ bit(bk), WHITE, -1);
      uint64_t ub = side_attacks(T, M, p, occ ^ bit(wk), BLACK, -1);
      wi = ((uw | T.king[wk]) >> bk) & 1;  // white to move could take the black king
      bi = ((ub | T.king[bk]) >> wk) & 1;
      // an illegal position is in wins(stm) from the start, which masks its
      // exit candidates in both the W sweep and verification
      if (M.synth) {
        if (!wi) synth_exits_pos(T, M, p, occ, WHITE, idx, ub, &xw0, &xl0, &xr0);
        if (!bi) synth_exits_pos(T, M, p, occ, BLACK, idx, uw, &xw1, &xl1, &xr1);
      }
      if (wi != bi) {
        int c = bi ? WHITE : BLACK;
        bool mate = !has_legal(T, M, p, occ, c, c == WHITE ? ub : uw);
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
    if (xrW) {  // the exit part of the static refutations (zero for a table)
      uint64_t r0 = __ballot(xr0), r1 = __ballot(xr1);
      if (lane == 0) {
        xrW[a] = r0;
        xrB[a] = r1;
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

// ------------------------------------------------- init by placements
// For an all-single material, the positions that share the kings and the
// pieces of every set but the last differ only in the last set's square t,
// and form a run of R = 62 - L consecutive indices (L = the last set; its
// digit is the rank of t among the R squares the others leave free).  The
// placement-wise init evaluates such a run (a "placement") per lane:
// legality and king safety become 64-square masks over t, computed once per
// placement, and each synthetic exit's index is a per-placement base plus
// the rank of the moved position's last square.


// The squares t of the last set's piece (slot jx) at which square y is
// attacked by colour c's other non-king pieces over occupancy occ (which
// leaves out slot jx): a piece at t blocks a slider's ray onto y exactly
// when t lies strictly between them.  All squares if a knight attacks y,
// none if nothing does.
__device__ __forceinline__ uint64_t fixed_attacks_on(const Tabs &T, const Mat &M, uint64_t p,
                                                     uint64_t occ, int y, int c, int jx)
{
  // the attackers found from y (a slider attacks y exactly when y's rays of
  // its kind reach it): two ray sets instead of one per piece
  uint64_t rq = 0, bq = 0, nn = 0;  // rooks and queens, bishops and queens, knights
  for (int i = 0; i < MAXP - 1; i++) {
    if (i >= M.nk[c]) break;
    int j = M.pcs[c][i];
    if (j == jx) continue;
    int pt = M.pt[j] & 7;
    uint64_t b = bit(sq_get(p, j));
    rq |= pt == ROOK || pt == QUEEN ? b : 0;
    bq |= pt == BISHOP || pt == QUEEN ? b : 0;
    nn |= pt == KNIGHT ? b : 0;
  }
  uint64_t atk = T.knight[y] & nn;
  if (rq) atk |= rook_att(y, occ) & rq;
  if (bq) atk |= bishop_att(y, occ) & bq;
  uint64_t inter = ~0ull;
  for (uint64_t b = atk; b; b &= b - 1) {
    int s = ctz64(b);
    inter &= (nn >> s) & 1 ? 0ull : between_sq(s, y);
  }
  return atk ? ~inter : 0ull;
}

// The squares t from which the last set's piece (type pt, colour c) would
// attack y, the others standing on occ: reverse attacks from y, y excluded.
__device__ __forceinline__ uint64_t last_attacks_on(const Tabs &T, int pt, int y, uint64_t occ)
{
  return piece_att(T, pt, y, occ) & ~bit(y);
}

// Is square y attacked by colour o, as a mask over t: o's other pieces
// (blockable at t) and, if the last piece is o's, that piece from t.
// occ: the occupancy without the last piece (and without a king lifted by
// the caller).
__device__ __forceinline__ uint64_t attacked_mask(const Tabs &T, const Mat &M, uint64_t p,
                                                  uint64_t occ, int y, int o, int jx, int cx,
                                                  int ptx)
{
  uint64_t a = fixed_attacks_on(T, M, p, occ, y, o, jx);
  if (cx == o) a |= last_attacks_on(T, ptx, y, occ);
  return a;
}

// sq_to_idx over the sets before set L of packed squares q (set `skip`
// left out, -1: none) with the factors of ii; *ob receives the occupancy
// of the kings and those sets, against which the last set's square is
// ranked
__device__ __forceinline__ uint64_t partial_idx(const Idx &ii, uint64_t q, int L, int skip,
                                                uint64_t *ob)
{
  uint64_t occ = bit(sq_get(q, 0)) | bit(sq_get(q, 1));
  uint64_t idx = 0;
  for (int k = 0; k < MAXS; k++) {
    if (k >= L) break;
    if (k == skip) continue;
    int s = sq_get(q, 2 + k);
    idx += (uint64_t)(uint32_t)(s - popc64(occ & (bit(s) - 1))) * ii.factor[k];
    occ |= bit(s);
  }
  *ob = occ;
  return idx;
}

// the square t moved by a king exit's symmetry (mirror mask m, then the
// a1-h8 flip)

// wins W/B, mates W/B, exit W candidates W/B, exit L candidates W/B, exit
// refutations W/B
constexpr int IB_PLANES = 10;
constexpr int IB_WORDS = 64;  // words per wave of one plane (R <= 62 of them used)

// calc_illegal_worker + calc_mate_worker (src/tb8gen.c:664-727) and the
// synthetic exits, one placement per lane, for an all-single material with
// at least 64 placements per slice.  A wave takes 64 consecutive placements
// = 64 R positions = R bitmap words of each plane (a "chunk"), assembles
// them in LDS and stores each word once.  Same outputs as k_init, and for
// a synthetic job with xrW/xrB the exit refutations: positions with a legal
// king exit whose value is not a win for the side to move after it (such a
// position is never lost, whatever the ply), and the mates.
__global__ __launch_bounds__(256, 4) void k_init_pl(const Mat *__restrict__ Mp, uint64_t *winsW,
                                                 uint64_t *winsB, uint64_t *mateW, uint64_t *mateB,
                                                 uint16_t *codesW, uint16_t *codesB,
                                                 uint64_t *kexW0, uint64_t *kexW1,
                                                 uint64_t *kexL0, uint64_t *kexL1, uint64_t *xrW,
                                                 uint64_t *xrB, int s0, int ns,
                                                 unsigned long long *cnt)
{
  __shared__ Tabs T;
  __shared__ Mat Ms;
  __shared__ uint64_t buf[4][IB_PLANES][IB_WORDS];
  load_tabs(T);
  {
    const uint32_t *a = reinterpret_cast<const uint32_t *>(Mp);
    uint32_t *b = reinterpret_cast<uint32_t *>(&Ms);
    for (int i = threadIdx.x; i < (int)(sizeof(Mat) / 4); i += blockDim.x) b[i] = a[i];
    __syncthreads();
  }
  const Mat &M = Ms;
  int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  uint64_t(*B)[IB_WORDS] = buf[wave];
  const int L = M.ii.numsets - 1, jx = 2 + L, R = 62 - L;
  const int cx = (M.pt[jx] >> 3) & 1, ptx = M.pt[jx];
  const uint64_t np = M.ii.size / (uint64_t)R;    // placements per slice
  const uint64_t cps = (M.words + R - 1) / R;     // chunks per slice (padding words included)
  uint64_t cw[4] = { 0, 0, 0, 0 };                // counts: illegal W, B, mates W, B
  for (uint64_t gc = (uint64_t)blockIdx.x * 4 + wave; gc < (uint64_t)ns * cps;
       gc += (uint64_t)gridDim.x * 4) {
    int s = s0 + (int)(gc / cps);
    uint64_t g = gc % cps;
    uint64_t P = g * 64 + lane;
    bool valid = P < np;
    // this lane's placement: the squares of every set but the last
    uint64_t p = kings_of(T, s);
    uint64_t occF = 0;
    if (valid) occF = idx_to_sq_1(M.ii, P * (uint64_t)R, &p) & ~bit(sq_get(p, jx));
    else occF = bit(sq_get(p, 0)) | bit(sq_get(p, 1));
    uint64_t F = ~occF;  // the R squares of the last piece
    int wk = sq_get(p, WHITE), bk = sq_get(p, BLACK);
    // illegal(c): c to move can take the other king
    uint64_t WI = attacked_mask(T, M, p, occF, bk, WHITE, jx, cx, ptx);
    uint64_t BI = attacked_mask(T, M, p, occF, wk, BLACK, jx, cx, ptx);
    if (T.king[wk] & bit(bk)) WI = BI = ~0ull;
    // king moves of each side that stay legal (own piece not on the
    // target, target not attacked with the king lifted), as masks over t,
    // OR-ed into the escapes for the mate test (named registers, not an
    // array indexed by the side, which would live in scratch memory)
    uint64_t esc0 = 0, esc1 = 0;
    for (int c = 0; c < 2; c++) {
      int o = c ^ 1, k = sq_get(p, c);
      uint64_t own = 0;
      for (int i = 0; i < MAXP - 1; i++) {
        if (i >= M.nk[c]) break;
        int j = M.pcs[c][i];
        if (j != jx) own |= bit(sq_get(p, j));
      }
      uint64_t ys = T.king[k] & ~own & ~T.king[sq_get(p, o)];
      uint64_t e = 0;
      while (ys) {
        int y = ctz64(ys);
        ys &= ys - 1;
        uint64_t ok = ~attacked_mask(T, M, p, occF ^ bit(k), y, o, jx, cx, ptx);
        if (cx == c) ok &= ~bit(y);
        e |= ok;
      }
      if (c == WHITE) esc0 = e;
      else esc1 = e;
    }
    // mates: in check, not illegal, no king move.  The other moves are
    // tried only for those t, by the full test, with the wave's candidates
    // queued in LDS (the wave's assembly buffer, not yet in use) so that
    // every lane tests one: a lane's own candidates number 0 to R.
    // the mates found stay in the wave's buffer (qm) until the exits are
    // done: they are not needed before, and registers are short there
    uint64_t *qm = reinterpret_cast<uint64_t *>(reinterpret_cast<uint8_t *>(B) + 2 * QCAP) + 64;
    {
      uint8_t *wb = reinterpret_cast<uint8_t *>(B);
      uint16_t *qe = reinterpret_cast<uint16_t *>(wb);          // QCAP entries: lane << 6 | t
      uint64_t *qp = reinterpret_cast<uint64_t *>(wb + 2 * QCAP);  // each lane's placement; qm: mates found, [side][lane]
      static_assert(2 * QCAP + 3 * 64 * 8 <= IB_PLANES * IB_WORDS * 8, "mate queue fits");
      qp[lane] = p;
      qm[lane] = 0;
      qm[64 + lane] = 0;
      for (int c = 0; c < 2; c++) {
        uint64_t cand = valid ? (c == WHITE ? BI & ~WI & ~esc0 : WI & ~BI & ~esc1) & F : 0;
        int tot;
        int off = wave_excl_scan(popc64(cand), lane, &tot);
        for (int base = 0; base < tot; base += QCAP) {
          fill_round(qe, cand, off, base, lane);
          wave_sync();
          int cntq = tot - base < QCAP ? tot - base : QCAP;
          for (int e = lane; e < cntq; e += 64) {
            uint16_t v = qe[e];
            int owner = v >> 6, t = v & 63;
            uint64_t pt_ = sq_set(qp[owner], jx, t), occ = 0;
            for (int i = 0; i < MAXP; i++)
              if (i < M.num) occ |= bit(sq_get(pt_, i));
            // no king move is legal here: only the other moves are tried
            if (!has_legal(T, M, pt_, occ, c, ~0ull))
              __hip_atomic_fetch_or(&qm[c * 64 + owner], bit(t), __ATOMIC_RELAXED,
                                    __HIP_MEMORY_SCOPE_WORKGROUP);
          }
          wave_sync();
        }
      }
    }
    // synthetic exits: each of the king's moves out of the slice, its
    // index a per-placement base plus the rank of the moved position's last
    // square; the candidates collect over squares t
    // per side: W candidates, L candidates, exit refutations (named
    // registers: an array indexed by the side would live in scratch memory)
    uint64_t xsW0 = 0, xsW1 = 0, xsL0 = 0, xsL1 = 0, xr0 = 0, xr1 = 0;
    if (M.synth && valid) {
      for (int c = 0; c < 2; c++) {
        int o = c ^ 1, k = sq_get(p, c);
        uint64_t ill = c == WHITE ? WI : BI;
        uint64_t own = 0, xrc = 0, xsw = 0, xsl = 0;
        for (int i = 0; i < MAXP - 1; i++) {
          if (i >= M.nk[c]) break;
          int j = M.pcs[c][i];
          if (j != jx) own |= bit(sq_get(p, j));
        }
        for (int e = 0; e < 8; e++) {
          if (e >= M.nexit[c]) break;
          int to = M.exit_to[c][e], m = M.exit_m[c][e], s2 = M.exit_s2[c][e];
          bool fl = M.exit_fl[c][e] != 0;
          if (own & bit(to)) continue;  // every t: an own piece stands there
          uint64_t go = ~attacked_mask(T, M, p, occF ^ bit(k), to, o, jx, cx, ptx) & ~ill & F;
          uint64_t xw = 0, xl = 0;  // this exit's W and L candidates
          if (cx == c) go &= ~bit(to);
          const uint64_t legal = go;
          uint64_t q0 = sq_set(p, c, to);
          uint64_t q = q0 ^ (0x0101010101010101ull * (uint64_t)m);
          if (fl) q = flip_diag_bytes(q);
          uint64_t ob, base, key;
          if (occF & bit(to)) {  // a capture of one of o's other pieces
            int jj = slot_of(p, to, M.col80[o]);
            int l = sq_get(M.set_of, jj);
            base = partial_idx(M.capt[l], q, L, l, &ob);
            key = synth_key(M.seedmix, o, 2, s2 * 8 + l);
          } else {
            base = partial_idx(M.ii, q, L, -1, &ob);
            key = synth_key(M.seedmix, o, 1, s2);
            if (cx == o && (go & bit(to))) {
              // the last piece (o's) taken on `to` itself: its set's
              // capture space, one index
              uint64_t ob2;
              uint64_t ci = partial_idx(M.capt[L], q, L, -1, &ob2);
              int v = synth_value_k(synth_key(M.seedmix, o, 2, s2 * 8 + L), ci);
              xw |= (uint64_t)(v == -2) << to;
              xl |= (uint64_t)(v == 2) << to;
              go &= ~bit(to);
            }
          }
          // the hash's high half decides each value; where it ties a
          // threshold's (about once in 2^31) the whole exit is evaluated
          // again with full hashes, off the common path.
          // The loop runs over the squares u of the moved position's frame
          // (the images of go's squares), whose rank among the free squares
          // gives the index, in 32-square halves (32-bit masks, the lowest
          // bit extracted in one step); the results are mapped back to t
          // once per exit.
          uint64_t gu = sym_bb(go, m, fl);
          uint32_t ob0 = (uint32_t)ob, ob1 = (uint32_t)(ob >> 32);
          uint32_t hwh[2] = { 0, 0 }, hlh[2] = { 0, 0 };
          uint32_t tmin = ~0u;  // 0 once a high half ties a threshold's
#pragma unroll
          for (int h = 0; h < 2; h++) {
            uint32_t w32 = 0, l32 = 0;
            uint32_t obh = h ? ob1 : ob0;
            // rank of the half's first square: 32 less the pieces below it
            uint32_t rb = h ? 32u - (uint32_t)__builtin_popcount(ob0) : 0u;
            for (uint32_t g2 = (uint32_t)(gu >> (32 * h)); g2; g2 &= g2 - 1) {
              uint32_t bt = g2 & (0u - g2);
              uint32_t r = rb + (uint32_t)__builtin_ctz(g2) - (uint32_t)__builtin_popcount(obh & (bt - 1));
              uint64_t z;
              uint32_t hh = synth_hi(key, base + r, &z);
              l32 |= hh < SYN_TW_HI ? bt : 0u;  // value 2
              uint32_t dw = hh - SYN_TW_HI;
              w32 |= dw < SYN_TL_HI - SYN_TW_HI ? bt : 0u;  // value -2
              tmin = min(tmin, min(dw, hh ^ SYN_TL_HI));
            }
            hwh[h] = w32;
            hlh[h] = l32;
          }
          bool tie = tmin == 0;
          uint64_t hw = sym_bb_inv(hwh[0] | (uint64_t)hwh[1] << 32, m, fl);
          uint64_t hl = sym_bb_inv(hlh[0] | (uint64_t)hlh[1] << 32, m, fl);
          if (__ballot(tie)) {
            hw = hl = 0;
            for (uint64_t g2 = go; g2; g2 &= g2 - 1) {
              int t = ctz64(g2);
              int u = sym_sq(t, m, fl);
              uint64_t idx = base + (uint64_t)(uint32_t)(u - popc64(ob & (bit(u) - 1)));
              int v = synth_value_k(key, idx);
              hw |= (uint64_t)(v == -2) << t;
              hl |= (uint64_t)(v == 2) << t;
            }
          }
          xw |= hw;
          xl |= hl;
          xsw |= xw;
          xsl |= xl;
          xrc |= legal & ~xl;  // a legal exit not into a win of the opponent
        }
        if (c == WHITE) {
          xsW0 = xsw;
          xsL0 = xsl;
          xr0 = xrc;
        } else {
          xsW1 = xsw;
          xsL1 = xsl;
          xr1 = xrc;
        }
      }
    }
    // rank-domain masks (bit r = position P R + r): the squares' bits
    // compressed through the free squares F (Hacker's Delight 7-4, the six
    // move masks of F computed once)
    uint64_t mv[6];
    {
      uint64_t mm = F, mk = ~F << 1;
#pragma unroll
      for (int i = 0; i < 6; i++) {
        uint64_t mp = mk ^ (mk << 1);
        mp ^= mp << 2;
        mp ^= mp << 4;
        mp ^= mp << 8;
        mp ^= mp << 16;
        mp ^= mp << 32;
        uint64_t v = mp & mm;
        mm = (mm ^ v) | (v >> (1 << i));
        mk &= ~mp;
        mv[i] = v;
      }
    }
    auto compress = [&](uint64_t x) {
      x &= F;
#pragma unroll
      for (int i = 0; i < 6; i++) {
        uint64_t t = x & mv[i];
        x = (x ^ t) | (t >> (1 << i));
      }
      return x;
    };
    uint64_t MW = qm[lane], MB = qm[64 + lane];
    wave_sync();  // the buffer is cleared next
    uint64_t rw = 0, rb = 0, rmw = 0, rmb = 0;
    uint64_t rx[4] = { 0, 0, 0, 0 }, rr[2] = { 0, 0 };
    if (valid) {
      rw = compress(WI);
      rb = compress(BI);
      rmw = compress(MW);
      rmb = compress(MB);
      if (M.synth) {
        rx[0] = compress(xsW0);
        rx[1] = compress(xsW1);
        rx[2] = compress(xsL0);
        rx[3] = compress(xsL1);
        // the mates join the refutations: lost already, they are never an
        // L_n candidate (a push-form step never makes them one, having no
        // legal move into its frontier), and the pull form then leaves
        // them out at every ply
        rr[0] = compress(xr0 | MW);
        rr[1] = compress(xr1 | MB);
      }
    }
    // assemble the chunk's words in LDS (a lane's R bits start at bit
    // lane * R) and store each word of each plane once
    for (int i = lane; i < IB_PLANES * IB_WORDS; i += 64) B[i / IB_WORDS][i % IB_WORDS] = 0;
    wave_sync();
    uint64_t rm[IB_PLANES] = { rw, rb, rmw, rmb, rx[0], rx[1], rx[2], rx[3], rr[0], rr[1] };
    int off = lane * R, w0 = off >> 6, sh = off & 63;
#pragma unroll
    for (int k = 0; k < IB_PLANES; k++) {
      uint64_t v = rm[k];
      if (!v) continue;
      atomicOr(reinterpret_cast<unsigned long long *>(&B[k][w0]), (unsigned long long)(v << sh));
      if (sh && sh + R > 64)
        atomicOr(reinterpret_cast<unsigned long long *>(&B[k][w0 + 1]),
                 (unsigned long long)(v >> (64 - sh)));
    }
    wave_sync();
    uint64_t wbase = phys(M, s) * M.words + g * (uint64_t)R;
    uint64_t nwc = M.words - g * (uint64_t)R < (uint64_t)R ? M.words - g * (uint64_t)R : (uint64_t)R;
    if ((uint64_t)lane < nwc) {
      winsW[wbase + lane] = B[0][lane];
      winsB[wbase + lane] = B[1][lane];
      mateW[wbase + lane] = B[2][lane];
      mateB[wbase + lane] = B[3][lane];
      if (M.synth) {
        kexW0[wbase + lane] = B[4][lane];
        kexW1[wbase + lane] = B[5][lane];
        kexL0[wbase + lane] = B[6][lane];
        kexL1[wbase + lane] = B[7][lane];
      }
      if (xrW) {
        xrW[wbase + lane] = B[8][lane];
        xrB[wbase + lane] = B[9][lane];
      }
    }
    // codes, 64 consecutive positions per store
    if (codesW) {
      uint64_t cbase = phys(M, s) * M.ii.size + g * 64 * (uint64_t)R;
      for (int k = 0; k < R; k++) {
        uint64_t i = cbase + (uint64_t)k * 64 + lane;
        if (g * 64 * (uint64_t)R + (uint64_t)k * 64 + lane >= M.ii.size) break;
        uint64_t a = B[0][k], b = B[1][k], ma = B[2][k], mb = B[3][k];
        codesW[i] = ((a >> lane) & 1) ? 3 : ((ma >> lane) & 1) ? 2 : 0;
        codesB[i] = ((b >> lane) & 1) ? 3 : ((mb >> lane) & 1) ? 2 : 0;
      }
    }
    wave_sync();  // the next chunk clears B
    cw[0] += popc64(rw);
    cw[1] += popc64(rb);
    cw[2] += popc64(rmw);
    cw[3] += popc64(rmb);
  }
  for (int k = 0; k < 4; k++) {
    uint64_t v = __ballot(cw[k] != 0) ? wave_sum64(cw[k]) : 0;
    if (lane == 0 && v) atomicAdd(cnt + k, (unsigned long long)v);
  }
}

// The capture part of the static refutations for any material (k_xr_pl's
// counterpart; k_init wrote the exit part): one position per lane, as in
// k_init, both sides; a capture by a non-king piece whose result's bit in
// sub[opp][TBG_SUB_WIN] is clear refutes (the captures of verify_pos,
// src/tb8gen.c:563-576).  nopen[c]: positions neither illegal for c to move
// nor refuted.
__global__ __launch_bounds__(256) void k_xr(const Mat *__restrict__ Mp,
                                            const uint64_t *__restrict__ subwinW,
                                            const uint64_t *__restrict__ subwinB,
                                            const uint64_t *__restrict__ winsW,
                                            const uint64_t *__restrict__ winsB, uint64_t *xrW,
                                            uint64_t *xrB, int s0, int ns,
                                            unsigned long long *nopen)
{
  __shared__ Tabs T;
  load_tabs(T);
  const Mat &M = *Mp;
  int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  uint64_t nw = (uint64_t)ns * M.words;
  uint64_t open[2] = { 0, 0 };
  bool stepping = M.all_single;
  int nsets = M.ii.numsets;
  for (uint64_t g0 = ((uint64_t)blockIdx.x * 4 + wave) * INIT_CHUNK; g0 < nw;
       g0 += (uint64_t)gridDim.x * 4 * INIT_CHUNK)
  for (uint64_t g = g0, D = 0, pk = 0; g < g0 + INIT_CHUNK && g < nw; g++) {
    int s = s0 + (int)(g / M.words);
    uint64_t w = g % M.words;
    uint64_t idx = w * 64 + lane;
    uint64_t a = phys(M, s) * M.words + w;
    uint64_t win[2] = { winsW[a], winsB[a] };
    bool ref[2] = { false, false };
    if (idx < M.ii.size) {
      uint64_t p, occ;
      if (stepping && g != g0 && w != 0) {
        p = pk;
        occ = step64(nsets, &D, &p);
      } else {
        p = kings_of(T, s);
        occ = idx_to_sq(T, M.ii, idx, &p);
        if (stepping) D = digits_of(p, nsets);
      }
      pk = p;
      for (int c = 0; c < 2; c++) {
        if ((win[c] >> lane) & 1) continue;  // illegal with c to move
        int o = c ^ 1;
        const uint64_t *subS = (o == WHITE ? subwinW : subwinB) + phys(M, s) * M.ctotal[o];
        uint64_t oppocc = 0;
        for (int t = 0; t < MAXP - 1; t++) {
          if (t >= M.nk[o]) break;
          oppocc |= bit(sq_get(p, M.pcs[o][t]));
        }
        bool r = false;
        for (int t = 0; t < MAXP - 1 && !r; t++) {
          if (t >= M.nk[c]) break;
          int j = M.pcs[c][t];
          uint64_t cap = piece_att(T, M.pt[j], sq_get(p, j), occ) & oppocc;
          while (cap && !r) {
            int to = ctz64(cap);
            cap &= cap - 1;
            int jj = slot_of(p, to, M.col80[o]);
            int l = sq_get(M.set_of, jj);
            uint64_t q = sq_set(p, j, to);
            q = sq_set(q, jj, sq_get(q, M.ii.last[l]));
            uint64_t ci = M.all_single ? capt_idx_1(M.capt[l], q, l) : sq_to_idx(T, M.capt[l], q);
            r = !bm_test(subS + M.coff[l], ci);
          }
        }
        ref[c] = r;
      }
    }
    uint64_t r0 = __ballot(ref[0]), r1 = __ballot(ref[1]);
    if (lane == 0) {
      uint64_t x0 = xrW[a] | r0, x1 = xrB[a] | r1;
      xrW[a] = x0;
      xrB[a] = x1;
      uint64_t b0 = w * 64;
      uint64_t vm = b0 >= M.ii.size ? 0 : M.ii.size - b0 >= 64 ? ~0ull : bit(M.ii.size - b0) - 1;
      open[0] += popc64(vm & ~x0 & ~win[0]);
      open[1] += popc64(vm & ~x1 & ~win[1]);
    }
  }
  if (lane == 0) {
    if (open[0]) atomicAdd(nopen + 0, (unsigned long long)open[0]);
    if (open[1]) atomicAdd(nopen + 1, (unsigned long long)open[1]);
  }
}

// The six move masks of Hacker's Delight 7-4 for compressing square masks
// through F (bit r of the result = the r-th square of F)
__device__ __forceinline__ void rank_masks(uint64_t F, uint64_t mv[6])
{
  uint64_t mm = F, mk = ~F << 1;
#pragma unroll
  for (int i = 0; i < 6; i++) {
    uint64_t mp = mk ^ (mk << 1);
    mp ^= mp << 2;
    mp ^= mp << 4;
    mp ^= mp << 8;
    mp ^= mp << 16;
    mp ^= mp << 32;
    uint64_t v = mp & mm;
    mm = (mm ^ v) | (v >> (1 << i));
    mk &= ~mp;
    mv[i] = v;
  }
}

__device__ __forceinline__ uint64_t compress_f(uint64_t x, uint64_t F, const uint64_t mv[6])
{
  x &= F;
#pragma unroll
  for (int i = 0; i < 6; i++) {
    uint64_t t = x & mv[i];
    x = (x ^ t) | (t >> (1 << i));
  }
  return x;
}

// nb <= 62 bits of the bitmap p from bit pos on
__device__ __forceinline__ uint64_t load_bits(const uint64_t *p, uint64_t pos, int nb)
{
  uint64_t w = pos >> 6;
  int sh = (int)(pos & 63);
  uint64_t v = gload(p, w) >> sh;
  if (sh + nb > 64) v |= gload(p, w + 1) << (64 - sh);
  return v & (bit(nb) - 1);
}

// The capture part of check_moves (src/tb8gen.c:563-576): a position, stm
// to move, with a capture by a non-king piece whose result is not in the
// capture-exit plane sub[opp][TBG_SUB_WIN] is not lost (at plies up to
// 100); the exits are fixed for the whole generation, so this is worked out
// once per position and OR-ed into xr, which k_init_pl wrote (zero, or the
// synthetic king exits' refutations).  Placement-parallel as k_init_pl: a
// lane holds every set's square but the last set's, whose R squares t are
// bit-parallel:
//  - a piece of stm (not the last set's) capturing one of the other sets'
//    pieces at d: the results for all t form one run of R + 1 bits of the
//    capture space (the last set's square ranked among F plus the vacated
//    square), valid where t does not block the move;
//  - the same piece capturing the last set's piece (t = d): one bit each;
//  - the last set's piece (stm's) capturing a piece at d: one result for
//    every t that reaches d.
// King captures are the verification's (real tables) or the exits' (k_init_pl).
// nopen[c] counts the positions, c to move, neither in wins (illegal) nor
// refuted.
// With sublossW/B (a synthetic job, whose king captures are exits), the same
// captures' results read from sub[opp][TBG_SUB_LOSS] also give the ply-1 W
// step's capture candidates (predecessors_sub from the loss plane,
// src/tb8gen.c:1046-1048, worked out forwards), OR-ed into candWW / candWB
// in place of that step's uncapture.
__global__ __launch_bounds__(256, 4) void k_xr_pl(const Mat *__restrict__ Mp,
                                                  const uint64_t *__restrict__ subwinW,
                                                  const uint64_t *__restrict__ subwinB,
                                                  const uint64_t *__restrict__ sublossW,
                                                  const uint64_t *__restrict__ sublossB,
                                                  const uint64_t *__restrict__ winsW,
                                                  const uint64_t *__restrict__ winsB,
                                                  uint64_t *xrW, uint64_t *xrB, uint64_t *candWW,
                                                  uint64_t *candWB, int s0, int ns,
                                                  unsigned long long *nopen)
{
  __shared__ Tabs T;
  __shared__ Mat Ms;
  __shared__ uint64_t buf[4][4][IB_WORDS];
  load_tabs(T);
  {
    const uint32_t *a = reinterpret_cast<const uint32_t *>(Mp);
    uint32_t *b = reinterpret_cast<uint32_t *>(&Ms);
    for (int i = threadIdx.x; i < (int)(sizeof(Mat) / 4); i += blockDim.x) b[i] = a[i];
    __syncthreads();
  }
  const Mat &M = Ms;
  int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  uint64_t(*B)[IB_WORDS] = buf[wave];
  const int L = M.ii.numsets - 1, jx = 2 + L, R = 62 - L;
  const int cx = (M.pt[jx] >> 3) & 1, ptx = M.pt[jx];
  const uint64_t fullR = bit(R) - 1;
  const uint64_t np = M.ii.size / (uint64_t)R;
  const uint64_t cps = (M.words + R - 1) / R;
  uint64_t open[2] = { 0, 0 };
  for (uint64_t gc = (uint64_t)blockIdx.x * 4 + wave; gc < (uint64_t)ns * cps;
       gc += (uint64_t)gridDim.x * 4) {
    int s = s0 + (int)(gc / cps);
    uint64_t g = gc % cps;
    uint64_t P = g * 64 + lane;
    bool valid = P < np;
    uint64_t p = kings_of(T, s);
    uint64_t occF = bit(sq_get(p, 0)) | bit(sq_get(p, 1));
    if (valid) occF = idx_to_sq_1(M.ii, P * (uint64_t)R, &p) & ~bit(sq_get(p, jx));
    const uint64_t F = ~occF;
    uint64_t mv[6];
    rank_masks(F, mv);
    // per side (named registers, not side-indexed arrays, which would live
    // in scratch memory): refuted; W1 capture candidates
    uint64_t rr0 = 0, rr1 = 0, rw0 = 0, rw1 = 0;
    const bool w1 = sublossW != nullptr;  // uniform
    for (int c = 0; c < 2 && valid; c++) {
      int o = c ^ 1;
      const uint64_t *subS = (o == WHITE ? subwinW : subwinB) + phys(M, s) * M.ctotal[o];
      const uint64_t *subL = w1 ? (o == WHITE ? sublossW : sublossB) + phys(M, s) * M.ctotal[o]
                                : nullptr;
      uint64_t opp = 0;
      for (int i = 0; i < MAXP - 1; i++) {
        if (i >= M.nk[o]) break;
        int j = M.pcs[o][i];
        if (j != jx) opp |= bit(sq_get(p, j));
      }
      uint64_t rk = 0, rsq = 0;  // refuted: rank domain, square domain
      uint64_t wk = 0, wsq = 0;  // a capture into a loss of o: the same
      for (int i = 0; i < MAXP - 1; i++) {
        if (i >= M.nk[c]) break;
        int j = M.pcs[c][i];
        if (j == jx) continue;
        int sj = sq_get(p, j);
        uint64_t a0 = piece_att(T, M.pt[j], sj, occF);
        uint64_t cap = a0 & opp;
        // t above the vacated square sj ranks one higher in the capture space
        uint64_t lowm = bit(popc64(F & (bit(sj) - 1))) - 1;
        while (cap) {
          int d = ctz64(cap);
          cap &= cap - 1;
          int l = sq_get(M.set_of, slot_of(p, d, M.col80[o]));
          uint64_t ob;
          uint64_t base = partial_idx(M.capt[l], sq_set(p, j, d), L, l, &ob);
          uint64_t seg = load_bits(subS + M.coff[l], base, R + 1);
          uint64_t segF = (seg & lowm) | ((seg >> 1) & ~lowm);
          uint64_t btw = between_sq(sj, d) & F;
          uint64_t ok = btw ? fullR & ~compress_f(btw, F, mv) : fullR;
          rk |= ok & ~segF;
          if (w1) {
            uint64_t sl = load_bits(subL + M.coff[l], base, R + 1);
            wk |= ok & ((sl & lowm) | ((sl >> 1) & ~lowm));
          }
        }
        if (cx == o) {  // the last set's piece taken on d, t = d
          uint64_t dd = a0 & F;
          // the capture-space index with j still on sj, then each d's from it:
          // j's rank change and the later sets' pieces it passes (no piece of
          // the last set in this space)
          uint64_t ob, c0 = 0, occb = 0;
          uint32_t ra = 0;
          Later LS = {};
          if (dd) {
            int kj = j - 2;  // j's set (all single)
            c0 = partial_idx(M.capt[L], p, L, -1, &ob);
            LS = later_sets(M.capt[L], p, kj);
            LS.n = L - 1 - kj;
            occb = occ_before(p, kj);
            ra = (uint32_t)(sj - popc64(occb & (bit(sj) - 1)));
          }
          while (dd) {
            int d = ctz64(dd);
            dd &= dd - 1;
            uint64_t ci = moved_idx_l(LS, c0, sj, d, occb, ra);
            if (!bm_test(subS + M.coff[L], ci)) rsq |= bit(d);
            if (w1 && bm_test(subL + M.coff[L], ci)) wsq |= bit(d);
          }
        }
      }
      if (cx == c) {  // the last set's piece takes the piece on d
        uint64_t oo = opp;
        while (oo) {
          int d = ctz64(oo);
          oo &= oo - 1;
          uint64_t from = piece_att(T, ptx, d, occF) & F;
          if (!from) continue;
          int l = sq_get(M.set_of, slot_of(p, d, M.col80[o]));
          uint64_t ob;
          uint64_t ci = partial_idx(M.capt[l], sq_set(p, jx, d), L, l, &ob);
          ci += (uint64_t)(uint32_t)(d - popc64(ob & (bit(d) - 1)));
          if (!bm_test(subS + M.coff[l], ci)) rsq |= from;
          if (w1 && bm_test(subL + M.coff[l], ci)) wsq |= from;
        }
      }
      uint64_t rrc = rk | compress_f(rsq, F, mv);
      uint64_t rwc = w1 ? wk | compress_f(wsq, F, mv) : 0;
      if (c == WHITE) {
        rr0 = rrc;
        rw0 = rwc;
      } else {
        rr1 = rrc;
        rw1 = rwc;
      }
    }
    // the chunk's R words of each plane: read, OR-ed, written back once
    uint64_t wbase = phys(M, s) * M.words + g * (uint64_t)R;
    uint64_t nwc = M.words - g * (uint64_t)R < (uint64_t)R ? M.words - g * (uint64_t)R : (uint64_t)R;
    B[0][lane] = (uint64_t)lane < nwc ? xrW[wbase + lane] : 0;
    B[1][lane] = (uint64_t)lane < nwc ? xrB[wbase + lane] : 0;
    if (w1) {
      B[2][lane] = (uint64_t)lane < nwc ? candWW[wbase + lane] : 0;
      B[3][lane] = (uint64_t)lane < nwc ? candWB[wbase + lane] : 0;
    }
    wave_sync();
    int off = lane * R, w0 = off >> 6, sh = off & 63;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      uint64_t v = k == 0 ? rr0 : k == 1 ? rr1 : k == 2 ? rw0 : rw1;
      if (!v) continue;
      atomicOr(reinterpret_cast<unsigned long long *>(&B[k][w0]), (unsigned long long)(v << sh));
      if (sh && sh + R > 64)
        atomicOr(reinterpret_cast<unsigned long long *>(&B[k][w0 + 1]),
                 (unsigned long long)(v >> (64 - sh)));
    }
    wave_sync();
    if ((uint64_t)lane < nwc) {
      xrW[wbase + lane] = B[0][lane];
      xrB[wbase + lane] = B[1][lane];
      if (w1) {
        candWW[wbase + lane] = B[2][lane];
        candWB[wbase + lane] = B[3][lane];
      }
      uint64_t b0 = (g * (uint64_t)R + lane) * 64;  // the word's first index in the slice
      uint64_t vm = b0 >= M.ii.size ? 0 : M.ii.size - b0 >= 64 ? ~0ull : bit(M.ii.size - b0) - 1;
      open[0] += popc64(vm & ~B[0][lane] & ~winsW[wbase + lane]);
      open[1] += popc64(vm & ~B[1][lane] & ~winsB[wbase + lane]);
    }
    wave_sync();  // the next chunk overwrites B
  }
  for (int k = 0; k < 2; k++) {
    uint64_t v = __ballot(open[k] != 0) ? wave_sum64(open[k]) : 0;
    if (lane == 0 && v) atomicAdd(nopen + k, (unsigned long long)v);
  }
}

// calc_sub_worker (src/tb8gen.c:280-314) for set k of colour c: classify
// every position with one piece of set k removed, c to move, by the
// smaller material's finished table (probe_wdl, src/probe.c:1677-1684).
__global__ __launch_bounds__(256) void k_resolve(const Mat *__restrict__ Mp,
                                                  const SubRef *__restrict__ Rp, int c, int k,
                                                  uint64_t *loss, uint64_t *bloss, uint64_t *cwin,
                                                  uint64_t *win, uint64_t *lwin, int s0, int ns)
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
          v = synth_value(synth_hash(M.seedmix, c, 0, k, idx));
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
    uint64_t b4 = __ballot(!ill && v == 2);
    if (lane == 0) {
      uint64_t a = phys(M, s) * M.ctotal[c] + M.coff[k] + w;
      loss[a] = b0;
      bloss[a] = b1;
      cwin[a] = b2;
      win[a] = b3;
      lwin[a] = b4;
    }
  }
}

// k_resolve for a material whose sets are all single, one capture-space
// word per lane: its 64 positions are consecutive indices, so the first is
// decoded in full and each next one is the last set's piece moved to the
// next free square (the last digit plus one); when there is none the digit
// carries and the index is decoded in full again, about once per 60
// positions.  The same classification per position as k_resolve.
__global__ __launch_bounds__(256) void k_resolve_w(const Mat *__restrict__ Mp,
                                                    const SubRef *__restrict__ Rp, int c, int k,
                                                    uint64_t *loss, uint64_t *bloss,
                                                    uint64_t *cwin, uint64_t *win,
                                                    uint64_t *lwin, int s0, int ns)
{
  __shared__ Tabs T;
  load_tabs(T);
  const Mat &M = *Mp;
  const SubRef &R = *Rp;
  const Idx &ci = M.capt[k];
  int m = M.ii.last[k];
  // the last set with a piece in this capture space, and that piece's slot
  int jl = 0;
  for (int j = 0; j < MAXS; j++) {
    if (j >= ci.numsets) break;
    if (ci.mult[j] > 0) jl = ci.last[j];
  }
  uint64_t cw = M.cwords[k];
  uint64_t nw = (uint64_t)ns * cw;
  uint64_t key = synth_key(M.seedmix, c, 0, k);
  for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < nw;
       g += (uint64_t)gridDim.x * blockDim.x) {
    int s = s0 + (int)(g / cw);
    uint64_t w = g % cw;
    uint64_t b0 = 0, b1 = 0, b2 = 0, b3 = 0, b4 = 0;
    uint64_t p = 0, occ = 0;
    bool have = false;
    for (int i = 0; i < 64; i++) {
      uint64_t idx = w * 64 + (uint64_t)i;
      if (idx >= ci.size) break;
      int t = sq_get(p, jl);
      uint64_t f = have ? ~occ & (~1ull << t) : 0;  // free squares above the last piece's
      if (f) {
        int t2 = ctz64(f);
        p = sq_set(p, jl, t2);
        occ = (occ ^ bit(t)) | bit(t2);
      } else {
        p = kings_of(T, s);
        occ = idx_to_sq(T, ci, idx, &p);
        have = true;
      }
      int v = 0;
      bool ill = attacked_by(T, M, p, occ, sq_get(p, c ^ 1), c, m);
      if (!ill) {
        if (M.synth) {  // synth_hash(M.seedmix, c, 0, k, idx), decided from its high half
          v = synth_value_k(key, idx);
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
      uint64_t b = bit(i);
      b0 |= !ill && v == -2 ? b : 0;
      b1 |= !ill && v == -1 ? b : 0;
      b2 |= ill || v >= 1 ? b : 0;
      b3 |= ill || v == 2 ? b : 0;
      b4 |= !ill && v == 2 ? b : 0;
    }
    uint64_t a = phys(M, s) * M.ctotal[c] + M.coff[k] + w;
    loss[a] = b0;
    bloss[a] = b1;
    cwin[a] = b2;
    win[a] = b3;
    lwin[a] = b4;
  }
}

// ------------------------------------------------------- Syzygy WDL files
// r-th free square (zero bit of occ) among the squares of `among`
__device__ __forceinline__ int nth_free(uint64_t occ, uint64_t among, int r)
{
  uint64_t f = ~occ & among;
  for (int i = 0; i < r; i++) f &= f - 1;
  return ctz64(f);
}

// Squares of the leading group for its digit q: decode_helper's PIECE_ENC
// branch (src/probe.c:884-946) for three unique pieces (the 31332
// placements: first piece in the a1-d1-d4 triangle, off the diagonal, or
// the later pieces' diagonal cases) or the two kings (slice numbering).
__device__ uint64_t syz_lead(const Tabs &T, const SyzLayout &L, uint32_t q, uint64_t *occ)
{
  const uint64_t DIAG = 0x8040201008040201ull;
  int p0, p1, p2 = 0;
  if (L.lead == 2) {
    p0 = T.kksq[q][0];
    p1 = T.kksq[q][1];
    *occ = bit(p0) | bit(p1);
    return (uint64_t)p0 | ((uint64_t)p1 << 8);
  }
  const uint32_t A = 6 * 63 * 62, B = 4 * 28 * 62, C = 4 * 7 * 28;
  if (q < A) {  // first piece off the diagonal
    uint32_t r = q / (63 * 62);
    q -= r * 63 * 62;
    p0 = L.inv_tri[r];
    r = q / 62;
    q -= r * 62;
    p1 = (int)r + ((int)r >= p0);
    p2 = nth_free(bit(p0) | bit(p1), ~0ull, (int)q);
  } else if (q < A + B) {  // first on the diagonal, second below it
    q -= A;
    uint32_t r = q / (28 * 62);
    q -= r * 28 * 62;
    p0 = L.inv_diag[r];
    r = q / 62;
    q -= r * 62;
    p1 = L.inv_lower[r];
    p2 = nth_free(bit(p0) | bit(p1), ~0ull, (int)q);
  } else if (q < A + B + C) {  // first two on the diagonal, third below it
    q -= A + B;
    uint32_t r = q / (7 * 28);
    q -= r * 7 * 28;
    p0 = L.inv_diag[r];
    r = q / 28;
    q -= r * 28;
    r += L.inv_diag[r] >= p0;
    p1 = L.inv_diag[r];
    p2 = L.inv_lower[q];
  } else {  // all three on the diagonal
    q -= A + B + C;
    uint32_t r = q / 42;
    q -= r * 42;
    p0 = L.inv_diag[r];
    r = q / 6;
    q -= r * 6;
    r += L.inv_diag[r] >= p0;
    p1 = L.inv_diag[r];
    p2 = nth_free(bit(p0) | bit(p1), DIAG, (int)q);
  }
  *occ = bit(p0) | bit(p1) | bit(p2);
  return (uint64_t)p0 | ((uint64_t)p1 << 8) | ((uint64_t)p2 << 16);
}

__global__ __launch_bounds__(256) void k_syz_gather(const Mat *__restrict__ Mp,
                                                     const uint16_t *__restrict__ codes,
                                                     SyzLayout L, uint64_t i0, uint64_t n,
                                                     uint8_t *out)
{
  __shared__ Tabs T;
  load_tabs(T);
  const Mat &M = *Mp;
  for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t I = i0 + k;
    uint64_t lead = I / L.lead_factor, rem = I - lead * L.lead_factor;
    uint64_t occ;
    uint64_t h = syz_lead(T, L, (uint32_t)lead, &occ);  // header squares, packed
    for (int g = 0; g < MAXP; g++) {
      if (g >= L.ngroups) break;
      uint32_t d = (uint32_t)((rem / L.gfactor[g]) % L.gsize[g]);
      int f = L.gfirst[g], m = L.gmult[g];
      uint64_t occ0 = occ;
      // unrank_binomial (src/probe.c:809-830): member i is the largest r
      // with C(r, i + 1) <= d, ranked among the squares free of occ0
      for (int i = MAXS - 1; i > 0; i--) {
        if (i >= m) continue;
        int r = i;
        while (r + 1 < 64 && T.binom[(i + 1) * 64 + r + 1] <= d) r++;
        d -= T.binom[(i + 1) * 64 + r];
        int s = nth_free(occ0, ~0ull, r);
        h = sq_set(h, f + i, s);
        occ |= bit(s);
      }
      int s = nth_free(occ0, ~0ull, (int)d);
      h = sq_set(h, f, s);
      occ |= bit(s);
    }
    uint64_t p = 0;
    for (int i = 0; i < MAXP; i++) {
      if (i >= L.num) break;
      p = sq_set(p, L.slot[i], sq_get(h, i));
    }
    uint64_t q;
    int s = normalize(T, p, &q);
    uint8_t v = 0xff;
    if (s >= 0) {
      uint16_t c = codes[(uint64_t)s * M.ii.size + sq_to_idx(T, M.ii, q)];
      if (c != 3) v = (uint8_t)(code_wdl(c) + 2);
    }
    out[k] = v;
  }
}

}  // namespace tbg

namespace tbg {
hipError_t upload_tabs(const Tabs *h) { return hipMemcpyToSymbol(HIP_SYMBOL(g_tabs), h, sizeof(Tabs)); }
}  // namespace tbg
