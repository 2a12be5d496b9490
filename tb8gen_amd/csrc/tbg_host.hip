// tbg_host.hip -- host side of the C ABI (include/tbg.h): material layout,
// resident device planes, kernel launches for the plain / batched / graph
// forms, and the generator's ply loop.
#include <stdarg.h>
#include <stddef.h>
#include <assert.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <filesystem>
#include <functional>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/tbg.h"
#include "tbg_kern.h"

using namespace tbg;

// Device memory of the entry point that allocates nothing: one copy per GPU,
// held until the process ends.  Every call on a GPU shares it.
static __device__ __align__(256) uint8_t scratch[TBG_NOALLOC_SCRATCH_BYTES];

namespace {

thread_local char g_errbuf[256];
int g_device = -1;
int g_cus = 256;
// bitmap planes per job: wins, prevW, prevL, newW, newL, the L and W steps'
// candidates and the static refutations, per side
constexpr uint64_t kPlanes = 16;
// plies a generation may take (the history's length)
constexpr int kMaxPly = 4096;

// all sets single and at least 64 placements (runs of the last set's R
// squares) per slice: the placement-parallel kernels apply
bool pl_layout(const Mat &M)
{
  uint64_t R = 62 - (uint64_t)(M.ii.numsets - 1);
  return M.all_single && M.ii.numsets >= 2 && M.ii.size / R >= 64;
}

uint64_t planes_of(const Mat &) { return kPlanes; }
// chunk-summary bytes per slice of a plane: one per 64 bitmap words
uint64_t sum_per_slice(const Mat &M) { return (M.words + 63) / 64; }
Tabs g_htabs;

// a synthetic job on the placement layout whose king moves all leave the
// slice: k_xr_pl also works out the ply-1 W step's capture candidates (the
// king's captures are exits, which k_init_pl covers), and that step runs no
// uncapture.  (A king move that the symmetries map back into the slice is
// a move within the table, and its captures are the uncapture's.)
bool w1_in_xr(const Mat &M)
{
  if (!M.synth || !pl_layout(M)) return false;
  const Tabs &T = g_htabs;
  int kq[2] = { T.kksq[M.synth_slice][0], T.kksq[M.synth_slice][1] };
  for (int stm = 0; stm < 2; stm++)
    if (popc64(T.king[kq[stm]] & ~T.king[kq[stm ^ 1]]) != M.nexit[stm]) return false;
  return true;
}
// geometry and scatter given to new contexts and servers (tbg_set_geometry(NULL, ..))
Geom g_geom = { 256, 4096, { 3, 16, 16, 16 } };
int g_scatter = TBG_SCATTER_BLOCK | TBG_SCATTER_COMBINE;
char g_arch[64] = "none";

// Scratch of tbg_generate_batch_noalloc: the batch's job-pointer array, its
// flags, its (job, side) pairs and the four prefix-sum arrays of its work
// units, at fixed offsets (256-byte aligned) inside `scratch` below.
constexpr int kNaMax = TBG_NOALLOC_MAX_BATCH;
constexpr size_t kNaJobs = 0;
constexpr size_t kNaFlags = kNaJobs + ((sizeof(JobDev *) * kNaMax + 255) & ~(size_t)255);
constexpr size_t kNaPairs = kNaFlags + 256;
constexpr size_t kNaPStride = (sizeof(uint64_t) * (2 * kNaMax + 1) + 255) & ~(size_t)255;
constexpr size_t kNaP = kNaPairs + ((sizeof(Pair) * 2 * kNaMax + 255) & ~(size_t)255);
constexpr size_t kNaBytes = kNaP + 4 * kNaPStride;
static_assert(kNaBytes <= TBG_NOALLOC_SCRATCH_BYTES, "no-allocation scratch layout");
// device address of `scratch` on each device (hipGetSymbolAddress at
// tbg_init: the array's name is not a device address in host code)
constexpr int kMaxDevices = 64;
uint8_t *g_scratch[kMaxDevices];

int fail(int code, const char *fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_errbuf, sizeof g_errbuf, fmt, ap);
  va_end(ap);
  return code;
}

#define HIPCHK(x)                                                                \
  do {                                                                           \
    hipError_t e_ = (x);                                                         \
    if (e_ != hipSuccess)                                                        \
      return fail(TBG_E_HIP, "%s: %s (%s:%d)", #x, hipGetErrorString(e_), __FILE__, \
                  __LINE__);                                                     \
  } while (0)

#define TRY(x)                   \
  do {                           \
    int r_ = (x);                \
    if (r_ != TBG_OK) return r_; \
  } while (0)

// ------------------------------------------------------------ tables
const uint8_t kInvTri[10] = { 1, 2, 3, 10, 11, 19, 0, 9, 18, 27 };  // src/probe.c:327-329

bool on_diag(int s) { return (s & 7) == (s >> 3); }
bool touching(int a, int b)
{
  int df = (a & 7) - (b & 7), dr = (a >> 3) - (b >> 3);
  return df >= -1 && df <= 1 && dr >= -1 && dr <= 1;
}

void build_tabs(Tabs &T)
{
  memset(&T, 0, sizeof T);
  for (int sq = 0; sq < 64; sq++) {
    int f = sq & 7, r = sq >> 3;
    for (int df = -2; df <= 2; df++)
      for (int dr = -2; dr <= 2; dr++) {
        int f2 = f + df, r2 = r + dr;
        if (f2 < 0 || f2 > 7 || r2 < 0 || r2 > 7 || (!df && !dr)) continue;
        int ad = df < 0 ? -df : df, ar = dr < 0 ? -dr : dr;
        if (ad <= 1 && ar <= 1) T.king[sq] |= 1ull << (r2 * 8 + f2);
        if ((ad == 1 && ar == 2) || (ad == 2 && ar == 1)) T.knight[sq] |= 1ull << (r2 * 8 + f2);
      }
  }
  for (int n = 0; n < 64; n++) T.binom[n] = 1;
  for (int k = 1; k < 8; k++) {
    T.binom[k * 64] = 0;
    for (int n = 1; n < 64; n++) {
      uint64_t v = (uint64_t)T.binom[(k - 1) * 64 + n - 1] + T.binom[k * 64 + n - 1];
      T.binom[k * 64 + n] = v > 0xffffffffull ? 0xffffffffu : (uint32_t)v;
    }
  }
  // slice numbering: valid (wk, bk) pairs in triangle order, pairs with both
  // kings on a1-h8 numbered last (src/probe.c:441-522)
  int n = 0, nd = 441;
  for (int i = 0; i < 10; i++) {
    int wk = kInvTri[i];
    for (int bk = 0; bk < 64; bk++) {
      int16_t v = -1;
      if (bk != wk && !touching(wk, bk) && !(on_diag(wk) && (bk & 7) < (bk >> 3)))
        v = (int16_t)((on_diag(wk) && on_diag(bk)) ? nd++ : n++);
      T.kkidx[i * 64 + bk] = v;
      if (v >= 0) {
        T.kksq[v][0] = (uint8_t)wk;
        T.kksq[v][1] = (uint8_t)bk;
      }
    }
  }
  for (int i = 0; i < 64; i++) T.tri[i] = 0;
  for (int i = 0; i < 10; i++) T.tri[kInvTri[i]] = (int8_t)i;
}

int kkmap(int wk, int bk)
{
  uint64_t q;
  uint64_t p = (uint64_t)wk | ((uint64_t)bk << 8);
  if (wk == bk || touching(wk, bk)) return -1;
  return normalize(g_htabs, p, &q);
}

// manager order (src/kslice.c:41-84): the slices of the black-king walk;
// white's order is its colour flip
void manager_order(int stm, int *out)
{
  int s = 0;
  for (int i = 0; i < 10; i++)
    for (int bk = 0; bk < 64; bk++) {
      int v = g_htabs.kkidx[i * 64 + bk];
      if (v < 0) continue;
      out[s++] = stm == BLACK ? v : kkmap(g_htabs.kksq[v][1], g_htabs.kksq[v][0]);
    }
}

// ------------------------------------------------------------ material
const char kPieceChar[] = " PNBRQK";
const int kPieceOrder[16] = { 0, 0, 3, 5, 7, 9, 1, 0, 0, 0, 4, 6, 8, 10, 2, 0 };  // tb8gen.c:1331

void calc_idx(Idx &ii)
{
  int n = 62;
  uint64_t sub[MAXS];
  for (int i = 0; i < ii.numsets; i++) {
    sub[i] = g_htabs.binom[ii.mult[i] * 64 + n];
    n -= ii.mult[i];
  }
  uint64_t f = 1;
  ii.small_q = 1;
  for (int i = 0; i < ii.numsets; i++) ii.small_q &= sub[i] < (1u << 21);
  for (int i = MAXS - 1; i >= 0; i--) {
    if (i >= ii.numsets) { ii.factor[i] = 0; ii.inv[i] = 0; ii.invf[i] = 0; continue; }
    assert(f < (1ull << 32));  // the sets after the first hold at most 5 of 8 pieces
    ii.factor[i] = (uint32_t)f;
    ii.inv[i] = 1.0 / (double)f;
    ii.invf[i] = 1.0f / (float)f;
    f *= sub[i];
  }
  ii.size = f;
}

int setup_mat(Mat &M, const char *name)
{
  memset(&M, 0, sizeof M);
  int pt[16], k = 0, color = 0, cnt[16] = { 0 };
  for (const char *s = name; *s; s++) {
    if (*s == 'v' && !color) { color = 8; continue; }
    int found = 0;
    for (int i = KNIGHT; i <= KING; i++)
      if (*s == kPieceChar[i]) {
        if (k >= MAXP) return fail(TBG_E_INVAL, "too many pieces in %s", name);
        cnt[i | color]++;
        pt[k++] = i | color;
        found = 1;
      }
    if (!found) return fail(TBG_E_INVAL, "bad piece '%c' in %s (pawnless only)", *s, name);
  }
  if (!color) return fail(TBG_E_INVAL, "material %s has no 'v'", name);
  if (k < 3) return fail(TBG_E_INVAL, "need at least 3 pieces");
  if (cnt[KING] != 1 || cnt[KING | 8] != 1) return fail(TBG_E_INVAL, "need one king per side");
  for (int i = 0; i < k; i++)
    for (int j = i + 1; j < k; j++)
      if (kPieceOrder[pt[i]] > kPieceOrder[pt[j]]) { int x = pt[i]; pt[i] = pt[j]; pt[j] = x; }
  M.num = k;
  for (int i = 0; i < MAXP; i++) M.pt[i] = i < k ? pt[i] : 0;
  for (int c = 0; c < 2; c++) {
    M.nk[c] = 0;
    M.col80[c] = 0;
    for (int i = 2; i < k; i++)
      if ((pt[i] >> 3) == c) {
        M.pcs[c][M.nk[c]++] = i;
        M.col80[c] |= 0x80ull << (8 * i);
      }
  }
  M.set_of = ~0ull;
  int ns = 0;
  for (int i = 2; i < k;) {
    int j = i;
    for (; i < k && pt[i] == pt[j]; i++) M.set_of = sq_set(M.set_of, i, ns);
    M.ii.first[ns] = j;
    M.ii.mult[ns] = i - j;
    M.ii.last[ns] = i - 1;
    M.setcol[ns] = pt[j] >> 3;
    M.setmask80[ns] = 0;
    for (int x = j; x < i; x++) M.setmask80[ns] |= 0x80ull << (8 * x);
    ns++;
  }
  M.numsets = M.ii.numsets = ns;
  M.all_single = 1;
  for (int x = 0; x < ns; x++) M.all_single &= M.ii.mult[x] == 1;
  calc_idx(M.ii);
  M.words = ((M.ii.size + 511) / 512) * 8;
  for (int s = 0; s < ns; s++) {
    M.capt[s] = M.ii;
    M.capt[s].mult[s]--;
    calc_idx(M.capt[s]);
    int c = M.setcol[s];
    M.cwords[s] = ((M.capt[s].size + 511) / 512) * 8;
    M.coff[s] = M.ctotal[c];
    M.ctotal[c] += M.cwords[s];
  }
  return TBG_OK;
}

// create_material_string (src/probe.c:1480-1499)
void material_string(const int *pt, int num, bool flip, char *out)
{
  int cnt[16] = { 0 };
  for (int i = 0; i < num; i++) cnt[pt[i] ^ (flip ? 8 : 0)]++;
  int j = 0;
  for (int i = KING; i >= KNIGHT; i--)
    while (cnt[i]--) out[j++] = kPieceChar[i];
  out[j++] = 'v';
  for (int i = KING; i >= KNIGHT; i--)
    while (cnt[i + 8]--) out[j++] = kPieceChar[i];
  out[j] = 0;
}

enum { CAT_PRED, CAT_VERIFY, CAT_SWEEP, CAT_UNCAP, CAT_INIT, CAT_RESOLVE, CAT_XR, CAT_N };

}  // namespace

// Device arrays of one launch configuration: the (job, side) pairs and, per
// kernel type, the prefix sums of their work units.
struct LaunchSet {
  Pair *pairs = nullptr;
  uint64_t *P[4] = { nullptr, nullptr, nullptr, nullptr };
  uint64_t total[4] = { 0, 0, 0, 0 };
  int np = 0;
  size_t cap = 0;  // pairs the device buffers hold
  size_t pred_lds = 0;
  bool pred_wide = false;  // some job scatters high-set unmoves to device memory
  uint64_t uncap_lds = 0;  // words of the largest whole-slice LDS bitmap of the uncapture
  bool verify_split = false;  // some job's slices take two-pass verification
  bool sums = false;          // some job keeps chunk summaries (the kernels' SUM form)
  Geom g;
  void release()
  {
    if (!cap) return;  // nothing allocated (or arrays owned elsewhere)
    (void)hipFree(pairs);
    for (auto &x : P) (void)hipFree(x);
    pairs = nullptr;
    for (auto &x : P) x = nullptr;
    cap = 0;
  }
};

struct tbg_ctx {
  char name[16];
  Mat h;
  Mat *dM = nullptr;
  SubRef hR[MAXS];
  SubRef *dR = nullptr;
  uint64_t nsl = 0, pw = 0;
  // device memory in two blocks: the state block (bitmap planes from its
  // start, `arena`, then the capture-exit planes, `subarena`, and the codes)
  // and the record block `rec` (material, capture tables, init counts,
  // flags, job record, its pointer, per-ply history)
  uint64_t *arena = nullptr, *subarena = nullptr;
  uint8_t *psum = nullptr;  // the planes' chunk summaries
  void *rec = nullptr;
  uint64_t *wins[2], *prevW[2], *prevL[2], *newW[2], *newL[2];  // canonical planes
  uint64_t *cand[2], *candW[2];
  uint64_t *xr[2] = { nullptr, nullptr };  // static refutations (captures, exits)
  bool xr_ok = false;                      // worked out since the last resolve and init
  uint64_t *sub[2][SUB_KINDS];
  uint16_t *codes[2] = { nullptr, nullptr };
  unsigned long long *dcnt = nullptr;  // init counts
  unsigned *dflags = nullptr;          // [0] jobs done, [1] uncapture ply pending
  unsigned *hflags = nullptr;          // pinned host copy of a batch's flags
  JobDev hj;                           // host mirror of the job
  JobDev *dj = nullptr;                // the job on the device
  JobDev **djp = nullptr;              // device array holding dj (a batch of one)
  long long *dhist = nullptr;
  std::vector<long long> hist;
  int max_ply = kMaxPly;
  bool ply_cap_asked = false;  // a bounded run (tune trials) may stop at max_ply
  // a server job: memory from the stream-ordered pool of the server's setup
  // stream (which it shares, so does not own), set up without host waits
  bool pooled = false;
  hipEvent_t ready = nullptr;  // pooled: set-up and init complete
  hipStream_t st = nullptr;
  int resolved = 0, inited = 0, generated = 0;
  uint64_t bytes = 0;        // device bytes held: both blocks
  uint64_t state_bytes = 0;  // the state block (what preemption moves to the host)
  int mgr[2][NSL];
  LaunchSet both, side[2], plain[2];
  bool prof = false;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> evs;
  double ms[CAT_N];
  uint64_t nl[CAT_N];
  uint64_t launches = 0, half_steps = 0;
  int scatter = g_scatter;
  Geom geom = g_geom;
  tbg_counts counts;
  double graph_ms = 0;
  uint64_t hs_generated = 0;  // half-steps executed by tbg_generate*
};

namespace {

// Block of the index a k_pred_j work-group owns: the whole slice if its
// bitmap fits the LDS budget, else the run of positions sharing the high
// digits 0..d-1 (factor[d-1] positions) for the smallest d that fits.
constexpr uint64_t kBlkMaxPos = 40 * 1024 * 8 - 128;

// A move of a piece of set k >= d changes digits k.. only (the digits
// before depend on the kings and the earlier sets), so it stays in its
// block: *blk_set = d.
void choose_block(const Mat &M, uint64_t *blk_pos, uint64_t *nblk, int *blk_set)
{
  if (M.ii.size <= kBlkMaxPos) { *blk_pos = M.ii.size; *nblk = 1; *blk_set = 0; return; }
  for (int d = 1; d <= M.ii.numsets; d++)
    if (M.ii.factor[d - 1] <= kBlkMaxPos) {
      *blk_pos = M.ii.factor[d - 1];
      *nblk = M.ii.size / M.ii.factor[d - 1];
      *blk_set = d;
      return;
    }
  *blk_pos = 1;
  *nblk = M.ii.size;
  *blk_set = M.ii.numsets;
}

// A table of few large slices (one block each) has few pred units per
// launch (462 per side), each long: the work-groups that draw a second unit
// set the launch's length while the rest of the chip idles.  Such a block's
// frontier is split into up to kMaxParts parts of at least one 64-word
// chunk per wave each, every part a unit of its own that scatters into its
// own LDS copy of the whole block.
constexpr int kMaxParts = 4;
int block_parts(uint64_t nblk, uint64_t blk_pos)
{
  if (nblk != 1) return 1;
  uint64_t chunks = (blk_pos + 64 * 64 - 1) / (64 * 64);
  uint64_t parts = chunks / (PB_THREADS / 64);
  return parts < 1 ? 1 : parts > (uint64_t)kMaxParts ? kMaxParts : (int)parts;
}

// Host side of a launch configuration: for np (job, side) pairs, the
// prefix sums of their work units per kernel type into P[kt * (np + 1) ..]
// and the launch parameters that depend on the jobs.  window_slices 0: each
// job's own slices (462, or 1 for a synthetic job).
void plan_launchset(LaunchSet &L, const JobDev *const *hjobs, const Mat *const *mats, int nj,
                    const Pair *pairs, int np, uint64_t window_slices, const Geom &g, uint64_t *P)
{
  L.np = np;
  L.g = g;
  uint64_t maxbw = 1;
  for (int kt = 0; kt < 4; kt++) {
    uint64_t *Pk = P + (size_t)kt * (np + 1);
    Pk[0] = 0;
    for (int i = 0; i < np; i++) {
      const Mat &M = *mats[pairs[i].job];
      uint64_t ns = window_slices ? window_slices : (M.synth ? 1 : NSL);
      const JobDev *J = hjobs[pairs[i].job];
      Pk[i + 1] = Pk[i] + ns * units_per_slice(M, J->nblk * J->blk_parts, kt, pairs[i].side, g);
    }
    L.total[kt] = Pk[np];
  }
  bool block = false;
  L.pred_wide = false;
  for (int i = 0; i < nj; i++) {
    const JobDev *j = hjobs[i];
    maxbw = j->blk_words > maxbw ? j->blk_words : maxbw;
    block |= (j->scatter & TBG_SCATTER_BLOCK) != 0;
    L.pred_wide |= !(j->scatter & TBG_SCATTER_BLOCK) || j->blk_set > 0;
  }
  // the uncapture privatises slices of up to 40 KB (those pred handles as
  // one block)
  L.uncap_lds = 0;
  L.verify_split = false;
  L.sums = false;
  for (int i = 0; i < nj; i++) {
    L.verify_split |= mats[i]->ii.size > VERIFY_SPLIT_SIZE;
    L.sums |= hjobs[i]->sum != nullptr;
    uint64_t nbw = (mats[i]->ii.size + 63) >> 6;
    if ((hjobs[i]->scatter & TBG_SCATTER_BLOCK) && hjobs[i]->nblk == 1 && nbw > L.uncap_lds)
      L.uncap_lds = nbw;
  }
  // the queues follow the block bitmap, which only the block scatter uses
  L.pred_lds = (block ? maxbw * 8 : 0) + PB_QUEUE_BYTES;
}

// Copies the pairs and the prefix sums of a planned launch set to its
// device arrays.  The sources may be pageable: the copies are staged
// before the calls return.
int upload_launchset(const LaunchSet &L, const Pair *pairs, const uint64_t *P, hipStream_t st)
{
  HIPCHK(hipMemcpyAsync(L.pairs, pairs, sizeof(Pair) * L.np, hipMemcpyHostToDevice, st));
  for (int kt = 0; kt < 4; kt++)
    HIPCHK(hipMemcpyAsync(L.P[kt], P + (size_t)kt * (L.np + 1), sizeof(uint64_t) * (L.np + 1),
                          hipMemcpyHostToDevice, st));
  return TBG_OK;
}

int build_launchset(LaunchSet &L, const std::vector<const JobDev *> &hjobs,
                    const std::vector<const Mat *> &mats, const std::vector<Pair> &pairs,
                    uint64_t window_slices, const Geom &g, hipStream_t st)
{
  if (pairs.size() > L.cap) L.release();  // else the device buffers are reused
  std::vector<uint64_t> P(4 * (pairs.size() + 1));
  plan_launchset(L, hjobs.data(), mats.data(), (int)hjobs.size(), pairs.data(), (int)pairs.size(),
                 window_slices, g, P.data());
  if (!L.cap) {
    L.cap = pairs.size();
    HIPCHK(hipMalloc(&L.pairs, sizeof(Pair) * L.cap));
    for (int kt = 0; kt < 4; kt++) HIPCHK(hipMalloc(&L.P[kt], sizeof(uint64_t) * (L.cap + 1)));
  }
  return upload_launchset(L, pairs.data(), P.data(), st);
}

// canonical job state: planes in their allocation order, ply 1, all steps
void reset_job(tbg_ctx *c)
{
  JobDev &J = c->hj;
  memset(&J, 0, sizeof J);
  J.M = c->dM;
  for (int s = 0; s < 2; s++) {
    J.wins[s] = c->wins[s];
    J.prevW[s] = c->prevW[s];
    J.prevL[s] = c->prevL[s];
    J.newW[s] = c->newW[s];
    J.newL[s] = c->newL[s];
    J.cand[s] = c->cand[s];
    J.candW[s] = c->candW[s];
    J.xr[s] = c->xr_ok ? c->xr[s] : nullptr;
    J.nopen = c->xr_ok ? c->dcnt + 4 : nullptr;
    J.codes[s] = c->codes[s];
    for (int k = 0; k < SUB_KINDS; k++) J.sub[s][k] = c->sub[s][k];
  }
  J.hist = c->dhist;
  J.sum = c->psum;
  J.arena = c->arena;
  J.pw = c->pw;
  J.spw = sum_per_slice(c->h);
  J.sum_pl = c->nsl * J.spw;
  J.sparse = 0;
  J.sum_plies = 0;  // summaries unknown until two plies have written them
  J.w1cap = c->xr_ok && w1_in_xr(c->h);
  J.n = 1;
  J.mask = 0xf;
  J.more = 0xf;
  J.scatter = c->scatter;
  J.max_ply = c->max_ply;
  choose_block(c->h, &J.blk_pos, &J.nblk, &J.blk_set);
  J.blk_words = (J.blk_pos + 63) / 64 + 1;
  J.blk_parts = block_parts(J.nblk, J.blk_pos);
}

int upload_job(tbg_ctx *c)
{
  HIPCHK(hipMemcpyAsync(c->dj, &c->hj, sizeof(JobDev), hipMemcpyHostToDevice, c->st));
  return TBG_OK;
}

int download_job(tbg_ctx *c)
{
  HIPCHK(hipMemcpyAsync(&c->hj, c->dj, sizeof(JobDev), hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return TBG_OK;
}

// the context's own launch sets (both sides, one side, one slice per
// launch for the plain form), with its geometry
int build_ctx_launchsets(tbg_ctx *c)
{
  std::vector<const JobDev *> hj{ &c->hj };
  std::vector<const Mat *> mats{ &c->h };
  int r = build_launchset(c->both, hj, mats, { { 0, 0 }, { 0, 1 } }, 0, c->geom, c->st);
  if (r != TBG_OK) return r;
  for (int s = 0; s < 2; s++) {
    r = build_launchset(c->side[s], hj, mats, { { 0, s } }, 0, c->geom, c->st);
    if (r != TBG_OK) return r;
    r = build_launchset(c->plain[s], hj, mats, { { 0, s } }, 1, c->geom, c->st);
    if (r != TBG_OK) return r;
  }
  return TBG_OK;
}

// Device memory of a context: hipMalloc, or for a pooled (server) context
// the stream-ordered pool of its stream (freed back to the pool, so
// admission after warm-up makes no driver allocation).
int dalloc(tbg_ctx *c, void **p, size_t bytes)
{
  HIPCHK(c->pooled ? hipMallocAsync(p, bytes, c->st) : hipMalloc(p, bytes));
  return TBG_OK;
}

void dfree(tbg_ctx *c, void *p)
{
  if (!p) return;
  if (c->pooled) (void)hipFreeAsync(p, c->st);
  else (void)hipFree(p);
}

// Byte layout of a context's two device blocks (every part 256-byte
// aligned).  State block: the bitmap planes, the capture-exit planes, the
// planes' chunk summaries, the codes of each side; record block: material,
// capture tables, counts (16 words), flags (4), job record, its pointer,
// the per-ply history.
struct CtxLayout {
  size_t arena, sub, sum, codes;             // state block parts (codes: per side)
  size_t o_sub, o_sum, o_codes, state;       // offsets, total
  size_t o_m, o_r, o_cnt, o_flags, o_j, o_jp, o_hist, rec;
};
constexpr size_t kPart = 256;
size_t part_up(size_t x) { return (x + kPart - 1) & ~(kPart - 1); }

// Words of one bitmap plane (all slices of a side); a plane of 8 MB or more
// starts on a 2 MB boundary of its block
uint64_t plane_words(const Mat &M, uint64_t nsl)
{
  uint64_t w = nsl * M.words;
  constexpr uint64_t kAlignWords = (2u << 20) / 8;
  return w >= 4 * kAlignWords ? (w + kAlignWords - 1) / kAlignWords * kAlignWords : w;
}

CtxLayout ctx_layout(const Mat &M, uint64_t nsl, int max_ply, bool with_codes)
{
  CtxLayout L;
  uint64_t pw = plane_words(M, nsl);
  L.arena = part_up(pw * 8 * planes_of(M));
  L.sub = part_up(SUB_KINDS * nsl * (M.ctotal[0] + M.ctotal[1]) * 8 + 64);
  L.codes = with_codes ? part_up(nsl * M.ii.size * 2) : 0;
  // chunk summaries serve the sparse plies of one-slice (synthetic) jobs
  L.sum = M.synth && nsl == 1 && M.words >= (1u << 20) ? part_up(planes_of(M) * sum_per_slice(M)) : 0;
  L.o_sub = L.arena;
  L.o_sum = L.o_sub + L.sub;
  L.o_codes = L.o_sum + L.sum;
  L.state = L.o_codes + 2 * L.codes;
  L.o_m = 0;
  L.o_r = part_up(sizeof(Mat));
  L.o_cnt = L.o_r + part_up(sizeof(SubRef) * MAXS);
  L.o_flags = L.o_cnt + part_up(16 * sizeof(unsigned long long));
  L.o_j = L.o_flags + part_up(4 * sizeof(unsigned));
  L.o_jp = L.o_j + part_up(sizeof(JobDev));
  L.o_hist = L.o_jp + part_up(sizeof(JobDev *));
  L.rec = L.o_hist + part_up(sizeof(long long) * 4 * (size_t)max_ply);
  return L;
}

CtxLayout ctx_layout(const tbg_ctx *c)
{
  return ctx_layout(c->h, c->nsl, c->max_ply, c->codes[0] != nullptr);
}

// the canonical planes in their allocation order, the capture-exit planes
// and the codes at their places in the state block at `state`
void carve_state(tbg_ctx *c, void *state, const CtxLayout &L, bool with_codes)
{
  char *b = static_cast<char *>(state);
  uint64_t *p = reinterpret_cast<uint64_t *>(b);
  c->arena = p;
  for (int s = 0; s < 2; s++) {
    c->wins[s] = p; p += c->pw;
    c->prevW[s] = p; p += c->pw;
    c->prevL[s] = p; p += c->pw;
    c->newW[s] = p; p += c->pw;
    c->newL[s] = p; p += c->pw;
    c->cand[s] = p; p += c->pw;
    c->candW[s] = p; p += c->pw;
  }
  for (int s = 0; s < 2; s++) {
    c->xr[s] = p;
    p += c->pw;
  }
  c->subarena = reinterpret_cast<uint64_t *>(b + L.o_sub);
  p = c->subarena;
  for (int col = 0; col < 2; col++)
    for (int k = 0; k < SUB_KINDS; k++) {
      c->sub[col][k] = p;
      p += c->nsl * c->h.ctotal[col];
    }
  for (int s = 0; s < 2; s++)
    c->codes[s] = with_codes ? reinterpret_cast<uint16_t *>(b + L.o_codes + s * L.codes) : nullptr;
  c->psum = L.sum ? reinterpret_cast<uint8_t *>(b + L.o_sum) : nullptr;
}

void carve_rec(tbg_ctx *c, void *rec, const CtxLayout &L)
{
  char *b = static_cast<char *>(rec);
  c->rec = rec;
  c->dM = reinterpret_cast<Mat *>(b + L.o_m);
  c->dR = reinterpret_cast<SubRef *>(b + L.o_r);
  c->dcnt = reinterpret_cast<unsigned long long *>(b + L.o_cnt);
  c->dflags = reinterpret_cast<unsigned *>(b + L.o_flags);
  c->dj = reinterpret_cast<JobDev *>(b + L.o_j);
  c->djp = reinterpret_cast<JobDev **>(b + L.o_jp);
  c->dhist = reinterpret_cast<long long *>(b + L.o_hist);
}

// ext: the stream of a pooled context (server jobs), else nullptr.  Two
// allocations; the planes and capture-exit planes start zero, the history
// at -1 (no step run).  A pooled context's record block is written by
// setup_pooled, in one copy; any other's here.
int ctx_alloc(tbg_ctx *c, bool with_codes, hipStream_t ext = nullptr)
{
  Mat &M = c->h;
  c->nsl = M.synth ? 1 : NSL;
  M.nslices_phys = c->nsl;
  c->pw = plane_words(M, c->nsl);
  if (ext) {
    c->st = ext;
    c->pooled = true;
  } else {
    HIPCHK(hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking));
  }
  CtxLayout L = ctx_layout(M, c->nsl, c->max_ply, with_codes);
  void *state = nullptr, *rec = nullptr;
  TRY(dalloc(c, &state, L.state));
  carve_state(c, state, L, with_codes);
  TRY(dalloc(c, &rec, L.rec));
  carve_rec(c, rec, L);
  c->state_bytes = L.state;
  c->bytes = L.state + L.rec;
  HIPCHK(hipMemsetAsync(state, 0, L.o_codes, c->st));
  if (!c->pooled) {
    HIPCHK(hipMemsetAsync(c->dhist, 0xff, sizeof(long long) * 4 * c->max_ply, c->st));
    HIPCHK(hipMemcpyAsync(c->dM, &c->h, sizeof(Mat), hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemsetAsync(c->dR, 0, sizeof(SubRef) * MAXS, c->st));
    HIPCHK(hipMemsetAsync(c->dcnt, 0, 16 * sizeof(unsigned long long), c->st));
    HIPCHK(hipMemcpyAsync(c->djp, &c->dj, sizeof(JobDev *), hipMemcpyHostToDevice, c->st));
  }
  c->hist.assign((size_t)4 * c->max_ply, -1);  // sized once: finish_job refills it in place
  manager_order(WHITE, c->mgr[WHITE]);
  manager_order(BLACK, c->mgr[BLACK]);
  for (int i = 0; i < CAT_N; i++) c->ms[i] = 0, c->nl[i] = 0;
  reset_job(c);
  if (c->pooled) return TBG_OK;  // the server batches it; no launch sets of its own, no wait
  TRY(build_ctx_launchsets(c));
  HIPCHK(hipStreamSynchronize(c->st));
  return TBG_OK;
}

// optional per-launch timing (not inside graph capture)
struct Timed {
  tbg_ctx *c;
  int cat;
  hipEvent_t a = nullptr, b = nullptr;
  Timed(tbg_ctx *c_, int cat_) : c(c_), cat(cat_)
  {
    if (!c) return;
    c->launches++;
    c->nl[cat]++;
    if (c->prof) {
      (void)hipEventCreate(&a);
      (void)hipEventCreate(&b);
      (void)hipEventRecord(a, c->st);
    }
  }
  ~Timed()
  {
    if (c && c->prof) {
      (void)hipEventRecord(b, c->st);
      c->evs.push_back({ cat, { a, b } });
    }
  }
};

void flush_timing(tbg_ctx *c)
{
  if (c->evs.empty()) return;
  (void)hipStreamSynchronize(c->st);
  for (auto &e : c->evs) {
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e.second.first, e.second.second);
    c->ms[e.first] += ms;
    (void)hipEventDestroy(e.second.first);
    (void)hipEventDestroy(e.second.second);
  }
  c->evs.clear();
}

unsigned cap_grid(uint64_t want, uint64_t cap) { return (unsigned)(want < 1 ? 1 : want < cap ? want : cap); }

// One launch of a job kernel.  kt: KT_*; kind: 0 L step, 1 W step.
// prof: context receiving timings (null inside graph capture).
// k_uncap_j's resident work-groups per CU (its launch bounds)
constexpr uint64_t kUncapResident = 5;

int launch(tbg_ctx *prof, JobDev *const *jobs, const LaunchSet &L, int kt, int kind, int sw0,
           const unsigned *flags, hipStream_t st, bool lean = false)
{
  uint64_t total = L.total[kt];
  if (!total) return TBG_OK;
  switch (kt) {
  case KT_PRED: {
    Timed t(prof, CAT_PRED);
    uint64_t units = kind == 2 ? 2 * total : total;  // kind 2: both steps
    auto k = L.pred_wide ? (L.sums ? k_pred_wide_sum_j : k_pred_wide_j)
                         : (L.sums ? k_pred_sum_j : k_pred_j);
    hipLaunchKernelGGL(k, dim3(cap_grid(units, (uint64_t)g_cus * L.g.wg[kt])), dim3(PB_THREADS),
                       L.pred_lds, st, jobs, L.pairs, L.P[kt], L.np, sw0, kind);
    break;
  }
  case KT_UNCAP: {
    Timed t(prof, CAT_UNCAP);
    uint64_t per_cu = lean && L.g.wg[kt] > kUncapResident ? kUncapResident : L.g.wg[kt];
    hipLaunchKernelGGL(k_uncap_j, dim3(cap_grid(total, (uint64_t)g_cus * per_cu)), dim3(256),
                       L.uncap_lds * 8, st, jobs, L.pairs, L.P[kt], L.np, sw0, kind, flags,
                       L.uncap_lds);
    break;
  }
  case KT_VERIFY: {
    Timed t(prof, CAT_VERIFY);
    auto k = L.verify_split ? (L.sums ? k_verify_j<true, true> : k_verify_j<true, false>)
                            : (L.sums ? k_verify_j<false, true> : k_verify_j<false, false>);
    hipLaunchKernelGGL(k, dim3(cap_grid(total, (uint64_t)g_cus * L.g.wg[kt])), dim3(256), 0, st,
                       jobs, L.pairs, L.P[kt], L.np, sw0, (int)L.g.verify_uw);
    break;
  }
  default: {
    Timed t(prof, CAT_SWEEP);
    hipLaunchKernelGGL(L.sums ? k_sweep_j<true> : k_sweep_j<false>,
                       dim3(cap_grid(total, (uint64_t)g_cus * L.g.wg[kt])), dim3(256), 0, st, jobs,
                       L.pairs, L.P[kt], L.np, sw0, (int)L.g.sweep_uw);
    break;
  }
  }
  HIPCHK(hipGetLastError());
  return TBG_OK;
}

// One whole ply of every job of a launch set (batched form; also what a
// graph captures): the L and W candidates of both sides (one pred launch),
// then verification, the sweep and the control kernel.
// lean_uncap: the host does not know whether some job is at ply 1 or 101
// (the kernel reads it from flags): the uncapture grids are capped at one
// round of resident work-groups, so that a ply without one costs little.
int launch_ply(tbg_ctx *prof, JobDev *const *jobs, int nj, const LaunchSet &L, unsigned *flags,
               hipStream_t st, bool edge = true, unsigned *done_out = nullptr,
               bool lean_uncap = false)
{
  // the L and W steps mark separate candidate planes from different
  // frontiers, so their uncaptures and preds all come first (the preds in
  // one launch); verification must see wins(opp) before this ply's sweep
  if (edge) TRY(launch(prof, jobs, L, KT_UNCAP, 0, 0, flags, st, lean_uncap));
  if (edge) TRY(launch(prof, jobs, L, KT_UNCAP, 1, 0, flags, st, lean_uncap));
  TRY(launch(prof, jobs, L, KT_PRED, 2, 0, flags, st));
  TRY(launch(prof, jobs, L, KT_VERIFY, 0, 0, flags, st));
  TRY(launch(prof, jobs, L, KT_SWEEP, 1, 0, flags, st));
  hipLaunchKernelGGL(k_ctrl, dim3(1), dim3(1024), 0, st, jobs, nj, flags, done_out);
  HIPCHK(hipGetLastError());
  return TBG_OK;
}

int check_stm(tbg_ctx *c, int stm)
{
  if (!c) return fail(TBG_E_INVAL, "null context");
  if (stm != 0 && stm != 1) return fail(TBG_E_INVAL, "side to move must be 0 or 1");
  if (!c->resolved || !c->inited) return fail(TBG_E_STATE, "resolve captures and init first");
  return TBG_OK;
}

// Host-driven single half-step (kind 0: L, 1: W) of ctx for stm at ply n.
int manual_step(tbg_ctx *c, int kind, int stm, int n, int more, uint64_t *count, bool plain)
{
  TRY(check_stm(c, stm));
  if (n < 1 || n >= c->max_ply) return fail(TBG_E_INVAL, "ply out of range");
  JobDev &J = c->hj;
  J.n = n;
  J.done = 0;
  J.sparse = 0;  // a host-driven step reads and writes every chunk
  J.mask = 1u << (kind * 2 + stm);
  J.more = more ? (kind == 0 ? 1u << (2 + stm) : 1u << stm) : 0u;
  for (int i = 0; i < 4; i++) J.cnt[i] = 0;
  TRY(upload_job(c));
  c->half_steps++;
  bool edge = n == 1 || n == DRAW_RULE + 1;
  if (!plain) {
    if (edge) TRY(launch(c, c->djp, c->side[stm], KT_UNCAP, kind, 0, nullptr, c->st));
    TRY(launch(c, c->djp, c->side[stm], KT_PRED, kind, 0, nullptr, c->st));
    TRY(launch(c, c->djp, c->side[stm], kind == 0 ? KT_VERIFY : KT_SWEEP, kind, 0, nullptr, c->st));
  } else {
    // one slice per launch in the reference's manager order (kslice.c:41-89)
    int ns = (int)c->nsl;
    for (int i = 0; i < ns; i++) {
      int s = c->h.synth ? c->h.synth_slice : c->mgr[stm][i];
      if (edge) TRY(launch(c, c->djp, c->plain[stm], KT_UNCAP, kind, s, nullptr, c->st));
      TRY(launch(c, c->djp, c->plain[stm], KT_PRED, kind, s, nullptr, c->st));
    }
    for (int i = 0; i < ns; i++) {
      int s = c->h.synth ? c->h.synth_slice : c->mgr[stm][i];
      TRY(launch(c, c->djp, c->plain[stm], kind == 0 ? KT_VERIFY : KT_SWEEP, kind, s, nullptr, c->st));
    }
  }
  unsigned long long v = 0;
  HIPCHK(hipMemcpyAsync(&v, reinterpret_cast<char *>(c->dj) + offsetof(JobDev, cnt) +
                                sizeof(unsigned long long) * (kind * 2 + stm),
                        sizeof v, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  if (count) *count = v;
  return TBG_OK;
}

int init_job_state(tbg_ctx *c)
{
  reset_job(c);
  TRY(upload_job(c));
  HIPCHK(hipMemsetAsync(c->dhist, 0xff, sizeof(long long) * 4 * c->max_ply, c->st));
  unsigned f[4] = { 0, 1, 0, 0 };
  HIPCHK(hipMemcpyAsync(c->dflags, f, sizeof f, hipMemcpyHostToDevice, c->st));
  return TBG_OK;
}

int zero_plane(tbg_ctx *c, uint64_t *p)
{
  HIPCHK(hipMemsetAsync(p, 0, c->pw * 8, c->st));
  return TBG_OK;
}

// illegal and mate positions, codes, job state; nothing read back (the
// counts stay in dcnt[0..3] until read_init_counts)
// The init kernels: illegal positions, mates and exit candidates, and with
// the capture exits resolved the static refutations; the counts go to
// dcnt[0..5], which must be zero.
int launch_init(tbg_ctx *c)
{
  Mat &M = c->h;
  int s0 = M.synth ? M.synth_slice : 0, ns = M.synth ? 1 : NSL;
  {
    // one placement per lane (pl_layout); with the capture exits resolved,
    // the static refutations of the verification after it
    uint64_t R = 62 - (uint64_t)(M.ii.numsets - 1);
    bool pl = pl_layout(M);
    uint64_t chunks = pl ? (uint64_t)ns * ((M.words + R - 1) / R)
                         : ((uint64_t)ns * M.words + 4 * INIT_CHUNK - 1) / (4 * INIT_CHUNK);
    dim3 grid = pl ? dim3(cap_grid((chunks + 3) / 4, (uint64_t)g_cus * 16))
                   : dim3(cap_grid(chunks, (uint64_t)g_cus * 16));
    c->xr_ok = false;
    {
      Timed t(c, CAT_INIT);
      if (pl)
        hipLaunchKernelGGL(k_init_pl, grid, dim3(256), 0, c->st, c->dM, c->wins[0], c->wins[1],
                           c->prevL[0], c->prevL[1], c->codes[0], c->codes[1], c->prevW[0],
                           c->prevW[1], c->newW[0], c->newW[1], c->xr[0], c->xr[1], s0, ns,
                           c->dcnt);
      else
        hipLaunchKernelGGL(k_init, grid, dim3(256), 0, c->st, c->dM, c->wins[0], c->wins[1],
                           c->prevL[0], c->prevL[1], c->codes[0], c->codes[1], c->prevW[0],
                           c->prevW[1], c->newW[0], c->newW[1], c->xr[0], c->xr[1], s0, ns,
                           c->dcnt);
      HIPCHK(hipGetLastError());
    }
    if (c->resolved) {
      Timed t(c, CAT_XR);
      bool w1 = w1_in_xr(M);
      if (pl)
        hipLaunchKernelGGL(k_xr_pl, grid, dim3(256), 0, c->st, c->dM, c->sub[0][TBG_SUB_WIN],
                           c->sub[1][TBG_SUB_WIN], w1 ? c->sub[0][TBG_SUB_LOSS] : nullptr,
                           w1 ? c->sub[1][TBG_SUB_LOSS] : nullptr, c->wins[0], c->wins[1],
                           c->xr[0], c->xr[1], c->candW[0], c->candW[1], s0, ns, c->dcnt + 4);
      else
        hipLaunchKernelGGL(k_xr, grid, dim3(256), 0, c->st, c->dM, c->sub[0][TBG_SUB_WIN],
                           c->sub[1][TBG_SUB_WIN], c->wins[0], c->wins[1], c->xr[0], c->xr[1],
                           s0, ns, c->dcnt + 4);
      HIPCHK(hipGetLastError());
      c->xr_ok = true;
    }
  }
  return TBG_OK;
}

int run_init_async(tbg_ctx *c)
{
  HIPCHK(hipMemsetAsync(c->dcnt, 0, 6 * sizeof(unsigned long long), c->st));
  TRY(launch_init(c));
  c->counts.nplies = 0;
  c->inited = 1;
  c->generated = 0;
  return init_job_state(c);
}

// the broken / mate counts of the last init, read on stream st
int read_init_counts(tbg_ctx *c, hipStream_t st)
{
  unsigned long long v[4];
  HIPCHK(hipMemcpyAsync(v, c->dcnt, sizeof v, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  c->counts.broken[0] = v[0];
  c->counts.broken[1] = v[1];
  c->counts.mates[0] = v[2];
  c->counts.mates[1] = v[3];
  return TBG_OK;
}

// the capture-exit classification of every set (k_resolve), from the
// capture tables at dR
int launch_resolve(tbg_ctx *c)
{
  const Mat &M = c->h;
  int s0 = M.synth ? M.synth_slice : 0, ns = M.synth ? 1 : NSL;
  for (int k = 0; k < M.numsets; k++) {
    int col = M.setcol[k];
    uint64_t words = (uint64_t)ns * M.cwords[k];
    Timed t(c, CAT_RESOLVE);
    // a word per lane when the words fill the chip four threads deep
    if (M.all_single && words >= (uint64_t)g_cus * 1024)
      hipLaunchKernelGGL(k_resolve_w, dim3(cap_grid((words + 255) / 256, (uint64_t)g_cus * 16)),
                         dim3(256), 0, c->st, c->dM, c->dR + k, col, k, c->sub[col][0],
                         c->sub[col][1], c->sub[col][2], c->sub[col][3], c->sub[col][SUB_LWIN], s0,
                         ns);
    else  // a word per wave
      hipLaunchKernelGGL(k_resolve, dim3(cap_grid((words + 3) / 4, (uint64_t)g_cus * 16)),
                         dim3(256), 0, c->st, c->dM, c->dR + k, col, k, c->sub[col][0],
                         c->sub[col][1], c->sub[col][2], c->sub[col][3], c->sub[col][SUB_LWIN], s0,
                         ns);
    HIPCHK(hipGetLastError());
  }
  return TBG_OK;
}

// Set-up of a pooled (server) context of a synthetic material, queued on
// its stream without a host wait or a pageable copy: the whole record block
// (material, capture tables, zero counts, flags, job record, its pointer, a
// history of -1: no step run) goes up in one copy from `stage` (pinned host
// memory of at least the layout's rec bytes, untouched until the copy has
// run), then the capture-exit classification and the init kernels.
int setup_pooled(tbg_ctx *c, void *stage)
{
  if (!c->pooled || !c->h.synth) return fail(TBG_E_STATE, "not a pooled synthetic context");
  memset(c->hR, 0, sizeof c->hR);  // hashed exits: no smaller tables
  c->resolved = 1;
  c->xr_ok = true;  // launch_init works the static refutations out
  c->counts.nplies = 0;
  c->inited = 1;
  c->generated = 0;
  reset_job(c);
  CtxLayout L = ctx_layout(c);
  char *h = static_cast<char *>(stage);
  memset(h, 0, L.o_hist);
  memcpy(h + L.o_m, &c->h, sizeof(Mat));
  memcpy(h + L.o_r, c->hR, sizeof(SubRef) * MAXS);
  unsigned f[4] = { 0, 1, 0, 0 };  // no job done; an uncapture ply pending
  memcpy(h + L.o_flags, f, sizeof f);
  memcpy(h + L.o_j, &c->hj, sizeof(JobDev));
  memcpy(h + L.o_jp, &c->dj, sizeof(JobDev *));
  memset(h + L.o_hist, 0xff, L.rec - L.o_hist);
  HIPCHK(hipMemcpyAsync(c->rec, h, L.rec, hipMemcpyHostToDevice, c->st));
  TRY(launch_resolve(c));
  return launch_init(c);
}

int run_init(tbg_ctx *c)
{
  TRY(run_init_async(c));
  return read_init_counts(c, c->st);
}

// plain form: the host reads the job's flags each ply and issues one launch
// per slice and phase for the steps that run
int run_plain(tbg_ctx *c)
{
  for (;;) {
    TRY(download_job(c));
    JobDev &J = c->hj;
    if (J.done) break;
    bool edge = J.n == 1 || J.n == DRAW_RULE + 1;
    int ns = (int)c->nsl;
    for (int kind = 0; kind < 2; kind++)
      for (int stm = 0; stm < 2; stm++) {
        if (!((J.mask >> (kind * 2 + stm)) & 1)) {
          TRY(zero_plane(c, kind == 0 ? J.newL[stm] : J.newW[stm]));
          continue;
        }
        for (int i = 0; i < ns; i++) {
          int s = c->h.synth ? c->h.synth_slice : c->mgr[stm][i];
          if (edge) TRY(launch(c, c->djp, c->plain[stm], KT_UNCAP, kind, s, nullptr, c->st));
          TRY(launch(c, c->djp, c->plain[stm], KT_PRED, kind, s, nullptr, c->st));
        }
        for (int i = 0; i < ns; i++) {
          int s = c->h.synth ? c->h.synth_slice : c->mgr[stm][i];
          TRY(launch(c, c->djp, c->plain[stm], kind == 0 ? KT_VERIFY : KT_SWEEP, kind, s, nullptr,
                     c->st));
        }
      }
    hipLaunchKernelGGL(k_ctrl, dim3(1), dim3(1024), 0, c->st, c->djp, 1, c->dflags, nullptr);
    HIPCHK(hipGetLastError());
  }
  return TBG_OK;
}

// Batched and graph forms over any set of jobs.  jobs: device array of nj
// JobDev; L: their launch set.  Counts executed half-steps into *hs.
int run_batch(tbg_ctx *prof, JobDev *const *jobs, int nj, const LaunchSet &L, unsigned *flags,
              hipStream_t st, bool graph, double *graph_ms)
{
  if (!graph) {
    // the flags come back into pinned memory: a direct copy, no staging
    if (!prof->hflags)
      HIPCHK(hipHostMalloc((void **)&prof->hflags, 2 * sizeof(unsigned), hipHostMallocDefault));
    unsigned *hf = prof->hflags;
    hf[1] = 1;
    for (;;) {
      // the host reads the flags after every ply, so it knows whether any
      // job is at ply 1 or 101 and needs the uncapture launches
      TRY(launch_ply(prof, jobs, nj, L, flags, st, hf[1] != 0));
      HIPCHK(hipMemcpyAsync(hf, flags, 2 * sizeof(unsigned), hipMemcpyDeviceToHost, st));
      HIPCHK(hipStreamSynchronize(st));
      if ((int)hf[0] >= nj) break;
    }
    return TBG_OK;
  }
  // capture one ply once, replay it; check for completion every 4 plies
  hipGraph_t g = nullptr;
  hipGraphExec_t ge = nullptr;
  HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  int r = launch_ply(nullptr, jobs, nj, L, flags, st);
  hipError_t e = hipStreamEndCapture(st, &g);
  if (r != TBG_OK) return r;
  HIPCHK(e);
  HIPCHK(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
  hipEvent_t a, b;
  HIPCHK(hipEventCreate(&a));
  HIPCHK(hipEventCreate(&b));
  HIPCHK(hipEventRecord(a, st));
  if (!prof->hflags)
    HIPCHK(hipHostMalloc((void **)&prof->hflags, 2 * sizeof(unsigned), hipHostMallocDefault));
  unsigned *hf = prof->hflags;
  for (;;) {
    for (int i = 0; i < 4; i++) HIPCHK(hipGraphLaunch(ge, st));
    HIPCHK(hipMemcpyAsync(hf, flags, 2 * sizeof(unsigned), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if ((int)hf[0] >= nj) break;
  }
  HIPCHK(hipEventRecord(b, st));
  HIPCHK(hipEventSynchronize(b));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, a, b));
  if (graph_ms) *graph_ms += ms;
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  (void)hipGraphExecDestroy(ge);
  (void)hipGraphDestroy(g);
  return TBG_OK;
}

// after a generation: pull the final job state and the per-ply counts
// st: the stream to read on (a pooled job's own stream is shared)
// the host's account of a finished job from its final job record (hj) and
// per-ply history (hist)
int finish_counts(tbg_ctx *c)
{
  // the reference's loop has no ply cap: reaching it with moves still to
  // resolve would leave a truncated table
  if (c->hj.done == 1 && c->hj.more && c->hj.n >= c->max_ply && !c->ply_cap_asked)
    return fail(TBG_E_STATE, "%s: ply limit %d reached with positions still resolving",
                c->name, c->max_ply);
  int last = 0;
  for (int n = 1; n < c->max_ply; n++)
    for (int i = 0; i < 4; i++)
      if (c->hist[(size_t)n * 4 + i] >= 0) last = n;
  c->counts.nplies = last;
  uint64_t hs = 0;
  for (int n = 1; n <= last; n++)
    for (int i = 0; i < 4; i++) hs += c->hist[(size_t)n * 4 + i] >= 0;
  c->hs_generated += hs;
  c->generated = 1;
  return TBG_OK;
}

int finish_job(tbg_ctx *c, hipStream_t st = nullptr)
{
  if (!st) st = c->st;
  HIPCHK(hipMemcpyAsync(&c->hj, c->dj, sizeof(JobDev), hipMemcpyDeviceToHost, st));
  c->hist.assign((size_t)4 * c->max_ply, -1);
  HIPCHK(hipMemcpyAsync(c->hist.data(), c->dhist, sizeof(long long) * 4 * c->max_ply,
                        hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  return finish_counts(c);
}

uint64_t executed_half_steps(const tbg_ctx *c)
{
  uint64_t hs = 0;
  for (size_t i = 4; i < c->hist.size(); i++) hs += c->hist[i] >= 0;
  return hs;
}

void emit_lines(const tbg_ctx *c, tbg_line_fn fn, void *user)
{
  char line[64];
  const char *nm[4] = { "broken_w", "broken_b", "l0_w", "l0_b" };
  uint64_t v[4] = { c->counts.broken[0], c->counts.broken[1], c->counts.mates[0], c->counts.mates[1] };
  for (int i = 0; i < 4; i++) {
    snprintf(line, sizeof line, "%s = %llu", nm[i], (unsigned long long)v[i]);
    fn(line, user);
  }
  for (int n = 1; n <= c->counts.nplies; n++)
    for (int i = 0; i < 4; i++) {
      long long x = c->hist[(size_t)n * 4 + i];
      if (x < 0) continue;
      snprintf(line, sizeof line, "%c%d_%c = %lld", i < 2 ? 'l' : 'w', n, "wb"[i & 1], x);
      fn(line, user);
    }
}

// The scatter mode of the job, host and device copies alike (only that
// field: a generation may be under way), so the pred LDS size of the launch
// sets rebuilt next matches what the kernel uses.
int push_scatter(tbg_ctx *c)
{
  c->hj.scatter = c->scatter;
  HIPCHK(hipMemcpyAsync(&c->dj->scatter, &c->hj.scatter, sizeof(int), hipMemcpyHostToDevice, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return TBG_OK;
}

}  // namespace

// ====================================================================
// C ABI
// ====================================================================

// slice-file codec (tbg_io.cpp)
namespace tbg {
int io_write(const char *path, const void *src, uint64_t bytes, int threads, std::string *err);
int io_read(const char *path, void *dst, uint64_t bytes, int threads, std::string *err);
}  // namespace tbg

namespace {

// <wk><bk> of slice s, as the reference names its slice files
std::string slice_name(int s)
{
  char b[8];
  int wk = g_htabs.kksq[s][0], bk = g_htabs.kksq[s][1];
  snprintf(b, sizeof b, "%c%c%c%c", 'a' + (wk & 7), '1' + (wk >> 3), 'a' + (bk & 7),
           '1' + (bk >> 3));
  return b;
}

struct FileTask {
  std::string path;
  uint8_t *ptr;
  uint64_t bytes;
};

// Writes (or reads) every task's file: a single large file gets the codec's
// threads, many small ones are spread over a pool of file threads.
int run_files(std::vector<FileTask> &tasks, bool write)
{
  std::atomic<size_t> next{ 0 };
  std::mutex mu;
  std::string err;
  auto one = [&](const FileTask &t, int threads) {
    std::string e;
    int r = write ? io_write(t.path.c_str(), t.ptr, t.bytes, threads, &e)
                  : io_read(t.path.c_str(), t.ptr, t.bytes, threads, &e);
    if (r) {
      std::lock_guard<std::mutex> g(mu);
      if (err.empty()) err = e;
    }
  };
  if (tasks.size() == 1) {
    one(tasks[0], 16);
  } else {
    std::vector<std::thread> pool;
    int nt = (int)std::min<size_t>(16, tasks.size());
    for (int t = 0; t < nt; t++)
      pool.emplace_back([&]() {
        for (size_t i; (i = next++) < tasks.size();) one(tasks[i], 1);
      });
    for (auto &t : pool) t.join();
  }
  if (!err.empty()) return fail(TBG_E_IO, "%s", err.c_str());
  return TBG_OK;
}

int make_dirs(const std::string &d)
{
  std::error_code ec;
  std::filesystem::create_directories(d, ec);
  if (ec) return fail(TBG_E_IO, "cannot create %s", d.c_str());
  return TBG_OK;
}

// One plane ([slices][words] words, or [slices][n] of another element
// size) between the device and per-slice files <dir>/<stm>/<slice>.
int plane_files(tbg_ctx *c, void *dev, uint64_t slice_bytes, const std::string &dir, int stm,
                bool write, std::vector<uint8_t> &buf)
{
  uint64_t ns = c->nsl, total = ns * slice_bytes;
  std::string d = dir + "/" + "wb"[stm];
  if (write) TRY(make_dirs(d));
  buf.resize(total);
  if (write) HIPCHK(hipMemcpy(buf.data(), dev, total, hipMemcpyDeviceToHost));
  std::vector<FileTask> tasks;
  for (uint64_t i = 0; i < ns; i++) {
    int s = c->h.synth ? c->h.synth_slice : (int)i;
    tasks.push_back({ d + "/" + slice_name(s), buf.data() + i * slice_bytes, slice_bytes });
  }
  TRY(run_files(tasks, write));
  if (!write) HIPCHK(hipMemcpy(dev, buf.data(), total, hipMemcpyHostToDevice));
  return TBG_OK;
}

// the reference's four capture-exit files, and the legal wins (an addition)
const char *kSubNames[SUB_KINDS] = { "loss", "bloss", "cwin", "win", "lwin" };

// the planes of a checkpoint, in a fixed order (both directions use it)
int checkpoint_planes(tbg_ctx *c, const std::string &root, int n, bool write)
{
  std::vector<uint8_t> buf;
  JobDev &J = c->hj;
  uint64_t sb = c->h.words * 8;
  for (int stm = 0; stm < 2; stm++) {
    TRY(plane_files(c, J.wins[stm], sb, root + "/0/wins", stm, write, buf));
    TRY(plane_files(c, J.prevW[stm], sb, root + "/" + std::to_string(n - 1) + "/W", stm, write, buf));
    TRY(plane_files(c, J.prevL[stm], sb, root + "/" + std::to_string(n - 1) + "/L", stm, write, buf));
    if (c->codes[stm])
      TRY(plane_files(c, c->codes[stm], c->h.ii.size * 2, root + "/codes", stm, write, buf));
  }
  for (int col = 0; col < 2; col++)
    if (c->h.ctotal[col])
      for (int k = 0; k < SUB_KINDS; k++)
        TRY(plane_files(c, c->sub[col][k], c->h.ctotal[col] * 8,
                        root + "/sub/" + kSubNames[k], col, write, buf));
  return TBG_OK;
}

}  // namespace

extern "C" {

int tbg_init(int device)
{
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return fail(TBG_E_NODEV, "no HIP device");
  if (device < 0 || device >= n) return fail(TBG_E_INVAL, "device %d out of range", device);
  HIPCHK(hipSetDevice(device));
  hipDeviceProp_t prop;
  HIPCHK(hipGetDeviceProperties(&prop, device));
  g_cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
  snprintf(g_arch, sizeof g_arch, "%s", prop.gcnArchName);
  for (char *q = g_arch; *q; q++)
    if (*q == ' ' || *q == ':') *q = '_';
  build_tabs(g_htabs);
  HIPCHK(upload_tabs(&g_htabs));
  if (device < kMaxDevices && !g_scratch[device]) {
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(TBG_E_HIP, "address of the no-allocation scratch: %s", hipGetErrorString(e));
    g_scratch[device] = static_cast<uint8_t *>(p);
  }
  HIPCHK(hipDeviceSynchronize());
  g_device = device;
  return TBG_OK;
}

const char *tbg_strerror(int err)
{
  switch (err) {
  case TBG_OK: return "ok";
  case TBG_E_INVAL: return g_errbuf[0] ? g_errbuf : "invalid argument";
  case TBG_E_NOMEM: return "out of memory";
  case TBG_E_HIP: return g_errbuf[0] ? g_errbuf : "HIP error";
  case TBG_E_NOSUB: return g_errbuf[0] ? g_errbuf : "missing smaller table";
  case TBG_E_STATE: return g_errbuf[0] ? g_errbuf : "call out of order";
  case TBG_E_BUDGET: return g_errbuf[0] ? g_errbuf : "device memory budget exceeded";
  case TBG_E_NODEV: return "no HIP device";
  case TBG_E_IO: return g_errbuf[0] ? g_errbuf : "slice file error";
  default: return "unknown error";
  }
}

static int create_common(tbg_ctx **out, const char *material, int synth, int slice, uint64_t seed,
                         bool keep_codes = false, hipStream_t ext = nullptr)
{
  if (!out || !material) return fail(TBG_E_INVAL, "null argument");
  *out = nullptr;
  if (g_device < 0) return fail(TBG_E_STATE, "tbg_init not called");
  tbg_ctx *c = new tbg_ctx();
  snprintf(c->name, sizeof c->name, "%s", material);
  int r = setup_mat(c->h, material);
  if (r != TBG_OK) { delete c; return r; }
  c->h.synth = synth;
  c->h.synth_slice = slice;
  c->h.seed = seed;
  c->h.seedmix = splitmix64(seed);
  if (synth) {
    const Tabs &T = g_htabs;
    int kq[2] = { T.kksq[slice][0], T.kksq[slice][1] };
    for (int stm = 0; stm < 2; stm++) {
      int from = kq[stm], oks = kq[stm ^ 1], n = 0;
      uint64_t b = T.king[from] & ~T.king[oks];
      while (b) {
        int to = ctz64(b);
        b &= b - 1;
        int wk = stm == WHITE ? to : oks, bk = stm == WHITE ? oks : to;
        uint64_t q;
        int s2 = normalize(T, (uint64_t)wk | ((uint64_t)bk << 8), &q);
        if (s2 == slice) continue;
        c->h.exit_to[stm][n] = (uint8_t)to;
        c->h.exit_m[stm][n] = (uint8_t)mirror_mask(wk);
        c->h.exit_fl[stm][n] = flip_test(wk, bk) ? 1 : 0;
        c->h.exit_s2[stm][n] = (int16_t)s2;
        n++;
      }
      c->h.nexit[stm] = n;
    }
  }
  r = ctx_alloc(c, !synth || keep_codes, ext);
  if (r != TBG_OK) {
    tbg_destroy(c);
    return r == TBG_E_HIP ? fail(TBG_E_NOMEM, "device allocation failed") : r;
  }
  *out = c;
  return TBG_OK;
}

int tbg_create(tbg_ctx **out, const char *material) { return create_common(out, material, 0, -1, 0); }

int tbg_create_synth(tbg_ctx **out, const char *material, int slice, uint64_t seed)
{
  return tbg_create_synth_ex(out, material, slice, seed, 0);
}

int tbg_create_synth_ex(tbg_ctx **out, const char *material, int slice, uint64_t seed,
                        unsigned flags)
{
  if (slice < 0 || slice >= 441) return fail(TBG_E_INVAL, "synthetic slice must be in [0, 441)");
  if (flags & ~(unsigned)TBG_SYNTH_CODES) return fail(TBG_E_INVAL, "unknown flags");
  return create_common(out, material, 1, slice, seed, (flags & TBG_SYNTH_CODES) != 0);
}

void tbg_destroy(tbg_ctx *c)
{
  if (!c) return;
  // a pooled context's stream is shared: its frees are stream-ordered and
  // its owner has already waited for the work that used the memory
  if (c->st && !c->pooled) (void)hipStreamSynchronize(c->st);
  flush_timing(c);
  c->both.release();
  for (int s = 0; s < 2; s++) c->side[s].release(), c->plain[s].release();
  dfree(c, c->arena);  // the state block
  dfree(c, c->rec);
  if (c->ready) (void)hipEventDestroy(c->ready);
  if (c->hflags) (void)hipHostFree(c->hflags);
  if (c->st && !c->pooled) (void)hipStreamDestroy(c->st);
  delete c;
}

int tbg_get_info(const tbg_ctx *c, tbg_info *info)
{
  if (!c || !info) return fail(TBG_E_INVAL, "null argument");
  memset(info, 0, sizeof *info);
  info->npieces = c->h.num;
  for (int i = 0; i < 8; i++) info->pt[i] = (uint8_t)c->h.pt[i];
  info->numsets = c->h.numsets;
  for (int k = 0; k < c->h.numsets; k++) {
    info->mult[k] = c->h.ii.mult[k];
    info->capt_size[k] = c->h.capt[k].size;
    info->sub_offset[k] = c->h.coff[k];
  }
  info->slice_size = c->h.ii.size;
  info->slice_words = c->h.words;
  info->sub_words[0] = c->h.ctotal[0];
  info->sub_words[1] = c->h.ctotal[1];
  info->device_bytes = c->bytes;
  info->synth_slice = c->h.synth ? c->h.synth_slice : -1;
  return TBG_OK;
}

int tbg_resolve_captures(tbg_ctx *c, tbg_ctx *const *subs, int nsubs)
{
  if (!c) return fail(TBG_E_INVAL, "null context");
  Mat &M = c->h;
  for (int k = 0; k < M.numsets; k++) {
    SubRef &R = c->hR[k];
    memset(&R, 0, sizeof R);
    int spt[MAXP], slots[MAXP], n = 0;
    for (int i = 0; i < M.num; i++)
      if (i != M.ii.last[k]) { spt[n] = M.pt[i]; slots[n] = i; n++; }
    if (M.synth || n == 2) continue;  // hashed exits / KvK
    char nm[16], fl[16];
    material_string(spt, n, false, nm);
    material_string(spt, n, true, fl);
    tbg_ctx *hit = nullptr;
    int flip = 0;
    for (int i = 0; i < nsubs && !hit; i++) {
      if (!subs[i]) continue;
      if (!strcmp(subs[i]->name, nm)) hit = subs[i];
      else if (!strcmp(subs[i]->name, fl)) { hit = subs[i]; flip = 1; }
    }
    if (!hit) return fail(TBG_E_NOSUB, "table %s (or %s) not supplied", nm, fl);
    if (!hit->codes[0] || !hit->generated) return fail(TBG_E_STATE, "table %s not generated", hit->name);
    R.present = 1;
    R.flip = flip;
    R.num = hit->h.num;
    R.ii = hit->h.ii;
    R.codes[0] = hit->codes[0];
    R.codes[1] = hit->codes[1];
    bool used[MAXP] = { false };
    R.map = 0;
    for (int j = 0; j < R.num; j++) {
      int want = hit->h.pt[j] ^ (flip ? 8 : 0);
      int src = -1;
      for (int i = 0; i < n; i++)
        if (!used[i] && spt[i] == want) { used[i] = true; src = slots[i]; break; }
      if (src < 0) return fail(TBG_E_NOSUB, "table %s does not match", hit->name);
      R.map = sq_set(R.map, j, src);
    }
  }
  HIPCHK(hipMemcpyAsync(c->dR, c->hR, sizeof(SubRef) * MAXS, hipMemcpyHostToDevice, c->st));
  TRY(launch_resolve(c));
  if (!c->pooled) HIPCHK(hipStreamSynchronize(c->st));  // a pooled job is waited for by its event
  c->resolved = 1;
  if (c->xr_ok) {  // worked out from the earlier capture exits: verify every move
    c->xr_ok = false;
    c->hj.xr[0] = c->hj.xr[1] = nullptr;
    c->hj.nopen = nullptr;
    if (c->hj.w1cap) {  // the ply-1 W capture candidates too: the uncapture makes them again
      c->hj.w1cap = 0;
      // still unconsumed unless a generation has finished since (its sweeps
      // leave the candidate planes clear)
      if (!c->generated)
        for (int s = 0; s < 2; s++) HIPCHK(hipMemsetAsync(c->candW[s], 0, c->pw * 8, c->st));
    }
    if (c->inited) TRY(upload_job(c));
  }
  return TBG_OK;
}

int tbg_init_illegal_mate(tbg_ctx *c, tbg_counts *counts)
{
  if (!c) return fail(TBG_E_INVAL, "null context");
  TRY(run_init(c));
  if (counts) *counts = c->counts;
  return TBG_OK;
}

int tbg_step_L(tbg_ctx *c, int stm, int n, int more, uint64_t *count)
{
  return manual_step(c, 0, stm, n, more, count, false);
}
int tbg_step_W(tbg_ctx *c, int stm, int n, int more, uint64_t *count)
{
  return manual_step(c, 1, stm, n, more, count, false);
}
int tbg_step_L_plain(tbg_ctx *c, int stm, int n, int more, uint64_t *count)
{
  return manual_step(c, 0, stm, n, more, count, true);
}
int tbg_step_W_plain(tbg_ctx *c, int stm, int n, int more, uint64_t *count)
{
  return manual_step(c, 1, stm, n, more, count, true);
}

int tbg_end_ply(tbg_ctx *c)
{
  if (!c) return fail(TBG_E_INVAL, "null context");
  JobDev &J = c->hj;
  for (int s = 0; s < 2; s++) {
    uint64_t *x = J.prevW[s];
    J.prevW[s] = J.newW[s];
    J.newW[s] = x;
    x = J.prevL[s];
    J.prevL[s] = J.newL[s];
    J.newL[s] = x;
  }
  return upload_job(c);
}

int tbg_generate_batch(tbg_ctx *const *cs, int nc, int form, tbg_counts *counts)
{
  if (!cs || nc < 1 || nc > 1024) return fail(TBG_E_INVAL, "bad batch");
  if (form < TBG_FORM_PLAIN || form > TBG_FORM_GRAPH) return fail(TBG_E_INVAL, "bad form");
  for (int i = 0; i < nc; i++) {
    if (!cs[i]) return fail(TBG_E_INVAL, "null context in batch");
    if (!cs[i]->resolved) return fail(TBG_E_STATE, "resolve captures of %s first", cs[i]->name);
    for (int j = 0; j < i; j++)
      if (cs[j] == cs[i]) return fail(TBG_E_INVAL, "context twice in a batch");
  }
  for (int i = 0; i < nc; i++) TRY(run_init(cs[i]));
  tbg_ctx *c0 = cs[0];
  if (form == TBG_FORM_PLAIN) {
    for (int i = 0; i < nc; i++) {
      TRY(run_plain(cs[i]));
      TRY(finish_job(cs[i]));
    }
  } else if (nc == 1) {
    TRY(run_batch(c0, c0->djp, 1, c0->both, c0->dflags, c0->st, form == TBG_FORM_GRAPH, &c0->graph_ms));
    TRY(finish_job(c0));
  } else {
    // the jobs of all contexts behind one pointer array, one launch set
    JobDev **dj = nullptr;
    unsigned *df = nullptr;
    HIPCHK(hipMalloc(&dj, sizeof(JobDev *) * nc));
    HIPCHK(hipMalloc(&df, 4 * sizeof(unsigned)));
    std::vector<const JobDev *> hj;
    std::vector<const Mat *> mats;
    std::vector<Pair> pairs;
    std::vector<JobDev *> ptrs;
    for (int i = 0; i < nc; i++) {
      HIPCHK(hipStreamSynchronize(cs[i]->st));
      ptrs.push_back(cs[i]->dj);
      hj.push_back(&cs[i]->hj);
      mats.push_back(&cs[i]->h);
      pairs.push_back({ i, 0 });
      pairs.push_back({ i, 1 });
    }
    HIPCHK(hipMemcpyAsync(dj, ptrs.data(), sizeof(JobDev *) * nc, hipMemcpyHostToDevice, c0->st));
    unsigned f[4] = { 0, 1, 0, 0 };
    HIPCHK(hipMemcpyAsync(df, f, sizeof f, hipMemcpyHostToDevice, c0->st));
    LaunchSet L;
    int r = build_launchset(L, hj, mats, pairs, 0, c0->geom, c0->st);
    if (r == TBG_OK)
      r = run_batch(c0, dj, nc, L, df, c0->st, form == TBG_FORM_GRAPH, &c0->graph_ms);
    (void)hipStreamSynchronize(c0->st);
    L.release();
    (void)hipFree(dj);
    (void)hipFree(df);
    if (r != TBG_OK) return r;
    for (int i = 0; i < nc; i++) TRY(finish_job(cs[i]));
  }
  if (counts)
    for (int i = 0; i < nc; i++) counts[i] = cs[i]->counts;
  return TBG_OK;
}

int tbg_generate_batch_noalloc(tbg_ctx *const *cs, int nc, tbg_counts *counts)
{
  if (!cs || nc < 1) return fail(TBG_E_INVAL, "bad batch");
  if (nc > TBG_NOALLOC_MAX_BATCH)
    return fail(TBG_E_INVAL, "batch of %d contexts: the no-allocation scratch holds at most %d",
                nc, TBG_NOALLOC_MAX_BATCH);
  if (g_device < 0 || g_device >= kMaxDevices || !g_scratch[g_device])
    return fail(TBG_E_STATE, "tbg_init not called");
  for (int i = 0; i < nc; i++) {
    if (!cs[i]) return fail(TBG_E_INVAL, "null context in batch");
    if (!cs[i]->resolved) return fail(TBG_E_STATE, "resolve captures of %s first", cs[i]->name);
    if (cs[i]->prof) return fail(TBG_E_STATE, "profiling of %s is on (it creates events)", cs[i]->name);
    if (cs[i]->pooled) return fail(TBG_E_STATE, "%s belongs to a server", cs[i]->name);
    for (int j = 0; j < i; j++)
      if (cs[j] == cs[i]) return fail(TBG_E_INVAL, "context twice in a batch");
  }
  for (int i = 0; i < nc; i++) TRY(run_init(cs[i]));
  tbg_ctx *c0 = cs[0];
  uint8_t *sp = g_scratch[g_device];
  JobDev **dj = reinterpret_cast<JobDev **>(sp + kNaJobs);
  unsigned *df = reinterpret_cast<unsigned *>(sp + kNaFlags);
  // host staging on the stack: no heap either
  JobDev *ptrs[kNaMax];
  const JobDev *hj[kNaMax];
  const Mat *mats[kNaMax];
  Pair pairs[2 * kNaMax];
  uint64_t P[4 * (2 * kNaMax + 1)];
  for (int i = 0; i < nc; i++) {
    HIPCHK(hipStreamSynchronize(cs[i]->st));
    ptrs[i] = cs[i]->dj;
    hj[i] = &cs[i]->hj;
    mats[i] = &cs[i]->h;
    pairs[2 * i] = { i, 0 };
    pairs[2 * i + 1] = { i, 1 };
  }
  LaunchSet L;  // its arrays live in the scratch: never released
  plan_launchset(L, hj, mats, nc, pairs, 2 * nc, 0, c0->geom, P);
  L.pairs = reinterpret_cast<Pair *>(sp + kNaPairs);
  for (int kt = 0; kt < 4; kt++) L.P[kt] = reinterpret_cast<uint64_t *>(sp + kNaP + kt * kNaPStride);
  L.cap = 2 * nc;
  // the scratch keeps the previous call's contents: every field read below
  // is written first (flags start at {0 done, 1 uncapture pending})
  unsigned f[4] = { 0, 1, 0, 0 };
  HIPCHK(hipMemcpyAsync(dj, ptrs, sizeof(JobDev *) * nc, hipMemcpyHostToDevice, c0->st));
  HIPCHK(hipMemcpyAsync(df, f, sizeof f, hipMemcpyHostToDevice, c0->st));
  TRY(upload_launchset(L, pairs, P, c0->st));
  unsigned hf[2] = { 0, 1 };
  for (;;) {
    TRY(launch_ply(nullptr, dj, nc, L, df, c0->st, hf[1] != 0));
    HIPCHK(hipMemcpyAsync(hf, df, sizeof hf, hipMemcpyDeviceToHost, c0->st));
    HIPCHK(hipStreamSynchronize(c0->st));
    if ((int)hf[0] >= nc) break;
  }
  for (int i = 0; i < nc; i++) TRY(finish_job(cs[i], c0->st));
  if (counts)
    for (int i = 0; i < nc; i++) counts[i] = cs[i]->counts;
  return TBG_OK;
}

int tbg_emit_lines(const tbg_ctx *c, tbg_line_fn fn, void *user)
{
  if (!c || !fn) return fail(TBG_E_INVAL, "null argument");
  if (!c->generated) return fail(TBG_E_STATE, "%s is not generated", c->name);
  emit_lines(c, fn, user);
  return TBG_OK;
}

int tbg_generate(tbg_ctx *c, int form, tbg_line_fn fn, void *user, tbg_counts *counts)
{
  if (!c) return fail(TBG_E_INVAL, "null context");
  tbg_ctx *one[1] = { c };
  TRY(tbg_generate_batch(one, 1, form, counts));
  if (fn) emit_lines(c, fn, user);
  return TBG_OK;
}

}  // extern "C"

struct tbg_registry {
  int form = TBG_FORM_GRAPH;
  std::vector<std::pair<std::string, tbg_ctx *>> tables;  // in generation order
};

namespace {

// colour orientation of generated smaller materials: the side with more
// pieces, then the stronger piece string (K Q R B N), is white
int piece_rank(char ch)
{
  const char *o = "KQRBN";
  const char *p = strchr(o, ch);
  return p ? (int)(p - o) : 99;
}

std::string sorted_side(std::string s)
{
  std::stable_sort(s.begin(), s.end(), [](char a, char b) { return piece_rank(a) < piece_rank(b); });
  return s;
}

bool side_ge(const std::string &a, const std::string &b)
{
  if (a.size() != b.size()) return a.size() > b.size();
  for (size_t i = 0; i < a.size(); i++)
    if (a[i] != b[i]) return piece_rank(a[i]) < piece_rank(b[i]);
  return true;
}

std::string canonical_material(const std::string &m)
{
  size_t v = m.find('v');
  std::string w = sorted_side(m.substr(0, v)), b = sorted_side(m.substr(v + 1));
  return side_ge(w, b) ? w + "v" + b : b + "v" + w;
}

// materials one capture away, without a bare pair of kings
std::vector<std::string> sub_material_names(const std::string &m)
{
  std::vector<std::string> out;
  size_t v = m.find('v');
  if (v == std::string::npos) return out;
  std::string w = m.substr(0, v), b = m.substr(v + 1);
  for (int white = 1; white >= 0; white--) {
    const std::string &side = white ? w : b, &other = white ? b : w;
    for (size_t i = 0; i < side.size(); i++) {
      if (side[i] == 'K') continue;
      std::string rest = side.substr(0, i) + side.substr(i + 1);
      std::string sub = white ? rest + "v" + other : other + "v" + rest;
      if (sub.size() - 1 <= 2) continue;
      std::string c = canonical_material(sub);
      if (std::find(out.begin(), out.end(), c) == out.end()) out.push_back(c);
    }
  }
  return out;
}

tbg_ctx *registry_find(tbg_registry *r, const std::string &m)
{
  for (auto &e : r->tables)
    if (e.first == m) return e.second;
  return nullptr;
}

}  // namespace

extern "C" {

int tbg_registry_create(tbg_registry **out, int form)
{
  if (!out) return fail(TBG_E_INVAL, "null argument");
  if (form < TBG_FORM_PLAIN || form > TBG_FORM_GRAPH) return fail(TBG_E_INVAL, "bad form");
  *out = new tbg_registry();
  (*out)->form = form;
  return TBG_OK;
}

int tbg_registry_get(tbg_registry *r, const char *material, tbg_ctx **out)
{
  if (!r || !material || !out) return fail(TBG_E_INVAL, "null argument");
  *out = nullptr;
  std::string want(material);
  if (tbg_ctx *c = registry_find(r, want)) {
    *out = c;
    return TBG_OK;
  }
  // post-order walk of the dependency graph: every smaller material first
  std::vector<std::string> need;
  std::vector<std::string> stack;
  std::function<int(const std::string &)> visit = [&](const std::string &m) -> int {
    if (registry_find(r, m) || std::find(need.begin(), need.end(), m) != need.end()) return TBG_OK;
    if (std::find(stack.begin(), stack.end(), m) != stack.end()) return TBG_OK;
    stack.push_back(m);
    for (const std::string &sub : sub_material_names(m)) TRY(visit(sub));
    stack.pop_back();
    need.push_back(m);
    return TBG_OK;
  };
  TRY(visit(want));
  // one batch per piece count, smallest first
  size_t lo = 99, hi = 0;
  for (auto &m : need) lo = std::min(lo, m.size()), hi = std::max(hi, m.size());
  for (size_t len = lo; len <= hi; len++) {
    std::vector<tbg_ctx *> batch;
    std::vector<std::string> names;
    int rc = TBG_OK;
    for (auto &m : need) {
      if (m.size() != len) continue;
      tbg_ctx *c = nullptr;
      rc = tbg_create(&c, m.c_str());
      if (rc != TBG_OK) break;
      batch.push_back(c);
      names.push_back(m);
    }
    std::vector<tbg_ctx *> subs;
    for (auto &e : r->tables) subs.push_back(e.second);
    for (size_t i = 0; i < batch.size() && rc == TBG_OK; i++)
      rc = tbg_resolve_captures(batch[i], subs.data(), (int)subs.size());
    if (rc == TBG_OK && !batch.empty())
      rc = tbg_generate_batch(batch.data(), (int)batch.size(), r->form, nullptr);
    if (rc != TBG_OK) {
      for (tbg_ctx *c : batch) tbg_destroy(c);
      return rc;
    }
    for (size_t i = 0; i < batch.size(); i++) r->tables.push_back({ names[i], batch[i] });
  }
  *out = registry_find(r, want);
  return *out ? TBG_OK : fail(TBG_E_STATE, "%s was not generated", material);
}

int tbg_registry_count(const tbg_registry *r)
{
  return r ? (int)r->tables.size() : 0;
}

int tbg_registry_entry(const tbg_registry *r, int i, const char **name, tbg_ctx **ctx)
{
  if (!r || i < 0 || i >= (int)r->tables.size()) return fail(TBG_E_INVAL, "no such registry entry");
  if (name) *name = r->tables[i].first.c_str();
  if (ctx) *ctx = r->tables[i].second;
  return TBG_OK;
}

void tbg_registry_destroy(tbg_registry *r)
{
  if (!r) return;
  for (auto &e : r->tables) tbg_destroy(e.second);
  delete r;
}

int tbg_file_write(const char *path, const void *data, uint64_t bytes)
{
  if (!path || (!data && bytes)) return fail(TBG_E_INVAL, "null argument");
  std::string e;
  if (io_write(path, data, bytes, 0, &e)) return fail(TBG_E_IO, "%s", e.c_str());
  return TBG_OK;
}

int tbg_file_read(const char *path, void *data, uint64_t bytes)
{
  if (!path || (!data && bytes)) return fail(TBG_E_INVAL, "null argument");
  std::string e;
  if (io_read(path, data, bytes, 0, &e)) return fail(TBG_E_IO, "%s", e.c_str());
  return TBG_OK;
}

int tbg_checkpoint(tbg_ctx *c, const char *root)
{
  if (!c || !root) return fail(TBG_E_INVAL, "null argument");
  if (!c->inited) return fail(TBG_E_STATE, "nothing to checkpoint: init or restore first");
  TRY(download_job(c));  // the current planes, ply and flags
  JobDev &J = c->hj;
  if (J.n < 2)
    return fail(TBG_E_STATE, "checkpoint after ply 1 (a synthetic job's exits are in scratch planes at ply 1)");
  std::string r(root);
  TRY(make_dirs(r));
  TRY(checkpoint_planes(c, r, J.n, true));
  std::vector<long long> hist((size_t)4 * J.n);
  HIPCHK(hipMemcpy(hist.data(), c->dhist, hist.size() * sizeof(long long), hipMemcpyDeviceToHost));
  unsigned long long cnt[4];
  HIPCHK(hipMemcpy(cnt, c->dcnt, sizeof cnt, hipMemcpyDeviceToHost));
  std::string path = r + "/tbg_state";
  FILE *F = fopen(path.c_str(), "w");
  if (!F) return fail(TBG_E_IO, "cannot write %s", path.c_str());
  fprintf(F, "tbg-checkpoint 1\nmaterial %s\nslice %d\nseed %llu\n", c->name,
          c->h.synth ? c->h.synth_slice : -1, (unsigned long long)c->h.seed);
  fprintf(F, "n %d\nmore %u\nmask %u\ndone %d\n", J.n, J.more, J.mask, J.done == 1 ? 1 : 0);
  fprintf(F, "init %llu %llu %llu %llu\n", cnt[0], cnt[1], cnt[2], cnt[3]);
  for (int n = 1; n < J.n; n++)
    fprintf(F, "ply %d %lld %lld %lld %lld\n", n, hist[n * 4], hist[n * 4 + 1], hist[n * 4 + 2],
            hist[n * 4 + 3]);
  if (fclose(F) != 0) return fail(TBG_E_IO, "cannot write %s", path.c_str());
  return TBG_OK;
}

int tbg_restore(tbg_ctx *c, const char *root)
{
  if (!c || !root) return fail(TBG_E_INVAL, "null argument");
  std::string r(root);
  std::string path = r + "/tbg_state";
  FILE *F = fopen(path.c_str(), "r");
  if (!F) return fail(TBG_E_IO, "cannot read %s", path.c_str());
  char mat[32] = { 0 };
  int ver = 0, slice = 0, n = 0, done = 0;
  unsigned long long seed = 0, cnt[4] = { 0, 0, 0, 0 };
  unsigned more = 0, mask = 0;
  bool ok = fscanf(F, "tbg-checkpoint %d material %31s slice %d seed %llu n %d more %u mask %u "
                      "done %d init %llu %llu %llu %llu",
                   &ver, mat, &slice, &seed, &n, &more, &mask, &done, &cnt[0], &cnt[1], &cnt[2],
                   &cnt[3]) == 12;
  std::vector<long long> hist((size_t)4 * c->max_ply, -1);
  int pn;
  long long h[4];
  while (ok && fscanf(F, " ply %d %lld %lld %lld %lld", &pn, &h[0], &h[1], &h[2], &h[3]) == 5) {
    if (pn < 1 || pn >= c->max_ply) { ok = false; break; }
    for (int i = 0; i < 4; i++) hist[(size_t)pn * 4 + i] = h[i];
  }
  fclose(F);
  if (!ok || ver != 1 || n < 2 || n >= c->max_ply) return fail(TBG_E_IO, "%s: bad state file", path.c_str());
  if (strcmp(mat, c->name) != 0 || slice != (c->h.synth ? c->h.synth_slice : -1) ||
      (c->h.synth && seed != c->h.seed))
    return fail(TBG_E_INVAL, "checkpoint of %s slice %d does not match this context", mat, slice);
  HIPCHK(hipStreamSynchronize(c->st));
  c->xr_ok = false;  // not in the files: the restored job verifies every move
  reset_job(c);  // the canonical planes: the files go to wins, prevW, prevL
  JobDev &J = c->hj;
  uint64_t pb = c->pw * 8;
  for (int s = 0; s < 2; s++) {
    HIPCHK(hipMemsetAsync(c->newW[s], 0, pb, c->st));
    HIPCHK(hipMemsetAsync(c->newL[s], 0, pb, c->st));
    HIPCHK(hipMemsetAsync(c->cand[s], 0, pb, c->st));
    HIPCHK(hipMemsetAsync(c->candW[s], 0, pb, c->st));
  }
  HIPCHK(hipStreamSynchronize(c->st));
  TRY(checkpoint_planes(c, r, n, false));
  J.n = n;
  J.more = more;
  J.mask = mask;
  J.done = done ? 1 : 0;
  HIPCHK(hipMemcpy(c->dhist, hist.data(), hist.size() * sizeof(long long), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->dcnt, cnt, sizeof cnt, hipMemcpyHostToDevice));
  c->counts.broken[0] = cnt[0];
  c->counts.broken[1] = cnt[1];
  c->counts.mates[0] = cnt[2];
  c->counts.mates[1] = cnt[3];
  c->resolved = 1;
  c->inited = 1;
  c->generated = done ? 1 : 0;
  TRY(upload_job(c));
  HIPCHK(hipStreamSynchronize(c->st));
  return TBG_OK;
}

int tbg_continue(tbg_ctx *c, int form, int stop_at, tbg_line_fn fn, void *user, tbg_counts *counts)
{
  if (!c) return fail(TBG_E_INVAL, "null context");
  if (form < TBG_FORM_PLAIN || form > TBG_FORM_GRAPH) return fail(TBG_E_INVAL, "bad form");
  if (stop_at < 0) return fail(TBG_E_INVAL, "stop_at must be >= 0");
  if (!c->inited) return fail(TBG_E_STATE, "init or restore first");
  TRY(download_job(c));
  JobDev &J = c->hj;
  if (J.done != 1 && !(stop_at > 0 && J.n >= stop_at)) {
    J.done = 0;
    J.stop_at = stop_at;
    TRY(upload_job(c));
    unsigned f[4] = { 0, (J.n == 1 || J.n == DRAW_RULE + 1) ? 1u : 0u, 0, 0 };
    HIPCHK(hipMemcpyAsync(c->dflags, f, sizeof f, hipMemcpyHostToDevice, c->st));
    if (form == TBG_FORM_PLAIN) TRY(run_plain(c));
    else TRY(run_batch(c, c->djp, 1, c->both, c->dflags, c->st, form == TBG_FORM_GRAPH, &c->graph_ms));
  }
  TRY(finish_job(c));
  c->generated = c->hj.done == 1;
  if (counts) *counts = c->counts;
  if (fn) emit_lines(c, fn, user);
  return TBG_OK;
}

int tbg_export_plies(tbg_ctx *c, const char *root)
{
  if (!c || !root) return fail(TBG_E_INVAL, "null argument");
  if (!c->generated || !c->codes[0] || !c->codes[1])
    return fail(TBG_E_STATE, "export needs a finished generation that keeps codes");
  std::string r(root);
  uint64_t *tmp = nullptr;
  HIPCHK(hipMalloc(&tmp, c->pw * 8));
  std::vector<uint8_t> buf;
  uint64_t sb = c->h.words * 8;
  int rc = TBG_OK;
  auto code_plane = [&](int stm, uint16_t code) -> int {
    hipLaunchKernelGGL(k_code_plane, dim3(cap_grid((c->pw + 3) / 4, (uint64_t)g_cus * 16)), dim3(256),
                       0, c->st, c->codes[stm], c->h.ii.size, c->h.words, (int)c->nsl, code, tmp);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->st));
    return TBG_OK;
  };
  for (int stm = 0; stm < 2 && rc == TBG_OK; stm++) {
    rc = plane_files(c, c->hj.wins[stm], sb, r + "/0/wins", stm, true, buf);
    if (rc == TBG_OK) rc = code_plane(stm, 2);  // mates: L_0
    if (rc == TBG_OK) rc = plane_files(c, tmp, sb, r + "/0/L", stm, true, buf);
    for (int n = 1; n <= c->counts.nplies && rc == TBG_OK; n++)
      for (int kind = 0; kind < 2 && rc == TBG_OK; kind++) {
        if (c->hist[(size_t)n * 4 + kind * 2 + stm] < 0) continue;  // the step did not run
        rc = code_plane(stm, (uint16_t)((n << 2) | (kind == 0 ? 2 : 1)));
        if (rc == TBG_OK)
          rc = plane_files(c, tmp, sb, r + "/" + std::to_string(n) + (kind == 0 ? "/L" : "/W"), stm,
                           true, buf);
      }
  }
  (void)hipFree(tmp);
  return rc;
}

}  // extern "C"

namespace tbg {
int rtbw_write(const char *path, const uint8_t *pieces, int num, uint8_t order, int nsides,
               uint8_t *const *vals, const uint64_t *n, int threads, std::string *err);
}

namespace {

uint64_t binom64(int n, int k)
{
  if (k < 0 || k > n) return 0;
  uint64_t r = 1;
  for (int i = 0; i < k; i++) r = r * (uint64_t)(n - i) / (uint64_t)(i + 1);
  return r;
}

// create_piece_string (src/probe.c:243-255): the idx-th multiset of n
// non-king pieces in the probing code's enumeration, strongest first
std::string syz_piece_string(int n, uint64_t idx)
{
  static const char pc[] = "PNBRQ";
  std::string s((size_t)n, ' ');
  for (int k = n - 1; k > 0; k--) {
    int l = 0;
    while (idx >= binom64(k + 1 + l, k + 1)) l++;
    idx -= binom64(k + l, k + 1);
    s[(size_t)(n - 1 - k)] = pc[l];
  }
  if (n) s[(size_t)(n - 1)] = pc[idx];
  return s;
}

// non-king pieces of colour c, strongest first
std::string side_pieces(const Mat &M, int c)
{
  std::string s;
  for (int t = QUEEN; t >= KNIGHT; t--)
    for (int i = 2; i < M.num; i++)
      if (M.pt[i] == (t | (c << 3))) s += kPieceChar[t];
  return s;
}

// The name under which the probing code looks for this material's WDL file
// (init_tablebases' enumeration, src/probe.c:304-318: the side with more
// pieces, or with equal counts the later multiset, is white) and whether
// its colours are the reverse of ours.  False beyond 7 pieces.
bool syzygy_name(const Mat &M, std::string *name, int *swap)
{
  std::string w = side_pieces(M, WHITE), b = side_pieces(M, BLACK);
  for (int p = 1; p <= 5; p++)
    for (int q = 0; q <= std::min(p, 5 - p); q++)
      for (int64_t k = (int64_t)binom64(p + 4, 4) - 1; k >= 0; k--) {
        std::string ws = syz_piece_string(p, (uint64_t)k);
        for (int64_t l = q < p ? (int64_t)binom64(q + 4, 4) - 1 : k; l >= 0; l--) {
          std::string bs = syz_piece_string(q, (uint64_t)l);
          if ((ws == w && bs == b) || (ws == b && bs == w)) {
            *name = "K" + ws + "vK" + bs;
            *swap = !(ws == w && bs == b);
            return true;
          }
        }
      }
  return false;
}

// Header piece order and index layout of the file (set_enc_info,
// src/probe.c:725-770): both kings first, then, unless every non-king
// piece has a twin, one unique piece to make the leading triple; then the
// remaining pieces in runs of like pieces, strongest first, white first.
// The leading group is the most significant digit (order = number of
// groups after it).
int syz_layout(const Mat &M, int swap, SyzLayout *L, uint8_t *pieces, uint8_t *order)
{
  memset(L, 0, sizeof *L);
  int hp[MAXP], n = 0;
  hp[n++] = KING;
  hp[n++] = KING | 8;
  for (int nc = 0; nc < 2; nc++)
    for (int t = QUEEN; t >= KNIGHT; t--)
      for (int i = 2; i < M.num; i++)
        if (M.pt[i] == (t | ((nc ^ swap) << 3))) hp[n++] = t | (nc << 3);
  int cnt[16] = { 0 };
  for (int i = 0; i < n; i++) cnt[hp[i]]++;
  int lead = 2;
  for (int i = 2; i < n; i++)
    if (cnt[hp[i]] == 1) {  // a unique piece joins the kings
      int u = hp[i];
      for (int j = i; j > 2; j--) hp[j] = hp[j - 1];
      hp[2] = u;
      lead = 3;
      break;
    }
  L->num = n;
  L->lead = lead;
  bool used[MAXP] = { false };
  for (int i = 0; i < n; i++) {
    pieces[i] = (uint8_t)hp[i];
    int want = hp[i] ^ (swap ? 8 : 0), slot = -1;
    for (int j = 0; j < M.num && slot < 0; j++)
      if (!used[j] && M.pt[j] == want) slot = j;
    if (slot < 0) return fail(TBG_E_INVAL, "internal error: header piece without a slot");
    used[slot] = true;
    L->slot[i] = slot;
  }
  uint64_t f = 1;
  int free_sq = 64 - lead;
  for (int i = lead; i < n;) {
    int g = L->ngroups++, m = 0;
    while (i + m < n && hp[i + m] == hp[i]) m++;
    L->gfirst[g] = i;
    L->gmult[g] = m;
    L->gfactor[g] = f;
    L->gsize[g] = binom64(free_sq, m);
    f *= L->gsize[g];
    free_sq -= m;
    i += m;
  }
  L->lead_factor = f;
  L->tb_size = f * (lead == 2 ? 462 : 31332);
  *order = (uint8_t)L->ngroups;
  for (int i = 0; i < 10; i++) L->inv_tri[i] = kInvTri[i];
  for (int i = 0; i < 8; i++) {
    L->inv_diag[i] = (uint8_t)(9 * i);        // a1 .. h8
    L->inv_diag[8 + i] = (uint8_t)(7 * (i + 1));  // h1 .. a8
  }
  int k = 0;
  for (int r = 0; r < 8; r++)  // squares below a1-h8, rank by rank, then a1-h8
    for (int fl = r + 1; fl < 8; fl++) L->inv_lower[k++] = (uint8_t)(r * 8 + fl);
  for (int i = 0; i < 8; i++) L->inv_lower[k++] = (uint8_t)(9 * i);
  return TBG_OK;
}

}  // namespace

extern "C" {

int tbg_rtbw_name(tbg_ctx *c, char *name, size_t cap)
{
  if (!c || !name) return fail(TBG_E_INVAL, "null argument");
  std::string nm;
  int swap;
  if (c->h.synth || !syzygy_name(c->h, &nm, &swap))
    return fail(TBG_E_INVAL, "%s has no WDL file name (at most 7 pieces, whole tables)", c->name);
  if (nm.size() + 1 > cap) return fail(TBG_E_INVAL, "name buffer too small");
  memcpy(name, nm.c_str(), nm.size() + 1);
  return TBG_OK;
}

int tbg_write_rtbw(tbg_ctx *c, const char *dir, char *path, size_t cap)
{
  if (!c || !dir) return fail(TBG_E_INVAL, "null argument");
  if (c->h.synth) return fail(TBG_E_INVAL, "a single-slice job is not a whole table");
  if (!c->generated || !c->codes[0] || !c->codes[1])
    return fail(TBG_E_STATE, "writing a table needs a finished generation that keeps codes");
  std::string nm;
  int swap;
  if (!syzygy_name(c->h, &nm, &swap))
    return fail(TBG_E_INVAL, "%s has no WDL file name (at most 7 pieces)", c->name);
  std::string file = std::string(dir) + "/" + nm + ".rtbw";
  if (path && file.size() + 1 > cap) return fail(TBG_E_INVAL, "path buffer too small");
  SyzLayout L;
  uint8_t pieces[MAXP], order;
  TRY(syz_layout(c->h, swap, &L, pieces, &order));
  // a material equal to its colour reversal has one sub-table (white to
  // move); otherwise one per side to move, in the name's colours
  int nsides = side_pieces(c->h, WHITE) == side_pieces(c->h, BLACK) ? 1 : 2;
  std::vector<uint8_t> vals[2];
  const uint64_t chunk = std::min<uint64_t>(L.tb_size, 1ull << 28);
  uint8_t *dbuf = nullptr;
  HIPCHK(hipMalloc(&dbuf, chunk));
  int rc = TBG_OK;
  for (int t = 0; t < nsides && rc == TBG_OK; t++) {
    L.stm = t ^ swap;
    vals[t].resize(L.tb_size);
    for (uint64_t i0 = 0; i0 < L.tb_size && rc == TBG_OK; i0 += chunk) {
      uint64_t n = std::min(chunk, L.tb_size - i0);
      hipLaunchKernelGGL(k_syz_gather, dim3(cap_grid((n + 255) / 256, (uint64_t)g_cus * 32)), dim3(256),
                         0, c->st, c->dM, c->codes[L.stm], L, i0, n, dbuf);
      hipError_t e = hipGetLastError();
      if (e == hipSuccess) e = hipMemcpyAsync(vals[t].data() + i0, dbuf, n, hipMemcpyDeviceToHost, c->st);
      if (e == hipSuccess) e = hipStreamSynchronize(c->st);
      if (e != hipSuccess) rc = fail(TBG_E_HIP, "HIP error: %s", hipGetErrorString(e));
    }
  }
  (void)hipFree(dbuf);
  if (rc != TBG_OK) return rc;
  uint8_t *vp[2] = { vals[0].data(), nsides == 2 ? vals[1].data() : nullptr };
  uint64_t ns[2] = { L.tb_size, L.tb_size };
  std::string e;
  if (rtbw_write(file.c_str(), pieces, L.num, order, nsides, vp, ns, 0, &e))
    return fail(TBG_E_IO, "%s", e.c_str());
  if (path) memcpy(path, file.c_str(), file.size() + 1);
  return TBG_OK;
}

int tbg_get_codes(tbg_ctx *c, int stm, int slice, uint16_t *dst, uint64_t n)
{
  if (!c || !dst || (stm & ~1)) return fail(TBG_E_INVAL, "bad argument");
  if (!c->codes[stm]) return fail(TBG_E_STATE, "context keeps no codes");
  uint64_t ps = c->h.synth ? 0 : (uint64_t)slice;
  if (c->h.synth ? slice != c->h.synth_slice : (slice < 0 || slice >= NSL))
    return fail(TBG_E_INVAL, "slice out of range");
  if (n > c->h.ii.size) return fail(TBG_E_INVAL, "too many codes");
  HIPCHK(hipMemcpyAsync(dst, c->codes[stm] + ps * c->h.ii.size, n * 2, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return TBG_OK;
}

int tbg_get_bitmap(tbg_ctx *c, int plane, int stm, int slice, uint64_t *dst, uint64_t nwords)
{
  if (!c || !dst || (stm & ~1)) return fail(TBG_E_INVAL, "bad argument");
  if (c->h.synth ? slice != c->h.synth_slice : (slice < 0 || slice >= NSL))
    return fail(TBG_E_INVAL, "slice out of range");
  if (nwords > c->h.words) return fail(TBG_E_INVAL, "too many words");
  uint64_t *p = plane == TBG_PLANE_WINS ? c->hj.wins[stm]
              : plane == TBG_PLANE_W    ? c->hj.prevW[stm]
              : plane == TBG_PLANE_L    ? c->hj.prevL[stm]
                                        : nullptr;
  if (!p) return fail(TBG_E_INVAL, "bad plane");
  uint64_t ps = c->h.synth ? 0 : (uint64_t)slice;
  HIPCHK(hipMemcpyAsync(dst, p + ps * c->h.words, nwords * 8, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return TBG_OK;
}

int tbg_get_sub_bitmap(tbg_ctx *c, int kind, int colour, int slice, uint64_t *dst, uint64_t nwords)
{
  if (!c || !dst || (colour & ~1) || kind < 0 || kind > 3) return fail(TBG_E_INVAL, "bad argument");
  if (c->h.synth ? slice != c->h.synth_slice : (slice < 0 || slice >= NSL))
    return fail(TBG_E_INVAL, "slice out of range");
  if (nwords > c->h.ctotal[colour]) return fail(TBG_E_INVAL, "too many words");
  uint64_t ps = c->h.synth ? 0 : (uint64_t)slice;
  HIPCHK(hipMemcpyAsync(dst, c->sub[colour][kind] + ps * c->h.ctotal[colour], nwords * 8,
                        hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return TBG_OK;
}

int tbg_set_scatter(tbg_ctx *c, int mode)
{
  if (!c) return fail(TBG_E_INVAL, "null context");
  if (mode & ~(TBG_SCATTER_BLOCK | TBG_SCATTER_COMBINE)) return fail(TBG_E_INVAL, "bad scatter mode");
  HIPCHK(hipStreamSynchronize(c->st));
  c->scatter = mode;
  TRY(push_scatter(c));
  return build_ctx_launchsets(c);
}

namespace {
int check_geometry(const tbg_geometry *g)
{
  if (!g) return fail(TBG_E_INVAL, "null geometry");
  if (g->verify_unit_words < 256 || g->verify_unit_words > 65536 || g->verify_unit_words % 256)
    return fail(TBG_E_INVAL, "verify_unit_words must be a multiple of 256 in [256, 65536]");
  if (g->sweep_unit_words < 512 || g->sweep_unit_words > 65536 || g->sweep_unit_words % 512)
    return fail(TBG_E_INVAL, "sweep_unit_words must be a multiple of 512 in [512, 65536]");
  const int wg[4] = { g->pred_wg_per_cu, g->uncap_wg_per_cu, g->verify_wg_per_cu, g->sweep_wg_per_cu };
  for (int x : wg)
    if (x < 1 || x > 64) return fail(TBG_E_INVAL, "work-groups per CU must be in [1, 64]");
  if (g->scatter & ~(TBG_SCATTER_BLOCK | TBG_SCATTER_COMBINE))
    return fail(TBG_E_INVAL, "bad scatter mode");
  return TBG_OK;
}

Geom to_geom(const tbg_geometry *g)
{
  Geom r;
  r.verify_uw = (uint32_t)g->verify_unit_words;
  r.sweep_uw = (uint32_t)g->sweep_unit_words;
  r.wg[KT_PRED] = (uint32_t)g->pred_wg_per_cu;
  r.wg[KT_UNCAP] = (uint32_t)g->uncap_wg_per_cu;
  r.wg[KT_VERIFY] = (uint32_t)g->verify_wg_per_cu;
  r.wg[KT_SWEEP] = (uint32_t)g->sweep_wg_per_cu;
  return r;
}

void from_geom(const Geom &r, int scatter, tbg_geometry *g)
{
  g->verify_unit_words = (int)r.verify_uw;
  g->sweep_unit_words = (int)r.sweep_uw;
  g->pred_wg_per_cu = (int)r.wg[KT_PRED];
  g->uncap_wg_per_cu = (int)r.wg[KT_UNCAP];
  g->verify_wg_per_cu = (int)r.wg[KT_VERIFY];
  g->sweep_wg_per_cu = (int)r.wg[KT_SWEEP];
  g->scatter = scatter;
}
}  // namespace

int tbg_set_geometry(tbg_ctx *c, const tbg_geometry *g)
{
  TRY(check_geometry(g));
  if (!c) {
    g_geom = to_geom(g);
    g_scatter = g->scatter;
    return TBG_OK;
  }
  HIPCHK(hipStreamSynchronize(c->st));
  c->geom = to_geom(g);
  c->scatter = g->scatter;
  TRY(push_scatter(c));
  return build_ctx_launchsets(c);
}

int tbg_get_geometry(const tbg_ctx *c, tbg_geometry *g)
{
  if (!g) return fail(TBG_E_INVAL, "null geometry");
  if (c) from_geom(c->geom, c->scatter, g);
  else from_geom(g_geom, g_scatter, g);
  return TBG_OK;
}

namespace {
// Candidate geometries of the autotuner.  Trial i runs candidate i of every
// kernel at once: the kernels of a ply run one after another on one stream,
// so each kernel's device time depends on its own parameters only, and
// each kernel's best candidate is read from its own time.
const tbg_geometry kTune[] = {
  // verify uw, sweep uw, pred wg, uncap wg, verify wg, sweep wg, scatter
  { 256, 4096, 3, 16, 16, 16, TBG_SCATTER_BLOCK | TBG_SCATTER_COMBINE },  // the built-in default
  { 1024, 2048, 2, 8, 16, 16, TBG_SCATTER_BLOCK | TBG_SCATTER_COMBINE },
  { 4096, 16384, 4, 32, 8, 8, TBG_SCATTER_BLOCK | TBG_SCATTER_COMBINE },
  { 512, 8192, 8, 16, 32, 16, TBG_SCATTER_GLOBAL },
  { 2048, 1024, 3, 4, 16, 32, TBG_SCATTER_GLOBAL | TBG_SCATTER_COMBINE },
  { 8192, 32768, 6, 16, 8, 8, TBG_SCATTER_BLOCK },
};
constexpr int kNTune = (int)(sizeof kTune / sizeof kTune[0]);

void tune_key(const tbg_ctx *c, int max_plies, char *k, size_t n)
{
  snprintf(k, n, "%s %d %s %d %d", c->name, c->h.synth ? c->h.synth_slice : -1, g_arch, g_cus,
           max_plies);
}

bool tune_lookup(const char *path, const char *key, tbg_geometry *g)
{
  FILE *f = path ? fopen(path, "r") : nullptr;
  if (!f) return false;
  char line[512];
  bool hit = false;
  while (fgets(line, sizeof line, f)) {
    char mat[64], arch[64], k[256];
    int sl, cus, mp;
    tbg_geometry t;
    double ms;
    if (sscanf(line, "tbg-geometry %63s %d %63s %d %d %d %d %d %d %d %d %d %lf", mat, &sl, arch, &cus,
               &mp, &t.verify_unit_words, &t.sweep_unit_words, &t.pred_wg_per_cu,
               &t.uncap_wg_per_cu, &t.verify_wg_per_cu, &t.sweep_wg_per_cu, &t.scatter, &ms) != 13)
      continue;
    snprintf(k, sizeof k, "%s %d %s %d %d", mat, sl, arch, cus, mp);
    if (strcmp(k, key) == 0 && check_geometry(&t) == TBG_OK) {
      *g = t;  // the last matching line wins
      hit = true;
    }
  }
  fclose(f);
  return hit;
}

// One timed generation (bounded to max_plies plies when > 0) with geometry
// g; per-category device times into ms[], their sum returned in *total.
int tune_trial(tbg_ctx *c, const tbg_geometry &g, int max_plies, double *ms, double *total)
{
  TRY(tbg_set_geometry(c, &g));
  TRY(tbg_set_profiling(c, 1));
  int saved = c->max_ply;
  if (max_plies > 0 && max_plies + 1 < saved) c->max_ply = max_plies + 1;
  c->ply_cap_asked = true;
  tbg_ctx *one[1] = { c };
  int r = tbg_generate_batch(one, 1, TBG_FORM_BATCHED, nullptr);
  c->max_ply = saved;
  c->ply_cap_asked = false;
  if (r != TBG_OK) return r;
  flush_timing(c);
  *total = 0;
  for (int i = 0; i < CAT_N; i++) {
    ms[i] = c->ms[i];
    *total += c->ms[i];
  }
  return TBG_OK;
}
}  // namespace

int tbg_autotune(tbg_ctx *c, int max_plies, const char *cache_path, tbg_geometry *out, double *out_ms)
{
  if (!c) return fail(TBG_E_INVAL, "null context");
  if (!c->resolved) return fail(TBG_E_STATE, "resolve captures of %s first", c->name);
  if (max_plies < 0) return fail(TBG_E_INVAL, "max_plies must be >= 0");
  char key[256];
  tune_key(c, max_plies, key, sizeof key);
  tbg_geometry best;
  if (tune_lookup(cache_path, key, &best)) {
    TRY(tbg_set_geometry(c, &best));
    if (out) *out = best;
    if (out_ms) *out_ms = -1;
    return TBG_OK;
  }
  bool prof = c->prof;
  double ms[kNTune][CAT_N], tot[kNTune];
  for (int i = 0; i < kNTune; i++) TRY(tune_trial(c, kTune[i], max_plies, ms[i], &tot[i]));
  // each kernel's fastest candidate
  auto argmin = [&](int cat) {
    int b = 0;
    for (int i = 1; i < kNTune; i++)
      if (ms[i][cat] < ms[b][cat]) b = i;
    return b;
  };
  int bp = argmin(CAT_PRED), bv = argmin(CAT_VERIFY), bs = argmin(CAT_SWEEP), bu = argmin(CAT_UNCAP);
  tbg_geometry mix;
  mix.verify_unit_words = kTune[bv].verify_unit_words;
  mix.verify_wg_per_cu = kTune[bv].verify_wg_per_cu;
  mix.sweep_unit_words = kTune[bs].sweep_unit_words;
  mix.sweep_wg_per_cu = kTune[bs].sweep_wg_per_cu;
  mix.pred_wg_per_cu = kTune[bp].pred_wg_per_cu;
  mix.scatter = kTune[bp].scatter;
  mix.uncap_wg_per_cu = kTune[bu].uncap_wg_per_cu;
  double mms[CAT_N], mtot;
  TRY(tune_trial(c, mix, max_plies, mms, &mtot));
  // the combination is confirmed against the best whole trial
  int bt = 0;
  for (int i = 1; i < kNTune; i++)
    if (tot[i] < tot[bt]) bt = i;
  best = mtot <= tot[bt] ? mix : kTune[bt];
  double best_ms = mtot <= tot[bt] ? mtot : tot[bt];
  TRY(tbg_set_geometry(c, &best));
  TRY(tbg_set_profiling(c, prof));
  if (max_plies > 0) c->generated = 0;  // a bounded trial leaves a partial table
  if (cache_path) {
    FILE *f = fopen(cache_path, "a");
    if (f) {
      fprintf(f, "tbg-geometry %s %d %d %d %d %d %d %d %.4f\n", key, best.verify_unit_words,
              best.sweep_unit_words, best.pred_wg_per_cu, best.uncap_wg_per_cu,
              best.verify_wg_per_cu, best.sweep_wg_per_cu, best.scatter, best_ms);
      fclose(f);
    }
  }
  if (out) *out = best;
  if (out_ms) *out_ms = best_ms;
  return TBG_OK;
}

int tbg_set_profiling(tbg_ctx *c, int on)
{
  if (!c) return fail(TBG_E_INVAL, "null context");
  flush_timing(c);
  c->prof = on != 0;
  for (int i = 0; i < CAT_N; i++) c->ms[i] = 0, c->nl[i] = 0;
  c->launches = c->half_steps = c->hs_generated = 0;
  c->graph_ms = 0;
  return TBG_OK;
}

int tbg_get_timing(const tbg_ctx *cc, tbg_timing *t)
{
  tbg_ctx *c = const_cast<tbg_ctx *>(cc);
  if (!c || !t) return fail(TBG_E_INVAL, "null argument");
  flush_timing(c);
  t->ms_pred = c->ms[CAT_PRED];
  t->ms_verify = c->ms[CAT_VERIFY];
  t->ms_sweep = c->ms[CAT_SWEEP];
  t->ms_uncap = c->ms[CAT_UNCAP];
  t->ms_init = c->ms[CAT_INIT];
  t->ms_resolve = c->ms[CAT_RESOLVE];
  t->ms_xr = c->ms[CAT_XR];
  t->ms_total = 0;
  for (int i = 0; i < CAT_N; i++) t->ms_total += c->ms[i];
  t->launches = c->launches;
  t->half_steps = c->half_steps + c->hs_generated;
  t->n_pred = c->nl[CAT_PRED];
  t->n_verify = c->nl[CAT_VERIFY];
  t->n_sweep = c->nl[CAT_SWEEP];
  t->n_uncap = c->nl[CAT_UNCAP];
  t->n_init = c->nl[CAT_INIT];
  t->n_resolve = c->nl[CAT_RESOLVE];
  t->n_xr = c->nl[CAT_XR];
  return TBG_OK;
}

}  // extern "C"

// ====================================================================
// Host self-test of the __host__ __device__ primitives (no GPU needed):
// attacks against ray walking, select_zero and div_fix against brute
// force, index round trips and bijection, slice symmetry invariants.
// ====================================================================
namespace {

uint64_t walk_att(int pt, int sq, uint64_t occ)
{
  static const int df[8] = { 1, -1, 0, 0, 1, 1, -1, -1 }, dr[8] = { 0, 0, 1, -1, 1, -1, 1, -1 };
  int d0 = (pt & 7) == BISHOP ? 4 : 0, d1 = (pt & 7) == ROOK ? 4 : 8;
  uint64_t a = 0;
  for (int d = d0; d < d1; d++)
    for (int f = (sq & 7) + df[d], r = (sq >> 3) + dr[d]; f >= 0 && f < 8 && r >= 0 && r < 8;
         f += df[d], r += dr[d]) {
      a |= 1ull << (r * 8 + f);
      if (occ & (1ull << (r * 8 + f))) break;
    }
  return a;
}

uint64_t rng_state = 0x2545f4914f6cdd1dull;
uint64_t rnd()
{
  rng_state ^= rng_state << 13;
  rng_state ^= rng_state >> 7;
  rng_state ^= rng_state << 17;
  return rng_state;
}

}  // namespace

namespace {
const Tabs &numbering_tabs()
{
  static const Tabs T = [] {
    Tabs t;
    build_tabs(t);
    return t;
  }();
  return T;
}
}  // namespace

extern "C" int tbg_king_slice(int wk, int bk)
{
  if (wk < 0 || wk > 63 || bk < 0 || bk > 63 || wk == bk || touching(wk, bk)) return -1;
  uint64_t q;
  return normalize(numbering_tabs(), (uint64_t)wk | ((uint64_t)bk << 8), &q);
}

extern "C" int tbg_slice_kings(int s, int *wk, int *bk)
{
  if (s < 0 || s >= NSL || !wk || !bk) return fail(TBG_E_INVAL, "no such slice");
  *wk = numbering_tabs().kksq[s][0];
  *bk = numbering_tabs().kksq[s][1];
  return TBG_OK;
}

extern "C" int tbg_host_selftest(char *msg, int len)
{
#define ST_FAIL(...)                       \
  do {                                     \
    if (msg) snprintf(msg, len, __VA_ARGS__); \
    return 1;                              \
  } while (0)
  build_tabs(g_htabs);
  const Tabs &T = g_htabs;
  for (int it = 0; it < 4000; it++) {
    uint64_t occ = rnd() & rnd();
    for (int sq = 0; sq < 64; sq++) {
      uint64_t o = occ | (1ull << sq);
      if (rook_att(sq, o) != walk_att(ROOK, sq, o)) ST_FAIL("rook attacks sq %d occ %llx", sq, (unsigned long long)o);
      if (bishop_att(sq, o) != walk_att(BISHOP, sq, o)) ST_FAIL("bishop attacks sq %d", sq);
    }
  }
  for (int it = 0; it < 20000; it++) {
    uint64_t occ = rnd() & rnd() & rnd();
    int nz = 64 - popc64(occ);
    int r = (int)(rnd() % (uint64_t)nz);
    int s = -1, seen = 0;
    for (int q = 0; q < 64; q++)
      if (!(occ >> q & 1) && seen++ == r) { s = q; break; }
    if (select_zero(occ, r) != s) ST_FAIL("select_zero");
  }
  // synth_value_k against synth_value(synth_hash(..)): random keys and
  // indices, and hashes whose high half ties a threshold's (preimages by
  // inverting splitmix64's steps)
  {
    auto inv_xs = [](uint64_t y, int sh) {
      uint64_t x = y;
      for (int i = 0; i < 64 / sh + 1; i++) x = y ^ (x >> sh);
      return x;
    };
    auto inv_mul = [](uint64_t a) {
      uint64_t x = a;
      for (int i = 0; i < 6; i++) x *= 2 - a * x;
      return x;
    };
    for (int it = 0; it < 400000; it++) {
      uint64_t seedmix = splitmix64(rnd());
      int stm = (int)(rnd() & 1), kind = (int)(rnd() % 3), region = (int)(rnd() % 4096);
      uint64_t key = synth_key(seedmix, stm, kind, region), idx;
      if (it & 1) {
        idx = rnd() & 0xfffffffffull;
      } else {  // an index whose hash ties a threshold in its high half
        uint64_t hi = (it & 2) ? (SYN_T_WIN >> 32) : (SYN_T_LOSS >> 32);
        uint64_t h = hi << 32 | (rnd() >> 32);
        uint64_t z = inv_xs(h, 31) * inv_mul(0x94d049bb133111ebull);
        z = inv_xs(z, 27) * inv_mul(0xbf58476d1ce4e5b9ull);
        uint64_t x = inv_xs(z, 30) - 0x9e3779b97f4a7c15ull;
        if (splitmix64(x) != h) ST_FAIL("splitmix64 preimage");
        idx = x ^ key;  // any 64-bit value: the function takes idx ^ key
      }
      int want = synth_value(splitmix64(idx ^ key));
      if ((idx >> 36) == 0 && want != synth_value(synth_hash(seedmix, stm, kind, region, idx)))
        ST_FAIL("synth_key");
      if (synth_value_k(key, idx) != want) ST_FAIL("synth_value_k");
    }
  }
  // square masks under the slice symmetries, against sym_sq per square
  for (int m : { 0, 7, 56, 63 })
    for (int fl = 0; fl < 2; fl++) {
      for (int t = 0; t < 64; t++) {
        int u = sym_sq(t, m, fl != 0);
        if (sym_bb(1ull << t, m, fl != 0) != 1ull << u) ST_FAIL("sym_bb m %d fl %d t %d", m, fl, t);
        if (sym_bb_inv(1ull << u, m, fl != 0) != 1ull << t) ST_FAIL("sym_bb_inv m %d fl %d", m, fl);
      }
      for (int it = 0; it < 1000; it++) {
        uint64_t x = rnd(), y = 0;
        for (int t = 0; t < 64; t++)
          if ((x >> t) & 1) y |= 1ull << sym_sq(t, m, fl != 0);
        if (sym_bb(x, m, fl != 0) != y || sym_bb_inv(y, m, fl != 0) != x) ST_FAIL("sym_bb mask");
      }
    }
  for (int it = 0; it < 200000; it++) {
    uint64_t d = 1 + rnd() % 0xffffffffull, n = rnd() % (1ull << 40), rem;
    uint64_t q = div_fix(n, d, 1.0 / (double)d, &rem);
    if (q != n / d || rem != n % d) ST_FAIL("div_fix %llu / %llu", (unsigned long long)n, (unsigned long long)d);
  }
  // 462 slices, KKMap invariant under the 8 board symmetries
  int cnt[NSL] = { 0 };
  for (int wk = 0; wk < 64; wk++)
    for (int bk = 0; bk < 64; bk++) {
      int s = kkmap(wk, bk);
      if (wk == bk || touching(wk, bk)) { if (s != -1) ST_FAIL("kkmap touching"); continue; }
      if (s < 0 || s >= NSL) ST_FAIL("kkmap range %d %d", wk, bk);
      cnt[s]++;
      for (int sym = 0; sym < 8; sym++) {
        auto tr = [&](int x) {
          int f = x & 7, r = x >> 3;
          if (sym & 1) f = 7 - f;
          if (sym & 2) r = 7 - r;
          if (sym & 4) { int t = f; f = r; r = t; }
          return r * 8 + f;
        };
        if (kkmap(tr(wk), tr(bk)) != s) ST_FAIL("kkmap symmetry");
      }
    }
  for (int s = 0; s < NSL; s++) {
    if (!cnt[s]) ST_FAIL("slice %d unused", s);
    int wk = T.kksq[s][0], bk = T.kksq[s][1];
    if (kkmap(wk, bk) != s) ST_FAIL("kksq/kkmap disagree at %d", s);
    if ((s >= 441) != (on_diag(wk) && on_diag(bk))) ST_FAIL("diagonal slice numbering");
  }
  // div_small against integer division around quotient boundaries, for
  // divisors up to 2^32 and quotients up to 2^21
  for (uint64_t d = 2; d < (1ull << 32); d = d * 3 / 2 + 1) {
    float inv = 1.0f / (float)d;
    for (uint64_t q = 0; q < (1u << 21); q = q * 2 + 1)
      for (int e = -1; e <= 1; e++) {
        int64_t n = (int64_t)(q * d) + e;
        if (n < 0 || (uint64_t)n / d >= (1u << 21)) continue;
        uint64_t r;
        uint32_t qq = div_small((uint64_t)n, (uint32_t)d, inv, &r);
        if (qq != (uint64_t)n / d || r != (uint64_t)n % d)
          ST_FAIL("div_small %lld / %llu", (long long)n, (unsigned long long)d);
      }
  }
  // index bijection (exhaustive on small materials) and round trips
  // KRRRRRvK: a set whose radix (C(62, 5)) takes the double-precision digits
  const char *mats[] = { "KRvK", "KRRvK", "KBNvK", "KQRvKR", "KNNNvKB", "KQRBvKQRN", "KRRRvKRR",
                         "KRRRRRvK" };
  for (const char *nm : mats) {
    Mat M;
    if (setup_mat(M, nm) != TBG_OK) ST_FAIL("setup %s", nm);
    std::vector<Idx> spaces(1, M.ii);
    for (int k = 0; k < M.numsets; k++) spaces.push_back(M.capt[k]);
    for (size_t sp = 0; sp < spaces.size(); sp++) {
      const Idx &ii = spaces[sp];
      bool exhaustive = ii.size <= 300000;
      for (int s = 0; s < NSL; s += exhaustive ? 37 : 1) {
        uint64_t kings = (uint64_t)T.kksq[s][0] | ((uint64_t)T.kksq[s][1] << 8);
        uint64_t nsamp = exhaustive ? ii.size : 200;
        for (uint64_t t = 0; t < nsamp; t++) {
          uint64_t idx = exhaustive ? t : rnd() % ii.size;
          uint64_t p = kings;
          uint64_t occ = idx_to_sq(T, ii, idx, &p);
          uint64_t chk = 0;
          int npl = 2;
          for (int k = 0; k < ii.numsets; k++) npl += ii.mult[k];
          for (int k = 0; k < ii.numsets; k++)
            for (int a = 0; a < ii.mult[k]; a++) {
              int q = sq_get(p, ii.first[k] + a);
              if (q > 63) ST_FAIL("%s: square out of range", nm);
              chk |= 1ull << q;
            }
          chk |= (1ull << T.kksq[s][0]) | (1ull << T.kksq[s][1]);
          if (chk != occ || popc64(occ) != npl) ST_FAIL("%s space %zu: occupancy idx %llu", nm, sp, (unsigned long long)idx);
          // members of a set ascend (unrank_binomial order)
          for (int k = 0; k < ii.numsets; k++)
            for (int a = 1; a < ii.mult[k]; a++)
              if (sq_get(p, ii.first[k] + a) <= sq_get(p, ii.first[k] + a - 1)) ST_FAIL("%s: set order", nm);
          if (sq_to_idx(T, ii, p) != idx) ST_FAIL("%s space %zu: round trip idx %llu slice %d", nm, sp, (unsigned long long)idx, s);
          bool single = true;
          for (int k = 0; k < ii.numsets; k++) single &= ii.mult[k] == 1;
          if (single && ii.small_q) {
            uint64_t p1 = kings;
            if (idx_to_sq_1(ii, idx, &p1) != occ || p1 != p) ST_FAIL("%s space %zu: idx_to_sq_1", nm, sp);
          }
          if (sp == 0 && M.all_single) {  // step_by against a full decode
            uint32_t deltas[4] = { 1, 63, 4097, (uint32_t)(rnd() % 100000) + 1 };
            for (uint32_t dl : deltas) {
              if (idx + dl >= ii.size) continue;
              uint64_t D = digits_of(p, ii.numsets), p1 = p, p2 = kings;
              uint64_t o1 = step_by(ii.numsets, &D, &p1, dl), o2 = idx_to_sq(T, ii, idx + dl, &p2);
              if (o1 != o2 || p1 != p2 || D != digits_of(p2, ii.numsets))
                ST_FAIL("%s: step_by %u from %llu", nm, dl, (unsigned long long)idx);
            }
          }
          if (sp == 0 && M.all_single && idx + 64 < ii.size) {  // step64 against a full decode
            uint64_t D = digits_of(p, ii.numsets), p1 = p, p2 = kings;
            uint64_t o1 = step64(ii.numsets, &D, &p1), o2 = idx_to_sq(T, ii, idx + 64, &p2);
            if (o1 != o2 || p1 != p2 || D != digits_of(p2, ii.numsets))
              ST_FAIL("%s: step64 from %llu", nm, (unsigned long long)idx);
          }
          // permuting members within a set does not change the index
          for (int k = 0; k < ii.numsets; k++)
            if (ii.mult[k] >= 2) {
              int f = ii.first[k];
              uint64_t p2 = sq_set(sq_set(p, f, sq_get(p, f + 1)), f + 1, sq_get(p, f));
              if (sq_to_idx(T, ii, p2) != idx) ST_FAIL("%s: set permutation", nm);
            }
        }
      }
    }
  }
  // incremental index of a moved piece equals the recomputed index
  {
    const char *dm[] = { "KQRvKR", "KQRBvKQRN", "KRvKN", "KQRBNvK" };
    for (const char *nm : dm) {
      Mat M;
      if (setup_mat(M, nm) != TBG_OK || !M.all_single) ST_FAIL("setup %s", nm);
      for (int it = 0; it < 20000; it++) {
        int s = (int)(rnd() % NSL);
        uint64_t idx = rnd() % M.ii.size;
        uint64_t p = (uint64_t)T.kksq[s][0] | ((uint64_t)T.kksq[s][1] << 8);
        uint64_t occ = idx_to_sq(T, M.ii, idx, &p);
        int kj = (int)(rnd() % (uint64_t)M.ii.numsets);
        int a = sq_get(p, 2 + kj);
        uint64_t ob = occ_before(p, kj);
        uint32_t ra = (uint32_t)(a - popc64(ob & (bit(a) - 1)));
        for (int b = 0; b < 64; b++) {
          if (occ & bit(b)) continue;
          uint64_t want = sq_to_idx(T, M.ii, sq_set(p, 2 + kj, b));
          if (moved_idx_l(later_sets(M.ii, p, kj), idx, a, b, ob, ra) != want)
            ST_FAIL("%s: moved_idx", nm);
        }
      }
    }
  }
  // a king move that needs no symmetry: incremental index equals re-indexing
  {
    const char *km[] = { "KQRvKR", "KQRBvKQRN", "KRvKN" };
    for (const char *nm : km) {
      Mat M;
      if (setup_mat(M, nm) != TBG_OK || !M.all_single) ST_FAIL("setup %s", nm);
      int tested = 0;
      for (int it = 0; it < 20000; it++) {
        int sl = (int)(rnd() % NSL);
        uint64_t idx = rnd() % M.ii.size;
        uint64_t p = (uint64_t)T.kksq[sl][0] | ((uint64_t)T.kksq[sl][1] << 8);
        uint64_t occ = idx_to_sq(T, M.ii, idx, &p);
        int stm = (int)(rnd() & 1);
        int from = sq_get(p, stm);
        uint64_t b = T.king[from] & ~occ & ~T.king[sq_get(p, stm ^ 1)];
        while (b) {
          int to = ctz64(b);
          b &= b - 1;
          uint64_t q0 = sq_set(p, stm, to), q;
          if (normalize(T, q0, &q) < 0 || q != q0) continue;
          tested++;
          if (king_moved_idx_a(all_sets(M.ii, q0), idx, from, to) != sq_to_idx(T, M.ii, q0))
            ST_FAIL("%s: king_moved_idx_a", nm);
          if (sq_to_idx_a(all_sets(M.ii, q0), q0) != sq_to_idx(T, M.ii, q0))
            ST_FAIL("%s: sq_to_idx_a", nm);
          if (sq_to_idx_1(M.ii, q0) != sq_to_idx(T, M.ii, q0)) ST_FAIL("%s: sq_to_idx_1", nm);
          for (int l = 0; l < M.numsets; l++)
            if (capt_idx_1(M.capt[l], q0, l) != sq_to_idx(T, M.capt[l], q0))
              ST_FAIL("%s: capt_idx_1 set %d", nm, l);
          if (king_moved_idx(M.ii, idx, q0, from, to) != sq_to_idx(T, M.ii, q0))
            ST_FAIL("%s: king_moved_idx", nm);
        }
      }
      if (tested < 1000) ST_FAIL("%s: too few identity king moves (%d)", nm, tested);
    }
  }
  // pred blocks: a move of a piece of a set >= blk_set stays in its block
  {
    const char *bm[] = { "KQRvKRN", "KQRBvKQRN", "KRRvKQ", "KQQvKRR", "KBBNvKN" };
    for (const char *nm : bm) {
      Mat M;
      if (setup_mat(M, nm) != TBG_OK) ST_FAIL("setup %s", nm);
      uint64_t bp, nb;
      int bs;
      choose_block(M, &bp, &nb, &bs);
      if (bp * nb != M.ii.size) ST_FAIL("%s: blocks do not tile the slice", nm);
      for (int it = 0; it < 20000; it++) {
        int sl = (int)(rnd() % NSL);
        uint64_t idx = rnd() % M.ii.size;
        uint64_t p = (uint64_t)T.kksq[sl][0] | ((uint64_t)T.kksq[sl][1] << 8);
        uint64_t occ = idx_to_sq(T, M.ii, idx, &p);
        int j = 2 + (int)(rnd() % (uint64_t)(M.num - 2));
        if (sq_get(M.set_of, j) < bs) continue;
        for (int b = 0; b < 64; b++) {
          if (occ & bit(b)) continue;
          uint64_t i2 = sq_to_idx(T, M.ii, sq_set(p, j, b));
          if (i2 / bp != idx / bp) ST_FAIL("%s: a move of set %d leaves its block", nm, sq_get(M.set_of, j));
        }
      }
    }
  }
  if (msg) snprintf(msg, len, "ok");
  return 0;
#undef ST_FAIL
}

// ====================================================================
// Served form: continuous-batching job server
// ====================================================================
namespace {

// Preemption: a resident context's state block (bitmap planes, capture-exit
// planes, codes) goes to pinned host memory in one copy, with the job
// record, and returns to the pool; the copy back rebases the job record's
// plane pointers onto the new block.  The record block stays resident.
// Runs on the context's own stream (the server's set-up stream), with the
// batch idle.
struct CtxImage {
  void *host = nullptr;
  size_t bytes = 0;
  void *state = nullptr;  // the released block
};

int ctx_spill(tbg_ctx *c, CtxImage &im)
{
  size_t total = c->state_bytes;
  if (!im.host) {
    HIPCHK(hipHostMalloc(&im.host, total, hipHostMallocDefault));
    im.bytes = total;
  }
  HIPCHK(hipMemcpyAsync(im.host, c->arena, total, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipMemcpyAsync(&c->hj, c->dj, sizeof(JobDev), hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  im.state = c->arena;
  dfree(c, c->arena);
  c->arena = c->subarena = nullptr;
  return TBG_OK;
}

template <class X> X *rebase(X *p, const void *old, size_t n, void *now)
{
  const char *a = reinterpret_cast<const char *>(p), *o = static_cast<const char *>(old);
  if (!p || a < o || a >= o + n) return p;
  return reinterpret_cast<X *>(static_cast<char *>(now) + (a - o));
}

int ctx_restore_image(tbg_ctx *c, CtxImage &im)
{
  CtxLayout L = ctx_layout(c->h, c->nsl, c->max_ply, c->codes[0] != nullptr);
  size_t n = c->state_bytes;
  void *state = nullptr;
  TRY(dalloc(c, &state, n));
  HIPCHK(hipMemcpyAsync(state, im.host, n, hipMemcpyHostToDevice, c->st));
  // the canonical planes at their places in the new block, and the job
  // record's (swapped) plane pointers moved with them
  carve_state(c, state, L, c->codes[0] != nullptr);
  JobDev &J = c->hj;
  for (int s = 0; s < 2; s++) {
    J.wins[s] = rebase(J.wins[s], im.state, n, state);
    J.prevW[s] = rebase(J.prevW[s], im.state, n, state);
    J.prevL[s] = rebase(J.prevL[s], im.state, n, state);
    J.newW[s] = rebase(J.newW[s], im.state, n, state);
    J.newL[s] = rebase(J.newL[s], im.state, n, state);
    J.cand[s] = rebase(J.cand[s], im.state, n, state);
    J.candW[s] = rebase(J.candW[s], im.state, n, state);
    J.xr[s] = rebase(J.xr[s], im.state, n, state);
    for (int k = 0; k < SUB_KINDS; k++) J.sub[s][k] = rebase(J.sub[s][k], im.state, n, state);
    J.codes[s] = c->codes[s];
  }
  J.sum = c->psum;
  J.arena = c->arena;
  HIPCHK(hipMemcpyAsync(c->dj, &c->hj, sizeof(JobDev), hipMemcpyHostToDevice, c->st));
  return TBG_OK;
}

// one king slice of a served job
struct SrvSlice {
  int slice = 0;
  tbg_ctx *ctx = nullptr;
  // pinned host memory, allocated at submission: the final codes of each
  // side (pinned[stm]), the staging of the set-up copy (stage) and the
  // landing place of the finished record (fin: counts, flags, job record,
  // its pointer and history, as in the record block from o_cnt on)
  void *host = nullptr;
  uint16_t *pinned[2] = { nullptr, nullptr };
  char *stage = nullptr, *fin = nullptr;
  CtxImage image;  // its device state while preempted
  std::vector<long long> hist;
  tbg_counts counts;
  bool done = false;    // finished: record and codes copied (or copying) to pinned memory
  bool landed = false;  // those copies complete
  bool narrow = false;  // the pinned codes are one byte each (every ply below 64)
  hipEvent_t ev = nullptr;
};

struct SrvJob {
  std::string material;
  tbg_job_spec spec;
  tbg_job_result res;
  int state = 0;  // 0 waiting, 1 resident, 2 preempted, 3 done
  std::vector<SrvSlice> sl;
  uint64_t bytes = 0;  // admission estimate, all slices
  int left = 0;        // slices whose codes have not landed
};

// a served slice's two device blocks (a synthetic context keeping codes)
CtxLayout served_layout(const Mat &M) { return ctx_layout(M, 1, kMaxPly, true); }

uint64_t estimate_bytes(const char *material, int *err)
{
  Mat M;
  *err = setup_mat(M, material);
  if (*err) return 0;
  CtxLayout L = served_layout(M);
  return L.state + L.rec;
}

// pinned bytes of a served slice (codes of both sides, staging, landing)
size_t served_host_bytes(const Mat &M, size_t *o_stage, size_t *o_fin)
{
  CtxLayout L = served_layout(M);
  size_t cb = part_up(M.ii.size * 2);
  *o_stage = 2 * cb;
  *o_fin = *o_stage + part_up(L.rec);
  return *o_fin + part_up(L.rec - L.o_cnt);
}

}  // namespace

struct tbg_server {
  uint64_t budget = 0;
  int max_active = 1;
  Geom geom = g_geom;
  std::deque<SrvJob> jobs;  // appended by submit (references stay valid)
  std::mutex mu;            // jobs' length, states and results; closed; serving
  std::condition_variable cv;
  bool closed = false, serving = false, started = false;
  std::chrono::steady_clock::time_point t0;
  // st: the batch; side: codes of finished slices; setup: job allocation,
  // init, preemption and restore
  hipStream_t st = nullptr, side = nullptr, setup = nullptr;
  double now_ms() const
  {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
};

namespace {

// the serving loop (tbg_server_serve)
struct Serve {
  tbg_server *s;
  bool plain;
  int max_slices;
  std::vector<int> waiting;                    // jobs in arrival order
  std::vector<std::pair<int, int>> active;     // (job, slice) in the batch
  std::vector<std::pair<int, int>> joining;    // set-up or restore queued
  std::vector<std::pair<int, int>> spilling;   // record and codes on their way to pinned memory
  std::vector<int> preempted;                  // jobs on the host, oldest first
  uint64_t used = 0, peak = 0;
  int peak_active = 0;
  double occ_integral = 0, last_t = 0;
  uint64_t plies = 0, done = 0, rejected = 0, spills = 0, restores = 0;
  double psp = 0, spill_ms = 0;

  SrvJob &J(int i) { return s->jobs[i]; }
  int slots_used() const { return (int)(active.size() + joining.size()); }

  void complete(int ji, int status)
  {
    std::lock_guard<std::mutex> lk(s->mu);
    SrvJob &j = J(ji);
    j.res.status = status;
    j.res.done_ms = s->now_ms();
    j.state = 3;
    s->cv.notify_all();
  }

  // set-up of one slice: two pool allocations, one copy, the capture-exit
  // and init kernels, all queued on the set-up stream
  int admit_slice(int ji, int k)
  {
    SrvJob &j = J(ji);
    SrvSlice &q = j.sl[k];
    tbg_ctx *c = nullptr;
    TRY(create_common(&c, j.material.c_str(), 1, q.slice, j.spec.seed, true, s->setup));
    q.ctx = c;
    TRY(setup_pooled(c, q.stage));
    HIPCHK(hipEventCreateWithFlags(&c->ready, hipEventDisableTiming));
    HIPCHK(hipEventRecord(c->ready, s->setup));
    used += c->bytes;
    joining.push_back({ ji, k });
    return TBG_OK;
  }

  int admit(int ji)
  {
    SrvJob &j = J(ji);
    j.res.admit_ms = s->now_ms();
    j.res.device_bytes = 0;
    for (int k = 0; k < (int)j.sl.size(); k++) {
      TRY(admit_slice(ji, k));
      j.res.device_bytes += j.sl[k].ctx->bytes;
      j.res.positions = j.sl[k].ctx->h.ii.size;
    }
    peak = used > peak ? used : peak;
    std::lock_guard<std::mutex> lk(s->mu);
    j.state = 1;
    return TBG_OK;
  }

  // with the batch idle: job ji's unfinished slices' state blocks to pinned
  // host memory
  int preempt(int ji)
  {
    SrvJob &j = J(ji);
    double t = s->now_ms();
    for (int k = 0; k < (int)j.sl.size(); k++) {
      SrvSlice &q = j.sl[k];
      if (q.done) continue;
      TRY(ctx_spill(q.ctx, q.image));
      used -= q.ctx->state_bytes;
      active.erase(std::remove(active.begin(), active.end(), std::make_pair(ji, k)), active.end());
    }
    j.res.spills++;
    spills++;
    spill_ms += s->now_ms() - t;
    preempted.push_back(ji);
    std::lock_guard<std::mutex> lk(s->mu);
    j.state = 2;
    return TBG_OK;
  }

  int restore(int ji)
  {
    SrvJob &j = J(ji);
    double t = s->now_ms();
    for (int k = 0; k < (int)j.sl.size(); k++) {
      SrvSlice &q = j.sl[k];
      if (q.done) continue;
      TRY(ctx_restore_image(q.ctx, q.image));
      HIPCHK(hipEventRecord(q.ctx->ready, s->setup));
      used += q.ctx->state_bytes;
      joining.push_back({ ji, k });
    }
    peak = used > peak ? used : peak;
    restores++;
    spill_ms += s->now_ms() - t;
    std::lock_guard<std::mutex> lk(s->mu);
    j.state = 1;
    return TBG_OK;
  }

  // a slice the batch finished (its done word from k_ctrl: the done value
  // and the last ply that found positions): its record and codes go to
  // pinned memory on the side stream, narrowed to a byte each when every
  // ply is below 64 (on the device, into the no longer needed planes, one
  // side after the other: the side stream orders each narrowing after the
  // previous copy).  The batch has completed the plies that touched it.
  int leave(std::pair<int, int> pr, unsigned word)
  {
    SrvSlice &q = J(pr.first).sl[pr.second];
    tbg_ctx *c = q.ctx;
    CtxLayout L = ctx_layout(c);
    uint64_t n = c->h.ii.size;
    int maxn = (int)(word >> 2);
    HIPCHK(hipMemcpyAsync(q.fin, reinterpret_cast<char *>(c->rec) + L.o_cnt, L.rec - L.o_cnt,
                          hipMemcpyDeviceToHost, s->side));
    q.narrow = maxn < 64 && L.arena >= n;
    for (int stm = 0; stm < 2; stm++) {
      if (q.narrow) {
        uint8_t *tmp = reinterpret_cast<uint8_t *>(c->arena);
        hipLaunchKernelGGL(k_codes8, dim3(cap_grid((n + 2047) / 2048, (uint64_t)g_cus * 8)),
                           dim3(256), 0, s->side, c->codes[stm], n, tmp);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(q.pinned[stm], tmp, n, hipMemcpyDeviceToHost, s->side));
      } else {
        HIPCHK(hipMemcpyAsync(q.pinned[stm], c->codes[stm], n * 2, hipMemcpyDeviceToHost, s->side));
      }
    }
    if (!q.ev) HIPCHK(hipEventCreateWithFlags(&q.ev, hipEventDisableTiming));
    HIPCHK(hipEventRecord(q.ev, s->side));
    q.done = true;
    spilling.push_back(pr);
    return TBG_OK;
  }

  // a left slice whose copies have completed: its account from the landed
  // record, its device memory back to the pool
  int land(std::pair<int, int> pr)
  {
    SrvJob &j = J(pr.first);
    SrvSlice &q = j.sl[pr.second];
    tbg_ctx *c = q.ctx;
    CtxLayout L = ctx_layout(c);
    const char *f = q.fin;
    const unsigned long long *cnt = reinterpret_cast<const unsigned long long *>(f);
    memcpy(&c->hj, f + (L.o_j - L.o_cnt), sizeof(JobDev));
    const long long *h = reinterpret_cast<const long long *>(f + (L.o_hist - L.o_cnt));
    c->hist.assign(h, h + (size_t)4 * c->max_ply);
    TRY(finish_counts(c));
    for (int i = 0; i < 2; i++) {
      c->counts.broken[i] = cnt[i];
      c->counts.mates[i] = cnt[2 + i];
    }
    q.hist = c->hist;
    q.counts = c->counts;
    j.res.nplies = std::max(j.res.nplies, c->counts.nplies);
    uint64_t hs = executed_half_steps(c);
    j.res.half_steps += hs;
    psp += (double)hs * (double)c->h.ii.size / 2.0;
    used -= c->bytes;
    tbg_destroy(c);
    q.ctx = nullptr;
    q.landed = true;
    if (--j.left == 0) {
      done++;
      complete(pr.first, TBG_OK);
    }
    return TBG_OK;
  }

  // admitted jobs larger than waiting job wi whose release lets it in
  // (largest first; a job is preempted at most twice), or none
  std::vector<int> victims(int wi)
  {
    std::vector<int> cand;
    SrvJob &w = J(wi);
    for (auto &a : active) {
      SrvJob &j = J(a.first);
      if (j.state != 1 || j.bytes <= w.bytes || j.res.spills >= 2) continue;
      if (std::find(cand.begin(), cand.end(), a.first) != cand.end()) continue;
      bool all_active = true;  // none of its slices still joining
      for (auto &q : joining) all_active &= q.first != a.first;
      if (all_active) cand.push_back(a.first);
    }
    std::sort(cand.begin(), cand.end(), [&](int x, int y) { return J(x).bytes > J(y).bytes; });
    std::vector<int> out;
    uint64_t freed = 0;
    for (int v : cand) {
      if (used - freed + w.bytes <= s->budget) break;
      out.push_back(v);
      for (auto &q : J(v).sl)
        if (!q.done) freed += q.ctx->state_bytes;
    }
    if (used - freed + w.bytes > s->budget) out.clear();
    return out;
  }
};

// The running batch's device arrays in one block, written in one copy per
// rebuild from a pinned mirror: the job pointers, the flags, the (job,
// side) pairs and the four prefix-sum arrays of the work units; the done
// words (written by k_ctrl) follow, outside the copy.
struct BatchArrays {
  size_t o_flags = 0, o_pairs = 0, o_P = 0, pstride = 0, o_done = 0, bytes = 0;
  char *dev = nullptr, *host = nullptr;
  unsigned *hflags = nullptr, *hdone = nullptr;  // pinned: what k_ctrl left

  int alloc(int m)
  {
    o_flags = part_up(sizeof(JobDev *) * m);
    o_pairs = o_flags + kPart;
    pstride = part_up(sizeof(uint64_t) * (2 * m + 1));
    o_P = o_pairs + part_up(sizeof(Pair) * 2 * m);
    o_done = o_P + 4 * pstride;
    bytes = o_done + part_up(sizeof(unsigned) * m);
    HIPCHK(hipMalloc(&dev, bytes));
    HIPCHK(hipHostMalloc((void **)&host, bytes, hipHostMallocDefault));
    HIPCHK(hipHostMalloc((void **)&hflags, 4 * sizeof(unsigned), hipHostMallocDefault));
    HIPCHK(hipHostMalloc((void **)&hdone, sizeof(unsigned) * m, hipHostMallocDefault));
    return TBG_OK;
  }
  void release()
  {
    (void)hipFree(dev);
    (void)hipHostFree(host);
    (void)hipHostFree(hflags);
    (void)hipHostFree(hdone);
  }
  JobDev **djobs() const { return reinterpret_cast<JobDev **>(dev); }
  unsigned *dflags() const { return reinterpret_cast<unsigned *>(dev + o_flags); }
  unsigned *ddone() const { return reinterpret_cast<unsigned *>(dev + o_done); }

  // the launch set of the slices `ctx` (both sides of each) planned into
  // the mirror and copied up on stream st; L's arrays point into the block
  int build(LaunchSet &L, const std::vector<tbg_ctx *> &ctx, const Geom &g, hipStream_t st)
  {
    int na = (int)ctx.size();
    std::vector<const JobDev *> hj(na);
    std::vector<const Mat *> mats(na);
    Pair *pairs = reinterpret_cast<Pair *>(host + o_pairs);
    JobDev **ptrs = reinterpret_cast<JobDev **>(host);
    for (int i = 0; i < na; i++) {
      ptrs[i] = ctx[i]->dj;
      hj[i] = &ctx[i]->hj;
      mats[i] = &ctx[i]->h;
      pairs[2 * i] = { i, 0 };
      pairs[2 * i + 1] = { i, 1 };
    }
    unsigned f[4] = { 0, 1, 0, 0 };  // no job done; a new slice may be at ply 1
    memcpy(host + o_flags, f, sizeof f);
    // plan_launchset writes the four prefix arrays back to back; each goes
    // to its own aligned place
    std::vector<uint64_t> P(4 * (2 * (size_t)na + 1));
    plan_launchset(L, hj.data(), mats.data(), na, pairs, 2 * na, 0, g, P.data());
    for (int kt = 0; kt < 4; kt++)
      memcpy(host + o_P + kt * pstride, P.data() + (size_t)kt * (2 * na + 1),
             sizeof(uint64_t) * (2 * na + 1));
    L.pairs = reinterpret_cast<Pair *>(dev + o_pairs);
    for (int kt = 0; kt < 4; kt++) L.P[kt] = reinterpret_cast<uint64_t *>(dev + o_P + kt * pstride);
    L.cap = 0;  // the arrays belong to this block: LaunchSet::release leaves them
    HIPCHK(hipMemcpyAsync(dev, host, o_done, hipMemcpyHostToDevice, st));
    return TBG_OK;
  }
};

}  // namespace

extern "C" {

int tbg_server_create(tbg_server **out, uint64_t budget_bytes, int max_active)
{
  if (!out || max_active < 1 || max_active > 1024) return fail(TBG_E_INVAL, "bad server arguments");
  if (g_device < 0) return fail(TBG_E_STATE, "tbg_init not called");
  tbg_server *s = new tbg_server();
  s->budget = budget_bytes;
  s->max_active = max_active;
  hipError_t e1 = hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking);
  hipError_t e2 = hipStreamCreateWithFlags(&s->side, hipStreamNonBlocking);
  hipError_t e3 = hipStreamCreateWithFlags(&s->setup, hipStreamNonBlocking);
  if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess) {
    tbg_server_destroy(s);
    return fail(TBG_E_HIP, "server streams");
  }
  // job memory comes from the device's stream-ordered pool; keep what it
  // has reserved rather than returning it at every synchronisation
  hipMemPool_t pool;
  if (hipDeviceGetDefaultMemPool(&pool, g_device) == hipSuccess) {
    uint64_t keep = ~0ull;
    (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
  }
  *out = s;
  return TBG_OK;
}

void tbg_server_destroy(tbg_server *s)
{
  if (!s) return;
  if (s->st) (void)hipStreamSynchronize(s->st);
  if (s->side) (void)hipStreamSynchronize(s->side);
  if (s->setup) (void)hipStreamSynchronize(s->setup);
  for (auto &j : s->jobs)
    for (auto &q : j.sl) {
      if (q.ctx) tbg_destroy(q.ctx);
      if (q.host) (void)hipHostFree(q.host);
      if (q.image.host) (void)hipHostFree(q.image.host);
      if (q.ev) (void)hipEventDestroy(q.ev);
    }
  if (s->setup) (void)hipStreamSynchronize(s->setup);  // the pooled frees
  if (s->st) (void)hipStreamDestroy(s->st);
  if (s->side) (void)hipStreamDestroy(s->side);
  if (s->setup) (void)hipStreamDestroy(s->setup);
  delete s;
}

int tbg_server_submit(tbg_server *s, const tbg_job_spec *spec, uint64_t *job_id)
{
  if (!s || !spec || !spec->material) return fail(TBG_E_INVAL, "null argument");
  if (spec->slice < 0 || spec->slice >= 441) return fail(TBG_E_INVAL, "slice must be in [0, 441)");
  int ns = spec->nslices ? spec->nslices : 1;
  if (ns < 1 || ns > TBG_JOB_MAX_SLICES)
    return fail(TBG_E_INVAL, "a job has 1 to %d king slices", TBG_JOB_MAX_SLICES);
  Mat M;
  int err = setup_mat(M, spec->material);
  if (err) return err;
  SrvJob j;
  j.material = spec->material;
  j.spec = *spec;
  j.spec.material = nullptr;
  j.spec.nslices = ns;
  memset(&j.res, 0, sizeof j.res);
  j.res.nslices = ns;
  j.bytes = estimate_bytes(spec->material, &err) * (uint64_t)ns;
  j.left = ns;
  j.sl.resize(ns);
  // each slice's pinned host memory: its codes, the set-up staging, the
  // landing place of its final record (a pinned allocation takes a
  // fraction of a millisecond, so it is made here, not while serving)
  size_t o_stage, o_fin, hb = served_host_bytes(M, &o_stage, &o_fin);
  size_t cb = part_up(M.ii.size * 2);
  auto free_host = [&]() {
    for (auto &q : j.sl)
      if (q.host) (void)hipHostFree(q.host);
  };
  for (int k = 0; k < ns; k++) {
    SrvSlice &q = j.sl[k];
    q.slice = (spec->slice + k) % 441;
    if (hipHostMalloc(&q.host, hb, hipHostMallocDefault) != hipSuccess) {
      q.host = nullptr;
      free_host();
      return fail(TBG_E_NOMEM, "pinned host memory for a job's results");
    }
    char *h = static_cast<char *>(q.host);
    q.pinned[0] = reinterpret_cast<uint16_t *>(h);
    q.pinned[1] = reinterpret_cast<uint16_t *>(h + cb);
    q.stage = h + o_stage;
    q.fin = h + o_fin;
  }
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->closed) {
    free_host();
    return fail(TBG_E_STATE, "the server takes no more jobs");
  }
  double a = spec->arrival_ms;
  if (a < 0) a = s->started ? s->now_ms() : 0.0;
  j.spec.arrival_ms = j.res.arrival_ms = a;
  s->jobs.push_back(std::move(j));
  if (job_id) *job_id = s->jobs.size() - 1;
  s->cv.notify_all();
  return TBG_OK;
}

int tbg_server_close(tbg_server *s)
{
  if (!s) return fail(TBG_E_INVAL, "null server");
  std::lock_guard<std::mutex> lk(s->mu);
  s->closed = true;
  s->cv.notify_all();
  return TBG_OK;
}

int tbg_server_serve(tbg_server *s, int form, tbg_server_stats *stats)
{
  if (!s) return fail(TBG_E_INVAL, "null server");
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->serving) return fail(TBG_E_STATE, "the server is already serving");
    s->serving = true;
    if (!s->started) {
      s->t0 = std::chrono::steady_clock::now();
      s->started = true;
    }
  }
  Serve v;
  v.s = s;
  v.plain = form == TBG_FORM_PLAIN;
  // the plain server runs one slice at a time; both check the device every
  // kPlies plies (a slice that finishes meanwhile idles, its kernels skip it)
  v.max_slices = v.plain ? 1 : s->max_active;
  const int kPlies = 4;
  size_t seen = 0;
  BatchArrays B;
  hipEvent_t evb = nullptr;
  int r = B.alloc(v.max_slices);
  if (r == TBG_OK && hipEventCreateWithFlags(&evb, hipEventDisableTiming) != hipSuccess)
    r = fail(TBG_E_HIP, "server event");
  LaunchSet L;
  bool changed = true, inflight = false;
  int blocked = -1;  // an arrived job that waits for memory
  // The loop keeps the device busy: the next group of plies is queued as
  // soon as the previous one has ended and the batch is updated; the host's
  // bookkeeping (landed results, admission, set-up) runs meanwhile.
  while (r == TBG_OK) {
    // 0. jobs submitted since the last look join the waiting list
    bool closed;
    {
      std::lock_guard<std::mutex> lk(s->mu);
      for (; seen < s->jobs.size(); seen++) {
        double a = s->jobs[seen].spec.arrival_ms;
        auto it = std::upper_bound(v.waiting.begin(), v.waiting.end(), a,
                                   [&](double x, int ji) { return x < s->jobs[ji].spec.arrival_ms; });
        v.waiting.insert(it, (int)seen);
      }
      closed = s->closed;
    }
    if (closed && !inflight && v.waiting.empty() && v.active.empty() && v.joining.empty() &&
        v.spilling.empty() && v.preempted.empty())
      break;
    double t = s->now_ms();
    v.occ_integral += (double)v.used * (t - v.last_t);
    v.last_t = t;
    // 1. the group in flight has ended: finished slices leave the batch,
    //    their record and codes on their way to pinned memory
    if (inflight) {
      HIPCHK(hipEventSynchronize(evb));
      inflight = false;
      if (B.hflags[0]) {
        std::vector<std::pair<int, int>> keep;
        for (int i = 0; i < (int)v.active.size() && r == TBG_OK; i++) {
          if (B.hdone[i] & 3) r = v.leave(v.active[i], B.hdone[i]);
          else keep.push_back(v.active[i]);
        }
        if (r != TBG_OK) break;
        v.active = keep;
        changed = true;
      }
    }
    // 2. a blocked job preempts larger admitted jobs, the batch being idle
    if (blocked >= 0 && !v.plain) {
      std::vector<int> vic = v.victims(blocked);
      if (!vic.empty()) {
        for (int ji : vic) {
          r = v.preempt(ji);
          if (r != TBG_OK) break;
        }
        if (r != TBG_OK) break;
        changed = true;
        v.waiting.erase(std::find(v.waiting.begin(), v.waiting.end(), blocked));
        r = v.admit(blocked);
        if (r != TBG_OK) break;
      }
    }
    blocked = -1;
    // 3. slices whose set-up (or restore) completed join the batch (the
    //    plain server: one slice at a time)
    for (size_t i = 0; i < v.joining.size();) {
      if (v.plain && !v.active.empty()) break;
      auto pr = v.joining[i];
      if (hipEventQuery(v.J(pr.first).sl[pr.second].ctx->ready) == hipSuccess) {
        v.active.push_back(pr);
        v.joining.erase(v.joining.begin() + i);
        changed = true;
      } else {
        i++;
      }
    }
    v.peak_active = (int)v.active.size() > v.peak_active ? (int)v.active.size() : v.peak_active;
    // 4. the next group: the batch rebuilt when its membership changed, then
    //    kPlies plies queued and the flags and done words read back
    int na = (int)v.active.size();
    if (na) {
      if (changed) {
        std::vector<tbg_ctx *> ctx;
        for (auto &pr : v.active) ctx.push_back(v.J(pr.first).sl[pr.second].ctx);
        r = B.build(L, ctx, s->geom, s->st);
        if (r != TBG_OK) break;
        changed = false;
      }
      for (int k = 0; k < kPlies && r == TBG_OK; k++) {
        r = launch_ply(nullptr, B.djobs(), na, L, B.dflags(), s->st, true, B.ddone(), true);
        v.plies++;
      }
      if (r != TBG_OK) break;
      HIPCHK(hipMemcpyAsync(B.hflags, B.dflags(), 4 * sizeof(unsigned), hipMemcpyDeviceToHost, s->st));
      HIPCHK(hipMemcpyAsync(B.hdone, B.ddone(), sizeof(unsigned) * na, hipMemcpyDeviceToHost, s->st));
      HIPCHK(hipEventRecord(evb, s->st));
      inflight = true;
    }
    // 5. host work while those plies run: landed results, admission (set-up
    //    and init queued on the set-up stream), restores
    for (size_t i = 0; i < v.spilling.size() && r == TBG_OK;) {
      auto pr = v.spilling[i];
      if (hipEventQuery(v.J(pr.first).sl[pr.second].ev) != hipSuccess) { i++; continue; }
      v.spilling.erase(v.spilling.begin() + i);
      r = v.land(pr);
    }
    if (r != TBG_OK) break;
    while (!v.waiting.empty()) {
      int ji = v.waiting.front();
      SrvJob &j = v.J(ji);
      if (j.spec.arrival_ms > s->now_ms()) break;
      if (j.bytes > s->budget) {
        v.waiting.erase(v.waiting.begin());
        v.rejected++;
        v.complete(ji, TBG_E_BUDGET);
        continue;
      }
      int ns = (int)j.sl.size();
      bool slots = v.plain ? v.active.empty() && v.joining.empty() && v.preempted.empty()
                           : v.slots_used() + ns <= v.max_slices;
      if (!slots) break;
      if (v.used + j.bytes > s->budget) {
        blocked = ji;
        break;
      }
      v.waiting.erase(v.waiting.begin());
      r = v.admit(ji);
      if (r != TBG_OK) break;
    }
    if (r != TBG_OK) break;
    // preempted jobs come back when they fit and no arrived job is blocked
    while (!v.plain && blocked < 0 && !v.preempted.empty()) {
      int ji = v.preempted.front();
      uint64_t need = 0;
      int ns = 0;
      for (auto &q : v.J(ji).sl)
        if (!q.done) need += q.ctx->state_bytes, ns++;
      if (v.used + need > s->budget || v.slots_used() + ns > v.max_slices) break;
      v.preempted.erase(v.preempted.begin());
      r = v.restore(ji);
      if (r != TBG_OK) break;
    }
    if (r != TBG_OK) break;
    // 6. nothing running: wait for a set-up, else a copy, else the next
    //    arrival or submission
    if (!inflight) {
      if (!v.joining.empty()) {
        auto pr = v.joining.front();
        HIPCHK(hipEventSynchronize(v.J(pr.first).sl[pr.second].ctx->ready));
      } else if (!v.spilling.empty()) {
        HIPCHK(hipStreamSynchronize(s->side));
      } else if (blocked < 0) {
        std::unique_lock<std::mutex> lk(s->mu);
        double wait = 50.0;  // look again for submissions at least this often
        if (!v.waiting.empty())
          wait = std::min(wait, s->jobs[v.waiting.front()].spec.arrival_ms - s->now_ms());
        if (wait > 0 && seen == s->jobs.size() && !s->closed)
          s->cv.wait_for(lk, std::chrono::microseconds((int64_t)(wait * 1000)));
        else if (wait > 0 && !v.waiting.empty())
          std::this_thread::sleep_for(std::chrono::microseconds((int64_t)(wait * 1000)));
      }
    }
  }
  (void)hipStreamSynchronize(s->st);
  (void)hipStreamSynchronize(s->side);
  (void)hipStreamSynchronize(s->setup);
  B.release();
  if (evb) (void)hipEventDestroy(evb);
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->serving = false;
  }
  if (r != TBG_OK) return r;
  if (stats) {
    stats->wall_ms = s->now_ms();
    stats->jobs_done = v.done;
    stats->jobs_rejected = v.rejected;
    stats->mean_resident_bytes = v.last_t > 0 ? v.occ_integral / v.last_t : 0;
    stats->peak_resident_bytes = v.peak;
    stats->peak_active = v.peak_active;
    stats->plies = v.plies;
    stats->position_side_plies = v.psp;
    stats->spills = v.spills;
    stats->restores = v.restores;
    stats->spill_ms = v.spill_ms;
  }
  return TBG_OK;
}

int tbg_server_run(tbg_server *s, int form, tbg_server_stats *stats)
{
  TRY(tbg_server_close(s));
  return tbg_server_serve(s, form, stats);
}

int tbg_job_poll(tbg_server *s, uint64_t id, tbg_job_result *out)
{
  if (!s || !out) return fail(TBG_E_INVAL, "null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  if (id >= s->jobs.size()) return fail(TBG_E_INVAL, "bad job id");
  if (s->jobs[id].state != 3) return TBG_PENDING;
  *out = s->jobs[id].res;
  return TBG_OK;
}

int tbg_server_result(tbg_server *s, uint64_t id, tbg_job_result *out)
{
  if (!s || !out) return fail(TBG_E_INVAL, "null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  if (id >= s->jobs.size()) return fail(TBG_E_INVAL, "bad job id");
  *out = s->jobs[id].res;
  return TBG_OK;
}

int tbg_server_codes_part(tbg_server *s, uint64_t id, int part, int stm, uint16_t *dst, uint64_t n)
{
  if (!s || !dst || (stm & ~1)) return fail(TBG_E_INVAL, "bad argument");
  std::lock_guard<std::mutex> lk(s->mu);
  if (id >= s->jobs.size()) return fail(TBG_E_INVAL, "bad job id");
  SrvJob &j = s->jobs[id];
  if (j.state != 3 || j.res.status != TBG_OK) return fail(TBG_E_STATE, "job has no result");
  if (part < 0 || part >= (int)j.sl.size()) return fail(TBG_E_INVAL, "no such slice of the job");
  if (n > j.res.positions) return fail(TBG_E_INVAL, "too many codes");
  const SrvSlice &q = j.sl[part];
  if (q.narrow) {  // one byte per code on the host: widened here
    const uint8_t *b = reinterpret_cast<const uint8_t *>(q.pinned[stm]);
    for (uint64_t i = 0; i < n; i++) dst[i] = b[i];
  } else {
    memcpy(dst, q.pinned[stm], n * 2);
  }
  return TBG_OK;
}

int tbg_server_codes(tbg_server *s, uint64_t id, int stm, uint16_t *dst, uint64_t n)
{
  return tbg_server_codes_part(s, id, 0, stm, dst, n);
}

int tbg_server_lines_part(tbg_server *s, uint64_t id, int part, tbg_line_fn fn, void *user)
{
  if (!s || !fn) return fail(TBG_E_INVAL, "bad argument");
  tbg_ctx tmp;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (id >= s->jobs.size()) return fail(TBG_E_INVAL, "bad job id");
    SrvJob &j = s->jobs[id];
    if (j.state != 3 || j.res.status != TBG_OK) return fail(TBG_E_STATE, "job has no result");
    if (part < 0 || part >= (int)j.sl.size()) return fail(TBG_E_INVAL, "no such slice of the job");
    tmp.counts = j.sl[part].counts;
    tmp.hist = j.sl[part].hist;
  }
  emit_lines(&tmp, fn, user);
  return TBG_OK;
}

int tbg_server_lines(tbg_server *s, uint64_t id, tbg_line_fn fn, void *user)
{
  return tbg_server_lines_part(s, id, 0, fn, user);
}

}  // extern "C"
