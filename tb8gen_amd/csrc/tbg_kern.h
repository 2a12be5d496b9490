// tbg_kern.h -- kernel entry points and the capture-table reference shared
// by the kernel and host translation units.
#pragma once
#include "tbg.h"
#include "tbg_dev.h"

namespace tbg {

// A finished smaller-material table that capture exits of set k land in.
struct SubRef {
  const uint16_t *codes[2];  // [462 * size] codes per side to move
  Idx ii;                    // its index layout
  uint64_t map;              // packed: its slot j <- our slot map[j]
  int num;                   // its piece count
  int flip;                  // 1: colours swapped (probe.c:1554-1565)
  int present;               // 0: a bare pair of kings (always a draw)
};

// Capture-exit planes per colour: the four of the C ABI (TBG_SUB_LOSS,
// _BLOSS, _CWIN, _WIN; an illegal result counts as a win, a capture into it
// being illegal) and the legal wins alone, which the ply-1 uncapture of the
// L step starts from (the wins plane is about half illegal positions).
constexpr int SUB_LWIN = 4;
constexpr int SUB_KINDS = 5;

// Device-resident state of one generation job (a material, or one king
// slice of one).  The kernels read the ply, the step mask and the plane
// pointers from here, so a whole ply can be captured in a hipGraph and
// replayed; k_ctrl advances it (the more_* logic of src/tb8gen.c:1406-1444).
struct JobDev {
  const Mat *M;
  uint64_t *wins[2], *prevW[2], *prevL[2], *newW[2], *newL[2];
  uint64_t *cand[2], *candW[2];  // candidates of the L and of the W step
  uint64_t *sub[2][SUB_KINDS];
  // static refutations of the L step (k_init_pl + k_xr_pl): a position with a
  // capture or synthetic exit whose result is not a win of the opponent
  // (within the 50-move rule), and for a synthetic job on the placement
  // layout the mates, lost already and never an L_n candidate; nullptr: not
  // worked out, verify every move
  uint64_t *xr[2];
  // per side: positions neither won (illegal) at init nor statically refuted,
  // what a pull-form L step verifies at most (k_xr_pl counts them)
  const unsigned long long *nopen;
  uint16_t *codes[2];
  long long *hist;            // [max_ply][4] counts per step, -1: step did not run
  unsigned long long cnt[4];  // this ply: L(W), L(B), W(W), W(B)
  int n, done;                // done: 1 finished, 2 paused at stop_at
  int maxn;                   // the last ply whose steps found positions (0: none)
  unsigned mask;              // steps of this ply: bit 0 L(W), 1 L(B), 2 W(W), 3 W(B)
  unsigned more;              // bit 0 more_ww, 1 more_wb, 2 more_lw, 3 more_lb
  unsigned pull;              // bit s: side s's L step of this ply in the pull form
  int scatter;                // TBG_SCATTER_*
  int max_ply;
  int stop_at;                // > 0: pause (done = 2) on reaching this ply
  uint64_t blk_pos, nblk, blk_words;
  int blk_set;                // pieces of sets >= blk_set move within a pred block
  int blk_parts;              // pred units per block (> 1 only for one block per slice)
  // Chunk summaries: one byte per 64 bitmap words of each of the arena's
  // planes (nonzero: a set bit may be there), plane after plane in arena
  // order, spw bytes per slice.  Every ply keeps those of the planes it
  // writes; in a sparse ply (sparse != 0) the kernels read and write only
  // the chunks whose summaries are set, and pred sets the summaries of the
  // candidate chunks it marks.
  uint8_t *sum;
  const uint64_t *arena;      // the planes' base; pw words per plane
  uint64_t pw, spw, sum_pl;   // spw: summary bytes per slice; sum_pl: per plane
  int sparse;
  int sum_plies;              // plies since the summaries were last unknown
  // the ply-1 W step's capture candidates are in candW already (k_xr_pl of a
  // synthetic job worked them out), so that step runs no uncapture
  int w1cap;
};

// the summary bytes of plane p (one of the job's arena planes)
TBG_HD uint8_t *plane_sum(const JobDev &J, const uint64_t *p)
{
  uint64_t k = (uint64_t)(p - J.arena) / J.pw;
  return J.sum + k * J.sum_pl;
}

// One (job, side to move) pair of a launch.
struct Pair {
  int job, side;
};

// Work units per slice of each launch type, one unit per work-group at a
// time: pred = one part of an index block (nblk = blocks x parts here);
// uncap = 256 bitmap words (64 per wave);
// verify and sweep = a number of words chosen by the launch geometry.
enum { KT_PRED = 0, KT_UNCAP = 1, KT_VERIFY = 2, KT_SWEEP = 3 };
constexpr uint64_t UNIT_WORDS = 256;

// Launch geometry of the ply kernels (tbg_geometry of the C-ABI).
struct Geom {
  uint32_t verify_uw;  // words per verify unit, multiple of 256
  uint32_t sweep_uw;   // words per sweep unit, multiple of 512
  uint32_t wg[4];      // grid cap per kernel type, work-groups per CU
};

// uncapture units: the capture spaces of the opponent's sets, back to back
TBG_HD uint64_t uncap_units(const Mat &M, int side)
{
  uint64_t u = 0;
  for (int k = 0; k < M.numsets; k++)
    if (M.setcol[k] != side) u += (M.cwords[k] + UNIT_WORDS - 1) / UNIT_WORDS;
  return u;
}

// Verification units of a slice of at least VERIFY_BIG_WORDS bitmap words
// (7 and 8 pieces) are at least VERIFY_BIG_UNIT words: most of its plies
// have few candidates, and a unit with none is then read with many loads in
// flight and its L words cleared in one go (small units kept one load per
// wave in flight).  Such a slice still has over a thousand units; a smaller
// one keeps the geometry's unit, or a launch would have too few units to
// fill the chip.  Only the two-pass form (slices above VERIFY_SPLIT_SIZE
// positions) handles large units.
constexpr uint64_t VERIFY_SPLIT_SIZE = 1ull << 22;  // positions per slice
constexpr uint32_t VERIFY_BIG_UNIT = 4096;
constexpr uint64_t VERIFY_BIG_WORDS = 1ull << 22;
TBG_HD uint32_t verify_unit(const Mat &M, uint32_t uw)
{
  return M.words >= VERIFY_BIG_WORDS && uw < VERIFY_BIG_UNIT ? VERIFY_BIG_UNIT : uw;
}

TBG_HD uint64_t units_per_slice(const Mat &M, uint64_t nblk, int kt, int side, const Geom &g)
{
  switch (kt) {
  case KT_PRED: return nblk;
  case KT_UNCAP: return uncap_units(M, side);
  case KT_VERIFY: {
    uint32_t uw = verify_unit(M, g.verify_uw);
    return (M.words + uw - 1) / uw;
  }
  default: return (M.words + g.sweep_uw - 1) / g.sweep_uw;
  }
}

constexpr int PB_THREADS = 512;
constexpr int PB_QUEUE_BYTES = 8 * 1024 * 2;

// _sum forms and SUM = true: the kernels that read and write chunk
// summaries (a launch with some job that keeps them); the others leave
// them out, registers and all
__global__ void k_pred_j(JobDev *const *jobs, const Pair *pairs, const uint64_t *P, int np, int sw0,
                         int kind);
__global__ void k_pred_sum_j(JobDev *const *jobs, const Pair *pairs, const uint64_t *P, int np,
                             int sw0, int kind);
__global__ void k_pred_wide_j(JobDev *const *jobs, const Pair *pairs, const uint64_t *P, int np,
                              int sw0, int kind);
__global__ void k_pred_wide_sum_j(JobDev *const *jobs, const Pair *pairs, const uint64_t *P, int np,
                                  int sw0, int kind);
__global__ void k_uncap_j(JobDev *const *jobs, const Pair *pairs, const uint64_t *P, int np, int sw0,
                          int kind, const unsigned *flags, uint64_t lds_words);
template <bool SPLIT, bool SUM>
__global__ __launch_bounds__(256, SPLIT ? 4 : 5) void k_verify_j(JobDev *const *jobs, const Pair *pairs,
                                                   const uint64_t *P, int np, int sw0, int uw);
template <bool SUM>
__global__ void k_sweep_j(JobDev *const *jobs, const Pair *pairs, const uint64_t *P, int np, int sw0,
                          int uw);
// flags[0]: jobs done; flags[1]: some job is at ply 1 or 101 (uncapture)
// done_out (nullable): done_out[j] = job j's done value | its maxn << 2
__global__ void k_ctrl(JobDev *const *jobs, int nj, unsigned *flags, unsigned *done_out);
// k_init: consecutive bitmap words per wave (positions step by 64 between
// them, for all-single materials without a full decode)
constexpr int INIT_CHUNK = 8;
__global__ void k_init(const Mat *Mp, uint64_t *winsW, uint64_t *winsB, uint64_t *mateW,
                       uint64_t *mateB, uint16_t *codesW, uint16_t *codesB, uint64_t *kexW0,
                       uint64_t *kexW1, uint64_t *kexL0, uint64_t *kexL1, uint64_t *xrW,
                       uint64_t *xrB, int s0, int ns, unsigned long long *cnt);
// the capture part of the static refutations for any material, OR-ed into
// xrW / xrB after k_init; the same launch shape as k_init
__global__ void k_xr(const Mat *Mp, const uint64_t *subwinW, const uint64_t *subwinB,
                     const uint64_t *winsW, const uint64_t *winsB, uint64_t *xrW, uint64_t *xrB,
                     int s0, int ns, unsigned long long *nopen);
// the same outputs for an all-single material of at least 64 placements
// per slice, one placement (every set's square but the last) per lane
__global__ void k_init_pl(const Mat *Mp, uint64_t *winsW, uint64_t *winsB, uint64_t *mateW,
                          uint64_t *mateB, uint16_t *codesW, uint16_t *codesB, uint64_t *kexW0,
                          uint64_t *kexW1, uint64_t *kexL0, uint64_t *kexL1, uint64_t *xrW,
                          uint64_t *xrB, int s0, int ns, unsigned long long *cnt);
// the capture part of the static refutations, OR-ed into xrW / xrB (which
// k_init_pl cleared or filled with the exit part), from the capture-exit
// planes sub[colour][TBG_SUB_WIN]; same launch shape as k_init_pl.  With
// sublossW/B (else null) also the ply-1 W step's capture candidates, OR-ed
// into candWW / candWB
__global__ void k_xr_pl(const Mat *Mp, const uint64_t *subwinW, const uint64_t *subwinB,
                        const uint64_t *sublossW, const uint64_t *sublossB,
                        const uint64_t *winsW, const uint64_t *winsB, uint64_t *xrW,
                        uint64_t *xrB, uint64_t *candWW, uint64_t *candWB, int s0, int ns,
                        unsigned long long *nopen);
__global__ void k_resolve(const Mat *Mp, const SubRef *Rp, int c, int k, uint64_t *loss,
                          uint64_t *bloss, uint64_t *cwin, uint64_t *win, uint64_t *lwin, int s0,
                          int ns);
// the same for all-single materials, one capture-space word (64 consecutive
// positions) per lane
__global__ void k_resolve_w(const Mat *Mp, const SubRef *Rp, int c, int k, uint64_t *loss,
                            uint64_t *bloss, uint64_t *cwin, uint64_t *win, uint64_t *lwin, int s0,
                            int ns);

// bitmap of the positions whose code is `code`, one bit per position, for
// `nslices` slices of `words` words (the reference's W_n / L_n files)
__global__ void k_code_plane(const uint16_t *codes, uint64_t size, uint64_t words, int nslices,
                             uint16_t code, uint64_t *out);

// codes narrowed to one byte each (every ply below 64)
__global__ void k_codes8(const uint16_t *codes, uint64_t n, uint8_t *out);

// One sub-table (one side to move) of a Syzygy WDL file, the format
// src/probe.c reads (encode, src/probe.c:603-687; set_enc_info, 725-770).
// Header piece i stands on our slot slot[i].  The leading group (both
// kings, or three unique pieces) is the most significant digit; group g
// after it has digit (rem / gfactor[g]) % gsize[g], members gfirst[g] ..
// gfirst[g] + gmult[g] - 1, ranked among the squares left free by the
// leading group and the groups before g.
struct SyzLayout {
  uint64_t tb_size;
  uint64_t lead_factor;
  uint64_t gfactor[MAXP];
  uint64_t gsize[MAXP];
  int num, lead, ngroups, stm;
  int gfirst[MAXP], gmult[MAXP];
  int slot[MAXP];
  uint8_t inv_tri[10], inv_diag[16], inv_lower[36];
};

// out[k] = WDL + 2 (0..4) of index i0 + k of sub-table L, or 0xff where the
// position is broken (kings touch) or illegal: a value the file may choose
// freely
__global__ void k_syz_gather(const Mat *Mp, const uint16_t *codes, SyzLayout L, uint64_t i0,
                             uint64_t n, uint8_t *out);

// copies the shared tables to the device (defined with the kernels)
hipError_t upload_tabs(const Tabs *h);

}  // namespace tbg
