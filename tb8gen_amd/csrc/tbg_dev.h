// tbg_dev.h -- position, index and move primitives of the king-slice engine.
//
// Everything here is __host__ __device__ so the host build of the library
// can self-test it against brute force (tbg_host_selftest) before a kernel
// ever runs.  A position keeps its (up to) 8 piece squares packed one per
// byte in a uint64_t ("sqp", slot i in byte i, slot order of the reference:
// WK, BK, then the sets, src/tb8gen.c:1331-1367).  Packed squares keep every
// slot access a shift, so per-lane state never spills to scratch memory,
// and the board symmetries become byte-parallel bit operations.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define TBG_HD __host__ __device__ __forceinline__

namespace tbg {

constexpr int MAXP = 8;      // pieces incl. kings (src/defs.h:11)
constexpr int MAXS = 6;      // sets of like non-king pieces (src/defs.h:12)
constexpr int NSL = 462;     // king slices
constexpr int DRAW_RULE = 100;

enum { WHITE = 0, BLACK = 1 };
enum { KNIGHT = 2, BISHOP = 3, ROOK = 4, QUEEN = 5, KING = 6 };

// Index layout of one material or capture space (src/tb8gen.h:4-13).
struct Idx {
  int numsets;
  int first[MAXS], mult[MAXS], last[MAXS];
  uint64_t size;
  uint32_t factor[MAXS];     // below 2^32 for up to 8 pieces (at most 61*60*59*58*57)
  double inv[MAXS];          // 1.0 / factor[k], for digit extraction
  float invf[MAXS];          // the same in single precision
  int small_q;               // every set's radix (its digit's range) is below 2^21
};

// Everything a kernel needs to know about a material.
struct Mat {
  int num, numsets;
  int pt[MAXP];              // piece code per slot (colour bit 8)
  int nk[2];                 // number of non-king pieces per colour
  int pcs[2][MAXP];          // their slots
  int setcol[MAXS];          // colour of each set
  uint64_t set_of;           // packed: set index of slot i in byte i (0xff for kings)
  uint64_t setmask80[MAXS];  // 0x80 in the byte of each slot of set k
  uint64_t col80[2];         // 0x80 in the byte of each non-king slot of colour c
  Idx ii, capt[MAXS];
  uint64_t words;            // 64-bit words per slice bitmap
  uint64_t cwords[MAXS], coff[MAXS], ctotal[2];
  uint64_t nslices_phys;     // 462, or 1 for a synthetic single-slice job
  int synth, synth_slice;
  uint64_t seed, seedmix;    // synthetic exits: seed and splitmix64(seed)
  int all_single;            // every set holds one piece (set k in slot 2 + k)
  // synthetic job: side c's king moves that leave the slice (the kings are
  // fixed, so each target's slice and symmetry are too): target square,
  // target slice, mirror mask and diagonal flip of the moved king pair
  int nexit[2];
  uint8_t exit_to[2][8], exit_m[2][8], exit_fl[2][8];
  int16_t exit_s2[2][8];
};

// Tables shared by all kernels (copied to LDS per work-group).
struct Tabs {
  uint64_t king[64];
  uint64_t knight[64];
  uint32_t binom[8 * 64];    // binom[k * 64 + n] = C(n, k)
  int16_t kkidx[10 * 64];    // slice numbering (src/probe.c:441-522)
  int8_t tri[64];            // white-king triangle square -> row of kkidx
  uint8_t kksq[NSL][2];      // slice -> (wk, bk)
};

// ---------------------------------------------------------------- bits
TBG_HD uint64_t bit(int sq) { return 1ull << sq; }
TBG_HD int ctz64(uint64_t x) { return __builtin_ctzll(x); }
TBG_HD int popc64(uint64_t x) { return __builtin_popcountll(x); }
TBG_HD int popc32(uint32_t x) { return __builtin_popcount(x); }
TBG_HD uint64_t brev64(uint64_t x) { return __builtin_bitreverse64(x); }

TBG_HD int sq_get(uint64_t p, int i) { return (int)((p >> (i << 3)) & 0xff); }
TBG_HD uint64_t sq_set(uint64_t p, int i, int v)
{
  int sh = i << 3;
  return (p & ~(0xffull << sh)) | ((uint64_t)v << sh);
}

// Slot holding square `to` among the slots selected by byte mask `valid`
// (0x80 in each candidate byte); -1 if none.
TBG_HD int slot_of(uint64_t p, int to, uint64_t valid80)
{
  const uint64_t lo7 = 0x7f7f7f7f7f7f7f7full;
  uint64_t x = p ^ (0x0101010101010101ull * (uint64_t)to);
  // per-byte "x != 0" without carries between bytes
  uint64_t nz = ((x & lo7) + lo7) | x;
  uint64_t z = ~nz & valid80;
  return z ? (ctz64(z) >> 3) : -1;
}

// ------------------------------------------------------------- attacks
// Line attacks by the o^(o-2r) identity in both directions (bit reversal
// turns the negative direction into a positive one); mask excludes the
// slider's own square.
TBG_HD uint64_t line_att(uint64_t occ, uint64_t r, uint64_t mask)
{
  uint64_t o = occ & mask;
  uint64_t f = o - r;
  uint64_t b = brev64(o) - brev64(r);
  return (f ^ brev64(b)) & mask;
}
TBG_HD uint64_t rook_att(int sq, uint64_t occ)
{
  uint64_t r = bit(sq);
  uint64_t file = (0x0101010101010101ull << (sq & 7)) ^ r;
  uint64_t rank = (0xffull << (sq & 56)) ^ r;
  return line_att(occ, r, file) | line_att(occ, r, rank);
}
TBG_HD uint64_t bishop_att(int sq, uint64_t occ)
{
  uint64_t r = bit(sq);
  int d = (sq & 7) - (sq >> 3);
  uint64_t diag = d >= 0 ? 0x8040201008040201ull >> (8 * d) : 0x8040201008040201ull << (-8 * d);
  int e = (sq & 7) + (sq >> 3) - 7;
  uint64_t anti = e >= 0 ? 0x0102040810204080ull << (8 * e) : 0x0102040810204080ull >> (-8 * e);
  return line_att(occ, r, diag ^ r) | line_att(occ, r, anti ^ r);
}
// piece_attacks / non_king_piece_attacks (src/movegen.h:79-117)
TBG_HD uint64_t piece_att(const Tabs &T, int pt, int sq, uint64_t occ)
{
  switch (pt & 7) {
  case KNIGHT: return T.knight[sq];
  case BISHOP: return bishop_att(sq, occ);
  case ROOK: return rook_att(sq, occ);
  case QUEEN: return rook_att(sq, occ) | bishop_att(sq, occ);
  default: return T.king[sq];
  }
}

// ------------------------------------------------------------- symmetry
// MirrorMask (src/tb8gen.c:75-76) and FlipTest (src/tb8gen.c:78-86)
TBG_HD int mirror_mask(int wk) { return ((wk & 4) ? 7 : 0) | ((wk & 32) ? 0x38 : 0); }
TBG_HD bool flip_test(int wk, int bk)
{
  const uint64_t A1D1D4 = 0x080c0e0full, A1D4 = 0x08040201ull, LOWER = 0x80c0e0f0f8fcfeffull;
  int m = mirror_mask(wk);
  int s1 = wk ^ m, s2 = bk ^ m;
  return !((A1D1D4 >> s1) & 1) || (((A1D4 >> s1) & 1) && !((LOWER >> s2) & 1));
}
// The square t of a position normalised by (m, fl): t ^ m, then for fl the
// a1-h8 flip (file and rank exchanged)
TBG_HD int sym_sq(int t, int m, bool fl)
{
  int u = t ^ m;
  return fl ? ((u & 7) << 3) | (u >> 3) : u;
}
// A square mask under t -> t ^ m for m in {0, 7, 56, 63} (mirror_mask): a
// file mirror reverses the bits of each byte, a rank mirror the byte order
TBG_HD uint64_t mirror_bb(uint64_t x, int m)
{
  uint32_t lo = (uint32_t)x, hi = (uint32_t)(x >> 32);
  if (m & 7) {
    lo = __builtin_bswap32(__builtin_bitreverse32(lo));
    hi = __builtin_bswap32(__builtin_bitreverse32(hi));
  }
  if (m & 56) {
    uint32_t t = __builtin_bswap32(lo);
    lo = __builtin_bswap32(hi);
    hi = t;
  }
  return lo | (uint64_t)hi << 32;
}
// A square mask flipped about the a1-h8 diagonal (three delta swaps: 4x4
// blocks, 2x2 blocks, single squares)
TBG_HD uint64_t transpose_bb(uint64_t x)
{
  uint64_t t = 0x0f0f0f0f00000000ull & (x ^ (x << 28));
  x ^= t ^ (t >> 28);
  t = 0x3333000033330000ull & (x ^ (x << 14));
  x ^= t ^ (t >> 14);
  t = 0x5500550055005500ull & (x ^ (x << 7));
  x ^= t ^ (t >> 7);
  return x;
}
// the masks of sym_sq's images of a mask's squares, and back
TBG_HD uint64_t sym_bb(uint64_t x, int m, bool fl)
{
  x = mirror_bb(x, m);
  return fl ? transpose_bb(x) : x;
}
TBG_HD uint64_t sym_bb_inv(uint64_t x, int m, bool fl) { return mirror_bb(fl ? transpose_bb(x) : x, m); }
// FlipDiag of every byte
TBG_HD uint64_t flip_diag_bytes(uint64_t p)
{
  return ((p & 0x0707070707070707ull) << 3) | ((p >> 3) & 0x0707070707070707ull);
}
// normalize (src/tb8gen.c:230-238) of packed squares; returns the slice
// KKMap[wk][bk] (-1 for kings that touch).
TBG_HD int normalize(const Tabs &T, uint64_t p, uint64_t *out)
{
  int wk = sq_get(p, 0), bk = sq_get(p, 1);
  int m = mirror_mask(wk);
  uint64_t q = p ^ (0x0101010101010101ull * (uint64_t)m);
  if (flip_test(wk, bk)) q = flip_diag_bytes(q);
  *out = q;
  return T.kkidx[T.tri[sq_get(q, 0)] * 64 + sq_get(q, 1)];
}
// mirror_diagonal of the non-king slots (src/tb8gen.c:240-244)
TBG_HD uint64_t mirror_nonkings(uint64_t p)
{
  uint64_t f = flip_diag_bytes(p);
  return (p & 0xffffull) | (f & ~0xffffull);
}

// ---------------------------------------------------------------- index
// r-th (0-based) zero bit of occ (the _pdep_u64 of src/tb8gen.c:117,121)
TBG_HD int select_zero(uint64_t occ, int r)
{
  uint64_t f = ~occ;
  uint32_t x = (uint32_t)f;
  int base = 0, c = popc32(x);
  if (r >= c) { r -= c; x = (uint32_t)(f >> 32); base = 32; }
  c = popc32(x & 0xffffu); if (r >= c) { r -= c; x >>= 16; base += 16; }
  c = popc32(x & 0xffu);   if (r >= c) { r -= c; x >>= 8;  base += 8; }
  c = popc32(x & 0xfu);    if (r >= c) { r -= c; x >>= 4;  base += 4; }
  c = popc32(x & 0x3u);    if (r >= c) { r -= c; x >>= 2;  base += 2; }
  c = (int)(x & 1u);       if (r >= c) { base += 1; }
  return base;
}

// floor(n / d) for n < 2^53 via one double multiply and a +-1 fix-up
TBG_HD uint64_t div_fix(uint64_t n, uint64_t d, double inv, uint64_t *rem)
{
  uint64_t q = (uint64_t)((double)n * inv);
  int64_t r = (int64_t)(n - q * d);
  if (r < 0) { q--; r += (int64_t)d; }
  else if ((uint64_t)r >= d) { q++; r -= (int64_t)d; }
  *rem = (uint64_t)r;
  return q;
}

// floor(n / d) for a quotient below 2^21 (a digit of the index: at most its
// set's radix).  n is converted in two halves (relative error below 2^-23),
// the product with the rounded 1/d adds 2^-23 more, so the estimate is
// within 2^21 * 2^-22 = 0.5 of n / d and one fix-up step either way gives
// the quotient; q * d is one 32 x 32 -> 64-bit multiply.
TBG_HD uint32_t div_small(uint64_t n, uint32_t d, float inv, uint64_t *rem)
{
  float fn = (float)(uint32_t)(n >> 32) * 4294967296.0f + (float)(uint32_t)n;
  uint32_t q = (uint32_t)(fn * inv);
  int64_t r = (int64_t)(n - (uint64_t)q * d);
  if (r < 0) { q--; r += (int64_t)d; }
  else if ((uint64_t)r >= d) { q++; r -= (int64_t)d; }
  *rem = (uint64_t)r;
  return q;
}

// sq_to_idx_helper (src/tb8gen.c:129-149).  Within a set the members need
// not be sorted: member a's place in sorted order is the number of members
// on lower squares.
TBG_HD uint64_t sq_to_idx(const Tabs &T, const Idx &ii, uint64_t p)
{
  uint64_t occ = bit(sq_get(p, 0)) | bit(sq_get(p, 1));
  uint64_t idx = 0;
  for (int k = 0; k < MAXS; k++) {
    if (k >= ii.numsets) break;
    int f = ii.first[k], m = ii.mult[k];
    if (m == 1) {
      int s = sq_get(p, f);
      uint32_t rank = (uint32_t)(s - popc64(occ & (bit(s) - 1)));
      idx += (uint64_t)rank * ii.factor[k];
      occ |= bit(s);
    } else {
      uint64_t setb = 0;
      for (int a = 0; a < MAXS; a++)
        if (a < m) setb |= bit(sq_get(p, f + a));
      uint32_t d = 0;
      for (int a = 0; a < MAXS; a++) {
        if (a >= m) break;
        int s = sq_get(p, f + a);
        int rank = s - popc64(occ & (bit(s) - 1));
        int pos = popc64(setb & (bit(s) - 1));
        d += T.binom[(pos + 1) * 64 + rank];
      }
      idx += (uint64_t)d * ii.factor[k];
      occ |= setb;
    }
  }
  return idx;
}

// idx_to_sq_init + idx_to_sq_unpack + unrank_binomial (src/tb8gen.c:105-175).
// p holds the kings in slots 0 and 1 on entry; returns the occupancy.
TBG_HD uint64_t idx_to_sq(const Tabs &T, const Idx &ii, uint64_t idx, uint64_t *pp)
{
  uint64_t p = *pp;
  uint64_t occ = bit(sq_get(p, 0)) | bit(sq_get(p, 1));
  for (int k = 0; k < MAXS; k++) {
    if (k >= ii.numsets) break;
    uint64_t d;
    if (ii.factor[k] == 1) {
      d = idx;
      idx = 0;
    } else if (ii.small_q) {
      uint64_t rem;
      d = div_small(idx, ii.factor[k], ii.invf[k], &rem);
      idx = rem;
    } else {
      uint64_t rem;
      d = div_fix(idx, ii.factor[k], ii.inv[k], &rem);
      idx = rem;
    }
    int f = ii.first[k], m = ii.mult[k];
    if (m == 0) continue;  // set emptied by a capture (unrank_binomial n == 0)
    if (m == 1) {
      int s = select_zero(occ, (int)d);
      p = sq_set(p, f, s);
      occ |= bit(s);
    } else {
      uint64_t occ0 = occ;
      uint32_t dd = (uint32_t)d;
      for (int i = MAXS - 1; i > 0; i--) {
        if (i >= m) continue;
        // largest r with C(r, i+1) <= dd
        int lo = i, hi = 62;
        while (hi - lo > 1) {
          int mid = (lo + hi) >> 1;
          if (T.binom[(i + 1) * 64 + mid] <= dd) lo = mid; else hi = mid;
        }
        dd -= T.binom[(i + 1) * 64 + lo];
        int s = select_zero(occ0, lo);
        p = sq_set(p, f + i, s);
        occ |= bit(s);
      }
      int s = select_zero(occ0, (int)dd);
      p = sq_set(p, f, s);
      occ |= bit(s);
    }
  }
  *pp = p;
  return occ;
}

// What moved_idx_l reads about the sets after kj (all sets single): their
// squares in position p and their factors (0 past the last set), and kj's
// own factor, gathered once per moving piece so that the loop over its
// targets reads no table (a table read there would also wait on every LDS
// atomic the loop has queued).
struct Later {
  uint32_t fj;
  uint32_t f[MAXS - 1];
  int s[MAXS - 1];
  int n;  // number of later sets
};
TBG_HD Later later_sets(const Idx &ii, uint64_t p, int kj)
{
  Later L;
  L.fj = ii.factor[kj];
  L.n = ii.numsets - 1 - kj;
  for (int k = 0; k < MAXS - 1; k++) {
    int kk = kj + 1 + k;
    L.f[k] = 0;
    L.s[k] = 0;
    if (kk < ii.numsets) {
      L.f[k] = ii.factor[kk];
      L.s[k] = sq_get(p, 2 + kk);
    }
  }
  return L;
}
// Index after the piece of set kj (all sets single) moves from a to b, from
// the index before: its own rank changes, and each later piece strictly
// between a and b shifts its rank by one (its set of lower occupied squares
// gains or loses a).  occb = kings + pieces of the sets before kj;
// rank_a = rank of a among the squares free of occb; L = later_sets(kj).
TBG_HD uint64_t moved_idx_l(const Later &L, uint64_t idx, int a, int b, uint64_t occb,
                            uint32_t rank_a)
{
  uint32_t rank_b = (uint32_t)(b - popc64(occb & (bit(b) - 1)));
  int lo = a < b ? a : b, hi = a < b ? b : a;
  uint32_t acc = 0;
#pragma unroll
  for (int k = 0; k < MAXS - 1; k++) {
    if (k >= L.n) break;  // the same for every lane: a scalar branch
    acc += (L.s[k] > lo && L.s[k] < hi) ? L.f[k] : 0u;
  }
  int64_t dr = (int64_t)(int32_t)(rank_b - rank_a) * (int64_t)L.fj;
  uint64_t r = idx + (uint64_t)dr;
  return a < b ? r + acc : r - acc;
}

// idx_to_sq of an all-single material whose digits take div_small
// (ii.small_q): set k's one piece is slot 2 + k, the last set's factor is 1.
// Written per set index so that the table reads can all be issued before
// the first is needed.
TBG_HD uint64_t idx_to_sq_1(const Idx &ii, uint64_t idx, uint64_t *pp)
{
  uint64_t p = *pp;
  uint64_t occ = bit(sq_get(p, 0)) | bit(sq_get(p, 1));
  int n = ii.numsets;
  uint32_t f[MAXS];
  float inv[MAXS];
  for (int k = 0; k < MAXS; k++) {
    f[k] = ii.factor[k];
    inv[k] = ii.invf[k];
  }
  for (int k = 0; k < MAXS; k++) {
    if (k >= n) break;
    uint64_t d = idx;
    if (k + 1 < n) d = div_small(idx, f[k], inv[k], &idx);
    int s = select_zero(occ, (int)d);
    p = sq_set(p, 2 + k, s);
    occ |= bit(s);
  }
  *pp = p;
  return occ;
}

// sq_to_idx of an all-single material (set k's piece in slot 2 + k),
// written per set index so that the factor reads can all be issued first
TBG_HD uint64_t sq_to_idx_1(const Idx &ii, uint64_t p)
{
  uint64_t occ = bit(sq_get(p, 0)) | bit(sq_get(p, 1));
  uint64_t idx = 0;
  int n = ii.numsets;
  uint32_t f[MAXS];
  for (int k = 0; k < MAXS; k++) f[k] = ii.factor[k];
  for (int k = 0; k < MAXS; k++) {
    if (k >= n) break;
    int s = sq_get(p, 2 + k);
    idx += (uint64_t)(uint32_t)(s - popc64(occ & (bit(s) - 1))) * f[k];
    occ |= bit(s);
  }
  return idx;
}

// sq_to_idx in the capture space of set l of an all-single material (its
// Idx: the material's, set l emptied): the other sets' pieces in slots
// 2 + k, set l's slot ignored
TBG_HD uint64_t capt_idx_1(const Idx &ci, uint64_t p, int l)
{
  uint64_t occ = bit(sq_get(p, 0)) | bit(sq_get(p, 1));
  uint64_t idx = 0;
  int n = ci.numsets;
  uint32_t f[MAXS];
  for (int k = 0; k < MAXS; k++) f[k] = ci.factor[k];
  for (int k = 0; k < MAXS; k++) {
    if (k >= n) break;
    if (k == l) continue;
    int s = sq_get(p, 2 + k);
    idx += (uint64_t)(uint32_t)(s - popc64(occ & (bit(s) - 1))) * f[k];
    occ |= bit(s);
  }
  return idx;
}

// Stepping through consecutive indices of an all-single material: set k's
// digit is the rank of its square among the 62 - k squares left by the kings
// and the sets before it (a radix of 62 - k), six bits per digit in D.
TBG_HD uint64_t digits_of(uint64_t p, int n)
{
  uint64_t occ = bit(sq_get(p, 0)) | bit(sq_get(p, 1));
  uint64_t D = 0;
  for (int k = 0; k < MAXS; k++) {
    if (k >= n) break;
    int s = sq_get(p, 2 + k);
    D |= (uint64_t)(s - popc64(occ & (bit(s) - 1))) << (6 * k);
    occ |= bit(s);
  }
  return D;
}
// The position of index + 64 from that of index (its digits D and squares
// p, both updated): 64 is added to the last digit with carries upwards, and
// the squares of the sets from the lowest changed digit on are placed again.
// Returns the occupancy.  Past the table's end the result is meaningless.
TBG_HD uint64_t step64(int n, uint64_t *dp, uint64_t *pp)
{
  uint64_t D = *dp, p = *pp;
  int add = 64, low = n - 1;
  for (int j = MAXS - 1; j >= 0; j--) {
    if (j >= n) continue;
    int r = 62 - j, d = (int)((D >> (6 * j)) & 63) + add;
    add = 0;
    while (d >= r) {  // twice at most: 64 < 2 r
      d -= r;
      add++;
    }
    D = (D & ~(63ull << (6 * j))) | ((uint64_t)d << (6 * j));
    low = j;
    if (!add) break;
  }
  uint64_t occ = bit(sq_get(p, 0)) | bit(sq_get(p, 1));
  for (int j = 0; j < MAXS; j++) {
    if (j >= n) break;
    if (j >= low) p = sq_set(p, 2 + j, select_zero(occ, (int)((D >> (6 * j)) & 63)));
    occ |= bit(sq_get(p, 2 + j));
  }
  *dp = D;
  *pp = p;
  return occ;
}

// The position of index + delta from that of index (digits D and squares
// p, both updated), for 0 < delta < 2^24: delta is added to the last digit
// and the carries go upwards (division by each set's radix), then the
// squares from the lowest changed digit on are placed again.  Returns the
// occupancy.  Past the table's end the result is meaningless.
TBG_HD uint64_t step_by(int n, uint64_t *dp, uint64_t *pp, uint32_t delta)
{
  uint64_t D = *dp, p = *pp;
  uint32_t add = delta;
  int low = n - 1;
#pragma unroll
  for (int j = MAXS - 1; j >= 0; j--) {
    if (j >= n) continue;
    uint32_t r = 62 - j, d = (uint32_t)((D >> (6 * j)) & 63) + add;
    add = d / r;  // a constant divisor per digit
    d -= add * r;
    D = (D & ~(63ull << (6 * j))) | ((uint64_t)d << (6 * j));
    low = j;
    if (!add) break;
  }
  uint64_t occ = bit(sq_get(p, 0)) | bit(sq_get(p, 1));
  for (int j = 0; j < MAXS; j++) {
    if (j >= n) break;
    if (j >= low) p = sq_set(p, 2 + j, select_zero(occ, (int)((D >> (6 * j)) & 63)));
    occ |= bit(sq_get(p, 2 + j));
  }
  *dp = D;
  *pp = p;
  return occ;
}

// Index after a king moves from `from` to `to` when the position needs no
// symmetry (normalize returned it unchanged), from the index before: the
// kings precede every set, so each piece strictly between the two squares
// shifts its rank by one and nothing else changes.  All sets single; p is
// the position before or after the move (the pieces are the same).
TBG_HD uint64_t king_moved_idx(const Idx &ii, uint64_t idx, uint64_t p, int from, int to)
{
  int lo = from < to ? from : to, hi = from < to ? to : from;
  uint32_t acc = 0;  // at most the sum of the factors: below 2^30
  for (int k = 0; k < MAXS; k++) {
    if (k >= ii.numsets) break;
    int s = sq_get(p, 2 + k);
    acc += (s > lo && s < hi) ? ii.factor[k] : 0u;
  }
  return from < to ? idx + acc : idx - acc;
}

// king_moved_idx with the sets' squares (of the position before or after
// the move) and factors gathered beforehand: a king move's rank shifts
// come from every set, each factor 0 past the last set
struct AllSets {
  uint32_t f[MAXS];
  int s[MAXS];
};
TBG_HD AllSets all_sets(const Idx &ii, uint64_t p)
{
  AllSets A;
  for (int k = 0; k < MAXS; k++) {
    A.f[k] = k < ii.numsets ? ii.factor[k] : 0u;
    A.s[k] = sq_get(p, 2 + k);
  }
  return A;
}
// sq_to_idx of an all-single material from its gathered factors
TBG_HD uint64_t sq_to_idx_a(const AllSets &A, uint64_t p)
{
  uint64_t occ = bit(sq_get(p, 0)) | bit(sq_get(p, 1));
  uint64_t idx = 0;
  for (int k = 0; k < MAXS; k++) {
    if (A.f[k] == 0) break;
    int s = sq_get(p, 2 + k);
    idx += (uint64_t)(uint32_t)(s - popc64(occ & (bit(s) - 1))) * A.f[k];
    occ |= bit(s);
  }
  return idx;
}
TBG_HD uint64_t king_moved_idx_a(const AllSets &A, uint64_t idx, int from, int to)
{
  int lo = from < to ? from : to, hi = from < to ? to : from;
  uint32_t acc = 0;
#pragma unroll
  for (int k = 0; k < MAXS; k++) {
    if (A.f[k] == 0) break;  // past the last set
    acc += (A.s[k] > lo && A.s[k] < hi) ? A.f[k] : 0u;
  }
  return from < to ? idx + acc : idx - acc;
}

// kings + pieces of the sets before kj (all sets single)
TBG_HD uint64_t occ_before(uint64_t p, int kj)
{
  uint64_t o = bit(sq_get(p, 0)) | bit(sq_get(p, 1));
  for (int k = 0; k < MAXS; k++) {
    if (k >= kj) break;
    o |= bit(sq_get(p, 2 + k));
  }
  return o;
}

// is `target` attacked by the non-king pieces of colour c (plus the king of
// c); slot `skip` (captured piece) ignored.  Mirrors is_attacked_by
// (src/movegen.h:180-188) over the packed position.
TBG_HD bool attacked_by(const Tabs &T, const Mat &M, uint64_t p, uint64_t occ, int target, int c,
                        int skip)
{
  uint64_t tb = bit(target);
  if (T.king[sq_get(p, c)] & tb) return true;
  for (int t = 0; t < MAXP - 1; t++) {
    if (t >= M.nk[c]) break;
    int j = M.pcs[c][t];
    if (j == skip) continue;
    if (piece_att(T, M.pt[j], sq_get(p, j), occ) & tb) return true;
  }
  return false;
}

// splitmix64 and the seeded exit hash of synthetic jobs
TBG_HD uint64_t splitmix64(uint64_t x)
{
  x += 0x9e3779b97f4a7c15ull;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
  return x ^ (x >> 31);
}
// One splitmix64 round over the packed key XOR the premixed seed
// (seedmix = splitmix64(seed), kept in the material).
TBG_HD uint64_t synth_hash(uint64_t seedmix, int stm, int kind, int region, uint64_t idx)
{
  uint64_t key = (idx & 0xfffffffffull) | ((uint64_t)(region & 0xfff) << 36) |
                 ((uint64_t)(kind & 3) << 48) | ((uint64_t)(stm & 1) << 50);
  return splitmix64(key ^ seedmix);
}
constexpr uint64_t SYN_T_WIN = 553402322211286548ull;    // 0.03 * 2^64
constexpr uint64_t SYN_T_LOSS = 1106804644422573097ull;  // 0.06 * 2^64
TBG_HD int synth_value(uint64_t h) { return h < SYN_T_WIN ? 2 : h < SYN_T_LOSS ? -2 : 0; }

// synth_hash's key without the index, XOR the premixed seed: the hash of
// index idx (below 2^36) is splitmix64(idx ^ key)
TBG_HD uint64_t synth_key(uint64_t seedmix, int stm, int kind, int region)
{
  return (((uint64_t)(region & 0xfff) << 36) | ((uint64_t)(kind & 3) << 48) |
          ((uint64_t)(stm & 1) << 50)) ^ seedmix;
}
// The high 32 bits of splitmix64(idx ^ key), and in *z the state before
// the last multiply (to finish the hash with).  They decide synth_value
// except where they equal a threshold's high half, SYN_TW_HI or SYN_TL_HI
// (about one hash in 2^31).
constexpr uint32_t SYN_TW_HI = (uint32_t)(SYN_T_WIN >> 32), SYN_TL_HI = (uint32_t)(SYN_T_LOSS >> 32);
TBG_HD uint32_t synth_hi(uint64_t key, uint64_t idx, uint64_t *zp)
{
  uint64_t z = (idx ^ key) + 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z ^= z >> 27;
  *zp = z;
  uint32_t zl = (uint32_t)z, zh = (uint32_t)(z >> 32);
  uint32_t ph = (uint32_t)(((uint64_t)zl * 0x133111ebu) >> 32) + zl * 0x94d049bbu + zh * 0x133111ebu;
  return ph ^ (ph >> 31);
}
// synth_value(splitmix64(idx ^ key)) from the high half of the hash: the
// last multiply's low half is needed only when the high half ties a
// threshold's
TBG_HD int synth_value_k(uint64_t key, uint64_t idx)
{
  uint64_t z;
  uint32_t hh = synth_hi(key, idx, &z);
  int v = hh < SYN_TW_HI ? 2 : hh < SYN_TL_HI ? -2 : 0;
  if (hh == SYN_TW_HI || hh == SYN_TL_HI) {
    uint64_t h = z * 0x94d049bb133111ebull;
    v = synth_value(h ^ (h >> 31));
  }
  return v;
}

// code <-> WDL (tb8gen.c:300-312 classes; W_n with n <= 100 is a win)
TBG_HD int code_wdl(uint16_t c)
{
  int tag = c & 3, n = c >> 2;
  if (tag == 1) return n <= DRAW_RULE ? 2 : 1;
  if (tag == 2) return n <= DRAW_RULE ? -2 : -1;
  return 0;
}

}  // namespace tbg
