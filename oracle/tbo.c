/*
 * tbo.c -- CPU oracle: plain-C restatement of the reference generator's
 * king-slice retrograde analysis (syzygy1/tb8gen).
 *
 * TEST INFRASTRUCTURE ONLY (see tbo.h).  Each function cites the reference
 * file:line it restates.  Differences from the reference that do not change
 * results:
 *   - all 462 slices of a material are resident at once; there is no
 *     manager window and no slice files (kslice.c:41-84, 306-379);
 *   - index digits are always recomputed by division instead of the
 *     incremental idx_to_sq_add (tb8gen.c:182-201), which avoids the uint32
 *     gap truncation the reference marks FIXME at tb8gen.c:188;
 *   - the r-th free square is found by a bit scan instead of _pdep_u64;
 *   - slider attacks come from ray tables instead of the fixed-shift magics
 *     (magic.c); both compute the same attack sets;
 *   - capture exits (probe_wdl, probe.c:1639-1684) are answered from this
 *     oracle's own tables of the smaller materials: those hold the exact WDL
 *     of every legal position, so max(best capture, table) = table value.
 * One deliberate semantic choice: check_king_moves (tb8gen.c:543-555) builds
 * the capture-table index of a king capture of a duplicated enemy set
 * without the fix-up check_moves applies (tb8gen.c:586-588).  By default
 * this oracle applies the fix-up (the chess-correct index); the reference's
 * arithmetic is available with tbo_set_ref_king_capture_index(1), to
 * measure the divergence.  Distinct-piece materials are unaffected.
 */
#define _GNU_SOURCE
#include "tbo.h"

#include <omp.h>
#include <stdatomic.h>
#include <stdbool.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef uint64_t bb_t;

enum { WHITE = 0, BLACK = 1 };
enum { KNIGHT = 2, BISHOP = 3, ROOK = 4, QUEEN = 5, KING = 6 }; /* types.h:17 */
#define DRAW_RULE 100                                           /* tb8gen.c:36 */
#define MAX_PIECES 8
#define MAX_SETS 6
#define NSLICES 462

static int g_threads = 1;
static int g_ref_kcap = 0;
static int g_keep_synth_codes = 1;

void tbo_set_threads(int n) { g_threads = n > 0 ? n : 1; }
void tbo_set_ref_king_capture_index(int on) { g_ref_kcap = on != 0; }
void tbo_set_keep_synth_codes(int on) { g_keep_synth_codes = on != 0; }

/* ------------------------------------------------------------------ */
/* Board geometry (movegen.c:14-73 semantics, own construction).       */

static bb_t KingAtt[64], KnightAtt[64];
static bb_t Ray[8][64];
static const int RayDF[8] = { 0, 1, 1, 1, 0, -1, -1, -1 };
static const int RayDR[8] = { 1, 1, 0, -1, -1, -1, 0, 1 };
/* ray directions 0..7: N, NE, E, SE, S, SW, W, NW; increasing square for
 * N, NE, E, NW (+8, +9, +1, +7) */
static const int RayUp[8] = { 1, 1, 1, 0, 0, 0, 0, 1 };

static inline bb_t bit(int sq) { return 1ULL << sq; }
static inline int lsb(bb_t b) { return __builtin_ctzll(b); }
static inline int msb(bb_t b) { return 63 - __builtin_clzll(b); }
static inline int popcnt(bb_t b) { return __builtin_popcountll(b); }
static inline int pop_lsb(bb_t *b) { int s = lsb(*b); *b &= *b - 1; return s; }

static bb_t ray_attacks(int sq, bb_t occ, int d0, int step)
{
  bb_t a = 0;
  for (int d = d0; d < 8; d += step) {
    bb_t r = Ray[d][sq], b = r & occ;
    if (b) r ^= Ray[d][RayUp[d] ? lsb(b) : msb(b)];
    a |= r;
  }
  return a;
}
static inline bb_t rook_att(int sq, bb_t occ) { return ray_attacks(sq, occ, 0, 2); }
static inline bb_t bishop_att(int sq, bb_t occ) { return ray_attacks(sq, occ, 1, 2); }

/* non_king_piece_attacks / piece_attacks (movegen.h:79-117) */
static inline bb_t piece_att(int pt, int sq, bb_t occ)
{
  switch (pt & 7) {
  case KNIGHT: return KnightAtt[sq];
  case BISHOP: return bishop_att(sq, occ);
  case ROOK: return rook_att(sq, occ);
  case QUEEN: return ray_attacks(sq, occ, 0, 1);
  default: return KingAtt[sq];
  }
}

/* ------------------------------------------------------------------ */
/* Slice tables (tb8gen.c:69-89, probe.c:310-340, 441-546).            */

static int16_t KKIdx[10][64];
static int16_t KKMap[64][64];
static bool FlipTest[64][64];
static uint8_t MirrorMask[64], FlipDiag[64];
static uint8_t KKSq[NSLICES][2];
static int8_t TriOf[64];
static uint64_t Binom[8][64];
/* The slice numbering follows probe.c:327-329 (InvTriangle): white king
 * squares b1 c1 d1 c2 d2 d3 (off-diagonal) then a1 b2 c3 d4. */
static const uint8_t InvTri[10] = { 1, 2, 3, 10, 11, 19, 0, 9, 18, 27 };

static int tables_ready;

static bool on_diag(int sq) { return (sq & 7) == (sq >> 3); }
static bool adjacent(int a, int b)
{
  int df = (a & 7) - (b & 7), dr = (a >> 3) - (b >> 3);
  return df >= -1 && df <= 1 && dr >= -1 && dr <= 1;
}

static void init_tables(void)
{
  if (tables_ready) return;
  for (int sq = 0; sq < 64; sq++) {
    int f = sq & 7, r = sq >> 3;
    KingAtt[sq] = KnightAtt[sq] = 0;
    for (int df = -2; df <= 2; df++)
      for (int dr = -2; dr <= 2; dr++) {
        int f2 = f + df, r2 = r + dr;
        if (f2 < 0 || f2 > 7 || r2 < 0 || r2 > 7 || (!df && !dr)) continue;
        int ad = abs(df), ar = abs(dr);
        if (ad <= 1 && ar <= 1) KingAtt[sq] |= bit(r2 * 8 + f2);
        if ((ad == 1 && ar == 2) || (ad == 2 && ar == 1)) KnightAtt[sq] |= bit(r2 * 8 + f2);
      }
    for (int d = 0; d < 8; d++) {
      bb_t b = 0;
      for (int f2 = f + RayDF[d], r2 = r + RayDR[d]; f2 >= 0 && f2 < 8 && r2 >= 0 && r2 < 8;
           f2 += RayDF[d], r2 += RayDR[d])
        b |= bit(r2 * 8 + f2);
      Ray[d][sq] = b;
    }
    FlipDiag[sq] = (uint8_t)(((sq & 7) << 3) | (sq >> 3));
    /* tb8gen.c:75-76 */
    MirrorMask[sq] = (uint8_t)(((sq & 0x04) ? 0x07 : 0) | ((sq & 0x20) ? 0x38 : 0));
  }

  /* KKIdx (probe.c:441-522): for each white-king triangle square in
   * InvTri order and each black-king square in ascending order, valid
   * pairs (not equal, not adjacent, and with a diagonal white king the
   * black king on or below a1-h8) are numbered 0..440; pairs with both
   * kings on a1-h8 are numbered 441..461 in the same order. */
  int n = 0, nd = 441;
  for (int i = 0; i < 10; i++) {
    int wk = InvTri[i];
    for (int bk = 0; bk < 64; bk++) {
      KKIdx[i][bk] = -1;
      if (bk == wk || adjacent(wk, bk)) continue;
      if (on_diag(wk) && (bk & 7) < (bk >> 3)) continue;
      KKIdx[i][bk] = (int16_t)((on_diag(wk) && on_diag(bk)) ? nd++ : n++);
    }
  }
  if (n != 441 || nd != 462) { fprintf(stderr, "tbo: KKIdx construction failed\n"); abort(); }
  for (int i = 0; i < 64; i++) TriOf[i] = -1;
  for (int i = 0; i < 10; i++) TriOf[InvTri[i]] = (int8_t)i;
  /* init_indices (probe.c:534-539) */
  for (int i = 0; i < 10; i++)
    for (int j = 0; j < 64; j++)
      if (KKIdx[i][j] >= 0) {
        KKSq[KKIdx[i][j]][0] = InvTri[i];
        KKSq[KKIdx[i][j]][1] = (uint8_t)j;
      }
  /* init_tables (tb8gen.c:69-89) */
  const bb_t A1D1D4 = 0x080c0e0fULL, A1D4 = 0x08040201ULL, LOWER = 0x80c0e0f0f8fcfeffULL;
  for (int i = 0; i < 64; i++)
    for (int j = 0; j < 64; j++) {
      int s1 = i ^ MirrorMask[i], s2 = j ^ MirrorMask[i];
      FlipTest[i][j] = false;
      if (!(bit(s1) & A1D1D4) || ((bit(s1) & A1D4) && !(bit(s2) & LOWER))) {
        FlipTest[i][j] = true;
        s1 = FlipDiag[s1];
        s2 = FlipDiag[s2];
      }
      KKMap[i][j] = KKIdx[TriOf[s1]][s2];
    }
  /* Binomial[k][n] = C(n, k) (probe.c:541-546) */
  for (int j = 0; j < 64; j++) Binom[0][j] = 1;
  for (int i = 1; i < 8; i++) {
    Binom[i][0] = 0;
    for (int j = 1; j < 64; j++) Binom[i][j] = Binom[i - 1][j - 1] + Binom[i][j - 1];
  }
  tables_ready = 1;
}

int tbo_kkidx(int tri, int bk) { init_tables(); return KKIdx[tri][bk]; }
int tbo_kkmap(int wk, int bk) { init_tables(); return KKMap[wk][bk]; }
void tbo_kksquare(int s, uint8_t *o) { init_tables(); o[0] = KKSq[s][0]; o[1] = KKSq[s][1]; }

/* ------------------------------------------------------------------ */
/* Index (tb8gen.c:92-228, tb8gen.h:4-13).                              */

typedef struct {
  int numsets;
  uint64_t size;
  uint64_t factor[MAX_SETS];
  uint32_t subfactor[MAX_SETS];
  int first[MAX_SETS], mult[MAX_SETS], last[MAX_SETS];
} IdxInfo;

/* calc_factors (tb8gen.c:209-228); the +64 overshoot guard on subfactor[0]
 * only serves the incremental digit add, which this oracle does not use. */
static void calc_factors(IdxInfo *ii)
{
  for (int i = 0; i < MAX_SETS; i++) ii->factor[i] = 0;
  for (int i = 0, n = 62; i < ii->numsets; i++) {
    ii->subfactor[i] = (uint32_t)Binom[ii->mult[i]][n]; /* subfactor() probe.c:715-723 */
    n -= ii->mult[i];
  }
  uint64_t f = 1;
  for (int i = ii->numsets - 1; i >= 0; i--) {
    ii->factor[i] = f;
    f *= ii->subfactor[i];
  }
  ii->size = f;
}

/* sq_to_idx_helper (tb8gen.c:129-149): sorts each set in place. */
static uint64_t sq_to_idx(const IdxInfo *ii, uint8_t *sq)
{
  uint64_t idx = 0;
  bb_t occ = bit(sq[0]) | bit(sq[1]);
  for (int k = 0; k < ii->numsets; k++) {
    int i = ii->first[k], m = ii->mult[k];
    for (int a = 0; a < m; a++)                   /* sort_squares 92-98 */
      for (int b = a + 1; b < m; b++)
        if (sq[i + a] > sq[i + b]) { uint8_t t = sq[i + a]; sq[i + a] = sq[i + b]; sq[i + b] = t; }
    uint64_t s = 0;
    bb_t occ2 = occ;
    for (int j = 0; j < m; j++) {
      int sqr = sq[i + j];
      int rank = sqr - popcnt(occ & (bit(sqr) - 1)); /* rank_among_free 100-103 */
      occ2 |= bit(sqr);
      s += Binom[j + 1][rank];
    }
    idx += s * ii->factor[k];
    occ = occ2;
  }
  return idx;
}

/* r-th set bit of b (the _pdep_u64(1 << r, b) of tb8gen.c:117,121). */
static inline int select_bit(bb_t b, int r)
{
  while (r-- > 0) b &= b - 1;
  return lsb(b);
}

/* unrank_binomial (tb8gen.c:105-126) */
static bb_t unrank_set(uint64_t d, int m, uint8_t *p, bb_t occ)
{
  if (m == 0) return occ;
  bb_t freeb = ~occ;
  for (int i = m - 1; i > 0; i--) {
    int r = i;
    while (d >= Binom[i + 1][r + 1]) r++;
    d -= Binom[i + 1][r];
    int sq = select_bit(freeb, r);
    p[i] = (uint8_t)sq;
    occ |= bit(sq);
  }
  int sq = select_bit(freeb, (int)d);
  p[0] = (uint8_t)sq;
  return occ | bit(sq);
}

/* idx_to_sq_init + idx_to_sq_unpack (tb8gen.c:161-175).  Kings must be in
 * sq[0], sq[1]. */
static bb_t idx_to_sq(const IdxInfo *ii, uint64_t idx, uint8_t *sq)
{
  bb_t occ = bit(sq[0]) | bit(sq[1]);
  for (int k = 0; k < ii->numsets; k++) {
    uint64_t d = idx / ii->factor[k];
    idx -= d * ii->factor[k];
    occ = unrank_set(d, ii->mult[k], sq + ii->first[k], occ);
  }
  return occ;
}

/* normalize / mirror_diagonal (tb8gen.c:230-244) */
static inline void normalize(const uint8_t *sq, uint8_t *sq2)
{
  uint8_t m = MirrorMask[sq[0]];
  for (int i = 0; i < MAX_PIECES; i++) sq2[i] = sq[i] ^ m;
  if (FlipTest[sq[0]][sq[1]])
    for (int i = 0; i < MAX_PIECES; i++) sq2[i] = FlipDiag[sq2[i]];
}
static inline void mirror_diagonal(uint8_t *sq)
{
  for (int i = 2; i < MAX_PIECES; i++) sq[i] = FlipDiag[sq[i]];
}

/* ------------------------------------------------------------------ */
/* Material and tables.                                                 */

typedef struct {
  char name[16];
  int num;
  uint8_t pt[MAX_PIECES];
  int8_t pcs[2][MAX_PIECES];
  int pc_to_set[MAX_PIECES];
  IdxInfo ii, capt[MAX_SETS];
  uint64_t size, words;
  uint64_t cwords[MAX_SETS], coff[MAX_SETS], ctotal[2];
  uint16_t *code[2];
  /* sub bitmaps [colour][kind] of 462 * ctotal[colour] words */
  uint64_t *sub[2][4];
  int nplies;
  int64_t cnt[2][2][TBO_MAX_PLY]; /* [kind L/W][stm][n] */
  int64_t broken[2], mates[2];
  /* synthetic single-slice mode */
  int synth;
  int synth_slice;
  uint64_t seed, seedmix;
} Table;

#define MAXTABLES 512
static Table *g_tab[MAXTABLES];
static int g_ntab;

static const char PieceChar[] = " PNBRQK";

/* piece_order (tb8gen.c:1331-1339), indexed by piece code */
static const int PieceOrder[16] = { 0, 0, 3, 5, 7, 9, 1, 0, 0, 0, 4, 6, 8, 10, 2, 0 };

/* Parse and lay out a material like main() (tb8gen.c:1284-1385) and
 * kslice_setup (kslice.c:98-127). */
static int setup_material(Table *t, const char *name)
{
  uint8_t pt[16];
  int color = 0, k = 0, cnt[16] = { 0 };
  for (const char *s = name; *s; s++) {
    if (*s == 'v' && !color) { color = 8; continue; }
    int found = 0;
    for (int i = 2; i <= KING; i++)
      if (*s == PieceChar[i]) {
        if (k >= MAX_PIECES) return -1;
        cnt[i | color]++;
        pt[k++] = (uint8_t)(i | color);
        found = 1;
        break;
      }
    if (!found) return -1;
  }
  if (!color || k < 3 || cnt[KING] != 1 || cnt[KING | 8] != 1) return -1;
  snprintf(t->name, sizeof t->name, "%s", name);
  t->num = k;
  for (int i = 0; i < k; i++)
    for (int j = i + 1; j < k; j++)
      if (PieceOrder[pt[i]] > PieceOrder[pt[j]]) { uint8_t x = pt[i]; pt[i] = pt[j]; pt[j] = x; }
  for (int i = 0; i < MAX_PIECES; i++) t->pt[i] = i < k ? pt[i] : 0;
  for (int c = 0; c < 2; c++) {
    int m = 0;
    for (int i = 0; i < k; i++)
      if ((pt[i] >> 3) == c) t->pcs[c][m++] = (int8_t)i;
    t->pcs[c][m] = -1;
  }
  int ns = 0;
  for (int i = 2; i < k;) {
    int j = i;
    for (; i < k && pt[i] == pt[j]; i++) t->pc_to_set[i] = ns;
    t->ii.first[ns] = j;
    t->ii.mult[ns] = i - j;
    t->ii.last[ns] = i - 1;
    ns++;
  }
  t->ii.numsets = ns;
  calc_factors(&t->ii);
  t->size = t->ii.size;
  t->words = ((t->size + 511) / 512) * 8; /* bits_to_aligned kslice.c:92-96 */
  t->ctotal[0] = t->ctotal[1] = 0;
  for (int s = 0; s < ns; s++) {
    t->capt[s] = t->ii;
    t->capt[s].mult[s]--;
    calc_factors(&t->capt[s]);
    int c = t->pt[t->ii.first[s]] >> 3;
    t->cwords[s] = ((t->capt[s].size + 511) / 512) * 8;
    t->coff[s] = t->ctotal[c];
    t->ctotal[c] += t->cwords[s];
  }
  return 0;
}

/* create_material_string (probe.c:1480-1499) */
static void material_string(const uint8_t *pt, int num, bool flip, char *str)
{
  int cnt[16] = { 0 };
  for (int i = 0; i < num; i++) cnt[pt[i] ^ (flip ? 8 : 0)]++;
  int j = 0;
  for (int i = KING; i >= KNIGHT; i--)
    while (cnt[i]--) str[j++] = PieceChar[i];
  str[j++] = 'v';
  for (int i = KING; i >= KNIGHT; i--)
    while (cnt[i + 8]--) str[j++] = PieceChar[i];
  str[j] = 0;
}

/* Preferred colour orientation for a generated sub-material: the side with
 * more pieces, then the stronger piece string, is white. */
static int side_cmp(const char *a, const char *b)
{
  size_t la = strlen(a), lb = strlen(b);
  if (la != lb) return la > lb ? 1 : -1;
  for (size_t i = 0; i < la; i++)
    if (a[i] != b[i]) {
      const char *ord = "NBRQK";
      return strchr(ord, a[i]) > strchr(ord, b[i]) ? 1 : -1;
    }
  return 0;
}

static void canonical_name(const char *in, char *out)
{
  char w[16] = { 0 }, b[16] = { 0 };
  const char *v = strchr(in, 'v');
  size_t lw = (size_t)(v - in);
  memcpy(w, in, lw); w[lw] = 0;
  strcpy(b, v + 1);
  if (side_cmp(w, b) >= 0) snprintf(out, 16, "%.7sv%.7s", w, b);
  else snprintf(out, 16, "%.7sv%.7s", b, w);
}

int tbo_find(const char *material)
{
  for (int i = 0; i < g_ntab; i++)
    if (!strcmp(g_tab[i]->name, material)) return i;
  return -1;
}

static inline bool bm_test(const uint64_t *p, uint64_t idx) { return (p[idx >> 6] >> (idx & 63)) & 1; }
static inline void bm_set(uint64_t *p, uint64_t idx) { p[idx >> 6] |= 1ULL << (idx & 63); }
static inline void bm_set_atomic(uint64_t *p, uint64_t idx)
{
  __atomic_fetch_or(&p[idx >> 6], 1ULL << (idx & 63), __ATOMIC_RELAXED);
}

/* is_attacked_by (movegen.h:180-188) */
static bool attacked_by(const uint8_t *pt, const uint8_t *sq, int num, bb_t occ, int target, int side)
{
  bb_t b = bit(target);
  for (int i = 0; i < num; i++)
    if ((pt[i] >> 3) == side && (piece_att(pt[i], sq[i], occ) & b)) return true;
  return false;
}

/* king slot is the colour itself for pawnless tables (movegen.h:173-177) */
static inline bool opp_king_attacked(const uint8_t *pt, const uint8_t *sq, int num, bb_t occ, int stm)
{
  return attacked_by(pt, sq, num, occ, sq[stm ^ 1], stm);
}

/* ------------------------------------------------------------------ */
/* Probe (probe.c:1522-1684) from this oracle's own tables.             */

static int code_to_wdl(uint16_t c)
{
  int tag = c & 3, n = c >> 2;
  if (tag == 1) return n <= DRAW_RULE ? 2 : 1;
  if (tag == 2) return n <= DRAW_RULE ? -2 : -1;
  return 0;
}

static uint16_t table_lookup(const Table *t, int stm, const uint8_t *sq_in)
{
  uint8_t sq[MAX_PIECES] = { 0 }, sq2[MAX_PIECES];
  memcpy(sq, sq_in, (size_t)t->num);
  normalize(sq, sq2);
  int s = KKMap[sq[0]][sq[1]];
  if (s < 0) return TBO_ILLEGAL;
  uint64_t idx = sq_to_idx(&t->ii, sq2);
  return t->code[stm][(uint64_t)s * t->size + idx];
}

/* Exact WDL of a legal position given as a piece list (types, squares). */
static int probe_wdl(const uint8_t *pt, const uint8_t *sq, int num, int stm)
{
  if (num == 2) return 0; /* KvK (probe.c:1525) */
  char name[16];
  material_string(pt, num, false, name);
  int h = tbo_find(name);
  bool flip = false;
  if (h < 0) {
    material_string(pt, num, true, name);
    h = tbo_find(name);
    flip = true;
  }
  if (h < 0) {
    fprintf(stderr, "tbo: missing table for probe\n");
    abort();
  }
  const Table *t = g_tab[h];
  /* list_squares (probe.c:1511-1520): place pieces into the table's slots */
  uint8_t tsq[MAX_PIECES];
  bool used[MAX_PIECES] = { false };
  for (int j = 0; j < t->num; j++) {
    int want = t->pt[j] ^ (flip ? 8 : 0);
    for (int i = 0; i < num; i++)
      if (!used[i] && pt[i] == want) {
        used[i] = true;
        tsq[j] = sq[i] ^ (flip ? 0x38 : 0);
        break;
      }
  }
  uint16_t c = table_lookup(t, stm ^ (flip ? 1 : 0), tsq);
  if (c == TBO_ILLEGAL) {
    fprintf(stderr, "tbo: probe of an illegal position\n");
    abort();
  }
  return code_to_wdl(c);
}

/* ------------------------------------------------------------------ */
/* Synthetic exits for single-slice mode.                               */

static inline uint64_t splitmix64(uint64_t x)
{
  x += 0x9e3779b97f4a7c15ULL;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
  return x ^ (x >> 31);
}

/* one splitmix64 round over the packed key XOR splitmix64(seed) */
static inline uint64_t synth_hash_mixed(uint64_t seedmix, int stm, int kind, int region, uint64_t idx)
{
  uint64_t key = (idx & 0xfffffffffULL) | ((uint64_t)(region & 0xfff) << 36) |
                 ((uint64_t)(kind & 3) << 48) | ((uint64_t)(stm & 1) << 50);
  return splitmix64(key ^ seedmix);
}

uint64_t tbo_synth_hash(uint64_t seed, int stm, int kind, int region, uint64_t idx)
{
  return synth_hash_mixed(splitmix64(seed), stm, kind, region, idx);
}
/* win for the side to move below T_WIN, loss in [T_WIN, T_LOSS) */
#define SYN_T_WIN 553402322211286548ULL   /* 0.03 * 2^64 */
#define SYN_T_LOSS 1106804644422573097ULL /* 0.06 * 2^64 */
static inline int synth_value(uint64_t h) { return h < SYN_T_WIN ? 2 : h < SYN_T_LOSS ? -2 : 0; }

/* ------------------------------------------------------------------ */
/* Stages.                                                              */

#define SLICE_LO(t) ((t)->synth ? (t)->synth_slice : 0)
#define SLICE_HI(t) ((t)->synth ? (t)->synth_slice + 1 : NSLICES)

static inline uint64_t *bm_slice(const Table *t, uint64_t *bm, int s)
{
  return bm + (t->synth ? 0 : (uint64_t)s * t->words);
}
static inline uint64_t *sub_slice(const Table *t, uint64_t *bm, int colour, int s)
{
  return bm + (t->synth ? 0 : (uint64_t)s * t->ctotal[colour]);
}
static inline bool in_table(const Table *t, int s) { return !t->synth || s == t->synth_slice; }

/* calc_sub_worker / calc_sub_kslices (tb8gen.c:280-342): classify every
 * position with one piece of set k (colour stm) removed, stm to move. */
static void calc_sub(Table *t, int stm)
{
  for (int k = 0; k < t->ii.numsets; k++) {
    if ((t->pt[t->ii.first[k]] >> 3) != stm) continue;
    const IdxInfo *ci = &t->capt[k];
    int m = t->ii.last[k], n = t->num - 1;
    uint8_t spt[MAX_PIECES];
    memcpy(spt, t->pt, MAX_PIECES);
    spt[m] = t->pt[n];
#pragma omp parallel for schedule(dynamic, 1) num_threads(g_threads)
    for (int s = SLICE_LO(t); s < SLICE_HI(t); s++) {
      uint64_t *p[4];
      for (int q = 0; q < 4; q++) p[q] = sub_slice(t, t->sub[stm][q], stm, s) + t->coff[k];
      uint8_t sq[MAX_PIECES] = { 0 };
      for (uint64_t idx = 0; idx < ci->size; idx++) {
        sq[0] = KKSq[s][0];
        sq[1] = KKSq[s][1];
        bb_t occ = idx_to_sq(ci, idx, sq);
        sq[m] = sq[n];
        if (opp_king_attacked(spt, sq, n, occ, stm)) {
          bm_set(p[2], idx);
          bm_set(p[3], idx);
          continue;
        }
        int v;
        if (t->synth)
          v = synth_value(synth_hash_mixed(t->seedmix, stm, 0, k, idx));
        else
          v = probe_wdl(spt, sq, n, stm);
        if (v == -2) bm_set(p[0], idx);
        else if (v == -1) bm_set(p[1], idx);
        else if (v >= 1) {
          bm_set(p[2], idx);
          if (v == 2) bm_set(p[3], idx);
        }
      }
    }
  }
}

/* has_legal_moves / has_legal_caps (movegen.c:75-111) */
static bool has_legal_move_or_cap(const uint8_t *pt, const uint8_t *sq_in, int num, bb_t occ, int stm)
{
  uint8_t sq[MAX_PIECES], ptc[MAX_PIECES];
  memcpy(sq, sq_in, MAX_PIECES);
  for (int i = 0; i < num; i++) {
    if ((pt[i] >> 3) != stm) continue;
    int from = sq[i];
    bb_t b = piece_att(pt[i], from, occ) & ~occ;
    while (b) {
      int to = pop_lsb(&b);
      sq[i] = (uint8_t)to;
      bool ok = !attacked_by(pt, sq, num, occ ^ bit(from) ^ bit(to), sq[stm], stm ^ 1);
      sq[i] = (uint8_t)from;
      if (ok) return true;
    }
  }
  for (int i = 0; i < num; i++) {
    if ((pt[i] >> 3) != stm) continue;
    int from = sq[i];
    bb_t b = piece_att(pt[i], from, occ) & occ;
    while (b) {
      int to = pop_lsb(&b);
      int j = 0;
      while (sq[j] != to) j++;
      if (!((pt[i] ^ pt[j]) & 8)) continue;
      /* remove piece j (do_capture movegen.h:219-234) */
      int n2 = 0;
      uint8_t sq2[MAX_PIECES];
      for (int x = 0; x < num; x++)
        if (x != j) { ptc[n2] = pt[x]; sq2[n2] = x == i ? (uint8_t)to : sq[x]; n2++; }
      int ks = -1;
      for (int x = 0; x < n2; x++)
        if (ptc[x] == (uint8_t)((stm << 3) | KING)) ks = x;
      if (!attacked_by(ptc, sq2, n2, occ ^ bit(from), sq2[ks], stm ^ 1)) return true;
    }
  }
  return false;
}

/* calc_illegal_and_mate (tb8gen.c:664-768): illegal(c) = c to move and the
 * enemy king attacked; mates = in check with neither a legal move nor a
 * legal capture.  Writes wins[c] := illegal(c) and L_0. */
static void calc_illegal_and_mate(Table *t, uint64_t **wins, uint64_t **L0)
{
  int64_t broken[2] = { 0, 0 }, mates[2] = { 0, 0 };
#pragma omp parallel for schedule(dynamic, 1) num_threads(g_threads) reduction(+ : broken[:2], mates[:2])
  for (int s = SLICE_LO(t); s < SLICE_HI(t); s++) {
    uint64_t *iw = bm_slice(t, wins[0], s), *ib = bm_slice(t, wins[1], s);
    uint64_t *lw = bm_slice(t, L0[0], s), *lb = bm_slice(t, L0[1], s);
    uint8_t sq[MAX_PIECES] = { 0 };
    for (uint64_t idx = 0; idx < t->size; idx++) {
      sq[0] = KKSq[s][0];
      sq[1] = KKSq[s][1];
      bb_t occ = idx_to_sq(&t->ii, idx, sq);
      bool w_ill = opp_king_attacked(t->pt, sq, t->num, occ, WHITE);
      bool b_ill = opp_king_attacked(t->pt, sq, t->num, occ, BLACK);
      if (w_ill) { bm_set(iw, idx); broken[0]++; }
      if (b_ill) { bm_set(ib, idx); broken[1]++; }
      if (w_ill != b_ill) {
        int c = b_ill ? WHITE : BLACK; /* the side in check, to move */
        if (!has_legal_move_or_cap(t->pt, sq, t->num, occ, c)) {
          bm_set(c ? lb : lw, idx);
          mates[c]++;
        }
      }
    }
  }
  for (int c = 0; c < 2; c++) { t->broken[c] = broken[c]; t->mates[c] = mates[c]; }
}

/* mark_king_unmoves (tb8gen.c:246-262) */
static void mark_king_unmoves(const Table *t, int stm, bb_t occ, uint8_t *sq, int s, uint64_t *cand)
{
  uint8_t sq2[MAX_PIECES];
  bb_t b = KingAtt[sq[stm]] & ~(occ | KingAtt[sq[stm ^ 1]]);
  while (b) {
    sq[stm] = (uint8_t)pop_lsb(&b);
    normalize(sq, sq2);
    int s2 = KKMap[sq2[0]][sq2[1]];
    if (!in_table(t, s2)) continue; /* synthetic mode: leaves the slice */
    uint64_t *p = bm_slice(t, cand, s2);
    bm_set_atomic(p, sq_to_idx(&t->ii, sq2));
    if (s < 441 && s2 >= 441) {
      mirror_diagonal(sq2);
      bm_set_atomic(p, sq_to_idx(&t->ii, sq2));
    }
  }
}

/* mark_unmoves (tb8gen.c:264-276) */
static void mark_unmoves(const Table *t, int k, uint64_t *p, bb_t occ, uint8_t *sq)
{
  uint8_t sq2[MAX_PIECES];
  bb_t b = piece_att(t->pt[k], sq[k], occ) & ~occ;
  while (b) {
    sq[k] = (uint8_t)pop_lsb(&b);
    memcpy(sq2, sq, MAX_PIECES);
    bm_set_atomic(p, sq_to_idx(&t->ii, sq2));
  }
}

/* predecessors_worker (tb8gen.c:484-527): stm unmoves from the frontier of
 * stm^1 positions src, marked into cand. */
static void predecessors(const Table *t, int stm, uint64_t *src, uint64_t *cand)
{
#pragma omp parallel for schedule(dynamic, 1) num_threads(g_threads)
  for (int s = SLICE_LO(t); s < SLICE_HI(t); s++) {
    const uint64_t *p = bm_slice(t, src, s);
    uint64_t *q = bm_slice(t, cand, s);
    uint8_t sq[MAX_PIECES] = { 0 };
    for (uint64_t w = 0; w < t->words; w++) {
      uint64_t x = p[w];
      while (x) {
        uint64_t idx = w * 64 + (uint64_t)pop_lsb(&x);
        if (idx >= t->size) break;
        sq[0] = KKSq[s][0];
        sq[1] = KKSq[s][1];
        bb_t occ = idx_to_sq(&t->ii, idx, sq);
        uint8_t tmp = sq[stm];
        mark_king_unmoves(t, stm, occ, sq, s, cand);
        sq[stm] = tmp;
        for (int i = 1; t->pcs[stm][i] >= 0; i++) {
          int j = t->pcs[stm][i];
          uint8_t tmp2 = sq[j];
          mark_unmoves(t, j, q, occ, sq);
          sq[j] = tmp2;
        }
      }
    }
  }
}

/* predecessors_sub_worker (tb8gen.c:346-414): uncapture from the stm^1
 * capture-exit bitmaps (src: kind region, colour stm^1).  With nsrc == 2
 * the source is src0 & ~src1 (tb8gen.c:964-969). */
static void predecessors_sub(const Table *t, int stm, uint64_t *src0, uint64_t *src1, bool legality,
                             uint64_t *cand)
{
  int o = stm ^ 1;
  for (int k = 0; k < t->ii.numsets; k++) {
    int m = t->ii.last[k];
    if ((t->pt[m] >> 3) == stm) continue;
    const IdxInfo *ci = &t->capt[k];
    int n = t->num - 1;
    uint8_t spt[MAX_PIECES];
    memcpy(spt, t->pt, MAX_PIECES);
    spt[m] = t->pt[n];
#pragma omp parallel for schedule(dynamic, 1) num_threads(g_threads)
    for (int s = SLICE_LO(t); s < SLICE_HI(t); s++) {
      const uint64_t *p0 = sub_slice(t, src0, o, s) + t->coff[k];
      const uint64_t *p1 = src1 ? sub_slice(t, src1, o, s) + t->coff[k] : NULL;
      uint64_t *q = bm_slice(t, cand, s);
      uint8_t sq[MAX_PIECES] = { 0 };
      for (uint64_t w = 0; w < t->cwords[k]; w++) {
        uint64_t x = p0[w] & (p1 ? ~p1[w] : ~0ULL);
        while (x) {
          uint64_t idx = w * 64 + (uint64_t)pop_lsb(&x);
          if (idx >= ci->size) break;
          sq[0] = KKSq[s][0];
          sq[1] = KKSq[s][1];
          bb_t occ = idx_to_sq(ci, idx, sq);
          if (legality) {
            uint8_t lsq[MAX_PIECES];
            memcpy(lsq, sq, MAX_PIECES);
            lsq[m] = lsq[n];
            if (opp_king_attacked(spt, lsq, n, occ, o)) continue;
          }
          /* uncapture by the king */
          sq[m] = sq[stm];
          mark_king_unmoves(t, stm, occ, sq, s, cand);
          sq[stm] = sq[m];
          /* uncapture by non-king pieces */
          for (int i = 1; t->pcs[stm][i] >= 0; i++) {
            int j = t->pcs[stm][i];
            sq[m] = sq[j];
            mark_unmoves(t, j, q, occ, sq);
            sq[j] = sq[m];
          }
        }
      }
    }
  }
}

/* Synthetic king-move exit value for the opponent o after stm's king went
 * from `from` to `to` (capture of slot j if j >= 0).  Returns 3 if the
 * move is illegal (stm's king attacked afterwards), else the WDL for o. */
static int synth_king_exit(const Table *t, int stm, const uint8_t *sq, bb_t occ, int from, int j,
                           const uint8_t *sq2, int s2)
{
  int o = stm ^ 1, to = sq[stm];
  if (j >= 0) {
    uint8_t xsq[MAX_PIECES], xpt[MAX_PIECES], csq[MAX_PIECES];
    int nx = 0;
    for (int i = 0; i < t->num; i++)
      if (i != j) { xpt[nx] = t->pt[i]; xsq[nx] = sq[i]; nx++; }
    if (attacked_by(xpt, xsq, nx, occ ^ bit(from), to, o)) return 3;
    int l = t->pc_to_set[j];
    memcpy(csq, sq2, MAX_PIECES);
    if (!g_ref_kcap) csq[j] = csq[t->ii.last[l]];
    return synth_value(synth_hash_mixed(t->seedmix, o, 2, s2 * 8 + l, sq_to_idx(&t->capt[l], csq)));
  }
  if (attacked_by(t->pt, sq, t->num, occ ^ bit(from) ^ bit(to), to, o)) return 3;
  uint8_t csq[MAX_PIECES];
  memcpy(csq, sq2, MAX_PIECES);
  return synth_value(synth_hash_mixed(t->seedmix, o, 1, s2, sq_to_idx(&t->ii, csq)));
}

/* check_king_moves (tb8gen.c:537-569) */
static bool check_king_moves(const Table *t, int stm, bb_t occ, uint8_t *sq, const uint64_t *wins_o,
                             const uint64_t *subo)
{
  uint8_t sq2[MAX_PIECES];
  int o = stm ^ 1, from = sq[stm];
  bb_t b = KingAtt[from] & ~KingAtt[sq[o]];
  bb_t caps = b & occ;
  while (caps) {
    int to = pop_lsb(&caps);
    int j = 0;
    while (sq[j] != to) j++;
    if ((t->pt[j] >> 3) == stm) continue;
    sq[stm] = (uint8_t)to;
    normalize(sq, sq2);
    int l = t->pc_to_set[j];
    int s2 = KKMap[sq2[0]][sq2[1]];
    if (!in_table(t, s2)) {
      int v = synth_king_exit(t, stm, sq, occ, from, j, sq2, s2);
      if (v == 3 || v == 2) continue;
      return false;
    }
    if (!g_ref_kcap) sq2[j] = sq2[t->ii.last[l]];
    const uint64_t *p = sub_slice(t, (uint64_t *)subo, o, s2) + t->coff[l];
    if (!bm_test(p, sq_to_idx(&t->capt[l], sq2))) return false;
  }
  b &= ~occ;
  while (b) {
    sq[stm] = (uint8_t)pop_lsb(&b);
    normalize(sq, sq2);
    int s2 = KKMap[sq2[0]][sq2[1]];
    if (!in_table(t, s2)) {
      int v = synth_king_exit(t, stm, sq, occ, from, -1, sq2, s2);
      if (v == 3 || v == 2) continue;
      return false;
    }
    if (!bm_test(bm_slice(t, (uint64_t *)wins_o, s2), sq_to_idx(&t->ii, sq2))) return false;
  }
  return true;
}

/* check_moves (tb8gen.c:571-605) */
static bool check_moves(const Table *t, int k, int s, const uint64_t *wins_s, bb_t occ, uint8_t *sq,
                        const uint64_t *subo)
{
  uint8_t sq2[MAX_PIECES];
  bb_t b = piece_att(t->pt[k], sq[k], occ);
  bb_t caps = b & occ;
  while (caps) {
    int to = pop_lsb(&caps);
    int j = 0;
    while (sq[j] != to) j++;
    if (!((t->pt[k] ^ t->pt[j]) & 8)) continue;
    memcpy(sq2, sq, MAX_PIECES);
    int l = t->pc_to_set[j];
    sq2[k] = (uint8_t)to;
    sq2[j] = sq2[t->ii.last[l]];
    const uint64_t *q = sub_slice(t, (uint64_t *)subo, (t->pt[j] >> 3), s) + t->coff[l];
    if (!bm_test(q, sq_to_idx(&t->capt[l], sq2))) return false;
  }
  b &= ~occ;
  while (b) {
    sq[k] = (uint8_t)pop_lsb(&b);
    memcpy(sq2, sq, MAX_PIECES);
    if (!bm_test(wins_s, sq_to_idx(&t->ii, sq2))) return false;
  }
  return true;
}

/* check_successors_worker (tb8gen.c:607-661): keep a candidate only if it
 * is legal and every move reaches wins(o) / the sub win bitmap.  Returns
 * the number of surviving candidates. */
static int64_t check_successors(const Table *t, int stm, uint64_t *cand, const uint64_t *wins_o,
                                const uint64_t *subo)
{
  int64_t total = 0;
#pragma omp parallel for schedule(dynamic, 1) num_threads(g_threads) reduction(+ : total)
  for (int s = SLICE_LO(t); s < SLICE_HI(t); s++) {
    uint64_t *p = bm_slice(t, cand, s);
    const uint64_t *q = bm_slice(t, (uint64_t *)wins_o, s);
    uint8_t sq[MAX_PIECES] = { 0 };
    for (uint64_t w = 0; w < t->words; w++) {
      uint64_t x = p[w], x2 = x;
      while (x) {
        int bt = pop_lsb(&x);
        uint64_t idx = w * 64 + (uint64_t)bt;
        if (idx >= t->size) { x2 &= ~(1ULL << bt); continue; }
        sq[0] = KKSq[s][0];
        sq[1] = KKSq[s][1];
        bb_t occ = idx_to_sq(&t->ii, idx, sq);
        bool ok = !opp_king_attacked(t->pt, sq, t->num, occ, stm);
        for (int i = 1; ok && t->pcs[stm][i] >= 0; i++) {
          int j = t->pcs[stm][i];
          uint8_t tmp = sq[j];
          ok = check_moves(t, j, s, q, occ, sq, subo);
          sq[j] = tmp;
        }
        if (ok) {
          uint8_t tmp = sq[stm];
          ok = check_king_moves(t, stm, occ, sq, wins_o, subo);
          sq[stm] = tmp;
        }
        if (!ok) x2 &= ~(1ULL << bt);
      }
      p[w] = x2;
      total += popcnt(x2);
    }
  }
  return total;
}

/* Synthetic mode, ply 1: positions with a legal king move or king capture
 * into an exit that is a loss (W_1 candidates) or a win (L_1 candidates)
 * for the opponent. */
static void synth_king_exits(const Table *t, int stm, uint64_t *candW, uint64_t *candL)
{
  int s = t->synth_slice, o = stm ^ 1;
#pragma omp parallel for schedule(dynamic, 64) num_threads(g_threads)
  for (uint64_t w = 0; w < t->words; w++) {
    uint8_t sq[MAX_PIECES] = { 0 }, sq2[MAX_PIECES];
    uint64_t ww = 0, wl = 0;
    for (int bt = 0; bt < 64; bt++) {
      uint64_t idx = w * 64 + (uint64_t)bt;
      if (idx >= t->size) break;
      sq[0] = KKSq[s][0];
      sq[1] = KKSq[s][1];
      bb_t occ = idx_to_sq(&t->ii, idx, sq);
      int from = sq[stm];
      bb_t b = KingAtt[from] & ~KingAtt[sq[o]];
      while (b) {
        int to = pop_lsb(&b);
        int j = -1;
        if (occ & bit(to)) {
          j = 0;
          while (sq[j] != to) j++;
          if ((t->pt[j] >> 3) == stm) continue;
        }
        sq[stm] = (uint8_t)to;
        normalize(sq, sq2);
        int s2 = KKMap[sq2[0]][sq2[1]];
        int v = s2 == s ? 3 : synth_king_exit(t, stm, sq, occ, from, j, sq2, s2);
        sq[stm] = (uint8_t)from;
        if (v == -2) ww |= 1ULL << bt;
        if (v == 2) wl |= 1ULL << bt;
      }
    }
    candW[w] = ww;
    candL[w] = wl;
  }
}

static uint64_t *xalloc(uint64_t words)
{
  uint64_t *p = calloc(words ? words : 1, sizeof(uint64_t));
  if (!p) { fprintf(stderr, "tbo: out of memory\n"); abort(); }
  return p;
}

static void record(Table *t, int stm, const uint64_t *bm, uint16_t code)
{
  if (!t->code[stm]) return;
#pragma omp parallel for schedule(dynamic, 1) num_threads(g_threads)
  for (int s = SLICE_LO(t); s < SLICE_HI(t); s++) {
    const uint64_t *p = bm_slice(t, (uint64_t *)bm, s);
    uint16_t *c = t->code[stm] + (t->synth ? 0 : (uint64_t)s * t->size);
    for (uint64_t w = 0; w < t->words; w++) {
      uint64_t x = p[w];
      while (x) {
        uint64_t idx = w * 64 + (uint64_t)pop_lsb(&x);
        if (idx < t->size) c[idx] = code;
      }
    }
  }
}

static int64_t count_bm(const Table *t, const uint64_t *bm)
{
  int64_t c = 0;
  uint64_t nw = t->synth ? t->words : t->words * NSLICES;
  for (uint64_t i = 0; i < nw; i++) c += popcnt(bm[i]);
  return c;
}

/* The generator proper: the stage order and ply loop of main()
 * (tb8gen.c:1390-1444) with calc_L_n / calc_W_n (tb8gen.c:943-1084). */
static double g_loop_seconds;
static int64_t g_half_steps;

static void run_generation(Table *t, FILE *log, int max_plies)
{
  uint64_t nw = t->synth ? t->words : t->words * NSLICES;
  uint64_t *wins[2], *prevW[2], *prevL[2], *newW[2], *newL[2], *cand[2];
  for (int c = 0; c < 2; c++) {
    wins[c] = xalloc(nw); prevW[c] = xalloc(nw); prevL[c] = xalloc(nw);
    newW[c] = xalloc(nw); newL[c] = xalloc(nw); cand[c] = xalloc(nw);
    for (int q = 0; q < 4; q++)
      t->sub[c][q] = xalloc(t->synth ? t->ctotal[c] : t->ctotal[c] * NSLICES);
  }
  calc_sub(t, WHITE);
  calc_sub(t, BLACK);
  calc_illegal_and_mate(t, wins, prevL);
  for (int c = 0; c < 2; c++) {
    record(t, c, wins[c], TBO_ILLEGAL);
    record(t, c, prevL[c], TBO_LOSS(0));
  }
  if (log) {
    fprintf(log, "broken_w = %ld\nbroken_b = %ld\nl0_w = %ld\nl0_b = %ld\n", (long)t->broken[0],
            (long)t->broken[1], (long)t->mates[0], (long)t->mates[1]);
    fflush(log);
  }
  for (int k = 0; k < 2; k++)
    for (int c = 0; c < 2; c++)
      for (int n = 0; n < TBO_MAX_PLY; n++) t->cnt[k][c][n] = -1;

  /* synthetic king-move exits feed ply 1 like uncaptures */
  uint64_t *kexW[2] = { NULL, NULL }, *kexL[2] = { NULL, NULL };
  if (t->synth)
    for (int c = 0; c < 2; c++) {
      kexW[c] = xalloc(nw);
      kexL[c] = xalloc(nw);
      synth_king_exits(t, c, kexW[c], kexL[c]);
    }

  bool more_ww = true, more_wb = true, more_lw = true, more_lb = true;
  int n;
  double t_loop0 = omp_get_wtime();
  g_half_steps = 0;
  for (n = 1;; n++) {
    if (n > DRAW_RULE + 1 && !(more_ww || more_wb || more_lw || more_lb)) break;
    if (max_plies > 0 && n > max_plies) break;
    if (n >= TBO_MAX_PLY) break;
    bool forced = n == DRAW_RULE + 1;
    bool more_l[2] = { more_lw, more_lb }, more_w[2] = { more_ww, more_wb };
    bool next_w_other[2], next_l_other[2];
    /* L steps: calc_L_n (tb8gen.c:943-1026) */
    for (int stm = 0; stm < 2; stm++) {
      int o = stm ^ 1;
      memset(newL[stm], 0, nw * 8);
      if (!(forced || more_l[stm])) { next_w_other[stm] = false; continue; }
      memset(cand[stm], 0, nw * 8);
      g_half_steps++;
      if (n == 1) {
        predecessors_sub(t, stm, t->sub[o][3], NULL, true, cand[stm]);
        if (t->synth)
          for (uint64_t i = 0; i < nw; i++) cand[stm][i] |= kexL[stm][i];
      } else if (n == DRAW_RULE + 1)
        predecessors_sub(t, stm, t->sub[o][2], t->sub[o][3], false, cand[stm]);
      if (more_l[stm] && n > 1) predecessors(t, stm, prevW[o], cand[stm]);
      int64_t cnt = check_successors(t, stm, cand[stm], wins[o], t->sub[o][n <= DRAW_RULE ? 3 : 2]);
      memcpy(newL[stm], cand[stm], nw * 8);
      record(t, stm, newL[stm], TBO_LOSS(n));
      t->cnt[0][stm][n] = cnt;
      if (log) { fprintf(log, "l%d_%c = %ld\n", n, "wb"[stm], (long)cnt); fflush(log); }
      next_w_other[stm] = cnt != 0;
    }
    /* W steps: calc_W_n (tb8gen.c:1028-1084) */
    for (int stm = 0; stm < 2; stm++) {
      int o = stm ^ 1;
      memset(newW[stm], 0, nw * 8);
      if (!(forced || more_w[stm])) { next_l_other[stm] = false; continue; }
      memset(cand[stm], 0, nw * 8);
      g_half_steps++;
      if (n == 1) {
        predecessors_sub(t, stm, t->sub[o][0], NULL, false, cand[stm]);
        if (t->synth)
          for (uint64_t i = 0; i < nw; i++) cand[stm][i] |= kexW[stm][i];
      } else if (n == DRAW_RULE + 1)
        predecessors_sub(t, stm, t->sub[o][1], NULL, false, cand[stm]);
      if (more_w[stm]) predecessors(t, stm, prevL[o], cand[stm]);
      for (uint64_t i = 0; i < nw; i++) {
        newW[stm][i] = cand[stm][i] & ~wins[stm][i];
        wins[stm][i] |= newW[stm][i];
      }
      int64_t cnt = count_bm(t, newW[stm]);
      record(t, stm, newW[stm], TBO_WIN(n));
      t->cnt[1][stm][n] = cnt;
      if (log) { fprintf(log, "w%d_%c = %ld\n", n, "wb"[stm], (long)cnt); fflush(log); }
      next_l_other[stm] = cnt != 0;
    }
    /* flag update (tb8gen.c:1416-1423): an L(stm) result enables
     * W(o) next ply, a W(stm) result enables L(o) next ply */
    more_wb = next_w_other[WHITE];
    more_ww = next_w_other[BLACK];
    more_lb = next_l_other[WHITE];
    more_lw = next_l_other[BLACK];
    for (int c = 0; c < 2; c++) {
      uint64_t *x = prevW[c]; prevW[c] = newW[c]; newW[c] = x;
      x = prevL[c]; prevL[c] = newL[c]; newL[c] = x;
    }
  }
  t->nplies = n - 1;
  g_loop_seconds = omp_get_wtime() - t_loop0;
  for (int c = 0; c < 2; c++) {
    free(wins[c]); free(prevW[c]); free(prevL[c]); free(newW[c]); free(newL[c]); free(cand[c]);
    free(kexW[c]); free(kexL[c]);
  }
}

static int add_table(Table *t)
{
  if (g_ntab >= MAXTABLES) return -1;
  g_tab[g_ntab] = t;
  return g_ntab++;
}

static int generate_rec(const char *material, FILE *log)
{
  init_tables();
  int h = tbo_find(material);
  if (h >= 0) return h;
  Table *t = calloc(1, sizeof(Table));
  if (!t) return -2;
  if (setup_material(t, material)) { free(t); return -1; }
  /* sub-materials first (dependency order) */
  for (int k = 0; k < t->ii.numsets; k++) {
    uint8_t spt[MAX_PIECES];
    int n = 0;
    for (int i = 0; i < t->num; i++)
      if (i != t->ii.last[k]) spt[n++] = t->pt[i];
    if (n == 2) continue;
    char nm[16], fl[16], cn[16];
    material_string(spt, n, false, nm);
    material_string(spt, n, true, fl);
    if (tbo_find(nm) >= 0 || tbo_find(fl) >= 0) continue;
    canonical_name(nm, cn);
    if (generate_rec(cn, NULL) < 0) { free(t); return -3; }
  }
  for (int c = 0; c < 2; c++) {
    t->code[c] = calloc(t->size * NSLICES, sizeof(uint16_t));
    if (!t->code[c]) { fprintf(stderr, "tbo: out of memory\n"); abort(); }
  }
  run_generation(t, log, 0);
  return add_table(t);
}

int tbo_generate(const char *material, const char *log_path)
{
  FILE *log = NULL;
  if (log_path) log = strcmp(log_path, "-") ? fopen(log_path, "w") : stdout;
  int h = generate_rec(material, log);
  if (log && log != stdout) fclose(log);
  return h;
}

int tbo_synth_slice(const char *material, int s, uint64_t seed, const char *log_path, int max_plies)
{
  init_tables();
  if (s < 0 || s >= 441) return -1;
  Table *t = calloc(1, sizeof(Table));
  if (!t) return -2;
  if (setup_material(t, material)) { free(t); return -1; }
  t->synth = 1;
  t->synth_slice = s;
  t->seed = seed;
  t->seedmix = splitmix64(seed);
  snprintf(t->name, sizeof t->name, "#%d", g_ntab);
  for (int c = 0; c < 2 && g_keep_synth_codes; c++) {
    t->code[c] = calloc(t->size, sizeof(uint16_t));
    if (!t->code[c]) { fprintf(stderr, "tbo: out of memory\n"); abort(); }
  }
  FILE *log = NULL;
  if (log_path) log = strcmp(log_path, "-") ? fopen(log_path, "w") : stdout;
  run_generation(t, log, max_plies);
  if (log && log != stdout) fclose(log);
  return add_table(t);
}

void tbo_free_all(void)
{
  for (int i = 0; i < g_ntab; i++) {
    Table *t = g_tab[i];
    for (int c = 0; c < 2; c++) {
      free(t->code[c]);
      for (int q = 0; q < 4; q++) free(t->sub[c][q]);
    }
    free(t);
  }
  g_ntab = 0;
}

/* ------------------------------------------------------------------ */
/* Accessors.                                                           */

static Table *T(int h) { return (h >= 0 && h < g_ntab) ? g_tab[h] : NULL; }
uint64_t tbo_slice_size(int h) { return T(h) ? T(h)->size : 0; }
int tbo_num_pieces(int h) { return T(h) ? T(h)->num : -1; }
void tbo_piece_types(int h, uint8_t *pt) { if (T(h)) memcpy(pt, T(h)->pt, MAX_PIECES); }
int tbo_num_sets(int h) { return T(h) ? T(h)->ii.numsets : -1; }
int tbo_capt_size(int h, int set, uint64_t *size)
{
  if (!T(h) || set < 0 || set >= T(h)->ii.numsets) return -1;
  *size = T(h)->capt[set].size;
  return 0;
}
const uint16_t *tbo_codes(int h, int stm) { return T(h) ? T(h)->code[stm & 1] : NULL; }
int tbo_num_plies(int h) { return T(h) ? T(h)->nplies : -1; }
int64_t tbo_count(int h, int kind, int stm, int n)
{
  if (!T(h) || n < 0 || n >= TBO_MAX_PLY) return -1;
  return T(h)->cnt[kind & 1][stm & 1][n];
}
int64_t tbo_broken(int h, int stm) { return T(h) ? T(h)->broken[stm & 1] : -1; }
int64_t tbo_mates(int h, int stm) { return T(h) ? T(h)->mates[stm & 1] : -1; }

uint64_t tbo_sq_to_idx(int h, const uint8_t *sq)
{
  uint8_t s[MAX_PIECES] = { 0 };
  memcpy(s, sq, (size_t)T(h)->num);
  return sq_to_idx(&T(h)->ii, s);
}

int tbo_idx_to_sq(int h, int s, uint64_t idx, uint8_t *sq)
{
  if (!T(h) || s < 0 || s >= NSLICES || idx >= T(h)->size) return -1;
  uint8_t x[MAX_PIECES] = { 0 };
  x[0] = KKSq[s][0];
  x[1] = KKSq[s][1];
  idx_to_sq(&T(h)->ii, idx, x);
  memcpy(sq, x, (size_t)T(h)->num);
  return 0;
}

int tbo_capt_idx_to_sq(int h, int set, int s, uint64_t idx, uint8_t *sq)
{
  if (!T(h) || set < 0 || set >= T(h)->ii.numsets || idx >= T(h)->capt[set].size) return -1;
  uint8_t x[MAX_PIECES] = { 0 };
  x[0] = KKSq[s][0];
  x[1] = KKSq[s][1];
  idx_to_sq(&T(h)->capt[set], idx, x);
  memcpy(sq, x, (size_t)T(h)->num);
  return 0;
}

uint64_t tbo_sub_words(int h, int colour) { return T(h) ? T(h)->ctotal[colour & 1] : 0; }
uint64_t tbo_sub_offset_words(int h, int set) { return T(h) ? T(h)->coff[set] : 0; }
int tbo_sub_bitmap(int h, int kind, int colour, int s, uint64_t *dst)
{
  Table *t = T(h);
  if (!t || !t->sub[colour & 1][kind & 3]) return -1;
  memcpy(dst, sub_slice(t, t->sub[colour & 1][kind & 3], colour & 1, s), t->ctotal[colour & 1] * 8);
  return 0;
}

int tbo_lookup(int h, int stm, const uint8_t *sq)
{
  if (!T(h) || T(h)->synth) return -1;
  return table_lookup(T(h), stm & 1, sq);
}

void tbo_lookup_bulk(int h, int stm, uint64_t n, const uint8_t *sq, uint16_t *out)
{
  Table *t = T(h);
  int num = t->num;
#pragma omp parallel for schedule(static) num_threads(g_threads)
  for (uint64_t i = 0; i < n; i++) {
    const uint8_t *p = sq + i * (uint64_t)num;
    bool bad = false;
    for (int a = 0; a < num && !bad; a++)
      for (int b = a + 1; b < num; b++)
        if (p[a] == p[b]) { bad = true; break; }
    out[i] = bad ? 0xffff : table_lookup(t, stm, p);
  }
}

/* CPU baseline sample: generate the smaller materials (untimed), then run
 * the material's sub stage, init and at most max_plies plies; reports the
 * seconds spent in the ply loop and the half-steps it executed.  The table
 * is discarded. */
int tbo_time_plies(const char *material, int max_plies, double *loop_seconds, int64_t *half_steps)
{
  init_tables();
  Table *t = calloc(1, sizeof(Table));
  if (!t) return -2;
  if (setup_material(t, material)) { free(t); return -1; }
  for (int k = 0; k < t->ii.numsets; k++) {
    uint8_t spt[MAX_PIECES];
    int n = 0;
    for (int i = 0; i < t->num; i++)
      if (i != t->ii.last[k]) spt[n++] = t->pt[i];
    if (n == 2) continue;
    char nm[16], fl[16], cn[16];
    material_string(spt, n, false, nm);
    material_string(spt, n, true, fl);
    if (tbo_find(nm) >= 0 || tbo_find(fl) >= 0) continue;
    canonical_name(nm, cn);
    if (generate_rec(cn, NULL) < 0) { free(t); return -3; }
  }
  run_generation(t, NULL, max_plies);
  if (loop_seconds) *loop_seconds = g_loop_seconds;
  if (half_steps) *half_steps = g_half_steps;
  for (int c = 0; c < 2; c++)
    for (int q = 0; q < 4; q++) free(t->sub[c][q]);
  free(t);
  return 0;
}

/* ------------------------------------------------------------------ */
/* Sampled forward check of a synthetic slice too large to generate     */
/* here (the 8-piece slice): plies 1 and 2 recomputed position by       */
/* position from the moves out of each sampled position, the way the     */
/* recurrences of calc_L_n / calc_W_n (tb8gen.c:943-1084) read forward.  */
/* The engine's own ply-1 planes stand in for the targets' ply-1 sets.   */

/* One legal move of stm out of position sq (slice s) and what it reaches:
 * kind 0 = quiet move inside the slice to index *tgt; kind 1 = a terminal
 * value *val for o (a capture inside the slice, or a king move or king
 * capture out of the slice, valued by the seeded hash).  Returns the
 * number of legal moves; fn receives each. */
typedef void (*move_fn)(void *u, int kind, uint64_t tgt, int val, const uint8_t *q, bb_t qocc);

static int synth_moves(const Table *t, int stm, int s, const uint8_t *sq_in, bb_t occ, move_fn fn,
                       void *u)
{
  int o = stm ^ 1, nlegal = 0;
  uint8_t sq[MAX_PIECES], sq2[MAX_PIECES];
  memcpy(sq, sq_in, MAX_PIECES);
  /* non-king pieces: captures (check_moves, tb8gen.c:571-590), then quiet moves */
  for (int i = 1; t->pcs[stm][i] >= 0; i++) {
    int k = t->pcs[stm][i], from = sq[k];
    bb_t b = piece_att(t->pt[k], from, occ);
    bb_t caps = b & occ;
    while (caps) {
      int to = pop_lsb(&caps), j = 0;
      while (sq[j] != to) j++;
      if (!((t->pt[k] ^ t->pt[j]) & 8)) continue;
      /* legal iff stm's king is not attacked once piece j is gone */
      uint8_t xpt[MAX_PIECES], xsq[MAX_PIECES];
      int nx = 0;
      for (int x = 0; x < t->num; x++)
        if (x != j) { xpt[nx] = t->pt[x]; xsq[nx] = x == k ? (uint8_t)to : sq[x]; nx++; }
      if (attacked_by(xpt, xsq, nx, occ ^ bit(from), sq[stm], o)) continue;
      nlegal++;
      int l = t->pc_to_set[j];
      memcpy(sq2, sq, MAX_PIECES);
      sq2[k] = (uint8_t)to;
      sq2[j] = sq2[t->ii.last[l]];
      uint64_t ci = sq_to_idx(&t->capt[l], sq2);
      fn(u, 1, 0, synth_value(synth_hash_mixed(t->seedmix, o, 0, l, ci)), NULL, 0);
    }
    bb_t q = b & ~occ;
    while (q) {
      int to = pop_lsb(&q);
      sq[k] = (uint8_t)to;
      bb_t occ2 = occ ^ bit(from) ^ bit(to);
      if (!attacked_by(t->pt, sq, t->num, occ2, sq[stm], o)) {
        nlegal++;
        memcpy(sq2, sq, MAX_PIECES);
        fn(u, 0, sq_to_idx(&t->ii, sq2), 0, sq, occ2);
      }
      sq[k] = (uint8_t)from;
    }
  }
  /* the king (check_king_moves, tb8gen.c:537-569) */
  int from = sq[stm];
  bb_t b = KingAtt[from] & ~KingAtt[sq[o]];
  while (b) {
    int to = pop_lsb(&b), j = -1;
    if (occ & bit(to)) {
      j = 0;
      while (sq[j] != to) j++;
      if ((t->pt[j] >> 3) == stm) continue;
    }
    sq[stm] = (uint8_t)to;
    normalize(sq, sq2);
    int s2 = KKMap[sq2[0]][sq2[1]];
    if (s2 != s) {
      int v = synth_king_exit(t, stm, sq, occ, from, j, sq2, s2);
      if (v != 3) { nlegal++; fn(u, 1, 0, v, NULL, 0); }
    } else if (j >= 0) {
      uint8_t xpt[MAX_PIECES], xsq[MAX_PIECES];
      int nx = 0;
      for (int x = 0; x < t->num; x++)
        if (x != j) { xpt[nx] = t->pt[x]; xsq[nx] = sq[x]; nx++; }
      if (!attacked_by(xpt, xsq, nx, occ ^ bit(from), to, o)) {
        nlegal++;
        int l = t->pc_to_set[j];
        sq2[j] = sq2[t->ii.last[l]];
        uint64_t ci = sq_to_idx(&t->capt[l], sq2);
        fn(u, 1, 0, synth_value(synth_hash_mixed(t->seedmix, o, 0, l, ci)), NULL, 0);
      }
    } else {
      bb_t occ2 = occ ^ bit(from) ^ bit(to);
      if (!attacked_by(t->pt, sq, t->num, occ2, to, o)) {
        nlegal++;
        fn(u, 0, sq_to_idx(&t->ii, sq2), 0, sq, occ2);
      }
    }
    sq[stm] = (uint8_t)from;
  }
  return nlegal;
}

typedef struct {
  const Table *t;
  int stm;
  const uint64_t *W1o, *L1o;
  /* any legal move to a loss for o: via a mate (L_0), a terminal -2, or
   * a quiet move into the engine's L_1(o); every legal move to a win for
   * o: terminal 2 (ply 1), or terminal 2 / quiet into W_1(o) (ply 2) */
  bool w1, all_win1, w2, all_win2, any_w1o;
} Fwd;

static void fwd_move(void *u, int kind, uint64_t tgt, int val, const uint8_t *q, bb_t qocc)
{
  Fwd *f = (Fwd *)u;
  const Table *t = f->t;
  int o = f->stm ^ 1;
  if (kind == 1) {
    if (val == -2) f->w1 = true;
    if (val != 2) f->all_win1 = f->all_win2 = false;
    return;
  }
  /* a quiet move inside the slice: o to move at q, which is legal for o */
  f->all_win1 = false;
  if (opp_king_attacked(t->pt, q, t->num, qocc, f->stm) &&
      !has_legal_move_or_cap(t->pt, q, t->num, qocc, o))
    f->w1 = true; /* q in L_0(o): a mate */
  if (bm_test(f->L1o, tgt)) f->w2 = true;
  if (bm_test(f->W1o, tgt)) f->any_w1o = true;
  else f->all_win2 = false;
}

static uint64_t rng_next(uint64_t *st)
{
  *st += 0x9e3779b97f4a7c15ULL;
  uint64_t z = *st;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

/* A random set bit of plane (a random word, then the next nonzero word at
 * most 2^20 words on, wrapping); 0 when none was found. */
static int pick_bit(const uint64_t *p, uint64_t words, uint64_t size, uint64_t *st, uint64_t *idx)
{
  uint64_t w = rng_next(st) % words;
  for (uint64_t k = 0; k < (1u << 20); k++) {
    uint64_t x = p[w];
    if (x) {
      int n = popcnt(x), r = (int)(rng_next(st) % (uint64_t)n);
      uint64_t i = w * 64 + (uint64_t)select_bit(x, r);
      if (i < size) { *idx = i; return 1; }
    }
    w = w + 1 == words ? 0 : w + 1;
  }
  return 0;
}

int64_t tbo_synth_check(const char *material, int s, uint64_t seed, const uint64_t *const *planes,
                        uint64_t nsamples, uint64_t sample_seed, uint64_t *checked,
                        int64_t *first_bad)
{
  init_tables();
  if (s < 0 || s >= 441) return -1;
  Table tab;
  memset(&tab, 0, sizeof tab);
  Table *t = &tab;
  if (setup_material(t, material)) return -1;
  t->synth = 1;
  t->synth_slice = s;
  t->seed = seed;
  t->seedmix = splitmix64(seed);
  /* planes: wins1[2], W1[2], L1[2] after ply 1, W2[2], L2[2] after ply 2 */
  const uint64_t *wins1[2] = { planes[0], planes[1] }, *W1[2] = { planes[2], planes[3] };
  const uint64_t *L1[2] = { planes[4], planes[5] }, *W2[2] = { planes[6], planes[7] };
  const uint64_t *L2[2] = { planes[8], planes[9] };
  int64_t bad = 0;
  uint64_t nchk = 0;
  if (first_bad) first_bad[0] = -1;
  /* sample list: per side, half uniform, half set bits of the engine's
   * W1, L1, W2, L2 planes of that side (an eighth each) */
  uint64_t total = 2 * nsamples;
  uint64_t *sidx = xalloc(total);
  uint64_t st = sample_seed;
  for (uint64_t i = 0; i < total; i++) {
    int stm = (int)(i / nsamples);
    uint64_t j = i % nsamples, idx = rng_next(&st) % t->size;
    if (j >= nsamples / 2) {
      const uint64_t *pl[4] = { W1[stm], L1[stm], W2[stm], L2[stm] };
      pick_bit(pl[j % 4], t->words, t->size, &st, &idx);
    }
    sidx[i] = idx;
  }
#pragma omp parallel for schedule(dynamic, 64) num_threads(g_threads) reduction(+ : bad, nchk)
  for (uint64_t i = 0; i < total; i++) {
    int stm = (int)(i / nsamples), o = stm ^ 1;
    uint64_t idx = sidx[i];
    uint8_t sq[MAX_PIECES] = { 0 };
    sq[0] = KKSq[s][0];
    sq[1] = KKSq[s][1];
    bb_t occ = idx_to_sq(&t->ii, idx, sq);
    bool ill = opp_king_attacked(t->pt, sq, t->num, occ, stm);
    Fwd f = { t, stm, W1[o], L1[o], false, true, false, true, false };
    int nl = ill ? 0 : synth_moves(t, stm, s, sq, occ, fwd_move, &f);
    bool w1 = !ill && f.w1;
    bool l1 = !ill && nl > 0 && f.all_win1;
    bool w2 = !ill && !w1 && f.w2;
    bool l2 = !ill && f.any_w1o && f.all_win2;
    bool exp[5] = { ill || w1, w1, l1, w2, l2 };
    const uint64_t *got[5] = { wins1[stm], W1[stm], L1[stm], W2[stm], L2[stm] };
    nchk++;
    for (int k = 0; k < 5; k++) {
      bool g = bm_test(got[k], idx);
      if (g != exp[k]) {
        bad++;
#pragma omp critical
        if (first_bad && first_bad[0] < 0) {
          first_bad[0] = stm;
          first_bad[1] = k;
          first_bad[2] = (int64_t)idx;
          first_bad[3] = g;
        }
      }
    }
  }
  free(sidx);
  if (checked) *checked = nchk;
  return bad;
}

/* CPU baseline sample of the init stage on a slice too large to generate
 * here: calc_illegal_and_mate (tb8gen.c:664-768) and the synthetic
 * king-move exits, position by position over indices [lo, lo + n) of slice
 * s, into local bitmaps.  Returns the wall seconds; counts[4] = illegal W,
 * illegal B, mates W, mates B in the range. */
double tbo_synth_init_time(const char *material, int s, uint64_t seed, uint64_t lo, uint64_t n,
                           int64_t *counts)
{
  init_tables();
  Table tab;
  memset(&tab, 0, sizeof tab);
  Table *t = &tab;
  if (s < 0 || s >= 441 || setup_material(t, material)) return -1.0;
  t->synth = 1;
  t->synth_slice = s;
  t->seed = seed;
  t->seedmix = splitmix64(seed);
  if (lo >= t->size) return -1.0;
  if (n > t->size - lo) n = t->size - lo;
  uint64_t nw = (n + 63) / 64;
  uint64_t *bm = xalloc(8 * nw);
  int64_t c0 = 0, c1 = 0, c2 = 0, c3 = 0;
  double t0 = omp_get_wtime();
#pragma omp parallel for schedule(dynamic, 64) num_threads(g_threads) reduction(+ : c0, c1, c2, c3)
  for (uint64_t w = 0; w < nw; w++) {
    uint8_t sq[MAX_PIECES] = { 0 }, sq2[MAX_PIECES];
    uint64_t out[8] = { 0 };
    for (int bt = 0; bt < 64; bt++) {
      uint64_t i = w * 64 + (uint64_t)bt;
      if (i >= n) break;
      sq[0] = KKSq[s][0];
      sq[1] = KKSq[s][1];
      bb_t occ = idx_to_sq(&t->ii, lo + i, sq);
      bool ill[2];
      for (int c = 0; c < 2; c++) ill[c] = opp_king_attacked(t->pt, sq, t->num, occ, c);
      if (ill[0]) { out[0] |= 1ULL << bt; c0++; }
      if (ill[1]) { out[1] |= 1ULL << bt; c1++; }
      if (ill[0] != ill[1]) {
        int c = ill[1] ? WHITE : BLACK;
        if (!has_legal_move_or_cap(t->pt, sq, t->num, occ, c)) {
          out[2 + c] |= 1ULL << bt;
          if (c) c3++; else c2++;
        }
      }
      /* the exits of synth_king_exits */
      for (int stm = 0; stm < 2; stm++) {
        int o = stm ^ 1, from = sq[stm];
        bb_t b = KingAtt[from] & ~KingAtt[sq[o]];
        while (b) {
          int to = pop_lsb(&b), j = -1;
          if (occ & bit(to)) {
            j = 0;
            while (sq[j] != to) j++;
            if ((t->pt[j] >> 3) == stm) continue;
          }
          sq[stm] = (uint8_t)to;
          normalize(sq, sq2);
          int s2 = KKMap[sq2[0]][sq2[1]];
          int v = s2 == s ? 3 : synth_king_exit(t, stm, sq, occ, from, j, sq2, s2);
          sq[stm] = (uint8_t)from;
          if (v == -2) out[4 + stm] |= 1ULL << bt;
          if (v == 2) out[6 + stm] |= 1ULL << bt;
        }
      }
    }
    for (int k = 0; k < 8; k++) bm[k * nw + w] = out[k];
  }
  double sec = omp_get_wtime() - t0;
  free(bm);
  if (counts) { counts[0] = c0; counts[1] = c1; counts[2] = c2; counts[3] = c3; }
  return sec;
}

int tbo_get_threads(void) { return g_threads; }

/* ------------------------------------------------------------------ */
/* Syzygy WDL files: a restatement of the reader the reference probes    */
/* its capture exits with (probe.c), for pawnless tables.  Used to read  */
/* the engine's final-table writer output.  The piece-encoding tables     */
/* (Triangle, OffDiag, Diag, Lower, probe.c:300-440) are built by rule.   */

static uint8_t SzTri[64], SzDiag[64], SzLower[64], SzInvDiag[16], SzInvLower[36];
static int8_t SzOffDiag[64];

static void sz_tables(void)
{
  static int ready;
  if (ready) return;
  init_tables();
  for (int sq = 0; sq < 64; sq++) {
    int f = sq & 7, r = sq >> 3;
    /* Triangle: the a1-d1-d4 class of sq under the board symmetries */
    int cf = f > 3 ? 7 - f : f, cr = r > 3 ? 7 - r : r;
    if (cr > cf) { int x = cf; cf = cr; cr = x; }
    SzTri[sq] = (uint8_t)TriOf[cr * 8 + cf];
    SzOffDiag[sq] = (int8_t)(r > f ? 1 : r < f ? -1 : 0);
    SzDiag[sq] = (uint8_t)(r == f ? f : r + f == 7 ? 8 + r : 0);
    /* Lower: squares below a1-h8 numbered rank by rank (0..27), squares
     * above it as their mirror image, a1-h8 itself 28..35 */
    int lf = r > f ? r : f, lr = r > f ? f : r;
    SzLower[sq] = (uint8_t)(lf == lr ? 28 + lf : lr * 7 - lr * (lr - 1) / 2 + (lf - lr - 1));
  }
  for (int i = 0; i < 8; i++) {
    SzInvDiag[i] = (uint8_t)(9 * i);
    SzInvDiag[8 + i] = (uint8_t)(7 * (i + 1));
  }
  for (int sq = 0; sq < 64; sq++)
    if ((sq >> 3) <= (sq & 7)) SzInvLower[SzLower[sq]] = (uint8_t)sq;
  ready = 1;
}

typedef struct {
  int constant, value;                 /* a constant sub-table (flag 0x80) */
  int block_size, idx_bits, min_len, max_len;
  uint32_t num_syms;
  uint64_t base[64];                   /* by code length */
  uint32_t offset[64];
  uint8_t symlen[4096];
  const uint8_t *sympat, *index_table, *size_table, *data;
  uint64_t tb_size;
  uint8_t pieces[MAX_PIECES], norm[MAX_PIECES];
  uint64_t factor[MAX_PIECES];
} SzSub;

typedef struct {
  uint8_t *file;
  uint64_t len;
  int num, kk_enc, symmetric, split;
  int wcnt[16];                        /* the name's piece counts, colour bit 8 */
  SzSub sub[2];
} SzFile;

#define SZ_MAX 16
static SzFile *g_sz[SZ_MAX];

static uint32_t rd16(const uint8_t *p) { return (uint32_t)p[0] | ((uint32_t)p[1] << 8); }
static uint32_t rd32(const uint8_t *p) { return rd16(p) | (rd16(p + 2) << 16); }

/* calc_symlen (probe.c:1045-1059) */
static void sz_symlen(SzSub *d, uint32_t s, bool *done)
{
  const uint8_t *w = d->sympat + 3 * s;
  uint32_t s2 = ((uint32_t)w[2] << 4) | (w[1] >> 4);
  if (s2 == 0xfff) d->symlen[s] = 0;
  else {
    uint32_t s1 = ((uint32_t)(w[1] & 0xf) << 8) | w[0];
    if (!done[s1]) sz_symlen(d, s1, done);
    if (!done[s2]) sz_symlen(d, s2, done);
    d->symlen[s] = (uint8_t)(d->symlen[s1] + d->symlen[s2] + 1);
  }
  done[s] = true;
}

/* set_enc_info (probe.c:725-770), pawnless */
static uint64_t sz_enc_info(SzSub *e, const SzFile *f, const uint8_t *tb, int shift)
{
  for (int i = 0; i < f->num; i++) {
    e->pieces[i] = (tb[i + 1] >> shift) & 0x0f;
    e->norm[i] = 0;
  }
  int order = (tb[0] >> shift) & 0x0f;
  int k = e->norm[0] = (uint8_t)(f->kk_enc ? 2 : 3);
  for (int i = k; i < f->num; i += e->norm[i])
    for (int j = i; j < f->num && e->pieces[j] == e->pieces[i]; j++) e->norm[i]++;
  int n = 64 - k;
  uint64_t fac = 1;
  for (int i = 0; k < f->num || i == order; i++) {
    if (i == order) {
      e->factor[0] = fac;
      fac *= f->kk_enc ? 462 : 31332;
    } else {
      e->factor[k] = fac;
      fac *= Binom[e->norm[k]][n];
      n -= e->norm[k];
      k += e->norm[k];
    }
  }
  return fac;
}

/* setup_pairs + setup_huffman (probe.c:1061-1208): parses one sub-table's
 * header at *pp; returns -1 for a form this restatement does not read
 * (rANS).  sizes[] = index, size-table and data bytes. */
static int sz_pairs(SzSub *d, const uint8_t **pp, uint64_t *sizes)
{
  const uint8_t *p = *pp;
  sizes[0] = sizes[1] = sizes[2] = 0;
  if (p[0] & 0x80) {
    d->constant = 1;
    d->value = p[1];
    *pp = p + 2;
    return 0;
  }
  if (p[0] & 0x40) return -1;
  d->constant = 0;
  d->block_size = p[1];
  d->idx_bits = p[2];
  uint32_t real_blocks = rd32(p + 4), nblocks = real_blocks + p[3];
  d->max_len = p[8];
  d->min_len = p[9];
  int h = d->max_len - d->min_len + 1;
  if (h < 1 || d->max_len > 32) return -1;
  uint64_t tmp[64] = { 0 };
  uint32_t off[64];
  for (int i = 0; i < h; i++) off[i] = rd16(p + 10 + 2 * i);
  d->num_syms = rd16(p + 10 + 2 * h);
  d->sympat = p + 12 + 2 * h;
  for (int i = h - 2; i >= 0; i--) tmp[i] = (tmp[i + 1] + off[i] - off[i + 1]) / 2;
  for (int i = 0; i < h; i++) {
    d->base[d->min_len + i] = tmp[i] << (64 - (d->min_len + i));
    d->offset[d->min_len + i] = off[i];
  }
  *pp = p + 12 + 2 * h + 3 * d->num_syms + (d->num_syms & 1);
  bool done[4096] = { false };
  for (uint32_t s = 0; s < d->num_syms; s++)
    if (!done[s]) sz_symlen(d, s, done);
  sizes[0] = 6 * ((d->tb_size + (1ULL << d->idx_bits) - 1) >> d->idx_bits);
  sizes[1] = 2ULL * nblocks;
  sizes[2] = (uint64_t)real_blocks << d->block_size;
  return 0;
}

/* 64 bits of the block's bit stream from bit position pos, most
 * significant first (read_be_u64, util.h:75) */
static uint64_t sz_bits(const uint8_t *p, uint64_t pos)
{
  const uint8_t *b = p + (pos >> 3);
  uint64_t x = 0;
  for (int i = 0; i < 8; i++) x = (x << 8) | b[i];
  int sh = (int)(pos & 7);
  return sh ? (x << sh) | (b[8] >> (8 - sh)) : x;
}

/* decompress_pairs / decompress_huff (probe.c:1388-1470): the stored byte
 * of position idx */
static int sz_value(const SzSub *d, uint64_t idx)
{
  if (d->constant) return d->value;
  uint32_t main_idx = (uint32_t)(idx >> d->idx_bits);
  int litidx = (int)(idx & ((1u << d->idx_bits) - 1)) - (int)(1u << (d->idx_bits - 1));
  uint32_t block = rd32(d->index_table + 6 * (uint64_t)main_idx);
  litidx += (int)rd16(d->index_table + 6 * (uint64_t)main_idx + 4);
  if (litidx < 0)
    while (litidx < 0) litidx += (int)rd16(d->size_table + 2 * (uint64_t)--block) + 1;
  else
    while (litidx > (int)rd16(d->size_table + 2 * (uint64_t)block))
      litidx -= (int)rd16(d->size_table + 2 * (uint64_t)block++) + 1;
  const uint8_t *ptr = d->data + ((uint64_t)block << d->block_size);
  uint64_t pos = 0;
  uint32_t sym;
  for (;;) {
    uint64_t code = sz_bits(ptr, pos);
    int l = d->min_len;
    while (code < d->base[l]) l++;
    sym = d->offset[l] + (uint32_t)((code - d->base[l]) >> (64 - l));
    if (litidx < (int)d->symlen[sym] + 1) break;
    litidx -= (int)d->symlen[sym] + 1;
    pos += (uint64_t)l;
  }
  while (d->symlen[sym] != 0) {
    const uint8_t *w = d->sympat + 3 * sym;
    uint32_t s1 = ((uint32_t)(w[1] & 0xf) << 8) | w[0];
    if (litidx < (int)d->symlen[s1] + 1) sym = s1;
    else {
      litidx -= (int)d->symlen[s1] + 1;
      sym = ((uint32_t)w[2] << 4) | (w[1] >> 4);
    }
  }
  return d->sympat[3 * sym];
}

/* Opens <path> as the WDL file of material `name` (e.g. "KQvKR", white as
 * the file's name has it): init_table (probe.c:1210-1345) and the entry
 * set-up of detect_tb (probe.c:143-191).  Returns a handle, or < 0. */
int tbo_rtbw_open(const char *path, const char *name)
{
  sz_tables();
  int h = -1;
  for (int i = 0; i < SZ_MAX; i++)
    if (!g_sz[i]) { h = i; break; }
  if (h < 0) return -2;
  FILE *F = fopen(path, "rb");
  if (!F) return -3;
  fseek(F, 0, SEEK_END);
  long len = ftell(F);
  fseek(F, 0, SEEK_SET);
  SzFile *f = calloc(1, sizeof(SzFile));
  f->file = calloc((size_t)len + 128, 1);  /* slack for reads past a block */
  f->len = (uint64_t)len;
  if (fread(f->file, 1, (size_t)len, F) != (size_t)len) { fclose(F); free(f->file); free(f); return -3; }
  fclose(F);
  int color = 0, cnt[16] = { 0 }, num = 0;
  for (const char *s = name; *s; s++) {
    if (*s == 'v') { color = 8; continue; }
    for (int i = KNIGHT; i <= KING; i++)
      if (*s == PieceChar[i]) { cnt[i | color]++; num++; }
  }
  memcpy(f->wcnt, cnt, sizeof cnt);
  f->num = num;
  int uniq = 0;
  for (int i = 0; i < 16; i++) uniq += cnt[i] == 1;
  f->kk_enc = uniq == 2;
  f->symmetric = 1;
  for (int i = 0; i < 8; i++) f->symmetric &= cnt[i] == cnt[i + 8];
  const uint8_t *d = f->file;
  if (rd32(d) != 0x5d23e871u) { free(f->file); free(f); return -4; }
  f->split = d[4] & 1;
  d += 5;
  uint64_t sz0 = sz_enc_info(&f->sub[0], f, d, 0);
  f->sub[0].tb_size = sz0;
  if (f->split) f->sub[1].tb_size = sz_enc_info(&f->sub[1], f, d, 4);
  d += f->num + 1;
  d += (uintptr_t)(d - f->file) & 1;
  uint64_t sizes[2][3];
  for (int t = 0; t < 1 + f->split; t++)
    if (sz_pairs(&f->sub[t], &d, sizes[t])) { free(f->file); free(f); return -5; }
  for (int t = 0; t < 1 + f->split; t++) { f->sub[t].index_table = d; d += sizes[t][0]; }
  for (int t = 0; t < 1 + f->split; t++) { f->sub[t].size_table = d; d += sizes[t][1]; }
  for (int t = 0; t < 1 + f->split; t++) {
    d = f->file + (((uint64_t)(d - f->file) + 63) & ~63ULL);
    f->sub[t].data = d;
    d += sizes[t][2];
  }
  if ((uint64_t)(d - f->file) > f->len) { free(f->file); free(f); return -6; }
  g_sz[h] = f;
  return h;
}

void tbo_rtbw_close(int h)
{
  if (h < 0 || h >= SZ_MAX || !g_sz[h]) return;
  free(g_sz[h]->file);
  free(g_sz[h]);
  g_sz[h] = NULL;
}

uint64_t tbo_rtbw_size(int h, int side)
{
  if (h < 0 || h >= SZ_MAX || !g_sz[h] || side < 0 || side > g_sz[h]->split) return 0;
  return g_sz[h]->sub[side].tb_size;
}

/* encode, PIECE_ENC (probe.c:603-687); p[] in header order, modified */
static uint64_t sz_encode(const SzSub *e, const SzFile *f, uint8_t *p)
{
  int n = f->num, k;
  uint64_t idx;
  if (p[0] & 0x04)
    for (int i = 0; i < n; i++) p[i] ^= 0x07;
  if (p[0] & 0x20)
    for (int i = 0; i < n; i++) p[i] ^= 0x38;
  for (int i = 0; i < n; i++)
    if (SzOffDiag[p[i]]) {
      if (SzOffDiag[p[i]] > 0 && i < (f->kk_enc ? 2 : 3))
        for (int j = 0; j < n; j++) p[j] = FlipDiag[p[j]];
      break;
    }
  if (f->kk_enc) {
    idx = (uint64_t)KKIdx[SzTri[p[0]]][p[1]];
    k = 2;
  } else {
    int s1 = p[1] > p[0], s2 = (p[2] > p[0]) + (p[2] > p[1]);
    if (SzOffDiag[p[0]])
      idx = (uint64_t)SzTri[p[0]] * 63 * 62 + (uint64_t)(p[1] - s1) * 62 + (uint64_t)(p[2] - s2);
    else if (SzOffDiag[p[1]])
      idx = 6 * 63 * 62 + (uint64_t)SzDiag[p[0]] * 28 * 62 + (uint64_t)SzLower[p[1]] * 62 +
            (uint64_t)(p[2] - s2);
    else if (SzOffDiag[p[2]])
      idx = 6 * 63 * 62 + 4 * 28 * 62 + (uint64_t)SzDiag[p[0]] * 7 * 28 +
            (uint64_t)(SzDiag[p[1]] - s1) * 28 + SzLower[p[2]];
    else
      idx = 6 * 63 * 62 + 4 * 28 * 62 + 4 * 7 * 28 + (uint64_t)SzDiag[p[0]] * 7 * 6 +
            (uint64_t)(SzDiag[p[1]] - s1) * 6 + (uint64_t)(SzDiag[p[2]] - s2);
    k = 3;
  }
  idx *= e->factor[0];
  bb_t occ = 0;
  for (int i = 0; i < k; i++) occ |= bit(p[i]);
  while (k < n) {
    int t = k + e->norm[k];
    for (int a = k; a < t; a++)
      for (int b = a + 1; b < t; b++)
        if (p[a] > p[b]) { uint8_t x = p[a]; p[a] = p[b]; p[b] = x; }
    uint64_t s = 0;
    for (int i = k; i < t; i++) s += Binom[i - k + 1][p[i] - popcnt(occ & (bit(p[i]) - 1))];
    idx += s * e->factor[k];
    for (; k < t; k++) occ |= bit(p[k]);
  }
  return idx;
}

/* decode_helper + decode, PIECE_ENC (probe.c:859-980): squares in header
 * order of sub-table index idx; 0, or -1 for an index outside the table */
static int sz_decode(const SzSub *e, const SzFile *f, uint64_t idx, uint8_t *p)
{
  if (idx >= e->tb_size) return -1;
  int n = f->num, lead = f->kk_enc ? 2 : 3;
  /* mixed radix in any digit order: digit = (idx / factor) % range */
  uint64_t q = (idx / e->factor[0]) % (lead == 2 ? 462 : 31332);
  bb_t occ = 0;
  if (lead == 2) {
    p[0] = KKSq[q][0];
    p[1] = KKSq[q][1];
    occ = bit(p[0]) | bit(p[1]);
  } else {
    const uint64_t A = 6 * 63 * 62, B = 4 * 28 * 62, C = 4 * 7 * 28;
    const bb_t DIAG = 0x8040201008040201ULL;
    uint64_t r;
    if (q < A) {
      r = q / (63 * 62); q -= r * 63 * 62;
      p[0] = InvTri[r];
      r = q / 62; q -= r * 62;
      p[1] = (uint8_t)(r + (r >= p[0]));
      occ = bit(p[0]) | bit(p[1]);
      p[2] = (uint8_t)select_bit(~occ, (int)q);
    } else if (q < A + B) {
      q -= A;
      r = q / (28 * 62); q -= r * 28 * 62;
      p[0] = SzInvDiag[r];
      r = q / 62; q -= r * 62;
      p[1] = SzInvLower[r];
      occ = bit(p[0]) | bit(p[1]);
      p[2] = (uint8_t)select_bit(~occ, (int)q);
    } else if (q < A + B + C) {
      q -= A + B;
      r = q / (7 * 28); q -= r * 7 * 28;
      p[0] = SzInvDiag[r];
      r = q / 28; q -= r * 28;
      r += SzInvDiag[r] >= p[0];
      p[1] = SzInvDiag[r];
      p[2] = SzInvLower[q];
    } else {
      q -= A + B + C;
      r = q / 42; q -= r * 42;
      p[0] = SzInvDiag[r];
      r = q / 6; q -= r * 6;
      r += SzInvDiag[r] >= p[0];
      p[1] = SzInvDiag[r];
      occ = bit(p[0]) | bit(p[1]);
      p[2] = (uint8_t)select_bit(~occ & DIAG, (int)q);
    }
    occ = bit(p[0]) | bit(p[1]) | bit(p[2]);
  }
  for (int k = lead; k < n; k += e->norm[k]) {
    uint64_t range = Binom[e->norm[k]][64 - popcnt(occ)];
    occ = unrank_set((idx / e->factor[k]) % range, e->norm[k], p + k, occ);
  }
  return 0;
}

/* probe_table's WDL path (probe.c:1522-1611) for a pawnless position:
 * pt[i] piece code (colour bit 8), sq[i] its square, i < num.  Returns
 * the stored WDL value (-2..2), or -9 on bad input. */
int tbo_rtbw_probe(int h, int stm, int num, const uint8_t *pt, const uint8_t *sq)
{
  if (h < 0 || h >= SZ_MAX || !g_sz[h] || num != g_sz[h]->num) return -9;
  const SzFile *f = g_sz[h];
  int cnt[16] = { 0 };
  for (int i = 0; i < num; i++) cnt[pt[i]]++;
  bool same = !memcmp(cnt, f->wcnt, sizeof cnt), flip, bside;
  if (f->symmetric) {
    flip = stm != WHITE;
    bside = false;
  } else {
    flip = !same;
    bside = (stm == WHITE) == flip;
  }
  const SzSub *e = &f->sub[f->split ? bside : 0];
  uint8_t p[MAX_PIECES];
  /* list_squares (probe.c:1513-1521) */
  for (int i = 0; i < num;) {
    int t = e->pieces[i] ^ (flip ? 8 : 0), i0 = i;
    for (int j = 0; j < num; j++)
      if (pt[j] == t) p[i++] = (uint8_t)(sq[j] ^ (flip ? 0x38 : 0));
    if (i == i0) return -9;
  }
  uint64_t idx = sz_encode(e, f, p);
  if (idx >= e->tb_size) return -9;
  return sz_value(e, idx) - 2;
}

/* Round trip of the index restatement: for every index of sub-table side
 * (step `stride`), decode then encode; returns the number of indices whose
 * position has distinct squares and encodes back to itself, and counts the
 * others in *bad (must stay 0). */
uint64_t tbo_rtbw_roundtrip(int h, int side, uint64_t stride, uint64_t *bad)
{
  *bad = 0;
  if (h < 0 || h >= SZ_MAX || !g_sz[h]) return 0;
  const SzFile *f = g_sz[h];
  const SzSub *e = &f->sub[side];
  uint64_t ok = 0, nbad = 0;
#pragma omp parallel for num_threads(g_threads) reduction(+ : ok, nbad) schedule(static)
  for (uint64_t i = 0; i < e->tb_size; i += stride) {
    uint8_t p[MAX_PIECES], p2[MAX_PIECES];
    sz_decode(e, f, i, p);
    memcpy(p2, p, sizeof p);
    if (sz_encode(e, f, p2) == i) ok++;
    else nbad++;
  }
  *bad = nbad;
  return ok;
}

/* A table's entries for positions decoded from the file index space:
 * squares in header order and header piece codes, for tests. */
int tbo_rtbw_decode(int h, int side, uint64_t idx, uint8_t *pt, uint8_t *sq)
{
  if (h < 0 || h >= SZ_MAX || !g_sz[h] || side < 0 || side > g_sz[h]->split) return -1;
  const SzFile *f = g_sz[h];
  if (sz_decode(&f->sub[side], f, idx, sq)) return -1;
  for (int i = 0; i < f->num; i++) pt[i] = f->sub[side].pieces[i];
  return 0;
}

/* Checks a WDL file against a table's codes (material as the codes are
 * laid out, reference slot order; codes[stm][slice * slice_size + idx]):
 * every stride-th position of every slice, both sides, legal ones only.
 * Returns the number checked; *bad counts mismatches, first_bad (5
 * entries, may be NULL) = stm, slice, idx, code, file value of the first. */
uint64_t tbo_rtbw_verify(int h, const char *material, const uint16_t *codes_w,
                         const uint16_t *codes_b, uint64_t stride, uint64_t *bad,
                         int64_t *first_bad)
{
  *bad = 0;
  Table *t = calloc(1, sizeof(Table));
  if (!t || setup_material(t, material)) { free(t); return 0; }
  const uint16_t *codes[2] = { codes_w, codes_b };
  uint64_t checked = 0, nbad = 0;
  int64_t fb[5] = { -1, -1, -1, -1, -1 };
  for (int stm = 0; stm < 2; stm++)
#pragma omp parallel for num_threads(g_threads) reduction(+ : checked, nbad) schedule(dynamic)
    for (int s = 0; s < NSLICES; s++)
      for (uint64_t idx = (uint64_t)s % stride; idx < t->size; idx += stride) {
        uint16_t c = codes[stm][(uint64_t)s * t->size + idx];
        if (c == TBO_ILLEGAL) continue;
        uint8_t sq[MAX_PIECES] = { 0 };
        sq[0] = KKSq[s][0];
        sq[1] = KKSq[s][1];
        idx_to_sq(&t->ii, idx, sq);
        int v = tbo_rtbw_probe(h, stm, t->num, t->pt, sq);
        checked++;
        if (v != code_to_wdl(c)) {
          nbad++;
#pragma omp critical
          if (fb[0] < 0) {
            fb[0] = stm; fb[1] = s; fb[2] = (int64_t)idx; fb[3] = c; fb[4] = v;
          }
        }
      }
  if (first_bad) memcpy(first_bad, fb, sizeof fb);
  free(t);
  *bad = nbad;
  return checked;
}

/* The rule-built encoding tables, for tests against the reference's values */
void tbo_rtbw_tables(uint8_t *tri, int8_t *offdiag, uint8_t *diag, uint8_t *lower)
{
  sz_tables();
  memcpy(tri, SzTri, 64);
  memcpy(offdiag, SzOffDiag, 64);
  memcpy(diag, SzDiag, 64);
  memcpy(lower, SzLower, 64);
}
