/*
 * dtzcheck.c -- DTZ50 successor-consistency check of finished tables.
 *
 * TEST INFRASTRUCTURE ONLY (loaded by tests/ through dtzcheck.py).  It pins
 * the 50-move handling of the generator -- the n = 1 and n = 101 capture
 * cases (src/tb8gen.c:964-970, 1046-1048), the cwin verification above
 * ply 100 (src/tb8gen.c:1010), the forced ply 101 (src/tb8gen.c:1426-1433)
 * and the cursed / blessed values above it -- without the oracle's or the
 * engine's recurrences: the codes of a table are data, and each position's
 * code is recomputed from its legal moves, made by the brute-force solver's
 * board walk (naive_moves.h), and the codes they reach:
 *
 *   illegal (3)    the side not to move is in check
 *   L_0            in check with no legal move
 *   W_n            n = min over the winning moves of: 1 for a capture into
 *                  a loss of the smaller table (-2), 101 for a capture into
 *                  a blessed loss (-1), m + 1 for a quiet move into L_m
 *   L_n            every legal move loses (at least one exists): n = max of
 *                  1 for a capture into a win (2), 101 for a capture into a
 *                  cursed win (1), m + 1 for a quiet move into W_m
 *   draw (0)       otherwise
 *
 * Smaller tables give WDL 2 for W_n with n <= 100, 1 above (cursed), -2 and
 * -1 likewise for L_n, 0 otherwise; a bare pair of kings is 0.  Since every
 * W_n / L_n label points to a strictly smaller one, a table that passes the
 * check at every position is the unique fixpoint of these rules.
 *
 * Table layout (the reference's, restated here from its rules): slots WK,
 * BK, then the pieces ordered N < B < R < Q, white before black within a
 * type (src/tb8gen.c:1331-1339); sets are runs of equal pieces; a set's
 * digit is sum_j C(rank_j, j + 1) over its squares in ascending order,
 * ranked among the squares free of the kings and the earlier sets
 * (src/tb8gen.c:129-149); factor[last] = 1 (src/tb8gen.c:209-228); king
 * slices numbered as src/probe.c:441-522 after the mirror and a1-h8 flip of
 * src/tb8gen.c:69-89, 230-238; codes[stm][slice * size + index].
 */
#include <omp.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "naive_moves.h"

#define NSL 462

typedef struct {
  int k, numsets;
  int col[8], typ[8];
  int first[6], mult[6];
  uint64_t factor[6], size;
  const uint16_t *codes[2];
} DcTab;

static int16_t KKIdx[10][64];
static int TriOf[64];
static uint8_t KKSq[NSL][2];
static uint64_t Binom[8][64];
static int ready;

static int adjacent(int a, int b)
{
  int df = (a & 7) - (b & 7), dr = (a >> 3) - (b >> 3);
  return df >= -1 && df <= 1 && dr >= -1 && dr <= 1;
}
static int on_diag(int s) { return (s & 7) == (s >> 3); }

static void init(void)
{
  if (ready) return;
  /* white king a1-d1-d4 triangle: b1 c1 d1 c2 d2 d3 off the diagonal, then
   * a1 b2 c3 d4 (src/probe.c:327-329); pairs numbered in that order, both
   * kings on a1-h8 after the first 441 */
  static const int tri[10] = { 1, 2, 3, 10, 11, 19, 0, 9, 18, 27 };
  int n = 0, nd = 441;
  for (int i = 0; i < 64; i++) TriOf[i] = -1;
  for (int i = 0; i < 10; i++) {
    int wk = tri[i];
    TriOf[wk] = i;
    for (int bk = 0; bk < 64; bk++) {
      KKIdx[i][bk] = -1;
      if (bk == wk || adjacent(wk, bk)) continue;
      if (on_diag(wk) && (bk & 7) < (bk >> 3)) continue;
      int s = on_diag(wk) && on_diag(bk) ? nd++ : n++;
      KKIdx[i][bk] = (int16_t)s;
      KKSq[s][0] = (uint8_t)wk;
      KKSq[s][1] = (uint8_t)bk;
    }
  }
  for (int j = 0; j < 64; j++) Binom[0][j] = 1;
  for (int i = 1; i < 8; i++) {
    Binom[i][0] = 0;
    for (int j = 1; j < 64; j++) Binom[i][j] = Binom[i - 1][j - 1] + Binom[i][j - 1];
  }
  ready = 1;
}

static int type_rank(int t)
{
  switch (t) {
  case 'N': return 0;
  case 'B': return 1;
  case 'R': return 2;
  case 'Q': return 3;
  default: return 4;
  }
}

static int setup(DcTab *T, const char *m)
{
  int c = 0, n = 0, col[8], typ[8];
  for (const char *s = m; *s; s++) {
    if (*s == 'v') { c = 1; continue; }
    if (!strchr("KQRBN", *s) || n >= 8) return -1;
    col[n] = c;
    typ[n] = *s;
    n++;
  }
  T->k = 0;
  for (int cc = 0; cc < 2; cc++)  /* the kings first: WK, BK */
    for (int i = 0; i < n; i++)
      if (typ[i] == 'K' && col[i] == cc) { T->col[T->k] = cc; T->typ[T->k] = 'K'; T->k++; }
  if (T->k != 2) return -1;
  for (int tr = 0; tr < 4; tr++)
    for (int cc = 0; cc < 2; cc++)
      for (int i = 0; i < n; i++)
        if (typ[i] != 'K' && col[i] == cc && type_rank(typ[i]) == tr) {
          T->col[T->k] = cc;
          T->typ[T->k] = typ[i];
          T->k++;
        }
  T->numsets = 0;
  for (int i = 2; i < T->k;) {
    int j = i;
    while (j < T->k && T->col[j] == T->col[i] && T->typ[j] == T->typ[i]) j++;
    T->first[T->numsets] = i;
    T->mult[T->numsets] = j - i;
    T->numsets++;
    i = j;
  }
  uint64_t f = 1;
  int free_sq = 62;
  uint64_t radix[6];
  for (int s = 0; s < T->numsets; s++) {
    radix[s] = Binom[T->mult[s]][free_sq];
    free_sq -= T->mult[s];
  }
  for (int s = T->numsets - 1; s >= 0; s--) {
    T->factor[s] = f;
    f *= radix[s];
  }
  T->size = f;
  return 0;
}

/* squares of (slice s, index idx) in slot order */
static void unrank(const DcTab *T, int s, uint64_t idx, int *sq)
{
  sq[0] = KKSq[s][0];
  sq[1] = KKSq[s][1];
  uint64_t occ = (1ULL << sq[0]) | (1ULL << sq[1]);
  for (int k = 0; k < T->numsets; k++) {
    uint64_t d = idx / T->factor[k];
    idx -= d * T->factor[k];
    int m = T->mult[k], f = T->first[k];
    uint64_t occ0 = occ;
    /* member i (descending): the largest r with C(r, i + 1) <= d */
    for (int i = m - 1; i >= 0; i--) {
      int r = i;
      while (r + 1 < 64 && Binom[i + 1][r + 1] <= d) r++;
      d -= Binom[i + 1][r];
      /* the r-th square free of occ0 */
      int x = -1, cnt = -1;
      while (cnt < r) {
        x++;
        if (!((occ0 >> x) & 1)) cnt++;
      }
      sq[f + i] = x;
      occ |= 1ULL << x;
    }
  }
}

/* code of a position given as n pieces (col, typ, square; captured pieces
 * skipped), side to move stm, in table T (either colour orientation); -1
 * if T does not hold that material */
static int lookup(const DcTab *T, int n, const int *col, const int *typ, const int *sq, int stm)
{
  int cnt[2][128] = { { 0 } }, need[2][128] = { { 0 } }, m = 0;
  for (int i = 0; i < n; i++)
    if (sq[i] >= 0) { cnt[col[i]][typ[i]]++; m++; }
  if (m != T->k) return -1;
  for (int i = 0; i < T->k; i++) need[T->col[i]][T->typ[i]]++;
  int flip = memcmp(cnt, need, sizeof cnt) != 0;
  if (flip) {
    for (int t = 0; t < 128; t++)
      if (cnt[0][t] != need[1][t] || cnt[1][t] != need[0][t]) return -1;
  }
  int s2[8], used[8] = { 0 };
  for (int j = 0; j < T->k; j++) {
    for (int i = 0; i < n; i++) {
      if (sq[i] < 0 || used[i] || (col[i] ^ flip) != T->col[j] || typ[i] != T->typ[j]) continue;
      used[i] = 1;
      s2[j] = flip ? sq[i] ^ 56 : sq[i];
      break;
    }
  }
  stm ^= flip;
  /* mirror and a1-h8 flip (src/tb8gen.c:69-89, 230-238) */
  int wk = s2[0], bk = s2[1];
  int mm = ((wk & 4) ? 7 : 0) | ((wk & 32) ? 0x38 : 0);
  int w1 = wk ^ mm, b1 = bk ^ mm;
  const uint64_t A1D1D4 = 0x080c0e0fULL, A1D4 = 0x08040201ULL, LOWER = 0x80c0e0f0f8fcfeffULL;
  int fl = !((A1D1D4 >> w1) & 1) || (((A1D4 >> w1) & 1) && !((LOWER >> b1) & 1));
  for (int j = 0; j < T->k; j++) {
    int x = s2[j] ^ mm;
    s2[j] = fl ? ((x & 7) << 3) | (x >> 3) : x;
  }
  if (TriOf[s2[0]] < 0) return -2;
  int s = KKIdx[TriOf[s2[0]]][s2[1]];
  if (s < 0) return -2;
  uint64_t idx = 0, occ = (1ULL << s2[0]) | (1ULL << s2[1]);
  for (int k = 0; k < T->numsets; k++) {
    int f = T->first[k], mlt = T->mult[k], v[6];
    for (int a = 0; a < mlt; a++) v[a] = s2[f + a];
    for (int a = 0; a < mlt; a++)
      for (int b = a + 1; b < mlt; b++)
        if (v[b] < v[a]) { int t = v[a]; v[a] = v[b]; v[b] = t; }
    uint64_t d = 0, occ2 = occ;
    for (int a = 0; a < mlt; a++) {
      int rank = v[a] - __builtin_popcountll(occ & ((1ULL << v[a]) - 1));
      d += Binom[a + 1][rank];
      occ2 |= 1ULL << v[a];
    }
    idx += d * T->factor[k];
    occ = occ2;
  }
  return T->codes[stm][(uint64_t)s * T->size + idx];
}

static int wdl(int code)
{
  int tag = code & 3, n = code >> 2;
  if (tag == 1) return n <= 100 ? 2 : 1;
  if (tag == 2) return n <= 100 ? -2 : -1;
  return 0;
}

typedef struct {
  const DcTab *T;
  const DcTab *subs;
  int nsub, o;
  int nmoves, wmin, lmax, all_lose, bad, via101;
} Acc;

static void on_move(void *u, int i, int to, int j, const int *sq, const int *board)
{
  (void)i;
  (void)to;
  (void)board;
  Acc *a = (Acc *)u;
  const DcTab *T = a->T;
  a->nmoves++;
  int code;
  if (j < 0) {
    code = lookup(T, T->k, T->col, T->typ, sq, a->o);
    if (code < 0 || code == 3) { a->bad = 1; return; }  /* a legal move must reach a legal position */
    int tag = code & 3, m = code >> 2;
    if (tag == 2 && m + 1 < a->wmin) a->wmin = m + 1;
    if (tag == 1) { if (m + 1 > a->lmax) a->lmax = m + 1; }
    else a->all_lose = 0;
    return;
  }
  int v = 0;
  if (T->k - 1 > 2) {
    code = -1;
    for (int t = 0; t < a->nsub && code == -1; t++)
      code = lookup(&a->subs[t], T->k, T->col, T->typ, sq, a->o);
    if (code < 0 || code == 3) { a->bad = 1; return; }
    v = wdl(code);
  }
  if (v == -2 && 1 < a->wmin) a->wmin = 1;
  if (v == -1 && 101 < a->wmin) a->wmin = 101;
  if (v == 2) { if (1 > a->lmax) a->lmax = 1; }
  else if (v == 1) { if (101 > a->lmax) a->lmax = 101; }
  else a->all_lose = 0;
  if (v == 1 || v == -1) a->via101 = 1;
}

/* The code the rules give to (stm, slice s, idx) of T. */
static int expected(const DcTab *T, const DcTab *subs, int nsub, int s, uint64_t idx, int stm,
                    int *via101)
{
  int sq[8], board[64];
  unrank(T, s, idx, sq);
  for (int x = 0; x < 64; x++) board[x] = -1;
  for (int i = 0; i < T->k; i++) board[sq[i]] = i;
  int o = stm ^ 1, ks = -1, ko = -1;
  for (int i = 0; i < T->k; i++)
    if (T->typ[i] == 'K') { if (T->col[i] == stm) ks = sq[i]; else ko = sq[i]; }
  if (nm_attacked(T->col, T->typ, board, ko, stm)) return 3;
  Acc a = { T, subs, nsub, o, 0, 1 << 30, 0, 1, 0, 0 };
  nm_legal_moves(T->col, T->typ, T->k, sq, board, stm, on_move, &a);
  if (a.bad) return -1;
  *via101 = a.via101;
  if (!a.nmoves) return nm_attacked(T->col, T->typ, board, ks, o) ? 2 : 0;
  if (a.wmin < (1 << 30)) return (a.wmin << 2) | 1;
  if (a.all_lose) return (a.lmax << 2) | 2;
  return 0;
}

/* Checks both sides of positions g = start + i * stride (i < count) of the
 * flattened (slice, index) space of `material` (codes cw, cb).  subs: the
 * tables one capture smaller (either colour orientation).  stats[8]:
 * positions x sides checked, illegal, W_n n <= 100, W_n n > 100, L_n
 * n <= 100, L_n n > 100 (L_0 included), draws, codes decided through a
 * capture into a cursed or blessed value.  Returns the mismatches;
 * first_bad[5] = stm, slice, index, table code, rules' code (-1: a legal
 * move reached an illegal or missing position). */
int64_t dc_check_list(const char *material, const uint16_t *cw, const uint16_t *cb, int nsub,
                      const char *const *sub_mat, const uint16_t *const *sub_cw,
                      const uint16_t *const *sub_cb, const uint64_t *list, uint64_t start,
                      uint64_t stride, uint64_t count, int threads, int64_t *first_bad,
                      uint64_t *stats);

int64_t dc_check(const char *material, const uint16_t *cw, const uint16_t *cb, int nsub,
                 const char *const *sub_mat, const uint16_t *const *sub_cw,
                 const uint16_t *const *sub_cb, uint64_t start, uint64_t stride, uint64_t count,
                 int threads, int64_t *first_bad, uint64_t *stats)
{
  return dc_check_list(material, cw, cb, nsub, sub_mat, sub_cw, sub_cb, NULL, start, stride,
                       count, threads, first_bad, stats);
}

/* dc_check over the positions list[0 .. count) (flattened indices) when
 * list is not NULL */
int64_t dc_check_list(const char *material, const uint16_t *cw, const uint16_t *cb, int nsub,
                      const char *const *sub_mat, const uint16_t *const *sub_cw,
                      const uint16_t *const *sub_cb, const uint64_t *list, uint64_t start,
                      uint64_t stride, uint64_t count, int threads, int64_t *first_bad,
                      uint64_t *stats)
{
  init();
  DcTab T, subs[8];
  if (setup(&T, material) || nsub > 8) return -1;
  T.codes[0] = cw;
  T.codes[1] = cb;
  for (int i = 0; i < nsub; i++) {
    if (setup(&subs[i], sub_mat[i])) return -1;
    subs[i].codes[0] = sub_cw[i];
    subs[i].codes[1] = sub_cb[i];
  }
  uint64_t total = (uint64_t)NSL * T.size;
  if (stride == 0) stride = 1;
  int64_t bad = 0;
  uint64_t st[8] = { 0 };
  if (first_bad) first_bad[0] = -1;
#pragma omp parallel for schedule(dynamic, 4096) num_threads(threads) reduction(+ : bad)
  for (uint64_t i = 0; i < count; i++) {
    uint64_t g = list ? list[i] : start + i * stride;
    if (g >= total) continue;
    int s = (int)(g / T.size);
    uint64_t idx = g % T.size;
    uint64_t loc[8] = { 0 };
    for (int stm = 0; stm < 2; stm++) {
      int via = 0;
      int e = expected(&T, subs, nsub, s, idx, stm, &via);
      int got = T.codes[stm][g];
      loc[0]++;
      if (e == 3) loc[1]++;
      else if ((e & 3) == 1) loc[(e >> 2) <= 100 ? 2 : 3]++;
      else if ((e & 3) == 2) loc[(e >> 2) <= 100 ? 4 : 5]++;
      else if (e == 0) loc[6]++;
      if (via && e > 0) loc[7]++;
      if (e != got) {
        bad++;
#pragma omp critical
        if (first_bad && first_bad[0] < 0) {
          first_bad[0] = stm;
          first_bad[1] = s;
          first_bad[2] = (int64_t)idx;
          first_bad[3] = got;
          first_bad[4] = e;
        }
      }
    }
    for (int k = 0; k < 8; k++)
      if (loc[k]) {
#pragma omp atomic
        st[k] += loc[k];
      }
  }
  if (stats) memcpy(stats, st, sizeof st);
  return bad;
}

/* slice size of a material in this layout (0 if malformed) */
uint64_t dc_slice_size(const char *material)
{
  init();
  DcTab T;
  return setup(&T, material) ? 0 : T.size;
}

/* squares (slot order) of (slice, index): for pinning this layout against
 * the oracle's */
int dc_unrank(const char *material, int s, uint64_t idx, int *sq)
{
  init();
  DcTab T;
  if (setup(&T, material) || s < 0 || s >= NSL || idx >= T.size) return -1;
  unrank(&T, s, idx, sq);
  return 0;
}
