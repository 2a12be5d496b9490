/*
 * naive.c -- independent brute-force DTZ solver used to pin the oracle.
 *
 * It shares no code with the oracle or the HIP path: positions are indexed
 * over the full board (64^k squares, no king-slice symmetry), moves are
 * generated by walking the board square by square, and every ply is
 * computed by a forward (pull) evaluation of every undecided position
 * rather than by unmove scatter plus verification.  The per-ply sets are
 * defined from the rules:
 *   illegal        side not to move is in check
 *   L_0            in check, no legal move or capture
 *   W_n            undecided legal position with a legal quiet move into
 *                  L_{n-1} of the opponent, or (n == 1) a legal capture into
 *                  a lost position of the smaller material, or (n == 101) a
 *                  legal capture into a blessed loss
 *   L_n            undecided legal position with a legal quiet move into
 *                  W_{n-1} of the opponent (n > 1), or a legal capture into a
 *                  won (n == 1) / cursed-won (n == 101) smaller position, and
 *                  whose every legal move reaches an opponent win: quiet
 *                  moves a W_m with m <= n-1, captures a win (n <= 100) or
 *                  a win or cursed win (n > 100)
 * WDL of a smaller position: W_n -> 2 (n <= 100) or 1, L_n -> -2 or -1,
 * otherwise 0; a lone pair of kings is 0.
 *
 * Usage: naive MATERIAL OUT.bin   (e.g. naive KRvK krvk.bin)
 * Output: two int16 arrays [stm][64^k] of codes (0 draw, 3 illegal,
 * n<<2|1 win, n<<2|2 loss, 0x7fff invalid placement); piece order = the
 * order of the letters in MATERIAL.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "naive_moves.h"

#define UNK (-1)
#define INVALID 0x7fff

typedef struct Solved {
  char key[16];
  int k;
  int col[8], typ[8]; /* typ: 'K','Q','R','B','N' */
  int16_t *v[2];
  struct Solved *next;
} Solved;

static Solved *g_solved;

/* Is square x attacked by colour `by`? */
static int attacked(const Solved *m, const int *board, int x, int by)
{
  return nm_attacked(m->col, m->typ, board, x, by);
}

static int king_of(const Solved *m, int c)
{
  for (int i = 0; i < m->k; i++)
    if (m->typ[i] == 'K' && m->col[i] == c) return i;
  return -1;
}

static uint64_t npos(int k)
{
  uint64_t n = 1;
  for (int i = 0; i < k; i++) n *= 64;
  return n;
}

static Solved *solve(const int *col, const int *typ, int k);

static int wdl_of(int16_t c)
{
  if (c == UNK || c == 0) return 0;
  int tag = c & 3, n = c >> 2;
  if (tag == 1) return n <= 100 ? 2 : 1;
  if (tag == 2) return n <= 100 ? -2 : -1;
  return 0;
}

/* list of legal moves: quiet targets (full index into same table) and
 * captures (WDL of the resulting smaller position, side to move = opp) */
typedef struct {
  int nq, nc;
  uint64_t q[256];
  int cw[64];
} Moves;

typedef struct {
  const Solved *m;
  Solved **sub;
  int o;
  Moves *mv;
} MoveCtx;

static void take_move(void *u, int i, int to, int j, const int *sq, const int *board)
{
  (void)i;
  (void)to;
  (void)board;
  MoveCtx *x = (MoveCtx *)u;
  const Solved *m = x->m;
  if (j < 0) {
    uint64_t q = 0;
    for (int a = m->k - 1; a >= 0; a--) q = q * 64 + (uint64_t)sq[a];
    x->mv->q[x->mv->nq++] = q;
  } else {
    int wdl = 0;
    if (x->sub[j]) {
      uint64_t q = 0;
      for (int a = m->k - 1; a >= 0; a--)
        if (a != j) q = q * 64 + (uint64_t)sq[a];
      wdl = wdl_of(x->sub[j]->v[x->o][q]);
    }
    x->mv->cw[x->mv->nc++] = wdl;
  }
}

static void gen_moves(const Solved *m, Solved **sub, uint64_t p, int stm, Moves *mv)
{
  int sq[8], board[64];
  for (int i = 0; i < 64; i++) board[i] = -1;
  uint64_t x = p;
  for (int i = 0; i < m->k; i++) { sq[i] = (int)(x & 63); x >>= 6; board[sq[i]] = i; }
  mv->nq = mv->nc = 0;
  MoveCtx ctx = { m, sub, stm ^ 1, mv };
  nm_legal_moves(m->col, m->typ, m->k, sq, board, stm, take_move, &ctx);
}

static Solved *solve(const int *col, const int *typ, int k)
{
  char key[16];
  for (int i = 0; i < k; i++) key[i] = (char)(col[i] ? typ[i] + 32 : typ[i]);
  key[k] = 0;
  for (Solved *s = g_solved; s; s = s->next)
    if (!strcmp(s->key, key)) return s;
  Solved *m = calloc(1, sizeof *m);
  strcpy(m->key, key);
  m->k = k;
  memcpy(m->col, col, sizeof(int) * k);
  memcpy(m->typ, typ, sizeof(int) * k);
  /* smaller materials reached by capturing piece j */
  Solved *sub[8] = { 0 };
  for (int j = 0; j < k; j++) {
    if (typ[j] == 'K') continue;
    int c2[8], t2[8], n2 = 0;
    for (int a = 0; a < k; a++)
      if (a != j) { c2[n2] = col[a]; t2[n2] = typ[a]; n2++; }
    sub[j] = n2 > 2 ? solve(c2, t2, n2) : NULL;
  }
  uint64_t N = npos(k);
  for (int c = 0; c < 2; c++) {
    m->v[c] = malloc(N * sizeof(int16_t));
    if (!m->v[c]) { fprintf(stderr, "oom\n"); exit(1); }
  }
  /* classify placements */
#pragma omp parallel for schedule(static)
  for (uint64_t p = 0; p < N; p++) {
    int sq[8], board[64], ok = 1;
    for (int i = 0; i < 64; i++) board[i] = -1;
    uint64_t x = p;
    for (int i = 0; i < k; i++) {
      sq[i] = (int)(x & 63);
      x >>= 6;
      if (board[sq[i]] >= 0) ok = 0;
      board[sq[i]] = i;
    }
    if (!ok) { m->v[0][p] = m->v[1][p] = INVALID; continue; }
    for (int c = 0; c < 2; c++) {
      int ko = king_of(m, c ^ 1);
      m->v[c][p] = attacked(m, board, sq[ko], c) ? 3 : UNK;
    }
  }
  /* L_0 */
#pragma omp parallel for schedule(dynamic, 4096)
  for (uint64_t p = 0; p < N; p++) {
    Moves mv;
    for (int c = 0; c < 2; c++) {
      if (m->v[c][p] != UNK) continue;
      /* in check? = the position with the other side to move is illegal */
      if (m->v[c ^ 1][p] != 3) continue;
      gen_moves(m, sub, p, c, &mv);
      if (mv.nq == 0 && mv.nc == 0) m->v[c][p] = 2; /* LOSS(0) */
    }
  }
  int16_t *nv[2];
  for (int c = 0; c < 2; c++) nv[c] = malloc(N * sizeof(int16_t));
  for (int n = 1; n < 4000; n++) {
    long changed = 0;
    for (int c = 0; c < 2; c++) memcpy(nv[c], m->v[c], N * sizeof(int16_t));
#pragma omp parallel for schedule(dynamic, 4096) reduction(+ : changed)
    for (uint64_t p = 0; p < N; p++) {
      Moves mv;
      for (int c = 0; c < 2; c++) {
        if (m->v[c][p] != UNK) continue;
        int o = c ^ 1;
        gen_moves(m, sub, p, c, &mv);
        int16_t lossprev = (int16_t)(((n - 1) << 2) | 2), winprev = (int16_t)(((n - 1) << 2) | 1);
        /* W_n */
        int w = 0;
        for (int i = 0; i < mv.nq && !w; i++) w = m->v[o][mv.q[i]] == lossprev;
        for (int i = 0; i < mv.nc && !w; i++)
          w = (n == 1 && mv.cw[i] == -2) || (n == 101 && mv.cw[i] == -1);
        if (w) { nv[c][p] = (int16_t)((n << 2) | 1); changed++; continue; }
        /* L_n */
        int cand = 0;
        for (int i = 0; i < mv.nq && !cand && n > 1; i++) cand = m->v[o][mv.q[i]] == winprev;
        for (int i = 0; i < mv.nc && !cand; i++)
          cand = (n == 1 && mv.cw[i] == 2) || (n == 101 && mv.cw[i] == 1);
        if (!cand) continue;
        int all = 1;
        for (int i = 0; i < mv.nq && all; i++) {
          int16_t t = m->v[o][mv.q[i]];
          all = t != UNK && (t & 3) == 1 && (t >> 2) <= n - 1;
        }
        for (int i = 0; i < mv.nc && all; i++) all = mv.cw[i] >= (n <= 100 ? 2 : 1);
        if (all) { nv[c][p] = (int16_t)((n << 2) | 2); changed++; }
      }
    }
    for (int c = 0; c < 2; c++) { int16_t *t = m->v[c]; m->v[c] = nv[c]; nv[c] = t; }
    if (n > 101 && !changed) break;
  }
  free(nv[0]);
  free(nv[1]);
  for (int c = 0; c < 2; c++)
    for (uint64_t p = 0; p < N; p++)
      if (m->v[c][p] == UNK) m->v[c][p] = 0;
  m->next = g_solved;
  g_solved = m;
  return m;
}

int main(int argc, char **argv)
{
  if (argc < 3) { fprintf(stderr, "usage: naive MATERIAL OUT.bin\n"); return 2; }
  int col[8], typ[8], k = 0, c = 0;
  for (const char *s = argv[1]; *s; s++) {
    if (*s == 'v') { c = 1; continue; }
    if (!strchr("KQRBN", *s) || k >= 8) { fprintf(stderr, "bad material\n"); return 2; }
    col[k] = c;
    typ[k] = *s;
    k++;
  }
  Solved *m = solve(col, typ, k);
  FILE *f = fopen(argv[2], "wb");
  if (!f) return 1;
  uint64_t N = npos(k);
  fwrite(m->v[0], sizeof(int16_t), N, f);
  fwrite(m->v[1], sizeof(int16_t), N, f);
  fclose(f);
  return 0;
}
