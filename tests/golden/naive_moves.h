/*
 * naive_moves.h -- the board walk of the independent brute-force solver
 * (naive.c), shared with the DTZ50 consistency checker (dtzcheck.c).
 *
 * Test infrastructure.  It shares no code with the oracle or the HIP path:
 * attacks are found by walking outward square by square, and moves by
 * making them on a 64-square board and testing the mover's king.
 * Pieces: col[i] 0 white / 1 black, typ[i] one of 'K','Q','R','B','N';
 * sq[i] the square of piece i (-1 once captured); board[x] the piece on x
 * or -1.
 */
#ifndef NAIVE_MOVES_H
#define NAIVE_MOVES_H

static const int DF[8] = { 1, -1, 0, 0, 1, 1, -1, -1 }; /* 0-3 straight, 4-7 diagonal */
static const int DR[8] = { 0, 0, 1, -1, 1, -1, 1, -1 };
static const int NF[8] = { 1, 2, 2, 1, -1, -2, -2, -1 };
static const int NR[8] = { 2, 1, -1, -2, -2, -1, 1, 2 };

static int onb(int f, int r) { return f >= 0 && f < 8 && r >= 0 && r < 8; }

/* Is square x attacked by colour `by`?  Walks outward from x. */
static int nm_attacked(const int *col, const int *typ, const int *board, int x, int by)
{
  int f = x & 7, r = x >> 3;
  for (int d = 0; d < 8; d++) {
    int cf = f + DF[d], cr = r + DR[d], dist = 1;
    for (; onb(cf, cr); cf += DF[d], cr += DR[d], dist++) {
      int j = board[cr * 8 + cf];
      if (j < 0) continue;
      if (col[j] == by) {
        int t = typ[j];
        if (t == 'Q' || (t == 'R' && d < 4) || (t == 'B' && d >= 4) || (t == 'K' && dist == 1))
          return 1;
      }
      break;
    }
  }
  for (int d = 0; d < 8; d++) {
    int cf = f + NF[d], cr = r + NR[d];
    if (!onb(cf, cr)) continue;
    int j = board[cr * 8 + cf];
    if (j >= 0 && col[j] == by && typ[j] == 'N') return 1;
  }
  return 0;
}

/* fn(u, mover, to, captured piece or -1, squares after, board after) for
 * each legal move (the mover's king not attacked afterwards) */
typedef void (*nm_move_fn)(void *u, int i, int to, int j, const int *sq, const int *board);

static void nm_try(const int *col, const int *typ, int k, int *sq, int *board, int i, int to,
                   int stm, nm_move_fn fn, void *u)
{
  int o = stm ^ 1, from = sq[i], j = board[to];
  if (j >= 0 && (col[j] == stm || typ[j] == 'K')) return;
  board[from] = -1;
  board[to] = i;
  sq[i] = to;
  if (j >= 0) sq[j] = -1;
  int ks = -1;
  for (int a = 0; a < k; a++)
    if (typ[a] == 'K' && col[a] == stm) ks = sq[a];
  if (!nm_attacked(col, typ, board, ks, o)) fn(u, i, to, j, sq, board);
  sq[i] = from;
  if (j >= 0) sq[j] = to;
  board[to] = j;
  board[from] = i;
}

static void nm_legal_moves(const int *col, const int *typ, int k, int *sq, int *board, int stm,
                           nm_move_fn fn, void *u)
{
  for (int i = 0; i < k; i++) {
    if (col[i] != stm) continue;
    int t = typ[i], f = sq[i] & 7, r = sq[i] >> 3;
    if (t == 'N') {
      for (int d = 0; d < 8; d++)
        if (onb(f + NF[d], r + NR[d]))
          nm_try(col, typ, k, sq, board, i, (r + NR[d]) * 8 + f + NF[d], stm, fn, u);
      continue;
    }
    int d0 = t == 'B' ? 4 : 0, d1 = t == 'R' ? 4 : 8;
    for (int d = d0; d < d1; d++) {
      for (int cf = f + DF[d], cr = r + DR[d]; onb(cf, cr); cf += DF[d], cr += DR[d]) {
        int to = cr * 8 + cf, occupied = board[to] >= 0;
        nm_try(col, typ, k, sq, board, i, to, stm, fn, u);
        if (occupied || t == 'K') break;
      }
    }
  }
}

#endif
