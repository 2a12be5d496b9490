/*
 * tbo.h -- CPU oracle for the king-slice retrograde analysis.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * generator's algorithm (syzygy1/tb8gen, src/tb8gen.c, src/kslice.c,
 * src/movegen.h, src/probe.c), used as the parity checker for the HIP path.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product library never links or calls it.
 *
 * Every table this oracle produces is a per-(side to move, king slice,
 * index) array of 16-bit codes:
 *     0                 draw (never entered a W_n or L_n set), legal
 *     3                 illegal (side not to move is in check)
 *     (n << 2) | 1      in W_n: side to move wins, conversion in n plies
 *     (n << 2) | 2      in L_n: side to move loses, conversion in n plies
 * W_n / L_n are exactly the per-ply sets the reference writes to its
 * n/W/<c>/<kk> and n/L/<c>/<kk> slice files (tb8gen.c:1075, 1017).
 */
#ifndef TBO_H
#define TBO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TBO_DRAW 0
#define TBO_ILLEGAL 3
#define TBO_WIN(n) ((uint16_t)(((n) << 2) | 1))
#define TBO_LOSS(n) ((uint16_t)(((n) << 2) | 2))
#define TBO_MAX_PLY 4096

/* Options, set before generating. */
void tbo_set_threads(int nthreads);
/* 1: reproduce the reference's king-capture index for duplicated enemy
 * sets (check_king_moves, tb8gen.c:543-555, which omits the fix-up that
 * check_moves applies at tb8gen.c:586-588).  0 (default): correct index. */
void tbo_set_ref_king_capture_index(int on);
/* 0: synthetic slices keep no codes (only their count lines); default 1. */
void tbo_set_keep_synth_codes(int on);

/* Generate a material (e.g. "KQRvKR") and, recursively, every smaller
 * material its captures reach.  Count lines in the reference's format
 * (tb8gen.c:764-767, 1024, 1082) go to log_path (NULL: none, "-": stdout).
 * Returns a handle >= 0, or a negative error. */
int tbo_generate(const char *material, const char *log_path);
int tbo_find(const char *material);
void tbo_free_all(void);

/* Table geometry. */
uint64_t tbo_slice_size(int h);          /* positions per king slice   */
int tbo_num_pieces(int h);
void tbo_piece_types(int h, uint8_t *pt); /* slot order, 8 entries     */
int tbo_num_sets(int h);
int tbo_capt_size(int h, int set, uint64_t *size);
const uint16_t *tbo_codes(int h, int stm); /* [462 * slice_size]     */
int tbo_num_plies(int h);
/* per-ply counts: kind 0 = L, 1 = W; returns -1 if the step did not run */
int64_t tbo_count(int h, int kind, int stm, int n);
int64_t tbo_broken(int h, int stm);
int64_t tbo_mates(int h, int stm);

/* Index helpers (reference layout), exported for the parity tests. */
int tbo_kkidx(int tri, int bk);
int tbo_kkmap(int wk, int bk);
void tbo_kksquare(int s, uint8_t *wk_bk);
uint64_t tbo_sq_to_idx(int h, const uint8_t *sq);       /* sq[num], normalised */
int tbo_idx_to_sq(int h, int s, uint64_t idx, uint8_t *sq); /* fills sq[num] */
int tbo_capt_idx_to_sq(int h, int set, int s, uint64_t idx, uint8_t *sq);

/* Sub (capture-exit) bitmaps, as the reference writes to sub/<kind>/<c>/<kk>:
 * kind 0 = loss, 1 = bloss, 2 = cwin, 3 = win.  Copies the colour's
 * concatenated per-set region of slice s (64-byte aligned sets). */
uint64_t tbo_sub_words(int h, int colour);
int tbo_sub_bitmap(int h, int kind, int colour, int s, uint64_t *dst);
uint64_t tbo_sub_offset_words(int h, int set);

/* Look up the code of an arbitrary position of the table's material:
 * sq[i] is the square of the piece in slot i (any order within a set). */
int tbo_lookup(int h, int stm, const uint8_t *sq);
/* Bulk form: n positions, each num squares in slot order. */
void tbo_lookup_bulk(int h, int stm, uint64_t n, const uint8_t *sq, uint16_t *out);

/* Synthetic single-slice mode (one king slice of a large material, capture
 * and king-move exits drawn from a seeded hash, see tbo.c).  Runs the full
 * fixpoint for slice s and returns the number of plies; codes are kept
 * only if keep_codes.  Count lines go to log_path. */
int tbo_synth_slice(const char *material, int s, uint64_t seed,
                    const char *log_path, int max_plies);
uint64_t tbo_synth_hash(uint64_t seed, int stm, int kind, int region, uint64_t idx);

/* Syzygy WDL files (the reader of probe.c, restated; pawnless only).
 * name: the file's material, white as in the file name. */
int tbo_rtbw_open(const char *path, const char *name);
void tbo_rtbw_close(int h);
uint64_t tbo_rtbw_size(int h, int side);
/* stored WDL (-2..2) of a position: pt[i] piece code, sq[i] square */
int tbo_rtbw_probe(int h, int stm, int num, const uint8_t *pt, const uint8_t *sq);
uint64_t tbo_rtbw_roundtrip(int h, int side, uint64_t stride, uint64_t *bad);
int tbo_rtbw_decode(int h, int side, uint64_t idx, uint8_t *pt, uint8_t *sq);
void tbo_rtbw_tables(uint8_t *tri, int8_t *offdiag, uint8_t *diag, uint8_t *lower);
uint64_t tbo_rtbw_verify(int h, const char *material, const uint16_t *codes_w,
                         const uint16_t *codes_b, uint64_t stride, uint64_t *bad,
                         int64_t *first_bad);

/* Sampled forward check of the first two plies of a synthetic slice
 * (tbo_synth_slice's exits) too large to generate here.  planes[10] are the
 * engine's bitmaps of the slice: wins[W,B], W_1[W,B], L_1[W,B] after ply 1,
 * W_2[W,B], L_2[W,B] after ply 2.  For nsamples seeded positions per side
 * (half uniform, half set bits of the engine's W_1, L_1, W_2, L_2), each of
 * the five memberships is recomputed from the legal moves out of the
 * position (targets' ply-1 sets read from the engine's planes, mates and
 * exits computed here).  Returns the number of mismatching memberships;
 * first_bad[4] = stm, plane (0 wins, 1 W_1, 2 L_1, 3 W_2, 4 L_2), index,
 * the engine's bit. */
int64_t tbo_synth_check(const char *material, int s, uint64_t seed, const uint64_t *const *planes,
                        uint64_t nsamples, uint64_t sample_seed, uint64_t *checked,
                        int64_t *first_bad);

/* CPU baseline sample of init (illegal, mates, synthetic exits) over
 * indices [lo, lo + n) of a synthetic slice; returns wall seconds. */
double tbo_synth_init_time(const char *material, int s, uint64_t seed, uint64_t lo, uint64_t n,
                           int64_t *counts);
/* CPU baseline sample (see tbo.c). */
int tbo_time_plies(const char *material, int max_plies, double *loop_seconds, int64_t *half_steps);
int tbo_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
