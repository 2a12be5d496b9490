/*
 * tbg.h -- C ABI of the MI355X king-slice retrograde engine (libtbg.so).
 *
 * Plain C types only.  The library owns all device memory; host buffers
 * passed to the get/put calls belong to the caller.  Every call returns an
 * int status: TBG_OK (0) or a negative TBG_E_* code (tbg_strerror).
 *
 * The reference (syzygy1/tb8gen) has no library boundary: its hot path sits
 * behind three internal interfaces that these entry points replace:
 *   - the worker interface run_threaded(func, work) over [begin, end)
 *     index units (src/threads.h:28-53, src/threads.c:266-289), used by
 *     predecessors / predecessors_sub / check_successors / calc_sub /
 *     calc_illegal / calc_mate (src/tb8gen.c:280-768);
 *   - the slice store src/kslice.h:75-98 (slot-addressed bitmaps, clear /
 *     and_not / or / count, file read/write);
 *   - capture resolution int probe_wdl(Position*, int, int)
 *     (src/probe.h:170, src/probe.c:1677-1684).
 * Bitmaps use the reference's bit order: bit idx&63 of little-endian
 * 64-bit word idx>>6 (= bit idx&7 of byte idx>>3, src/kslice.h:30-55),
 * each slice padded to 64 bytes (src/kslice.c:92-96).
 */
#ifndef TBG_H
#define TBG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TBG_OK 0
#define TBG_E_INVAL (-1)   /* bad argument (material string, side, slice, size) */
#define TBG_E_NOMEM (-2)   /* device or host allocation failed            */
#define TBG_E_HIP (-3)     /* a HIP runtime call or kernel launch failed  */
#define TBG_E_NOSUB (-4)   /* a smaller material's table is not supplied  */
#define TBG_E_STATE (-5)   /* call out of order (e.g. step before init)   */
#define TBG_E_BUDGET (-6)  /* job exceeds the device-memory budget        */
#define TBG_E_NODEV (-7)   /* no HIP device                               */
#define TBG_E_IO (-8)      /* a slice file could not be written or read   */

#define TBG_NSLICES 462
#define TBG_DRAW_RULE 100 /* src/tb8gen.c:36 */

/* Position codes of a finished table (one uint16 per side/slice/index):
 * 0 draw, 3 illegal, (n << 2) | 1 in W_n, (n << 2) | 2 in L_n. */
#define TBG_CODE_DRAW 0
#define TBG_CODE_ILLEGAL 3

/* which bitmap plane tbg_get_bitmap returns */
#define TBG_PLANE_WINS 0   /* illegal + W_1..W_n so far   (0/wins/<c>)      */
#define TBG_PLANE_W 1      /* W_n of the last completed ply (n/W/<c>)      */
#define TBG_PLANE_L 2      /* L_n of the last completed ply (n/L/<c>)      */

/* sub (capture exit) kinds, src/tb8gen.c:321-325 */
#define TBG_SUB_LOSS 0
#define TBG_SUB_BLOSS 1
#define TBG_SUB_CWIN 2
#define TBG_SUB_WIN 3

/* generation forms (tbg_generate) */
#define TBG_FORM_PLAIN 0   /* one slice per launch, reference manager order */
#define TBG_FORM_BATCHED 1 /* one launch per phase over all 462 slices      */
#define TBG_FORM_GRAPH 2   /* batched, each ply replayed from a hipGraph with
                              device-side step flags                        */

typedef struct tbg_ctx tbg_ctx;

typedef struct tbg_info {
  int npieces;
  uint8_t pt[8];            /* piece codes in slot order, src/types.h:19-22 */
  int numsets;
  int mult[6];
  uint64_t slice_size;      /* positions per king slice (IdxInfo.size)      */
  uint64_t slice_words;     /* 64-bit words per slice bitmap (64 B aligned) */
  uint64_t capt_size[6];    /* positions per slice with one piece of set k removed */
  uint64_t sub_words[2];    /* words per slice of a colour's sub region     */
  uint64_t sub_offset[6];   /* word offset of set k inside its colour region */
  uint64_t device_bytes;    /* device memory held by the context            */
  int synth_slice;          /* -1, or the single slice of a synthetic job   */
} tbg_info;

/* Per-step counts, in the order the reference prints them. */
typedef struct tbg_counts {
  uint64_t broken[2];       /* broken_w / broken_b  (tb8gen.c:764-765) */
  uint64_t mates[2];        /* l0_w / l0_b          (tb8gen.c:766-767) */
  int nplies;               /* last ply executed                       */
} tbg_counts;

/* Callback receiving each count line as the reference formats it, e.g.
 * "l17_w = 1234" (tb8gen.c:1024, 1082). */
typedef void (*tbg_line_fn)(const char *line, void *user);

/* Library setup: selects the HIP device.  Replaces init_threads /
 * init_movegen / init_tables (tb8gen.c:1279-1282). */
int tbg_init(int device);
const char *tbg_strerror(int err);

/* Lays out a material (tb8gen.c:1284-1385, kslice.c:98-127) and allocates
 * its resident device state.  material e.g. "KQRvKR". */
int tbg_create(tbg_ctx **out, const char *material);
/* Synthetic single-slice job: only king slice `slice` (0 <= slice < 441) of
 * `material`; capture and king-move exits come from a seeded splitmix64
 * hash (win below 0.03, loss in [0.03, 0.06)). */
int tbg_create_synth(tbg_ctx **out, const char *material, int slice, uint64_t seed);
/* tbg_create_synth with options: TBG_SYNTH_CODES keeps the per-position
 * codes of the slice (2 bytes per position and side, read by tbg_get_codes)
 * as a whole-table job does; without it only the bitmaps are kept. */
#define TBG_SYNTH_CODES 1u
int tbg_create_synth_ex(tbg_ctx **out, const char *material, int slice, uint64_t seed,
                        unsigned flags);
void tbg_destroy(tbg_ctx *ctx);
int tbg_get_info(const tbg_ctx *ctx, tbg_info *info);

/* Capture exits for both colours: calc_sub_kslices (tb8gen.c:317-342)
 * with probe_wdl answered from the finished device tables of the smaller
 * materials in subs[] (either colour orientation).  KvK needs no table. */
int tbg_resolve_captures(tbg_ctx *ctx, tbg_ctx *const *subs, int nsubs);
/* Illegal positions and mates: calc_illegal_and_mate (tb8gen.c:730-768). */
int tbg_init_illegal_mate(tbg_ctx *ctx, tbg_counts *counts);

/* One half-step over all slices of the material (batched form):
 * calc_L_n (tb8gen.c:943-1026) and calc_W_n (tb8gen.c:1028-1084).
 * `more` is the reference's more_l / more_w flag.  *count = |L_n| / |W_n|. */
int tbg_step_L(tbg_ctx *ctx, int stm, int n, int more, uint64_t *count);
int tbg_step_W(tbg_ctx *ctx, int stm, int n, int more, uint64_t *count);
/* Plain form of the same half-steps: one launch per slice and phase, in the
 * reference's manager order (kslice.c:41-89). */
int tbg_step_L_plain(tbg_ctx *ctx, int stm, int n, int more, uint64_t *count);
int tbg_step_W_plain(tbg_ctx *ctx, int stm, int n, int more, uint64_t *count);
/* Ends ply n: this ply's W_n / L_n become the next ply's frontiers. */
int tbg_end_ply(tbg_ctx *ctx);

/* The whole generator for one material after tbg_resolve_captures:
 * illegal/mate, then the ply loop of main() (tb8gen.c:1406-1444) with its
 * more_* flags, emitting the reference's count lines through `fn`. */
int tbg_generate(tbg_ctx *ctx, int form, tbg_line_fn fn, void *user, tbg_counts *counts);

/* Several materials generated together (batched and graph forms: every
 * launch covers all of them; plain form: one after the other).  Each
 * context must have its capture exits resolved; they must not depend on
 * each other.  counts receives nc entries (or is null). */
int tbg_generate_batch(tbg_ctx *const *ctxs, int nc, int form, tbg_counts *counts);
/* tbg_generate_batch in the batched form for callers that cannot allocate
 * device memory (tight serving loops, harnesses that forbid hipMalloc).  No
 * counterpart in the reference.  The batch's per-call device state (job
 * pointers, flags, work-unit prefix sums) lives in a fixed library array of
 * TBG_NOALLOC_SCRATCH_BYTES per GPU instead of memory allocated per call; it
 * calls no hipMalloc, hipFree or other allocating routine.  At most
 * TBG_NOALLOC_MAX_BATCH contexts (TBG_E_INVAL beyond); profiling must be off.
 * Every call on one GPU shares that array: calls that overlap on one GPU
 * (from several host threads) race and must not be made. */
#define TBG_NOALLOC_MAX_BATCH 256
#define TBG_NOALLOC_SCRATCH_BYTES 32768
int tbg_generate_batch_noalloc(tbg_ctx *const *ctxs, int nc, tbg_counts *counts);
/* The count lines of a generated context, in the reference's order. */
int tbg_emit_lines(const tbg_ctx *ctx, tbg_line_fn fn, void *user);

/* ---------------------------------------------------------------------
 * Served form: a continuous-batching job server.  A job is one or several
 * king slices of a material, each a synthetic slice job (tbg_create_synth:
 * seeded hash exits, codes kept).  The serving loop admits arrived jobs in
 * FIFO order while the device-memory budget and the slot limit allow, runs
 * one ply of every admitted slice per launch sequence (slices join and
 * leave the running batch at ply boundaries), and copies each finished
 * slice's codes to pinned host memory on a side stream.  When the oldest
 * waiting job does not fit, admitted jobs larger than it are preempted: at
 * a ply boundary their whole device state (bitmap planes, capture exits,
 * codes, job record) is copied to pinned host memory and their device
 * memory released; they are copied back and rejoin once memory allows,
 * bit-exact.  With TBG_FORM_PLAIN it serves one job at a time instead (the
 * comparison).  No counterpart in the reference, which generates one
 * material per run.
 *
 * Jobs may be submitted before or while the loop runs (tbg_server_serve on
 * one host thread, tbg_server_submit / tbg_job_poll from others).
 * --------------------------------------------------------------------- */
typedef struct tbg_server tbg_server;
typedef struct tbg_job_spec {
  const char *material;  /* e.g. "KQRvKQR"                                  */
  int slice;             /* first king slice, 0 <= slice < 441              */
  uint64_t seed;         /* exit hash seed (the same for every slice)       */
  double arrival_ms;     /* arrival after the start of serving; < 0: the
                            moment of submission                            */
  int nslices;           /* king slices slice, slice + 1, ... (mod 441); 0: 1,
                            at most TBG_JOB_MAX_SLICES                      */
} tbg_job_spec;
#define TBG_JOB_MAX_SLICES 16
typedef struct tbg_job_result {
  int status;            /* TBG_OK, TBG_E_BUDGET (never fits), ...          */
  int nplies;            /* the largest of the job's slices                 */
  double arrival_ms, admit_ms, done_ms; /* done = all codes on the host     */
  uint64_t positions;    /* per side and slice                              */
  uint64_t device_bytes;
  uint64_t half_steps;   /* summed over the job's slices                    */
  int nslices;
  int spills;            /* times the job was preempted to host memory      */
} tbg_job_result;
typedef struct tbg_server_stats {
  double wall_ms;
  uint64_t jobs_done, jobs_rejected;
  double mean_resident_bytes;  /* time-averaged device bytes of admitted jobs */
  uint64_t peak_resident_bytes;
  int peak_active;             /* slices in the running batch                 */
  uint64_t plies;              /* launch sequences run                        */
  double position_side_plies;  /* sum over jobs of half-steps x positions / 2 */
  uint64_t spills, restores;   /* preemptions to pinned host memory and back  */
  double spill_ms;             /* host wall time spent spilling and restoring */
} tbg_server_stats;
/* tbg_job_poll: the job is not finished yet (not an error) */
#define TBG_PENDING 1
int tbg_server_create(tbg_server **out, uint64_t budget_bytes, int max_active);
void tbg_server_destroy(tbg_server *srv);
/* Thread-safe; may be called while tbg_server_serve runs. */
int tbg_server_submit(tbg_server *srv, const tbg_job_spec *spec, uint64_t *job_id);
/* The serving loop: runs until tbg_server_close has been called and every
 * submitted job is done (thread-safe to call from another thread). */
int tbg_server_serve(tbg_server *srv, int form, tbg_server_stats *stats);
int tbg_server_close(tbg_server *srv);
/* tbg_server_close, then tbg_server_serve: every job submitted beforehand. */
int tbg_server_run(tbg_server *srv, int form, tbg_server_stats *stats);
/* TBG_OK and the result once the job is done (or rejected), TBG_PENDING
 * before; thread-safe. */
int tbg_job_poll(tbg_server *srv, uint64_t job_id, tbg_job_result *out);
int tbg_server_result(tbg_server *srv, uint64_t job_id, tbg_job_result *out);
/* codes of slice number `part` (0 .. nslices - 1) of a finished job */
int tbg_server_codes(tbg_server *srv, uint64_t job_id, int stm, uint16_t *dst, uint64_t n);
int tbg_server_codes_part(tbg_server *srv, uint64_t job_id, int part, int stm, uint16_t *dst,
                          uint64_t n);
int tbg_server_lines(tbg_server *srv, uint64_t job_id, tbg_line_fn fn, void *user);
int tbg_server_lines_part(tbg_server *srv, uint64_t job_id, int part, tbg_line_fn fn, void *user);

/* Data plane (device -> host, synchronous). */
int tbg_get_codes(tbg_ctx *ctx, int stm, int slice, uint16_t *dst, uint64_t n);
int tbg_get_bitmap(tbg_ctx *ctx, int plane, int stm, int slice, uint64_t *dst, uint64_t nwords);
int tbg_get_sub_bitmap(tbg_ctx *ctx, int kind, int colour, int slice, uint64_t *dst,
                       uint64_t nwords);

/* Predecessor scatter strategy (the reference scatters with byte-wide
 * relaxed atomic ORs, src/kslice.h:40-45):
 *   TBG_SCATTER_GLOBAL: one 64-bit device atomic OR per unmove;
 *   TBG_SCATTER_BLOCK: a work-group owns a block of the index and
 *     ORs same-block unmoves into an LDS bitmap, merged once per word;
 *   | TBG_SCATTER_COMBINE: the unmoves that still need device atomics (king
 *     unmoves, pieces whose moves leave the block) are combined per word
 *     across a wave first (a segmented OR over lanes whose targets share a
 *     word: sort-then-OR without the sort), one atomic per run.
 * Default: TBG_SCATTER_BLOCK | TBG_SCATTER_COMBINE (the autotuner also tries
 * the others). */
#define TBG_SCATTER_GLOBAL 0
#define TBG_SCATTER_BLOCK 1
#define TBG_SCATTER_COMBINE 2
int tbg_set_scatter(tbg_ctx *ctx, int mode);

/* Material scheduler: generates a material and every smaller material its
 * captures reach, smallest first, each piece count of the dependency graph
 * as one batch (tbg_generate_batch), and keeps the tables resident for the
 * materials asked for later.  Smaller materials are generated in one colour
 * orientation (the side with more, then stronger, pieces white);
 * tbg_resolve_captures maps colour-swapped names onto them (probe.c:
 * 1554-1565).  This replaces the pre-built Syzygy files the reference's
 * probe_wdl reads for capture exits (probe.c:83-97, 1522-1684).  The
 * registry owns the contexts it returns. */
typedef struct tbg_registry tbg_registry;
int tbg_registry_create(tbg_registry **out, int form);
int tbg_registry_get(tbg_registry *reg, const char *material, tbg_ctx **out);
int tbg_registry_count(const tbg_registry *reg);  /* materials resident */
/* the i-th resident material, in generation order (smallest first) */
int tbg_registry_entry(const tbg_registry *reg, int i, const char **name, tbg_ctx **ctx);
void tbg_registry_destroy(tbg_registry *reg);

/* Slice files in the reference's format (src/util.c:179-343): frames of
 * {compressed size, size, offset} (three little-endian uint64) followed by
 * the zstd level-6 compression of at most 50 KiB of the slice's bytes, in
 * any order.  Host only: no device or tbg_init needed.  Replace
 * write_data / read_data of src/util.c. */
int tbg_file_write(const char *path, const void *data, uint64_t bytes);
int tbg_file_read(const char *path, void *data, uint64_t bytes);

/* Checkpoint and resume at ply granularity, in the reference's slice-file
 * tree under root (names of src/kslice.c:306-315, <wk><bk> per slice):
 *   0/wins/<w|b>/..     wins(stm): the illegal positions and W_1..W_{n-1}
 *   <n-1>/W/<w|b>/..    W_{n-1}, the frontier of the next L step
 *   <n-1>/L/<w|b>/..    L_{n-1}, the frontier of the next W step
 *   sub/<loss|bloss|cwin|win>/<w|b>/..  capture exits (by colour)
 * plus this library's own codes/<w|b>/.. (uint16 per position) and
 * tbg_state (ply, step flags, count history).  Directories are created.
 * tbg_checkpoint needs a context past ply 1 (after tbg_continue stopped at
 * a ply, or a finished generation).  tbg_restore loads one into a fresh
 * context of the same material (and slice, seed), replacing capture
 * resolution and init; tbg_continue then runs on from the saved ply. */
int tbg_checkpoint(tbg_ctx *ctx, const char *root);
int tbg_restore(tbg_ctx *ctx, const char *root);

/* Runs the ply loop from the context's current state (after
 * tbg_init_illegal_mate, tbg_restore or an earlier stop) until the fixpoint,
 * or until ply `stop_at` is reached when stop_at > 0 (the state is then at
 * that ply boundary and can be checkpointed or continued).  Lines of the
 * plies run so far go to fn.  Continues the loop of src/tb8gen.c:1406-1444. */
int tbg_continue(tbg_ctx *ctx, int form, int stop_at, tbg_line_fn fn, void *user,
                 tbg_counts *counts);

/* The reference's intermediate files of a finished generation (with
 * codes): 0/wins, 0/L (mates) and, for every executed half-step, n/W and
 * n/L per side and slice, derived on the device from the codes. */
int tbg_export_plies(tbg_ctx *ctx, const char *root);

/* Final table writer: the finished table as a Syzygy WDL file
 * <dir>/<name>.rtbw, in the format the reference's probe_wdl maps and
 * decodes (src/probe.c:1210-1345 layout, 603-687 index, 1388-1456
 * decompression), so the reference can resolve the captures of a larger
 * material from this engine's tables.  Fills in the merge() stub of
 * src/tb8gen.c:1233-1235,1448-1452 for WDL.  Values: 2 win, 1 cursed win
 * (conversion after more than 100 plies), 0 draw, -1 blessed loss, -2 loss,
 * stored as value + 2; broken and illegal positions hold the sub-table's
 * most frequent value.  name is the file's material name in the probing
 * code's colour convention (src/probe.c:304-318); at most 7 pieces.  path
 * (nullable, cap bytes) receives the file's path. */
int tbg_rtbw_name(tbg_ctx *ctx, char *name, size_t cap);
int tbg_write_rtbw(tbg_ctx *ctx, const char *dir, char *path, size_t cap);

/* Launch geometry of the ply kernels.  Results never depend on it; only
 * speed does.  tbg_autotune picks one by measurement; tbg_set_geometry with
 * a NULL context sets the default that later contexts and servers start
 * from (a batch runs with the geometry of its first context). */
typedef struct tbg_geometry {
  int verify_unit_words;  /* bitmap words per verification work unit: multiple of 256, <= 65536;
                             slices of 2^22 words or more use at least 4096 */
  int sweep_unit_words;   /* bitmap words per W-sweep work unit: multiple of 512, <= 65536 */
  int pred_wg_per_cu;     /* grid caps, work-groups per CU, 1..64 */
  int uncap_wg_per_cu;
  int verify_wg_per_cu;
  int sweep_wg_per_cu;
  int scatter;            /* TBG_SCATTER_* */
} tbg_geometry;
int tbg_set_geometry(tbg_ctx *ctx, const tbg_geometry *g);
int tbg_get_geometry(const tbg_ctx *ctx, tbg_geometry *g);

/* Autotuner: times a fixed set of candidate geometries on ctx's own
 * generation (captures resolved first; each trial regenerates the table,
 * stopping after max_plies plies when max_plies > 0) and applies the
 * fastest: each kernel's parameters from its own device time, the
 * combination confirmed by one more trial against the best single
 * candidate.  cache_path (nullable) is a text file of earlier choices keyed
 * by material, slice, GPU architecture, CU count and max_plies: a hit
 * applies the stored geometry without trials (*ms = -1), a miss appends the
 * choice.  After a bounded run the table must be generated again.  No
 * counterpart in the reference, which has one fixed thread pool. */
int tbg_autotune(tbg_ctx *ctx, int max_plies, const char *cache_path, tbg_geometry *out,
                 double *ms);

/* Last kernel timings of the context (HIP events on its stream). */
typedef struct tbg_timing {
  double ms_total;          /* device time since tbg_set_profiling(ctx, 1)  */
  double ms_pred, ms_verify, ms_sweep, ms_uncap, ms_init, ms_resolve;
  uint64_t launches;
  uint64_t half_steps;      /* half-steps executed (L or W, one side)       */
  uint64_t n_pred, n_verify, n_sweep, n_uncap, n_init, n_resolve; /* launches */
  double ms_xr;             /* the static refutations' kernel (after init)  */
  uint64_t n_xr;
} tbg_timing;
int tbg_get_timing(const tbg_ctx *ctx, tbg_timing *t);
int tbg_set_profiling(tbg_ctx *ctx, int on);

/* King-slice numbering (host only, no device needed).  tbg_king_slice:
 * the slice of white king wk and black king bk after the mirror and a1-h8
 * flip, as KKMap[wk][bk] (src/tb8gen.c:38, 69-89); -1 for squares outside
 * 0..63, equal or adjacent kings.  tbg_slice_kings: the two squares that
 * number slice s (0 <= s < 462), as KKSquare[s] (src/probe.c:524-538);
 * TBG_E_INVAL otherwise.  Slice s names the files of a slice-file tree. */
int tbg_king_slice(int wk, int bk);
int tbg_slice_kings(int s, int *wk, int *bk);

/* Self-test of the index, symmetry and attack primitives compiled for the
 * host (no device needed).  0 on success; msg receives a description. */
int tbg_host_selftest(char *msg, int len);

#ifdef __cplusplus
}
#endif
#endif
