/*
 * swg.h -- C ABI of the MI355X (gfx950) SALSA pairwise-distance library
 * (libswg.so).  Plain C types only: pointers, sizes, int/float scalars.
 *
 * The library replaces the per-block alignment loop of the reference's C#
 * programs (DSC-SPIDAL/csharp, SalsaTPL/...):
 *
 *   delegate int ComputeBlockHandler(Block block, PartialMatrix<TDistance> matrix,
 *                                    bool isDiagonal);
 *     Salsa.SmithWatermanMPI/Program.cs:561 (implementation :203-257)
 *     Salsa.NeedlemanWunschTPL/Program.cs:977 (implementation :206-267)
 *     Salsa.SmithWatermanMS/Program.cs:1018 (adds `int rank`, impl :265-387)
 *
 * whose inner call chain is SmithWatermanGotoh.Align -> Construct -> Traceback
 * (Salsa.CoreTPL/Bio/Algorithms/SmithWaterman/SmithWatermanGotoh.cs:52-332)
 * followed by ComputeDistance (Salsa.SmithWatermanMPI/Program.cs:259-314).
 *
 * Conventions: every int-returning call returns SWG_OK (0) on success and a
 * negative SWG_E* code on failure; swg_last_error() then returns a message
 * (thread-local).  A context is used by one host thread at a time.
 */
#ifndef SWG_H
#define SWG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- return codes ---------------------------------------------------- */
#define SWG_OK                 0
#define SWG_E_ARGUMENT       (-1)  /* bad argument (reference: ArgumentException / SmithWatermanGotohException on null) */
#define SWG_E_SYMBOL         (-2)  /* residue >= 127 (reference: IndexOutOfRangeException at scores[c1,c2]) */
#define SWG_E_NONINTEGRAL    (-3)  /* scoring matrix entry or gap penalty is not an integer: integer kernels refuse it */
#define SWG_E_DEVICE         (-4)  /* HIP runtime error */
#define SWG_E_UNSUPPORTED    (-5)  /* configuration not supported by the integer kernels */
#define SWG_E_JC_UNDEFINED   (-6)  /* Jukes-Cantor log(0): reference throws Exception (Alignment.cs:342-345) */
#define SWG_E_CAPACITY       (-7)  /* request exceeds a stated size limit */

/* ---- enums mirroring the reference ----------------------------------- */
/* Salsa.CoreTPL/Bio/Algorithms/AlignmentType.cs:3-15 */
#define SWG_ALIGN_PROTEIN 0
#define SWG_ALIGN_NUCLEIC 1
#define SWG_DIST_PERCENT_IDENTITY 0
#define SWG_DIST_KIMURA2          1
#define SWG_DIST_JUKES_CANTOR     2
#define SWG_DIST_MINMAX_NORM      3
/* local (SmithWatermanGotoh, SWMPI/SWMS) or global (NeedlemanWunschTPL) */
#define SWG_MODE_SW_LOCAL  0
#define SWG_MODE_NW_GLOBAL 1
/* TDistance: Int16 (USE_INT16 builds) or Double (Salsa.SmithWatermanMPI/Program.cs:16-22) */
#define SWG_TDIST_INT16  0
#define SWG_TDIST_DOUBLE 1
/* diagonal-block semantics: SWMPI/NW compute j<i with s1 = the row (larger
 * index) and mirror (Salsa.SmithWatermanMPI/Program.cs:209-231); SWMS computes
 * every (i,j) of the block (Salsa.SmithWatermanMS/Program.cs:341-375) */
#define SWG_DIAG_LOWER_MIRROR 0
#define SWG_DIAG_FULL         1
/* which program's ComputeDistance helpers a local alignment follows: SWMPI
 * (PairwiseAlignment, Salsa.CoreTPL/Bio/Algorithms/Alignment.cs:187-348) or
 * SWMS (Salsa.SmithWatermanMS/Program.cs:417-668: PercentSimilarity counts
 * non-gap columns scoring > 0 only, Kimura2 clamped to 10, JukesCantor 10
 * instead of an exception, MinMaxNormScore).  The global mode always follows
 * Salsa.NeedlemanWunschTPL/Program.cs:269-300,547-700. */
#define SWG_PROGRAM_SWMPI 0
#define SWG_PROGRAM_SWMS  1
/* schedules */
#define SWG_SCHED_PERSISTENT     0  /* length-bucketed tiles, device work queue, hipGraph per block,
                                       variant (int32 one-pass / int16 two-pass) autotuned */
#define SWG_SCHED_PLAIN          1  /* one launch per block, natural order, fixed strip height, int32 */
#define SWG_SCHED_PERSISTENT_I32 2  /* persistent, forced int32 one-pass variant */
#define SWG_SCHED_PERSISTENT_I16 3  /* persistent, forced int16 score pass + bounded int32 re-run */
#define SWG_SCHED_PERSISTENT_LPP16 4 /* persistent, 16 lanes per pair (anti-diagonal, DPP row shifts),
                                       4 pairs per wave: latency when few row sequences are in flight */
#define SWG_SCHED_PERSISTENT_LPS16 5 /* persistent, the int16 score pass with 16 lanes per pair (8 pairs per
                                       wave) + the bounded int32 re-run, where the column sequences use
                                       at most 7 symbols, score + gap_open lies in [0, 255] for them and
                                       gap_open >= 1 (e.g. DNA with EDNAFULL), else as
                                       _PERSISTENT_LPP16.  The automatic choice for thin blocks and
                                       small serving batches */

#define SWG_MATRIX_SIZE 127      /* ScoringMatrix.SIZE, Salsa.CoreTPL/Bio/Algorithms/ScoringMatrix.cs:34 */
#define SWG_MAX_SEQ_LEN 32767    /* gap-size arrays are short in the reference (SmithWatermanGotoh.cs:83-84) */

typedef struct swg_ctx swg_ctx;

typedef struct swg_params {
    int32_t mode;            /* SWG_MODE_* */
    int32_t alignment_type;  /* SWG_ALIGN_*  (SmithWatermanSection.AlignmentType) */
    int32_t distance_fn;     /* SWG_DIST_*   (SmithWatermanSection.DistanceFunctionType); nucleic
                                alignments: PercentIdentity, Kimura2 or JukesCantor
                                (ComputeDistance, Salsa.SmithWatermanMPI/Program.cs:259-314),
                                MinMaxNormScore with SWG_PROGRAM_SWMS or the global mode;
                                protein: PercentSimilarity whatever is set */
    int32_t tdistance;       /* SWG_TDIST_* */
    int32_t diag_mode;       /* SWG_DIAG_* */
    int32_t schedule;        /* SWG_SCHED_* */
    float   gap_open;        /* _gapOpen  (float in SWMPI, Program.cs:58) */
    float   gap_ext;         /* _gapExtension (Program.cs:59) */
    const float* scores;     /* ScoringMatrix.Scores: 127*127 row-major, row = s1 symbol (ScoringMatrix.cs:155,188) */
    int32_t device;          /* HIP device ordinal */
    int32_t program;         /* SWG_PROGRAM_* (local mode) */
} swg_params;

/* Per-pair statistics (optional output, same row-major position as the
 * distance): the reference's PairwiseAlignment.Score, Sequence1.Length and
 * Identity (nucleic) or Similarity (protein).  With Kimura2 / JukesCantor the
 * counts behind those distances (Alignment.cs:254-348) instead: columns =
 * length - gapCount, count = transitions (Kimura2) or differenceCount
 * (JukesCantor), reserved = transversions (Kimura2). */
typedef struct swg_pair_stats {
    int32_t score;    /* INT32_MIN where the reference's Score is -Infinity (empty sequence) */
    int32_t columns;  /* aligned length */
    int32_t count;    /* identity (nucleic) or similarity (protein) */
    int32_t reserved;
} swg_pair_stats;

/* Create a context: validates the scoring matrix and gap penalties
 * (integral, as the integer kernels require: SWG_E_NONINTEGRAL; negative
 * penalties: SWG_E_UNSUPPORTED) and selects the device.
 * Replaces `new SmithWatermanGotoh(type)` + `ScoringMatrix.Load(name)`
 * (Salsa.SmithWatermanMPI/Program.cs:206-207). */
int  swg_create(swg_ctx** ctx, const swg_params* p);

/* Hand over the sequence set (the reference's static _sequences,
 * Salsa.SmithWatermanMPI/Program.cs:65,94-96): residues concatenated,
 * upper-case ASCII, offsets[n+1].  Copied and kept resident in HBM.
 * Kimura2 / JukesCantor refuse the gap character '-' as a residue
 * (SWG_E_UNSUPPORTED). */
int  swg_set_sequences(swg_ctx* ctx, const uint8_t* residues, const int64_t* offsets, int32_t n);

/* ComputeBlock (Salsa.SmithWatermanMPI/Program.cs:203-257): align every pair
 * of the block [row_start,row_end] x [col_start,col_end] (inclusive, global
 * sequence indices, as Range.StartIndex/EndIndex) and write TDistance values
 * into `out` (host memory, row-major, element (i,j) at
 * (i-row_start)*out_ld + (j-col_start)).  On a diagonal block with
 * SWG_DIAG_LOWER_MIRROR only j<i is aligned (s1 = sequence i) and both (i,j)
 * and (j,i) are written; (i,i) is written as 0.  `stats` (optional, host) gets
 * one swg_pair_stats per computed (i,j) at the same position.  *pairs_done
 * gets the reference's return value (number of alignments).  JukesCantor:
 * SWG_E_JC_UNDEFINED where a pair's d <= double.Epsilon (the reference throws;
 * that pair's distance is written as 0, the others as usual). */
int  swg_compute_block(swg_ctx* ctx, int32_t row_start, int32_t row_end,
                       int32_t col_start, int32_t col_end, int32_t is_diagonal,
                       void* out, int64_t out_ld, swg_pair_stats* stats, int64_t* pairs_done);

/* Same as swg_compute_block, but `dev_out`/`dev_stats` are device pointers
 * and the call only enqueues work on the context's own stream (except the
 * autotuner's first blocks of a new key, which run each candidate variant
 * synchronously on the real work).  Completes at swg_synchronize(). */
int  swg_compute_block_device(swg_ctx* ctx, int32_t row_start, int32_t row_end,
                              int32_t col_start, int32_t col_end, int32_t is_diagonal,
                              void* dev_out, int64_t out_ld, swg_pair_stats* dev_stats,
                              int64_t* pairs_done);
/* Wait for the enqueued blocks.  JukesCantor: SWG_E_JC_UNDEFINED if a pair
 * of any block since the last check had d <= double.Epsilon. */
int  swg_synchronize(swg_ctx* ctx);

/* Allocation-free variant of swg_compute_block_device.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, semantics and completion (swg_synchronize) as
 * swg_compute_block_device, local alignment only (global mode:
 * SWG_E_UNSUPPORTED); it runs the 16-lanes-per-pair int32 kernel launched
 * directly on the context stream and calls no hipMalloc, hipFree or other
 * allocating routine.  Every device buffer its kernels need besides the
 * caller's inputs and outputs (descriptor, work counters, task lists, count
 * words of the word-mode distances, row-boundary scratch) is carved from one
 * library-static device array of SWG_NOALLOC_SCRATCH_BYTES per GPU, reserved
 * when the library's code object is loaded on that GPU and held until the
 * process ends.  A block whose task lists (~8 bytes per 4 pairs) and count
 * words (4 bytes per pair and plane: Jukes-Cantor and MinMaxNormScore 1,
 * Kimura2 2) leave no room for one workgroup's boundary rows
 * ((longest sequence + 32) x 64 bytes) fails with SWG_E_CAPACITY; otherwise
 * as many workgroups run as fit, up to the resident maximum.  Every call on a
 * GPU shares that GPU's array and leaves its contents behind (the work
 * counters are reset by a fill kernel at the start of each call): calls on
 * one context are ordered by its stream, but calls that overlap on one GPU
 * from different contexts or threads race. */
#define SWG_NOALLOC_SCRATCH_BYTES (64ull << 20)
int  swg_compute_block_device_noalloc(swg_ctx* ctx, int32_t row_start, int32_t row_end,
                                      int32_t col_start, int32_t col_end, int32_t is_diagonal,
                                      void* dev_out, int64_t out_ld, swg_pair_stats* dev_stats,
                                      int64_t* pairs_done);

/* Schedule control (SWG_SCHED_*) and strip height of the int32 alignment
 * kernels (one-pass 8, 16 or 32; the two-pass re-run 16 or 32; the global
 * kernel 8 or 16; for 16 lanes per pair, 8 selects 8 rows per lane, anything
 * else 4); 0 = the defaults (16; 4 rows per lane), or autotuned with
 * SWG_SCHED_PERSISTENT.  A non-zero value limits the autotuner to the choice
 * of variant. */
int  swg_set_schedule(swg_ctx* ctx, int32_t schedule, int32_t strip_rows);
/* Tuning knobs: SWG_OPT_STREAM_CHUNK_BYTES = device output per streamed row
 * chunk of swg_compute_block (blocks larger than one chunk are computed chunk
 * by chunk while the previous chunk is copied to pinned host memory;
 * default 64 MiB). */
#define SWG_OPT_STREAM_CHUNK_BYTES 1
/* SWG_OPT_KERNEL_TIMING = 1: the following blocks run a one-thread stamp
 * kernel (the device's constant-rate wall clock) before their kernels and
 * after each of them, inside the block's captured graph, for
 * swg_last_kernel_times; 0 (default) leaves the stamps out. */
#define SWG_OPT_KERNEL_TIMING 2
/* SWG_OPT_SCORE_STRIP_ROWS = rows per strip of the int16 score pass (16, 24,
 * 32 or 40; 0 = autotuned, first tried: 32 in the max3 form, else 24).  Like swg_set_schedule's strip_rows, a
 * fixed value limits the autotuner to the choice of variant. */
#define SWG_OPT_SCORE_STRIP_ROWS 3
/* SWG_OPT_TUNE_MIN_CELLS = DP cells a block needs for the autotuner to time
 * it (default 5e10); smaller blocks run the key's current candidate untimed. */
#define SWG_OPT_TUNE_MIN_CELLS 4
/* SWG_OPT_REVERSE_STRIP_ROWS = rows per strip of the two-pass schedule's
 * reverse score pass (8 or 16; 0 = default 16: measured 54 against 84 ms on
 * a 3125 x 3125 block of the 50,000-sequence DNA workload). */
#define SWG_OPT_REVERSE_STRIP_ROWS 5
/* SWG_OPT_REVERSE_ROW_CAP = reverse rows the two-pass schedule's reverse pass
 * evaluates per pair at most (a multiple of 16; default 512); a pair whose
 * band of possible alignment starts reaches past it is re-run over its whole
 * prefix (exact either way; bounds the pass's longest tasks). */
#define SWG_OPT_REVERSE_ROW_CAP 6
/* SWG_OPT_LPS_LANES = lanes per pair of the int16 lanes-per-pair score pass
 * (SWG_SCHED_PERSISTENT_LPS16): 0 (default) chooses per block -- 32 when the
 * block's 16-lane waves would not fill the GPU (a lone query), else 16 --;
 * 16 or 32 fix it. */
#define SWG_OPT_LPS_LANES 7
/* SWG_OPT_SCORE_MAX3 = 1 (default): the int16 score pass of the two-pass
 * schedule runs its max3 form (values biased by gap_open, three-way maxima in
 * one packed instruction) where score + gap_open >= 0 for every pair of the
 * set's residue codes; 0 runs the plain form.  Results are identical. */
#define SWG_OPT_SCORE_MAX3 8
int  swg_set_option(swg_ctx* ctx, int32_t option, int64_t value);
/* Timing of the last block's kernels (HIP events on the context stream
 * around its graph or launch): milliseconds, DP cells, strip height of the
 * int32 alignment kernel and the variant that ran (1 plain, 2 int32 one-pass,
 * 3 int16 score pass + bounded int32 re-run, 4 sixteen lanes per pair: the
 * SWG_SCHED_* value). */
int  swg_last_kernel_stats(swg_ctx* ctx, double* kernel_ms, double* cells, int32_t* strip_rows,
                           int32_t* variant);

/* Launch configuration of the last block (no counterpart in the reference,
 * which runs one fixed loop per block, Salsa.SmithWatermanMPI/Program.cs:
 * 203-257,618-642): the kernel variant and its tile shape, and whether the
 * autotuner chose it.  With SWG_SCHED_PERSISTENT the autotuner searches, per
 * key (residue alphabet, mean-length class, block-size class), on the first
 * blocks of the key (each run once, on its real work): the variant (int16
 * two-pass, int32 one-pass, 16 lanes per pair) at its default strips, then
 * the winner's first strip (score pass 24/32/40, one-pass 8/16/32, rows per lane
 * 4/8, global kernel 8/16), then the two-pass re-run strip (16/32); the
 * highest cells/ms wins.  A lower-mirror diagonal block, which cannot be
 * split into trial chunks, runs a candidate only as its key's first block,
 * later ones the fastest configuration timed so far.  Mean-length classes are
 * a factor of sqrt(2) wide, centred on the first block of each class.  Thin blocks (too few rows to fill the GPU one lane
 * per pair) always take 16 lanes per pair: the int16 score pass so
 * (_PERSISTENT_LPS16) where its alphabet and scores allow, else the int32
 * kernel (_PERSISTENT_LPP16). */
typedef struct swg_block_config {
    int32_t variant;          /* SWG_SCHED_PLAIN, _PERSISTENT_I32, _PERSISTENT_I16, _PERSISTENT_LPP16
                                 or _PERSISTENT_LPS16 */
    int32_t strip_rows;       /* rows per strip of the first kernel (score pass, one-pass or global
                                 kernel) or rows per lane (16 lanes per pair) */
    int32_t rerun_strip_rows; /* two-pass: rows per strip of the int32 re-run; 0 otherwise */
    int32_t pairs_per_wave;   /* 128 score pass (two pairs per lane), 64 one-pass / global, 4 sixteen
                                 lanes per pair, 8 score pass with sixteen lanes per pair */
    int32_t tuning;           /* 0 fixed by the schedule (or a thin block), 1 an autotuner trial,
                                 2 the autotuner's choice, 3 the fastest configuration timed so
                                 far (a diagonal block while its key's search goes on) */
    int32_t trials;           /* configurations the autotuner has timed for this block's key */
} swg_block_config;
int  swg_last_block_config(swg_ctx* ctx, swg_block_config* out);

/* Kernels of a block (no counterpart in the reference: per-kernel
 * accounting for roofline measurements). */
#define SWG_KERNEL_ONE_PASS      1  /* int32 one-pass alignment (persistent or plain) */
#define SWG_KERNEL_SCORE16       2  /* int16x2 score pass */
#define SWG_KERNEL_SORT          3  /* sort of the pairs by first-maximum cell (3 small kernels) */
#define SWG_KERNEL_RERUN_TRACK   4  /* re-run of flagged pairs (saturation, empty sequence) */
#define SWG_KERNEL_RERUN_CAPTURE 5  /* re-run up to the known first-maximum cell */
#define SWG_KERNEL_LPP16         6  /* 16 lanes per pair */
#define SWG_KERNEL_GLOBAL        7  /* global (NW) alignment */
#define SWG_KERNEL_REVERSE       8  /* reverse score pass: where the first-maximum alignments can start */
#define SWG_KERNEL_RECT64        9  /* re-run of the big rectangles, one wave per pair */
#define SWG_KERNEL_DISTANCE     10  /* Jukes-Cantor / Kimura2 distances from the carried counts
                                       (Kimura2 runs the statistics kernels twice before it) */
#define SWG_KERNEL_LPS16        11  /* int16x2 score pass, 16 lanes per pair */
typedef struct swg_kernel_time {
    int32_t kernel;     /* SWG_KERNEL_* */
    int32_t reserved;
    double  ms;         /* wall-clock time from the previous kernel's end (block start for the first) */
    double  cells;      /* DP cells the kernel is charged with: the block's cells (one-pass, score
                           pass); the rectangles up to each pair's first-maximum block (reverse
                           pass); the rectangles from each pair's earliest optimal start to that
                           block (capture re-run: those of at most 65536 cells; big-rectangle
                           re-run: the others); the flagged pairs' cells (tracking re-run);
                           0 for the sorts */
    double  cells_run;  /* DP cells the kernel evaluated, idle lanes, pad rows and columns included */
} swg_kernel_time;
/* After a block run with SWG_OPT_KERNEL_TIMING (synchronizes the context
 * stream): up to `max` entries for the last block's kernels; *n = count. */
int  swg_last_kernel_times(swg_ctx* ctx, swg_kernel_time* out, int32_t max, int32_t* n);

const char* swg_last_error(void);
void swg_destroy(swg_ctx* ctx);

/* ---- serving (no counterpart in the reference, which only runs batch
 * jobs): query sequences aligned SW-Gotoh against a device-resident panel
 * with continuous batching.  s1 = the query, s2 = each panel sequence; the
 * distances follow the context semantics above.  Thread-safe: any host thread
 * may submit and wait; one internal thread issues all device work. ---- */
typedef struct swg_server swg_server;

/* panel: residues/offsets[npanel+1] as swg_set_sequences.  Queries may use
 * the panel's symbols and every symbol the scoring matrix scores.  Up to
 * `lanes` batches of at most `max_batch` queries (each <= max_query_len
 * residues) are in flight at once; a batch is admitted only while the summed
 * per-query device footprint stays within hbm_budget_bytes: per panel pair
 * sizeof(TDistance), 16 with statistics, 18 for the two-pass schedule's best
 * cells and sort slots, 4 per count-word plane (Jukes-Cantor 1, Kimura2 2)
 * and ~1 for the task tiles. */
int  swg_serve_create(swg_server** server, const swg_params* p, const uint8_t* panel_residues,
                      const int64_t* panel_offsets, int32_t npanel, int32_t max_query_len,
                      int32_t max_batch, int32_t lanes, int64_t hbm_budget_bytes, int32_t want_stats);
/* Queue one query; `out` (npanel TDistance) and `stats` (npanel entries or
 * NULL) are written when it completes and must stay valid until then. */
int  swg_serve_submit(swg_server* server, const uint8_t* query, int32_t len, void* out,
                      swg_pair_stats* stats, int64_t* ticket);
/* Wait up to timeout_ms (-1: forever) for a ticket; *done = 1 when finished. */
int  swg_serve_wait(swg_server* server, int64_t ticket, int32_t timeout_ms, int32_t* done);
/* Submit, admission and completion times of a ticket in ms since the server
 * started (-1 while not reached). */
int  swg_serve_times(swg_server* server, int64_t ticket, double* submit_ms, double* admit_ms, double* done_ms);
/* Forget a finished ticket. */
int  swg_serve_release(swg_server* server, int64_t ticket);
/* Completed queries, DP cells issued and the server clock (ms). */
int  swg_serve_stats(swg_server* server, int64_t* completed, double* cells, double* now_ms);
void swg_serve_destroy(swg_server* server);

#ifdef __cplusplus
}
#endif
#endif
