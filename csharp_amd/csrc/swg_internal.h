// Internal interface between the host scheduler (swg_api.hip) and the HIP
// kernels (swg_kernels.hip).  Not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace swg {

constexpr int kWave = 64;             // lanes per wavefront = pairs per task
constexpr int kMaxCodes = 64;         // residue codes incl. the pad code
constexpr int kNeg = -(1 << 30);      // integer stand-in for -Infinity
constexpr int kPadScore = -(1 << 20); // score of the pad code: never on an optimal path

// Per-block description, read by the kernels from device memory so that one
// captured graph can be replayed for every block.
struct BlockDesc {
    int32_t ntasks;
    int32_t mirror;        // 1: also write (b,a) (diagonal block, lower-mirror semantics)
    int64_t r0, c0, ld;    // output position of pair (a,b): (a-r0)*ld + (b-c0)
    void* out;             // TDistance array (int16 or double)
    int32_t* stats;        // optional swg_pair_stats (4 x int32) at the same position
    int32_t rows;          // rows of the block (score-pass desc; output desc of word-mode distances)
    int32_t wr, wc;        // score-pass desc: bucket size (first-maximum rows, columns) of the re-run sort
    int32_t aux_ld;        // output desc: row pitch of aux (the block's columns)
    uint32_t* aux;         // output desc: count words of the block's pairs, plane p of pair (a,b) at
                           // p*rows*aux_ld + (a-r0)*aux_ld + (b-c0)
};

// What the carried count word T = (hi << 16) | lo holds.  hi grows by one per
// diagonal column and by gap_inc >> 16 per gap column; lo by the profile's
// count flag.  For kDistFraction the alignment kernels write the distance
// 1 - lo/hi themselves; for the others they store T into aux plane
// `aux_plane` and distance_kernel turns the planes into the distance.
enum DistKind : int32_t {
    kDistFraction = 0,     // hi = columns, lo = identity / similarity
    kDistJukesCantor = 1,  // hi = diagonal columns, lo = differences (plane 0)
    kDistKimura = 2,       // hi = diagonal columns, lo = transitions (plane 0) or transversions (plane 1)
    kDistMinMax = 3,       // the alignment score (plane 0); T as for kDistFraction
};

// Read-only data of a context (sequence set + scoring).
// Bytes a sequence of L residues occupies in the residue store: its codes,
// then pad codes to the end of the word after its last one, so that every
// word from 0 to (L + 3) / 4 can be read and holds pad codes past L.
inline int64_t seq_span(int64_t L) { return ((L + 3) & ~int64_t(3)) + 4; }

struct SeqSet {
    const uint8_t* res;    // residue codes; each sequence starts 4-byte aligned, seq_span bytes
    const int64_t* off;    // start of each sequence in res
    const int32_t* len;
    int64_t pad_off;       // offset of an all-pad run of length >= maxlen + 16
    const int2* prof;      // [ncode][ncode]: (score + gap_open, carried-count increment)
    int32_t ncode;         // includes the pad code (= ncode-1)
    int32_t gap_open, gap_ext;
    int32_t tdist;         // 0 int16, 1 double
    int32_t nw_region;     // global variant: distance over the aligned region (nucleic identity)
    int32_t gap_inc;       // increment of the count word per gap column (0x10000 or 0)
    int32_t dist_kind;     // DistKind
    int32_t aux_plane;     // word-mode distances: the aux plane this pass stores
    int32_t* undefined;    // set to 1 where Jukes-Cantor is undefined (d <= double.Epsilon)
    int32_t* undefined_rows;  // optional: also set [a - r0] for the row sequence a of such a pair
    int32_t bio_helpers;   // 1: the SWMS / NW helpers (Kimura2 clamped to 10, Jukes-Cantor 10 where undefined)
    int32_t minmax_global; // MinMaxNormScore's minimum: 0 (local) or 2*open + (La+Lb)*ext (global)
    const int32_t* self_score;  // per sequence: sum of M[c][c] over its residues (MinMaxNormScore)
};

// Task lists: task t aligns s1 = task_row[t] against the 64 sequences of
// tile task_tile[t] (tile_col[tile*64 + lane], -1 = idle lane).
struct TaskSet {
    const int32_t* task_row;
    const int32_t* task_tile;
    const int32_t* tile_col;
    const int32_t* tile_maxlen;   // max (limited) length over the tile's lanes
    const BlockDesc* desc;
    int32_t* queue;               // [0]: next task (persistent schedule)
};

// Score-only pass (int16 x 2 packed): task t aligns s1 = task_row[t] against
// the 128 sequences of tile task_tile[t] (lane l: tile_col[tile*128 + l] in
// the low half, tile_col[tile*128 + 64 + l] in the high half) and writes, per
// pair, the first-maximum row bi and the first column jb of the kFirstMaxBlock-column block
// holding the first maximum, (bi << 16) | jb, into best_cell at the pair's
// position in the block, or kCellTrack | (La << 16) | Lb when the score may
// have saturated or a sequence is empty.
struct ScoreTaskSet {
    const int32_t* task_row;
    const int32_t* task_tile;
    const int32_t* tile_col;
    const int32_t* tile_maxlen;
    const BlockDesc* desc;        // ntasks, r0, c0, ld (of best_cell)
    int32_t* queue;
    uint32_t* best_cell;
    int16_t* best_score;          // the first-maximum score (valid where best_cell is not flagged)
};

// Columns per first-maximum block of the score pass: best_cell names the
// block, the re-run finds the exact column inside it.
constexpr int kFirstMaxBlock = 16;

// best_cell flag: the pair's score may have saturated int16, or a sequence
// is empty; the re-run then tracks the first maximum over the whole matrix.
constexpr uint32_t kCellTrack = 0x80000000u;

// Sort of the block's pairs by first-maximum cell for the re-run: buckets of
// wr x wc cells on a kBucketDim x kBucketDim grid, plus two lists, each
// padded to whole 64-lane tiles: flagged pairs (the last bucket) and, in the
// sort by re-run rectangle, rectangles of more than SortSet::big_rect cells,
// which one wave per pair re-runs: kBigRect for throughput, kBigRectThin for
// thin blocks, where the longest re-run of one lane bounds the latency.  The
// big list is kBigClasses buckets (from kBigBucket) by size, class c holding
// rectangles of (2^c, 2^(c+1)] x big_rect cells (the last class the rest),
// placed largest class first, so that the one-wave-per-pair queue starts the
// longest rectangles first and ends with short ones.
constexpr int kBucketDim = 128;
constexpr int kBigClasses = 8;
constexpr int kBigBucket = kBucketDim * kBucketDim;
constexpr int kNumBuckets = kBucketDim * kBucketDim + kBigClasses + 1;
constexpr int kFlaggedBucket = kNumBuckets - 1;
constexpr int64_t kBigRect = 65536;
constexpr int64_t kBigRectThin = 8192;

// meta words (int32 index; the 64-bit counters at meta64[k] = meta + 2k)
constexpr int kMetaFlaggedTiles = 0;   // tiles of the flagged list (slots first)
constexpr int kMetaTiles = 1;          // all tiles
constexpr int kMetaRectCells = 1;      // meta64: cells of the capture rectangles
constexpr int kMetaFlaggedCells = 2;   // meta64: cells of the flagged pairs
constexpr int kMetaCaptureRun = 3;     // meta64: cells evaluated by the capture re-run
constexpr int kMetaFlaggedRun = 4;     // meta64: cells evaluated by the flagged re-run
constexpr int kMetaSlots = 10;         // slots in use (flagged list padded, then the buckets)
constexpr int kMetaBigTiles = 11;      // tiles of the flagged and big lists (the buckets' first tile)
constexpr int kMetaBigCells = 6;       // meta64: cells of the big rectangles
constexpr int kMetaBigRun = 7;         // meta64: cells evaluated by the big-rectangle re-run
constexpr int kMetaWords = 64;         // zeroed with the bucket counts before every sort

struct SortSet {
    const uint32_t* best_cell;    // [rows * ld] of the score-pass desc
    const uint32_t* start_cell;   // null: key = (first-maximum score, min(bi, wend)) for the reverse pass;
                                  // else key = the re-run rectangle [si..bi] x [sj..min(jb+kFirstMaxBlock-1, Lb)]
    const int16_t* best_score;    // the first-maximum score of each pair (reverse-pass key)
    int64_t big_rect;             // rectangles of more cells go to the one-wave-per-pair re-run
    const BlockDesc* desc;        // desc[0]: output (mirror), desc[1]: score pass (r0, c0, ld, rows, wr, wc)
    const int32_t* len;           // sequence lengths
    int32_t* meta;                // kMetaWords words, then the bucket counts
    int32_t* count;               // [kNumBuckets] pairs per bucket (zeroed before the sort)
    int32_t* cursor;              // [kNumBuckets]
    int32_t* base;                // [kNumBuckets] first slot of the bucket
    uint32_t* sorted;             // slot -> pair position in best_cell, ~0u = idle lane
};

// Reverse pass: for the pairs of tiles [meta[0], meta[1]) of `sorted`, the
// topmost row and leftmost column at which an alignment ending in the
// first-maximum block with the first-maximum score can start:
// start_cell = (si << 16) | sj.
struct ReverseSet {
    const uint32_t* sorted;
    const uint32_t* best_cell;
    const int16_t* best_score;
    int32_t row_cap;              // reverse rows evaluated at most (a multiple of the strip height); a pair
                                  // whose band is still live there gets its whole prefix as rectangle
    uint32_t* start_cell;
    const BlockDesc* desc;        // as SortSet
    int32_t* meta;
    int32_t* queue;
};

// Re-run: tile t of `sorted` (64 pairs, one per lane; tiles
// [meta[kMetaBigTiles], meta[kMetaTiles]) capture the statistics at the known
// first-maximum cell, tiles [0, meta[kMetaFlaggedTiles]) track the first
// maximum).  start_cell (optional): the capture re-run covers rows si..bi and
// columns sj.. only.  The big-rectangle re-run takes the pairs of tiles
// [meta[kMetaFlaggedTiles], meta[kMetaBigTiles]), one per wave.
struct RerunSet {
    const uint32_t* sorted;
    const uint32_t* best_cell;
    const uint32_t* start_cell;
    const BlockDesc* desc;        // as SortSet
    int32_t* meta;
    int32_t* queue;
};

struct Scratch {
    int4* bnd;                    // per-slot row-boundary buffer
    int64_t bnd_stride;           // in int4, per slot
};

// Launch the local-alignment kernel.  strip_rows in {8,16,32};
// persistent: grid = nslots and tasks are drawn from queue[0];
// otherwise grid = nslots with static round-robin over tasks.
hipError_t launch_sw(const SeqSet& S, const TaskSet& T, const Scratch& W, int strip_rows,
                     bool persistent, int nslots, hipStream_t stream);
// Global kernel; region: carry the aligned-region word (nucleic
// PercentIdentity, or statistics requested).
hipError_t launch_nw(const SeqSet& S, const TaskSet& T, const Scratch& W, int strip_rows, bool region,
                     bool persistent, int nslots, hipStream_t stream);
// Score-only int16x2 pass; strip_rows in {16, 24, 32, 40}; mx3: the biased
// max3 form (needs score + gap_open >= 0 for every pair of the set's codes).
hipError_t launch_score16(const SeqSet& S, const ScoreTaskSet& T, const Scratch& W, int strip_rows, bool mx3,
                          int nslots, hipStream_t stream);
// Score-only int16x2 pass with `lanes` = 16 or 32 lanes per pair (thin
// blocks): tiles of lps_tile(lanes) = 2 * 64 / lanes column sequences (lane
// group g: tile_col[tile*T + g] in the low half, tile_col[tile*T + T/2 + g]
// in the high half), the same best_cell / best_score output; strip_rows
// (rows per lane) 8; column sequences coded < 8 with scores plus gap_open in
// [0, 255], gap_open >= 1 and a pad code < 8 or >= 12 (the host checks this:
// swg_ctx::lps_ok).
constexpr int lps_tile(int lanes) { return 2 * kWave / lanes; }
// mx3: its max3 form (always possible where the table holds the scores).
hipError_t launch_lps16(const SeqSet& S, const ScoreTaskSet& T, const Scratch& W, int strip_rows, int lanes,
                        bool mx3, int nslots, hipStream_t stream);
int lps16_blocks_per_cu(int strip_rows, int lanes, bool mx3);
// Sort of the score pass's first-maximum cells (count, scan, scatter).
hipError_t launch_sort(const SortSet& P, int grid, hipStream_t stream);
// Re-run of the sorted pairs; strip_rows in {16, 32}; track = the flagged list.
hipError_t launch_rerun(const SeqSet& S, const RerunSet& T, const Scratch& W, int strip_rows, bool track,
                        int nslots, hipStream_t stream);
int rerun_blocks_per_cu(int strip_rows, int ncode);
// Re-run of the big rectangles, one wave per pair.
hipError_t launch_rect64(const SeqSet& S, const RerunSet& T, const Scratch& W, int nslots, hipStream_t stream);
int rect64_blocks_per_cu(int ncode);
// Reverse pass; strip_rows in {8, 16}; mx3: its max3 form, values biased by
// kRevBias (needs score + gap_open >= 0 for the set's codes and row_cap times
// the largest score plus max(gap_open, gap_ext) below kRevBias).
constexpr int kRevBias = 16384;
hipError_t launch_reverse(const SeqSet& S, const ReverseSet& T, const Scratch& W, int strip_rows, bool mx3,
                          int nslots, hipStream_t stream);
int reverse_blocks_per_cu(int strip_rows, int ncode, bool mx3);
// Sixteen-lanes-per-pair kernel: tiles of kLppTile column sequences;
// strip_rows (rows per lane) in {4, 8}.
constexpr int kLppTile = 4;
hipError_t launch_lpp16(const SeqSet& S, const TaskSet& T, const Scratch& W, int strip_rows, bool persistent,
                        int nslots, hipStream_t stream);
int lpp16_blocks_per_cu(int strip_rows, int ncode);
// Occupancy: resident workgroups (one wave each) per CU for a variant.
int sw_blocks_per_cu(int strip_rows, bool nw);
int score16_blocks_per_cu(int strip_rows, int ncode, bool mx3);
int nw_blocks_per_cu(int strip_rows, bool region);
// Distances of a block from the count words in aux (Jukes-Cantor, Kimura2,
// MinMaxNormScore): desc[0] is the output desc (rows, aux_ld = columns); the
// grid (256-thread workgroups) strides over the block, whatever its size.
hipError_t launch_distance(const SeqSet& S, const BlockDesc* desc, int grid, hipStream_t stream);
// (i, i) := 0 (distance and statistics) when desc[0] is a lower-mirror
// diagonal block (desc[0].rows entries); nothing otherwise.
hipError_t launch_zero_diagonal(int tdist, const BlockDesc* desc, int grid, hipStream_t stream);
// Kernel-timing stamp: p[k] = the device's constant-rate wall clock, taken
// when every earlier kernel on the stream has finished (a node of the block's
// graph like any other kernel).
hipError_t launch_stamp(unsigned long long* p, int k, hipStream_t stream);
// Fill kernels (graph-capturable replacement for hipMemsetAsync).
hipError_t launch_fill_i32(int32_t* p, int32_t value, int64_t n, hipStream_t stream);
hipError_t launch_fill_bytes_zero(void* p, int64_t nbytes, hipStream_t stream);

}  // namespace swg
