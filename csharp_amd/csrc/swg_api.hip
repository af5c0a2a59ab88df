// Host side of libswg.so: the C ABI declared in include/swg.h, the
// block -> task scheduler (length buckets, work queue, hipGraph per block,
// autotuner of the kernel variant and its tile shape) and device-memory
// management.
#include "../../include/swg.h"
#include "swg_internal.h"

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <deque>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

using namespace swg;

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg)
{
    g_last_error = msg;
    return code;
}

#define SWG_HIP(expr)                                                                     \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess)                                                             \
            return fail(SWG_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;  // capacity in elements
    hipError_t reserve(size_t want)
    {
        if (want <= n) return hipSuccess;
        if (p) { hipError_t e = hipFree(p); if (e != hipSuccess) return e; p = nullptr; n = 0; }
        size_t bytes = ((want * sizeof(T)) + 255) & ~size_t(255);
        hipError_t e = hipMalloc(reinterpret_cast<void**>(&p), bytes);
        if (e == hipSuccess) n = bytes / sizeof(T);
        return e;
    }
    void release() { if (p) (void)hipFree(p); p = nullptr; n = 0; }
};

// Device scratch of the allocation-free entry point
// (swg_compute_block_device_noalloc): one file-scope array per GPU, reserved
// when the code object is loaded and held until the process ends.  Host code
// never passes `scratch` itself to a kernel (on the host the name is only a
// stand-in): swg_create resolves the array's device address on each device
// once (hipGetSymbolAddress, with the device current) into g_scratch_ptr.
static __device__ __align__(256) unsigned char scratch[SWG_NOALLOC_SCRATCH_BYTES];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kMaxDevices] = {};

// Resolve this device's scratch address (the device must be current).
int scratch_setup(int device)
{
    if (device < 0 || device >= kMaxDevices) return fail(SWG_E_ARGUMENT, "device ordinal out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return SWG_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(SWG_E_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[device] = static_cast<unsigned char*>(p);
    return SWG_OK;
}

unsigned char* scratch_ptr(int device)
{
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return g_scratch_ptr[device];
}

// Kernel variants a block can run with.
enum Variant : int {
    kVarPlain = 0,     // sw_kernel<16, false>: natural order, static assignment
    kVarI32 = 1,       // sw_kernel<R, true>: one pass, int32, statistics carried
    kVarI16 = 2,       // score16_kernel (int16x2) + sort by first-maximum cell + rerun_kernel
    kVarLpp = 3,       // lpp16_kernel: 16 lanes per pair, 4 pairs per wave
    kVarLps = 4,       // lps16_kernel (int16x2, 16 lanes per pair, 8 pairs per wave) + the two-pass re-run
    kNumVariants = 5,
};

constexpr int kScoreTile = 2 * kWave;  // pairs per score-pass task

// The score pass, the sorts, the reverse pass and the re-runs.
bool two_pass(int v) { return v == kVarI16 || v == kVarLps; }

// Work counters of a block (int32 words of BlockBufs::queue, zeroed by a fill
// kernel at the start of every block): one per persistent launch.
constexpr int kQueueFirst = 0;       // first kernel (one-pass, 16-lane or score pass)
constexpr int kQueueTrack = 16;      // flagged re-run
constexpr int kQueueCapture = 32;    // capture re-run
constexpr int kQueueReverse = 48;    // reverse pass
constexpr int kQueueRect = 64;       // big-rectangle re-run
constexpr int kQueueSecond = 80;     // Kimura2 second pass: + kQueueFirst / kQueueTrack / kQueueCapture / kQueueRect
constexpr int kQueueWords = 160;

// One set of per-block device buffers; two of them alternate so that the
// host can plan block k+1 while block k runs.
struct BlockBufs {
    // alignment tasks (single pass, or the bounded second pass)
    DevBuf<int32_t> task_row, task_tile, tile_col, tile_maxlen;
    // score pass
    DevBuf<int32_t> s_task_row, s_task_tile, s_tile_col, s_tile_maxlen;
    DevBuf<uint32_t> best_cell;
    DevBuf<int16_t> best_score;
    DevBuf<uint32_t> start_cell;
    DevBuf<uint32_t> sorted, sorted2;    // reverse-pass / re-run order (slot -> pair position)
    DevBuf<int32_t> sortbuf, sortbuf2;   // meta words, then bucket count, cursor, base
    DevBuf<int32_t> queue;        // work counters, see kQueue*
    DevBuf<BlockDesc> desc;       // [0]: output desc, [1]: score-pass desc
    DevBuf<uint32_t> aux;         // Kimura2: first-pass count word per pair of the block
    std::vector<int32_t> h_task_row, h_task_tile, h_tile_col, h_tile_maxlen;
    hipEvent_t done = nullptr;
    bool pending = false;
    // one captured graph per variant and grid size class (the whole device or
    // a serving lane's share), keyed by every baked launch argument; with
    // kernel timing, the graph's stamp kernels and the kernel ids they follow
    static constexpr int kMaxGraphKernels = 12;
    hipGraphExec_t graph[kNumVariants][2] = {};
    std::vector<unsigned char> graph_key[kNumVariants][2];
    int32_t graph_kids[kNumVariants][2][kMaxGraphKernels] = {};
    int graph_kn[kNumVariants][2] = {};

    void destroy_graphs()
    {
        for (auto& per_variant : graph)
            for (auto& gexec : per_variant) {
                if (gexec) (void)hipGraphExecDestroy(gexec);
                gexec = nullptr;
            }
    }

    // device memory and the completion event (graphs are destroyed by the owner)
    void release()
    {
        task_row.release(); task_tile.release(); tile_col.release(); tile_maxlen.release();
        s_task_row.release(); s_task_tile.release(); s_tile_col.release(); s_tile_maxlen.release();
        best_cell.release(); best_score.release(); start_cell.release(); sorted.release(); sorted2.release();
        sortbuf.release(); sortbuf2.release(); queue.release(); desc.release(); aux.release();
        if (done) (void)hipEventDestroy(done);
        done = nullptr;
    }
};

struct Plan {
    int64_t pairs = 0;
    double cells = 0.0;      // DP cells of the block's pairs
    double cells_run = 0.0;  // DP cells the (first) kernel evaluates, pad lanes and columns included
    int32_t ntasks = 0;      // tasks of the (first) kernel
};

// A block's launch configuration: the kernel variant and its tile shape --
// rows per strip of the first kernel (score pass, one-pass, global kernel) or
// rows per lane (16 lanes per pair), and rows per strip of the int32 re-run
// (two-pass variant only).  The pairs per wave follow from the variant: 128
// for the score pass (two per lane), 64 for the one-pass and global kernels,
// 4 for 16 lanes per pair.
struct VariantCfg {
    int v = kVarI32;
    int s1 = 16;
    int s2 = 0;
    int lanes = 16;  // kVarLps: lanes per pair (16 or 32)
    bool operator==(const VariantCfg& o) const { return v == o.v && s1 == o.s1 && s2 == o.s2 && lanes == o.lanes; }
};

// Autotuner record of one key (see choose_variant): the configurations
// measured so far and their rates, the candidates of the current stage not
// yet measured, and the choice once every stage is done.
struct TuneEntry {
    // stage: the search stage the trial belonged to (0: the variants)
    struct Trial { VariantCfg cfg; double rate; int stage; };
    int64_t key = 0;       // residue alphabet and block-size class
    double lclass = 0.0;   // length class of the key's first block (see find_entry)
    int stage = 0;
    bool done = false;
    VariantCfg best{};
    std::vector<Trial> trials;
    std::vector<VariantCfg> pending;
};

// An execution lane: a stream, two alternating buffer sets and a boundary
// scratch partition.  A context has one lane; a server has several lanes whose
// batches run concurrently on the device, each on its share of the CUs.
struct Lane {
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    BlockBufs bufs[2];
    int cur = 0;
    DevBuf<int4> bnd;
    int64_t bnd_stride = 0;
    int bnd_slots = 0;
    int share = 1;  // the lane's persistent grids use 1/share of the resident workgroups
    // Jukes-Cantor undefined: [0] sticky flag of the lane, [1 + row - r0] per
    // row of a serving batch
    DevBuf<int32_t> undef;
    bool undef_rows = false;
};

}  // namespace

struct swg_ctx {
    swg_params p{};
    int device = 0;
    int num_cu = 0;

    int32_t iscores[SWG_MATRIX_SIZE * SWG_MATRIX_SIZE];
    int32_t gap_open = 0, gap_ext = 0;
    bool count_similarity = false;  // protein: PercentSimilarity
    int32_t dist_kind = kDistFraction;  // DistKind
    bool bio_helpers = false;       // SWMS / NW ComputeDistance helpers (SWG_PROGRAM_SWMS or global mode)
    bool int16_ok = false;          // scores + penalties small enough for the int16 score pass
    bool lps_ok = false;            // and the 16-lanes-per-pair score pass (<= 8 codes, score + open in [0, 255])
    bool mx3_ok = false;            // and the score pass's max3 form (score + open >= 0 for the set's codes)
    int32_t max_score = 0;          // the largest score of the set's codes (the reverse pass's max3 form)
    bool nw = false;                // global alignment (NeedlemanWunschTPL)

    // sequences
    int32_t nseq = 0;
    int32_t maxlen = 0;
    std::vector<int32_t> h_len;
    std::vector<int64_t> h_off;
    int16_t code_of[256];
    int32_t ncode = 0;  // incl. pad
    DevBuf<uint8_t> res;
    DevBuf<int64_t> off;
    DevBuf<int32_t> len;
    DevBuf<int2> prof;
    DevBuf<int32_t> self_score;  // MinMaxNormScore: sum of M[c][c] per sequence
    int64_t pad_off = 0;

    Lane main;      // the context's own execution lane
    int32_t slot_maxlen = 0;  // longest sequence the scratch must hold (>= maxlen)

    // outputs staged on device for the host-memory entry point
    DevBuf<uint8_t> out_stage;
    DevBuf<int32_t> stats_stage;
    // streaming of large blocks: row chunks computed into alternating device
    // buffers while a copy stream moves the previous chunk to pinned memory
    hipStream_t copy_stream = nullptr;
    hipEvent_t ev_comp[2] = {nullptr, nullptr}, ev_copy[2] = {nullptr, nullptr};
    DevBuf<uint8_t> stream_dev[2];
    uint8_t* stream_host[2] = {nullptr, nullptr};
    size_t stream_host_bytes = 0;
    size_t stream_chunk_bytes = size_t(64) << 20;

    // column tiles of the last off-diagonal block (a server's panel is the
    // same column range for every batch): valid for (c0, c1, width) of this
    // sequence set
    struct ColTiles {
        int32_t c0 = -1, c1 = -1, width = 0;
        std::vector<int32_t> tcol, tmax;
        double sum_cols = 0.0;
    } col_tiles;

    int32_t schedule = SWG_SCHED_PERSISTENT;
    int32_t strip_rows = 0;   // int32 kernels' strip (re-run strip of the two-pass); 0 = autotuned
    int32_t score_strip = 0;  // score pass strip; 0 = autotuned
    int32_t reverse_strip = 16;  // reverse pass strip (SWG_OPT_REVERSE_STRIP_ROWS)
    int32_t reverse_row_cap = 512;  // reverse rows per pair at most (SWG_OPT_REVERSE_ROW_CAP)
    int32_t lps_lanes = 0;          // lanes per pair of the lanes-per-pair score pass; 0 = by block (SWG_OPT_LPS_LANES)
    int32_t score_max3 = 1;         // the score pass's max3 form where the set allows it (SWG_OPT_SCORE_MAX3)
    double tune_min_cells = 5e10;  // smaller blocks run the key's current candidate untimed
    // autotuner cache (entries are never removed while the deque lives)
    std::deque<TuneEntry> tuned;
    double last_kernel_ms = 0.0, last_cells = 0.0;
    int32_t last_strip = 0;
    int32_t last_variant = 0;
    VariantCfg last_cfg{};
    int32_t last_tuning = 0;   // 0 fixed, 1 autotuner trial, 2 autotuner choice, 3 fastest so far
    int32_t last_trials = 0;   // trials measured for the last block's key

    // kernel-timing mode: a stamp kernel before the block's kernels and after
    // each of them (inside the block's graph), wall-clock ticks at wall_khz
    bool kernel_timing = false;
    static constexpr int kMaxKernels = BlockBufs::kMaxGraphKernels;
    DevBuf<unsigned long long> kt_stamp;   // kMaxKernels + 1 stamps
    int64_t wall_khz = 0;
    int32_t kt_ids[kMaxKernels] = {};
    int kt_n = 0;
    Plan kt_plan;
    const int32_t* kt_meta = nullptr;   // device meta words of the two-pass sorts: by first-maximum cell
    const int32_t* kt_meta2 = nullptr;  // and by re-run rectangle
};

namespace {

bool is_integral(float v) { return std::isfinite(v) && std::floor(v) == v && std::fabs(v) < 32768.0f; }

int elem_size(const swg_ctx* c) { return c->p.tdistance == SWG_TDIST_DOUBLE ? 8 : 2; }

// The configuration of variant v with its default strips, or those the
// caller fixed (swg_set_schedule's strip_rows, SWG_OPT_SCORE_STRIP_ROWS).
// The score pass runs its max3 form where the sequence set allows it,
// unless SWG_OPT_SCORE_MAX3 turned it off.  (The lanes-per-pair score pass
// always can: its byte table holds score + gap_open >= 0.)
bool use_mx3(const swg_ctx* c) { return c->mx3_ok && c->score_max3 != 0; }

// The reverse pass's max3 form keeps its values biased by kRevBias: a value
// clamped at the floor must stay dead over the band's at most row_cap rows.
bool use_rev_mx3(const swg_ctx* c)
{
    return use_mx3(c) && (int64_t)c->reverse_row_cap * c->max_score + std::max(c->gap_open, c->gap_ext) < kRevBias;
}

VariantCfg default_cfg(const swg_ctx* c, int v)
{
    VariantCfg f;
    f.v = v;
    f.s2 = 0;
    if (c->nw) f.s1 = c->strip_rows == 8 ? 8 : 16;
    else if (v == kVarPlain) f.s1 = 16;
    else if (v == kVarLpp) f.s1 = c->strip_rows == 8 ? 8 : 4;
    else if (v == kVarLps) {
        f.s1 = 8;                                     // rows per lane of the score pass
        f.s2 = c->strip_rows == 32 ? 32 : 16;         // re-run
    }
    else if (v == kVarI32) f.s1 = c->strip_rows ? c->strip_rows : 16;
    else {
        // score pass: 32 rows at two waves per SIMD in the max3 form, else 24
        // rows at three
        f.s1 = c->score_strip ? c->score_strip : use_mx3(c) ? 32 : 24;
        f.s2 = c->strip_rows == 32 ? 32 : 16;         // re-run
    }
    return f;
}

int pairs_per_wave(const VariantCfg& f)
{
    return f.v == kVarI16 ? kScoreTile : f.v == kVarLpp ? kLppTile : f.v == kVarLps ? lps_tile(f.lanes) : kWave;
}

// Resident workgroups of the configuration's kernels on the lane's share of
// the CUs: n1 first kernel, n2 re-run, n3 reverse pass and n4 big-rectangle
// re-run (two-pass variant).
int slots_for(const swg_ctx* c, const Lane& ln, const VariantCfg& f, int* n1, int* n2, int* n3 = nullptr,
              int* n4 = nullptr, bool nw_region = true)
{
    const int v = f.v, s1 = f.s1, s2 = f.s2;
    int a = c->nw        ? nw_blocks_per_cu(s1, nw_region)
          : v == kVarI16 ? score16_blocks_per_cu(s1, c->ncode, use_mx3(c))
          : v == kVarLpp ? lpp16_blocks_per_cu(s1, c->ncode)
          : v == kVarLps ? lps16_blocks_per_cu(s1, f.lanes, c->score_max3 != 0)
                         : sw_blocks_per_cu(s1, false);
    int b = two_pass(v) ? rerun_blocks_per_cu(s2, c->ncode) : 0;
    int r = two_pass(v) ? reverse_blocks_per_cu(c->reverse_strip, c->ncode, use_rev_mx3(c)) : 0;
    int q = two_pass(v) ? rect64_blocks_per_cu(c->ncode) : 0;
    if (a <= 0 || (two_pass(v) && (b <= 0 || r <= 0 || q <= 0))) return fail(SWG_E_DEVICE, "occupancy query failed");
    *n1 = std::max(1, a * c->num_cu / ln.share);
    *n2 = std::max(1, b * c->num_cu / ln.share);
    if (n3) *n3 = std::max(1, r * c->num_cu / ln.share);
    if (n4) *n4 = std::max(1, q * c->num_cu / ln.share);
    return SWG_OK;
}

// Boundary scratch of a lane for `slots` resident workgroups (HBM
// budget-capped: 24 GB over all lanes).
int ensure_scratch(swg_ctx* c, Lane& ln, int* slots)
{
    // the global kernel keeps two more words per column (second half of the slot)
    const int64_t stride = (int64_t)(std::max(c->maxlen, c->slot_maxlen) + 16) * kWave * (c->nw ? 2 : 1);
    const double budget = 24.0 * (1u << 30) / ln.share;
    const int step = std::max(1, c->num_cu / ln.share);
    while (*slots > step && (double)(*slots) * (double)stride * sizeof(int4) > budget) *slots -= step;
    if (ln.bnd.p && ln.bnd_stride == stride && ln.bnd_slots >= *slots) return SWG_OK;
    if (ln.bnd_stride != stride) ln.bnd.release();
    SWG_HIP(ln.bnd.reserve((size_t)(*slots) * (size_t)stride));
    ln.bnd_stride = stride;
    ln.bnd_slots = *slots;
    return SWG_OK;
}

// ---- task planning ------------------------------------------------------

struct BlockGeom {
    int32_t r0, r1, c0, c1;
    bool diag_lower;
};

// Rows/columns of the block sorted by length, longest first (persistent
// variants) or natural order (plain).
void order_block(const swg_ctx* c, const BlockGeom& g, bool sorted, std::vector<int32_t>& rows,
                 std::vector<int32_t>& cols)
{
    const std::vector<int32_t>& L = c->h_len;
    rows.resize((size_t)(g.r1 - g.r0 + 1));
    std::iota(rows.begin(), rows.end(), g.r0);
    cols.resize((size_t)(g.c1 - g.c0 + 1));
    std::iota(cols.begin(), cols.end(), g.c0);
    if (sorted) {
        auto longer = [&](int32_t x, int32_t y) { return L[x] > L[y]; };
        std::stable_sort(rows.begin(), rows.end(), longer);
        std::stable_sort(cols.begin(), cols.end(), longer);
    }
}

// Tiles of `width` column sequences; task = (row sequence, tile).  Off-
// diagonal blocks share one tile set across rows; a lower-mirror diagonal
// block gets per-row tiles of its columns j < i.  Tasks are emitted row by
// row, rows longest first, so the queue drains long work first.
void plan_tiles(swg_ctx* c, const BlockGeom& g, bool sorted, int width, std::vector<int32_t>& trow,
                std::vector<int32_t>& ttile, std::vector<int32_t>& tcol, std::vector<int32_t>& tmax, Plan& plan)
{
    const std::vector<int32_t>& L = c->h_len;
    trow.clear(); ttile.clear(); tcol.clear(); tmax.clear();
    std::vector<int32_t> rows, cols;
    swg_ctx::ColTiles& ct = c->col_tiles;
    const bool cached = sorted && !g.diag_lower && ct.c0 == g.c0 && ct.c1 == g.c1 && ct.width == width;
    if (cached) {
        // rows only: the columns' sorted tiles are those of the last block
        BlockGeom gr = g;
        gr.c0 = 0; gr.c1 = -1;
        order_block(c, gr, sorted, rows, cols);
    } else {
        order_block(c, g, sorted, rows, cols);
    }
    auto emit_tile = [&](const int32_t* ids, int n) -> int32_t {
        const int32_t t = (int32_t)tmax.size();
        int32_t mx = 0;
        for (int k = 0; k < width; ++k) {
            const int32_t id = k < n ? ids[k] : -1;
            tcol.push_back(id);
            if (id >= 0) mx = std::max(mx, L[id]);
        }
        tmax.push_back(mx);
        return t;
    };
    plan.pairs = 0;
    plan.cells = 0.0;
    plan.cells_run = 0.0;
    if (!g.diag_lower) {
        double sum_cols = 0.0;
        if (cached) {
            tcol = ct.tcol;
            tmax = ct.tmax;
            sum_cols = ct.sum_cols;
        } else {
            for (size_t k = 0; k < cols.size(); k += width)
                emit_tile(&cols[k], (int)std::min<size_t>(width, cols.size() - k));
            for (int32_t x : cols) sum_cols += L[x];
            if (sorted) {
                ct.c0 = g.c0; ct.c1 = g.c1; ct.width = width;
                ct.tcol = tcol; ct.tmax = tmax; ct.sum_cols = sum_cols;
            }
        }
        const int32_t ntiles = (int32_t)tmax.size();
        for (int32_t a : rows) {
            for (int32_t t = 0; t < ntiles; ++t) { trow.push_back(a); ttile.push_back(t); }
            plan.cells += (double)L[a] * sum_cols;
        }
        plan.pairs = (int64_t)rows.size() * (int64_t)(g.c1 - g.c0 + 1);
    } else {
        std::vector<int32_t> valid;
        valid.reserve(cols.size());
        for (int32_t a : rows) {
            valid.clear();
            for (int32_t x : cols) if (x < a) valid.push_back(x);
            for (size_t k = 0; k < valid.size(); k += width) {
                trow.push_back(a);
                ttile.push_back(emit_tile(&valid[k], (int)std::min<size_t>(width, valid.size() - k)));
            }
            for (int32_t x : valid) plan.cells += (double)L[a] * L[x];
            plan.pairs += (int64_t)valid.size();
        }
    }
    plan.ntasks = (int32_t)trow.size();
    for (size_t k = 0; k < trow.size(); ++k)
        plan.cells_run += (double)L[trow[k]] * (double)((tmax[ttile[k]] + 3) & ~3) * width;
}

template <typename T>
int upload(hipStream_t st, DevBuf<T>& d, const std::vector<T>& h)
{
    SWG_HIP(d.reserve(std::max<size_t>(h.size(), 1)));
    if (!h.empty()) SWG_HIP(hipMemcpyAsync(d.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, st));
    return SWG_OK;
}

// The kernels' view of the sequence set.  Kimura2 runs the statistics
// kernels twice: the first pass (this SeqSet) stores the transitions' count
// words into aux plane 0, the second (`second`: the profile's second table)
// the transversions' into plane 1.
SeqSet make_seqset(const swg_ctx* c, const Lane& ln, SeqSet* second)
{
    SeqSet S{};
    S.res = c->res.p; S.off = c->off.p; S.len = c->len.p; S.pad_off = c->pad_off;
    S.prof = c->prof.p; S.ncode = c->ncode;
    S.gap_open = c->gap_open; S.gap_ext = c->gap_ext;
    S.tdist = c->p.tdistance == SWG_TDIST_DOUBLE ? 1 : 0;
    S.nw_region = c->nw && !c->count_similarity ? 1 : 0;
    S.gap_inc = c->dist_kind == kDistFraction || c->dist_kind == kDistMinMax ? 0x10000 : 0;
    S.dist_kind = c->dist_kind;
    S.aux_plane = 0;
    S.undefined = ln.undef.p;
    S.undefined_rows = ln.undef_rows ? ln.undef.p + 1 : nullptr;
    S.bio_helpers = c->bio_helpers ? 1 : 0;
    S.minmax_global = c->nw ? 1 : 0;
    S.self_score = c->self_score.p;
    *second = S;
    if (c->dist_kind == kDistKimura) {
        second->aux_plane = 1;
        second->prof = c->prof.p + (size_t)c->ncode * c->ncode;
    }
    return S;
}

template <typename T>
void append_key(std::vector<unsigned char>& k, const T& v)
{
    const unsigned char* p = reinterpret_cast<const unsigned char*>(&v);
    k.insert(k.end(), p, p + sizeof(T));
}

// Plan, upload and launch one block with configuration f.
int run_variant(swg_ctx* c, Lane& ln, BlockBufs& B, const BlockGeom& g, const VariantCfg& f, void* dev_out,
                int64_t ld, int32_t* dev_stats, Plan& plan)
{
    const int v = f.v, s1 = f.s1, s2 = f.s2;
    // the global kernel carries the aligned-region word only where it is read:
    // nucleic PercentIdentity, or the statistics (their fourth field)
    const bool nw_region =
        c->nw && ((!c->count_similarity && c->dist_kind == kDistFraction) || dev_stats != nullptr);
    int n1, n2, n3, n4;
    int rc = slots_for(c, ln, f, &n1, &n2, &n3, &n4, nw_region);
    if (rc) return rc;
    int want = std::max(std::max(n1, n2), two_pass(v) ? std::max(n3, n4) : 0);
    rc = ensure_scratch(c, ln, &want);
    if (rc) return rc;
    n1 = std::min(n1, want);
    n2 = std::min(n2, want);
    n3 = std::min(n3, want);
    n4 = std::min(n4, want);

    const hipStream_t st = ln.stream;
    BlockDesc desc[2] = {};
    if (two_pass(v)) {
        plan_tiles(c, g, true, v == kVarLps ? lps_tile(f.lanes) : kScoreTile, B.h_task_row, B.h_task_tile, B.h_tile_col,
                   B.h_tile_maxlen, plan);
        if ((rc = upload(st, B.s_task_row, B.h_task_row)) || (rc = upload(st, B.s_task_tile, B.h_task_tile)) ||
            (rc = upload(st, B.s_tile_col, B.h_tile_col)) || (rc = upload(st, B.s_tile_maxlen, B.h_tile_maxlen)))
            return rc;
        const size_t cells = (size_t)(g.r1 - g.r0 + 1) * (size_t)(g.c1 - g.c0 + 1);
        // slots: every pair, the paddings of the flagged and big lists and the last tile's
        if (cells + 3 * (size_t)kWave >= (size_t(1) << 31))
            return fail(SWG_E_CAPACITY, "block too large for the two-pass schedule (2^31 pairs)");
        SWG_HIP(B.best_cell.reserve(std::max<size_t>(cells, 1)));
        SWG_HIP(B.best_score.reserve(std::max<size_t>(cells, 1)));
        SWG_HIP(B.start_cell.reserve(std::max<size_t>(cells, 1)));
        SWG_HIP(B.sorted.reserve(cells + 3 * (size_t)kWave));
        SWG_HIP(B.sorted2.reserve(cells + 3 * (size_t)kWave));
        SWG_HIP(B.sortbuf.reserve(kMetaWords + 3 * kNumBuckets));
        SWG_HIP(B.sortbuf2.reserve(kMetaWords + 3 * kNumBuckets));
        // bucket size: the block's longest sequences over kBucketDim buckets
        // (slots: every pair, the flagged list's padding and the last tile's)
        int32_t max_r = 1, max_c = 1;
        for (int32_t i = g.r0; i <= g.r1; ++i) max_r = std::max(max_r, c->h_len[i]);
        for (int32_t j = g.c0; j <= g.c1; ++j) max_c = std::max(max_c, c->h_len[j]);
        desc[0].ntasks = 0;
        desc[1].ntasks = plan.ntasks;
        desc[1].r0 = g.r0; desc[1].c0 = g.c0; desc[1].ld = g.c1 - g.c0 + 1;
        desc[1].rows = g.r1 - g.r0 + 1;
        desc[1].wr = (max_r + kBucketDim - 1) / kBucketDim;
        desc[1].wc = (max_c + kBucketDim - 1) / kBucketDim;
    } else {
        plan_tiles(c, g, v != kVarPlain, v == kVarLpp ? kLppTile : kWave, B.h_task_row, B.h_task_tile, B.h_tile_col,
                   B.h_tile_maxlen, plan);
        if ((rc = upload(st, B.task_row, B.h_task_row)) || (rc = upload(st, B.task_tile, B.h_task_tile)) ||
            (rc = upload(st, B.tile_col, B.h_tile_col)) || (rc = upload(st, B.tile_maxlen, B.h_tile_maxlen)))
            return rc;
        desc[0].ntasks = plan.ntasks;
    }
    desc[0].mirror = g.diag_lower ? 1 : 0;
    desc[0].r0 = g.r0; desc[0].c0 = g.c0; desc[0].ld = ld;
    desc[0].out = dev_out; desc[0].stats = dev_stats;
    const bool kimura = c->dist_kind == kDistKimura;
    const bool words = c->dist_kind != kDistFraction;  // distance_kernel after the alignment kernels
    const int brows = g.r1 - g.r0 + 1, bcols = g.c1 - g.c0 + 1;
    desc[0].rows = brows;
    if (words) {
        SWG_HIP(B.aux.reserve(std::max<size_t>((size_t)(kimura ? 2 : 1) * brows * bcols, 1)));
        desc[0].aux = B.aux.p;
        desc[0].aux_ld = bcols;
    }
    SWG_HIP(B.desc.reserve(2));
    SWG_HIP(B.queue.reserve(kQueueWords));
    SWG_HIP(hipMemcpyAsync(B.desc.p, desc, sizeof(desc), hipMemcpyHostToDevice, st));

    SeqSet S2;
    const SeqSet S = make_seqset(c, ln, &S2);
    Scratch W;
    W.bnd = ln.bnd.p;
    W.bnd_stride = ln.bnd_stride;
    TaskSet T{};
    T.task_row = B.task_row.p; T.task_tile = B.task_tile.p; T.tile_col = B.tile_col.p;
    T.tile_maxlen = B.tile_maxlen.p; T.desc = B.desc.p; T.queue = B.queue.p;
    ScoreTaskSet ST{};
    SortSet P{}, P2{};
    ReverseSet RV{};
    RerunSet RT{}, RC{}, RB{};
    const int sort_grid = std::max(1, c->num_cu * 2 / ln.share);
    // the distance and diagonal kernels stride over the block from its
    // descriptor: a fixed grid, so that a captured graph fits every block size
    const int pair_grid = std::max(1, c->num_cu * 4 / ln.share);
    const int tdist = c->p.tdistance == SWG_TDIST_DOUBLE ? 1 : 0;
    if (two_pass(v)) {
        ST.task_row = B.s_task_row.p; ST.task_tile = B.s_task_tile.p; ST.tile_col = B.s_tile_col.p;
        ST.tile_maxlen = B.s_tile_maxlen.p; ST.desc = B.desc.p + 1; ST.queue = B.queue.p; ST.best_cell = B.best_cell.p;
        ST.best_score = B.best_score.p;
        // sort 1: by first-maximum cell, for the reverse pass
        P.best_cell = B.best_cell.p; P.start_cell = nullptr; P.desc = B.desc.p; P.len = c->len.p;
        P.best_score = B.best_score.p;
        P.big_rect = v == kVarLps ? kBigRectThin : kBigRect;
        P.meta = B.sortbuf.p; P.count = B.sortbuf.p + kMetaWords;
        P.cursor = P.count + kNumBuckets; P.base = P.count + 2 * kNumBuckets; P.sorted = B.sorted.p;
        RV.sorted = B.sorted.p; RV.best_cell = B.best_cell.p; RV.best_score = B.best_score.p;
        RV.start_cell = B.start_cell.p; RV.desc = B.desc.p; RV.meta = P.meta; RV.queue = B.queue.p + kQueueReverse;
        RV.row_cap = c->reverse_row_cap;
        // sort 2: by re-run rectangle
        P2 = P;
        P2.start_cell = B.start_cell.p;
        P2.meta = B.sortbuf2.p; P2.count = B.sortbuf2.p + kMetaWords;
        P2.cursor = P2.count + kNumBuckets; P2.base = P2.count + 2 * kNumBuckets; P2.sorted = B.sorted2.p;
        RT.sorted = B.sorted2.p; RT.best_cell = B.best_cell.p; RT.start_cell = nullptr; RT.desc = B.desc.p;
        RT.meta = P2.meta;
        RC = RT;
        RC.start_cell = B.start_cell.p;
        RT.queue = B.queue.p + kQueueTrack;
        RC.queue = B.queue.p + kQueueCapture;
        RB = RC;
        RB.queue = B.queue.p + kQueueRect;
    }
    // Kimura2 second pass: the same launches on fresh work counters
    TaskSet T2 = T;
    T2.queue = B.queue.p + kQueueSecond + kQueueFirst;
    RerunSet RT2 = RT, RC2 = RC, RB2 = RB;
    RT2.queue = B.queue.p + kQueueSecond + kQueueTrack;
    RC2.queue = B.queue.p + kQueueSecond + kQueueCapture;
    RB2.queue = B.queue.p + kQueueSecond + kQueueRect;

    // the block's kernels in stream order; with `mark`, a stamp kernel runs
    // first and after each kernel (kernel-timing mode)
    const bool timing = c->kernel_timing && &ln == &c->main;
    int32_t ids[BlockBufs::kMaxGraphKernels] = {};
    int nids = 0;
    auto launch_all = [&](bool mark) -> hipError_t {
        int nk = 0;
        auto after = [&](int kernel_id) -> hipError_t {
            if (!mark) return hipSuccess;
            if (nk >= BlockBufs::kMaxGraphKernels) return hipErrorInvalidValue;
            ids[nk++] = kernel_id;
            return launch_stamp(c->kt_stamp.p, nk, st);
        };
        hipError_t e = mark ? launch_stamp(c->kt_stamp.p, 0, st) : hipSuccess;
        if (e == hipSuccess) e = launch_fill_i32(B.queue.p, 0, kQueueWords, st);
        if (e == hipSuccess) e = launch_zero_diagonal(tdist, B.desc.p, pair_grid, st);
        if (e != hipSuccess) return e;
        const bool persist = v != kVarPlain;
        if (c->nw) {
            e = launch_nw(S, T, W, s1, nw_region, persist, n1, st);
            if (e == hipSuccess) e = after(SWG_KERNEL_GLOBAL);
            if (e == hipSuccess && kimura) e = launch_nw(S2, T2, W, s1, nw_region, persist, n1, st);
            if (e == hipSuccess && kimura) e = after(SWG_KERNEL_GLOBAL);
        } else if (v == kVarI32 || v == kVarPlain) {
            e = launch_sw(S, T, W, s1, persist, n1, st);
            if (e == hipSuccess) e = after(SWG_KERNEL_ONE_PASS);
            if (e == hipSuccess && kimura) e = launch_sw(S2, T2, W, s1, persist, n1, st);
            if (e == hipSuccess && kimura) e = after(SWG_KERNEL_ONE_PASS);
        } else if (v == kVarLpp) {
            e = launch_lpp16(S, T, W, s1, true, n1, st);
            if (e == hipSuccess) e = after(SWG_KERNEL_LPP16);
            if (e == hipSuccess && kimura) e = launch_lpp16(S2, T2, W, s1, true, n1, st);
            if (e == hipSuccess && kimura) e = after(SWG_KERNEL_LPP16);
        } else {
            e = launch_fill_i32(P.meta, 0, kMetaWords + kNumBuckets, st);
            if (e == hipSuccess) e = launch_fill_i32(P2.meta, 0, kMetaWords + kNumBuckets, st);
            if (e == hipSuccess)
                e = v == kVarLps ? launch_lps16(S, ST, W, s1, f.lanes, c->score_max3 != 0, n1, st) : launch_score16(S, ST, W, s1, use_mx3(c), n1, st);
            if (e == hipSuccess) e = after(v == kVarLps ? SWG_KERNEL_LPS16 : SWG_KERNEL_SCORE16);
            if (e == hipSuccess) e = launch_sort(P, sort_grid, st);
            if (e == hipSuccess) e = after(SWG_KERNEL_SORT);
            if (e == hipSuccess) e = launch_reverse(S, RV, W, c->reverse_strip, use_rev_mx3(c), n3, st);
            if (e == hipSuccess) e = after(SWG_KERNEL_REVERSE);
            if (e == hipSuccess) e = launch_sort(P2, sort_grid, st);
            if (e == hipSuccess) e = after(SWG_KERNEL_SORT);
            if (e == hipSuccess) e = launch_rerun(S, RT, W, s2, true, n2, st);
            if (e == hipSuccess) e = after(SWG_KERNEL_RERUN_TRACK);
            if (e == hipSuccess) e = launch_rerun(S, RC, W, s2, false, n2, st);
            if (e == hipSuccess) e = after(SWG_KERNEL_RERUN_CAPTURE);
            if (e == hipSuccess) e = launch_rect64(S, RB, W, n4, st);
            if (e == hipSuccess) e = after(SWG_KERNEL_RECT64);
            if (kimura) {
                if (e == hipSuccess) e = launch_rerun(S2, RT2, W, s2, true, n2, st);
                if (e == hipSuccess) e = after(SWG_KERNEL_RERUN_TRACK);
                if (e == hipSuccess) e = launch_rerun(S2, RC2, W, s2, false, n2, st);
                if (e == hipSuccess) e = after(SWG_KERNEL_RERUN_CAPTURE);
                if (e == hipSuccess) e = launch_rect64(S2, RB2, W, n4, st);
                if (e == hipSuccess) e = after(SWG_KERNEL_RECT64);
            }
        }
        if (words) {
            if (e == hipSuccess) e = launch_distance(S, B.desc.p, pair_grid, st);
            if (e == hipSuccess) e = after(SWG_KERNEL_DISTANCE);
        }
        nids = nk;
        return e;
    };

    // capture (or re-capture) before the timing events so that host-side
    // capture time is not counted as device time
    const int gs = ln.share == 1 ? 0 : 1;
    if (plan.ntasks > 0 && v != kVarPlain) {
        // every pointer and scalar baked into the graph is part of its key:
        // re-capture whenever one changed (new sequences, grown buffers)
        std::vector<unsigned char> key;
        append_key(key, S); append_key(key, T); append_key(key, W); append_key(key, ST); append_key(key, P);
        append_key(key, RT); append_key(key, RC); append_key(key, sort_grid); append_key(key, P2);
        append_key(key, RV); append_key(key, n3); append_key(key, RB); append_key(key, n4);
        append_key(key, s1); append_key(key, s2); append_key(key, n1); append_key(key, n2);
        append_key(key, S2); append_key(key, T2); append_key(key, RT2); append_key(key, RC2); append_key(key, RB2);
        append_key(key, timing); append_key(key, c->kt_stamp.p); append_key(key, pair_grid); append_key(key, tdist);
        append_key(key, c->reverse_strip); append_key(key, nw_region); append_key(key, f.lanes);
        append_key(key, use_mx3(c)); append_key(key, c->score_max3); append_key(key, use_rev_mx3(c));
        hipGraphExec_t& gexec = B.graph[v][gs];
        std::vector<unsigned char>& gkey = B.graph_key[v][gs];
        if (!gexec || key != gkey) {
            if (gexec) { (void)hipGraphExecDestroy(gexec); gexec = nullptr; }
            hipGraph_t gr = nullptr;
            SWG_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            const hipError_t e = launch_all(timing);
            hipError_t e2 = hipStreamEndCapture(st, &gr);
            if (e != hipSuccess || e2 != hipSuccess) {
                if (gr) (void)hipGraphDestroy(gr);
                return fail(SWG_E_DEVICE, std::string("graph capture failed: ") +
                                              hipGetErrorString(e != hipSuccess ? e : e2));
            }
            hipError_t e3 = hipGraphInstantiate(&gexec, gr, nullptr, nullptr, 0);
            (void)hipGraphDestroy(gr);
            if (e3 != hipSuccess) { gexec = nullptr; return fail(SWG_E_DEVICE, "hipGraphInstantiate failed"); }
            gkey = key;
            B.graph_kn[v][gs] = nids;
            std::copy(ids, ids + nids, B.graph_kids[v][gs]);
        }
    }
    if (timing) {
        c->kt_n = 0;
        c->kt_plan = plan;
        c->kt_meta = two_pass(v) ? P.meta : nullptr;
        c->kt_meta2 = two_pass(v) ? P2.meta : nullptr;
    }
    SWG_HIP(hipEventRecord(ln.ev0, st));
    if (plan.ntasks == 0 && g.diag_lower) SWG_HIP(launch_zero_diagonal(tdist, B.desc.p, pair_grid, st));
    if (plan.ntasks > 0) {
        if (v == kVarPlain) {
            SWG_HIP(launch_all(timing));
        } else {
            SWG_HIP(hipGraphLaunch(B.graph[v][gs], st));
            nids = B.graph_kn[v][gs];
            std::copy(B.graph_kids[v][gs], B.graph_kids[v][gs] + nids, ids);
        }
        if (timing) {
            c->kt_n = nids;
            std::copy(ids, ids + nids, c->kt_ids);
        }
    }
    SWG_HIP(hipEventRecord(ln.ev1, st));
    SWG_HIP(hipEventRecord(B.done, st));
    B.pending = true;
    c->last_strip = two_pass(v) ? s2 : s1;
    c->last_variant = v;
    c->last_cfg = f;
    return SWG_OK;
}

BlockBufs& next_bufs(Lane& ln, int* rc)
{
    BlockBufs& B = ln.bufs[ln.cur];
    ln.cur ^= 1;
    *rc = SWG_OK;
    if (B.pending) {
        hipError_t e = hipEventSynchronize(B.done);
        if (e != hipSuccess) *rc = fail(SWG_E_DEVICE, std::string("hipEventSynchronize: ") + hipGetErrorString(e));
        B.pending = false;
    }
    return B;
}

// Autotuner.  Per key (residue alphabet, mean-length class, block-size
// class) the first blocks of the key run candidate configurations, each once
// and synchronously, on real work (nothing is recomputed); the one with the
// highest cells/ms wins.  A lower-mirror diagonal block runs a candidate only
// as its key's first block; later ones, while the search goes on, take the
// fastest configuration timed so far.  The search is staged, one dimension at a time:
//   stage 0  the variant at its default tile shape: int16 two-pass (score
//            strip 32 in the max3 form, else 24; re-run strip 16), int32
//            one-pass (strip 16), 16 lanes per pair (4 rows per lane) --
//            lanes per pair 1 or 16, int16 or int32; the two-pass first for
//            alphabets of <= 8 codes, the one-pass first for larger ones, so
//            that a run of a single block gets the variant that measured
//            faster on such data (the host entry point times the variants
//            after the first on trial chunks, see trial_rows);
//   stage 1  the winner's first strip: score strips 24, 32 and 40 (16, never
//            faster on measured data, only when fixed), int32 strips 8 and 32,
//            rows per lane 8 <-> 4, the global kernel's 8 <-> 16;
//   stage 2  the two-pass re-run strip 16 <-> 32.
// Strips fixed by the caller (swg_set_schedule's strip_rows,
// SWG_OPT_SCORE_STRIP_ROWS) end the search after stage 0.

int64_t tune_key(const swg_ctx* c, const BlockGeom& g)
{
    // two size classes only: the host entry point splits a big block into
    // row chunks of different sizes, and each chunk should not re-open the
    // trials (the first chunks time one variant each, the rest use the faster)
    const double pairs = (double)(g.r1 - g.r0 + 1) * (double)(g.c1 - g.c0 + 1) * (g.diag_lower ? 0.5 : 1.0);
    const int size_class = pairs >= 4e6 ? 1 : 0;
    return ((int64_t)c->ncode << 40) | (int64_t)size_class;
}

// 2 log2(1 + mean length) over the block's row and column sequences: one
// unit is a factor of sqrt(2) in length.
double length_class(const swg_ctx* c, const BlockGeom& g)
{
    const std::vector<int32_t>& L = c->h_len;
    double sum = 0.0;
    int64_t cnt = 0;
    for (int32_t i = g.r0; i <= g.r1; ++i) { sum += L[i]; ++cnt; }
    for (int32_t i = g.c0; i <= g.c1; ++i) { sum += L[i]; ++cnt; }
    return cnt ? 2.0 * std::log2(1.0 + sum / (double)cnt) : 0.0;
}

// The autotuner record of block g: among the records of its alphabet and
// size class, the one whose first block's length class is nearest, if less
// than half a unit away; else a new record (create) centred on this block.
// Classes centred on the blocks seen, not fixed bins, keep blocks of nearly
// equal mean length together: the 16 row ranges of a 50,000-sequence set
// have means of 833-865 residues, across the bin edge 2 log2(1 + m) = 19.5.
TuneEntry* find_entry(swg_ctx* c, const BlockGeom& g, bool create)
{
    const int64_t key = tune_key(c, g);
    const double lc = length_class(c, g);
    TuneEntry* t = nullptr;
    double dist = 0.5;
    for (auto& e : c->tuned)
        if (e.key == key && std::fabs(e.lclass - lc) < dist) { t = &e; dist = std::fabs(e.lclass - lc); }
    if (!t && create) {
        c->tuned.emplace_back();
        t = &c->tuned.back();
        t->key = key;
        t->lclass = lc;
    }
    return t;
}


// Candidates of the entry's current stage (see above).
std::vector<VariantCfg> stage_candidates(const swg_ctx* c, const TuneEntry& t)
{
    std::vector<VariantCfg> out;
    if (t.stage == 0) {
        if (c->nw) { out.push_back(default_cfg(c, kVarI32)); return out; }
        const bool small_alphabet = c->ncode <= 8;
        if (c->int16_ok && small_alphabet) out.push_back(default_cfg(c, kVarI16));
        out.push_back(default_cfg(c, kVarI32));
        if (c->int16_ok && !small_alphabet) out.push_back(default_cfg(c, kVarI16));
        out.push_back(default_cfg(c, kVarLpp));
        return out;
    }
    if (c->strip_rows != 0 || c->score_strip != 0) return out;
    const VariantCfg b = t.best;
    VariantCfg f = b;
    // the strip stages start with the incumbent again when several variants
    // were timed: their rates come from trial chunks and other blocks (a
    // diagonal block, a block computed whole), not comparable with the
    // strips' few percent
    int variants_timed = 0;
    for (const auto& x : t.trials) variants_timed += x.stage == 0 ? 1 : 0;
    if (t.stage == 1 && variants_timed >= 2) out.push_back(b);
    if (t.stage == 1) {
        if (c->nw) { f.s1 = b.s1 == 16 ? 8 : 16; out.push_back(f); }
        else if (b.v == kVarI16)
            for (int r : {24, 32, 40})
                if (r != b.s1) { f.s1 = r; out.push_back(f); }
        else if (b.v == kVarLpp) { f.s1 = b.s1 == 4 ? 8 : 4; out.push_back(f); }
        else if (b.v == kVarI32)
            for (int r : {8, 16, 32})
                if (r != b.s1) { f.s1 = r; out.push_back(f); }
    } else if (t.stage == 2 && b.v == kVarI16 && !c->nw) {
        f.s2 = b.s2 == 16 ? 32 : 16;
        out.push_back(f);
    }
    return out;
}

// Move the entry on once its stage's candidates are all measured: the best
// so far (after the variant stage, among the strip stages' trials only),
// then the next stage's candidates not yet timed in the strip stages, or
// done.
void tune_advance(const swg_ctx* c, TuneEntry& t)
{
    while (t.pending.empty() && !t.done) {
        double best_rate = -1.0;
        for (const auto& x : t.trials)
            if (x.rate > best_rate && (t.stage <= 1 || x.stage >= 1)) { best_rate = x.rate; t.best = x.cfg; }
        if (t.stage >= 3) { t.done = true; break; }
        for (const VariantCfg& f : stage_candidates(c, t)) {
            bool tried = false;
            for (const auto& x : t.trials) tried = tried || (x.cfg == f && t.stage >= 1 && x.stage >= 1);
            if (!tried) t.pending.push_back(f);
        }
        ++t.stage;
    }
}

// Too few row sequences to fill the GPU with one-lane-per-pair waves: the
// 16-lanes-per-pair kernel gives each pair 16 lanes instead.
bool thin_block(const swg_ctx* c, const Lane& ln, const BlockGeom& g)
{
    const double rows = g.r1 - g.r0 + 1, cols = g.c1 - g.c0 + 1;
    const double waves = g.diag_lower ? rows * std::ceil(cols / 2.0 / kWave) : rows * std::ceil(cols / kWave);
    int n1 = 0, n2 = 0;
    if (slots_for(c, ln, default_cfg(c, kVarI32), &n1, &n2) != SWG_OK) return false;
    return waves < 2.0 * n1;
}

// The lanes-per-pair score pass for block g: 32 lanes per pair when its
// 16-lane waves would not fill the lane's resident slots (a lone query: each
// wave then runs half as many stripes), else 16 (less stripe padding).
VariantCfg lps_cfg(const swg_ctx* c, const Lane& ln, const BlockGeom& g)
{
    VariantCfg f = default_cfg(c, kVarLps);
    const double rows = g.r1 - g.r0 + 1, cols = g.c1 - g.c0 + 1;
    const double tile = lps_tile(16);
    const double waves = g.diag_lower ? rows * std::ceil(cols / 2.0 / tile) : rows * std::ceil(cols / tile);
    const int slots = std::max(1, lps16_blocks_per_cu(f.s1, 16, c->score_max3 != 0) * c->num_cu / ln.share);
    if (waves < slots) f.lanes = 32;
    if (c->lps_lanes) f.lanes = c->lps_lanes;
    return f;
}

// The configuration of a block: fixed by the schedule, the 16-lane kernel for
// a thin block, or the autotuner's (*entry: its record; *trial: the
// configuration is a candidate still to be measured).
int choose_variant(swg_ctx* c, const Lane& ln, const BlockGeom& g, VariantCfg* cfg, TuneEntry** entry, bool* trial)
{
    *entry = nullptr;
    *trial = false;
    if (c->schedule == SWG_SCHED_PLAIN) { *cfg = default_cfg(c, kVarPlain); return SWG_OK; }
    if (c->nw && (c->schedule == SWG_SCHED_PERSISTENT_I16 || c->schedule == SWG_SCHED_PERSISTENT_LPP16 ||
                  c->schedule == SWG_SCHED_PERSISTENT_LPS16))
        return fail(SWG_E_UNSUPPORTED, "the global alignment has only the int32 one-pass kernel");
    if (c->schedule == SWG_SCHED_PERSISTENT_LPP16) { *cfg = default_cfg(c, kVarLpp); return SWG_OK; }
    if (c->schedule == SWG_SCHED_PERSISTENT_LPS16) {
        *cfg = c->lps_ok ? lps_cfg(c, ln, g) : default_cfg(c, kVarLpp);
        return SWG_OK;
    }
    if (c->schedule == SWG_SCHED_PERSISTENT_I32) { *cfg = default_cfg(c, kVarI32); return SWG_OK; }
    if (c->schedule == SWG_SCHED_PERSISTENT_I16 && c->int16_ok) { *cfg = default_cfg(c, kVarI16); return SWG_OK; }
    if (!c->nw && thin_block(c, ln, g)) { *cfg = c->lps_ok ? lps_cfg(c, ln, g) : default_cfg(c, kVarLpp); return SWG_OK; }
    TuneEntry* t = find_entry(c, g, true);
    tune_advance(c, *t);
    *entry = t;
    if (t->done) { *cfg = t->best; return SWG_OK; }
    if (g.diag_lower && !t->trials.empty()) {
        // a lower-mirror triangle cannot be split into trial chunks: rather
        // than run a candidate that may be several times slower on all of
        // it, it takes the fastest configuration timed so far
        double best_rate = -1.0;
        for (const auto& x : t->trials)
            if (x.rate > best_rate) { best_rate = x.rate; *cfg = x.cfg; }
        return SWG_OK;
    }
    *cfg = t->pending.front();
    *trial = true;
    return SWG_OK;
}

// True while the autotuner has timed the first of its kernel variants for
// the key of block g and still has others to time (its next block of at
// least tune_min_cells runs one).
bool tuning_pending(const swg_ctx* c, const Lane& ln, const BlockGeom& g)
{
    if (c->schedule != SWG_SCHED_PERSISTENT || (!c->nw && thin_block(c, ln, g))) return false;
    const TuneEntry* e = find_entry(const_cast<swg_ctx*>(c), g, false);
    return e && !e->done && e->stage <= 1 && !e->trials.empty();
}

// Rows of a trial chunk: the variants after the first one (the expected
// best, timed on a whole chunk) may run several times slower, so the host
// entry point times them on chunks of about kTrialPairs pairs (still in the
// large size class of tune_key) rather than on whole ones; their rates only
// decide the variant, whose differences are large.  The strip stages, whose
// candidates differ by a few percent, are timed on whole chunks and compared
// among themselves (a small chunk's end of the persistent queue weighs more,
// and differently for each strip height).
constexpr double kTrialPairs = 4.5e6;
int32_t trial_rows(int64_t cols) { return (int32_t)std::max<int64_t>(1, (int64_t)std::ceil(kTrialPairs / cols)); }

// Enqueue one block on a lane's stream into device buffers.
int enqueue_block(swg_ctx* c, Lane& ln, int32_t r0, int32_t r1, int32_t c0, int32_t c1, int32_t is_diag,
                  void* dev_out, int64_t ld, int32_t* dev_stats, int64_t* pairs_done)
{
    if (!c->res.p) return fail(SWG_E_ARGUMENT, "swg_set_sequences has not been called");
    if (r0 < 0 || c0 < 0 || r1 < r0 - 1 || c1 < c0 - 1 || r1 >= c->nseq || c1 >= c->nseq)
        return fail(SWG_E_ARGUMENT, "block range outside the sequence set");
    BlockGeom g;
    g.r0 = r0; g.r1 = r1; g.c0 = c0; g.c1 = c1;
    g.diag_lower = is_diag && c->p.diag_mode == SWG_DIAG_LOWER_MIRROR;
    if (g.diag_lower && (r0 != c0 || r1 != c1))
        return fail(SWG_E_ARGUMENT, "a diagonal block must have equal row and column ranges");
    if (ld < (int64_t)(c1 - c0 + 1)) return fail(SWG_E_ARGUMENT, "out_ld smaller than the block width");
    if (r1 < r0 || c1 < c0) {
        if (pairs_done) *pairs_done = 0;
        return SWG_OK;
    }
    VariantCfg cfg;
    TuneEntry* entry = nullptr;
    bool trial = false;
    int rc = choose_variant(c, ln, g, &cfg, &entry, &trial);
    if (rc) return rc;
    BlockBufs& B = next_bufs(ln, &rc);
    if (rc) return rc;
    Plan plan;
    rc = run_variant(c, ln, B, g, cfg, dev_out, ld, dev_stats, plan);
    if (rc) return rc;
    c->last_cells = plan.cells;
    if (trial && plan.cells >= c->tune_min_cells) {
        SWG_HIP(hipStreamSynchronize(ln.stream));
        B.pending = false;
        float ms = 0.0f;
        SWG_HIP(hipEventElapsedTime(&ms, ln.ev0, ln.ev1));
        c->last_kernel_ms = ms;
        entry->trials.push_back({cfg, plan.cells / std::max(1e-3, (double)ms), entry->stage - 1});
        entry->pending.erase(entry->pending.begin());
        tune_advance(c, *entry);
    }
    c->last_tuning = trial ? 1 : !entry ? 0 : entry->done ? 2 : 3;
    c->last_trials = entry ? (int32_t)entry->trials.size() : 0;
    if (pairs_done) *pairs_done = plan.pairs;
    return SWG_OK;
}

// Residue coding + device upload of a sequence set, optionally followed by
// `nslots` empty slots of capacity `slot_cap` residues (serving queries are
// written into them later).  With `matrix_symbols`, every symbol that has a
// non-zero score also gets a code so that later queries can use it; the
// set's own symbols are then coded first (the slots are only ever row
// sequences, so the column alphabet stays the set's).
int load_sequences(swg_ctx* c, const uint8_t* residues, const int64_t* offsets, int32_t n, int32_t nslots,
                   int32_t slot_cap, bool matrix_symbols)
{
    if (!c || n < 0 || nslots < 0 || slot_cap < 0 || (n > 0 && (!residues || !offsets)))
        return fail(SWG_E_ARGUMENT, "null argument");
    if (slot_cap > SWG_MAX_SEQ_LEN) return fail(SWG_E_CAPACITY, "query capacity above 32767 residues");
    SWG_HIP(hipSetDevice(c->device));
    SWG_HIP(hipStreamSynchronize(c->main.stream));
    for (auto& B : c->main.bufs) B.pending = false;
    c->tuned.clear();
    c->col_tiles = swg_ctx::ColTiles();
    // residue codes: distinct symbols of the set (and of the matrix), in ASCII order, + pad
    bool seen[256] = {false};
    int32_t maxlen = slot_cap;
    for (int32_t i = 0; i < n; ++i) {
        const int64_t L = offsets[i + 1] - offsets[i];
        if (L < 0) return fail(SWG_E_ARGUMENT, "offsets must be non-decreasing");
        if (L > SWG_MAX_SEQ_LEN) return fail(SWG_E_CAPACITY, "sequence longer than 32767 residues");
        maxlen = std::max<int32_t>(maxlen, (int32_t)L);
        for (int64_t k = offsets[i]; k < offsets[i + 1]; ++k) seen[residues[k]] = true;
    }
    if (c->nw) {
        // global scores have no 0 floor: every value must stay far from the
        // int32 stand-in for -Infinity
        int64_t maxabs = 0;
        for (int k = 0; k < SWG_MATRIX_SIZE * SWG_MATRIX_SIZE; ++k) maxabs = std::max<int64_t>(maxabs, std::abs(c->iscores[k]));
        const int64_t bound = 2 * ((int64_t)maxlen + 8) * (std::abs((int64_t)c->gap_open) + std::abs((int64_t)c->gap_ext) + maxabs);
        if (bound >= (int64_t(1) << 28))
            return fail(SWG_E_CAPACITY, "global alignment scores could exceed the int32 range at this sequence length");
    }
    if (matrix_symbols)
        for (int x = 0; x < SWG_MATRIX_SIZE; ++x)
            for (int y = 0; y < SWG_MATRIX_SIZE; ++y)
                if (c->iscores[x * SWG_MATRIX_SIZE + y] != 0) seen[x] = seen[y] = true;
    if (c->dist_kind == kDistJukesCantor || c->dist_kind == kDistKimura) {
        // Kimura2 / Jukes-Cantor count a column holding the gap character as a
        // gap whatever the pointer: residues '-' would need per-residue gap
        // increments, which the kernels do not carry
        for (int32_t i = 0; i < n; ++i)
            for (int64_t k = offsets[i]; k < offsets[i + 1]; ++k)
                if (residues[k] == '-')
                    return fail(SWG_E_UNSUPPORTED, "Kimura2/JukesCantor with the gap character '-' in a sequence");
        seen[(int)'-'] = false;
    }
    bool in_set[256] = {false};
    for (int32_t i = 0; i < n; ++i)
        for (int64_t k = offsets[i]; k < offsets[i + 1]; ++k) in_set[residues[k]] = true;
    int32_t ncode = 0;
    for (int s = 0; s < 256; ++s) c->code_of[s] = -1;
    for (int pass = 0; pass < 2; ++pass)
        for (int s = 0; s < 256; ++s) {
            if (!seen[s] || c->code_of[s] >= 0 || (pass == 0 && !in_set[s])) continue;
            if (s >= SWG_MATRIX_SIZE) return fail(SWG_E_SYMBOL, "residue symbol >= 127 (the reference indexes a 127x127 matrix)");
            c->code_of[s] = (int16_t)ncode++;
        }
    int32_t set_codes = 0;  // codes of the set's own symbols: 0 .. set_codes - 1
    for (int s = 0; s < 256; ++s)
        if (in_set[s] && c->code_of[s] >= 0) ++set_codes;
    if (ncode + 1 > kMaxCodes) return fail(SWG_E_CAPACITY, "more than 63 distinct residue symbols");
    const int32_t pad = ncode;
    c->ncode = ncode + 1;
    std::vector<uint8_t> sym_of(c->ncode, 0);
    for (int s = 0; s < 256; ++s) if (c->code_of[s] >= 0) sym_of[c->code_of[s]] = (uint8_t)s;

    // profile table (score + gap_open, increment of the packed count word);
    // Kimura2 has a second table for its second pass
    const int ntables = c->dist_kind == kDistKimura ? 2 : 1;
    std::vector<int2> prof((size_t)ntables * c->ncode * c->ncode);
    for (int tb = 0; tb < ntables; ++tb)
        for (int x = 0; x < c->ncode; ++x)
            for (int y = 0; y < c->ncode; ++y) {
                int2 v;
                if (x == pad || y == pad) { v.x = kPadScore; v.y = 0; }
                else {
                    const int sx = sym_of[x], sy = sym_of[y];
                    const int sc = c->iscores[sx * SWG_MATRIX_SIZE + sy];
                    int cnt = (sx == sy) ? 1 : 0;  // identity: c1 == c2 (SmithWatermanGotoh.cs:284)
                    if (c->count_similarity && sx != sy && sc > 0) cnt = 1;  // similarity: scores[c1,c2] > 0 (:290)
                    // SWMS and global protein distance: non-gap columns scoring > 0 only
                    // (Salsa.SmithWatermanMS/Program.cs:521-541, Salsa.NeedlemanWunschTPL/Program.cs:583-604)
                    if (c->bio_helpers && c->count_similarity) cnt = (sc > 0 && sx != '-' && sy != '-') ? 1 : 0;
                    // Jukes-Cantor: differences; Kimura2: transitions (A<->G,
                    // C<->T), then transversions (Alignment.cs:268-283, 328-337)
                    const bool ts = (sx == 'A' && sy == 'G') || (sx == 'G' && sy == 'A') ||
                                    (sx == 'C' && sy == 'T') || (sx == 'T' && sy == 'C');
                    if (c->dist_kind == kDistJukesCantor) cnt = sx != sy ? 1 : 0;
                    if (c->dist_kind == kDistKimura) cnt = tb == 0 ? (ts ? 1 : 0) : (sx != sy && !ts ? 1 : 0);
                    v.x = sc + c->gap_open;
                    v.y = 0x10000 + cnt;
                }
                prof[((size_t)tb * c->ncode + x) * c->ncode + y] = v;
            }

    // the 16-lanes-per-pair score pass keeps, per row, score + gap_open of
    // column codes 0..7 as bytes and scores the pad code -gap_open (exact
    // only for gap_open >= 1: pad cells then stay below every real
    // maximum); its column sequences are the set's, whose codes must be
    // < 8, and the pad code must be < 8 or >= 12 (it selects the byte 0)
    c->lps_ok = c->int16_ok && !c->nw && c->gap_open >= 1 && set_codes <= 7 && (pad < 8 || pad >= 12);
    for (int x = 0; x < pad && c->lps_ok; ++x)
        for (int y = 0; y < set_codes; ++y)
            if (prof[(size_t)x * c->ncode + y].x < 0 || prof[(size_t)x * c->ncode + y].x > 255) c->lps_ok = false;
    // the score pass's max3 form keeps every value biased by gap_open and
    // >= 0: score + gap_open >= 0 for every pair of codes
    c->mx3_ok = c->int16_ok && !c->nw;
    c->max_score = 0;
    for (int x = 0; x < pad; ++x)
        for (int y = 0; y < pad; ++y) {
            if (prof[(size_t)x * c->ncode + y].x < 0) c->mx3_ok = false;
            c->max_score = std::max(c->max_score, prof[(size_t)x * c->ncode + y].x - c->gap_open);
        }

    // residue codes, each sequence 4-byte aligned and followed by pad codes
    // up to the end of its next word (seq_span), then the slots, then a tail
    // of pad codes
    const int32_t ntot = n + nslots;
    std::vector<int64_t> off(std::max(ntot, 1));
    int64_t total = 0;
    for (int32_t i = 0; i < n; ++i) {
        off[i] = total;
        total += seq_span(offsets[i + 1] - offsets[i]);
    }
    for (int32_t q = 0; q < nslots; ++q) {
        off[n + q] = total;
        total += seq_span(slot_cap);
    }
    const int64_t pad_off = total;
    total += maxlen + 64;
    std::vector<uint8_t> codes((size_t)total, (uint8_t)pad);
    std::vector<int32_t> lens(std::max(ntot, 1), 0);
    for (int32_t i = 0; i < n; ++i) {
        const int64_t L = offsets[i + 1] - offsets[i];
        lens[i] = (int32_t)L;
        for (int64_t k = 0; k < L; ++k) codes[off[i] + k] = (uint8_t)c->code_of[residues[offsets[i] + k]];
    }
    SWG_HIP(c->res.reserve(codes.size()));
    SWG_HIP(c->off.reserve(off.size()));
    SWG_HIP(c->len.reserve(lens.size()));
    SWG_HIP(c->prof.reserve(prof.size()));
    SWG_HIP(hipMemcpy(c->res.p, codes.data(), codes.size(), hipMemcpyHostToDevice));
    SWG_HIP(hipMemcpy(c->off.p, off.data(), off.size() * 8, hipMemcpyHostToDevice));
    SWG_HIP(hipMemcpy(c->len.p, lens.data(), lens.size() * 4, hipMemcpyHostToDevice));
    SWG_HIP(hipMemcpy(c->prof.p, prof.data(), prof.size() * sizeof(int2), hipMemcpyHostToDevice));
    if (c->dist_kind == kDistMinMax) {
        // si.Sum(b => mat[b, b]) (Salsa.SmithWatermanMS/Program.cs:547-548)
        std::vector<int32_t> self((size_t)std::max(ntot, 1), 0);
        for (int32_t i = 0; i < n; ++i)
            for (int64_t k = offsets[i]; k < offsets[i + 1]; ++k)
                self[i] += c->iscores[residues[k] * SWG_MATRIX_SIZE + residues[k]];
        SWG_HIP(c->self_score.reserve(self.size()));
        SWG_HIP(hipMemcpy(c->self_score.p, self.data(), self.size() * 4, hipMemcpyHostToDevice));
    }
    c->pad_off = pad_off;
    c->nseq = ntot;
    c->h_len.assign(lens.begin(), lens.begin() + ntot);
    c->h_off.assign(off.begin(), off.begin() + ntot);
    if (maxlen != c->maxlen) { c->main.bnd.release(); c->main.bnd_slots = 0; }
    c->maxlen = maxlen;
    c->slot_maxlen = slot_cap;
    return SWG_OK;
}

// Jukes-Cantor pairs with d <= double.Epsilon since the last check of the
// main lane (the lane's stream must be idle): the reference throws
// Exception("Jukes and Cantor Distance Undefined - Log(Zero)") for them.
int take_undefined(swg_ctx* c)
{
    if (c->dist_kind != kDistJukesCantor || c->bio_helpers || !c->main.undef.p) return SWG_OK;
    int32_t flag = 0;
    SWG_HIP(hipMemcpy(&flag, c->main.undef.p, sizeof(flag), hipMemcpyDeviceToHost));
    if (!flag) return SWG_OK;
    SWG_HIP(hipMemset(c->main.undef.p, 0, sizeof(flag)));
    return fail(SWG_E_JC_UNDEFINED, "Jukes and Cantor Distance Undefined - Log(Zero)");
}

}  // namespace

extern "C" {

const char* swg_last_error(void) { return g_last_error.c_str(); }

int swg_create(swg_ctx** out, const swg_params* p)
{
    if (!out || !p) return fail(SWG_E_ARGUMENT, "null argument");
    *out = nullptr;
    if (!p->scores) return fail(SWG_E_ARGUMENT, "scores is null");
    if (p->mode != SWG_MODE_SW_LOCAL && p->mode != SWG_MODE_NW_GLOBAL) return fail(SWG_E_ARGUMENT, "bad mode");
    if (p->alignment_type != SWG_ALIGN_PROTEIN && p->alignment_type != SWG_ALIGN_NUCLEIC)
        return fail(SWG_E_ARGUMENT, "bad alignment_type");
    if (p->tdistance != SWG_TDIST_INT16 && p->tdistance != SWG_TDIST_DOUBLE) return fail(SWG_E_ARGUMENT, "bad tdistance");
    if (p->diag_mode != SWG_DIAG_LOWER_MIRROR && p->diag_mode != SWG_DIAG_FULL) return fail(SWG_E_ARGUMENT, "bad diag_mode");
    if (p->schedule < SWG_SCHED_PERSISTENT || p->schedule > SWG_SCHED_PERSISTENT_LPS16)
        return fail(SWG_E_ARGUMENT, "bad schedule");
    if (p->program != SWG_PROGRAM_SWMPI && p->program != SWG_PROGRAM_SWMS) return fail(SWG_E_ARGUMENT, "bad program");
    if (p->distance_fn < SWG_DIST_PERCENT_IDENTITY || p->distance_fn > SWG_DIST_MINMAX_NORM)
        return fail(SWG_E_ARGUMENT, "bad distance_fn");
    if (p->alignment_type == SWG_ALIGN_NUCLEIC && p->distance_fn == SWG_DIST_MINMAX_NORM &&
        p->mode == SWG_MODE_SW_LOCAL && p->program != SWG_PROGRAM_SWMS)
        return fail(SWG_E_UNSUPPORTED, "MinMaxNormScore is a distance of the SWMS and global programs");

    if (!is_integral(p->gap_open) || !is_integral(p->gap_ext))
        return fail(SWG_E_NONINTEGRAL, "gap penalties must be integers below 32768 for the integer kernels");
    // penalties are costs subtracted from the score (SmithWatermanGotoh.cs:159-187);
    // the kernels keep the horizontal gap state as max(H, 0) with an unsigned
    // saturating subtract of gap_ext and run pad rows/columns as DP borders,
    // both of which need gap_open >= 0 and gap_ext >= 0
    if (p->gap_open < 0.0f || p->gap_ext < 0.0f)
        return fail(SWG_E_UNSUPPORTED, "negative gap penalties are not supported (gap_open and gap_ext must be >= 0)");
    for (int k = 0; k < SWG_MATRIX_SIZE * SWG_MATRIX_SIZE; ++k)
        if (!is_integral(p->scores[k]))
            return fail(SWG_E_NONINTEGRAL, "scoring matrix entries must be integers below 32768 for the integer kernels");

    swg_ctx* c = new swg_ctx();
    c->p = *p;
    c->p.scores = nullptr;
    for (int k = 0; k < SWG_MATRIX_SIZE * SWG_MATRIX_SIZE; ++k) c->iscores[k] = (int32_t)p->scores[k];
    c->gap_open = (int32_t)p->gap_open;
    c->gap_ext = (int32_t)p->gap_ext;
    c->count_similarity = p->alignment_type == SWG_ALIGN_PROTEIN;
    // ComputeDistance (Salsa.SmithWatermanMPI/Program.cs:259-285): the
    // distance function applies to nucleic alignments; protein uses
    // PercentSimilarity whatever it says
    if (p->alignment_type == SWG_ALIGN_NUCLEIC && p->distance_fn == SWG_DIST_KIMURA2) c->dist_kind = kDistKimura;
    if (p->alignment_type == SWG_ALIGN_NUCLEIC && p->distance_fn == SWG_DIST_JUKES_CANTOR)
        c->dist_kind = kDistJukesCantor;
    if (p->alignment_type == SWG_ALIGN_NUCLEIC && p->distance_fn == SWG_DIST_MINMAX_NORM) c->dist_kind = kDistMinMax;
    c->bio_helpers = p->mode == SWG_MODE_NW_GLOBAL || p->program == SWG_PROGRAM_SWMS;
    c->nw = p->mode == SWG_MODE_NW_GLOBAL;
    c->schedule = p->schedule;
    {
        // the int16 score pass holds score + gap_open and the pad score in int16
        int maxabs = 0;
        for (int k = 0; k < SWG_MATRIX_SIZE * SWG_MATRIX_SIZE; ++k) maxabs = std::max(maxabs, std::abs(c->iscores[k]));
        c->int16_ok = maxabs + c->gap_open <= 8192 && c->gap_ext <= 8192;
    }
    c->device = p->device;

    auto bail = [&](hipError_t e, const char* what) {
        std::string m = std::string(what) + ": " + hipGetErrorString(e);
        swg_destroy(c);
        return fail(SWG_E_DEVICE, m);
    };
    hipError_t e = hipSetDevice(c->device);
    if (e != hipSuccess) return bail(e, "hipSetDevice");
    {
        const int rc = scratch_setup(c->device);
        if (rc) { std::string m = g_last_error; swg_destroy(c); return fail(rc, m); }
    }
    hipDeviceProp_t prop;
    e = hipGetDeviceProperties(&prop, c->device);
    if (e != hipSuccess) return bail(e, "hipGetDeviceProperties");
    c->num_cu = prop.multiProcessorCount;
    e = hipStreamCreateWithFlags(&c->main.stream, hipStreamNonBlocking);
    if (e != hipSuccess) return bail(e, "hipStreamCreate");
    if ((e = hipEventCreate(&c->main.ev0)) != hipSuccess) return bail(e, "hipEventCreate");
    if ((e = hipEventCreate(&c->main.ev1)) != hipSuccess) return bail(e, "hipEventCreate");
    if ((e = c->kt_stamp.reserve(swg_ctx::kMaxKernels + 1)) != hipSuccess) return bail(e, "hipMalloc");
    {
        int khz = 0;
        if ((e = hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, c->device)) != hipSuccess)
            return bail(e, "hipDeviceGetAttribute");
        c->wall_khz = khz;
    }
    if ((e = c->main.undef.reserve(1)) != hipSuccess) return bail(e, "hipMalloc");
    if ((e = hipMemset(c->main.undef.p, 0, sizeof(int32_t))) != hipSuccess) return bail(e, "hipMemset");
    for (auto& B : c->main.bufs)
        if ((e = hipEventCreateWithFlags(&B.done, hipEventDisableTiming)) != hipSuccess) return bail(e, "hipEventCreate");
    *out = c;
    return SWG_OK;
}

int swg_set_sequences(swg_ctx* c, const uint8_t* residues, const int64_t* offsets, int32_t n)
{
    return load_sequences(c, residues, offsets, n, 0, 0, false);
}

int swg_compute_block_device(swg_ctx* c, int32_t r0, int32_t r1, int32_t c0, int32_t c1, int32_t is_diag,
                             void* dev_out, int64_t ld, swg_pair_stats* dev_stats, int64_t* pairs_done)
{
    if (!c || !dev_out) return fail(SWG_E_ARGUMENT, "null argument");
    SWG_HIP(hipSetDevice(c->device));
    return enqueue_block(c, c->main, r0, r1, c0, c1, is_diag, dev_out, ld, reinterpret_cast<int32_t*>(dev_stats),
                         pairs_done);
}

int swg_synchronize(swg_ctx* c)
{
    if (!c) return fail(SWG_E_ARGUMENT, "null argument");
    SWG_HIP(hipStreamSynchronize(c->main.stream));
    for (auto& B : c->main.bufs) B.pending = false;
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, c->main.ev0, c->main.ev1) == hipSuccess) c->last_kernel_ms = ms;
    return take_undefined(c);
}

}  // extern "C"

namespace {

// Carve `bytes` (256-byte aligned) from the scratch array; null when it does
// not fit.
unsigned char* carve(unsigned char* base, size_t* used, size_t bytes)
{
    const size_t at = *used;
    const size_t want = (bytes + 255) & ~size_t(255);
    if (want > SWG_NOALLOC_SCRATCH_BYTES - at) return nullptr;
    *used = at + want;
    return base + at;
}

// swg_compute_block_device_noalloc: the 16-lanes-per-pair int32 kernel
// (statistics carried, one pass) launched directly on the context stream,
// with the block's task lists, descriptor, work counters, count words and
// row-boundary scratch all carved from the device's scratch array.
int enqueue_block_noalloc(swg_ctx* c, int32_t r0, int32_t r1, int32_t c0, int32_t c1, int32_t is_diag,
                          void* dev_out, int64_t ld, int32_t* dev_stats, int64_t* pairs_done)
{
    if (!c->res.p) return fail(SWG_E_ARGUMENT, "swg_set_sequences has not been called");
    if (c->nw) return fail(SWG_E_UNSUPPORTED, "the allocation-free entry point runs the local alignment only");
    if (r0 < 0 || c0 < 0 || r1 < r0 - 1 || c1 < c0 - 1 || r1 >= c->nseq || c1 >= c->nseq)
        return fail(SWG_E_ARGUMENT, "block range outside the sequence set");
    BlockGeom g;
    g.r0 = r0; g.r1 = r1; g.c0 = c0; g.c1 = c1;
    g.diag_lower = is_diag && c->p.diag_mode == SWG_DIAG_LOWER_MIRROR;
    if (g.diag_lower && (r0 != c0 || r1 != c1))
        return fail(SWG_E_ARGUMENT, "a diagonal block must have equal row and column ranges");
    if (ld < (int64_t)(c1 - c0 + 1)) return fail(SWG_E_ARGUMENT, "out_ld smaller than the block width");
    if (r1 < r0 || c1 < c0) {
        if (pairs_done) *pairs_done = 0;
        return SWG_OK;
    }
    unsigned char* base = scratch_ptr(c->device);
    if (!base) return fail(SWG_E_DEVICE, "scratch address of this device was not resolved");
    int rc;
    Lane& ln = c->main;
    BlockBufs& B = next_bufs(ln, &rc);  // waits until this set's host task lists are free again
    if (rc) return rc;
    const VariantCfg f = default_cfg(c, kVarLpp);
    const int v = f.v, s1 = f.s1;
    Plan plan;
    plan_tiles(c, g, true, kLppTile, B.h_task_row, B.h_task_tile, B.h_tile_col, B.h_tile_maxlen, plan);

    // scratch layout: descriptor, work counters, task lists, count words,
    // then as many workgroups' boundary rows as fit (at most one per
    // resident workgroup)
    const bool kimura = c->dist_kind == kDistKimura;
    const bool words = c->dist_kind != kDistFraction;
    const int brows = g.r1 - g.r0 + 1, bcols = g.c1 - g.c0 + 1;
    size_t used = 0;
    auto* d_desc = reinterpret_cast<BlockDesc*>(carve(base, &used, sizeof(BlockDesc)));
    auto* d_queue = reinterpret_cast<int32_t*>(carve(base, &used, kQueueWords * sizeof(int32_t)));
    auto* d_task_row = reinterpret_cast<int32_t*>(carve(base, &used, B.h_task_row.size() * sizeof(int32_t)));
    auto* d_task_tile = reinterpret_cast<int32_t*>(carve(base, &used, B.h_task_tile.size() * sizeof(int32_t)));
    auto* d_tile_col = reinterpret_cast<int32_t*>(carve(base, &used, B.h_tile_col.size() * sizeof(int32_t)));
    auto* d_tile_max = reinterpret_cast<int32_t*>(carve(base, &used, B.h_tile_maxlen.size() * sizeof(int32_t)));
    uint32_t* d_aux = nullptr;
    if (words)
        d_aux = reinterpret_cast<uint32_t*>(
            carve(base, &used, (size_t)(kimura ? 2 : 1) * (size_t)brows * (size_t)bcols * sizeof(uint32_t)));
    // lpp16_kernel touches boundary columns 1 .. maxlen + 26, one int4 per
    // pair of its wave (kLppTile pairs)
    const int64_t stride = (int64_t)(c->maxlen + 32) * kLppTile;
    const size_t slot_bytes = (size_t)stride * sizeof(int4);
    const bool fits = d_desc && d_queue && d_task_row && d_task_tile && d_tile_col && d_tile_max && (!words || d_aux);
    const int64_t room = fits ? (int64_t)((SWG_NOALLOC_SCRATCH_BYTES - used) / slot_bytes) : 0;
    if (room < 1)
        return fail(SWG_E_CAPACITY, "block too large for the allocation-free entry point: its task lists" +
                                        std::string(words ? ", count words" : "") +
                                        " and one workgroup's boundary rows exceed the " +
                                        std::to_string(SWG_NOALLOC_SCRATCH_BYTES >> 20) + " MiB device scratch");
    int4* d_bnd = reinterpret_cast<int4*>(base + used);
    const int per_cu = lpp16_blocks_per_cu(s1, c->ncode);
    if (per_cu <= 0) return fail(SWG_E_DEVICE, "occupancy query failed");
    const int nslots = (int)std::max<int64_t>(1, std::min<int64_t>(room, (int64_t)per_cu * c->num_cu));

    BlockDesc desc = {};
    desc.ntasks = plan.ntasks;
    desc.mirror = g.diag_lower ? 1 : 0;
    desc.r0 = g.r0; desc.c0 = g.c0; desc.ld = ld;
    desc.out = dev_out; desc.stats = dev_stats;
    desc.rows = brows;
    if (words) { desc.aux = d_aux; desc.aux_ld = bcols; }

    const hipStream_t st = ln.stream;
    SWG_HIP(hipMemcpyAsync(d_desc, &desc, sizeof(desc), hipMemcpyHostToDevice, st));
    auto put = [&](int32_t* dst, const std::vector<int32_t>& h) -> hipError_t {
        return h.empty() ? hipSuccess
                         : hipMemcpyAsync(dst, h.data(), h.size() * sizeof(int32_t), hipMemcpyHostToDevice, st);
    };
    SWG_HIP(put(d_task_row, B.h_task_row));
    SWG_HIP(put(d_task_tile, B.h_task_tile));
    SWG_HIP(put(d_tile_col, B.h_tile_col));
    SWG_HIP(put(d_tile_max, B.h_tile_maxlen));

    SeqSet S2;
    const SeqSet S = make_seqset(c, ln, &S2);
    Scratch W;
    W.bnd = d_bnd;
    W.bnd_stride = stride;
    TaskSet T{};
    T.task_row = d_task_row; T.task_tile = d_task_tile; T.tile_col = d_tile_col; T.tile_maxlen = d_tile_max;
    T.desc = d_desc; T.queue = d_queue + kQueueFirst;
    TaskSet T2 = T;
    T2.queue = d_queue + kQueueSecond + kQueueFirst;
    const int pair_grid = std::max(1, c->num_cu * 4);
    const int tdist = c->p.tdistance == SWG_TDIST_DOUBLE ? 1 : 0;

    SWG_HIP(hipEventRecord(ln.ev0, st));
    SWG_HIP(launch_zero_diagonal(tdist, d_desc, pair_grid, st));
    if (plan.ntasks > 0) {
        // the work counters start at zero on every call (a fill kernel, no memset)
        SWG_HIP(launch_fill_i32(d_queue, 0, kQueueWords, st));
        SWG_HIP(launch_lpp16(S, T, W, s1, true, nslots, st));
        if (kimura) SWG_HIP(launch_lpp16(S2, T2, W, s1, true, nslots, st));
        if (words) SWG_HIP(launch_distance(S, d_desc, pair_grid, st));
    }
    SWG_HIP(hipEventRecord(ln.ev1, st));
    SWG_HIP(hipEventRecord(B.done, st));
    B.pending = true;
    c->last_strip = s1;
    c->last_variant = v;
    c->last_cfg = f;
    c->last_tuning = 0;
    c->last_trials = 0;
    c->last_cells = plan.cells;
    c->kt_n = 0;  // no per-kernel stamps on this path
    if (pairs_done) *pairs_done = plan.pairs;
    return SWG_OK;
}

}  // namespace

extern "C" {

int swg_compute_block_device_noalloc(swg_ctx* c, int32_t r0, int32_t r1, int32_t c0, int32_t c1, int32_t is_diag,
                                     void* dev_out, int64_t ld, swg_pair_stats* dev_stats, int64_t* pairs_done)
{
    if (!c || !dev_out) return fail(SWG_E_ARGUMENT, "null argument");
    SWG_HIP(hipSetDevice(c->device));
    return enqueue_block_noalloc(c, r0, r1, c0, c1, is_diag, dev_out, ld, reinterpret_cast<int32_t*>(dev_stats),
                                 pairs_done);
}

}  // extern "C"

namespace {

// A block whose output exceeds one chunk is computed in row chunks: chunk k
// runs on the compute stream into device buffer k%2 while the copy stream
// moves chunk k-1 to pinned memory and the host copies chunk k-2 into `out`.
// While the autotuner times configurations for the chunks' key, the chunks
// are trial chunks (trial_rows).
int compute_block_streamed(swg_ctx* c, int32_t r0, int32_t r1, int32_t c0, int32_t c1, void* out, int64_t out_ld,
                           int64_t* pairs_done)
{
    const int es = elem_size(c);
    const int64_t cols = c1 - c0 + 1;
    const int64_t row_bytes = cols * es;
    const int32_t chunk_rows = (int32_t)std::max<int64_t>(1, (int64_t)c->stream_chunk_bytes / row_bytes);
    const size_t chunk_bytes = (size_t)chunk_rows * row_bytes;
    if (!c->copy_stream) {
        SWG_HIP(hipStreamCreateWithFlags(&c->copy_stream, hipStreamNonBlocking));
        for (int b = 0; b < 2; ++b) {
            SWG_HIP(hipEventCreateWithFlags(&c->ev_comp[b], hipEventDisableTiming));
            SWG_HIP(hipEventCreateWithFlags(&c->ev_copy[b], hipEventDisableTiming));
            SWG_HIP(hipEventRecord(c->ev_copy[b], c->copy_stream));
        }
    }
    if (c->stream_host_bytes < chunk_bytes) {
        for (int b = 0; b < 2; ++b) {
            if (c->stream_host[b]) { SWG_HIP(hipHostFree(c->stream_host[b])); c->stream_host[b] = nullptr; }
            SWG_HIP(hipHostMalloc(reinterpret_cast<void**>(&c->stream_host[b]), chunk_bytes, 0));
        }
        c->stream_host_bytes = chunk_bytes;
    }
    for (int b = 0; b < 2; ++b) SWG_HIP(c->stream_dev[b].reserve(chunk_bytes));
    const int32_t trial = std::min(chunk_rows, trial_rows(cols));
    int64_t total_pairs = 0;
    double cells = 0.0;
    int32_t span[2][2] = {};  // rows of the chunk in buffer b
    auto consume = [&](int32_t k) -> int {
        const int b = k & 1;
        const int32_t a0 = span[b][0], a1 = span[b][1];
        SWG_HIP(hipEventSynchronize(c->ev_copy[b]));
        uint8_t* dst = static_cast<uint8_t*>(out) + (int64_t)(a0 - r0) * out_ld * es;
        for (int32_t i = 0; i < a1 - a0 + 1; ++i)
            std::memcpy(dst + (int64_t)i * out_ld * es, c->stream_host[b] + (int64_t)i * row_bytes, row_bytes);
        return SWG_OK;
    };
    int32_t k = 0;
    for (int32_t a0 = r0; a0 <= r1; a0 = span[k & 1][1] + 1, ++k) {
        const int b = k & 1;
        BlockGeom g;
        g.r0 = a0; g.r1 = std::min(r1, a0 + chunk_rows - 1); g.c0 = c0; g.c1 = c1; g.diag_lower = false;
        int32_t a1 = g.r1;
        if (tuning_pending(c, c->main, g)) {
            a1 = std::min(r1, a0 + trial - 1);
            if (r1 - a1 < trial) a1 = g.r1;  // no remainder below the trial size
        }
        // device buffer b is free once chunk k-2's copy finished
        SWG_HIP(hipStreamWaitEvent(c->main.stream, c->ev_copy[b], 0));
        // (chunk k-2's rows were consumed before this chunk was planned)
        span[b][0] = a0;
        span[b][1] = a1;
        int64_t p = 0;
        int rc = enqueue_block(c, c->main, a0, a1, c0, c1, 0, c->stream_dev[b].p, cols, nullptr, &p);
        if (rc) return rc;
        total_pairs += p;
        cells += c->last_cells;
        SWG_HIP(hipEventRecord(c->ev_comp[b], c->main.stream));
        SWG_HIP(hipStreamWaitEvent(c->copy_stream, c->ev_comp[b], 0));
        SWG_HIP(hipMemcpyAsync(c->stream_host[b], c->stream_dev[b].p, (size_t)(a1 - a0 + 1) * row_bytes,
                               hipMemcpyDeviceToHost, c->copy_stream));
        SWG_HIP(hipEventRecord(c->ev_copy[b], c->copy_stream));
        if (k >= 1 && (rc = consume(k - 1))) return rc;
    }
    int rc = consume(k - 1);
    if (rc) return rc;
    SWG_HIP(hipStreamSynchronize(c->main.stream));
    for (auto& B : c->main.bufs) B.pending = false;
    c->last_cells = cells;
    if (pairs_done) *pairs_done = total_pairs;
    return take_undefined(c);
}

}  // namespace

extern "C" {

int swg_compute_block(swg_ctx* c, int32_t r0, int32_t r1, int32_t c0, int32_t c1, int32_t is_diag,
                      void* out, int64_t out_ld, swg_pair_stats* stats, int64_t* pairs_done)
{
    if (!c || !out) return fail(SWG_E_ARGUMENT, "null argument");
    if (r1 < r0 || c1 < c0) { if (pairs_done) *pairs_done = 0; return SWG_OK; }
    SWG_HIP(hipSetDevice(c->device));
    const int64_t rows = r1 - r0 + 1, cols = c1 - c0 + 1;
    if (out_ld < cols) return fail(SWG_E_ARGUMENT, "out_ld smaller than the block width");
    const int es = elem_size(c);
    const bool mirror = is_diag && c->p.diag_mode == SWG_DIAG_LOWER_MIRROR;
    if (!mirror && !stats && c->res.p && r0 >= 0 && c0 >= 0 && r1 < c->nseq && c1 < c->nseq) {
        // in chunks: output larger than one chunk, or a block the autotuner
        // would time whole while trial chunks would do
        BlockGeom g;
        g.r0 = r0; g.r1 = r1; g.c0 = c0; g.c1 = c1; g.diag_lower = false;
        if ((size_t)(rows * cols * es) > c->stream_chunk_bytes ||
            (rows >= 2 * (int64_t)trial_rows(cols) && tuning_pending(c, c->main, g)))
            return compute_block_streamed(c, r0, r1, c0, c1, out, out_ld, pairs_done);
    }
    const size_t out_bytes = (size_t)(rows * cols * es + 3) & ~size_t(3);
    SWG_HIP(c->out_stage.reserve(out_bytes));
    if (stats) SWG_HIP(c->stats_stage.reserve((size_t)rows * cols * 4));
    SWG_HIP(launch_fill_bytes_zero(c->out_stage.p, (int64_t)out_bytes, c->main.stream));
    if (stats) SWG_HIP(launch_fill_i32(c->stats_stage.p, 0, rows * cols * 4, c->main.stream));
    int rc = enqueue_block(c, c->main, r0, r1, c0, c1, is_diag, c->out_stage.p, cols,
                           stats ? c->stats_stage.p : nullptr, pairs_done);
    if (rc) return rc;
    // one flat copy when the caller's rows are contiguous (a strided copy
    // moves the block row by row)
    if (out_ld == cols)
        SWG_HIP(hipMemcpyAsync(out, c->out_stage.p, (size_t)(rows * cols * es), hipMemcpyDeviceToHost,
                               c->main.stream));
    else
        SWG_HIP(hipMemcpy2DAsync(out, (size_t)out_ld * es, c->out_stage.p, (size_t)cols * es, (size_t)cols * es,
                                 (size_t)rows, hipMemcpyDeviceToHost, c->main.stream));
    if (stats)
        SWG_HIP(hipMemcpyAsync(stats, c->stats_stage.p, (size_t)rows * cols * 16, hipMemcpyDeviceToHost, c->main.stream));
    return swg_synchronize(c);
}

int swg_set_schedule(swg_ctx* c, int32_t schedule, int32_t strip_rows)
{
    if (!c) return fail(SWG_E_ARGUMENT, "null argument");
    if (schedule < SWG_SCHED_PERSISTENT || schedule > SWG_SCHED_PERSISTENT_LPS16)
        return fail(SWG_E_ARGUMENT, "bad schedule");
    if (strip_rows != 0 && strip_rows != 8 && strip_rows != 16 && strip_rows != 32)
        return fail(SWG_E_ARGUMENT, "strip_rows must be 0 (default), 8, 16 or 32");
    c->schedule = schedule;
    if (strip_rows != c->strip_rows) c->tuned.clear();  // the search depends on the fixed strips
    c->strip_rows = strip_rows;
    return SWG_OK;
}

int swg_set_option(swg_ctx* c, int32_t option, int64_t value)
{
    if (!c) return fail(SWG_E_ARGUMENT, "null argument");
    if (option == SWG_OPT_STREAM_CHUNK_BYTES) {
        if (value < 1) return fail(SWG_E_ARGUMENT, "chunk size must be positive");
        c->stream_chunk_bytes = (size_t)value;
        return SWG_OK;
    }
    if (option == SWG_OPT_KERNEL_TIMING) {
        if (value != 0 && value != 1) return fail(SWG_E_ARGUMENT, "kernel timing is 0 or 1");
        c->kernel_timing = value == 1;
        return SWG_OK;
    }
    if (option == SWG_OPT_SCORE_STRIP_ROWS) {
        if (value != 0 && value != 16 && value != 24 && value != 32 && value != 40)
            return fail(SWG_E_ARGUMENT, "score-pass strip rows are 0 (autotuned), 16, 24, 32 or 40");
        if (value != c->score_strip) c->tuned.clear();
        c->score_strip = (int32_t)value;
        return SWG_OK;
    }
    if (option == SWG_OPT_REVERSE_STRIP_ROWS) {
        if (value != 0 && value != 8 && value != 16)
            return fail(SWG_E_ARGUMENT, "reverse-pass strip rows are 0 (default 16), 8 or 16");
        c->reverse_strip = value == 0 ? 16 : (int32_t)value;
        return SWG_OK;
    }
    if (option == SWG_OPT_SCORE_MAX3) {
        if (value != 0 && value != 1) return fail(SWG_E_ARGUMENT, "the score pass's max3 form: 0 (off) or 1 (on)");
        if (value != c->score_max3) c->tuned.clear();  // its default strip and rates differ
        c->score_max3 = (int32_t)value;
        return SWG_OK;
    }
    if (option == SWG_OPT_LPS_LANES) {
        if (value != 0 && value != 16 && value != 32)
            return fail(SWG_E_ARGUMENT, "lanes per pair of the lanes-per-pair score pass: 0 (by block), 16 or 32");
        c->lps_lanes = (int32_t)value;
        return SWG_OK;
    }
    if (option == SWG_OPT_REVERSE_ROW_CAP) {
        if (value < 16 || value > 32768 || value % 16)
            return fail(SWG_E_ARGUMENT, "the reverse-pass row cap is a multiple of 16 in [16, 32768]");
        c->reverse_row_cap = (int32_t)value;
        return SWG_OK;
    }
    if (option == SWG_OPT_TUNE_MIN_CELLS) {
        if (value < 0) return fail(SWG_E_ARGUMENT, "the tuning threshold is a cell count >= 0");
        c->tune_min_cells = (double)value;
        return SWG_OK;
    }
    return fail(SWG_E_ARGUMENT, "unknown option");
}

int swg_last_block_config(swg_ctx* c, swg_block_config* out)
{
    if (!c || !out) return fail(SWG_E_ARGUMENT, "null argument");
    const VariantCfg& f = c->last_cfg;
    out->variant = f.v == kVarPlain ? SWG_SCHED_PLAIN
                 : f.v == kVarI32   ? SWG_SCHED_PERSISTENT_I32
                 : f.v == kVarLpp   ? SWG_SCHED_PERSISTENT_LPP16
                 : f.v == kVarLps   ? SWG_SCHED_PERSISTENT_LPS16
                                    : SWG_SCHED_PERSISTENT_I16;
    out->strip_rows = f.s1;
    out->rerun_strip_rows = two_pass(f.v) ? f.s2 : 0;
    out->pairs_per_wave = pairs_per_wave(f);
    out->tuning = c->last_tuning;
    out->trials = c->last_trials;
    return SWG_OK;
}

int swg_last_kernel_times(swg_ctx* c, swg_kernel_time* out, int32_t max, int32_t* n)
{
    if (!c || !n || (max > 0 && !out)) return fail(SWG_E_ARGUMENT, "null argument");
    *n = 0;
    SWG_HIP(hipSetDevice(c->device));
    SWG_HIP(hipStreamSynchronize(c->main.stream));
    uint64_t meta64[kMetaWords / 2] = {}, meta64b[kMetaWords / 2] = {};
    if (c->kt_meta) SWG_HIP(hipMemcpy(meta64, c->kt_meta, sizeof(meta64), hipMemcpyDeviceToHost));
    if (c->kt_meta2) SWG_HIP(hipMemcpy(meta64b, c->kt_meta2, sizeof(meta64b), hipMemcpyDeviceToHost));
    const int cnt = std::min<int>(c->kt_n, max);
    unsigned long long stamp[swg_ctx::kMaxKernels + 1] = {};
    if (cnt > 0) SWG_HIP(hipMemcpy(stamp, c->kt_stamp.p, (size_t)(c->kt_n + 1) * sizeof(stamp[0]), hipMemcpyDeviceToHost));
    if (cnt > 0 && c->wall_khz <= 0) return fail(SWG_E_DEVICE, "unknown wall-clock rate");
    for (int k = 0; k < cnt; ++k) {
        swg_kernel_time& t = out[k];
        t.kernel = c->kt_ids[k];
        t.reserved = 0;
        t.ms = (double)(int64_t)(stamp[k + 1] - stamp[k]) / (double)c->wall_khz;
        t.cells = 0.0;
        t.cells_run = 0.0;
        switch (t.kernel) {
        case SWG_KERNEL_ONE_PASS: case SWG_KERNEL_GLOBAL: case SWG_KERNEL_LPP16: case SWG_KERNEL_SCORE16:
        case SWG_KERNEL_LPS16:
            t.cells = c->kt_plan.cells;
            t.cells_run = c->kt_plan.cells_run;
            break;
        case SWG_KERNEL_REVERSE:
            t.cells = (double)meta64[kMetaRectCells];
            t.cells_run = (double)meta64[kMetaCaptureRun];
            break;
        case SWG_KERNEL_RERUN_TRACK:
            t.cells = (double)meta64b[kMetaFlaggedCells];
            t.cells_run = (double)meta64b[kMetaFlaggedRun];
            break;
        case SWG_KERNEL_RERUN_CAPTURE:
            t.cells = (double)meta64b[kMetaRectCells];
            t.cells_run = (double)meta64b[kMetaCaptureRun];
            break;
        case SWG_KERNEL_RECT64:
            t.cells = (double)meta64b[kMetaBigCells];
            t.cells_run = (double)meta64b[kMetaBigRun];
            break;
        default: break;
        }
        // Kimura2 runs each statistics kernel twice; their run counters hold both passes
        if (c->dist_kind == kDistKimura && (t.kernel == SWG_KERNEL_RERUN_TRACK ||
                                            t.kernel == SWG_KERNEL_RERUN_CAPTURE || t.kernel == SWG_KERNEL_RECT64))
            t.cells_run *= 0.5;
    }
    *n = cnt;
    return SWG_OK;
}

int swg_last_kernel_stats(swg_ctx* c, double* kernel_ms, double* cells, int32_t* strip_rows, int32_t* variant)
{
    if (!c) return fail(SWG_E_ARGUMENT, "null argument");
    if (kernel_ms) *kernel_ms = c->last_kernel_ms;
    if (cells) *cells = c->last_cells;
    if (strip_rows) *strip_rows = c->last_strip;
    if (variant)
        *variant = c->last_variant == kVarPlain ? SWG_SCHED_PLAIN
                 : c->last_variant == kVarI32   ? SWG_SCHED_PERSISTENT_I32
                 : c->last_variant == kVarLpp   ? SWG_SCHED_PERSISTENT_LPP16
                 : c->last_variant == kVarLps   ? SWG_SCHED_PERSISTENT_LPS16
                                                : SWG_SCHED_PERSISTENT_I16;
    return SWG_OK;
}

void swg_destroy(swg_ctx* c)
{
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->main.stream) (void)hipStreamSynchronize(c->main.stream);
    for (auto& B : c->main.bufs) {
        B.destroy_graphs();
        B.release();
    }
    c->res.release(); c->off.release(); c->len.release(); c->prof.release(); c->main.bnd.release();
    c->self_score.release();
    c->main.undef.release();
    c->out_stage.release(); c->stats_stage.release();
    if (c->copy_stream) (void)hipStreamSynchronize(c->copy_stream);
    for (int b = 0; b < 2; ++b) {
        c->stream_dev[b].release();
        if (c->stream_host[b]) (void)hipHostFree(c->stream_host[b]);
        if (c->ev_comp[b]) (void)hipEventDestroy(c->ev_comp[b]);
        if (c->ev_copy[b]) (void)hipEventDestroy(c->ev_copy[b]);
    }
    if (c->copy_stream) (void)hipStreamDestroy(c->copy_stream);
    if (c->main.ev0) (void)hipEventDestroy(c->main.ev0);
    if (c->main.ev1) (void)hipEventDestroy(c->main.ev1);
    c->kt_stamp.release();
    if (c->main.stream) (void)hipStreamDestroy(c->main.stream);
    delete c;
}

}  // extern "C"

// ===========================================================================
// Serving: continuous batching of query sequences against a device-resident
// panel.  A scheduler thread owns every HIP call of the server.  Each lane
// (stream + buffers + its own boundary scratch) runs one batch at a time:
// when a lane is free, the pending queries (FIFO) that fit the lane's batch
// size and the HBM budget are admitted, written into query slots of the
// resident sequence store, aligned against the whole panel as one block
// (rows = the batch's slots, columns = the panel) and copied back through
// pinned memory.  Lanes run concurrently, so a batch admitted while another
// is on the GPU fills the device as the earlier one drains.  Every pair is an
// independent alignment, so a query's distances do not depend on the batch
// it shared.
// ===========================================================================

#include <condition_variable>
#include <memory>
#include <thread>
#include <unordered_map>

namespace {

using Clock = std::chrono::steady_clock;

struct ServeRequest {
    int64_t ticket = 0;
    std::vector<uint8_t> codes;   // residue codes
    void* out = nullptr;          // caller's host buffer: npanel TDistance
    swg_pair_stats* stats = nullptr;
    Clock::time_point t_submit, t_admit, t_done;
    int state = 0;                // 0 pending, 1 running, 2 done, < 0 error code
    int64_t footprint = 0;
};

struct ServeLane {
    Lane lane;
    std::vector<std::shared_ptr<ServeRequest>> batch;
    DevBuf<uint8_t> dout;
    DevBuf<int32_t> dstats;
    uint8_t* h_out = nullptr;     // pinned staging
    int32_t* h_stats = nullptr;
    uint8_t* h_codes = nullptr;   // pinned staging for query residues + lengths
    int32_t* h_undef = nullptr;   // pinned copy of the lane's Jukes-Cantor flags
    hipEvent_t done = nullptr;
    bool active = false;
    int64_t footprint = 0;
    int32_t slot_base = 0;        // first query slot (sequence index) of this lane
};

}  // namespace

struct swg_server {
    swg_ctx* c = nullptr;
    int32_t npanel = 0, max_batch = 0, max_qlen = 0, nlanes = 0;
    int32_t sched = SWG_SCHED_PERSISTENT;  // batch kernel policy (swg_params.schedule)
    bool want_stats = true;
    int64_t budget = 0, used = 0;
    std::vector<std::unique_ptr<ServeLane>> lanes;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<std::shared_ptr<ServeRequest>> pending;
    std::unordered_map<int64_t, std::shared_ptr<ServeRequest>> tickets;
    int64_t next_ticket = 1;
    bool stop = false;
    int error = SWG_OK;
    std::string error_msg;
    std::thread worker;
    Clock::time_point t0;
    int64_t completed = 0;
    double cells = 0.0;
};

namespace {

int64_t query_footprint(const swg_server* s, int32_t qlen)
{
    // device bytes per panel pair: the distance and statistics, the two-pass
    // schedule's best cell (4), best score (2), start cell (4) and two sort
    // slots (2 x 4), the count words of the word-mode distances (Jukes-Cantor
    // one plane, Kimura2 two, 4 bytes each) and ~1 byte of task tiles; plus
    // the query's residues
    const int planes = s->c->dist_kind == kDistKimura ? 2 : s->c->dist_kind == kDistFraction ? 0 : 1;
    const int64_t per_pair = elem_size(s->c) + (s->want_stats ? 16 : 0) + 18 + 4 * planes + 1;
    return per_pair * s->npanel + qlen;
}

// Start a batch on lane `L` (worker thread, lock held by the caller only
// while taking requests).
int launch_batch(swg_server* s, ServeLane& L)
{
    swg_ctx* c = s->c;
    const int nb = (int)L.batch.size();
    const hipStream_t st = L.lane.stream;
    // query residues into the lane's slots, lengths into the length table
    // (each with the pad codes of its span: a slot keeps an earlier query's
    // codes past that)
    const int64_t stride = seq_span(s->max_qlen);
    int64_t pos = 0;
    std::vector<int32_t> lens(nb);
    for (int k = 0; k < nb; ++k) {
        const auto& q = L.batch[k];
        lens[k] = (int32_t)q->codes.size();
        std::memcpy(L.h_codes + pos, q->codes.data(), q->codes.size());
        std::memset(L.h_codes + pos + lens[k], c->ncode - 1, (size_t)(seq_span(lens[k]) - lens[k]));
        c->h_len[L.slot_base + k] = lens[k];
        pos += stride;
    }
    std::memcpy(L.h_codes + pos, lens.data(), nb * sizeof(int32_t));
    for (int k = 0; k < nb; ++k)
        SWG_HIP(hipMemcpyAsync(c->res.p + c->h_off[L.slot_base + k], L.h_codes + k * stride,
                               (size_t)seq_span(lens[k]), hipMemcpyHostToDevice, st));
    SWG_HIP(hipMemcpyAsync(c->len.p + L.slot_base, L.h_codes + pos, nb * sizeof(int32_t), hipMemcpyHostToDevice, st));
    if (c->dist_kind == kDistJukesCantor)
        SWG_HIP(launch_fill_i32(L.lane.undef.p, 0, 1 + nb, st));
    // batches use the int16 score pass with 16 lanes per pair where the
    // alphabet allows (its waves hold 8 pairs: a query against the panel
    // fills the GPU, and a deep queue's batches end evenly); else batches too
    // small to fill the GPU one lane per pair use the int32 16-lanes-per-pair
    // kernel and larger ones the int16 two-pass (or int32) variant -- no
    // timing trials on the serving path
    // a partial batch that starts while another lane's batch is on the
    // device gets a 1/lanes share of the resident slots, so that neither
    // lane's later kernels wait for the other's persistent grid to drain
    // (latency under moderate load); alone, or full (a deep queue: the
    // device saturated), the whole device (throughput)
    bool others = false;
    for (auto& Lp : s->lanes) others = others || (Lp.get() != &L && Lp->active);
    L.lane.share = others && nb < s->max_batch ? s->nlanes : 1;
    BlockGeom g;
    g.r0 = L.slot_base; g.r1 = L.slot_base + nb - 1; g.c0 = 0; g.c1 = s->npanel - 1; g.diag_lower = false;
    const bool i16 = c->int16_ok && s->sched != SWG_SCHED_PERSISTENT_I32;
    const int thin = c->lps_ok ? SWG_SCHED_PERSISTENT_LPS16 : SWG_SCHED_PERSISTENT_LPP16;
    c->schedule = c->nw                                        ? SWG_SCHED_PERSISTENT_I32
                : s->sched == SWG_SCHED_PERSISTENT_LPP16        ? SWG_SCHED_PERSISTENT_LPP16
                : s->sched == SWG_SCHED_PERSISTENT_LPS16        ? thin
                : s->sched == SWG_SCHED_PERSISTENT_I16 && i16   ? SWG_SCHED_PERSISTENT_I16
                : s->sched == SWG_SCHED_PERSISTENT && (c->lps_ok || thin_block(c, L.lane, g)) ? thin
                : i16                                           ? SWG_SCHED_PERSISTENT_I16
                                                                : SWG_SCHED_PERSISTENT_I32;
    int64_t pairs = 0;
    int rc = enqueue_block(c, L.lane, L.slot_base, L.slot_base + nb - 1, 0, s->npanel - 1, 0, L.dout.p, s->npanel,
                           s->want_stats ? L.dstats.p : nullptr, &pairs);
    if (rc) return rc;
    s->cells += c->last_cells;
    SWG_HIP(hipMemcpyAsync(L.h_out, L.dout.p, (size_t)nb * s->npanel * elem_size(c), hipMemcpyDeviceToHost, st));
    if (s->want_stats)
        SWG_HIP(hipMemcpyAsync(L.h_stats, L.dstats.p, (size_t)nb * s->npanel * 16, hipMemcpyDeviceToHost, st));
    if (c->dist_kind == kDistJukesCantor)
        SWG_HIP(hipMemcpyAsync(L.h_undef, L.lane.undef.p, (size_t)(1 + nb) * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    SWG_HIP(hipEventRecord(L.done, st));
    L.active = true;
    return SWG_OK;
}

void finish_batch(swg_server* s, ServeLane& L)
{
    const int es = elem_size(s->c);
    const auto now = Clock::now();
    std::lock_guard<std::mutex> lk(s->mu);
    for (size_t k = 0; k < L.batch.size(); ++k) {
        auto& q = L.batch[k];
        std::memcpy(q->out, L.h_out + k * (size_t)s->npanel * es, (size_t)s->npanel * es);
        if (q->stats && s->want_stats)
            std::memcpy(q->stats, L.h_stats + k * (size_t)s->npanel * 4, (size_t)s->npanel * 16);
        q->t_done = now;
        q->state = 2;
        if (s->c->dist_kind == kDistJukesCantor && L.h_undef[1 + k]) {
            // a pair of this query has no Jukes-Cantor distance: the query fails
            q->state = SWG_E_JC_UNDEFINED;
            s->error_msg = "Jukes and Cantor Distance Undefined - Log(Zero)";
        }
        ++s->completed;
    }
    s->used -= L.footprint;
    L.footprint = 0;
    L.batch.clear();
    L.active = false;
    for (auto& B : L.lane.bufs) B.pending = false;
}

void serve_loop(swg_server* s)
{
    (void)hipSetDevice(s->c->device);
    for (;;) {
        bool progressed = false;
        // completions
        for (auto& Lp : s->lanes) {
            ServeLane& L = *Lp;
            if (!L.active) continue;
            hipError_t q = hipEventQuery(L.done);
            if (q == hipSuccess) { finish_batch(s, L); progressed = true; }
            else if (q != hipErrorNotReady) {
                std::lock_guard<std::mutex> lk(s->mu);
                s->error = SWG_E_DEVICE;
                s->error_msg = std::string("batch failed: ") + hipGetErrorString(q);
                for (auto& r : L.batch) r->state = SWG_E_DEVICE;
                s->used -= L.footprint;  // the failed batch's HBM share is free again
                L.footprint = 0;
                L.batch.clear();
                L.active = false;
                s->cv.notify_all();
            }
        }
        // admissions: FIFO, one batch per free lane
        for (auto& Lp : s->lanes) {
            ServeLane& L = *Lp;
            if (L.active) continue;
            {
                std::lock_guard<std::mutex> lk(s->mu);
                while (!s->pending.empty() && (int)L.batch.size() < s->max_batch) {
                    auto& q = s->pending.front();
                    if (s->used + L.footprint + q->footprint > s->budget && !L.batch.empty()) break;
                    if (s->used + L.footprint + q->footprint > s->budget && s->used > 0) break;
                    L.footprint += q->footprint;
                    q->t_admit = Clock::now();
                    q->state = 1;
                    L.batch.push_back(q);
                    s->pending.pop_front();
                }
                s->used += L.footprint;
            }
            if (L.batch.empty()) continue;
            int rc = launch_batch(s, L);
            progressed = true;
            if (rc) {
                std::lock_guard<std::mutex> lk(s->mu);
                s->error = rc;
                s->error_msg = swg_last_error();
                for (auto& r : L.batch) r->state = rc;
                s->used -= L.footprint;
                L.footprint = 0;
                L.batch.clear();
            }
        }
        std::unique_lock<std::mutex> lk(s->mu);
        if (progressed) s->cv.notify_all();
        bool any_active = false;
        for (auto& Lp : s->lanes) any_active = any_active || Lp->active;
        if (s->stop && !any_active) break;
        if (!progressed) {
            if (any_active) s->cv.wait_for(lk, std::chrono::microseconds(100));
            else s->cv.wait_for(lk, std::chrono::milliseconds(5), [&] { return s->stop || !s->pending.empty(); });
        }
    }
}

double ms_since(const swg_server* s, Clock::time_point t)
{
    return std::chrono::duration<double, std::milli>(t - s->t0).count();
}

}  // namespace

extern "C" {

int swg_serve_create(swg_server** out, const swg_params* p, const uint8_t* panel_residues,
                     const int64_t* panel_offsets, int32_t npanel, int32_t max_query_len, int32_t max_batch,
                     int32_t lanes, int64_t hbm_budget_bytes, int32_t want_stats)
{
    if (!out || !p || npanel < 1 || max_query_len < 1 || max_batch < 1 || lanes < 1 || lanes > 8 ||
        hbm_budget_bytes <= 0)
        return fail(SWG_E_ARGUMENT, "bad server arguments");
    *out = nullptr;
    if (p->alignment_type == SWG_ALIGN_NUCLEIC && p->distance_fn == SWG_DIST_MINMAX_NORM)
        return fail(SWG_E_UNSUPPORTED, "the server does not implement MinMaxNormScore");
    if (p->schedule == SWG_SCHED_PLAIN || p->schedule < SWG_SCHED_PERSISTENT || p->schedule > SWG_SCHED_PERSISTENT_LPS16)
        return fail(SWG_E_ARGUMENT, "server schedule: SWG_SCHED_PERSISTENT, _PERSISTENT_I32, _I16, _LPP16 or _LPS16");
    swg_params q = *p;
    q.schedule = SWG_SCHED_PERSISTENT_I32;
    swg_ctx* c = nullptr;
    int rc = swg_create(&c, &q);
    if (rc) return rc;
    rc = load_sequences(c, panel_residues, panel_offsets, npanel, lanes * max_batch, max_query_len, true);
    if (rc) { swg_destroy(c); return rc; }
    auto* s = new swg_server();
    s->c = c;
    s->npanel = npanel; s->max_batch = max_batch; s->max_qlen = max_query_len; s->nlanes = lanes;
    s->sched = p->schedule;
    s->want_stats = want_stats != 0;
    s->budget = hbm_budget_bytes;
    s->t0 = Clock::now();
    auto bail = [&](const char* what, hipError_t e) {
        std::string m = std::string(what) + ": " + hipGetErrorString(e);
        swg_serve_destroy(s);
        return fail(SWG_E_DEVICE, m);
    };
    const int es = elem_size(c);
    const size_t nout = (size_t)max_batch * npanel;
    for (int i = 0; i < lanes; ++i) {
        auto L = std::make_unique<ServeLane>();
        hipError_t e;
        if ((e = hipStreamCreateWithFlags(&L->lane.stream, hipStreamNonBlocking)) != hipSuccess) return bail("stream", e);
        if ((e = hipEventCreate(&L->lane.ev0)) != hipSuccess) return bail("event", e);
        if ((e = hipEventCreate(&L->lane.ev1)) != hipSuccess) return bail("event", e);
        if ((e = hipEventCreateWithFlags(&L->done, hipEventDisableTiming)) != hipSuccess) return bail("event", e);
        for (auto& B : L->lane.bufs)
            if ((e = hipEventCreateWithFlags(&B.done, hipEventDisableTiming)) != hipSuccess) return bail("event", e);
        if ((e = L->dout.reserve(nout * es)) != hipSuccess) return bail("hipMalloc", e);
        if (s->want_stats && (e = L->dstats.reserve(nout * 4)) != hipSuccess) return bail("hipMalloc", e);
        if ((e = hipHostMalloc(reinterpret_cast<void**>(&L->h_out), nout * es, 0)) != hipSuccess) return bail("hipHostMalloc", e);
        if (s->want_stats &&
            (e = hipHostMalloc(reinterpret_cast<void**>(&L->h_stats), nout * 16, 0)) != hipSuccess)
            return bail("hipHostMalloc", e);
        const size_t codes_bytes = (size_t)max_batch * (size_t)seq_span(max_query_len) + max_batch * 4;
        if ((e = hipHostMalloc(reinterpret_cast<void**>(&L->h_codes), codes_bytes, 0)) != hipSuccess)
            return bail("hipHostMalloc", e);
        if ((e = hipHostMalloc(reinterpret_cast<void**>(&L->h_undef), (size_t)(1 + max_batch) * 4, 0)) != hipSuccess)
            return bail("hipHostMalloc", e);
        if ((e = L->lane.undef.reserve(1 + max_batch)) != hipSuccess) return bail("hipMalloc", e);
        if ((e = hipMemset(L->lane.undef.p, 0, (size_t)(1 + max_batch) * 4)) != hipSuccess) return bail("hipMemset", e);
        L->lane.undef_rows = true;
        L->slot_base = npanel + i * max_batch;
        s->lanes.push_back(std::move(L));
    }
    s->worker = std::thread(serve_loop, s);
    *out = s;
    return SWG_OK;
}

int swg_serve_submit(swg_server* s, const uint8_t* query, int32_t len, void* out, swg_pair_stats* stats,
                     int64_t* ticket)
{
    if (!s || (!query && len > 0) || !out || !ticket || len < 0) return fail(SWG_E_ARGUMENT, "null argument");
    if (len > s->max_qlen) return fail(SWG_E_CAPACITY, "query longer than the server's max_query_len");
    auto r = std::make_shared<ServeRequest>();
    r->codes.resize(len);
    for (int32_t k = 0; k < len; ++k) {
        const int16_t code = s->c->code_of[query[k]];
        if (code < 0) return fail(SWG_E_SYMBOL, "query symbol outside the panel/matrix alphabet");
        r->codes[k] = (uint8_t)code;
    }
    r->out = out;
    r->stats = stats;
    r->footprint = query_footprint(s, len);
    if (r->footprint > s->budget) return fail(SWG_E_CAPACITY, "one query exceeds the HBM budget");
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) return fail(SWG_E_ARGUMENT, "server is stopping");
    r->ticket = s->next_ticket++;
    r->t_submit = Clock::now();
    s->tickets[r->ticket] = r;
    s->pending.push_back(r);
    *ticket = r->ticket;
    s->cv.notify_all();
    return SWG_OK;
}

int swg_serve_wait(swg_server* s, int64_t ticket, int32_t timeout_ms, int32_t* done)
{
    if (!s || !done) return fail(SWG_E_ARGUMENT, "null argument");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->tickets.find(ticket);
    if (it == s->tickets.end()) return fail(SWG_E_ARGUMENT, "unknown ticket");
    auto r = it->second;
    auto ready = [&] { return r->state == 2 || r->state < 0; };
    if (timeout_ms < 0) s->cv.wait(lk, ready);
    else s->cv.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready);
    *done = r->state == 2 ? 1 : 0;
    if (r->state < 0) return fail(r->state, "query failed: " + s->error_msg);
    return SWG_OK;
}

int swg_serve_times(swg_server* s, int64_t ticket, double* submit_ms, double* admit_ms, double* done_ms)
{
    if (!s) return fail(SWG_E_ARGUMENT, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->tickets.find(ticket);
    if (it == s->tickets.end()) return fail(SWG_E_ARGUMENT, "unknown ticket");
    const auto& r = *it->second;
    if (submit_ms) *submit_ms = ms_since(s, r.t_submit);
    if (admit_ms) *admit_ms = r.state >= 1 ? ms_since(s, r.t_admit) : -1.0;
    if (done_ms) *done_ms = r.state == 2 ? ms_since(s, r.t_done) : -1.0;
    return SWG_OK;
}

int swg_serve_release(swg_server* s, int64_t ticket)
{
    if (!s) return fail(SWG_E_ARGUMENT, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->tickets.find(ticket);
    if (it == s->tickets.end()) return fail(SWG_E_ARGUMENT, "unknown ticket");
    if (it->second->state == 0 || it->second->state == 1) return fail(SWG_E_ARGUMENT, "query still in flight");
    s->tickets.erase(it);
    return SWG_OK;
}

int swg_serve_stats(swg_server* s, int64_t* completed, double* cells, double* now_ms)
{
    if (!s) return fail(SWG_E_ARGUMENT, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    if (completed) *completed = s->completed;
    if (cells) *cells = s->cells;
    if (now_ms) *now_ms = ms_since(s, Clock::now());
    return SWG_OK;
}

void swg_serve_destroy(swg_server* s)
{
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        s->cv.notify_all();
    }
    if (s->worker.joinable()) s->worker.join();
    (void)hipSetDevice(s->c ? s->c->device : 0);
    for (auto& Lp : s->lanes) {
        ServeLane& L = *Lp;
        if (L.lane.stream) (void)hipStreamSynchronize(L.lane.stream);
        for (auto& B : L.lane.bufs) {
            B.destroy_graphs();
            B.release();
        }
        L.lane.bnd.release();
        L.dout.release(); L.dstats.release();
        if (L.h_out) (void)hipHostFree(L.h_out);
        if (L.h_stats) (void)hipHostFree(L.h_stats);
        if (L.h_codes) (void)hipHostFree(L.h_codes);
        if (L.h_undef) (void)hipHostFree(L.h_undef);
        L.lane.undef.release();
        if (L.done) (void)hipEventDestroy(L.done);
        if (L.lane.ev0) (void)hipEventDestroy(L.lane.ev0);
        if (L.lane.ev1) (void)hipEventDestroy(L.lane.ev1);
        if (L.lane.stream) (void)hipStreamDestroy(L.lane.stream);
    }
    if (s->c) swg_destroy(s->c);
    delete s;
}

}  // extern "C"
