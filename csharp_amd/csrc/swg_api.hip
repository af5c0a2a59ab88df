// Host side of libswg.so: the C ABI declared in include/swg.h, the
// block -> task scheduler (length buckets, work queue, hipGraph per block,
// strip-height autotuner) and device-memory management.
#include "../../include/swg.h"
#include "swg_internal.h"

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
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

// One set of per-block device buffers; two of them alternate so that the
// host can plan block k+1 while block k runs.
struct BlockBufs {
    DevBuf<int32_t> task_row, task_tile, tile_col, tile_maxlen, queue;
    DevBuf<BlockDesc> desc;
    std::vector<int32_t> h_task_row, h_task_tile, h_tile_col, h_tile_maxlen;
    hipEvent_t done = nullptr;
    bool pending = false;
    // captured graph (persistent schedule) for this buffer set
    hipGraphExec_t graph = nullptr;
    unsigned char graph_key[sizeof(SeqSet) + sizeof(TaskSet) + sizeof(Scratch) + 2 * sizeof(int)] = {0};
};

}  // namespace

struct swg_ctx {
    swg_params p{};
    int device = 0;
    int num_cu = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev_k0 = nullptr, ev_k1 = nullptr;

    int32_t iscores[SWG_MATRIX_SIZE * SWG_MATRIX_SIZE];
    int32_t gap_open = 0, gap_ext = 0;
    bool count_similarity = false;  // protein: PercentSimilarity

    // sequences
    int32_t nseq = 0;
    int32_t maxlen = 0;
    std::vector<int32_t> h_len;
    int16_t code_of[256];
    int32_t ncode = 0;  // incl. pad
    DevBuf<uint8_t> res;
    DevBuf<int64_t> off;
    DevBuf<int32_t> len;
    DevBuf<int2> prof;
    int64_t pad_off = 0;

    // scratch
    DevBuf<int4> bnd;
    int64_t bnd_stride = 0;
    int nslots = 0;
    int slots_strip = 0;

    // outputs staged on device for the host-memory entry point
    DevBuf<uint8_t> out_stage;
    DevBuf<int32_t> stats_stage;

    BlockBufs bufs[2];
    int cur = 0;

    int32_t schedule = SWG_SCHED_PERSISTENT;
    int32_t strip_rows = 0;  // 0 = autotune
    int32_t tuned_strip = 0;
    double last_kernel_ms = 0.0, last_cells = 0.0;
    int32_t last_strip = 0;
};

namespace {

bool is_integral(float v) { return std::isfinite(v) && std::floor(v) == v && std::fabs(v) < 32768.0f; }

int elem_size(const swg_ctx* c) { return c->p.tdistance == SWG_TDIST_DOUBLE ? 8 : 2; }

// Number of resident workgroups for the kernel variant and the boundary
// buffer they need (one slot per resident wave; HBM budget-capped).
int ensure_slots(swg_ctx* c, int strip)
{
    if (c->nslots > 0 && c->slots_strip == strip && c->bnd.p) return SWG_OK;
    int per_cu = sw_blocks_per_cu(strip, false);
    if (per_cu <= 0) return fail(SWG_E_DEVICE, "occupancy query failed");
    int slots = per_cu * c->num_cu;
    const int64_t stride = (int64_t)(c->maxlen + 16) * kWave;
    const double budget = 24.0 * (1u << 30);
    while (slots > c->num_cu && (double)slots * (double)stride * sizeof(int4) > budget) slots -= c->num_cu;
    SWG_HIP(c->bnd.reserve((size_t)slots * (size_t)stride));
    c->bnd_stride = stride;
    c->nslots = slots;
    c->slots_strip = strip;
    return SWG_OK;
}

// ---- task planning ------------------------------------------------------

struct Plan {
    int64_t pairs = 0;
    double cells = 0.0;
};

// Tiles of 64 column sequences and tasks (row sequence, tile).
//   persistent: columns sorted by length (longest first) so that the lanes of
//     a wave do similar work; tasks ordered by estimated cost, longest first,
//     drained from a device queue.
//   plain: columns in natural order, tasks row-major, static assignment.
void plan_block(swg_ctx* c, BlockBufs& B, int32_t r0, int32_t r1, int32_t c0, int32_t c1,
                bool diag_lower, bool sorted, Plan& plan)
{
    const std::vector<int32_t>& L = c->h_len;
    auto& trow = B.h_task_row; auto& ttile = B.h_task_tile;
    auto& tcol = B.h_tile_col; auto& tmax = B.h_tile_maxlen;
    trow.clear(); ttile.clear(); tcol.clear(); tmax.clear();
    plan = Plan{};

    std::vector<int32_t> cols((size_t)(c1 - c0 + 1));
    std::iota(cols.begin(), cols.end(), c0);
    if (sorted)
        std::stable_sort(cols.begin(), cols.end(), [&](int32_t x, int32_t y) { return L[x] > L[y]; });

    auto emit_tile = [&](const int32_t* ids, int n) -> int32_t {
        const int32_t t = (int32_t)tmax.size();
        int32_t mx = 0;
        for (int k = 0; k < kWave; ++k) {
            const int32_t id = k < n ? ids[k] : -1;
            tcol.push_back(id);
            if (id >= 0) mx = std::max(mx, L[id]);
        }
        tmax.push_back(mx);
        return t;
    };

    std::vector<int32_t> rows((size_t)(r1 - r0 + 1));
    std::iota(rows.begin(), rows.end(), r0);
    if (sorted)
        std::stable_sort(rows.begin(), rows.end(), [&](int32_t x, int32_t y) { return L[x] > L[y]; });

    if (!diag_lower) {
        std::vector<int32_t> tiles;
        double sum_cols = 0.0;
        for (size_t k = 0; k < cols.size(); k += kWave) {
            const int n = (int)std::min<size_t>(kWave, cols.size() - k);
            tiles.push_back(emit_tile(&cols[k], n));
        }
        for (int32_t x : cols) sum_cols += L[x];
        for (int32_t a : rows) {
            for (int32_t t : tiles) { trow.push_back(a); ttile.push_back(t); }
            plan.cells += (double)L[a] * sum_cols;
        }
        plan.pairs = (int64_t)rows.size() * (int64_t)cols.size();
    } else {
        std::vector<int32_t> valid;
        valid.reserve(cols.size());
        for (int32_t a : rows) {
            valid.clear();
            for (int32_t x : cols) if (x < a) valid.push_back(x);
            for (size_t k = 0; k < valid.size(); k += kWave) {
                const int n = (int)std::min<size_t>(kWave, valid.size() - k);
                trow.push_back(a);
                ttile.push_back(emit_tile(&valid[k], n));
            }
            for (int32_t x : valid) plan.cells += (double)L[a] * L[x];
            plan.pairs += (int64_t)valid.size();
        }
    }

    if (sorted) {
        // longest-processing-time-first order for the work queue
        const size_t nt = trow.size();
        std::vector<uint32_t> order(nt);
        std::iota(order.begin(), order.end(), 0u);
        std::vector<int64_t> cost(nt);
        for (size_t k = 0; k < nt; ++k) cost[k] = (int64_t)L[trow[k]] * (int64_t)tmax[ttile[k]];
        std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) { return cost[x] > cost[y]; });
        std::vector<int32_t> r2(nt), t2(nt);
        for (size_t k = 0; k < nt; ++k) { r2[k] = trow[order[k]]; t2[k] = ttile[order[k]]; }
        trow.swap(r2); ttile.swap(t2);
    }
}

int upload_plan(swg_ctx* c, BlockBufs& B, const BlockDesc& desc)
{
    const size_t nt = B.h_task_row.size();
    const size_t ntile = B.h_tile_maxlen.size();
    SWG_HIP(B.task_row.reserve(std::max<size_t>(nt, 1)));
    SWG_HIP(B.task_tile.reserve(std::max<size_t>(nt, 1)));
    SWG_HIP(B.tile_col.reserve(std::max<size_t>(ntile * kWave, 1)));
    SWG_HIP(B.tile_maxlen.reserve(std::max<size_t>(ntile, 1)));
    SWG_HIP(B.queue.reserve(64));
    SWG_HIP(B.desc.reserve(1));
    if (nt) {
        SWG_HIP(hipMemcpyAsync(B.task_row.p, B.h_task_row.data(), nt * 4, hipMemcpyHostToDevice, c->stream));
        SWG_HIP(hipMemcpyAsync(B.task_tile.p, B.h_task_tile.data(), nt * 4, hipMemcpyHostToDevice, c->stream));
    }
    if (ntile) {
        SWG_HIP(hipMemcpyAsync(B.tile_col.p, B.h_tile_col.data(), ntile * kWave * 4, hipMemcpyHostToDevice, c->stream));
        SWG_HIP(hipMemcpyAsync(B.tile_maxlen.p, B.h_tile_maxlen.data(), ntile * 4, hipMemcpyHostToDevice, c->stream));
    }
    SWG_HIP(hipMemcpyAsync(B.desc.p, &desc, sizeof(BlockDesc), hipMemcpyHostToDevice, c->stream));
    // the host vectors must stay alive until the copies complete: callers
    // synchronise on B.done before replanning this buffer set
    return SWG_OK;
}

SeqSet make_seqset(const swg_ctx* c)
{
    SeqSet S;
    S.res = c->res.p; S.off = c->off.p; S.len = c->len.p; S.pad_off = c->pad_off;
    S.prof = c->prof.p; S.ncode = c->ncode;
    S.gap_open = c->gap_open; S.gap_ext = c->gap_ext;
    S.tdist = c->p.tdistance == SWG_TDIST_DOUBLE ? 1 : 0;
    return S;
}

TaskSet make_taskset(BlockBufs& B)
{
    TaskSet T;
    T.task_row = B.task_row.p; T.task_tile = B.task_tile.p;
    T.tile_col = B.tile_col.p; T.tile_maxlen = B.tile_maxlen.p;
    T.desc = B.desc.p; T.queue = B.queue.p;
    return T;
}

int pick_strip(swg_ctx* c)
{
    if (c->schedule == SWG_SCHED_PLAIN) return 16;
    if (c->strip_rows) return c->strip_rows;
    if (c->tuned_strip) return c->tuned_strip;
    return 16;
}

// Enqueue one block on the context stream into device buffers.
int enqueue_block(swg_ctx* c, int32_t r0, int32_t r1, int32_t c0, int32_t c1, int32_t is_diag,
                  void* dev_out, int64_t ld, int32_t* dev_stats, int64_t* pairs_done)
{
    if (!c->res.p) return fail(SWG_E_ARGUMENT, "swg_set_sequences has not been called");
    if (r0 < 0 || c0 < 0 || r1 < r0 - 1 || c1 < c0 - 1 || r1 >= c->nseq || c1 >= c->nseq)
        return fail(SWG_E_ARGUMENT, "block range outside the sequence set");
    const bool diag_lower = is_diag && c->p.diag_mode == SWG_DIAG_LOWER_MIRROR;
    if (diag_lower && (r0 != c0 || r1 != c1))
        return fail(SWG_E_ARGUMENT, "a diagonal block must have equal row and column ranges");
    if (ld < (int64_t)(c1 - c0 + 1)) return fail(SWG_E_ARGUMENT, "out_ld smaller than the block width");

    const bool persistent = c->schedule == SWG_SCHED_PERSISTENT;
    const int strip = pick_strip(c);
    int rc = ensure_slots(c, strip);
    if (rc) return rc;

    BlockBufs& B = c->bufs[c->cur];
    c->cur ^= 1;
    if (B.pending) { SWG_HIP(hipEventSynchronize(B.done)); B.pending = false; }

    Plan plan;
    plan_block(c, B, r0, r1, c0, c1, diag_lower, persistent, plan);
    BlockDesc desc{};
    desc.ntasks = (int32_t)B.h_task_row.size();
    desc.mirror = diag_lower ? 1 : 0;
    desc.r0 = r0; desc.c0 = c0; desc.ld = ld;
    desc.out = dev_out; desc.stats = dev_stats;
    rc = upload_plan(c, B, desc);
    if (rc) return rc;

    const SeqSet S = make_seqset(c);
    const TaskSet T = make_taskset(B);
    Scratch W; W.bnd = c->bnd.p; W.bnd_stride = c->bnd_stride;

    SWG_HIP(hipEventRecord(c->ev_k0, c->stream));
    if (desc.ntasks > 0) {
        if (c->p.mode == SWG_MODE_NW_GLOBAL) {
            SWG_HIP(launch_nw(S, T, W, strip, persistent, c->nslots, c->stream));
        } else if (persistent) {
            // the graph bakes every pointer and scalar of the launch: re-capture
            // whenever one of them changed (new sequences, grown buffers, other strip)
            unsigned char key[sizeof(B.graph_key)];
            std::memset(key, 0, sizeof(key));
            unsigned char* kp = key;
            std::memcpy(kp, &S, sizeof(S)); kp += sizeof(S);
            std::memcpy(kp, &T, sizeof(T)); kp += sizeof(T);
            std::memcpy(kp, &W, sizeof(W)); kp += sizeof(W);
            std::memcpy(kp, &strip, sizeof(int)); kp += sizeof(int);
            std::memcpy(kp, &c->nslots, sizeof(int));
            if (!B.graph || std::memcmp(key, B.graph_key, sizeof(key)) != 0) {
                if (B.graph) { (void)hipGraphExecDestroy(B.graph); B.graph = nullptr; }
                hipGraph_t g = nullptr;
                SWG_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
                hipError_t e1 = launch_fill_i32(B.queue.p, 0, 1, c->stream);
                hipError_t e2 = launch_sw(S, T, W, strip, true, c->nslots, c->stream);
                hipError_t e3 = hipStreamEndCapture(c->stream, &g);
                if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess) {
                    if (g) (void)hipGraphDestroy(g);
                    return fail(SWG_E_DEVICE, "graph capture of the block kernel failed");
                }
                hipError_t e4 = hipGraphInstantiate(&B.graph, g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                if (e4 != hipSuccess) { B.graph = nullptr; return fail(SWG_E_DEVICE, "hipGraphInstantiate failed"); }
                std::memcpy(B.graph_key, key, sizeof(key));
            }
            SWG_HIP(hipGraphLaunch(B.graph, c->stream));
        } else {
            SWG_HIP(launch_sw(S, T, W, strip, false, c->nslots, c->stream));
        }
    }
    SWG_HIP(hipEventRecord(c->ev_k1, c->stream));
    SWG_HIP(hipEventRecord(B.done, c->stream));
    B.pending = true;
    c->last_cells = plan.cells;
    c->last_strip = strip;
    if (pairs_done) *pairs_done = plan.pairs;
    return SWG_OK;
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
    if (p->schedule != SWG_SCHED_PERSISTENT && p->schedule != SWG_SCHED_PLAIN) return fail(SWG_E_ARGUMENT, "bad schedule");
    if (p->alignment_type == SWG_ALIGN_NUCLEIC && p->distance_fn != SWG_DIST_PERCENT_IDENTITY)
        return fail(SWG_E_UNSUPPORTED, "only PercentIdentity is implemented for nucleic alignments");
    if (p->mode == SWG_MODE_NW_GLOBAL)
        return fail(SWG_E_UNSUPPORTED, "the global (NW) kernel is not available in this build");
    if (!is_integral(p->gap_open) || !is_integral(p->gap_ext))
        return fail(SWG_E_NONINTEGRAL, "gap penalties must be integers below 32768 for the integer kernels");
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
    c->schedule = p->schedule;
    c->device = p->device;

    auto bail = [&](hipError_t e, const char* what) {
        std::string m = std::string(what) + ": " + hipGetErrorString(e);
        swg_destroy(c);
        return fail(SWG_E_DEVICE, m);
    };
    hipError_t e = hipSetDevice(c->device);
    if (e != hipSuccess) return bail(e, "hipSetDevice");
    hipDeviceProp_t prop;
    e = hipGetDeviceProperties(&prop, c->device);
    if (e != hipSuccess) return bail(e, "hipGetDeviceProperties");
    c->num_cu = prop.multiProcessorCount;
    e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) return bail(e, "hipStreamCreate");
    if ((e = hipEventCreate(&c->ev_k0)) != hipSuccess) return bail(e, "hipEventCreate");
    if ((e = hipEventCreate(&c->ev_k1)) != hipSuccess) return bail(e, "hipEventCreate");
    for (auto& B : c->bufs)
        if ((e = hipEventCreateWithFlags(&B.done, hipEventDisableTiming)) != hipSuccess) return bail(e, "hipEventCreate");
    *out = c;
    return SWG_OK;
}

int swg_set_sequences(swg_ctx* c, const uint8_t* residues, const int64_t* offsets, int32_t n)
{
    if (!c || n < 0 || (n > 0 && (!residues || !offsets))) return fail(SWG_E_ARGUMENT, "null argument");
    SWG_HIP(hipSetDevice(c->device));
    SWG_HIP(hipStreamSynchronize(c->stream));
    for (auto& B : c->bufs) B.pending = false;
    // residue codes: distinct symbols of the set, in ASCII order, + pad
    bool seen[256] = {false};
    int32_t maxlen = 0;
    for (int32_t i = 0; i < n; ++i) {
        const int64_t L = offsets[i + 1] - offsets[i];
        if (L < 0) return fail(SWG_E_ARGUMENT, "offsets must be non-decreasing");
        if (L > SWG_MAX_SEQ_LEN) return fail(SWG_E_CAPACITY, "sequence longer than 32767 residues");
        maxlen = std::max<int32_t>(maxlen, (int32_t)L);
        for (int64_t k = offsets[i]; k < offsets[i + 1]; ++k) seen[residues[k]] = true;
    }
    int32_t ncode = 0;
    for (int s = 0; s < 256; ++s) {
        c->code_of[s] = -1;
        if (seen[s]) {
            if (s >= SWG_MATRIX_SIZE) return fail(SWG_E_SYMBOL, "residue symbol >= 127 (the reference indexes a 127x127 matrix)");
            c->code_of[s] = (int16_t)ncode++;
        }
    }
    if (ncode + 1 > kMaxCodes) return fail(SWG_E_CAPACITY, "more than 63 distinct residue symbols");
    const int32_t pad = ncode;
    c->ncode = ncode + 1;
    std::vector<uint8_t> sym_of(c->ncode, 0);
    for (int s = 0; s < 256; ++s) if (c->code_of[s] >= 0) sym_of[c->code_of[s]] = (uint8_t)s;

    // profile table (score + gap_open, increment of the packed count word)
    std::vector<int2> prof((size_t)c->ncode * c->ncode);
    for (int x = 0; x < c->ncode; ++x)
        for (int y = 0; y < c->ncode; ++y) {
            int2 v;
            if (x == pad || y == pad) { v.x = kPadScore; v.y = 0; }
            else {
                const int sx = sym_of[x], sy = sym_of[y];
                const int sc = c->iscores[sx * SWG_MATRIX_SIZE + sy];
                int cnt = (sx == sy) ? 1 : 0;  // identity: c1 == c2 (SmithWatermanGotoh.cs:284)
                if (c->count_similarity && sx != sy && sc > 0) cnt = 1;  // similarity: scores[c1,c2] > 0 (:290)
                v.x = sc + c->gap_open;
                v.y = 0x10000 + cnt;
            }
            prof[(size_t)x * c->ncode + y] = v;
        }

    // residue codes, each sequence 4-byte aligned, tail of pad codes
    std::vector<int64_t> off(std::max(n, 1));
    int64_t total = 0;
    for (int32_t i = 0; i < n; ++i) {
        off[i] = total;
        total += ((offsets[i + 1] - offsets[i]) + 3) & ~int64_t(3);
    }
    const int64_t pad_off = total;
    total += maxlen + 64;
    std::vector<uint8_t> codes((size_t)total, (uint8_t)pad);
    std::vector<int32_t> lens(std::max(n, 1), 0);
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
    c->pad_off = pad_off;
    c->nseq = n;
    c->h_len.assign(lens.begin(), lens.begin() + n);
    if (maxlen != c->maxlen) { c->bnd.release(); c->nslots = 0; }
    c->maxlen = maxlen;
    return SWG_OK;
}

int swg_compute_block_device(swg_ctx* c, int32_t r0, int32_t r1, int32_t c0, int32_t c1, int32_t is_diag,
                             void* dev_out, int64_t ld, swg_pair_stats* dev_stats, int64_t* pairs_done)
{
    if (!c || !dev_out) return fail(SWG_E_ARGUMENT, "null argument");
    SWG_HIP(hipSetDevice(c->device));
    return enqueue_block(c, r0, r1, c0, c1, is_diag, dev_out, ld, reinterpret_cast<int32_t*>(dev_stats), pairs_done);
}

int swg_synchronize(swg_ctx* c)
{
    if (!c) return fail(SWG_E_ARGUMENT, "null argument");
    SWG_HIP(hipStreamSynchronize(c->stream));
    for (auto& B : c->bufs) B.pending = false;
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, c->ev_k0, c->ev_k1) == hipSuccess) c->last_kernel_ms = ms;
    return SWG_OK;
}

int swg_compute_block(swg_ctx* c, int32_t r0, int32_t r1, int32_t c0, int32_t c1, int32_t is_diag,
                      void* out, int64_t out_ld, swg_pair_stats* stats, int64_t* pairs_done)
{
    if (!c || !out) return fail(SWG_E_ARGUMENT, "null argument");
    if (r1 < r0 || c1 < c0) { if (pairs_done) *pairs_done = 0; return SWG_OK; }
    SWG_HIP(hipSetDevice(c->device));
    const int64_t rows = r1 - r0 + 1, cols = c1 - c0 + 1;
    if (out_ld < cols) return fail(SWG_E_ARGUMENT, "out_ld smaller than the block width");
    const int es = elem_size(c);
    const size_t out_bytes = (size_t)(rows * cols * es + 3) & ~size_t(3);
    SWG_HIP(c->out_stage.reserve(out_bytes));
    if (stats) SWG_HIP(c->stats_stage.reserve((size_t)rows * cols * 4));
    SWG_HIP(launch_fill_bytes_zero(c->out_stage.p, (int64_t)out_bytes, c->stream));
    if (stats) SWG_HIP(launch_fill_i32(c->stats_stage.p, 0, rows * cols * 4, c->stream));
    int rc = enqueue_block(c, r0, r1, c0, c1, is_diag, c->out_stage.p, cols, stats ? c->stats_stage.p : nullptr,
                           pairs_done);
    if (rc) return rc;
    SWG_HIP(hipMemcpy2DAsync(out, (size_t)out_ld * es, c->out_stage.p, (size_t)cols * es, (size_t)cols * es,
                             (size_t)rows, hipMemcpyDeviceToHost, c->stream));
    if (stats)
        SWG_HIP(hipMemcpyAsync(stats, c->stats_stage.p, (size_t)rows * cols * 16, hipMemcpyDeviceToHost, c->stream));
    return swg_synchronize(c);
}

int swg_set_schedule(swg_ctx* c, int32_t schedule, int32_t strip_rows)
{
    if (!c) return fail(SWG_E_ARGUMENT, "null argument");
    if (schedule != SWG_SCHED_PERSISTENT && schedule != SWG_SCHED_PLAIN) return fail(SWG_E_ARGUMENT, "bad schedule");
    if (strip_rows != 0 && strip_rows != 8 && strip_rows != 16 && strip_rows != 32)
        return fail(SWG_E_ARGUMENT, "strip_rows must be 0 (autotune), 8, 16 or 32");
    c->schedule = schedule;
    c->strip_rows = strip_rows;
    return SWG_OK;
}

int swg_last_kernel_stats(swg_ctx* c, double* kernel_ms, double* cells, int32_t* strip_rows)
{
    if (!c) return fail(SWG_E_ARGUMENT, "null argument");
    if (kernel_ms) *kernel_ms = c->last_kernel_ms;
    if (cells) *cells = c->last_cells;
    if (strip_rows) *strip_rows = c->last_strip;
    return SWG_OK;
}

void swg_destroy(swg_ctx* c)
{
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto& B : c->bufs) {
        if (B.graph) (void)hipGraphExecDestroy(B.graph);
        B.task_row.release(); B.task_tile.release(); B.tile_col.release(); B.tile_maxlen.release();
        B.queue.release(); B.desc.release();
        if (B.done) (void)hipEventDestroy(B.done);
    }
    c->res.release(); c->off.release(); c->len.release(); c->prof.release(); c->bnd.release();
    c->out_stage.release(); c->stats_stage.release();
    if (c->ev_k0) (void)hipEventDestroy(c->ev_k0);
    if (c->ev_k1) (void)hipEventDestroy(c->ev_k1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

}  // extern "C"
