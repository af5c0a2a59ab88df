// HIP kernels for gfx950 (CDNA4): all-pairs Smith-Waterman-Gotoh with the
// traceback statistics carried through the DP.
//
// Work unit ("task"): one wavefront aligns one sequence s1 (uniform across
// the wave, the rows of the DP) against 64 sequences s2 (one per lane, the
// columns).  Because s1 is wave-uniform, the substitution scores of a strip of
// R rows against every residue code form a small query profile that is staged
// in LDS once per strip and read with 128-bit LDS loads indexed by the lane's
// column residue.  The row above each strip (V-o, G, T, TG per column) lives in
// a per-workgroup boundary buffer in HBM laid out [column][lane] so that every
// access is one coalesced 1 KiB wave transaction; per cell that is 32/R bytes.
//
// Recurrence: SmithWatermanGotoh.Construct
// (Salsa.CoreTPL/Bio/Algorithms/SmithWaterman/SmithWatermanGotoh.cs:147-215);
// traceback statistics: SmithWatermanGotoh.Traceback (:262-315) rewritten as a
// forward recurrence: T(i,j) = statistics of the traceback path that starts by
// reading the pointer at (i,j), TG/TH those of the vertical/horizontal gap
// states.  T packs (columns << 16) | count, count = identity (nucleic) or
// similarity (protein); the increments come from the profile.
#include "swg_internal.h"

#include <climits>

namespace swg {

namespace {

__device__ __forceinline__ int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }

// Distance epilogue: ComputeDistance (Salsa.SmithWatermanMPI/Program.cs:259-314)
// with PairwiseAlignment.PercentIdentity/PercentSimilarity (Alignment.cs:187-195):
//   pct = count / (columns * 1.0f)     binary32, correctly rounded
//   Int16:  (short)((1.0f - pct) * short.MaxValue)   (x64 cvttss2si; NaN -> 0)
//   Double: (double)(1.0f - pct)
// The double quotient rounded to float is the correctly rounded binary32
// quotient (53 >= 2*24+2).  Built with -ffp-contract=off.
__device__ __forceinline__ void write_distance(const SeqSet& S, const BlockDesc& D, int a, int b,
                                               int score, int Tb)
{
    const uint32_t cols = (uint32_t)Tb >> 16;
    const uint32_t cnt = (uint32_t)Tb & 0xFFFFu;
    const float pct = (float)((double)cnt / (double)cols);
    const float rr = 1.0f - pct;
    const int64_t pos = (int64_t)(a - D.r0) * D.ld + (b - D.c0);
    const int64_t mpos = (int64_t)(b - D.r0) * D.ld + (a - D.c0);
    if (S.tdist) {
        const double dv = (double)rr;
        static_cast<double*>(D.out)[pos] = dv;
        if (D.mirror) static_cast<double*>(D.out)[mpos] = dv;
    } else {
        const float x = rr * 32767.0f;
        const int xi = (x == x) ? (int)x : 0;
        const short sv = (short)xi;
        static_cast<short*>(D.out)[pos] = sv;
        if (D.mirror) static_cast<short*>(D.out)[mpos] = sv;
    }
    if (D.stats) {
        reinterpret_cast<int4*>(D.stats)[pos] = make_int4(score, (int)cols, (int)cnt, 0);
        if (D.mirror) reinterpret_cast<int4*>(D.stats)[mpos] = make_int4(score, (int)cols, (int)cnt, 0);
    }
}

// Stage the strip's query profile: lprof[c2][r] = prof[s1[i0+r]][c2] for
// every residue code c2 (rows past the end of s1 use the pad code).
template <int R>
__device__ __forceinline__ void stage_profile(int4* lprof, const SeqSet& S, const uint8_t* s1,
                                              int La, int i0, int lane)
{
    const int ncode = S.ncode, pad = ncode - 1;
    int2* lp = reinterpret_cast<int2*>(lprof);
    for (int idx = lane; idx < ncode * R; idx += kWave) {
        const int c2 = idx / R, r = idx % R;
        const int row = i0 + r;
        const int c1 = row < La ? (int)s1[row] : pad;
        lp[c2 * (R + 2) + r] = S.prof[c1 * ncode + c2];
    }
}

template <int R, bool PERSIST>
__global__ __launch_bounds__(64) void sw_kernel(SeqSet S, TaskSet T, Scratch W)
{
    static_assert(R % 2 == 0, "strip height must be even");
    __shared__ int4 lprof[kMaxCodes * (R / 2 + 1)];

    const int lane = threadIdx.x;
    int4* __restrict__ bnd = W.bnd + (int64_t)blockIdx.x * W.bnd_stride;
    const int ntasks = T.desc->ntasks;
    const int o = S.gap_open, e = S.gap_ext;
    const int pad = S.ncode - 1;

    for (int iter = 0;; ++iter) {
        int t;
        if (PERSIST) {
            int v = 0;
            if (lane == 0) v = atomicAdd(T.queue, 1);
            t = uni(v);
        } else {
            t = blockIdx.x + iter * (int)gridDim.x;
        }
        if (t >= ntasks) break;

        const int a = uni(T.task_row[t]);
        const int tile = uni(T.task_tile[t]);
        const int maxLb = uni(T.tile_maxlen[tile]);
        const int b = T.tile_col[(int64_t)tile * kWave + lane];
        const int La = uni(S.len[a]);
        const int Lb = b >= 0 ? S.len[b] : 0;
        const uint8_t* s1 = S.res + S.off[a];
        const uint32_t* s2w = reinterpret_cast<const uint32_t*>(S.res + (b >= 0 ? S.off[b] : S.pad_off));
        const int ncols = (maxLb + 3) & ~3;

        // row 0 of the DP: V = 0 (t = -o), G = -inf, T = 0 (SmithWatermanGotoh.cs:131-139)
        for (int j = 1; j <= ncols; ++j) bnd[(int64_t)j * kWave + lane] = make_int4(-o, kNeg, 0, 0);

        int best = INT_MIN, bestT = 0;
        for (int i0 = 0; i0 < La; i0 += R) {
            __syncthreads();
            stage_profile<R>(lprof, S, s1, La, i0, lane);
            __syncthreads();

            int tr[R], Tr[R], Hr[R], THr[R], rb[R], rT[R];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                tr[r] = -o; Tr[r] = 0; Hr[r] = kNeg; THr[r] = 0; rb[r] = INT_MIN; rT[r] = 0;
            }
            int dtb = -o, dTb = 0;  // V[i0][0]-o and T[i0][0] (border: 0, STOP)

            uint32_t nw = s2w[0];
            int4 nb0 = bnd[1 * kWave + lane], nb1 = bnd[2 * kWave + lane];
            int4 nb2 = bnd[3 * kWave + lane], nb3 = bnd[4 * kWave + lane];
            for (int j = 1; j <= ncols; j += 4) {
                const uint32_t w = nw;
                int4 cb[4] = {nb0, nb1, nb2, nb3};
                nw = s2w[(j + 3) >> 2];
                const int64_t nxt = (int64_t)(j + 4) * kWave + lane;
                nb0 = bnd[nxt]; nb1 = bnd[nxt + kWave]; nb2 = bnd[nxt + 2 * kWave]; nb3 = bnd[nxt + 3 * kWave];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int jj = j + q;
                    int c2 = (int)((w >> (8 * q)) & 0xFFu);
                    c2 = jj > Lb ? pad : c2;
                    int tu = cb[q].x, Gu = cb[q].y, Tu = cb[q].z, TGu = cb[q].w;
                    int dt = dtb, dT = dTb;
                    dtb = tu; dTb = Tu;
                    const int4* pr = lprof + c2 * (R / 2 + 1);
#pragma unroll
                    for (int r2 = 0; r2 < R / 2; ++r2) {
                        const int4 pp = pr[r2];
#pragma unroll
                        for (int h = 0; h < 2; ++h) {
                            const int r = 2 * r2 + h;
                            const int sc = h ? pp.z : pp.x;
                            const int inc = h ? pp.w : pp.y;
                            // vertical gap (SmithWatermanGotoh.cs:159-172)
                            const int g1 = Gu - e;
                            const bool cg = g1 > tu;
                            const int G = cg ? g1 : tu;
                            const int TG = (cg ? TGu : Tu) + 0x10000;
                            // horizontal gap (:174-187)
                            const int h1 = Hr[r] - e;
                            const bool ch = h1 > tr[r];
                            Hr[r] = ch ? h1 : tr[r];
                            THr[r] = (ch ? THr[r] : Tr[r]) + 0x10000;
                            // diagonal and V (:154-157, :190)
                            const int f = dt + sc;
                            int V = max(max(f, G), Hr[r]);
                            V = max(V, 0);
                            // pointer order STOP > DIAGONAL > UP > LEFT (:193-208)
                            const int Td = dT + inc;
                            int x = (V == G) ? TG : THr[r];
                            x = (V == f) ? Td : x;
                            const int Tn = (V == 0) ? 0 : x;
                            // first strict maximum of this row (:211-214)
                            const bool cb2 = V > rb[r];
                            rb[r] = cb2 ? V : rb[r];
                            rT[r] = cb2 ? Tn : rT[r];
                            const int tn = V - o;
                            dt = tr[r]; dT = Tr[r];
                            tr[r] = tn; Tr[r] = Tn;
                            tu = tn; Gu = G; Tu = Tn; TGu = TG;
                        }
                    }
                    bnd[(int64_t)jj * kWave + lane] = make_int4(tu, Gu, Tu, TGu);
                }
            }
            // rows in order: first strict maximum in row-major order
#pragma unroll
            for (int r = 0; r < R; ++r) {
                if (i0 + r < La && rb[r] > best) { best = rb[r]; bestT = rT[r]; }
            }
        }

        if (b >= 0) {
            int score = best, Tb = bestT;
            if (La == 0 || Lb == 0) { score = INT_MIN; Tb = 0; }
            write_distance(S, *T.desc, a, b, score, Tb);
        }
    }
}

__global__ void fill_i32_kernel(int32_t* p, int32_t v, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

template <int R, bool P>
hipError_t launch_one(const SeqSet& S, const TaskSet& T, const Scratch& W, int nslots, hipStream_t st)
{
    hipLaunchKernelGGL((sw_kernel<R, P>), dim3(nslots), dim3(kWave), 0, st, S, T, W);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_sw(const SeqSet& S, const TaskSet& T, const Scratch& W, int strip_rows,
                     bool persistent, int nslots, hipStream_t st)
{
    switch (strip_rows) {
    case 8:  return persistent ? launch_one<8, true>(S, T, W, nslots, st)  : launch_one<8, false>(S, T, W, nslots, st);
    case 16: return persistent ? launch_one<16, true>(S, T, W, nslots, st) : launch_one<16, false>(S, T, W, nslots, st);
    case 32: return persistent ? launch_one<32, true>(S, T, W, nslots, st) : launch_one<32, false>(S, T, W, nslots, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_nw(const SeqSet&, const TaskSet&, const Scratch&, int, bool, int, hipStream_t)
{
    return hipErrorNotSupported;
}

int sw_blocks_per_cu(int strip_rows, bool /*nw*/)
{
    int nb = 0;
    hipError_t err = hipErrorInvalidValue;
    switch (strip_rows) {
    case 8:  err = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, sw_kernel<8, true>, kWave, 0); break;
    case 16: err = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, sw_kernel<16, true>, kWave, 0); break;
    case 32: err = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, sw_kernel<32, true>, kWave, 0); break;
    default: break;
    }
    return err == hipSuccess ? nb : 0;
}

hipError_t launch_fill_i32(int32_t* p, int32_t value, int64_t n, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(fill_i32_kernel, dim3((unsigned)blocks), dim3(256), 0, st, p, value, n);
    return hipGetLastError();
}

hipError_t launch_fill_bytes_zero(void* p, int64_t nbytes, hipStream_t st)
{
    // callers allocate with 4-byte granularity
    return launch_fill_i32(static_cast<int32_t*>(p), 0, (nbytes + 3) / 4, st);
}

}  // namespace swg
