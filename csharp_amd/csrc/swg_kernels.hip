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
// similarity (protein); the increments come from the profile.  For the
// Jukes-Cantor and Kimura2 distances the high half counts diagonal columns
// only (gap columns add 0) and the count is differences, transitions or
// transversions.
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
__device__ __forceinline__ void write_fraction(const SeqSet& S, const BlockDesc& D, int a, int b, uint32_t cnt,
                                               uint32_t den, int4 st)
{
    const float pct = (float)((double)cnt / (double)den);
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
        reinterpret_cast<int4*>(D.stats)[pos] = st;
        if (D.mirror) reinterpret_cast<int4*>(D.stats)[mpos] = st;
    }
}

// double -> Int16 as the x64 CLR converts it (conv.i2: cvttsd2si, then the
// low 16 bits; NaN and values outside the int32 range give 0x80000000)
__device__ __forceinline__ short f64_to_i16(double x)
{
    const int xi = (x > -2147483649.0 && x < 2147483648.0) ? (int)x : INT_MIN;
    return (short)xi;
}

// double.Epsilon of the CLR: the smallest positive subnormal
constexpr double kDoubleEpsilon = 4.9406564584124654e-324;

__device__ __forceinline__ void write_distance(const SeqSet& S, const BlockDesc& D, int a, int b,
                                               int score, int Tb)
{
    const uint32_t hi = (uint32_t)Tb >> 16, lo = (uint32_t)Tb & 0xFFFFu;
    if (S.dist_kind == kDistFraction) {
        write_fraction(S, D, a, b, lo, hi, make_int4(score, (int)hi, (int)lo, 0));
        return;
    }
    // word mode: distance_kernel computes the distance
    D.aux[((int64_t)S.aux_plane * D.rows + (a - D.r0)) * D.aux_ld + (b - D.c0)] =
        S.dist_kind == kDistMinMax ? (uint32_t)score : (uint32_t)Tb;
    if (D.stats) {
        const int4 st = make_int4(score, (int)hi, (int)lo, 0);
        reinterpret_cast<int4*>(D.stats)[(int64_t)(a - D.r0) * D.ld + (b - D.c0)] = st;
        if (D.mirror) reinterpret_cast<int4*>(D.stats)[(int64_t)(b - D.r0) * D.ld + (a - D.c0)] = st;
    }
}

// Stage the strip's query profile: lprof[c2][r] = prof[s1[i0+r]][c2] for
// every residue code c2 (rows past the end of s1 use the pad code).
// __builtin_amdgcn_sched_barrier mask: VALU and SALU instructions may be
// scheduled across the barrier, memory instructions may not
constexpr int kSchedAluOnly = 0x0006;

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
    if ((int)blockIdx.x >= ntasks) return;  // more workgroups than tasks: the first ntasks drain the queue
    const int o = S.gap_open, e = S.gap_ext;
    const int ginc = S.gap_inc;  // count-word increment per gap column
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
                tr[r] = -o; Tr[r] = 0; Hr[r] = 0; THr[r] = 0; rb[r] = INT_MIN; rT[r] = 0;
            }
            int dtb = -o, dTb = 0;  // V[i0][0]-o and T[i0][0] (border: 0, STOP)

            // residue words: columns j .. j+3 and the next block, loaded a
            // block ahead (word (Lb + 3) / 4 holds only pad codes, and the
            // words before it pad codes past Lb)
            const int lastW = (Lb + 3) >> 2;
            uint32_t wcur = s2w[0], wnx = s2w[min(1, lastW)];
            __builtin_amdgcn_sched_barrier(kSchedAluOnly);  // in the loop's order
            // boundary words of the next two columns, each reloaded two
            // columns ahead once its column is done
            int4 bw[2] = {bnd[1 * kWave + lane], bnd[2 * kWave + lane]};
            // profile words (score, increment) of the current column, two rows
            // per int4, each reloaded with the next column's once its rows
            // are done
            int4 pp[R / 2];
            {
                const int c2 = 1 > Lb ? pad : (int)(wcur & 0xFFu);
#pragma unroll
                for (int r2 = 0; r2 < R / 2; ++r2) pp[r2] = lprof[c2 * (R / 2 + 1) + r2];
            }
            __builtin_amdgcn_sched_barrier(kSchedAluOnly);
            for (int j = 1; j <= ncols; j += 4) {
                const uint32_t wc = wcur;
                wcur = wnx;
                wnx = s2w[min(((j + 3) >> 2) + 1, lastW)];
                __builtin_amdgcn_sched_barrier(kSchedAluOnly);
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int jj = j + q;
                    // code of column jj + 1 (its profile words are loaded
                    // during this column)
                    int c2n = (int)(((q < 3 ? wc : wcur) >> (8 * ((q + 1) & 3))) & 0xFFu);
                    c2n = jj + 1 > Lb ? pad : c2n;
                    const int4* prn = lprof + c2n * (R / 2 + 1);
                    int tu = bw[q & 1].x, Gu = bw[q & 1].y, Tu = bw[q & 1].z, TGu = bw[q & 1].w;
                    int dt = dtb, dT = dTb;
                    dtb = tu; dTb = Tu;
#pragma unroll
                    for (int r2 = 0; r2 < R / 2; ++r2) {
                        const int4 pq = pp[r2];
#pragma unroll
                        for (int h = 0; h < 2; ++h) {
                            const int r = 2 * r2 + h;
                            const int sc = h ? pq.z : pq.x;
                            const int inc = h ? pq.w : pq.y;
                            // vertical gap (SmithWatermanGotoh.cs:159-172)
                            const int g1 = Gu - e;
                            const bool cg = g1 > tu;
                            const int G = cg ? g1 : tu;
                            const int TG = (cg ? TGu : Tu) + ginc;
                            // horizontal gap (:174-187)
                            // H is kept as max(H, 0) (an unsigned saturating subtract): V = max(f, G, H, 0)
                            // = max(f, G, H'); the gap run it carries matters only where H >= 0, where H' = H
                            const int h1 = (int)__builtin_elementwise_sub_sat((unsigned)Hr[r], (unsigned)e);
                            const bool ch = h1 > tr[r];
                            Hr[r] = ch ? h1 : tr[r];
                            THr[r] = (ch ? THr[r] : Tr[r]) + ginc;
                            // diagonal and V (:154-157, :190)
                            const int f = dt + sc;
                            int V = max(max(f, G), Hr[r]);
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
                        pp[r2] = prn[r2];
                        // the loads stay a column ahead of their use (only ALU
                        // instructions may cross)
                        if ((r2 & 1) == 1) __builtin_amdgcn_sched_barrier(kSchedAluOnly);
                    }
                    bnd[(int64_t)jj * kWave + lane] = make_int4(tu, Gu, Tu, TGu);
                    bw[q & 1] = bnd[(int64_t)(jj + 2) * kWave + lane];
                    __builtin_amdgcn_sched_barrier(kSchedAluOnly);
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

// ---------------------------------------------------------------------------
// Sixteen lanes per pair: a wave aligns s1 (uniform) against 4 sequences s2,
// one per 16-lane DPP row.  Lane k of a group owns R consecutive DP rows of a
// stripe of 16*R rows and runs one column behind lane k-1 (anti-diagonal
// wavefront): at step t it computes column t-k+1 of its rows, taking the row
// above its first row (V-o, G, T, TG) and the column residue from lane k-1
// with a DPP row shift, while lane 0 reads the previous stripe's bottom row
// from the boundary buffer and lane 15 writes this stripe's.  Four times the
// waves per sequence s1 of the one-lane-per-pair kernel, for latency when few
// sequences s1 are in flight.
//
// Lanes before their first column ("pre-columns", j < 1) run pad cells whose
// value is 0 with T = 0, i.e. the state of DP column 0 (this needs
// gap_open >= 0 and gap_ext >= 0); lanes past the last column run pad cells,
// which can never hold a new row maximum (see sw_kernel).
constexpr int kLppPairs = 4;   // pairs per wave
constexpr int kLpp = 16;       // lanes per pair

__device__ __forceinline__ int row_shr1(int old, int src)
{
    return __builtin_amdgcn_update_dpp(old, src, 0x111, 0xF, 0xF, false);  // row_shr:1
}

template <int R, bool PERSIST>
__global__ __launch_bounds__(64) void lpp16_kernel(SeqSet S, TaskSet T, Scratch W)
{
    constexpr int H = kLpp * R;   // stripe height
    constexpr int PS = H / 2 + 1; // int4 per residue code in the stripe profile
    extern __shared__ int4 dyn_lds[];
    int4* lprof = dyn_lds;                                             // ncode * PS int4
    int2* lmerge = reinterpret_cast<int2*>(dyn_lds + S.ncode * PS);    // 4 groups * H (best, T)

    const int lane = threadIdx.x;
    const int g = lane >> 4, k = lane & 15;
    int4* __restrict__ bnd = W.bnd + (int64_t)blockIdx.x * W.bnd_stride;
    const int ntasks = T.desc->ntasks;
    if ((int)blockIdx.x >= ntasks) return;  // more workgroups than tasks: the first ntasks drain the queue
    const int o = S.gap_open, e = S.gap_ext;
    const int ginc = S.gap_inc;  // count-word increment per gap column
    const int pad = S.ncode - 1;

    for (int iter = 0;; ++iter) {
        int t_task;
        if (PERSIST) {
            int v = 0;
            if (lane == 0) v = atomicAdd(T.queue, 1);
            t_task = uni(v);
        } else {
            t_task = blockIdx.x + iter * (int)gridDim.x;
        }
        if (t_task >= ntasks) break;

        const int a = uni(T.task_row[t_task]);
        const int tile = uni(T.task_tile[t_task]);
        const int maxLb = uni(T.tile_maxlen[tile]);
        const int b = T.tile_col[(int64_t)tile * kLppPairs + g];
        const int La = uni(S.len[a]);
        const int Lb = b >= 0 ? S.len[b] : 0;
        const uint8_t* s1 = S.res + S.off[a];
        const uint32_t* s2w = reinterpret_cast<const uint32_t*>(S.res + (b >= 0 ? S.off[b] : S.pad_off));
        const int nsteps = (maxLb + (kLpp - 1) + 3) & ~3;  // lane 15 reaches column maxLb

        // row 0 of the DP (SmithWatermanGotoh.cs:131-139), also beyond the last
        // column so that lane 0 always reads defined values
        for (int j = 1 + lane; j <= nsteps + 8; j += kWave)
            for (int q = 0; q < kLppPairs; ++q) bnd[(int64_t)j * kLppPairs + q] = make_int4(-o, kNeg, 0, 0);
        __syncthreads();  // other lanes read these columns

        int best = INT_MIN, bestT = 0;
        for (int i0 = 0; i0 < La; i0 += H) {
            __syncthreads();
            // stage the stripe profile: lprof[c2][row], rows past s1 use the pad code
            {
                int2* lp = reinterpret_cast<int2*>(lprof);
                for (int idx = lane; idx < S.ncode * H; idx += kWave) {
                    const int c2 = idx / H, r = idx % H;
                    const int row = i0 + r;
                    const int c1 = row < La ? (int)s1[row] : pad;
                    lp[c2 * (H + 2) + r] = S.prof[c1 * S.ncode + c2];
                }
            }
            __syncthreads();

            int tr[R], Tr[R], Hr[R], THr[R], rb[R], rT[R];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                tr[r] = -o; Tr[r] = 0; Hr[r] = 0; THr[r] = 0; rb[r] = INT_MIN; rT[r] = 0;
            }
            int ot = -o, oG = kNeg, oT = 0, oTG = 0;  // this lane's last-row output of the previous step
            int pt = -o, pT = 0;                      // up-values received at the previous step (diagonal)
            int cprev = pad;

            uint32_t nw = s2w[0];
            int4 nb0 = bnd[1 * kLppPairs + g], nb1 = bnd[2 * kLppPairs + g];
            int4 nb2 = bnd[3 * kLppPairs + g], nb3 = bnd[4 * kLppPairs + g];
            for (int t = 0; t < nsteps; t += 4) {
                const uint32_t cw = nw;
                int4 cb[4] = {nb0, nb1, nb2, nb3};
                nw = s2w[(t + 4) >> 2];
                const int64_t nxt = (int64_t)(t + 5) * kLppPairs + g;
                nb0 = bnd[nxt]; nb1 = bnd[nxt + kLppPairs]; nb2 = bnd[nxt + 2 * kLppPairs];
                nb3 = bnd[nxt + 3 * kLppPairs];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int tt = t + q;
                    int c0 = (int)((cw >> (8 * q)) & 0xFFu);
                    c0 = tt + 1 > Lb ? pad : c0;
                    const int c2 = row_shr1(c0, cprev);
                    cprev = c2;
                    int tu = row_shr1(cb[q].x, ot), Gu = row_shr1(cb[q].y, oG);
                    int Tu = row_shr1(cb[q].z, oT), TGu = row_shr1(cb[q].w, oTG);
                    int dt = pt, dT = pT;
                    pt = tu; pT = Tu;
                    const int4* pr = lprof + c2 * PS + k * (R / 2);
#pragma unroll
                    for (int r2 = 0; r2 < R / 2; ++r2) {
                        const int4 pp = pr[r2];
#pragma unroll
                        for (int h = 0; h < 2; ++h) {
                            const int r = 2 * r2 + h;
                            const int sc = h ? pp.z : pp.x;
                            const int inc = h ? pp.w : pp.y;
                            const int g1 = Gu - e;
                            const bool cg = g1 > tu;
                            const int G = cg ? g1 : tu;
                            const int TG = (cg ? TGu : Tu) + ginc;
                            // H is kept as max(H, 0) (an unsigned saturating subtract): V = max(f, G, H, 0)
                            // = max(f, G, H'); the gap run it carries matters only where H >= 0, where H' = H
                            const int h1 = (int)__builtin_elementwise_sub_sat((unsigned)Hr[r], (unsigned)e);
                            const bool ch = h1 > tr[r];
                            Hr[r] = ch ? h1 : tr[r];
                            THr[r] = (ch ? THr[r] : Tr[r]) + ginc;
                            const int f = dt + sc;
                            int V = max(max(f, G), Hr[r]);
                            const int Td = dT + inc;
                            int x = (V == G) ? TG : THr[r];
                            x = (V == f) ? Td : x;
                            const int Tn = (V == 0) ? 0 : x;
                            const bool cb2 = V > rb[r];
                            rb[r] = cb2 ? V : rb[r];
                            rT[r] = cb2 ? Tn : rT[r];
                            const int tn = V - o;
                            dt = tr[r]; dT = Tr[r];
                            tr[r] = tn; Tr[r] = Tn;
                            tu = tn; Gu = G; Tu = Tn; TGu = TG;
                        }
                    }
                    ot = tu; oG = Gu; oT = Tu; oTG = TGu;
                    const int j = tt - k + 1;
                    if (k == kLpp - 1 && j >= 1) bnd[(int64_t)j * kLppPairs + g] = make_int4(ot, oG, oT, oTG);
                }
            }
            // first strict maximum in row-major order over the stripe's rows
#pragma unroll
            for (int r = 0; r < R; ++r) lmerge[g * H + k * R + r] = make_int2(rb[r], rT[r]);
            __syncthreads();
            if (k == 0) {
                for (int r = 0; r < H && i0 + r < La; ++r) {
                    const int2 m = lmerge[g * H + r];
                    if (m.x > best) { best = m.x; bestT = m.y; }
                }
            }
        }
        if (k == 0 && b >= 0) {
            int score = best, Tb = bestT;
            if (La == 0 || Lb == 0) { score = INT_MIN; Tb = 0; }
            write_distance(S, *T.desc, a, b, score, Tb);
        }
    }
}

// ---------------------------------------------------------------------------
// Global (Needleman-Wunsch) alignment, one lane per pair, same strip layout
// as sw_kernel.  Recurrence of SmithWatermanGotoh.Construct without the 0
// floor, borders V[i][0] = -(o+(i-1)e) and V[0][j] = -(o+(j-1)e) as single gap
// runs back to the origin, result at (m, n), DIAGONAL > UP > LEFT.
// The distance of Salsa.NeedlemanWunschTPL/Program.cs:269-300 needs, besides
// (columns, count), the aligned region [first non-gap, last non-gap]: its
// length is columns - lead - trail, where lead is the border run the path
// starts with and trail the run of same-direction gap columns it ends with.
// A second packed word carries them: T2 = (lead << 16) | (up << 15) | run,
// `run` = length of the gap run the traceback starts with (up = vertical).
// REGION = false leaves that word out (protein PercentSimilarity over all
// columns, the word-mode distances, no statistics requested): 12 of the
// kernel's operations per cell and a third of its boundary traffic.
constexpr uint32_t kUpBit = 0x8000u;

__device__ __forceinline__ int nw_border(int o, int e, int k) { return -(o + (k - 1) * e); }

template <int R, bool PERSIST, bool REGION>
__global__ __launch_bounds__(64) void nw_kernel(SeqSet S, TaskSet T, Scratch W)
{
    __shared__ int4 lprof[kMaxCodes * (R / 2 + 1)];
    const int lane = threadIdx.x;
    // per slot: (t, G, T1, TG1) then (T2, TG2), each [column][lane]
    int4* __restrict__ bnd = W.bnd + (int64_t)blockIdx.x * W.bnd_stride;
    int2* __restrict__ bnd2 = reinterpret_cast<int2*>(bnd + W.bnd_stride / 2);
    const int ntasks = T.desc->ntasks;
    if ((int)blockIdx.x >= ntasks) return;  // more workgroups than tasks: the first ntasks drain the queue
    const int o = S.gap_open, e = S.gap_ext;
    const int ginc = S.gap_inc;  // count-word increment per gap column
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

        // row 0: a horizontal gap run of j columns from the origin
        for (int j = 1; j <= ncols; ++j) {
            bnd[(int64_t)j * kWave + lane] = make_int4(nw_border(o, e, j) - o, kNeg, j * ginc, 0);
            if (REGION) bnd2[(int64_t)j * kWave + lane] = make_int2((j << 16) | j, 0);
        }

        // result cell (La, Lb): border cells when a sequence is empty
        int capV = 0;
        uint32_t cap1 = 0, cap2 = 0;
        if (La == 0 && Lb > 0) { capV = nw_border(o, e, Lb); cap1 = (uint32_t)(Lb * ginc); cap2 = ((uint32_t)Lb << 16) | Lb; }
        if (Lb == 0 && La > 0) {
            capV = nw_border(o, e, La); cap1 = (uint32_t)(La * ginc); cap2 = ((uint32_t)La << 16) | kUpBit | La;
        }

        for (int i0 = 0; i0 < La; i0 += R) {
            __syncthreads();
            stage_profile<R>(lprof, S, s1, La, i0, lane);
            __syncthreads();
            const int rcap = La - 1 - i0;  // row of the result cell in this strip (if < R)

            int tr[R], Tr1[R], Tr2[R], Hr[R], TH1[R], TH2[R];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int i = i0 + r + 1;  // column-0 border: a vertical run of i rows
                tr[r] = nw_border(o, e, i) - o; Tr1[r] = i * ginc; Tr2[r] = (i << 16) | (int)kUpBit | i;
                Hr[r] = kNeg; TH1[r] = 0; TH2[r] = 0;
            }
            int dtb, dTb1, dTb2;  // V[i0][0] for the first row's diagonal
            if (i0 == 0) { dtb = -o; dTb1 = 0; dTb2 = 0; }
            else { dtb = nw_border(o, e, i0) - o; dTb1 = i0 * ginc; dTb2 = (i0 << 16) | (int)kUpBit | i0; }

            // residue words: columns j .. j+3 and the next block, loaded a
            // block ahead (word (Lb + 3) / 4 holds only pad codes, and the
            // words before it pad codes past Lb)
            const int lastW = (Lb + 3) >> 2;
            uint32_t wcur = s2w[0], wnx = s2w[min(1, lastW)];
            __builtin_amdgcn_sched_barrier(kSchedAluOnly);  // in the loop's order
            // boundary words of the next two columns, each reloaded two
            // columns ahead once its column is done
            int4 bw[2] = {bnd[(int64_t)1 * kWave + lane], bnd[(int64_t)2 * kWave + lane]};
            int2 bw2[2] = {make_int2(0, 0), make_int2(0, 0)};
            if (REGION) { bw2[0] = bnd2[(int64_t)1 * kWave + lane]; bw2[1] = bnd2[(int64_t)2 * kWave + lane]; }
            // profile words (score, increment) of the current column, two rows
            // per int4, each reloaded with the next column's once its rows
            // are done
            int4 pp[R / 2];
            {
                const int c2 = 1 > Lb ? pad : (int)(wcur & 0xFFu);
#pragma unroll
                for (int r2 = 0; r2 < R / 2; ++r2) pp[r2] = lprof[c2 * (R / 2 + 1) + r2];
            }
            __builtin_amdgcn_sched_barrier(kSchedAluOnly);

            for (int j = 1; j <= ncols; j += 4) {
                const uint32_t wc = wcur;
                wcur = wnx;
                wnx = s2w[min(((j + 3) >> 2) + 1, lastW)];
                __builtin_amdgcn_sched_barrier(kSchedAluOnly);
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int jj = j + q;
                    // code of column jj + 1 (its profile words are loaded
                    // during this column)
                    int c2n = (int)(((q < 3 ? wc : wcur) >> (8 * ((q + 1) & 3))) & 0xFFu);
                    c2n = jj + 1 > Lb ? pad : c2n;
                    const int4* prn = lprof + c2n * (R / 2 + 1);
                    const int4 bu = bw[q & 1];
                    const int2 bu2 = bw2[q & 1];
                    int tu = bu.x, Gu = bu.y, Tu1 = bu.z, TGu1 = bu.w, Tu2 = bu2.x, TGu2 = bu2.y;
                    int dt = dtb, dT1 = dTb1, dT2 = dTb2;
                    dtb = tu; dTb1 = Tu1; dTb2 = Tu2;
                    const bool hitcol = jj == Lb;
#pragma unroll
                    for (int r2 = 0; r2 < R / 2; ++r2) {
                        const int4 pq = pp[r2];
#pragma unroll
                        for (int h = 0; h < 2; ++h) {
                            const int r = 2 * r2 + h;
                            const int sc = h ? pq.z : pq.x;
                            const int inc = h ? pq.w : pq.y;
                            // vertical gap: extend TG or open from T (keeping its run if vertical)
                            const int g1 = Gu - e;
                            const bool cg = g1 > tu;
                            const int G = cg ? g1 : tu;
                            const int TG1 = (cg ? TGu1 : Tu1) + ginc;
                            int TG2 = 0;
                            if (REGION) {
                                const int baseG = (Tu2 & (int)kUpBit) ? (Tu2 ^ (int)kUpBit) : (Tu2 & (int)0xFFFF0000u);
                                TG2 = (cg ? TGu2 : baseG) + 1;
                            }
                            // horizontal gap
                            const int h1 = Hr[r] - e;
                            const bool ch = h1 > tr[r];
                            Hr[r] = ch ? h1 : tr[r];
                            TH1[r] = (ch ? TH1[r] : Tr1[r]) + ginc;
                            if (REGION) {
                                const int baseH = (Tr2[r] & (int)kUpBit) ? (Tr2[r] & (int)0xFFFF0000u) : Tr2[r];
                                TH2[r] = (ch ? TH2[r] : baseH) + 1;
                            }
                            const int f = dt + sc;
                            const int V = max(max(f, G), Hr[r]);
                            int x1 = (V == G) ? TG1 : TH1[r];
                            int x2 = 0;
                            if (REGION) x2 = (V == G) ? (TG2 | (int)kUpBit) : TH2[r];
                            const bool dg = V == f;
                            x1 = dg ? dT1 + inc : x1;
                            if (REGION) x2 = dg ? (dT2 & (int)0xFFFF0000u) : x2;
                            const int tn = V - o;
                            if (r == rcap) {
                                capV = hitcol ? V : capV;
                                cap1 = hitcol ? (uint32_t)x1 : cap1;
                                cap2 = hitcol ? (uint32_t)x2 : cap2;
                            }
                            dt = tr[r]; dT1 = Tr1[r]; dT2 = Tr2[r];
                            tr[r] = tn; Tr1[r] = x1; Tr2[r] = x2;
                            tu = tn; Gu = G; Tu1 = x1; TGu1 = TG1; Tu2 = x2; TGu2 = TG2;
                        }
                        pp[r2] = prn[r2];
                        // the loads stay a column ahead of their use (only ALU
                        // instructions may cross)
                        if ((r2 & 1) == 1) __builtin_amdgcn_sched_barrier(kSchedAluOnly);
                    }
                    bnd[(int64_t)jj * kWave + lane] = make_int4(tu, Gu, Tu1, TGu1);
                    if (REGION) bnd2[(int64_t)jj * kWave + lane] = make_int2(Tu2, TGu2);
                    bw[q & 1] = bnd[(int64_t)(jj + 2) * kWave + lane];
                    if (REGION) bw2[q & 1] = bnd2[(int64_t)(jj + 2) * kWave + lane];
                    __builtin_amdgcn_sched_barrier(kSchedAluOnly);
                }
            }
        }

        if (b >= 0) {
            const uint32_t cols = cap1 >> 16, cnt = cap1 & 0xFFFFu;
            if (!REGION) cap2 = 0;  // region not carried: reported as the whole alignment
            const uint32_t lead = cap2 >> 16, trail = cap2 & 0x7FFFu;
            // region = last - first + 1 over the non-gap indices; with both
            // sequences empty both indices are -1 and the region counts 1
            int region = (La == 0 && Lb == 0) ? 1 : (La == 0 || Lb == 0) ? 0 : (int)cols - (int)lead - (int)trail;
            if (region < 0) region = 0;
            // nucleic: identity over the aligned region (Program.cs:547-571);
            // protein: positive-score columns over all columns (:583-604)
            if (S.dist_kind != kDistFraction) {
                // word mode: the score (MinMaxNormScore) or the count word
                // for distance_kernel, and the statistics
                const BlockDesc& D = *T.desc;
                D.aux[((int64_t)S.aux_plane * D.rows + (a - D.r0)) * D.aux_ld + (b - D.c0)] =
                    S.dist_kind == kDistMinMax ? (uint32_t)capV : cap1;
                if (D.stats) {
                    const int4 st = S.dist_kind == kDistMinMax ? make_int4(capV, (int)cols, (int)cnt, region)
                                                               : make_int4(capV, (int)cols, (int)cnt, 0);
                    reinterpret_cast<int4*>(D.stats)[(int64_t)(a - D.r0) * D.ld + (b - D.c0)] = st;
                    if (D.mirror) reinterpret_cast<int4*>(D.stats)[(int64_t)(b - D.r0) * D.ld + (a - D.c0)] = st;
                }
            } else {
                write_fraction(S, *T.desc, a, b, cnt, S.nw_region ? (uint32_t)region : cols,
                               make_int4(capV, (int)cols, (int)cnt, region));
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Score-only pass, two pairs per lane in packed int16 halves.
//
// Same strip structure and query profile as sw_kernel, but only the score
// recurrence (SmithWatermanGotoh.cs:154-190) and the first strict maximum
// (:211-214) are evaluated: per row, the packed row maximum and the column of
// its first occurrence (v_pk_* ops, one instruction for two cells).  The
// diagonal add saturates, so a score that would exceed 32767 pins at 32767 and
// the pair is sent to the exact int32 re-run over its whole matrix.  Every
// other operation only lowers values, and all values stay >= -(gap_open+16384).
typedef short s16x2 __attribute__((ext_vector_type(2)));
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ s16x2 v2(uint32_t x) { return __builtin_bit_cast(s16x2, x); }
__device__ __forceinline__ uint32_t u2(s16x2 v) { return __builtin_bit_cast(uint32_t, v); }
__device__ __forceinline__ uint32_t splat16(int x) { return ((uint32_t)x & 0xFFFFu) * 0x10001u; }

constexpr int kPad16 = -16384;
constexpr int kSat16 = 32766;  // best >= this: the pair may have saturated

// The max3 form of the score pass (MX3): every value is kept biased by
// gap_open, so that all of them are >= 0 -- t = V - o + o = V, G + o,
// H' + o, the diagonal candidate V_diag + score + o (the profile holds
// score + o, required >= 0 for the set's codes, and 0 for the pad code:
// a pad score of -o) -- and then, as int16 bit patterns in [0, 0x7BFF], they
// order exactly as the fp16 numbers they spell (non-negative fp16 patterns
// below the infinity are ordered like integers; f16 denormals are not
// flushed in these kernels).  gfx950's v_pk_maximum3_f16 is then an exact
// integer max of three: V = max3(f, G, H') in one packed operation instead
// of two, and the row's running maximum takes two columns per operation
// (max3(bm, t of column j-1, t of column j) every other column).  A value
// past 0x7BFF is an infinity or a positive NaN pattern; the first such value
// comes from exact operands, every later operation keeps a result >= 0x7C00
// - o, and the pair is flagged (best >= kSatMx3(o)) for the int32 re-run:
// the numbers below the threshold are exact.
typedef _Float16 h16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t max3b(uint32_t a, uint32_t b, uint32_t c)
{
    const h16x2 r = __builtin_elementwise_maximum(
        __builtin_elementwise_maximum(__builtin_bit_cast(h16x2, a), __builtin_bit_cast(h16x2, b)),
        __builtin_bit_cast(h16x2, c));
    return __builtin_bit_cast(uint32_t, r);
}
__device__ __forceinline__ int kSatMx3(int o) { return 0x7C00 - o; }

template <int R, bool MX3>
__device__ __forceinline__ void stage_profile16(int4* lprof, const SeqSet& S, const uint8_t* s1, int La, int i0,
                                                int lane)
{
    const int ncode = S.ncode, pad = ncode - 1;
    short* lp = reinterpret_cast<short*>(lprof);
    for (int idx = lane; idx < ncode * R; idx += kWave) {
        const int c2 = idx / R, r = idx % R;
        const int row = i0 + r;
        const int c1 = row < La ? (int)s1[row] : pad;
        const int v = S.prof[c1 * ncode + c2].x;
        const int lo = MX3 ? 0 : kPad16;  // the pad score (MX3: -gap_open, biased to 0)
        lp[c2 * (R + 8) + r] = (short)(v < lo ? lo : v);
    }
}

// Joint profile for small alphabets: entry (ca, cc) of row r holds the
// packed pair (score of code ca | score of code cc << 16), i.e. both halves'
// scores of one lane in one word.  Layout [r][ca][cc] with 8 words per ca and
// a row stride of 64 words: the byte offset of a pair in its row is
// ca * 32 + cc * 4 < 256, so the offsets of four columns are the bytes of
// (word of A's codes << 5) | (word of B's codes << 2), and a lane reads two
// rows with one ds_read2st64_b32.
constexpr int kJointMaxCodes = 8;
constexpr int kJointRow = 64;  // words per row (= kJointMaxCodes^2)

template <int R, bool MX3>
__device__ __forceinline__ void stage_joint16(uint32_t* jp, const SeqSet& S, const uint8_t* s1, int La, int i0,
                                              int lane)
{
    const int ncode = S.ncode, pad = ncode - 1, npr = ncode * ncode;
    for (int idx = lane; idx < npr * R; idx += kWave) {
        const int r = idx / npr, pr = idx - r * npr;
        const int ca = pr / ncode, cc = pr - ca * ncode;
        const int row = i0 + r;
        const int c1 = row < La ? (int)s1[row] : pad;
        int va = S.prof[c1 * ncode + ca].x, vc = S.prof[c1 * ncode + cc].x;
        const int lo = MX3 ? 0 : kPad16;
        va = va < lo ? lo : va;
        vc = vc < lo ? lo : vc;
        jp[r * kJointRow + ca * kJointMaxCodes + cc] = ((uint32_t)va & 0xFFFFu) | ((uint32_t)vc << 16);
    }
}

// First maximum: per row only the maximum of the current kFirstMaxBlock-column
// block is kept; once per block the rows' block maxima are folded, in row
// order, into the strip's first maximum (value, row, block), which the end of
// the strip folds into the pair's.  Row-major order decides: a larger value,
// or an equal value in an earlier row (a later block of the same row never
// replaces an equal value).  The re-run finds the exact column inside the
// block.  best_cell holds (bi << 16) | jb, jb = the block's first column.
// Rows past La (the last strip's padding) never exceed the valid rows above
// them (their values come only from vertical and horizontal gaps of those
// rows and a diagonal add of kPad16), and they follow every valid row, so
// they cannot take the first maximum.
// The profile words of a column are loaded one column ahead into the
// registers of the rows just computed (no second set); with three row states
// per row the kernel runs score16_waves(R) waves per SIMD.
constexpr int score16_waves(int R) { return R <= 16 ? 4 : R <= 24 ? 3 : 2; }

template <int N>
struct StripRows {
    static constexpr int value = N;
};

template <int R, bool JOINT, bool MX3>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(score16_waves(R), score16_waves(R)))) void
score16_kernel(SeqSet S, ScoreTaskSet T, Scratch W)
{
    static_assert(R % 8 == 0, "strip height must be a multiple of 8");
    constexpr int PS = R / 8 + 1;  // int4 per residue code (R int16 + 8 int16 of padding)
    constexpr int LDS_INT4 = JOINT ? R * kJointRow / 4 : kMaxCodes * PS;
    __shared__ int4 lprof[LDS_INT4];

    const int lane = threadIdx.x;
    int2* __restrict__ bnd = reinterpret_cast<int2*>(W.bnd + (int64_t)blockIdx.x * W.bnd_stride);
    const BlockDesc* D = T.desc;
    const int ntasks = D->ntasks;
    if ((int)blockIdx.x >= ntasks) return;  // more workgroups than tasks: the first ntasks drain the queue
    const int o = S.gap_open, e = S.gap_ext;
    const s16x2 O2 = v2(splat16(o)), E2 = v2(splat16(e));
    const uint32_t NEG2 = splat16(kPad16), M1 = splat16(-1);
    // t = V - o, G and H' at DP row 0 / column 0 (MX3: biased by o)
    const uint32_t T0 = MX3 ? 0u : splat16(-o), G0 = MX3 ? 0u : NEG2, H0 = MX3 ? splat16(o) : 0u;
    const int sat = MX3 ? kSatMx3(o) : kSat16;

    for (;;) {
        int v = 0;
        if (lane == 0) v = atomicAdd(T.queue, 1);
        const int t = uni(v);
        if (t >= ntasks) break;

        const int a = uni(T.task_row[t]);
        const int tile = uni(T.task_tile[t]);
        const int maxLb = uni(T.tile_maxlen[tile]);
        const int bA = T.tile_col[(int64_t)tile * (2 * kWave) + lane];
        const int bB = T.tile_col[(int64_t)tile * (2 * kWave) + kWave + lane];
        const int La = uni(S.len[a]);
        const int LbA = bA >= 0 ? S.len[bA] : 0;
        const int LbB = bB >= 0 ? S.len[bB] : 0;
        const uint8_t* s1 = S.res + S.off[a];
        const uint32_t* wA = reinterpret_cast<const uint32_t*>(S.res + (bA >= 0 ? S.off[bA] : S.pad_off));
        const uint32_t* wB = reinterpret_cast<const uint32_t*>(S.res + (bB >= 0 ? S.off[bB] : S.pad_off));
        // word (L + 3) / 4 of a column sequence holds only pad codes (and the
        // words before it pad codes past L): reading word min(k, lastA)
        // gives the pad code for every column past the sequence's end
        const int lastA = (LbA + 3) >> 2, lastB = (LbB + 3) >> 2;
        const int ncols = (maxLb + 3) & ~3;

        for (int j = 1; j <= ncols; ++j) bnd[(int64_t)j * kWave + lane] = make_int2((int)T0, (int)G0);

        // the pair's first maximum: value (-1: none yet), row (1-based) and
        // first column of its block, one pair per 16-bit half
        uint32_t GV = M1, GR = 0u, GJ = 0u;
        // one strip of RR rows from row i0: R rows, and for the last strip
        // of the row sequence the next multiple of 8 rows at or above what
        // is left (fewer pad rows)
        auto strip = [&](auto rr, const int i0) {
            constexpr int RR = decltype(rr)::value;
            __syncthreads();
            if (JOINT) stage_joint16<RR, MX3>(reinterpret_cast<uint32_t*>(lprof), S, s1, La, i0, lane);
            else stage_profile16<RR, MX3>(lprof, S, s1, La, i0, lane);
            __syncthreads();

            uint32_t tr[RR], Hr[RR], bm[RR];
#pragma unroll
            for (int r = 0; r < RR; ++r) { tr[r] = T0; Hr[r] = H0; bm[r] = NEG2; }
            const uint32_t row1 = splat16(i0 + 1);  // row of strip row 0
            uint32_t dtb = T0;

            // residue words: current block, next block (needed by the
            // look-ahead load of the last column) and, in the loop, the one
            // after; loaded before the boundary words, in the loop's order,
            // so that the loop waits for each load only as long as it must
            uint32_t nwa = wA[0], nwb = wB[0];
            uint32_t n2a = wA[min(1, lastA)], n2b = wB[min(1, lastB)];
            __builtin_amdgcn_sched_barrier(kSchedAluOnly);
            // boundary words of the next four columns, each reloaded four
            // columns ahead once its column is done
            int2 cb[4] = {bnd[1 * kWave + lane], bnd[2 * kWave + lane], bnd[3 * kWave + lane],
                          bnd[4 * kWave + lane]};
            __builtin_amdgcn_sched_barrier(kSchedAluOnly);
            // profile words of the current column: group g = rows 4g .. 4g+3
            // (JOINT: the joint entries at byte offset ca of each row; else
            // int4 pairs (code ca, code cc) of 8 rows, shared by groups 2m and
            // 2m+1); load(g, ...) replaces the group's words, once the group
            // has been computed, with those of the next column
            int4 cur[RR / 4];
            auto load = [&](int g, uint32_t ca, uint32_t cc) {
                if (JOINT) {
                    const uint32_t* pj =
                        reinterpret_cast<const uint32_t*>(reinterpret_cast<const char*>(lprof) + ca);
                    cur[g] = make_int4((int)pj[(4 * g) * kJointRow], (int)pj[(4 * g + 1) * kJointRow],
                                       (int)pj[(4 * g + 2) * kJointRow], (int)pj[(4 * g + 3) * kJointRow]);
                } else if (g & 1) {
                    cur[g - 1] = lprof[ca * (RR / 8 + 1) + (g >> 1)];
                    cur[g] = lprof[cc * (RR / 8 + 1) + (g >> 1)];
                }
            };
            // JOINT: the byte offsets of the word's four columns
            auto joint_offsets = [](uint32_t wa, uint32_t wb) { return (wa << 5) | (wb << 2); };
            uint32_t nxo = JOINT ? joint_offsets(nwa, nwb) : 0u;
#pragma unroll
            for (int g = 0; g < RR / 4; ++g) {
                if (JOINT) load(g, nxo & 0xFFu, 0u);
                else load(g, nwa & 0xFFu, nwb & 0xFFu);
            }
            for (int j = 1; j <= ncols; j += 4) {
                const uint32_t cwa = nwa, cwb = nwb;
                const uint32_t cxo = nxo;
                nwa = n2a;
                nwb = n2b;
                if (JOINT) nxo = joint_offsets(nwa, nwb);
                n2a = wA[min((j + 7) >> 2, lastA)];
                n2b = wB[min((j + 7) >> 2, lastB)];
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int jj = j + q;
                    const int sh = 8 * ((q + 1) & 3);
                    // codes (JOINT: byte offset) of column jj + 1
                    const uint32_t ca = JOINT ? (((q < 3 ? cxo : nxo) >> sh) & 0xFFu)
                                              : (((q < 3 ? cwa : nwa) >> sh) & 0xFFu);
                    const uint32_t cc = JOINT ? 0u : (((q < 3 ? cwb : nwb) >> sh) & 0xFFu);
                    uint32_t tu = (uint32_t)cb[q].x, Gu = (uint32_t)cb[q].y;
                    uint32_t dt = dtb;
                    dtb = tu;
#pragma unroll
                    for (int r4 = 0; r4 < RR / 4; ++r4) {
                        uint32_t sp4[4];
                        if (JOINT) {
                            const int4 x = cur[r4];
                            sp4[0] = (uint32_t)x.x; sp4[1] = (uint32_t)x.y; sp4[2] = (uint32_t)x.z; sp4[3] = (uint32_t)x.w;
                        } else {
                            // rows 4*r4 .. 4*r4+3 of pair A (low) and pair B (high)
                            const int4 xa = cur[2 * (r4 >> 1)], xc = cur[2 * (r4 >> 1) + 1];
                            const uint32_t wa0 = (r4 & 1) ? (uint32_t)xa.z : (uint32_t)xa.x;
                            const uint32_t wa1 = (r4 & 1) ? (uint32_t)xa.w : (uint32_t)xa.y;
                            const uint32_t wc0 = (r4 & 1) ? (uint32_t)xc.z : (uint32_t)xc.x;
                            const uint32_t wc1 = (r4 & 1) ? (uint32_t)xc.w : (uint32_t)xc.y;
                            sp4[0] = __builtin_amdgcn_perm(wc0, wa0, 0x05040100u);
                            sp4[1] = __builtin_amdgcn_perm(wc0, wa0, 0x07060302u);
                            sp4[2] = __builtin_amdgcn_perm(wc1, wa1, 0x05040100u);
                            sp4[3] = __builtin_amdgcn_perm(wc1, wa1, 0x07060302u);
                        }
#pragma unroll
                        for (int k = 0; k < 4; ++k) {
                            const int r = 4 * r4 + k;
                            const s16x2 G = __builtin_elementwise_max(v2(Gu) - E2, v2(tu));
                            // the horizontal state is kept as H' = max(H, 0), which
                            // carries the 0 floor of V: an unsigned saturating
                            // subtract gives max(H' - e, 0) = max(H - e, 0), so
                            // H' = max(max(H' - e, 0), t) = max(H, 0) and
                            // V = max(f, G, H, 0) = max(f, G, H').  MX3: the
                            // floor of H' + o is o, the third operand of a max3
                            const uint32_t hs = u2(__builtin_bit_cast(s16x2, __builtin_elementwise_sub_sat(
                                __builtin_bit_cast(u16x2, Hr[r]), __builtin_bit_cast(u16x2, E2))));
                            const uint32_t H = MX3 ? max3b(hs, tr[r], u2(O2))
                                                   : u2(__builtin_elementwise_max(v2(hs), v2(tr[r])));
                            const s16x2 f = __builtin_elementwise_add_sat(v2(dt), v2(sp4[k]));
                            // G last: the row-to-row chain (t of the row above ->
                            // G -> V -> t) is three dependent operations
                            const s16x2 V = MX3 ? v2(max3b(u2(f), H, u2(G)))
                                                : __builtin_elementwise_max(__builtin_elementwise_max(f, v2(H)), G);
                            const s16x2 tn = V - O2;
                            // the row maximum: of V, or (MX3) of t = V, two
                            // columns at once after every second column
                            if (!MX3) bm[r] = u2(__builtin_elementwise_max(v2(bm[r]), V));
                            else if (q & 1) bm[r] = max3b(bm[r], tr[r], u2(tn));
                            dt = tr[r];
                            tr[r] = u2(tn);
                            Hr[r] = H;
                            tu = u2(tn);
                            Gu = u2(G);
                        }
                        load(r4, ca, cc);
                        // the loads stay here (a column ahead of their use):
                        // only ALU instructions may cross
                        __builtin_amdgcn_sched_barrier(kSchedAluOnly);
                    }
                    bnd[(int64_t)jj * kWave + lane] = make_int2((int)tu, (int)Gu);
                    cb[q] = bnd[(int64_t)(jj + 4) * kWave + lane];
                    __builtin_amdgcn_sched_barrier(kSchedAluOnly);
                }
                // after the last 4-column step of a kFirstMaxBlock-column block (or of the
                // row): the rows' block maxima, first row first, then the
                // pair's first maximum.  A mask per half is the sign of a
                // difference (no wrap: values >= -1, block maxima >= 0, rows
                // in [0, 32767]), spread to 16 bits; a bitfield insert
                // selects with it.
                static_assert(kFirstMaxBlock % 4 == 0 && (kFirstMaxBlock & (kFirstMaxBlock - 1)) == 0,
                              "a block is a power-of-two number of 4-column steps");
                if (((j - 1) & (kFirstMaxBlock - 1)) == kFirstMaxBlock - 4 || j + 4 > ncols) {
                    const uint32_t JB = splat16(((j - 1) & ~(kFirstMaxBlock - 1)) + 1);
                    uint32_t BV = bm[0], BR = 0u;
                    bm[0] = NEG2;
#pragma unroll
                    for (int r = 1; r < RR; ++r) {
                        const uint32_t d = u2(v2(BV) - v2(bm[r]));
                        uint32_t mk;  // bm[r] > BV
                        asm("v_pk_ashrrev_i16 %0, %1, %2" : "=v"(mk) : "s"(0x000F000Fu), "v"(d));
                        BR = (mk & splat16(r)) | (~mk & BR);
                        BV = u2(__builtin_elementwise_max(v2(BV), v2(bm[r])));
                        bm[r] = NEG2;
                    }
                    // (BV, row, JB) replaces (GV, GR, GJ) if BV is larger, or
                    // equal in an earlier row -- only ever a row of this strip,
                    // as earlier strips hold smaller rows (BV - GV can wrap
                    // only where GV - BV is negative, which selects by itself)
                    BR = u2(v2(BR) + v2(row1));
                    uint32_t gt, lt, rl;
                    asm("v_pk_ashrrev_i16 %0, %1, %2" : "=v"(gt) : "s"(0x000F000Fu), "v"(u2(v2(GV) - v2(BV))));
                    asm("v_pk_ashrrev_i16 %0, %1, %2" : "=v"(lt) : "s"(0x000F000Fu), "v"(u2(v2(BV) - v2(GV))));
                    asm("v_pk_ashrrev_i16 %0, %1, %2" : "=v"(rl) : "s"(0x000F000Fu), "v"(u2(v2(BR) - v2(GR))));
                    const uint32_t m = gt | (~lt & rl);
                    GR = (m & BR) | (~m & GR);
                    GJ = (m & JB) | (~m & GJ);
                    GV = u2(__builtin_elementwise_max(v2(GV), v2(BV)));
                }
            }
        };
        int i0 = 0;
        for (; La - i0 > R - 8; i0 += R) strip(StripRows<R>{}, i0);
        const int q8 = (La - i0 + 7) >> 3;  // wave-uniform: 0 .. R / 8 - 1 blocks of 8 rows left
        if (q8 == 1) strip(StripRows<8>{}, i0);
        if constexpr (R >= 24) if (q8 == 2) strip(StripRows<16>{}, i0);
        if constexpr (R >= 32) if (q8 == 3) strip(StripRows<24>{}, i0);
        if constexpr (R >= 40) if (q8 == 4) strip(StripRows<32>{}, i0);
        static_assert(R <= 40, "the last strip covers at most 40 rows");

        const int bestA = (short)(GV & 0xFFFFu), bestB = (short)(GV >> 16);
        const uint32_t cellA = ((GR & 0xFFFFu) << 16) | (GJ & 0xFFFFu), cellB = (GR & 0xFFFF0000u) | (GJ >> 16);
        if (bA >= 0) {
            const uint32_t c = (La == 0 || LbA == 0 || bestA >= sat) ? (kCellTrack | ((uint32_t)La << 16) | (uint32_t)LbA) : cellA;
            const int64_t pos = (int64_t)(a - D->r0) * D->ld + (bA - D->c0);
            T.best_cell[pos] = c;
            T.best_score[pos] = (int16_t)bestA;
        }
        if (bB >= 0) {
            const uint32_t c = (La == 0 || LbB == 0 || bestB >= sat) ? (kCellTrack | ((uint32_t)La << 16) | (uint32_t)LbB) : cellB;
            const int64_t pos = (int64_t)(a - D->r0) * D->ld + (bB - D->c0);
            T.best_cell[pos] = c;
            T.best_score[pos] = (int16_t)bestB;
        }
    }
}

// ---------------------------------------------------------------------------
// Score pass with G = 16 or 32 lanes per pair, for blocks with few row
// sequences (serving batches, small blocks): the score16_kernel recurrence,
// saturation rule and first-maximum output, laid out like lpp16_kernel.  A
// wave aligns s1 (uniform) against 2 * 64 / G sequences s2: one G-lane group
// per one (G = 16) or two (G = 32) DPP rows, two sequences per group in the
// int16 halves (pair A low, pair B high).  Lane k of a group owns R
// consecutive rows of a stripe of G * R rows and runs one column behind lane
// k-1: at step t it computes column j = t - k + 1 of its rows, taking the row
// above its first row (V - o and G of lane k-1's last row, computed at step
// t - 1) and the selector of column j's codes with a DPP shift by one lane
// (row_shr for G = 16; wave_shr for G = 32, the group's first lanes then
// taking their own values with a select), while lane 0 reads the previous
// stripe's bottom row from the boundary buffer and lane G - 1 writes this
// stripe's.  G = 32 halves a wave's stripes, for a query alone on the GPU
// (its waves then finish sooner, at the cost of taller stripe padding).
//
// Scores come from registers: per row, the (score + gap_open) of column codes
// 0..7 as the bytes of two words (the column sequences' codes are < 8 and
// their scores plus gap_open lie in [0, 255]; the row may hold any code), and
// one v_perm per row picks both halves' bytes into the int16 halves: the
// selector is the pair of column codes, min(code, 12) per half, so that a
// pad code of 12 or more selects the byte 0.  The pad code scores 0, i.e.
// -gap_open: pad columns
// (past a sequence's end), pad rows (past s1's end) and the pre-columns a
// lane runs before its first column hold values at most M - gap_open, M the
// pair's maximum over its real cells (every path into them takes a step
// scoring at most -gap_open), so they never take the first maximum; a
// pre-column cell is exactly DP column 0 (V = 0, t = -o, H' = 0).
//
// The first maximum is folded per lane once every kFirstMaxBlock steps: lane
// k's block is then the columns s - k + 1 .. s - k + kFirstMaxBlock of the
// steps s .. s + kFirstMaxBlock - 1 (the first column clamped to 1; the
// reverse pass and the re-run take a block at any column), and at the end
// the G lanes' first maxima are merged: the larger value, then the earlier
// row.

// The value lane k-1 of the lane's LN-lane group holds; `old` in the
// group's first lane.
template <int LN>
__device__ __forceinline__ uint32_t group_shr1(uint32_t old, uint32_t src, bool first)
{
    if (LN == 16) return (uint32_t)__builtin_amdgcn_update_dpp((int)old, (int)src, 0x111, 0xF, 0xF, false);
    const uint32_t v = (uint32_t)__builtin_amdgcn_update_dpp((int)old, (int)src, 0x138, 0xF, 0xF, false);  // wave_shr:1
    return first ? old : v;
}

template <int R>
__device__ __forceinline__ void lps_fold(uint32_t (&bm)[R], uint32_t& GV, uint32_t& GR, uint32_t& GJ, uint32_t row1,
                                         uint32_t JB)
{
    const uint32_t NEG2 = splat16(kPad16);
    uint32_t BV = bm[0], BR = 0u;
    bm[0] = NEG2;
#pragma unroll
    for (int r = 1; r < R; ++r) {
        uint32_t mk;  // bm[r] > BV
        asm("v_pk_ashrrev_i16 %0, %1, %2" : "=v"(mk) : "s"(0x000F000Fu), "v"(u2(v2(BV) - v2(bm[r]))));
        BR = (mk & splat16(r)) | (~mk & BR);
        BV = u2(__builtin_elementwise_max(v2(BV), v2(bm[r])));
        bm[r] = NEG2;
    }
    // (BV, row, JB) replaces (GV, GR, GJ) if BV is larger, or equal in an
    // earlier row (the lane's rows of earlier stripes are earlier rows)
    BR = u2(v2(BR) + v2(row1));
    uint32_t gt, lt, rl;
    asm("v_pk_ashrrev_i16 %0, %1, %2" : "=v"(gt) : "s"(0x000F000Fu), "v"(u2(v2(GV) - v2(BV))));
    asm("v_pk_ashrrev_i16 %0, %1, %2" : "=v"(lt) : "s"(0x000F000Fu), "v"(u2(v2(BV) - v2(GV))));
    asm("v_pk_ashrrev_i16 %0, %1, %2" : "=v"(rl) : "s"(0x000F000Fu), "v"(u2(v2(BR) - v2(GR))));
    const uint32_t m = gt | (~lt & rl);
    GR = (m & BR) | (~m & GR);
    GJ = (m & JB) | (~m & GJ);
    GV = u2(__builtin_elementwise_max(v2(GV), v2(BV)));
}

template <int R, int LN, bool MX3>
__global__ __launch_bounds__(64) void lps16_kernel(SeqSet S, ScoreTaskSet T, Scratch W)
{
    static_assert(LN == 16 || LN == 32, "16 or 32 lanes per pair");
    constexpr int NG = kWave / LN;  // lane groups per wave
    constexpr int TILE = 2 * NG;   // column sequences per task
    // [c1][c2], 64 x 8: score + gap_open, 0 for the pad code (and codes past it)
    __shared__ uint32_t tab[kMaxCodes * kJointMaxCodes / 4];
    __shared__ uint32_t mrg[3 * kWave];  // per lane: first maximum (value, row, block)

    const int lane = threadIdx.x;
    const int g = lane / LN, k = lane % LN;
    const bool first = k == 0;
    const int ncode = S.ncode, pad = ncode - 1;
    for (int x = lane; x < kMaxCodes * kJointMaxCodes; x += kWave) {
        const int c1 = x >> 3, c2 = x & 7;
        const uint32_t b = c1 < pad && c2 < pad ? (uint32_t)S.prof[c1 * ncode + c2].x & 0xFFu : 0u;
        reinterpret_cast<uint8_t*>(tab)[x] = (uint8_t)b;
    }
    __syncthreads();

    int2* __restrict__ bnd = reinterpret_cast<int2*>(W.bnd + (int64_t)blockIdx.x * W.bnd_stride);
    const BlockDesc* D = T.desc;
    const int ntasks = D->ntasks;
    if ((int)blockIdx.x >= ntasks) return;  // more workgroups than tasks: the first ntasks drain the queue
    const int o = S.gap_open, e = S.gap_ext;
    const s16x2 O2 = v2(splat16(o)), E2 = v2(splat16(e));
    const uint32_t NEG2 = splat16(kPad16), M1 = splat16(-1);
    // t = V - o, G and H' at DP row 0 / column 0 (MX3: biased by o, see
    // score16_kernel; the table's bytes are score + o >= 0 and 0 for pads)
    const uint32_t T0 = MX3 ? 0u : splat16(-o), G0 = MX3 ? 0u : NEG2, H0 = MX3 ? splat16(o) : 0u;
    const int sat = MX3 ? kSatMx3(o) : kSat16;
    const uint32_t padsel = (uint32_t)min(pad, 12) * 0x00010001u | 0x0C000C00u;
    const u16x2 C12 = __builtin_bit_cast(u16x2, splat16(12));

    for (;;) {
        int v = 0;
        if (lane == 0) v = atomicAdd(T.queue, 1);
        const int t_task = uni(v);
        if (t_task >= ntasks) break;

        const int a = uni(T.task_row[t_task]);
        const int tile = uni(T.task_tile[t_task]);
        const int maxLb = uni(T.tile_maxlen[tile]);
        const int bA = T.tile_col[(int64_t)tile * TILE + g];
        const int bB = T.tile_col[(int64_t)tile * TILE + NG + g];
        const int La = uni(S.len[a]);
        const int LbA = bA >= 0 ? S.len[bA] : 0;
        const int LbB = bB >= 0 ? S.len[bB] : 0;
        const uint8_t* s1 = S.res + S.off[a];
        const uint32_t* wA = reinterpret_cast<const uint32_t*>(S.res + (bA >= 0 ? S.off[bA] : S.pad_off));
        const uint32_t* wB = reinterpret_cast<const uint32_t*>(S.res + (bB >= 0 ? S.off[bB] : S.pad_off));
        // word (L + 3) / 4 holds only pad codes (and the words before it pad
        // codes past L): word min(k, last) gives the pad code past the end
        const int lastA = (LbA + 3) >> 2, lastB = (LbB + 3) >> 2;
        const int nsteps = (maxLb + (LN - 1) + 3) & ~3;  // lane LN - 1 reaches column maxLb

        // row 0 of the DP (t = -o, G = -inf), also past the last column so
        // that lane 0's look-ahead reads defined values
        for (int j = 1 + lane; j <= nsteps + 8; j += kWave)
            for (int q = 0; q < NG; ++q) bnd[(int64_t)j * NG + q] = make_int2((int)T0, (int)G0);
        __syncthreads();  // other lanes read these columns

        uint32_t GV = M1, GR = 0u, GJ = 0u;
        for (int i0 = 0; i0 < La; i0 += LN * R) {
            // this lane's rows: score bytes of codes 0..7 (pad rows: 0)
            uint32_t tlo[R], thi[R];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int row = i0 + k * R + r;
                const int c1 = row < La ? (int)s1[row] : pad;
                tlo[r] = tab[2 * c1];
                thi[r] = tab[2 * c1 + 1];
            }
            uint32_t tr[R], Hr[R], bm[R];
#pragma unroll
            for (int r = 0; r < R; ++r) { tr[r] = T0; Hr[r] = H0; bm[r] = NEG2; }
            const uint32_t row1 = splat16(i0 + k * R + 1);  // (1-based) row of this lane's first row
            uint32_t ot = T0, oG = G0;  // this lane's last row (t, G) of the previous step
            uint32_t pt = T0;             // the row above, received at the previous step (diagonal of row 0)
            uint32_t sel = padsel;        // this lane's column codes as a byte selector

            uint32_t nwa = wA[0], nwb = wB[0];
            int2 nb[4] = {bnd[1 * NG + g], bnd[2 * NG + g], bnd[3 * NG + g], bnd[4 * NG + g]};
            for (int t = 0; t < nsteps; t += 4) {
                const uint32_t cwa = nwa, cwb = nwb;
                const int2 cb[4] = {nb[0], nb[1], nb[2], nb[3]};
                nwa = wA[min((t >> 2) + 1, lastA)];
                nwb = wB[min((t >> 2) + 1, lastB)];
                const int64_t nxt = (int64_t)(t + 5) * NG + g;
#pragma unroll
                for (int q = 0; q < 4; ++q) nb[q] = bnd[nxt + q * NG];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int tt = t + q;
                    // lane 0: the codes of column tt + 1 (byte q of the words) as
                    // a selector of bytes (cA, 0, cB, 0); lane k: lane k-1's
                    const uint32_t s0 =
                        __builtin_bit_cast(uint32_t, __builtin_elementwise_min(__builtin_bit_cast(u16x2,
                            __builtin_amdgcn_perm(cwb, cwa, 0x0C040C00u + (uint32_t)q * 0x00010001u)), C12)) |
                        0x0C000C00u;
                    sel = group_shr1<LN>(s0, sel, first);
                    uint32_t tu = group_shr1<LN>((uint32_t)cb[q].x, ot, first);
                    uint32_t Gu = group_shr1<LN>((uint32_t)cb[q].y, oG, first);
                    uint32_t dt = pt;
                    pt = tu;
#pragma unroll
                    for (int r = 0; r < R; ++r) {
                        const uint32_t sp = __builtin_amdgcn_perm(thi[r], tlo[r], sel);
                        const s16x2 G = __builtin_elementwise_max(v2(Gu) - E2, v2(tu));
                        // H' = max(H, 0), see score16_kernel (MX3: its form)
                        const uint32_t hs = u2(__builtin_bit_cast(s16x2, __builtin_elementwise_sub_sat(
                            __builtin_bit_cast(u16x2, Hr[r]), __builtin_bit_cast(u16x2, E2))));
                        const uint32_t H = MX3 ? max3b(hs, tr[r], u2(O2))
                                               : u2(__builtin_elementwise_max(v2(hs), v2(tr[r])));
                        const s16x2 f = __builtin_elementwise_add_sat(v2(dt), v2(sp));
                        const s16x2 V = MX3 ? v2(max3b(u2(f), H, u2(G)))
                                            : __builtin_elementwise_max(__builtin_elementwise_max(f, v2(H)), G);
                        const s16x2 tn = V - O2;
                        if (!MX3) bm[r] = u2(__builtin_elementwise_max(v2(bm[r]), V));
                        else if (q & 1) bm[r] = max3b(bm[r], tr[r], u2(tn));
                        dt = tr[r];
                        tr[r] = u2(tn);
                        Hr[r] = H;
                        tu = u2(tn);
                        Gu = u2(G);
                    }
                    ot = tu;
                    oG = Gu;
                    // lane LN - 1: this stripe's last row at its column tt - (LN - 2)
                    if (k == LN - 1 && tt >= LN - 1) bnd[(int64_t)(tt - (LN - 2)) * NG + g] = make_int2((int)ot, (int)oG);
                }
                if (((t + 3) & (kFirstMaxBlock - 1)) == kFirstMaxBlock - 1 || t + 4 >= nsteps)
                    lps_fold<R>(bm, GV, GR, GJ, row1, splat16(max(1, ((t + 3) & ~(kFirstMaxBlock - 1)) - k + 1)));
            }
        }

        // merge the group's LN lanes: the larger value, then the earlier row
        mrg[lane] = GV;
        mrg[kWave + lane] = GR;
        mrg[2 * kWave + lane] = GJ;
        __syncthreads();
        if (k == 0) {
            int best[2] = {-1, -1}, brow[2] = {0, 0}, bcol[2] = {0, 0};
            for (int x = 0; x < LN; ++x) {
                const uint32_t xv = mrg[lane + x], xr = mrg[kWave + lane + x], xj = mrg[2 * kWave + lane + x];
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const int val = (short)(xv >> (16 * h)), row = (int)((xr >> (16 * h)) & 0xFFFFu);
                    if (val > best[h] || (val == best[h] && row < brow[h])) {
                        best[h] = val;
                        brow[h] = row;
                        bcol[h] = (int)((xj >> (16 * h)) & 0xFFFFu);
                    }
                }
            }
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int b = h ? bB : bA, Lb = h ? LbB : LbA;
                if (b < 0) continue;
                const uint32_t cell = (La == 0 || Lb == 0 || best[h] >= sat)
                                          ? (kCellTrack | ((uint32_t)La << 16) | (uint32_t)Lb)
                                          : (((uint32_t)brow[h] << 16) | (uint32_t)bcol[h]);
                const int64_t pos = (int64_t)(a - D->r0) * D->ld + (b - D->c0);
                T.best_cell[pos] = cell;
                T.best_score[pos] = (int16_t)best[h];
            }
        }
        __syncthreads();  // mrg is reused by the next task
    }
}

__device__ __forceinline__ int wave_max(int v)
{
#pragma unroll
    for (int s = 32; s > 0; s >>= 1) v = max(v, __shfl_xor(v, s, kWave));
    return v;
}

// ---------------------------------------------------------------------------
// Re-run of the two-pass schedule, pairs sorted by first-maximum cell.
//
// sort_count / sort_scan / sort_scatter put the block's pairs into buckets of
// similar first-maximum cell (bi, bj): flagged pairs (kCellTrack) first,
// padded to whole 64-lane tiles, then the buckets in serpentine order from
// the largest (rows of buckets from the last, alternately right-to-left and
// left-to-right), packed without padding: a tile that straddles two buckets
// holds neighbours in both coordinates, and the re-run's queue drains the
// long work first.
// Bucket coordinate of a rectangle side x >= 1: linear in steps of w, or
// (re-run rectangles, mostly small) 2-cell steps up to 64 and steps of
// about (w * kBucketDim) / 96 beyond.
__device__ __forceinline__ int side_bucket(int x, int w, bool fine)
{
    x = max(x - 1, 0);
    if (!fine) return min(x / w, kBucketDim - 1);
    if (x < 64) return x >> 1;
    const int wq = max(1, (w * kBucketDim + 95) / 96);
    return min(32 + (x - 64) / wq, kBucketDim - 1);
}

// Bucket of a first-maximum score S (the reverse pass's key): one per score
// below 96, then four per doubling, the last bucket holding the rest.
__device__ __forceinline__ int score_bucket(int s)
{
    if (s < 96) return max(s, 0);
    const int lg = 31 - __clz((unsigned)(s / 96));             // floor(log2(S / 96))
    const int sub = (int)(((int64_t)(s - (96 << lg)) * 4) / (96 << lg));  // quarter of the octave
    return min(96 + 4 * lg + sub, kBucketDim - 1);
}

// Bucket of the pair at `pos` and the DP cells of the rectangle its next pass
// evaluates (the reverse pass: bi x wend; the re-run: (bi-si+1) x (wend-sj+1);
// a flagged pair: La x Lb).
// The reverse pass's cost per pair is its band: it starts at the
// first-maximum block and runs back along the alignment for about as many
// rows as the score S allows (capped by bi and wend), wherever the block
// lies; its tiles are therefore keyed on (S, min(bi, wend)), so that the
// pairs of a wave have bands of similar length and the queue starts with
// the longest.  The re-run's cost is its rectangle, which is its key.
__device__ __forceinline__ int pair_bucket(const SortSet& P, const BlockDesc& D1, int64_t pos, int Lb,
                                           uint64_t* cells, bool* flagged)
{
    const uint32_t x = P.best_cell[pos];
    if (x & kCellTrack) {
        *flagged = true;
        *cells = (uint64_t)((x >> 16) & 0x7FFFu) * (uint64_t)Lb;
        return kFlaggedBucket;
    }
    *flagged = false;
    const int bi = (int)((x >> 16) & 0x7FFFu), jb = (int)(x & 0xFFFFu);
    const int wend = min(jb + kFirstMaxBlock - 1, Lb);
    int h = bi, w = wend;
    if (P.start_cell) {
        const uint32_t st = P.start_cell[pos];
        h = bi - (int)(st >> 16) + 1;
        w = wend - (int)(st & 0xFFFFu) + 1;
    }
    *cells = (uint64_t)h * (uint64_t)w;
    if (!P.start_cell)
        return score_bucket((int)P.best_score[pos]) * kBucketDim + side_bucket(min(bi, wend), max(D1.wr, D1.wc), false);
    if ((int64_t)*cells > P.big_rect) {
        // size class floor(log2((cells - 1) / big_rect)): (2^c, 2^(c+1)] x big_rect cells
        const int c = 63 - __clzll((*cells - 1) / (uint64_t)P.big_rect);
        return kBigBucket + min(c, kBigClasses - 1);
    }
    return side_bucket(h, D1.wr, true) * kBucketDim + side_bucket(w, D1.wc, true);
}

// Slot order: rank 0 = the flagged list, ranks 1 .. kBigClasses = the big
// rectangles, largest class first, then the serpentine.
__device__ __forceinline__ int bucket_at_rank(int q)
{
    if (q == 0) return kFlaggedBucket;
    if (q <= kBigClasses) return kBigBucket + kBigClasses - q;
    const int rr = q - 1 - kBigClasses, band = rr / kBucketDim, pos = rr - band * kBucketDim;
    const int qi = kBucketDim - 1 - band;
    const int qj = (band & 1) ? pos : kBucketDim - 1 - pos;
    return qi * kBucketDim + qj;
}

// Pairs of the block: row `row` of the score-pass desc holds columns
// 0 .. ncols(row)-1 (a lower-mirror diagonal block only j < i).
__device__ __forceinline__ int64_t row_pairs(const BlockDesc& D1, bool lower, int row)
{
    if (!lower) return D1.ld;
    const int64_t n = D1.r0 + row - D1.c0;
    return n < 0 ? 0 : (n < D1.ld ? n : D1.ld);
}


__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v)
{
#pragma unroll
    for (int s = 32; s > 0; s >>= 1) {
        const uint32_t lo = __shfl_xor((uint32_t)v, s, kWave), hi = __shfl_xor((uint32_t)(v >> 32), s, kWave);
        v += ((uint64_t)hi << 32) | lo;
    }
    return v;
}

// Also accumulates, for the accounting of the next pass, the cells of the
// pairs' rectangles (pair_bucket) and of the flagged pairs.
// Counts per bucket are gathered in LDS per workgroup and added to the
// global counts once per non-empty bucket: most re-run rectangles are small
// and fall into a few hundred buckets, so global atomics per pair would
// serialise on them.
constexpr int kSortThreads = 512;

// The sort's work items: a row's pairs in chunks of kSortChunk, so that a
// block of a few rows (a serving query against its panel) spreads over the
// GPU's workgroups; fn(row, column) for each pair of this workgroup's items,
// in the same order for every call within a workgroup.
constexpr int kSortChunk = 2048;
template <typename F>
__device__ __forceinline__ void for_each_sort_pair(const BlockDesc& D1, bool lower, F&& fn)
{
    const int64_t nchunk = (D1.ld + kSortChunk - 1) / kSortChunk;
    const int64_t nitems = (int64_t)D1.rows * nchunk;
    for (int64_t it = blockIdx.x; it < nitems; it += (int64_t)gridDim.x) {
        const int row = (int)(it / nchunk);
        const int64_t c_lo = (it - (int64_t)row * nchunk) * kSortChunk;
        const int64_t n = row_pairs(D1, lower, row);
        const int64_t c_hi = n < c_lo + kSortChunk ? n : c_lo + kSortChunk;
        for (int64_t c = c_lo + threadIdx.x; c < c_hi; c += kSortThreads) fn(row, c);
    }
}

__global__ __launch_bounds__(kSortThreads) void sort_count_kernel(SortSet P)
{
    __shared__ int hist[kNumBuckets];
    const BlockDesc D1 = P.desc[1];
    const bool lower = P.desc[0].mirror != 0;
    for (int k = threadIdx.x; k < kNumBuckets; k += kSortThreads) hist[k] = 0;
    __syncthreads();
    uint64_t rect = 0, flagged = 0, big = 0;
    for_each_sort_pair(D1, lower, [&](int row, int64_t c) {
        uint64_t cells;
        bool fl;
        const int k = pair_bucket(P, D1, (int64_t)row * D1.ld + c, P.len[D1.c0 + c], &cells, &fl);
        atomicAdd(&hist[k], 1);
        if (fl) flagged += cells;
        else if (k >= kBigBucket && k < kBigBucket + kBigClasses) big += cells;
        else rect += cells;
    });
    __syncthreads();
    for (int k = threadIdx.x; k < kNumBuckets; k += kSortThreads)
        if (hist[k]) atomicAdd(&P.count[k], hist[k]);
    // the cell totals: summed over the workgroup first, then one atomic per
    // workgroup and total, none for a zero (every wave of every workgroup
    // adding to the same three words serialised on them: 165 us of a
    // 20,000-pair block's sort)
    __shared__ uint64_t wsum[3][kSortThreads / kWave];
    rect = wave_sum_u64(rect);
    flagged = wave_sum_u64(flagged);
    big = wave_sum_u64(big);
    if ((threadIdx.x & (kWave - 1)) == 0) {
        wsum[0][threadIdx.x / kWave] = rect;
        wsum[1][threadIdx.x / kWave] = flagged;
        wsum[2][threadIdx.x / kWave] = big;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        uint64_t t = 0;
        for (int w = 0; w < kSortThreads / kWave; ++w) t += wsum[threadIdx.x][w];
        unsigned long long* m = reinterpret_cast<unsigned long long*>(P.meta);
        const int slot = threadIdx.x == 0 ? kMetaRectCells : threadIdx.x == 1 ? kMetaFlaggedCells : kMetaBigCells;
        if (t) atomicAdd(&m[slot], (unsigned long long)t);
    }
}

// Pairs in the big list (all its size classes).
__device__ __forceinline__ int big_count(const SortSet& P)
{
    int n = 0;
#pragma unroll
    for (int c = 0; c < kBigClasses; ++c) n += P.count[kBigBucket + c];
    return n;
}

__device__ __forceinline__ int whole_tiles(int n) { return (n + kWave - 1) / kWave * kWave; }

// Slots of rank q: the flagged list and the big list are padded to whole
// tiles (the big list after its last, smallest, class), the buckets are not.
__device__ __forceinline__ int rank_slots(const SortSet& P, int q)
{
    const int n = P.count[bucket_at_rank(q)];
    if (q == 0) return whole_tiles(n);
    if (q == kBigClasses) {
        const int nb = big_count(P);
        return n + whole_tiles(nb) - nb;
    }
    return n;
}

// One workgroup of kScanThreads: exclusive scan of the slot counts in rank
// order, kScanPer consecutive ranks per thread (their counts loaded at once,
// then a scan over the threads: waves by shuffles, the wave totals in LDS).
constexpr int kScanThreads = 1024;
constexpr int kScanPer = (kNumBuckets + kScanThreads - 1) / kScanThreads;

__global__ __launch_bounds__(kScanThreads) void sort_scan_kernel(SortSet P)
{
    __shared__ int wsum[kScanThreads / kWave];
    const int tid = threadIdx.x, lane = tid & (kWave - 1), w = tid / kWave;
    const int q0 = tid * kScanPer;
    int cnt[kScanPer];
    int sum = 0;
#pragma unroll
    for (int u = 0; u < kScanPer; ++u) {
        cnt[u] = q0 + u < kNumBuckets ? rank_slots(P, q0 + u) : 0;
        sum += cnt[u];
    }
    // inclusive prefix over the wave's lanes, then over the waves
    int incl = sum;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const int y = __shfl_up(incl, d, kWave);
        if (lane >= d) incl += y;
    }
    if (lane == kWave - 1) wsum[w] = incl;
    __syncthreads();
    int before = 0, total = 0;
    for (int x = 0; x < kScanThreads / kWave; ++x) {
        before += x < w ? wsum[x] : 0;
        total += wsum[x];
    }
    int slot = before + incl - sum;
#pragma unroll
    for (int u = 0; u < kScanPer; ++u) {
        if (q0 + u < kNumBuckets) {
            const int k = bucket_at_rank(q0 + u);
            P.base[k] = slot;
            P.cursor[k] = 0;
        }
        slot += cnt[u];
    }
    if (tid == 0) {
        const int tf = (P.count[kFlaggedBucket] + kWave - 1) / kWave;
        P.meta[kMetaFlaggedTiles] = tf;
        P.meta[kMetaBigTiles] = tf + whole_tiles(big_count(P)) / kWave;
        P.meta[kMetaTiles] = (total + kWave - 1) / kWave;
        P.meta[kMetaSlots] = total;
    }
}

// Each workgroup counts its pairs per bucket in LDS, reserves one slot range
// per non-empty bucket, then places its pairs with LDS atomics.
__global__ __launch_bounds__(kSortThreads) void sort_scatter_kernel(SortSet P)
{
    __shared__ int hist[kNumBuckets];
    const BlockDesc D1 = P.desc[1];
    const bool lower = P.desc[0].mirror != 0;
    // idle lanes: the last tiles of the flagged and big lists, and the last tile
    if (blockIdx.x == 0 && threadIdx.x < kWave) {
        const int nf = P.count[kFlaggedBucket];
        const int endf = (nf + kWave - 1) / kWave * kWave;
        if (nf + (int)threadIdx.x < endf) P.sorted[nf + threadIdx.x] = ~0u;
        const int nbig = big_count(P);  // the big list starts where the flagged list's tiles end
        const int nb = endf + nbig;
        const int endb = endf + whole_tiles(nbig);
        if (nb + (int)threadIdx.x < endb) P.sorted[nb + threadIdx.x] = ~0u;
        const int n = P.meta[kMetaSlots];
        const int end = (n + kWave - 1) / kWave * kWave;
        if (n + (int)threadIdx.x < end) P.sorted[n + threadIdx.x] = ~0u;
    }
    for (int k = threadIdx.x; k < kNumBuckets; k += kSortThreads) hist[k] = 0;
    __syncthreads();
    for_each_sort_pair(D1, lower, [&](int row, int64_t c) {
        uint64_t cells;
        bool fl;
        atomicAdd(&hist[pair_bucket(P, D1, (int64_t)row * D1.ld + c, P.len[D1.c0 + c], &cells, &fl)], 1);
    });
    __syncthreads();
    for (int k = threadIdx.x; k < kNumBuckets; k += kSortThreads)
        if (hist[k]) hist[k] = P.base[k] + atomicAdd(&P.cursor[k], hist[k]);
    __syncthreads();
    for_each_sort_pair(D1, lower, [&](int row, int64_t c) {
        uint64_t cells;
        bool fl;
        const int64_t pos = (int64_t)row * D1.ld + c;
        const int k = pair_bucket(P, D1, pos, P.len[D1.c0 + c], &cells, &fl);
        P.sorted[atomicAdd(&hist[k], 1)] = (uint32_t)pos;
    });
}

// Each lane aligns its own pair: s1 differs per lane, so the substitution
// scores come from the [c1][c2] profile table staged in LDS.  A lane's rows
// are bottom-aligned to the wave's last row Rmax (max bi rounded up to whole
// strips): its first Rmax - bi rows are pad rows, whose cells equal the DP's
// row 0 (V = 0, T = 0, and G = -gap_open below them; this needs gap_open >= 0
// and gap_ext >= 0), so every lane's row bi is the last strip's bottom row,
// which the boundary buffer holds afterwards.  The score pass gives the first
// column jb of the kFirstMaxBlock-column block holding the first maximum of
// row bi, which is the maximum of that row: its first strict maximum in that
// block is the first maximum (columns past the sequence end hold pad cells,
// never a new maximum).  Every value in the rectangle
// [1..bi] x [1..jb+kFirstMaxBlock-1] depends only on cells above and
// to the left, so it equals the full DP's, and the T word captured there is
// that of the reference's traceback from its first maximum
// (SmithWatermanGotoh.cs:211-229).
// TRACK (flagged pairs): bi = La, bj = Lb and the first maximum is tracked as
// in sw_kernel; pad rows hold V = 0, which can only be a first maximum when
// every cell is 0, and then T = 0 as at (1, 1).
template <int R, bool TRACK>
__global__ __launch_bounds__(64) void rerun_kernel(SeqSet S, RerunSet T, Scratch W)
{
    extern __shared__ int2 tab[];  // [ncode][ncode]
    const int lane = threadIdx.x;
    const int ncode = S.ncode, pad = ncode - 1;
    for (int k = lane; k < ncode * ncode; k += kWave) tab[k] = S.prof[k];
    __syncthreads();

    int4* __restrict__ bnd = W.bnd + (int64_t)blockIdx.x * W.bnd_stride;
    const BlockDesc* D = T.desc;
    const int64_t r0 = D[1].r0, c0 = D[1].c0;
    const uint32_t ld1 = (uint32_t)D[1].ld;
    const int o = S.gap_open, e = S.gap_ext;
    const int ginc = S.gap_inc;  // count-word increment per gap column
    const int t_lo = TRACK ? 0 : T.meta[kMetaBigTiles];
    const int t_hi = TRACK ? T.meta[kMetaFlaggedTiles] : T.meta[kMetaTiles];
    if ((int)blockIdx.x >= t_hi - t_lo) return;  // more workgroups than tiles: the first ones drain the queue
    // cells evaluated, added to meta once after the task loop: a second
    // lane-0 branch inside the loop lets the compiler merge it with the
    // queue's and leave lanes 1..63 in the loop without lane 0
    uint64_t run = 0;

    for (;;) {
        int v = 0;
        if (lane == 0) v = atomicAdd(T.queue, 1);
        const int t = uni(v) + t_lo;
        if (t >= t_hi) break;

        const uint32_t p = T.sorted[(int64_t)t * kWave + lane];
        const bool live = p != ~0u;
        // the lane's rectangle: rows si..bi of s1, columns sj..sj+w-1 of s2
        int a = 0, b = 0, La = 0, Lb = 0, bi = 0, bj = 0, si = 1, sj = 1, w = 0, h = 0;
        if (live) {
            a = (int)(r0 + p / ld1);
            b = (int)(c0 + p % ld1);
            La = S.len[a];
            Lb = S.len[b];
            if (TRACK) {
                bi = La; bj = Lb; w = Lb;
            } else {
                const uint32_t cell = T.best_cell[p];
                bi = (int)((cell >> 16) & 0x7FFFu);
                bj = (int)(cell & 0xFFFFu);
                if (T.start_cell) {
                    const uint32_t st = T.start_cell[p];
                    si = (int)(st >> 16);
                    sj = (int)(st & 0xFFFFu);
                }
                w = min(bj + kFirstMaxBlock - 1, Lb) - sj + 1;
            }
            h = bi - si + 1;
        }
        // capture: the rows end at a strip boundary, so row bi is the last
        // strip's bottom row, which the boundary buffer holds afterwards
        const int Rmax0 = uni(wave_max(h)), Cmax = uni(wave_max(w));
        const int Rmax = TRACK ? Rmax0 : (Rmax0 + R - 1) / R * R;
        const int ro = Rmax - h;  // pad rows before the lane's first row
        const uint8_t* s1 = S.res + (live ? S.off[a] + si - 1 : S.pad_off);
        // s2 from column sj: 4-byte words from the aligned address below it,
        // shifted by the misalignment
        const int64_t c2b = live ? S.off[b] + sj - 1 : S.pad_off;
        const uint32_t* s2w = reinterpret_cast<const uint32_t*>(S.res + (c2b & ~int64_t(3)));
        const int sh = (int)(c2b & 3);
        const int ncols = (Cmax + 3) & ~3;

        for (int j = 1; j <= ncols; ++j) bnd[(int64_t)j * kWave + lane] = make_int4(-o, kNeg, 0, 0);

        int best = INT_MIN, bestT = 0;
        for (int i0 = 0; i0 < Rmax; i0 += R) {
            int rowb[R];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int k = i0 + r - ro;
                rowb[r] = ((k >= 0 && k < h) ? (int)s1[k] : pad) * ncode;
            }
            int tr[R], Tr[R], Hr[R], THr[R];
            int rb[TRACK ? R : 1], rT[TRACK ? R : 1];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                tr[r] = -o; Tr[r] = 0; Hr[r] = 0; THr[r] = 0;
                if (TRACK) { rb[r] = INT_MIN; rT[r] = 0; }
            }
            int dtb = -o, dTb = 0;

            // residue words (4 columns each): current, next, and the one after
            uint32_t a1 = s2w[1], a2 = s2w[2];
            uint32_t nw = __builtin_amdgcn_alignbyte(a1, s2w[0], sh);
            int4 nb0 = bnd[1 * kWave + lane], nb1 = bnd[2 * kWave + lane];
            int4 nb2 = bnd[3 * kWave + lane], nb3 = bnd[4 * kWave + lane];
            // the table entries of a column are loaded one column ahead
            int2 cur[R];
            {
                const int c2 = w >= 1 ? (int)(nw & 0xFFu) : pad;
#pragma unroll
                for (int r = 0; r < R; ++r) cur[r] = tab[rowb[r] + c2];
            }
            uint32_t n2 = __builtin_amdgcn_alignbyte(a2, a1, sh);  // the block after next
            for (int j = 1; j <= ncols; j += 4) {
                const uint32_t wd = nw;
                int4 cb[4] = {nb0, nb1, nb2, nb3};
                nw = n2;
                a1 = a2;
                a2 = s2w[((j + 7) >> 2) + 1];
                n2 = __builtin_amdgcn_alignbyte(a2, a1, sh);
                const int64_t nxt = (int64_t)(j + 4) * kWave + lane;
                nb0 = bnd[nxt]; nb1 = bnd[nxt + kWave]; nb2 = bnd[nxt + 2 * kWave]; nb3 = bnd[nxt + 3 * kWave];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int jj = j + q;
                    int c2n = (int)(((q < 3 ? wd : nw) >> (8 * ((q + 1) & 3))) & 0xFFu);
                    c2n = jj + 1 > w ? pad : c2n;
                    int2 nxtp[R];
#pragma unroll
                    for (int r = 0; r < R; ++r) nxtp[r] = tab[rowb[r] + c2n];
                    // keep the look-ahead loads ahead of this column's cells
                    __builtin_amdgcn_sched_barrier(0);
                    int tu = cb[q].x, Gu = cb[q].y, Tu = cb[q].z, TGu = cb[q].w;
                    int dt = dtb, dT = dTb;
                    dtb = tu; dTb = Tu;
#pragma unroll
                    for (int r = 0; r < R; ++r) {
                        const int sc = cur[r].x, inc = cur[r].y;
                        const int g1 = Gu - e;
                        const bool cg = g1 > tu;
                        const int G = cg ? g1 : tu;
                        const int TG = (cg ? TGu : Tu) + ginc;
                        // H is kept as max(H, 0) (an unsigned saturating subtract): V = max(f, G, H, 0)
                        // = max(f, G, H'); the gap run it carries matters only where H >= 0, where H' = H
                        const int h1 = (int)__builtin_elementwise_sub_sat((unsigned)Hr[r], (unsigned)e);
                        const bool ch = h1 > tr[r];
                        Hr[r] = ch ? h1 : tr[r];
                        THr[r] = (ch ? THr[r] : Tr[r]) + ginc;
                        const int f = dt + sc;
                        int V = max(max(f, G), Hr[r]);
                        const int Td = dT + inc;
                        int x = (V == G) ? TG : THr[r];
                        x = (V == f) ? Td : x;
                        const int Tn = (V == 0) ? 0 : x;
                        if (TRACK) {
                            const bool cb2 = V > rb[r];
                            rb[r] = cb2 ? V : rb[r];
                            rT[r] = cb2 ? Tn : rT[r];
                        }
                        const int tn = V - o;
                        dt = tr[r]; dT = Tr[r];
                        tr[r] = tn; Tr[r] = Tn;
                        tu = tn; Gu = G; Tu = Tn; TGu = TG;
                    }
                    bnd[(int64_t)jj * kWave + lane] = make_int4(tu, Gu, Tu, TGu);
#pragma unroll
                    for (int r = 0; r < R; ++r) cur[r] = nxtp[r];
                }
            }
            if (TRACK) {
#pragma unroll
                for (int r = 0; r < R; ++r)
                    if (i0 + r < Rmax && rb[r] > best) { best = rb[r]; bestT = rT[r]; }
            }
        }

        if (!TRACK) {
            // first strict maximum of row bi in columns jb .. jb + kFirstMaxBlock - 1
            // (rectangle columns jb - sj + 1 .., those inside it)
#pragma unroll
            for (int q = 0; q < kFirstMaxBlock; ++q) {
                const int col = bj - sj + 1 + q;
                if (col >= 1 && col <= w) {
                    const int4 x = bnd[(int64_t)col * kWave + lane];
                    const int V = x.x + o;
                    if (V > best) { best = V; bestT = x.z; }
                }
            }
        }
        if (live) {
            int score = best, Tb = bestT;
            if (La == 0 || Lb == 0) { score = INT_MIN; Tb = 0; }
            write_distance(S, D[0], a, b, score, Tb);
        }
        run += (uint64_t)Rmax * (uint64_t)ncols * kWave;
    }
    if (lane == 0 && run)
        atomicAdd(reinterpret_cast<unsigned long long*>(T.meta) + (TRACK ? kMetaFlaggedRun : kMetaCaptureRun),
                  (unsigned long long)run);
}

// ---------------------------------------------------------------------------
// Re-run of the big rectangles (more than SortSet::big_rect cells), one wave per pair.
// One lane per pair would leave a single lane with millions of cells while the
// rest of the GPU idles.  Lane k owns R consecutive rows of a stripe of 64*R
// rows of the rectangle and runs one column behind lane k-1 (anti-diagonal
// wavefront, as lpp16_kernel with the whole wave on one pair): the row above
// its first row (V-o, G, T, TG) and the column residue come from lane k-1
// with a DPP wave shift, lane 0 reads the previous stripe's bottom row from
// the boundary buffer and lane 63 writes this stripe's.  Pre-columns and pad
// rows/columns as in lpp16_kernel.
// The local DP of the rectangle [si..bi] x [sj..wend] is run with its first
// strict maximum tracked over all of it: the rectangle's values are at most
// the full DP's and equal to them along the reference's traceback path, so
// every cell before (bi, bj) in row-major order stays below S and the first
// maximum is (bi, bj), with the reference's statistics.
constexpr int kRectLanes = 64;

__device__ __forceinline__ int wave_shr1(int old, int src)
{
    return __builtin_amdgcn_update_dpp(old, src, 0x138, 0xF, 0xF, false);  // wave_shr:1
}

template <int R>
__global__ __launch_bounds__(64) void rect64_kernel(SeqSet S, RerunSet T, Scratch W)
{
    constexpr int H = kRectLanes * R;  // stripe height
    constexpr int PS = H / 2 + 1;      // int4 per residue code in the stripe profile
    extern __shared__ int4 dyn_lds[];
    int4* lprof = dyn_lds;                                           // ncode * PS int4
    int2* lmerge = reinterpret_cast<int2*>(dyn_lds + S.ncode * PS);  // H (best, T)

    const int lane = threadIdx.x;
    int4* __restrict__ bnd = W.bnd + (int64_t)blockIdx.x * W.bnd_stride;  // [column]
    const BlockDesc* D = T.desc;
    const int64_t r0 = D[1].r0, c0 = D[1].c0;
    const uint32_t ld1 = (uint32_t)D[1].ld;
    const int o = S.gap_open, e = S.gap_ext;
    const int ginc = S.gap_inc;  // count-word increment per gap column
    const int pad = S.ncode - 1;
    const int64_t s_lo = (int64_t)T.meta[kMetaFlaggedTiles] * kWave;
    const int64_t s_hi = (int64_t)T.meta[kMetaBigTiles] * kWave;
    if ((int64_t)blockIdx.x >= s_hi - s_lo) return;  // more workgroups than pairs: the first ones drain the queue
    uint64_t run = 0;  // cells evaluated, added to meta once after the task loop

    for (;;) {
        int v = 0;
        if (lane == 0) v = atomicAdd(T.queue, 1);
        const int64_t slot = s_lo + uni(v);
        if (slot >= s_hi) break;
        const uint32_t p = (uint32_t)uni((int)T.sorted[slot]);
        if (p == ~0u) continue;  // the list's padding

        const int a = (int)(r0 + p / ld1), b = (int)(c0 + p % ld1);
        const uint32_t cell = (uint32_t)uni((int)T.best_cell[p]), st = (uint32_t)uni((int)T.start_cell[p]);
        const int bi = (int)((cell >> 16) & 0x7FFFu), jb = (int)(cell & 0xFFFFu);
        const int si = (int)(st >> 16), sj = (int)(st & 0xFFFFu);
        const int wend = min(jb + kFirstMaxBlock - 1, uni(S.len[b]));
        const int h = bi - si + 1, w = wend - sj + 1;
        const uint8_t* s1 = S.res + S.off[a] + si - 1;
        const int64_t c2b = S.off[b] + sj - 1;
        const uint32_t* s2w = reinterpret_cast<const uint32_t*>(S.res + (c2b & ~int64_t(3)));
        const int sh = (int)(c2b & 3);
        const int nsteps = (w + (kRectLanes - 1) + 3) & ~3;  // lane 63 reaches column w

        // row 0 of the rectangle (V = 0, G = -inf, T = 0), also past the last
        // column so that lane 0 always reads defined values
        for (int j = 1 + lane; j <= nsteps + 8; j += kWave) bnd[j] = make_int4(-o, kNeg, 0, 0);
        __syncthreads();  // other lanes read these columns

        int best = INT_MIN, bestT = 0;
        for (int i0 = 0; i0 < h; i0 += H) {
            __syncthreads();
            {
                int2* lp = reinterpret_cast<int2*>(lprof);
                for (int idx = lane; idx < S.ncode * H; idx += kWave) {
                    const int c2 = idx / H, r = idx % H;
                    const int row = i0 + r;
                    const int c1 = row < h ? (int)s1[row] : pad;
                    lp[c2 * (H + 2) + r] = S.prof[c1 * S.ncode + c2];
                }
            }
            __syncthreads();

            int tr[R], Tr[R], Hr[R], THr[R], rb[R], rT[R];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                tr[r] = -o; Tr[r] = 0; Hr[r] = 0; THr[r] = 0; rb[r] = INT_MIN; rT[r] = 0;
            }
            int ot = -o, oG = kNeg, oT = 0, oTG = 0;  // this lane's last-row output of the previous step
            int pt = -o, pT = 0;                      // up-values received at the previous step (diagonal)
            int cprev = pad;

            uint32_t wlo = s2w[0], whi = s2w[1];
            int4 nb0 = bnd[1], nb1 = bnd[2], nb2 = bnd[3], nb3 = bnd[4];
            for (int t = 0; t < nsteps; t += 4) {
                const uint32_t cw = __builtin_amdgcn_alignbyte(whi, wlo, sh);  // columns t+1 .. t+4
                wlo = whi;
                whi = s2w[(t >> 2) + 2];
                int4 cb[4] = {nb0, nb1, nb2, nb3};
                nb0 = bnd[t + 5]; nb1 = bnd[t + 6]; nb2 = bnd[t + 7]; nb3 = bnd[t + 8];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int tt = t + q;
                    int cl = (int)((cw >> (8 * q)) & 0xFFu);
                    cl = tt + 1 > w ? pad : cl;
                    const int c2 = wave_shr1(cl, cprev);
                    cprev = c2;
                    int tu = wave_shr1(cb[q].x, ot), Gu = wave_shr1(cb[q].y, oG);
                    int Tu = wave_shr1(cb[q].z, oT), TGu = wave_shr1(cb[q].w, oTG);
                    int dt = pt, dT = pT;
                    pt = tu; pT = Tu;
                    const int4* pr = lprof + c2 * PS + lane * (R / 2);
#pragma unroll
                    for (int r2 = 0; r2 < R / 2; ++r2) {
                        const int4 pp = pr[r2];
#pragma unroll
                        for (int hh = 0; hh < 2; ++hh) {
                            const int r = 2 * r2 + hh;
                            const int sc = hh ? pp.z : pp.x;
                            const int inc = hh ? pp.w : pp.y;
                            const int g1 = Gu - e;
                            const bool cg = g1 > tu;
                            const int G = cg ? g1 : tu;
                            const int TG = (cg ? TGu : Tu) + ginc;
                            // H is kept as max(H, 0) (an unsigned saturating subtract): V = max(f, G, H, 0)
                            // = max(f, G, H'); the gap run it carries matters only where H >= 0, where H' = H
                            const int h1 = (int)__builtin_elementwise_sub_sat((unsigned)Hr[r], (unsigned)e);
                            const bool ch = h1 > tr[r];
                            Hr[r] = ch ? h1 : tr[r];
                            THr[r] = (ch ? THr[r] : Tr[r]) + ginc;
                            const int f = dt + sc;
                            int V = max(max(f, G), Hr[r]);
                            const int Td = dT + inc;
                            int x = (V == G) ? TG : THr[r];
                            x = (V == f) ? Td : x;
                            const int Tn = (V == 0) ? 0 : x;
                            const bool cb2 = V > rb[r];
                            rb[r] = cb2 ? V : rb[r];
                            rT[r] = cb2 ? Tn : rT[r];
                            const int tn = V - o;
                            dt = tr[r]; dT = Tr[r];
                            tr[r] = tn; Tr[r] = Tn;
                            tu = tn; Gu = G; Tu = Tn; TGu = TG;
                        }
                    }
                    ot = tu; oG = Gu; oT = Tu; oTG = TGu;
                    const int j = tt - lane + 1;
                    if (lane == kRectLanes - 1 && j >= 1) bnd[j] = make_int4(ot, oG, oT, oTG);
                }
            }
            // first strict maximum in row-major order over the stripe's rows
#pragma unroll
            for (int r = 0; r < R; ++r) lmerge[lane * R + r] = make_int2(rb[r], rT[r]);
            __syncthreads();
            if (lane == 0) {
                for (int r = 0; r < H && i0 + r < h; ++r) {
                    const int2 m = lmerge[r];
                    if (m.x > best) { best = m.x; bestT = m.y; }
                }
            }
            run += (uint64_t)H * (uint64_t)nsteps;
        }
        if (lane == 0) write_distance(S, D[0], a, b, best, bestT);
    }
    if (lane == 0 && run) atomicAdd(reinterpret_cast<unsigned long long*>(T.meta) + kMetaBigRun, (unsigned long long)run);
}

// ---------------------------------------------------------------------------
// Reverse pass: where can an optimal alignment ending in the first-maximum
// block start?  Per pair (one per lane), the DP of the reversed prefixes
// s1[1..bi] and s2[1..wend], wend = min(jb + kFirstMaxBlock - 1, Lb), anchored
// at the block: reverse row 1 is s1[bi], reverse column 1 is s2[wend]; the
// border cell (0, c) is 0 for c < wend - jb + 1 (the block's columns) and
// -inf elsewhere, and there is no 0 floor.  V'(i', c') is then the best score
// of an alignment path from the forward cell (bi - i' + 1, wend - c' + 1) to a
// cell of the block.  Every forward V is at most the first-maximum score S, so
// V' <= S, and the start of the reference's traceback path reaches S: it lies
// in a reverse row and column at most the last ones where V' reaches S.  With
// si = bi - hit_row + 1 and sj = wend - hit_col + 1, the rectangle
// [si..bi] x [sj..wend] holds every optimal path into the block, and the local
// DP over it (0 border) gives the reference's statistics: on the path each
// value equals the full DP's (its prefix lies in the rectangle and no value can
// exceed the full DP's), and so does every comparison the traceback depends on
// there (the diagonal candidate of a DIAGONAL cell and the gap chain of a gap
// run lie on the path; the other candidates can only be lower).
// No hit (only possible when S = 0): the whole prefix rectangle.
// Row cap: the pass stops after T.row_cap reverse rows; a pair whose band is
// still live there could start anywhere above, so it gets its whole prefix
// [1..bi] x [1..wend] as rectangle -- always a valid one (its values are the
// full DP's).  This bounds the pass's longest tasks, the pairs with the
// longest alignments, whose rectangle is often their prefix anyway.
//
// Two pairs per lane in the int16 halves of one register (v_pk_* with
// saturation): pair A = slot 128t + lane, pair B = slot 128t + 64 + lane of
// the sorted list.  Saturation cannot create a hit: a path through a value
// clamped at -32768 would need a segment scoring >= 32768 + S to climb back
// to S, more than any local alignment of the pair scores (S < 32766).
// Each lane's two pairs have their own row sequences, so the scores come from
// a table indexed by residue codes: for alphabets of at most
// kRevJointMaxCodes codes (JOINT) a joint table jt[ra][rb][ca][cb] holds both
// halves' scores packed, one 32-bit read per row and column; otherwise the
// [c1][c2] table, one 16-bit read per half and a pack.
constexpr int kRevJointMaxCodes = 7;  // 7^4 words = 9.6 KB of LDS

// LDS of the reverse pass: the [c1][c2] table (int16), padded to a word,
// then the joint table
size_t reverse_lds_bytes(int ncode)
{
    const size_t n2 = (size_t)ncode * ncode;
    return ((n2 + 1) & ~size_t(1)) * sizeof(short) + (ncode <= kRevJointMaxCodes ? n2 * n2 * sizeof(uint32_t) : 0);
}

// Waves per SIMD of the reverse pass: three for 8-row strips with the joint
// table (158 VGPRs), else two.
constexpr int reverse_waves(int R, bool joint) { return R == 8 && joint ? 3 : 2; }

// The max3 form of the reverse pass (MX3): every value biased by kRevBias
// (-inf = 0), kept in [0, 0x7BFF] with unsigned saturating subtracts, the
// pad score -32768 (a saturating add leaves a negative pattern, below every
// other operand of a max3), three-way maxima as v_pk_maximum3_f16 on the
// int16 patterns (see score16_kernel), and the row and column maxima taken
// in the t = V' - o domain, two columns (rows) per instruction.  A value
// clamped at 0 stands for a true value below -kRevBias; along the band's at
// most row_cap rows a path climbs by at most row_cap times the largest
// score, which the host keeps below kRevBias - max(o, e), so nothing derived
// from a clamped value reaches the live threshold min(0, o - e) or S: every
// live cell is computed exactly.  A pair whose S + kRevBias passes 0x7BFF
// takes its whole prefix as rectangle (always valid).
template <int R, bool JOINT, bool MX3>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(reverse_waves(R, JOINT), reverse_waves(R, JOINT))))
void reverse_kernel(SeqSet S, ReverseSet T, Scratch W)
{
    extern __shared__ short tab16[];  // [ncode][ncode]: score + gap_open, clamped to int16; then jt
    const int lane = threadIdx.x;
    const int ncode = S.ncode, pad = ncode - 1, ncode2 = ncode * ncode;
    for (int k = lane; k < ncode2; k += kWave) {
        const int v = S.prof[k].x;
        tab16[k] = MX3 ? (short)(v < 0 ? -32768 : v) : (short)(v < kPad16 ? kPad16 : v);
    }
    uint32_t* jt = reinterpret_cast<uint32_t*>(tab16 + ((ncode2 + 1) & ~1));
    if (JOINT) {
        __syncthreads();
        for (int k = lane; k < ncode2 * ncode2; k += kWave) {
            const int rr = k / ncode2, cc = k - rr * ncode2;   // rr = ra * ncode + rb, cc = ca * ncode + cb
            const int ra = rr / ncode, rb = rr - ra * ncode, ca = cc / ncode, cb = cc - ca * ncode;
            jt[k] = (uint32_t)(uint16_t)tab16[ra * ncode + ca] | ((uint32_t)(uint16_t)tab16[rb * ncode + cb] << 16);
        }
    }
    __syncthreads();

    int2* __restrict__ bnd = reinterpret_cast<int2*>(W.bnd + (int64_t)blockIdx.x * W.bnd_stride);
    const BlockDesc* D = T.desc;
    const int64_t r0 = D[1].r0, c0 = D[1].c0;
    const uint32_t ld1 = (uint32_t)D[1].ld;
    const int o = S.gap_open, e = S.gap_ext;
    const s16x2 O2 = v2(splat16(o)), E2 = v2(splat16(e));
    // -inf and t = V' - o of V' = 0 (MX3: biased)
    const uint32_t NEG2 = MX3 ? 0u : 0x80008000u, T0 = MX3 ? splat16(kRevBias - o) : splat16(-o);
    const int64_t slot0 = (int64_t)T.meta[kMetaFlaggedTiles] * kWave, slot1 = (int64_t)T.meta[kMetaTiles] * kWave;
    const int ntasks = (int)((slot1 - slot0 + 2 * kWave - 1) / (2 * kWave));
    if ((int)blockIdx.x >= ntasks) return;  // more workgroups than tasks: the first ntasks drain the queue
    uint64_t run = 0;  // cells evaluated, added to meta after the task loop

    for (;;) {
        int v = 0;
        if (lane == 0) v = atomicAdd(T.queue, 1);
        const int t = uni(v);
        if (t >= ntasks) break;

        // per half (h = 0: pair A, 1: pair B)
        int a[2], b[2], bi[2], wend[2], anc[2], best[2], sh[2];
        bool live[2];
        uint32_t p[2];
        int64_t wi[2];
#pragma unroll
        for (int hh = 0; hh < 2; ++hh) {
            const int64_t slot = slot0 + (int64_t)t * 2 * kWave + hh * kWave + lane;
            p[hh] = slot < slot1 ? T.sorted[slot] : ~0u;
            live[hh] = p[hh] != ~0u;
            a[hh] = 0; b[hh] = 0; bi[hh] = 0; wend[hh] = 0; anc[hh] = 0; best[hh] = 32767;
            if (live[hh]) {
                a[hh] = (int)(r0 + p[hh] / ld1);
                b[hh] = (int)(c0 + p[hh] % ld1);
                const uint32_t cell = T.best_cell[p[hh]];
                bi[hh] = (int)((cell >> 16) & 0x7FFFu);
                const int jb = (int)(cell & 0xFFFFu);
                wend[hh] = min(jb + kFirstMaxBlock - 1, (int)S.len[b[hh]]);
                anc[hh] = wend[hh] - jb + 1;
                best[hh] = T.best_score[p[hh]];
                if (MX3 && best[hh] + kRevBias > 0x7BFF) {
                    // outside the biased range: no reverse rows, the whole prefix
                    bi[hh] = 0;
                    anc[hh] = 0;
                }
            }
            // reverse columns c'..c'+3 are s2[wend-c'-3 .. wend-c'] (0-based),
            // one misaligned word; columns past wend are pad, word indices
            // below the buffer clamp to 0
            const int64_t top = live[hh] ? S.off[b[hh]] + wend[hh] - 4 : S.pad_off;
            sh[hh] = (int)(top & 3);
            wi[hh] = top >> 2;
        }
        const int Rmax = uni(wave_max(max(bi[0], bi[1]))), Cmax = uni(wave_max(max(wend[0], wend[1])));
        const int ncols = (Cmax + 3) & ~3;
        const uint8_t* s1a = S.res + (live[0] ? S.off[a[0]] : S.pad_off);
        const uint8_t* s1b = S.res + (live[1] ? S.off[a[1]] : S.pad_off);
        const uint32_t* s2w = reinterpret_cast<const uint32_t*>(S.res);
        // the hit value: S (MX3: t = S - o, biased)
        const int sb0 = MX3 ? best[0] + kRevBias - o : best[0], sb1 = MX3 ? best[1] + kRevBias - o : best[1];
        const uint32_t S2 = ((uint32_t)sb0 & 0xFFFFu) | ((uint32_t)sb1 << 16);
        const uint32_t BI2 = ((uint32_t)bi[0] & 0xFFFFu) | ((uint32_t)bi[1] << 16);
        const uint32_t WE2 = ((uint32_t)wend[0] & 0xFFFFu) | ((uint32_t)wend[1] << 16);

        // reverse row 0: V' = 0 on each pair's block columns (t = V' - o), else
        // -inf; only columns 0 .. kFirstMaxBlock - 1 hold a 0
        for (int j = 1; j <= kFirstMaxBlock; ++j) {
            const uint32_t tl = j < anc[0] ? (T0 & 0xFFFFu) : (NEG2 & 0xFFFFu);
            const uint32_t th = j < anc[1] ? (T0 & 0xFFFF0000u) : (NEG2 & 0xFFFF0000u);
            bnd[(int64_t)j * kWave + lane] = make_int2((int)(tl | th), (int)NEG2);
        }

        // Band (X-drop).  Call a cell live when V' >= min(0, o - e).  Every
        // cell x of an optimal path from a hit to the block is live: the part
        // of the path before x is a local alignment ending at a forward cell
        // whose value is at most S, and the part from x on scores at most
        // V'(x); the two parts add up to S, or to S + o - e when x splits a
        // gap run (both parts then charge the run's opening), so V'(x) >=
        // min(0, o - e).  So a cell none of whose three predecessors is live
        // is left out (-inf), which keeps every hit and, values only
        // dropping, creates none; and once a row has no live cell, no later
        // row holds a hit (the path from such a hit would cross that row at a
        // live cell).  Per 16-row strip the wave evaluates columns jlo .. jend:
        // jlo = the first 4-column group with a live cell in the row above the
        // strip (earlier columns have no live predecessor, by induction along
        // the row), and the strip stops after a group with no live cell in any
        // row that lies past the last live column of the row above (later
        // columns again have no live predecessor).  bnd holds the row above for
        // columns jlo .. jtop; later columns read -inf.
        // live: V' >= min(0, o - e); MX3: t = V' - o + kRevBias >= kRevBias - max(o, e)
        const uint32_t LIVE2 = MX3 ? splat16(kRevBias - max(o, e)) : splat16(min(0, o - e));
        int jlo = 1, jtop = kFirstMaxBlock, live_hi = kFirstMaxBlock - 1;
        uint32_t done = 0;  // per half: 0xFFFF once a row without live cells (or past bi) was seen
        uint32_t hitr = 0, hitc = 0;
        int rows_done = 0;  // reverse rows evaluated
        for (int i0 = 0; i0 < Rmax && i0 < T.row_cap; i0 += R) {
            // per row: JOINT, the table row of the codes pair (ra, rb); else
            // the [c1] rows of both halves
            int rba[R], rbb[R];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int ka = bi[0] - 1 - (i0 + r), kb = bi[1] - 1 - (i0 + r);
                const int ra = ka >= 0 ? (int)s1a[ka] : pad, rb = kb >= 0 ? (int)s1b[kb] : pad;
                if (JOINT) {
                    rba[r] = (ra * ncode + rb) * ncode2;
                    rbb[r] = 0;
                } else {
                    rba[r] = ra * ncode;
                    rbb[r] = rb * ncode;
                }
            }
            uint32_t tr[R], Hr[R], rm[R];
#pragma unroll
            for (int r = 0; r < R; ++r) { tr[r] = NEG2; Hr[r] = NEG2; rm[r] = NEG2; }
            uint32_t dtb = i0 == 0 ? T0 : NEG2;  // V'(i0, jlo - 1) - o (jlo = 1 in the first strip)

            // residue words of the current and the next 4 columns (byte 3 = the
            // first of the 4), rolling down through s2
            int64_t wa = wi[0] - ((jlo - 1) >> 2), wb = wi[1] - ((jlo - 1) >> 2);
            uint32_t loa = s2w[wa < 0 ? 0 : wa], hia = s2w[wa + 1 < 0 ? 0 : wa + 1];
            uint32_t lob = s2w[wb < 0 ? 0 : wb], hib = s2w[wb + 1 < 0 ? 0 : wb + 1];
            uint32_t wda = __builtin_amdgcn_alignbyte(hia, loa, sh[0]);
            uint32_t wdb = __builtin_amdgcn_alignbyte(hib, lob, sh[1]);
            hia = loa; hib = lob; --wa; --wb;
            loa = s2w[wa < 0 ? 0 : wa];
            lob = s2w[wb < 0 ? 0 : wb];
            uint32_t wdaN = __builtin_amdgcn_alignbyte(hia, loa, sh[0]);
            uint32_t wdbN = __builtin_amdgcn_alignbyte(hib, lob, sh[1]);
            // both pairs' scores of the first column, packed (the loop reloads
            // them row by row, a column ahead)
            auto col_scores = [&](int ca, int cb, uint32_t* out) {
                if (JOINT) {
                    const int cp = ca * ncode + cb;
#pragma unroll
                    for (int r = 0; r < R; ++r) out[r] = jt[rba[r] + cp];
                } else {
#pragma unroll
                    for (int r = 0; r < R; ++r)
                        out[r] = (uint32_t)(uint16_t)tab16[rba[r] + ca] | ((uint32_t)(uint16_t)tab16[rbb[r] + cb] << 16);
                }
            };
            uint32_t cur[R];
            col_scores(wend[0] >= jlo ? (int)(wda >> 24) : pad, wend[1] >= jlo ? (int)(wdb >> 24) : pad, cur);
            // boundary words of the next four columns, each reloaded four
            // columns ahead once its column is done (columns past jtop are
            // stale and read as -inf); the residue words one block further
            // ahead than they are needed
            int2 bw[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) bw[q] = bnd[(int64_t)(jlo + q) * kWave + lane];
            uint32_t loaP = s2w[wa - 1 < 0 ? 0 : wa - 1], lobP = s2w[wb - 1 < 0 ? 0 : wb - 1];
            __builtin_amdgcn_sched_barrier(kSchedAluOnly);
            int lo_next = 0, hi_next = 0, jend = jlo - 1;
            for (int j = jlo; j <= ncols; j += 4) {
                uint32_t gm = NEG2, lm = NEG2;  // group maxima: all rows, last row
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int jj = j + q;
                    int ca = (int)(((q < 3 ? wda : wdaN) >> (8 * (3 - ((q + 1) & 3)))) & 0xFFu);
                    int cb = (int)(((q < 3 ? wdb : wdbN) >> (8 * (3 - ((q + 1) & 3)))) & 0xFFu);
                    ca = jj + 1 > wend[0] ? pad : ca;
                    cb = jj + 1 > wend[1] ? pad : cb;
                    const int cp = ca * ncode + cb;
                    uint32_t tu = jj <= jtop ? (uint32_t)bw[q].x : NEG2, Gu = jj <= jtop ? (uint32_t)bw[q].y : NEG2;
                    uint32_t dt = dtb;
                    dtb = tu;
                    s16x2 cm = v2(NEG2);
#pragma unroll
                    for (int r = 0; r < R; ++r) {
                        const uint32_t sp = cur[r];
                        uint32_t tn;
                        s16x2 G, H;
                        if (MX3) {
                            // unsigned saturating subtracts: the floor 0 is -inf
                            G = __builtin_elementwise_max(
                                __builtin_bit_cast(s16x2, __builtin_elementwise_sub_sat(
                                    __builtin_bit_cast(u16x2, Gu), __builtin_bit_cast(u16x2, E2))), v2(tu));
                            H = __builtin_elementwise_max(
                                __builtin_bit_cast(s16x2, __builtin_elementwise_sub_sat(
                                    __builtin_bit_cast(u16x2, Hr[r]), __builtin_bit_cast(u16x2, E2))), v2(tr[r]));
                            const s16x2 f = __builtin_elementwise_add_sat(v2(dt), v2(sp));
                            const uint32_t V = max3b(u2(f), u2(H), u2(G));
                            tn = __builtin_bit_cast(uint32_t, __builtin_elementwise_sub_sat(
                                __builtin_bit_cast(u16x2, V), __builtin_bit_cast(u16x2, O2)));
                            // maxima of t: the row's over two columns, the
                            // column's over two rows (tu: the row above's t)
                            if (q & 1) rm[r] = max3b(rm[r], tr[r], tn);
                            if (r & 1) cm = v2(max3b(u2(cm), tu, tn));
                            if (r == R - 1) lm = u2(__builtin_elementwise_max(v2(lm), v2(tn)));
                        } else {
                            G = __builtin_elementwise_max(__builtin_elementwise_sub_sat(v2(Gu), E2), v2(tu));
                            H = __builtin_elementwise_max(__builtin_elementwise_sub_sat(v2(Hr[r]), E2), v2(tr[r]));
                            const s16x2 f = __builtin_elementwise_add_sat(v2(dt), v2(sp));
                            const s16x2 V = __builtin_elementwise_max(__builtin_elementwise_max(f, H), G);
                            rm[r] = u2(__builtin_elementwise_max(v2(rm[r]), V));
                            cm = __builtin_elementwise_max(cm, V);
                            if (r == R - 1) lm = u2(__builtin_elementwise_max(v2(lm), V));
                            tn = u2(__builtin_elementwise_sub_sat(V, O2));
                        }
                        dt = tr[r];
                        tr[r] = tn;
                        Hr[r] = u2(H);
                        tu = tn;
                        Gu = u2(G);
                        // the next column's scores of this row, into the
                        // register just read; the loads stay a column ahead
                        // of their use (only ALU instructions may cross)
                        if (JOINT) cur[r] = jt[rba[r] + cp];
                        else cur[r] = (uint32_t)(uint16_t)tab16[rba[r] + ca] | ((uint32_t)(uint16_t)tab16[rbb[r] + cb] << 16);
                        if ((r & 3) == 3) __builtin_amdgcn_sched_barrier(kSchedAluOnly);
                    }
                    bnd[(int64_t)jj * kWave + lane] = make_int2((int)tu, (int)Gu);
                    bw[q] = bnd[(int64_t)(jj + 4) * kWave + lane];
                    __builtin_amdgcn_sched_barrier(kSchedAluOnly);
                    // column hit: cm == S (cm <= S, so S - cm >= 0 is 0 only on a
                    // hit) inside the pair's columns; the largest column wins
                    const uint32_t JJ = splat16(jj);
                    uint32_t miss, out;
                    const uint32_t dc = u2(__builtin_elementwise_sub_sat(cm, v2(S2)));   // <= 0
                    const uint32_t dw = u2(__builtin_elementwise_sub_sat(v2(WE2), v2(JJ)));  // >= 0 inside
                    asm("v_pk_ashrrev_i16 %0, %1, %2" : "=v"(miss) : "s"(0x000F000Fu), "v"(dc));
                    asm("v_pk_ashrrev_i16 %0, %1, %2" : "=v"(out) : "s"(0x000F000Fu), "v"(dw));
                    hitc = u2(__builtin_elementwise_max(v2(hitc), v2(JJ & ~miss & ~out)));
                    gm = u2(__builtin_elementwise_max(v2(gm), cm));
                }
                wda = wdaN; wdb = wdbN;
                hia = loa; hib = lob; loa = loaP; lob = lobP; --wa; --wb;
                loaP = s2w[wa - 1 < 0 ? 0 : wa - 1];
                lobP = s2w[wb - 1 < 0 ? 0 : wb - 1];
                __builtin_amdgcn_sched_barrier(kSchedAluOnly);
                wdaN = __builtin_amdgcn_alignbyte(hia, loa, sh[0]);
                wdbN = __builtin_amdgcn_alignbyte(hib, lob, sh[1]);
                jend = j + 3;
                // halves with V' >= -o: V' - (-o) >= 0, sign mask 0
                uint32_t dg, dl;
                const uint32_t xg = u2(__builtin_elementwise_sub_sat(v2(gm), v2(LIVE2)));
                const uint32_t xl = u2(__builtin_elementwise_sub_sat(v2(lm), v2(LIVE2)));
                asm("v_pk_ashrrev_i16 %0, %1, %2" : "=v"(dg) : "s"(0x000F000Fu), "v"(xg));
                asm("v_pk_ashrrev_i16 %0, %1, %2" : "=v"(dl) : "s"(0x000F000Fu), "v"(xl));
                const bool any_live = __builtin_amdgcn_ballot_w64(dg != 0xFFFFFFFFu) != 0;
                const bool any_last = __builtin_amdgcn_ballot_w64(dl != 0xFFFFFFFFu) != 0;
                if (any_last) {
                    if (lo_next == 0) lo_next = j;
                    hi_next = j + 3;
                }
                if (!any_live && j + 3 > live_hi) break;
            }
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const uint32_t RR = splat16(i0 + r + 1);
                uint32_t miss, out, dead;
                const uint32_t dr = u2(__builtin_elementwise_sub_sat(v2(rm[r]), v2(S2)));
                const uint32_t dv = u2(__builtin_elementwise_sub_sat(v2(BI2), v2(RR)));
                const uint32_t dd = u2(__builtin_elementwise_sub_sat(v2(rm[r]), v2(LIVE2)));
                asm("v_pk_ashrrev_i16 %0, %1, %2" : "=v"(miss) : "s"(0x000F000Fu), "v"(dr));
                asm("v_pk_ashrrev_i16 %0, %1, %2" : "=v"(out) : "s"(0x000F000Fu), "v"(dv));
                asm("v_pk_ashrrev_i16 %0, %1, %2" : "=v"(dead) : "s"(0x000F000Fu), "v"(dd));
                hitr = u2(__builtin_elementwise_max(v2(hitr), v2(RR & ~miss & ~out)));
                done |= dead | out;
            }
            run += 2ull * R * (uint64_t)(jend - jlo + 1) * kWave;
            rows_done = i0 + R;
            // no live cell in the strip's last row, or every pair past a row
            // without one: no later row holds a hit
            if (lo_next == 0 || __builtin_amdgcn_ballot_w64(done != 0xFFFFFFFFu) == 0) break;
            jlo = lo_next;
            live_hi = hi_next;
            jtop = jend;
        }
#pragma unroll
        for (int hh = 0; hh < 2; ++hh) {
            if (live[hh]) {
                const int hr = (int)((hitr >> (16 * hh)) & 0xFFFFu), hc = (int)((hitc >> (16 * hh)) & 0xFFFFu);
                // stopped by the row cap with live cells left: the whole prefix
                const bool capped = bi[hh] > rows_done && ((done >> (16 * hh)) & 0xFFFFu) != 0xFFFFu;
                const bool found = hr > 0 && hc > 0 && !capped;
                const int si = found ? bi[hh] - hr + 1 : 1;
                const int sj = found ? wend[hh] - hc + 1 : 1;
                T.start_cell[p[hh]] = ((uint32_t)si << 16) | (uint32_t)sj;
            }
        }
    }
    // once, after the task loop (see rerun_kernel)
    if (lane == 0 && run) atomicAdd(reinterpret_cast<unsigned long long*>(T.meta) + kMetaCaptureRun, (unsigned long long)run);
}

// Jukes-Cantor, Kimura2 and MinMaxNormScore distances of a block from the
// words the alignment kernels stored (count words with hi = length - gapCount
// of the aligned pair, or the score):
//   PairwiseAlignment.ComputeJukesCantorDistance (Alignment.cs:314-348)
//   PairwiseAlignment.ComputeKimuraDistance (Alignment.cs:254-302)
//   with the SWMS / NW helpers: Kimura2 above 10 is 10, Jukes-Cantor where
//   undefined is 10 (Salsa.SmithWatermanMS/Program.cs:555-668,
//   Salsa.NeedlemanWunschTPL/Program.cs:606-700), and ComputeMinMaxNormScore
//   (Salsa.SmithWatermanMS/Program.cs:544-553: minimum 0;
//   Salsa.NeedlemanWunschTPL/Program.cs:573-581: 2*open + (|si|+|sj|)*ext)
// in binary64 (built with -ffp-contract=off), then Int16 as (short)(d * 32767)
// or Double as is (the Double builds have no MinMaxNormScore case: 0).
// A fixed grid strides over the block's rows x columns, both read from the
// descriptor, so that one captured launch serves blocks of every size.
constexpr int kDistThreads = 256;

__device__ __forceinline__ void distance_of_pair(const SeqSet& S, const BlockDesc& D, int row, int col);

__global__ __launch_bounds__(kDistThreads) void distance_kernel(SeqSet S, const BlockDesc* desc)
{
    const BlockDesc D = desc[0];
    const int64_t n = (int64_t)D.rows * D.aux_ld;
    for (int64_t k = (int64_t)blockIdx.x * kDistThreads + threadIdx.x; k < n;
         k += (int64_t)gridDim.x * kDistThreads)
        distance_of_pair(S, D, (int)(k / D.aux_ld), (int)(k % D.aux_ld));
}

// (i, i) of a lower-mirror diagonal block: the reference leaves it at the
// matrix's initial 0 (Salsa.SmithWatermanMPI/Program.cs:209-231 aligns j < i
// only); written here so that a reused device buffer holds no stale value.
__global__ __launch_bounds__(kDistThreads) void zero_diagonal_kernel(int tdist, const BlockDesc* desc)
{
    const BlockDesc D = desc[0];
    if (!D.mirror) return;
    for (int64_t i = (int64_t)blockIdx.x * kDistThreads + threadIdx.x; i < D.rows;
         i += (int64_t)gridDim.x * kDistThreads) {
        const int64_t pos = i * D.ld + i;
        if (tdist) static_cast<double*>(D.out)[pos] = 0.0;
        else static_cast<short*>(D.out)[pos] = 0;
        if (D.stats) reinterpret_cast<int4*>(D.stats)[pos] = make_int4(0, 0, 0, 0);
    }
}

__device__ __forceinline__ void distance_of_pair(const SeqSet& S, const BlockDesc& D, int row, int col)
{
    const int a = (int)D.r0 + row, b = (int)D.c0 + col;
    if (D.mirror && b >= a) return;  // lower-mirror block: pairs j < i only
    const int64_t k = (int64_t)row * D.aux_ld + col;
    const uint32_t w = D.aux[k];
    const uint32_t hi = w >> 16, lo = w & 0xFFFFu;
    double v;
    if (S.dist_kind == kDistMinMax) {
        const int64_t s = (int32_t)w;
        const int64_t La = S.len[a], Lb = S.len[b];
        const int64_t maxS = max(S.self_score[a], S.self_score[b]);
        const int64_t minS = S.minmax_global ? 2 * (int64_t)S.gap_open + (La + Lb) * (int64_t)S.gap_ext : 0;
        v = 1.0 - (double)(s - minS) * 1.0 / (double)(maxS - minS);
        // the Double builds have no MinMaxNormScore case; a local pair with an
        // empty sequence has no score (-Infinity): both give 0
        if (S.tdist || s == INT_MIN) v = S.tdist ? 0.0 : __builtin_nan("");
    } else if (S.dist_kind == kDistJukesCantor) {
        const double d = 1.0 - ((4.0 / 3.0) * ((double)lo / (double)hi));
        if (d <= kDoubleEpsilon) {
            v = 10.0;  // SWMS / NW: the artificial distance
            if (!S.bio_helpers) {  // SWMPI: the reference throws here
                v = 0.0;
                *S.undefined = 1;
                if (S.undefined_rows) S.undefined_rows[row] = 1;
            }
        } else {
            v = -0.75 * log(d);
        }
    } else {
        const uint32_t tv = D.aux[(int64_t)D.rows * D.aux_ld + k] & 0xFFFFu;
        const double P = (double)lo / (double)hi;
        const double Q = (double)tv / (double)hi;
        if (1.0 - (2.0 * P + Q) <= kDoubleEpsilon) v = 10.0;
        else if (1.0 - (2.0 * Q) <= kDoubleEpsilon) v = 10.0;
        else v = (-0.5 * log(1.0 - 2.0 * P - Q) - 0.25 * log(1.0 - 2.0 * Q));
        if (S.bio_helpers && v > 10.0) v = 10.0;
        if (D.stats) {
            // the second pass left (score, hi, transversions, 0)
            const int64_t pos = (int64_t)row * D.ld + col;
            int4* st = reinterpret_cast<int4*>(D.stats);
            const int4 x = make_int4(st[pos].x, (int)hi, (int)lo, (int)tv);
            st[pos] = x;
            if (D.mirror) st[(int64_t)(b - D.r0) * D.ld + (a - D.c0)] = x;
        }
    }
    const int64_t pos = (int64_t)row * D.ld + col;
    const int64_t mpos = (int64_t)(b - D.r0) * D.ld + (a - D.c0);
    if (S.tdist) {
        static_cast<double*>(D.out)[pos] = v;
        if (D.mirror) static_cast<double*>(D.out)[mpos] = v;
    } else {
        const short sv = f64_to_i16(v * 32767.0);
        static_cast<short*>(D.out)[pos] = sv;
        if (D.mirror) static_cast<short*>(D.out)[mpos] = sv;
    }
}

__global__ void stamp_kernel(unsigned long long* p, int k)
{
    if (threadIdx.x == 0) p[k] = wall_clock64();
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

template <int R, bool REGION>
void launch_nw_r(const SeqSet& S, const TaskSet& T, const Scratch& W, bool persistent, int nslots, hipStream_t st)
{
    if (persistent) hipLaunchKernelGGL((nw_kernel<R, true, REGION>), dim3(nslots), dim3(kWave), 0, st, S, T, W);
    else hipLaunchKernelGGL((nw_kernel<R, false, REGION>), dim3(nslots), dim3(kWave), 0, st, S, T, W);
}

hipError_t launch_nw(const SeqSet& S, const TaskSet& T, const Scratch& W, int strip_rows, bool region, bool persistent,
                     int nslots, hipStream_t st)
{
    switch (strip_rows) {
    case 8:
        if (region) launch_nw_r<8, true>(S, T, W, persistent, nslots, st);
        else launch_nw_r<8, false>(S, T, W, persistent, nslots, st);
        break;
    case 16:
        if (region) launch_nw_r<16, true>(S, T, W, persistent, nslots, st);
        else launch_nw_r<16, false>(S, T, W, persistent, nslots, st);
        break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

int nw_blocks_per_cu(int strip_rows, bool region)
{
    int nb = 0;
    hipError_t err = hipErrorInvalidValue;
    switch (strip_rows) {
    case 8:
        err = region ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, nw_kernel<8, true, true>, kWave, 0)
                     : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, nw_kernel<8, true, false>, kWave, 0);
        break;
    case 16:
        err = region ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, nw_kernel<16, true, true>, kWave, 0)
                     : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, nw_kernel<16, true, false>, kWave, 0);
        break;
    default: break;
    }
    return err == hipSuccess ? nb : 0;
}

size_t lpp16_lds_bytes(int strip_rows, int ncode)
{
    const int H = kLpp * strip_rows;
    return (size_t)ncode * (H / 2 + 1) * sizeof(int4) + (size_t)kLppPairs * H * sizeof(int2);
}

hipError_t launch_lpp16(const SeqSet& S, const TaskSet& T, const Scratch& W, int strip_rows, bool persistent,
                        int nslots, hipStream_t st)
{
    const size_t lds = lpp16_lds_bytes(strip_rows, S.ncode);
    switch (strip_rows) {
    case 4:
        if (persistent) hipLaunchKernelGGL((lpp16_kernel<4, true>), dim3(nslots), dim3(kWave), lds, st, S, T, W);
        else hipLaunchKernelGGL((lpp16_kernel<4, false>), dim3(nslots), dim3(kWave), lds, st, S, T, W);
        break;
    case 8:
        if (persistent) hipLaunchKernelGGL((lpp16_kernel<8, true>), dim3(nslots), dim3(kWave), lds, st, S, T, W);
        else hipLaunchKernelGGL((lpp16_kernel<8, false>), dim3(nslots), dim3(kWave), lds, st, S, T, W);
        break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

int lpp16_blocks_per_cu(int strip_rows, int ncode)
{
    int nb = 0;
    hipError_t err = hipErrorInvalidValue;
    const size_t lds = lpp16_lds_bytes(strip_rows, ncode);
    switch (strip_rows) {
    case 4: err = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, lpp16_kernel<4, true>, kWave, lds); break;
    case 8: err = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, lpp16_kernel<8, true>, kWave, lds); break;
    default: break;
    }
    return err == hipSuccess ? nb : 0;
}

hipError_t launch_score16(const SeqSet& S, const ScoreTaskSet& T, const Scratch& W, int strip_rows, bool mx3,
                          int nslots, hipStream_t st)
{
    const bool joint = S.ncode <= kJointMaxCodes;
#define SWG_SCORE16(RR)                                                                                           \
    if (joint && mx3) hipLaunchKernelGGL((score16_kernel<RR, true, true>), dim3(nslots), dim3(kWave), 0, st, S, T, W); \
    else if (joint) hipLaunchKernelGGL((score16_kernel<RR, true, false>), dim3(nslots), dim3(kWave), 0, st, S, T, W); \
    else if (mx3) hipLaunchKernelGGL((score16_kernel<RR, false, true>), dim3(nslots), dim3(kWave), 0, st, S, T, W);  \
    else hipLaunchKernelGGL((score16_kernel<RR, false, false>), dim3(nslots), dim3(kWave), 0, st, S, T, W);
    switch (strip_rows) {
    case 16: SWG_SCORE16(16) break;
    case 24: SWG_SCORE16(24) break;
    case 32: SWG_SCORE16(32) break;
    case 40: SWG_SCORE16(40) break;
    default: return hipErrorInvalidValue;
    }
#undef SWG_SCORE16
    return hipGetLastError();
}

hipError_t launch_lps16(const SeqSet& S, const ScoreTaskSet& T, const Scratch& W, int strip_rows, int lanes,
                        bool mx3, int nslots, hipStream_t st)
{
    if (strip_rows != 8 || (lanes != 16 && lanes != 32)) return hipErrorInvalidValue;
    if (lanes == 16 && mx3) hipLaunchKernelGGL((lps16_kernel<8, 16, true>), dim3(nslots), dim3(kWave), 0, st, S, T, W);
    else if (lanes == 16) hipLaunchKernelGGL((lps16_kernel<8, 16, false>), dim3(nslots), dim3(kWave), 0, st, S, T, W);
    else if (mx3) hipLaunchKernelGGL((lps16_kernel<8, 32, true>), dim3(nslots), dim3(kWave), 0, st, S, T, W);
    else hipLaunchKernelGGL((lps16_kernel<8, 32, false>), dim3(nslots), dim3(kWave), 0, st, S, T, W);
    return hipGetLastError();
}

int lps16_blocks_per_cu(int strip_rows, int lanes, bool mx3)
{
    int nb = 0;
    hipError_t err = hipErrorInvalidValue;
    if (strip_rows == 8 && lanes == 16)
        err = mx3 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, lps16_kernel<8, 16, true>, kWave, 0)
                  : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, lps16_kernel<8, 16, false>, kWave, 0);
    else if (strip_rows == 8 && lanes == 32)
        err = mx3 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, lps16_kernel<8, 32, true>, kWave, 0)
                  : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, lps16_kernel<8, 32, false>, kWave, 0);
    return err == hipSuccess ? nb : 0;
}

hipError_t launch_sort(const SortSet& P, int grid, hipStream_t st)
{
    hipLaunchKernelGGL(sort_count_kernel, dim3(grid), dim3(kSortThreads), 0, st, P);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(sort_scan_kernel, dim3(1), dim3(kScanThreads), 0, st, P);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(sort_scatter_kernel, dim3(grid), dim3(kSortThreads), 0, st, P);
    return hipGetLastError();
}

hipError_t launch_rerun(const SeqSet& S, const RerunSet& T, const Scratch& W, int strip_rows, bool track,
                        int nslots, hipStream_t st)
{
    const size_t lds = (size_t)S.ncode * S.ncode * sizeof(int2);
    if (track) {
        // the flagged list (rare) always runs 16-row strips
        hipLaunchKernelGGL((rerun_kernel<16, true>), dim3(nslots), dim3(kWave), lds, st, S, T, W);
        return hipGetLastError();
    }
    switch (strip_rows) {
    case 16: hipLaunchKernelGGL((rerun_kernel<16, false>), dim3(nslots), dim3(kWave), lds, st, S, T, W); break;
    case 32: hipLaunchKernelGGL((rerun_kernel<32, false>), dim3(nslots), dim3(kWave), lds, st, S, T, W); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// Rows per lane of the big-rectangle re-run: 8 when the stripe profile of
// 512 rows fits in 64 KiB of LDS with the merge rows, else 4.
static int rect64_rows(int ncode) { return ncode <= 12 ? 8 : 4; }

static size_t rect64_lds(int ncode, int R)
{
    return (size_t)ncode * (size_t)(kRectLanes * R / 2 + 1) * sizeof(int4) + (size_t)kRectLanes * R * sizeof(int2);
}

hipError_t launch_rect64(const SeqSet& S, const RerunSet& T, const Scratch& W, int nslots, hipStream_t st)
{
    const int R = rect64_rows(S.ncode);
    const size_t lds = rect64_lds(S.ncode, R);
    if (R == 8) hipLaunchKernelGGL((rect64_kernel<8>), dim3(nslots), dim3(kWave), lds, st, S, T, W);
    else hipLaunchKernelGGL((rect64_kernel<4>), dim3(nslots), dim3(kWave), lds, st, S, T, W);
    return hipGetLastError();
}

int rect64_blocks_per_cu(int ncode)
{
    const int R = rect64_rows(ncode);
    const size_t lds = rect64_lds(ncode, R);
    int nb = 0;
    const hipError_t err = R == 8 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, rect64_kernel<8>, kWave, lds)
                                  : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, rect64_kernel<4>, kWave, lds);
    return err == hipSuccess ? nb : 0;
}

hipError_t launch_reverse(const SeqSet& S, const ReverseSet& T, const Scratch& W, int strip_rows, bool mx3,
                          int nslots, hipStream_t st)
{
    const size_t lds = reverse_lds_bytes(S.ncode);
    const bool joint = S.ncode <= kRevJointMaxCodes;
#define SWG_REVERSE(RR)                                                                                            \
    if (joint && mx3) hipLaunchKernelGGL((reverse_kernel<RR, true, true>), dim3(nslots), dim3(kWave), lds, st, S, T, W); \
    else if (joint) hipLaunchKernelGGL((reverse_kernel<RR, true, false>), dim3(nslots), dim3(kWave), lds, st, S, T, W); \
    else if (mx3) hipLaunchKernelGGL((reverse_kernel<RR, false, true>), dim3(nslots), dim3(kWave), lds, st, S, T, W);  \
    else hipLaunchKernelGGL((reverse_kernel<RR, false, false>), dim3(nslots), dim3(kWave), lds, st, S, T, W);
    switch (strip_rows) {
    case 8: SWG_REVERSE(8) break;
    case 16: SWG_REVERSE(16) break;
    default: return hipErrorInvalidValue;
    }
#undef SWG_REVERSE
    return hipGetLastError();
}

int reverse_blocks_per_cu(int strip_rows, int ncode, bool mx3)
{
    const size_t lds = reverse_lds_bytes(ncode);
    const bool joint = ncode <= kRevJointMaxCodes;
    int nb = 0;
    hipError_t err = hipErrorInvalidValue;
#define SWG_REVERSE_OCC(RR)                                                                                        \
    err = joint && mx3 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, reverse_kernel<RR, true, true>, kWave, lds)  \
        : joint        ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, reverse_kernel<RR, true, false>, kWave, lds) \
        : mx3          ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, reverse_kernel<RR, false, true>, kWave, lds) \
                       : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, reverse_kernel<RR, false, false>, kWave, lds);
    switch (strip_rows) {
    case 8: SWG_REVERSE_OCC(8) break;
    case 16: SWG_REVERSE_OCC(16) break;
    default: break;
    }
#undef SWG_REVERSE_OCC
    return err == hipSuccess ? nb : 0;
}

int rerun_blocks_per_cu(int strip_rows, int ncode)
{
    const size_t lds = (size_t)ncode * ncode * sizeof(int2);
    int nb = 0;
    hipError_t err = hipErrorInvalidValue;
    switch (strip_rows) {
    case 16: err = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, rerun_kernel<16, false>, kWave, lds); break;
    case 32: err = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, rerun_kernel<32, false>, kWave, lds); break;
    default: break;
    }
    return err == hipSuccess ? nb : 0;
}

int score16_blocks_per_cu(int strip_rows, int ncode, bool mx3)
{
    const bool joint = ncode <= kJointMaxCodes;
    int nb = 0;
    hipError_t err = hipErrorInvalidValue;
#define SWG_SCORE16_OCC(RR)                                                                          \
    err = joint && mx3 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, score16_kernel<RR, true, true>, kWave, 0)   \
        : joint        ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, score16_kernel<RR, true, false>, kWave, 0)  \
        : mx3          ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, score16_kernel<RR, false, true>, kWave, 0)  \
                       : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, score16_kernel<RR, false, false>, kWave, 0);
    switch (strip_rows) {
    case 16: SWG_SCORE16_OCC(16) break;
    case 24: SWG_SCORE16_OCC(24) break;
    case 32: SWG_SCORE16_OCC(32) break;
    case 40: SWG_SCORE16_OCC(40) break;
    default: break;
    }
#undef SWG_SCORE16_OCC
    return err == hipSuccess ? nb : 0;
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

hipError_t launch_distance(const SeqSet& S, const BlockDesc* desc, int grid, hipStream_t st)
{
    if (grid <= 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(distance_kernel, dim3((unsigned)grid), dim3(kDistThreads), 0, st, S, desc);
    return hipGetLastError();
}

hipError_t launch_zero_diagonal(int tdist, const BlockDesc* desc, int grid, hipStream_t st)
{
    if (grid <= 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(zero_diagonal_kernel, dim3((unsigned)grid), dim3(kDistThreads), 0, st, tdist, desc);
    return hipGetLastError();
}

hipError_t launch_stamp(unsigned long long* p, int k, hipStream_t st)
{
    hipLaunchKernelGGL(stamp_kernel, dim3(1), dim3(kWave), 0, st, p, k);
    return hipGetLastError();
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
