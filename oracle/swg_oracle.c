/*
 * CPU oracle for the SALSA pairwise-distance path -- TEST INFRASTRUCTURE ONLY.
 * Used by tests/ as the checker, by __graft_entry__.smoke() as the checker and
 * by bench.py's cpu_baseline leg.  The product never links or calls it.
 *
 * Build: see oracle/Makefile (gcc -O2 -ffp-contract=off, no fast-math: the
 * distance epilogue depends on per-operation IEEE binary32 rounding).
 *
 * Pinning of each function (what checks it beyond its own restated lines):
 *   oracle_sw_literal   brute-force enumeration on tiny inputs, closed forms,
 *                       hand-worked tie-break vectors (tests/test_oracle.py)
 *   oracle_sw_carry     agreement with oracle_sw_literal on random and
 *                       adversarial pairs (tests/test_oracle.py)
 *   oracle_nw_literal   brute-force global enumeration on tiny inputs;
 *                       its tie order and end-gap handling follow this file's
 *                       own definition (the reference's NW aligner is closed):
 *                       "parity unpinned" against the reference
 *   epilogues           exhaustive (columns <= 4000) table against an
 *                       independent numpy float32 evaluation, hand values
 */
#include "swg_oracle.h"

#include <math.h>
#include <float.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* Directions enum, Salsa.CoreTPL/Bio/Algorithms/SmithWaterman/Directions.cs:26-32 */
enum { DIR_STOP = 0, DIR_LEFT = 1, DIR_DIAGONAL = 2, DIR_UP = 3 };
#define GAP_CHAR '-'  /* PairwiseAlignment.GAP, Alignment.cs:36 */

static float max4f(float a, float b, float c, float d)
{
    /* SmithWatermanGotoh.Max (SmithWatermanGotoh.cs:337-340) */
    float ab = a > b ? a : b;
    float cd = c > d ? c : d;
    return ab > cd ? ab : cd;
}

static int check_symbols(const uint8_t* s, int32_t n)
{
    for (int32_t k = 0; k < n; ++k)
        if (s[k] >= ORACLE_MATRIX_SIZE) return -1;
    return 0;
}

static int is_transition(char a, char b)
{
    /* Alignment.cs:274-275 */
    return (a == 'A' && b == 'G') || (a == 'G' && b == 'A') ||
           (a == 'C' && b == 'T') || (a == 'T' && b == 'C');
}

/* Column statistics of an aligned pair, as PairwiseAlignment computes them
 * over Sequence1/Sequence2 (Alignment.cs:261-284). */
static void kimura_counts(const char* r1, const char* r2, int32_t len, oracle_result* res)
{
    int32_t ts = 0, tv = 0, gaps = 0;
    for (int32_t k = 0; k < len; ++k) {
        char a = r1[k], b = r2[k];
        if (a != b) {
            if (a == GAP_CHAR || b == GAP_CHAR) { ++gaps; continue; }
            if (is_transition(a, b)) ++ts; else ++tv;
        }
    }
    res->transitions = ts;
    res->transversions = tv;
    res->kimura_gaps = gaps;
}

/* Literal restatement of SmithWatermanGotoh.Align, Construct and Traceback,
 * Salsa.CoreTPL/Bio/Algorithms/SmithWaterman/SmithWatermanGotoh.cs:52-332. */
int oracle_sw_literal(const uint8_t* s1, int32_t n1, const uint8_t* s2, int32_t n2,
                      const float* scores, float o, float e,
                      oracle_result* res, char* aln1, char* aln2)
{
    if (check_symbols(s1, n1) || check_symbols(s2, n2)) return -1;
    const int32_t m = n1 + 1, n = n2 + 1;
    const size_t mn = (size_t)m * (size_t)n;

    /* :70-92 -- pointers zero-initialised (STOP), borders STOP, sizes 1 */
    uint8_t* pointers = (uint8_t*)calloc(mn, 1);
    int16_t* sv = (int16_t*)malloc(mn * sizeof(int16_t));
    int16_t* sh = (int16_t*)malloc(mn * sizeof(int16_t));
    float* g = (float*)malloc((size_t)n * sizeof(float));
    float* v = (float*)malloc((size_t)n * sizeof(float));
    char* r1 = (char*)malloc((size_t)(n1 + n2) + 1);
    char* r2 = (char*)malloc((size_t)(n1 + n2) + 1);
    if (!pointers || !sv || !sh || !g || !v || !r1 || !r2) {
        free(pointers); free(sv); free(sh); free(g); free(v); free(r1); free(r2);
        return -2;
    }
    for (size_t k = 0; k < mn; ++k) sv[k] = sh[k] = 1;

    /* Construct, :119-219 */
    g[0] = -INFINITY;
    float h = -INFINITY;
    v[0] = 0.0f;
    for (int32_t j = 1; j < n; ++j) { g[j] = -INFINITY; v[j] = 0.0f; }

    int32_t cell_row = 0, cell_col = 0;
    float cell_score = -INFINITY;  /* Cell(), Cell.cs:39-44 */

    for (int32_t i = 1, k = n; i < m; ++i, k += n) {
        h = -INFINITY;
        float vdiag = v[0];
        for (int32_t j = 1, l = k + 1; j < n; ++j, ++l) {
            float sim = scores[(size_t)s1[i - 1] * ORACLE_MATRIX_SIZE + s2[j - 1]];
            float f = vdiag + sim;
            float g1 = g[j] - e;
            float g2 = v[j] - o;
            if (g1 > g2) {
                g[j] = g1;
                sv[l] = (int16_t)(sv[l - n] + 1);
            } else {
                g[j] = g2;
            }
            float h1 = h - e;
            float h2 = v[j - 1] - o;
            if (h1 > h2) {
                h = h1;
                sh[l] = (int16_t)(sh[l - 1] + 1);
            } else {
                h = h2;
            }
            vdiag = v[j];
            v[j] = max4f(f, g[j], h, 0.0f);
            if (v[j] == 0.0f)       pointers[l] = DIR_STOP;
            else if (v[j] == f)     pointers[l] = DIR_DIAGONAL;
            else if (v[j] == g[j])  pointers[l] = DIR_UP;
            else                    pointers[l] = DIR_LEFT;
            if (v[j] > cell_score) { cell_row = i; cell_col = j; cell_score = v[j]; }
        }
    }

    /* Traceback, :229-332 */
    int32_t len1 = 0, identity = 0, similarity = 0, gaps = 0, possim = 0;
    int32_t i = cell_row, j = cell_col;
    size_t k = (size_t)i * n;
    int going = 1;
    while (going) {
        switch (pointers[k + j]) {
        case DIR_UP: {
            int32_t len = sv[k + j];
            for (int32_t l = 0; l < len; ++l) {
                r1[len1] = (char)s1[--i]; r2[len1] = GAP_CHAR; ++len1;
                k -= n; ++gaps;
            }
            break;
        }
        case DIR_DIAGONAL: {
            char c1 = (char)s1[--i];
            char c2 = (char)s2[--j];
            k -= n;
            r1[len1] = c1; r2[len1] = c2; ++len1;
            if (c1 == c2) { ++identity; ++similarity; }
            else if (scores[(size_t)(uint8_t)c1 * ORACLE_MATRIX_SIZE + (uint8_t)c2] > 0) ++similarity;
            if (c1 != GAP_CHAR && c2 != GAP_CHAR && scores[(size_t)(uint8_t)c1 * ORACLE_MATRIX_SIZE + (uint8_t)c2] > 0)
                ++possim;
            break;
        }
        case DIR_LEFT: {
            int32_t len = sh[k + j];
            for (int32_t l = 0; l < len; ++l) {
                r1[len1] = GAP_CHAR; r2[len1] = (char)s2[--j]; ++len1;
                ++gaps;
            }
            break;
        }
        default:
            going = 0;
        }
    }

    memset(res, 0, sizeof(*res));
    res->score = cell_score;
    res->best_row = cell_row;
    res->best_col = cell_col;
    res->start1 = i;
    res->start2 = j;
    res->columns = len1;
    res->identity = identity;
    res->similarity = similarity;
    res->gaps = gaps;
    res->pos_similarity = possim;
    kimura_counts(r1, r2, len1, res);
    if (aln1 && aln2) {  /* Reverse(), :345-355 */
        for (int32_t q = 0; q < len1; ++q) { aln1[q] = r1[len1 - 1 - q]; aln2[q] = r2[len1 - 1 - q]; }
    }
    free(pointers); free(sv); free(sh); free(g); free(v); free(r1); free(r2);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Count-carrying integer DP.  T(i,j) holds the statistics of the traceback
 * path that starts by reading the pointer at (i,j); TG/TH hold those of the
 * vertical/horizontal gap states.  Follows the recurrence of
 * SmithWatermanGotoh.cs:147-215 and the traceback rules of :262-315. */
typedef struct { int32_t cols, ident, sim, gaps, ts, tv, kgaps; } cstats;

static cstats cs_zero(void) { cstats z = {0, 0, 0, 0, 0, 0, 0}; return z; }

/* a gap column pairs `residue` with '-': a Kimura gap unless the residue is '-' */
static cstats cs_gapcol(cstats t, uint8_t residue)
{
    t.cols += 1; t.gaps += 1;
    if (residue != GAP_CHAR) t.kgaps += 1;
    return t;
}

static cstats cs_diag(cstats t, uint8_t c1, uint8_t c2, const int32_t* M)
{
    t.cols += 1;
    if (c1 == c2) { t.ident += 1; t.sim += 1; }
    else {
        if (M[(size_t)c1 * ORACLE_MATRIX_SIZE + c2] > 0) t.sim += 1;
        if (c1 != GAP_CHAR && c2 != GAP_CHAR) {
            if (is_transition((char)c1, (char)c2)) t.ts += 1; else t.tv += 1;
        } else {
            t.kgaps += 1;
        }
    }
    return t;
}

int oracle_sw_carry(const uint8_t* s1, int32_t n1, const uint8_t* s2, int32_t n2,
                    const int32_t* M, int32_t o, int32_t e, oracle_result* res)
{
    if (check_symbols(s1, n1) || check_symbols(s2, n2)) return -1;
    const int32_t NEG = INT32_MIN / 2;
    const size_t n = (size_t)n2 + 1;
    int32_t* Vp = (int32_t*)malloc(n * sizeof(int32_t));
    int32_t* Gp = (int32_t*)malloc(n * sizeof(int32_t));
    cstats* Tp = (cstats*)malloc(n * sizeof(cstats));
    cstats* TGp = (cstats*)malloc(n * sizeof(cstats));
    if (!Vp || !Gp || !Tp || !TGp) { free(Vp); free(Gp); free(Tp); free(TGp); return -2; }
    for (size_t j = 0; j < n; ++j) { Vp[j] = 0; Gp[j] = NEG; Tp[j] = cs_zero(); TGp[j] = cs_zero(); }

    int found = 0;
    int32_t best = 0, bi = 0, bj = 0;
    cstats bestT = cs_zero();

    for (int32_t i = 1; i <= n1; ++i) {
        const uint8_t c1 = s1[i - 1];
        int32_t Vleft = 0, H = NEG, Vdiag = Vp[0];
        cstats Tleft = cs_zero(), TH = cs_zero(), Tdiag = Tp[0];
        for (int32_t j = 1; j <= n2; ++j) {
            const uint8_t c2 = s2[j - 1];
            const int32_t f = Vdiag + M[(size_t)c1 * ORACLE_MATRIX_SIZE + c2];
            int32_t G; cstats TG;
            const int32_t g1 = Gp[j] - e, g2 = Vp[j] - o;
            if (g1 > g2) { G = g1; TG = TGp[j]; } else { G = g2; TG = Tp[j]; }
            TG = cs_gapcol(TG, c1);
            const int32_t h1 = H - e, h2 = Vleft - o;
            if (h1 > h2) { H = h1; } else { H = h2; TH = Tleft; }
            TH = cs_gapcol(TH, c2);
            int32_t V = f > G ? f : G;
            if (H > V) V = H;
            if (0 > V) V = 0;
            cstats T;
            if (V == 0)      T = cs_zero();
            else if (V == f) T = cs_diag(Tdiag, c1, c2, M);
            else if (V == G) T = TG;
            else             T = TH;
            if (!found || V > best) { found = 1; best = V; bestT = T; bi = i; bj = j; }
            Vdiag = Vp[j]; Tdiag = Tp[j];
            Vp[j] = V; Gp[j] = G; Tp[j] = T; TGp[j] = TG;
            Vleft = V; Tleft = T;
        }
    }
    memset(res, 0, sizeof(*res));
    res->score = found ? (float)best : -INFINITY;
    res->best_row = bi; res->best_col = bj;
    res->columns = bestT.cols; res->identity = bestT.ident; res->similarity = bestT.sim;
    res->gaps = bestT.gaps; res->transitions = bestT.ts; res->transversions = bestT.tv;
    res->kimura_gaps = bestT.kgaps;
    res->start1 = -1; res->start2 = -1;  /* not carried */
    free(Vp); free(Gp); free(Tp); free(TGp);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Global variant (semantics defined here; see header). */
int oracle_nw_literal(const uint8_t* s1, int32_t n1, const uint8_t* s2, int32_t n2,
                      const float* scores, float o, float e,
                      oracle_result* res, char* aln1, char* aln2)
{
    if (check_symbols(s1, n1) || check_symbols(s2, n2)) return -1;
    const int32_t m = n1 + 1, n = n2 + 1;
    const size_t mn = (size_t)m * (size_t)n;
    uint8_t* pointers = (uint8_t*)calloc(mn, 1);
    int32_t* sv = (int32_t*)malloc(mn * sizeof(int32_t));
    int32_t* sh = (int32_t*)malloc(mn * sizeof(int32_t));
    float* g = (float*)malloc((size_t)n * sizeof(float));
    float* v = (float*)malloc((size_t)n * sizeof(float));
    char* r1 = (char*)malloc((size_t)(n1 + n2) + 1);
    char* r2 = (char*)malloc((size_t)(n1 + n2) + 1);
    if (!pointers || !sv || !sh || !g || !v || !r1 || !r2) {
        free(pointers); free(sv); free(sh); free(g); free(v); free(r1); free(r2);
        return -2;
    }
    for (size_t k = 0; k < mn; ++k) sv[k] = sh[k] = 1;
    /* border runs are single gaps back to the origin */
    for (int32_t i = 1; i < m; ++i) { pointers[(size_t)i * n] = DIR_UP; sv[(size_t)i * n] = i; }
    for (int32_t j = 1; j < n; ++j) { pointers[j] = DIR_LEFT; sh[j] = j; }

    g[0] = -INFINITY;
    v[0] = 0.0f;
    for (int32_t j = 1; j < n; ++j) { g[j] = -INFINITY; v[j] = -(o + (float)(j - 1) * e); }
    for (int32_t i = 1, k = n; i < m; ++i, k += n) {
        float h = -INFINITY;
        float vdiag = v[0];
        v[0] = -(o + (float)(i - 1) * e);
        for (int32_t j = 1, l = k + 1; j < n; ++j, ++l) {
            float sim = scores[(size_t)s1[i - 1] * ORACLE_MATRIX_SIZE + s2[j - 1]];
            float f = vdiag + sim;
            float g1 = g[j] - e, g2 = v[j] - o;
            if (g1 > g2) { g[j] = g1; sv[l] = sv[l - n] + 1; } else { g[j] = g2; }
            float h1 = h - e, h2 = v[j - 1] - o;
            if (h1 > h2) { h = h1; sh[l] = sh[l - 1] + 1; } else { h = h2; }
            vdiag = v[j];
            float best = f > g[j] ? f : g[j];
            if (h > best) best = h;
            v[j] = best;
            if (v[j] == f)         pointers[l] = DIR_DIAGONAL;
            else if (v[j] == g[j]) pointers[l] = DIR_UP;
            else                   pointers[l] = DIR_LEFT;
        }
    }
    float score = v[n - 1];

    int32_t len1 = 0, identity = 0, similarity = 0, gaps = 0, nwsim = 0;
    int32_t i = m - 1, j = n - 1;
    size_t k = (size_t)i * n;
    int going = 1;
    while (going) {
        switch (pointers[k + j]) {
        case DIR_UP: {
            int32_t len = sv[k + j];
            for (int32_t l = 0; l < len; ++l) {
                r1[len1] = (char)s1[--i]; r2[len1] = GAP_CHAR; ++len1; k -= n; ++gaps;
            }
            break;
        }
        case DIR_DIAGONAL: {
            char c1 = (char)s1[--i], c2 = (char)s2[--j];
            k -= n;
            r1[len1] = c1; r2[len1] = c2; ++len1;
            float sc = scores[(size_t)(uint8_t)c1 * ORACLE_MATRIX_SIZE + (uint8_t)c2];
            if (c1 == c2) { ++identity; ++similarity; } else if (sc > 0) ++similarity;
            if (sc > 0) ++nwsim;
            break;
        }
        case DIR_LEFT: {
            int32_t len = sh[k + j];
            for (int32_t l = 0; l < len; ++l) {
                r1[len1] = GAP_CHAR; r2[len1] = (char)s2[--j]; ++len1; ++gaps;
            }
            break;
        }
        default:
            going = 0;
        }
    }
    memset(res, 0, sizeof(*res));
    /* forward order */
    for (int32_t q = 0; q < len1 / 2; ++q) {
        char t = r1[q]; r1[q] = r1[len1 - 1 - q]; r1[len1 - 1 - q] = t;
        t = r2[q]; r2[q] = r2[len1 - 1 - q]; r2[len1 - 1 - q] = t;
    }
    /* ComputePercentIdentity, Salsa.NeedlemanWunschTPL/Program.cs:547-571 */
    int32_t fa = -1, fb = -1, la = -1, lb = -1;
    for (int32_t q = 0; q < len1; ++q) {
        if (r1[q] != GAP_CHAR) { if (fa < 0) fa = q; la = q; }
        if (r2[q] != GAP_CHAR) { if (fb < 0) fb = q; lb = q; }
    }
    int32_t first = fa > fb ? fa : fb;
    int32_t last = la < lb ? la : lb;
    int32_t rid = 0;
    for (int32_t q = first; q <= last && q >= 0; ++q)
        if (r1[q] == r2[q]) ++rid;
    res->score = score;
    res->best_row = m - 1; res->best_col = n - 1;
    res->start1 = i; res->start2 = j;
    res->columns = len1; res->identity = identity; res->similarity = similarity; res->gaps = gaps;
    res->nw_region_len = last - first + 1;
    res->nw_region_identity = rid;
    res->nw_similarity = nwsim;
    kimura_counts(r1, r2, len1, res);
    if (aln1 && aln2) { memcpy(aln1, r1, (size_t)len1); memcpy(aln2, r2, (size_t)len1); }
    free(pointers); free(sv); free(sh); free(g); free(v); free(r1); free(r2);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* ComputeDistance epilogues, Salsa.SmithWatermanMPI/Program.cs:259-314 and
 * PairwiseAlignment.PercentIdentity/PercentSimilarity, Alignment.cs:187-195. */

static int32_t cvtt_f32_i32(float x)
{
    /* x64 cvttss2si: truncation; NaN and out-of-range give INT32_MIN */
    if (!(x > -2147483904.0f && x < 2147483648.0f)) return INT32_MIN;
    return (int32_t)x;
}

static int32_t cvtt_f64_i32(double x)
{
    if (!(x > -2147483649.0 && x < 2147483648.0)) return INT32_MIN;
    return (int32_t)x;
}

int16_t oracle_dist_i16_fraction(int32_t count, int32_t columns)
{
    volatile float pct = (float)count / ((float)columns * 1.0f);
    volatile float r = 1.0f - pct;
    volatile float x = r * (float)32767;  /* short.MaxValue promoted to float */
    return (int16_t)(uint16_t)((uint32_t)cvtt_f32_i32(x) & 0xFFFFu);
}

double oracle_dist_f64_fraction(int32_t count, int32_t columns)
{
    volatile float pct = (float)count / ((float)columns * 1.0f);
    volatile float r = 1.0f - pct;
    return (double)r;
}

int16_t oracle_f64_to_i16(double x)
{
    return (int16_t)(uint16_t)((uint32_t)cvtt_f64_i32(x) & 0xFFFFu);
}

double oracle_kimura2(int32_t columns, int32_t gaps, int32_t ts, int32_t tv)
{
    double length = columns, gapCount = gaps;
    double P = ts / (length - gapCount);
    double Q = tv / (length - gapCount);
    if (1.0 - (2.0 * P + Q) <= DBL_TRUE_MIN) return 10.0;
    if (1.0 - (2.0 * Q) <= DBL_TRUE_MIN) return 10.0;
    return (-0.5 * log(1.0 - 2.0 * P - Q) - 0.25 * log(1.0 - 2.0 * Q));
}

double oracle_jukes_cantor(int32_t columns, int32_t gaps, int32_t differences, int32_t* undefined)
{
    double length = columns, gapCount = gaps, diff = differences;
    double d = 1.0 - ((4.0 / 3.0) * (diff / (length - gapCount)));
    *undefined = 0;
    if (d <= DBL_TRUE_MIN) { *undefined = 1; return 0.0; }
    return (-0.75 * log(d));
}

/* ------------------------------------------------------------------------- */
typedef struct {
    const uint8_t* residues; const int64_t* offsets; const int32_t* lengths;
    const int32_t* a; const int32_t* b; int64_t npairs;
    const float* scores; float o, e;
    int32_t nucleic, dist_fn, tdist_double, global;
    void* out; int32_t* out_stats;
    int64_t next; pthread_mutex_t mu;
    int32_t err;
    int32_t program;
} mt_job;

/* sum of M[c][c] over a sequence: si.Sum(b => mat[b, b]) */
static int64_t self_score(const uint8_t* s, int32_t n, const float* M)
{
    int64_t t = 0;
    for (int32_t k = 0; k < n; ++k) t += (int64_t)M[(size_t)s[k] * ORACLE_MATRIX_SIZE + s[k]];
    return t;
}

/* ComputeMinMaxNormScore (Salsa.SmithWatermanMS/Program.cs:544-553,
 * Salsa.NeedlemanWunschTPL/Program.cs:573-581) and the Int16 distance
 * (TDistance)((1.0 - norm) * TDistance.MaxValue) */
static int16_t minmax_i16(int64_t s, int64_t self_i, int64_t self_j, int64_t min_s)
{
    const int64_t max_s = self_i > self_j ? self_i : self_j;
    const double norm = (double)(s - min_s) * 1.0 / (double)(max_s - min_s);
    return oracle_f64_to_i16((1.0 - norm) * 32767);
}

static void* mt_worker(void* arg)
{
    mt_job* J = (mt_job*)arg;
    const int64_t chunk = 64;
    for (;;) {
        pthread_mutex_lock(&J->mu);
        int64_t lo = J->next; J->next += chunk;
        pthread_mutex_unlock(&J->mu);
        if (lo >= J->npairs) break;
        int64_t hi = lo + chunk < J->npairs ? lo + chunk : J->npairs;
        for (int64_t p = lo; p < hi; ++p) {
            int32_t ia = J->a[p], ib = J->b[p];
            oracle_result r;
            int rc = J->global
                ? oracle_nw_literal(J->residues + J->offsets[ia], J->lengths[ia], J->residues + J->offsets[ib],
                                    J->lengths[ib], J->scores, J->o, J->e, &r, NULL, NULL)
                : oracle_sw_literal(J->residues + J->offsets[ia], J->lengths[ia], J->residues + J->offsets[ib],
                                    J->lengths[ib], J->scores, J->o, J->e, &r, NULL, NULL);
            if (rc != 0) { J->err = -1; continue; }
            double d = 0.0; int16_t d16 = 0;
            const uint8_t* sa = J->residues + J->offsets[ia];
            const uint8_t* sb = J->residues + J->offsets[ib];
            if (J->global) {
                /* Salsa.NeedlemanWunschTPL/Program.cs:269-300 with
                 * ComputePercentIdentity (:547-571) / ComputePercentSimilarity (:583-604) */
                int32_t cnt = J->nucleic ? r.nw_region_identity : r.nw_similarity;
                int32_t den = J->nucleic ? r.nw_region_len : r.columns;
                d16 = oracle_dist_i16_fraction(cnt, den);
                d = oracle_dist_f64_fraction(cnt, den);
                if (J->nucleic && J->dist_fn == 3) {
                    const int64_t min_s = 2 * (int64_t)J->o + ((int64_t)J->lengths[ia] + J->lengths[ib]) * (int64_t)J->e;
                    d16 = minmax_i16((int64_t)r.score, self_score(sa, J->lengths[ia], J->scores),
                                     self_score(sb, J->lengths[ib], J->scores), min_s);
                    d = 0.0;
                }
                int kim = J->nucleic && (J->dist_fn == 1 || J->dist_fn == 2);
                if (kim) {
                    /* ComputeKimuraDistance / ComputeJukesCantorDistance of the
                     * global program (:606-700): clamp at 10, 10 where undefined */
                    if (J->dist_fn == 1) {
                        d = oracle_kimura2(r.columns, r.kimura_gaps, r.transitions, r.transversions);
                        if (d > 10.0) d = 10.0;
                    } else {
                        int32_t und = 0;
                        d = oracle_jukes_cantor(r.columns, r.kimura_gaps, r.transitions + r.transversions, &und);
                        if (und) d = 10.0;
                    }
                    d16 = oracle_f64_to_i16(d * 32767);
                }
                if (J->tdist_double) ((double*)J->out)[p] = d; else ((int16_t*)J->out)[p] = d16;
                if (J->out_stats) {
                    J->out_stats[4 * p + 0] = (int32_t)r.score;
                    J->out_stats[4 * p + 1] = kim ? r.columns - r.kimura_gaps : r.columns;
                    J->out_stats[4 * p + 2] = !kim ? cnt : J->dist_fn == 1 ? r.transitions
                                                                         : r.transitions + r.transversions;
                    J->out_stats[4 * p + 3] = !kim ? r.nw_region_len : J->dist_fn == 1 ? r.transversions : 0;
                }
                continue;
            }
            if (!J->nucleic) {
                const int32_t sim = J->program ? r.pos_similarity : r.similarity;
                d16 = oracle_dist_i16_fraction(sim, r.columns);
                d = oracle_dist_f64_fraction(sim, r.columns);
            } else if (J->dist_fn == 3) {
                d16 = isinf(r.score) ? 0
                    : minmax_i16((int64_t)r.score, self_score(sa, J->lengths[ia], J->scores),
                                 self_score(sb, J->lengths[ib], J->scores), 0);
                d = 0.0;
            } else if (J->dist_fn == 0) {
                d16 = oracle_dist_i16_fraction(r.identity, r.columns);
                d = oracle_dist_f64_fraction(r.identity, r.columns);
            } else if (J->dist_fn == 1) {
                d = oracle_kimura2(r.columns, r.kimura_gaps, r.transitions, r.transversions);
                if (J->program && d > 10.0) d = 10.0;  /* Salsa.SmithWatermanMS/Program.cs:600-604 */
                d16 = oracle_f64_to_i16(d * 32767);
            } else {
                int32_t und = 0;
                d = oracle_jukes_cantor(r.columns, r.kimura_gaps, r.transitions + r.transversions, &und);
                if (und && J->program) d = 10.0;       /* Salsa.SmithWatermanMS/Program.cs:659-664 */
                else if (und) J->err = -3;
                d16 = oracle_f64_to_i16(d * 32767);
            }
            if (J->tdist_double) ((double*)J->out)[p] = d; else ((int16_t*)J->out)[p] = d16;
            if (J->out_stats) {
                const int kim = J->nucleic && (J->dist_fn == 1 || J->dist_fn == 2);
                J->out_stats[4 * p + 0] = isinf(r.score) ? INT32_MIN : (int32_t)r.score;
                J->out_stats[4 * p + 1] = kim ? r.columns - r.kimura_gaps : r.columns;
                J->out_stats[4 * p + 2] = !kim ? r.identity
                                        : J->dist_fn == 1 ? r.transitions : r.transitions + r.transversions;
                J->out_stats[4 * p + 3] = !kim ? (J->program ? r.pos_similarity : r.similarity)
                                        : J->dist_fn == 1 ? r.transversions : 0;
            }
        }
    }
    return NULL;
}

int oracle_sw_pairs_mt(const uint8_t* residues, const int64_t* offsets, const int32_t* lengths,
                       const int32_t* a, const int32_t* b, int64_t npairs,
                       const float* scores, float o, float e,
                       int32_t nucleic, int32_t dist_fn, int32_t tdist_double,
                       void* out, int32_t* out_stats, int32_t threads)
{
    return oracle_pairs_mt(residues, offsets, lengths, a, b, npairs, scores, o, e, dist_fn < 0, 0, nucleic,
                           dist_fn < 0 ? 0 : dist_fn, tdist_double, out, out_stats, threads);
}

int oracle_pairs_mt(const uint8_t* residues, const int64_t* offsets, const int32_t* lengths,
                    const int32_t* a, const int32_t* b, int64_t npairs,
                    const float* scores, float o, float e,
                    int32_t global, int32_t program, int32_t nucleic, int32_t dist_fn, int32_t tdist_double,
                    void* out, int32_t* out_stats, int32_t threads)
{
    mt_job J;
    J.residues = residues; J.offsets = offsets; J.lengths = lengths;
    J.a = a; J.b = b; J.npairs = npairs; J.scores = scores; J.o = o; J.e = e;
    J.nucleic = nucleic; J.dist_fn = dist_fn; J.tdist_double = tdist_double; J.global = global;
    J.program = program;
    J.out = out; J.out_stats = out_stats; J.next = 0; J.err = 0;
    pthread_mutex_init(&J.mu, NULL);
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    pthread_t tid[256];
    for (int t = 0; t < threads; ++t) pthread_create(&tid[t], NULL, mt_worker, &J);
    for (int t = 0; t < threads; ++t) pthread_join(tid[t], NULL);
    pthread_mutex_destroy(&J.mu);
    return J.err;
}
