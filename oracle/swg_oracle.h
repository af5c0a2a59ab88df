/*
 * CPU oracle for the SALSA pairwise-distance path -- TEST INFRASTRUCTURE ONLY.
 *
 * Nothing in the product (the csharp_amd package, its HIP library or its C ABI)
 * links, loads or calls this code.  It is used by tests/ (as the checker), by
 * __graft_entry__.smoke() (as the checker) and by bench.py's cpu_baseline leg
 * (the reported CPU baseline, kind "port").
 *
 * Every function restates a routine of the reference (DSC-SPIDAL/csharp,
 * SalsaTPL/...) and cites the file:line it follows.
 */
#ifndef SWG_ORACLE_H
#define SWG_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORACLE_MATRIX_SIZE 127  /* ScoringMatrix.SIZE, Salsa.CoreTPL/Bio/Algorithms/ScoringMatrix.cs:34 */

typedef struct oracle_result {
    float   score;          /* PairwiseAlignment.Score (Alignment.cs:125) */
    int32_t best_row;       /* Cell.Row of the traceback start (Cell.cs:52) */
    int32_t best_col;       /* Cell.Column */
    int32_t start1;         /* PairwiseAlignment.Start1 (SmithWatermanGotoh.cs:326) */
    int32_t start2;         /* PairwiseAlignment.Start2 (SmithWatermanGotoh.cs:327) */
    int32_t columns;        /* Sequence1.Length: aligned length incl. gap columns */
    int32_t identity;       /* SmithWatermanGotoh.cs:284-288 */
    int32_t similarity;     /* SmithWatermanGotoh.cs:284-294 (NAligner rule) */
    int32_t gaps;           /* SmithWatermanGotoh.cs:273,307 */
    int32_t transitions;    /* Alignment.cs:274-278 (A<->G, C<->T) */
    int32_t transversions;  /* Alignment.cs:279-282 */
    /* global (NW) variant only: Salsa.NeedlemanWunschTPL/Program.cs:547-604 */
    int32_t nw_region_len;      /* lastNonGap - firstNonGap + 1 */
    int32_t nw_region_identity; /* identity counted inside that region */
    int32_t nw_similarity;      /* non-gap columns with M[a][b] > 0 */
    int32_t kimura_gaps;    /* gapCount of Alignment.cs:270-273, 330-333: columns whose
                               two characters differ and one is the gap '-' */
    int32_t pos_similarity; /* local: non-gap columns with M[a][b] > 0
                               (Salsa.SmithWatermanMS/Program.cs:521-541) */
} oracle_result;

/* Literal restatement of SmithWatermanGotoh.Align
 * (Salsa.CoreTPL/Bio/Algorithms/SmithWaterman/SmithWatermanGotoh.cs:52-332):
 * float DP, O(mn) direction matrix and gap-size matrices, pointer traceback.
 * s1/s2 are uppercase ASCII codes < 127; scores is the 127x127 row-major float
 * table of ScoringMatrix.Load (row = s1 symbol).  aln1/aln2 (optional, capacity
 * n1+n2) receive the aligned sequences in forward order.  Returns 0, or -1 on
 * a symbol >= 127 (the reference's IndexOutOfRangeException). */
int oracle_sw_literal(const uint8_t* s1, int32_t n1, const uint8_t* s2, int32_t n2,
                      const float* scores, float gap_open, float gap_ext,
                      oracle_result* res, char* aln1, char* aln2);

/* Second, independent CPU implementation: integer DP that carries the
 * traceback statistics through the recurrence (no O(mn) memory).  It must
 * agree with oracle_sw_literal on score, columns, identity, similarity,
 * gaps, transitions and transversions. */
int oracle_sw_carry(const uint8_t* s1, int32_t n1, const uint8_t* s2, int32_t n2,
                    const int32_t* iscores, int32_t gap_open, int32_t gap_ext,
                    oracle_result* res);

/* Global (Needleman-Wunsch) variant with the same recurrence minus the 0 floor,
 * affine borders V[i][0] = -(o+(i-1)e), result cell (m-1,n-1), DIAG > UP > LEFT
 * priority and strict '>' for gap extension.  The reference's NW aligner is in
 * a closed assembly; these semantics are defined here (parity unpinned).
 * Fills columns, identity, gaps and the nw_* region fields used by
 * Salsa.NeedlemanWunschTPL/Program.cs:547-604. */
int oracle_nw_literal(const uint8_t* s1, int32_t n1, const uint8_t* s2, int32_t n2,
                      const float* scores, float gap_open, float gap_ext,
                      oracle_result* res, char* aln1, char* aln2);

/* ---- ComputeDistance epilogues (Salsa.SmithWatermanMPI/Program.cs:259-314) ---- */

/* x64 binary32 chain: (short)((1.0f - count/(columns*1.0f)) * short.MaxValue)
 * with cvttss2si semantics (NaN / out of range -> 0x80000000, low 16 bits kept). */
int16_t oracle_dist_i16_fraction(int32_t count, int32_t columns);
/* Double build: (double)(1.0f - count/(columns*1.0f)) */
double  oracle_dist_f64_fraction(int32_t count, int32_t columns);
/* PairwiseAlignment.ComputeKimuraDistance (Alignment.cs:254-302). */
double  oracle_kimura2(int32_t columns, int32_t gaps, int32_t transitions, int32_t transversions);
/* PairwiseAlignment.ComputeJukesCantorDistance (Alignment.cs:314-348);
 * returns 0 and sets *undefined=1 where the reference throws. */
double  oracle_jukes_cantor(int32_t columns, int32_t gaps, int32_t identity, int32_t* undefined);
/* (short)(double) with x64 cvttsd2si semantics. */
int16_t oracle_f64_to_i16(double x);

/* ---- multi-threaded block driver used as the CPU baseline ---- */
/* Aligns pairs (a[k], b[k]) (s1 = seq a[k]) with oracle_sw_literal on
 * `threads` POSIX threads and writes the Int16 or Double distance of
 * Salsa.SmithWatermanMPI/Program.cs:259-314 for distance function `dist_fn`
 * (0 PercentIdentity, 1 Kimura2, 2 JukesCantor) and alignment type
 * `nucleic` (0 = protein -> PercentSimilarity).  `out_stats` (optional) gets
 * (score, columns, identity, similarity) as 4 int32 per pair; for Kimura2
 * (score, columns - gapCount, transitions, transversions), for JukesCantor
 * (score, columns - gapCount, differences, 0).
 * dist_fn = -1 selects the global aligner oracle_nw_literal and the distance
 * of Salsa.NeedlemanWunschTPL/Program.cs:269-300 (nucleic: identity over the
 * aligned region; protein: positive-score columns over all columns);
 * out_stats then gets (score, columns, count, region length). */
int oracle_sw_pairs_mt(const uint8_t* residues, const int64_t* offsets, const int32_t* lengths,
                       const int32_t* a, const int32_t* b, int64_t npairs,
                       const float* scores, float gap_open, float gap_ext,
                       int32_t nucleic, int32_t dist_fn, int32_t tdist_double,
                       void* out, int32_t* out_stats, int32_t threads);

/* Same driver with the program's ComputeDistance helpers: `global` selects
 * oracle_nw_literal and Salsa.NeedlemanWunschTPL/Program.cs:269-300,547-700;
 * otherwise program 0 = SWMPI (as oracle_sw_pairs_mt), 1 = SWMS
 * (Salsa.SmithWatermanMS/Program.cs:417-668: PercentSimilarity over non-gap
 * columns scoring > 0, Kimura2 clamped to 10, JukesCantor 10 where
 * undefined).  dist_fn 3 = MinMaxNormScore (SWMS: minimum 0; global:
 * 2*open + (|si|+|sj|)*ext), Int16 only (the Double builds return 0). */
int oracle_pairs_mt(const uint8_t* residues, const int64_t* offsets, const int32_t* lengths,
                    const int32_t* a, const int32_t* b, int64_t npairs,
                    const float* scores, float gap_open, float gap_ext,
                    int32_t global, int32_t program, int32_t nucleic, int32_t dist_fn, int32_t tdist_double,
                    void* out, int32_t* out_stats, int32_t threads);

#ifdef __cplusplus
}
#endif
#endif
