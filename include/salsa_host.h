/*
 * salsa_host.h -- C++ host layer above the swg.h boundary (libsalsa_host.so).
 *
 * The reference's host is C# (.NET 4.0) and no CLR exists on this system, so
 * this library restates the host-side pieces of the path in C++ with the
 * reference's names and semantics; a C# deployment keeps its own host and
 * calls swg.h through P/Invoke instead.
 *
 *  - RangePartitioner / BlockPartitioner  (Salsa.CoreTPL/RangePartitioner.cs:17-55,
 *                                          Salsa.CoreTPL/BlockPartitioner.cs:10-28)
 *  - MPIComputeLowerTriangularBlocked.Compute + checkerboard ownership
 *                                          (Salsa.SmithWatermanMPI/Program.cs:574-642)
 *  - Scatter mirror and WriteFullMatrixOnRank0 layout
 *                                          (Salsa.SmithWatermanMPI/Program.cs:645-677,786-888)
 *  - SWMS block files                      (Salsa.SmithWatermanMS/Program.cs:1081-1160)
 *  - ScoringMatrix.Load                    (Salsa.CoreTPL/Bio/Algorithms/ScoringMatrix.cs:145-196)
 *  - synthetic sequence generators for the benchmark workloads (no reference counterpart)
 */
#ifndef SALSA_HOST_H
#define SALSA_HOST_H

#include <stdint.h>
#include "swg.h"

#ifdef __cplusplus
extern "C" {
#endif

/* RangePartitioner.Partition(startIndex, length, numPartitions): the first
 * length % n ranges get one extra element; ends are inclusive. */
int salsa_range_partition(int32_t start, int32_t length, int32_t nparts, int32_t* starts, int32_t* ends);

/* Checkerboard ownership of block (rank, j) (Program.cs:628-639): computed iff
 * rank == j, or rank > j and rank+j even, or rank < j and rank+j odd. */
int salsa_block_is_computed(int32_t rank, int32_t j);

/* MPIComputeLowerTriangularBlocked.Compute for one rank of `world`: calls
 * ComputeBlock (swg_compute_block) for every owned block of the rank's row
 * stripe, writing into `partial` (PartialMatrix rows x n, row-major, ld n).
 * elem_size = sizeof(TDistance): 2 (Int16) or 8 (Double). */
int salsa_compute_rank(swg_ctx* ctx, int32_t n, int32_t world, int32_t rank, void* partial,
                       int32_t elem_size, int64_t* pairs);

/* Every rank's Compute followed by the Scatter mirror, on one process: the
 * n x n matrix that WriteFullMatrixOnRank0 writes (row-major TDistance). */
int salsa_compute_full_matrix(swg_ctx* ctx, int32_t n, int32_t world, void* full, int32_t elem_size,
                              int64_t* pairs);

/* Raw row-major matrix file (MatrixBinaryWriter / WriteFullMatrixOnRank0). */
int salsa_write_matrix(const char* path, const void* data, int64_t nbytes);

/* The index file rank 0 writes beside the matrix (WriteIndexFile,
 * Salsa.SmithWatermanMPI/Program.cs:194,316-331): line i is
 * "<i>\t<labels[i]>\t<offsets[i+1]-offsets[i]>", UTF-8 labels, line end
 * "\r\n" (crlf = 1, the reference's Windows StreamWriter) or "\n". */
int salsa_write_index_file(const char* path, const char* const* labels, const int64_t* offsets, int32_t n,
                           int32_t crlf);

/* One SWMS block file "<rank>-<j>.bin" (header of Program.cs:1116-1160 then
 * the block's TDistance values row-major, little endian). names[4]: sequence
 * names at row start, row end, column start, column end. */
int salsa_write_swms_block(const char* path, const char* project, int32_t rank, const char* sysname,
                           const char* fasta, int32_t row_block, int32_t col_block,
                           int32_t row_start, int32_t row_end, int32_t col_start, int32_t col_end,
                           const char* const* names, const void* data, int64_t nelem, int32_t elem_size);

/* SWMS rank driver: MPIComputeLowerTriangularBlocked.Compute(lastSavedBlock)
 * with WriteBlocks and the done file (Salsa.SmithWatermanMS/Program.cs:
 * 1164-1227, 1081-1160, 188-201).  Computes the rank's checkerboard blocks j >
 * last_saved_block in order, keeps unsaved blocks in memory, writes
 * "<block_dir>/<rank>-<j>.bin" for every unsaved owned block each time
 * block_write_frequency blocks have been computed (and for the rest at the
 * end), then "<done_dir>/<rank>.txt" ("Done: Rank <rank> done.").  labels[n]
 * name the sequences in the block headers.  stop_after >= 0 stops after that
 * many computed blocks without the final flush or done file (failure
 * injection for restart tests; returns SWG_E_CAPACITY).  *blocks_saved gets
 * the last saved block index (the restart file's lastSavedBlock). */
int salsa_swms_compute_rank(swg_ctx* ctx, int32_t n, int32_t world, int32_t rank, int32_t last_saved_block,
                            int32_t block_write_frequency, const char* block_dir, const char* done_dir,
                            const char* project, const char* sysname, const char* fasta,
                            const char* const* labels, int32_t elem_size, int32_t stop_after,
                            int32_t* last_saved, int64_t* pairs);

/* MatrixBinaryReader.ReadInt16 / ReadDouble(fileName, startRowIndex, rowCount)
 * (Salsa.CoreTPL/Blas/MatrixBinaryReader.cs:224-311): rows start_row ..
 * start_row+row_count-1 of a row-major file of n_cols TDistance per row, as the
 * downstream programs (PWC, MDS) load the matrix this path writes.  elem_size
 * 2 (Int16) or 8 (Double); a short file is an error. */
int salsa_read_matrix_rows(const char* path, int32_t n_cols, int64_t start_row, int32_t row_count,
                           int32_t elem_size, void* out);

/* MatrixBinaryReader.ReadInt16(path, prefix, ext, startIndex, endIndex, ranges)
 * (MatrixBinaryReader.cs:191-217): the partial stripe files
 * "<dir>/<prefix>_<i><ext>" (SWMPI's per-rank PartialMatrixBinaryWriter output,
 * Salsa.SmithWatermanMPI/Program.cs:145-151) for i = start_index..end_index,
 * range i holding range_ends[i]-range_starts[i]+1 rows, concatenated into out. */
int salsa_read_partial_stripes(const char* dir, const char* prefix, const char* ext, int32_t start_index,
                               int32_t end_index, const int32_t* range_starts, const int32_t* range_ends,
                               int32_t n_cols, int32_t elem_size, void* out);

/* The distance a downstream consumer derives from a stored Int16 value.
 * consumer 0: PWCParallelism.getDistanceValue (Salsa.PairwiseClusteringTPL/
 * PWCParallelism.cs:211-218): value / 32767.0.  consumer 1:
 * SALSAParallelism.getDistanceValue (MDSasCHISQ/SALSAParallelism.cs:318-370,
 * Int16 build): 32767 is undefined when distance_cut > 0 and 1.0 otherwise,
 * a negative value is undefined; undefined is -1.0 if undefined_value < 0,
 * else undefined_value (squared when processing_option == 2); a defined value
 * is value / 32767.0, squared when processing_option == 2. */
double salsa_distance_value(int32_t value, int32_t consumer, double distance_cut, double undefined_value,
                            int32_t processing_option);

/* ReadRestartFile (Salsa.SmithWatermanMS/Program.cs:245-262): the first line
 * is the world size; line 1+k holds "<original rank> <last saved block>" for
 * the k-th restarted process. */
int salsa_read_restart_file(const char* path, int32_t process_rank, int32_t* world, int32_t* rank,
                            int32_t* last_saved_block);

/* ScoringMatrix.Load from a matrix text file into a 127x127 float table. */
int salsa_load_scoring_matrix(const char* path, float* scores);

/* ---- synthetic workloads ---------------------------------------------- */
/* Length models: 0 uniform integer [p0,p1]; 1 bimodal normal: weight p0 of
 * N(p1,p2), else N(p3,p4), rounded and clipped to [p5,p6]; 2 log-normal:
 * round(exp(N(ln p0, p1))) clipped to [p2,p3]. */
#define SALSA_LEN_UNIFORM 0
#define SALSA_LEN_BIMODAL 1
#define SALSA_LEN_LOGNORMAL 2

/* DNA: n sequences mutated from n_anc uniform-ACGT ancestors.  Returns the
 * number of residue bytes (offsets[n+1] filled) or -needed when cap is too
 * small. */
int64_t salsa_synth_dna(uint64_t seed, int32_t n, int32_t n_anc, double sub_rate, double indel_rate,
                        int32_t len_kind, const double* len_params, uint8_t* out, int64_t cap, int64_t* offsets);
/* Protein: residues i.i.d. over the 20 amino acids with background
 * frequencies from a seeded Dirichlet(alpha=1). */
int64_t salsa_synth_protein(uint64_t seed, int32_t n, int32_t len_kind, const double* len_params,
                            uint8_t* out, int64_t cap, int64_t* offsets);

#ifdef __cplusplus
}
#endif
#endif
