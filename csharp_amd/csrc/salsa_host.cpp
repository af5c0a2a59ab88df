// C++ restatement of the reference's host-side pieces of the distance path
// (see include/salsa_host.h for the file:line of each).
#include "salsa_host.h"

#include <cctype>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

extern "C" {

int salsa_range_partition(int32_t start, int32_t length, int32_t nparts, int32_t* starts, int32_t* ends)
{
    // RangePartitioner.Partition (Salsa.CoreTPL/RangePartitioner.cs:17-55)
    if (nparts < 1 || !starts || !ends) return SWG_E_ARGUMENT;
    const int32_t chunk = length / nparts;
    int32_t rem = length % nparts;
    for (int32_t i = 0; i < nparts; ++i) {
        if (rem > 0) {
            starts[i] = start; ends[i] = start + chunk;
            start += chunk + 1;
            --rem;
        } else {
            starts[i] = start; ends[i] = start + chunk - 1;
            start += chunk;
        }
    }
    return SWG_OK;
}

int salsa_block_is_computed(int32_t rank, int32_t j)
{
    // MPIComputeLowerTriangularBlocked.Compute (Salsa.SmithWatermanMPI/Program.cs:628-639)
    if (rank == j) return 1;
    const bool odd = ((rank + j) & 1) == 1;
    if (rank > j && !odd) return 1;
    if (rank < j && odd) return 1;
    return 0;
}

}  // extern "C"

// Computes the owned blocks of `rank` into its stripe (rows of the rank x n).
static int compute_rank_into(swg_ctx* ctx, int32_t n, int32_t world, int32_t rank, uint8_t* stripe,
                             int64_t ld_elems, int elem, int64_t* pairs)
{
    std::vector<int32_t> rs(world), re(world);
    int rc = salsa_range_partition(0, n, world, rs.data(), re.data());
    if (rc) return rc;
    int64_t total = 0;
    for (int32_t j = 0; j < world; ++j) {
        if (!salsa_block_is_computed(rank, j)) continue;
        // the block's (row_start, col_start) element inside the stripe
        uint8_t* base = stripe + ((int64_t)0 * ld_elems + rs[j]) * elem;
        int64_t done = 0;
        rc = swg_compute_block(ctx, rs[rank], re[rank], rs[j], re[j], rank == j ? 1 : 0, base, ld_elems, nullptr,
                               &done);
        if (rc) return rc;
        total += done;
    }
    if (pairs) *pairs = total;
    return SWG_OK;
}

extern "C" {

int salsa_compute_rank(swg_ctx* ctx, int32_t n, int32_t world, int32_t rank, void* partial, int32_t elem,
                       int64_t* pairs)
{
    if (!ctx || !partial || world < 1 || rank < 0 || rank >= world || (elem != 2 && elem != 8))
        return SWG_E_ARGUMENT;
    return compute_rank_into(ctx, n, world, rank, static_cast<uint8_t*>(partial), n, elem, pairs);
}

int salsa_compute_full_matrix(swg_ctx* ctx, int32_t n, int32_t world, void* full, int32_t elem, int64_t* pairs)
{
    if (!ctx || !full || world < 1 || (elem != 2 && elem != 8)) return SWG_E_ARGUMENT;
    std::vector<int32_t> rs(world), re(world);
    int rc = salsa_range_partition(0, n, world, rs.data(), re.data());
    if (rc) return rc;
    uint8_t* F = static_cast<uint8_t*>(full);
    std::memset(F, 0, (size_t)n * n * elem);
    int64_t total = 0;
    for (int32_t r = 0; r < world; ++r) {
        int64_t done = 0;
        rc = compute_rank_into(ctx, n, world, r, F + (int64_t)rs[r] * n * elem, n, elem, &done);
        if (rc) return rc;
        total += done;
    }
    // Scatter (Program.cs:645-677): every computed off-diagonal block (r,j)
    // is transposed into block (j,r) of the owner of rows j
    for (int32_t r = 0; r < world; ++r)
        for (int32_t j = 0; j < world; ++j) {
            if (r == j || !salsa_block_is_computed(r, j)) continue;
            for (int32_t i = rs[r]; i <= re[r]; ++i)
                for (int32_t k = rs[j]; k <= re[j]; ++k)
                    std::memcpy(F + ((int64_t)k * n + i) * elem, F + ((int64_t)i * n + k) * elem, elem);
        }
    if (pairs) *pairs = total;
    return SWG_OK;
}

int salsa_write_matrix(const char* path, const void* data, int64_t nbytes)
{
    FILE* f = std::fopen(path, "wb");
    if (!f) return SWG_E_ARGUMENT;
    const size_t w = std::fwrite(data, 1, (size_t)nbytes, f);
    std::fclose(f);
    return w == (size_t)nbytes ? SWG_OK : SWG_E_ARGUMENT;
}

// WriteIndexFile (Salsa.SmithWatermanMPI/Program.cs:316-331): one line per
// sequence, "<index>\t<label>\t<residue count>", as StreamWriter.WriteLine
// writes it (UTF-8 without a byte-order mark; the line end is the platform's,
// "\r\n" on the Windows builds of the reference).
int salsa_write_index_file(const char* path, const char* const* labels, const int64_t* offsets, int32_t n,
                           int32_t crlf)
{
    if (!path || n < 0 || (n > 0 && (!labels || !offsets))) return SWG_E_ARGUMENT;
    FILE* f = std::fopen(path, "wb");
    if (!f) return SWG_E_ARGUMENT;
    bool ok = true;
    for (int32_t i = 0; i < n && ok; ++i)
        ok = std::fprintf(f, "%d\t%s\t%lld%s", i, labels[i] ? labels[i] : "",
                          (long long)(offsets[i + 1] - offsets[i]), crlf ? "\r\n" : "\n") > 0;
    ok = (std::fclose(f) == 0) && ok;
    return ok ? SWG_OK : SWG_E_ARGUMENT;
}

// MatrixBinaryReader.ReadInt16/ReadDouble(fileName, startRowIndex, rowCount)
// (Salsa.CoreTPL/Blas/MatrixBinaryReader.cs:224-311, 290-309): seek to
// startRow * n_cols elements, then row_count rows of n_cols TDistance.
int salsa_read_matrix_rows(const char* path, int32_t n_cols, int64_t start_row, int32_t row_count, int32_t elem,
                           void* out)
{
    if (!path || !out || n_cols < 0 || start_row < 0 || row_count < 0 || (elem != 2 && elem != 8))
        return SWG_E_ARGUMENT;
    FILE* f = std::fopen(path, "rb");
    if (!f) return SWG_E_ARGUMENT;
    const int64_t off = start_row * (int64_t)n_cols * elem;
    const size_t want = (size_t)row_count * (size_t)n_cols * (size_t)elem;
    size_t got = 0;
    if (fseeko(f, (off_t)off, SEEK_SET) == 0) got = std::fread(out, 1, want, f);
    std::fclose(f);
    return got == want ? SWG_OK : SWG_E_ARGUMENT;  // a short file is an error (ReadInt16 throws at EOF)
}

// MatrixBinaryReader.ReadInt16(path, prefix, ext, startIndex, endIndex, ranges)
// (MatrixBinaryReader.cs:191-217): the partial stripes "<prefix>_<i><ext>"
// that SWMPI writes per rank (Salsa.SmithWatermanMPI/Program.cs:145-151,
// PartialMatrixBinaryWriter: the rank's rows x N, row-major), for i =
// start_index..end_index, ranges[i] rows each, one after the other.
int salsa_read_partial_stripes(const char* dir, const char* prefix, const char* ext, int32_t start_index,
                               int32_t end_index, const int32_t* range_starts, const int32_t* range_ends,
                               int32_t n_cols, int32_t elem, void* out)
{
    if (!dir || !prefix || !ext || !range_starts || !range_ends || !out || start_index < 0 || end_index < start_index)
        return SWG_E_ARGUMENT;
    uint8_t* dst = static_cast<uint8_t*>(out);
    for (int32_t i = start_index; i <= end_index; ++i) {
        const int32_t rows = range_ends[i] - range_starts[i] + 1;
        const std::string name = std::string(dir) + "/" + prefix + "_" + std::to_string(i) + ext;
        const int rc = salsa_read_matrix_rows(name.c_str(), n_cols, 0, rows, elem, dst);
        if (rc) return rc;
        dst += (size_t)rows * (size_t)n_cols * (size_t)elem;
    }
    return SWG_OK;
}

// The consumers' distance of a stored Int16 value.
//  consumer 0: PWCParallelism.getDistanceValue
//    (Salsa.PairwiseClusteringTPL/PWCParallelism.cs:211-218): value / 32767.0.
//  consumer 1: SALSAParallelism.getDistanceValue (MDSasCHISQ/SALSAParallelism.cs:
//    318-370, Int16 build): 32767 means "undefined" when distance_cut > 0 (then
//    -1.0 if undefined_value < 0, else undefined_value, squared when
//    processing_option == 2) and 1.0 otherwise; a negative value is undefined
//    the same way; else value / 32767.0, squared when processing_option == 2.
double salsa_distance_value(int32_t value, int32_t consumer, double distance_cut, double undefined_value,
                            int32_t processing_option)
{
    const double scale = 32767.0;
    if (consumer == 0) return value / scale;
    auto undefined = [&]() {
        if (undefined_value < 0.0) return -1.0;
        return processing_option == 2 ? undefined_value * undefined_value : undefined_value;
    };
    if (value == 32767) return distance_cut > 0.0 ? undefined() : 1.0;
    if (value < 0) return undefined();
    const double d = value / scale;
    return processing_option == 2 ? d * d : d;
}

static void put_u16(std::vector<uint8_t>& b, uint32_t v) { b.push_back(v & 0xFF); b.push_back((v >> 8) & 0xFF); }
static void put_u32(std::vector<uint8_t>& b, uint32_t v)
{
    for (int k = 0; k < 4; ++k) b.push_back((v >> (8 * k)) & 0xFF);
}
static void put_str16(std::vector<uint8_t>& b, const char* s)
{
    const size_t n = s ? std::strlen(s) : 0;
    put_u16(b, (uint32_t)n);
    for (size_t k = 0; k < n; ++k) b.push_back((uint8_t)s[k]);
}

int salsa_write_swms_block(const char* path, const char* project, int32_t rank, const char* sysname,
                           const char* fasta, int32_t row_block, int32_t col_block,
                           int32_t row_start, int32_t row_end, int32_t col_start, int32_t col_end,
                           const char* const* names, const void* data, int64_t nelem, int32_t elem_size)
{
    // WriteRowBlockMetaData (Salsa.SmithWatermanMS/Program.cs:1116-1160)
    std::vector<uint8_t> h;
    put_str16(h, project);
    put_u16(h, (uint32_t)rank);
    put_str16(h, sysname);
    put_str16(h, fasta);
    put_u16(h, (uint32_t)row_block);
    put_u16(h, (uint32_t)col_block);
    put_u32(h, (uint32_t)row_start);
    put_u32(h, (uint32_t)row_end);
    put_u32(h, (uint32_t)col_start);
    put_u32(h, (uint32_t)col_end);
    for (int k = 0; k < 4; ++k) put_str16(h, names ? names[k] : "");
    FILE* f = std::fopen(path, "wb");
    if (!f) return SWG_E_ARGUMENT;
    bool ok = std::fwrite(h.data(), 1, h.size(), f) == h.size();
    // WriteBlocks (:1102-1108): the block's values row-major
    ok = ok && std::fwrite(data, (size_t)elem_size, (size_t)nelem, f) == (size_t)nelem;
    std::fclose(f);
    return ok ? SWG_OK : SWG_E_ARGUMENT;
}

// ScoringMatrixTokenizer with the default delimiters (whitespace)
static std::vector<std::string> tokens(const std::string& line)
{
    std::vector<std::string> t;
    std::string cur;
    for (char ch : line) {
        if (ch == ' ' || ch == '\t' || ch == '\n' || ch == '\r' || ch == '\f') {
            if (!cur.empty()) { t.push_back(cur); cur.clear(); }
        } else {
            cur.push_back(ch);
        }
    }
    if (!cur.empty()) t.push_back(cur);
    return t;
}

static std::string trim(const std::string& s)
{
    size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char)s[a])) ++a;
    while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
    return s.substr(a, b - a);
}

int salsa_load_scoring_matrix(const char* path, float* scores)
{
    // ScoringMatrix.Load (Salsa.CoreTPL/Bio/Algorithms/ScoringMatrix.cs:145-196)
    if (!path || !scores) return SWG_E_ARGUMENT;
    FILE* f = std::fopen(path, "rb");
    if (!f) return SWG_E_ARGUMENT;
    std::vector<std::string> lines;
    {
        std::string cur;
        int ch;
        while ((ch = std::fgetc(f)) != EOF) {
            if (ch == '\n') { lines.push_back(cur); cur.clear(); }
            else cur.push_back((char)ch);
        }
        if (!cur.empty()) lines.push_back(cur);
        std::fclose(f);
    }
    const int SIZE = SWG_MATRIX_SIZE;
    std::memset(scores, 0, sizeof(float) * SIZE * SIZE);
    char acids[SWG_MATRIX_SIZE];
    std::memset(acids, 0, sizeof(acids));
    size_t li = 0;
    // skip comment lines (:167)
    while (li < lines.size()) {
        const std::string t = trim(lines[li]);
        if (t.empty() || t[0] != '#') break;
        ++li;
    }
    if (li >= lines.size()) return SWG_E_ARGUMENT;
    {
        const std::vector<std::string> hdr = tokens(trim(lines[li++]));
        if ((int)hdr.size() > SIZE) return SWG_E_ARGUMENT;
        for (size_t j = 0; j < hdr.size(); ++j) acids[j] = hdr[j][0];
    }
    for (; li < lines.size(); ++li) {
        const std::vector<std::string> tk = tokens(trim(lines[li]));
        if (tk.empty()) return SWG_E_ARGUMENT;  // the reference throws on an empty line
        const unsigned char acid = (unsigned char)tk[0][0];
        if (acid >= SIZE) return SWG_E_SYMBOL;
        size_t next = 1;
        for (int i = 0; i < SIZE; ++i) {
            if (acids[i] != 0) {
                if (next >= tk.size()) return SWG_E_ARGUMENT;
                const unsigned char col = (unsigned char)acids[i];
                if (col >= SIZE) return SWG_E_SYMBOL;
                scores[acid * SIZE + col] = std::strtof(tk[next++].c_str(), nullptr);
            }
        }
    }
    return SWG_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// SWMS: Compute(lastSavedBlock) + WriteBlocks + done file
// (Salsa.SmithWatermanMS/Program.cs:1164-1227, 1081-1160, 188-201)

#include <memory>

namespace {

struct UnsavedBlock {
    int32_t j;
    std::vector<uint8_t> data;
};

int write_unsaved(std::vector<UnsavedBlock>& unsaved, int32_t rank, const std::vector<int32_t>& rs,
                  const std::vector<int32_t>& re, const char* block_dir, const char* project, const char* sysname,
                  const char* fasta, const char* const* labels, int32_t elem)
{
    // WriteBlocks walks j from lastUnsavedBlock down to lastSavedBlock+1
    for (auto it = unsaved.rbegin(); it != unsaved.rend(); ++it) {
        const int32_t j = it->j;
        const std::string path = std::string(block_dir) + "/" + std::to_string(rank) + "-" + std::to_string(j) + ".bin";
        const char* names[4] = {labels ? labels[rs[rank]] : "", labels ? labels[re[rank]] : "",
                                labels ? labels[rs[j]] : "", labels ? labels[re[j]] : ""};
        const int64_t nelem = (int64_t)(re[rank] - rs[rank] + 1) * (re[j] - rs[j] + 1);
        int rc = salsa_write_swms_block(path.c_str(), project, rank, sysname, fasta, rank, j, rs[rank], re[rank],
                                        rs[j], re[j], names, it->data.data(), nelem, elem);
        if (rc) return rc;
    }
    unsaved.clear();
    return SWG_OK;
}

}  // namespace

extern "C" int salsa_swms_compute_rank(swg_ctx* ctx, int32_t n, int32_t world, int32_t rank, int32_t last_saved_block,
                                       int32_t block_write_frequency, const char* block_dir, const char* done_dir,
                                       const char* project, const char* sysname, const char* fasta,
                                       const char* const* labels, int32_t elem, int32_t stop_after,
                                       int32_t* last_saved, int64_t* pairs)
{
    if (!ctx || !block_dir || world < 1 || rank < 0 || rank >= world || (elem != 2 && elem != 8))
        return SWG_E_ARGUMENT;
    std::vector<int32_t> rs(world), re(world);
    int rc = salsa_range_partition(0, n, world, rs.data(), re.data());
    if (rc) return rc;
    int64_t total = 0;
    int32_t computed = 0, unsaved_count = 0, saved = last_saved_block;
    std::vector<UnsavedBlock> unsaved;
    const int64_t rows = re[rank] - rs[rank] + 1;
    for (int32_t j = last_saved_block + 1; j < world; ++j) {
        if (!salsa_block_is_computed(rank, j)) continue;
        const int64_t cols = re[j] - rs[j] + 1;
        UnsavedBlock ub;
        ub.j = j;
        ub.data.assign((size_t)(rows * cols * elem), 0);
        int64_t done = 0;
        rc = swg_compute_block(ctx, rs[rank], re[rank], rs[j], re[j], rank == j ? 1 : 0, ub.data.data(), cols,
                               nullptr, &done);
        if (rc) return rc;
        total += done;
        unsaved.push_back(std::move(ub));
        ++computed;
        ++unsaved_count;
        if (block_write_frequency > 0 && unsaved_count == block_write_frequency) {
            rc = write_unsaved(unsaved, rank, rs, re, block_dir, project, sysname, fasta, labels, elem);
            if (rc) return rc;
            unsaved_count = 0;
            saved = j;
        }
        if (stop_after >= 0 && computed >= stop_after) {
            if (last_saved) *last_saved = saved;
            if (pairs) *pairs = total;
            return SWG_E_CAPACITY;
        }
    }
    if (block_write_frequency > 0 && unsaved_count > 0) {
        rc = write_unsaved(unsaved, rank, rs, re, block_dir, project, sysname, fasta, labels, elem);
        if (rc) return rc;
        saved = world - 1;
    }
    if (done_dir) {
        const std::string path = std::string(done_dir) + "/" + std::to_string(rank) + ".txt";
        FILE* f = std::fopen(path.c_str(), "w");
        if (!f) return SWG_E_ARGUMENT;
        std::fprintf(f, "Done: Rank %d done.\n", rank);
        std::fclose(f);
    }
    if (last_saved) *last_saved = saved;
    if (pairs) *pairs = total;
    return SWG_OK;
}

extern "C" int salsa_read_restart_file(const char* path, int32_t process_rank, int32_t* world, int32_t* rank,
                                       int32_t* last_saved_block)
{
    FILE* f = path ? std::fopen(path, "r") : nullptr;
    if (!f || !world || !rank || !last_saved_block) { if (f) std::fclose(f); return SWG_E_ARGUMENT; }
    int w = 0;
    int ok = std::fscanf(f, " %d", &w) == 1;
    int r = -1, ls = -1;
    for (int32_t k = 0; ok && k <= process_rank; ++k) ok = std::fscanf(f, " %d %d", &r, &ls) == 2;
    std::fclose(f);
    if (!ok) return SWG_E_ARGUMENT;
    *world = w;
    *rank = r;
    *last_saved_block = ls;
    return SWG_OK;
}
