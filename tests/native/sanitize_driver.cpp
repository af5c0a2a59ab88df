// Driver of the host-code sanitizer build (tests/native/Makefile: AddressSanitizer
// + UndefinedBehaviorSanitizer): runs the CPU oracle and the parts of the host
// library that need no GPU on random and edge-case inputs, with a few
// invariants checked on the way, so that out-of-bounds accesses, leaks and
// undefined behaviour in that code surface as sanitizer reports.
//   sanitize_driver <repository root> <scratch directory>
#include "salsa_host.h"
#include "swg_oracle.h"

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#define CHECK(cond)                                                                 \
    do {                                                                            \
        if (!(cond)) {                                                              \
            std::fprintf(stderr, "%s:%d: check failed: %s\n", __FILE__, __LINE__, #cond); \
            std::exit(1);                                                           \
        }                                                                           \
    } while (0)

namespace {

struct Set {
    std::vector<uint8_t> res;
    std::vector<int64_t> off;
    std::vector<int32_t> len;
    int32_t n() const { return (int32_t)len.size(); }
};

Set make_set(const std::vector<std::string>& seqs)
{
    Set s;
    s.off.push_back(0);
    for (const auto& x : seqs) {
        s.res.insert(s.res.end(), x.begin(), x.end());
        s.off.push_back((int64_t)s.res.size());
        s.len.push_back((int32_t)x.size());
    }
    if (s.res.empty()) s.res.push_back(0);  // a valid pointer for the empty set
    return s;
}

Set synth_dna(uint64_t seed, int32_t n, double lo, double hi)
{
    const double p[2] = {lo, hi};
    std::vector<int64_t> off(n + 1);
    const int64_t need = -salsa_synth_dna(seed, n, 8, 0.1, 0.02, 0, p, nullptr, 0, off.data());
    CHECK(need >= 0);
    std::vector<uint8_t> res((size_t)need + 1);
    CHECK(salsa_synth_dna(seed, n, 8, 0.1, 0.02, 0, p, res.data(), need, off.data()) == need);
    std::vector<std::string> seqs;
    for (int32_t i = 0; i < n; ++i) seqs.emplace_back(res.begin() + off[i], res.begin() + off[i + 1]);
    return make_set(seqs);
}

Set synth_protein(uint64_t seed, int32_t n)
{
    const double p[4] = {120, 0.5, 1, 400};
    std::vector<int64_t> off(n + 1);
    const int64_t need = -salsa_synth_protein(seed, n, 2, p, nullptr, 0, off.data());
    CHECK(need >= 0);
    std::vector<uint8_t> res((size_t)need + 1);
    CHECK(salsa_synth_protein(seed, n, 2, p, res.data(), need, off.data()) == need);
    std::vector<std::string> seqs;
    for (int32_t i = 0; i < n; ++i) seqs.emplace_back(res.begin() + off[i], res.begin() + off[i + 1]);
    return make_set(seqs);
}

void oracle_pairs(const Set& s, const float* M, float o, float e, bool nucleic)
{
    std::vector<int32_t> iM(127 * 127);
    for (int k = 0; k < 127 * 127; ++k) iM[k] = (int32_t)M[k];
    std::mt19937 rng(7);
    // literal vs carried recurrence, with the aligned strings
    for (int t = 0; t < 150; ++t) {
        const int32_t a = rng() % s.n(), b = rng() % s.n();
        const uint8_t* s1 = s.res.data() + s.off[a];
        const uint8_t* s2 = s.res.data() + s.off[b];
        std::vector<char> al1(s.len[a] + s.len[b] + 1), al2(s.len[a] + s.len[b] + 1);
        oracle_result lit{}, car{}, glob{};
        CHECK(oracle_sw_literal(s1, s.len[a], s2, s.len[b], M, o, e, &lit, al1.data(), al2.data()) == 0);
        CHECK(oracle_sw_carry(s1, s.len[a], s2, s.len[b], iM.data(), (int32_t)o, (int32_t)e, &car) == 0);
        CHECK(lit.columns == car.columns && lit.identity == car.identity && lit.similarity == car.similarity);
        CHECK(lit.columns <= s.len[a] + s.len[b]);
        CHECK(oracle_nw_literal(s1, s.len[a], s2, s.len[b], M, o, e, &glob, al1.data(), al2.data()) == 0);
        CHECK(glob.columns >= std::max(s.len[a], s.len[b]));
    }
    // the threaded block driver over every distance kind and program
    const int64_t np = 400;
    std::vector<int32_t> A(np), B(np);
    for (int64_t k = 0; k < np; ++k) { A[k] = rng() % s.n(); B[k] = rng() % s.n(); }
    std::vector<int16_t> d16(np);
    std::vector<double> d64(np);
    std::vector<int32_t> st(np * 4);
    for (int glob = 0; glob < 2; ++glob)
        for (int program = 0; program < 2; ++program)
            for (int dist = 0; dist < 4; ++dist) {
                if (dist == 3 && !glob && program == 0) continue;  // MinMaxNormScore: SWMS and global only
                for (int tdouble = 0; tdouble < 2; ++tdouble) {
                    void* out = tdouble ? (void*)d64.data() : (void*)d16.data();
                    const int rc = oracle_pairs_mt(s.res.data(), s.off.data(), s.len.data(), A.data(), B.data(), np,
                                                   M, o, e, glob, program, nucleic ? 1 : 0, dist, tdouble, out,
                                                   st.data(), 4);
                    CHECK(rc == 0 || rc == -3);  // -3: a Jukes-Cantor pair the reference throws on
                }
            }
}

void epilogues()
{
    for (int32_t L = 0; L <= 300; ++L)
        for (int32_t c = 0; c <= L; ++c) {
            const int16_t v = oracle_dist_i16_fraction(c, L);
            const double f = oracle_dist_f64_fraction(c, L);
            CHECK(L == 0 || (v >= 0 && f >= 0.0 && f <= 1.0));
        }
    int32_t undef = 0;
    for (int32_t L = 1; L <= 60; ++L)
        for (int32_t t = 0; t <= L; t += 3)
            for (int32_t v = 0; t + v <= L; v += 5) {
                (void)oracle_kimura2(L, 0, t, v);
                (void)oracle_jukes_cantor(L, 0, L - t, &undef);
            }
    CHECK(oracle_f64_to_i16(std::nan("")) == 0 && oracle_f64_to_i16(1e300) == 0);
}

void partitions()
{
    for (int32_t n = 0; n <= 60; ++n)
        for (int32_t P = 1; P <= 9; ++P) {
            std::vector<int32_t> st(P), en(P);
            CHECK(salsa_range_partition(0, n, P, st.data(), en.data()) == 0 || n < P);
        }
    for (int32_t r = 0; r < 12; ++r)
        for (int32_t j = 0; j < 12; ++j)
            if (r != j) CHECK(salsa_block_is_computed(r, j) != salsa_block_is_computed(j, r));
}

void files(const std::string& dir, const Set& s)
{
    // full matrix and row-range reads
    const int32_t n = 37;
    std::vector<int16_t> m((size_t)n * n);
    for (size_t k = 0; k < m.size(); ++k) m[k] = (int16_t)(k * 7919);
    const std::string mp = dir + "/m.bin";
    CHECK(salsa_write_matrix(mp.c_str(), m.data(), (int64_t)m.size() * 2) == 0);
    std::vector<int16_t> rows((size_t)5 * n);
    CHECK(salsa_read_matrix_rows(mp.c_str(), n, 30, 5, 2, rows.data()) == 0);
    CHECK(std::memcmp(rows.data(), m.data() + (size_t)30 * n, rows.size() * 2) == 0);
    CHECK(salsa_read_matrix_rows(mp.c_str(), n, 35, 5, 2, rows.data()) != 0);  // past the end
    // partial stripes "<prefix>_<i><ext>"
    std::vector<int32_t> st(3), en(3);
    CHECK(salsa_range_partition(0, n, 3, st.data(), en.data()) == 0);
    for (int i = 0; i < 3; ++i) {
        const std::string p = dir + "/dist_" + std::to_string(i) + ".bin";
        CHECK(salsa_write_matrix(p.c_str(), m.data() + (size_t)st[i] * n, (int64_t)(en[i] - st[i] + 1) * n * 2) == 0);
    }
    std::vector<int16_t> back((size_t)n * n);
    CHECK(salsa_read_partial_stripes(dir.c_str(), "dist", ".bin", 0, 2, st.data(), en.data(), n, 2, back.data()) == 0);
    CHECK(back == m);
    // SWMS block file, restart file, index file
    const char* names[4] = {"r0", "r1", "c0", "c1"};
    CHECK(salsa_write_swms_block((dir + "/0-1.bin").c_str(), "proj", 0, "sys", "x.fasta", 0, 1, 0, 9, 10, 36, names,
                                 m.data(), 10 * 27, 2) == 0);
    const std::string rp = dir + "/restart.txt";
    FILE* f = std::fopen(rp.c_str(), "w");
    CHECK(f);
    std::fprintf(f, "3\n0 4\n1 7\n2 -1\n");
    std::fclose(f);
    int32_t w = 0, r = 0, last = 0;
    CHECK(salsa_read_restart_file(rp.c_str(), 1, &w, &r, &last) == 0 && w == 3 && r == 1 && last == 7);
    std::vector<std::string> labels;
    std::vector<const char*> lp;
    for (int32_t i = 0; i < s.n(); ++i) labels.push_back("seq" + std::to_string(i));
    for (const auto& x : labels) lp.push_back(x.c_str());
    CHECK(salsa_write_index_file((dir + "/index.txt").c_str(), lp.data(), s.off.data(), s.n(), 1) == 0);
    for (int v = -40000; v <= 40000; v += 997) {
        (void)salsa_distance_value(v, 0, -1.0, -1.0, 0);
        (void)salsa_distance_value(v, 1, 0.5, 2.0, 1);
    }
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc < 3) {
        std::fprintf(stderr, "usage: sanitize_driver <repository root> <scratch directory>\n");
        return 2;
    }
    const std::string root = argv[1], dir = argv[2];
    std::vector<float> ednafull(127 * 127), blosum(127 * 127);
    CHECK(salsa_load_scoring_matrix((root + "/csharp_amd/matrices/EDNAFULL").c_str(), ednafull.data()) == 0);
    CHECK(salsa_load_scoring_matrix((root + "/csharp_amd/matrices/BLOSUM62").c_str(), blosum.data()) == 0);
    CHECK(ednafull['A' * 127 + 'A'] == 5.0f && blosum['W' * 127 + 'W'] == 11.0f);

    Set dna = synth_dna(3, 120, 0, 180);
    std::vector<std::string> extra = {"", "A", "ACACACACAC", "NNNNACGTNN", std::string(200, 'T')};
    for (const auto& x : extra) {
        dna.res.insert(dna.res.end(), x.begin(), x.end());
        dna.off.push_back((int64_t)dna.res.size());
        dna.len.push_back((int32_t)x.size());
    }
    const Set prot = synth_protein(4, 80);
    oracle_pairs(dna, ednafull.data(), 16.0f, 4.0f, true);
    oracle_pairs(prot, blosum.data(), 10.0f, 1.0f, false);
    epilogues();
    partitions();
    files(dir, dna);
    std::printf("sanitize driver: ok\n");
    return 0;
}
