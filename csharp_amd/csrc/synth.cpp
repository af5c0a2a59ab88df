// Deterministic synthetic sequence sets for the benchmark workloads.
// PRNG: xoshiro256** seeded through splitmix64 (stable across platforms and
// library versions, unlike numpy's distribution streams).
#include "salsa_host.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace {

struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed)
    {
        uint64_t x = seed;
        for (int k = 0; k < 4; ++k) {  // splitmix64
            x += 0x9E3779B97F4A7C15ull;
            uint64_t z = x;
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            s[k] = z ^ (z >> 31);
        }
    }
    static uint64_t rotl(uint64_t v, int k) { return (v << k) | (v >> (64 - k)); }
    uint64_t next()
    {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
        s[2] ^= t; s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }  // [0,1)
    uint32_t below(uint32_t n) { return (uint32_t)(((next() >> 32) * (uint64_t)n) >> 32); }
    double normal()
    {  // Box-Muller, one value per call
        double u1 = uniform(), u2 = uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
    int geometric(double p)
    {  // number of trials until the first success, >= 1
        int k = 1;
        while (uniform() >= p) ++k;
        return k;
    }
};

int32_t draw_length(Rng& r, int kind, const double* p)
{
    double v = 0.0;
    double lo = 1, hi = 1;
    switch (kind) {
    case SALSA_LEN_UNIFORM:
        return (int32_t)p[0] + (int32_t)r.below((uint32_t)(p[1] - p[0] + 1));
    case SALSA_LEN_BIMODAL:
        v = (r.uniform() < p[0]) ? p[1] + p[2] * r.normal() : p[3] + p[4] * r.normal();
        lo = p[5]; hi = p[6];
        break;
    default:
        v = std::exp(std::log(p[0]) + p[1] * r.normal());
        lo = p[2]; hi = p[3];
        break;
    }
    v = std::floor(v + 0.5);
    return (int32_t)std::min(hi, std::max(lo, v));
}

int32_t max_length(int kind, const double* p)
{
    switch (kind) {
    case SALSA_LEN_UNIFORM: return (int32_t)p[1];
    case SALSA_LEN_BIMODAL: return (int32_t)p[6];
    default: return (int32_t)p[3];
    }
}

const char kDNA[4] = {'A', 'C', 'G', 'T'};
const char kAA[21] = "ARNDCQEGHILKMFPSTWYV";

}  // namespace

extern "C" {

// Each sequence: pick an ancestor, take its prefix of the drawn length, then
// walk it: with probability indel_rate an indel event (insertion or deletion
// with equal probability, length ~ Geometric(0.5)); otherwise with
// probability sub_rate a substitution by a different base; otherwise a copy.
int64_t salsa_synth_dna(uint64_t seed, int32_t n, int32_t n_anc, double sub_rate, double indel_rate,
                        int32_t len_kind, const double* len_params, uint8_t* out, int64_t cap, int64_t* offsets)
{
    if (n < 0 || n_anc < 1 || !len_params || !offsets) return 0;
    Rng r(seed);
    const int32_t amax = max_length(len_kind, len_params) + 64;
    std::vector<std::vector<uint8_t>> anc((size_t)n_anc, std::vector<uint8_t>((size_t)amax));
    for (auto& a : anc)
        for (auto& ch : a) ch = (uint8_t)kDNA[r.below(4)];
    int64_t pos = 0;
    std::vector<uint8_t> seq;
    for (int32_t i = 0; i < n; ++i) {
        const auto& a = anc[r.below((uint32_t)n_anc)];
        const int32_t L = draw_length(r, len_kind, len_params);
        seq.clear();
        for (int32_t k = 0; k < L;) {
            const double u = r.uniform();
            if (u < indel_rate) {
                const int g = r.geometric(0.5);
                if (r.uniform() < 0.5) {
                    for (int q = 0; q < g; ++q) seq.push_back((uint8_t)kDNA[r.below(4)]);
                } else {
                    k += g;
                }
            } else if (u < indel_rate + sub_rate) {
                const int cur = (int)(std::find(kDNA, kDNA + 4, (char)a[k]) - kDNA);
                seq.push_back((uint8_t)kDNA[(cur + 1 + (int)r.below(3)) & 3]);
                ++k;
            } else {
                seq.push_back(a[k]);
                ++k;
            }
        }
        offsets[i] = pos;
        if (out && pos + (int64_t)seq.size() <= cap) std::memcpy(out + pos, seq.data(), seq.size());
        pos += (int64_t)seq.size();
    }
    offsets[n] = pos;
    return (!out || pos > cap) ? -pos : pos;
}

int64_t salsa_synth_protein(uint64_t seed, int32_t n, int32_t len_kind, const double* len_params,
                            uint8_t* out, int64_t cap, int64_t* offsets)
{
    if (n < 0 || !len_params || !offsets) return 0;
    Rng r(seed);
    // Dirichlet(alpha = 1) = normalised Exp(1) draws
    double w[20], sum = 0.0;
    for (int k = 0; k < 20; ++k) {
        double u = r.uniform();
        if (u < 1e-300) u = 1e-300;
        w[k] = -std::log(u);
        sum += w[k];
    }
    double cdf[20], acc = 0.0;
    for (int k = 0; k < 20; ++k) { acc += w[k] / sum; cdf[k] = acc; }
    cdf[19] = 1.0;
    int64_t pos = 0;
    for (int32_t i = 0; i < n; ++i) {
        const int32_t L = draw_length(r, len_kind, len_params);
        offsets[i] = pos;
        for (int32_t k = 0; k < L; ++k) {
            const double u = r.uniform();
            int c = 0;
            while (c < 19 && u >= cdf[c]) ++c;
            if (out && pos < cap) out[pos] = (uint8_t)kAA[c];
            ++pos;
        }
    }
    offsets[n] = pos;
    return (!out || pos > cap) ? -pos : pos;
}

}  // extern "C"
