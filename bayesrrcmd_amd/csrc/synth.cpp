// synth.cpp -- deterministic synthetic PLINK genotypes for tests and benchmarks.
//
// Per SNP j: allele frequency p_j ~ U(maf_lo, maf_hi) of the counted allele,
// Hardy-Weinberg genotypes (dosage 2 w.p. p^2, 1 w.p. 2p(1-p), 0 otherwise),
// optional missing calls with probability miss_rate, and columns with no
// genotype 1 or 2 redrawn (RaggedSparseMarker rejects them,
// raggedsparsemarker.cpp:68-77). Output is the SNP-major .bed payload (no magic
// bytes): ceil(N/4) bytes per column, PLINK 2-bit codes, LSB first.
// Every value is a function of (seed, j, i) through Philox4x32-10, so any
// slice of the matrix can be regenerated independently.
#include <cstdint>
#include <algorithm>
#include <cstring>
#include <utility>
#include <thread>
#include <vector>

#include "brr_rng.h"
#include "brr_synth.h"

namespace {

constexpr uint32_t kStreamMaf = 10;
constexpr uint32_t kStreamGeno = 11;

void gen_column(uint64_t seed, uint32_t j, uint32_t N, double maf_lo, double maf_hi, double miss_rate,
                uint8_t* col, double* maf_out)
{
    const uint64_t nbytes = ((uint64_t)N + 3) / 4;
    for (uint32_t attempt = 0; attempt < 1000; ++attempt) {
        const brr::U4 f = brr::draw_block(seed, j, attempt, kStreamMaf, 0);
        const double p = maf_lo + (maf_hi - maf_lo) * brr::unit_u53(f.x[0], f.x[1]);
        const double pm = miss_rate;
        // thresholds on a 32-bit uniform: missing, then dosage 2, then 1
        const double t_miss = pm * 4294967296.0;
        const double t2 = t_miss + (1.0 - pm) * p * p * 4294967296.0;
        const double t1 = t2 + (1.0 - pm) * 2.0 * p * (1.0 - p) * 4294967296.0;
        std::memset(col, 0, nbytes);
        uint64_t n12 = 0;
        for (uint32_t i4 = 0; i4 < N; i4 += 4) {
            const brr::U4 x = brr::draw_block(seed, i4 >> 2, j, kStreamGeno | (attempt << 8), 0);
            uint8_t byte = 0;
            for (uint32_t e = 0; e < 4 && i4 + e < N; ++e) {
                const double u = (double)x.x[e];
                unsigned v;
                if (u < t_miss) v = 1u;             // missing
                else if (u < t2) { v = 0u; ++n12; } // dosage 2
                else if (u < t1) { v = 2u; ++n12; } // dosage 1
                else v = 3u;                        // dosage 0
                byte |= (uint8_t)(v << (2 * e));
            }
            col[i4 >> 2] = byte;
        }
        if (n12 > 0) {
            if (maf_out) *maf_out = p;
            return;
        }
    }
    if (maf_out) *maf_out = -1.0;
}

} // namespace

extern "C" int brr_synth_bed(uint64_t seed, uint32_t num_inds, uint32_t num_snps, double maf_lo,
                             double maf_hi, double miss_rate, uint8_t* bed, uint64_t column_stride,
                             double* maf, uint32_t threads)
{
    if (!bed || num_inds < 2 || column_stride < ((uint64_t)num_inds + 3) / 4) return 1;
    if (!(maf_lo > 0.0) || !(maf_hi <= 1.0) || maf_lo > maf_hi || miss_rate < 0.0 || miss_rate >= 1.0) return 1;
    if (threads == 0) threads = 1;
    if (threads > num_snps) threads = num_snps;
    std::vector<std::thread> pool;
    for (uint32_t t = 0; t < threads; ++t) {
        pool.emplace_back([=]() {
            const uint32_t a = (uint32_t)((uint64_t)num_snps * t / threads);
            const uint32_t b = (uint32_t)((uint64_t)num_snps * (t + 1) / threads);
            for (uint32_t j = a; j < b; ++j)
                gen_column(seed, j, num_inds, maf_lo, maf_hi, miss_rate, bed + (size_t)j * column_stride,
                           maf ? maf + j : nullptr);
        });
    }
    for (auto& th : pool) th.join();
    if (maf)
        for (uint32_t j = 0; j < num_snps; ++j)
            if (maf[j] < 0.0) return 2;
    return 0;
}

// y(:, t) = X beta(:, t) for T effect vectors over the standardised columns
// (value lut[j][dosage], dosage 3 = missing), threads over individual ranges.
extern "C" int brr_synth_xb(const uint8_t* bed, uint64_t column_stride, uint32_t num_inds,
                            uint32_t num_snps, const double* lut, uint32_t num_traits,
                            const double* beta, double* out, uint32_t threads)
{
    if (!bed || !lut || !beta || !out || num_traits == 0 || column_stride < ((uint64_t)num_inds + 3) / 4)
        return 1;
    const uint32_t T = num_traits;
    std::vector<std::pair<uint32_t, uint32_t>> nz;     // (marker, trait) with beta != 0
    for (uint32_t j = 0; j < num_snps; ++j)
        for (uint32_t t = 0; t < T; ++t)
            if (beta[(size_t)j * T + t] != 0.0) nz.emplace_back(j, t);
    std::memset(out, 0, sizeof(double) * num_inds * T);
    if (threads == 0) threads = 1;
    const uint32_t chunks = (num_inds + 1023) / 1024;
    if (threads > chunks) threads = chunks;
    std::vector<std::thread> pool;
    for (uint32_t th = 0; th < threads; ++th) {
        pool.emplace_back([=, &nz]() {
            for (uint32_t c = th; c < chunks; c += threads) {
                const uint32_t i0 = c * 1024, i1 = std::min<uint32_t>(num_inds, i0 + 1024);
                for (const auto& e : nz) {
                    const uint8_t* col = bed + (size_t)e.first * column_stride;
                    const double* lt = lut + 4 * (size_t)e.first;
                    // 2-bit code v -> dosage: 0->2, 1->missing, 2->1, 3->0
                    const double wv[4] = {lt[2], lt[3], lt[1], lt[0]};
                    const double b = beta[(size_t)e.first * T + e.second];
                    for (uint32_t i = i0; i < i1; ++i)
                        out[(size_t)i * T + e.second] += b * wv[(col[i >> 2] >> (2 * (i & 3))) & 3];
                }
            }
        });
    }
    for (auto& t : pool) t.join();
    return 0;
}
