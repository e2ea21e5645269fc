// host_markers.cpp -- MarkerBuilder semantics on the host, for the device arenas.
//
// The reference builds one Marker per SNP column from the PLINK bytes
// (preprocessgraph.cpp:24-60): the 2-bit decode rule allele1 = !b[k],
// allele2 = !b[k+1] (LSB first), then DenseMarkerBuilder standardises the
// column (densemarkerbuilder.cpp:23-60) and RaggedSparseMarkerBuilder collects
// the ascending index lists of genotype 1, 2 and missing with their statistics
// (raggedsparsemarkerbuilder.cpp:25-55, sparsemarker.cpp:8-19). The device
// keeps dense columns as their 2-bit codes plus a 4-entry table of the
// standardised values, which reproduces every element of the reference's
// FP64 column exactly.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "brr_hip.h"

namespace {

// 2-bit PLINK value -> dosage class (0, 1, 2; 3 = missing)
inline int dosage_class(unsigned v) { return v == 0u ? 2 : v == 1u ? 3 : v == 2u ? 1 : 0; }

inline unsigned code_at(const uint8_t* col, uint32_t i) { return (col[i >> 2] >> ((i & 3u) * 2u)) & 3u; }

void dense_lut_one(const uint8_t* col, uint32_t N, double* lut)
{
    // processAllele: sum of non-missing dosages, missing individuals recorded
    double sum = 0.0;
    uint32_t nmiss = 0;
    for (uint32_t i = 0; i < N; ++i) {
        const int d = dosage_class(code_at(col, i));
        if (d == 3) { ++nmiss; continue; }
        sum += (double)d;
    }
    // endColumn: missing -> mean of the rest; centre by the column mean; scale
    // to squared norm N-1. Sums in individual order.
    const double mean = sum / ((double)N - (double)nmiss);
    const double value[4] = {0.0, 1.0, 2.0, mean};
    double s = 0.0;
    for (uint32_t i = 0; i < N; ++i) s += value[dosage_class(code_at(col, i))];
    const double cmean = s / (double)N;
    double centred[4];
    for (int d = 0; d < 4; ++d) centred[d] = value[d] - cmean;
    double sqn = 0.0;
    for (uint32_t i = 0; i < N; ++i) {
        const double x = centred[dosage_class(code_at(col, i))];
        sqn += x * x;
    }
    const double sigma = 1.0 / (std::sqrt(sqn / ((double)N - 1.0)));
    for (int d = 0; d < 4; ++d) lut[d] = centred[d] * sigma;
}

template <typename F>
void parallel_for(uint32_t n, uint32_t threads, F&& f)
{
    if (threads <= 1 || n < 2) {
        for (uint32_t j = 0; j < n; ++j) f(j);
        return;
    }
    if (threads > n) threads = n;
    std::vector<std::thread> pool;
    pool.reserve(threads);
    for (uint32_t t = 0; t < threads; ++t) {
        pool.emplace_back([&, t]() {
            const uint32_t a = (uint32_t)((uint64_t)n * t / threads);
            const uint32_t b = (uint32_t)((uint64_t)n * (t + 1) / threads);
            for (uint32_t j = a; j < b; ++j) f(j);
        });
    }
    for (auto& th : pool) th.join();
}

} // namespace

extern "C" {

int brr_bed_check_magic(const uint8_t* h)
{
    if (!h) return BRR_HIP_EINVAL;
    return (h[0] == 0x6c && h[1] == 0x1b && h[2] == 0x01) ? BRR_HIP_OK : BRR_HIP_EINVAL;
}

int brr_dense_lut(const uint8_t* column, uint32_t num_inds, double* lut)
{
    if (!column || !lut || num_inds < 2) return BRR_HIP_EINVAL;
    dense_lut_one(column, num_inds, lut);
    return BRR_HIP_OK;
}

int brr_dense_luts(const uint8_t* bed, uint64_t stride, uint32_t N, uint32_t M, double* lut,
                   uint32_t threads)
{
    if (!bed || !lut || N < 2 || stride < ((uint64_t)N + 3) / 4) return BRR_HIP_EINVAL;
    parallel_for(M, threads, [&](uint32_t j) { dense_lut_one(bed + (size_t)j * stride, N, lut + 4 * (size_t)j); });
    return BRR_HIP_OK;
}

int brr_ragged_count(const uint8_t* bed, uint64_t stride, uint32_t N, uint32_t M, uint64_t* ones_off,
                     uint64_t* twos_off, uint64_t* miss_off, uint32_t threads)
{
    if (!bed || !ones_off || !twos_off || !miss_off || stride < ((uint64_t)N + 3) / 4) return BRR_HIP_EINVAL;
    parallel_for(M, threads, [&](uint32_t j) {
        const uint8_t* col = bed + (size_t)j * stride;
        uint64_t c[4] = {0, 0, 0, 0};
        for (uint32_t i = 0; i < N; ++i) ++c[code_at(col, i)];
        ones_off[j] = c[2];   // 2-bit value 2 = heterozygote (allele1=1, allele2=0)
        twos_off[j] = c[0];   // value 0 = homozygote counted allele (1,1)
        miss_off[j] = c[1];   // value 1 = missing (0,1)
    });
    // exclusive prefix sums into M+1 offsets
    uint64_t a = 0, b = 0, m = 0;
    for (uint32_t j = 0; j < M; ++j) {
        const uint64_t ca = ones_off[j], cb = twos_off[j], cm = miss_off[j];
        ones_off[j] = a; twos_off[j] = b; miss_off[j] = m;
        a += ca; b += cb; m += cm;
    }
    ones_off[M] = a; twos_off[M] = b; miss_off[M] = m;
    return BRR_HIP_OK;
}

int brr_ragged_fill(const uint8_t* bed, uint64_t stride, uint32_t N, uint32_t M, const uint64_t* ones_off,
                    const uint64_t* twos_off, const uint64_t* miss_off, int32_t* ones, int32_t* twos,
                    int32_t* miss, double* stats, uint32_t threads)
{
    if (!bed || !ones_off || !twos_off || !miss_off || !stats || stride < ((uint64_t)N + 3) / 4)
        return BRR_HIP_EINVAL;
    std::vector<uint8_t> invalid(M, 0);
    parallel_for(M, threads, [&](uint32_t j) {
        const uint8_t* col = bed + (size_t)j * stride;
        int32_t* o = ones + ones_off[j];
        int32_t* t = twos + twos_off[j];
        int32_t* ms = miss + miss_off[j];
        uint64_t no = 0, nt = 0, nm = 0;
        double mean = 0.0, sqrdZ = 0.0, Zsum = 0.0;
        for (uint32_t i = 0; i < N; ++i) {
            const unsigned v = code_at(col, i);
            const unsigned allele1 = !(v & 1u);
            const unsigned allele2 = !((v >> 1) & 1u);
            // SparseMarker::updateStatistics
            if (allele1 == 0 && allele2 == 1) {
            } else if (allele1 == 1 || allele2 == 1) {
                const double value = (double)(allele1 + allele2);
                mean += value;
                sqrdZ += value * value;
                Zsum += value;
            }
            if (allele1 == 0 && allele2 == 1) ms[nm++] = (int32_t)i;
            else if (allele1 == 1 && allele2 == 0) o[no++] = (int32_t)i;
            else if (allele1 == 1 && allele2 == 1) t[nt++] = (int32_t)i;
        }
        const double numInds = (double)N;
        mean /= numInds;
        const double sd = std::sqrt((sqrdZ - 2.0 * mean * Zsum + numInds * mean * mean) / (numInds - 1.0));
        double* st = stats + 4 * (size_t)j;
        st[0] = mean; st[1] = sd; st[2] = sqrdZ; st[3] = Zsum;
        if (no == 0 && nt == 0) invalid[j] = 1;
    });
    for (uint32_t j = 0; j < M; ++j)
        if (invalid[j]) return BRR_HIP_EINVAL;
    return BRR_HIP_OK;
}

} // extern "C"
