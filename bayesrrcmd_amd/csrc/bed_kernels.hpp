// bed_kernels.hpp -- a PLINK .bed payload decoded on the device.
//
// The reference turns every SNP column of the .bed file into a Marker on the
// host (preprocessgraph.cpp:24-60): the 2-bit decode, then DenseMarkerBuilder
// (densemarkerbuilder.cpp:23-60) or RaggedSparseMarkerBuilder with
// SparseMarker::updateStatistics (raggedsparsemarkerbuilder.cpp:25-55,
// sparsemarker.cpp:8-19). The device holds a column as its 2-bit codes, so the
// decode reduces to: repack each column to a 16-byte-aligned stride, count the
// four code values, and derive the per-marker table (dense) or statistics
// (ragged) from those counts and, for the dense table, from the column sums the
// builder forms in individual order.
//
// 2-bit values (LSB first): 0 = genotype 2, 1 = missing, 2 = genotype 1,
// 3 = genotype 0.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace brr {

constexpr int kBedThreads = 256;

__device__ __forceinline__ uint32_t bed_wave_sum_u32(uint32_t v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += (uint32_t)__shfl_xor((int)v, o, 64);
    return v;
}

// One workgroup per marker j. src is the packed payload (column j at byte
// j * cb, cb = ceil(N/4)) read as words; the caller provides at least 20
// readable bytes past M * cb. Output column j is col_words words at
// codes + j * col_words: the column's cb bytes, then zeros. counts[4j + v] =
// number of individuals i < N whose value is v (pad bits ignored). When stats is
// not null, stats[4j..4j+3] = {mean, sd, sqrdZ, Zsum} of SparseMarker and the
// first marker without genotype 1 or 2 calls lowers *bad. eigen = 1: the
// EigenSparse statistics instead (missing calls imputed at the mean,
// eigensparsemarkerbuilder.cpp:41-59), and only a marker without any nonzero
// call (eigensparsemarker.cpp:60-71) lowers *bad.
static __global__ void __launch_bounds__(kBedThreads)
bed_unpack_kernel(const uint32_t* __restrict__ src, uint64_t cb, uint32_t N, uint64_t col_words,
                  uint32_t* __restrict__ codes, uint32_t* __restrict__ counts, double* __restrict__ stats,
                  uint32_t* __restrict__ bad, int eigen)
{
    __shared__ uint32_t red[3][kBedThreads / 64];
    const uint32_t j = blockIdx.x;
    const int tid = threadIdx.x;
    const uint64_t base = (uint64_t)j * cb;
    uint4* dst = reinterpret_cast<uint4*>(codes + (size_t)j * col_words);
    const uint32_t nvec = (uint32_t)(col_words / 4u);
    uint32_t c0 = 0, c1 = 0, c2 = 0;
    for (uint32_t t = (uint32_t)tid; t < nvec; t += kBedThreads) {
        const uint64_t s = base + 16ull * t;                 // first payload byte of this 16-byte chunk
        const uint32_t* wp = src + (s >> 2);
        const uint32_t sh = 8u * (uint32_t)(s & 3u);
        uint32_t w[5];
#pragma unroll
        for (int k = 0; k < 5; ++k) w[k] = wp[k];
        uint32_t out[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            uint32_t x = (uint32_t)((((uint64_t)w[k + 1] << 32) | (uint64_t)w[k]) >> sh);
            const int64_t inside = (int64_t)cb - (int64_t)(16ull * t + 4u * (uint32_t)k);   // bytes of x in the column
            if (inside <= 0) x = 0u;
            else if (inside < 4) x &= (1u << (8 * (uint32_t)inside)) - 1u;
            out[k] = x;
            const int64_t npos = (int64_t)N - (int64_t)(64ull * t + 16u * (uint32_t)k);    // individuals < N in x
            uint32_t pm = 0x55555555u;
            if (npos <= 0) pm = 0u;
            else if (npos < 16) pm &= (1u << (2 * (uint32_t)npos)) - 1u;
            const uint32_t lo = x & pm, hi = (x >> 1) & pm;
            c0 += (uint32_t)__builtin_popcount(pm & ~(lo | hi));
            c1 += (uint32_t)__builtin_popcount(lo & ~hi);
            c2 += (uint32_t)__builtin_popcount(hi & ~lo);
        }
        dst[t] = make_uint4(out[0], out[1], out[2], out[3]);
    }
    c0 = bed_wave_sum_u32(c0);
    c1 = bed_wave_sum_u32(c1);
    c2 = bed_wave_sum_u32(c2);
    if ((tid & 63) == 0) { red[0][tid >> 6] = c0; red[1][tid >> 6] = c1; red[2][tid >> 6] = c2; }
    __syncthreads();
    if (tid == 0) {
        uint32_t n[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) n[q] = red[q][0] + red[q][1] + red[q][2] + red[q][3];
        counts[4 * (size_t)j + 0] = n[0];
        counts[4 * (size_t)j + 1] = n[1];
        counts[4 * (size_t)j + 2] = n[2];
        counts[4 * (size_t)j + 3] = N - n[0] - n[1] - n[2];
        if (stats) {
            // SparseMarker::updateStatistics sums the values 1 and 2 in
            // individual order; the sums are integers, so the counts give them
            // exactly. The rest is the same expression in the same order.
            const uint64_t ones = n[2], twos = n[0], nm = n[1];
            const double numInds = (double)N;
            double Zsum = (double)(ones + 2u * twos);
            double sqrdZ = (double)(ones + 4u * twos);
            const double mean = Zsum / numInds;
            if (eigen) {
                sqrdZ += (mean * mean) * (double)nm;
                Zsum += mean * (double)nm;
            }
            const double sd = sqrt((sqrdZ - 2.0 * mean * Zsum + numInds * mean * mean) / (numInds - 1.0));
            double* st = stats + 4 * (size_t)j;
            st[0] = mean;
            st[1] = sd;
            st[2] = sqrdZ;
            st[3] = Zsum;
            if (ones == 0 && twos == 0 && (!eigen || nm == 0)) atomicMin(bad, j);
        }
    }
}

// The running FP64 sum of tab(v) over the values v of one column's individuals
// 0..N-1, added one at a time in individual order. tab(v) = tabl[v * kBedThreads]:
// the calling thread's entries in LDS, one plane per value so that the lanes of
// a wave read distinct banks whatever their values.
__device__ __forceinline__ double bed_ordered_sum(const uint4* col, uint32_t N, const double* tabl)
{
    double acc = 0.0;
    const uint32_t nfull = N / 64u;
    uint4 next = nfull ? col[0] : make_uint4(0u, 0u, 0u, 0u);
    for (uint32_t q = 0; q < nfull; ++q) {
        const uint4 w = next;
        if (q + 1u < nfull) next = col[q + 1u];
        const uint32_t xs[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
        for (int k = 0; k < 4; ++k)
#pragma unroll
            for (uint32_t e = 0; e < 16u; ++e) acc = acc + tabl[((xs[k] >> (2u * e)) & 3u) * kBedThreads];
    }
    uint32_t rest = N - 64u * nfull;
    if (rest) {
        const uint4 w = col[nfull];
        const uint32_t xs[4] = {w.x, w.y, w.z, w.w};
        for (int k = 0; k < 4 && rest; ++k) {
            const uint32_t n = rest < 16u ? rest : 16u;
            for (uint32_t e = 0; e < n; ++e) acc = acc + tabl[((xs[k] >> (2u * e)) & 3u) * kBedThreads];
            rest -= n;
        }
    }
    return acc;
}

// DenseMarkerBuilder::endColumn from the counts and the codes: one lane per
// marker. The table holds the standardised value of dosage 0, 1, 2 and of a
// missing call. The mean of the non-missing dosages and, without missing calls,
// the column mean are integer sums (exact from the counts); with missing calls
// the column sum, and always the squared norm, are summed in individual order.
static __global__ void __launch_bounds__(kBedThreads)
bed_dense_table_kernel(const uint32_t* __restrict__ codes, uint64_t col_words, uint32_t N, uint32_t M,
                       const uint32_t* __restrict__ counts, double* __restrict__ lut)
{
    __shared__ double tab[4 * kBedThreads];
    const uint32_t j = blockIdx.x * kBedThreads + threadIdx.x;
    if (j >= M) return;
    double* tabl = tab + threadIdx.x;                 // this thread's entries, value v at tabl[v * kBedThreads]
    const uint4* col = reinterpret_cast<const uint4*>(codes + (size_t)j * col_words);
    const uint32_t n2 = counts[4 * (size_t)j + 0];   // value 0: dosage 2
    const uint32_t nm = counts[4 * (size_t)j + 1];   // value 1: missing
    const uint32_t n1 = counts[4 * (size_t)j + 2];   // value 2: dosage 1
    const double numInds = (double)N;
    const double sum = (double)((uint64_t)n1 + 2u * (uint64_t)n2);
    const double mean = sum / (numInds - (double)nm);
    double s = sum;                                   // integer terms: the index-order sum is exact
    if (nm) {
        tabl[0 * kBedThreads] = 2.0;
        tabl[1 * kBedThreads] = mean;
        tabl[2 * kBedThreads] = 1.0;
        tabl[3 * kBedThreads] = 0.0;
        s = bed_ordered_sum(col, N, tabl);
    }
    const double cmean = s / numInds;
    const double d0 = 0.0 - cmean, d1 = 1.0 - cmean, d2 = 2.0 - cmean, dm = mean - cmean;
    tabl[0 * kBedThreads] = d2 * d2;
    tabl[1 * kBedThreads] = dm * dm;
    tabl[2 * kBedThreads] = d1 * d1;
    tabl[3 * kBedThreads] = d0 * d0;
    const double sqn = bed_ordered_sum(col, N, tabl);
    const double sigma = 1.0 / (sqrt(sqn / (numInds - 1.0)));
    double* o = lut + 4 * (size_t)j;
    o[0] = d0 * sigma;
    o[1] = d1 * sigma;
    o[2] = d2 * sigma;
    o[3] = dm * sigma;
}

} // namespace brr
