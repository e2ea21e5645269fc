/*
 * markers.c -- CPU ORACLE (test infrastructure only): genotype decoding and
 * the Dense / RaggedSparse marker builders of the reference.
 */
#include "brr_oracle.h"

#include <math.h>
#include <string.h>

int orc_check_bed_magic(const uint8_t* h)
{
    /* preprocessgraph.cpp:195-200 */
    return (h[0] == 0x6c && h[1] == 0x1b && h[2] == 0x01) ? 0 : -1;
}

void orc_decode_bed_column(const uint8_t* col, uint32_t N, uint8_t* a1, uint8_t* a2)
{
    /* preprocessgraph.cpp:34-57: for each byte, bits k = 0,2,4,6 (LSB first):
     * allele1 = !b[k], allele2 = !b[k+1]; pad bits after the last individual
     * are ignored. */
    uint32_t i = 0;
    size_t byte = 0;
    while (i < N) {
        const uint8_t b = col[byte++];
        unsigned k = 0;
        while (k < 7 && i < N) {
            a1[i] = (uint8_t)!((b >> k) & 1u); ++k;
            a2[i] = (uint8_t)!((b >> k) & 1u); ++k;
            ++i;
        }
    }
}

void orc_dense_build(const uint8_t* col, uint32_t N, double* Cx)
{
    /* DenseMarkerBuilder::processAllele (densemarkerbuilder.cpp:23-41) and
     * endColumn (:43-60). */
    uint8_t a1, a2;
    double sum = 0.0;
    uint32_t nmiss = 0;
    for (uint32_t i = 0; i < N; ++i) {
        const uint8_t b = col[i >> 2];
        const unsigned k = (i & 3u) * 2u;
        a1 = (uint8_t)!((b >> k) & 1u);
        a2 = (uint8_t)!((b >> (k + 1)) & 1u);
        if (a1 == 0 && a2 == 1) {
            Cx[i] = -1.0;          /* placeholder, replaced below */
            ++nmiss;
        } else {
            const double value = (double)(a1 + a2);
            Cx[i] = value;
            sum += value;
        }
    }
    const double mean = sum / ((double)N - (double)nmiss);
    if (nmiss) {
        for (uint32_t i = 0; i < N; ++i) {
            const uint8_t b = col[i >> 2];
            const unsigned k = (i & 3u) * 2u;
            if (((b >> k) & 3u) == 1u) Cx[i] = mean;
        }
    }
    /* snpData.array() -= snpData.mean(); */
    double s = 0.0;
    for (uint32_t i = 0; i < N; ++i) s += Cx[i];
    const double cmean = s / (double)N;
    for (uint32_t i = 0; i < N; ++i) Cx[i] -= cmean;
    double sqn = 0.0;
    for (uint32_t i = 0; i < N; ++i) sqn += Cx[i] * Cx[i];
    const double sigma = 1.0 / (sqrt(sqn / ((double)N - 1.0)));
    for (uint32_t i = 0; i < N; ++i) Cx[i] *= sigma;
}

void orc_dense_lut(const uint8_t* col, uint32_t N, double lut[4])
{
    /* Same arithmetic as orc_dense_build, with the per-element values collected
     * per dosage class. Sums run in individual order, as there. */
    double sum = 0.0;
    uint32_t nmiss = 0;
    for (uint32_t i = 0; i < N; ++i) {
        const unsigned code = (col[i >> 2] >> ((i & 3u) * 2u)) & 3u;
        if (code == 1u) { ++nmiss; continue; }
        const double value = (double)((code == 0u) ? 2 : (code == 2u) ? 1 : 0);
        sum += value;
    }
    const double mean = sum / ((double)N - (double)nmiss);
    const double raw[4] = {0.0, 1.0, 2.0, mean};   /* dosage 0,1,2, missing */
    double s = 0.0;
    for (uint32_t i = 0; i < N; ++i) {
        const unsigned code = (col[i >> 2] >> ((i & 3u) * 2u)) & 3u;
        const int d = (code == 0u) ? 2 : (code == 2u) ? 1 : (code == 3u) ? 0 : 3;
        s += raw[d];
    }
    const double cmean = s / (double)N;
    double centred[4];
    for (int d = 0; d < 4; ++d) centred[d] = raw[d] - cmean;
    double sqn = 0.0;
    for (uint32_t i = 0; i < N; ++i) {
        const unsigned code = (col[i >> 2] >> ((i & 3u) * 2u)) & 3u;
        const int d = (code == 0u) ? 2 : (code == 2u) ? 1 : (code == 3u) ? 0 : 3;
        sqn += centred[d] * centred[d];
    }
    const double sigma = 1.0 / (sqrt(sqn / ((double)N - 1.0)));
    for (int d = 0; d < 4; ++d) lut[d] = centred[d] * sigma;
}

int orc_ragged_build(const uint8_t* col, uint32_t N,
                     int32_t* ones, uint64_t* n_ones,
                     int32_t* twos, uint64_t* n_twos,
                     int32_t* miss, uint64_t* n_miss,
                     double stats[4])
{
    /* RaggedSparseMarkerBuilder::processAllele (raggedsparsemarkerbuilder.cpp:25-41)
     * with SparseMarker::updateStatistics (sparsemarker.cpp:8-19) and endColumn
     * (raggedsparsemarkerbuilder.cpp:43-55). */
    double mean = 0.0, sqrdZ = 0.0, Zsum = 0.0;
    uint64_t no = 0, nt = 0, nm = 0;
    for (uint32_t i = 0; i < N; ++i) {
        const uint8_t b = col[i >> 2];
        const unsigned k = (i & 3u) * 2u;
        const unsigned allele1 = !((b >> k) & 1u);
        const unsigned allele2 = !((b >> (k + 1)) & 1u);
        /* updateStatistics */
        if (allele1 == 0 && allele2 == 1) {
            /* missing: no statistics */
        } else if (allele1 == 1 || allele2 == 1) {
            const double value = (double)(allele1 + allele2);
            mean += value;
            sqrdZ += value * value;
            Zsum += value;
        }
        if (allele1 == 0 && allele2 == 1) miss[nm++] = (int32_t)i;
        else if (allele1 == 1 && allele2 == 0) ones[no++] = (int32_t)i;
        else if (allele1 == 1 && allele2 == 1) twos[nt++] = (int32_t)i;
    }
    const double numInds = (double)N;
    mean /= numInds;
    const double sd = sqrt((sqrdZ - 2.0 * mean * Zsum + numInds * mean * mean) / (numInds - 1.0));
    stats[0] = mean; stats[1] = sd; stats[2] = sqrdZ; stats[3] = Zsum;
    *n_ones = no; *n_twos = nt; *n_miss = nm;
    return (no == 0 && nt == 0) ? 0 : 1;
}

size_t orc_ragged_serialise(const double stats[4],
                            const int32_t* ones, uint64_t n_ones,
                            const int32_t* twos, uint64_t n_twos,
                            const int32_t* miss, uint64_t n_miss, uint8_t* buf)
{
    const size_t total = 4 * sizeof(double) + 3 * sizeof(uint64_t)
        + sizeof(int32_t) * (size_t)(n_ones + n_twos + n_miss);
    if (!buf) return total;
    size_t off = 0;
    memcpy(buf + off, stats, 4 * sizeof(double)); off += 4 * sizeof(double);
    const int32_t* lists[3] = {ones, twos, miss};
    const uint64_t counts[3] = {n_ones, n_twos, n_miss};
    for (int l = 0; l < 3; ++l) {
        memcpy(buf + off, &counts[l], sizeof(uint64_t)); off += sizeof(uint64_t);
        if (counts[l]) {
            memcpy(buf + off, lists[l], sizeof(int32_t) * (size_t)counts[l]);
            off += sizeof(int32_t) * (size_t)counts[l];
        }
    }
    return off;
}

int orc_eigen_build(const uint8_t* col, uint32_t N,
                    int32_t* ones, uint64_t* n_ones,
                    int32_t* twos, uint64_t* n_twos,
                    int32_t* miss, uint64_t* n_miss,
                    double stats[4])
{
    /* EigenSparseMarkerBuilder::processAllele (eigensparsemarkerbuilder.cpp:22-39):
     * the same statistics as the ragged builder (SparseMarker::updateStatistics)
     * and the same calls kept (1, 2, missing); endColumn (:41-78) then adds the
     * mean-imputed missing calls to sqrdZ and Zsum before the sd. */
    double mean = 0.0, sqrdZ = 0.0, Zsum = 0.0;
    uint64_t no = 0, nt = 0, nm = 0;
    for (uint32_t i = 0; i < N; ++i) {
        const uint8_t b = col[i >> 2];
        const unsigned k = (i & 3u) * 2u;
        const unsigned allele1 = !((b >> k) & 1u);
        const unsigned allele2 = !((b >> (k + 1)) & 1u);
        if (allele1 == 0 && allele2 == 1) {
            miss[nm++] = (int32_t)i;
        } else if (allele1 == 1 || allele2 == 1) {
            const double value = (double)(allele1 + allele2);
            mean += value;
            sqrdZ += value * value;
            Zsum += value;
            if (allele1 + allele2 == 1) ones[no++] = (int32_t)i;
            else twos[nt++] = (int32_t)i;
        }
    }
    const double numInds = (double)N;
    mean /= numInds;
    sqrdZ += (mean * mean) * (double)nm;
    Zsum += mean * (double)nm;
    const double sd = sqrt((sqrdZ - 2.0 * mean * Zsum + numInds * mean * mean) / (numInds - 1.0));
    stats[0] = mean; stats[1] = sd; stats[2] = sqrdZ; stats[3] = Zsum;
    *n_ones = no; *n_twos = nt; *n_miss = nm;
    /* EigenSparseMarker::isValid (eigensparsemarker.cpp:60-71): some nonzero */
    return (no + nt + nm) > 0 ? 1 : 0;
}

size_t orc_eigen_serialise(const double stats[4],
                           const int32_t* ones, uint64_t n_ones,
                           const int32_t* twos, uint64_t n_twos,
                           const int32_t* miss, uint64_t n_miss, uint8_t* buf)
{
    /* EigenSparseMarker::write (eigensparsemarker.cpp:41-58): the SparseMarker
     * statistics, an Eigen::Index (int64) count, the values (f64) and the inner
     * indices (int32) of Zg in index order; a missing call's value is the mean
     * (eigensparsemarkerbuilder.cpp:67-74). */
    const uint64_t count = n_ones + n_twos + n_miss;
    const size_t total = 4 * sizeof(double) + sizeof(int64_t) + count * (sizeof(double) + sizeof(int32_t));
    if (!buf) return total;
    size_t off = 0;
    memcpy(buf + off, stats, 4 * sizeof(double)); off += 4 * sizeof(double);
    const int64_t c64 = (int64_t)count;
    memcpy(buf + off, &c64, sizeof(int64_t)); off += sizeof(int64_t);
    uint8_t* vals = buf + off;
    uint8_t* idxs = buf + off + count * sizeof(double);
    uint64_t a = 0, b = 0, m = 0;
    for (uint64_t q = 0; q < count; ++q) {
        const int64_t ia = a < n_ones ? ones[a] : INT64_MAX;
        const int64_t ib = b < n_twos ? twos[b] : INT64_MAX;
        const int64_t im = m < n_miss ? miss[m] : INT64_MAX;
        double v;
        int32_t idx;
        if (ia < ib && ia < im) { v = 1.0; idx = (int32_t)ia; ++a; }
        else if (ib < im) { v = 2.0; idx = (int32_t)ib; ++b; }
        else { v = stats[0]; idx = (int32_t)im; ++m; }
        memcpy(vals + q * sizeof(double), &v, sizeof(double));
        memcpy(idxs + q * sizeof(int32_t), &idx, sizeof(int32_t));
    }
    return total;
}
