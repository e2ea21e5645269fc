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

int brr_eigen_fill(const uint8_t* bed, uint64_t stride, uint32_t N, uint32_t M, const uint64_t* ones_off,
                   const uint64_t* twos_off, const uint64_t* miss_off, int32_t* ones, int32_t* twos,
                   int32_t* miss, double* stats, uint32_t threads)
{
    // EigenSparseMarkerBuilder (eigensparsemarkerbuilder.cpp:22-78): the ragged
    // lists; the statistics of SparseMarker::updateStatistics, then the missing
    // calls imputed at the mean in sqrdZ and Zsum before the sd.
    if (!bed || !ones_off || !twos_off || !miss_off || !stats || N < 2 || stride < ((uint64_t)N + 3) / 4)
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
            if (v == 1u) { ms[nm++] = (int32_t)i; continue; }
            if (v == 3u) continue;
            const double value = v == 2u ? 1.0 : 2.0;
            mean += value;
            sqrdZ += value * value;
            Zsum += value;
            if (v == 2u) o[no++] = (int32_t)i;
            else t[nt++] = (int32_t)i;
        }
        const double numInds = (double)N;
        mean /= numInds;
        sqrdZ += (mean * mean) * (double)nm;
        Zsum += mean * (double)nm;
        const double sd = std::sqrt((sqrdZ - 2.0 * mean * Zsum + numInds * mean * mean) / (numInds - 1.0));
        double* st = stats + 4 * (size_t)j;
        st[0] = mean; st[1] = sd; st[2] = sqrdZ; st[3] = Zsum;
        if (no + nt + nm == 0) invalid[j] = 1;   // EigenSparseMarker::isValid
    });
    for (uint32_t j = 0; j < M; ++j)
        if (invalid[j]) return BRR_HIP_EINVAL;
    return BRR_HIP_OK;
}

} // extern "C"

// ---------------------------------------------------------------------------
// The reference's preprocessed files, read straight into the device arenas.
// A preprocessed file is the markers' serialised bytes back to back; its index
// holds {pos, compressedSize, originalSize} (three u64) per marker
// (preprocessgraph.cpp:92-126, compression.cpp:137-154). With --compress each
// marker is one zlib stream (deflateInit level -1, Z_FINISH, compression.cpp:5-20).
#include <zlib.h>

namespace {

// The bytes of marker j: a view into data (uncompressed) or inflated into buf.
// Returns false on a bad index entry or a failed inflate.
bool marker_bytes(const uint8_t* data, uint64_t data_bytes, const uint64_t* index, uint32_t j, int compressed,
                  std::vector<uint8_t>& buf, const uint8_t** out, uint64_t* size)
{
    const uint64_t pos = index[3 * (size_t)j], csize = index[3 * (size_t)j + 1], osize = index[3 * (size_t)j + 2];
    if (pos > data_bytes || csize > data_bytes - pos) return false;
    if (!compressed) {
        if (csize != osize) return false;
        *out = data + pos;
        *size = csize;
        return true;
    }
    if (osize > (uint64_t)UINT32_MAX || csize > (uint64_t)UINT32_MAX) return false;
    buf.resize(osize ? osize : 1);
    z_stream s;
    std::memset(&s, 0, sizeof(s));
    if (inflateInit(&s) != Z_OK) return false;
    s.next_in = const_cast<Bytef*>(data + pos);
    s.avail_in = (uInt)csize;
    s.next_out = buf.data();
    s.avail_out = (uInt)osize;
    const int rc = inflate(&s, Z_FINISH);
    const uint64_t got = osize - s.avail_out;
    inflateEnd(&s);
    if (rc != Z_STREAM_END || got != osize) return false;
    *out = buf.data();
    *size = osize;
    return true;
}

// RaggedSparseMarker layout (raggedsparsemarker.cpp:17-66, sparsemarker.cpp:21-58):
// {mean, sd, sqrdZ, Zsum} f64, then for ones, twos, missing a u64 count and
// that many int32 indices. Sets the three counts and the list offsets in bytes.
bool ragged_layout(const uint8_t* p, uint64_t size, uint64_t n[3], uint64_t at[3])
{
    uint64_t off = 4 * sizeof(double);
    for (int l = 0; l < 3; ++l) {
        if (size < off + sizeof(uint64_t)) return false;
        std::memcpy(&n[l], p + off, sizeof(uint64_t));
        off += sizeof(uint64_t);
        if (n[l] > (size - off) / sizeof(int32_t)) return false;
        at[l] = off;
        off += n[l] * sizeof(int32_t);
    }
    return off == size;
}

// EigenSparseMarker layout (eigensparsemarker.cpp:17-58): {mean, sd, sqrdZ,
// Zsum} f64, an int64 count, count f64 values, count int32 indices. Sets the
// count and the byte offsets of the values and of the indices.
bool eigen_layout(const uint8_t* p, uint64_t size, uint64_t* count, uint64_t* vals_at, uint64_t* idx_at)
{
    const uint64_t head = 4 * sizeof(double) + sizeof(int64_t);
    if (size < head) return false;
    int64_t c = 0;
    std::memcpy(&c, p + 4 * sizeof(double), sizeof(int64_t));
    if (c < 0 || (uint64_t)c > (size - head) / (sizeof(double) + sizeof(int32_t))) return false;
    *count = (uint64_t)c;
    *vals_at = head;
    *idx_at = head + (uint64_t)c * sizeof(double);
    return *idx_at + (uint64_t)c * sizeof(int32_t) == size;
}

// The class of each nonzero of a serialised EigenSparseMarker: 0 = genotype 1
// (value 1), 1 = genotype 2 (value 2), 2 = missing (any other value: the
// builder stores the mean there, eigensparsemarkerbuilder.cpp:67-74).
inline int eigen_class(double v) { return v == 1.0 ? 0 : v == 2.0 ? 1 : 2; }

} // namespace

extern "C" {

int brr_eigen_read_counts(const uint8_t* data, uint64_t data_bytes, const uint64_t* index, uint32_t M,
                          int compressed, uint64_t* ones_off, uint64_t* twos_off, uint64_t* miss_off,
                          uint32_t threads)
{
    if (!data || !index || !ones_off || !twos_off || !miss_off || M == 0) return BRR_HIP_EINVAL;
    std::vector<uint8_t> bad(M, 0);
    uint64_t* offs[3] = {ones_off, twos_off, miss_off};
    parallel_for(M, threads, [&](uint32_t j) {
        std::vector<uint8_t> buf;
        const uint8_t* p;
        uint64_t size, count, va, ia;
        if (!marker_bytes(data, data_bytes, index, j, compressed, buf, &p, &size) ||
            !eigen_layout(p, size, &count, &va, &ia)) {
            bad[j] = 1;
            return;
        }
        uint64_t n[3] = {0, 0, 0};
        for (uint64_t q = 0; q < count; ++q) {
            double v;
            std::memcpy(&v, p + va + q * sizeof(double), sizeof(double));
            ++n[eigen_class(v)];
        }
        for (int l = 0; l < 3; ++l) offs[l][j + 1] = n[l];
    });
    for (uint32_t j = 0; j < M; ++j)
        if (bad[j]) return BRR_HIP_EINVAL;
    for (int l = 0; l < 3; ++l) {
        offs[l][0] = 0;
        for (uint32_t j = 0; j < M; ++j) offs[l][j + 1] += offs[l][j];
    }
    return BRR_HIP_OK;
}

int brr_eigen_read_fill(const uint8_t* data, uint64_t data_bytes, const uint64_t* index, uint32_t N, uint32_t M,
                        int compressed, const uint64_t* ones_off, const uint64_t* twos_off,
                        const uint64_t* miss_off, int32_t* ones, int32_t* twos, int32_t* miss, double* stats,
                        uint32_t threads)
{
    if (!data || !index || !ones_off || !twos_off || !miss_off || !stats || M == 0) return BRR_HIP_EINVAL;
    const uint64_t* offs[3] = {ones_off, twos_off, miss_off};
    int32_t* lists[3] = {ones, twos, miss};
    for (int l = 0; l < 3; ++l)
        if (offs[l][M] && !lists[l]) return BRR_HIP_EINVAL;
    std::vector<uint8_t> bad(M, 0);
    parallel_for(M, threads, [&](uint32_t j) {
        std::vector<uint8_t> buf;
        const uint8_t* p;
        uint64_t size, count, va, ia;
        if (!marker_bytes(data, data_bytes, index, j, compressed, buf, &p, &size) ||
            !eigen_layout(p, size, &count, &va, &ia)) {
            bad[j] = 1;
            return;
        }
        std::memcpy(stats + 4 * (size_t)j, p, 4 * sizeof(double));
        uint64_t n[3] = {0, 0, 0};
        int64_t prev = -1;
        for (uint64_t q = 0; q < count; ++q) {
            double v;
            int32_t idx;
            std::memcpy(&v, p + va + q * sizeof(double), sizeof(double));
            std::memcpy(&idx, p + ia + q * sizeof(int32_t), sizeof(int32_t));
            if (idx <= prev || idx >= (int64_t)N) { bad[j] = 1; return; }
            prev = idx;
            const int l = eigen_class(v);
            if (offs[l][j] + n[l] >= offs[l][j + 1]) { bad[j] = 1; return; }
            lists[l][offs[l][j] + n[l]++] = idx;
        }
        for (int l = 0; l < 3; ++l)
            if (n[l] != offs[l][j + 1] - offs[l][j]) bad[j] = 1;
    });
    for (uint32_t j = 0; j < M; ++j)
        if (bad[j]) return BRR_HIP_EINVAL;
    return BRR_HIP_OK;
}

int brr_ragged_read_counts(const uint8_t* data, uint64_t data_bytes, const uint64_t* index, uint32_t M,
                           int compressed, uint64_t* ones_off, uint64_t* twos_off, uint64_t* miss_off,
                           uint32_t threads)
{
    if (!data || !index || !ones_off || !twos_off || !miss_off || M == 0) return BRR_HIP_EINVAL;
    std::vector<uint8_t> bad(M, 0);
    uint64_t* offs[3] = {ones_off, twos_off, miss_off};
    parallel_for(M, threads, [&](uint32_t j) {
        std::vector<uint8_t> buf;
        const uint8_t* p;
        uint64_t size, n[3], at[3];
        if (!marker_bytes(data, data_bytes, index, j, compressed, buf, &p, &size) || !ragged_layout(p, size, n, at)) {
            bad[j] = 1;
            return;
        }
        for (int l = 0; l < 3; ++l) offs[l][j + 1] = n[l];
    });
    for (uint32_t j = 0; j < M; ++j)
        if (bad[j]) return BRR_HIP_EINVAL;
    for (int l = 0; l < 3; ++l) {
        offs[l][0] = 0;
        for (uint32_t j = 0; j < M; ++j) offs[l][j + 1] += offs[l][j];
    }
    return BRR_HIP_OK;
}

int brr_ragged_read_fill(const uint8_t* data, uint64_t data_bytes, const uint64_t* index, uint32_t N, uint32_t M,
                         int compressed, const uint64_t* ones_off, const uint64_t* twos_off,
                         const uint64_t* miss_off, int32_t* ones, int32_t* twos, int32_t* miss, double* stats,
                         uint32_t threads)
{
    if (!data || !index || !ones_off || !twos_off || !miss_off || !stats || M == 0) return BRR_HIP_EINVAL;
    const uint64_t* offs[3] = {ones_off, twos_off, miss_off};
    int32_t* lists[3] = {ones, twos, miss};
    for (int l = 0; l < 3; ++l)
        if (offs[l][M] && !lists[l]) return BRR_HIP_EINVAL;
    std::vector<uint8_t> bad(M, 0);
    parallel_for(M, threads, [&](uint32_t j) {
        std::vector<uint8_t> buf;
        const uint8_t* p;
        uint64_t size, n[3], at[3];
        if (!marker_bytes(data, data_bytes, index, j, compressed, buf, &p, &size) || !ragged_layout(p, size, n, at)) {
            bad[j] = 1;
            return;
        }
        std::memcpy(stats + 4 * (size_t)j, p, 4 * sizeof(double));
        for (int l = 0; l < 3; ++l) {
            if (n[l] != offs[l][j + 1] - offs[l][j]) { bad[j] = 1; return; }
            int32_t* dst = lists[l] + offs[l][j];
            std::memcpy(dst, p + at[l], n[l] * sizeof(int32_t));
            int64_t prev = -1;
            for (uint64_t q = 0; q < n[l]; ++q) {
                if (dst[q] <= prev || dst[q] >= (int64_t)N) { bad[j] = 1; return; }
                prev = dst[q];
            }
        }
    });
    for (uint32_t j = 0; j < M; ++j)
        if (bad[j]) return BRR_HIP_EINVAL;
    return BRR_HIP_OK;
}

int brr_dense_read_ppbed(const uint8_t* data, uint64_t data_bytes, const uint64_t* index, uint32_t N, uint32_t M,
                         int compressed, uint8_t* bed, double* lut, uint32_t threads)
{
    // DenseMarker (densemarker.cpp:34-49): N doubles, the standardised column.
    // Its at most four distinct values become the column's table; individual i
    // gets the 2-bit code of its value's slot, so lut[slot(code_i)] == Cx[i]
    // bit for bit (slots 0, 1, 2, 3 take codes 11, 10, 00, 01: the table order
    // of brr_hip_upload_dense).
    if (!data || !index || !bed || !lut || N == 0 || M == 0) return BRR_HIP_EINVAL;
    const uint64_t stride = ((uint64_t)N + 3) / 4;
    std::vector<uint8_t> bad(M, 0);
    parallel_for(M, threads, [&](uint32_t j) {
        std::vector<uint8_t> buf;
        const uint8_t* p;
        uint64_t size;
        if (!marker_bytes(data, data_bytes, index, j, compressed, buf, &p, &size) ||
            size != (uint64_t)N * sizeof(double)) {
            bad[j] = 1;
            return;
        }
        static const unsigned kSlotCode[4] = {3u, 2u, 0u, 1u};
        uint64_t vals[4];
        int nv = 0;
        uint8_t* col = bed + (size_t)j * stride;
        std::memset(col, 0, stride);
        for (uint32_t i = 0; i < N; ++i) {
            uint64_t b;
            std::memcpy(&b, p + (size_t)i * sizeof(double), sizeof(double));
            int s = 0;
            while (s < nv && vals[s] != b) ++s;
            if (s == nv) {
                if (nv == 4) { bad[j] = 1; return; }
                vals[nv++] = b;
            }
            col[i >> 2] |= (uint8_t)(kSlotCode[s] << (2 * (i & 3u)));
        }
        // pad calls of the last byte: genotype 0 (code 11), as in a .bed
        for (uint32_t i = N; i < stride * 4; ++i) col[i >> 2] |= (uint8_t)(3u << (2 * (i & 3u)));
        double* lt = lut + 4 * (size_t)j;
        for (int s = 0; s < 4; ++s) {
            double v = 0.0;
            if (s < nv) std::memcpy(&v, &vals[s], sizeof(double));
            lt[s] = v;
        }
    });
    for (uint32_t j = 0; j < M; ++j)
        if (bad[j]) return BRR_HIP_EINVAL;
    return BRR_HIP_OK;
}

} // extern "C"
