/*
 * brr_synth.h -- deterministic synthetic PLINK genotype generator exported by
 * libbrr_hip.so for tests and benchmarks. It has no counterpart in the
 * reference, which reads real .bed files (src/preprocessgraph.cpp:152-244).
 */
#ifndef BRR_SYNTH_H
#define BRR_SYNTH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Fills num_snps SNP-major columns of ceil(num_inds/4) PLINK bytes at
 * bed + j * column_stride. maf (may be NULL) receives each column's counted-
 * allele frequency. Returns 0 on success, 1 on bad arguments, 2 if a column
 * stayed monomorphic after 1000 redraws. */
int brr_synth_bed(uint64_t seed, uint32_t num_inds, uint32_t num_snps, double maf_lo, double maf_hi,
                  double miss_rate, uint8_t* bed, uint64_t column_stride, double* maf, uint32_t threads);

/* Genetic values of synthetic phenotypes: out (num_inds x num_traits, trait
 * fastest) = X beta, X the standardised columns (lut: num_snps x 4 doubles for
 * dosage 0, 1, 2, missing, as brr_dense_luts returns them) and beta
 * num_snps x num_traits (mostly zero). Returns 0, or 1 on bad arguments. */
int brr_synth_xb(const uint8_t* bed, uint64_t column_stride, uint32_t num_inds, uint32_t num_snps,
                 const double* lut, uint32_t num_traits, const double* beta, double* out, uint32_t threads);

#ifdef __cplusplus
}
#endif
#endif
