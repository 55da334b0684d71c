/*
 * types.h -- scalar types of the GaDE hot path (tools/types.h:28-39).
 * csr_data_t is complex128 stored as interleaved (re, im) doubles;
 * csr_index_t is int32.
 */
#ifndef GADE_TYPES_H
#define GADE_TYPES_H

#include <stdint.h>

#ifdef __cplusplus
#include <complex>
typedef std::complex<double> csr_data_t;
#else
#include <complex.h>
typedef double complex csr_data_t;
#endif
typedef csr_data_t cdouble_t;
typedef int csr_index_t;
typedef csr_index_t dimType;
typedef csr_index_t indexType;

#endif
