// sfgs_prims.hpp -- device-wide primitives (exclusive scan, stable radix sort).
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>

namespace sfgs {
hipError_t scan_int(void *tmp, size_t &bytes, const int *in, int *out, int n, hipStream_t s);
hipError_t scan_ll(void *tmp, size_t &bytes, const long long *in, long long *out, int n, hipStream_t s);
hipError_t sort_pairs_i32(void *tmp, size_t &bytes, const int *kin, int *kout, const int *vin, int *vout,
                          int n, int end_bit, hipStream_t s);
hipError_t sort_pairs_u64(void *tmp, size_t &bytes, const unsigned long long *kin, unsigned long long *kout,
                          const int *vin, int *vout, int n, int end_bit, hipStream_t s);
}
