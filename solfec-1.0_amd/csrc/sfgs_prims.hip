// sfgs_prims.hip -- device-wide scan and stable radix sort (rocPRIM via hipCUB),
// kept in their own translation unit because they are slow to compile.
#include <hipcub/hipcub.hpp>
#include "sfgs_prims.hpp"

namespace sfgs {

hipError_t scan_int(void *tmp, size_t &bytes, const int *in, int *out, int n, hipStream_t s)
{
  return hipcub::DeviceScan::ExclusiveSum(tmp, bytes, in, out, n, s);
}

hipError_t scan_ll(void *tmp, size_t &bytes, const long long *in, long long *out, int n, hipStream_t s)
{
  return hipcub::DeviceScan::ExclusiveSum(tmp, bytes, in, out, n, s);
}

hipError_t sort_pairs_i32(void *tmp, size_t &bytes, const int *kin, int *kout, const int *vin, int *vout,
                          int n, int end_bit, hipStream_t s)
{
  return hipcub::DeviceRadixSort::SortPairs(tmp, bytes, kin, kout, vin, vout, n, 0, end_bit, s);
}

hipError_t sort_pairs_u64(void *tmp, size_t &bytes, const unsigned long long *kin, unsigned long long *kout,
                          const int *vin, int *vout, int n, int end_bit, hipStream_t s)
{
  return hipcub::DeviceRadixSort::SortPairs(tmp, bytes, kin, kout, vin, vout, n, 0, end_bit, s);
}

} // namespace sfgs
