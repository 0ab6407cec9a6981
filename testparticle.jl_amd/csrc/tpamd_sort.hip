// tpamd_sort.hip -- radix partition of the adjoint's dL/dB records by cell
// bin (hipCUB over rocPRIM), kept in its own translation unit so that the
// pusher source does not carry rocPRIM's templates.
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include "tpamd_adjoint.h"

namespace tpamd {

// Sorts (key, record) pairs on key bits [begin_bit, end_bit).  With
// d_temp == nullptr only the temporary storage size is written.
template <typename R>
hipError_t adj_sort_records(void* d_temp, size_t& temp_bytes, const unsigned* keys_in,
                            unsigned* keys_out, const AdjRec<R>* vals_in, AdjRec<R>* vals_out,
                            long long n, int begin_bit, int end_bit, hipStream_t stream)
{
    return hipcub::DeviceRadixSort::SortPairs(d_temp, temp_bytes, keys_in, keys_out, vals_in,
                                              vals_out, (size_t)n, begin_bit, end_bit, stream);
}

template hipError_t adj_sort_records<float>(void*, size_t&, const unsigned*, unsigned*,
                                            const AdjRec<float>*, AdjRec<float>*, long long, int,
                                            int, hipStream_t);
template hipError_t adj_sort_records<double>(void*, size_t&, const unsigned*, unsigned*,
                                             const AdjRec<double>*, AdjRec<double>*, long long,
                                             int, int, hipStream_t);

}  // namespace tpamd
