// scan.h -- device-wide exclusive prefix sum (scan.hip).
#pragma once

#include <hip/hip_runtime.h>

namespace gs {

// Largest n scan_exclusive accepts.
long long scan_max_elements();
// Bytes of temporary storage scan_exclusive needs for n elements.
size_t scan_temp_bytes(long long n);
// out[i] = sum(in[0..i-1]); *total_out (device pointer, may be NULL) = sum(in).
// in == out is allowed.
hipError_t scan_exclusive(hipStream_t stream, const unsigned *in, unsigned *out, long long n,
                          unsigned *temp, unsigned *total_out);

}  // namespace gs
