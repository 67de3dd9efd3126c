// Random-gather rate of the L2 over a table of the DFIRE table's size: the
// ceiling for the scoring kernel's data-dependent table reads.  Each thread
// draws indices with an xorshift hash (no index loads), gathers eight floats
// at a time, sums them and writes one value.
#include "dev_common.h"

namespace sd {

__global__ __launch_bounds__(256) void gather_probe_kernel(const float *table, unsigned n,
                                                           int per_thread, float *out) {
    const unsigned tid = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned h = tid * 2654435761u + 0x9e3779b9u;
    float acc = 0.f;
    for (int i = 0; i < per_thread; i += 8) {
        float v[8];
#pragma unroll
        for (int k = 0; k < 8; k++) {
            h ^= h << 13;
            h ^= h >> 17;
            h ^= h << 5;
            v[k] = gld(table, __umulhi(h, n));  // uniform in [0, n)
        }
#pragma unroll
        for (int k = 0; k < 8; k++) acc += v[k];
    }
    out[tid] = acc;
}

hipError_t launch_gather_probe(const float *table, unsigned n, int per_thread, float *out,
                               int blocks, hipStream_t s) {
    hipLaunchKernelGGL(gather_probe_kernel, dim3(blocks), dim3(256), 0, s, table, n, per_thread,
                       out);
    return hipGetLastError();
}

}  // namespace sd
