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

// The receptor displacement as a GEMM, the form the FP32 kernels do not use
// (they displace each near receptor atom inside the scoring wave): for one
// swarm, D[g][q] = sum_m e[g][m] M[m][q] for its glowworms g and the 192
// coordinates q = comp * 64 + atom of one receptor tile that some glowworm of
// the swarm can reach, FP32 on the matrix cores.  One wave per (swarm, tile,
// 32 glowworms): six 32 x 32 tiles of v_mfma_f32_32x32x2f32 (two modes per
// instruction; A = the glowworms' extents, B = the tile's modes from the
// [m][3][slot] FP32 array), D written to HBM as [item][32 g][192 q].
typedef float floatx16 __attribute__((ext_vector_type(16)));

__global__ __launch_bounds__(64) void anm_gemm_probe_kernel(const DevComplex *cx, const double *pos,
                                                            int P, int G, const int *pairs,
                                                            size_t first, int chunks, float *out) {
    const size_t item = first + blockIdx.x;
    const size_t pr = item / (size_t)chunks;
    const int ch = (int)(item % (size_t)chunks);
    const int s = pairs[3 * pr], k = pairs[3 * pr + 1], t = pairs[3 * pr + 2];
    const DevComplex &c = cx[k];
    const int lane = threadIdx.x;
    const int row = lane & 31, kk = lane >> 5;
    const int g = ch * 32 + row;
    const int slots = c.ntiles * kTile;
    const double *e = pos + ((size_t)s * G + min(g, G - 1)) * P + 7;
    floatx16 acc[6];
#pragma unroll
    for (int o = 0; o < 6; o++)
#pragma unroll
        for (int v = 0; v < 16; v++) acc[o][v] = 0.f;
    for (int m0 = 0; m0 < c.nmr; m0 += 2) {
        const int m = m0 + kk;
        const float av = (g < G && m < c.nmr) ? (float)e[m] : 0.f;
#pragma unroll
        for (int o = 0; o < 6; o++) {
            const int q = 32 * o + row;  // B column of this lane: comp q / 64, atom q % 64
            const float bv = m < c.nmr ? gld(c.rmod32, (size_t)(m * 3 + (q >> 6)) * slots +
                                                         t * kTile + (q & 63))
                                       : 0.f;
            acc[o] = __builtin_amdgcn_mfma_f32_32x32x2f32(av, bv, acc[o], 0, 0, 0);
        }
    }
    // accumulator v of lane l: row 8 (v / 4) + 4 (l / 32) + v % 4, column l % 32
    float *D = out + (size_t)blockIdx.x * 32 * 192;
#pragma unroll
    for (int o = 0; o < 6; o++)
#pragma unroll
        for (int v = 0; v < 16; v++) {
            const int i = 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
            D[(size_t)i * 192 + 32 * o + row] = acc[o][v];
        }
}

hipError_t launch_anm_gemm_probe(const DevComplex *cx, const double *pos, int P, int G,
                                 const int *pairs, size_t first, size_t count, int chunks,
                                 float *out, hipStream_t s) {
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(anm_gemm_probe_kernel, dim3((unsigned)count), dim3(64), 0, s, cx, pos, P, G,
                       pairs, first, chunks, out);
    return hipGetLastError();
}

hipError_t launch_gather_probe(const float *table, unsigned n, int per_thread, float *out,
                               int blocks, hipStream_t s) {
    hipLaunchKernelGGL(gather_probe_kernel, dim3(blocks), dim3(256), 0, s, table, n, per_thread,
                       out);
    return hipGetLastError();
}

}  // namespace sd
