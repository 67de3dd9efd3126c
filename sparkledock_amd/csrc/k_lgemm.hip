// The ligand's normal-mode displacement of every glowworm of a run as an
// FP32 matrix-core GEMM, (poses x modes) . (modes x 3 Nl): for complex k
// (nm ligand modes, nlb blocks of 64 atom slots)
//
//     D[g][b][c][s] = sum_m e[g][m] M[b][m][c][s],
//
// e the glowworm's ligand extents rounded to FP32 and M the complex's FP32
// ligand modes by block slot.  The reference adds the modes to each ligand
// atom of each pose in FP64 (prepare_gso_tasks.cc:487-504); its GPU path
// materialises every pose's displaced ligand (prepare_gso_task.cu:268-333).
// Every ligand atom of every glowworm is scored each step, so this product
// is needed whole; the FP32 scoring kernels read D in their set-up, add it to
// the FP64 coordinates, rotate and round to their grid, and their guard band
// covers D's rounding (a sum of nm FP32 products in any order of correctly
// rounded operations is within (nm + 2) 2^-24 S of the exact one, S the
// pose's ligand displacement bound; the band takes twice that).
//
// One wave per (32 glowworms of one complex, one ligand block): the block's
// 3 x 64 coordinates are six 32 x 32 output tiles, each nm / 2 steps of
// v_mfma_f32_32x32x2f32 (A: the 32 glowworms' extents of two modes, B: the
// two modes of 32 coordinates), written to HBM as rows of 128 bytes.
#include "dev_common.h"

namespace sd {

typedef float floatx16 __attribute__((ext_vector_type(16)));

__global__ __launch_bounds__(64) void ligand_disp_kernel(LigDispArgs a) {
    const long long item = (long long)blockIdx.x;
    // the complex of this item: the last k with item0[k] <= item
    int lo = 0, hi = a.n_cx;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (a.item0[mid] <= item)
            lo = mid;
        else
            hi = mid;
    }
    const int k = lo;
    const DevComplex &c = a.cx[k];
    const int nlb = c.nlblocks, nm = c.nml;
    const long long local = item - a.item0[k];
    const int b = (int)(local % nlb);
    const long long g0 = a.glow0[k] + 32 * (local / nlb);  // first glowworm of the tile
    const long long gend = a.glow0[k + 1];
    const int lane = threadIdx.x, row = lane & 31, kk = lane >> 5;
    // A operand of lane l: row l % 32 (glowworm g0 + row), mode m0 + l / 32
    const long long g = g0 + row;
    const double *e = a.pos + (size_t)(g < gend ? g : gend - 1) * a.P + 7 + a.nmr;
    const float *M = c.lblkm + (size_t)b * 3 * nm * kBlock;
    floatx16 acc[6];
#pragma unroll
    for (int o = 0; o < 6; o++)
#pragma unroll
        for (int v = 0; v < 16; v++) acc[o][v] = 0.f;
    // Every load is unconditional (a mode index past nm is clamped to the last
    // mode, whose finite values then meet an A of zero), so the loads of two
    // steps (four modes) are issued together: one round trip per two steps
    // instead of one per MFMA.
    for (int m0 = 0; m0 < nm; m0 += 4) {
        int mc[2];
        double ev[2];
        float bv[2][6];
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int m = m0 + 2 * h + kk;
            mc[h] = m < nm ? m : nm - 1;
            ev[h] = gld(e, mc[h]);
#pragma unroll
            for (int o = 0; o < 6; o++) {
                // B operand of lane l: mode m0 + 2h + l / 32, column q = 32 o + l % 32
                // (coordinate q / 64 of slot q % 64; M row 3 m + coordinate)
                const int q = 32 * o + row;
                bv[h][o] = gld(M, (size_t)(3 * mc[h] + (q >> 6)) * kBlock + (q & 63));
            }
        }
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int m = m0 + 2 * h + kk;
            const float av = (g < gend && m < nm) ? (float)ev[h] : 0.f;
#pragma unroll
            for (int o = 0; o < 6; o++)
                acc[o] = __builtin_amdgcn_mfma_f32_32x32x2f32(av, bv[h][o], acc[o], 0, 0, 0);
        }
    }
    // accumulator v of lane l: row 8 (v / 4) + 4 (l / 32) + v % 4, column l % 32
    float *D = a.ldisp + a.ldisp_base[k] + (long long)b * 3 * kBlock;
    const long long stride = (long long)nlb * 3 * kBlock;
#pragma unroll
    for (int v = 0; v < 16; v++) {
        const long long gi = g0 + 8 * (v >> 2) + 4 * kk + (v & 3);
        if (gi < gend) {
            float *Dg = D + gi * stride;
#pragma unroll
            for (int o = 0; o < 6; o++) Dg[32 * o + row] = acc[o][v];
        }
    }
}

hipError_t launch_ligand_disp(const LigDispArgs &a, hipStream_t s) {
    if (a.n_items <= 0) return hipSuccess;
    if (a.n_items > 0x7fffffffLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL(ligand_disp_kernel, dim3((unsigned)a.n_items), dim3(64), 0, s, a);
    return hipGetLastError();
}

}  // namespace sd
