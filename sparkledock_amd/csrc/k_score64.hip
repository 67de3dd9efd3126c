// Plain FP64 scoring: one workgroup per glowworm, brute force over all
// receptor x ligand pairs (the per-complex baseline form; its reference design
// is the calculate_dfire kernel of source/dock/src/cuda/gpu_baseline:492).
//
// Layout per workgroup:
//   LDS  ligand pose in FP64 SoA (x[nl], y[nl], z[nl]) + type*20 [nl]
//        + two 256-entry row-sum buffers
//   lane one receptor row i at a time (i = chunk*256 + tid), its FP64
//        coordinates displaced in registers; ligand atoms are LDS broadcasts
//
// The energy is summed exactly as prepare_gso_tasks.cc:1539-1570 does: each
// row sum v_i runs over j in order (one lane), and thread 0 adds the v_i in
// row order (double-buffered, one barrier per chunk), so the FP64 energy is
// bit-identical to the CPU reference.  The epilogue applies the score affine
// map and the luciferin update (prepare_gso_tasks.cc:1571, 558).
#include "dev_common.h"

namespace sd {

constexpr int kScore64Threads = 256;

__global__ __launch_bounds__(kScore64Threads) void score64_kernel(ScoreArgs a) {
    extern __shared__ double lds[];
    const int gid = blockIdx.x;
    if (gid >= a.n_glow) return;
    const int tid = threadIdx.x;
    const int swarm = gid / a.G;
    const DevComplex &c = a.cx[a.swarm_cx[swarm]];
    const int nl = c.nl, nr = c.nr;
    double *lx = lds;
    double *ly = lx + nl;
    double *lz = ly + nl;
    double *vbuf = lz + nl;                       // [2][256]
    int *lt = (int *)(vbuf + 2 * kScore64Threads);  // [nl]

    const double *pose = a.pos + (size_t)gid * a.P;
    double inv[4];
    quat_inverse(pose + 3, inv);
    for (int j = tid; j < nl; j += kScore64Threads) {
        double x, y, z;
        ligand_atom64(c, pose, inv, j, x, y, z);
        lx[j] = x;
        ly[j] = y;
        lz[j] = z;
        lt[j] = gld(c.ltype, j);
    }
    if (tid == 0) {
        double r[3];
        ref_point64(c, pose, inv, r);
        a.ref[(size_t)gid * 3 + 0] = r[0];
        a.ref[(size_t)gid * 3 + 1] = r[1];
        a.ref[(size_t)gid * 3 + 2] = r[2];
    }
    __syncthreads();

    const double *ext = pose + 7;
    const double *table = c.table64;
    double energy = 0.0;  // thread 0 only
    long long npairs = 0;
    const int nchunks = (nr + kScore64Threads - 1) / kScore64Threads;
    for (int ch = 0; ch < nchunks; ch++) {
        const int i = ch * kScore64Threads + tid;
        double v = 0.0;
        if (i < nr) {
            double rx, ry, rz;
            receptor_atom64(c, ext, i, rx, ry, rz);
            const int ra = gld(c.rtype, i);
            for (int j = 0; j < nl; j++) {
                double val;
                if (pair_term64(rx, ry, rz, lx[j], ly[j], lz[j], ra, lt[j], table, val)) {
                    v += val;
                    npairs++;
                }
            }
        }
        double *vb = vbuf + (ch & 1) * kScore64Threads;
        vb[tid] = v;
        __syncthreads();
        if (tid == 0) {
            const int rows = min(kScore64Threads, nr - ch * kScore64Threads);
            for (int t = 0; t < rows; t++) energy += vb[t];
        }
    }

    // pair count: wave reduction, then across the 4 waves through LDS
    for (int off = 32; off > 0; off >>= 1) npairs += __shfl_down(npairs, off, 64);
    __syncthreads();
    long long *wsum = (long long *)vbuf;
    if ((tid & 63) == 0) wsum[tid >> 6] = npairs;
    __syncthreads();
    if (tid == 0) {
        long long tot = 0;
        for (int w = 0; w < kScore64Threads / 64; w++) tot += wsum[w];
        const double s = score_of_energy(energy);
        a.score[gid] = s;
        if (a.pairs) a.pairs[gid] = tot;
        if (a.update_luciferin) {
            a.luciferin[gid] = (1.0 - a.rho) * a.luciferin[gid] + a.gamma * s;
            if (a.step) a.step[gid] += 1;
        }
    }
}

hipError_t launch_score64(const ScoreArgs &a, int max_nl, hipStream_t s) {
    if (a.n_glow <= 0) return hipSuccess;
    const size_t lds = sizeof(double) * (3 * (size_t)max_nl + 2 * kScore64Threads) +
                       sizeof(int) * (size_t)max_nl;
    if (lds > 64 * 1024) {
        const hipError_t e = hipFuncSetAttribute((const void *)score64_kernel,
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(score64_kernel, dim3(a.n_glow), dim3(kScore64Threads), lds, s, a);
    return hipGetLastError();
}

}  // namespace sd
