// FP32 guard-band scoring (the batched form).
//
// Work unit: one wavefront per (glowworm, ligand block of 64 atoms), taken
// from a work list that keeps the blocks of one glowworm adjacent.  Lane l
// owns ligand atom lperm[block*64 + l], transformed in FP64 in the reference
// order (prepare_gso_tasks.cc:487-526) and kept in registers.  The receptor is
// cut into compact tiles of 64 atoms (recursive bisection) with bounding spheres; a tile is
// visited only when its sphere, padded by the cutoff, the block sphere and the
// normal-mode displacement bounds of this pose, reaches the block.  Inside a
// visited tile each lane displaces one receptor atom (FP32 modes) and the
// wave keeps only atoms that can reach the block (ballot).  Those atoms are
// broadcast one by one with readlane and every lane forms the pair with its
// ligand atom.
//
// Distances: both atoms are placed on a 2^-20 A fixed-point grid relative to
// the tile centre (FP64 coordinates rounded once), so coordinate differences
// are exact integers and convert to FP32 exactly below 16 A.  d2 in FP32 then
// differs from the FP64 reference d2 by at most 9.5e-5 A^2 (see the bound in
// the design notes), under the 1e-4 A^2 default band.  A pair whose FP32 d2 lies
// within the band of a bin boundary k^2/4 (k = 0..30, which includes d2 = 0
// and the 225 A^2 cutoff) is recomputed in FP64 from the reference-order
// coordinates, so every bin equals the reference bin.
//
// Energies: table values are added per lane with a compensated (Kahan) FP32
// sum; lanes are combined in FP64 and written as one partial per (glowworm,
// block); finish32 adds a glowworm's partials in block order in FP64.
#include "dev_common.h"

namespace sd {

constexpr int kScore32Threads = 256;
constexpr float kQ = 1048576.0f;              // 2^20 units per A
constexpr float kInvQ2x2 = 1.0f / 524288.0f;  // sqrt(d2q) * 2^-19 = 2*sqrt(d2)
constexpr float kQ2over4 = 274877906944.0f;   // 2^38 = 2^40 / 4
constexpr float kQ2 = 1099511627776.0f;       // 2^40
constexpr int kFar = 1 << 29;

__global__ __launch_bounds__(kScore32Threads) void score32_kernel(ScoreArgs a) {
    const int lane = threadIdx.x & 63;
    const int wv = blockIdx.x * (kScore32Threads / 64) + (threadIdx.x >> 6);
    if (wv >= a.n_work) return;
    const int gid = a.work[2 * wv];
    const int b = a.work[2 * wv + 1];
    const DevComplex &c = a.cx[a.swarm_cx[gid / a.G]];
    const double *pose = a.pos + (size_t)gid * a.P;
    const double *rext = pose + 7;
    const double *lext = pose + 7 + c.nmr;
    double inv[4];
    quat_inverse(pose + 3, inv);

    // this lane's ligand atom, FP64 in reference order
    const int j = c.lperm[b * kBlock + lane];
    double Lx = 0, Ly = 0, Lz = 0;
    int lb = 0;
    if (j >= 0) {
        ligand_atom64(c, pose, inv, j, Lx, Ly, Lz);
        lb = c.ltype[j];
    }
    // block sphere in the docked frame and displacement bounds of this pose
    const float4 blk = c.lblocks[b];
    double cbx = blk.x, cby = blk.y, cbz = blk.z;
    rotate_point(pose + 3, inv, cbx, cby, cbz);
    cbx += pose[0];
    cby += pose[1];
    cbz += pose[2];
    float dr = 0.f, dl = 0.f;
    for (int m = 0; m < c.nmr; m++) dr += fabsf((float)rext[m]) * c.rmode_max[m];
    for (int m = 0; m < c.nml; m++) dl += fabsf((float)lext[m]) * c.lmode_max[m];
    dr = dr * 1.0001f + 1e-3f;
    dl = dl * 1.0001f + 1e-3f;
    const float reach_atom = 15.0f + blk.w + dl + dr + 0.01f;
    const float fcx = (float)cbx, fcy = (float)cby, fcz = (float)cbz;

    const float band = a.band;
    const float cutq = (225.0f + band) * kQ2;
    const float bandq = band * kQ2;
    const int slots = c.ntiles * kTile;

    float ksum = 0.f, kcomp = 0.f;
    int npairs = 0;
    unsigned int nguard = 0;
    for (int t = 0; t < c.ntiles; t++) {
        const float4 tl = c.tiles[t];
        const float tx = tl.x - fcx, ty = tl.y - fcy, tz = tl.z - fcz;
        const float lim = reach_atom + tl.w;
        if (tx * tx + ty * ty + tz * tz > lim * lim) continue;
        const int slot = t * kTile + lane;
        const double rx = c.rrel[slot], ry = c.rrel[slots + slot], rz = c.rrel[2 * slots + slot];
        const float ax = (float)rx + tx, ay = (float)ry + ty, az = (float)rz + tz;
        const bool near = ax * ax + ay * ay + az * az <= reach_atom * reach_atom;
        unsigned long long mask = __ballot(near);
        if (!mask) continue;
        int Rqx = kFar, Rqy = kFar, Rqz = kFar, ra = 0;
        if (near) {
            float Dx = 0.f, Dy = 0.f, Dz = 0.f;
            for (int m = 0; m < c.nmr; m++) {
                const float e = (float)rext[m];
                const float *M = c.rmod32 + (size_t)m * 3 * slots;
                Dx = __builtin_fmaf(M[slot], e, Dx);
                Dy = __builtin_fmaf(M[slots + slot], e, Dy);
                Dz = __builtin_fmaf(M[2 * slots + slot], e, Dz);
            }
            Rqx = (int)rint((rx + (double)Dx) * (double)kQ);
            Rqy = (int)rint((ry + (double)Dy) * (double)kQ);
            Rqz = (int)rint((rz + (double)Dz) * (double)kQ);
            ra = c.rtype_s[slot];
        }
        int Lqx = kFar, Lqy = kFar, Lqz = kFar;
        if (j >= 0) {
            Lqx = (int)rint((Lx - (double)tl.x) * (double)kQ);
            Lqy = (int)rint((Ly - (double)tl.y) * (double)kQ);
            Lqz = (int)rint((Lz - (double)tl.z) * (double)kQ);
        }
        while (mask) {
            const int i = __ffsll((long long)mask) - 1;
            mask &= mask - 1;
            const int sx = __builtin_amdgcn_readlane(Rqx, i);
            const int sy = __builtin_amdgcn_readlane(Rqy, i);
            const int sz = __builtin_amdgcn_readlane(Rqz, i);
            const int sa = __builtin_amdgcn_readlane(ra, i);
            const float fx = (float)(Lqx - sx);
            const float fy = (float)(Lqy - sy);
            const float fz = (float)(Lqz - sz);
            const float d2q = __builtin_fmaf(fz, fz, __builtin_fmaf(fy, fy, fx * fx));
            if (d2q <= cutq) {
                const float s = __builtin_sqrtf(d2q) * kInvQ2x2;
                const float kk = __builtin_rintf(s);
                const float bq = kk * kk * kQ2over4;
                float val = 0.f;
                bool hit;
                if (fabsf(d2q - bq) < bandq) {
                    // within the guard band of a boundary: decide in FP64
                    double Rx, Ry, Rz, v64 = 0.0;
                    receptor_atom64(c, rext, c.rperm[t * kTile + i], Rx, Ry, Rz);
                    hit = pair_term64(Rx, Ry, Rz, Lx, Ly, Lz, sa, lb, c.table, v64);
                    val = (float)v64;
                    nguard++;
                } else {
                    const int d = (int)s - 1;
                    val = c.table[sa + lb + bin_of_d(d)];
                    hit = true;
                }
                if (hit) {
                    const float y = val - kcomp;
                    const float tsum = ksum + y;
                    kcomp = (tsum - ksum) - y;
                    ksum = tsum;
                    npairs++;
                }
            }
        }
    }
    double e = (double)ksum - (double)kcomp;
    long long np = npairs;
    for (int off = 32; off > 0; off >>= 1) {
        e += __shfl_xor(e, off, 64);
        np += __shfl_xor(np, off, 64);
    }
    unsigned int ng = nguard;
    for (int off = 32; off > 0; off >>= 1) ng += __shfl_xor(ng, off, 64);
    if (lane == 0) {
        a.partial[(size_t)gid * a.max_lblocks + b] = e;
        a.partial_pairs[(size_t)gid * a.max_lblocks + b] = np;
        if (a.guard_count && ng) atomicAdd(a.guard_count, (unsigned long long)ng);
    }
}

// Per glowworm: energy = partials in block order (FP64), score, luciferin,
// reference point (prepare_gso_tasks.cc:513-526, 558, 1571).
__global__ __launch_bounds__(256) void finish32_kernel(ScoreArgs a) {
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= a.n_glow) return;
    const DevComplex &c = a.cx[a.swarm_cx[gid / a.G]];
    const double *part = a.partial + (size_t)gid * a.max_lblocks;
    const long long *pp = a.partial_pairs + (size_t)gid * a.max_lblocks;
    double e = 0.0;
    long long np = 0;
    for (int b = 0; b < c.nlblocks; b++) {
        e += part[b];
        np += pp[b];
    }
    const double s = score_of_energy(e);
    a.score[gid] = s;
    if (a.pairs) a.pairs[gid] = np;
    const double *pose = a.pos + (size_t)gid * a.P;
    double inv[4], r[3];
    quat_inverse(pose + 3, inv);
    ref_point64(c, pose, inv, r);
    a.ref[(size_t)gid * 3 + 0] = r[0];
    a.ref[(size_t)gid * 3 + 1] = r[1];
    a.ref[(size_t)gid * 3 + 2] = r[2];
    if (a.update_luciferin) {
        a.luciferin[gid] = (1.0 - a.rho) * a.luciferin[gid] + a.gamma * s;
        if (a.step) a.step[gid] += 1;
    }
}

__global__ void advance_kernel(int *k) {
    if (threadIdx.x == 0) *k += 1;
}

hipError_t launch_score32(const ScoreArgs &a, hipStream_t s) {
    if (a.n_work <= 0) return hipSuccess;
    const int waves_per_wg = kScore32Threads / 64;
    const int grid = (a.n_work + waves_per_wg - 1) / waves_per_wg;
    hipLaunchKernelGGL(score32_kernel, dim3(grid), dim3(kScore32Threads), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_finish32(const ScoreArgs &a, hipStream_t s) {
    if (a.n_glow <= 0) return hipSuccess;
    hipLaunchKernelGGL(finish32_kernel, dim3((a.n_glow + 255) / 256), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_advance(int *d_k, hipStream_t s) {
    hipLaunchKernelGGL(advance_kernel, dim3(1), dim3(64), 0, s, d_k);
    return hipGetLastError();
}

}  // namespace sd
