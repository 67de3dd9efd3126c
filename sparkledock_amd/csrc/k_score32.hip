// FP32 guard-band scoring (the batched form).
//
// Work unit: one wavefront per (glowworm, ligand block of 64 atoms), taken
// from a work list that keeps the blocks of one glowworm adjacent.  Lane l
// owns ligand atom lperm[block*64 + l], transformed in FP64 in the reference
// order (prepare_gso_tasks.cc:487-526) and kept in registers.  The receptor is
// cut into compact tiles of 64 atoms (recursive bisection) with bounding
// spheres; a tile is visited only when its sphere, padded by the cutoff, the
// block sphere and the normal-mode displacement bounds of this pose, reaches
// the block.  In a visited tile each lane displaces one receptor atom (FP32
// modes) and the wave keeps only atoms that can reach the block (ballot).
//
// Phase A (filter): the kept receptor atoms are broadcast one at a time with
// readlane; every lane forms d2 with its ligand atom and the pairs inside the
// cutoff (plus the guard band) are appended to a per-wave queue in LDS
// (ballot + mbcnt compaction: d2, table row, pair tag).
// Phase B (evaluate): when the queue fills, every lane takes queue entries
// (all 64 lanes busy), finds the bin, issues the table gathers of several
// entries back to back and only then adds them (compensated FP32 sums), so
// the L2 latency of one gather overlaps the others.
//
// Distances: both atoms are placed on a 2^-20 A fixed-point grid relative to
// the tile centre (FP64 coordinates rounded once), so coordinate differences
// are exact integers and convert to FP32 exactly below 16 A.  d2 in FP32 then
// differs from the FP64 reference d2 by at most 9.5e-5 A^2, under the 1e-4
// A^2 default band.  A pair whose FP32 d2 lies within the band of a boundary
// k^2/4 (k = 0..30: d2 = 0, every bin edge and the 225 A^2 cutoff) is
// recomputed in FP64 from the reference-order coordinates, so every bin and
// the in/out decision equal the reference's.  A pair outside every band is
// inside iff d2 < 225, and its bin is floor(2 sqrt(d2)) mapped through
// dist_to_bins (raw v_sqrt_f32, 1 ulp, cannot cross an integer there).
//
// Energies: per-lane compensated (Kahan) FP32 sums; lanes are combined in
// FP64 and written as one partial per (glowworm, block); finish32 adds a
// glowworm's partials in block order in FP64.
#include "dev_common.h"

namespace sd {

constexpr int kScore32Threads = 256;
constexpr int kWaves = kScore32Threads / 64;
constexpr int kQueue = 256;                   // queued pairs per wave (phase A -> B)
constexpr int kUnroll = kQueue / 64;          // gathers in flight per lane in phase B
constexpr float kQ = 1048576.0f;              // 2^20 units per A
constexpr float kInvQ2x2 = 1.0f / 524288.0f;  // sqrt(d2q) * 2^-19 = 2*sqrt(d2)
constexpr float kQ2over4 = 274877906944.0f;   // 2^38 = 2^40 / 4
constexpr float kQ2 = 1099511627776.0f;       // 2^40
constexpr float k225q = 225.0f * kQ2;         // the cutoff in 2^-40 A^2 units (exact)
constexpr int kFar = 1 << 29;
constexpr int kZeroIdx = SDGSO_TABLE_LEN + SDGSO_NBINS - 1;  // a zero of the table pad

struct Kahan {
    float sum = 0.f, comp = 0.f;
    __device__ __forceinline__ void add(float v, bool take) {
        const float y = v - comp;
        const float t = sum + y;
        const float c2 = (t - sum) - y;
        sum = take ? t : sum;
        comp = take ? c2 : comp;
    }
};

// FP64 decision for a queued pair inside the guard band of a boundary.
// The pair is rebuilt from scratch (pose transform of both atoms), so no
// per-lane state crosses the call.
// Returns (table value, 1) for a pair inside (0, 225], else (0, 0).
__device__ __noinline__ float2 guard_pair64(const DevComplex &c, const double *pose,
                                            int lig_atom, int rec_atom) {
    double inv[4];
    quat_inverse(pose + 3, inv);
    double Rx, Ry, Rz, Lx, Ly, Lz, v64 = 0.0;
    receptor_atom64(c, pose + 7, rec_atom, Rx, Ry, Rz);
    ligand_atom64(c, pose, inv, lig_atom, Lx, Ly, Lz);
    const bool hit = pair_term64(Rx, Ry, Rz, Lx, Ly, Lz, gld(c.rtype, rec_atom),
                                 gld(c.ltype, lig_atom), c.table, v64);
    return make_float2(hit ? (float)v64 : 0.f, hit ? 1.f : 0.f);
}

__global__ __launch_bounds__(kScore32Threads) void score32_kernel(ScoreArgs a) {
    __shared__ float q_d2[kWaves][kQueue];
    __shared__ int q_row[kWaves][kQueue];
    __shared__ int q_tag[kWaves][kQueue];
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int wv = blockIdx.x * kWaves + w;
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
    const int j = gld(c.lperm, b * kBlock + lane);
    double Lx = 0, Ly = 0, Lz = 0;
    int lb = 0;
    if (j >= 0) {
        ligand_atom64(c, pose, inv, j, Lx, Ly, Lz);
        lb = gld(c.ltype, j);
    }
    // block sphere in the docked frame and displacement bounds of this pose
    const float4 blk = gld(c.lblocks, b);
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
    const float cutq = (225.0f + 2.0f * band) * kQ2;
    const float bandq = band * kQ2;
    const int slots = c.ntiles * kTile;
    float *qd = q_d2[w];
    int *qr = q_row[w];
    int *qt = q_tag[w];

    Kahan acc;
    int npairs = 0;
    unsigned int nguard = 0;
    int qn = 0;  // wave-uniform queue length

    // Phase B: evaluate the queued pairs with all lanes.
    auto flush = [&]() {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        float v[kUnroll];
        bool take[kUnroll], near[kUnroll];
        int tag[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; u++) {
            const int e = u * 64 + lane;
            const bool valid = e < qn;
            const float d2q = valid ? qd[e] : 0.f;
            const int row = valid ? qr[e] : 0;
            tag[u] = valid ? qt[e] : 0;
            const float s = __builtin_amdgcn_sqrtf(d2q) * kInvQ2x2;
            const float kk = __builtin_rintf(s);
            const float diff = __builtin_fmaf(kk * kk, -kQ2over4, d2q);
            near[u] = valid && fabsf(diff) < bandq;
            take[u] = valid && !near[u] && d2q <= k225q;
            const int idx = take[u] ? row + bin_of_d((int)s - 1) : kZeroIdx;
            v[u] = gld(c.table, idx);
        }
#pragma unroll
        for (int u = 0; u < kUnroll; u++) {
            acc.add(v[u], take[u]);
            npairs += take[u] ? 1 : 0;
            if (near[u]) {
                const int t = tag[u];
                const float2 r = guard_pair64(c, pose, gld(c.lperm, b * kBlock + (t >> 24)),
                                              gld(c.rperm, t & 0xffffff));
                const bool hit = r.y != 0.f;
                acc.add(r.x, hit);
                npairs += hit ? 1 : 0;
                nguard++;
            }
        }
        __builtin_amdgcn_wave_barrier();
    };

    for (int t = 0; t < c.ntiles; t++) {
        const float4 tl = gld(c.tiles, t);
        const float tx = tl.x - fcx, ty = tl.y - fcy, tz = tl.z - fcz;
        const float lim = reach_atom + tl.w;
        if (tx * tx + ty * ty + tz * tz > lim * lim) continue;
        const int slot = t * kTile + lane;
        const double rx = gld(c.rrel, slot), ry = gld(c.rrel, slots + slot),
                     rz = gld(c.rrel, 2 * slots + slot);
        const float ax = (float)rx + tx, ay = (float)ry + ty, az = (float)rz + tz;
        const bool nearb = ax * ax + ay * ay + az * az <= reach_atom * reach_atom;
        unsigned long long mask = __ballot(nearb);
        if (!mask) continue;
        int Rqx = kFar, Rqy = kFar, Rqz = kFar, ra = 0;
        if (nearb) {
            float Dx = 0.f, Dy = 0.f, Dz = 0.f;
            for (int m = 0; m < c.nmr; m++) {
                const float e = (float)rext[m];
                const float *M = c.rmod32 + (size_t)m * 3 * slots;
                Dx = __builtin_fmaf(gld(M, slot), e, Dx);
                Dy = __builtin_fmaf(gld(M, slots + slot), e, Dy);
                Dz = __builtin_fmaf(gld(M, 2 * slots + slot), e, Dz);
            }
            Rqx = (int)rint((rx + (double)Dx) * (double)kQ);
            Rqy = (int)rint((ry + (double)Dy) * (double)kQ);
            Rqz = (int)rint((rz + (double)Dz) * (double)kQ);
            ra = gld(c.rtype_s, slot);
        }
        int Lqx = kFar, Lqy = kFar, Lqz = kFar;
        if (j >= 0) {
            Lqx = (int)rint((Lx - (double)tl.x) * (double)kQ);
            Lqy = (int)rint((Ly - (double)tl.y) * (double)kQ);
            Lqz = (int)rint((Lz - (double)tl.z) * (double)kQ);
        }
        const int row_l = lb;
        const int tag_l = (lane << 24) | (t * kTile);
        // Phase A: filter the kept receptor atoms against all 64 ligand lanes
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
            const bool in = d2q <= cutq;
            const unsigned long long m = __ballot(in);
            if (m) {
                if (in) {
                    const int pos = qn + __builtin_amdgcn_mbcnt_hi(
                                             (unsigned)(m >> 32),
                                             __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
                    qd[pos] = d2q;
                    qr[pos] = sa + row_l;
                    qt[pos] = tag_l + i;
                }
                qn += __popcll(m);
                if (qn > kQueue - 64) {
                    flush();
                    qn = 0;
                }
            }
        }
    }
    if (qn) flush();

    double e = (double)acc.sum - (double)acc.comp;
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
    const int grid = (a.n_work + kWaves - 1) / kWaves;
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
