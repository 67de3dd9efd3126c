// Plain FP64 scoring: one workgroup per glowworm, brute force over all
// receptor x ligand pairs (the per-complex baseline form; its reference design
// is the calculate_dfire kernel of source/dock/src/cuda/gpu_baseline:492).
//
// Layout per workgroup:
//   LDS  ligand pose in FP64 in groups of four atoms (x[4], y[4], z[4],
//        type*20[4]: 112 bytes), nl padded to nlp = a multiple of 16 with
//        far-away atoms, + two 256-entry row-sum buffers
//   lane two receptor rows at a time (base + tid and base + 256 + tid) while
//        512-row chunks are full, then one (base + tid), their FP64
//        coordinates displaced in registers; ligand atoms are LDS broadcasts,
//        two atoms' coordinates per 16-byte read, a group's seven reads from
//        one address, each atom read once for the lane's rows
//
// A pair's table value is a buffer load whose offset lies past the table
// (the load returns +0.0) unless the pair is inside (0, 225]; the loads go
// into a ring of eight and each is added to its row sum a group of eight
// pairs later, so a gather's latency is not waited for at its pair.  The additions stay in
// pair order, and adding +0.0 leaves an FP64 sum unchanged (a row sum that
// starts at +0.0 is never -0.0).
// The energy is summed exactly as prepare_gso_tasks.cc:1539-1570 does: each
// row sum v_i runs over j in order (one lane), and thread 0 adds the v_i in
// row order (double-buffered, one barrier per chunk), so the FP64 energy is
// bit-identical to the CPU reference.  The epilogue applies the score affine
// map and the luciferin update (prepare_gso_tasks.cc:1571, 558).
#include "dev_common.h"

namespace sd {

// sqrt(x) for 0 < x <= 225, equal to the correctly rounded sqrt for
// x >= 2^-100: v_rsq_f64, then the Goldschmidt iteration for the root and
// half its reciprocal with two Newton corrections of the root (P. Markstein,
// "IA-64 and Elementary Functions", 2000), the steps the AMDGPU backend takes
// for an FP64 sqrt less its range scaling (for x < 2^-767) and its zero and
// infinity cases.  Smaller x are raised to 2^-100: every x < 4 has bin 0.
__device__ __forceinline__ double sqrt_in_cutoff(double x) {
    // one v_max_f64 (x is a sum of products, never a signalling NaN)
    asm("v_max_f64 %0, %1, %2" : "=v"(x) : "v"(x), "s"(0x1p-100));
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y, h = y * 0.5;
    const double r = __builtin_fma(-h, g, 0.5);
    g = __builtin_fma(g, r, g);
    h = __builtin_fma(h, r, h);
    double e = __builtin_fma(-g, g, x);
    g = __builtin_fma(e, h, g);
    e = __builtin_fma(-g, g, x);
    return __builtin_fma(e, h, g);
}

// five waves per SIMD: 96 VGPRs (98 unbounded; two spilled outside the pair loop)
__global__ __launch_bounds__(kScore64Threads) __attribute__((amdgpu_waves_per_eu(5, 8))) void score64_kernel(ScoreArgs a) {
    extern __shared__ __align__(16) double lds[];
    const int gid = blockIdx.x;
    if (gid >= a.n_glow) return;
    const int tid = threadIdx.x;
    const int swarm = gid / a.G;
    const DevComplex &c = a.cx[a.swarm_cx[swarm]];
    const int nl = c.nl, nr = c.nr;
    const int nlp = score64_padded_nl(nl);
    // ligand atoms in groups of four, 112 bytes each: x[4], y[4], z[4] in FP64
    // and the four types, so one address and immediate offsets read a group
    double *la = lds;                               // [nlp / 4][14]
    double *vbuf = la + 14 * (nlp / 4);             // [2][256]

    const double *pose = a.pos + (size_t)gid * a.P;
    double inv[4];
    quat_inverse(pose + 3, inv);
    for (int j = tid; j < nlp; j += kScore64Threads) {
        double x = 1e30, y = 1e30, z = 1e30;  // padding: no pair within the cutoff
        int t = 0;
        if (j < nl) {
            ligand_atom64(c, pose, inv, j, x, y, z);
            t = gld(c.ltype, j);
        }
        double *q = la + 14 * (j >> 2) + (j & 3);
        q[0] = x;
        q[4] = y;
        q[8] = z;
        ((int *)(la + 14 * (j >> 2) + 12))[j & 3] = t;
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
    const __amdgpu_buffer_rsrc_t tbl = __builtin_amdgcn_make_buffer_rsrc(
        (void *)c.table64, (short)0, (SDGSO_TABLE_LEN + 20) * 8, 0x00020000);
    constexpr unsigned kNoLoad64 = 0x80000000u;  // past the table: the load returns +0.0
    double energy = 0.0;  // thread 0 only
    // a lane's pairs: at most ceil(nr / 256) * nlp, far below 2^32
    unsigned lane_pairs = 0;
    int nchunk = 0;  // 256-row halves so far: the row-sum buffer alternates
    // One chunk of R * 256 receptor rows from `base`: lane tid takes rows
    // base + r * 256 + tid (r < R), so that one broadcast read of a ligand
    // atom serves R rows.  A group is eight pairs: 8 / R atoms x R rows.
    auto chunk = [&](auto rows_per_lane, int base) {
        constexpr int R = decltype(rows_per_lane)::value;
        // rows past nr sit far away and find no pair, so every lane runs the loop
        double rx[R], ry[R], rz[R], v[R];
        int ra[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int i = base + r * kScore64Threads + tid;
            rx[r] = -1e30;
            ry[r] = -1e30;
            rz[r] = -1e30;
            ra[r] = 0;
            v[r] = 0.0;
            if (i < nr) {
                receptor_atom64(c, ext, i, rx[r], ry[r], rz[r]);
                ra[r] = gld(c.rtype, i);
            }
        }
        double ring[8];
#pragma unroll
        for (int k = 0; k < 8; k++) ring[k] = 0.0;
        // The table offsets of two groups of eight pairs, the group in
        // flight and the next one.  A register holds kNoLoad64 until a pair
        // within the cutoff writes it and goes back to kNoLoad64 once its
        // load has been added, so a pair outside the cutoff costs nothing.
        unsigned offA[8], offB[8];
#pragma unroll
        for (int k = 0; k < 8; k++) offA[k] = offB[k] = kNoLoad64;
        // the offsets of a group, slot 4 * (h + r) + k for atom 4 h + k of
        // the group and row r (no memory access under a branch); returns
        // whether any lane has a pair of them within the cutoff
        auto offsets = [&](const double *g, unsigned(&off)[8]) {
            unsigned long long near_lanes = 0;  // lanes with a pair within the cutoff
#pragma unroll
            for (int h = 0; h < 2 / R; h++) {
                const double *q = g + 14 * h;
                const double2 x01 = *(const double2 *)(q + 0), x23 = *(const double2 *)(q + 2);
                const double2 y01 = *(const double2 *)(q + 4), y23 = *(const double2 *)(q + 6);
                const double2 z01 = *(const double2 *)(q + 8), z23 = *(const double2 *)(q + 10);
                const int4 t4 = *(const int4 *)(q + 12);
                const double xs[4] = {x01.x, x01.y, x23.x, x23.y};
                const double ys[4] = {y01.x, y01.y, y23.x, y23.y};
                const double zs[4] = {z01.x, z01.y, z23.x, z23.y};
                const int ts[4] = {t4.x, t4.y, t4.z, t4.w};
#pragma unroll
                for (int r = 0; r < R; r++) {
#pragma unroll
                    for (int k = 0; k < 4; k++) {
                        // pair_bin64 with the test d2 > 0 (coincident atoms,
                        // rare) under the cutoff branch: one compare less per pair
                        const double dx = rx[r] - xs[k];
                        const double dy = ry[r] - ys[k];
                        const double dz = rz[r] - zs[k];
                        const double d2 = dx * dx + dy * dy + dz * dz;
                        const bool near = d2 <= kCutoff2;
                        near_lanes |= __builtin_amdgcn_ballot_w64(near);
                        if (near) {
                            const int d = (int)(sqrt_in_cutoff(d2) * 2.0 - 1.0);
                            const bool in = d2 > 0.0;
                            off[4 * (h + r) + k] =
                                in ? 8u * (unsigned)(ra[r] + ts[k] + bin_of_d(d)) : kNoLoad64;
                            lane_pairs += in ? 1u : 0u;
                        }
                    }
                }
            }
            return near_lanes != 0;
        };
        // In pair order within each row, the values of the previous group
        // into the row sums (their offsets back to kNoLoad64), and this
        // group's loads into the freed slots.  The empty asm statements order
        // them: the adds after this group's arithmetic (a slot's load then
        // has a whole group behind it), the loads after the adds (a load
        // reuses its slot's registers).  A group with no pair within the
        // cutoff in any lane issues no loads (they would all return +0.0),
        // and its adds are skipped a group later.
        bool pending = false;  // the ring holds the loads of the previous group
        auto issue = [&](unsigned(&off)[8], unsigned(&prev)[8], bool hit) {
#pragma unroll
            for (int k = 0; k < 8; k++) asm volatile("" : "+v"(ring[k]) : "v"(off[7]));
            if (pending) {
#pragma unroll
                for (int k = 0; k < 8; k++) v[k / (8 / R)] += ring[k];
#pragma unroll
                for (int k = 0; k < 8; k++) prev[k] = kNoLoad64;
            }
            if (hit) {
#pragma unroll
                for (int k = 0; k < 8; k++) {
                    asm volatile("" : "+v"(off[k]) : "v"(v[R - 1]));
                    ring[k] = __builtin_bit_cast(
                        double, __builtin_amdgcn_raw_buffer_load_b64(tbl, off[k], 0, 0));
                }
            }
            pending = hit;
        };
        // nothing in flight on entry (the rows' loads done), so that inside
        // the loop only the ring's loads are waited for
        __builtin_amdgcn_s_waitcnt(0);
        // the byte offset of the sixteen atoms in a VGPR: their 28 reads
        // take it with immediate offsets, one add per sixteen atoms
        unsigned gb = 0;
        for (int j0 = 0; j0 < nlp; j0 += 16, gb += 16 * 28) {
            asm("" : "+v"(gb));
            const double *g16 = (const double *)((const char *)la + gb);
#pragma unroll
            for (int gi = 0; gi < 2 * R; gi += 2) {
                const bool hitA = offsets(g16 + 14 * (2 / R) * gi, offA);
                issue(offA, offB, hitA);
                const bool hitB = offsets(g16 + 14 * (2 / R) * (gi + 1), offB);
                issue(offB, offA, hitB);
            }
        }
        if (pending) {
#pragma unroll
            for (int k = 0; k < 8; k++) v[k / (8 / R)] += ring[k];
        }
        // each 256-row half through the double buffer in turn: a half is
        // written only after the barrier that follows thread 0's sum of the
        // half before the previous one
#pragma unroll
        for (int r = 0; r < R; r++) {
            double *vb = vbuf + (nchunk & 1) * kScore64Threads;
            vb[tid] = v[r];
            __syncthreads();
            if (tid == 0) {
                const int rows = min(kScore64Threads, nr - base - r * kScore64Threads);
                for (int t = 0; t < rows; t++) energy += vb[t];
            }
            nchunk++;
        }
    };
    // chunks of 512 rows while they are full, then of 256
    int base = 0;
    for (; nr - base >= 2 * kScore64Threads; base += 2 * kScore64Threads)
        chunk(std::integral_constant<int, 2>{}, base);
    for (; base < nr; base += kScore64Threads) chunk(std::integral_constant<int, 1>{}, base);

    // pair count: wave reduction, then across the 4 waves through LDS
    long long npairs = lane_pairs;
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
    const size_t lds = score64_lds_bytes(max_nl);
    if (lds > 64 * 1024) {
        const hipError_t e = hipFuncSetAttribute((const void *)score64_kernel,
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(score64_kernel, dim3(a.n_glow), dim3(kScore64Threads), lds, s, a);
    return hipGetLastError();
}

}  // namespace sd
