// FP32 guard-band scoring (the batched form).
//
// Work unit: one wavefront per (glowworm, ligand block of 64 atoms), or per
// (glowworm, pair of neighbouring blocks) in the two-block variants, taken
// from a work list that keeps the blocks of one glowworm adjacent, in list
// order or in chunks per XCD (ScoreArgs::xcd_chunk).  Lane l owns ligand
// atom lperm[block*64 + l]: its FP64 coordinates plus its FP32 mode
// displacement (summed in the set-up, or read from the run's ligand
// displacement GEMM, ScoreArgs::ldisp), rotated by the pose's FP64 matrix
// and translated (prepare_gso_tasks.cc:487-526 up to rounding), then rounded
// once to a 2^-20 A fixed-point grid whose origin is the (first) block's
// centre.
//
// Receptor tiles: compact groups of 64 atoms (recursive bisection) with
// bounding spheres.  The wave tests 64 tile spheres per instruction against
// the block sphere, padded by the cutoff and this pose's normal-mode
// displacement bounds, lists the candidates in LDS and visits them; the
// receptor atoms that can reach the block are displaced (FP32 modes), rounded
// to the same grid and pre-tested against the block's ligand atoms.
//
// Kernel: score32b; near atoms are batched, displaced by cooperative lanes,
// box-tested, packed into LDS and pre-tested with broadcast reads (FP32 or
// half-precision MFMA); passing pairs are queued as one word and evaluated
// 64 per round (described at the kernel).  The run's autotuner picks a
// variant per complex: one or two ligand blocks per wave ("bcast",
// "bcast2"), the pre-test form, and how waves get their work items (one item
// per wave from the hardware dispatcher, or a grid of resident waves taking
// items from a queue, "-queue").
//
// Evaluate (all 64 lanes busy): exact d2 from the fixed-point coordinates
// (integer differences are exact and convert to FP32 exactly below 16 A), the
// bin, and the table value; several gathers are in flight per lane before
// they are added into the lane's FP64 sum.  The FP32 d2 differs from the FP64 reference
// d2 by at most 9.04e-5 + 3.1e-6 (k_r S_r + k_l S_l) A^2 (S the pose's
// receptor and ligand displacement bounds, k the FP32 roundings per mode
// term); each wave widens the run's band (default 1e-4 A^2) to that bound.  A
// pair within the band of a boundary k^2/4 (k = 0..30: d2 = 0, every bin edge
// and the 225 A^2 cutoff) is recomputed in FP64 from the reference-order
// coordinates, so every bin and in/out decision equals the reference's.
// Outside every band a pair is inside iff d2 < 225 and its bin is
// floor(2 sqrt(d2)) through dist_to_bins (raw v_sqrt_f32, 1 ulp, cannot cross
// an integer there).
//
// Lane sums are combined in FP64 and written as one partial per
// (glowworm, block); finish32 adds a glowworm's partials in block order in
// FP64, so the energy keeps its relative accuracy when the terms cancel.
#include <string>
#include <type_traits>

#include "dev_common.h"

namespace sd {

typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef _Float16 half4 __attribute__((ext_vector_type(4)));

constexpr int kMaxTiles = kMaxTiles32;        // receptor tiles per complex (host-checked)
constexpr float kQ = 1048576.0f;              // 2^20 units per A
constexpr float kInvQ = 1.0f / 1048576.0f;
constexpr float kInvQ2x2 = 1.0f / 524288.0f;  // sqrt(d2q) * 2^-19 = 2*sqrt(d2)
constexpr float kQ2over4 = 274877906944.0f;   // 2^38 = 2^40 / 4
constexpr float kQ2 = 1099511627776.0f;       // 2^40
constexpr float k225q = 225.0f * kQ2;         // the cutoff in 2^-40 A^2 units (exact)
constexpr int kFar = 1 << 29;                 // far-away fixed-point coordinate (512 A)
// The table gathers read tablep ([a][bin][b]): row a*kPRow from the receptor
// atom, b from the ligand atom, then bin*168.

// Each lane adds its table values (FP32, so exact in FP64) into one FP64
// sum: a convert and an FP64 add per term, which issue at the FP32 rate on
// this GPU.  A lane's n terms err by at most n 2^-53 sum |T|, so a sum whose
// terms cancel keeps its relative accuracy.

// x/y/z += M[i][m][.] * e_m for m in mode order, one mode at a time (the
// reference order of prepare_gso_tasks.cc:465-504, as modes_add64), from the
// atom-major copy of the modes: atom i's values are consecutive, so after the
// first miss the loads of the later modes hit the cache.
__device__ __forceinline__ void modes_add64_am(const double *modes_am, int i, const double *ext,
                                               int nm, double &x, double &y, double &z) {
    const double *M = modes_am + (size_t)i * nm * 3;
#pragma unroll 1
    for (int m = 0; m < nm; m++) {
        const double e = ext[m];
        x += gld(M, 3 * m) * e;
        y += gld(M, 3 * m + 1) * e;
        z += gld(M, 3 * m + 2) * e;
    }
}

// FP64 decision for a pair inside the guard band of a boundary, rebuilt from
// the pose in the reference's order: returns (FP64 table value, 1) for a
// pair inside (0, 225], else (0, 0).  Called from the evaluation loop of the
// FP32 kernels by the lanes that hold such a pair, for about 4e-5 of the
// pairs, so it is kept out of line and lean: its mode sums run one mode at a
// time, which leaves most of the caller's registers untouched across the
// call.
__device__ __noinline__ double2 guard_pair64(const DevComplex &c, const double *pose,
                                             int lig_atom, int rec_atom) {
    if (lig_atom < 0 || rec_atom < 0) return make_double2(0.0, 0.0);
    double Rx = gld(c.r0, rec_atom), Ry = gld(c.r0, c.nr + rec_atom),
           Rz = gld(c.r0, 2 * c.nr + rec_atom);
    modes_add64_am(c.rmodes_am, rec_atom, pose + 7, c.nmr, Rx, Ry, Rz);
    double Lx = gld(c.l0, lig_atom), Ly = gld(c.l0, c.nl + lig_atom),
           Lz = gld(c.l0, 2 * c.nl + lig_atom);
    modes_add64_am(c.lmodes_am, lig_atom, pose + 7 + c.nmr, c.nml, Lx, Ly, Lz);
    double inv[4];
    quat_inverse(pose + 3, inv);
    rotate_point(pose + 3, inv, Lx, Ly, Lz);
    Lx += pose[0];
    Ly += pose[1];
    Lz += pose[2];
    double v64 = 0.0;
    const bool hit = pair_term64(Rx, Ry, Rz, Lx, Ly, Lz, gld(c.rtype, rec_atom),
                                 gld(c.ltype, lig_atom), c.table, v64);
    return make_double2(hit ? v64 : 0.0, hit ? 1.0 : 0.0);
}

__device__ __forceinline__ int qround(double v) { return (int)rint(v * (double)kQ); }

__device__ __forceinline__ void wave_sync_lds() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// bin_of_d in one v_med3_i32: max(0, min(d - 2, (d + 11) >> 1)) equals
// med3(d - 2, 0, (d + 11) >> 1) for d >= -1, since (d + 11) >> 1 > 0 there:
// d - 2 < 0 gives 0 both ways, and otherwise the median of 0 <= d - 2 and
// the positive (d + 11) >> 1 is their minimum.
__device__ __forceinline__ int bin_med3(int d) {
    int r;
    asm("v_med3_i32 %0, %1, 0, %2" : "=v"(r) : "v"(d - 2), "v"((d + 11) >> 1));
    return r;
}

// Classification of one exact fixed-point d2: bin of a pair outside every
// guard band (take), or a pair inside a band (near: decided in FP64).
struct PairClass {
    bool take, near;
    int bin;
};

__device__ __forceinline__ PairClass classify(float d2q, float bandq) {
    const float s = __builtin_amdgcn_sqrtf(d2q) * kInvQ2x2;
    const float kk = __builtin_rintf(s);
    const float diff = __builtin_fmaf(kk * kk, -kQ2over4, d2q);
    PairClass p;
    // boundaries are k = 0..30; a pair beyond 225 + band is simply out
    p.near = s < 30.5f && fabsf(diff) < bandq;
    p.take = !p.near && d2q <= k225q;
    p.bin = bin_med3((int)s - 1);
    return p;
}

__device__ __forceinline__ float d2_fixed(int4 a, int4 b) {
    const float fx = (float)(a.x - b.x);
    const float fy = (float)(a.y - b.y);
    const float fz = (float)(a.z - b.z);
    return __builtin_fmaf(fz, fz, __builtin_fmaf(fy, fy, fx * fx));
}

// Per-wave state shared by both filter variants.
struct WaveCtx {
    int lane, gid, b, nblk, ncand;
    int t_scan;       // next tile to scan for candidates
    const double *pose;
    double cb[3];     // block centre (fixed-point origin)
    float fc[3];
    float reach;      // cutoff + block radius + displacement bounds
    float dlr;        // bound of the mode displacements dl + dr (A)
    float dl;         // bound of the ligand's mode displacements (A)
    float band;       // this wave's guard band (A^2): the run's, widened to the error bound
    Rot64 rot;        // the pose's rotation matrix (set-up only)
    float rmax;       // bound of |r| for the receptor atoms that pass the atom test
    int4 L;           // this lane's ligand atom: fixed point + type*20
    int lig;          // its original index (-1 = padding)
};

// Candidate tiles from tile t0 on: 64 sphere tests per instruction, listed
// in LDS (at most `cap` entries); advances t0 past the tiles tested.  With
// kTwo a tile is also a candidate when it reaches the second block's sphere
// (centre fb, reach rb; rb < 0: no second block).
template <bool kTwo = false>
__device__ __forceinline__ int scan_candidates(int ntiles, const float4 *tiles, const WaveCtx &x,
                                               unsigned short *ct, int &t0, int cap,
                                               const float *fb = nullptr, float rb = -1.f) {
    int n = 0;
    for (; t0 < ntiles && n <= cap - 64; t0 += 64) {
        const int t = t0 + x.lane;
        bool ok = false;
        if (t < ntiles) {
            const float4 tl = gld(tiles, t);
            const float tx = tl.x - x.fc[0], ty = tl.y - x.fc[1], tz = tl.z - x.fc[2];
            const float lim = x.reach + tl.w;
            ok = tx * tx + ty * ty + tz * tz <= lim * lim;
            if constexpr (kTwo) {
                const float ux = tl.x - fb[0], uy = tl.y - fb[1], uz = tl.z - fb[2];
                const float limb = rb + tl.w;
                ok = ok || (rb >= 0.f && ux * ux + uy * uy + uz * uz <= limb * limb);
            }
        }
        const unsigned long long m = __ballot(ok);
        if (ok)
            ct[n + __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32),
                                             __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u))] =
                (unsigned short)t;
        n += __popcll(m);
    }
    wave_sync_lds();
    return n;
}

__device__ __forceinline__ int scan_candidates(const DevComplex &c, const WaveCtx &x,
                                               unsigned short *ct, int &t0, int cap) {
    return scan_candidates(c.ntiles, c.tiles, x, ct, t0, cap);
}

// Set-up common to the variants: pose, block sphere, ligand atom, receptor
// extents (LDS), candidate tiles (LDS; left to the caller when scan is
// false).  have_d: dA holds this lane's slot of block b's ligand
// displacement from the run's GEMM.  Returns false for an idle wave.
__device__ __forceinline__ bool wave_setup(const ScoreArgs &a, const DevComplex &c, int gid,
                                           int b, bool have_d, float3 dA, float *ef,
                                           unsigned short *ct, WaveCtx &x, int cap = kMaxTiles,
                                           bool scan = true) {
    x.lane = threadIdx.x & 63;
    x.gid = gid;
    x.b = b;
    x.pose = a.pos + (size_t)gid * a.P;
    const double *rext = x.pose + 7;
    const double *lext = x.pose + 7 + c.nmr;
    x.rot = rot_of(x.pose + 3);
    const float4 blk = gld(c.lblocks, b);
    x.cb[0] = blk.x;
    x.cb[1] = blk.y;
    x.cb[2] = blk.z;
    rot_apply(x.rot, x.cb[0], x.cb[1], x.cb[2]);
    x.cb[0] += x.pose[0];
    x.cb[1] += x.pose[1];
    x.cb[2] += x.pose[2];
    x.lig = gld(c.lperm, b * kBlock + x.lane);
    // padding lanes sit at -512 A on every axis, padding receptor rows at +512 A
    x.L = make_int4(-kFar, -kFar, -kFar, 0);
    if (x.lig >= 0) {
        double px, py, pz;
        ligand_slot_grid(c, x.pose, x.rot, b, x.lane, have_d, dA, px, py, pz);
        x.L = make_int4(qround(px - x.cb[0]), qround(py - x.cb[1]), qround(pz - x.cb[2]),
                        gld(c.ltype_b, x.lig));
    }
    x.nblk = __popcll(__ballot(x.lig >= 0));
    if (x.lane < c.nmr) ef[x.lane] = (float)rext[x.lane];
    float dr = 0.f, dl = 0.f;
    for (int m = 0; m < c.nmr; m++) dr += fabsf((float)rext[m]) * c.rmode_max[m];
    for (int m = 0; m < c.nml; m++) dl += fabsf((float)lext[m]) * c.lmode_max[m];
    dr = dr * 1.0001f + 1e-3f;
    dl = dl * 1.0001f + 1e-3f;
    // Guard band that keeps this wave's bins exact.  Per axis the fixed-point
    // coordinates differ from the reference's FP64 ones by at most 1.01 grid
    // units (two roundings to 2^-20 A, FP64 noise) plus the FP32 displacement
    // errors k u S of both atoms, u = 2^-24, S the displacement bound (dr,
    // dl) and k the roundings per mode term (FP32 mode and extent, and the
    // sum: at most 11 for the broadcast variant's in-lane FMAs and group sum,
    // nm + 2 for a sum of nm products in any order, as the ligand's).  The
    // ligand's from the run's matrix-core GEMM takes 2 (nm + 2): the bound
    // holds for any order of correctly rounded steps, and the factor covers
    // accumulation steps that round less tightly.  With |dx| + |dy| + |dz| <=
    // 15 sqrt(3) and 3 u 225 for the FP32 sum of squares,
    // |d2 - d2_ref| <= 9.04e-5 + 3.098e-6 (k_r dr + k_l dl) A^2.
    const int k_l = have_d ? 2 * c.nml + 4 : c.nml + 2;
    x.band = fmaxf(a.band, 9.1e-5f + 3.11e-6f * ((float)max(11, c.nmr + 2) * dr +
                                                   (float)k_l * dl));
    x.dlr = dl + dr;
    x.dl = dl;
    x.reach = 15.0f + blk.w + dl + dr + 0.01f;
    x.rmax = x.reach + dr + 0.01f;
    x.fc[0] = (float)x.cb[0];
    x.fc[1] = (float)x.cb[1];
    x.fc[2] = (float)x.cb[2];
    x.t_scan = 0;
    x.ncand = scan ? scan_candidates(c, x, ct, x.t_scan, cap) : 0;
    return true;
}

// The wave's energy and pair count go to block x.b's partial; a second block
// b2 (>= 0) of a two-block wave gets zeros.
__device__ __forceinline__ void write_partials(const ScoreArgs &a, const WaveCtx &x,
                                               double acc, int npairs,
                                               unsigned int nguard, long long n_cand,
                                               long long n_near, long long n_queued,
                                               int b2 = -1) {
    double e = acc;
    long long np = npairs;
    for (int off = 32; off > 0; off >>= 1) {
        e += __shfl_xor(e, off, 64);
        np += __shfl_xor(np, off, 64);
    }
    unsigned int ng = nguard;
    for (int off = 32; off > 0; off >>= 1) ng += __shfl_xor(ng, off, 64);
    if (x.lane == 0) {
        a.partial[(size_t)x.b * a.n_glow + x.gid] = e;
        a.partial_pairs[(size_t)x.b * a.n_glow + x.gid] = np;
        if (b2 >= 0) {
            a.partial[(size_t)b2 * a.n_glow + x.gid] = 0.0;
            a.partial_pairs[(size_t)b2 * a.n_glow + x.gid] = 0;
        }
        if (a.guard_count && ng) atomicAdd(a.guard_count, (unsigned long long)ng);
        if (a.stats) {
            // counters spread over kStatSlots slots each, so that waves rarely collide
            unsigned long long *st = a.stats + (blockIdx.x % kStatSlots);
            atomicAdd(st + 0 * kStatSlots, 1ull);
            atomicAdd(st + 1 * kStatSlots, n_cand ? 1ull : 0ull);
            atomicAdd(st + 2 * kStatSlots, (unsigned long long)n_cand);
            atomicAdd(st + 3 * kStatSlots, (unsigned long long)n_near);
            atomicAdd(st + 4 * kStatSlots, (unsigned long long)n_queued);
            atomicAdd(st + 5 * kStatSlots, (unsigned long long)np);
        }
    }
}

// ---------------------------------------------------------------- broadcast variant
// Near receptor atoms are collected across the visited tiles (their slots in a
// pending list) and displaced 64 at a time: eight lanes read an atom's mode
// line together and sum their products across lanes, then each lane rounds
// one atom to the block's grid.  The displaced atoms are packed into LDS and
// the wave walks them with broadcast LDS reads, four per round.  Each lane
// pre-tests its ligand atom with the FP32 form W = |r|^2 - 2 r.l against
// T - |l|^2 (A^2), and the pairs that pass are queued as one word: the LDS
// byte offsets of the packed receptor atom (high half) and of the ligand atom
// (low half).  The exact fixed-point d2 is formed in the evaluate phase, 64
// queued pairs per round, every lane with a pair (a partial round is padded
// with a far-away atom).  A round's table gather is a buffer load whose offset
// is past the table's end for a pair that does not count (the load returns 0
// without a memory access); its value is added kUnroll rounds later, so that
// kUnroll rounds of gathers are in flight per wave, also across the pre-test
// work between two flushes.  Only whole groups of kUnroll rounds are
// evaluated until the wave's last batch; the packed atoms are double-buffered
// so that the pairs left over from one batch wait in the queue for the next.
//
// Pre-test bound: with u = 2^-24, every pair that can count (d2 < 225 + band)
// has |l| <= rho - 15 and |r| <= rho, rho the wave's reach; converting the
// fixed-point values to float, squaring, the three rounded FMAs of W and the
// rounded T - |l|^2 stay within 28 u rho^2 of the exact d2, so
// T = 225 + 2 band + 2^-18 rho^2 + 10^-3 keeps every such pair (extra pairs
// are rejected by the exact evaluation).
//
// MFMA pre-test (kMfma): V of 32 x 32 pairs is one v_mfma_f32_32x32x8f16 with
// rows (x, y, z, n_hi | n_lo, 1, 1, 0) of the packed receptor atoms and
// columns (-2x, -2y, -2z, 1 | 1, m_hi, m_lo, 0) of the ligand atoms, n and m
// the squared norms split into two halves.  Half precision rounds each
// coordinate within 2^-11 relative, so the cross term is within
// 2^-9 |r||l| <= 2^-9 rho (rho - 15); the split norms and the FP32
// accumulation add less than 2^-18 rho^2.  Pairs with
// V <= 225 + 2 band + 2^-8 rho (rho - 15) + 2^-17 rho^2 + 10^-3 are queued.
// Rows past the batch and padding ligand lanes carry a norm of 60000, so they
// never pass; a wave whose atoms could exceed the half range (|r| >= 250 A)
// takes the FP32 pre-test.
// Two ligand blocks per wave (kLB = 2): the wave takes blocks b and b + 1 of
// one glowworm (a 128-atom ligand block, two atoms per lane), so the
// candidate scan, the tile visits and the displacement of the receptor atoms
// near both blocks are shared.  Both blocks' atoms are rounded to the grid of
// block b's centre.  A pending atom carries a near flag per block; a batch is
// packed in the order [near b only | near both | near b + 1 only], and the
// pre-test makes one pass per block over its range, so every atom is tested
// only against the blocks it is near.  rho, the bound of |r| and of |l| + 15
// about the origin, takes the second block's sphere into account.
// Queue room: kUnroll rounds of 64 pairs left over from a flush, plus the
// pairs of four pre-tested atoms (or of one MFMA 32 x 32 group).
template <int kUnroll, bool kMfma>
constexpr int queue_len() { return 64 * kUnroll + (kMfma ? 1024 : 256); }
constexpr int kCandB = 128;     // candidate-list chunk of the broadcast variant
constexpr int kDispRounds = 8;  // displacement rounds whose mode-line loads are in flight together
constexpr int kRtRows = 2 * 64 + 1;        // two batches of packed atoms and one far-away atom
constexpr unsigned kNoLoad = 0x80000000u;  // buffer offset past the table: the load returns 0

// f(std::integral_constant<int, I>) for I = I0 .. N-1, unrolled.
template <int I, int N, class F>
__device__ __forceinline__ void static_for(F &&f) {
    if constexpr (I < N) {
        f(std::integral_constant<int, I>{});
        static_for<I + 1, N>(f);
    }
}

// v from the lane the DPP control kCtrl selects (0 where it selects none).
template <int kCtrl, int kRowMask = 0xF>
__device__ __forceinline__ float dpp_f32(float v) {
    return __builtin_bit_cast(
        float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), kCtrl, kRowMask, 0xF, false));
}

// Scans within rows of 16 lanes, each step v += v' * g with v' the value of
// the lane s below in the same row, as one v_fmac_f32_dpp per step (from
// dpp_f32 and an fma the compiler builds a v_mov_b32 of zero, a
// v_mov_b32_dpp and a v_fmac_f32).  A lane whose source lies outside its
// row keeps v, as adding 0 * g would.  Each s_nop gives the two wait states
// a DPP read needs after a VALU write of its source.
//
// acc + x(s = 1) gb, then steps s = 1 (g1) and s = 2 (g2)
__device__ __forceinline__ float dpp_scan_x12(float acc, float x, float gb, float g1, float g2) {
    asm("s_nop 1\n\t"
        "v_fmac_f32_dpp %0, %1, %2 row_shr:1 row_mask:0xf bank_mask:0xf\n\t"
        "s_nop 1\n\t"
        "v_fmac_f32_dpp %0, %0, %3 row_shr:1 row_mask:0xf bank_mask:0xf\n\t"
        "s_nop 1\n\t"
        "v_fmac_f32_dpp %0, %0, %4 row_shr:2 row_mask:0xf bank_mask:0xf"
        : "+v"(acc)
        : "v"(x), "v"(gb), "v"(g1), "v"(g2));
    return acc;
}

// steps s = 1 (g1), 2 (g2) and 4 (g4)
__device__ __forceinline__ float dpp_scan_124(float v, float g1, float g2, float g4) {
    asm("s_nop 1\n\t"
        "v_fmac_f32_dpp %0, %0, %1 row_shr:1 row_mask:0xf bank_mask:0xf\n\t"
        "s_nop 1\n\t"
        "v_fmac_f32_dpp %0, %0, %2 row_shr:2 row_mask:0xf bank_mask:0xf\n\t"
        "s_nop 1\n\t"
        "v_fmac_f32_dpp %0, %0, %3 row_shr:4 row_mask:0xf bank_mask:0xf"
        : "+v"(v)
        : "v"(g1), "v"(g2), "v"(g4));
    return v;
}

// float4 at byte offset off from p (a 32-bit offset on the scalar base).
__device__ __forceinline__ float4 gld_at(const float4 *p, unsigned off) {
    typedef float f4 __attribute__((ext_vector_type(4)));
    const f4 v = *(const __attribute__((address_space(1))) f4 *)(
        (const __attribute__((address_space(1))) char *)p + off);
    return make_float4(v.x, v.y, v.z, v.w);
}

// Sum of v over aligned groups of kC lanes (8, 16 or 32), exact in the
// group's last lane: quad swaps, then row_shr:4 (8-lane sums in lanes 4-7 and
// 12-15 of a row), row_shr:8 (16-lane sums in lanes 12-15) and row_bcast:15
// into rows 1 and 3 (32-lane sums in lanes 16-31 of each half).
template <int kC>
__device__ __forceinline__ float group_sum(float v) {
    static_assert(kC == 8 || kC == 16 || kC == 32, "group of 8, 16 or 32 lanes");
    v += dpp_f32<0xB1>(v);  // quad_perm [1, 0, 3, 2]
    v += dpp_f32<0x4E>(v);  // quad_perm [2, 3, 0, 1]
    v += dpp_f32<0x114>(v);
    if constexpr (kC >= 16) v += dpp_f32<0x118>(v);
    if constexpr (kC >= 32) v += dpp_f32<0x142, 0xA>(v);
    return v;
}

// Values the whole wave holds (scalar registers).
__device__ __forceinline__ float uniform_f32(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, v)));
}
__device__ __forceinline__ double uniform_f64(double v) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)u);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(u >> 32));
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

template <class T>
__device__ __forceinline__ T *uniform_ptr(T *p) {
    const unsigned long long v = (unsigned long long)p;
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
    return (T *)(((unsigned long long)hi << 32) | lo);
}

// Work items per atomic of the work-queue form (kPersist = 2).
constexpr int kQChunk = 8;

// kPersist: 0, one work item per wave (the hardware dispatcher balances the
// waves); 1, a grid of resident waves, wave k taking items k, k + W, k + 2W,
// ... (W waves); 2, a grid of resident waves taking kQChunk consecutive items
// per atomic from a counter that starts at zero.
template <int kUnroll, int kMinWaves, bool kMfma, int kWpb, int kTilesB, int kC, int kLB = 1,
          bool kByComp = false, int kPersist = 0>
__global__ __launch_bounds__(64 * kWpb, kMinWaves) void score32b_kernel(ScoreArgs a) {
    static_assert(kLB == 1 || (kLB == 2 && !kMfma), "one block, or two with the FP32 pre-test");
    constexpr int kPend = 64 * kTilesB + 64;  // pending near slots: a batch's worth + one round
    constexpr int kQueue = queue_len<kUnroll, kMfma>();
    constexpr int kGroup = 64 * kUnroll;  // pairs of one group of rounds
    // a pending entry: receptor slot (low 16 bits) and, with two blocks, near
    // flags (bits 30-31, which the 32-bit product of the entry with a mode
    // record's size in bytes drops, so that product is the record's offset)
    using PdT = std::conditional_t<kLB == 2, unsigned, unsigned short>;
    __shared__ int4 s_rt[kWpb][kRtRows];     // packed near receptor atoms: fixed point + row
    __shared__ float4 s_rf[kWpb][64];        //   this batch's atoms in A: x, y, z, |r|^2
    __shared__ unsigned short s_rs[kWpb][2 * 64];  // receptor slots of the packed atoms
    __shared__ PdT s_pd[kWpb][kPend];        // near slots waiting for displacement
    __shared__ int4 s_lg[kWpb][64 * kLB];    // the blocks' ligand atoms: fixed point + type*20
    __shared__ int s_q[kWpb][kQueue + 64];   // queued pairs (LDS offsets of both atoms), 64 dummy slots
    __shared__ float s_ef[kWpb][SDGSO_MAX_MODES];
    __shared__ unsigned short s_ct[kWpb][kCandB];  // candidate tiles, scanned in chunks
    __shared__ float4 s_box[kWpb][5];        // the box test's parameters (read once per batch)
    const int w = kWpb == 1 ? 0 : __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    // one work item: (glowworm, first ligand block) wv of the list
    auto item = [&](const int wv) {
    const int gid = __builtin_amdgcn_readfirstlane(a.work[2 * wv]);
    const int b = __builtin_amdgcn_readfirstlane(a.work[2 * wv + 1]);
    const int kx = __builtin_amdgcn_readfirstlane(a.swarm_cx[gid / a.G]);
    const DevComplex &c = a.cx[kx];
    // the fields the tile visits and the evaluation read, in scalar registers
    const int ntiles = __builtin_amdgcn_readfirstlane(c.ntiles);
    const int nmr = __builtin_amdgcn_readfirstlane(c.nmr);
    const float4 *tiles = uniform_ptr(c.tiles);
    const float4 *rpos32 = uniform_ptr(c.rpos32);
    const double *rrel4 = uniform_ptr(c.rrel4);
    const float4 *rmodA = uniform_ptr((const float4 *)c.rmodA);
    const __amdgpu_buffer_rsrc_t tbl = __builtin_amdgcn_make_buffer_rsrc(
        (void *)uniform_ptr(c.tablep), (short)0, kTableLenP * 4, 0x00020000);
    int4 *rt = s_rt[w];
    float4 *rf = s_rf[w];
    unsigned short *rs = s_rs[w];
    PdT *pd = s_pd[w];
    int4 *lg = s_lg[w];
    int *qq = s_q[w];
    const unsigned q_dummy = (unsigned)(kQueue + (int)(threadIdx.x & 63)) << 2;  // this lane's dummy slot (bytes)
    float *ef = s_ef[w];
    unsigned short *ct = s_ct[w];
    // this lane's slots of the wave's ligand blocks in the run's ligand
    // displacement GEMM output, if any, loaded first: they come from HBM
    const bool have_d = a.ldisp != nullptr;
    float3 dA = make_float3(0.f, 0.f, 0.f), dB = dA;
    if (have_d) {
        const int nlb = __builtin_amdgcn_readfirstlane(c.nlblocks);
        const float *pA = uniform_ptr(a.ldisp + a.ldisp_base[kx] +
                                      ((long long)gid * nlb + b) * 3 * kBlock);
        const int l = threadIdx.x & 63;
        dA = make_float3(gld(pA, l), gld(pA, kBlock + l), gld(pA, 2 * kBlock + l));
        if (kLB == 2 && b + 1 < nlb)
            dB = make_float3(gld(pA, 3 * kBlock + l), gld(pA, 4 * kBlock + l),
                             gld(pA, 5 * kBlock + l));
    }
    WaveCtx x;
    wave_setup(a, c, gid, b, have_d, dA, ef, ct, x, kCandB, kLB == 1);
    const int lane = x.lane;
    lg[lane] = x.L;
    // second block (kLB = 2): b + 1 when the glowworm has it; its sphere (fb,
    // reach_b < 0 without it) and this lane's atom on block b's grid
    int b2 = -1;
    float fb[3] = {0.f, 0.f, 0.f};
    float reach_b = -1.f;
    int4 LB = make_int4(-kFar, -kFar, -kFar, 0);
    int lig_b = -1;
    float rho = x.reach;  // bound of |r| and of |l| + 15 about the origin
    if constexpr (kLB == 2) {
        b2 = b + 1 < __builtin_amdgcn_readfirstlane(c.nlblocks) ? b + 1 : -1;
        if (b2 >= 0) {
            const float4 blk = gld(c.lblocks, b2);
            double cx = blk.x, cy = blk.y, cz = blk.z;
            rot_apply(x.rot, cx, cy, cz);
            cx += x.pose[0];
            cy += x.pose[1];
            cz += x.pose[2];
            lig_b = gld(c.lperm, b2 * kBlock + lane);
            if (lig_b >= 0) {
                double px, py, pz;
                ligand_slot_grid(c, x.pose, x.rot, b2, lane, have_d, dB, px, py, pz);
                LB = make_int4(qround(px - x.cb[0]), qround(py - x.cb[1]), qround(pz - x.cb[2]),
                               gld(c.ltype_b, lig_b));
            }
            fb[0] = uniform_f32((float)cx);
            fb[1] = uniform_f32((float)cy);
            fb[2] = uniform_f32((float)cz);
            reach_b = uniform_f32(15.0f + blk.w + x.dlr + 0.01f);
            const float dx = fb[0] - x.fc[0], dy = fb[1] - x.fc[1], dz = fb[2] - x.fc[2];
            rho = fmaxf(rho, __builtin_sqrtf(dx * dx + dy * dy + dz * dz) * 1.0001f + reach_b + 0.01f);
        }
        lg[64 + lane] = LB;
        x.ncand = scan_candidates<true>(ntiles, tiles, x, ct, x.t_scan, kCandB, fb, reach_b);
    }
    // Box test of a packed atom r (about block b's centre): u = R^T r is the
    // atom in the ligand frame about that centre, and a pair with an atom of
    // a block can count only if u lies within 15 A + the ligand displacement
    // bound (+ 0.01 A for the FP32 rounding of u, at most 1e-4 A) of the
    // block's box.  Block b + 1's box is offset by the difference of centres.
    // Kept in LDS: the rows of R^T, (15 + dl + 0.01)^2, the offset of the
    // centres and both boxes' half extents (-1e30 without a second block, a
    // box that no atom is near).
    float4 *box = s_box[w];
    if (lane == 0) {
        const float4 hA = gld(c.lhalf, b);
        const float box_lim = 15.0f + x.dl + 0.01f;
        float4 hB = make_float4(-1e30f, -1e30f, -1e30f, 0.f);
        float4 off = make_float4(0.f, 0.f, 0.f, 0.f);
        if (kLB == 2 && b2 >= 0) {
            hB = gld(c.lhalf, b2);
            const float4 ca = gld(c.lblocks, b), cbb = gld(c.lblocks, b2);
            off = make_float4(ca.x - cbb.x, ca.y - cbb.y, ca.z - cbb.z, 0.f);
        }
        const double *m = x.rot.m;
        box[0] = make_float4((float)m[0], (float)m[3], (float)m[6], box_lim * box_lim);
        box[1] = make_float4((float)m[1], (float)m[4], (float)m[7], off.y);
        box[2] = make_float4((float)m[2], (float)m[5], (float)m[8], off.z);
        box[3] = make_float4(hA.x, hA.y, hA.z, off.x);
        box[4] = make_float4(hB.x, hB.y, hB.z, 0.f);
    }
    // the set-up's wave-uniform results, in scalar registers
    const double cb0 = uniform_f64(x.cb[0]), cb1 = uniform_f64(x.cb[1]), cb2 = uniform_f64(x.cb[2]);
    const float fc0 = uniform_f32(x.fc[0]), fc1 = uniform_f32(x.fc[1]), fc2 = uniform_f32(x.fc[2]);
    const float reach = uniform_f32(x.reach);
    rho = uniform_f32(rho);
    if (lane == 0) rt[2 * 64] = make_int4(kFar, kFar, kFar, 0);
    wave_sync_lds();
    // queue words: LDS byte offsets from s_rt[0][0] (high half) and s_lg[0][0] (low half)
    const char *rt_base = (const char *)&s_rt[0][0];
    const char *lg_base = (const char *)&s_lg[0][0];
    const unsigned rt0 = (unsigned)(w * kRtRows) * 16u;
    const unsigned lane_lo = (unsigned)(w * 64 * kLB + lane) * 16u;
    const int null_tag = (int)(((rt0 + 2 * 64 * 16u) << 16) | (unsigned)(w * 64 * kLB) * 16u);
    const float band = uniform_f32(x.band);
    const float bandq = band * kQ2;
    const float thr = 225.0f + 2.0f * band + 0x1p-18f * rho * rho + 1e-3f;
    const float reach2 = reach * reach;
    const float reach2b = reach_b >= 0.f ? reach_b * reach_b : -1.f;
    // this lane's pre-test operands: -2 l and T - |l|^2 (padding lanes never pass)
    const float lx = (float)x.L.x * kInvQ, ly = (float)x.L.y * kInvQ, lz = (float)x.L.z * kInvQ;
    const float m2x = -2.f * lx, m2y = -2.f * ly, m2z = -2.f * lz;
    const float tl = x.lig >= 0 ? thr - __builtin_fmaf(lz, lz, __builtin_fmaf(ly, ly, lx * lx))
                                : -1e30f;
    // and for the second block's atom
    const float lbx = (float)LB.x * kInvQ, lby = (float)LB.y * kInvQ, lbz = (float)LB.z * kInvQ;
    const float m2xb = -2.f * lbx, m2yb = -2.f * lby, m2zb = -2.f * lbz;
    const float tlb = lig_b >= 0
                          ? thr - __builtin_fmaf(lbz, lbz, __builtin_fmaf(lby, lby, lbx * lbx))
                          : -1e30f;
    // MFMA pre-test: B operands of the two 32-atom column groups, and the threshold
    const bool mfma_ok = kMfma && uniform_f32(x.rmax) < 250.0f;
    const float thr16 = 225.0f + 2.0f * band + 0x1p-8f * reach * (reach - 15.0f) +
                        0x1p-17f * reach * reach + 1e-3f;
    half4 bop[2];
    if (kMfma) {
#pragma unroll
        for (int g = 0; g < 2; g++) {
            const int j = 32 * g + (lane & 31);
            const int4 q = lg[j];
            const bool real = gld(c.lperm, x.b * kBlock + j) >= 0;
            const float px = real ? (float)q.x * kInvQ : 0.f, py = real ? (float)q.y * kInvQ : 0.f,
                        pz = real ? (float)q.z * kInvQ : 0.f;
            const float m = real ? __builtin_fmaf(pz, pz, __builtin_fmaf(py, py, px * px)) : 60000.f;
            const _Float16 mh = (_Float16)m;
            const _Float16 ml = (_Float16)(m - (float)mh);
            if (lane < 32)
                bop[g] = half4{(_Float16)(-2.f * px), (_Float16)(-2.f * py), (_Float16)(-2.f * pz),
                               (_Float16)1.f};
            else
                bop[g] = half4{(_Float16)1.f, mh, ml, (_Float16)0.f};
        }
    }

    double acc = 0.0;
    int npairs = 0;          // in-cutoff pairs evaluated by this lane
    unsigned int nguard = 0;
    int qn = 0;
    int old = 0;             // queued pairs of the previous batch (front of the queue)
    int p = 0;               // buffer of the batch being packed
    long long n_near = 0, n_queued = 0;
    // table values in flight: slot j holds the load of the last round that used it
    float ring[kUnroll];
#pragma unroll
    for (int j = 0; j < kUnroll; j++) ring[j] = 0.f;

    // One round: the 64 queued pairs from entry e0, table load into ring slot j
    // (whose previous load is added first).
    // A round's inputs: the queue word and both packed atoms.
    struct RoundIn {
        int tag;
        int4 r, l;
    };
    auto fetch_atoms = [&](RoundIn &in) {
        in.r = *(const int4 *)(rt_base + ((unsigned)in.tag >> 16));
        in.l = *(const int4 *)(lg_base + (in.tag & 0xffff));
    };
    auto compute = [&](auto jc, const RoundIn &in) {
        constexpr int j = decltype(jc)::value;
        const int tag = in.tag;
        const int4 r = in.r, l = in.l;
        const PairClass pc = classify(d2_fixed(l, r), bandq);
        // r.w and l.w are byte offsets of the receptor type's row and the
        // ligand type's column; a bin is 4 * 168 bytes
        const unsigned off =
            pc.take ? __umul24((unsigned)pc.bin, 4u * SDGSO_NTYPES) + (unsigned)(r.w + l.w)
                    : kNoLoad;
        acc += (double)ring[j];
        ring[j] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(tbl, off, 0, 0));
        npairs += pc.take ? 1 : 0;
        if (pc.near) {
            // ligand slot li of block x.b (li >= 64: slot li - 64 of block x.b + 1)
            const int li = (int)((unsigned)(tag & 0xffff) >> 4) - w * 64 * kLB;
            const int ki = (int)((unsigned)tag >> 20) - w * kRtRows;
            const double2 g2 = guard_pair64(c, x.pose, gld(c.lperm, x.b * kBlock + li),
                                            gld(c.rperm, (int)rs[ki]));
            acc += g2.x;
            npairs += g2.y != 0.0 ? 1 : 0;
            nguard++;
        }
    };
    // One round: the 64 queued pairs from entry e0, table load into ring slot j
    // (whose previous load is added first).
    auto round = [&](auto jc, int e0) {
        RoundIn in;
        in.tag = qq[e0 + lane];
        fetch_atoms(in);
        compute(jc, in);
    };
    // Whole groups of kUnroll rounds: every slot is reloaded in every group, so
    // a slot's load has the other kUnroll - 1 loads behind it when it is added.
    // The LDS reads run ahead of the arithmetic: the queue word two rounds
    // ahead and the atoms one round ahead.
    auto eval_groups = [&](int ng) {
        for (int g = 0; g < ng; g++) {
            const int e = kGroup * g;
            RoundIn in[kUnroll];
            in[0].tag = qq[e + lane];
            if constexpr (kUnroll > 1) in[1].tag = qq[e + 64 + lane];
            fetch_atoms(in[0]);
            static_for<0, kUnroll>([&](auto jc) {
                constexpr int j = decltype(jc)::value;
                if constexpr (j + 2 < kUnroll) in[j + 2].tag = qq[e + 64 * (j + 2) + lane];
                if constexpr (j + 1 < kUnroll) fetch_atoms(in[j + 1]);
                compute(jc, in[j]);
            });
        }
    };
    // Rounds 0..nround-1, for the last flush of a batch or of the wave.
    auto eval_tail = [&](int nround) {
        for (int r0 = 0; r0 < nround; r0 += kUnroll)
            static_for<0, kUnroll>([&](auto jc) {
                if (r0 + decltype(jc)::value < nround) round(jc, 64 * (r0 + decltype(jc)::value));
            });
    };
    // Evaluate the whole groups of the queue; the rest (< kGroup) moves to the front.
    auto flush_groups = [&]() {
        const int ng = qn / kGroup;
        if (ng == 0) return;
        n_queued += kGroup * ng;
        wave_sync_lds();
        eval_groups(ng);
        const int rem = qn - kGroup * ng;
        int v[kUnroll];
#pragma unroll
        for (int j = 0; j < kUnroll; j++)
            v[j] = 64 * j + lane < rem ? qq[kGroup * ng + 64 * j + lane] : 0;
        wave_sync_lds();
#pragma unroll
        for (int j = 0; j < kUnroll; j++)
            if (64 * j + lane < rem) qq[64 * j + lane] = v[j];
        qn = rem;
        old = max(0, old - kGroup * ng);
    };
    // Evaluate everything (after flush_groups: qn < kGroup), padding the last round.
    auto flush_all = [&]() {
        if (qn == 0) return;
        n_queued += qn;
        const int pad = (-qn) & 63;
        if (lane < pad) qq[qn + lane] = null_tag;
        wave_sync_lds();
        eval_tail((qn + pad) >> 6);
        wave_sync_lds();
        qn = 0;
        old = 0;
    };

    // Displace the pending atoms pd[0..n) (n <= 64, one per lane), pack them
    // into buffer p, queue their pairs that pass the pre-test and evaluate the
    // whole rounds.  Pairs of the previous batch (buffer 1-p) that are still
    // queued at the end are evaluated too, since the next batch overwrites it.
    auto batch = [&](int n) {
        old = qn;
        // this lane's atom: slot, near flags, FP64 coordinates and type row
        // (loads first)
        int slot = 0;
        unsigned fl = 0;
        double2 rxy = make_double2(0.0, 0.0), rzw = make_double2(0.0, 0.0);
        if (lane < n) {
            const unsigned e = pd[lane];
            slot = (int)(e & 0xffffu);
            fl = e >> 30;
            rxy = gld2(rrel4, 2 * (size_t)slot);
            rzw = gld2(rrel4, 2 * (size_t)slot + 1);
        }
        // displacement D = sum_m M_m e_m of the batch's atoms, kC lanes per atom: in
        // round k lane l takes float4 j = l % kC of atom k * 64 / kC + l / kC (an
        // atom's modes are kC float4 in whole 128-byte lines, so a round reads
        // 64 / kC lines), weighs its four floats by their modes' extents per
        // component, and the group's sums reach its last lane; they are staged
        // in rf, which the packing below overwrites.  The loads of kDispRounds
        // rounds are issued together.
        float4 D = make_float4(0.f, 0.f, 0.f, 0.f);
        if constexpr (kByComp) {
            // component-major modes: float4 j of an atom holds modes
            // 4*pos .. 4*pos + 3 of coordinate comp (q float4 per coordinate);
            // each lane weighs its four floats by their extents, and a scan over
            // the q lanes of a coordinate (shifts 1, 2, 4 within the row, each
            // lane adding only from lanes of its own coordinate) leaves the
            // coordinate's sum in its last lane, which stages it in rf
            const int j = lane & (kC - 1);
            const unsigned j16 = 16u * (unsigned)j;  // byte offset of float4 j
            const int q = (nmr + 3) >> 2;
            const int comp = (j >= q ? 1 : 0) + (j >= 2 * q ? 1 : 0) + (j >= 3 * q ? 1 : 0);
            const int pos = j - comp * q;
            float w4[4];
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const int m = 4 * pos + t;
                w4[t] = comp < 3 && m < nmr ? ef[min(m, SDGSO_MAX_MODES - 1)] : 0.f;
            }
            const float g1 = pos >= 1 ? 1.f : 0.f, g2 = pos >= 2 ? 1.f : 0.f,
                        g4 = pos >= 4 ? 1.f : 0.f;
            const bool last = comp < 3 && pos == q - 1;
            constexpr int kApr = 64 / kC;  // atoms per round
            const int nround = (n + kApr - 1) / kApr;
            for (int k0 = 0; k0 < nround; k0 += kDispRounds) {
                float4 qv[kDispRounds];
                int av[kDispRounds];
#pragma unroll
                for (int r = 0; r < kDispRounds; r++) {
                    av[r] = (k0 + r) * kApr + lane / kC;
                    qv[r] = (k0 + r < nround && av[r] < n)
                                ? gld_at(rmodA, (unsigned)pd[av[r]] * (16u * kC) + j16)
                                : make_float4(0.f, 0.f, 0.f, 0.f);
                }
#pragma unroll
                for (int r = 0; r < kDispRounds; r++) {
                    if (k0 + r >= nround) break;
                    float v = qv[r].x * w4[0];
                    v = __builtin_fmaf(qv[r].y, w4[1], v);
                    v = __builtin_fmaf(qv[r].z, w4[2], v);
                    v = __builtin_fmaf(qv[r].w, w4[3], v);
                    // q >= 3 here (11 to 20 modes); with q <= 4 every g4 is 0
                    v = dpp_scan_124(v, g1, g2, g4);
                    if (last && av[r] < n) ((float *)&rf[av[r]])[comp] = v;
                }
            }
            wave_sync_lds();
            if (lane < n) D = rf[lane];
        } else if (kC == 8 && nmr > 0) {
            // packed by coordinate (up to 10 modes): coordinate k holds floats
            // [k cs, (k + 1) cs), cs = max(nmr, 4), so the float4 of lane j
            // holds modes of its first coordinate ka and maybe of ka + 1.  The
            // lane weighs its floats into a (coordinate ka) and b (ka + 1); b
            // moves to lane j + 1, whose first coordinate is ka + 1 (a
            // coordinate spans at least four floats); then a scan over the
            // lanes of one first coordinate (at most three: shifts 1 and 2)
            // leaves each coordinate's sum in its last lane
            const int j = lane & 7;
            const unsigned j16 = 16u * (unsigned)j;  // byte offset of float4 j
            const int cs = nmr > 4 ? nmr : 4;
            const int ka = (4 * j) / cs;
            float wa[4], wb[4];
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const int f = 4 * j + t, k = f / cs, m = f - k * cs;
                const float e = k < 3 && m < nmr ? ef[min(m, SDGSO_MAX_MODES - 1)] : 0.f;
                wa[t] = k == ka ? e : 0.f;
                wb[t] = k == ka + 1 ? e : 0.f;
            }
            const float gb = j >= 1 && (4 * j - 1) / cs == ka && (4 * j - 4) / cs != ka ? 1.f : 0.f;
            const float g1 = j >= 1 && (4 * j - 4) / cs == ka ? 1.f : 0.f;
            const float g2 = j >= 2 && (4 * j - 8) / cs == ka ? 1.f : 0.f;
            const bool last = ka < 3 && (j == 7 || (4 * j + 4) / cs != ka);
            constexpr int kApr = 8;  // atoms per round
            const int nround = (n + kApr - 1) / kApr;
            for (int k0 = 0; k0 < nround; k0 += kDispRounds) {
                float4 qv[kDispRounds];
                int av[kDispRounds];
#pragma unroll
                for (int r = 0; r < kDispRounds; r++) {
                    av[r] = (k0 + r) * kApr + lane / 8;
                    qv[r] = (k0 + r < nround && av[r] < n)
                                ? gld_at(rmodA, (unsigned)pd[av[r]] * (16u * 8u) + j16)
                                : make_float4(0.f, 0.f, 0.f, 0.f);
                }
#pragma unroll
                for (int r = 0; r < kDispRounds; r++) {
                    if (k0 + r >= nround) break;
                    float va = qv[r].x * wa[0];
                    va = __builtin_fmaf(qv[r].y, wa[1], va);
                    va = __builtin_fmaf(qv[r].z, wa[2], va);
                    va = __builtin_fmaf(qv[r].w, wa[3], va);
                    float vb = qv[r].y * wb[1];
                    vb = __builtin_fmaf(qv[r].z, wb[2], vb);
                    vb = __builtin_fmaf(qv[r].w, wb[3], vb);
                    const float v = dpp_scan_x12(va, vb, gb, g1, g2);
                    if (last && av[r] < n) ((float *)&rf[av[r]])[ka] = v;
                }
            }
            wave_sync_lds();
            if (lane < n) D = rf[lane];
        } else if (nmr > 0) {
            const int j = lane & (kC - 1);
            const unsigned j16 = 16u * (unsigned)j;  // byte offset of float4 j
            float wx[4], wy[4], wz[4];
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const int f = 4 * j + t, m = f / 3, comp = f - 3 * m;
                const float e = f < 3 * nmr ? ef[min(m, SDGSO_MAX_MODES - 1)] : 0.f;
                wx[t] = comp == 0 ? e : 0.f;
                wy[t] = comp == 1 ? e : 0.f;
                wz[t] = comp == 2 ? e : 0.f;
            }
            constexpr int kApr = 64 / kC;  // atoms per round
            const int nround = (n + kApr - 1) / kApr;
            for (int k0 = 0; k0 < nround; k0 += kDispRounds) {
                float4 q[kDispRounds];
                int av[kDispRounds];
#pragma unroll
                for (int r = 0; r < kDispRounds; r++) {
                    av[r] = (k0 + r) * kApr + lane / kC;
                    q[r] = (k0 + r < nround && av[r] < n)
                               ? gld_at(rmodA, (unsigned)pd[av[r]] * (16u * kC) + j16)
                               : make_float4(0.f, 0.f, 0.f, 0.f);
                }
#pragma unroll
                for (int r = 0; r < kDispRounds; r++) {
                    if (k0 + r >= nround) break;
                    float px = q[r].x * wx[0], py = q[r].x * wy[0], pz = q[r].x * wz[0];
                    px = __builtin_fmaf(q[r].y, wx[1], px);
                    py = __builtin_fmaf(q[r].y, wy[1], py);
                    pz = __builtin_fmaf(q[r].y, wz[1], pz);
                    px = __builtin_fmaf(q[r].z, wx[2], px);
                    py = __builtin_fmaf(q[r].z, wy[2], py);
                    pz = __builtin_fmaf(q[r].z, wz[2], pz);
                    px = __builtin_fmaf(q[r].w, wx[3], px);
                    py = __builtin_fmaf(q[r].w, wy[3], py);
                    pz = __builtin_fmaf(q[r].w, wz[3], pz);
                    px = group_sum<kC>(px);
                    py = group_sum<kC>(py);
                    pz = group_sum<kC>(pz);
                    if (j == kC - 1 && av[r] < n) rf[av[r]] = make_float4(px, py, pz, 0.f);
                }
            }
            wave_sync_lds();
            if (lane < n) D = rf[lane];
        }
        // this lane's atom on the block's grid, and in FP32
        int4 R = make_int4(kFar, kFar, kFar, 0);
        float fx = 0.f, fy = 0.f, fz = 0.f;
        if (lane < n) {
            R = make_int4(qround((rxy.x - cb0) + (double)D.x), qround((rxy.y - cb1) + (double)D.y),
                          qround((rzw.x - cb2) + (double)D.z), (int)rzw.y);
            fx = (float)R.x * kInvQ;
            fy = (float)R.y * kInvQ;
            fz = (float)R.z * kInvQ;
        }
        // box test (see the set-up): the near flags of atoms that cannot
        // count with a block are cleared
        {
            const float4 q0 = box[0], q1 = box[1], q2 = box[2], hA = box[3];
            const float ux = __builtin_fmaf(q0.x, fx, __builtin_fmaf(q0.y, fy, q0.z * fz));
            const float uy = __builtin_fmaf(q1.x, fx, __builtin_fmaf(q1.y, fy, q1.z * fz));
            const float uz = __builtin_fmaf(q2.x, fx, __builtin_fmaf(q2.y, fy, q2.z * fz));
            auto in_box = [&](float vx, float vy, float vz, float4 h) {
                const float ex = fmaxf(fabsf(vx) - h.x, 0.f);
                const float ey = fmaxf(fabsf(vy) - h.y, 0.f);
                const float ez = fmaxf(fabsf(vz) - h.z, 0.f);
                return __builtin_fmaf(ez, ez, __builtin_fmaf(ey, ey, ex * ex)) <= q0.w;
            };
            unsigned keep = in_box(ux, uy, uz, hA) ? 1u : 0u;
            if constexpr (kLB == 2) {
                const float4 hB = box[4];
                if (in_box(ux + hA.w, uy + q1.w, uz + q2.w, hB)) keep |= 2u;
                fl &= keep;
            } else {
                fl = keep;
            }
            if (lane >= n) fl = 0u;
        }
        // packing position: the atoms near the block first, [0, a_end), then
        // the others; with two blocks the order is [near block b only | near
        // both | near block b + 1 only | near neither], so that each block's
        // atoms form one range, [0, a_end) and [b_begin, b_end)
        const unsigned long long m_a = __ballot(fl == 1u);
        const unsigned long long m_ab = __ballot(fl == 3u);
        const unsigned long long m_b = __ballot(fl == 2u);
        const unsigned long long m_0 = __ballot(lane < n && fl == 0u);
        const int n_a = __popcll(m_a), n_ab = __popcll(m_ab), n_b = __popcll(m_b);
        const unsigned long long mine = fl == 1u ? m_a : (fl == 3u ? m_ab : (fl == 2u ? m_b : m_0));
        const int pos = (fl == 1u ? 0 : (fl == 3u ? n_a : (fl == 2u ? n_a + n_ab : n_a + n_ab + n_b))) +
                        (int)__builtin_amdgcn_mbcnt_hi((unsigned)(mine >> 32),
                                                       __builtin_amdgcn_mbcnt_lo((unsigned)mine, 0u));
        const int a_end = n_a + n_ab, b_begin = n_a, b_end = n_a + n_ab + n_b;
        wave_sync_lds();  // every lane has read its staged displacement
        if (lane < n) {
            rt[64 * p + pos] = R;
            rf[pos] = make_float4(fx, fy, fz, __builtin_fmaf(fz, fz, __builtin_fmaf(fy, fy, fx * fx)));
            rs[64 * p + pos] = (unsigned short)slot;
        }
        // a round past the batch's end reads a far-away atom
        if (kLB == 2 && lane >= n) rf[lane] = make_float4(0.f, 0.f, 0.f, 1e30f);
        wave_sync_lds();
        const unsigned hib = rt0 + 64u * 16u * (unsigned)p;  // offset of this buffer's atom 0
        if (mfma_ok) {
            // lane l: row 8*(v/4) + 4*(l/32) + v%4 of accumulator v, column l%32
            for (int h = 0; h * 32 < a_end; h++) {
                const int ri = 32 * h + (lane & 31);
                const float4 r = rf[ri];
                half4 av;
                if (ri < a_end) {
                    const _Float16 nh = (_Float16)r.w;
                    if (lane < 32)
                        av = half4{(_Float16)r.x, (_Float16)r.y, (_Float16)r.z, nh};
                    else
                        av = half4{(_Float16)(r.w - (float)nh), (_Float16)1.f, (_Float16)1.f,
                                   (_Float16)0.f};
                } else {
                    av = lane < 32 ? half4{(_Float16)0.f, (_Float16)0.f, (_Float16)0.f,
                                           (_Float16)60000.f}
                                   : half4{(_Float16)0.f, (_Float16)0.f, (_Float16)0.f,
                                           (_Float16)0.f};
                }
#pragma unroll 1
                for (int g = 0; g < 2; g++) {
                    if (32 * g >= x.nblk) break;
                    if (qn > kQueue - 1024) flush_groups();
                    floatx16 d = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f,
                                  0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
                    d = __builtin_amdgcn_mfma_f32_32x32x8f16(av, g ? bop[1] : bop[0], d, 0, 0, 0);
                    const unsigned tagb =
                        ((hib + 16u * (unsigned)(32 * h + 4 * (lane >> 5))) << 16) |
                        ((unsigned)(w * 64 + 32 * g + (lane & 31)) * 16u);
#pragma unroll
                    for (int v = 0; v < 16; v++) {
                        const bool hit = d[v] <= thr16;
                        const unsigned long long m = __ballot(hit);
                        if (m) {
                            const unsigned qb = __builtin_amdgcn_readfirstlane((unsigned)qn << 2);
                            if (hit) {
                                const unsigned rank = __builtin_amdgcn_mbcnt_hi(
                                    (unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
                                *(int *)((char *)qq + ((rank << 2) + qb)) =
                                    (int)(tagb + ((16u * (unsigned)(8 * (v >> 2) + (v & 3))) << 16));
                            }
                            qn += __popcll(m);
                        }
                    }
                }
            }
        } else if constexpr (kLB == 2) {
            // a pass per block over its range, four packed atoms per round;
            // the ranges are widened to whole rounds (an atom outside a
            // block's reach has no pair that counts with it, and rf[n..63]
            // holds a far-away atom that never passes)
            auto pretest = [&](int i_begin, int i_end, float ox, float oy, float oz, float ot,
                               unsigned olo) {
                // queue word of atom i0 + j: tr + (j << 20), tr kept per round
                unsigned tr = ((hib + 16u * (unsigned)i_begin) << 16) | olo;
                for (int i0 = i_begin; i0 < i_end; i0 += 4, tr += 4u << 20) {
                    float4 r[4];
#pragma unroll
                    for (int j = 0; j < 4; j++) r[j] = rf[i0 + j];
                    unsigned long long m[4];
                    bool in[4];
#pragma unroll
                    for (int j = 0; j < 4; j++) {
                        const float wv2 = __builtin_fmaf(
                            r[j].x, ox, __builtin_fmaf(r[j].y, oy, __builtin_fmaf(r[j].z, oz, r[j].w)));
                        in[j] = wv2 <= ot;
                        m[j] = __ballot(in[j]);
                    }
                    // for an atom with passing pairs every lane stores: a
                    // passing pair at the queue's end (qb, bytes) plus its
                    // rank, the others into the wave's dummy slots past the
                    // queue, so there is no exec mask change per atom
#pragma unroll
                    for (int j = 0; j < 4; j++) {
                        if (m[j]) {
                            const unsigned qb = __builtin_amdgcn_readfirstlane((unsigned)qn << 2);
                            const unsigned rank = __builtin_amdgcn_mbcnt_hi(
                                (unsigned)(m[j] >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m[j], 0u));
                            // the address is formed in every lane (the empty asm
                            // keeps it from moving under an exec mask), then one
                            // v_cndmask on the compare's lane mask picks it
                            unsigned hit_addr = (rank << 2) + qb;
                            asm volatile("" : "+v"(hit_addr));
                            const unsigned addr = in[j] ? hit_addr : q_dummy;
                            *(int *)((char *)qq + addr) = (int)(tr + ((unsigned)j << 20));
                            qn += __popcll(m[j]);
                        }
                    }
                    if (qn > kQueue - 4 * 64) flush_groups();
                }
            };
            pretest(0, a_end, m2x, m2y, m2z, tl, lane_lo);
            pretest(b_begin & ~3, b_end, m2xb, m2yb, m2zb, tlb, lane_lo + 64u * 16u);
        } else {
            // four packed atoms per round: their broadcast reads are in flight together
            for (int i0 = 0; i0 < a_end; i0 += 4) {
                float4 r[4];
#pragma unroll
                for (int j = 0; j < 4; j++) r[j] = rf[min(i0 + j, 63)];
                unsigned long long m[4];
                bool in[4];
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const float wv2 = __builtin_fmaf(
                        r[j].x, m2x, __builtin_fmaf(r[j].y, m2y, __builtin_fmaf(r[j].z, m2z, r[j].w)));
                    in[j] = wv2 <= tl;
                    m[j] = i0 + j < a_end ? __ballot(in[j]) : 0ull;
                }
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    if (m[j]) {
                        const unsigned hi = __builtin_amdgcn_readfirstlane(
                            (hib + 16u * (unsigned)(i0 + j)) << 16);
                        const unsigned qb = __builtin_amdgcn_readfirstlane((unsigned)qn << 2);
                        if (in[j]) {
                            const unsigned rank = __builtin_amdgcn_mbcnt_hi(
                                (unsigned)(m[j] >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m[j], 0u));
                            *(int *)((char *)qq + ((rank << 2) + qb)) = (int)(hi | lane_lo);
                        }
                        qn += __popcll(m[j]);
                    }
                }
                if (qn > kQueue - 4 * 64) flush_groups();
            }
        }
        flush_groups();
        if (old > 0) flush_all();
        p ^= 1;
    };

    // Visit the candidate tiles (chunks of kCandB): near test of every slot
    // against the block sphere; near slots join the pending list.
    int npd = 0;
    long long n_cand = 0;
    // near test of this lane's atom of tile t (position pp); near slots are
    // appended (with two blocks, with a near flag per block)
    auto visit = [&](int t, float4 pp) {
        const float ax = pp.x - fc0;
        const float ay = pp.y - fc1;
        const float az = pp.z - fc2;
        bool nearb = ax * ax + ay * ay + az * az <= reach2;
        unsigned fl = 0;
        if constexpr (kLB == 2) {
            const float bx = pp.x - fb[0];
            const float by = pp.y - fb[1];
            const float bz = pp.z - fb[2];
            const bool nb = bx * bx + by * by + bz * bz <= reach2b;
            fl = (nearb ? 1u : 0u) | (nb ? 2u : 0u);
            nearb = nearb || nb;
        }
        const unsigned long long m = __ballot(nearb);
        if (!m) return;
        if (nearb)
            pd[npd + __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32),
                                               __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u))] =
                (PdT)((unsigned)(t * kTile + lane) | (fl << 30));
        npd += __popcll(m);
        n_near += __popcll(m);
    };
    for (;;) {
        const int nc = x.ncand;
        n_cand += nc;
        // the chunk's tile list in registers; tile k is lane k of c0 / c1
        const int c0 = ct[lane], c1 = ct[64 + lane];
        auto tile_of = [&](int k) {
            k = min(k, nc - 1);
            return k < 64 ? __builtin_amdgcn_readlane(c0, k) : __builtin_amdgcn_readlane(c1, k - 64);
        };
        for (int k = 0; k < nc; k += kTilesB) {
            // kTilesB tiles per round, their loads in flight together (a tile past
            // the list's end re-reads the last one and is not visited)
            int tt[kTilesB];
            float4 pp[kTilesB];
#pragma unroll
            for (int j = 0; j < kTilesB; j++) {
                tt[j] = tile_of(k + j);
                pp[j] = gld(rpos32, tt[j] * kTile + lane);
            }
#pragma unroll
            for (int j = 0; j < kTilesB; j++)
                if (k + j < nc) visit(tt[j], pp[j]);
            while (npd >= 64) {
                wave_sync_lds();
                batch(64);
                // move the remaining pending slots to the front
                const int rem = npd - 64;
                PdT v[kTilesB];
#pragma unroll
                for (int j = 0; j < kTilesB; j++)
                    v[j] = 64 * j + lane < rem ? pd[64 + 64 * j + lane] : (PdT)0;
                wave_sync_lds();
#pragma unroll
                for (int j = 0; j < kTilesB; j++)
                    if (64 * j + lane < rem) pd[64 * j + lane] = v[j];
                npd = rem;
            }
        }
        if (x.t_scan >= ntiles) break;
        __builtin_amdgcn_wave_barrier();  // every lane has read the list before it is rewritten
        x.ncand = scan_candidates<kLB == 2>(ntiles, tiles, x, ct, x.t_scan, kCandB, fb, reach_b);
    }
    if (npd) {
        wave_sync_lds();
        batch(npd);
    }
    flush_all();
#pragma unroll
    for (int j = 0; j < kUnroll; j++) acc += (double)ring[j];
    write_partials(a, x, acc, npairs, nguard, n_cand, n_near, n_queued, b2);
    };
    if constexpr (kPersist == 0) {
        int wv = blockIdx.x * kWpb + w;
        const int K = a.xcd_chunk;
        if (K > 0 && wv < a.n_work / (8 * K) * (8 * K)) {
            const int i = wv >> 3;
            wv = ((i / K) * 8 + (wv & 7)) * K + i % K;
        }
        if (wv < a.n_work) item(wv);
    } else if constexpr (kPersist == 1) {
        for (int wv = blockIdx.x * kWpb + w; wv < a.n_work; wv += gridDim.x * kWpb) item(wv);
    } else {
        for (;;) {
            unsigned v = 0;
            if ((threadIdx.x & 63) == 0) v = atomicAdd(a.counter, (unsigned)kQChunk);
            const int q0 = (int)__builtin_amdgcn_readfirstlane(v);
            if (q0 >= a.n_work) return;
            for (int wv = q0; wv < min(q0 + kQChunk, a.n_work); wv++) item(wv);
        }
    }
}

// Per glowworm: energy = partials in block order (FP64), score, luciferin,
// reference point (prepare_gso_tasks.cc:513-526, 558, 1571).  The partials
// are block-major, so the threads of a wave read consecutive glowworms.
__global__ __launch_bounds__(256) void finish32_kernel(ScoreArgs a) {
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= a.n_glow) return;
    const DevComplex &c = a.cx[a.swarm_cx[gid / a.G]];
    double e = 0.0;
    long long np = 0;
    for (int b = 0; b < c.nlblocks; b++) {
        e += a.partial[(size_t)b * a.n_glow + gid];
        np += a.partial_pairs[(size_t)b * a.n_glow + gid];
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

// Variant table of the FP32 scoring kernel: filter x (rounds of table
// gathers in flight, min waves/SIMD).
namespace {
struct Variant {
    const char *name;
    // for 8, 16 and 32 float4 of modes per receptor slot, mode-major, and for
    // 16 component-major (modes_by_component)
    void (*kernel[4])(ScoreArgs);
    int waves;                     // waves per workgroup
    int blocks = 1;                // ligand blocks per wave (work list of block pairs for 2)
    int persist = 0;               // kPersist of the kernel: a grid of resident waves if > 0
};
#define SD_BCAST(R, O, M, T) \
    { score32b_kernel<R, O, M, 1, T, 8>, score32b_kernel<R, O, M, 1, T, 16>, \
      score32b_kernel<R, O, M, 1, T, 32>, score32b_kernel<R, O, M, 1, T, 16, 1, true> }
#define SD_BCAST2(R, O, T) \
    { score32b_kernel<R, O, false, 1, T, 8, 2>, score32b_kernel<R, O, false, 1, T, 16, 2>, \
      score32b_kernel<R, O, false, 1, T, 32, 2>, score32b_kernel<R, O, false, 1, T, 16, 2, true> }
#define SD_BCAST2P(R, O, T, P) \
    { score32b_kernel<R, O, false, 1, T, 8, 2, false, P>, score32b_kernel<R, O, false, 1, T, 16, 2, false, P>, \
      score32b_kernel<R, O, false, 1, T, 32, 2, false, P>, score32b_kernel<R, O, false, 1, T, 16, 2, true, P> }

const Variant kVariants[] = {
    {"bcast-mfma16/ring2/occ4", SD_BCAST(2, 4, true, 4), 1},
    {"bcast/ring4/occ4", SD_BCAST(4, 4, false, 4), 1},
    {"bcast2/ring4/occ4", SD_BCAST2(4, 4, 4), 1, 2},
    {"bcast2-queue/ring4/occ4", SD_BCAST2P(4, 4, 4, 2), 1, 2, 2},
};
constexpr int kNumVariants = sizeof(kVariants) / sizeof(kVariants[0]);
}  // namespace

int score32_variants() { return kNumVariants; }

// The variant of a run that has not been tuned: 128-atom ligand blocks, the
// fastest on almost every complex of the measured workloads.
int score32_default_variant() {
    for (int v = 0; v < kNumVariants; v++)
        if (std::string(kVariants[v].name) == "bcast2/ring4/occ4") return v;
    return 0;
}

int score32_variant_blocks(int v) { return (v >= 0 && v < kNumVariants) ? kVariants[v].blocks : 1; }

const char *score32_variant_name(int v) {
    return (v >= 0 && v < kNumVariants) ? kVariants[v].name : "?";
}

__global__ void zero_counter_kernel(unsigned *c) {
    if (threadIdx.x == 0) *c = 0u;
}

// Resident one-wave workgroups of a persistent kernel: occupancy x CUs.
static int resident_grid(void (*kernel)(ScoreArgs), int threads) {
    int dev = 0, cus = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, (const void *)kernel, threads, 0) !=
            hipSuccess)
        return 0;
    return cus * (per > 1 ? per : 1);
}

hipError_t launch_score32(const ScoreArgs &a, int variant, hipStream_t s) {
    if (a.n_work <= 0) return hipSuccess;
    if (variant < 0 || variant >= kNumVariants) return hipErrorInvalidValue;
    if (kVariants[variant].persist == 2 && !a.counter) return hipErrorInvalidValue;
    // A launch holds fewer than 2^32 threads (2^26 waves); longer work lists
    // are cut into chunks (the kernels address partials by glowworm and
    // block, so a chunk only needs the offset work pointer).
    // a multiple of every variant's waves per workgroup; tests/test_gpu_configs.py
    // checks the glowworms on both sides of this cut
    constexpr int kChunk = (1 << 26) - 4;
    const int wpb = kVariants[variant].waves;
    const int kc = mode_chunks(a.nmr);
    const int ki = modes_by_component(a.nmr) ? 3 : (kc == 8 ? 0 : (kc == 16 ? 1 : 2));
    void (*kernel)(ScoreArgs) = kVariants[variant].kernel[ki];
    const int persist = kVariants[variant].persist;
    const int resident = persist ? resident_grid(kernel, 64 * wpb) : 0;
    if (persist && resident <= 0) return hipErrorInvalidConfiguration;
    for (int base = 0; base < a.n_work; base += kChunk) {
        ScoreArgs part = a;
        part.work = a.work + 2 * (size_t)base;
        part.n_work = min(kChunk, a.n_work - base);
        int grid = (part.n_work + wpb - 1) / wpb;
        if (persist) {
            grid = min(grid, resident);
            if (persist == 2) {
                hipLaunchKernelGGL(zero_counter_kernel, dim3(1), dim3(64), 0, s, a.counter);
                const hipError_t e = hipGetLastError();
                if (e != hipSuccess) return e;
            }
        }
        hipLaunchKernelGGL(kernel, dim3(grid), dim3(64 * wpb), 0, s,
                           part);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
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
