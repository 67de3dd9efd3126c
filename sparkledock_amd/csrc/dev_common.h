// Device helpers that restate the reference's FP64 pose arithmetic in its
// operation order, and the FP32 form's grid-position helpers (rotation matrix,
// FP32 mode sums) that are close to it within stated bounds.  Every
// translation unit that includes this file is built with -ffp-contract=off,
// so no multiply-add is fused unless it is written as an fma.
#pragma once

#include <hip/hip_runtime.h>

#include "sd_internal.h"

namespace sd {

// Load through the global address space.  Pointers that come out of a
// descriptor in memory are generic to the compiler, which then emits flat
// loads (counted on both the vector-memory and LDS counters); every array the
// kernels read lives in HBM.
template <class T>
__device__ __forceinline__ T gld(const T *p, size_t i) {
    return ((const __attribute__((address_space(1))) T *)p)[i];
}

__device__ __forceinline__ float4 gld(const float4 *p, size_t i) {
    typedef float f4 __attribute__((ext_vector_type(4)));
    const f4 v = ((const __attribute__((address_space(1))) f4 *)p)[i];
    return make_float4(v.x, v.y, v.z, v.w);
}

// Two consecutive doubles p[2i], p[2i+1] in one 16-byte load.
__device__ __forceinline__ double2 gld2(const double *p, size_t i) {
    typedef double d2 __attribute__((ext_vector_type(2)));
    const d2 v = ((const __attribute__((address_space(1))) d2 *)p)[i];
    return make_double2(v.x, v.y);
}

// inverse = conj(q) / |q|^2 (prepare_gso_tasks.cc:288-302)
__device__ __forceinline__ void quat_inverse(const double q[4], double inv[4]) {
    const double n2 = q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
    inv[0] = q[0] / n2;
    inv[1] = -q[1] / n2;
    inv[2] = -q[2] / n2;
    inv[3] = -q[3] / n2;
}

// q * (0, p) * inv, keeping the zero terms and the left-to-right order of
// prepare_gso_tasks.cc:304-329.
__device__ __forceinline__ void rotate_point(const double q[4], const double inv[4], double &x,
                                             double &y, double &z) {
    const double a0 = 0.0, a1 = x, a2 = y, a3 = z;
    const double b0 = q[0] * a0 - q[1] * a1 - q[2] * a2 - q[3] * a3;
    const double b1 = q[0] * a1 + q[1] * a0 + q[2] * a3 - q[3] * a2;
    const double b2 = q[0] * a2 - q[1] * a3 + q[2] * a0 + q[3] * a1;
    const double b3 = q[0] * a3 + q[1] * a2 - q[2] * a1 + q[3] * a0;
    x = b0 * inv[1] + b1 * inv[0] + b2 * inv[3] - b3 * inv[2];
    y = b0 * inv[2] - b1 * inv[3] + b2 * inv[0] + b3 * inv[1];
    z = b0 * inv[3] + b1 * inv[2] - b2 * inv[1] + b3 * inv[0];
}

// Rotation matrix of p -> q (0, p) q^-1 for a quaternion of any norm.  The
// FP32 kernels rotate with it only the positions that they round to their
// 2^-20 A grid, where its last-bit differences from rotate_point's order are
// far below a grid unit; every pair decided in FP64 goes through
// rotate_point.
struct Rot64 {
    double m[9];
};

__device__ __forceinline__ Rot64 rot_of(const double q[4]) {
    const double w = q[0], x = q[1], y = q[2], z = q[3];
    const double ww = w * w, xx = x * x, yy = y * y, zz = z * z;
    const double s = 1.0 / (ww + xx + yy + zz), s2 = 2.0 * s;
    Rot64 r;
    r.m[0] = (ww + xx - yy - zz) * s;
    r.m[1] = (x * y - w * z) * s2;
    r.m[2] = (x * z + w * y) * s2;
    r.m[3] = (x * y + w * z) * s2;
    r.m[4] = (ww - xx + yy - zz) * s;
    r.m[5] = (y * z - w * x) * s2;
    r.m[6] = (x * z - w * y) * s2;
    r.m[7] = (y * z + w * x) * s2;
    r.m[8] = (ww - xx - yy + zz) * s;
    return r;
}

__device__ __forceinline__ void rot_apply(const Rot64 &r, double &x, double &y, double &z) {
    const double a = x, b = y, c = z;
    x = fma(r.m[0], a, fma(r.m[1], b, r.m[2] * c));
    y = fma(r.m[3], a, fma(r.m[4], b, r.m[5] * c));
    z = fma(r.m[6], a, fma(r.m[7], b, r.m[8] * c));
}

// dist_to_bins[d] - 1 (interface.h:30-35) for d in [0, 29]:
// 0,0,0,1,...,13 (d = 3..15), then pairs 13,14,14,...,19,19, then 20 at d = 29.
__device__ __forceinline__ int bin_of_d(int d) { return max(0, min(d - 2, (d + 11) >> 1)); }

// x/y/z += M[m][.][i] * e_m for m = 0..nm-1 in mode order (FP64, product
// rounded, then added: prepare_gso_tasks.cc:465-504).  Modes are SoA
// [nm][3][n]; the loads of four modes are issued before their adds so that
// their latencies overlap, while the additions keep the reference order.
__device__ __forceinline__ void modes_add64(const double *modes, int n, int i, const double *ext,
                                            int nm, double &x, double &y, double &z) {
    int m = 0;
    for (; m + 4 <= nm; m += 4) {
        double mx[4], my[4], mz[4], e[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const double *M = modes + (size_t)(m + k) * 3 * n;
            mx[k] = gld(M, i);
            my[k] = gld(M, n + i);
            mz[k] = gld(M, 2 * n + i);
            e[k] = ext[m + k];
        }
#pragma unroll
        for (int k = 0; k < 4; k++) {
            x += mx[k] * e[k];
            y += my[k] * e[k];
            z += mz[k] * e[k];
        }
    }
    for (; m < nm; m++) {
        const double *M = modes + (size_t)m * 3 * n;
        const double e = ext[m];
        x += gld(M, i) * e;
        y += gld(M, n + i) * e;
        z += gld(M, 2 * n + i) * e;
    }
}

// Receptor atom i displaced by the normal modes, FP64, reference order
// (prepare_gso_tasks.cc:465-482; masked atoms have zero mode rows).
__device__ __forceinline__ void receptor_atom64(const DevComplex &c, const double *ext, int i,
                                                double &x, double &y, double &z) {
    x = gld(c.r0, i);
    y = gld(c.r0, c.nr + i);
    z = gld(c.r0, 2 * c.nr + i);
    modes_add64(c.rmodes, c.nr, i, ext, c.nmr, x, y, z);
}

// Ligand atom j of a pose: modes, rotation, translation
// (prepare_gso_tasks.cc:487-526).
__device__ __forceinline__ void ligand_atom64(const DevComplex &c, const double *pose,
                                              const double inv[4], int j, double &x, double &y,
                                              double &z) {
    x = gld(c.l0, j);
    y = gld(c.l0, c.nl + j);
    z = gld(c.l0, 2 * c.nl + j);
    modes_add64(c.lmodes, c.nl, j, pose + 7 + c.nmr, c.nml, x, y, z);
    rotate_point(pose + 3, inv, x, y, z);
    x += pose[0];
    y += pose[1];
    z += pose[2];
}

// The same for slot `lane` of ligand block b, read from the block-slot copy
// (every load of a wave is contiguous), for a position that the FP32 kernels
// round to their grid: the mode displacement, FP32, is added to the FP64
// coordinates, which the pose's matrix R then rotates.  The displacement is
// d (have_d: this slot's values from the run's ligand displacement GEMM) or
// formed here as an FP32 FMA chain in mode order over the FP32 modes and
// extents.  Either is within (nml + 2) u S of the exact one, u = 2^-24 and S
// the pose's ligand displacement bound.
__device__ __forceinline__ void ligand_slot_grid(const DevComplex &c, const double *pose,
                                                 const Rot64 &R, int b, int lane, bool have_d,
                                                 float3 d, double &x, double &y, double &z) {
    const int nm = have_d ? 0 : c.nml;
    const double *B = c.lblk + (size_t)b * 3 * kBlock + lane;
    const float *BM = c.lblkm + (size_t)b * 3 * nm * kBlock + lane;
    x = gld(B, 0);
    y = gld(B, kBlock);
    z = gld(B, 2 * kBlock);
    const double *ext = pose + 7 + c.nmr;
    float dx = d.x, dy = d.y, dz = d.z;
    int m = 0;
    for (; m + 4 <= nm; m += 4) {
        float mx[4], my[4], mz[4], e[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const float *M = BM + (size_t)(3 * (m + k)) * kBlock;
            mx[k] = gld(M, 0);
            my[k] = gld(M, kBlock);
            mz[k] = gld(M, 2 * kBlock);
            e[k] = (float)ext[m + k];
        }
#pragma unroll
        for (int k = 0; k < 4; k++) {
            dx = __builtin_fmaf(mx[k], e[k], dx);
            dy = __builtin_fmaf(my[k], e[k], dy);
            dz = __builtin_fmaf(mz[k], e[k], dz);
        }
    }
    for (; m < nm; m++) {
        const float *M = BM + (size_t)(3 * m) * kBlock;
        const float e = (float)ext[m];
        dx = __builtin_fmaf(gld(M, 0), e, dx);
        dy = __builtin_fmaf(gld(M, kBlock), e, dy);
        dz = __builtin_fmaf(gld(M, 2 * kBlock), e, dz);
    }
    x += (double)dx;
    y += (double)dy;
    z += (double)dz;
    rot_apply(R, x, y, z);
    x += pose[0];
    y += pose[1];
    z += pose[2];
}

// Ligand reference point of a pose (rotation + translation, no modes).
__device__ __forceinline__ void ref_point64(const DevComplex &c, const double *pose,
                                            const double inv[4], double out[3]) {
    double x = c.lref[0], y = c.lref[1], z = c.lref[2];
    rotate_point(pose + 3, inv, x, y, z);
    out[0] = x + pose[0];
    out[1] = y + pose[1];
    out[2] = z + pose[2];
}

// Bin argument d of one pair in FP64, (int)(sqrt(d2) * 2 - 1)
// (prepare_gso_tasks.cc:1542-1556); false when the pair is outside (0, 225].
__device__ __forceinline__ bool pair_bin64(double rx, double ry, double rz, double lx, double ly,
                                           double lz, int &d) {
    const double dx = rx - lx;
    const double dy = ry - ly;
    const double dz = rz - lz;
    const double d2 = dx * dx + dy * dy + dz * dz;
    if (d2 <= kCutoff2 && d2 > 0.0) {
        d = (int)(sqrt(d2) * 2.0 - 1.0);
        return true;
    }
    return false;
}

// DFIRE contribution of one pair in FP64 (prepare_gso_tasks.cc:1542-1566);
// returns false when the pair is outside (0, 225].  The table is the FP64
// one (FP64 form) or its FP32 copy (the FP32 form's guard-band pairs, whose
// other terms come from the same copy).
template <class TableT>
__device__ __forceinline__ bool pair_term64(double rx, double ry, double rz, double lx, double ly,
                                            double lz, int ra, int lb, const TableT *table,
                                            double &val) {
    int d;
    if (pair_bin64(rx, ry, rz, lx, ly, lz, d)) {
        val = (double)gld(table, ra + lb + bin_of_d(d));
        return true;
    }
    return false;
}

__device__ __forceinline__ double score_of_energy(double e) { return (e * 0.0157 - 4.7) * -1; }

}  // namespace sd
