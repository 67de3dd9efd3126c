// Neighbour selection and movement: one workgroup per swarm, one lane per
// glowworm (prepare_gso_tasks.cc:584-814; the reference GPU form is
// cal_move_neighbors + move_position_v2, prepare_gso_task.cu:791-1136).
//
// LDS holds the swarm's luciferin and reference points (broadcast reads in
// the O(G) neighbour scan), 32 B per glowworm, so several swarms share a CU.
// The neighbour list is never stored: pass 1 counts the neighbours and sums
// the luciferin differences in index order, pass 2 walks the roulette with
// the same per-neighbour probabilities (prepare_gso_tasks.cc:1424-1488).
// Each lane then copies its chosen neighbour's pre-move pose to the run's
// [P][N] array a.mp (consecutive lanes, consecutive addresses), and after a
// barrier every lane updates its own pose in place from that copy.
#include "dev_common.h"

namespace sd {

__device__ void slerp64(double *self, double *other, double t) {
    // prepare_gso_tasks.cc:1370-1421
    double sn = 0, on = 0;
    for (int i = 0; i < 4; i++) {
        sn += self[i] * self[i];
        on += other[i] * other[i];
    }
    sn = sqrt(sn);
    on = sqrt(on);
    for (int i = 0; i < 4; i++) {
        self[i] /= sn;
        other[i] /= on;
    }
    double dot = 0;
    for (int i = 0; i < 4; i++) dot += self[i] * other[i];
    if (dot < 0) {
        dot = dot * -1;
        for (int i = 0; i < 4; i++) self[i] = -self[i];
    }
    if (dot > 0.9995) {
        for (int i = 0; i < 4; i++) self[i] = self[i] + t * (other[i] - self[i]);
        double rn = 0;
        for (int i = 0; i < 4; i++) rn += self[i] * self[i];
        rn = sqrt(rn);
        for (int i = 0; i < 4; i++) self[i] /= rn;
    } else {
        dot = (dot < 1) ? dot : 1;
        dot = (dot > -1) ? dot : -1;
        const double omega = acos(dot);
        const double so = sin(omega);
        const double c0 = sin((1.0 - t) * omega) / so;
        const double c1 = sin(t * omega) / so;
        for (int i = 0; i < 4; i++) self[i] = c0 * self[i] + c1 * other[i];
    }
}

// other's value i at other[i * stride]
__device__ void move_extent64(double *self, const double *other, size_t stride, int n,
                              double step) {
    // prepare_gso_tasks.cc:1352-1368
    double nn = 0;
    for (int i = 0; i < n; i++) {
        const double d = other[i * stride] - self[i];
        nn += d * d;
    }
    nn = sqrt(nn);
    if (nn >= 1e-8) {
        for (int i = 0; i < n; i++) {
            double d = other[i * stride] - self[i];
            d *= step / nn;
            self[i] += d;
        }
    }
}

__global__ __launch_bounds__(1024) void select_move_kernel(SelectMoveArgs a) {
    extern __shared__ double lds[];
    const int s = blockIdx.x;
    const int g = threadIdx.x;
    const int G = a.G, P = a.P;
    double *lu = lds;               // [G]
    double *rf = lu + G;            // [G][3]
    const size_t base = (size_t)s * G;
    const size_t N = (size_t)a.n_swarms * G;
    if (g < G) {
        lu[g] = a.luciferin[base + g];
        rf[3 * g + 0] = a.ref[(base + g) * 3 + 0];
        rf[3 * g + 1] = a.ref[(base + g) * 3 + 1];
        rf[3 * g + 2] = a.ref[(base + g) * 3 + 2];
    }
    __syncthreads();

    int sel = g;
    int nn = 0;
    if (g < G) {
        if (a.do_select) {
            const double lg = lu[g];
            const double vis = a.vision[base + g];
            const double r2 = vis * vis;
            const double gx = rf[3 * g], gy = rf[3 * g + 1], gz = rf[3 * g + 2];
            // pass 1: neighbour count and sum of luciferin differences, index order
            double total = 0;
            for (int n = 0; n < G; n++) {
                const double ln = lu[n];
                if (n != g && lg < ln) {
                    double dist = 0;
                    const double d0 = gx - rf[3 * n], d1 = gy - rf[3 * n + 1],
                                 d2 = gz - rf[3 * n + 2];
                    dist += d0 * d0;
                    dist += d1 * d1;
                    dist += d2 * d2;
                    if (dist < r2) {
                        total += ln - lg;
                        nn++;
                    }
                }
            }
            // pass 2: roulette over p_k = diff_k / total; k clamped to [1, nn]
            if (nn > 0) {
                const int k = a.d_k ? *a.d_k : a.k;
                const double u = a.stream[(size_t)s * a.n_steps * G + (size_t)k * G + g];
                double sum = 0.0;
                int first = -1, last = -1, chosen = -1;
                for (int n = 0; n < G; n++) {
                    const double ln = lu[n];
                    if (n != g && lg < ln) {
                        double dist = 0;
                        const double d0 = gx - rf[3 * n], d1 = gy - rf[3 * n + 1],
                                     d2 = gz - rf[3 * n + 2];
                        dist += d0 * d0;
                        dist += d1 * d1;
                        dist += d2 * d2;
                        if (dist < r2) {
                            if (first < 0) first = n;
                            last = n;
                            if (chosen < 0 && sum < u) {
                                sum += (ln - lg) / total;
                                if (!(sum < u)) chosen = n;
                            }
                        }
                    }
                }
                sel = !(0.0 < u) ? first : (chosen >= 0 ? chosen : last);
            }
            a.nnei[base + g] = nn;
            a.select[base + g] = sel;
        } else {
            sel = a.select[base + g];
            nn = a.nnei[base + g];
        }
    }
    if (!a.do_move) return;
    // copy the chosen neighbour's pre-move pose; nobody writes pos before the
    // barrier, and each lane reads back only its own copy
    if (g < G) {
        const double *src = a.pos + (base + sel) * P;
        for (int p = 0; p < P; p++) a.mp[(size_t)p * N + base + g] = src[p];
    }
    __syncthreads();
    if (g >= G) return;
    double mp[7];
    for (int p = 0; p < 7; p++) mp[p] = a.mp[(size_t)p * N + base + g];
    const double *mpx = a.mp + 7 * N + base + g;  // extent i at mpx[i * N]
    double *cur = a.pos + (base + g) * P;
    const int moved = (sel != g);
    a.moved[base + g] = moved;
    if (moved) {
        // prepare_gso_tasks.cc:758-794
        double delta[3];
        double n = 0;
        for (int d = 0; d < 3; d++) {
            delta[d] = mp[d] - cur[d];
            n += delta[d] * delta[d];
        }
        n = sqrt(n);
        if (n >= 1e-8) {
            for (int d = 0; d < 3; d++) {
                delta[d] *= (a.step_t / n);
                cur[d] += delta[d];
            }
        }
        double q[4], o[4];
        for (int i = 0; i < 4; i++) {
            q[i] = cur[3 + i];
            o[i] = mp[3 + i];
        }
        slerp64(q, o, a.step_r);
        for (int i = 0; i < 4; i++) cur[3 + i] = q[i];
        move_extent64(cur + 7, mpx, N, a.nmr, a.step_nm);
        move_extent64(cur + 7 + a.nmr, mpx + (size_t)a.nmr * N, N, a.nml, a.step_nm);
    }
    // prepare_gso_tasks.cc:806-808
    double v = a.vision[base + g] + a.beta * (a.max_neighbors - nn);
    if (v < 0.0) v = 0;
    if (v > a.max_vision) v = a.max_vision;
    a.vision[base + g] = v;
}

hipError_t launch_select_move(const SelectMoveArgs &a, hipStream_t s) {
    if (a.n_swarms <= 0) return hipSuccess;
    if (a.G < 1 || a.G > 1024 || (a.do_move && !a.mp)) return hipErrorInvalidValue;
    const int threads = ((a.G + 63) / 64) * 64;
    const size_t lds = sizeof(double) * 4 * (size_t)a.G;  // at most 32 KB
    hipLaunchKernelGGL(select_move_kernel, dim3(a.n_swarms), dim3(threads), lds, s, a);
    return hipGetLastError();
}

}  // namespace sd
