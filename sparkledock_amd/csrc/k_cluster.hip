// Final coordinates and clustering of docked poses (the output side of the
// GSO run).
//
// coords_kernel: chosen receptor and ligand atoms of a range of glowworms in
// FP64, in the order of the reference's final-pose code (receptor modes, then
// ligand modes, rotation and translation; prepare_gso_tasks.cc:882-920), one
// thread per (glowworm, atom).
//
// rmsd_kernel: for every swarm of a chunk, the G x G matrix of backbone RMSDs
// between its glowworms, each entry summed over atoms in order as cal_rmsd
// does (prepare_gso_tasks.cc:1258-1271): dx^2, dy^2, dz^2 per atom, then
// divided by the atom count and square-rooted.
#include "dev_common.h"

namespace sd {

__global__ __launch_bounds__(256) void coords_kernel(CoordsArgs a) {
    const int atom = blockIdx.x * blockDim.x + threadIdx.x;
    const int gl = blockIdx.y;  // glowworm within the range
    const int n = a.n_rec + a.n_lig;
    if (atom >= n) return;
    const DevComplex &c = *a.cx;
    const double *pose = a.pos + (size_t)(a.first + gl) * a.P;
    double x, y, z;
    if (atom < a.n_rec) {
        receptor_atom64(c, pose + 7, a.rec_idx ? a.rec_idx[atom] : atom, x, y, z);
    } else {
        double inv[4];
        quat_inverse(pose + 3, inv);
        const int j = atom - a.n_rec;
        ligand_atom64(c, pose, inv, a.lig_idx ? a.lig_idx[j] : j, x, y, z);
    }
    double *o = a.out + ((size_t)gl * n + atom) * 3;
    o[0] = x;
    o[1] = y;
    o[2] = z;
}

__global__ __launch_bounds__(256) void rmsd_kernel(const double *bb, int nb, int G, int n_swarms,
                                                   double *rmsd) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t per = (size_t)G * G;
    if (t >= per * n_swarms) return;
    const size_t s = t / per;
    const int g = (int)((t % per) / G), h = (int)(t % G);
    const double *A = bb + (s * G + g) * (size_t)nb * 3;
    const double *B = bb + (s * G + h) * (size_t)nb * 3;
    double sum = 0.0;
    for (int i = 0; i < nb; i++) {
        const double dx = A[3 * i] - B[3 * i];
        const double dy = A[3 * i + 1] - B[3 * i + 1];
        const double dz = A[3 * i + 2] - B[3 * i + 2];
        sum += dx * dx;
        sum += dy * dy;
        sum += dz * dz;
    }
    sum /= nb;
    rmsd[t] = sqrt(sum);
}

hipError_t launch_coords(const CoordsArgs &a, int count, hipStream_t s) {
    const int n = a.n_rec + a.n_lig;
    if (count <= 0 || n <= 0) return hipSuccess;
    hipLaunchKernelGGL(coords_kernel, dim3((n + 255) / 256, count), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_rmsd(const double *bb, int nb, int G, int n_swarms, double *rmsd, hipStream_t s) {
    const size_t total = (size_t)G * G * n_swarms;
    if (total == 0) return hipSuccess;
    hipLaunchKernelGGL(rmsd_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, bb, nb,
                       G, n_swarms, rmsd);
    return hipGetLastError();
}

}  // namespace sd
