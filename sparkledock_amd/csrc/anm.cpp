// Normal modes of a structure by the anisotropic network model
// (source/dock/src/cal_anm.cc:61-219, build_hessian1D 307-359, calModes
// 403-514): the Calpha Hessian (15 A cutoff, unit spring constant), its
// lowest non-zero eigenpairs on the GPU (rocSOLVER syevd, of which the first
// n_modes + 6 are kept, as the reference's index range), each residue's vector given to all
// its atoms, every mode normalised over the 3N components and scaled so that
// random extents drawn as in the reference give the requested RMSD.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <cmath>
#include <cstring>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "sd_internal.h"

namespace sd {
// Structure access (structure.cpp).
int structure_atoms(sdgso_structure s);
const char *structure_atom_name(sdgso_structure s, int i);
void structure_residues(sdgso_structure s, std::vector<int> &first, std::vector<int> &count);
}  // namespace sd

using namespace sd;

namespace {

void hip_ok(hipError_t e, const char *what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}
void rb_ok(rocblas_status s, const char *what) {
    if (s != rocblas_status_success)
        throw std::runtime_error(std::string(what) + " failed (rocBLAS status " + std::to_string((int)s) + ")");
}

// The Hessian of the elastic network of n nodes (build_hessian1D): for every
// pair within the cutoff, the 3 x 3 super-element -g d d^T / |d|^2 off the
// diagonal and subtracted from both diagonal blocks.
void build_hessian(const std::vector<double> &c, int n, double cutoff, double g, std::vector<double> &H) {
    const int dof = 3 * n;
    H.assign((size_t)dof * dof, 0.0);
    const double cut2 = cutoff * cutoff;
    for (int i = 0; i < n; i++)
        for (int j = i + 1; j < n; j++) {
            const double d[3] = {c[3 * j] - c[3 * i], c[3 * j + 1] - c[3 * i + 1], c[3 * j + 2] - c[3 * i + 2]};
            double d2 = 0.0;
            for (int k = 0; k < 3; k++) d2 += d[k] * d[k];
            if (d2 > cut2) continue;
            double se[3][3];
            for (int k = 0; k < 3; k++)
                for (int l = 0; l < 3; l++) se[k][l] = d[k] * d[l] * (-g / d2);
            for (int k = 0; k < 3; k++)
                for (int l = 0; l < 3; l++) {
                    H[(size_t)(3 * i + k) * dof + (3 * j + l)] = se[k][l];
                    H[(size_t)(3 * j + l) * dof + (3 * i + k)] = se[k][l];
                    H[(size_t)(3 * i + k) * dof + (3 * i + l)] -= se[k][l];
                    H[(size_t)(3 * j + l) * dof + (3 * j + k)] -= se[k][l];
                }
        }
}

// The lowest `count` eigenpairs of the symmetric matrix H (dof x dof), in
// increasing order: values[count], vectors column-major [count][dof].
void lowest_eigenpairs(std::vector<double> &H, int dof, int count, std::vector<double> &values,
                       std::vector<double> &vectors) {
    rocblas_handle h = nullptr;
    rb_ok(rocblas_create_handle(&h), "rocblas_create_handle");
    double *dA = nullptr, *dW = nullptr, *dZ = nullptr;
    rocblas_int *dinfo = nullptr;
    auto cleanup = [&] {
        if (dA) (void)hipFree(dA);
        if (dW) (void)hipFree(dW);
        if (dZ) (void)hipFree(dZ);
        if (dinfo) (void)hipFree(dinfo);
        rocblas_destroy_handle(h);
    };
    try {
        const size_t nn = (size_t)dof * dof;
        hip_ok(hipMalloc(&dA, nn * sizeof(double)), "hipMalloc");
        hip_ok(hipMalloc(&dW, (size_t)dof * sizeof(double)), "hipMalloc");
        hip_ok(hipMalloc(&dZ, (size_t)dof * sizeof(double)), "hipMalloc");  // tridiagonal workspace
        hip_ok(hipMalloc(&dinfo, sizeof(rocblas_int)), "hipMalloc");
        hip_ok(hipMemcpy(dA, H.data(), nn * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy");
        // all eigenpairs (divide and conquer); the eigenvectors overwrite A
        rb_ok(rocsolver_dsyevd(h, rocblas_evect_original, rocblas_fill_upper, dof, dA, dof, dW, dZ,
                               dinfo),
              "rocsolver_dsyevd");
        hip_ok(hipDeviceSynchronize(), "rocsolver_dsyevd");
        rocblas_int info = 0;
        hip_ok(hipMemcpy(&info, dinfo, sizeof(info), hipMemcpyDeviceToHost), "hipMemcpy");
        if (info != 0)
            throw std::runtime_error("eigensolver did not converge (info " + std::to_string(info) + ")");
        values.resize(count);
        vectors.resize((size_t)count * dof);
        hip_ok(hipMemcpy(values.data(), dW, (size_t)count * sizeof(double), hipMemcpyDeviceToHost), "hipMemcpy");
        hip_ok(hipMemcpy(vectors.data(), dA, (size_t)count * dof * sizeof(double), hipMemcpyDeviceToHost),
               "hipMemcpy");
    } catch (...) {
        cleanup();
        throw;
    }
    cleanup();
}

}  // namespace

int sdgso_structure_anm(sdgso_structure s, double *xyz, int n_modes, double rmsd, int n_draws,
                        double *modes, double *eigenvalues) {
    try {
        if (!s || !xyz || !modes || n_modes <= 0 || n_draws <= 0 || !(rmsd > 0))
            throw std::runtime_error("bad arguments");
        const int N = structure_atoms(s);
        // re-centre on the mean of all atoms (cal_anm.cc:63, norm 516-540)
        double mean[3] = {0, 0, 0};
        for (int i = 0; i < N; i++)
            for (int k = 0; k < 3; k++) mean[k] += xyz[3 * (size_t)i + k];
        for (int k = 0; k < 3; k++) mean[k] = mean[k] / N;
        for (int i = 0; i < N; i++)
            for (int k = 0; k < 3; k++) xyz[3 * (size_t)i + k] -= mean[k];
        // network nodes: CA atoms, else C4 (cal_anm.cc:64-78)
        std::vector<int> node;
        for (const char *want : {"CA", "C4"}) {
            for (int i = 0; i < N; i++)
                if (std::strcmp(structure_atom_name(s, i), want) == 0) node.push_back(i);
            if (!node.empty()) break;
        }
        if (node.empty()) throw std::runtime_error("no CA or C4 atoms for the elastic network");
        const int n = (int)node.size(), dof = 3 * n;
        if (n_modes + 6 > dof) throw std::runtime_error("more modes than the network has");
        // every residue maps its atoms to its own node (cal_anm.cc:104-127)
        std::vector<int> rfirst, rcount;
        structure_residues(s, rfirst, rcount);
        if ((int)rfirst.size() != n)
            throw std::runtime_error("the elastic network needs one node atom per residue (" +
                                     std::to_string(rfirst.size()) + " residues, " +
                                     std::to_string(n) + " nodes)");
        std::vector<int> node_of(N, -1);
        for (int r = 0; r < n; r++) {
            if (node[r] < rfirst[r] || node[r] >= rfirst[r] + rcount[r])
                throw std::runtime_error("residue " + std::to_string(r) + " has no node atom");
            for (int i = rfirst[r]; i < rfirst[r] + rcount[r]; i++) node_of[i] = r;
        }
        std::vector<double> c(3 * (size_t)n), H;
        for (int r = 0; r < n; r++)
            for (int k = 0; k < 3; k++) c[3 * r + k] = xyz[3 * (size_t)node[r] + k];
        build_hessian(c, n, 15.0, 1.0, H);
        // the first n_modes + 6 eigenpairs; zero modes (< 1e-6) are skipped (calModes)
        std::vector<double> val, vec;
        lowest_eigenpairs(H, dof, n_modes + 6, val, vec);
        int nz = 0;
        for (double v : val)
            if (v < 1e-6) nz++;
        if (nz + n_modes > n_modes + 6)
            throw std::runtime_error("more zero modes than expected (" + std::to_string(nz) + ")");
        std::vector<double> ev(n_modes);
        std::vector<double> M((size_t)n_modes * N * 3);
        for (int m = 0; m < n_modes; m++) {
            ev[m] = val[nz + m];
            const double *z = vec.data() + (size_t)(nz + m) * dof;
            // a deterministic sign: the largest component (first of equals) positive
            int imax = 0;
            for (int i = 1; i < dof; i++)
                if (std::fabs(z[i]) > std::fabs(z[imax])) imax = i;
            const double sg = z[imax] < 0 ? -1.0 : 1.0;
            double *out = M.data() + (size_t)m * N * 3;
            for (int i = 0; i < N; i++)
                for (int k = 0; k < 3; k++) out[3 * (size_t)i + k] = sg * z[3 * node_of[i] + k];
            double ss = 0.0;
            for (size_t t = 0; t < (size_t)N * 3; t++) ss += std::pow(out[t], 2);
            ss = std::sqrt(ss);
            for (size_t t = 0; t < (size_t)N * 3; t++) out[t] /= ss;
        }
        // scale: random extents N(0, 1) of the reference's generator give the RMSD
        // (cal_anm.cc:167-199)
        std::default_random_engine gen;
        gen.seed(324324u);
        std::normal_distribution<double> dist(0.0, 1.0);
        double coef = 0.0;
        for (int i = 0; i < n_draws; i++) {
            double sum = 0.0;
            for (int j = 0; j < n_modes; j++) {
                const double v = dist(gen);
                const double var = 1.0 / ev[j];
                sum += v * v * var;
            }
            coef += std::sqrt(sum);
        }
        coef = coef / n_draws;
        const double scale_item = std::sqrt((double)N) * rmsd / coef;
        for (int m = 0; m < n_modes; m++) {
            const double sc = scale_item * std::sqrt(1.0 / ev[m]);
            for (size_t t = 0; t < (size_t)N * 3; t++) M[(size_t)m * N * 3 + t] *= sc;
        }
        std::memcpy(modes, M.data(), M.size() * sizeof(double));
        if (eigenvalues) std::memcpy(eigenvalues, ev.data(), n_modes * sizeof(double));
        return SDGSO_OK;
    } catch (const std::bad_alloc &) {
        set_error("host allocation failed");
        return SDGSO_ENOMEM;
    } catch (const std::exception &e) {
        set_error(e.what());
        return SDGSO_EINVAL;
    }
}
