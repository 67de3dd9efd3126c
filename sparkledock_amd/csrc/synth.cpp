// Seeded synthetic inputs (host only): complexes, DFIRE-like table, initial
// poses.  Everything derives from std::mt19937_64 with self-written uniform
// and normal transforms, so a seed gives the same inputs on every platform
// with IEEE double arithmetic.
#include <cmath>
#include <cstdint>
#include <random>
#include <vector>

#include "../../include/sdgso.h"

namespace {

struct Rng {
    std::mt19937_64 g;
    explicit Rng(uint64_t seed) : g(seed) {}
    double uniform() { return (double)(g() >> 11) * (1.0 / 9007199254740992.0); }  // [0,1)
    double normal() {
        // Box-Muller; u1 in (0,1]
        const double u1 = 1.0 - uniform();
        const double u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
    int below(int n) { return (int)(uniform() * n); }
};

constexpr double kPi = 3.14159265358979323846;

}  // namespace

extern "C" {

int sdgso_synth_molecule(uint64_t seed, int n_atoms, int n_modes, double *xyz, int *type,
                         double *modes) {
    if (n_atoms <= 0 || n_modes < 0) return SDGSO_EINVAL;
    Rng rng(seed);
    const double radius = std::cbrt(n_atoms * 11.4);
    const double step = 3.8;
    std::vector<double> p(3 * (size_t)n_atoms);
    double cur[3] = {0, 0, 0};
    for (int i = 0; i < n_atoms; i++) {
        double nxt[3] = {cur[0], cur[1], cur[2]};
        bool ok = false;
        for (int tr = 0; tr < 64 && !ok; tr++) {
            double d[3] = {rng.normal(), rng.normal(), rng.normal()};
            const double n = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
            for (int k = 0; k < 3; k++) nxt[k] = cur[k] + d[k] * step / n;
            ok = nxt[0] * nxt[0] + nxt[1] * nxt[1] + nxt[2] * nxt[2] <= radius * radius;
        }
        if (!ok) {  // step back toward the centre
            const double n = std::sqrt(cur[0] * cur[0] + cur[1] * cur[1] + cur[2] * cur[2]);
            for (int k = 0; k < 3; k++) nxt[k] = cur[k] - cur[k] * step / n;
        }
        for (int k = 0; k < 3; k++) {
            cur[k] = nxt[k];
            p[3 * (size_t)i + k] = cur[k];
        }
    }
    double mean[3] = {0, 0, 0};
    for (int i = 0; i < n_atoms; i++)
        for (int k = 0; k < 3; k++) mean[k] += p[3 * (size_t)i + k];
    for (int k = 0; k < 3; k++) mean[k] /= n_atoms;
    for (int i = 0; i < n_atoms; i++)
        for (int k = 0; k < 3; k++) p[3 * (size_t)i + k] -= mean[k];
    if (xyz)
        for (size_t i = 0; i < p.size(); i++) xyz[i] = p[i];
    std::vector<int> t(n_atoms);
    for (int i = 0; i < n_atoms; i++) t[i] = rng.below(SDGSO_NTYPES);
    if (type)
        for (int i = 0; i < n_atoms; i++) type[i] = t[i];
    if (n_modes > 0) {
        // orthonormal random 3N-vectors: Gaussian draws, modified Gram-Schmidt
        const size_t L = 3 * (size_t)n_atoms;
        std::vector<double> m((size_t)n_modes * L);
        for (int a = 0; a < n_modes; a++) {
            double *v = &m[(size_t)a * L];
            for (size_t i = 0; i < L; i++) v[i] = rng.normal();
            for (int pass = 0; pass < 2; pass++)
                for (int b = 0; b < a; b++) {
                    const double *w = &m[(size_t)b * L];
                    double dot = 0;
                    for (size_t i = 0; i < L; i++) dot += v[i] * w[i];
                    for (size_t i = 0; i < L; i++) v[i] -= dot * w[i];
                }
            double n = 0;
            for (size_t i = 0; i < L; i++) n += v[i] * v[i];
            n = std::sqrt(n);
            for (size_t i = 0; i < L; i++) v[i] /= n;
        }
        if (modes)
            for (size_t i = 0; i < m.size(); i++) modes[i] = m[i];
    }
    return SDGSO_OK;
}

int sdgso_synth_table(uint64_t seed, float *table) {
    if (!table) return SDGSO_EINVAL;
    Rng rng(seed);
    for (int i = 0; i < SDGSO_TABLE_LEN; i++) table[i] = (float)(rng.normal() * 0.5);
    return SDGSO_OK;
}

int sdgso_synth_poses(uint64_t seed, int n_swarms, int n_glowworms, int n_rec_modes,
                      int n_lig_modes, double centre_radius, double swarm_radius, double *pos) {
    if (n_swarms <= 0 || n_glowworms <= 0 || n_rec_modes < 0 || n_lig_modes < 0 || !pos)
        return SDGSO_EINVAL;
    Rng rng(seed);
    const int P = 7 + n_rec_modes + n_lig_modes;
    const double golden = kPi * (3.0 - std::sqrt(5.0));
    for (int s = 0; s < n_swarms; s++) {
        // swarm centres on a Fibonacci sphere around the receptor
        const double z = 1.0 - 2.0 * (s + 0.5) / n_swarms;
        const double rr = std::sqrt(std::max(0.0, 1.0 - z * z));
        const double ph = golden * s;
        const double c[3] = {centre_radius * rr * std::cos(ph), centre_radius * rr * std::sin(ph),
                             centre_radius * z};
        for (int g = 0; g < n_glowworms; g++) {
            double *p = pos + ((size_t)s * n_glowworms + g) * P;
            // translation in the swarm ball by rejection (cal_anm.cc:845-859)
            double x, y, w;
            do {
                x = (2 * rng.uniform() - 1) * swarm_radius;
                y = (2 * rng.uniform() - 1) * swarm_radius;
                w = (2 * rng.uniform() - 1) * swarm_radius;
            } while (x * x + y * y + w * w > swarm_radius * swarm_radius);
            p[0] = c[0] + x;
            p[1] = c[1] + y;
            p[2] = c[2] + w;
            // uniform random unit quaternion (Shoemake; cal_anm.cc:863-870)
            const double u1 = rng.uniform(), u2 = rng.uniform(), u3 = rng.uniform();
            p[3] = std::sqrt(1 - u1) * std::sin(2 * kPi * u2);
            p[4] = std::sqrt(1 - u1) * std::cos(2 * kPi * u2);
            p[5] = std::sqrt(u1) * std::sin(2 * kPi * u3);
            p[6] = std::sqrt(u1) * std::cos(2 * kPi * u3);
            for (int m = 0; m < n_rec_modes + n_lig_modes; m++) p[7 + m] = 2 * rng.uniform() - 1;
        }
    }
    return SDGSO_OK;
}

}  // extern "C"
