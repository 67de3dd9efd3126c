// Starting poses of the swarms around the receptor surface
// (source/dock/src/cal_anm.cc calculate_surface_point 541-773 and
// calculate_start_pos 775-900, source/dock/src/kmeans.cc 126-243): swarm
// centres from k-means++ over the atoms of surface residues, 100 golden-spiral
// points around each centroid, points near another centroid or within 3 A of
// a receptor atom dropped, and a second k-means over what remains; then per
// glowworm a translation drawn uniformly in the swarm sphere, a uniform
// random rotation and normally distributed mode extents.  The random draws
// use the same C++ standard-library engines and distributions as the
// reference, in the same order.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "sd_internal.h"

namespace sd {
int structure_atoms(sdgso_structure s);
void structure_residues(sdgso_structure s, std::vector<int> &first, std::vector<int> &count);
const char *structure_residue_name(sdgso_structure s, int r);
}  // namespace sd

using namespace sd;

namespace {

inline double dist3(const double *a, const double *b) {
    return std::sqrt((a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]) +
                     (a[2] - b[2]) * (a[2] - b[2]));
}

// k-means++ seeding (kmeans.cc:126-187): the first centre uniformly at random,
// every further one by the distance-weighted draw over the points in order.
void kmeans_init(double *c, const double *p, int n, int K, uint32_t seed) {
    std::mt19937 gen(seed);
    std::uniform_int_distribution<int> pick(0, n - 1);
    const int id = pick(gen);
    for (int i = 0; i < K; i++)
        for (int d = 0; d < 3; d++) c[i * 3 + d] = p[id * 3 + d];
    std::vector<double> near((size_t)n);
    for (int k = 1; k < K; k++) {
        double sum = 0;
        for (int q = 0; q < n; q++) {
            double best = std::numeric_limits<double>::max();
            for (int t = 0; t < k; t++) {
                const double d = dist3(p + 3 * (size_t)q, c + 3 * t);
                if (d < best) best = d;
            }
            sum += best;
            near[q] = best;
        }
        std::uniform_real_distribution<float> unit(0.0, 1.0);
        const double prob = unit(gen);
        sum = sum * prob;
        for (int q = 0; q < n; q++) {
            sum = sum - near[q];
            if (sum > 0) continue;
            for (int d = 0; d < 3; d++) c[k * 3 + d] = p[q * 3 + d];
            break;
        }
    }
}

// Lloyd iterations (kmeans.cc:191-243): nearest centre, then centre = mean.
void kmeans_run(int *group, const double *p, double *c, int iters, int n, int K) {
    std::vector<double> acc(3 * (size_t)K);
    std::vector<int> cnt(K);
    for (int it = 0; it < iters; it++) {
        for (int q = 0; q < n; q++) {
            double best = std::numeric_limits<float>::max();
            int id = -1;
            for (int k = 0; k < K; k++) {
                const double d = dist3(c + 3 * k, p + 3 * (size_t)q);
                if (d < best) {
                    best = d;
                    id = k;
                }
            }
            group[q] = id;
        }
        std::fill(acc.begin(), acc.end(), 0.0);
        std::fill(cnt.begin(), cnt.end(), 0);
        for (int q = 0; q < n; q++) {
            const int g = group[q];
            if (g < 0) continue;
            for (int d = 0; d < 3; d++) acc[3 * g + d] += p[3 * (size_t)q + d];
            cnt[g]++;
        }
        for (int k = 0; k < K; k++)
            for (int d = 0; d < 3; d++) c[3 * k + d] = acc[3 * k + d] / (1.0 * cnt[k]);
    }
}

// cal_anm.cc:977-989
void points_on_sphere(double *out, int n) {
    const double pi = 3.14159265358979323846;
    const double inc = pi * (3.0 - std::sqrt(5.0));
    const double off = 2.0 / n;
    for (int i = 0; i < n; i++) {
        const double y = i * off - 1.0 + (off / 2.0);
        const double r = std::sqrt(1 - y * y);
        const double phi = i * inc;
        out[i * 3] = std::cos(phi) * r;
        out[i * 3 + 1] = y;
        out[i * 3 + 2] = std::sin(phi) * r;
    }
}

bool surface_residue(const std::string &r) {
    static const char *kNames[] = {"ARG", "ASN", "ASP", "ASX", "CSO", "GLN", "GLU", "GLX", "GLY",
                                   "HID", "HIE", "HIP", "IS",  "HSD", "HSE", "HSP", "LYS", "PHD",
                                   "PRO", "PTR", "SEP", "SER", "THR", "TPO", "TYR", "HIS"};
    for (const char *k : kNames)
        if (r == k) return true;
    return false;
}

template <class F>
int guard(F &&f) {
    try {
        f();
        return SDGSO_OK;
    } catch (const std::bad_alloc &) {
        set_error("host allocation failed");
        return SDGSO_ENOMEM;
    } catch (const std::exception &e) {
        set_error(e.what());
        return SDGSO_EINVAL;
    }
}

}  // namespace

int sdgso_kmeans(const double *points, int n, int k, uint32_t seed, int iterations, double *centres,
                 int *group) {
    return guard([&] {
        if (!points || !centres || n <= 0 || k <= 0 || iterations < 0)
            throw std::runtime_error("bad arguments");
        kmeans_init(centres, points, n, k, seed);
        std::vector<int> g((size_t)n);
        kmeans_run(group ? group : g.data(), points, centres, iterations, n, k);
    });
}

int sdgso_starting_poses(sdgso_structure receptor, const double *rec_xyz, int n_lig,
                         const double *lig_xyz, const double *rec_translation, int swarms,
                         int glowworms, double swarm_radius, uint32_t start_points_seed,
                         uint32_t anm_seed, int n_modes_rec, int n_modes_lig, double *centres,
                         double *pos) {
    return guard([&] {
        if (!receptor || !rec_xyz || !lig_xyz || !rec_translation || !pos || swarms <= 0 ||
            glowworms <= 0 || n_lig <= 0 || !(swarm_radius > 0) || n_modes_rec < 0 || n_modes_lig < 0)
            throw std::runtime_error("bad arguments");
        const int nr = structure_atoms(receptor);
        // ligand diameter: the largest pair distance (cal_anm.cc:552-572, cal_dist 1043-1058)
        double lig_diam = 0;
        for (int i = 0; i < n_lig; i++)
            for (int j = i + 1; j < n_lig; j++) {
                double d = 0.0;
                for (int c = 0; c < 3; c++) {
                    const double diff = lig_xyz[3 * (size_t)i + c] - lig_xyz[3 * (size_t)j + c];
                    d += diff * diff;
                }
                d = std::sqrt(d);
                if (lig_diam < d) lig_diam = d;
            }
        const double surf_dist = lig_diam < 10.0 * 2 ? 10.0 : lig_diam / 4.0;
        // atoms of surface residues (select_surface, cal_anm.cc:991-1022)
        std::vector<int> rfirst, rcount;
        structure_residues(receptor, rfirst, rcount);
        std::vector<double> surf;
        for (size_t r = 0; r < rfirst.size(); r++)
            if (surface_residue(structure_residue_name(receptor, (int)r)))
                for (int i = rfirst[r]; i < rfirst[r] + rcount[r]; i++)
                    for (int c = 0; c < 3; c++) surf.push_back(rec_xyz[3 * (size_t)i + c]);
        const int ns = (int)(surf.size() / 3);
        if (ns <= swarms)
            throw std::runtime_error("more swarms (" + std::to_string(swarms) +
                                     ") than surface atoms (" + std::to_string(ns) + ")");
        std::vector<double> cent(3 * (size_t)swarms);
        std::vector<int> grp((size_t)ns);
        kmeans_init(cent.data(), surf.data(), ns, swarms, 324324u);
        kmeans_run(grp.data(), surf.data(), cent.data(), 100, ns, swarms);
        // 100 points around every centroid (cal_anm.cc:646-662)
        const int np = 100;
        std::vector<double> samp((size_t)swarms * np * 3);
        for (int i = 0; i < swarms; i++) {
            double *cs = samp.data() + (size_t)i * np * 3;
            points_on_sphere(cs, np);
            for (int j = 0; j < np; j++)
                for (int d = 0; d < 3; d++) cs[j * 3 + d] = cs[j * 3 + d] * surf_dist + cent[i * 3 + d];
        }
        // drop points within the surface distance of a neighbouring centroid (:664-704)
        std::vector<char> drop((size_t)swarms * np, 0);
        for (int i = 0; i < swarms; i++)
            for (int j = 0; j < swarms; j++) {
                if (j == i || !(dist3(&cent[i * 3], &cent[j * 3]) <= 20.0)) continue;
                for (int q = 0; q < np; q++)
                    if (dist3(&samp[((size_t)i * np + q) * 3], &cent[j * 3]) <= surf_dist)
                        drop[(size_t)i * np + q] = 1;
            }
        std::vector<double> kept;
        for (size_t t = 0; t < (size_t)swarms * np; t++)
            if (!drop[t])
                for (int d = 0; d < 3; d++) kept.push_back(samp[t * 3 + d]);
        // drop points within 3 A of a receptor atom (:709-741)
        std::vector<double> outer;
        for (size_t t = 0; t < kept.size() / 3; t++) {
            int count = 0;
            for (int a = 0; a < nr; a++)
                if (dist3(&kept[t * 3], rec_xyz + 3 * (size_t)a) <= 3.0) count++;
            if (count == 0)
                for (int d = 0; d < 3; d++) outer.push_back(kept[t * 3 + d]);
        }
        const int no = (int)(outer.size() / 3);
        if (no <= swarms)
            throw std::runtime_error("too few surface points (" + std::to_string(no) + ") for " +
                                     std::to_string(swarms) + " swarms");
        std::vector<double> sc(3 * (size_t)swarms);
        std::vector<int> og((size_t)no);
        kmeans_init(sc.data(), outer.data(), no, swarms, 324324u);
        kmeans_run(og.data(), outer.data(), sc.data(), 500, no, swarms);
        for (int i = 0; i < swarms; i++)
            for (int d = 0; d < 3; d++) sc[i * 3 + d] += rec_translation[d];
        if (centres) std::memcpy(centres, sc.data(), sc.size() * sizeof(double));
        // poses (calculate_start_pos, cal_anm.cc:800-898)
        const double pi = 3.14159265358979323846;
        const int P = 7 + n_modes_rec + n_modes_lig;
        std::mt19937 gen(start_points_seed);
        std::uniform_real_distribution<double> unit(0.0, 1.0);
        std::default_random_engine nm_gen(anm_seed);
        std::normal_distribution<double> nm(0.0, 1.0);
        const double r2 = swarm_radius * swarm_radius;
        for (int i = 0; i < swarms; i++)
            for (int j = 0; j < glowworms; j++) {
                double *p = pos + ((size_t)i * glowworms + j) * P;
                double g[3];
                for (;;) {
                    const double v0 = unit(gen), v1 = unit(gen), v2 = unit(gen);
                    const double x = (2 * v0 - 1) * swarm_radius;
                    const double y = (2 * v1 - 1) * swarm_radius;
                    const double z = (2 * v2 - 1) * swarm_radius;
                    if (x * x + y * y + z * z <= r2) {
                        g[0] = x;
                        g[1] = y;
                        g[2] = z;
                        break;
                    }
                }
                for (int d = 0; d < 3; d++) p[d] = sc[i * 3 + d] + g[d];
                const double u1 = unit(gen), u2 = unit(gen), u3 = unit(gen);
                p[3] = std::sqrt(1 - u1) * std::sin(2 * pi * u2);
                p[4] = std::sqrt(1 - u1) * std::cos(2 * pi * u2);
                p[5] = std::sqrt(u1) * std::sin(2 * pi * u3);
                p[6] = std::sqrt(u1) * std::cos(2 * pi * u3);
                for (int m = 0; m < n_modes_rec; m++) p[7 + m] = nm(nm_gen);
                for (int m = 0; m < n_modes_lig; m++) p[7 + n_modes_rec + m] = nm(nm_gen);
            }
    });
}
