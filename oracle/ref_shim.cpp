// Test-only bridge to the reference's own CPU functions, compiled against the
// reference sources where they lie (oracle/Makefile, target `ref`) and linked
// with oracle/_ref/libref_gso.so built from
// source/dock/src/prepare_gso_tasks.cc.  It exposes extern "C" wrappers so the
// oracle tests can compare the C restatement with the reference bit for bit.
// Nothing in the product links this file.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "complex.h"
#include "fastdfire.h"
#include "prepare_gso_tasks.h"
#include "swarm_centers.h"

template <class T>
static bool rd(FILE *f, T *p, size_t n) {
    return std::fread(p, sizeof(T), n, f) == n;
}

extern "C" {

// cdfire_calculate_dfire (prepare_gso_tasks.cc:1522-1615): returns the score.
// g and s are passed as 0 so that the reference's debug print (g==3 && s==3)
// stays silent.
double ref_dfire_score(const double *table, const int *rt, const int *lt, const double *r_xyz,
                       const double *l_xyz, int nr, int nl) {
    FastDifire fd;
    fd.difire_energy = const_cast<double *>(table);
    double e = 0.0;
    cdfire_calculate_dfire(&e, 0, nullptr, nullptr, const_cast<int *>(rt), const_cast<int *>(lt),
                           fd, const_cast<double *>(r_xyz), const_cast<double *>(l_xyz), nr, nl, 0,
                           0);
    fd.difire_energy = nullptr;
    return e;
}

// rotate (prepare_gso_tasks.cc:277-338)
void ref_rotate(double *xyz, const double *q, int n) { rotate(xyz, const_cast<double *>(q), n); }

// slerp (prepare_gso_tasks.cc:1370-1421)
void ref_slerp(double *self, double *other, double t) { slerp(self, other, t); }

// move_extent (prepare_gso_tasks.cc:1352-1368)
void ref_move_extent(double *self, double *other, int n, double step) {
    move_extent(self, other, n, step);
}

// compute_probability_moving_from_neighbors (prepare_gso_tasks.cc:1463-1488)
void ref_probabilities(double *prob, int *nei, double *lucif, int nnei, double cur) {
    compute_probability_moving_from_neighbors(prob, nei, lucif, nnei, cur);
}

// select_random_neighbor (prepare_gso_tasks.cc:1424-1459); swarm = 1 keeps
// the reference's debug print (swarm==0 && n==56) silent.
int ref_select(double u, int *nei, double *prob, int nnei, int self, int G) {
    std::vector<int> sel(G, -1);
    select_random_neighbor(sel.data(), u, nei, prob, 0, self, nnei, 1);
    return sel[self];
}

// ---- whole-run harness over cal_gso_tasks_cpu (prepare_gso_tasks.cc:359-822) ----
// The reference runs its compiled STEPS (3) with 10+10 modes and then calls
// exit(0) (line 822) without writing output; an atexit handler saves
// centers.pos.  Call it in a child process.
static double *g_pos = nullptr;
static size_t g_pos_len = 0;
static const char *g_out = nullptr;

static void dump_pos() {
    FILE *f = std::fopen(g_out, "wb");
    if (!f) return;
    std::fwrite(g_pos, sizeof(double), g_pos_len, f);
    std::fclose(f);
}

// Input layout (all little-endian): int32 nr, nl, S, G; then
// r_xyz f64[nr*3], r_modes f64[10*nr*3], r_mask i32[nr], r_type i32[nr],
// l_xyz f64[nl*3], l_modes f64[10*nl*3], l_mask i32[nl], l_type i32[nl],
// lig_ref f64[3], table f64[168*168*20], pos f64[S*G*27].
int ref_cpu_run(const char *in_path, const char *out_path) {
    FILE *f = std::fopen(in_path, "rb");
    if (!f) return 1;
    int hdr[4];
    if (!rd(f, hdr, 4)) return 2;
    const int nr = hdr[0], nl = hdr[1], S = hdr[2], G = hdr[3];
    const int P = 27;
    Complex rec, lig;
    rec.num_atoms = nr;
    lig.num_atoms = nl;
    rec.atom_coordinates = new double[nr * 3];
    rec.modes = new double[10 * nr * 3];
    rec.mask = new int[nr];
    int *rtype = new int[nr];
    lig.atom_coordinates = new double[nl * 3];
    lig.modes = new double[10 * nl * 3];
    lig.mask = new int[nl];
    int *ltype = new int[nl];
    double *lref = new double[3];
    FastDifire fd;
    fd.difire_energy = new double[168 * 168 * 20];
    SwarmCenters centers;
    centers.swarms = S;
    centers.pos_len = G * P;
    centers.anm_rec = 10;
    centers.anm_lig = 10;
    centers.swarm_start = 0;
    centers.swarm_chunk = S;
    centers.pos = new double[(size_t)S * G * P];
    bool ok = rd(f, rec.atom_coordinates, nr * 3) && rd(f, rec.modes, 10 * nr * 3) &&
              rd(f, rec.mask, nr) && rd(f, rtype, nr) && rd(f, lig.atom_coordinates, nl * 3) &&
              rd(f, lig.modes, 10 * nl * 3) && rd(f, lig.mask, nl) && rd(f, ltype, nl) &&
              rd(f, lref, 3) && rd(f, fd.difire_energy, 168 * 168 * 20) &&
              rd(f, centers.pos, (size_t)S * G * P);
    std::fclose(f);
    if (!ok) return 3;
    g_pos = centers.pos;
    g_pos_len = (size_t)S * G * P;
    g_out = out_path;
    std::atexit(dump_pos);
    cal_gso_tasks_cpu(centers, rec, lig, fd, 0, 0.5, 0.5, true, 0.5, 10, 10, false, S, G, nullptr,
                      nullptr, lref, nullptr, rtype, ltype);
    return 4;  // not reached: the reference exits inside
}
}
