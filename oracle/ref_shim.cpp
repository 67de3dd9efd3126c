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

#include <fstream>
#include <random>
#include <string>

#include "complex.h"
#include "fastdfire.h"
#include "kmeans.h"
#include "read_input_structure.h"
#include "prepare_gso_tasks.h"
#include "swarm_centers.h"

// Defined in the reference sources without a header declaration.
void minimum_volume_ellipsoid(double *center, double *poles, double *atom_coordinates,
                              double precision, int num_atoms);
void init_center_translate(double *center_coord, double *coord, vector<string> element,
                           int *structure_size, const int num_atom, const int num_structure);
void get_nm_mask(int *mask, vector<string> residue_names, map<string, string> STANDARD_TYPES,
                 map<string, string> MODIFIED_TYPES, string *DNA_STANDARD_TYPES,
                 string *RNA_STANDARD_TYPES, int *residue_atom_number, int *residue_size);

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

// libstdc++'s std::default_random_engine + std::normal_distribution<double>,
// the generator cal_anm.cc:166-184 draws its extents from (the third-party
// arithmetic the ANM restatement follows).
void ref_std_normal(unsigned seed, int n, double mu, double sigma, double *out) {
    std::default_random_engine g;
    g.seed(seed);
    std::normal_distribution<double> d(mu, sigma);
    for (int i = 0; i < n; i++) out[i] = d(g);
}

// kmeans_init_centers3 + run_kmeans3 (kmeans.cc:126-243)
void ref_kmeans(const double *points, int n, int k, int seed, int iters, double *centres, int *group) {
    std::vector<double> p(points, points + 3 * (size_t)n);
    kmeans_init_centers3(centres, p.data(), n, k, seed);
    run_kmeans3(group, p.data(), centres, iters, n, k);
}

// cal_rmsd (prepare_gso_tasks.cc:1258-1271)
double ref_rmsd(const double *a, const double *b, int n) {
    return cal_rmsd(const_cast<double *>(a), const_cast<double *>(b), n, n);
}

// minimum_volume_ellipsoid (prepare_gso_tasks.cc:106-267): the centre
void ref_mvee(const double *xyz, int n, double tol, double *centre) {
    double poles[18];
    std::vector<double> c(xyz, xyz + 3 * (size_t)n);
    minimum_volume_ellipsoid(centre, poles, c.data(), tol, n);
}

// ---- structures read by the reference's own reader (read_input_structure.cc) ----
void *ref_pdb_open(const char *path, int noxt, int noh, int now) {
    Complex *c = new Complex();
    read_input_structure(*c, path, noxt != 0, noh != 0, now != 0, false, (int)std::strlen(path));
    return c;
}
void ref_pdb_close(void *h) { delete static_cast<Complex *>(h); }
int ref_pdb_natoms(void *h) { return static_cast<Complex *>(h)->num_atoms; }
int ref_pdb_nresidues(void *h) { return static_cast<Complex *>(h)->num_res_type; }
void ref_pdb_coords(void *h, double *xyz) {
    Complex *c = static_cast<Complex *>(h);
    std::memcpy(xyz, c->atom_coordinates, sizeof(double) * 3 * c->num_atoms);
}
// residue_atom_number[0..nres]: the first atom of every residue and the end
void ref_pdb_residue_bounds(void *h, int *b) {
    Complex *c = static_cast<Complex *>(h);
    std::memcpy(b, c->residue_atom_number, sizeof(int) * (c->num_res_type + 1));
}
void ref_pdb_atom_name(void *h, int i, char *out) {
    std::snprintf(out, 8, "%s", static_cast<Complex *>(h)->atom_name[i].c_str());
}
// get_docking_model (prepare_gso_tasks.cc:31-69): the first num_atoms types
void ref_pdb_types(void *h, int *types) {
    Complex *c = static_cast<Complex *>(h);
    FastDifire fd;
    std::vector<int> all((size_t)c->chain.size() * c->residue_atom_number[c->num_res_type] + 1, -1);
    get_docking_model(all.data(), *c, fd);
    std::memcpy(types, all.data(), sizeof(int) * c->num_atoms);
}
// get_nm_mask (complex.cc:149-209)
void ref_pdb_mask(void *h, int *mask) {
    Complex *c = static_cast<Complex *>(h);
    get_nm_mask(c->mask, c->residue_name, c->STANDARD_TYPES, c->MODIFIED_TYPES,
                c->DNA_STANDARD_TYPES, c->RNA_STANDARD_TYPES, c->residue_atom_number,
                c->residue_size);
    std::memcpy(mask, c->mask, sizeof(int) * c->num_atoms);
}
// init_center_translate (initialization.cc:3-31)
void ref_pdb_centre(void *h, double *t) {
    Complex *c = static_cast<Complex *>(h);
    init_center_translate(t, c->atom_coordinates, c->elements, c->structure_size, c->num_atoms,
                          c->num_structures);
}
// write_pdb (prepare_gso_tasks.cc:1279-1349) with the given coordinates
void ref_pdb_write(void *h, const double *xyz, const char *path) {
    Complex *c = static_cast<Complex *>(h);
    std::ofstream f(path);
    write_pdb(*c, f, const_cast<double *>(xyz));
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
