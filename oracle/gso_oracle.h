/*
 * CPU FP64 oracle for the SparkleDock GSO step: TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference CPU path
 * (source/dock/src/prepare_gso_tasks.cc, cal_gso_tasks_cpu, lines 359-822)
 * used as the correctness checker for the HIP kernels and as the CPU baseline
 * timed beside them.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product library never links it.
 *
 * Every function follows the reference's FP64 operation order; build with
 * -ffp-contract=off so that no multiply-add is fused.
 *
 * Parity pins (what checks each function):
 *   or_rotate, or_dfire_energy, or_dist_bin, or_slerp, or_move_extent,
 *   or_probabilities, or_select: bit-exact against the compiled reference
 *     functions (oracle/_ref) and against committed golden vectors; plus
 *     closed-form/invariant tests (rotation norms, one-hot tables, boundary
 *     +-1ulp bins, roulette hand cases).
 *   or_gso_step: bit-exact against the reference's own cal_gso_tasks_cpu run
 *     for its compiled 3 steps (10+10 modes) through oracle/_ref.
 *   or_stream: against the C++11 mt19937 known answer and libstdc++'s
 *     uniform_real_distribution output.
 *   or_rmsd: bit-exact against the reference's cal_rmsd; closed forms
 *     (identical sets, a rigid translation).
 *   or_bsas: hand-built pose sets whose clusters follow from the distances
 *     (the reference's loop is unreachable code, after its exit(0)).
 *   or_mvee_centre: against the compiled reference's
 *     minimum_volume_ellipsoid within 1e-9 A (its Eigen 4x4 inverse rounds
 *     differently); translation equivariance; symmetric point sets.
 */
#ifndef SD_GSO_ORACLE_H
#define SD_GSO_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_NTYPES 168
#define OR_NBINS 20
/* flat table length plus the 20-entry zero pad that the d2 == 225 bin (20)
 * can reach for types (167,167) through the reference's flat index. */
#define OR_TABLE_LEN (OR_NTYPES * OR_NTYPES * OR_NBINS + OR_NBINS)

typedef struct {
    int nr, nl;              /* receptor / ligand atoms */
    int nmr, nml;            /* receptor / ligand normal modes */
    const double *r_xyz;     /* [nr][3] */
    const double *l_xyz;     /* [nl][3] */
    const int *r_type;       /* [nr] in [0,168) */
    const int *l_type;       /* [nl] */
    const int *r_mask;       /* [nr] 1 = displaced by modes */
    const int *l_mask;       /* [nl] */
    const double *r_modes;   /* [nmr][nr][3] */
    const double *l_modes;   /* [nml][nl][3] */
    const double *lig_ref;   /* [3] ligand reference point */
    const double *table;     /* [OR_TABLE_LEN] */
} or_complex;

typedef struct {
    double rho, gamma, beta;
    double vision0, max_vision, luciferin0;
    int max_neighbors;
    double step_t, step_r, step_nm;
} or_params;

typedef struct {
    int S, G, P;             /* swarms, glowworms per swarm, pose length */
    double *pos;             /* [S][G][P] */
    double *luciferin;       /* [S][G] */
    double *score;           /* [S][G] */
    double *vision;          /* [S][G] */
    double *ref;             /* [S][G][3] ligand reference point of the current pose */
    int *nnei;               /* [S][G] */
    int *moved;              /* [S][G] */
    int *step;               /* [S][G] */
    int *select;             /* [S][G] neighbour chosen in the last step */
    double *move_pos;        /* [S][G][P] scratch: copy of the chosen neighbour's pose */
    int64_t *pairs;          /* [S][G] in-cutoff pair count of the last scoring, or NULL */
} or_state;

void or_default_params(or_params *p);

/* distance -> DFIRE bin (interface.h:30-35, prepare_gso_tasks.cc:1553-1556).
 * Returns -1 for excluded pairs (d2 > 225 or d2 == 0), else 0..20. */
int or_dist_bin(double d2);

/* Quaternion sandwich rotation of n points in place (prepare_gso_tasks.cc:277-338). */
void or_rotate(double *xyz, const double *q, int n);

/* Raw DFIRE sum E = sum_i sum_j T[...] in reference order (prepare_gso_tasks.cc:1522-1570).
 * *pairs (optional) receives the in-cutoff pair count. */
double or_dfire_energy(const double *r_xyz, int nr, const int *r_type,
                       const double *l_xyz, int nl, const int *l_type,
                       const double *table, int64_t *pairs);

/* score = (E*0.0157 - 4.7)*-1 (prepare_gso_tasks.cc:1571) */
double or_score_from_energy(double e);

/* Build the receptor and ligand coordinates and the ligand reference point of one
 * pose (prepare_gso_tasks.cc:457-526).  R [nr][3], L [nl][3], ref [3]. */
void or_pose_coords(const or_complex *c, const double *pose, double *R, double *L, double *ref);

/* Score of one pose; optional pair count. */
double or_score_pose(const or_complex *c, const double *pose, double *ref_out, int64_t *pairs);

void or_slerp(double *self, double *other, double t);                    /* :1370-1421 */
void or_move_extent(double *self, const double *other, int n, double step); /* :1352-1368 */
void or_probabilities(double *prob, const int *nei, const double *lucif, int nnei,
                      double current);                                   /* :1463-1488 */
/* Roulette choice (prepare_gso_tasks.cc:1424-1459); index clamped to [1, nnei]. */
int or_select(double u, const int *nei, const double *prob, int nnei, int self);

/* State management. */
int or_state_alloc(or_state *st, int S, int G, int P, int with_pairs);
void or_state_free(or_state *st);
void or_state_init(or_state *st, const or_params *p, const double *pos);

/* One GSO step k over all swarms (prepare_gso_tasks.cc:429-819).
 * u is the [S][n_steps][G] stream.  rescore_all != 0 scores every glowworm
 * (the reference GPU policy); 0 skips unmoved ones (the reference CPU policy);
 * the results are identical.  threads <= 0 uses the OpenMP default. */
void or_gso_step(const or_complex *c, const or_params *p, or_state *st,
                 const double *u, int n_steps, int k, int rescore_all, int threads);

/* Phases of one step, exposed for the single-phase parity tests. */
void or_phase_score(const or_complex *c, const or_params *p, or_state *st, int rescore_all, int threads);
void or_phase_select(const or_params *p, or_state *st, const double *u, int n_steps, int k, int threads);
void or_phase_move(const or_params *p, or_state *st, int nmr, int nml, int threads);

/* The reference's random stream: u[s][k][g] = successive
 * uniform_real_distribution<double>(0,1) draws of mt19937(seed), swarm-major
 * (prepare_gso_tasks.cc:411-427). */
void or_stream(uint32_t seed, int S, int n_steps, int G, double *u);

/* Raw mt19937 outputs (for the known-answer test). */
void or_mt19937(uint32_t seed, int n, uint32_t *out);

/* Backbone RMSD of two coordinate sets of n atoms, summed per atom as
 * dx^2, dy^2, dz^2 in order (cal_rmsd, prepare_gso_tasks.cc:1258-1271). */
double or_rmsd(const double *a, const double *b, int n);

/* Greedy clustering of G poses (prepare_gso_tasks.cc:975-1080): coords is
 * [G][n][3] backbone coordinates, score [G]; poses in decreasing score order
 * (ties by index) join the first cluster whose first member is within
 * cutoff, else found a new one.  cluster[G] receives cluster numbers in order
 * of foundation, rep[G] the representative of each pose's cluster.  Build
 * decisions: every pose is counted once (the reference adds the top pose to
 * its own cluster a second time); ties keep index order (the reference's
 * std::sort leaves their order unspecified). */
void or_bsas(const double *coords, const double *score, int G, int n, double cutoff,
             int *cluster, int *rep);

/* Centre of the minimum-volume enclosing ellipsoid (Khachiyan, stopped at a
 * weight update <= tol; prepare_gso_tasks.cc:106-267), with a cofactor 4x4
 * inverse. */
void or_mvee_centre(const double *xyz, int n, double tol, double *centre);

#ifdef __cplusplus
}
#endif
#endif
