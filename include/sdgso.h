/*
 * sdgso.h: C-ABI of the MI355X (gfx950) GSO docking hot path.
 *
 * One library, libsdgso.so, owns all device memory behind opaque handles and
 * takes plain host pointers and sizes.  Every call returns an int status
 * (SDGSO_OK = 0); sdgso_last_error() returns the message of the calling
 * thread's last failure.  Handles are not thread-safe: use one per thread.
 *
 * Reference interfaces replaced (paths relative to the reference root):
 *   sdgso_complex_create  <- the per-complex inputs prepare_gso_tasks() hands to
 *                            the GSO engine: Complex coordinates/modes/mask,
 *                            FastDifire table, dfire_objects_* types and the
 *                            ligand reference point
 *                            (source/dock/src/prepare_gso_tasks.cc:1151-1256;
 *                             source/dock/include/prepare_gso_tasks.h:8-25)
 *   sdgso_score_poses     <- cdfire_calculate_dfire(), the scoring plugin, with
 *                            the pose transform in front of it
 *                            (prepare_gso_tasks.cc:457-545, 1522-1615;
 *                             prepare_gso_tasks.h:36-40)
 *   sdgso_run_*           <- cal_gso_tasks_gpu()/cal_gso_tasks_cpu(): the GSO
 *                            step loop over every glowworm of every swarm
 *                            (source/dock/src/cuda/prepare_gso_task.cu:2479-2972;
 *                             prepare_gso_tasks.cc:359-822; prepare_gso_tasks.h:18-25)
 */
#ifndef SDGSO_H
#define SDGSO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SDGSO_NTYPES 168
#define SDGSO_NBINS 20
#define SDGSO_TABLE_LEN (SDGSO_NTYPES * SDGSO_NTYPES * SDGSO_NBINS)
#define SDGSO_MAX_MODES 32

enum {
    SDGSO_OK = 0,
    SDGSO_EINVAL = 1,  /* bad argument (size, range, null pointer) */
    SDGSO_EHIP = 2,    /* HIP runtime failure or no usable gfx950 device */
    SDGSO_ENOMEM = 3,  /* device or host allocation failed */
    SDGSO_ESTATE = 4   /* call out of order (e.g. stepping past the stream) */
};

/* Scoring forms. */
enum {
    SDGSO_FP64 = 0,       /* plain FP64 brute force over all Nr*Nl pairs */
    SDGSO_FP32_GUARD = 1  /* FP32 distances on receptor tiles, FP64 recompute of
                             pairs within the guard band of a bin boundary,
                             FP32 table values summed in FP64 */
};

/* GSO constants (prepare_gso_tasks.cc:386-400; steps :783, 791-794). */
typedef struct {
    double rho, gamma, beta;
    double vision0, max_vision, luciferin0;
    int max_neighbors;
    double step_translation, step_rotation, step_nmodes;
} sdgso_params;

typedef struct sdgso_complex_s *sdgso_complex;
typedef struct sdgso_run_s *sdgso_run;

const char *sdgso_last_error(void);
int sdgso_version(void);
void sdgso_default_params(sdgso_params *p);

/* Receptor/ligand of one docking problem.
 *   rec_xyz [n_rec][3] f64, rec_type [n_rec] in [0,168), rec_mask [n_rec] (1 =
 *   moved by the normal modes, as complex.cc:149-209 sets it),
 *   rec_modes [n_rec_modes][n_rec][3] f64 (NULL when n_rec_modes == 0); the
 *   same for the ligand; lig_ref [3] the ligand reference point;
 *   table [168*168*20] f64, index a*3360 + b*20 + bin, as
 *   FastDifire::difire_energy holds it (fastdfire.h:93, filled at
 *   prepare_gso_tasks.cc:1175-1189).  The FP64 form gathers these values.  The
 *   FP32 guard form gathers them rounded to nearest FP32: each term is within
 *   2^-24 |T| of the FP64 value, so its energy is within 2^-24 * sum |T| (over
 *   in-cutoff pairs) of the FP64 form's, on top of its summation error; a
 *   table of FP32 values (the synthetic one) loses nothing. */
int sdgso_complex_create(int n_rec, const double *rec_xyz, const int *rec_type,
                         const int *rec_mask, int n_rec_modes, const double *rec_modes,
                         int n_lig, const double *lig_xyz, const int *lig_type,
                         const int *lig_mask, int n_lig_modes, const double *lig_modes,
                         const double *lig_ref, const double *table, sdgso_complex *out);
int sdgso_complex_destroy(sdgso_complex c);

/* Score n_poses poses [n][7+nr+nl] = (tx,ty,tz, qw,qx,qy,qz, rec extents,
 * lig extents); out_score [n]; optional out_ref [n][3] ligand reference points,
 * optional out_pairs [n] in-cutoff pair counts (NULL to skip).
 * guard_band in A^2 (SDGSO_FP32_GUARD only; <= 0 selects 1e-4). */
int sdgso_score_poses(sdgso_complex c, int n_poses, const double *poses, int form,
                      double guard_band, double *out_score, double *out_ref,
                      int64_t *out_pairs);

/* The scoring call without allocation.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  It scores
 * as sdgso_score_poses does, with the same kernels, but poses [n][7+nr+nl],
 * out_score [n], out_ref [n][3] and out_pairs [n] are DEVICE pointers
 * (out_ref and out_pairs may be NULL), the work is enqueued on hip_stream
 * (a hipStream_t; NULL = the default stream) without waiting for it, and the
 * scratch the kernels need comes from one static device array of
 * SDGSO_NOALLOC_SCRATCH_MB MiB per GPU, held until the process ends.  A call
 * whose scratch would exceed it (per pose 4 bytes, 24 more without out_ref,
 * and in the FP32 form 16 * nb + 8 * ceil(nb / 2) more, nb = ceil(nl / 64)
 * ligand blocks; plus 256-byte alignment of each region) fails with
 * SDGSO_EINVAL and a
 * message giving the most poses that fit.  Every call on a GPU shares that
 * GPU's array: calls that overlap on one GPU race. */
#define SDGSO_NOALLOC_SCRATCH_MB 64
int sdgso_score_poses_noalloc(sdgso_complex c, int n_poses, const double *poses, int form,
                              double guard_band, double *out_score, double *out_ref,
                              int64_t *out_pairs, void *hip_stream);

/* A GSO run over one or more complexes sharing the glowworm count and pose
 * length.  Complex k owns swarms [sum_{j<k} n_swarms[j], ...).
 *   pos    [sum S][G][P] initial poses
 *   stream [sum S][n_steps][G] the uniform draws of each swarm and step
 *          (each complex's segment is its own mt19937 stream, swarm-major)
 * The run keeps all state resident on the device. */
int sdgso_run_create(int n_complex, const sdgso_complex *complexes, const int *n_swarms,
                     int n_glowworms, const double *pos, const double *stream, int n_steps,
                     const sdgso_params *params, int form, double guard_band, sdgso_run *out);
int sdgso_run_destroy(sdgso_run r);

/* Attach a HIP stream (hipStream_t) for all later launches of this run; NULL =
 * the library's own stream. */
int sdgso_run_set_stream(sdgso_run r, void *hip_stream);

/* Advance n steps (asynchronous on the run's stream).  use_graph != 0 replays
 * the step as a captured hipGraph (captured at the first such call; while
 * profiling is enabled, as two graphs, the scoring launches and the rest of
 * the step, with the timing events between them, and a third in front of
 * them when the ligand displacement GEMM is on). */
int sdgso_run_steps(sdgso_run r, int n, int use_graph);

/* Phases of one step, for single-phase tests: bit 0 score + luciferin,
 * bit 1 neighbour selection, bit 2 move + vision; the step counter advances
 * when bit 2 runs. */
int sdgso_run_phases(sdgso_run r, int mask);

int sdgso_run_sync(sdgso_run r);

/* Number of steps taken so far. */
int sdgso_run_step_index(sdgso_run r);

/* Copy state back (synchronises).  Any pointer may be NULL.
 *   pos [N][P], luciferin/score/vision [N], nnei/moved/select [N],
 *   ref [N][3], pairs [N] (in-cutoff pairs of the last scoring), N = sum S * G. */
int sdgso_run_read(sdgso_run r, double *pos, double *luciferin, double *score, double *vision,
                   int *nnei, int *moved, int *select, double *ref, int64_t *pairs);

/* Overwrite state (synchronises).  NULL pointers leave that field unchanged. */
int sdgso_run_write(sdgso_run r, const double *pos, const double *luciferin, const double *score,
                    const double *vision, const int *moved, const double *ref);

/* Per-phase HIP-event timing on the run's stream, for eager steps and graph
 * replays alike.  enable != 0 starts recording; read returns totals in
 * milliseconds and step counts since the last enable: times[0] the scoring
 * launches, times[1] the rest of the step (the FP32 form's per-glowworm
 * epilogue, selection + movement), times[2] the ligand displacement GEMM
 * (0 when it is off); times_ms[3], launches[3]. */
int sdgso_run_profile(sdgso_run r, int enable);
int sdgso_run_profile_read(sdgso_run r, double *times_ms, int64_t *launches);

/* Guard-band statistics of the FP32 form since the last enable of profiling:
 * stats[0] pairs recomputed in FP64. */
int sdgso_run_guard_stats(sdgso_run r, int64_t *stats);

/* Work order of the FP32 scoring launches (no counterpart in the reference):
 * xcd_chunk = 0 hands the (glowworm, ligand block) work items to waves in
 * list order; xcd_chunk = K > 0 in chunks of K items per XCD of the MI355X
 * (workgroup w runs on XCD w % 8, so XCD x takes chunks x, x + 8, ...: the
 * receptor region that one XCD's L2 holds is that of fewer swarms).  Results
 * do not depend on it.  A run starts in list order; sdgso_run_tune times a
 * scoring pass of the chosen variants in each order (K = 8192) and keeps the
 * faster.  sdgso_run_work_order returns the order and tune_ms[2], the last
 * tune's milliseconds per pass in list order and chunked (either pointer may
 * be NULL). */
int sdgso_run_set_work_order(sdgso_run r, int xcd_chunk);
int sdgso_run_work_order(sdgso_run r, int *xcd_chunk, double *tune_ms);

/* The ligand's normal-mode displacement as an FP32 matrix-core GEMM,
 * (poses x modes) . (modes x 3 Nl), the form the reference's GPU path
 * materialises per pose (prepare_gso_task.cu:268-333; the CPU path adds the
 * modes per atom, prepare_gso_tasks.cc:487-504).  When it is on, an FP32 run
 * with ligand modes runs it before each scoring pass, by
 * v_mfma_f32_32x32x2f32 over every glowworm of the run, and the scoring
 * kernels read its output instead of summing the ligand modes themselves; it
 * needs 768 bytes of device memory per (glowworm, 64-atom ligand block).  A
 * run starts with it off.  sdgso_run_tune times a scoring pass of the chosen
 * kernel variants with it (the GEMM included) and without it and keeps the
 * faster (it stays off when its output does not fit in device memory);
 * sdgso_run_set_ligand_gemm turns it on (SDGSO_ENOMEM when its output does
 * not fit) or off (enable = 0); the FP64 form and runs without ligand modes
 * never use it.  sdgso_run_ligand_gemm reports whether it is on, the bytes it
 * holds and tune_ms[2], the last tune's milliseconds per scoring pass with
 * and without it (zeros before a tune); any pointer may be NULL.
 * sdgso_run_ligand_displacements runs it on the current poses and returns
 * the FP32 displacements of glowworms [first, first + count) of one complex,
 * out[count][Nl][3] by original atom (SDGSO_ESTATE when it is off). */
int sdgso_run_set_ligand_gemm(sdgso_run r, int enable);
int sdgso_run_ligand_gemm(sdgso_run r, int *enabled, int64_t *bytes, double *tune_ms);
int sdgso_run_ligand_displacements(sdgso_run r, int first, int count, float *out);

/* Autotuner of the FP32 guard-band scoring kernel (no counterpart in the
 * reference, whose tile sizes are compile-time constants,
 * prepare_gso_task.cu:2503-2514).  sdgso_tune_variants() is the number of
 * compiled variants (ligand-block size x filter form x gathers in flight per
 * lane x register/occupancy target x work distribution); a name starting
 * with "bcast2" is a variant whose wave takes a 128-atom ligand block (two
 * 64-atom blocks of one glowworm), the others take 64 atoms; "-queue" names
 * the persistent form (a grid of resident waves taking 8 work items at a time
 * from an atomic counter), the others take one work item per wave.  sdgso_run_tune scores the current poses of
 * each complex of the run with each variant `repeats` times (luciferin and
 * step counters untouched) and keeps, per complex, the fastest for later
 * steps: times_ms[n_complex][sdgso_tune_variants()] receives the
 * milliseconds per launch of each complex's share and chosen[n_complex] the
 * choices (either may be NULL).  Complexes that share a variant share a
 * launch.  sdgso_run_set_variant uses one variant for every complex (e.g. one
 * chosen by an earlier sdgso_run_tune of the same workload) without timing;
 * sdgso_run_variants reads the per-complex choices. */
int sdgso_tune_variants(void);
const char *sdgso_tune_variant_name(int variant);
int sdgso_run_tune(sdgso_run r, int repeats, double *times_ms, int *chosen);
int sdgso_run_set_variant(sdgso_run r, int variant);
int sdgso_run_variants(sdgso_run r, int *variants);

/* Filter statistics of the FP32 guard-band scoring kernel (no counterpart in
 * the reference, which scores every pair): scores the run's current poses
 * once with the chosen variant, leaves the run state unchanged, and
 * returns stats[6] = {waves (glowworm x ligand block, or x pair of blocks
 * for a two-block variant), waves with a candidate
 * receptor tile, candidate tiles visited, receptor atoms passing the atom
 * test, pairs queued for exact evaluation, pairs inside the cutoff}. */
int sdgso_run_filter_stats(sdgso_run r, int64_t *stats);

/* Random-gather rate of the device over a float table of `table_floats`
 * entries (no counterpart in the reference): the measured ceiling of the
 * scoring kernels' table reads.  Returns gathers per second in *rate. */
int sdgso_probe_gather(int64_t table_floats, double *rate);

/* The receptor's normal-mode displacement as an FP32 matrix-core GEMM, timed
 * on the run's current poses (no counterpart in the reference, a measurement
 * of the form the FP32 kernels do not take): for every swarm, the
 * displacement of every receptor atom in the tiles that some glowworm of the
 * swarm can reach, for each of its glowworms, by v_mfma_f32_32x32x2f32, and
 * written to HBM.  result[5] = {milliseconds, bytes written, matrix-core
 * flops issued, displacements computed, (swarm, tile) pairs}.  The run must
 * be in the FP32 form and have receptor modes; its state is unchanged. */
int sdgso_probe_anm_gemm(sdgso_run r, double *result);

/* ---- output side of a run ---- */

/* Final coordinates of glowworms [first, first + count) of one complex of the
 * run, from the current poses in FP64 in the reference's order (receptor
 * modes; ligand modes, rotation, translation: prepare_gso_tasks.cc:882-920):
 * rec_xyz[count][Nr][3], lig_xyz[count][Nl][3] (either may be NULL). */
int sdgso_run_final_coords(sdgso_run r, int first, int count, double *rec_xyz, double *lig_xyz);

/* Clustering of every swarm of complex `complex` of the run (BSAS,
 * prepare_gso_tasks.cc:961-1080): glowworms in decreasing order of their last
 * score (ties by index) join the first cluster whose representative (its
 * first member) is within `cutoff` RMSD over the chosen backbone atoms of the
 * receptor and the ligand (the reference uses CA and P atoms and 4 A), else
 * found a new cluster.  cluster[S][G] receives each glowworm's cluster number
 * within its swarm (in order of foundation) and representative[S][G] (may be
 * NULL) the representative's glowworm index within the swarm. */
int sdgso_run_cluster(sdgso_run r, int complex, const int *rec_backbone, int n_rec_backbone,
                      const int *lig_backbone, int n_lig_backbone, double cutoff, int *cluster,
                      int *representative);

/* ---- molecular structures (host only) ---- */

/* A structure read from a PDB file: the first model's ATOM/HETATM records,
 * coordinates rounded through float as the reference stores them, residues
 * split where (name, number, insertion) changes, chains where the chain id
 * changes (read_input_structure.cc:131-307).  Flags drop OXT atoms, atoms
 * whose name starts with H, and HOH residues (the setup's noxt / noh / now).
 * An element outside C N O H S P CL MG FE PB SE F is an error (the
 * reference exits). */
typedef struct sdgso_structure_s *sdgso_structure;
#define SDGSO_PDB_NO_OXT 1
#define SDGSO_PDB_NO_H 2
#define SDGSO_PDB_NO_WATER 4
int sdgso_structure_read_pdb(const char *path, int flags, sdgso_structure *out);
/* One CA atom of residue ALA i+1 in chain A per coordinate triple: the
 * metadata of a synthetic Calpha chain, for writing and clustering. */
int sdgso_structure_synthetic(int n_atoms, const double *xyz, sdgso_structure *out);
int sdgso_structure_destroy(sdgso_structure s);
int sdgso_structure_size(sdgso_structure s, int *n_atoms, int *n_residues, int *n_chains);
int sdgso_structure_coords(sdgso_structure s, double *xyz);
/* name[8], resname[8], chain[4] receive NUL-terminated strings (may be NULL). */
int sdgso_structure_atom(sdgso_structure s, int atom, char *name, char *resname, int *resseq,
                         char *chain);
/* DFIRE atom types in [0, 168) (fastdfire.h tables as used by
 * get_docking_model, prepare_gso_tasks.cc:31-69).  An atom outside the
 * tables is an error (the reference reads an undefined type). */
int sdgso_structure_dfire_types(sdgso_structure s, int *types);
/* 1 for atoms of protein, modified-protein, DNA and RNA residues, else 0
 * (get_nm_mask, complex.cc:149-209). */
int sdgso_structure_nm_mask(sdgso_structure s, int *mask);
/* Indices of the backbone atoms named CA or P, in file order (the atoms the
 * clustering RMSD uses, prepare_gso_tasks.cc:1017-1050); *n receives the
 * count (index may be NULL to count only). */
int sdgso_structure_backbone(sdgso_structure s, int *index, int *n);
/* The translation that centres the structure: minus the mean of the non-H
 * atoms (init_center_translate, initialization.cc:3-31). */
int sdgso_structure_centre(sdgso_structure s, double *translation);
/* Atoms in the reference's write_pdb line format (prepare_gso_tasks.cc:
 * 1279-1349), serial numbers from 1, coordinates from xyz (NULL: the
 * structure's own); append != 0 appends to the file. */
int sdgso_structure_write_pdb(sdgso_structure s, const double *xyz, const char *path, int append);

/* Normal modes by the anisotropic network model (cal_anm.cc:61-219): xyz
 * [N][3] (in/out) are the structure's centred coordinates, re-centred in
 * place on the mean of all atoms as the reference does before the modes;
 * the network nodes are the CA atoms (else C4), one per residue, with a 15 A
 * cutoff; the lowest n_modes non-zero modes (eigenvalue >= 1e-6) are given
 * to every atom of their residue, normalised over the 3N components and
 * scaled by sqrt(N) * rmsd / coef * sqrt(1 / eigenvalue), coef the mean norm
 * of n_draws random extents (libstdc++ default_random_engine seeded 324324,
 * N(0, 1)) weighted by 1 / eigenvalue.  modes[n_modes][N][3]; eigenvalues
 * [n_modes] (may be NULL).  Each mode's sign is fixed so that its largest
 * component is positive (the reference takes the eigensolver's).  Uses the
 * GPU (rocSOLVER). */
int sdgso_structure_anm(sdgso_structure s, double *xyz, int n_modes, double rmsd, int n_draws,
                        double *modes, double *eigenvalues);

/* k-means++ seeding from std::mt19937(seed) (kmeans_init_centers3,
 * kmeans.cc:126-187) followed by `iterations` Lloyd steps (run_kmeans3,
 * kmeans.cc:191-243) over points[n][3]: centres[k][3], group[n] (may be
 * NULL). */
int sdgso_kmeans(const double *points, int n, int k, uint32_t seed, int iterations, double *centres,
                 int *group);

/* Starting poses of `swarms` swarms of `glowworms` glowworms around the
 * receptor (calculate_surface_point and calculate_start_pos,
 * cal_anm.cc:541-900): rec_xyz and lig_xyz are the structures' coordinates
 * before centring, rec_translation the receptor's centring translation.
 * Swarm centres come from k-means over the atoms of surface residues, 100
 * sphere points around each centroid at a distance set by the ligand's
 * diameter, without points near another centroid or within 3 A of the
 * receptor, and a second k-means; each glowworm gets a translation uniform in
 * the swarm_radius sphere and a uniform rotation (std::mt19937
 * start_points_seed) and N(0, 1) mode extents (std::default_random_engine
 * anm_seed).  centres[swarms][3] (may be NULL); pos[swarms][glowworms][7 +
 * n_modes_rec + n_modes_lig].  The number of swarms must be given (the
 * reference's count from the receptor's SASA needs FreeSASA). */
int sdgso_starting_poses(sdgso_structure receptor, const double *rec_xyz, int n_lig,
                         const double *lig_xyz, const double *rec_translation, int swarms,
                         int glowworms, double swarm_radius, uint32_t start_points_seed,
                         uint32_t anm_seed, int n_modes_rec, int n_modes_lig, double *centres,
                         double *pos);

/* Centre of the minimum-volume enclosing ellipsoid of the points
 * (Khachiyan's iteration stopped when the weight update is at most
 * `tolerance`; the reference uses 0.01), the ligand reference point of
 * minimum_volume_ellipsoid, prepare_gso_tasks.cc:106-267. */
int sdgso_reference_point(int n_atoms, const double *xyz, double tolerance, double *centre);

/* ---- host utilities (no device use) ---- */

/* The reference's random stream: u[s][k][g] = successive
 * uniform_real_distribution<double>(0,1) draws from mt19937(seed)
 * (prepare_gso_tasks.cc:411-427). */
int sdgso_stream(uint32_t seed, int n_swarms, int n_steps, int n_glowworms, double *u);

/* Seeded synthetic complex: random-walk chains (3.8 A steps) in globules of
 * radius (N*11.4)^(1/3), uniform types over 168, orthonormal random modes.
 * Any output pointer may be NULL. */
int sdgso_synth_molecule(uint64_t seed, int n_atoms, int n_modes, double *xyz, int *type,
                         double *modes);
/* 168*168*20 f32 table drawn N(0,1)*0.5. */
int sdgso_synth_table(uint64_t seed, float *table);
/* Initial poses [S][G][7+nr+nl] around swarm centres on a sphere of radius
 * centre_radius (translations in a swarm_radius ball, Shoemake quaternions,
 * extents U[-1,1]; cal_anm.cc:845-892). */
int sdgso_synth_poses(uint64_t seed, int n_swarms, int n_glowworms, int n_rec_modes,
                      int n_lig_modes, double centre_radius, double swarm_radius, double *pos);

#ifdef __cplusplus
}
#endif
#endif
