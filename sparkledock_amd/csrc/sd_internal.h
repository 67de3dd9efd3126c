// Internal declarations shared by the HIP kernels and the C-ABI layer of
// libsdgso.so.  Device-side data layouts (all in HBM, owned by the library):
//
//   complex (one per docking problem)
//     r0      f64 [3][nr]          receptor coordinates, SoA
//     rmodes  f64 [nmr][3][nr]     receptor modes, SoA, mask folded in (zeros)
//     l0      f64 [3][nl]          ligand coordinates, SoA
//     lmodes  f64 [nml][3][nl]
//     rmodes_am, lmodes_am  f64 [n][nm][3]: the same, atom-major (guard band)
//     rtype   i32 [nr]             type * 3360 (row offset into the table)
//     ltype   i32 [nl]             type * 20
//     table64 f64 [168*168*20+20]  DFIRE table + 20 zeros (bin 20 of (167,167))
//     table   f32 [168*168*20+20]  the same rounded to FP32 (FP32 guard form)
//     tablep  f32 [168*3528+64]    the FP32 values as [a][bin 0..20][b] (FP32 form)
//     ltype_b i32 [nl]             4 * type b (byte offset into a tablep row)
//   receptor tiles (FP32 guard form): 64 atoms per tile, grouped by bisection
//     tiles   f32x4 [ntiles]       tile centre (x,y,z) and radius
//     rmod32  f32 [nmr][3][ntiles*64]      modes by slot (the GEMM probe)
//     rpos32  f32x4 [ntiles*64]            FP32 coordinates (atom test)
//     rrel4   f64 [ntiles*64][4]           FP64 coordinates, tablep row's byte offset 4*type*3528
//     rmodA   f32 [ntiles*64][rmod_stride] modes by slot (displacement)
//     rperm   i32 [ntiles*64]      original atom index (-1 = padding)
//   ligand blocks (FP32 guard form): 64 atoms per block, grouped by bisection
//     lperm   i32 [nlblocks*64]    original atom index (-1 = padding)
//     lblocks f32x4 [nlblocks]     block centre in the ligand frame, radius
//     lhalf   f32x4 [nlblocks]     half extents of the block's box about that centre
//     lblk    f64 [nlblocks][3][64]        coordinates of the block slots
//     lblkm   f32 [nlblocks][3*nml][64]    modes of the block slots (FP32)
//       (row 3*m + c: coordinate c of mode m)
//   run state, N = sum(swarms) * G glowworms
//     pos [N][P] f64, luciferin/score/vision [N] f64, ref [N][3] f64,
//     nnei/moved/select/step [N] i32, pairs [N] i64,
//     stream [sum S][n_steps][G] f64
//     ldisp  f32, per glowworm [nlblocks][3][64]: the ligand's FP32 mode
//            displacement by block slot (FP32 form, ligand modes > 0)
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "../../include/sdgso.h"

namespace sd {

constexpr int kTableLen = SDGSO_TABLE_LEN + SDGSO_NBINS;
// The FP32 kernels' copy of the table, [a][bin][b]: entry a*kPRow + bin*168 +
// b holds the reference's flat entry a*3360 + b*20 + bin, for bin 0..20 (bin
// 20 is the reference's read past a row at d2 == 225, interface.h:35), and
// zeros follow.  With the ligand type fastest, the pairs of one receptor atom
// in one distance shell with ligand atoms of nearby types share 128-byte
// lines, which one gather instruction fetches once.
constexpr int kPRow = (SDGSO_NBINS + 1) * SDGSO_NTYPES;           // 3528
constexpr int kTableLenP = SDGSO_NTYPES * kPRow + 64;             // + zeros
constexpr int kZeroIdxP = SDGSO_NTYPES * kPRow;                   // a zero
constexpr int kTile = 64;              // receptor atoms per tile = wavefront width
constexpr int kBlock = 64;             // ligand atoms per block = wavefront width
constexpr double kQuantScale = 1048576.0;  // 2^20: fixed-point units per A
constexpr double kCutoff2 = 225.0;
constexpr int kMaxTiles32 = 1024;      // FP32 form: receptor tiles per complex (65536 atoms)
constexpr int kStatSlots = 256;        // slots per filter-statistics counter
constexpr int kScore64Threads = 256;   // FP64 form: threads per workgroup (one glowworm)

// FP64 form: ligand atoms padded to a multiple of 16 (far-away atoms), and its
// LDS: FP64 x, y, z and an int type per padded atom, two 256-row sum buffers.
inline __host__ __device__ int score64_padded_nl(int nl) { return (nl + 15) & ~15; }
inline size_t score64_lds_bytes(int max_nl) {
    const size_t nlp = (size_t)score64_padded_nl(max_nl);
    return sizeof(double) * (3 * nlp + 2 * kScore64Threads) + sizeof(int) * nlp;
}

// float4 per receptor slot that hold its FP32 modes (rmodA): 3*nm floats in
// whole 128-byte lines, rounded up to 8, 16 or 32 float4 (nm <= 32).
inline __host__ __device__ int mode_chunks(int nm) {
    const int lines = (3 * nm + 31) / 32;
    return lines <= 1 ? 8 : (lines <= 2 ? 16 : 32);
}

// Order of a slot's FP32 modes in rmodA.
//  - Up to 10 modes (8 float4 per slot): by coordinate, packed, [x of modes
//    0..nm-1 | y | z] with each coordinate in max(nm, 4) consecutive floats,
//    so a float4 holds modes of at most two coordinates.
//  - 11 to 20 modes (16 float4): component-major, each coordinate in
//    ceil(nm / 4) whole float4 (zeros past nm), so every float4 holds modes of
//    one coordinate; the FP32 kernels take a variant for that order.
//  - More: mode-major, x, y, z of mode 0, then of mode 1, and so on.
inline __host__ __device__ bool modes_by_component(int nm) { return nm >= 11 && nm <= 20; }
inline __host__ __device__ int mode_slot_index(int nm, int m, int k) {
    if (nm <= 10) return k * (nm > 4 ? nm : 4) + m;
    if (nm <= 20) return k * 4 * ((nm + 3) / 4) + m;
    return 3 * m + k;
}

// Per-complex device descriptor (plain data; lives in a device array).
struct DevComplex {
    int nr, nl, nmr, nml;
    const double *r0;
    const double *rmodes;
    const double *l0;
    const double *lmodes;
    const double *rmodes_am;  // [nr][nmr][3] the receptor modes atom-major (FP32 form's guard)
    const double *lmodes_am;  // [nl][nml][3]
    const int *rtype;
    const int *ltype;
    const float *table;      // FP32 copy (FP32 form, its guard-band pairs)
    const double *table64;   // the caller's FP64 values (FP64 form)
    const float *tablep;     // FP32 copy in the [a][bin][b] order (FP32 form)
    const int *ltype_b;      // [nl] 4 * ligand type b (byte offset into tablep rows)
    double lref[3];
    // receptor tiles
    int ntiles;
    const float4 *tiles;
    const float *rmod32;     // [nmr][3][ntiles*64] modes (f32; the GEMM probe's B)
    // per slot, for gathers by slot: near-test coordinates (f32, absolute,
    // w = 0), coordinates and table row (f64 absolute x, y, z and the row
    // type * kPRow as a double) and modes (f32, in the order
    // modes_by_component(nmr) gives, padded to 4 * mode_chunks(nmr) =
    // rmod_stride floats)
    const float4 *rpos32;
    const double *rrel4;
    const float *rmodA;
    int rmod_stride;
    const int *rperm;
    float rmode_max[SDGSO_MAX_MODES];  // max over atoms of |M_m,i|
    // ligand blocks
    int nlblocks;
    const int *lperm;
    const float4 *lblocks;
    const float4 *lhalf;     // [nlblocks] half extents of the block's box (ligand frame)
    const double *lblk;      // [nlblocks][3][64] coordinates by block slot
    const float *lblkm;      // [nlblocks][3*nml][64] modes by block slot, FP32
    float lmode_max[SDGSO_MAX_MODES];
};

// Arguments of one scoring launch over a set of glowworms.
struct ScoreArgs {
    const DevComplex *cx;   // [n_complex]
    const int *swarm_cx;    // [n_swarms] complex of each swarm
    int G, P;
    int n_glow;             // glowworms in this launch
    const double *pos;      // [n_glow][P]
    double *score;          // [n_glow]
    double *luciferin;      // [n_glow] (updated when update_luciferin)
    double *ref;            // [n_glow][3]
    int64_t *pairs;         // [n_glow] or null
    int *step;              // [n_glow] or null
    double rho, gamma;
    int update_luciferin;
    // FP32 guard form
    float band;             // guard band (A^2)
    unsigned long long *guard_count;  // FP64 recomputes (optional)
    double *partial;        // [max_lblocks][n_glow] per-block energies (FP32 form)
    long long *partial_pairs;  // [max_lblocks][n_glow] per-block pair counts
    int max_lblocks;
    int nmr;                // receptor modes of every complex of the launch
    const int *work;        // FP32 form: (glowworm, ligand block) pairs, or (glowworm, first
                            // block of a pair) for two-block variants; null otherwise
    int n_work;
    // FP32 form, optional: filter statistics [6][kStatSlots] (waves, waves
    // with candidate tiles, candidate tiles, near receptor atoms, queued
    // pairs, in-cutoff pairs), each the sum over its slots
    unsigned long long *stats;
    // FP32 form, work-queue variants: the queue's counter (one word)
    unsigned *counter;
    // FP32 form, optional: the ligand's FP32 mode displacement of every
    // (glowworm, ligand block slot) from ligand_disp_kernel, [3][64] floats
    // per (glowworm, block); glowworm gid of complex k (a.cx[k], nlb blocks)
    // starts at ldisp + ldisp_base[k] + gid * nlb * 192.  Null: the scoring
    // set-up sums the FP32 ligand modes itself.
    const float *ldisp;
    const long long *ldisp_base;
    // FP32 form, one-item-per-wave variants: 0 takes the work items in list
    // order; K > 0 hands them out in chunks of K per XCD (workgroup w runs on
    // XCD w % 8, so the i-th wave an XCD receives takes item (i / K * 8 +
    // w % 8) * K + i % K; the last 8 K items of a launch stay in order)
    int xcd_chunk;
};

// The ligand displacement GEMM of a run: for every complex k of the run,
// items [item0[k], item0[k + 1]) are (32-glowworm tile, ligand block) pairs,
// tile-major, over the complex's glowworms [glow0[k], glow0[k + 1]).
struct LigDispArgs {
    const DevComplex *cx;   // [n_cx]
    int n_cx;
    const double *pos;      // [N][P] run poses
    int P, nmr;
    const long long *item0; // [n_cx + 1]
    const int *glow0;       // [n_cx + 1]
    float *ldisp;           // as ScoreArgs::ldisp
    const long long *ldisp_base;
    long long n_items;
};

struct SelectMoveArgs {
    int G, P, nmr, nml;
    int n_swarms;
    double *pos;            // [N][P]
    const double *luciferin;
    const double *ref;
    double *vision;
    int *nnei;
    int *moved;
    int *select;
    const double *stream;   // [n_swarms][n_steps][G]
    int n_steps;
    int k;                  // step index into the stream (when d_k is null)
    int *d_k;               // device-resident step index (graph replay)
    double beta, max_vision;
    int max_neighbors;
    double step_t, step_r, step_nm;
    int do_select, do_move;
    double *mp;             // [P][N] each glowworm's copy of its chosen neighbour's pre-move pose
};

// Final coordinates of chosen atoms for a range of glowworms of one complex.
struct CoordsArgs {
    const DevComplex *cx;   // the complex (device)
    const double *pos;      // run poses [N][P]
    int P;
    int first;              // first glowworm (global index)
    const int *rec_idx;     // [n_rec] receptor atoms, or null for 0..n_rec-1
    int n_rec;
    const int *lig_idx;     // [n_lig] ligand atoms, or null for 0..n_lig-1
    int n_lig;
    double *out;            // [count][n_rec + n_lig][3]
};

// Launchers (defined in the .hip files).
hipError_t launch_score64(const ScoreArgs &a, int max_nl, hipStream_t s);
hipError_t launch_score32(const ScoreArgs &a, int variant, hipStream_t s);
int score32_variants();
const char *score32_variant_name(int v);
int score32_variant_blocks(int v);  // ligand blocks per wave: 1, or 2 (work list of block pairs)
int score32_default_variant();
hipError_t launch_finish32(const ScoreArgs &a, hipStream_t s);
hipError_t launch_ligand_disp(const LigDispArgs &a, hipStream_t s);
hipError_t launch_select_move(const SelectMoveArgs &a, hipStream_t s);
hipError_t launch_advance(int *d_k, hipStream_t s);
hipError_t launch_coords(const CoordsArgs &a, int count, hipStream_t s);
hipError_t launch_rmsd(const double *bb, int nb, int G, int n_swarms, double *rmsd, hipStream_t s);
hipError_t launch_gather_probe(const float *table, unsigned n, int per_thread, float *out,
                               int blocks, hipStream_t s);
hipError_t launch_anm_gemm_probe(const DevComplex *cx, const double *pos, int P, int G,
                                 const int *pairs, size_t first, size_t count, int chunks,
                                 float *out, hipStream_t s);

void set_error(const std::string &msg);

// Checks for a usable gfx950 device and makes it current (SDGSO_OK or an
// error status with the message set).
int use_device();

// What other translation units need of a complex: its device-resident
// descriptor and sizes.
struct ComplexView {
    const DevComplex *dev;
    int nr, nl, nmr, nml, nlblocks, ntiles;
};
bool complex_view(sdgso_complex c, ComplexView *v);

}  // namespace sd
