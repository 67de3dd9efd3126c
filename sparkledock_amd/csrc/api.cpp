// C-ABI entry points of libsdgso.so (declared in include/sdgso.h).
//
// Host responsibilities: argument validation, the per-complex preprocessing
// (SoA layouts, mask folded into the modes, receptor tiles and ligand blocks
// grouped by recursive bisection, for the FP32 guard form), device residency of every array, the
// step sequence on one HIP stream, optional hipGraph replay of a step, and
// HIP-event timing of the phases.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdint>
#include <cstring>
#include <memory>
#include <random>
#include <string>
#include <vector>

#include "sd_internal.h"

namespace sd {

static thread_local std::string g_error;

void set_error(const std::string &msg) { g_error = msg; }

struct Failure {
    int code;
};

[[noreturn]] static void fail(int code, const std::string &msg) {
    set_error(msg);
    throw Failure{code};
}

#define HIPCHK(expr)                                                                     \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess)                                                            \
            fail(SDGSO_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_));       \
    } while (0)

template <class F>
static int guarded(F &&f) {
    try {
        f();
        return SDGSO_OK;
    } catch (const Failure &e) {
        return e.code;
    } catch (const std::bad_alloc &) {
        set_error("host allocation failed");
        return SDGSO_ENOMEM;
    } catch (const std::exception &e) {
        set_error(e.what());
        return SDGSO_EINVAL;
    }
}

// ---------------------------------------------------------------- device setup
static int g_device = -1;

static void ensure_device() {
    if (g_device >= 0) {
        HIPCHK(hipSetDevice(g_device));
        return;
    }
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
        fail(SDGSO_EHIP, "no HIP device visible (libsdgso needs an MI355X / gfx950)");
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, dev));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        fail(SDGSO_EHIP, std::string("device is ") + prop.gcnArchName +
                             "; libsdgso is built for gfx950 only");
    g_device = dev;
}

template <class T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { reset(); }
    void reset() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void alloc(size_t count) {
        reset();
        if (count == 0) return;
        if (hipMalloc(&p, count * sizeof(T)) != hipSuccess) {
            p = nullptr;
            // clear the runtime's last error, which a later launch check would report
            (void)hipGetLastError();
            fail(SDGSO_ENOMEM, "hipMalloc of " + std::to_string(count * sizeof(T)) + " bytes failed");
        }
        n = count;
    }
    void upload(const T *src, size_t count) {
        alloc(count);
        if (count) HIPCHK(hipMemcpy(p, src, count * sizeof(T), hipMemcpyHostToDevice));
    }
    void upload(const std::vector<T> &v) { upload(v.data(), v.size()); }
};

// Recursive coordinate bisection into groups of `size` atoms: split the set
// along its longest extent so that the lower part holds a whole number of
// groups; every group is full except the last.  Gives compact groups (small
// bounding spheres) where a space-filling-curve order of a sparse chain does
// not.
static void rcb(const std::vector<double> &xyz, std::vector<int> &idx, size_t lo, size_t hi,
                int size, std::vector<int> &out) {
    const size_t n = hi - lo;
    const size_t k = (n + size - 1) / size;
    if (k <= 1) {
        out.insert(out.end(), idx.begin() + lo, idx.begin() + hi);
        return;
    }
    double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
    for (size_t i = lo; i < hi; i++)
        for (int c = 0; c < 3; c++) {
            mn[c] = std::min(mn[c], xyz[3 * (size_t)idx[i] + c]);
            mx[c] = std::max(mx[c], xyz[3 * (size_t)idx[i] + c]);
        }
    int ax = 0;
    for (int c = 1; c < 3; c++)
        if (mx[c] - mn[c] > mx[ax] - mn[ax]) ax = c;
    const size_t cut = lo + (k / 2) * size;
    std::nth_element(idx.begin() + lo, idx.begin() + cut, idx.begin() + hi, [&](int a, int b) {
        const double va = xyz[3 * (size_t)a + ax], vb = xyz[3 * (size_t)b + ax];
        return va < vb || (va == vb && a < b);
    });
    rcb(xyz, idx, lo, cut, size, out);
    rcb(xyz, idx, cut, hi, size, out);
}

static std::vector<int> group_order(const std::vector<double> &xyz, int n, int size) {
    std::vector<int> idx(n), out;
    for (int i = 0; i < n; i++) idx[i] = i;
    out.reserve(n);
    rcb(xyz, idx, 0, n, size, out);
    return out;
}

// Ligand atoms in block order: 128-atom groups by recursive bisection, each
// halved the same way, so that blocks 2j and 2j + 1 (64 atoms each) are the
// two halves of one compact group, which the two-block kernels take together.
// Atoms within each group of `size` in order of DFIRE type (then index): a
// round of the FP32 form's evaluation gathers table entries a*3360 + b*20 +
// bin for 64 pairs, and pairs of equal or neighbouring types share 128-byte
// table lines that one load instruction fetches once.  The groups' atom sets,
// and so their spheres and boxes, do not change.
static void sort_groups_by_type(std::vector<int> &order, const int *type, int size) {
    for (size_t lo = 0; lo < order.size(); lo += size) {
        const size_t hi = std::min(order.size(), lo + (size_t)size);
        std::sort(order.begin() + lo, order.begin() + hi, [&](int a, int b) {
            return type[a] < type[b] || (type[a] == type[b] && a < b);
        });
    }
}

static std::vector<int> block_order(const std::vector<double> &xyz, int n) {
    const std::vector<int> g = group_order(xyz, n, 2 * kBlock);
    std::vector<int> out;
    out.reserve(n);
    for (size_t lo = 0; lo < g.size(); lo += 2 * kBlock) {
        const size_t hi = std::min(g.size(), lo + 2 * (size_t)kBlock);
        std::vector<int> idx(g.begin() + lo, g.begin() + hi);
        rcb(xyz, idx, 0, idx.size(), kBlock, out);
    }
    return out;
}

}  // namespace sd

using namespace sd;

// ---------------------------------------------------------------- complex
struct sdgso_complex_s {
    int nr = 0, nl = 0, nmr = 0, nml = 0;
    // host copies for the measurement probes: receptor tiles (centre,
    // radius) and the ligand's largest |xyz| about its frame's origin
    std::vector<float4> h_tiles;
    double lig_rmax = 0;
    DevBuf<double> r0, rmodes, l0, lmodes;
    DevBuf<double> rmodes_am, lmodes_am;  // the same modes atom-major, [n][nm][3]
    DevBuf<int> rtype, ltype;
    DevBuf<double> table64;    // the caller's table (FP64 form)
    DevBuf<float> table;       // the same rounded to FP32 (FP32 guard form)
    DevBuf<float> tablep;      // the FP32 values as [a][bin][b] (FP32 form's gathers)
    DevBuf<int> ltype_b;       // ligand types
    DevBuf<float4> tiles, lblocks, lhalf;
    DevBuf<float> rmod32;      // [nmr][3][ntiles*64]
    DevBuf<float4> rpos32;     // [ntiles*64] f32 coordinates (near test)
    DevBuf<double> rrel4;      // [ntiles*64][4] f64 absolute coordinates, type row
    DevBuf<float> rmodA;       // [ntiles*64][rmod_stride] modes by slot
    int rmod_stride = 0;
    DevBuf<int> rperm, lperm;
    DevBuf<double> lblk;       // [nlblocks][3][64] ligand coordinates by block slot
    DevBuf<float> lblkm;       // [nlblocks][3*nml][64] ligand modes by block slot, FP32
    DevComplex dev{};          // descriptor with device pointers
    DevBuf<DevComplex> dself;  // the same descriptor in device memory
    double max_rec_extent = 0; // |xyz| maxima, for the guard-band bound
};

struct sdgso_run_s {
    int n_complex = 0;
    std::vector<sdgso_complex> cxs;
    std::vector<int> n_swarms;
    int S = 0, G = 0, P = 0, N = 0, nmr = 0, nml = 0;
    int n_steps = 0, k = 0;
    int form = SDGSO_FP64;
    float band = 1e-4f;
    sdgso_params params{};
    int max_nl = 0, max_lblocks = 0;
    DevBuf<DevComplex> dcx;
    DevBuf<int> swarm_cx;
    DevBuf<double> pos, luciferin, score, vision, ref, stream, partial;
    DevBuf<double> mp;  // select/move: the chosen neighbours' pre-move poses, [P][N]
    DevBuf<int> nnei, moved, select, step, work, work2;
    DevBuf<int64_t> pairs;
    DevBuf<long long> partial_pairs;
    DevBuf<int> d_k;
    DevBuf<unsigned long long> guard_count;
    DevBuf<unsigned> counter;  // the work-queue variants' counter
    hipStream_t own_stream = nullptr;
    hipStream_t stream_h = nullptr;
    hipGraphExec_t graph = nullptr;        // one step
    hipGraphExec_t graph_score = nullptr;  // profiled replays: the scoring launches,
    hipGraphExec_t graph_rest = nullptr;   //   then the rest of the step
    bool profiling = false;
    // FP32 form: scoring-kernel variant of each complex (sdgso_run_tune) and
    // the complex's range [work0[k], work0[k+1]) of the work list (work2_0:
    // of the list of block pairs that two-block variants read)
    std::vector<int> variants;
    std::vector<size_t> work0, work2_0;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_score, ev_move, ev_lgemm;
    double t_score = 0, t_move = 0, t_lgemm = 0;
    int64_t n_score = 0, n_move = 0, n_lgemm = 0;
    // FP32 form with ligand modes: the ligand displacement GEMM that runs
    // before each scoring pass when lgemm is set.  ldisp holds its output,
    // [glowworm][block][3][64] f32 (complex k's glowworm gid at
    // ldisp_base[k] + gid * nlblocks_k * 192); the GEMM's items of complex k
    // are [lg_item0[k], lg_item0[k + 1]) over glowworms [lg_glow0[k],
    // lg_glow0[k + 1]).
    bool lgemm = false;
    DevBuf<float> ldisp;
    DevBuf<long long> ldisp_base, lg_item0;
    DevBuf<int> lg_glow0;
    long long lg_items = 0;
    hipGraphExec_t graph_lgemm = nullptr;  // profiled replays: the GEMM alone
    double lgemm_tune_ms[2] = {0, 0};      // last tune: scoring pass with, without it
    // FP32 form: the scoring launches' work order (ScoreArgs::xcd_chunk), set
    // by the tuner; its last timing of a scoring pass in list order and
    // XCD-chunked
    int xcd_chunk = 0;
    double order_tune_ms[2] = {0, 0};
};

// The chunk of work items per XCD that the tuner tries against list order.
constexpr int kXcdChunk = 8192;

static void check_table_and_types(const double *table, const int *rt, int nr, const int *lt, int nl) {
    if (!table) fail(SDGSO_EINVAL, "table is null");
    for (int i = 0; i < SDGSO_TABLE_LEN; i++)
        if (!std::isfinite(table[i]))
            fail(SDGSO_EINVAL, "table value at " + std::to_string(i) + " is not finite");
    for (int i = 0; i < nr; i++)
        if (rt[i] < 0 || rt[i] >= SDGSO_NTYPES)
            fail(SDGSO_EINVAL, "receptor type out of [0,168) at atom " + std::to_string(i));
    for (int i = 0; i < nl; i++)
        if (lt[i] < 0 || lt[i] >= SDGSO_NTYPES)
            fail(SDGSO_EINVAL, "ligand type out of [0,168) at atom " + std::to_string(i));
}

// SoA [3][n] coordinates, [nm][3][n] modes with the mask folded in (zero rows
// for atoms the modes do not move: R += 0*e leaves R unchanged exactly).
static void soa_inputs(int n, const double *xyz, const int *mask, int nm, const double *modes,
                       std::vector<double> &soa, std::vector<double> &msoa) {
    soa.assign(3 * (size_t)n, 0.0);
    for (int i = 0; i < n; i++)
        for (int c = 0; c < 3; c++) soa[(size_t)c * n + i] = xyz[3 * (size_t)i + c];
    msoa.assign((size_t)nm * 3 * n, 0.0);
    for (int m = 0; m < nm; m++)
        for (int i = 0; i < n; i++) {
            if (mask && mask[i] != 1) continue;
            for (int c = 0; c < 3; c++)
                msoa[((size_t)m * 3 + c) * n + i] = modes[((size_t)m * n + i) * 3 + c];
        }
}

static void build_tiles(sdgso_complex c, const double *xyz, const int *type, const std::vector<double> &msoa) {
    const int n = c->nr, nm = c->nmr;
    std::vector<double> v(xyz, xyz + 3 * (size_t)n);
    std::vector<int> order = group_order(v, n, kTile);
    sort_groups_by_type(order, type, kTile);
    const int ntiles = (n + kTile - 1) / kTile;
    const size_t slots = (size_t)ntiles * kTile;
    std::vector<float4> tiles(ntiles);
    std::vector<double> rrel(3 * slots, 0.0);
    std::vector<float> m32((size_t)nm * 3 * slots, 0.0f);
    std::vector<int> ts(slots, 0), perm(slots, -1);
    // modes of a slot: 3*nm floats in mode_chunks(nm) float4, in the order
    // mode_slot_index gives
    const int stride = 4 * mode_chunks(nm);
    std::vector<float4> p32(slots);
    std::vector<double> r4(4 * slots, 0.0);
    std::vector<float> mA((size_t)stride * slots, 0.0f);
    for (int t = 0; t < ntiles; t++) {
        const int a0 = t * kTile, a1 = std::min(n, a0 + kTile);
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        for (int s = a0; s < a1; s++)
            for (int k = 0; k < 3; k++) {
                lo[k] = std::min(lo[k], xyz[3 * (size_t)order[s] + k]);
                hi[k] = std::max(hi[k], xyz[3 * (size_t)order[s] + k]);
            }
        // centre rounded to a float so that kernels and host agree exactly
        float cf[3];
        for (int k = 0; k < 3; k++) cf[k] = (float)(0.5 * (lo[k] + hi[k]));
        double rad = 0;
        for (int s = a0; s < a1; s++) {
            double d2 = 0;
            for (int k = 0; k < 3; k++) {
                const double d = xyz[3 * (size_t)order[s] + k] - (double)cf[k];
                d2 += d * d;
            }
            rad = std::max(rad, std::sqrt(d2));
        }
        tiles[t] = make_float4(cf[0], cf[1], cf[2], (float)(rad * (1 + 1e-6) + 1e-4));
        for (int s = a0; s < t * kTile + kTile; s++) {
            if (s < a1) {
                const int i = order[s];
                perm[s] = i;
                ts[s] = 4 * type[i] * kPRow;  // byte offset of the type's tablep row
                for (int k = 0; k < 3; k++) rrel[(size_t)k * slots + s] = xyz[3 * (size_t)i + k] - (double)cf[k];
                for (int m = 0; m < nm; m++)
                    for (int k = 0; k < 3; k++) {
                        const float mv = (float)msoa[((size_t)m * 3 + k) * n + i];
                        m32[((size_t)m * 3 + k) * slots + s] = mv;
                        mA[(size_t)s * stride + mode_slot_index(nm, m, k)] = mv;
                    }
            } else {
                // padding slot: far away, type 0, zero modes
                for (int k = 0; k < 3; k++) rrel[(size_t)k * slots + s] = 1.0e4;
            }
            // absolute coordinates (the input's, or far away for padding) and
            // the type row, exact as a double
            for (int k = 0; k < 3; k++)
                r4[4 * (size_t)s + k] = s < a1 ? xyz[3 * (size_t)order[s] + k]
                                               : rrel[(size_t)k * slots + s] + (double)cf[k];
            r4[4 * (size_t)s + 3] = (double)ts[s];
            p32[s] = make_float4((float)r4[4 * (size_t)s], (float)r4[4 * (size_t)s + 1],
                                 (float)r4[4 * (size_t)s + 2], 0.0f);
        }
    }
    for (int m = 0; m < nm; m++) {
        double mx = 0;
        for (int i = 0; i < n; i++) {
            double d2 = 0;
            for (int k = 0; k < 3; k++) {
                const double x = msoa[((size_t)m * 3 + k) * n + i];
                d2 += x * x;
            }
            mx = std::max(mx, std::sqrt(d2));
        }
        c->dev.rmode_max[m] = (float)(mx * (1 + 1e-5));
    }
    c->tiles.upload(tiles);
    c->h_tiles = tiles;
    c->rmod32.upload(m32);
    c->rpos32.upload(p32);
    c->rrel4.upload(r4);
    c->rmodA.upload(mA);
    c->rmod_stride = stride;
    c->rperm.upload(perm);
    c->dev.ntiles = ntiles;
}

static void build_lblocks(sdgso_complex c, const double *xyz, const int *type,
                          const std::vector<double> &msoa) {
    const int n = c->nl, nm = c->nml;
    std::vector<double> v(xyz, xyz + 3 * (size_t)n);
    std::vector<int> order = block_order(v, n);
    sort_groups_by_type(order, type, kBlock);
    const int nb = (n + kBlock - 1) / kBlock;
    std::vector<float4> blocks(nb), half(nb);
    std::vector<int> perm((size_t)nb * kBlock, -1);
    for (int b = 0; b < nb; b++) {
        const int a0 = b * kBlock, a1 = std::min(n, a0 + kBlock);
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        for (int s = a0; s < a1; s++)
            for (int k = 0; k < 3; k++) {
                lo[k] = std::min(lo[k], xyz[3 * (size_t)order[s] + k]);
                hi[k] = std::max(hi[k], xyz[3 * (size_t)order[s] + k]);
            }
        float cf[3], hf[3];
        for (int k = 0; k < 3; k++) {
            cf[k] = (float)(0.5 * (lo[k] + hi[k]));
            // the box about the (float) centre, rounded outwards
            const double h = std::max(hi[k] - (double)cf[k], (double)cf[k] - lo[k]);
            hf[k] = (float)(h * (1 + 1e-6) + 1e-4);
        }
        half[b] = make_float4(hf[0], hf[1], hf[2], 0.f);
        double rad = 0;
        for (int s = a0; s < a1; s++) {
            double d2 = 0;
            for (int k = 0; k < 3; k++) {
                const double d = xyz[3 * (size_t)order[s] + k] - (double)cf[k];
                d2 += d * d;
            }
            rad = std::max(rad, std::sqrt(d2));
            perm[s] = order[s];
        }
        blocks[b] = make_float4(cf[0], cf[1], cf[2], (float)(rad * (1 + 1e-6) + 1e-4));
    }
    for (int m = 0; m < nm; m++) {
        double mx = 0;
        for (int i = 0; i < n; i++) {
            double d2 = 0;
            for (int k = 0; k < 3; k++) {
                const double x = msoa[((size_t)m * 3 + k) * n + i];
                d2 += x * x;
            }
            mx = std::max(mx, std::sqrt(d2));
        }
        c->dev.lmode_max[m] = (float)(mx * (1 + 1e-5));
    }
    // coordinates (FP64) and modes (FP32) of every block slot, slot-contiguous
    // (padding: zeros)
    std::vector<double> lb((size_t)nb * 3 * kBlock, 0.0);
    std::vector<float> lbm((size_t)nb * 3 * nm * kBlock, 0.0f);
    for (int s = 0; s < nb * kBlock; s++) {
        const int i = perm[s];
        if (i < 0) continue;
        double *B = lb.data() + (size_t)(s / kBlock) * 3 * kBlock + s % kBlock;
        float *BM = lbm.data() + (size_t)(s / kBlock) * 3 * nm * kBlock + s % kBlock;
        for (int k = 0; k < 3; k++) B[(size_t)k * kBlock] = xyz[3 * (size_t)i + k];
        for (int m = 0; m < nm; m++)
            for (int k = 0; k < 3; k++)
                BM[(size_t)(3 * m + k) * kBlock] = (float)msoa[((size_t)m * 3 + k) * n + i];
    }
    c->lblocks.upload(blocks);
    c->lhalf.upload(half);
    c->lperm.upload(perm);
    c->lblk.upload(lb);
    c->lblkm.upload(lbm);
    c->dev.nlblocks = nb;
}

namespace sd {

int use_device() { return guarded([] { ensure_device(); }); }

bool complex_view(sdgso_complex c, ComplexView *v) {
    if (!c || !v) return false;
    v->dev = c->dself.p;
    v->nr = c->nr;
    v->nl = c->nl;
    v->nmr = c->nmr;
    v->nml = c->nml;
    v->nlblocks = c->dev.nlblocks;
    v->ntiles = c->dev.ntiles;
    return true;
}

}  // namespace sd

extern "C" {

const char *sdgso_last_error(void) { return g_error.c_str(); }

int sdgso_version(void) { return 101; }

void sdgso_default_params(sdgso_params *p) {
    if (!p) return;
    p->rho = 0.4;
    p->gamma = 0.6;
    p->beta = 0.08;
    p->vision0 = 0.2;
    p->max_vision = 5.0;
    p->luciferin0 = 5.0;
    p->max_neighbors = 5;
    p->step_translation = 0.5;
    p->step_rotation = 0.5;
    p->step_nmodes = 0.5;
}

int sdgso_complex_create(int n_rec, const double *rec_xyz, const int *rec_type, const int *rec_mask,
                         int n_rec_modes, const double *rec_modes, int n_lig, const double *lig_xyz,
                         const int *lig_type, const int *lig_mask, int n_lig_modes,
                         const double *lig_modes, const double *lig_ref, const double *table,
                         sdgso_complex *out) {
    return guarded([&] {
        if (!out) fail(SDGSO_EINVAL, "out is null");
        *out = nullptr;
        if (n_rec <= 0 || n_lig <= 0) fail(SDGSO_EINVAL, "atom counts must be positive");
        if (n_rec_modes < 0 || n_rec_modes > SDGSO_MAX_MODES || n_lig_modes < 0 ||
            n_lig_modes > SDGSO_MAX_MODES)
            fail(SDGSO_EINVAL, "mode counts must be in [0, 32]");
        if (!rec_xyz || !lig_xyz || !rec_type || !lig_type || !lig_ref)
            fail(SDGSO_EINVAL, "null coordinate/type/reference pointer");
        if ((n_rec_modes && !rec_modes) || (n_lig_modes && !lig_modes))
            fail(SDGSO_EINVAL, "modes pointer is null");
        check_table_and_types(table, rec_type, n_rec, lig_type, n_lig);
        ensure_device();
        std::unique_ptr<sdgso_complex_s> c(new sdgso_complex_s);
        c->nr = n_rec;
        c->nl = n_lig;
        c->nmr = n_rec_modes;
        c->nml = n_lig_modes;
        std::vector<double> rsoa, rmsoa, lsoa, lmsoa;
        soa_inputs(n_rec, rec_xyz, rec_mask, n_rec_modes, rec_modes, rsoa, rmsoa);
        soa_inputs(n_lig, lig_xyz, lig_mask, n_lig_modes, lig_modes, lsoa, lmsoa);
        c->r0.upload(rsoa);
        c->rmodes.upload(rmsoa);
        c->l0.upload(lsoa);
        c->lmodes.upload(lmsoa);
        // atom-major copies for the FP32 form's FP64 guard-band recompute: an
        // atom's modes in consecutive lines
        auto atom_major = [](const std::vector<double> &msoa, int n, int nm) {
            std::vector<double> am(msoa.size());
            for (int m = 0; m < nm; m++)
                for (int k = 0; k < 3; k++)
                    for (int i = 0; i < n; i++)
                        am[((size_t)i * nm + m) * 3 + k] = msoa[((size_t)m * 3 + k) * n + i];
            return am;
        };
        c->rmodes_am.upload(atom_major(rmsoa, n_rec, n_rec_modes));
        c->lmodes_am.upload(atom_major(lmsoa, n_lig, n_lig_modes));
        std::vector<int> rt(n_rec), lt(n_lig);
        for (int i = 0; i < n_rec; i++) rt[i] = rec_type[i] * SDGSO_NTYPES * SDGSO_NBINS;
        for (int i = 0; i < n_lig; i++) lt[i] = lig_type[i] * SDGSO_NBINS;
        c->rtype.upload(rt);
        c->ltype.upload(lt);
        // both copies carry the 20-entry zero pad (bin 20 of types (167, 167))
        std::vector<double> tab64(kTableLen, 0.0);
        std::memcpy(tab64.data(), table, sizeof(double) * SDGSO_TABLE_LEN);
        c->table64.upload(tab64);
        std::vector<float> tab(kTableLen, 0.0f);
        for (int i = 0; i < SDGSO_TABLE_LEN; i++) tab[i] = (float)table[i];
        c->table.upload(tab);
        std::vector<float> tabp(kTableLenP, 0.0f);
        for (int ta = 0; ta < SDGSO_NTYPES; ta++)
            for (int bin = 0; bin <= SDGSO_NBINS; bin++)
                for (int tb = 0; tb < SDGSO_NTYPES; tb++)
                    tabp[(size_t)ta * kPRow + bin * SDGSO_NTYPES + tb] =
                        tab[(size_t)ta * SDGSO_NTYPES * SDGSO_NBINS + tb * SDGSO_NBINS + bin];
        c->tablep.upload(tabp);
        std::vector<int> lb4(n_lig);  // byte offset of each ligand type in a tablep row
        for (int i = 0; i < n_lig; i++) lb4[i] = 4 * lig_type[i];
        c->ltype_b.upload(lb4);
        DevComplex &d = c->dev;
        d.nr = n_rec;
        d.nl = n_lig;
        d.nmr = n_rec_modes;
        d.nml = n_lig_modes;
        for (int i = 0; i < n_lig; i++)
            c->lig_rmax = std::max(c->lig_rmax, std::sqrt(lig_xyz[3 * (size_t)i] * lig_xyz[3 * (size_t)i] +
                                                          lig_xyz[3 * (size_t)i + 1] * lig_xyz[3 * (size_t)i + 1] +
                                                          lig_xyz[3 * (size_t)i + 2] * lig_xyz[3 * (size_t)i + 2]));
        build_tiles(c.get(), rec_xyz, rec_type, rmsoa);
        build_lblocks(c.get(), lig_xyz, lig_type, lmsoa);
        d.r0 = c->r0.p;
        d.rmodes = c->rmodes.p;
        d.l0 = c->l0.p;
        d.lmodes = c->lmodes.p;
        d.rmodes_am = c->rmodes_am.p;
        d.lmodes_am = c->lmodes_am.p;
        d.rtype = c->rtype.p;
        d.ltype = c->ltype.p;
        d.table = c->table.p;
        d.table64 = c->table64.p;
        d.tablep = c->tablep.p;
        d.ltype_b = c->ltype_b.p;
        for (int k = 0; k < 3; k++) d.lref[k] = lig_ref[k];
        d.tiles = c->tiles.p;
        d.rmod32 = c->rmod32.p;
        d.rpos32 = c->rpos32.p;
        d.rrel4 = c->rrel4.p;
        d.rmodA = c->rmodA.p;
        d.rmod_stride = c->rmod_stride;
        d.rperm = c->rperm.p;
        d.lperm = c->lperm.p;
        d.lblocks = c->lblocks.p;
        d.lhalf = c->lhalf.p;
        d.lblk = c->lblk.p;
        d.lblkm = c->lblkm.p;
        c->dself.upload(&d, 1);
        *out = c.release();
    });
}

int sdgso_complex_destroy(sdgso_complex c) {
    return guarded([&] {
        if (!c) return;
        if (g_device >= 0) (void)hipSetDevice(g_device);
        delete c;
    });
}

}  // extern "C"

// ---------------------------------------------------------------- runs
static void profile_pair(sdgso_run r, std::vector<std::pair<hipEvent_t, hipEvent_t>> &v,
                         std::pair<hipEvent_t, hipEvent_t> &out) {
    HIPCHK(hipEventCreate(&out.first));
    HIPCHK(hipEventCreate(&out.second));
    v.push_back(out);
    (void)r;
}

static ScoreArgs score_args(sdgso_run r) {
    ScoreArgs a{};
    a.cx = r->dcx.p;
    a.swarm_cx = r->swarm_cx.p;
    a.G = r->G;
    a.P = r->P;
    a.n_glow = r->N;
    a.pos = r->pos.p;
    a.score = r->score.p;
    a.luciferin = r->luciferin.p;
    a.ref = r->ref.p;
    a.pairs = r->pairs.p;
    a.step = r->step.p;
    a.rho = r->params.rho;
    a.gamma = r->params.gamma;
    a.update_luciferin = 1;
    a.band = r->band;
    a.guard_count = r->guard_count.p;
    a.partial = r->partial.p;
    a.partial_pairs = r->partial_pairs.p;
    a.max_lblocks = r->max_lblocks;
    a.nmr = r->nmr;
    a.work = r->work.p;
    a.n_work = (int)r->work.n;
    a.counter = r->counter.p;
    a.ldisp = r->lgemm ? r->ldisp.p : nullptr;
    a.ldisp_base = r->ldisp_base.p;
    a.xcd_chunk = r->xcd_chunk;
    return a;
}

// The ligand displacement GEMM over every glowworm of the run (FP32 form,
// when enabled): it reads the current poses, so it runs before each scoring
// pass over them.
static void enqueue_lgemm(sdgso_run r) {
    if (!r->lgemm) return;
    LigDispArgs g{};
    g.cx = r->dcx.p;
    g.n_cx = r->n_complex;
    g.pos = r->pos.p;
    g.P = r->P;
    g.nmr = r->nmr;
    g.item0 = r->lg_item0.p;
    g.glow0 = r->lg_glow0.p;
    g.ldisp = r->ldisp.p;
    g.ldisp_base = r->ldisp_base.p;
    g.n_items = r->lg_items;
    HIPCHK(launch_ligand_disp(g, r->stream_h));
}

// Switches the ligand displacement GEMM on (allocating its output) or off
// (freeing it).  Runs without ligand modes, and the FP64 form, never use it.
static void set_lgemm(sdgso_run r, bool on) {
    r->lgemm = false;
    r->ldisp.reset();
    if (!on || r->form != SDGSO_FP32_GUARD || r->nml == 0) return;
    std::vector<long long> base, item0(1, 0);
    std::vector<int> glow0(1, 0);
    size_t floats = 0;
    for (int k = 0; k < r->n_complex; k++) {
        const long long nb = r->cxs[k]->dev.nlblocks;
        const long long ng = (long long)r->n_swarms[k] * r->G;
        base.push_back((long long)floats - (long long)glow0.back() * nb * 3 * kBlock);
        floats += (size_t)(ng * nb * 3 * kBlock);
        item0.push_back(item0.back() + (ng + 31) / 32 * nb);
        glow0.push_back(glow0.back() + (int)ng);
    }
    if (item0.back() > 0x7fffffffLL) fail(SDGSO_EINVAL, "too many ligand GEMM items");
    r->ldisp.alloc(floats);
    r->ldisp_base.upload(base);
    r->lg_item0.upload(item0);
    r->lg_glow0.upload(glow0);
    r->lg_items = item0.back();
    r->lgemm = true;
}

// The work of complexes [k0, k1) in the list that variant v reads: (glowworm,
// block) pairs, or (glowworm, first block of a pair) for two-block variants.
static void set_work(sdgso_run r, int v, int k0, int k1, ScoreArgs &a) {
    const bool two = score32_variant_blocks(v) == 2;
    const std::vector<size_t> &w0 = two ? r->work2_0 : r->work0;
    a.work = (two ? r->work2.p : r->work.p) + 2 * w0[k0];
    a.n_work = (int)(w0[k1] - w0[k0]);
}

// FP32 scoring of the work list: consecutive complexes that use the same
// variant share a launch.
static void launch_score32_run(sdgso_run r, const ScoreArgs &a) {
    const int n = (int)r->variants.size();
    for (int k0 = 0; k0 < n;) {
        int k1 = k0 + 1;
        while (k1 < n && r->variants[k1] == r->variants[k0]) k1++;
        ScoreArgs part = a;
        set_work(r, r->variants[k0], k0, k1, part);
        HIPCHK(launch_score32(part, r->variants[k0], r->stream_h));
        k0 = k1;
    }
}

// The scoring phase: the scoring kernel (score64, or score32 and its
// per-glowworm epilogue finish32), luciferin included.
static void enqueue_score(sdgso_run r) {
    const ScoreArgs a = score_args(r);
    if (r->form == SDGSO_FP64) {
        HIPCHK(launch_score64(a, r->max_nl, r->stream_h));
    } else {
        enqueue_lgemm(r);
        launch_score32_run(r, a);
        HIPCHK(launch_finish32(a, r->stream_h));
    }
}

static void enqueue_select_move(sdgso_run r, int do_select, int do_move) {
    SelectMoveArgs m{};
    m.G = r->G;
    m.P = r->P;
    m.nmr = r->nmr;
    m.nml = r->nml;
    m.n_swarms = r->S;
    m.pos = r->pos.p;
    m.luciferin = r->luciferin.p;
    m.ref = r->ref.p;
    m.vision = r->vision.p;
    m.nnei = r->nnei.p;
    m.moved = r->moved.p;
    m.select = r->select.p;
    m.stream = r->stream.p;
    m.n_steps = r->n_steps;
    m.k = r->k;
    m.d_k = r->d_k.p;
    m.beta = r->params.beta;
    m.max_vision = r->params.max_vision;
    m.max_neighbors = r->params.max_neighbors;
    m.step_t = r->params.step_translation;
    m.step_r = r->params.step_rotation;
    m.step_nm = r->params.step_nmodes;
    m.do_select = do_select;
    m.do_move = do_move;
    m.mp = r->mp.p;
    HIPCHK(launch_select_move(m, r->stream_h));
}

// A step in three parts: the ligand displacement GEMM (FP32 form, when
// enabled), the scoring launches, and the rest (the FP32 form's
// per-glowworm epilogue, selection + movement, the device step counter).
static void enqueue_score_launches(sdgso_run r, const ScoreArgs &a) {
    if (r->form == SDGSO_FP64)
        HIPCHK(launch_score64(a, r->max_nl, r->stream_h));
    else
        launch_score32_run(r, a);
}

static void enqueue_rest(sdgso_run r, const ScoreArgs &a) {
    if (r->form != SDGSO_FP64) HIPCHK(launch_finish32(a, r->stream_h));
    enqueue_select_move(r, 1, 1);
    HIPCHK(launch_advance(r->d_k.p, r->stream_h));
}

// Events around the parts of the next step (profiling).
static void step_events(sdgso_run r, std::pair<hipEvent_t, hipEvent_t> &e0,
                        std::pair<hipEvent_t, hipEvent_t> &e1,
                        std::pair<hipEvent_t, hipEvent_t> &e2) {
    if (r->lgemm) profile_pair(r, r->ev_lgemm, e0);
    profile_pair(r, r->ev_score, e1);
    profile_pair(r, r->ev_move, e2);
}

static void enqueue_step(sdgso_run r, bool timed) {
    std::pair<hipEvent_t, hipEvent_t> e0{}, e1{}, e2{};
    if (timed) step_events(r, e0, e1, e2);
    const ScoreArgs a = score_args(r);
    if (timed && r->lgemm) HIPCHK(hipEventRecord(e0.first, r->stream_h));
    enqueue_lgemm(r);
    if (timed && r->lgemm) HIPCHK(hipEventRecord(e0.second, r->stream_h));
    if (timed) HIPCHK(hipEventRecord(e1.first, r->stream_h));
    enqueue_score_launches(r, a);
    if (timed) {
        HIPCHK(hipEventRecord(e1.second, r->stream_h));
        HIPCHK(hipEventRecord(e2.first, r->stream_h));
    }
    enqueue_rest(r, a);
    if (timed) HIPCHK(hipEventRecord(e2.second, r->stream_h));
}

// Captures part of a step (0: all of it, 1: the scoring launches, 2: the
// rest, 3: the ligand displacement GEMM) as an executable graph.  The stream
// index lives on the device (advanced by the last node), so every replay is
// the same graph.
static hipGraphExec_t capture_step(sdgso_run r, int part) {
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamBeginCapture(r->stream_h, hipStreamCaptureModeThreadLocal));
    try {
        const ScoreArgs a = score_args(r);
        if (part == 0 || part == 3) enqueue_lgemm(r);
        if (part == 0 || part == 1) enqueue_score_launches(r, a);
        if (part == 0 || part == 2) enqueue_rest(r, a);
    } catch (...) {
        // leave the stream out of capture mode before reporting
        if (hipStreamEndCapture(r->stream_h, &g) == hipSuccess && g) (void)hipGraphDestroy(g);
        throw;
    }
    HIPCHK(hipStreamEndCapture(r->stream_h, &g));
    hipGraphExec_t ex = nullptr;
    const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) fail(SDGSO_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    return ex;
}

static void drop_graph(sdgso_run r) {
    for (hipGraphExec_t *g : {&r->graph, &r->graph_score, &r->graph_rest, &r->graph_lgemm}) {
        if (*g) (void)hipGraphExecDestroy(*g);
        *g = nullptr;
    }
}

static void harvest_profile(sdgso_run r) {
    for (auto &e : r->ev_lgemm) {
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, e.first, e.second));
        r->t_lgemm += ms;
        r->n_lgemm++;
        (void)hipEventDestroy(e.first);
        (void)hipEventDestroy(e.second);
    }
    for (auto &e : r->ev_score) {
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, e.first, e.second));
        r->t_score += ms;
        r->n_score++;
        (void)hipEventDestroy(e.first);
        (void)hipEventDestroy(e.second);
    }
    for (auto &e : r->ev_move) {
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, e.first, e.second));
        r->t_move += ms;
        r->n_move++;
        (void)hipEventDestroy(e.first);
        (void)hipEventDestroy(e.second);
    }
    r->ev_lgemm.clear();
    r->ev_score.clear();
    r->ev_move.clear();
}

extern "C" {

int sdgso_run_create(int n_complex, const sdgso_complex *complexes, const int *n_swarms,
                     int n_glowworms, const double *pos, const double *stream, int n_steps,
                     const sdgso_params *params, int form, double guard_band, sdgso_run *out) {
    return guarded([&] {
        if (!out) fail(SDGSO_EINVAL, "out is null");
        *out = nullptr;
        if (n_complex <= 0 || !complexes || !n_swarms) fail(SDGSO_EINVAL, "no complexes");
        if (n_glowworms <= 0 || n_glowworms > 1024)
            fail(SDGSO_EINVAL, "glowworms per swarm must be in [1, 1024]");
        if (n_steps < 0 || (n_steps > 0 && !stream)) fail(SDGSO_EINVAL, "stream missing");
        if (!pos) fail(SDGSO_EINVAL, "pos is null");
        if (form != SDGSO_FP64 && form != SDGSO_FP32_GUARD) fail(SDGSO_EINVAL, "unknown form");
        ensure_device();
        std::unique_ptr<sdgso_run_s> r(new sdgso_run_s);
        r->n_complex = n_complex;
        r->G = n_glowworms;
        r->form = form;
        r->band = (float)(guard_band > 0 ? guard_band : 1e-4);
        if (params)
            r->params = *params;
        else
            sdgso_default_params(&r->params);
        std::vector<DevComplex> descs;
        std::vector<int> scx;
        for (int k = 0; k < n_complex; k++) {
            sdgso_complex c = complexes[k];
            if (!c) fail(SDGSO_EINVAL, "null complex handle");
            if (n_swarms[k] <= 0) fail(SDGSO_EINVAL, "swarm counts must be positive");
            if (k == 0) {
                r->nmr = c->nmr;
                r->nml = c->nml;
            } else if (c->nmr != r->nmr || c->nml != r->nml) {
                fail(SDGSO_EINVAL, "all complexes of a run need the same mode counts");
            }
            r->cxs.push_back(c);
            r->n_swarms.push_back(n_swarms[k]);
            descs.push_back(c->dev);
            for (int s = 0; s < n_swarms[k]; s++) scx.push_back(k);
            r->max_nl = std::max(r->max_nl, c->nl);
            r->max_lblocks = std::max(r->max_lblocks, c->dev.nlblocks);
            if (form == SDGSO_FP32_GUARD && c->dev.ntiles > kMaxTiles32)
                fail(SDGSO_EINVAL, "receptor too large for the FP32 form (" +
                                       std::to_string(c->nr) + " atoms; at most " +
                                       std::to_string(kMaxTiles32 * kTile) + ")");
            r->S += n_swarms[k];
        }
        r->P = 7 + r->nmr + r->nml;
        const size_t N = (size_t)r->S * r->G;
        if (N > (size_t)1 << 30) fail(SDGSO_EINVAL, "too many glowworms");
        r->N = (int)N;
        r->n_steps = n_steps;
        // LDS limits of the kernels
        const size_t lds64 = score64_lds_bytes(r->max_nl);
        if (form == SDGSO_FP64 && lds64 > 160 * 1024)
            fail(SDGSO_EINVAL, "ligand too large for the FP64 form (" + std::to_string(r->max_nl) + " atoms)");
        r->dcx.upload(descs);
        r->swarm_cx.upload(scx);
        r->pos.upload(pos, N * r->P);
        r->mp.alloc(N * r->P);
        std::vector<double> init((size_t)N);
        std::fill(init.begin(), init.end(), r->params.luciferin0);
        r->luciferin.upload(init);
        std::fill(init.begin(), init.end(), 0.0);
        r->score.upload(init);
        std::fill(init.begin(), init.end(), r->params.vision0);
        r->vision.upload(init);
        std::vector<double> zeros3(3 * N, 0.0);
        r->ref.upload(zeros3);
        std::vector<int> zi(N, 0);
        r->nnei.upload(zi);
        r->moved.upload(zi);
        r->step.upload(zi);
        for (size_t i = 0; i < N; i++) zi[i] = (int)(i % r->G);
        r->select.upload(zi);
        std::vector<int64_t> z64(N, 0);
        r->pairs.upload(z64);
        if (n_steps > 0) r->stream.upload(stream, (size_t)r->S * n_steps * r->G);
        std::vector<unsigned long long> zc(4, 0);
        r->guard_count.upload(zc);
        std::vector<int> k0(1, 0);
        r->d_k.upload(k0);
        if (form == SDGSO_FP32_GUARD) {
            r->partial.alloc(N * (size_t)r->max_lblocks);
            r->partial_pairs.alloc(N * (size_t)r->max_lblocks);
            // work list: (glowworm, ligand block) pairs, blocks of one glowworm
            // adjacent; work2: (glowworm, even block), the block pairs (b, b + 1)
            std::vector<int> work, work2;
            size_t total = 0;
            for (int k = 0; k < n_complex; k++)
                total += (size_t)n_swarms[k] * r->G * r->cxs[k]->dev.nlblocks;
            if (total > (size_t)INT32_MAX / 2)
                fail(SDGSO_EINVAL, "run too large: " + std::to_string(total) +
                                       " (glowworm, ligand block) pairs");
            work.reserve(2 * total);
            work2.reserve(total + 2 * N);
            r->work0.assign(1, 0);
            r->work2_0.assign(1, 0);
            size_t g0 = 0;
            for (int k = 0; k < n_complex; k++) {
                const int nb = r->cxs[k]->dev.nlblocks;
                const size_t ng = (size_t)n_swarms[k] * r->G;
                for (size_t g = 0; g < ng; g++)
                    for (int b = 0; b < nb; b++) {
                        work.push_back((int)(g0 + g));
                        work.push_back(b);
                        if (b % 2 == 0) {
                            work2.push_back((int)(g0 + g));
                            work2.push_back(b);
                        }
                    }
                g0 += ng;
                r->work0.push_back(work.size() / 2);
                r->work2_0.push_back(work2.size() / 2);
            }
            r->work.upload(work);
            r->work.n = work.size() / 2;
            r->work2.upload(work2);
            r->work2.n = work2.size() / 2;
            r->variants.assign(n_complex, score32_default_variant());
            r->counter.alloc(1);
        }
        HIPCHK(hipStreamCreateWithFlags(&r->own_stream, hipStreamNonBlocking));
        r->stream_h = r->own_stream;
        *out = r.release();
    });
}

int sdgso_run_destroy(sdgso_run r) {
    return guarded([&] {
        if (!r) return;
        if (g_device >= 0) (void)hipSetDevice(g_device);
        if (r->stream_h) (void)hipStreamSynchronize(r->stream_h);
        drop_graph(r);
        for (auto &e : r->ev_lgemm) {
            (void)hipEventDestroy(e.first);
            (void)hipEventDestroy(e.second);
        }
        for (auto &e : r->ev_score) {
            (void)hipEventDestroy(e.first);
            (void)hipEventDestroy(e.second);
        }
        for (auto &e : r->ev_move) {
            (void)hipEventDestroy(e.first);
            (void)hipEventDestroy(e.second);
        }
        if (r->own_stream) (void)hipStreamDestroy(r->own_stream);
        delete r;
    });
}

int sdgso_run_set_stream(sdgso_run r, void *hip_stream) {
    return guarded([&] {
        if (!r) fail(SDGSO_EINVAL, "null run");
        HIPCHK(hipStreamSynchronize(r->stream_h));
        drop_graph(r);
        r->stream_h = hip_stream ? (hipStream_t)hip_stream : r->own_stream;
    });
}

int sdgso_run_steps(sdgso_run r, int n, int use_graph) {
    return guarded([&] {
        if (!r) fail(SDGSO_EINVAL, "null run");
        if (n < 0) fail(SDGSO_EINVAL, "negative step count");
        if (r->k + n > r->n_steps)
            fail(SDGSO_ESTATE, "stepping past the random stream (" + std::to_string(r->n_steps) +
                                   " steps provided)");
        ensure_device();
        if (use_graph && n > 0) {
            if (!r->profiling) {
                // one step captured once, replayed n times
                if (!r->graph) r->graph = capture_step(r, 0);
                for (int i = 0; i < n; i++) {
                    HIPCHK(hipGraphLaunch(r->graph, r->stream_h));
                    r->k++;
                }
                return;
            }
            // profiled replays: the step as two graphs (three with the
            // ligand displacement GEMM) with the timing events between them
            // (fresh events per step, read back later)
            if (r->lgemm && !r->graph_lgemm) r->graph_lgemm = capture_step(r, 3);
            if (!r->graph_score) r->graph_score = capture_step(r, 1);
            if (!r->graph_rest) r->graph_rest = capture_step(r, 2);
            for (int i = 0; i < n; i++) {
                std::pair<hipEvent_t, hipEvent_t> e0{}, e1{}, e2{};
                step_events(r, e0, e1, e2);
                if (r->lgemm) {
                    HIPCHK(hipEventRecord(e0.first, r->stream_h));
                    HIPCHK(hipGraphLaunch(r->graph_lgemm, r->stream_h));
                    HIPCHK(hipEventRecord(e0.second, r->stream_h));
                }
                HIPCHK(hipEventRecord(e1.first, r->stream_h));
                HIPCHK(hipGraphLaunch(r->graph_score, r->stream_h));
                HIPCHK(hipEventRecord(e1.second, r->stream_h));
                HIPCHK(hipEventRecord(e2.first, r->stream_h));
                HIPCHK(hipGraphLaunch(r->graph_rest, r->stream_h));
                HIPCHK(hipEventRecord(e2.second, r->stream_h));
                r->k++;
            }
            return;
        }
        for (int i = 0; i < n; i++) {
            enqueue_step(r, r->profiling);
            r->k++;
        }
    });
}

int sdgso_run_phases(sdgso_run r, int mask) {
    return guarded([&] {
        if (!r) fail(SDGSO_EINVAL, "null run");
        ensure_device();
        if ((mask & 2) && r->k >= r->n_steps) fail(SDGSO_ESTATE, "no stream left for selection");
        if (mask & 1) enqueue_score(r);
        if (mask & 6) enqueue_select_move(r, (mask & 2) ? 1 : 0, (mask & 4) ? 1 : 0);
        if (mask & 4) {
            HIPCHK(launch_advance(r->d_k.p, r->stream_h));
            r->k++;
        }
    });
}

int sdgso_run_sync(sdgso_run r) {
    return guarded([&] {
        if (!r) fail(SDGSO_EINVAL, "null run");
        HIPCHK(hipStreamSynchronize(r->stream_h));
    });
}

int sdgso_run_step_index(sdgso_run r) { return r ? r->k : -1; }

int sdgso_run_read(sdgso_run r, double *pos, double *luciferin, double *score, double *vision,
                   int *nnei, int *moved, int *select, double *ref, int64_t *pairs) {
    return guarded([&] {
        if (!r) fail(SDGSO_EINVAL, "null run");
        HIPCHK(hipStreamSynchronize(r->stream_h));
        const size_t N = r->N;
        auto get = [&](void *dst, const void *src, size_t bytes) {
            if (dst) HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
        };
        get(pos, r->pos.p, N * r->P * sizeof(double));
        get(luciferin, r->luciferin.p, N * sizeof(double));
        get(score, r->score.p, N * sizeof(double));
        get(vision, r->vision.p, N * sizeof(double));
        get(nnei, r->nnei.p, N * sizeof(int));
        get(moved, r->moved.p, N * sizeof(int));
        get(select, r->select.p, N * sizeof(int));
        get(ref, r->ref.p, 3 * N * sizeof(double));
        get(pairs, r->pairs.p, N * sizeof(int64_t));
    });
}

int sdgso_run_write(sdgso_run r, const double *pos, const double *luciferin, const double *score,
                    const double *vision, const int *moved, const double *ref) {
    return guarded([&] {
        if (!r) fail(SDGSO_EINVAL, "null run");
        HIPCHK(hipStreamSynchronize(r->stream_h));
        const size_t N = r->N;
        auto put = [&](void *dst, const void *src, size_t bytes) {
            if (src) HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
        };
        put(r->pos.p, pos, N * r->P * sizeof(double));
        put(r->luciferin.p, luciferin, N * sizeof(double));
        put(r->score.p, score, N * sizeof(double));
        put(r->vision.p, vision, N * sizeof(double));
        put(r->moved.p, moved, N * sizeof(int));
        put(r->ref.p, ref, 3 * N * sizeof(double));
    });
}

int sdgso_run_profile(sdgso_run r, int enable) {
    return guarded([&] {
        if (!r) fail(SDGSO_EINVAL, "null run");
        HIPCHK(hipStreamSynchronize(r->stream_h));
        harvest_profile(r);
        r->t_score = r->t_move = r->t_lgemm = 0;
        r->n_score = r->n_move = r->n_lgemm = 0;
        r->profiling = enable != 0;
        HIPCHK(hipMemset(r->guard_count.p, 0, sizeof(unsigned long long) * 4));
    });
}

int sdgso_run_profile_read(sdgso_run r, double *times_ms, int64_t *launches) {
    return guarded([&] {
        if (!r) fail(SDGSO_EINVAL, "null run");
        HIPCHK(hipStreamSynchronize(r->stream_h));
        harvest_profile(r);
        if (times_ms) {
            times_ms[0] = r->t_score;
            times_ms[1] = r->t_move;
            times_ms[2] = r->t_lgemm;
        }
        if (launches) {
            launches[0] = r->n_score;
            launches[1] = r->n_move;
            launches[2] = r->n_lgemm;
        }
    });
}

int sdgso_run_guard_stats(sdgso_run r, int64_t *stats) {
    return guarded([&] {
        if (!r || !stats) fail(SDGSO_EINVAL, "null argument");
        HIPCHK(hipStreamSynchronize(r->stream_h));
        unsigned long long h[4];
        HIPCHK(hipMemcpy(h, r->guard_count.p, sizeof(h), hipMemcpyDeviceToHost));
        stats[0] = (int64_t)h[0];
    });
}

static int complex_of(sdgso_run r, int gid, int *first);

int sdgso_run_set_work_order(sdgso_run r, int xcd_chunk) {
    return guarded([&] {
        if (!r) fail(SDGSO_EINVAL, "null run");
        if (xcd_chunk < 0 || xcd_chunk > (1 << 20)) fail(SDGSO_EINVAL, "xcd_chunk out of [0, 2^20]");
        HIPCHK(hipStreamSynchronize(r->stream_h));
        drop_graph(r);
        r->xcd_chunk = xcd_chunk;
    });
}

int sdgso_run_work_order(sdgso_run r, int *xcd_chunk, double *tune_ms) {
    return guarded([&] {
        if (!r) fail(SDGSO_EINVAL, "null run");
        if (xcd_chunk) *xcd_chunk = r->xcd_chunk;
        if (tune_ms) {
            tune_ms[0] = r->order_tune_ms[0];
            tune_ms[1] = r->order_tune_ms[1];
        }
    });
}

int sdgso_run_set_ligand_gemm(sdgso_run r, int enable) {
    return guarded([&] {
        if (!r) fail(SDGSO_EINVAL, "null run");
        ensure_device();
        HIPCHK(hipStreamSynchronize(r->stream_h));
        drop_graph(r);
        set_lgemm(r, enable != 0);
    });
}

int sdgso_run_ligand_gemm(sdgso_run r, int *enabled, int64_t *bytes, double *tune_ms) {
    return guarded([&] {
        if (!r) fail(SDGSO_EINVAL, "null run");
        if (enabled) *enabled = r->lgemm ? 1 : 0;
        if (bytes) *bytes = (int64_t)(r->ldisp.n * sizeof(float));
        if (tune_ms) {
            tune_ms[0] = r->lgemm_tune_ms[0];
            tune_ms[1] = r->lgemm_tune_ms[1];
        }
    });
}

int sdgso_run_ligand_displacements(sdgso_run r, int first, int count, float *out) {
    return guarded([&] {
        if (!r || !out) fail(SDGSO_EINVAL, "null argument");
        if (!r->lgemm) fail(SDGSO_ESTATE, "the run's ligand displacement GEMM is off");
        if (count <= 0) return;
        if (first < 0 || first + count > r->N) fail(SDGSO_EINVAL, "glowworm range out of the run");
        int g0 = 0;
        const int k = complex_of(r, first, &g0);
        if (complex_of(r, first + count - 1, nullptr) != k)
            fail(SDGSO_EINVAL, "glowworm range spans two complexes");
        ensure_device();
        enqueue_lgemm(r);
        HIPCHK(hipStreamSynchronize(r->stream_h));
        const sdgso_complex c = r->cxs[k];
        const int nb = c->dev.nlblocks, nl = c->nl;
        std::vector<int> perm((size_t)nb * kBlock);
        HIPCHK(hipMemcpy(perm.data(), c->lperm.p, perm.size() * sizeof(int), hipMemcpyDeviceToHost));
        long long base = 0;
        HIPCHK(hipMemcpy(&base, r->ldisp_base.p + k, sizeof(base), hipMemcpyDeviceToHost));
        const size_t per = (size_t)nb * 3 * kBlock;
        std::vector<float> h(per * count);
        HIPCHK(hipMemcpy(h.data(), r->ldisp.p + base + (long long)first * (long long)per,
                         h.size() * sizeof(float), hipMemcpyDeviceToHost));
        for (int g = 0; g < count; g++)
            for (int b = 0; b < nb; b++)
                for (int sl = 0; sl < kBlock; sl++) {
                    const int atom = perm[(size_t)b * kBlock + sl];
                    if (atom < 0 || atom >= nl) continue;
                    for (int comp = 0; comp < 3; comp++)
                        out[((size_t)g * nl + atom) * 3 + comp] =
                            h[g * per + (size_t)b * 3 * kBlock + comp * kBlock + sl];
                }
    });
}

int sdgso_tune_variants(void) { return score32_variants(); }

const char *sdgso_tune_variant_name(int v) { return score32_variant_name(v); }

int sdgso_run_tune(sdgso_run r, int repeats, double *times_ms, int *chosen) {
    return guarded([&] {
        if (!r) fail(SDGSO_EINVAL, "null run");
        if (repeats <= 0) repeats = 1;
        ensure_device();
        const int nv = score32_variants(), nc = r->n_complex;
        if (r->form != SDGSO_FP32_GUARD) {
            if (times_ms)
                for (int i = 0; i < nc * nv; i++) times_ms[i] = 0;
            if (chosen)
                for (int k = 0; k < nc; k++) chosen[k] = 0;
            return;
        }
        HIPCHK(hipStreamSynchronize(r->stream_h));
        // score the current poses of each complex with every variant; luciferin
        // and step are left alone (the next step rescores these poses anyway)
        ScoreArgs a = score_args(r);
        a.update_luciferin = 0;
        enqueue_lgemm(r);  // the ligand displacements of these poses, for every launch below
        hipEvent_t e0, e1;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        for (int k = 0; k < nc; k++) {
            ScoreArgs part = a;
            double best = 1e300;
            int best_v = 0;
            for (int v = 0; v < nv; v++) {
                set_work(r, v, k, k + 1, part);
                HIPCHK(launch_score32(part, v, r->stream_h));  // warm-up
                HIPCHK(hipEventRecord(e0, r->stream_h));
                for (int i = 0; i < repeats; i++) HIPCHK(launch_score32(part, v, r->stream_h));
                HIPCHK(hipEventRecord(e1, r->stream_h));
                HIPCHK(hipEventSynchronize(e1));
                float ms = 0;
                HIPCHK(hipEventElapsedTime(&ms, e0, e1));
                const double t = ms / repeats;
                if (times_ms) times_ms[(size_t)k * nv + v] = t;
                if (t < best) {
                    best = t;
                    best_v = v;
                }
            }
            r->variants[k] = best_v;
            if (chosen) chosen[k] = best_v;
        }
        // the work order: a scoring pass of the chosen variants in list order
        // and one XCD-chunked; the faster stays
        {
            double t[2] = {0, 0};
            for (int pass = 0; pass < 2; pass++) {
                r->xcd_chunk = pass ? kXcdChunk : 0;
                ScoreArgs b = score_args(r);
                b.update_luciferin = 0;
                for (int i = 0; i <= repeats; i++) {  // i = 0: warm-up
                    if (i == 1) HIPCHK(hipEventRecord(e0, r->stream_h));
                    launch_score32_run(r, b);
                }
                HIPCHK(hipEventRecord(e1, r->stream_h));
                HIPCHK(hipEventSynchronize(e1));
                float ms = 0;
                HIPCHK(hipEventElapsedTime(&ms, e0, e1));
                t[pass] = ms / repeats;
            }
            r->order_tune_ms[0] = t[0];
            r->order_tune_ms[1] = t[1];
            r->xcd_chunk = t[1] < t[0] ? kXcdChunk : 0;
        }
        // the ligand displacement GEMM: a scoring pass of the chosen variants
        // with it (the GEMM included) and one without it; the faster stays
        r->lgemm_tune_ms[0] = r->lgemm_tune_ms[1] = 0;
        if (r->nml > 0) {
            const bool was = r->lgemm;
            bool can = true;
            try {
                if (!was) set_lgemm(r, true);
            } catch (const Failure &f) {
                // its output does not fit in device memory: it stays off
                if (f.code != SDGSO_ENOMEM) throw;
                set_lgemm(r, false);
                can = false;
            }
            if (can) {
                double t[2] = {0, 0};
                for (int pass = 0; pass < 2; pass++) {
                    ScoreArgs b = score_args(r);
                    b.update_luciferin = 0;
                    if (pass == 1) b.ldisp = nullptr;
                    for (int i = 0; i <= repeats; i++) {  // i = 0: warm-up
                        if (i == 1) HIPCHK(hipEventRecord(e0, r->stream_h));
                        if (pass == 0) enqueue_lgemm(r);
                        launch_score32_run(r, b);
                    }
                    HIPCHK(hipEventRecord(e1, r->stream_h));
                    HIPCHK(hipEventSynchronize(e1));
                    float ms = 0;
                    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
                    t[pass] = ms / repeats;
                }
                r->lgemm_tune_ms[0] = t[0];
                r->lgemm_tune_ms[1] = t[1];
                if (t[1] <= t[0]) set_lgemm(r, false);
            }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        HIPCHK(launch_finish32(a, r->stream_h));
        HIPCHK(hipStreamSynchronize(r->stream_h));
        drop_graph(r);
    });
}

int sdgso_run_set_variant(sdgso_run r, int variant) {
    return guarded([&] {
        if (!r) fail(SDGSO_EINVAL, "null run");
        if (variant < 0 || variant >= score32_variants())
            fail(SDGSO_EINVAL, "variant out of [0, " + std::to_string(score32_variants()) + ")");
        HIPCHK(hipStreamSynchronize(r->stream_h));
        drop_graph(r);
        std::fill(r->variants.begin(), r->variants.end(), variant);
    });
}

int sdgso_run_variants(sdgso_run r, int *variants) {
    return guarded([&] {
        if (!r || !variants) fail(SDGSO_EINVAL, "null argument");
        for (int k = 0; k < r->n_complex; k++)
            variants[k] = r->form == SDGSO_FP32_GUARD ? r->variants[k] : 0;
    });
}

int sdgso_run_filter_stats(sdgso_run r, int64_t *stats) {
    return guarded([&] {
        if (!r || !stats) fail(SDGSO_EINVAL, "null argument");
        if (r->form != SDGSO_FP32_GUARD) fail(SDGSO_EINVAL, "filter statistics need the FP32 form");
        ensure_device();
        HIPCHK(hipStreamSynchronize(r->stream_h));
        DevBuf<unsigned long long> st;
        std::vector<unsigned long long> h(6 * kStatSlots, 0);
        st.upload(h);
        ScoreArgs a = score_args(r);
        a.update_luciferin = 0;
        a.guard_count = nullptr;
        a.stats = st.p;
        // the per-block partials are scratch: no epilogue, the run state stays as it was
        enqueue_lgemm(r);
        launch_score32_run(r, a);
        HIPCHK(hipStreamSynchronize(r->stream_h));
        HIPCHK(hipMemcpy(h.data(), st.p, h.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
        for (int i = 0; i < 6; i++) {
            unsigned long long sum = 0;
            for (int k = 0; k < kStatSlots; k++) sum += h[(size_t)i * kStatSlots + k];
            stats[i] = (int64_t)sum;
        }
    });
}

int sdgso_probe_gather(int64_t table_floats, double *rate) {
    return guarded([&] {
        if (table_floats <= 0 || table_floats > (1 << 30) || !rate)
            fail(SDGSO_EINVAL, "bad probe arguments");
        ensure_device();
        DevBuf<float> table, out;
        std::vector<float> ones((size_t)table_floats, 1.0f);
        table.upload(ones);
        int dev = 0, cus = 0;
        HIPCHK(hipGetDevice(&dev));
        HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        const int blocks = cus * 8, per_thread = 2048;
        out.alloc((size_t)blocks * 256);
        hipStream_t s = nullptr;
        HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        hipEvent_t e0, e1;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        HIPCHK(launch_gather_probe(table.p, (unsigned)table_floats, per_thread, out.p, blocks, s));
        HIPCHK(hipEventRecord(e0, s));
        HIPCHK(launch_gather_probe(table.p, (unsigned)table_floats, per_thread, out.p, blocks, s));
        HIPCHK(hipEventRecord(e1, s));
        HIPCHK(hipEventSynchronize(e1));
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        (void)hipStreamDestroy(s);
        *rate = (double)blocks * 256 * per_thread / (ms * 1e-3);
    });
}

int sdgso_probe_anm_gemm(sdgso_run r, double *result) {
    return guarded([&] {
        if (!r || !result) fail(SDGSO_EINVAL, "null argument");
        if (r->form != SDGSO_FP32_GUARD) fail(SDGSO_EINVAL, "the probe needs a run in the FP32 form");
        if (r->nmr <= 0) fail(SDGSO_EINVAL, "the probe needs receptor modes");
        ensure_device();
        HIPCHK(hipStreamSynchronize(r->stream_h));
        const int G = r->G, P = r->P;
        std::vector<double> pos((size_t)r->N * P);
        HIPCHK(hipMemcpy(pos.data(), r->pos.p, pos.size() * sizeof(double), hipMemcpyDeviceToHost));
        // per swarm: the tiles any of its glowworms can reach (translation
        // spread + the ligand's extent + its and the receptor's mode
        // displacement bounds + 15 A + the tile's radius)
        std::vector<int> pairs;  // (swarm, complex, tile)
        int s0 = 0;
        for (int k = 0; k < r->n_complex; k++) {
            const sdgso_complex c = r->cxs[k];
            for (int s = s0; s < s0 + r->n_swarms[k]; s++) {
                double cen[3] = {0, 0, 0}, spread = 0, dl = 0, dr = 0;
                for (int g = 0; g < G; g++)
                    for (int d = 0; d < 3; d++) cen[d] += pos[((size_t)s * G + g) * P + d] / G;
                for (int g = 0; g < G; g++) {
                    const double *q = pos.data() + ((size_t)s * G + g) * P;
                    double t2 = 0, er = 0, el = 0;
                    for (int d = 0; d < 3; d++) t2 += (q[d] - cen[d]) * (q[d] - cen[d]);
                    for (int m = 0; m < c->nmr; m++) er += std::fabs(q[7 + m]) * c->dev.rmode_max[m];
                    for (int m = 0; m < c->nml; m++) el += std::fabs(q[7 + c->nmr + m]) * c->dev.lmode_max[m];
                    spread = std::max(spread, std::sqrt(t2));
                    dr = std::max(dr, er);
                    dl = std::max(dl, el);
                }
                const double reach = spread + c->lig_rmax + dl + 15.0 + dr + 0.01;
                for (int t = 0; t < (int)c->h_tiles.size(); t++) {
                    const float4 tl = c->h_tiles[t];
                    const double dx = tl.x - cen[0], dy = tl.y - cen[1], dz = tl.z - cen[2];
                    const double lim = reach + tl.w;
                    if (dx * dx + dy * dy + dz * dz <= lim * lim) {
                        pairs.push_back(s);
                        pairs.push_back(k);
                        pairs.push_back(t);
                    }
                }
            }
            s0 += r->n_swarms[k];
        }
        const size_t npairs = pairs.size() / 3;
        const int chunks = (G + 31) / 32;
        // D of a (swarm, tile, 32-glowworm chunk): 32 x 192 floats; written to
        // a buffer of at most 4 GiB that later launches overwrite
        const size_t per_item = (size_t)32 * 192 * sizeof(float);
        const size_t items = npairs * chunks;
        const size_t cap = std::max<size_t>(1, std::min<size_t>(items, ((size_t)4 << 30) / per_item));
        DevBuf<int> dpairs;
        dpairs.upload(pairs);
        DevBuf<float> out;
        out.alloc(cap * 32 * 192);
        hipEvent_t e0, e1;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        // warm-up on the first items, then the timed pass over all of them
        HIPCHK(launch_anm_gemm_probe(r->dcx.p, r->pos.p, P, G, dpairs.p, 0, std::min(items, cap),
                                     chunks, out.p, r->stream_h));
        HIPCHK(hipEventRecord(e0, r->stream_h));
        for (size_t i0 = 0; i0 < items; i0 += cap)
            HIPCHK(launch_anm_gemm_probe(r->dcx.p, r->pos.p, P, G, dpairs.p, i0, std::min(cap, items - i0),
                                         chunks, out.p, r->stream_h));
        HIPCHK(hipEventRecord(e1, r->stream_h));
        HIPCHK(hipEventSynchronize(e1));
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        const int kpad = (r->nmr + 1) / 2 * 2;
        result[0] = ms;
        result[1] = (double)items * per_item;                      // bytes written
        result[2] = 2.0 * kpad * 192.0 * 32.0 * (double)items;     // MFMA flops issued
        result[3] = (double)npairs * 64.0 * G;                     // displacements computed
        result[4] = (double)npairs;                                // (swarm, tile) pairs
    });
}

// ---------------------------------------------------------------- output side
// The complex that owns glowworm `gid` of a run, and its first glowworm.
static int complex_of(sdgso_run r, int gid, int *first) {
    int g0 = 0;
    for (int k = 0; k < r->n_complex; k++) {
        const int ng = r->n_swarms[k] * r->G;
        if (gid < g0 + ng) {
            if (first) *first = g0;
            return k;
        }
        g0 += ng;
    }
    fail(SDGSO_EINVAL, "glowworm index out of range");
}

int sdgso_run_final_coords(sdgso_run r, int first, int count, double *rec_xyz, double *lig_xyz) {
    return guarded([&] {
        if (!r) fail(SDGSO_EINVAL, "null run");
        if (count <= 0) return;
        if (first < 0 || first + count > r->N) fail(SDGSO_EINVAL, "glowworm range out of the run");
        int g0 = 0;
        const int k = complex_of(r, first, &g0);
        if (complex_of(r, first + count - 1, nullptr) != k)
            fail(SDGSO_EINVAL, "glowworm range spans two complexes");
        ensure_device();
        HIPCHK(hipStreamSynchronize(r->stream_h));
        const int nr = r->cxs[k]->nr, nl = r->cxs[k]->nl, n = nr + nl;
        const int chunk = (int)std::max<size_t>(1, std::min<size_t>(count, (256u << 20) / (24 * (size_t)n)));
        DevBuf<double> out;
        out.alloc((size_t)chunk * n * 3);
        std::vector<double> h((size_t)chunk * n * 3);
        for (int c0 = 0; c0 < count; c0 += chunk) {
            const int cn = std::min(chunk, count - c0);
            CoordsArgs a{};
            a.cx = r->dcx.p + k;
            a.pos = r->pos.p;
            a.P = r->P;
            a.first = first + c0;
            a.n_rec = nr;
            a.n_lig = nl;
            a.out = out.p;
            HIPCHK(launch_coords(a, cn, r->stream_h));
            HIPCHK(hipStreamSynchronize(r->stream_h));
            HIPCHK(hipMemcpy(h.data(), out.p, (size_t)cn * n * 3 * sizeof(double), hipMemcpyDeviceToHost));
            for (int g = 0; g < cn; g++) {
                const double *src = h.data() + (size_t)g * n * 3;
                if (rec_xyz)
                    std::memcpy(rec_xyz + ((size_t)(c0 + g) * nr) * 3, src, sizeof(double) * 3 * nr);
                if (lig_xyz)
                    std::memcpy(lig_xyz + ((size_t)(c0 + g) * nl) * 3, src + 3 * (size_t)nr,
                                sizeof(double) * 3 * nl);
            }
        }
    });
}

int sdgso_run_cluster(sdgso_run r, int complex, const int *rec_backbone, int n_rec_backbone,
                      const int *lig_backbone, int n_lig_backbone, double cutoff, int *cluster,
                      int *representative) {
    return guarded([&] {
        if (!r) fail(SDGSO_EINVAL, "null run");
        if (complex < 0 || complex >= r->n_complex) fail(SDGSO_EINVAL, "complex index out of range");
        const sdgso_complex cx = r->cxs[complex];
        const int nb = n_rec_backbone + n_lig_backbone;
        if (n_rec_backbone < 0 || n_lig_backbone < 0 || nb <= 0)
            fail(SDGSO_EINVAL, "no backbone atoms");
        if ((n_rec_backbone && !rec_backbone) || (n_lig_backbone && !lig_backbone))
            fail(SDGSO_EINVAL, "null backbone index array");
        for (int i = 0; i < n_rec_backbone; i++)
            if (rec_backbone[i] < 0 || rec_backbone[i] >= cx->nr) fail(SDGSO_EINVAL, "receptor backbone index out of range");
        for (int i = 0; i < n_lig_backbone; i++)
            if (lig_backbone[i] < 0 || lig_backbone[i] >= cx->nl) fail(SDGSO_EINVAL, "ligand backbone index out of range");
        if (!cluster) fail(SDGSO_EINVAL, "null output");
        ensure_device();
        HIPCHK(hipStreamSynchronize(r->stream_h));
        const int G = r->G, S = r->n_swarms[complex];
        int g0 = 0;
        for (int k = 0; k < complex; k++) g0 += r->n_swarms[k] * G;
        std::vector<double> score((size_t)S * G);
        HIPCHK(hipMemcpy(score.data(), r->score.p + g0, score.size() * sizeof(double), hipMemcpyDeviceToHost));
        DevBuf<int> ridx, lidx;
        if (n_rec_backbone) ridx.upload(rec_backbone, n_rec_backbone);
        if (n_lig_backbone) lidx.upload(lig_backbone, n_lig_backbone);
        // swarms per chunk: coordinates and RMSD matrices within a fixed budget
        const size_t per_swarm = std::max((size_t)G * nb * 24, (size_t)G * G * 16);
        const int chunk = (int)std::max<size_t>(1, std::min<size_t>(S, (512u << 20) / per_swarm));
        DevBuf<double> bb, rm;
        bb.alloc((size_t)chunk * G * nb * 3);
        rm.alloc((size_t)chunk * G * G);
        std::vector<double> h((size_t)chunk * G * G);
        std::vector<int> order(G);
        for (int s0 = 0; s0 < S; s0 += chunk) {
            const int cs = std::min(chunk, S - s0);
            CoordsArgs a{};
            a.cx = r->dcx.p + complex;
            a.pos = r->pos.p;
            a.P = r->P;
            a.first = g0 + s0 * G;
            a.rec_idx = ridx.p;
            a.n_rec = n_rec_backbone;
            a.lig_idx = lidx.p;
            a.n_lig = n_lig_backbone;
            a.out = bb.p;
            HIPCHK(launch_coords(a, cs * G, r->stream_h));
            HIPCHK(launch_rmsd(bb.p, nb, G, cs, rm.p, r->stream_h));
            HIPCHK(hipStreamSynchronize(r->stream_h));
            HIPCHK(hipMemcpy(h.data(), rm.p, (size_t)cs * G * G * sizeof(double), hipMemcpyDeviceToHost));
            for (int s = 0; s < cs; s++) {
                const double *M = h.data() + (size_t)s * G * G;
                const double *sc = score.data() + (size_t)(s0 + s) * G;
                int *cl = cluster + (size_t)(s0 + s) * G;
                int *rep = representative ? representative + (size_t)(s0 + s) * G : nullptr;
                // decreasing score, ties by index (prepare_gso_tasks.cc:1273-1275, 987)
                for (int g = 0; g < G; g++) order[g] = g;
                std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return sc[x] > sc[y]; });
                // greedy: join the first cluster whose representative is within the cutoff
                // (prepare_gso_tasks.cc:1004-1080)
                std::vector<int> reps;
                for (int g : order) {
                    int c = 0;
                    for (; c < (int)reps.size(); c++)
                        if (M[(size_t)reps[c] * G + g] <= cutoff) break;
                    if (c == (int)reps.size()) reps.push_back(g);
                    cl[g] = c;
                    if (rep) rep[g] = reps[c];
                }
            }
        }
    });
}

// ---------------------------------------------------------------- scoring plugin
int sdgso_score_poses(sdgso_complex c, int n_poses, const double *poses, int form,
                      double guard_band, double *out_score, double *out_ref, int64_t *out_pairs) {
    return guarded([&] {
        if (!c) fail(SDGSO_EINVAL, "null complex");
        if (n_poses < 0 || (n_poses > 0 && (!poses || !out_score)))
            fail(SDGSO_EINVAL, "bad pose buffers");
        if (n_poses == 0) return;
        // a run of n_poses single-glowworm swarms, scored once without
        // luciferin; the scoring set-up sums the ligand modes itself (no
        // displacement GEMM), as in sdgso_score_poses_noalloc
        sdgso_run r = nullptr;
        sdgso_complex cs[1] = {c};
        int ns[1] = {n_poses};
        int rc = sdgso_run_create(1, cs, ns, 1, poses, nullptr, 0, nullptr, form, guard_band, &r);
        if (rc != SDGSO_OK) throw Failure{rc};
        std::unique_ptr<sdgso_run_s, int (*)(sdgso_run)> hold(r, sdgso_run_destroy);
        ScoreArgs a = score_args(r);
        a.update_luciferin = 0;
        if (form == SDGSO_FP64) {
            HIPCHK(launch_score64(a, r->max_nl, r->stream_h));
        } else {
            launch_score32_run(r, a);
            HIPCHK(launch_finish32(a, r->stream_h));
        }
        rc = sdgso_run_read(r, nullptr, nullptr, out_score, nullptr, nullptr, nullptr, nullptr,
                            out_ref, out_pairs);
        if (rc != SDGSO_OK) throw Failure{rc};
    });
}

// ---------------------------------------------------------------- host utilities
int sdgso_stream(uint32_t seed, int n_swarms, int n_steps, int n_glowworms, double *u) {
    return guarded([&] {
        if (n_swarms < 0 || n_steps < 0 || n_glowworms < 0) fail(SDGSO_EINVAL, "negative size");
        if (!u && (size_t)n_swarms * n_steps * n_glowworms) fail(SDGSO_EINVAL, "u is null");
        // prepare_gso_tasks.cc:409-427: one engine, drawn serially, swarm-major
        std::mt19937 gen(seed);
        std::uniform_real_distribution<double> dist(0.0, 1.0);
        const size_t n = (size_t)n_swarms * n_steps * n_glowworms;
        for (size_t i = 0; i < n; i++) u[i] = dist(gen);
    });
}

}  // extern "C"
