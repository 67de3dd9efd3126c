// sdgso_score_poses_noalloc: the scoring plugin call (sdgso_score_poses) for
// callers that cannot allocate device memory, such as code under HIP graph
// capture or a tight serving loop.  Poses and outputs are the caller's device
// buffers and the launches go on the caller's stream; the scratch the
// kernels need (the complex index of each pose, the FP32 form's work list of
// (pose, ligand block) items and its per-item partial energies and pair
// counts, and the reference points when the caller does not want them) lives
// in one static device array of kScratchBytes per GPU, which the library
// holds until the process ends.
//
// Every call on a GPU uses that GPU's one array: calls that overlap on a GPU
// (two host threads, or two streams) race on it.  The array keeps the last
// call's contents; each call writes every scratch value it reads (the setup
// kernel writes the complex indices and the work list, the scoring kernel
// every partial that the epilogue adds), so nothing needs clearing.
#include <algorithm>
#include <mutex>
#include <string>

#include "dev_common.h"

namespace sd {

constexpr size_t kScratchBytes = (size_t)SDGSO_NOALLOC_SCRATCH_MB << 20;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Device address of `scratch` on each device, looked up once per device with
// that device current (the array's name is only a host-side stand-in).
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mutex;
static void *g_scratch[kMaxDevices];

static int scratch_address(int dev, unsigned char **out) {
    if (dev < 0 || dev >= kMaxDevices) {
        set_error("device index out of range for the scratch table");
        return SDGSO_EHIP;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
            return SDGSO_EHIP;
        }
        g_scratch[dev] = p;
    }
    *out = (unsigned char *)g_scratch[dev];
    return SDGSO_OK;
}

// Complex index 0 for every pose, and the work list: item i is (pose
// i / per_pose, block (i % per_pose) * stride), stride 2 for a two-block
// variant (its waves take blocks b and b + 1).
__global__ void noalloc_setup_kernel(int *swarm_cx, int n_poses, int *work, long long n_work,
                                     int per_pose, int stride) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n_poses) swarm_cx[i] = 0;
    if (i < n_work) {
        work[2 * i] = (int)(i / per_pose);
        work[2 * i + 1] = (int)(i % per_pose) * stride;
    }
}

static size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

}  // namespace sd

using namespace sd;

extern "C" int sdgso_score_poses_noalloc(sdgso_complex c, int n_poses, const double *poses,
                                         int form, double guard_band, double *out_score,
                                         double *out_ref, int64_t *out_pairs, void *hip_stream) {
    ComplexView v{};
    if (!complex_view(c, &v)) {
        set_error("null complex");
        return SDGSO_EINVAL;
    }
    if (n_poses < 0 || (n_poses > 0 && (!poses || !out_score))) {
        set_error("bad pose buffers");
        return SDGSO_EINVAL;
    }
    if (form != SDGSO_FP64 && form != SDGSO_FP32_GUARD) {
        set_error("unknown form");
        return SDGSO_EINVAL;
    }
    if (n_poses == 0) return SDGSO_OK;
    // the limits sdgso_run_create applies to the same complex
    if (form == SDGSO_FP64 && score64_lds_bytes(v.nl) > 160 * 1024) {
        set_error("ligand too large for the FP64 form (" + std::to_string(v.nl) + " atoms)");
        return SDGSO_EINVAL;
    }
    if (form == SDGSO_FP32_GUARD && v.ntiles > kMaxTiles32) {
        set_error("receptor too large for the FP32 form (" + std::to_string(v.nr) + " atoms)");
        return SDGSO_EINVAL;
    }
    // scratch layout: complex index per pose | reference points (when the
    // caller passes none) | FP32 form: work list, partial energies, partial
    // pair counts; each region 256-byte aligned
    const int variant = score32_default_variant();
    const int per_pose = score32_variant_blocks(variant) == 2 ? (v.nlblocks + 1) / 2 : v.nlblocks;
    const long long n_work = form == SDGSO_FP32_GUARD ? (long long)n_poses * per_pose : 0;
    const size_t o_cx = 0;
    const size_t o_ref = o_cx + align256(sizeof(int) * (size_t)n_poses);
    const size_t o_work = o_ref + (out_ref ? 0 : align256(3 * sizeof(double) * (size_t)n_poses));
    const size_t o_part = o_work + align256(2 * sizeof(int) * (size_t)n_work);
    const size_t n_part = form == SDGSO_FP32_GUARD ? (size_t)n_poses * v.nlblocks : 0;
    const size_t o_pp = o_part + align256(sizeof(double) * n_part);
    const size_t need = o_pp + align256(sizeof(long long) * n_part);
    if (need > kScratchBytes || n_work > (long long)INT32_MAX) {
        const size_t per = need / (size_t)n_poses + 1;
        set_error("sdgso_score_poses_noalloc: " + std::to_string(n_poses) + " poses of this complex need " +
                  std::to_string(need) + " bytes of scratch; the limit is " +
                  std::to_string(kScratchBytes) + " (SDGSO_NOALLOC_SCRATCH_MB MiB): pass at most " +
                  std::to_string(kScratchBytes / per) + " poses per call");
        return SDGSO_EINVAL;
    }
    int rc = use_device();
    if (rc != SDGSO_OK) return rc;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        set_error(std::string("hipGetDevice: ") + hipGetErrorString(e));
        return SDGSO_EHIP;
    }
    unsigned char *base = nullptr;
    rc = scratch_address(dev, &base);
    if (rc != SDGSO_OK) return rc;
    hipStream_t s = (hipStream_t)hip_stream;

    int *swarm_cx = (int *)(base + o_cx);
    int *work = (int *)(base + o_work);
    const long long n_items = std::max<long long>(n_poses, n_work);
    hipLaunchKernelGGL(noalloc_setup_kernel, dim3((unsigned)((n_items + 255) / 256)), dim3(256), 0,
                       s, swarm_cx, n_poses, work, n_work, per_pose,
                       score32_variant_blocks(variant));
    e = hipGetLastError();

    ScoreArgs a{};
    a.cx = v.dev;
    a.swarm_cx = swarm_cx;
    a.G = 1;
    a.P = 7 + v.nmr + v.nml;
    a.n_glow = n_poses;
    a.pos = poses;
    a.score = out_score;
    a.ref = out_ref ? out_ref : (double *)(base + o_ref);
    a.pairs = out_pairs;
    a.update_luciferin = 0;
    a.band = (float)(guard_band > 0 ? guard_band : 1e-4);
    a.nmr = v.nmr;
    if (e == hipSuccess) {
        if (form == SDGSO_FP64) {
            e = launch_score64(a, v.nl, s);
        } else {
            a.partial = (double *)(base + o_part);
            a.partial_pairs = (long long *)(base + o_pp);
            a.max_lblocks = v.nlblocks;
            a.work = work;
            a.n_work = (int)n_work;
            e = launch_score32(a, variant, s);
            if (e == hipSuccess) e = launch_finish32(a, s);
        }
    }
    if (e != hipSuccess) {
        set_error(std::string("sdgso_score_poses_noalloc launch: ") + hipGetErrorString(e));
        return SDGSO_EHIP;
    }
    return SDGSO_OK;
}
