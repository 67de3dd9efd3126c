"""sdgso_score_poses_noalloc (the scoring call that allocates nothing) through
the C-ABI, against sdgso_score_poses on the same poses: scores, reference
points and pair counts bit-identical (the same kernels on the same inputs),
in both forms, over two calls in a row on one GPU, the second smaller, so
that it runs on the scratch the first one left behind.

Device buffers come from the HIP runtime libsdgso.so itself is linked
against (looked up by its soname once the library is loaded)."""
import ctypes

import numpy as np
import pytest

import gpu_cases as GC
from sparkledock_amd import sdgso

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def hip():
    sdgso.lib()
    H = ctypes.CDLL("libamdhip64.so.7")
    H.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    H.hipFree.argtypes = [ctypes.c_void_p]
    H.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    H.hipDeviceSynchronize.argtypes = []
    return H


class Dev:
    """A device copy of a numpy array."""

    def __init__(self, H, host):
        self.H, self.host = H, np.ascontiguousarray(host)
        self.p = ctypes.c_void_p()
        assert H.hipMalloc(ctypes.byref(self.p), max(1, self.host.nbytes)) == 0
        assert H.hipMemcpy(self.p, self.host.ctypes.data, self.host.nbytes, H2D) == 0

    def get(self):
        out = np.empty_like(self.host)
        assert self.H.hipMemcpy(out.ctypes.data, self.p, out.nbytes, D2H) == 0
        return out

    def free(self):
        self.H.hipFree(self.p)


def noalloc(cx, poses, form, with_ref=True, with_pairs=True):
    H = hip()
    n = len(poses)
    p = Dev(H, poses)
    # outputs start as NaN / -1 so that a value the call does not write shows
    s = Dev(H, np.full(n, np.nan))
    ref = Dev(H, np.full((n, 3), np.nan))
    pairs = Dev(H, np.full(n, -1, np.int64))
    try:
        rc = sdgso.lib().sdgso_score_poses_noalloc(cx.h, n, p.p, form, 1e-4, s.p,
                                                    ref.p if with_ref else None,
                                                    pairs.p if with_pairs else None, None)
        assert H.hipDeviceSynchronize() == 0
        if rc != 0:
            raise sdgso.SdgsoError(rc, sdgso.lib().sdgso_last_error().decode())
        return s.get(), ref.get(), pairs.get()
    finally:
        for b in (p, s, ref, pairs):
            b.free()


@pytest.mark.parametrize("form", [sdgso.FP64, sdgso.FP32_GUARD])
def test_noalloc_equals_score_poses(form):
    c = GC.case(51, 700, 300, 10, 10, 4, 24)
    pr = GC.Pair(c)
    poses = c["pos"].reshape(-1, pr.P)
    for sub in (poses, poses[37:57]):        # 96 poses, then 20 others
        s0, ref0, p0 = pr.gpu.score(sub, form)
        s, ref, p = noalloc(pr.gpu, sub, form)
        assert p0.min() > 0
        assert np.array_equal(p, p0) and np.array_equal(s, s0) and np.array_equal(ref, ref0)
    # optional outputs left out: the scores are unchanged
    s, ref, p = noalloc(pr.gpu, poses[:10], form, with_ref=False, with_pairs=False)
    assert np.array_equal(s, pr.gpu.score(poses[:10], form)[0])
    assert np.all(np.isnan(ref)) and np.all(p == -1)


def test_noalloc_rejects_calls_beyond_the_scratch():
    """A call whose scratch would pass the static array's size is refused with
    a message, before anything runs; the buffers are real, of the full size."""
    c = GC.case(52, 200, 300, 0, 0, 1, 8)
    pr = GC.Pair(c)
    n = 600000   # 5 ligand blocks: 132 bytes of scratch per pose, 79 MB > 64 MiB
    poses = np.tile(c["pos"].reshape(-1, pr.P), (n // 8, 1))
    with pytest.raises(sdgso.SdgsoError) as e:
        noalloc(pr.gpu, poses, sdgso.FP32_GUARD, with_ref=False)
    assert e.value.code == 1 and "scratch" in str(e.value) and "at most" in str(e.value)
    # the FP64 form needs 28 bytes per pose: the same call fits
    s, _, p = noalloc(pr.gpu, poses[:40000], sdgso.FP64, with_ref=False)
    s0, _, p0 = pr.gpu.score(poses[:8], sdgso.FP64)
    assert np.array_equal(s[:8], s0) and np.array_equal(p[:8], p0)
