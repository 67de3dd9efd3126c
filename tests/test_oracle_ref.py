"""The CPU oracle against the reference's own compiled functions, live, on
fresh random inputs (only where oracle/_ref was built: `make -C oracle ref`
needs the reference sources)."""
import ctypes
import tempfile

import numpy as np
import pytest

import oracle_lib as O
import synth_np as SN

pytestmark = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built")

D = ctypes.POINTER(ctypes.c_double)
I = ctypes.POINTER(ctypes.c_int)


def p(a, t=D):
    return a.ctypes.data_as(t)


def test_dfire_random_full_type_range():
    R, L = O.ref(), O.lib()
    rng = np.random.default_rng(99)
    tab = O.pad_table((rng.normal(size=168 * 168 * 20) * 0.5).astype(np.float32))
    for _ in range(10):
        nr, nl = rng.integers(50, 300), rng.integers(30, 200)
        r = np.ascontiguousarray(rng.normal(size=(nr, 3)) * 9)
        l = np.ascontiguousarray(rng.normal(size=(nl, 3)) * 9 + 5)
        rt = rng.integers(0, 168, nr).astype(np.int32)
        lt = rng.integers(0, 167, nl).astype(np.int32)
        want = R.ref_dfire_score(p(tab), p(rt, I), p(lt, I), p(r), p(l), int(nr), int(nl))
        e = L.or_dfire_energy(p(r), int(nr), p(rt, I), p(l), int(nl), p(lt, I), p(tab), None)
        assert L.or_score_from_energy(e) == want


def test_rotate_random():
    R, L = O.ref(), O.lib()
    rng = np.random.default_rng(7)
    for _ in range(50):
        q = np.ascontiguousarray(rng.normal(size=4) * rng.uniform(0.1, 5))
        x = np.ascontiguousarray(rng.normal(size=(30, 3)) * 40)
        a, b = x.copy(), x.copy()
        R.ref_rotate(p(a), p(q), 30)
        L.or_rotate(p(b), p(q), 30)
        assert np.array_equal(a, b)


@pytest.mark.parametrize("seed,radius", [(21, 1.0), (22, 1.5), (23, 3.0)])
def test_three_steps_of_cal_gso_tasks_cpu(seed, radius):
    S, G = 3, 40
    c = SN.make_case(seed, 220, 110, 10, 10, S, G, mask_holes=True, swarm_radius=radius)
    cx = O.Complex(c["r_xyz"], c["r_type"], c["l_xyz"], c["l_type"], c["table"], c["lig_ref"],
                   c["r_modes"], c["l_modes"], c["r_mask"], c["l_mask"])
    with tempfile.TemporaryDirectory() as d:
        want, _ = O.ref_cpu_run(cx, c["pos"], S, G, d)
    st = O.run(cx, O.default_params(), c["pos"], O.stream(0, S, 3, G), 3, S, G, rescore_all=False)
    assert np.array_equal(st.get()["pos"], want)
