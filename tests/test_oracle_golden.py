"""The CPU oracle against golden vectors written by the reference's own
compiled functions (tests/golden/make_golden.py).  Bit-exact throughout."""
import ctypes
import os

import numpy as np
import pytest

import oracle_lib as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "golden_ref.npz")
D = ctypes.POINTER(ctypes.c_double)
I = ctypes.POINTER(ctypes.c_int)


def p(a, t=D):
    return a.ctypes.data_as(t)


@pytest.fixture(scope="module")
def g():
    return dict(np.load(GOLD))


def full_table(lo, blk):
    t = np.zeros((168, 168, 20))
    K = blk.shape[0]
    t[lo:lo + K, lo:lo + K] = blk
    return O.pad_table(t.reshape(-1))


def test_rotate(g):
    L = O.lib()
    for k, q in enumerate(g["rot_q"]):
        x = g["rot_pts"].copy()
        L.or_rotate(p(x), p(np.ascontiguousarray(q)), len(x))
        assert np.array_equal(x, g["rot_out"][k]), k


def test_dfire(g):
    L = O.lib()
    T = full_table(int(g["df_lo"]), g["df_block"])
    ro = lo = 0
    for k, (nr, nl) in enumerate(zip(g["df_nr"], g["df_nl"])):
        r = np.ascontiguousarray(g["df_r"][ro:ro + nr])
        rt = np.ascontiguousarray(g["df_rt"][ro:ro + nr].astype(np.int32))
        l = np.ascontiguousarray(g["df_l"][lo:lo + nl])
        lt = np.ascontiguousarray(g["df_lt"][lo:lo + nl].astype(np.int32))
        e = L.or_dfire_energy(p(r), int(nr), p(rt, I), p(l), int(nl), p(lt, I), p(T), None)
        assert L.or_score_from_energy(e) == g["df_score"][k], k
        ro += nr
        lo += nl


def test_slerp(g):
    L = O.lib()
    for k in range(len(g["sl_a"])):
        a = g["sl_a"][k].copy()
        b = g["sl_b"][k].copy()
        L.or_slerp(p(a), p(b), float(g["sl_t"][k]))
        assert np.array_equal(a, g["sl_out"][k]), k
        assert np.array_equal(b, g["sl_other"][k]), k


def test_move_extent(g):
    L = O.lib()
    off = 0
    for n in g["me_n"]:
        s = np.ascontiguousarray(g["me_self"][off:off + n].copy())
        o = np.ascontiguousarray(g["me_other"][off:off + n].copy())
        L.or_move_extent(p(s) if n else None, p(o) if n else None, int(n), 0.5)
        assert np.array_equal(s, g["me_out"][off:off + n])
        off += n


def test_probabilities_and_select(g):
    L = O.lib()
    off = 0
    for k in range(len(g["pr_self"])):
        n = int(g["pr_nnei"][k])
        nei = np.ascontiguousarray(g["pr_nei"][off:off + n].astype(np.int32))
        prob = np.zeros(max(n, 1))
        lu = np.ascontiguousarray(g["pr_lucif"][k])
        me = int(g["pr_self"][k])
        L.or_probabilities(p(prob), p(nei, I) if n else None, p(lu), n, float(lu[me]))
        assert np.array_equal(prob[:n], g["pr_prob"][off:off + n]), k
        sel = L.or_select(float(g["pr_u"][k]), p(nei, I) if n else None, p(prob), n, me)
        assert sel == g["pr_sel"][k], k
        off += n


def test_three_reference_steps(g):
    """Whole GSO steps: the reference's cal_gso_tasks_cpu ran its 3 compiled steps
    (10+10 modes, masked atoms, dense swarms so that glowworms move)."""
    S, G = int(g["run_S"]), int(g["run_G"])
    cx = O.Complex(g["run_r_xyz"], g["run_r_type"], g["run_l_xyz"], g["run_l_type"],
                   full_table(int(g["run_lo"]), g["run_block"])[:168 * 168 * 20], g["run_lig_ref"],
                   g["run_r_modes"], g["run_l_modes"], g["run_r_mask"], g["run_l_mask"])
    u = O.stream(0, S, 3, G)
    for rescore_all in (False, True):
        st = O.run(cx, O.default_params(), g["run_pos_in"], u, 3, S, G, rescore_all=rescore_all)
        out = st.get()
        assert np.array_equal(out["pos"], g["run_pos_out"])
    assert np.any(g["run_pos_out"] != g["run_pos_in"])
