"""Write the golden vectors that pin the CPU oracle, from the reference's own
functions compiled into oracle/_ref (run `make -C oracle ref` first; needs the
reference sources, so this runs only where they are mounted).

Output: tests/golden/golden_ref.npz: inputs and reference outputs of
  rotate                (prepare_gso_tasks.cc:277-338)
  cdfire_calculate_dfire (prepare_gso_tasks.cc:1522-1615), incl. bin boundaries
  slerp                 (prepare_gso_tasks.cc:1370-1421)
  move_extent           (prepare_gso_tasks.cc:1352-1368)
  compute_probability_moving_from_neighbors + select_random_neighbor
                        (prepare_gso_tasks.cc:1424-1488)
  cal_gso_tasks_cpu     (prepare_gso_tasks.cc:359-822), its 3 compiled steps
Tables are stored as one K x K x 20 block of types [lo, lo+K); the rest of the
168 x 168 x 20 table is zero.

Usage: python tests/golden/make_golden.py
"""
import ctypes
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_lib as O  # noqa: E402
import synth_np as SN  # noqa: E402

D = ctypes.POINTER(ctypes.c_double)
I = ctypes.POINTER(ctypes.c_int)


def p(a, t=D):
    return a.ctypes.data_as(t)


def block_table(rng, lo, K):
    blk = (rng.normal(size=(K, K, 20)) * 0.5).astype(np.float32)
    return blk


def expand_table(lo, blk):
    K = blk.shape[0]
    t = np.zeros((168, 168, 20), np.float32)
    t[lo:lo + K, lo:lo + K, :] = blk
    return t.reshape(-1)


def main():
    if not O.ref_available():
        sys.exit("oracle/_ref/libref_shim.so missing: run `make -C oracle ref`")
    R = O.ref()
    rng = np.random.default_rng(20261203)
    out = {}

    # ---- rotate: 48 points x 24 quaternions (unit, non-unit, negative parts)
    pts = rng.normal(size=(48, 3)) * 20.0
    qs = rng.normal(size=(24, 4))
    qs[:12] /= np.linalg.norm(qs[:12], axis=1, keepdims=True)
    qs[12:18] *= 3.7
    qs[18] = [1, 0, 0, 0]
    qs[19] = [0, 0, 0, 1]
    qs[20] = [np.cos(np.pi / 4), 0, 0, np.sin(np.pi / 4)]
    rot = np.zeros((24, 48, 3))
    for k in range(24):
        x = pts.copy()
        R.ref_rotate(p(x), p(np.ascontiguousarray(qs[k])), 48)
        rot[k] = x
    out.update(rot_pts=pts, rot_q=qs, rot_out=rot)

    # ---- dfire: random clouds and boundary placements
    lo, K = 150, 18
    blk = block_table(rng, lo, K)
    T = expand_table(lo, blk).astype(np.float64)
    cases_r, cases_l, cases_rt, cases_lt, energies = [], [], [], [], []
    for c in range(8):
        nr, nl = int(rng.integers(30, 90)), int(rng.integers(20, 60))
        r = rng.normal(size=(nr, 3)) * 6.0
        l = rng.normal(size=(nl, 3)) * 6.0 + rng.normal(size=3) * 4.0
        rt = rng.integers(lo, lo + K, nr).astype(np.int32)
        lt = rng.integers(lo, lo + K - 1, nl).astype(np.int32)  # keep (167,167) out of bin 20
        cases_r.append(r)
        cases_l.append(l)
        cases_rt.append(rt)
        cases_lt.append(lt)
    # boundary case: ligand atoms on the x axis at sqrt(k^2/4) and one ulp either side
    bounds = np.array([k * k / 4.0 for k in range(1, 31)])
    xs = []
    for b in bounds:
        x = np.sqrt(b)
        xs += [np.nextafter(x, 0), x, np.nextafter(x, 100)]
    xs += [14.999999999, 15.0, 15.000000001, 0.1, 0.49999999]
    l = np.zeros((len(xs), 3))
    l[:, 0] = xs
    r = np.zeros((1, 3))
    cases_r.append(r)
    cases_l.append(l)
    cases_rt.append(np.array([lo + 1], np.int32))
    cases_lt.append(rng.integers(lo, lo + K - 1, len(xs)).astype(np.int32))
    # diagonal placements (three non-zero components) near boundaries
    dirs = rng.normal(size=(40, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    rad = np.sqrt(bounds[rng.integers(0, 30, 40)]) * (1 + rng.normal(size=40) * 1e-15)
    l = dirs * rad[:, None] + np.array([1.5, -2.25, 0.75])
    cases_r.append(np.array([[1.5, -2.25, 0.75]]))
    cases_l.append(l)
    cases_rt.append(np.array([lo + 3], np.int32))
    cases_lt.append(rng.integers(lo, lo + K - 1, 40).astype(np.int32))
    for r, l, rt, lt in zip(cases_r, cases_l, cases_rt, cases_lt):
        r = np.ascontiguousarray(r)
        l = np.ascontiguousarray(l)
        energies.append(R.ref_dfire_score(p(T), p(rt, I), p(lt, I), p(r), p(l), len(r), len(l)))
    out.update(
        df_lo=lo, df_block=blk,
        df_nr=np.array([len(r) for r in cases_r]), df_nl=np.array([len(l) for l in cases_l]),
        df_r=np.concatenate(cases_r), df_l=np.concatenate(cases_l),
        df_rt=np.concatenate(cases_rt), df_lt=np.concatenate(cases_lt),
        df_score=np.array(energies))

    # ---- slerp
    n = 40
    a = rng.normal(size=(n, 4))
    b = rng.normal(size=(n, 4))
    b[:8] = a[:8] + rng.normal(size=(8, 4)) * 1e-3   # dot > 0.9995 branch
    b[8:16] = -a[8:16] + rng.normal(size=(8, 4)) * 0.2  # negative dot
    a[16:20] *= 5.0
    sl_out = np.zeros((n, 4))
    sl_other = np.zeros((n, 4))
    ts = np.where(np.arange(n) < 30, 0.5, rng.random(n))
    for k in range(n):
        x = a[k].copy()
        y = b[k].copy()
        R.ref_slerp(p(x), p(y), float(ts[k]))
        sl_out[k] = x
        sl_other[k] = y
    out.update(sl_a=a, sl_b=b, sl_t=ts, sl_out=sl_out, sl_other=sl_other)

    # ---- move_extent
    me_n = np.array([0, 1, 3, 10, 10, 20, 20, 10])
    me_self, me_other, me_out = [], [], []
    for k, m in enumerate(me_n):
        s_ = rng.uniform(-1, 1, m)
        o_ = rng.uniform(-1, 1, m)
        if k == 4:
            o_ = s_ + 1e-10  # below the 1e-8 threshold
        x = np.ascontiguousarray(s_.copy())
        R.ref_move_extent(p(x), p(np.ascontiguousarray(o_.copy())), int(m), 0.5)
        me_self.append(s_)
        me_other.append(o_)
        me_out.append(x)
    out.update(me_n=me_n, me_self=np.concatenate(me_self), me_other=np.concatenate(me_other),
               me_out=np.concatenate(me_out))

    # ---- probabilities + roulette
    G = 30
    ncase = 60
    pr_lucif = rng.normal(size=(ncase, G)) * 3 + 10
    pr_self = rng.integers(0, G, ncase)
    pr_u = rng.random(ncase)
    pr_u[0] = 2.0 ** -60  # u == 0 reads neighbors[-1] in the reference (undefined)
    pr_u[1] = 1.0 - 2 ** -53
    pr_nnei, pr_nei, pr_prob, pr_sel = [], [], [], []
    for k in range(ncase):
        g = pr_self[k]
        cand = [n_ for n_ in range(G) if n_ != g and pr_lucif[k, g] < pr_lucif[k, n_]]
        m = int(rng.integers(0, min(len(cand), 8) + 1)) if cand else 0
        nei = np.array(sorted(rng.choice(cand, m, replace=False)) if m else [], np.int32)
        if k == 2 and len(cand) >= 2:  # exact cumulative tie: u equals the first probability
            nei = np.array(cand[:2], np.int32)
        prob = np.zeros(max(len(nei), 1))
        R.ref_probabilities(p(prob), p(nei, I) if len(nei) else None,
                            p(np.ascontiguousarray(pr_lucif[k])), len(nei), float(pr_lucif[k, g]))
        if k == 2 and len(nei) == 2:
            pr_u[k] = prob[0]
        sel = R.ref_select(float(pr_u[k]), p(nei, I) if len(nei) else None, p(prob), len(nei),
                           int(g), G)
        pr_nnei.append(len(nei))
        pr_nei.append(nei)
        pr_prob.append(prob[:len(nei)])
        pr_sel.append(sel)
    out.update(pr_G=G, pr_lucif=pr_lucif, pr_self=pr_self, pr_u=pr_u, pr_nnei=np.array(pr_nnei),
               pr_nei=np.concatenate(pr_nei), pr_prob=np.concatenate(pr_prob),
               pr_sel=np.array(pr_sel))

    # ---- cal_gso_tasks_cpu: 3 compiled steps, 10+10 modes, dense swarms
    S, G3 = 3, 36
    case = SN.make_case(7, 160, 90, 10, 10, S, G3, mask_holes=True, swarm_radius=1.2)
    lo3, K3 = 0, 24
    blk3 = block_table(rng, lo3, K3)
    case["r_type"] = rng.integers(lo3, lo3 + K3, 160).astype(np.int32)
    case["l_type"] = rng.integers(lo3, lo3 + K3, 90).astype(np.int32)
    cx = O.Complex(case["r_xyz"], case["r_type"], case["l_xyz"], case["l_type"],
                   expand_table(lo3, blk3), case["lig_ref"], case["r_modes"], case["l_modes"],
                   case["r_mask"], case["l_mask"])
    with tempfile.TemporaryDirectory() as d:
        pos3, _ = O.ref_cpu_run(cx, case["pos"], S, G3, d)
    out.update(run_S=S, run_G=G3, run_lo=lo3, run_block=blk3, run_pos_in=case["pos"],
               run_pos_out=pos3, **{"run_" + k: case[k] for k in
                                    ("r_xyz", "l_xyz", "r_type", "l_type", "r_modes", "l_modes",
                                     "r_mask", "l_mask", "lig_ref")})
    moved = int(np.any(pos3 != case["pos"], axis=2).sum())
    path = os.path.join(HERE, "golden_ref.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes; run3 glowworms moved:", moved)


if __name__ == "__main__":
    main()
