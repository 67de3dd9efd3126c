"""The ligand displacement GEMM of an FP32 run (v_mfma_f32_32x32x2f32 over
every glowworm) through the C-ABI.

- Its FP32 displacements against the exact FP64 sum of the same modes and
  extents: within (nml + 2) 2^-24 of the sum of |e_m M_m| per coordinate,
  the bound of a sum of nml FP32 products (modes and extents rounded to FP32)
  in any order of correctly rounded steps; the scoring kernels' guard band
  assumes twice that.
- Runs with it on and off scored against the oracle (pair counts equal,
  energies within the FP32 bound of the parity tests), with two complexes of
  different ligand block counts and 37 glowworms per swarm, so that the
  GEMM's 32-glowworm tiles span swarms and end inside a complex.
- The per-phase timing reports it, and switching it off removes it.
"""
import numpy as np
import pytest

import gpu_cases as GC
import oracle_lib as O
from sparkledock_amd import sdgso

pytestmark = pytest.mark.gpu

U = 2.0 ** -24
REL_E = 1e-9
FLOOR_E = 1e-12


@pytest.mark.parametrize("nml,ext", [(1, 1.0), (2, 1.0), (7, 3.0), (10, 1.0), (11, 5.0),
                                     (20, 1.0), (32, 25.0)])
def test_displacements_within_the_fp32_bound(nml, ext):
    S, G, nmr = 3, 37, 3
    c = GC.case(200 + nml, 300, 260, nmr, nml, S, G)
    c["pos"][..., 7 + nmr:] *= ext
    pr = GC.Pair(c)
    run = sdgso.Run(pr.gpu, S, G, c["pos"], None, 0, form=sdgso.FP32_GUARD)
    assert run.ligand_gemm() == (False, 0)  # off until switched on or tuned on
    run.set_ligand_gemm(True)
    on, nbytes = run.ligand_gemm()
    assert on and nbytes == S * G * 5 * 3 * 64 * 4  # 260 atoms: 5 blocks of 64 slots
    D = run.ligand_displacements(0, S * G).astype(np.float64)
    e = c["pos"].reshape(S * G, -1)[:, 7 + nmr:]
    M = c["l_modes"] * c["l_mask"][None, :, None]  # the mask zeroes a row's modes
    exact = np.einsum("gm,mac->gac", e, M)
    mag = np.einsum("gm,mac->gac", np.abs(e), np.abs(M))
    err = np.abs(D - exact)
    bound = (nml + 2) * U * mag * (1 + 1e-6)
    assert np.all(err <= bound), float(np.max(err / np.maximum(bound, 1e-300)))
    assert np.all(D[:, c["l_mask"] == 0] == 0.0)
    # a sub-range of one complex reads the same values
    assert np.array_equal(run.ligand_displacements(40, 5), D[40:45].astype(np.float32))
    run.close()


@pytest.mark.parametrize("nml", [0, 3, 10, 20])
def test_scores_with_and_without_the_gemm(nml):
    cs = [GC.case(300 + nml, 600, 150, 10, nml, 2, 37),
          GC.case(400 + nml, 500, 330, 10, nml, 3, 37)]
    prs = [GC.Pair(c) for c in cs]
    pos = np.concatenate([c["pos"].reshape(-1) for c in cs])
    run = sdgso.Run([p.gpu for p in prs], [2, 3], 37, pos, None, 0, form=sdgso.FP32_GUARD)
    run.set_ligand_gemm(True)
    assert run.ligand_gemm()[0] == (nml > 0)
    run.phases(1)
    a = run.read()
    run.set_ligand_gemm(False)
    assert run.ligand_gemm() == (False, 0)
    run.phases(1)
    b = run.read()
    run.close()
    assert np.array_equal(a["pairs"], b["pairs"])
    g0 = 0
    for pr, c in zip(prs, cs):
        poses = c["pos"].reshape(-1, pr.P)
        n = len(poses)
        s0, ref0, p0 = pr.oracle_scores(poses)
        assert np.count_nonzero(p0) > n // 2
        assert np.array_equal(a["pairs"][g0:g0 + n], p0)
        assert np.array_equal(a["ref"][g0:g0 + n], ref0)
        denom = np.maximum(GC.sum_abs_terms(pr, poses), 1.0)
        e0 = (4.7 - s0) / 0.0157
        for st in (a, b):
            e = (4.7 - st["score"][g0:g0 + n]) / 0.0157
            assert np.all(np.abs(e - e0) <= REL_E * np.abs(e0) + FLOOR_E * denom)
        g0 += n


def test_profile_times_the_gemm_and_off_is_off():
    S, G, steps = 2, 40, 6
    c = GC.case(77, 400, 200, 4, 6, S, G)
    pr = GC.Pair(c)
    run = sdgso.Run(pr.gpu, S, G, c["pos"], O.stream(0, S, steps, G), steps,
                    form=sdgso.FP32_GUARD)
    run.set_ligand_gemm(True)
    run.profile(True)
    run.steps(2, graph=True)
    run.steps(1)
    p = run.profile_read()
    assert p["lgemm_launches"] == 3 and p["score_launches"] == 3 and p["lgemm_ms"] > 0
    run.set_ligand_gemm(False)
    run.profile(True)
    run.steps(2, graph=True)
    p = run.profile_read()
    assert p["lgemm_launches"] == 0 and p["lgemm_ms"] == 0 and p["score_launches"] == 2
    with pytest.raises(sdgso.SdgsoError) as e:
        run.ligand_displacements(0, 1)
    assert e.value.code == 4
    run.set_ligand_gemm(True)
    assert run.ligand_gemm()[0]
    run.steps(1)
    run.close()


def test_fp64_form_and_no_ligand_modes_take_no_gemm():
    c = GC.case(78, 300, 120, 5, 4, 1, 8)
    pr = GC.Pair(c)
    run = sdgso.Run(pr.gpu, 1, 8, c["pos"], None, 0, form=sdgso.FP64)
    assert run.ligand_gemm() == (False, 0)
    run.set_ligand_gemm(True)  # no effect on the FP64 form
    assert run.ligand_gemm() == (False, 0)
    run.close()
    c = GC.case(79, 300, 120, 5, 0, 1, 8)
    pr = GC.Pair(c)
    run = sdgso.Run(pr.gpu, 1, 8, c["pos"], None, 0, form=sdgso.FP32_GUARD)
    run.set_ligand_gemm(True)
    assert run.ligand_gemm() == (False, 0)
    run.close()


def score32_gemm(cx, poses, variant):
    """FP32 scores and pair counts of `poses` with one kernel variant and the
    ligand displacement GEMM on: single-glowworm swarms, scoring phase only."""
    run = sdgso.Run(cx, len(poses), 1, poses, None, 0, form=sdgso.FP32_GUARD)
    run.set_variant(variant)
    run.set_ligand_gemm(True)
    assert run.ligand_gemm()[0]
    run.phases(1)
    st = run.read()
    run.close()
    return st["score"], st["pairs"]


@pytest.mark.parametrize("nmr,nml", [(10, 10), (3, 20), (0, 32), (21, 1), (16, 11)])
def test_every_variant_with_the_gemm_against_oracle(nmr, nml):
    c = GC.case(500 + nml, 600, 280, nmr, nml, 3, 12)
    pr = GC.Pair(c)
    poses = c["pos"].reshape(-1, pr.P)
    s0, _, p0 = pr.oracle_scores(poses)
    denom = np.maximum(GC.sum_abs_terms(pr, poses), 1.0)
    e0 = (4.7 - s0) / 0.0157
    assert p0.sum() > 0
    for v in range(len(sdgso.tune_variants())):
        s, p = score32_gemm(pr.gpu, poses, v)
        assert np.array_equal(p, p0), v
        e = (4.7 - s) / 0.0157
        assert np.all(np.abs(e - e0) <= REL_E * np.abs(e0) + FLOOR_E * denom), v


def test_bins_exact_with_large_ligand_displacements_from_the_gemm():
    """Ligand atoms that ten modes with extents of tens of A move onto every
    bin boundary +- delta: per-bin pair counts through one-hot tables equal
    the oracle's with the GEMM's displacements, for every variant."""
    from test_gpu_parity import large_ligand_displacement_case
    c = large_ligand_displacement_case()
    pr = GC.Pair(c)
    poses = c["pos"].reshape(-1, pr.P)
    want = []
    for b in range(20):
        t = np.zeros((168, 168, 20), np.float32)
        t[:, :, b] = 1.0
        want.append(round((4.7 - pr.with_table(t.reshape(-1)).score_pose(poses[0])[0]) / 0.0157))
    want = np.array(want)
    assert want.sum() > 1000
    for v in range(len(sdgso.tune_variants())):
        got = []
        for b in range(20):
            t = np.zeros((168, 168, 20), np.float32)
            t[:, :, b] = 1.0
            cx = pr.gpu_with_table(t.reshape(-1))
            s, _ = score32_gemm(cx, poses, v)
            got.append(int(np.rint((4.7 - s[0]) / 0.0157)))
            cx.close()
        assert np.array_equal(np.array(got), want), v


def test_tuner_keeps_the_faster_form():
    S, G = 4, 50
    c = GC.case(81, 900, 700, 10, 10, S, G)
    pr = GC.Pair(c)
    run = sdgso.Run(pr.gpu, S, G, c["pos"], None, 0, form=sdgso.FP32_GUARD)
    assert run.ligand_gemm_tune_ms() == (0.0, 0.0)
    run.tune()
    t_on, t_off = run.ligand_gemm_tune_ms()
    assert t_on > 0 and t_off > 0
    assert run.ligand_gemm()[0] == (t_on < t_off)
    run.phases(1)
    p = run.read()["pairs"]
    assert np.array_equal(p, pr.oracle_scores(c["pos"].reshape(-1, pr.P))[2])
    run.close()


def test_gemm_that_does_not_fit_is_refused_and_leaves_no_error():
    """A run whose ligand displacement output would exceed the GPU's memory
    (4 M glowworms x 100 ligand blocks x 768 B = 307 GB): switching the GEMM
    on fails with SDGSO_ENOMEM and leaves it off; the run still scores, and
    a later run on the same thread launches without a stale runtime error."""
    rng = np.random.default_rng(93)
    nr, nl, S, G = 64, 6400, 4000, 1000
    r = rng.normal(size=(nr, 3)) * 4.0
    lg = rng.normal(size=(nl, 3)) * 12.0
    table = (rng.normal(size=168 * 168 * 20) * 0.5).astype(np.float32)
    lmodes = rng.normal(size=(1, nl, 3)) * 0.01
    cx = sdgso.Complex(r, rng.integers(0, 168, nr).astype(np.int32), lg,
                       rng.integers(0, 168, nl).astype(np.int32), table, np.zeros(3),
                       None, lmodes)
    pose = np.array([30.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 0.5])
    run = sdgso.Run(cx, S, G, np.tile(pose, S * G), None, 0, form=sdgso.FP32_GUARD)
    with pytest.raises(sdgso.SdgsoError) as e:
        run.set_ligand_gemm(True)
    assert e.value.code == 3
    assert run.ligand_gemm() == (False, 0)
    run.close()
    cx.close()
    c = GC.case(94, 300, 150, 2, 3, 2, 8)
    pr = GC.Pair(c)
    poses = c["pos"].reshape(-1, pr.P)
    small = sdgso.Run(pr.gpu, 16, 1, poses, None, 0, form=sdgso.FP32_GUARD)
    small.set_ligand_gemm(True)
    small.phases(1)
    assert np.array_equal(small.read()["pairs"], pr.oracle_scores(poses)[2])
    small.close()
