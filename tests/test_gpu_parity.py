"""HIP kernels against the CPU oracle, through the C-ABI.

Bars: FP64 form bit-exact (scores, pair counts, reference points); FP32 guard
form with identical DFIRE bins (per-bin pair counts equal) and energy within
1e-9 of the oracle relative to |E| plus 1e-12 of the sum of |T| over in-cutoff
pairs (FP64 lane sums of the FP32 table values keep the relative accuracy of
a sum whose terms cancel; the floor covers E near 0); GSO
choices (neighbour counts, selections, moved flags) bit-exact from identical
states; poses equal up to the slerp's acos/sin (GPU libm vs glibc): 1e-14.
"""
import numpy as np
import pytest

import gpu_cases as GC
import oracle_lib as O
from sparkledock_amd import sdgso

pytestmark = pytest.mark.gpu

REL_E = 1e-9      # FP32 guard energy, relative to |E| ...
FLOOR_E = 1e-12   # ... plus this much of sum |T| over in-cutoff pairs
POS_TOL = 1e-13


def assert_energy(e, e0, denom, variant):
    """FP32 guard energies `e` against the oracle's `e0`; `denom` is sum |T|."""
    err = np.abs(e - e0)
    assert np.all(err <= REL_E * np.abs(e0) + FLOOR_E * np.maximum(denom, 1.0)), \
        (variant, float(np.max(err / np.maximum(np.abs(e0), 1e-300))))


def poses_of(case):
    return case["pos"].reshape(-1, case["pos"].shape[-1])


VARIANTS = list(range(len(sdgso.tune_variants())))


def score32(cx, poses, variant):
    """FP32 guard-band scores and pair counts of `poses` with one kernel
    variant: a run of single-glowworm swarms, scoring phase only."""
    n = len(poses)
    run = sdgso.Run(cx, n, 1, poses, None, 0, form=sdgso.FP32_GUARD)
    run.set_variant(variant)
    run.phases(1)
    st = run.read()
    run.close()
    return st["score"], st["pairs"]


@pytest.mark.parametrize("seed,nmr,nml", [(1, 10, 10), (2, 0, 0), (3, 4, 7)])
def test_score_fp64_bit_exact(seed, nmr, nml):
    c = GC.case(seed, 500, 260, nmr, nml, 4, 16)
    pr = GC.Pair(c)
    poses = poses_of(c)
    s, ref, pairs = pr.gpu.score(poses, sdgso.FP64)
    s0, ref0, pairs0 = pr.oracle_scores(poses)
    assert pairs0.min() > 0
    assert np.array_equal(pairs, pairs0)
    assert np.array_equal(ref, ref0)
    assert np.array_equal(s, s0)


@pytest.mark.parametrize("seed,nmr,nml", [(4, 10, 10), (5, 0, 0), (6, 20, 20), (9, 24, 6), (10, 32, 1),
                                          (11, 1, 0), (12, 8, 3), (13, 11, 2), (14, 16, 5),
                                          (15, 21, 0), (16, 9, 9), (17, 2, 1), (18, 3, 4),
                                          (19, 5, 2), (22, 6, 6), (23, 7, 7)])
def test_score_fp32_guard(seed, nmr, nml):
    """Receptor mode counts cover every mode-line width of the broadcast
    variant's cooperative displacement, 8 float4 (up to 10 modes), 16 (up to
    21) and 32 (up to 32), and none, and the three orders of a slot's modes:
    packed by coordinate for 1 to 10 modes (a float4 with modes of two
    coordinates for 3, 5, 6, 7, 9 and 10), component-major with 3, 4 and 5
    float4 per coordinate (11, 16, 20 modes) and mode-major (21, 24, 32)."""
    c = GC.case(seed, 700, 300, nmr, nml, 4, 16)
    pr = GC.Pair(c)
    poses = poses_of(c)
    s, ref, pairs = pr.gpu.score(poses, sdgso.FP32_GUARD)
    s0, ref0, pairs0 = pr.oracle_scores(poses)
    assert np.array_equal(pairs, pairs0)
    assert np.array_equal(ref, ref0)
    denom = GC.sum_abs_terms(pr, poses)
    e0 = (4.7 - s0) / 0.0157
    for v in VARIANTS:
        if v:
            s, pairs = score32(pr.gpu, poses, v)
            assert np.array_equal(pairs, pairs0), v
        e = (4.7 - s) / 0.0157
        assert_energy(e, e0, denom, v)


@pytest.mark.parametrize("nmr", list(range(0, 33)))
def test_score_fp32_every_receptor_mode_count(nmr):
    """Every receptor mode count the kernels take (0 to 32, each with its
    record width and order: packed by coordinate up to 10, component-major
    11 to 20, mode-major above) and a ligand mode count that varies with it,
    every variant against the oracle: pair counts equal, energies within the
    FP32 bound."""
    nml = (7 * nmr + 3) % 33
    c = GC.case(100 + nmr, 500, 200, nmr, nml, 2, 8)
    pr = GC.Pair(c)
    poses = poses_of(c)
    s0, _, pairs0 = pr.oracle_scores(poses)
    denom = GC.sum_abs_terms(pr, poses)
    e0 = (4.7 - s0) / 0.0157
    assert pairs0.sum() > 0
    for v in VARIANTS:
        s, pairs = score32(pr.gpu, poses, v)
        assert np.array_equal(pairs, pairs0), (nmr, nml, v)
        assert_energy((4.7 - s) / 0.0157, e0, denom, v)


def per_bin_counts(pr, poses, form, variant=None):
    """Pair count of every bin through one-hot tables (exact integer sums);
    `variant` picks an FP32 kernel variant (default: the plugin call)."""
    out = []
    for b in range(20):
        t = np.zeros((168, 168, 20), np.float32)
        t[:, :, b] = 1.0
        cx = pr.gpu_with_table(t.reshape(-1))
        if variant is None:
            s, _, _ = cx.score(poses, form)
        else:
            s, _ = score32(cx, poses, variant)
        out.append(np.rint((4.7 - s) / 0.0157))
        cx.close()
    return np.array(out)


@pytest.mark.parametrize("nl", [40, 64, 128, 129, 200, 256])
def test_ligand_block_counts(nl):
    """Ligands of one to four 64-atom blocks, full and partial: the two-block
    variants pair blocks (0, 1), (2, 3) and leave a last odd block alone;
    every variant matches the oracle's pair counts and energies."""
    c = GC.case(20 + nl, 600, nl, 6, 4, 3, 8)
    pr = GC.Pair(c)
    poses = poses_of(c)
    s0, _, pairs0 = pr.oracle_scores(poses)
    assert pairs0.sum() > 0
    denom = GC.sum_abs_terms(pr, poses)
    e0 = (4.7 - s0) / 0.0157
    for v in VARIANTS:
        s, pairs = score32(pr.gpu, poses, v)
        assert np.array_equal(pairs, pairs0), v
        e = (4.7 - s) / 0.0157
        assert_energy(e, e0, denom, v)


def test_energy_with_terms_of_both_signs_over_eight_decades():
    """Table values of both signs spread over eight orders of magnitude: the
    large terms swamp the small ones in a plain or Kahan FP32 sum (errors of
    about 2^-24 of the running sum).  The FP64 lane sums still hold every
    variant's energy to the oracle's FP64 sum within 1e-9 |E| + 1e-12 sum |T|."""
    c = GC.case(31, 700, 300, 6, 4, 3, 16)
    rng = np.random.default_rng(31)
    n = len(c["table"])
    c["table"] = (rng.normal(size=n) * 10.0 ** rng.uniform(-4.0, 4.0, n)).astype(np.float32)
    pr = GC.Pair(c)
    poses = poses_of(c)
    s0, _, pairs0 = pr.oracle_scores(poses)
    assert pairs0.min() > 0
    denom = GC.sum_abs_terms(pr, poses)
    e0 = (4.7 - s0) / 0.0157
    for v in VARIANTS:
        s, pairs = score32(pr.gpu, poses, v)
        assert np.array_equal(pairs, pairs0), v
        assert_energy((4.7 - s) / 0.0157, e0, denom, v)


def test_fp64_table_values():
    """The table enters in FP64, as the reference's FastDifire holds it
    (fastdfire.h:93): with values that FP32 cannot represent (decimals of four
    places, as a DCparams text file gives them), the FP64 form is bit-exact to
    the oracle.  The FP32 form gathers the values rounded to FP32: its energy
    is within its summation tolerance of the oracle on the rounded table, and
    within 2^-24 sum |T| more of the oracle on the FP64 table."""
    c = GC.case(33, 700, 300, 10, 10, 3, 16)
    rng = np.random.default_rng(33)
    c["table"] = np.round(rng.normal(size=168 * 168 * 20) * 0.5, 4)
    t32 = c["table"].astype(np.float32)
    assert np.mean(t32.astype(np.float64) != c["table"]) > 0.99
    pr = GC.Pair(c)
    poses = poses_of(c)
    s0, ref0, pairs0 = pr.oracle_scores(poses)
    s32tab, _, _ = pr.oracle_scores(poses, table=t32)
    assert pairs0.min() > 0 and not np.array_equal(s0, s32tab)
    s, ref, pairs = pr.gpu.score(poses, sdgso.FP64)
    assert np.array_equal(pairs, pairs0)
    assert np.array_equal(ref, ref0)
    assert np.array_equal(s, s0)
    denom = GC.sum_abs_terms(pr, poses)
    e0, e0_32 = (4.7 - s0) / 0.0157, (4.7 - s32tab) / 0.0157
    for v in VARIANTS:
        s, pairs = score32(pr.gpu, poses, v)
        assert np.array_equal(pairs, pairs0), v
        e = (4.7 - s) / 0.0157
        assert_energy(e, e0_32, denom, v)
        assert np.all(np.abs(e - e0) <= REL_E * np.abs(e0) + (FLOOR_E + 2.0 ** -24) * denom), v


def test_bins_exact_on_lattice_boundaries():
    """Lattice geometry puts many pairs exactly on k^2/4 boundaries, on the
    225 A^2 cutoff and at d2 = 0: both forms must bin them as the oracle."""
    c = GC.grid_case(7)
    pr = GC.Pair(c)
    poses = poses_of(c)
    want = []
    for b in range(20):
        t = np.zeros((168, 168, 20), np.float32)
        t[:, :, b] = 1.0
        cx = pr.with_table(t.reshape(-1))
        want.append([round((4.7 - cx.score_pose(q)[0]) / 0.0157) for q in poses])
    want = np.array(want)
    for form in (sdgso.FP64, sdgso.FP32_GUARD):
        got = per_bin_counts(pr, poses, form)
        assert np.array_equal(got, want), form
    for v in VARIANTS[1:]:
        assert np.array_equal(per_bin_counts(pr, poses, None, v), want), v
    # the full table too (FP64: bit-exact; FP32: same pairs)
    s0, _, p0 = pr.oracle_scores(poses)
    s, _, p = pr.gpu.score(poses, sdgso.FP64)
    assert np.array_equal(s, s0) and np.array_equal(p, p0)
    s, _, p = pr.gpu.score(poses, sdgso.FP32_GUARD)
    assert np.array_equal(p, p0)


def boundary_sweep_case(seed=17):
    """One receptor atom at the origin and ligand atoms at d2 = k^2/4 + delta
    for every boundary k and offsets delta across the guard band, including
    the window between the band edge and 225 + band rounded to FP32."""
    rng = np.random.default_rng(seed)
    deltas = np.concatenate([np.linspace(-2.5e-4, 2.5e-4, 21), np.linspace(0.85e-4, 1.25e-4, 17),
                             -np.linspace(0.85e-4, 1.25e-4, 17), [0.0, 1e-9, -1e-9]])
    d2 = np.array([k * k / 4.0 + d for k in range(2, 31) for d in deltas])
    d2 = np.concatenate([d2, 225.0 + np.linspace(0.5e-4, 1.5e-4, 101)])
    n = len(d2)
    dirs = rng.normal(size=(n, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    c = dict(r_xyz=np.array([[0.0, 0.0, 0.0], [400.0, 0.0, 0.0]]),
             l_xyz=dirs * np.sqrt(d2)[:, None],
             r_type=np.array([5, 9], np.int32), l_type=rng.integers(0, 167, n).astype(np.int32),
             table=(rng.normal(size=168 * 168 * 20) * 0.5).astype(np.float32),
             r_modes=np.zeros((0, 2, 3)), l_modes=np.zeros((0, n, 3)),
             r_mask=np.ones(2, np.int32), l_mask=np.ones(n, np.int32),
             lig_ref=np.zeros(3))
    c["pos"] = np.array([[[0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0]]])
    return c


def test_fp32_guard_sweep_across_every_boundary():
    c = boundary_sweep_case()
    pr = GC.Pair(c)
    poses = poses_of(c)
    s0, _, p0 = pr.oracle_scores(poses)
    for form in (sdgso.FP64, sdgso.FP32_GUARD):
        s, _, p = pr.gpu.score(poses, form)
        assert np.array_equal(p, p0), form
        if form == sdgso.FP64:
            assert np.array_equal(s, s0)
        else:
            assert abs(s[0] - s0[0]) < 1e-5
    want = []
    for b in range(20):
        t = np.zeros((168, 168, 20), np.float32)
        t[:, :, b] = 1.0
        want.append(round((4.7 - pr.with_table(t.reshape(-1)).score_pose(poses[0])[0]) / 0.0157))
    got = per_bin_counts(pr, poses, sdgso.FP32_GUARD)[:, 0]
    assert np.array_equal(got, np.array(want))
    for v in VARIANTS[1:]:
        s, p = score32(pr.gpu, poses, v)
        assert np.array_equal(p, p0) and abs(s[0] - s0[0]) < 1e-5, v
        assert np.array_equal(per_bin_counts(pr, poses, None, v)[:, 0], np.array(want)), v


def test_tiny_distances_take_bin_zero():
    """Ligand atoms from 0.3 A down to 1e-160 A (d2 subnormal) and at 0 from a
    receptor atom: every 0 < d2 < 4 has bin 0 (sqrt(d2)*2 - 1 rounds to -1
    below about 3e-33, which the oracle maps to bin 0), d2 = 0 is no pair.
    The FP64 kernel's root skips the range scaling below 2^-767 and raises d2
    below 2^-100 to 2^-100; both forms must still match the oracle."""
    rng = np.random.default_rng(23)
    dist = np.array([0.3, 1e-1, 1e-5, 1e-10, 1e-16, 1e-17, 1e-20, 1e-30, 1e-54, 1e-60, 1e-100,
                     1e-154, 1e-155, 1e-160, 0.0, 1.9, 2.1, 7.49, 7.51])
    n = len(dist)
    dirs = rng.normal(size=(n, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    c = dict(r_xyz=np.array([[0.0, 0.0, 0.0], [400.0, 0.0, 0.0]]),
             l_xyz=dirs * dist[:, None],
             r_type=np.array([5, 9], np.int32), l_type=rng.integers(0, 167, n).astype(np.int32),
             table=(rng.normal(size=168 * 168 * 20) * 0.5).astype(np.float32),
             r_modes=np.zeros((0, 2, 3)), l_modes=np.zeros((0, n, 3)),
             r_mask=np.ones(2, np.int32), l_mask=np.ones(n, np.int32),
             lig_ref=np.zeros(3))
    c["pos"] = np.array([[[0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0]]])
    pr = GC.Pair(c)
    poses = poses_of(c)
    s0, _, p0 = pr.oracle_scores(poses)
    assert p0[0] == n - 1  # every atom but the one at the origin
    s, _, p = pr.gpu.score(poses, sdgso.FP64)
    assert np.array_equal(p, p0) and np.array_equal(s, s0)
    s, _, p = pr.gpu.score(poses, sdgso.FP32_GUARD)
    assert np.array_equal(p, p0) and abs(s[0] - s0[0]) < 1e-5
    want = []
    for b in range(20):
        t = np.zeros((168, 168, 20), np.float32)
        t[:, :, b] = 1.0
        want.append(round((4.7 - pr.with_table(t.reshape(-1)).score_pose(poses[0])[0]) / 0.0157))
    assert want[0] == 15  # the fifteen atoms with 0 < d < 2 A
    for form in (sdgso.FP64, sdgso.FP32_GUARD):
        assert np.array_equal(per_bin_counts(pr, poses, form)[:, 0], np.array(want)), form


def large_displacement_case(seed=19, nm=10, ext=25.0):
    """Receptor atom 0 displaced by ten modes with extents of tens of A, so
    that its FP32 displacement is off by a few 1e-6 A and d2 by more than the
    default 1e-4 A^2 band, and ligand atoms at d2 = k^2/4 + delta from its
    displaced position for every boundary k, delta across +-3e-3 and finely
    across +-4e-4."""
    rng = np.random.default_rng(seed)
    deltas = np.concatenate([np.linspace(-3e-3, 3e-3, 41), np.linspace(-4e-4, 4e-4, 161)])
    d2 = np.array([k * k / 4.0 + d for k in range(2, 31) for d in deltas])
    n = len(d2)
    r_modes = np.zeros((nm, 2, 3))
    r_modes[:, 0, :] = rng.normal(size=(nm, 3)) * 0.6
    e = rng.uniform(-ext, ext, nm)
    R = np.zeros(3)
    for m in range(nm):  # the reference's order: product rounded, then added
        R = R + r_modes[m, 0] * e[m]
    dirs = rng.normal(size=(n, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    c = dict(r_xyz=np.array([[0.0, 0.0, 0.0], [400.0, 0.0, 0.0]]),
             l_xyz=R + dirs * np.sqrt(d2)[:, None],
             r_type=np.array([5, 9], np.int32), l_type=rng.integers(0, 167, n).astype(np.int32),
             table=(rng.normal(size=168 * 168 * 20) * 0.5).astype(np.float32),
             r_modes=r_modes, l_modes=np.zeros((0, n, 3)),
             r_mask=np.ones(2, np.int32), l_mask=np.ones(n, np.int32),
             lig_ref=np.zeros(3))
    c["pos"] = np.concatenate([[0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0], e])[None, None, :]
    return c


@pytest.mark.parametrize("nm", [10, 20, 5])
def test_bins_exact_with_large_displacements(nm):
    """Each wave widens its guard band to its displacement bound, so bins
    stay exact when the FP32 mode sum is off by more than the default band;
    10 and 5 modes take the mode-major displacement, 20 the component-major
    one."""
    c = large_displacement_case(nm=nm)
    pr = GC.Pair(c)
    poses = poses_of(c)
    s0, _, p0 = pr.oracle_scores(poses)
    want = []
    for b in range(20):
        t = np.zeros((168, 168, 20), np.float32)
        t[:, :, b] = 1.0
        want.append(round((4.7 - pr.with_table(t.reshape(-1)).score_pose(poses[0])[0]) / 0.0157))
    want = np.array(want)
    assert want.sum() == p0[0] > 1000
    for v in VARIANTS:
        s, p = score32(pr.gpu, poses, v)
        assert np.array_equal(p, p0) and abs(s[0] - s0[0]) < 1e-5, v
        assert np.array_equal(per_bin_counts(pr, poses, None, v)[:, 0], want), v


def large_ligand_displacement_case(seed=21, nm=10, ext=25.0):
    """The ligand side of the case above: ten ligand modes with extents of
    tens of A move every ligand atom onto d2 = k^2/4 + delta from a receptor
    atom at the origin (the FP32 ligand mode sum is off by a few 1e-6 A)."""
    rng = np.random.default_rng(seed)
    deltas = np.concatenate([np.linspace(-3e-3, 3e-3, 21), np.linspace(-4e-4, 4e-4, 81)])
    d2 = np.array([k * k / 4.0 + d for k in range(2, 31) for d in deltas])
    n = len(d2)
    l_modes = rng.normal(size=(nm, n, 3)) * 0.6
    e = rng.uniform(-ext, ext, nm)
    dirs = rng.normal(size=(n, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    target = dirs * np.sqrt(d2)[:, None]
    D = np.zeros((n, 3))
    for m in range(nm):  # the reference's order: product rounded, then added
        D = D + l_modes[m] * e[m]
    c = dict(r_xyz=np.array([[0.0, 0.0, 0.0], [400.0, 0.0, 0.0]]),
             l_xyz=target - D,
             r_type=np.array([5, 9], np.int32), l_type=rng.integers(0, 167, n).astype(np.int32),
             table=(rng.normal(size=168 * 168 * 20) * 0.5).astype(np.float32),
             r_modes=np.zeros((0, 2, 3)), l_modes=l_modes,
             r_mask=np.ones(2, np.int32), l_mask=np.ones(n, np.int32),
             lig_ref=np.zeros(3))
    c["pos"] = np.concatenate([[0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0], e])[None, None, :]
    return c


def test_bins_exact_with_large_ligand_displacements():
    """The band also covers the FP32 ligand mode sum: bins stay exact at every
    boundary for ligand atoms that large ligand modes displace."""
    c = large_ligand_displacement_case()
    pr = GC.Pair(c)
    poses = poses_of(c)
    s0, _, p0 = pr.oracle_scores(poses)
    want = []
    for b in range(20):
        t = np.zeros((168, 168, 20), np.float32)
        t[:, :, b] = 1.0
        want.append(round((4.7 - pr.with_table(t.reshape(-1)).score_pose(poses[0])[0]) / 0.0157))
    want = np.array(want)
    assert want.sum() == p0[0] > 1000
    for v in VARIANTS:
        s, p = score32(pr.gpu, poses, v)
        assert np.array_equal(p, p0) and abs(s[0] - s0[0]) < 1e-5, v
        assert np.array_equal(per_bin_counts(pr, poses, None, v)[:, 0], want), v


def test_bins_exact_random_fp32():
    c = GC.case(8, 900, 400, 10, 10, 3, 8)
    pr = GC.Pair(c)
    poses = poses_of(c)
    a = per_bin_counts(pr, poses, sdgso.FP64)
    b = per_bin_counts(pr, poses, sdgso.FP32_GUARD)
    assert np.array_equal(a, b)
    assert a.sum() > 10000
    for v in VARIANTS[1:]:
        assert np.array_equal(per_bin_counts(pr, poses, None, v), a), v


def make_run(pr, c, steps, form=sdgso.FP64, stream=None):
    S, G = c["pos"].shape[:2]
    u = O.stream(0, S, steps, G) if stream is None else stream
    return sdgso.Run(pr.gpu, S, G, c["pos"], u, steps, form=form), u


def test_single_step_choices_from_identical_state():
    """Neighbour lists, roulette choices and moves from one identical state."""
    c = GC.case(9, 300, 150, 10, 10, 4, 64, swarm_radius=2.0)
    pr = GC.Pair(c)
    S, G = 4, 64
    steps = 8
    u = O.stream(0, S, steps, G)
    prm = O.default_params()
    st = O.run(pr.ora, prm, c["pos"], u, 5, S, G)  # a state with spread vision ranges
    ref_state = st.get()
    run, _ = make_run(pr, c, steps, stream=u)
    run.steps(5)  # advance the device stream index to k = 5
    run.write(pos=ref_state["pos"], luciferin=ref_state["luciferin"], score=ref_state["score"],
              vision=ref_state["vision"], moved=ref_state["moved"], ref=ref_state["ref"])
    # oracle: select + move of step 5 from the same state
    O.lib().or_phase_select(prm, st.st, O._p(np.ascontiguousarray(u)), steps, 5, 0)
    O.lib().or_phase_move(prm, st.st, 10, 10, 0)
    want = st.get()
    run.phases(2 | 4)
    got = run.read()
    assert np.array_equal(got["nnei"], want["nnei"].ravel())
    assert np.array_equal(got["select"], want["select"].ravel())
    assert np.array_equal(got["moved"], want["moved"].ravel())
    assert np.array_equal(got["vision"], want["vision"].ravel())
    assert want["moved"].sum() > 20
    gp = got["pos"].reshape(want["pos"].shape)
    # translation and extents are exact; the quaternion goes through acos/sin
    assert np.array_equal(gp[..., :3], want["pos"][..., :3])
    assert np.array_equal(gp[..., 7:], want["pos"][..., 7:])
    assert np.allclose(gp[..., 3:7], want["pos"][..., 3:7], rtol=0, atol=POS_TOL)


def test_fp64_run_tracks_oracle():
    c = GC.case(10, 300, 160, 10, 10, 3, 48, swarm_radius=2.5)
    pr = GC.Pair(c)
    S, G, steps = 3, 48, 10
    run, u = make_run(pr, c, steps)
    st = O.State(S, G, pr.P, O.default_params(), c["pos"])
    moved = 0
    for k in range(steps):
        run.steps(1)
        O.lib().or_gso_step(pr.ora.c, O.default_params(), st.st, O._p(np.ascontiguousarray(u)),
                            steps, k, 1, 0)
        g, w = run.read(), st.get()
        assert np.array_equal(g["select"], w["select"].ravel()), k
        assert np.array_equal(g["nnei"], w["nnei"].ravel()), k
        assert np.allclose(g["pos"].reshape(w["pos"].shape), w["pos"], rtol=0, atol=1e-11), k
        assert np.allclose(g["score"], w["score"].ravel(), rtol=1e-12, atol=1e-12), k
        assert np.allclose(g["luciferin"], w["luciferin"].ravel(), rtol=1e-12, atol=1e-12), k
        moved += int(g["moved"].sum())
    assert moved > 50


def test_graph_replay_equals_eager():
    c = GC.case(11, 400, 200, 10, 10, 3, 32, swarm_radius=3.0)
    pr = GC.Pair(c)
    outs = []
    for form in (sdgso.FP64, sdgso.FP32_GUARD):
        for graph in (False, True):
            run, _ = make_run(pr, c, 6, form=form)
            run.steps(2, graph=graph)
            run.sync()
            run.steps(4, graph=graph)
            outs.append(run.read())
            run.close()
    for a, b in ((outs[0], outs[1]), (outs[2], outs[3])):
        for key in ("pos", "luciferin", "score", "select", "pairs", "vision"):
            assert np.array_equal(a[key], b[key]), key


def test_every_graph_replay_equals_eager():
    """The captured step checked replay by replay, with host reads and syncs
    between replays: a run of replays only, and a run that alternates replays
    with eager steps (both advance the same device-resident step index), each
    equal after every step to an eager run from the same start."""
    c = GC.case(21, 400, 200, 10, 10, 3, 32, swarm_radius=3.0)
    pr = GC.Pair(c)
    keys = ("pos", "luciferin", "score", "select", "pairs", "vision", "moved", "nnei")
    for form in (sdgso.FP64, sdgso.FP32_GUARD):
        eager, _ = make_run(pr, c, 5, form=form)
        replays, _ = make_run(pr, c, 5, form=form)
        mixed, _ = make_run(pr, c, 5, form=form)
        # profiled replays: the step as two graphs with timing events between
        profiled, _ = make_run(pr, c, 5, form=form)
        profiled.profile(True)
        for k in range(5):
            eager.steps(1)
            replays.steps(1, graph=True)
            mixed.steps(1, graph=(k % 2 == 0))
            profiled.steps(1, graph=True)
            want = eager.read()
            for run in (replays, mixed, profiled):
                got = run.read()
                for key in keys:
                    assert np.array_equal(got[key], want[key]), (form, k, key)
        prof = profiled.profile_read()
        assert prof["score_launches"] == prof["move_launches"] == 5
        assert prof["score_ms"] > 0 and prof["move_ms"] > 0
        for run in (eager, replays, mixed, profiled):
            run.close()


def test_batch_equals_solo_runs():
    """Two complexes of different sizes in one run give the results of two solo runs."""
    ca = GC.case(12, 350, 180, 10, 10, 2, 32, swarm_radius=3.0)
    cb = GC.case(13, 520, 140, 10, 10, 3, 32, swarm_radius=3.0)
    pa, pb = GC.Pair(ca), GC.Pair(cb)
    steps = 4
    ua, ub = O.stream(0, 2, steps, 32), O.stream(0, 3, steps, 32)
    for form in (sdgso.FP64, sdgso.FP32_GUARD):
        ra = sdgso.Run(pa.gpu, 2, 32, ca["pos"], ua, steps, form=form)
        rb = sdgso.Run(pb.gpu, 3, 32, cb["pos"], ub, steps, form=form)
        both = sdgso.Run([pa.gpu, pb.gpu], [2, 3], 32,
                         np.concatenate([ca["pos"].ravel(), cb["pos"].ravel()]),
                         np.concatenate([ua.ravel(), ub.ravel()]), steps, form=form)
        for r in (ra, rb, both):
            r.steps(steps)
        a, b, ab = ra.read(), rb.read(), both.read()
        for key in ("pos", "score", "luciferin", "select", "pairs"):
            assert np.array_equal(np.concatenate([a[key], b[key]]), ab[key]), (form, key)


def test_fp32_run_close_to_fp64():
    c = GC.case(14, 600, 250, 10, 10, 4, 40, swarm_radius=4.0)
    pr = GC.Pair(c)
    r64, _ = make_run(pr, c, 5, form=sdgso.FP64)
    r32, _ = make_run(pr, c, 5, form=sdgso.FP32_GUARD)
    r64.steps(1)
    r32.steps(1)
    a, b = r64.read(), r32.read()
    assert np.array_equal(a["pairs"], b["pairs"])
    assert np.allclose(b["score"], a["score"], rtol=1e-9, atol=1e-9)


def test_step_past_stream_is_an_error():
    c = GC.case(15, 100, 60, 0, 0, 1, 8)
    pr = GC.Pair(c)
    run, _ = make_run(pr, c, 2)
    run.steps(2)
    with pytest.raises(sdgso.SdgsoError) as e:
        run.steps(1)
    assert e.value.code == 4


def test_filter_stats_consistent_with_pairs():
    """The filter counters of a scoring pass agree with the exact pair counts
    of the same poses, for every kernel variant; the pass changes no state."""
    c = GC.case(16, 900, 260, 10, 10, 3, 24, swarm_radius=3.0)
    pr = GC.Pair(c)
    _, _, pairs64 = pr.gpu.score(poses_of(c), sdgso.FP64)
    run, _ = make_run(pr, c, 2, form=sdgso.FP32_GUARD)
    before = run.read()
    n_blocks = -(-260 // 64)
    for v, name in enumerate(sdgso.tune_variants()):
        run.set_variant(v)
        st = run.filter_stats()
        # a two-block variant's waves take blocks (0, 1), (2, 3) and (4)
        two = name.startswith("bcast2")
        assert st["waves"] == 3 * 24 * (-(-n_blocks // 2) if two else n_blocks)
        assert 0 < st["waves_with_candidates"] <= st["waves"]
        assert st["near_atoms"] <= 64 * st["candidate_tiles"]
        assert st["cutoff_pairs"] == int(pairs64.sum())
        assert st["cutoff_pairs"] <= st["queued_pairs"] <= (128 if two else 64) * st["near_atoms"]
    after = run.read()
    for key in before:
        assert np.array_equal(before[key], after[key]), key


def dense_case(seed=18, nr=20000, nl=64, n_poses=3):
    """A receptor packed into a 12 A ball, so that every one of its 313 tiles
    is a candidate of the ligand block (more than one candidate-list chunk)."""
    rng = np.random.default_rng(seed)
    r = rng.normal(size=(nr, 3))
    r *= (12.0 * rng.random(nr) ** (1 / 3) / np.linalg.norm(r, axis=1))[:, None]
    lg = rng.normal(size=(nl, 3)) * 2.0
    c = dict(r_xyz=r, l_xyz=lg, r_type=rng.integers(0, 168, nr).astype(np.int32),
             l_type=rng.integers(0, 168, nl).astype(np.int32),
             table=(rng.normal(size=168 * 168 * 20) * 0.5).astype(np.float32),
             r_modes=np.zeros((0, nr, 3)), l_modes=np.zeros((0, nl, 3)),
             r_mask=np.ones(nr, np.int32), l_mask=np.ones(nl, np.int32), lig_ref=np.zeros(3))
    q = rng.normal(size=(n_poses, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    t = rng.normal(size=(n_poses, 3)) * 1.0
    c["pos"] = np.concatenate([t, q], axis=1).reshape(1, n_poses, 7)
    return c


def test_every_tile_a_candidate():
    c = dense_case()
    pr = GC.Pair(c)
    poses = poses_of(c)
    s0, _, p0 = pr.oracle_scores(poses)
    s64, _, p64 = pr.gpu.score(poses, sdgso.FP64)
    assert np.array_equal(p64, p0) and np.array_equal(s64, s0)
    denom = GC.sum_abs_terms(pr, poses)
    e0 = (4.7 - s0) / 0.0157
    for v in VARIANTS:
        s, p = score32(pr.gpu, poses, v)
        assert np.array_equal(p, p0), v
        assert_energy((4.7 - s) / 0.0157, e0, denom, v)
    run = sdgso.Run(pr.gpu, 1, len(poses), poses, None, 0, form=sdgso.FP32_GUARD)
    st = run.filter_stats()
    assert st["candidate_tiles"] == len(poses) * 313


def test_tune_per_complex():
    """The autotuner picks a variant per complex; the tuned batch gives the
    FP64 form's pair counts and scores within the FP32 tolerance."""
    ca = GC.case(19, 420, 200, 10, 10, 2, 24, swarm_radius=3.0)
    cb = GC.case(20, 900, 90, 10, 10, 3, 24, swarm_radius=3.0)
    pa, pb = GC.Pair(ca), GC.Pair(cb)
    pos = [np.concatenate([ca["pos"].ravel(), cb["pos"].ravel()])]
    u = np.concatenate([O.stream(0, 2, 2, 24).ravel(), O.stream(0, 3, 2, 24).ravel()])
    runs = {}
    for form in (sdgso.FP64, sdgso.FP32_GUARD):
        runs[form] = sdgso.Run([pa.gpu, pb.gpu], [2, 3], 24, pos[0], u, 2, form=form)
    chosen, times = runs[sdgso.FP32_GUARD].tune()
    assert len(chosen) == 2 and chosen == runs[sdgso.FP32_GUARD].variants()
    names = sdgso.tune_variants()
    assert sorted(times) == sorted(names)
    for k in range(2):
        assert times[names[chosen[k]]][k] == min(t[k] for t in times.values())
    for r in runs.values():
        r.steps(1)
    a, b = runs[sdgso.FP64].read(), runs[sdgso.FP32_GUARD].read()
    assert np.array_equal(a["pairs"], b["pairs"])
    assert np.allclose(b["score"], a["score"], rtol=1e-9, atol=1e-9)
    runs[sdgso.FP32_GUARD].set_variant(len(names) - 1)
    assert runs[sdgso.FP32_GUARD].variants() == [len(names) - 1] * 2


def test_anm_gemm_probe_counts():
    """The matrix-core GEMM probe covers, for every swarm, the tiles its
    glowworms can reach (at least every tile that holds a pair), 64 atoms x
    G glowworms per (swarm, tile), and leaves the run's state unchanged."""
    c = GC.case(61, 900, 260, 10, 10, 3, 40, swarm_radius=3.0)
    pr = GC.Pair(c)
    run, _ = make_run(pr, c, 2, form=sdgso.FP32_GUARD)
    before = run.read()
    res = run.probe_anm_gemm()
    assert res["ms"] > 0 and res["swarm_tiles"] >= 3
    assert res["displacements"] == res["swarm_tiles"] * 64 * 40
    assert res["bytes_written"] == res["swarm_tiles"] * 2 * 32 * 192 * 4
    after = run.read()
    for key in before:
        assert np.array_equal(before[key], after[key]), key
