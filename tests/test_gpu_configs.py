"""The HIP path against the CPU oracle at the shapes of the benchmark
workloads (sparkledock_amd/workloads.py), through the C-ABI.

  c0  757/481 atoms, 130 swarms x 200 glowworms, no modes, 20 steps, plain
      FP64: the whole run, step by step from the oracle's state;
  c1, c2, c3  the FP32 guard form over every complex at full size (one
      tuned scoring launch over all swarms): sampled glowworms of every
      complex against the oracle (pair counts, per-bin counts through one-hot
      tables, energies), and a full step of two swarms at G = 200 (the
      four-wave select/move workgroup) from an identical state;
  c4  32 complexes, 300 glowworms, 20+20 modes: a full FP64 step of one
      swarm of the complex with the most pairs (26417/2298 atoms: the FP64
      kernel's ligand and row sums in 67 KB of LDS, above the 64 KB default, and a
      select/move workgroup of 300 x 47-value poses), and the whole FP32 scoring
      launch with a one-block variant, whose 130 M-wave work list is cut at
      2^26 - 4 waves, checked on the glowworms at the cut, and with the
      two-block variant in the tuner's c4 configuration (XCD-chunked order,
      ligand displacement GEMM).

Reference semantics: prepare_gso_tasks.cc:359-822 (the step),
1522-1615 (scoring), 584-814 (selection and movement); the launch the
reference makes per swarm, prepare_gso_task.cu:791-897, uses one block of G
threads.  Bars as in test_gpu_parity.py: FP64 form bit-exact, FP32 form with
equal bins and energy within 1e-9 |E| + 1e-12 sum |T|, choices bit-exact from
identical states, poses equal up to the slerp's acos/sin.
"""
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle_lib as O
from sparkledock_amd import sdgso, workloads

pytestmark = pytest.mark.gpu

REL_E = 1e-9
FLOOR_E = 1e-12
POS_TOL = 1e-13
WORKERS = max(1, min(16, len(os.sched_getaffinity(0))))
CHUNK_WAVES = (1 << 26) - 4   # waves per scoring launch of the FP32 kernels


def say(*a):
    print(*a, flush=True)


def oracle_complex(inp, table=None):
    return O.Complex(inp["r_xyz"], inp["r_type"], inp["l_xyz"], inp["l_type"],
                     inp["table"] if table is None else table, inp["lig_ref"], inp["r_modes"],
                     inp["l_modes"], inp["r_mask"], inp["l_mask"])


def oracle_scores(cx, poses):
    """(scores, pair counts) of poses by the oracle, poses in parallel."""
    with ThreadPoolExecutor(WORKERS) as ex:
        out = list(ex.map(cx.score_pose, list(poses)))
    return np.array([o[0] for o in out]), np.array([o[2] for o in out], np.int64)


def onehot(b):
    t = np.zeros((168, 168, 20), np.float32)
    t[:, :, b] = 1.0
    return t.reshape(-1)


def counts_of(scores):
    return np.rint((4.7 - np.asarray(scores)) / 0.0157).astype(np.int64)


def check_energies(inp, poses, got_score, got_pairs, what, more=()):
    """Sampled poses of one complex: pair counts equal to the oracle's and
    energies within the FP32 guard form's bound; `more` holds further
    (scores, pair counts) of the same poses to check the same way."""
    s0, p0 = oracle_scores(oracle_complex(inp), poses)
    sa, _ = oracle_scores(oracle_complex(inp, np.abs(inp["table"])), poses)
    denom = np.maximum((4.7 - sa) / 0.0157, 1.0)
    e0 = (4.7 - s0) / 0.0157
    for sc, pc in [(got_score, got_pairs)] + list(more):
        assert np.array_equal(pc, p0), what
        err = np.abs((4.7 - np.asarray(sc)) / 0.0157 - e0)
        assert np.all(err <= REL_E * np.abs(e0) + FLOOR_E * denom), \
            (what, float(np.max(err / np.maximum(np.abs(e0), 1e-300))))
    return int(p0.sum())


def check_bins(inp, poses, variant, what):
    """Per-bin pair counts of the poses through one-hot tables: the FP32
    kernel variant against the oracle."""
    want, got = [], []
    for b in range(20):
        tab = onehot(b)
        s0, _ = oracle_scores(oracle_complex(inp, tab), poses)
        want.append(counts_of(s0))
        cx = sdgso.Complex(inp["r_xyz"], inp["r_type"], inp["l_xyz"], inp["l_type"], tab,
                           inp["lig_ref"], inp["r_modes"], inp["l_modes"], inp["r_mask"],
                           inp["l_mask"])
        run = sdgso.Run(cx, len(poses), 1, poses, None, 0, form=sdgso.FP32_GUARD)
        run.set_variant(variant)
        run.phases(1)
        got.append(counts_of(run.read()["score"]))
        run.close()
        cx.close()
    assert np.array_equal(np.array(got), np.array(want)), what


def spread_state(rng, n, vision_hi=5.0):
    """Vision ranges across [0, vision_hi] and luciferin around 5, so that a
    selection step has many neighbours to choose from."""
    return rng.uniform(0.0, vision_hi, n), 5.0 + rng.normal(size=n)


# ---------------------------------------------------------------- c0
def test_c0_whole_run_step_by_step():
    """c0 in full, plain FP64: 130 x 200 glowworms, 7-value poses (no modes),
    20 steps.  Before every step the device takes the oracle's state; the
    oracle re-scores only moved glowworms (the reference CPU policy,
    prepare_gso_tasks.cc:434-436), the device re-scores every glowworm.
    Scores, luciferin, vision, neighbour counts, selections and moved flags
    bit-exact every step; translations exact, quaternions within 1e-13."""
    wl = workloads.build("c0")
    S, G, steps = wl["n_swarms"][0], wl["G"], wl["steps"]
    inp = wl["inputs"][0]
    ora = oracle_complex(inp)
    P = ora.P
    assert (S, G, P, steps) == (130, 200, 7, 20)
    prm = O.default_params()
    st = O.State(S, G, P, prm, wl["pos"])
    u = np.ascontiguousarray(wl["stream"])
    run = sdgso.Run(wl["complexes"], wl["n_swarms"], G, wl["pos"], wl["stream"], steps,
                    form=sdgso.FP64)
    moved = 0
    for k in range(steps):
        w = st.get()
        if k:
            run.write(pos=w["pos"], luciferin=w["luciferin"], score=w["score"],
                      vision=w["vision"], moved=w["moved"], ref=w["ref"])
        run.steps(1)
        O.lib().or_gso_step(ora.c, prm, st.st, O._p(u), steps, k, 0, WORKERS)
        g, w = run.read(), st.get()
        for key in ("score", "luciferin", "vision", "nnei", "select", "moved"):
            assert np.array_equal(g[key], w[key].ravel()), (k, key)
        assert np.array_equal(g["ref"], w["ref"].reshape(-1, 3)), k
        gp, wp = g["pos"].reshape(S, G, P), w["pos"]
        assert np.array_equal(gp[..., :3], wp[..., :3]), k
        assert np.allclose(gp[..., 3:7], wp[..., 3:7], rtol=0, atol=POS_TOL), k
        moved += int(w["moved"].sum())
        say("c0 step %d: %d moved" % (k, int(w["moved"].sum())))
    assert moved > 1000


# ---------------------------------------------------------------- c1, c2, c3
@pytest.mark.parametrize("name", ["c1", "c2", "c3"])
def test_fp32_workload_sampled_against_oracle(name):
    """Every complex of the workload at full size in one run, the autotuner's
    variants, one step: the scores and pair counts of the step's scoring
    launch, for 8 sampled glowworms of every complex, against the oracle;
    per-bin counts of 2 of them through one-hot tables with the complex's
    tuned variant."""
    wl = workloads.build(name, steps=1)
    G = wl["G"]
    run = sdgso.Run(wl["complexes"], wl["n_swarms"], G, wl["pos"], wl["stream"], 1,
                    form=sdgso.FP32_GUARD)
    chosen, _ = run.tune()
    run.steps(1)
    got = run.read()
    names = sdgso.tune_variants()
    pos = wl["pos"].reshape(-1, wl["complexes"][0].P)
    rng = np.random.default_rng(11)
    g0 = 0
    for k, (inp, S) in enumerate(zip(wl["inputs"], wl["n_swarms"])):
        idx = g0 + np.sort(rng.choice(S * G, size=8, replace=False))
        what = "%s complex %d (%d/%d, %s)" % (name, k, len(inp["r_xyz"]), len(inp["l_xyz"]),
                                               names[chosen[k]])
        n = check_energies(inp, pos[idx], got["score"][idx], got["pairs"][idx], what)
        check_bins(inp, pos[idx[:2]], chosen[k], what)
        say("%s: %d pairs checked" % (what, n))
        g0 += S * G
    assert g0 == len(got["score"])


def two_swarm_state(name, form, seed):
    """Two swarms of the workload's first complex at G = 200 with spread
    vision ranges and luciferin, as a device run and an oracle state."""
    wl = workloads.build(name, swarm_scale=2.0 / workloads.spec(name)["complexes"][0][2], steps=1)
    inp = wl["inputs"][0]
    S, G = wl["n_swarms"][0], wl["G"]
    assert (S, G) == (2, 200)
    ora = oracle_complex(inp)
    prm = O.default_params()
    st = O.State(S, G, ora.P, prm, wl["pos"])
    vis, luc = spread_state(np.random.default_rng(seed), S * G)
    st.set("vision", vis)
    st.set("luciferin", luc)
    run = sdgso.Run(wl["complexes"][0], S, G, wl["pos"], wl["stream"], 1, form=form)
    run.write(vision=vis, luciferin=luc)
    return wl, ora, prm, st, run


def test_g200_full_step_two_swarms_fp64():
    """A whole FP64 step (score, luciferin, select, move) of two c1 swarms at
    G = 200, 27-value poses, from the same state on both sides: bit-exact."""
    wl, ora, prm, st, run = two_swarm_state("c1", sdgso.FP64, 3)
    S, G, P = 2, 200, ora.P
    run.steps(1)
    O.lib().or_gso_step(ora.c, prm, st.st, O._p(np.ascontiguousarray(wl["stream"])), 1, 0, 1,
                        WORKERS)
    g, w = run.read(), st.get()
    for key in ("score", "luciferin", "vision", "nnei", "select", "moved", "pairs"):
        assert np.array_equal(g[key], w[key].ravel()), key
    gp, wp = g["pos"].reshape(S, G, P), w["pos"]
    assert np.array_equal(gp[..., :3], wp[..., :3])
    assert np.array_equal(gp[..., 7:], wp[..., 7:])
    assert np.allclose(gp[..., 3:7], wp[..., 3:7], rtol=0, atol=POS_TOL)
    assert w["moved"].sum() > 100 and w["nnei"].max() > 5


def test_g200_full_step_two_swarms_fp32():
    """The same step in the FP32 guard form: its scoring phase against the
    oracle's (pair counts equal, energies within the bound), then selection
    and movement from the oracle's luciferin: bit-exact choices."""
    wl, ora, prm, st, run = two_swarm_state("c1", sdgso.FP32_GUARD, 4)
    S, G, P = 2, 200, ora.P
    run.phases(1)
    O.lib().or_phase_score(ora.c, prm, st.st, 1, WORKERS)
    g, w = run.read(), st.get()
    assert np.array_equal(g["pairs"], w["pairs"].ravel())
    sa, _ = oracle_scores(oracle_complex(wl["inputs"][0], np.abs(wl["inputs"][0]["table"])),
                          wl["pos"].reshape(-1, P))
    denom = np.maximum((4.7 - sa) / 0.0157, 1.0)
    e0, e = (4.7 - w["score"].ravel()) / 0.0157, (4.7 - g["score"]) / 0.0157
    assert np.all(np.abs(e - e0) <= REL_E * np.abs(e0) + FLOOR_E * denom)
    run.write(luciferin=w["luciferin"], score=w["score"])
    run.phases(2 | 4)
    O.lib().or_phase_select(prm, st.st, O._p(np.ascontiguousarray(wl["stream"])), 1, 0, WORKERS)
    O.lib().or_phase_move(prm, st.st, 10, 10, WORKERS)
    g, w = run.read(), st.get()
    for key in ("vision", "nnei", "select", "moved"):
        assert np.array_equal(g[key], w[key].ravel()), key
    gp, wp = g["pos"].reshape(S, G, P), w["pos"]
    assert np.array_equal(gp[..., :3], wp[..., :3])
    assert np.array_equal(gp[..., 7:], wp[..., 7:])
    assert np.allclose(gp[..., 3:7], wp[..., 3:7], rtol=0, atol=POS_TOL)
    assert w["moved"].sum() > 100


# ---------------------------------------------------------------- c4
def c4_complex(k, S):
    nr, nl, _ = workloads.c4_sizes()[k]
    inp = workloads.complex_inputs(k, nr, nl, 20, 20, sdgso.synth_table(workloads.TABLE_SEED))
    pos = workloads.pose_inputs(k, S, 300, nr, nl, 20, 20)
    return inp, pos


def test_c4_largest_complex_fp64_step_g300():
    """One swarm of the c4 complex with the most pairs (26417/2298 atoms,
    20+20 modes, 47-value poses) at G = 300: a whole FP64 step from the same
    state with spread vision ranges.  The FP64 kernel holds 2298 ligand atoms
    and its row sums in 67 KB of LDS, and select/move moves 300 poses of 47 values; every output
    bit-exact but the quaternions (1e-13)."""
    k = 16
    inp, pos = c4_complex(k, 1)
    assert (len(inp["r_xyz"]), len(inp["l_xyz"])) == (26417, 2298)
    S, G, P = 1, 300, 47
    cx = sdgso.Complex(inp["r_xyz"], inp["r_type"], inp["l_xyz"], inp["l_type"], inp["table"],
                       inp["lig_ref"], inp["r_modes"], inp["l_modes"], inp["r_mask"],
                       inp["l_mask"])
    u = sdgso.stream(0, S, 1, G)
    ora = oracle_complex(inp)
    prm = O.default_params()
    st = O.State(S, G, P, prm, pos)
    vis, luc = spread_state(np.random.default_rng(5), S * G)
    st.set("vision", vis)
    st.set("luciferin", luc)
    run = sdgso.Run(cx, S, G, pos, u, 1, form=sdgso.FP64)
    run.write(vision=vis, luciferin=luc)
    run.steps(1)
    O.lib().or_gso_step(ora.c, prm, st.st, O._p(np.ascontiguousarray(u)), 1, 0, 1, WORKERS)
    g, w = run.read(), st.get()
    for key in ("score", "luciferin", "vision", "nnei", "select", "moved", "pairs"):
        assert np.array_equal(g[key], w[key].ravel()), key
    gp, wp = g["pos"].reshape(S, G, P), w["pos"]
    assert np.array_equal(gp[..., :3], wp[..., :3])
    assert np.array_equal(gp[..., 7:], wp[..., 7:])
    assert np.allclose(gp[..., 3:7], wp[..., 3:7], rtol=0, atol=POS_TOL)
    assert w["moved"].sum() > 50 and w["pairs"].min() > 0


def test_c4_whole_launch_across_the_chunk_cut():
    """All 32 c4 complexes (22772 swarms x 300 glowworms) scored in one run
    with the one-block variant bcast/ring4/occ4: 130 M (glowworm, block)
    waves, cut into launches at 2^26 - 4 waves (here between two glowworms
    of complex 16).  The glowworms on either side of the cut, and one sampled
    glowworm of every complex, against the oracle; every glowworm's pair
    count equal to that of the two-block default variant's single launch, and
    of that variant in the tuner's c4 configuration (work items in chunks of
    8192 per XCD, the ligand displacement GEMM on: 100 GB of output), whose
    sampled glowworms are checked against the oracle too."""
    wl = workloads.build("c4", steps=1)
    G = wl["G"]
    names = sdgso.tune_variants()
    nblk = [-(-len(inp["l_xyz"]) // 64) for inp in wl["inputs"]]
    waves = np.concatenate([np.full(S * G, nb, np.int64) for S, nb in zip(wl["n_swarms"], nblk)])
    first = np.concatenate([[0], np.cumsum(waves)])
    assert first[-1] > CHUNK_WAVES
    run = sdgso.Run(wl["complexes"], wl["n_swarms"], G, wl["pos"], wl["stream"], 1,
                    form=sdgso.FP32_GUARD)
    run.set_variant(names.index("bcast/ring4/occ4"))
    run.phases(1)
    one = run.read()
    run.set_variant(names.index("bcast2/ring4/occ4"))
    run.phases(1)
    two = run.read()
    run.set_work_order(8192)
    run.set_ligand_gemm(True)
    assert run.ligand_gemm()[0]
    run.phases(1)
    tuned = run.read()
    run.close()
    assert np.array_equal(one["pairs"], two["pairs"])
    assert np.allclose(one["score"], two["score"], rtol=1e-9, atol=1e-9)
    assert np.array_equal(tuned["pairs"], two["pairs"])
    assert np.allclose(tuned["score"], two["score"], rtol=1e-9, atol=1e-9)
    # the glowworm whose waves hold the cut, and its neighbours
    cut = int(np.searchsorted(first, CHUNK_WAVES, side="right") - 1)
    say("cut at wave %d: glowworm %d has waves [%d, %d)" % (CHUNK_WAVES, cut, first[cut],
                                                             first[cut + 1]))
    assert first[cut] <= CHUNK_WAVES < first[cut + 1]
    sel = {cut - 1, cut, cut + 1}
    rng = np.random.default_rng(12)
    g0 = 0
    owner = []
    for k, S in enumerate(wl["n_swarms"]):
        sel.add(g0 + int(rng.integers(S * G)))
        owner.append((g0, g0 + S * G))
        g0 += S * G
    pos = wl["pos"].reshape(-1, wl["complexes"][0].P)
    for k, (a, b) in enumerate(owner):
        idx = np.array(sorted(g for g in sel if a <= g < b))
        if len(idx):
            check_energies(wl["inputs"][k], pos[idx], one["score"][idx], one["pairs"][idx],
                           "c4 complex %d glowworms %s" % (k, idx.tolist()),
                           more=[(tuned["score"][idx], tuned["pairs"][idx])])


def test_chunk_cut_inside_a_glowworm():
    """A launch cut that falls inside one glowworm's blocks: 704 ligand atoms
    (11 blocks per glowworm; 2^26 - 4 is not a multiple of 11) and 6.1 M
    glowworms, 67.2 M one-block waves.  Glowworm 6100805 has blocks 0-4 in
    the first launch and 5-10 in the second; its partials meet in the
    epilogue.  Its score and those of its neighbours against the oracle, and
    every glowworm's pair count against the two-block variant (36.6 M waves,
    one launch).  The glowworms repeat 64 distinct poses."""
    rng = np.random.default_rng(40)
    nr, nl, G, S = 150, 704, 64, 95400
    r = rng.normal(size=(nr, 3)) * 6.0
    lg = rng.normal(size=(nl, 3)) * 7.0
    inp = dict(r_xyz=r, l_xyz=lg, r_type=rng.integers(0, 168, nr).astype(np.int32),
               l_type=rng.integers(0, 168, nl).astype(np.int32),
               table=(rng.normal(size=168 * 168 * 20) * 0.5).astype(np.float32),
               r_modes=np.zeros((0, nr, 3)), l_modes=np.zeros((0, nl, 3)),
               r_mask=np.ones(nr, np.int32), l_mask=np.ones(nl, np.int32), lig_ref=np.zeros(3))
    q = rng.normal(size=(G, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    base = np.concatenate([rng.normal(size=(G, 3)) * 4.0, q], axis=1)
    pos = np.tile(base, (S, 1))
    cut = CHUNK_WAVES // 11
    assert 11 * cut < CHUNK_WAVES < 11 * (cut + 1) < 11 * S * G
    cx = sdgso.Complex(inp["r_xyz"], inp["r_type"], inp["l_xyz"], inp["l_type"], inp["table"],
                       inp["lig_ref"])
    names = sdgso.tune_variants()
    run = sdgso.Run(cx, S, G, pos, None, 0, form=sdgso.FP32_GUARD)
    out = {}
    for v in ("bcast/ring4/occ4", "bcast2/ring4/occ4"):
        run.set_variant(names.index(v))
        run.phases(1)
        st = run.read()
        out[v] = (st["score"], st["pairs"])
    run.close()
    a, b = out["bcast/ring4/occ4"], out["bcast2/ring4/occ4"]
    assert np.array_equal(a[1], b[1])
    assert np.allclose(a[0], b[0], rtol=1e-9, atol=1e-9)
    idx = np.array([0, cut - 1, cut, cut + 1, S * G - 1])
    n = check_energies(inp, pos[idx], a[0][idx], a[1][idx], "glowworms %s" % idx.tolist())
    assert n > 0
    # every copy of a pose scores alike, across the cut too
    assert np.array_equal(a[0].reshape(S, G), np.tile(a[0][:G], (S, 1)))
