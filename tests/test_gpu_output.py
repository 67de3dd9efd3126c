"""The output side of a run on the GPU: final FP64 coordinates of every
glowworm (prepare_gso_tasks.cc:882-920) and the per-swarm clustering
(prepare_gso_tasks.cc:961-1080), against the oracle's pose transform,
cal_rmsd restatement and BSAS pass on the same state."""
import numpy as np
import pytest

import gpu_cases as GC
import oracle_lib as O
from sparkledock_amd import sdgso

pytestmark = pytest.mark.gpu


def two_complex_run(steps=3):
    ca = GC.case(31, 380, 150, 10, 10, 2, 24, swarm_radius=2.0)
    cb = GC.case(32, 520, 110, 10, 10, 3, 24, swarm_radius=2.0)
    pa, pb = GC.Pair(ca), GC.Pair(cb)
    u = np.concatenate([O.stream(0, 2, steps, 24).ravel(), O.stream(0, 3, steps, 24).ravel()])
    run = sdgso.Run([pa.gpu, pb.gpu], [2, 3], 24,
                    np.concatenate([ca["pos"].ravel(), cb["pos"].ravel()]), u, steps,
                    form=sdgso.FP64)
    run.steps(steps)
    return run, (pa, pb)


def test_final_coords_equal_oracle():
    run, pairs = two_complex_run()
    st = run.read()
    # a range in the second complex (glowworms 48.. of the run)
    first, count = 48 + 5, 30
    rec, lig = run.final_coords(first, count)
    for i in range(count):
        R, L, _ = pairs[1].ora.pose_coords(st["pos"][first + i])
        assert np.array_equal(rec[i], R) and np.array_equal(lig[i], L), i
    with pytest.raises(sdgso.SdgsoError):
        run.final_coords(40, 20)  # spans both complexes


@pytest.mark.parametrize("cutoff", [4.0, 1.5])
def test_cluster_equals_oracle(cutoff):
    run, pairs = two_complex_run()
    st = run.read()
    G = 24
    for k, (pr, S, g0) in enumerate(((pairs[0], 2, 0), (pairs[1], 3, 48))):
        rb = np.arange(0, pr.ora.nr, 3, dtype=np.int32)
        lb = np.arange(1, pr.ora.nl, 2, dtype=np.int32)
        cl, rep = run.cluster(k, rb, lb, cutoff)
        for s in range(S):
            coords = []
            for g in range(G):
                R, L, _ = pr.ora.pose_coords(st["pos"][g0 + s * G + g])
                coords.append(np.concatenate([R[rb], L[lb]]))
            want_cl, want_rep = O.bsas(np.array(coords), st["score"][g0 + s * G:g0 + (s + 1) * G],
                                       cutoff)
            assert np.array_equal(cl[s], want_cl), (k, s)
            assert np.array_equal(rep[s], want_rep), (k, s)
    with pytest.raises(sdgso.SdgsoError):
        run.cluster(0, np.array([10**6], np.int32), np.array([0], np.int32))


def test_final_pdb_round_trip(tmp_path):
    run, pairs = two_complex_run(steps=2)
    pr = pairs[0]
    rec, lig = run.final_coords(7, 1)
    rs = sdgso.Structure.synthetic(pr.case["r_xyz"])
    ls = sdgso.Structure.synthetic(pr.case["l_xyz"])
    path = tmp_path / "lightdock_7.pdb"
    rs.write_pdb(path, rec[0])
    ls.write_pdb(path, lig[0], append=True)
    back = sdgso.Structure.read_pdb(path)
    assert back.n_atoms == pr.ora.nr + pr.ora.nl
    assert np.max(np.abs(back.coords() - np.concatenate([rec[0], lig[0]]))) <= 5.1e-4
