"""The lightdock driver end to end (`lightdock -f setup.json -s STEPS`, the
reference CLI of source/simulation/run_simulation.cc): it writes one
swarm_<i>/gso_<STEPS>.out per swarm in the reference line format
(prepare_gso_task.cu:2868-2888), the values equal the same run through the
C ABI, and in the FP64 form its files equal the oracle's run byte for byte;
it clusters every swarm into swarm_<i>/cluster.repr
(prepare_gso_tasks.cc:1084-1110) and optionally writes the final poses as
PDB files; it reads PDB inputs (the 2X9A benchmark complex, tests/golden/pdb)."""
import json
import math
import os
import subprocess

import numpy as np
import pytest

from sparkledock_amd import ranking, sdgso

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "sparkledock_amd", "lightdock")
PDB_DIR = os.path.join(ROOT, "tests", "golden", "pdb")


def read_clusters(path):
    out = []
    for line in open(path):
        c, size, score, rep, name = line.rstrip("\n").split(":")
        assert name == "lightdock_%s.pdb" % rep
        out.append((int(c), int(size), float(score), int(rep)))
    return out


@pytest.mark.parametrize("form", ["fp64", "fp32_guard"])
def test_lightdock_writes_swarm_files(tmp_path, form):
    nr, nl, S, G, nm, steps, seed = 400, 150, 3, 24, 4, 5, 2
    setup = {"swarms": S, "glowworms": G, "use_anm": True, "anm_rec": nm, "anm_lig": nm,
             "swarm_radius": 3.0, "synthetic_receptor_atoms": nr, "synthetic_ligand_atoms": nl,
             "synthetic_seed": seed, "form": form, "rho": 0.4, "gamma": 0.6}
    (tmp_path / "setup.json").write_text(json.dumps(setup))
    r = subprocess.run([DRIVER, "-f", "setup.json", "-s", str(steps), "-l", "1"], cwd=tmp_path,
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "simulation execution:" in r.stdout
    res = ranking.read_run(str(tmp_path), steps)
    assert sorted(res) == list(range(S))
    for s in range(S):
        lines = (tmp_path / ("swarm_%d" % s) / ("gso_%d.out" % steps)).read_text().splitlines()
        assert lines[0] == ranking.HEADER
        assert len(lines) == G + 1
        assert all(line.endswith("\t") for line in lines[1:])

    # the same run through the C ABI with the inputs the driver derives
    rx, rt, rm = sdgso.synth_molecule(1000 * seed + 1, nr, nm)
    lx, lt, lm = sdgso.synth_molecule(1000 * seed + 2, nl, nm)
    table = sdgso.synth_table(7)
    centre = math.pow(nr * 11.4, 1 / 3) + 0.5 * math.pow(nl * 11.4, 1 / 3)
    pos = sdgso.synth_poses(1000 * seed + 3, S, G, nm, nm, centre, 3.0)
    ref = [0.0, 0.0, 0.0]  # the driver's centroid: sequential sum of x_i / n
    for i in range(nl):
        for k in range(3):
            ref[k] += float(lx[i, k]) / nl
    cx = sdgso.Complex(rx, rt, lx, lt, table, ref, rm, lm)
    run = sdgso.Run(cx, S, G, pos, sdgso.stream(0, S, steps, G), steps,
                    form=sdgso.FP64 if form == "fp64" else sdgso.FP32_GUARD)
    run.steps(steps, graph=True)
    got = run.read()
    for s in range(S):
        for g, rec in enumerate(res[s]):
            i = s * G + g
            pose, rec_id, lig_id, lucif, nnei, vis, score = rec
            assert (rec_id, lig_id) == (0, 0)
            assert np.allclose(pose, got["pos"][i], atol=1.5e-6)
            assert abs(lucif - got["luciferin"][i]) < 1.5e-6
            assert nnei == got["nnei"][i]
            assert abs(vis - got["vision"][i]) < 1.5e-6
            assert abs(score - got["score"][i]) < 1.5e-6
    # clusters: every glowworm once, representatives and sizes from the same clustering
    cl, rep = run.cluster(0, np.arange(nr, dtype=np.int32), np.arange(nl, dtype=np.int32), 4.0)
    for s in range(S):
        rows = read_clusters(tmp_path / ("swarm_%d" % s) / "cluster.repr")
        assert [r[0] for r in rows] == list(range(len(rows)))
        assert sum(r[1] for r in rows) == G
        for c, size, score, r in rows:
            assert size == int(np.sum(cl[s] == c)) and rep[s][cl[s] == c][0] == r
            assert abs(score - got["score"][s * G + r]) <= 1e-5 * max(1.0, abs(score))


def reference_line(pose, lucif, nnei, vis, score):
    """One glowworm's line as the reference writer builds it
    (prepare_gso_task.cu:2868-2888): std::to_string of every value, which
    prints a double as "%f"."""
    return ("(" + ", ".join("%f" % v for v in pose) + ")\t0\t0\t" + "%f\t%d\t%f\t%f\t" %
            (lucif, nnei, vis, score))


def test_lightdock_files_equal_oracle_run(tmp_path):
    """The driver's files against the oracle: lightdock in the FP64 form, and
    the oracle's run of the same inputs (the CPU policy, re-scoring moved
    glowworms only), formatted by the reference writer's rules; every line of
    every swarm_<i>/gso_<S>.out equal byte for byte."""
    import oracle_lib as O
    nr, nl, S, G, nm, steps, seed = 400, 150, 3, 24, 4, 6, 5
    setup = {"swarms": S, "glowworms": G, "use_anm": True, "anm_rec": nm, "anm_lig": nm,
             "swarm_radius": 3.0, "synthetic_receptor_atoms": nr, "synthetic_ligand_atoms": nl,
             "synthetic_seed": seed, "form": "fp64"}
    (tmp_path / "setup.json").write_text(json.dumps(setup))
    r = subprocess.run([DRIVER, "-f", "setup.json", "-s", str(steps)], cwd=tmp_path,
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    # the inputs the driver derives from the set-up (host functions only)
    rx, rt, rm = sdgso.synth_molecule(1000 * seed + 1, nr, nm)
    lx, lt, lm = sdgso.synth_molecule(1000 * seed + 2, nl, nm)
    centre = math.pow(nr * 11.4, 1 / 3) + 0.5 * math.pow(nl * 11.4, 1 / 3)
    pos = sdgso.synth_poses(1000 * seed + 3, S, G, nm, nm, centre, 3.0)
    ref = [0.0, 0.0, 0.0]
    for i in range(nl):
        for k in range(3):
            ref[k] += float(lx[i, k]) / nl
    ora = O.Complex(rx, rt, lx, lt, sdgso.synth_table(7), ref, rm, lm)
    st = O.run(ora, O.default_params(), pos, O.stream(0, S, steps, G), steps, S, G,
               rescore_all=False).get()
    assert st["moved"].sum() > 0
    for s in range(S):
        got = (tmp_path / ("swarm_%d" % s) / ("gso_%d.out" % steps)).read_bytes().decode()
        want = ranking.HEADER + "\n" + "".join(
            reference_line(st["pos"][s, g], st["luciferin"][s, g], st["nnei"][s, g],
                           st["vision"][s, g], st["score"][s, g]) + "\n" for g in range(G))
        assert got == want, s


def test_lightdock_usage_errors(tmp_path):
    r = subprocess.run([DRIVER], cwd=tmp_path, stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                       text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr
    (tmp_path / "setup.json").write_text(json.dumps({"swarms": 1, "glowworms": 4}))
    r = subprocess.run([DRIVER, "-f", "setup.json", "-s", "1"], cwd=tmp_path,
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=60)
    assert r.returncode == 1 and "synthetic" in r.stderr


@pytest.mark.parametrize("use_anm", [False, True])
def test_lightdock_pdb_inputs(tmp_path, use_anm):
    """The 2X9A complex from its PDB files: typing, mask, centring, normal
    modes and the reference point feed the run; results equal the same run
    through the C ABI; the final poses are written as PDB files."""
    S, G, steps = 2, 12, 3
    nm = 10 if use_anm else 0
    setup = {"swarms": S, "glowworms": G, "use_anm": use_anm, "anm_rec": 10, "anm_lig": 10,
             "anm_rec_rmsd": 0.5, "anm_lig_rmsd": 0.5, "swarm_radius": 5.0,
             "receptor_pdb": os.path.join(PDB_DIR, "2X9A_r_u.pdb"),
             "ligand_pdb": os.path.join(PDB_DIR, "2X9A_l_u.pdb"),
             "noh": True, "noxt": True, "now": True, "starting_points_seed": 11,
             "form": "fp32_guard", "write_pdbs": True}
    (tmp_path / "setup.json").write_text(json.dumps(setup))
    r = subprocess.run([DRIVER, "-f", "setup.json", "-s", str(steps)], cwd=tmp_path,
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    res = ranking.read_run(str(tmp_path), steps)
    assert sorted(res) == list(range(S))
    # the same inputs through the C ABI
    rs = sdgso.Structure.read_pdb(setup["receptor_pdb"])
    ls = sdgso.Structure.read_pdb(setup["ligand_pdb"])
    rx = rs.coords() + rs.centre()
    lx = ls.coords() + ls.centre()
    rm = lm = None
    if use_anm:
        rx, rm, _ = rs.anm(rx, nm, 0.5, S * G)
        lx, lm, _ = ls.anm(lx, nm, 0.5, S * G)
    ref = sdgso.reference_point(lx)
    table = sdgso.synth_table(7)
    cx = sdgso.Complex(rx, rs.dfire_types(), lx, ls.dfire_types(), table, ref, rm, lm,
                       rs.nm_mask(), ls.nm_mask())
    _, pos = sdgso.starting_poses(rs, rs.coords(), ls.coords(), rs.centre(), S, G, 5.0, 11,
                                  324324, nm, nm)
    run = sdgso.Run(cx, S, G, pos, sdgso.stream(0, S, steps, G), steps, form=sdgso.FP32_GUARD)
    run.steps(steps, graph=True)
    got = run.read()
    for s in range(S):
        for g, rec in enumerate(res[s]):
            assert np.allclose(rec[0], got["pos"][s * G + g], atol=1.5e-6)
            assert abs(rec[6] - got["score"][s * G + g]) < 1.5e-6
        rows = read_clusters(tmp_path / ("swarm_%d" % s) / "cluster.repr")
        assert sum(r[1] for r in rows) == G
    # the final poses as PDB files: receptor then ligand, at the run's final coordinates
    rc, lc = run.final_coords(G + 3, 1)
    back = sdgso.Structure.read_pdb(tmp_path / "swarm_1" / "lightdock_3.pdb")
    assert back.n_atoms == rs.n_atoms + ls.n_atoms
    assert np.max(np.abs(back.coords() - np.concatenate([rc[0], lc[0]]))) <= 5.1e-4
