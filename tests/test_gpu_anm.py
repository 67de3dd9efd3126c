"""Normal modes of real structures (the 2X9A fixture) by the library's ANM
(rocSOLVER eigenpairs) against the numpy restatement of cal_anm.cc
(oracle/anm_oracle.py, LAPACK eigenpairs)."""
import importlib.util
import os

import numpy as np
import pytest

from sparkledock_amd import sdgso

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PDB_DIR = os.path.join(ROOT, "tests", "golden", "pdb")


def anm_oracle():
    spec = importlib.util.spec_from_file_location("anm_oracle",
                                                  os.path.join(ROOT, "oracle", "anm_oracle.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def residues_of(s):
    out = []
    last = None
    for i in range(s.n_atoms):
        _, res, num, chain = s.atom(i)
        if (res, num) != last:
            out.append([i, 0])
            last = (res, num)
        out[-1][1] += 1
    return [tuple(r) for r in out]


@pytest.mark.parametrize("name,n_modes", [("2X9A_r_u.pdb", 10), ("2X9A_l_u.pdb", 10),
                                          ("2X9A_l_u.pdb", 20)])
def test_anm_matches_oracle(name, n_modes):
    A = anm_oracle()
    s = sdgso.Structure.read_pdb(os.path.join(PDB_DIR, name))
    xyz = s.coords() + s.centre()
    names = [s.atom(i)[0] for i in range(s.n_atoms)]
    x, modes, ev = s.anm(xyz, n_modes, rmsd=0.5, n_draws=21)
    xo, mo, evo = A.anm(xyz, names, residues_of(s), n_modes, 0.5, 21)
    assert np.array_equal(x, xo)
    assert np.all(ev > 1e-6) and np.all(np.diff(ev) >= 0)
    assert np.allclose(ev, evo, rtol=1e-9, atol=0)
    assert np.max(np.abs(modes - mo)) <= 1e-6 * np.max(np.abs(mo))
    # every atom of a residue moves with its node; masked-free structure moves everywhere
    assert np.all(np.linalg.norm(modes, axis=2).max(axis=0) > 0)


def test_anm_errors():
    s = sdgso.Structure.synthetic(np.random.default_rng(1).normal(size=(5, 3)) * 3)
    with pytest.raises(sdgso.SdgsoError):
        s.anm(s.coords(), 20)  # 15 degrees of freedom cannot hold 20 + 6 modes
