"""Structure input/output (PDB reading, DFIRE typing, normal-mode mask,
centring, PDB writing), the ligand reference point and the clustering
oracle, on the host (no GPU).

The structure functions are compared with the reference's own reader
(read_input_structure.cc, complex.cc, initialization.cc) and writer
(write_pdb, prepare_gso_tasks.cc:1279-1349) compiled into oracle/_ref, on the
ten benchmark complexes the reference ships (bm/<id>/<id>_{r,l}_u.pdb, the
complex sizes of the batched workload).  Those comparisons run where the
reference sources are present and are skipped elsewhere.
"""
import glob
import math
import os

import numpy as np
import pytest

import oracle_lib as O
from sparkledock_amd import sdgso

BM = "/root/reference/bm"
PDBS = sorted(glob.glob(os.path.join(BM, "*", "*_u.pdb")))
need_ref = pytest.mark.skipif(not (PDBS and os.path.exists(O.REF_SHIM_SO)),
                              reason="reference sources / oracle/_ref not present")


@need_ref
@pytest.mark.parametrize("path", PDBS, ids=[os.path.basename(p) for p in PDBS])
def test_pdb_reading_matches_reference(path):
    ours = sdgso.Structure.read_pdb(path)
    ref = O.RefStructure(path)
    try:
        assert ours.n_atoms == ref.n
        assert ours.n_residues == ref.nres
        assert np.array_equal(ours.coords(), ref.coords())
        names = ref.names()
        assert [ours.atom(i)[0] for i in range(ours.n_atoms)] == names
        bounds = ref.residue_bounds()
        # residue boundaries: an atom starts a residue where the reference's does
        starts = {int(b) for b in bounds[:-1]}
        ours_starts = set()
        last = None
        for i in range(ours.n_atoms):
            name, res, num, chain = ours.atom(i)
            key = (res, num)
            if key != last:
                ours_starts.add(i)
                last = key
        assert starts == ours_starts
        assert np.array_equal(ours.dfire_types(), ref.types())
        assert np.array_equal(ours.nm_mask(), ref.mask())
        assert np.array_equal(ours.centre(), ref.centre())
        assert len(ours.backbone()) == sum(1 for n in names if n in ("CA", "P"))
    finally:
        ref.close()
        ours.close()


@need_ref
@pytest.mark.parametrize("path", PDBS[::4], ids=[os.path.basename(p) for p in PDBS[::4]])
def test_pdb_writer_matches_reference(path, tmp_path):
    ours = sdgso.Structure.read_pdb(path)
    ref = O.RefStructure(path)
    rng = np.random.default_rng(3)
    xyz = ours.coords() + rng.normal(size=(ours.n_atoms, 3)) * 40.0
    xyz[::7] *= -1.0
    a, b = tmp_path / "ours.pdb", tmp_path / "ref.pdb"
    ours.write_pdb(a, xyz)
    ref.write(xyz, str(b))
    assert a.read_bytes() == b.read_bytes()
    # and the written file reads back to the same atoms within the 3 printed decimals
    back = sdgso.Structure.read_pdb(a)
    assert back.n_atoms == ours.n_atoms
    assert np.max(np.abs(back.coords() - xyz)) <= 5.1e-4
    ref.close()
    ours.close()


def test_pdb_fields_and_filters(tmp_path):
    lines = [
        "HEADER    TEST",
        "MODEL        1",
        "ATOM      1  N   ALA A   1      11.104   6.134  -6.504  1.00  0.00           N",
        "ATOM      2  CA  ALA A   1      11.639   6.071  -5.147  1.00 15.25           C",
        "ATOM      3  H   ALA A   1      12.000   6.000  -5.000  1.00  0.00           H",
        "ATOM      4  CB  ALA A   1      12.500   7.200  -4.800  0.50  9.99           C",
        "ATOM      5  OXT ALA A   1      13.000   7.000  -4.000  1.00  0.00           O",
        "HETATM    6  O   HOH W 101       1.000   2.000   3.000  1.00  0.00           O",
        "ATOM      7  N   GLY B   2A     -1.250  -2.500 100.125                       N",
        "ATOM      8  CA  GLY B   2A     -1.750  -2.000 101.000  1.00  0.00           C",
        "MODEL        2",
        "ATOM      9  CA  GLY B   3      -9.000  -9.000  -9.000  1.00  0.00           C",
    ]
    p = tmp_path / "t.pdb"
    p.write_text("\n".join(lines) + "\n")
    s = sdgso.Structure.read_pdb(p)
    assert s.n_atoms == 5 and s.n_residues == 2 and s.n_chains == 2
    assert [s.atom(i)[0] for i in range(5)] == ["N", "CA", "CB", "N", "CA"]
    assert s.atom(3) == ("N", "GLY", 2, "B")
    x = s.coords()
    assert x[1, 0] == np.float64(np.float32(11.639))
    assert list(s.dfire_types()) == [74, 75, 78, 79, 80]
    assert list(s.nm_mask()) == [1] * 5
    assert list(s.backbone()) == [1, 4]
    t = s.centre()
    assert np.allclose(t, -x.mean(axis=0), rtol=0, atol=1e-12)
    # everything kept without the filters (hydrogen, OXT, water)
    s2 = sdgso.Structure.read_pdb(p, noxt=False, noh=False, now=False)
    assert s2.n_atoms == 8 and s2.n_residues == 3
    with pytest.raises(sdgso.SdgsoError):
        s2.dfire_types()  # H and OXT have no DFIRE type, water neither
    m = s2.nm_mask()
    assert list(m) == [1, 1, 1, 1, 1, 0, 1, 1]
    # written lines keep the reference's columns
    out = tmp_path / "w.pdb"
    s.write_pdb(out)
    w = out.read_text().splitlines()
    assert w[0] == "ATOM      1  N   ALA A   1      11.104   6.134  -6.504  1.00  0.00           N"
    assert w[3] == "ATOM      4  N   GLY B   2A     -1.250  -2.500 100.125  1.00  0.00           N"
    s.close()
    s2.close()


def test_synthetic_structure_metadata():
    xyz = np.arange(12, dtype=np.float64).reshape(4, 3)
    s = sdgso.Structure.synthetic(xyz)
    assert s.n_atoms == 4 and s.n_residues == 4
    assert s.atom(2) == ("CA", "ALA", 3, "A")
    assert list(s.backbone()) == [0, 1, 2, 3]
    assert list(s.dfire_types()) == [75] * 4
    assert np.array_equal(s.coords(), xyz)


def test_bad_pdb_inputs(tmp_path):
    with pytest.raises(sdgso.SdgsoError):
        sdgso.Structure.read_pdb(tmp_path / "missing.pdb")
    p = tmp_path / "x.pdb"
    p.write_text("ATOM      1  QQ  ALA A   1      11.104   6.134  -6.504  1.00  0.00\n")
    with pytest.raises(sdgso.SdgsoError):
        sdgso.Structure.read_pdb(p)
    p.write_text("REMARK nothing\n")
    with pytest.raises(sdgso.SdgsoError):
        sdgso.Structure.read_pdb(p)


# ---------------------------------------------------------------- reference point

def test_reference_point_closed_forms():
    # the vertices of a box: the enclosing ellipsoid is centred on the box
    v = np.array([[x, y, z] for x in (-1.0, 3.0) for y in (2.0, 4.0) for z in (-7.0, -5.0)])
    for c in (sdgso.reference_point(v, 1e-7), O.mvee_centre(v, 1e-7)):
        assert np.allclose(c, [1.0, 3.0, -6.0], atol=1e-6)
    rng = np.random.default_rng(4)
    pts = rng.normal(size=(300, 3)) * [6.0, 3.0, 2.0]
    a = sdgso.reference_point(pts)
    b = sdgso.reference_point(pts + [10.0, -4.0, 2.5])
    assert np.allclose(b - a, [10.0, -4.0, 2.5], atol=1e-9)
    assert np.allclose(a, O.mvee_centre(pts), atol=1e-9)


@need_ref
def test_reference_point_matches_reference():
    small = sorted(glob.glob(os.path.join(BM, "*", "*_l_u.pdb")), key=os.path.getsize)[:3]
    for path in small:
        s = sdgso.Structure.read_pdb(path)
        xyz = s.coords() + s.centre()
        want = np.zeros(3)
        O.ref().ref_mvee(O._p(np.ascontiguousarray(xyz)), len(xyz), 0.01, O._p(want))
        assert np.allclose(sdgso.reference_point(xyz), want, rtol=0, atol=1e-9), path
        assert np.allclose(O.mvee_centre(xyz), want, rtol=0, atol=1e-9), path
        s.close()


# ---------------------------------------------------------------- clustering oracle

@need_ref
def test_rmsd_matches_reference():
    rng = np.random.default_rng(5)
    for n in (1, 7, 300):
        a, b = rng.normal(size=(n, 3)) * 20, rng.normal(size=(n, 3)) * 20
        want = O.ref().ref_rmsd(O._p(a), O._p(b), n)
        assert O.rmsd(a, b) == want


def test_rmsd_closed_forms():
    rng = np.random.default_rng(6)
    a = rng.normal(size=(50, 3))
    assert O.rmsd(a, a) == 0.0
    assert abs(O.rmsd(a, a + [3.0, 4.0, 0.0]) - 5.0) < 1e-12


def test_bsas_hand_cases():
    rng = np.random.default_rng(7)
    base = rng.normal(size=(20, 3)) * 5
    # three far groups of poses, small jitter inside a group
    groups = [0, 1, 2, 0, 1, 2, 0, 1, 2, 0]
    coords = np.array([base + 50.0 * g + rng.normal(size=base.shape) * 0.1 for g in groups])
    score = rng.normal(size=len(groups))
    cl, rep = O.bsas(coords, score)
    for g in range(3):
        members = [i for i, x in enumerate(groups) if x == g]
        best = max(members, key=lambda i: (score[i], -i))
        assert all(rep[i] == best for i in members)
        assert len({cl[i] for i in members}) == 1
    assert sorted(set(cl)) == [0, 1, 2]
    # a chain of poses 3 A apart: the greedy pass in score order decides
    chain = np.array([base + [3.0 * i, 0.0, 0.0] for i in range(5)])
    sc = np.array([5.0, 4.0, 3.0, 2.0, 1.0])  # best first: 0 founds, 1 joins, 2 founds, 3 joins 2, 4 founds
    cl, rep = O.bsas(chain, sc)
    assert list(cl) == [0, 0, 1, 1, 2] and list(rep) == [0, 0, 2, 2, 4]
    # equal scores keep index order
    cl, rep = O.bsas(chain, np.zeros(5))
    assert list(rep) == [0, 0, 2, 2, 4]


# ---------------------------------------------------------------- normal-mode oracle

def _anm_oracle():
    import importlib.util
    spec = importlib.util.spec_from_file_location(
        "anm_oracle", os.path.join(os.path.dirname(O.__file__), "..", "oracle", "anm_oracle.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def test_minstd_known_answer():
    A = _anm_oracle()
    g = A.MinStdRand0(1)
    for _ in range(9999):
        g()
    assert g() == 1043618065  # C++11 [rand.predef]: 10000th output of minstd_rand0


@need_ref
def test_normal_distribution_matches_libstdcxx():
    A = _anm_oracle()
    want = np.zeros(501)
    O.ref().ref_std_normal(324324, len(want), 0.0, 1.0, O._p(want))
    g = A.MinStdRand0(324324)
    nd = A.NormalDistribution(0.0, 1.0)
    assert np.array_equal(np.array([nd(g) for _ in range(len(want))]), want)


def test_anm_hessian_properties():
    A = _anm_oracle()
    rng = np.random.default_rng(8)
    c = rng.normal(size=(30, 3)) * 3.0  # dense enough that the network is rigid
    H = A.hessian(c)
    assert np.array_equal(H, H.T)
    assert np.allclose(H.sum(axis=1), 0.0, atol=1e-12)
    w = np.linalg.eigvalsh(H)
    assert np.sum(np.abs(w) < 1e-8) == 6  # rigid-body modes of a connected network
    # a rigid translation is a zero mode exactly
    t = np.tile([0.3, -1.2, 0.7], 30)
    assert np.allclose(H @ t, 0.0, atol=1e-12)


# ---------------------------------------------------------------- starting poses

PDB_FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "pdb")


@need_ref
def test_kmeans_matches_reference():
    rng = np.random.default_rng(9)
    for n, k, it in ((500, 7, 100), (2000, 40, 30), (64, 63, 5)):
        pts = rng.normal(size=(n, 3)) * 12.0
        c, g = sdgso.kmeans(pts, k, seed=324324, iterations=it)
        wc = np.zeros((k, 3))
        wg = np.zeros(n, np.int32)
        O.ref().ref_kmeans(O._p(pts), n, k, 324324, it, O._p(wc), O._p(wg, O._I))
        assert np.array_equal(c, wc) and np.array_equal(g, wg)


def _sphere_points(n):
    """cal_anm.cc:977-989"""
    inc = math.pi * (3.0 - math.sqrt(5.0))
    off = 2.0 / n
    out = np.zeros((n, 3))
    for i in range(n):
        y = i * off - 1.0 + (off / 2.0)
        r = math.sqrt(1 - y * y)
        phi = i * inc
        out[i] = (math.cos(phi) * r, y, math.sin(phi) * r)
    return out


def _dist3(a, b):
    d = a - b
    return np.sqrt((d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2])


@need_ref
def test_starting_poses_restated():
    """calculate_surface_point / calculate_start_pos restated with the
    reference's k-means, the oracle's mt19937 stream and libstdc++ normals."""
    rs = sdgso.Structure.read_pdb(os.path.join(PDB_FIX, "2X9A_r_u.pdb"))
    ls = sdgso.Structure.read_pdb(os.path.join(PDB_FIX, "2X9A_l_u.pdb"))
    rx, lx, t = rs.coords(), ls.coords(), rs.centre()
    S, G, radius, nm = 12, 9, 10.0, 3
    centres, pos = sdgso.starting_poses(rs, rx, lx, t, S, G, radius, 1234, 77, nm, nm)
    # ligand diameter -> surface distance
    diam = max(_dist3(lx[i], lx[i + 1:]).max() for i in range(len(lx) - 1))
    sd = 10.0 if diam < 20.0 else diam / 4.0
    surface = {"ARG", "ASN", "ASP", "ASX", "CSO", "GLN", "GLU", "GLX", "GLY", "HID", "HIE", "HIP",
               "IS", "HSD", "HSE", "HSP", "LYS", "PHD", "PRO", "PTR", "SEP", "SER", "THR", "TPO",
               "TYR", "HIS"}
    surf = np.array([rx[i] for i in range(rs.n_atoms) if rs.atom(i)[1] in surface])

    def ref_kmeans(p, k, it):
        c = np.zeros((k, 3))
        g = np.zeros(len(p), np.int32)
        p = np.ascontiguousarray(p)
        O.ref().ref_kmeans(O._p(p), len(p), k, 324324, it, O._p(c), O._p(g, O._I))
        return c

    cent = ref_kmeans(surf, S, 100)
    sph = _sphere_points(100)
    samp = np.array([[sph[j] * sd + cent[i] for j in range(100)] for i in range(S)])
    drop = np.zeros((S, 100), bool)
    for i in range(S):
        for j in range(S):
            if j != i and _dist3(cent[i], cent[j]) <= 20.0:
                drop[i] |= _dist3(samp[i], cent[j]) <= sd
    kept = samp[~drop]
    outer = np.array([p for p in kept if not np.any(_dist3(rx, p) <= 3.0)])
    want_c = ref_kmeans(outer, S, 500) + t
    assert np.array_equal(centres, want_c)
    # poses from the mt19937 stream and libstdc++ normals
    u = iter(O.stream(1234, 1, 1, 20 * S * G).ravel())
    nrm = np.zeros(2 * nm * S * G)
    O.ref().ref_std_normal(77, len(nrm), 0.0, 1.0, O._p(nrm))
    nrm = iter(nrm)
    for i in range(S):
        for j in range(G):
            while True:
                x, y, z = ((2 * next(u) - 1) * radius for _ in range(3))
                if x * x + y * y + z * z <= radius * radius:
                    break
            u1, u2, u3 = next(u), next(u), next(u)
            q = [math.sqrt(1 - u1) * math.sin(2 * math.pi * u2), math.sqrt(1 - u1) * math.cos(2 * math.pi * u2),
                 math.sqrt(u1) * math.sin(2 * math.pi * u3), math.sqrt(u1) * math.cos(2 * math.pi * u3)]
            want = [want_c[i][0] + x, want_c[i][1] + y, want_c[i][2] + z] + q + [next(nrm) for _ in range(2 * nm)]
            assert np.array_equal(pos[i, j], np.array(want)), (i, j)
    rs.close()
    ls.close()


def test_starting_poses_properties():
    rs = sdgso.Structure.read_pdb(os.path.join(PDB_FIX, "2X9A_r_u.pdb"))
    ls = sdgso.Structure.read_pdb(os.path.join(PDB_FIX, "2X9A_l_u.pdb"))
    c, pos = sdgso.starting_poses(rs, rs.coords(), ls.coords(), rs.centre(), 20, 15, 8.0)
    assert pos.shape == (20, 15, 7)
    assert np.allclose(np.linalg.norm(pos[:, :, 3:7], axis=2), 1.0, atol=1e-12)
    assert np.all(np.linalg.norm(pos[:, :, :3] - c[:, None, :], axis=2) <= 8.0 + 1e-12)
    # centres lie outside the centred receptor's atoms on average distance
    rc = rs.coords() + rs.centre()
    assert np.min([np.min(np.linalg.norm(rc - x, axis=1)) for x in c]) > 1.0
    with pytest.raises(sdgso.SdgsoError):
        sdgso.starting_poses(rs, rs.coords(), ls.coords(), rs.centre(), 10**5, 2)
