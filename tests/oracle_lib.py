"""ctypes bindings to the CPU oracle (oracle/liboracle.so) and, where it was
built, the reference bridge (oracle/_ref/libref_shim.so).

Test infrastructure only: the oracle is the checker, never the product path.
"""
import ctypes
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SHIM_SO = os.path.join(ROOT, "oracle", "_ref", "libref_shim.so")

NTYPES = 168
NBINS = 20
TABLE_LEN = NTYPES * NTYPES * NBINS

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)
_I64 = ctypes.POINTER(ctypes.c_int64)
_U32 = ctypes.POINTER(ctypes.c_uint32)


def _p(a, t=_D):
    return a.ctypes.data_as(t) if a is not None else None


class OrComplex(ctypes.Structure):
    _fields_ = [
        ("nr", ctypes.c_int), ("nl", ctypes.c_int),
        ("nmr", ctypes.c_int), ("nml", ctypes.c_int),
        ("r_xyz", _D), ("l_xyz", _D),
        ("r_type", _I), ("l_type", _I),
        ("r_mask", _I), ("l_mask", _I),
        ("r_modes", _D), ("l_modes", _D),
        ("lig_ref", _D), ("table", _D),
    ]


class OrParams(ctypes.Structure):
    _fields_ = [
        ("rho", ctypes.c_double), ("gamma", ctypes.c_double), ("beta", ctypes.c_double),
        ("vision0", ctypes.c_double), ("max_vision", ctypes.c_double),
        ("luciferin0", ctypes.c_double), ("max_neighbors", ctypes.c_int),
        ("step_t", ctypes.c_double), ("step_r", ctypes.c_double), ("step_nm", ctypes.c_double),
    ]


class OrState(ctypes.Structure):
    _fields_ = [
        ("S", ctypes.c_int), ("G", ctypes.c_int), ("P", ctypes.c_int),
        ("pos", _D), ("luciferin", _D), ("score", _D), ("vision", _D), ("ref", _D),
        ("nnei", _I), ("moved", _I), ("step", _I), ("select", _I),
        ("move_pos", _D), ("pairs", _I64),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle/liboracle.so is not built (run `make -C oracle`)")
        L = ctypes.CDLL(ORACLE_SO)
        L.or_dist_bin.restype = ctypes.c_int
        L.or_dist_bin.argtypes = [ctypes.c_double]
        L.or_rotate.argtypes = [_D, _D, ctypes.c_int]
        L.or_dfire_energy.restype = ctypes.c_double
        L.or_dfire_energy.argtypes = [_D, ctypes.c_int, _I, _D, ctypes.c_int, _I, _D, _I64]
        L.or_score_from_energy.restype = ctypes.c_double
        L.or_score_from_energy.argtypes = [ctypes.c_double]
        L.or_pose_coords.argtypes = [ctypes.POINTER(OrComplex), _D, _D, _D, _D]
        L.or_score_pose.restype = ctypes.c_double
        L.or_score_pose.argtypes = [ctypes.POINTER(OrComplex), _D, _D, _I64]
        L.or_slerp.argtypes = [_D, _D, ctypes.c_double]
        L.or_move_extent.argtypes = [_D, _D, ctypes.c_int, ctypes.c_double]
        L.or_probabilities.argtypes = [_D, _I, _D, ctypes.c_int, ctypes.c_double]
        L.or_select.restype = ctypes.c_int
        L.or_select.argtypes = [ctypes.c_double, _I, _D, ctypes.c_int, ctypes.c_int]
        L.or_default_params.argtypes = [ctypes.POINTER(OrParams)]
        L.or_rmsd.restype = ctypes.c_double
        L.or_rmsd.argtypes = [_D, _D, ctypes.c_int]
        L.or_bsas.argtypes = [_D, _D, ctypes.c_int, ctypes.c_int, ctypes.c_double, _I, _I]
        L.or_mvee_centre.argtypes = [_D, ctypes.c_int, ctypes.c_double, _D]
        L.or_state_alloc.restype = ctypes.c_int
        L.or_state_alloc.argtypes = [ctypes.POINTER(OrState), ctypes.c_int, ctypes.c_int,
                                     ctypes.c_int, ctypes.c_int]
        L.or_state_free.argtypes = [ctypes.POINTER(OrState)]
        L.or_state_init.argtypes = [ctypes.POINTER(OrState), ctypes.POINTER(OrParams), _D]
        L.or_gso_step.argtypes = [ctypes.POINTER(OrComplex), ctypes.POINTER(OrParams),
                                  ctypes.POINTER(OrState), _D, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_int, ctypes.c_int]
        L.or_phase_score.argtypes = [ctypes.POINTER(OrComplex), ctypes.POINTER(OrParams),
                                     ctypes.POINTER(OrState), ctypes.c_int, ctypes.c_int]
        L.or_phase_select.argtypes = [ctypes.POINTER(OrParams), ctypes.POINTER(OrState), _D,
                                      ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.or_phase_move.argtypes = [ctypes.POINTER(OrParams), ctypes.POINTER(OrState),
                                    ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.or_stream.argtypes = [ctypes.c_uint32, ctypes.c_int, ctypes.c_int, ctypes.c_int, _D]
        L.or_mt19937.argtypes = [ctypes.c_uint32, ctypes.c_int, _U32]
        _lib = L
    return _lib


def default_params(**over):
    p = OrParams()
    lib().or_default_params(ctypes.byref(p))
    for k, v in over.items():
        setattr(p, k, v)
    return p


def pad_table(table):
    """FP64 copy of a flat 168*168*20 table plus the 20-entry zero pad."""
    t = np.zeros(TABLE_LEN + NBINS, dtype=np.float64)
    t[:TABLE_LEN] = np.asarray(table, dtype=np.float64).reshape(-1)
    return t


class Complex:
    """Owns the numpy arrays behind an OrComplex."""

    def __init__(self, r_xyz, r_type, l_xyz, l_type, table, lig_ref, r_modes=None, l_modes=None,
                 r_mask=None, l_mask=None):
        self.r_xyz = np.ascontiguousarray(r_xyz, dtype=np.float64).reshape(-1, 3)
        self.l_xyz = np.ascontiguousarray(l_xyz, dtype=np.float64).reshape(-1, 3)
        self.nr, self.nl = len(self.r_xyz), len(self.l_xyz)
        self.r_type = np.ascontiguousarray(r_type, dtype=np.int32)
        self.l_type = np.ascontiguousarray(l_type, dtype=np.int32)
        self.r_modes = (np.zeros((0, self.nr, 3)) if r_modes is None
                        else np.ascontiguousarray(r_modes, dtype=np.float64))
        self.l_modes = (np.zeros((0, self.nl, 3)) if l_modes is None
                        else np.ascontiguousarray(l_modes, dtype=np.float64))
        self.nmr, self.nml = self.r_modes.shape[0], self.l_modes.shape[0]
        self.r_mask = (np.ones(self.nr, np.int32) if r_mask is None
                       else np.ascontiguousarray(r_mask, dtype=np.int32))
        self.l_mask = (np.ones(self.nl, np.int32) if l_mask is None
                       else np.ascontiguousarray(l_mask, dtype=np.int32))
        self.lig_ref = np.ascontiguousarray(lig_ref, dtype=np.float64).reshape(3)
        self.table = pad_table(table)
        self.P = 7 + self.nmr + self.nml
        c = OrComplex()
        c.nr, c.nl, c.nmr, c.nml = self.nr, self.nl, self.nmr, self.nml
        c.r_xyz, c.l_xyz = _p(self.r_xyz), _p(self.l_xyz)
        c.r_type, c.l_type = _p(self.r_type, _I), _p(self.l_type, _I)
        c.r_mask, c.l_mask = _p(self.r_mask, _I), _p(self.l_mask, _I)
        c.r_modes, c.l_modes = _p(self.r_modes), _p(self.l_modes)
        c.lig_ref, c.table = _p(self.lig_ref), _p(self.table)
        self.c = c

    def score_pose(self, pose):
        pose = np.ascontiguousarray(pose, dtype=np.float64)
        ref = np.zeros(3)
        pairs = np.zeros(1, np.int64)
        s = lib().or_score_pose(ctypes.byref(self.c), _p(pose), _p(ref), _p(pairs, _I64))
        return s, ref, int(pairs[0])

    def pose_coords(self, pose):
        pose = np.ascontiguousarray(pose, dtype=np.float64)
        R = np.zeros((self.nr, 3))
        L = np.zeros((self.nl, 3))
        ref = np.zeros(3)
        lib().or_pose_coords(ctypes.byref(self.c), _p(pose), _p(R), _p(L), _p(ref))
        return R, L, ref


class State:
    """Owns an OrState."""

    def __init__(self, S, G, P, params, pos, with_pairs=True):
        self.st = OrState()
        self.S, self.G, self.P = S, G, P
        if lib().or_state_alloc(ctypes.byref(self.st), S, G, P, 1 if with_pairs else 0) != 0:
            raise MemoryError("or_state_alloc")
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1)
        assert pos.size == S * G * P
        lib().or_state_init(ctypes.byref(self.st), ctypes.byref(params), _p(pos))

    def __del__(self):
        try:
            lib().or_state_free(ctypes.byref(self.st))
        except Exception:
            pass

    def _arr(self, name, n, dt):
        ptr = getattr(self.st, name)
        return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dt, copy=True)

    def get(self):
        n = self.S * self.G
        out = {
            "pos": self._arr("pos", n * self.P, np.float64).reshape(self.S, self.G, self.P),
            "luciferin": self._arr("luciferin", n, np.float64).reshape(self.S, self.G),
            "score": self._arr("score", n, np.float64).reshape(self.S, self.G),
            "vision": self._arr("vision", n, np.float64).reshape(self.S, self.G),
            "ref": self._arr("ref", n * 3, np.float64).reshape(self.S, self.G, 3),
            "nnei": self._arr("nnei", n, np.int32).reshape(self.S, self.G),
            "moved": self._arr("moved", n, np.int32).reshape(self.S, self.G),
            "step": self._arr("step", n, np.int32).reshape(self.S, self.G),
            "select": self._arr("select", n, np.int32).reshape(self.S, self.G),
        }
        if self.st.pairs:
            out["pairs"] = self._arr("pairs", n, np.int64).reshape(self.S, self.G)
        return out

    def set(self, name, values):
        ptr = getattr(self.st, name)
        dst = np.ctypeslib.as_array(ptr, shape=(np.asarray(values).size,))
        dst[:] = np.asarray(values).reshape(-1)


def run(cx, params, pos, u, n_steps, S, G, rescore_all=True, threads=0, state=None):
    """Run n_steps GSO steps with the oracle; returns the final State."""
    st = state or State(S, G, cx.P, params, pos)
    u = np.ascontiguousarray(u, dtype=np.float64)
    for k in range(n_steps):
        lib().or_gso_step(ctypes.byref(cx.c), ctypes.byref(params), ctypes.byref(st.st), _p(u),
                          n_steps, k, 1 if rescore_all else 0, threads)
    return st


def stream(seed, S, n_steps, G):
    u = np.zeros(S * n_steps * G, dtype=np.float64)
    lib().or_stream(seed, S, n_steps, G, _p(u))
    return u.reshape(S, n_steps, G)


def mt19937(seed, n):
    out = np.zeros(n, dtype=np.uint32)
    lib().or_mt19937(seed, n, _p(out, _U32))
    return out


# ---------------------------------------------------------------- reference bridge
_ref = None


def ref_available():
    return os.path.exists(REF_SHIM_SO)


def ref():
    global _ref
    if _ref is None:
        R = ctypes.CDLL(REF_SHIM_SO, mode=os.RTLD_LAZY)
        R.ref_dfire_score.restype = ctypes.c_double
        R.ref_dfire_score.argtypes = [_D, _I, _I, _D, _D, ctypes.c_int, ctypes.c_int]
        R.ref_rotate.argtypes = [_D, _D, ctypes.c_int]
        R.ref_slerp.argtypes = [_D, _D, ctypes.c_double]
        R.ref_move_extent.argtypes = [_D, _D, ctypes.c_int, ctypes.c_double]
        R.ref_probabilities.argtypes = [_D, _I, _D, ctypes.c_int, ctypes.c_double]
        R.ref_select.restype = ctypes.c_int
        R.ref_select.argtypes = [ctypes.c_double, _I, _D, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        R.ref_std_normal.argtypes = [ctypes.c_uint, ctypes.c_int, ctypes.c_double, ctypes.c_double, _D]
        R.ref_kmeans.argtypes = [_D, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _D, _I]
        R.ref_rmsd.restype = ctypes.c_double
        R.ref_rmsd.argtypes = [_D, _D, ctypes.c_int]
        R.ref_mvee.argtypes = [_D, ctypes.c_int, ctypes.c_double, _D]
        R.ref_pdb_open.restype = ctypes.c_void_p
        R.ref_pdb_open.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        R.ref_pdb_close.argtypes = [ctypes.c_void_p]
        R.ref_pdb_natoms.restype = ctypes.c_int
        R.ref_pdb_natoms.argtypes = [ctypes.c_void_p]
        R.ref_pdb_nresidues.restype = ctypes.c_int
        R.ref_pdb_nresidues.argtypes = [ctypes.c_void_p]
        for name, t in (("ref_pdb_coords", _D), ("ref_pdb_residue_bounds", _I),
                        ("ref_pdb_types", _I), ("ref_pdb_mask", _I), ("ref_pdb_centre", _D)):
            getattr(R, name).argtypes = [ctypes.c_void_p, t]
        R.ref_pdb_atom_name.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_char_p]
        R.ref_pdb_write.argtypes = [ctypes.c_void_p, _D, ctypes.c_char_p]
        _ref = R
    return _ref


def rmsd(a, b):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    return lib().or_rmsd(_p(a), _p(b), len(a))


def bsas(coords, score, cutoff=4.0):
    """Oracle clustering of one swarm: coords [G][n][3], score [G] ->
    (cluster [G], representative [G])."""
    coords = np.ascontiguousarray(coords, dtype=np.float64)
    score = np.ascontiguousarray(score, dtype=np.float64)
    G, n = coords.shape[:2]
    cl = np.zeros(G, np.int32)
    rep = np.zeros(G, np.int32)
    lib().or_bsas(_p(coords), _p(score), G, n, float(cutoff), _p(cl, _I), _p(rep, _I))
    return cl, rep


def mvee_centre(xyz, tol=0.01):
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    c = np.zeros(3)
    lib().or_mvee_centre(_p(xyz), len(xyz), float(tol), _p(c))
    return c


class RefStructure:
    """A structure read by the reference's own reader (oracle/_ref)."""

    def __init__(self, path, noxt=True, noh=True, now=True):
        self.h = ref().ref_pdb_open(path.encode(), int(noxt), int(noh), int(now))
        self.n = ref().ref_pdb_natoms(self.h)
        self.nres = ref().ref_pdb_nresidues(self.h)

    def coords(self):
        x = np.zeros((self.n, 3))
        ref().ref_pdb_coords(self.h, _p(x))
        return x

    def residue_bounds(self):
        b = np.zeros(self.nres + 1, np.int32)
        ref().ref_pdb_residue_bounds(self.h, _p(b, _I))
        return b

    def names(self):
        buf = ctypes.create_string_buffer(8)
        out = []
        for i in range(self.n):
            ref().ref_pdb_atom_name(self.h, i, buf)
            out.append(buf.value.decode())
        return out

    def types(self):
        t = np.zeros(self.n, np.int32)
        ref().ref_pdb_types(self.h, _p(t, _I))
        return t

    def mask(self):
        m = np.zeros(self.n, np.int32)
        ref().ref_pdb_mask(self.h, _p(m, _I))
        return m

    def centre(self):
        c = np.zeros(3)
        ref().ref_pdb_centre(self.h, _p(c))
        return c

    def write(self, xyz, path):
        xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        ref().ref_pdb_write(self.h, _p(xyz), path.encode())

    def close(self):
        if self.h:
            ref().ref_pdb_close(self.h)
            self.h = None


def ref_cpu_run(cx, pos, S, G, workdir):
    """Run the reference's cal_gso_tasks_cpu (3 compiled steps, 10+10 modes) in a
    child process; returns the final poses [S][G][27]."""
    assert cx.nmr == 10 and cx.nml == 10
    inp = os.path.join(workdir, "ref_in.bin")
    out = os.path.join(workdir, "ref_out.bin")
    with open(inp, "wb") as f:
        f.write(np.array([cx.nr, cx.nl, S, G], dtype=np.int32).tobytes())
        f.write(cx.r_xyz.tobytes())
        f.write(cx.r_modes.tobytes())
        f.write(cx.r_mask.tobytes())
        f.write(cx.r_type.tobytes())
        f.write(cx.l_xyz.tobytes())
        f.write(cx.l_modes.tobytes())
        f.write(cx.l_mask.tobytes())
        f.write(cx.l_type.tobytes())
        f.write(cx.lig_ref.tobytes())
        f.write(cx.table[:TABLE_LEN].tobytes())
        f.write(np.ascontiguousarray(pos, dtype=np.float64).tobytes())
    code = ("import ctypes, os, sys; l = ctypes.CDLL(sys.argv[1], mode=os.RTLD_LAZY); "
            "l.ref_cpu_run(sys.argv[2].encode(), sys.argv[3].encode())")
    env = dict(os.environ, NUM_THREADS="1")
    r = subprocess.run([sys.executable, "-c", code, REF_SHIM_SO, inp, out], env=env,
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE, timeout=600)
    if r.returncode != 0 or not os.path.exists(out):
        raise RuntimeError("reference run failed: rc=%d %s" % (r.returncode, r.stderr[-2000:]))
    return np.fromfile(out, dtype=np.float64).reshape(S, G, 27), r.stdout.decode(errors="replace")
