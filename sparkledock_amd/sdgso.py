"""ctypes binding of libsdgso.so (include/sdgso.h).

The library is the product path: every call here goes to the HIP kernels.  If
libsdgso.so is missing, `lib()` raises; there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsdgso.so")

FP64 = 0
FP32_GUARD = 1
NTYPES = 168
NBINS = 20
TABLE_LEN = NTYPES * NTYPES * NBINS

_D = ctypes.POINTER(ctypes.c_double)
_F = ctypes.POINTER(ctypes.c_float)
_I = ctypes.POINTER(ctypes.c_int)
_I64 = ctypes.POINTER(ctypes.c_int64)
_VP = ctypes.c_void_p


class Params(ctypes.Structure):
    _fields_ = [
        ("rho", ctypes.c_double), ("gamma", ctypes.c_double), ("beta", ctypes.c_double),
        ("vision0", ctypes.c_double), ("max_vision", ctypes.c_double),
        ("luciferin0", ctypes.c_double), ("max_neighbors", ctypes.c_int),
        ("step_translation", ctypes.c_double), ("step_rotation", ctypes.c_double),
        ("step_nmodes", ctypes.c_double),
    ]


class SdgsoError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("sdgso error %d: %s" % (code, msg))
        self.code = code


# (name, restype, argtypes): every symbol include/sdgso.h declares
SIGNATURES = [
    ("sdgso_last_error", ctypes.c_char_p, []),
    ("sdgso_version", ctypes.c_int, []),
    ("sdgso_default_params", None, [ctypes.POINTER(Params)]),
    ("sdgso_complex_create", ctypes.c_int,
     [ctypes.c_int, _D, _I, _I, ctypes.c_int, _D, ctypes.c_int, _D, _I, _I, ctypes.c_int, _D, _D,
      _D, ctypes.POINTER(_VP)]),
    ("sdgso_complex_destroy", ctypes.c_int, [_VP]),
    ("sdgso_score_poses", ctypes.c_int,
     [_VP, ctypes.c_int, _D, ctypes.c_int, ctypes.c_double, _D, _D, _I64]),
    # device pointers and a hipStream_t
    ("sdgso_score_poses_noalloc", ctypes.c_int,
     [_VP, ctypes.c_int, _VP, ctypes.c_int, ctypes.c_double, _VP, _VP, _VP, _VP]),
    ("sdgso_run_create", ctypes.c_int,
     [ctypes.c_int, ctypes.POINTER(_VP), _I, ctypes.c_int, _D, _D, ctypes.c_int,
      ctypes.POINTER(Params), ctypes.c_int, ctypes.c_double, ctypes.POINTER(_VP)]),
    ("sdgso_run_destroy", ctypes.c_int, [_VP]),
    ("sdgso_run_set_stream", ctypes.c_int, [_VP, _VP]),
    ("sdgso_run_steps", ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_int]),
    ("sdgso_run_phases", ctypes.c_int, [_VP, ctypes.c_int]),
    ("sdgso_run_sync", ctypes.c_int, [_VP]),
    ("sdgso_run_step_index", ctypes.c_int, [_VP]),
    ("sdgso_run_read", ctypes.c_int, [_VP, _D, _D, _D, _D, _I, _I, _I, _D, _I64]),
    ("sdgso_run_write", ctypes.c_int, [_VP, _D, _D, _D, _D, _I, _D]),
    ("sdgso_run_profile", ctypes.c_int, [_VP, ctypes.c_int]),
    ("sdgso_run_profile_read", ctypes.c_int, [_VP, _D, _I64]),
    ("sdgso_run_guard_stats", ctypes.c_int, [_VP, _I64]),
    ("sdgso_run_set_work_order", ctypes.c_int, [_VP, ctypes.c_int]),
    ("sdgso_run_work_order", ctypes.c_int, [_VP, _I, _D]),
    ("sdgso_run_set_ligand_gemm", ctypes.c_int, [_VP, ctypes.c_int]),
    ("sdgso_run_ligand_gemm", ctypes.c_int, [_VP, _I, _I64, _D]),
    ("sdgso_run_ligand_displacements", ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_int, _F]),
    ("sdgso_tune_variants", ctypes.c_int, []),
    ("sdgso_tune_variant_name", ctypes.c_char_p, [ctypes.c_int]),
    ("sdgso_run_tune", ctypes.c_int, [_VP, ctypes.c_int, _D, _I]),
    ("sdgso_run_filter_stats", ctypes.c_int, [_VP, _I64]),
    ("sdgso_run_set_variant", ctypes.c_int, [_VP, ctypes.c_int]),
    ("sdgso_run_variants", ctypes.c_int, [_VP, _I]),
    ("sdgso_probe_gather", ctypes.c_int, [ctypes.c_int64, _D]),
    ("sdgso_probe_anm_gemm", ctypes.c_int, [_VP, _D]),
    ("sdgso_run_final_coords", ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_int, _D, _D]),
    ("sdgso_run_cluster", ctypes.c_int,
     [_VP, ctypes.c_int, _I, ctypes.c_int, _I, ctypes.c_int, ctypes.c_double, _I, _I]),
    ("sdgso_structure_read_pdb", ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(_VP)]),
    ("sdgso_structure_synthetic", ctypes.c_int, [ctypes.c_int, _D, ctypes.POINTER(_VP)]),
    ("sdgso_structure_destroy", ctypes.c_int, [_VP]),
    ("sdgso_structure_size", ctypes.c_int, [_VP, _I, _I, _I]),
    ("sdgso_structure_coords", ctypes.c_int, [_VP, _D]),
    ("sdgso_structure_atom", ctypes.c_int,
     [_VP, ctypes.c_int, ctypes.c_char_p, ctypes.c_char_p, _I, ctypes.c_char_p]),
    ("sdgso_structure_dfire_types", ctypes.c_int, [_VP, _I]),
    ("sdgso_structure_nm_mask", ctypes.c_int, [_VP, _I]),
    ("sdgso_structure_backbone", ctypes.c_int, [_VP, _I, _I]),
    ("sdgso_structure_centre", ctypes.c_int, [_VP, _D]),
    ("sdgso_structure_write_pdb", ctypes.c_int, [_VP, _D, ctypes.c_char_p, ctypes.c_int]),
    ("sdgso_reference_point", ctypes.c_int, [ctypes.c_int, _D, ctypes.c_double, _D]),
    ("sdgso_kmeans", ctypes.c_int,
     [_D, ctypes.c_int, ctypes.c_int, ctypes.c_uint32, ctypes.c_int, _D, _I]),
    ("sdgso_starting_poses", ctypes.c_int,
     [_VP, _D, ctypes.c_int, _D, _D, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_uint32,
      ctypes.c_uint32, ctypes.c_int, ctypes.c_int, _D, _D]),
    ("sdgso_structure_anm", ctypes.c_int,
     [_VP, _D, ctypes.c_int, ctypes.c_double, ctypes.c_int, _D, _D]),
    ("sdgso_stream", ctypes.c_int,
     [ctypes.c_uint32, ctypes.c_int, ctypes.c_int, ctypes.c_int, _D]),
    ("sdgso_synth_molecule", ctypes.c_int,
     [ctypes.c_uint64, ctypes.c_int, ctypes.c_int, _D, _I, _D]),
    ("sdgso_synth_table", ctypes.c_int, [ctypes.c_uint64, _F]),
    ("sdgso_synth_poses", ctypes.c_int,
     [ctypes.c_uint64, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
      ctypes.c_double, _D]),
]

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libsdgso.so is not built: run `python -m sparkledock_amd.build` "
                               "(there is no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise SdgsoError(rc, lib().sdgso_last_error().decode(errors="replace"))


def _p(a, t=_D):
    return None if a is None else a.ctypes.data_as(t)


def default_params(**over):
    p = Params()
    lib().sdgso_default_params(ctypes.byref(p))
    for k, v in over.items():
        setattr(p, k, v)
    return p


class Complex:
    """One docking problem resident on the device."""

    def __init__(self, r_xyz, r_type, l_xyz, l_type, table, lig_ref, r_modes=None, l_modes=None,
                 r_mask=None, l_mask=None):
        r_xyz = np.ascontiguousarray(r_xyz, dtype=np.float64).reshape(-1, 3)
        l_xyz = np.ascontiguousarray(l_xyz, dtype=np.float64).reshape(-1, 3)
        self.nr, self.nl = len(r_xyz), len(l_xyz)
        r_type = np.ascontiguousarray(r_type, dtype=np.int32)
        l_type = np.ascontiguousarray(l_type, dtype=np.int32)
        r_mask = None if r_mask is None else np.ascontiguousarray(r_mask, dtype=np.int32)
        l_mask = None if l_mask is None else np.ascontiguousarray(l_mask, dtype=np.int32)
        self.nmr = 0 if r_modes is None else int(np.asarray(r_modes).shape[0])
        self.nml = 0 if l_modes is None else int(np.asarray(l_modes).shape[0])
        rm = None if not self.nmr else np.ascontiguousarray(r_modes, dtype=np.float64)
        lm = None if not self.nml else np.ascontiguousarray(l_modes, dtype=np.float64)
        ref = np.ascontiguousarray(lig_ref, dtype=np.float64).reshape(3)
        # FP64, as the reference's table; an FP32 table converts exactly
        tab = np.ascontiguousarray(table, dtype=np.float64).reshape(-1)
        if tab.size != TABLE_LEN:
            raise ValueError("table must have 168*168*20 entries")
        self.P = 7 + self.nmr + self.nml
        h = _VP()
        _check(lib().sdgso_complex_create(
            self.nr, _p(r_xyz), _p(r_type, _I), _p(r_mask, _I), self.nmr, _p(rm),
            self.nl, _p(l_xyz), _p(l_type, _I), _p(l_mask, _I), self.nml, _p(lm),
            _p(ref), _p(tab), ctypes.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().sdgso_complex_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def score(self, poses, form=FP64, band=1e-4):
        poses = np.ascontiguousarray(poses, dtype=np.float64).reshape(-1, self.P)
        n = len(poses)
        s = np.zeros(n)
        ref = np.zeros((n, 3))
        pairs = np.zeros(n, np.int64)
        _check(lib().sdgso_score_poses(self.h, n, _p(poses), form, band, _p(s), _p(ref),
                                       _p(pairs, _I64)))
        return s, ref, pairs


class Run:
    """A GSO run over one or more complexes, state resident on the device."""

    def __init__(self, complexes, n_swarms, G, pos, stream, n_steps, params=None, form=FP64,
                 band=1e-4):
        if isinstance(complexes, Complex):
            complexes = [complexes]
            n_swarms = [n_swarms]
        self.complexes = list(complexes)
        self.n_swarms = [int(s) for s in n_swarms]
        self.S = sum(self.n_swarms)
        self.G = int(G)
        self.P = self.complexes[0].P
        self.N = self.S * self.G
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1)
        if pos.size != self.N * self.P:
            raise ValueError("pos has %d values, expected %d" % (pos.size, self.N * self.P))
        st = None if stream is None else np.ascontiguousarray(stream, dtype=np.float64).reshape(-1)
        if st is not None and st.size != self.S * n_steps * self.G:
            raise ValueError("stream has %d values, expected %d" % (st.size, self.S * n_steps * self.G))
        hs = (_VP * len(self.complexes))(*[c.h for c in self.complexes])
        ns = np.ascontiguousarray(self.n_swarms, dtype=np.int32)
        prm = params or default_params()
        h = _VP()
        _check(lib().sdgso_run_create(len(self.complexes), hs, _p(ns, _I), self.G, _p(pos), _p(st),
                                      int(n_steps), ctypes.byref(prm), form, band, ctypes.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().sdgso_run_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_stream(self, hip_stream_ptr):
        _check(lib().sdgso_run_set_stream(self.h, hip_stream_ptr))

    def steps(self, n, graph=False):
        _check(lib().sdgso_run_steps(self.h, int(n), 1 if graph else 0))

    def phases(self, mask):
        _check(lib().sdgso_run_phases(self.h, int(mask)))

    def sync(self):
        _check(lib().sdgso_run_sync(self.h))

    @property
    def k(self):
        return lib().sdgso_run_step_index(self.h)

    def read(self):
        N, P = self.N, self.P
        out = dict(pos=np.zeros((N, P)), luciferin=np.zeros(N), score=np.zeros(N),
                   vision=np.zeros(N), nnei=np.zeros(N, np.int32), moved=np.zeros(N, np.int32),
                   select=np.zeros(N, np.int32), ref=np.zeros((N, 3)),
                   pairs=np.zeros(N, np.int64))
        _check(lib().sdgso_run_read(self.h, _p(out["pos"]), _p(out["luciferin"]), _p(out["score"]),
                                    _p(out["vision"]), _p(out["nnei"], _I), _p(out["moved"], _I),
                                    _p(out["select"], _I), _p(out["ref"]), _p(out["pairs"], _I64)))
        return out

    def write(self, pos=None, luciferin=None, score=None, vision=None, moved=None, ref=None):
        def arr(x, dt):
            return None if x is None else np.ascontiguousarray(x, dtype=dt).reshape(-1)
        pos, luciferin, score, vision, ref = (arr(pos, np.float64), arr(luciferin, np.float64),
                                              arr(score, np.float64), arr(vision, np.float64),
                                              arr(ref, np.float64))
        moved = arr(moved, np.int32)
        _check(lib().sdgso_run_write(self.h, _p(pos), _p(luciferin), _p(score), _p(vision),
                                     _p(moved, _I), _p(ref)))

    def profile(self, enable=True):
        _check(lib().sdgso_run_profile(self.h, 1 if enable else 0))

    def profile_read(self):
        t = np.zeros(3)
        n = np.zeros(3, np.int64)
        _check(lib().sdgso_run_profile_read(self.h, _p(t), _p(n, _I64)))
        return dict(score_ms=t[0], move_ms=t[1], lgemm_ms=t[2], score_launches=int(n[0]),
                    move_launches=int(n[1]), lgemm_launches=int(n[2]))

    def set_work_order(self, xcd_chunk):
        """Scoring work order: 0 list order, K > 0 chunks of K items per XCD."""
        _check(lib().sdgso_run_set_work_order(self.h, int(xcd_chunk)))

    def work_order(self):
        """(xcd_chunk, (ms per scoring pass in list order, chunked) of the
        last tune)."""
        k = np.zeros(1, np.int32)
        t = np.zeros(2)
        _check(lib().sdgso_run_work_order(self.h, _p(k, _I), _p(t)))
        return int(k[0]), (float(t[0]), float(t[1]))

    def set_ligand_gemm(self, enable):
        """Switch the ligand displacement GEMM (FP32 form) on or off; a run
        starts with it off and the tuner keeps it when it is faster."""
        _check(lib().sdgso_run_set_ligand_gemm(self.h, 1 if enable else 0))

    def ligand_gemm(self):
        """(on, bytes of its output) of the ligand displacement GEMM."""
        e = np.zeros(1, np.int32)
        b = np.zeros(1, np.int64)
        _check(lib().sdgso_run_ligand_gemm(self.h, _p(e, _I), _p(b, _I64), None))
        return bool(e[0]), int(b[0])

    def ligand_gemm_tune_ms(self):
        """The last tune's ms per scoring pass with and without the GEMM."""
        t = np.zeros(2)
        _check(lib().sdgso_run_ligand_gemm(self.h, None, None, _p(t)))
        return float(t[0]), float(t[1])

    def ligand_displacements(self, first, count):
        """The GEMM's FP32 ligand displacements of glowworms [first, first +
        count) (one complex) for the current poses: [count][Nl][3]."""
        k, g0 = 0, 0
        while first >= g0 + self.n_swarms[k] * self.G:
            g0 += self.n_swarms[k] * self.G
            k += 1
        out = np.zeros((count, self.complexes[k].nl, 3), np.float32)
        _check(lib().sdgso_run_ligand_displacements(self.h, int(first), int(count), _p(out, _F)))
        return out

    def tune(self, repeats=1):
        """Time every FP32 scoring-kernel variant on the current poses of each
        complex and keep the fastest per complex; returns (chosen variant per
        complex, {variant name: [ms per launch of each complex's share]})."""
        nv = lib().sdgso_tune_variants()
        nc = len(self.complexes)
        t = np.zeros(nc * nv)
        ch = np.zeros(nc, np.int32)
        _check(lib().sdgso_run_tune(self.h, int(repeats), _p(t), _p(ch, _I)))
        names = tune_variants()
        t = t.reshape(nc, nv)
        return ch.tolist(), {names[v]: t[:, v].tolist() for v in range(nv)}

    def variants(self):
        """The FP32 scoring-kernel variant of each complex."""
        ch = np.zeros(len(self.complexes), np.int32)
        _check(lib().sdgso_run_variants(self.h, _p(ch, _I)))
        return ch.tolist()

    def set_variant(self, variant):
        _check(lib().sdgso_run_set_variant(self.h, int(variant)))

    def guard_stats(self):
        s = np.zeros(1, np.int64)
        _check(lib().sdgso_run_guard_stats(self.h, _p(s, _I64)))
        return dict(recomputed=int(s[0]))

    def final_coords(self, first, count):
        """Final FP64 coordinates of glowworms [first, first + count) (one
        complex): (receptor [count][Nr][3], ligand [count][Nl][3])."""
        k = 0
        g0 = 0
        while first >= g0 + self.n_swarms[k] * self.G:
            g0 += self.n_swarms[k] * self.G
            k += 1
        cx = self.complexes[k]
        rec = np.zeros((count, cx.nr, 3))
        lig = np.zeros((count, cx.nl, 3))
        _check(lib().sdgso_run_final_coords(self.h, int(first), int(count), _p(rec), _p(lig)))
        return rec, lig

    def cluster(self, complex_index, rec_backbone, lig_backbone, cutoff=4.0):
        """Clusters of every swarm of one complex: (cluster [S][G],
        representative [S][G]), both within the swarm."""
        rb = np.ascontiguousarray(rec_backbone, dtype=np.int32)
        lb = np.ascontiguousarray(lig_backbone, dtype=np.int32)
        S = self.n_swarms[complex_index]
        cl = np.zeros((S, self.G), np.int32)
        rep = np.zeros((S, self.G), np.int32)
        _check(lib().sdgso_run_cluster(self.h, int(complex_index), _p(rb, _I), len(rb), _p(lb, _I),
                                       len(lb), float(cutoff), _p(cl, _I), _p(rep, _I)))
        return cl, rep

    def probe_anm_gemm(self):
        """Time the receptor displacement as an FP32 matrix-core GEMM over the
        tiles each swarm can reach (a measurement; state unchanged)."""
        r = np.zeros(5)
        _check(lib().sdgso_probe_anm_gemm(self.h, _p(r)))
        return dict(ms=r[0], bytes_written=r[1], flops=r[2], displacements=r[3], swarm_tiles=r[4])

    def filter_stats(self):
        """Counts of one FP32 scoring pass over the current poses: waves
        (glowworm x ligand block, or x block pair for a two-block variant),
        waves with candidate tiles, candidate tiles,
        receptor atoms passing the atom test, queued pairs, in-cutoff pairs."""
        s = np.zeros(6, np.int64)
        _check(lib().sdgso_run_filter_stats(self.h, _p(s, _I64)))
        keys = ("waves", "waves_with_candidates", "candidate_tiles", "near_atoms", "queued_pairs",
                "cutoff_pairs")
        return dict(zip(keys, (int(v) for v in s)))


PDB_NO_OXT, PDB_NO_H, PDB_NO_WATER = 1, 2, 4


class Structure:
    """A molecular structure (atoms with PDB metadata), host only."""

    def __init__(self, handle):
        self.h = handle
        n = ctypes.c_int(0)
        r = ctypes.c_int(0)
        c = ctypes.c_int(0)
        _check(lib().sdgso_structure_size(self.h, ctypes.byref(n), ctypes.byref(r), ctypes.byref(c)))
        self.n_atoms, self.n_residues, self.n_chains = n.value, r.value, c.value

    @classmethod
    def read_pdb(cls, path, noxt=True, noh=True, now=True):
        h = _VP()
        flags = (PDB_NO_OXT if noxt else 0) | (PDB_NO_H if noh else 0) | (PDB_NO_WATER if now else 0)
        _check(lib().sdgso_structure_read_pdb(str(path).encode(), flags, ctypes.byref(h)))
        return cls(h)

    @classmethod
    def synthetic(cls, xyz):
        xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        h = _VP()
        _check(lib().sdgso_structure_synthetic(len(xyz), _p(xyz), ctypes.byref(h)))
        return cls(h)

    def coords(self):
        x = np.zeros((self.n_atoms, 3))
        _check(lib().sdgso_structure_coords(self.h, _p(x)))
        return x

    def atom(self, i):
        name = ctypes.create_string_buffer(8)
        res = ctypes.create_string_buffer(8)
        chain = ctypes.create_string_buffer(4)
        num = ctypes.c_int(0)
        _check(lib().sdgso_structure_atom(self.h, int(i), name, res, ctypes.byref(num), chain))
        return name.value.decode(), res.value.decode(), num.value, chain.value.decode()

    def dfire_types(self):
        t = np.zeros(self.n_atoms, np.int32)
        _check(lib().sdgso_structure_dfire_types(self.h, _p(t, _I)))
        return t

    def nm_mask(self):
        m = np.zeros(self.n_atoms, np.int32)
        _check(lib().sdgso_structure_nm_mask(self.h, _p(m, _I)))
        return m

    def backbone(self):
        n = ctypes.c_int(0)
        _check(lib().sdgso_structure_backbone(self.h, None, ctypes.byref(n)))
        idx = np.zeros(n.value, np.int32)
        _check(lib().sdgso_structure_backbone(self.h, _p(idx, _I), ctypes.byref(n)))
        return idx

    def centre(self):
        t = np.zeros(3)
        _check(lib().sdgso_structure_centre(self.h, _p(t)))
        return t

    def anm(self, xyz, n_modes, rmsd=0.5, n_draws=1):
        """Normal modes (needs the GPU): returns (re-centred coordinates [N][3],
        modes [n_modes][N][3], eigenvalues [n_modes])."""
        x = np.array(xyz, dtype=np.float64, order="C", copy=True)
        modes = np.zeros((n_modes, self.n_atoms, 3))
        ev = np.zeros(n_modes)
        _check(lib().sdgso_structure_anm(self.h, _p(x), int(n_modes), float(rmsd), int(n_draws),
                                         _p(modes), _p(ev)))
        return x, modes, ev

    def write_pdb(self, path, xyz=None, append=False):
        x = None if xyz is None else np.ascontiguousarray(xyz, dtype=np.float64)
        _check(lib().sdgso_structure_write_pdb(self.h, _p(x), str(path).encode(), 1 if append else 0))

    def close(self):
        if getattr(self, "h", None):
            lib().sdgso_structure_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def kmeans(points, k, seed=324324, iterations=100):
    """k-means++ seeding and Lloyd iterations: (centres [k][3], group [n])."""
    p = np.ascontiguousarray(points, dtype=np.float64)
    c = np.zeros((k, 3))
    g = np.zeros(len(p), np.int32)
    _check(lib().sdgso_kmeans(_p(p), len(p), int(k), int(seed), int(iterations), _p(c), _p(g, _I)))
    return c, g


def starting_poses(receptor, rec_xyz, lig_xyz, rec_translation, swarms, glowworms,
                   swarm_radius=10.0, start_points_seed=324324, anm_seed=324324, n_modes_rec=0,
                   n_modes_lig=0):
    """Swarm centres [swarms][3] and poses [swarms][glowworms][P] around the
    receptor surface (coordinates before centring)."""
    rx = np.ascontiguousarray(rec_xyz, dtype=np.float64)
    lx = np.ascontiguousarray(lig_xyz, dtype=np.float64)
    t = np.ascontiguousarray(rec_translation, dtype=np.float64)
    P = 7 + n_modes_rec + n_modes_lig
    c = np.zeros((swarms, 3))
    pos = np.zeros((swarms, glowworms, P))
    _check(lib().sdgso_starting_poses(receptor.h, _p(rx), len(lx), _p(lx), _p(t), int(swarms),
                                      int(glowworms), float(swarm_radius), int(start_points_seed),
                                      int(anm_seed), int(n_modes_rec), int(n_modes_lig), _p(c), _p(pos)))
    return c, pos


def reference_point(xyz, tolerance=0.01):
    """Centre of the minimum-volume enclosing ellipsoid (the ligand
    reference point)."""
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    c = np.zeros(3)
    _check(lib().sdgso_reference_point(len(xyz), _p(xyz), float(tolerance), _p(c)))
    return c


def probe_gather(table_floats=TABLE_LEN):
    """Random-gather rate (gathers/s) of the device over a float table."""
    r = ctypes.c_double(0)
    _check(lib().sdgso_probe_gather(int(table_floats), ctypes.byref(r)))
    return r.value


def tune_variants():
    """Names of the compiled FP32 scoring-kernel variants, by index."""
    return [lib().sdgso_tune_variant_name(v).decode() for v in range(lib().sdgso_tune_variants())]


def stream(seed, S, n_steps, G):
    u = np.zeros(S * n_steps * G)
    _check(lib().sdgso_stream(seed, S, n_steps, G, _p(u)))
    return u.reshape(S, n_steps, G)


def synth_molecule(seed, n_atoms, n_modes):
    xyz = np.zeros((n_atoms, 3))
    t = np.zeros(n_atoms, np.int32)
    m = np.zeros((n_modes, n_atoms, 3)) if n_modes else None
    _check(lib().sdgso_synth_molecule(seed, n_atoms, n_modes, _p(xyz), _p(t, _I), _p(m)))
    return xyz, t, m


def synth_table(seed):
    t = np.zeros(TABLE_LEN, np.float32)
    _check(lib().sdgso_synth_table(seed, _p(t, _F)))
    return t


def synth_poses(seed, S, G, nmr, nml, centre_radius, swarm_radius=10.0):
    pos = np.zeros((S, G, 7 + nmr + nml))
    _check(lib().sdgso_synth_poses(seed, S, G, nmr, nml, float(centre_radius), float(swarm_radius),
                                   _p(pos)))
    return pos
