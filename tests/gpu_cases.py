"""Shared inputs of the GPU parity tests: one seeded case, built both as a
device complex (the product path) and as an oracle complex (the checker)."""
import numpy as np

import oracle_lib as O
import synth_np as SN
from sparkledock_amd import sdgso


class Pair:
    def __init__(self, case):
        self.case = case
        c = case
        self.gpu = sdgso.Complex(c["r_xyz"], c["r_type"], c["l_xyz"], c["l_type"], c["table"],
                                 c["lig_ref"], c["r_modes"] if c["r_modes"].shape[0] else None,
                                 c["l_modes"] if c["l_modes"].shape[0] else None, c["r_mask"],
                                 c["l_mask"])
        self.ora = O.Complex(c["r_xyz"], c["r_type"], c["l_xyz"], c["l_type"], c["table"],
                             c["lig_ref"], c["r_modes"], c["l_modes"], c["r_mask"], c["l_mask"])
        self.P = self.ora.P

    def oracle_scores(self, poses, table=None):
        cx = self.ora if table is None else self.with_table(table)
        out = [cx.score_pose(q) for q in poses.reshape(-1, self.P)]
        return (np.array([o[0] for o in out]), np.array([o[1] for o in out]),
                np.array([o[2] for o in out], dtype=np.int64))

    def with_table(self, table):
        c = self.case
        return O.Complex(c["r_xyz"], c["r_type"], c["l_xyz"], c["l_type"], table, c["lig_ref"],
                         c["r_modes"], c["l_modes"], c["r_mask"], c["l_mask"])

    def gpu_with_table(self, table):
        c = self.case
        return sdgso.Complex(c["r_xyz"], c["r_type"], c["l_xyz"], c["l_type"], table,
                             c["lig_ref"], c["r_modes"] if c["r_modes"].shape[0] else None,
                             c["l_modes"] if c["l_modes"].shape[0] else None, c["r_mask"],
                             c["l_mask"])


def case(seed, nr, nl, nmr, nml, S, G, swarm_radius=10.0, mask_holes=True):
    return SN.make_case(seed, nr, nl, nmr, nml, S, G, mask_holes=mask_holes,
                        swarm_radius=swarm_radius)


def grid_case(seed, nr=400, nl=300, S=2, G=32):
    """Atoms on a half-angstrom lattice, identity rotations and lattice
    translations, no modes: every d2 is a multiple of 1/4 and thousands of
    pairs sit exactly on bin boundaries, on the cutoff and at d2 = 0."""
    rng = np.random.default_rng(seed)
    r = rng.integers(-16, 16, size=(nr, 3)) * 0.5
    l = rng.integers(-12, 12, size=(nl, 3)) * 0.5
    c = dict(r_xyz=r.astype(np.float64), l_xyz=l.astype(np.float64),
             r_type=rng.integers(0, 168, nr).astype(np.int32),
             l_type=rng.integers(0, 168, nl).astype(np.int32),
             table=(rng.normal(size=168 * 168 * 20) * 0.5).astype(np.float32),
             r_modes=np.zeros((0, nr, 3)), l_modes=np.zeros((0, nl, 3)),
             r_mask=np.ones(nr, np.int32), l_mask=np.ones(nl, np.int32),
             lig_ref=np.array([0.5, -1.0, 0.25]))
    pos = np.zeros((S, G, 7))
    pos[:, :, 0:3] = rng.integers(-8, 8, size=(S, G, 3)) * 0.5
    pos[:, :, 3] = 1.0
    c["pos"] = pos
    return c


def sum_abs_terms(pair, poses):
    """Sum over in-cutoff pairs of |T| for each pose (the conditioning-aware
    denominator of the FP32 energy error)."""
    tab = np.abs(pair.case["table"]).astype(np.float32)
    cx = pair.with_table(tab)
    return np.array([(4.7 - cx.score_pose(q)[0]) / 0.0157 for q in poses.reshape(-1, pair.P)])
