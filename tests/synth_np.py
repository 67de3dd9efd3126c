"""Small seeded synthetic complexes in numpy, for the oracle tests.

The shapes follow the workload description: random-walk C-alpha chains with
3.8 A steps packed in a globule of radius (N*11.4)^(1/3), uniform DFIRE types
over 168, an N(0,1)*0.5 float32 table, orthonormal random modes, extents in
[-1, 1], translations in a 10 A ball around swarm centres just outside the
receptor and Shoemake unit quaternions (cal_anm.cc:845-875).
"""
import numpy as np

NTYPES = 168
NBINS = 20


def globule(rng, n, step=3.8):
    radius = (n * 11.4) ** (1.0 / 3.0)
    xyz = np.zeros((n, 3))
    p = np.zeros(3)
    for i in range(n):
        for _ in range(64):
            d = rng.normal(size=3)
            d *= step / np.linalg.norm(d)
            q = p + d
            if np.linalg.norm(q) <= radius:
                break
        else:
            q = p - d if np.linalg.norm(p - d) <= radius else p * 0.5
        p = q
        xyz[i] = p
    return xyz - xyz.mean(axis=0)


def modes(rng, n_modes, n_atoms):
    if n_modes == 0:
        return np.zeros((0, n_atoms, 3))
    a = rng.normal(size=(3 * n_atoms, n_modes))
    q, _ = np.linalg.qr(a)
    return np.ascontiguousarray(q.T.reshape(n_modes, n_atoms, 3))


def table(rng):
    return (rng.normal(size=NTYPES * NTYPES * NBINS) * 0.5).astype(np.float32)


def quaternions(rng, n):
    u1, u2, u3 = rng.random(n), rng.random(n), rng.random(n)
    return np.stack([np.sqrt(1 - u1) * np.sin(2 * np.pi * u2), np.sqrt(1 - u1) * np.cos(2 * np.pi * u2),
                     np.sqrt(u1) * np.sin(2 * np.pi * u3), np.sqrt(u1) * np.cos(2 * np.pi * u3)], axis=1)


def poses(rng, S, G, nmr, nml, rec_radius, lig_radius, swarm_radius=10.0):
    P = 7 + nmr + nml
    out = np.zeros((S, G, P))
    dirs = rng.normal(size=(S, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    centres = dirs * (rec_radius + 0.5 * lig_radius)
    for s in range(S):
        t = np.zeros((G, 3))
        k = 0
        while k < G:
            x = (2 * rng.random(3) - 1) * swarm_radius
            if x @ x <= swarm_radius ** 2:
                t[k] = x
                k += 1
        out[s, :, 0:3] = centres[s] + t
        out[s, :, 3:7] = quaternions(rng, G)
        out[s, :, 7:] = rng.uniform(-1, 1, size=(G, nmr + nml))
    return out


def make_case(seed, nr, nl, nmr, nml, S, G, mask_holes=False, swarm_radius=10.0):
    rng = np.random.default_rng(seed)
    r = globule(rng, nr)
    l = globule(rng, nl)
    case = dict(
        r_xyz=r, l_xyz=l,
        r_type=rng.integers(0, NTYPES, nr).astype(np.int32),
        l_type=rng.integers(0, NTYPES, nl).astype(np.int32),
        table=table(rng),
        r_modes=modes(rng, nmr, nr), l_modes=modes(rng, nml, nl),
        r_mask=np.ones(nr, np.int32), l_mask=np.ones(nl, np.int32),
        lig_ref=l.mean(axis=0) + rng.normal(size=3) * 0.1,
    )
    if mask_holes:
        case["r_mask"][rng.random(nr) < 0.1] = 0
        case["l_mask"][rng.random(nl) < 0.1] = 0
    rr = (nr * 11.4) ** (1 / 3)
    lr = (nl * 11.4) ** (1 / 3)
    case["pos"] = poses(rng, S, G, nmr, nml, rr, lr, swarm_radius)
    return case
