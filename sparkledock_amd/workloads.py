"""The benchmark workloads (synthetic, seeded) and their construction.

Workloads (receptor atoms, ligand atoms, swarms) per complex:
  c0  one complex 757/481, 130 swarms x 200 glowworms, no modes (7-D poses),
      20 steps: the reference's CPU-runnable case, plain FP64
  c1  one complex 3002/1274, 402 x 200, 10+10 modes (27-D), 100 steps,
      FP32 guard band 1e-4 A^2
  c2  one complex 17307/1125, 1245 x 200, 10+10 modes, 100 steps
  c3  ten complexes of uneven sizes in one run, 200 glowworms, 10+10 modes,
      100 steps, step replayed as a hipGraph
  c4  32 complexes, receptor atoms log-uniform in [500, 30000], ligand atoms
      log-uniform in [300, 3000], swarms max(50, round(Nr/14)), 300 glowworms,
      20+20 modes, 50 steps
Inputs: random-walk chains in globules of radius (N*11.4)^(1/3) A, uniform
types over 168, one N(0,1)*0.5 float32 table, orthonormal random modes,
extents U[-1,1]; swarm centres on a Fibonacci sphere of radius R_rec + R_lig/2
with translations in a 10 A ball and uniform random quaternions; the random
stream of each complex is mt19937(0) as in a solo run.  The ligand reference
point is the ligand centroid.
"""
import math

import numpy as np

from . import sdgso

TABLE_SEED = 7
C3 = [(3002, 1274, 402), (3052, 1658, 336), (2039, 1440, 271), (2516, 1061, 298), (757, 481, 130),
      (2030, 1256, 294), (6143, 633, 620), (17307, 1125, 1245), (6032, 1766, 695), (6058, 1138, 558)]


def c4_sizes(seed=4):
    rng = np.random.default_rng(seed)
    nr = np.exp(rng.uniform(math.log(500), math.log(30000), 32)).round().astype(int)
    nl = np.exp(rng.uniform(math.log(300), math.log(3000), 32)).round().astype(int)
    return [(int(a), int(b), max(50, int(round(a / 14)))) for a, b in zip(nr, nl)]


CONFIGS = {
    "c0": dict(complexes=[(757, 481, 130)], G=200, modes=(0, 0), steps=20, form=sdgso.FP64),
    "c1": dict(complexes=[(3002, 1274, 402)], G=200, modes=(10, 10), steps=100,
               form=sdgso.FP32_GUARD),
    "c2": dict(complexes=[(17307, 1125, 1245)], G=200, modes=(10, 10), steps=100,
               form=sdgso.FP32_GUARD),
    "c3": dict(complexes=C3, G=200, modes=(10, 10), steps=100, form=sdgso.FP32_GUARD),
    "c4": dict(complexes=None, G=300, modes=(20, 20), steps=50, form=sdgso.FP32_GUARD),
}


def radius(n):
    return (n * 11.4) ** (1.0 / 3.0)


def complex_inputs(index, nr, nl, nmr, nml, table=None):
    """Host arrays of synthetic complex `index` (seeds derive from the index)."""
    r_xyz, r_type, r_modes = sdgso.synth_molecule(1000 * index + 1, nr, nmr)
    l_xyz, l_type, l_modes = sdgso.synth_molecule(1000 * index + 2, nl, nml)
    if table is None:
        table = sdgso.synth_table(TABLE_SEED)
    return dict(r_xyz=r_xyz, r_type=r_type, r_modes=r_modes, l_xyz=l_xyz, l_type=l_type,
                l_modes=l_modes, table=table, lig_ref=l_xyz.mean(axis=0),
                r_mask=np.ones(nr, np.int32), l_mask=np.ones(nl, np.int32))


def pose_inputs(index, S, G, nr, nl, nmr, nml):
    return sdgso.synth_poses(1000 * index + 3, S, G, nmr, nml, radius(nr) + 0.5 * radius(nl), 10.0)


def spec(name, swarm_scale=1.0, steps=None):
    cfg = dict(CONFIGS[name])
    comps = cfg["complexes"] if cfg["complexes"] is not None else c4_sizes()
    if swarm_scale != 1.0:
        comps = [(a, b, max(1, int(round(s * swarm_scale)))) for a, b, s in comps]
    cfg["complexes"] = comps
    if steps is not None:
        cfg["steps"] = steps
    return cfg


def build(name, swarm_scale=1.0, steps=None, form=None, device=True):
    """Device complexes, poses and streams of a workload.

    Returns dict(complexes=[sdgso.Complex], n_swarms, G, pos, stream, steps,
    form, inputs=[host dicts]); with device=False the complexes are not
    created on the device (complexes is empty).
    """
    cfg = spec(name, swarm_scale, steps)
    nmr, nml = cfg["modes"]
    table = sdgso.synth_table(TABLE_SEED)
    cxs, ns, pos, st, inputs = [], [], [], [], []
    for k, (nr, nl, S) in enumerate(cfg["complexes"]):
        inp = complex_inputs(k, nr, nl, nmr, nml, table)
        inputs.append(inp)
        if device:
            cxs.append(sdgso.Complex(inp["r_xyz"], inp["r_type"], inp["l_xyz"], inp["l_type"],
                                     table, inp["lig_ref"], inp["r_modes"], inp["l_modes"],
                                     inp["r_mask"], inp["l_mask"]))
        ns.append(S)
        pos.append(pose_inputs(k, S, cfg["G"], nr, nl, nmr, nml).reshape(-1))
        st.append(sdgso.stream(0, S, cfg["steps"], cfg["G"]).reshape(-1))
    return dict(complexes=cxs, n_swarms=ns, G=cfg["G"], pos=np.concatenate(pos),
                stream=np.concatenate(st), steps=cfg["steps"],
                form=cfg["form"] if form is None else form, inputs=inputs, spec=cfg)
