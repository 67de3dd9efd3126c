"""Ranking of docking results across swarms.

Reads the swarm_<i>/gso_<steps>.out files written by the lightdock driver
(line format of source/dock/src/cuda/prepare_gso_task.cu:2868-2888) the way
analysis/lgd_rank.py:365-409 parses them, ranks glowworms by scoring or by
luciferin in decreasing order (lgd_rank.py:311-336), and compares rankings
(top-N overlap).
"""
import glob
import os
import re

import numpy as np

HEADER = "#Coordinates  RecID  LigID  Luciferin  Neighbor's number  Vision Range  Scoring"


def parse_line(line):
    """One glowworm line: (pose list, rec_id, lig_id, luciferin, neighbours,
    vision range, scoring)."""
    first = line.index("(")
    last = line.index(")")
    pose = [float(v) for v in line[first + 1:last].split(",")]
    rest = line[last + 1:].split()
    if len(rest) >= 6:
        return (pose, int(rest[0]), int(rest[1]), float(rest[2]), int(rest[3]), float(rest[4]),
                float(rest[5]))
    return (pose, 0, 0, float(rest[0]), int(rest[1]), float(rest[2]), float(rest[3]))


def read_gso_out(path):
    """Glowworm records of one swarm file, in file order."""
    out = []
    with open(path) as f:
        for line in f:
            if line.startswith("#") or "(" not in line:
                continue
            out.append(parse_line(line))
    return out


def read_run(directory, steps):
    """All swarms of a run directory: dict swarm -> list of records."""
    res = {}
    for d in glob.glob(os.path.join(directory, "swarm_*")):
        m = re.search(r"swarm_(\d+)$", d)
        p = os.path.join(d, "gso_%d.out" % steps)
        if m and os.path.exists(p):
            res[int(m.group(1))] = read_gso_out(p)
    return res


def rank(values, k=None):
    """Indices of `values` in decreasing order (ties by index), first k."""
    values = np.asarray(values, dtype=np.float64).ravel()
    order = np.lexsort((np.arange(values.size), -values))
    return order if k is None else order[:k]


def top_overlap(values_a, values_b, k=100):
    """Fraction of the top-k glowworms (by decreasing value) shared by two runs
    over the same (swarm, glowworm) index space."""
    ta = set(rank(values_a, k).tolist())
    tb = set(rank(values_b, k).tolist())
    return len(ta & tb) / float(k)
