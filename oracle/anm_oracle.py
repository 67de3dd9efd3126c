"""Normal modes by the anisotropic network model: TEST INFRASTRUCTURE ONLY.

A numpy restatement of source/dock/src/cal_anm.cc (cal_anm 61-219,
build_hessian1D 307-359, calModes 403-514, norm 516-540) used as the checker
of the library's sdgso_structure_anm.  Only tests/ load it.

The reference file cannot be compiled here (it includes freesasa.h and
mpi.h, neither present), so this restatement is pinned by: the Hessian's
closed properties (symmetric, six zero modes of a rigid network, row sums
zero), numpy's LAPACK eigensolver, and libstdc++'s generator and normal
distribution restated below (pinned against the library, which calls
libstdc++ itself).  Eigenvector signs are solver-specific; both this
restatement and the library fix each mode's largest component positive, so
the reference's cuSOLVER signs are not reproduced ("sign parity unpinned").
"""
import math

import numpy as np


class MinStdRand0:
    """std::minstd_rand0: x <- 16807 x mod (2^31 - 1)."""

    M = 2147483647

    def __init__(self, seed):
        self.x = seed % self.M or 1

    def __call__(self):
        self.x = (16807 * self.x) % self.M
        return self.x


def generate_canonical(g):
    """libstdc++ generate_canonical<double, 53> over a minstd engine: two draws
    (log2 of the range 2^31 - 2 is below 31), sum (x - 1) * R^i, divided by R^2."""
    R = 2147483646.0
    s = 0.0
    t = 1.0
    for _ in range(2):
        s += float(g() - 1) * t
        t *= R
    r = s / t
    if r >= 1.0:
        r = math.nextafter(1.0, 0.0)
    return r


class NormalDistribution:
    """libstdc++ normal_distribution: Marsaglia's polar method, the second
    value of each pair saved for the next call."""

    def __init__(self, mu=0.0, sigma=1.0):
        self.mu, self.sigma = mu, sigma
        self.saved = None

    def __call__(self, g):
        if self.saved is not None:
            v, self.saved = self.saved, None
            return v * self.sigma + self.mu
        while True:
            x = 2.0 * generate_canonical(g) - 1.0
            y = 2.0 * generate_canonical(g) - 1.0
            r2 = x * x + y * y
            if r2 <= 1.0 and r2 != 0.0:
                break
        mult = math.sqrt(-2.0 * math.log(r2) / r2)
        self.saved = x * mult
        return y * mult * self.sigma + self.mu


def hessian(c, cutoff=15.0, gamma=1.0):
    """build_hessian1D (cal_anm.cc:307-359) for node coordinates c [n][3]."""
    n = len(c)
    dof = 3 * n
    H = np.zeros((dof, dof))
    cut2 = cutoff * cutoff
    for i in range(n):
        for j in range(i + 1, n):
            d = [c[j][0] - c[i][0], c[j][1] - c[i][1], c[j][2] - c[i][2]]
            d2 = 0.0
            for k in range(3):
                d2 += d[k] * d[k]
            if d2 > cut2:
                continue
            for k in range(3):
                for l in range(3):
                    se = d[k] * d[l] * (-gamma / d2)
                    H[3 * i + k, 3 * j + l] = se
                    H[3 * j + l, 3 * i + k] = se
                    H[3 * i + k, 3 * i + l] -= se
                    H[3 * j + l, 3 * j + k] -= se
    return H


def anm(xyz, names, residues, n_modes, rmsd, n_draws):
    """(re-centred xyz, modes [n_modes][N][3], eigenvalues [n_modes]) as
    cal_anm computes them; residues is a list of (first atom, atom count)."""
    x = np.array(xyz, dtype=np.float64)
    N = len(x)
    mean = [0.0, 0.0, 0.0]
    for i in range(N):
        for k in range(3):
            mean[k] += x[i, k]
    mean = [m / N for m in mean]
    x = x - np.array(mean)
    node = [i for i in range(N) if names[i] == "CA"] or [i for i in range(N) if names[i] == "C4"]
    n = len(node)
    assert len(residues) == n
    node_of = np.zeros(N, np.int64)
    for r, (f, cnt) in enumerate(residues):
        assert f <= node[r] < f + cnt
        node_of[f:f + cnt] = r
    H = hessian([x[i] for i in node])
    w, v = np.linalg.eigh(H)
    w, v = w[:n_modes + 6], v[:, :n_modes + 6]
    nz = int(np.sum(w < 1e-6))
    ev = w[nz:nz + n_modes]
    modes = np.zeros((n_modes, N, 3))
    for m in range(n_modes):
        z = v[:, nz + m]
        if z[int(np.argmax(np.abs(z)))] < 0:
            z = -z
        out = z.reshape(n, 3)[node_of]
        modes[m] = out / math.sqrt(float(np.sum(out.ravel() ** 2)))
    g = MinStdRand0(324324)
    nd = NormalDistribution(0.0, 1.0)
    coef = 0.0
    for _ in range(n_draws):
        s = 0.0
        for j in range(n_modes):
            val = nd(g)
            s += val * val * (1.0 / ev[j])
        coef += math.sqrt(s)
    coef = coef / n_draws
    item = math.sqrt(float(N)) * rmsd / coef
    for m in range(n_modes):
        modes[m] *= item * math.sqrt(1.0 / ev[m])
    return x, modes, ev
