"""Closed-form and invariant pins of the CPU oracle, independent of its own
formulas: integer geometry for bins, one-hot tables for sums, rotation
invariants, slerp/extent closed forms, hand-built roulette cases and the
C++11 MT19937 known answer."""
import ctypes
import math

import numpy as np
import pytest

import oracle_lib as O

D = ctypes.POINTER(ctypes.c_double)
I = ctypes.POINTER(ctypes.c_int)


def p(a, t=D):
    return a.ctypes.data_as(t)


# dist_to_bins of the reference (interface.h:30-35), written out as the bin
# edges it implies: bin b covers 2*sqrt(d2) in [EDGE[b], EDGE[b+1]).
EDGES = [0, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 18, 20, 22, 24, 26, 28, 30, 31]


def exact_bin_from_4d2(m4):
    """Bin of an exactly representable d2 = m4/4 (m4 integer), from integer math:
    2*sqrt(d2) = sqrt(m4) and floor(sqrt(m4)) = isqrt(m4)."""
    f = math.isqrt(m4)
    if m4 == 0 or m4 > 900:
        return -1
    for b in range(21):
        if EDGES[b] <= f < EDGES[b + 1]:
            return b
    raise AssertionError(m4)


def test_bins_on_quarter_grid():
    """Every d2 = m/4 (exact in binary) up to and past the cutoff."""
    L = O.lib()
    for m4 in range(0, 1000):
        assert L.or_dist_bin(m4 / 4.0) == exact_bin_from_4d2(m4), m4


def ref_formula_bin(d2):
    """(int)(sqrt(d2)*2.0 - 1.0) -> dist_to_bins - 1, evaluated in IEEE double
    (Python floats, correctly rounded sqrt), for the +-1 ulp cases where the
    rounding of sqrt, not the exact square root, decides the bin."""
    if not (0.0 < d2 <= 225.0):
        return -1
    d = int(math.sqrt(d2) * 2.0 - 1.0)
    f = d + 1
    for b in range(21):
        if EDGES[b] <= f < EDGES[b + 1]:
            return b
    raise AssertionError(d2)


def test_bins_one_ulp_around_boundaries():
    L = O.lib()
    flips = 0
    for k in range(1, 31):
        b = k * k / 4.0
        for d2 in (np.nextafter(b, 0), b, np.nextafter(b, 1e9)):
            assert L.or_dist_bin(float(d2)) == ref_formula_bin(float(d2)), (k, d2)
        # exact arithmetic would give the interval below; sqrt rounding may not
        flips += ref_formula_bin(float(np.nextafter(b, 0))) != exact_bin_from_4d2(max(k * k - 1, 1))
    assert flips > 0  # the reference formula really does differ from exact math here
    assert L.or_dist_bin(225.0) == 20
    assert L.or_dist_bin(np.nextafter(225.0, 1e9)) == -1
    assert L.or_dist_bin(0.0) == -1
    assert L.or_dist_bin(1e-300) == 0


def one_hot_table(bin_, a=None, b=None):
    t = np.zeros((168, 168, 20))
    if bin_ < 20:
        if a is None:
            t[:, :, bin_] = 1.0
        else:
            t[a, b, bin_] = 1.0
    return O.pad_table(t.reshape(-1))


def test_energy_counts_pairs_per_bin_on_integer_grid():
    """Atoms on a half-angstrom grid: every d2 is an exact multiple of 1/4, so
    the pair count of each bin follows from integer arithmetic."""
    rng = np.random.default_rng(3)
    r = rng.integers(-20, 20, size=(60, 3)) * 0.5
    l = rng.integers(-20, 20, size=(47, 3)) * 0.5
    # three pairs exactly at the 15 A cutoff (bin 20)
    l = np.concatenate([l, r[:3] + np.array([[15.0, 0, 0], [9.0, 12.0, 0], [5.0, 10.0, 10.0]])])
    rt = np.zeros(60, np.int32)
    lt = np.zeros(50, np.int32)
    m4 = ((4 * (r[:, None, :] - l[None, :, :]) ** 2).sum(-1)).round().astype(int)
    expect = np.zeros(21, int)
    for v in m4.ravel():
        b = exact_bin_from_4d2(int(v))
        if b >= 0:
            expect[b] += 1
    L = O.lib()
    for bin_ in range(20):
        T = one_hot_table(bin_)
        e = L.or_dfire_energy(p(r), 60, p(rt, I), p(l), 50, p(lt, I), p(T), None)
        # bin-20 pairs read T[0][1][0], which the bin-0 one-hot table also sets
        assert e == expect[bin_] + (expect[20] if bin_ == 0 else 0), bin_
    # bin 20 (d2 == 225) reads T[a][b+1][0] through the flat index
    t = np.zeros((168, 168, 20))
    t[0, 1, 0] = 1.0
    e = L.or_dfire_energy(p(r), 60, p(rt, I), p(l), 50, p(lt, I), p(O.pad_table(t.reshape(-1))), None)
    assert e == expect[20]
    assert expect[20] > 0 and expect[:20].sum() > 100


def test_energy_type_indexing():
    """A table that is one only at (a, b, bin) picks exactly those pairs."""
    rng = np.random.default_rng(5)
    r = rng.integers(-10, 10, size=(40, 3)) * 0.5
    l = rng.integers(-10, 10, size=(30, 3)) * 0.5
    rt = rng.integers(160, 168, 40).astype(np.int32)
    lt = rng.integers(0, 8, 30).astype(np.int32)
    L = O.lib()
    m4 = ((4 * (r[:, None, :] - l[None, :, :]) ** 2).sum(-1)).round().astype(int)
    for a, b, bin_ in [(163, 2, 5), (167, 7, 0), (160, 0, 19)]:
        T = one_hot_table(bin_, a, b)
        e = L.or_dfire_energy(p(r), 40, p(rt, I), p(l), 30, p(lt, I), p(T), None)
        want = sum(1 for i in range(40) for j in range(30)
                   if rt[i] == a and lt[j] == b and exact_bin_from_4d2(int(m4[i, j])) == bin_)
        assert e == want


def test_score_affine():
    L = O.lib()
    assert L.or_score_from_energy(0.0) == 4.7
    assert L.or_score_from_energy(100.0) == -(100.0 * 0.0157 - 4.7)


def rot(q, pts):
    x = np.ascontiguousarray(pts, dtype=np.float64).copy()
    O.lib().or_rotate(p(x), p(np.ascontiguousarray(q, dtype=np.float64)), len(x))
    return x


def test_rotation_invariants():
    rng = np.random.default_rng(11)
    pts = rng.normal(size=(100, 3)) * 30
    for _ in range(20):
        q = rng.normal(size=4)
        x = rot(q, pts)
        assert np.allclose(np.linalg.norm(x, axis=1), np.linalg.norm(pts, axis=1), rtol=1e-14)
        # -q gives bit-identical results (every product changes sign twice)
        assert np.array_equal(rot(-q, pts), x)
        # scaling q leaves the rotation unchanged (inverse = conj/|q|^2)
        assert np.allclose(rot(3.5 * q, pts), x, rtol=0, atol=1e-12)
        # agrees with the rotation matrix of the unit quaternion
        w, a, b, c = q / np.linalg.norm(q)
        R = np.array([[1 - 2 * (b * b + c * c), 2 * (a * b - w * c), 2 * (a * c + w * b)],
                      [2 * (a * b + w * c), 1 - 2 * (a * a + c * c), 2 * (b * c - w * a)],
                      [2 * (a * c - w * b), 2 * (b * c + w * a), 1 - 2 * (a * a + b * b)]])
        assert np.allclose(x, pts @ R.T, rtol=0, atol=1e-11)


def test_rotation_closed_forms():
    s = math.sqrt(0.5)
    x = rot([s, 0, 0, s], [[1.0, 0, 0]])  # 90 degrees about z
    assert np.allclose(x, [[0, 1, 0]], atol=1e-15)
    assert np.array_equal(rot([1, 0, 0, 0], [[1.5, -2.0, 3.25]]), [[1.5, -2.0, 3.25]])


def slerp(a, b, t):
    a = np.array(a, dtype=np.float64)
    b = np.array(b, dtype=np.float64)
    O.lib().or_slerp(p(a), p(b), t)
    return a


def test_slerp_closed_forms():
    rng = np.random.default_rng(2)
    for _ in range(20):
        a = rng.normal(size=4)
        b = rng.normal(size=4)
        an, bn = a / np.linalg.norm(a), b / np.linalg.norm(b)
        if an @ bn < 0:
            an = -an
        if an @ bn > 0.9995:
            continue
        half = (an + bn) / np.linalg.norm(an + bn)
        assert np.allclose(slerp(a, b, 0.5), half, atol=1e-14)
        assert np.allclose(slerp(a, b, 0.0), an, atol=1e-14)
    # lerp branch: nearly parallel inputs give the normalised midpoint
    a = np.array([1.0, 0, 0, 0])
    b = np.array([1.0, 1e-3, 0, 0])
    m = (a + b / np.linalg.norm(b)) / 2
    assert np.allclose(slerp(a, b, 0.5), m / np.linalg.norm(m), atol=1e-15)
    # antipodal sign flip: q and -q' are treated as the same rotation
    assert np.allclose(slerp([1, 0, 0, 0], [-1, 0, 0, 0], 0.5), [-1, 0, 0, 0])


def test_move_extent_closed_form():
    L = O.lib()
    rng = np.random.default_rng(4)
    for n in (1, 10, 20):
        s = rng.uniform(-1, 1, n)
        o = s + rng.normal(size=n)
        x = s.copy()
        L.or_move_extent(p(x), p(o), n, 0.5)
        d = x - s
        assert abs(np.linalg.norm(d) - 0.5) < 1e-14
        assert np.allclose(d / 0.5, (o - s) / np.linalg.norm(o - s), atol=1e-14)
    s = np.zeros(10)
    x = s.copy()
    L.or_move_extent(p(x), p(s + 1e-10), 10, 0.5)  # below the 1e-8 threshold
    assert np.array_equal(x, s)


def select(u, nei, lucif, me):
    nei = np.asarray(nei, np.int32)
    lucif = np.asarray(lucif, np.float64)
    prob = np.zeros(max(len(nei), 1))
    L = O.lib()
    L.or_probabilities(p(prob), p(nei, I), p(lucif), len(nei), float(lucif[me]))
    return L.or_select(float(u), p(nei, I), p(prob), len(nei), me), prob[:len(nei)]


def test_roulette_hand_cases():
    lu = [1.0, 2.0, 4.0, 5.0]
    # probabilities proportional to luciferin differences, in list order
    _, prob = select(0.5, [1, 2, 3], lu, 0)
    assert np.allclose(prob, [1 / 8, 3 / 8, 4 / 8])
    assert select(0.1, [1, 2, 3], lu, 0)[0] == 1
    assert select(0.125, [1, 2, 3], lu, 0)[0] == 1   # cumulative tie: stops at >= u
    assert select(0.13, [1, 2, 3], lu, 0)[0] == 2
    assert select(0.99, [1, 2, 3], lu, 0)[0] == 3
    assert select(0.7, [], lu, 0)[0] == 0             # no neighbours: itself
    # build decisions for the reference's undefined cases: u == 0 -> first,
    # rounding left sum < u -> last
    assert select(0.0, [2, 3], lu, 0)[0] == 2
    assert select(1.5, [2, 3], lu, 0)[0] == 3


def test_mt19937_known_answer():
    # C++11 [rand.predef]: the 10000th output of default-constructed mt19937
    assert O.mt19937(5489, 10000)[-1] == 4123659995


def test_stream_matches_numpy_mt19937():
    """libstdc++ generate_canonical<double,53>: (x0 + x1*2^32) / 2^64 from two
    raw 32-bit draws of the init_genrand(seed) engine."""
    bg = np.random.MT19937()
    bg._legacy_seeding(0)
    raw = bg.random_raw(2 * 2 * 3 * 5).astype(np.float64)
    want = (raw[0::2] + raw[1::2] * 4294967296.0) / 18446744073709551616.0
    got = O.stream(0, 2, 3, 5).ravel()
    assert np.array_equal(got, want)
