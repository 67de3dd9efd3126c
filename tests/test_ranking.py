"""Ranking helpers (host only): the gso_<steps>.out line parser follows
analysis/lgd_rank.py:365-409, the order follows lgd_rank.py:311-336."""
import numpy as np

from sparkledock_amd import ranking


def test_parse_reference_line_format():
    # a line as prepare_gso_task.cu:2868-2888 writes it (std::to_string = %f)
    line = "(1.500000, -2.000000, 3.250000, 1.000000, 0.000000, 0.000000, 0.000000)\t0\t0\t7.125000\t3\t0.440000\t-12.500000\t"
    pose, rec, lig, lucif, nnei, vis, score = ranking.parse_line(line)
    assert pose == [1.5, -2.0, 3.25, 1.0, 0.0, 0.0, 0.0]
    assert (rec, lig, lucif, nnei, vis, score) == (0, 0, 7.125, 3, 0.44, -12.5)
    # the default-solution form without receptor/ligand ids
    pose, rec, lig, lucif, nnei, vis, score = ranking.parse_line("(0.0, 1.0)\t2.5\t1\t0.2\t-1.0")
    assert (rec, lig, lucif, nnei, vis, score) == (0, 0, 2.5, 1, 0.2, -1.0)


def test_rank_and_overlap():
    v = np.array([0.5, 3.0, 3.0, -1.0, 2.0])
    assert ranking.rank(v).tolist() == [1, 2, 4, 0, 3]  # decreasing, ties by index
    a = np.arange(1000.0)
    b = a.copy()
    b[-10:] = -1.0  # the 10 best of a are the 10 worst of b
    assert ranking.top_overlap(a, a, 100) == 1.0
    assert ranking.top_overlap(a, b, 100) == 0.9


def test_percent_f_equals_c_printf():
    """The driver test formats the oracle's values with Python's "%f" as the
    reference's std::to_string(double) does (C++11 [string.conversions]:
    sprintf "%f"): the two agree on random values of every magnitude, on
    halfway cases of the sixth decimal, on signed zeros and tiny negatives."""
    import ctypes
    libc = ctypes.CDLL(None)
    buf = ctypes.create_string_buffer(512)
    rng = np.random.default_rng(3)
    vals = list(rng.normal(size=2000) * 10.0 ** rng.uniform(-9, 9, 2000))
    vals += [0.0, -0.0, -1e-7, 1e-7, 0.5e-6, -0.5e-6, 1.5e-6, 2.5e-6, 0.0000005, 123456.0000005,
             1e300, -1e300, 5e-324]
    for v in vals:
        libc.snprintf(buf, 512, b"%f", ctypes.c_double(v))
        assert buf.value.decode() == "%f" % v, v
