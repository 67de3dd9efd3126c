"""libsdgso.so on a host without a GPU: it loads, exports every symbol that
include/sdgso.h declares, validates arguments, reports a missing device as an
error, and its host utilities (random stream, synthetic inputs) match their
specifications."""
import os
import re

import numpy as np
import pytest

import oracle_lib as O
from sparkledock_amd import sdgso

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    hdr = open(os.path.join(ROOT, "include", "sdgso.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    return sorted(set(re.findall(r"\b(sdgso_[a-z_0-9]+)\s*\(", hdr)))


def test_exports_every_declared_symbol():
    L = sdgso.lib()
    names = declared_symbols()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n
    # and the binding covers each of them
    bound = {s[0] for s in sdgso.SIGNATURES}
    assert set(names) == bound


def test_version_and_params():
    assert sdgso.lib().sdgso_version() >= 100
    p = sdgso.default_params()
    o = O.default_params()
    assert (p.rho, p.gamma, p.beta, p.vision0, p.max_vision, p.luciferin0, p.max_neighbors) == \
        (o.rho, o.gamma, o.beta, o.vision0, o.max_vision, o.luciferin0, o.max_neighbors)
    assert (p.step_translation, p.step_rotation, p.step_nmodes) == (0.5, 0.5, 0.5)


def test_argument_errors_before_device():
    xyz = np.zeros((4, 3))
    t = np.zeros(4, np.int32)
    tab = np.zeros(sdgso.TABLE_LEN, np.float32)
    with pytest.raises(sdgso.SdgsoError) as e:
        sdgso.Complex(xyz, np.full(4, 200, np.int32), xyz, t, tab, [0, 0, 0])
    assert e.value.code == 1 and "type" in str(e.value)
    with pytest.raises(ValueError):
        sdgso.Complex(xyz, t, xyz, t, tab[:10], [0, 0, 0])


@pytest.mark.skipif(os.path.exists("/dev/kfd") and os.access("/dev/kfd", os.R_OK),
                    reason="a GPU is present")
def test_missing_device_is_an_error():
    xyz = np.zeros((4, 3))
    t = np.zeros(4, np.int32)
    with pytest.raises(sdgso.SdgsoError) as e:
        sdgso.Complex(xyz, t, xyz, t, np.zeros(sdgso.TABLE_LEN, np.float32), [0, 0, 0])
    assert e.value.code == 2


def test_stream_matches_oracle_and_libstdcxx():
    """The library draws with std::mt19937 + uniform_real_distribution; the
    oracle restates both (MT19937 and generate_canonical)."""
    for seed, S, K, G in [(0, 3, 5, 7), (123, 2, 4, 200)]:
        a = sdgso.stream(seed, S, K, G)
        b = O.stream(seed, S, K, G)
        assert np.array_equal(a, b)


def test_synthetic_molecule_properties():
    n = 3002
    xyz, t, m = sdgso.synth_molecule(11, n, 10)
    r = (n * 11.4) ** (1 / 3)
    assert abs(xyz.mean(axis=0)).max() < 1e-9
    assert np.linalg.norm(xyz, axis=1).max() < 2 * r
    steps = np.linalg.norm(np.diff(xyz, axis=0), axis=1)
    assert np.median(steps) == pytest.approx(3.8, abs=1e-9)
    assert t.min() >= 0 and t.max() < 168 and len(np.unique(t)) > 150
    mm = m.reshape(10, -1)
    assert np.abs(mm @ mm.T - np.eye(10)).max() < 1e-12
    x2, t2, m2 = sdgso.synth_molecule(11, n, 10)
    assert np.array_equal(xyz, x2) and np.array_equal(t, t2) and np.array_equal(m, m2)


def test_synthetic_table_and_poses():
    tab = sdgso.synth_table(7)
    assert tab.dtype == np.float32 and tab.size == 168 * 168 * 20
    assert abs(tab.mean()) < 0.01 and tab.std() == pytest.approx(0.5, rel=0.01)
    pos = sdgso.synth_poses(3, 20, 50, 10, 10, 40.0, 10.0)
    q = pos[:, :, 3:7]
    assert np.allclose(np.linalg.norm(q, axis=2), 1.0, atol=1e-12)
    centres = pos[:, :, :3].mean(axis=1)
    assert np.allclose(np.linalg.norm(centres, axis=1), 40.0, atol=3.0)
    spread = np.linalg.norm(pos[:, :, :3] - pos[:, :1, :3] * 0 - centres[:, None, :] * 0, axis=2)
    assert spread.shape == (20, 50)
    assert np.abs(pos[:, :, 7:]).max() <= 1.0


def test_device_bin_formula_matches_dist_to_bins():
    """dev_common.h computes dist_to_bins[d] - 1 (interface.h:30-35) as
    max(0, min(d - 2, (d + 11) >> 1)); check it for every reachable d (-1..29;
    -1 arises for s < 1 where the kernels use (int)s - 1)."""
    dist_to_bins = [1, 1, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 14, 15, 15, 16, 16, 17,
                    17, 18, 18, 19, 19, 20, 20, 21]
    src = open(os.path.join(ROOT, "sparkledock_amd", "csrc", "dev_common.h")).read()
    assert "return max(0, min(d - 2, (d + 11) >> 1));" in src
    for d in range(-1, 30):
        want = dist_to_bins[max(d, 0)] - 1
        assert max(0, min(d - 2, (d + 11) >> 1)) == want, d
