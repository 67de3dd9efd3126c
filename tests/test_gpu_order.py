"""Work order of the FP32 scoring launches through the C-ABI: the waves take
their (glowworm, ligand block) items in list order or in chunks per XCD, and
every order scores every glowworm identically (scores and pair counts bit for
bit, pair counts equal to the oracle's); the tuner times both orders and
keeps the faster."""
import numpy as np
import pytest

import gpu_cases as GC
from sparkledock_amd import sdgso

pytestmark = pytest.mark.gpu


def test_every_work_order_scores_identically():
    S, G = 6, 64
    c = GC.case(91, 800, 300, 10, 10, S, G)   # 5 ligand blocks: 384 items of one block, 192 pairs
    pr = GC.Pair(c)
    poses = c["pos"].reshape(-1, pr.P)
    _, _, p0 = pr.oracle_scores(poses)
    names = sdgso.tune_variants()
    for v in range(len(names)):
        got = []
        for k in (0, 1, 3, 16, 8192):
            run = sdgso.Run(pr.gpu, S, G, c["pos"], None, 0, form=sdgso.FP32_GUARD)
            run.set_variant(v)
            run.set_work_order(k)
            assert run.work_order()[0] == k
            run.phases(1)
            st = run.read()
            run.close()
            got.append((st["score"], st["pairs"]))
        for s, p in got:
            assert np.array_equal(p, p0), names[v]
            assert np.array_equal(s, got[0][0]), names[v]


def test_tuner_keeps_the_faster_order():
    S, G = 4, 50
    c = GC.case(92, 900, 500, 10, 10, S, G)
    pr = GC.Pair(c)
    run = sdgso.Run(pr.gpu, S, G, c["pos"], None, 0, form=sdgso.FP32_GUARD)
    assert run.work_order() == (0, (0.0, 0.0))
    run.tune()
    k, (t_list, t_chunk) = run.work_order()
    assert t_list > 0 and t_chunk > 0
    assert k == (8192 if t_chunk < t_list else 0)
    with pytest.raises(sdgso.SdgsoError):
        run.set_work_order(-1)
    run.close()
