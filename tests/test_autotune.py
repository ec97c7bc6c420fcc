"""The autotuner (miniappsforhpe_amd.autotune) picks a (z depth, strip) candidate by timing and leaves the propagator
and the batch in a state whose steps give the same bits as the default schedule."""
import numpy as np
import pytest
import torch

from test_parity_fused import interior_random, setup
from test_parity_type2 import to_gpu

pytestmark = pytest.mark.gpu


def test_tune_propagator(rng, cuda):
    from miniappsforhpe_amd import autotune

    G, OG, m, sp, P = setup((150, 33, 40), rng, cuda)
    init = [to_gpu(interior_random(G, rng), cuda) for _ in range(4)]

    def steps():
        P.reset()
        P.s1[0].copy_(init[0]); P.s2[0].copy_(init[1]); P.u.copy_(init[2]); P.w.copy_(init[3])
        P.prime()
        for _ in range(2):
            P.step_fwd_fused()
        P.step_adj_fused()
        torch.cuda.synchronize()
        return [t.clone() for t in (*P.sigma, P.u, P.w)]

    P.zblock, P.strip = 64, 0
    ref = steps()
    res = autotune.tune_propagator(P, zblocks=(16, 64), strips=(0, 1), reps=1)
    print("autotune:", res, flush=True)
    assert (res["zblock"], res["strip"]) in {(16, 0), (16, 1), (40, 0), (40, 1)}
    assert (P.zblock, P.strip) == (res["zblock"], res["strip"])
    assert len(res["ms"]) == 4 and all(v > 0 for v in res["ms"].values())
    for a, b in zip(ref, steps()):
        assert torch.equal(a, b)
