"""The issue-cost probe behind the pair kernels' issue-bound estimates:
plausible clocks per wave64 instruction, and argument checks."""
import numpy as np
import pytest

import collop

pytestmark = pytest.mark.gpu


def test_issue_probe(gpu_ctx):
    clk, ms = np.zeros(1), np.zeros(1)
    res = {}
    for kind in (0, 5):
        gpu_ctx.check(gpu_ctx.lib.collop_issue_probe(gpu_ctx.h, kind, 2000, collop._lib.dptr(clk),
                                                     collop._lib.dptr(ms)))
        res[kind] = float(clk[0])
    print("clocks per wave64 instruction: fma_f32 %.2f, rsq_f64 %.2f" % (res[0], res[5]))
    # a wave64 instruction on a 16-lane SIMD takes at least 4 clocks; the
    # transcendental FP64 reciprocal square root is slower than an FMA
    assert 2.0 < res[0] < 16.0 and res[5] > res[0]
    with pytest.raises(collop.CollopError):
        gpu_ctx.check(gpu_ctx.lib.collop_issue_probe(gpu_ctx.h, 9, 10, collop._lib.dptr(clk), None))
