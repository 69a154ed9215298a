"""The learned policy's predictor on the GPU (FP16 weights, FP32 accumulation,
MFMA for the 64x64 layer) against the FP64 reference in oracle/: outputs and
every gradient of the loss."""
import numpy as np
import pytest

import oracle
from tests.conftest import announce

pytestmark = pytest.mark.gpu


def _params(rng, scale=0.3):
    shapes = [(8, 64), (64,), (64, 64), (64,), (64, 2), (2,)]
    return [rng.normal(0, scale, s).astype(np.float32) for s in shapes]


def _half(ps):
    # the device keeps W1/W2/W3 in FP16 (biases FP32): compare against the same rounding
    return [p.astype(np.float16).astype(np.float32) if p.ndim == 2 else p for p in ps]


@pytest.mark.parametrize("rows", [1, 31, 32, 100, 1000])
def test_mlp_forward_and_gradients_match_fp64(rows):
    import dycgraph_amd.engine as E
    announce(f"mlp rows={rows}")
    rng = np.random.default_rng(rows)
    ps = _params(rng)
    x = np.abs(rng.normal(0, 2, (rows, 8))).astype(np.float32)
    y = np.stack([rng.integers(0, 2, rows), np.abs(rng.normal(2, 1, rows))], 1).astype(np.float32)
    out, grads = E.mlp_eval(x, ps, y)
    ref_out, _, ref_g = oracle.mlp_forward_backward(x, y, *_half(ps))
    # FP16 operands (activations and layer-2 gradients) bound the agreement
    assert np.max(np.abs(out - ref_out)) <= 2e-2 * max(1.0, np.max(np.abs(ref_out)))
    for name, g, rg in zip(["W1", "b1", "W2", "b2", "W3", "b3"], grads, ref_g):
        err = np.max(np.abs(g - rg))
        scale = max(np.max(np.abs(rg)), 1e-3)
        assert err <= 3e-2 * scale, (name, err, scale)
