"""The GPU RMAT generator returns exactly the CPU generator's edge set."""
import numpy as np
import pytest

from tests.conftest import announce

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("scale,edges", [(10, 5000), (16, 1 << 20), (12, 100)])
def test_gpu_rmat_matches_cpu(scale, edges):
    import dycgraph_amd as d
    announce(f"gen rmat scale={scale} edges={edges}")
    s1, d1 = d.gen_rmat(scale, edges, seed=4, gpu=False)
    s2, d2 = d.gen_rmat(scale, edges, seed=4, gpu=True)
    assert np.array_equal(s1, s2) and np.array_equal(d1, d2)
