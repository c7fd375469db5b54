"""The benchmark's cube-as-mesh workload builds the same lists as the
oracle's cube mesh (which the unstructured parity tests run)."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))


def test_cube_lists_match_oracle(oracle):
    import bench
    for nx in (1, 4, 7):
        a, b = bench.cube_lists(nx), oracle.cube_mesh(nx)
        assert set(a) == set(b)
        for k in a:
            assert np.array_equal(np.asarray(a[k]), np.asarray(b[k])), (nx, k)
            assert np.asarray(a[k]).dtype == np.asarray(b[k]).dtype, (nx, k)
