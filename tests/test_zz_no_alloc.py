"""hd_norm_and_error_noalloc (the entry point that allocates nothing; its
scratch is a fixed device array) against hd_norm_and_error, through the C
ABI: two calls in a row, the second on a smaller mesh, so the second reads
only the partials of its own launch out of the scratch the first filled.

Bar: both norms within 1e-13 relative of hd_norm_and_error's (FP64 sums in
a different order: block partials on the device instead of per-cell
partials summed on the host)."""

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu
H = None


def setup_module(module):
    global H
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import hyperdeal_amd
    H = hyperdeal_amd
    H.load_library()


CASES = [
    # dims, degree, n_cells (first call larger than the scratch's block count,
    # the second smaller than it)
    ((3, 3), 3, [4, 4, 4, 4, 4, 4], (2, 2), 3, [3, 3, 3, 2]),
    ((2, 2), 5, [8, 8, 8, 8], (1, 1), 2, [5, 3]),
]


@pytest.mark.parametrize("d1,k1,n1,d2,k2,n2", CASES)
def test_noalloc_norm_equals_norm_and_error(d1, k1, n1, d2, k2, n2):
    print("case: noalloc norm", d1, k1, n1, "then", d2, k2, n2, flush=True)
    out = torch.full((2,), -1.0, dtype=torch.float64, device="cuda")
    for dims, k, n in ((d1, k1, n1), (d2, k2, n2)):
        D = sum(dims)
        a = [0.3 * (d + 1) for d in range(D)]
        mf = H.MatrixFree(dims[0], dims[1], k, n, [0.0] * D, [1.0] * D)
        u = mf.initialize_dof_vector()
        H.VectorTools.interpolate(mf, u, H.FUNC_PRODUCT_SINES, a, 0.0)
        H.VectorTools.add_noise(mf, u, 20261203, 1e-2)
        t = 0.013
        ref = H.VectorTools.norm_and_error(mf, u, H.FUNC_PRODUCT_SINES, a, t)
        H.VectorTools.norm_and_error_noalloc(mf, u, H.FUNC_PRODUCT_SINES, out,
                                             a, t)
        got = out.cpu().numpy()
        assert abs(got[0] - ref[0]) <= 1e-13 * ref[0]
        assert abs(got[1] - ref[1]) <= 1e-13 * ref[1]
        assert ref[1] > 0.0
