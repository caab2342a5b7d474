"""The GPU's matrix-free operators against matrices assembled in the
reference's CSR format (panopaea/src/sparse.rs:41-129): derivative_0_primal
against the incidence matrix of dec/grid.rs:373-407 (bit-exact: one
subtraction per edge), the pressure operator against dt (n_c I - adjacency)
(bit-exact on integer-valued data, where every order of summation is exact;
to rounding otherwise)."""
import numpy as np
import pytest

import panopaea_amd as pa
from oracle import oracle as o

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_gpu_d0_primal_matches_incidence_matrix(ctx, dtype):
    h, w = 12, 20
    g = pa.Grid2d((h, w), dtype)
    v = np.random.default_rng(7).standard_normal(g.num_elem_0()).astype(dtype)
    e = g.new_simplex_1()
    g.derivative_0_primal(e, g.new_simplex_0().upload(v))
    assert np.array_equal(e.download(), o.d0_primal_matrix(h, w, dtype).mul_vec(v))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("dims,bc", [((16, 12), (0,) * 6), ((16, 12), (1, 0, 0, 1, 0, 0)),
                                     ((6, 8, 10), (0, 0, 0, 0, 0, 1)), ((6, 8, 10), (1, 1, 0, 0, 1, 0))])
def test_gpu_poisson_matches_csr(ctx, dtype, dims, bc):
    G = pa.Grid2d if len(dims) == 2 else pa.Grid3d
    g = G(dims, dtype, bc=bc)
    P = o.params(dims, bc=bc, dt=0.25)
    A = o.poisson_matrix(P, dtype)
    n = int(np.prod(dims))
    op = pa.PoissonOperator(g, 0.25)
    q = g.new_simplex_2()
    rng = np.random.default_rng(11)
    p = rng.integers(-100, 100, n).astype(dtype)
    op(q, g.new_simplex_2().upload(p))
    assert np.array_equal(q.download(), A.mul_vec(p))
    p = rng.standard_normal(n).astype(dtype)
    op(q, g.new_simplex_2().upload(p))
    want = A.mul_vec(p)
    tol = 1e-5 if dtype == np.float32 else 1e-13
    assert np.abs(q.download() - want).max() <= tol * np.abs(want).max()
