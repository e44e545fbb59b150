"""Host helpers of the drop-in that the reference exports beside the solvers
(gradient_domain/__init__.py): sparse_laplacian (operators.py:110-136) and
manufactured_problem (benchmark.py:65-95).  They build matrices and inputs on the host;
checked against the oracle and against the dense operator they define."""

import numpy as np
import pytest

import gdpoisson
from oracle import generators as G
from oracle import operators as O


@pytest.mark.parametrize("shape", [(63, 63), (47, 61), (1, 5), (127, 31)])
@pytest.mark.parametrize("seed", [0, 3])
def test_manufactured_problem_is_the_oracle_s(shape, seed):
    b, x = gdpoisson.manufactured_problem(shape, seed=seed)
    ob, ox = G.manufactured_problem(shape, seed=seed)
    assert np.array_equal(b, ob) and np.array_equal(x, ox)
    # the solution is exact for the zero-outside five-point operator
    assert np.abs(O.laplacian(x) - b).max() < 1e-12


@pytest.mark.parametrize("shape,h", [((7, 9), 1.0), ((13, 19), 0.5), ((1, 4), 2.0),
                                     ((5, 1), 1.0)])
def test_sparse_laplacian_is_the_dense_operator(shape, h):
    """test_operators.py:60-71: the sparse matrix applies the dense operator and is
    symmetric negative definite."""
    a = gdpoisson.sparse_laplacian(shape, h)
    assert a.shape == (shape[0] * shape[1],) * 2
    assert (a != O.sparse_laplacian(shape, h)).nnz == 0
    u = np.random.default_rng(11).standard_normal(shape)
    assert np.allclose((a @ u.reshape(-1)).reshape(shape), O.laplacian(u, h))
    d = a.toarray()
    assert np.array_equal(d, d.T) and np.linalg.eigvalsh(d).max() < 0.0
