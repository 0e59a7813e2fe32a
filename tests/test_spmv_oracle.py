"""The SpMV oracle (orc_spmv) pinned on the CPU: scipy's CSR product (duplicates add up) and
closed forms.  With the reference's inputs (data[i] = i % 13 + 0.5, x[i] = i % 17 + 0.5) every
product is a multiple of 1/4 and every row sum is far below 2^53, so y is exact in any order."""
import numpy as np
import scipy.sparse as sp

from conftest import random_csr


def test_matches_scipy(orc):
    rng = np.random.default_rng(3)
    row, col, _ = random_csr(rng, 500, 4000, 1, 1)
    data, x = orc.spmv_reference_values(500, len(col))
    A = sp.csr_matrix((data, col, row), shape=(500, 500))
    assert np.array_equal(orc.spmv(row, col, data, x), A @ x)        # exact: quarter-integers
    d2 = rng.standard_normal(len(col))
    x2 = rng.standard_normal(500)
    assert np.allclose(orc.spmv(row, col, d2, x2), sp.csr_matrix((d2, col, row), shape=(500, 500)) @ x2, rtol=1e-12)


def test_closed_forms(orc):
    n = 9
    row = np.arange(n + 1, dtype=np.int32)
    col = np.arange(n, dtype=np.int32)                                     # identity
    x = np.arange(n) * 1.5
    assert np.array_equal(orc.spmv(row, col, np.ones(n), x), x)
    row = np.array([0, 3, 3], np.int32)                                    # a row of three, an empty row
    assert np.array_equal(orc.spmv(row, np.array([1, 1, 0], np.int32), np.array([2.0, 3.0, 4.0]), np.array([10.0, 1.0])),
                          [45.0, 0.0])
