"""CsrMatrix's accessors in the Python mirror (crates/solidfem-solve/src/csr.rs:74-168) on a host
matrix built from the oracle's assembly: size, nonzeros, row, get and diagonal read the arrays the
matrix was made from, and read() returns them unchanged."""
import numpy as np

import refcpu
import solidfem_amd as sf
from workloads import loaded


def test_accessors_read_the_arrays_the_matrix_was_made_from():
    rs, cols, vals = refcpu.Problem(loaded([2.0, 1.0, 1.0], [2, 1, 1])).csr()
    m = sf.CsrMatrix.new(None, rs, cols, vals)
    r_rs, r_cols, r_vals = m.read()
    assert np.array_equal(r_rs, rs) and np.array_equal(r_cols, cols)
    assert np.array_equal(r_vals.view(np.uint64), vals.view(np.uint64))
    assert m.size == len(rs) - 1 and m.nonzeros() == len(vals)
    for i in (0, 5, len(rs) - 2):
        c, v = m.row(i)
        assert np.array_equal(c, cols[rs[i]:rs[i + 1]]) and np.array_equal(v, vals[rs[i]:rs[i + 1]])
        assert m.get(i, int(c[1])) == v[1] and m.get(i, 10 ** 6) == 0.0
    dense = np.zeros((m.size, m.size))
    for i in range(m.size):
        dense[i, cols[rs[i]:rs[i + 1]]] = vals[rs[i]:rs[i + 1]]
    assert np.array_equal(m.diagonal(), np.diag(dense))
