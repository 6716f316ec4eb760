"""Error measures of an approximate solve against another (test infrastructure).

For SPD A and two approximations x1, x2 of A x = b with residuals r_i = b - A x_i,
x1 - x2 = A^-1 (r2 - r1), so

    ||x1 - x2||_A  <=  ||r1||_{A^-1} + ||r2||_{A^-1},      ||r||_{A^-1} = sqrt(r . A^-1 r).

The residuals are formed in long double (64-bit significand), so the rounding of b - A x in f64,
which is the size of the residual itself at the rounding floor, does not enter them. A^-1 r comes
from the oracle's multigrid-preconditioned conjugate gradient to a relative residual of 1e-6: from
a zero start its energy r . y_k approaches r . A^-1 r from below, short by ||y - y_k||_A^2, at
most kappa * 1e-12 of it relative (about 1e-4 for the plates' condition numbers), well inside
the 1% slack the callers allow.
"""
import numpy as np


def _rows(rs, cols, vals, x):
    prod = vals.astype(np.longdouble) * np.asarray(x, dtype=np.float64)[cols].astype(np.longdouble)
    out = np.zeros(len(rs) - 1, dtype=np.longdouble)
    nonempty = np.diff(rs) > 0
    out[nonempty] = np.add.reduceat(prod, rs[:-1][nonempty])
    return out


def residual(rs, cols, vals, b, x):
    """b - A x, summed in long double, rounded once to f64."""
    return (np.asarray(b, dtype=np.longdouble) - _rows(rs, cols, vals, x)).astype(np.float64)


def energy(rs, cols, vals, v):
    """||v||_A = sqrt(v . A v) in long double."""
    v = np.asarray(v, dtype=np.float64)
    return float(np.sqrt(np.dot(v.astype(np.longdouble), _rows(rs, cols, vals, v))))


def inverse_norm(oracle_multigrid, r, tolerance=1e-6):
    """||r||_{A^-1} from the oracle's solve of A y = r."""
    if not np.any(r):
        return 0.0
    y, res = oracle_multigrid.solve_rhs(r, tolerance)
    assert res.status == 0, f"the oracle's inverse solve failed ({res.status})"
    return float(np.sqrt(max(0.0, float(np.dot(r.astype(np.longdouble), y.astype(np.longdouble))))))
