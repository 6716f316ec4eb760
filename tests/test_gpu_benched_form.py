"""The benchmarked solver forms against the oracle's FP64 solve, end to end.

The benchmark solves the jittered plate with TREE dot products and the explicit-inverse coarse
solve, with the FP32 V-cycle (Precision::Single) and in plain FP64 (Precision::Double), to
rtol 1e-8. Here both forms, at two sizes the oracle finishes in seconds, are compared with the
oracle's Precision::Double solve in the reference's own order (serial dot products, the reference
coarse solve), which is the reference's answer to the bit.

Tolerance, derived from the attained residuals (tests/errors.py): the energy-norm difference of
two approximations is at most the sum of their residuals' A^-1 norms, so

    ||x_gpu - x_oracle||_A / ||x_oracle||_A  <=  1.01 (||r_gpu||_{A^-1} + ||r_oracle||_{A^-1}) / ||x_oracle||_A

with the residuals formed in long double. The displacement error is also held to the 1e-7 of the
target (max |x_gpu - x_oracle| / max |x_oracle|).
"""
import numpy as np
import pytest

import errors
import refcpu
import solidfem_amd as sf
from workloads import plate_config

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("divisions", [(24, 24, 2), (48, 48, 2)], ids=["24x24x2", "48x48x2"])
def test_the_benched_forms_agree_with_the_oracles_fp64_solve(ctx, divisions):
    w = plate_config(divisions=divisions)
    ref = refcpu.Problem(w)
    rs, cols, vals = ref.csr()
    b = ref.rhs()
    fm = ref.free_map()
    oracle_mg = refcpu.Multigrid(ref, refcpu.options(precision=0, coarse_dofs=1500, threads=0))
    disp, res = oracle_mg.solve(1e-8)
    assert res.status == 0
    x_o = disp.reshape(-1, 3)[fm != 0xFFFFFFFF].reshape(-1)
    norm_o = errors.energy(rs, cols, vals, x_o)
    r_o = errors.inverse_norm(oracle_mg, errors.residual(rs, cols, vals, b, x_o))

    options = sf.PcgOptions(tolerance=1e-8, dot_order=sf.DotOrder.TREE)
    builder = sf.StudyBuilder.new(ctx, sf.Model.from_workload(w), options, sf.MultigridOptions(precision=sf.Precision.DOUBLE))
    builder.step(2 ** 62)
    study = builder.finish()
    mg64 = study.multigrid().with_precision(sf.Precision.DOUBLE, coarse_solve=1)
    mg32 = mg64.with_precision(sf.Precision.SINGLE)
    for name, mg in (("fp64", mg64), ("fp32", mg32)):
        pcg = sf.Pcg.new(ctx, study.system(), mg, study.rhs(), options)
        p = pcg.step(2 ** 62)
        assert p.outcome == sf.StepOutcome.CONVERGED
        x = pcg.estimate()
        e_energy = errors.energy(rs, cols, vals, x - x_o) / norm_o
        r_x = errors.inverse_norm(oracle_mg, errors.residual(rs, cols, vals, b, x))
        bound = 1.01 * (r_x + r_o) / norm_o
        e_disp = float(np.abs(x - x_o).max() / np.abs(x_o).max())
        print(f"{divisions} {name}: {p.iteration} iterations (oracle {res.iterations}), residual {p.relative_residual:.2e} "
              f"floor {p.at_rounding_floor}; energy error {e_energy:.3e} <= bound {bound:.3e}; displacement error {e_disp:.3e}")
        assert e_energy <= bound, (name, e_energy, bound)
        assert e_disp <= 1e-7, (name, e_disp)
