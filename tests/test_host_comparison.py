"""The product's mesh comparison (sf_mesh_comparison_between, host code) against the oracle's, and
the reference's verdict and safety-factor rules (crates/solidfem-elastic/src/comparison.rs:143-222,
crates/solidfem/tests/validation/comparison.rs), on the oracle's solves and recoveries of the
reference's cantilevers and tension bar. Given the same arrays the two comparisons must agree bit
for bit: peaks, holders, element counts, volumes, change and distance.
"""
import math

import numpy as np
import pytest

import refcpu
import solidfem_amd as sf
from test_oracle_comparison import CANTILEVER, away, clamped_bar

YIELD = 100.0


def bits(x):
    return np.float64(x).view(np.uint64)


def product_side(model, case):
    _, nodal, points = case
    return (model.mesh(), nodal, points)


@pytest.fixture(scope="module")
def cases():
    out = {n: clamped_bar(CANTILEVER, (4 * n, n, n), 1000.0, 0.3, (0.0, 0.0, -1.0)) for n in (1, 2, 4, 8)}
    out["base"] = clamped_bar(CANTILEVER, (16, 4, 4), 1000.0, 0.3, (0.0, 0.0, -1.0))
    out["near"] = clamped_bar(CANTILEVER, (20, 5, 5), 1000.0, 0.3, (0.0, 0.0, -1.0))
    out["tension-coarse"] = clamped_bar((20.0, 1.0, 1.0), (10, 1, 1), 2100.0, 0.0, (1000.0, 0.0, 0.0))
    out["tension-fine"] = clamped_bar((20.0, 1.0, 1.0), (20, 2, 2), 2100.0, 0.0, (1000.0, 0.0, 0.0))
    return out


def compare_both(coarse, fine, inside=None):
    pc = (sf.Model.from_workload(coarse[0].workload), coarse)
    pf = (sf.Model.from_workload(fine[0].workload), fine)
    product = sf.MeshComparison.between(product_side(*pc), product_side(*pf), inside)
    oracle = refcpu.compare(coarse, fine, inside)
    assert oracle["status"] == 0
    for name, side in (("coarse", product.coarse()), ("fine", product.fine())):
        o = oracle[name]
        assert bits(side.peak.von_mises) == bits(o["von_mises"]) and np.array_equal(side.peak.position, o["position"])
        assert (side.peak.at_point, side.peak.index, side.elements) == (o["at_point"], o["index"], o["elements"])
        assert bits(side.peak_element_volume) == bits(o["volume"])
    assert bits(product.peak_change()) == bits(oracle["peak_change"])
    assert bits(product.peak_distance()) == bits(oracle["peak_distance"])
    return product


PAIRS = [(1, 2), (2, 4), (4, 8), ("base", "near"), ("base", 8), ("tension-coarse", "tension-fine")]


@pytest.mark.parametrize("pair", PAIRS, ids=lambda p: f"{p[0]}-{p[1]}")
@pytest.mark.parametrize("region", ["whole", "away"])
def test_the_comparison_is_the_oracles(cases, pair, region):
    compare_both(cases[pair[0]], cases[pair[1]], None if region == "whole" else away)


def test_the_tension_bar_is_settled_and_its_safety_factors_agree(cases):
    c = compare_both(cases["tension-coarse"], cases["tension-fine"])
    assert c.peak_element_ratio() == 8.0
    assert c.verdict(2e-5 * 1.001) == sf.Verdict.SETTLED
    assert c.verdict(math.nan) == sf.Verdict.STILL_MOVING
    stress, yield_strength = 1000.0, 250.0
    for factor in c.safety_factors(yield_strength):
        assert abs(factor - yield_strength / stress) / (yield_strength / stress) < 1e-5 * 1.001
    assert abs(c.safety_factor_change(yield_strength)) < 2e-5 * 1.001


def assert_safety_factors(c):
    """validation/comparison.rs:70-95: yield over each peak, and their change opposite the peak's."""
    coarse, fine = c.safety_factors(YIELD)
    assert bits(coarse) == bits(YIELD / c.coarse().peak.von_mises) and bits(fine) == bits(YIELD / c.fine().peak.von_mises)
    change = c.safety_factor_change(YIELD)
    assert bits(change) == bits((fine - coarse) / fine) and change * c.peak_change() < 0.0


def test_verdicts_of_the_cantilevers(cases):
    for a, b in [(1, 2), (2, 4), (4, 8)]:
        c = compare_both(cases[a], cases[b], away)
        assert_safety_factors(c)
        assert c.element_ratio() == 8.0 and c.peak_element_ratio() == 8.0
        assert c.verdict(math.nan) == sf.Verdict.STILL_MOVING
        whole = compare_both(cases[a], cases[b])
        assert_safety_factors(whole)
        assert whole.verdict(0.05) == sf.Verdict.STILL_MOVING
    small_step = compare_both(cases["base"], cases["near"], away)
    assert small_step.element_ratio() < sf.MIN_ELEMENT_RATIO
    assert small_step.verdict(0.003) == sf.Verdict.MESHES_TOO_ALIKE
    assert small_step.verdict(math.nan) == sf.Verdict.MESHES_TOO_ALIKE
    halving = compare_both(cases["base"], cases[8], away)
    assert halving.verdict(0.003) == sf.Verdict.STILL_MOVING


def test_refusals_are_the_oracles(cases):
    model = {k: sf.Model.from_workload(cases[k][0].workload) for k in (1, 2)}
    side = {k: product_side(model[k], cases[k]) for k in (1, 2)}
    with pytest.raises(sf.ModelError) as e:
        sf.MeshComparison.between(side[1], side[2], lambda p: p[0] > 10.0)
    assert e.value.status == 59 and e.value.a == sf_compared("coarse")
    with pytest.raises(sf.ModelError) as e:
        sf.MeshComparison.between(side[2], side[1], away)
    assert e.value.status == 72 and (e.value.a, e.value.b) == (96, 12)
    c = sf.MeshComparison.between(side[1], side[2])
    for bad in (0.0, -1.0, math.inf, math.nan):
        with pytest.raises(sf.ModelError) as e:
            c.safety_factors(bad)
        assert e.value.status == 73
    # arrays that do not match the mesh
    mesh, nodal, points = side[1]
    with pytest.raises(sf.ModelError) as e:
        sf.MeshComparison.between((mesh, nodal[:-1], points), side[2])
    assert e.value.status == 58 and e.value.a == sf_compared("coarse")
    with pytest.raises(sf.ModelError) as e:
        sf.MeshComparison.between(side[1], (side[2][0], side[2][1], side[2][2].reshape(-1, 10)[:-1]))
    assert e.value.status == 58 and e.value.a == sf_compared("fine")
    # a region callback that raises is reported
    with pytest.raises(KeyError):
        sf.MeshComparison.between(side[1], side[2], lambda p: {}["x"])


def sf_compared(name):
    return {"coarse": 0, "fine": 1}[name]


def test_each_refusal_is_named_as_the_reference_names_it():
    """validation/comparison.rs:420-530, through the product and the oracle."""
    size = (2.0, 1.0, 1.0)
    coarse = clamped_bar(size, (2, 1, 1), 1000.0, 0.0, (0.0, 0.0, -1.0))
    fine = clamped_bar(size, (4, 2, 2), 1000.0, 0.0, (0.0, 0.0, -1.0))
    mesh = {id(c[0]): sf.Model.from_workload(c[0].workload).mesh() for c in (coarse, fine)}

    def product(a, b, inside=None, mesh_of=None):
        ma = mesh[id(a[0])] if mesh_of is None else mesh_of[0]
        mb = mesh[id(b[0])] if mesh_of is None else mesh_of[1]
        with pytest.raises(sf.ModelError) as e:
            sf.MeshComparison.between((ma, a[1], a[2]), (mb, b[1], b[2]), inside)
        return e.value

    def both(a, b, inside=None):
        e = product(a, b, inside)
        o = refcpu.compare(a, b, inside)
        assert o["status"] == e.status and o["a"] == e.a
        return e

    e = both(fine, coarse)
    assert e.status == 72 and (e.a, e.b) == (96, 12)
    assert both(coarse, coarse).status == 72
    # the clamp face holds nodes but no integration point
    e = both(coarse, fine, lambda p: p[0] == 0.0)
    assert e.status == 59 and e.a == 0
    assert both(coarse, fine, lambda p: p[0] > 10.0).status == 59
    first = coarse[2].reshape(-1, 10)[0, :3].copy()
    e = both(coarse, fine, lambda p: np.array_equal(p[:3], first))
    assert e.status == 59 and e.a == 1
    unstressed = (fine[0], fine[1].copy(), fine[2].copy())
    unstressed[1][:, 6] = 0.0
    unstressed[2][..., 9] = 0.0
    e = both(coarse, unstressed)
    assert e.status == 71 and e.a == 1
    broken = (coarse[0], coarse[1].copy(), coarse[2])
    broken[1][0, 6] = math.inf
    e = both(broken, fine)
    assert e.status == 70 and e.a == 0 and e.value == math.inf
    nan = (coarse[0], coarse[1], coarse[2].copy())
    nan[2].reshape(-1, 10)[5, 9] = math.nan
    e = both(nan, fine)
    assert e.status == 70 and e.a == 0 and math.isnan(e.value)
    # the coarse solution with the fine mesh; a fine solution one point short
    e = product(coarse, fine, mesh_of=(mesh[id(fine[0])], mesh[id(fine[0])]))
    assert e.status == 58 and e.a == 0
    short = (fine[0], fine[1], fine[2].reshape(-1, 10)[:-1])
    e = product(coarse, short)
    assert e.status == 58 and e.a == 1
    c = sf.MeshComparison.between((mesh[id(coarse[0])], coarse[1], coarse[2]), (mesh[id(fine[0])], fine[1], fine[2]))
    for bad in (0.0, -250.0, math.inf, math.nan):
        for call in (c.safety_factors, c.safety_factor_change):
            with pytest.raises(sf.ModelError) as err:
                call(bad)
            assert err.value.status == 73 and (bits(err.value.value) == bits(bad) or math.isnan(bad))
