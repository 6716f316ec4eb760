"""The peak stress (crates/solidfem-elastic/src/results.rs:149-226) on the host, against the oracle.

The oracle's peak is pinned by what the reference's own tests ask of it
(crates/solidfem/tests/validation/peak.rs:147-166 and comparison.rs:535-599): over a region it is
the larger of the largest node and the largest integration point inside, bit for bit (a brute-force
numpy maximum here); the position it reports holds that value; of equal values the first node,
then the first point, holds it; an empty region has no peak. A NaN is kept (results.rs:167-169):
the first NaN met is the peak, with its position. The product's peak (sf_peak_stress, whose region
is a callback) must give the oracle's value, position and holder for regions the oracle receives as
masks. The recovered fields come from the oracle's recovery of a seeded random field on a
cantilever (E 1000, nu 0.3, as the reference's tests use); the peak reads only those values.
"""
import numpy as np
import pytest

import refcpu
import solidfem_amd as sf
from workloads import loaded


@pytest.fixture(scope="module")
def field():
    w = loaded([2.0, 1.0, 1.0], [4, 2, 2])
    w.youngs, w.poisson = 1000.0, 0.3
    p = refcpu.Problem(w)
    u = 1e-3 * np.random.default_rng(5).standard_normal((p.node_count, 3))
    points, nodal = p.recover(u)
    return p.nodes(), nodal[:, 6].copy(), points.reshape(-1, 10).copy()


REGIONS = {
    "whole": lambda p: True,
    "far half": lambda p: p[0] >= 1.0,
    "away from the clamp": lambda p: p[0] > 0.25,
    "a slab": lambda p: 0.3 <= p[2] <= 0.7,
}


def masks(x, points, inside):
    return (np.array([inside(p) for p in x], dtype=bool), np.array([inside(p[:3]) for p in points], dtype=bool))


def brute_force(x, vm, points, inside):
    """largest_point.max(largest_node) of peak.rs:102-122 over the region."""
    node_in, point_in = masks(x, points, inside)
    values = np.concatenate([vm[node_in], points[point_in, 9]])
    return values.max() if values.size else None


@pytest.mark.parametrize("name", list(REGIONS))
def test_oracle_peak_is_the_largest_value_inside_and_its_position_holds_it(field, name):
    x, vm, points = field
    inside = REGIONS[name]
    node_in, point_in = masks(x, points, inside)
    value, position, at_point, index = refcpu.peak(x, vm, points, node_in, point_in)
    assert np.float64(value).view(np.uint64) == np.float64(brute_force(x, vm, points, inside)).view(np.uint64)
    assert inside(position)
    if at_point:
        assert points[index, 9] == value and np.array_equal(points[index, :3], position)
    else:
        assert vm[index] == value and np.array_equal(x[index], position)


def test_oracle_peak_of_equal_values_is_the_first_node_then_the_first_point(field):
    x, vm, points = field
    ones_n, ones_p = np.ones_like(vm), points.copy()
    ones_p[:, 9] = 1.0
    value, position, at_point, index = refcpu.peak(x, ones_n, ones_p)
    assert (value, at_point, index) == (1.0, False, 0) and np.array_equal(position, x[0])
    # with every node outside, the first integration point holds the tie
    value, position, at_point, index = refcpu.peak(x, ones_n, ones_p, np.zeros(len(x), bool), None)
    assert (at_point, index) == (True, 0) and np.array_equal(position, ones_p[0, :3])


def test_oracle_peak_keeps_the_first_nan(field):
    x, vm, points = field
    vm2, p2 = vm.copy(), points.copy()
    vm2[7] = np.nan
    p2[3, 9] = np.nan
    value, position, at_point, index = refcpu.peak(x, vm2, p2)
    assert np.isnan(value) and (at_point, index) == (False, 7) and np.array_equal(position, x[7])
    vm2[7] = vm[7]
    value, position, at_point, index = refcpu.peak(x, vm2, p2)
    assert np.isnan(value) and (at_point, index) == (True, 3)
    # a NaN after the largest finite value still replaces it
    vm3 = vm.copy()
    vm3[-1] = np.nan
    assert np.isnan(refcpu.peak(x, vm3, points)[0])


def test_oracle_peak_of_an_empty_region_is_none(field):
    x, vm, points = field
    assert refcpu.peak(x, vm, points, np.zeros(len(x), bool), np.zeros(len(points), bool)) is None


@pytest.mark.parametrize("name", list(REGIONS))
def test_peak_stress_is_the_oracle_peak(field, name):
    x, vm, points = field
    inside = REGIONS[name]
    node_in, point_in = masks(x, points, inside)
    value, position, at_point, index = refcpu.peak(x, vm, points, node_in, point_in)
    peak = sf.peak_stress(x, vm, points, inside)
    assert np.float64(peak.von_mises).view(np.uint64) == np.float64(value).view(np.uint64)
    assert np.array_equal(peak.position, position) and (peak.at_point, peak.index) == (at_point, index)
    assert sf.peak_von_mises(x, vm, points, inside) == value


def test_peak_stress_edge_cases_match_the_oracle(field):
    x, vm, points = field
    assert sf.peak_stress(x, vm, points, lambda p: p[0] < -1.0) is None
    assert sf.peak_von_mises(x, vm, points, lambda p: p[0] < -1.0) is None
    whole = sf.peak_von_mises(x, vm, points)
    assert whole >= sf.peak_von_mises(x, vm, points, REGIONS["far half"])
    ones_n, ones_p = np.ones_like(vm), points.copy()
    ones_p[:, 9] = 1.0
    on_nodes = {tuple(p) for p in x}
    peak = sf.peak_stress(x, ones_n, ones_p, lambda p: tuple(p) not in on_nodes)
    ref = refcpu.peak(x, ones_n, ones_p, np.zeros(len(x), bool), np.array([tuple(p[:3]) not in on_nodes for p in ones_p]))
    assert (peak.at_point, peak.index) == (ref[2], ref[3])
    vm2, p2 = vm.copy(), points.copy()
    vm2[7], p2[3, 9] = np.nan, np.nan
    peak = sf.peak_stress(x, vm2, p2)
    assert np.isnan(peak.von_mises) and (peak.at_point, peak.index) == (False, 7)
    # empty candidate lists
    assert sf.peak_stress(np.empty((0, 3)), np.empty(0), np.empty((0, 10))) is None


def test_a_region_callback_that_raises_is_reported(field):
    x, vm, points = field

    def broken(p):
        raise KeyError("region")

    with pytest.raises(KeyError):
        sf.peak_stress(x, vm, points, broken)
