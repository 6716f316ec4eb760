"""The solve server on the GPU: shared resident systems, continuous batching, HBM admission.

Every answer must be bit for bit the answer of the same problem solved alone by a StaticStudy
with the same options (crates/solidfem-elastic/src/study.rs:191-372), whatever else is in flight.
"""
import dataclasses

import numpy as np
import pytest

import solidfem_amd as sf
from workloads import block_config, stretch_components, stretched_cube

pytestmark = pytest.mark.gpu

OPTIONS = dict(precision=sf.Precision.SINGLE, coarse_solve=1, tolerance=1e-8, dot_order=sf.DotOrder.TREE)


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def alone(ctx, w):
    """The displacements of `w` from a StaticStudy built and solved with the server's options."""
    b = sf.StudyBuilder.new(ctx, sf.Model.from_workload(w),
                            sf.PcgOptions(tolerance=OPTIONS["tolerance"], dot_order=OPTIONS["dot_order"]),
                            sf.MultigridOptions(precision=OPTIONS["precision"], coarse_solve=OPTIONS["coarse_solve"]))
    b.step(2 ** 62)
    s = b.finish()
    p = s.solve()
    assert p.outcome == sf.StepOutcome.CONVERGED
    return s.displacements(), p.iteration


def server(**kw):
    return sf.Server.new(0, sf.ServerOptions(**{**OPTIONS, **kw}))


def test_a_served_answer_equals_the_study_alone(ctx):
    w = block_config(2, divisions=(6, 4, 4))
    srv = server()
    rid = srv.submit(w)
    st = srv.wait(rid, 120.0)
    assert st.state == sf.ServeState.DONE, st
    ref, iterations = alone(ctx, w)
    assert st.iterations == iterations
    assert np.array_equal(bits(srv.result(rid)), bits(ref))
    assert st.latency_seconds >= st.solve_seconds > 0 and st.build_seconds > 0 and not st.cache_hit


def test_requests_in_a_mix_equal_themselves_alone(ctx):
    # seven load cases of one block share a system; three other systems differ in material,
    # size or mesh density; everything is in flight at once
    shared = [block_config(c, divisions=(6, 4, 4)) for c in range(7)]
    others = [dataclasses.replace(block_config(1, divisions=(6, 4, 4)), youngs=120000.0),
              dataclasses.replace(block_config(4, divisions=(6, 4, 4)), size=(30.0, 40.0, 40.0)),
              block_config(5, divisions=(8, 4, 4))]
    workloads = shared[:3] + others[:1] + shared[3:] + others[1:]
    srv = server(max_columns=4, slice_iterations=3)
    ids = [srv.submit(w) for w in workloads]
    states = [srv.wait(i, 300.0) for i in ids]
    assert all(s.state == sf.ServeState.DONE for s in states), states
    for w, rid, st in zip(workloads, ids, states):
        ref, iterations = alone(ctx, w)
        assert st.iterations == iterations
        assert np.array_equal(bits(srv.result(rid)), bits(ref)), w
    stats = srv.stats()
    assert stats["builds"] == 4 and stats["cache_hits"] == len(workloads) - 4
    assert stats["completed"] == len(workloads) and stats["failed"] == 0
    assert 0 < stats["p50_seconds"] <= stats["p99_seconds"] <= stats["max_seconds"]


def test_admission_refuses_and_evicts_by_the_budget(ctx):
    small = block_config(0, divisions=(6, 4, 4))
    tiny = server(hbm_budget=1 << 20)
    rid = tiny.submit(small)
    st = tiny.wait(rid, 60.0)
    assert st.state == sf.ServeState.FAILED and st.error == 60
    with pytest.raises(sf.SolidfemError) as e:
        tiny.result(rid)
    assert e.value.status == 60
    # a budget that holds one of these systems at a time (each reserves about 8 MiB with two
    # columns): the idle one is evicted for the next
    one = server(hbm_budget=12 << 20, max_columns=2)
    first = one.submit(small)
    assert one.wait(first, 120.0).state == sf.ServeState.DONE
    other = dataclasses.replace(small, youngs=90000.0)
    second = one.submit(other)
    assert one.wait(second, 120.0).state == sf.ServeState.DONE
    stats = one.stats()
    assert stats["evictions"] == 1 and stats["resident_systems"] == 1 and stats["hbm_peak"] <= 12 << 20
    ref, _ = alone(ctx, other)
    assert np.array_equal(bits(one.result(second)), bits(ref))


def test_a_request_without_load_fails_by_name(ctx):
    srv = server()
    loaded = block_config(0, divisions=(6, 4, 4))
    unloaded = dataclasses.replace(loaded, load_value=17.0)  # no boundary face lies on z = 17
    first = srv.submit(unloaded)
    assert srv.wait(first, 60.0).error == 44
    ok = srv.submit(loaded)
    assert srv.wait(ok, 120.0).state == sf.ServeState.DONE
    again = srv.submit(unloaded)  # now the system is resident: the load alone is refused
    st = srv.wait(again, 60.0)
    assert st.state == sf.ServeState.FAILED and st.error == 44
    with pytest.raises(sf.SolidfemError) as e:
        srv.status(10 ** 9)
    assert e.value.status == 61
    srv.release(ok)
    with pytest.raises(sf.SolidfemError):
        srv.result(ok)


def text_model(text, youngs=70000.0, poisson=0.33, traction=(0.0, 0.0, -1.0)):
    """A model of a mesh sent as text: fixed at x = 0, loaded on the faces at x = max."""
    mesh = sf.QuadraticTetMesh.from_linear(sf.TetMesh.read_text(text))
    model = sf.Model.new(mesh, youngs, poisson)
    nodes = model.mesh().nodes()
    for n in np.flatnonzero(np.abs(nodes[:, 0]) < 1e-9):
        model.fix_node(int(n))
    xmax = nodes[:, 0].max()
    for face in model.mesh().boundary_faces():
        if np.all(np.abs(nodes[face[2:5], 0] - xmax) < 1e-9):
            model.add_traction(face, traction)
    return model


def model_alone(ctx, model):
    b = sf.StudyBuilder.new(ctx, model, sf.PcgOptions(tolerance=OPTIONS["tolerance"], dot_order=OPTIONS["dot_order"]),
                            sf.MultigridOptions(precision=OPTIONS["precision"], coarse_solve=OPTIONS["coarse_solve"]))
    b.step(2 ** 62)
    s = b.finish()
    p = s.solve()
    assert p.outcome == sf.StepOutcome.CONVERGED
    return s.displacements(), p.iteration


def test_models_of_meshes_sent_as_text_equal_themselves_alone(ctx):
    # an unstructured-looking mesh (a box with its interior vertices moved) sent as text: three
    # load cases share its system, another material makes a second, and the stretched cube with
    # prescribed components a third, all in flight together
    box = sf.box_tet_mesh((0.0, 0.0, 0.0), (24.0, 4.0, 4.0), (12, 3, 3))
    pos, tets = box.positions(), box.tets()
    inside = np.all((pos > 1e-9) & (pos < np.array([24.0, 4.0, 4.0]) - 1e-9), axis=1)
    pos[inside] += np.random.default_rng(9).uniform(-0.2, 0.2, (inside.sum(), 3))
    text = sf.TetMesh.text_of(pos, tets)
    models = [text_model(text, traction=t) for t in ((0.0, 0.0, -1.0), (0.0, 0.5, 0.0), (0.2, -0.3, 0.7))]
    models.append(text_model(text, youngs=3500.0, poisson=0.36))
    cube = sf.Model.from_workload(stretched_cube((2, 2, 2)))
    for n, a, v in stretch_components(cube.mesh().nodes()):
        cube.prescribe_displacement(n, a, v)
    models.append(cube)
    srv = server(max_columns=4, slice_iterations=2)
    ids = [srv.submit_model(m) for m in models]
    states = [srv.wait(i, 300.0) for i in ids]
    for k, (m, rid, st) in enumerate(zip(models, ids, states)):
        assert st.state == sf.ServeState.DONE, (k, st)
        ref, iterations = model_alone(ctx, m)
        assert st.iterations == iterations, k
        assert np.array_equal(bits(srv.result(rid)), bits(ref)), k
    assert sum(s.cache_hit for s in states) == 2  # the second and third load cases of the shared mesh
    assert srv.stats()["builds"] == 3
