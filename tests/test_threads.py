"""Thread-count invariance (crates/solidfem/tests/validation/multigrid_threads.rs:108-160).

* The oracle's row-chunk threading (WorkerPool, balanced_bounds and for_rows of
  crates/solidfem-solve/src/parallel.rs): a solve on 1, 2 and 7 threads, in all three precisions,
  and on the ceiling of 64, gives the same iterations and displacement bits.
* The product's host setup pool: the hierarchy built with 1, 3 and 8 host threads is the same to
  the bit (each count in a process of its own, since the pool is sized once per process).
"""
import hashlib
import os
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

import refcpu
from workloads import FINGERPRINT_CANTILEVER, FINGERPRINT_PLATE, block_config, loaded, plate_config

ROOT = Path(__file__).resolve().parent.parent


def solve(w, precision, threads):
    p = refcpu.Problem(w)
    mg = refcpu.Multigrid(p, refcpu.options(precision=precision, coarse_dofs=100, threads=threads))
    disp, r = mg.solve(1e-9)
    return mg.threads, r.iterations, refcpu.fingerprint(disp)


@pytest.mark.parametrize("name", ["cantilever", "plate"])
@pytest.mark.parametrize("precision", [0, 1, 2], ids=["double", "single-storage", "single"])
def test_the_oracle_has_the_same_bits_on_every_thread_count(name, precision):
    size, div = (FINGERPRINT_CANTILEVER if name == "cantilever" else FINGERPRINT_PLATE)[:2]
    w = loaded(size, div)
    used, iterations, fp = solve(w, precision, 1)
    assert used == 1
    for threads in (2, 7):
        used, it, f = solve(w, precision, threads)
        assert used == threads, "the hierarchy must be large enough to use the workers"
        assert (it, f) == (iterations, fp), f"{threads} threads"
    if precision == 0:
        # the Double path is the reference's own fingerprint
        expected = FINGERPRINT_CANTILEVER if name == "cantilever" else FINGERPRINT_PLATE
        assert (iterations, fp) == (expected[2], expected[3])


def test_an_unbounded_thread_count_starts_the_ceiling_and_keeps_the_bits():
    w = loaded(*FINGERPRINT_CANTILEVER[:2])
    _, it1, fp1 = solve(w, 1, 1)
    used, it, fp = solve(w, 1, 10 ** 6)
    assert used == 64 and (it, fp) == (it1, fp1)


DIGEST = r'''
import hashlib, sys
sys.path.insert(0, sys.argv[1]); sys.path.insert(0, sys.argv[1] + "/tests")
import numpy as np, refcpu, solidfem_amd as sf
from workloads import block_config, plate_config
h = hashlib.sha256()
for w in (plate_config(divisions=(16, 16, 2)), block_config(2, divisions=(8, 6, 6))):
    ref = refcpu.Problem(w)
    rs, cols, vals = ref.csr()
    csr = sf.CsrMatrix.new(None, rs, cols, vals)
    setup = sf.MultigridSetup.new(None, csr, ref.rigid_modes(), ref.free_positions(), sf.MultigridOptions(coarse_dofs=100))
    while not setup.step(2 ** 62):
        pass
    for level in range(setup.level_count()):
        m = setup.level_matrix(level)
        for k in ("starts", "columns", "values"):
            h.update(np.ascontiguousarray(m[k]).tobytes())
        h.update(setup.level_inverse_diagonal(level).tobytes())
        h.update(np.array(setup.smoother_spectrum(level)).tobytes())
        if level + 1 < setup.level_count():
            p = setup.level_matrix(level, prolongator=True)
            for k in ("starts", "columns", "values"):
                h.update(np.ascontiguousarray(p[k]).tobytes())
            h.update(setup.aggregates(level).tobytes())
    f = setup.coarse_factor()
    if f is not None:
        h.update(f.tobytes())
print(h.hexdigest())
'''


def test_the_product_host_setup_has_the_same_bits_on_every_thread_count():
    digests = {}
    for threads in (1, 3, 8):
        env = dict(os.environ, SOLIDFEM_HOST_THREADS=str(threads), OMP_NUM_THREADS=str(threads))
        out = subprocess.run([sys.executable, "-c", DIGEST, str(ROOT)], env=env, capture_output=True, text=True, timeout=600)
        assert out.returncode == 0, out.stderr[-2000:]
        digests[threads] = out.stdout.strip().splitlines()[-1]
    assert len(set(digests.values())) == 1, digests
