/* The oracle's self-check against the reference's golden fingerprints
 * (crates/solidfem/tests/validation/multigrid_precision.rs:24-65, 134-250). Test
 * infrastructure: prints one line per case and exits non-zero on a mismatch. */
#include <stdio.h>
#include <stdlib.h>
#include "refcpu.h"

static void loaded(ref_workload *w, double sx, double sy, double sz, int dx, int dy, int dz) {
    ref_default_workload(w);
    w->size[0] = sx; w->size[1] = sy; w->size[2] = sz;
    w->divisions[0] = dx; w->divisions[1] = dy; w->divisions[2] = dz;
    w->clamp_axis = 0; w->clamp_value = 0.0;
    w->load_axis = 0; w->load_value = sx;
    w->traction[0] = 0.0; w->traction[1] = 3.0; w->traction[2] = -7.0;
    w->youngs = 200000.0; w->poisson = 0.3;
}

static int run(const char *name, double sx, double sy, double sz, int dx, int dy, int dz, int precision,
               unsigned long long want_it, unsigned long long want_hash) {
    ref_workload w; ref_mg_options o; ref_result r;
    loaded(&w, sx, sy, sz, dx, dy, dz);
    ref_default_mg_options(&o);
    o.coarse_dofs = 100;
    o.precision = precision;
    ref_solve(&w, &o, 1e-9, NULL, &r);
    int ok = r.status == 0 && r.iterations == want_it && (want_hash == 0 || r.fingerprint == want_hash);
    printf("%-22s precision=%d status=%d iterations=%llu hash=%llu floor=%d rel=%.3e levels=%lld oc=%.3f %s\n", name,
           precision, r.status, (unsigned long long)r.iterations, (unsigned long long)r.fingerprint, r.at_rounding_floor,
           r.relative_residual, (long long)r.levels, r.operator_complexity, ok ? "OK" : "MISMATCH");
    return ok ? 0 : 1;
}

int main(void) {
    int bad = 0;
    bad += run("cantilever", 10, 1, 1, 12, 3, 3, 0, 61, 4639424340864456019ull);
    bad += run("plate", 20, 20, 1, 8, 8, 2, 0, 85, 5373319653865270473ull);
    bad += run("cantilever", 10, 1, 1, 12, 3, 3, 1, 61, 0);
    bad += run("plate", 20, 20, 1, 8, 8, 2, 1, 85, 0);
    bad += run("cantilever", 10, 1, 1, 12, 3, 3, 2, 65, 0);
    bad += run("plate", 20, 20, 1, 8, 8, 2, 2, 90, 0);
    return bad ? 1 : 0;
}
