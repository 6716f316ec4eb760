// The element functions the reference exports for callers (crates/solidfem-elastic/src/lib.rs:56-71):
// the Tet10 shape functions, the integration rules, point_geometry, strain_at, stress_at and
// von_mises, on the host for one point at a time. The device kernels evaluate the same formulas in
// the same order for every element; these are the same operations for a caller's own element work
// (a patch test, a custom quantity at a point).
#include <cmath>

#include "sf_host.hpp"

namespace sf {
namespace {

// crates/solidfem-elastic/src/shape.rs:9-20, :22-31
const int EDGES[6][2] = {{0, 1}, {1, 2}, {2, 0}, {0, 3}, {1, 3}, {2, 3}};
const double DL[4][3] = {{-1.0, -1.0, -1.0}, {1.0, 0.0, 0.0}, {0.0, 1.0, 0.0}, {0.0, 0.0, 1.0}};

void barycentric(const double n[3], double l[4]) {
    l[0] = 1.0 - n[0] - n[1] - n[2];
    l[1] = n[0];
    l[2] = n[1];
    l[3] = n[2];
}

// quadrature.rs:44-77
constexpr double A4 = 0.1381966011250105, B4 = 0.5854101966249685;
constexpr double A14 = 0.3108859192633006, B14 = 1.0 - 3.0 * A14, W14A = 0.11268792571801585;
constexpr double C14 = 0.09273525031089123, D14 = 1.0 - 3.0 * C14, W14B = 0.07349304311636195;
constexpr double E14 = 0.04550370412564965, F14 = 0.5 - E14, W14C = 0.042546020777081466;
const double RULE4[4][4] = {{A4, A4, A4, 0.25}, {B4, A4, A4, 0.25}, {A4, B4, A4, 0.25}, {A4, A4, B4, 0.25}};
const double RULE14[14][4] = {{A14, A14, A14, W14A}, {B14, A14, A14, W14A}, {A14, B14, A14, W14A}, {A14, A14, B14, W14A},
                              {C14, C14, C14, W14B}, {D14, C14, C14, W14B}, {C14, D14, C14, W14B}, {C14, C14, D14, W14B},
                              {E14, E14, F14, W14C}, {E14, F14, E14, W14C}, {F14, E14, E14, W14C}, {E14, F14, F14, W14C},
                              {F14, E14, F14, W14C}, {F14, F14, E14, W14C}};

}  // namespace
}  // namespace sf

using namespace sf;

extern "C" {

void sf_shape_values(const double natural[3], double out[10]) {
    // shape.rs:35-45
    double l[4];
    barycentric(natural, l);
    for (int i = 0; i < 4; ++i) out[i] = l[i] * (2.0 * l[i] - 1.0);
    for (int s = 0; s < 6; ++s) out[4 + s] = 4.0 * l[EDGES[s][0]] * l[EDGES[s][1]];
}

void sf_shape_derivatives(const double natural[3], double out[30]) {
    // shape.rs:47-67: out[3 a + axis] = d N_a / d natural[axis]
    double l[4];
    barycentric(natural, l);
    for (int axis = 0; axis < 3; ++axis) {
        for (int i = 0; i < 4; ++i) out[3 * i + axis] = (4.0 * l[i] - 1.0) * DL[i][axis];
        for (int s = 0; s < 6; ++s) {
            const int a = EDGES[s][0], b = EDGES[s][1];
            out[3 * (4 + s) + axis] = 4.0 * (DL[a][axis] * l[b] + l[a] * DL[b][axis]);
        }
    }
}

uint64_t sf_quadrature_points(int32_t rule, double *points) {
    // Quadrature::points (quadrature.rs:44-77): per point r, s, t and the weight (weights sum to 1)
    const double(*table)[4] = rule == 14 ? RULE14 : rule == 4 ? RULE4 : nullptr;
    const uint64_t n = rule == 14 ? 14 : rule == 4 ? 4 : 0;
    if (points)
        for (uint64_t q = 0; q < n; ++q)
            for (int k = 0; k < 4; ++k) points[4 * q + k] = table[q][k];
    return n;
}

sf_status sf_point_geometry(const double nodes[30], const double natural[3], sf_geometry *out) {
    // point_geometry (element.rs:25-80): the Jacobian and the position summed over the nodes in
    // order, the cofactors, the determinant (refused when not positive, or NaN), the gradients
    if (!nodes || !natural || !out) return SF_ERR_INVALID_ARGUMENT;
    double d[30], v[10];
    sf_shape_derivatives(natural, d);
    sf_shape_values(natural, v);
    double j[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, position[3] = {0, 0, 0};
    for (int a = 0; a < 10; ++a)
        for (int i = 0; i < 3; ++i) {
            for (int k = 0; k < 3; ++k) j[i][k] += d[3 * a + i] * nodes[3 * a + k];
            position[i] += v[a] * nodes[3 * a + i];
        }
    const double cof[3][3] = {
        {j[1][1] * j[2][2] - j[1][2] * j[2][1], j[0][2] * j[2][1] - j[0][1] * j[2][2], j[0][1] * j[1][2] - j[0][2] * j[1][1]},
        {j[1][2] * j[2][0] - j[1][0] * j[2][2], j[0][0] * j[2][2] - j[0][2] * j[2][0], j[0][2] * j[1][0] - j[0][0] * j[1][2]},
        {j[1][0] * j[2][1] - j[1][1] * j[2][0], j[0][1] * j[2][0] - j[0][0] * j[2][1], j[0][0] * j[1][1] - j[0][1] * j[1][0]}};
    const double det = j[0][0] * cof[0][0] + j[0][1] * cof[1][0] + j[0][2] * cof[2][0];
    out->det_j = det;
    for (int i = 0; i < 3; ++i) out->position[i] = position[i];
    if (std::isnan(det) || det <= 0.0) return SF_ERR_MODEL_INVERTED_JACOBIAN;
    for (int a = 0; a < 10; ++a)
        for (int k = 0; k < 3; ++k)
            out->gradients[3 * a + k] = (cof[k][0] * d[3 * a] + cof[k][1] * d[3 * a + 1] + cof[k][2] * d[3 * a + 2]) / det;
    return SF_OK;
}

void sf_strain_at(const sf_geometry *g, const double u[30], double strain[6]) {
    // strain_at (element.rs:135-171): h[i][j] = du_i/dx_j over the nodes in order, then
    // [xx, yy, zz, 2xy, 2yz, 2zx]
    double h[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int a = 0; a < 10; ++a)
        for (int i = 0; i < 3; ++i)
            for (int k = 0; k < 3; ++k) h[i][k] += u[3 * a + i] * g->gradients[3 * a + k];
    strain[0] = h[0][0];
    strain[1] = h[1][1];
    strain[2] = h[2][2];
    strain[3] = h[0][1] + h[1][0];
    strain[4] = h[1][2] + h[2][1];
    strain[5] = h[2][0] + h[0][2];
}

void sf_stress_at(double youngs, double poisson, const double strain[6], double stress[6]) {
    // stress_at (element.rs:174-186) with Material::lame (material.rs:24-31)
    const double lambda = youngs * poisson / ((1.0 + poisson) * (1.0 - 2.0 * poisson)), mu = youngs / (2.0 * (1.0 + poisson));
    const double trace = strain[0] + strain[1] + strain[2];
    for (int i = 0; i < 3; ++i) stress[i] = lambda * trace + 2.0 * mu * strain[i];
    for (int i = 3; i < 6; ++i) stress[i] = mu * strain[i];
}

double sf_von_mises(const double stress[6]) {
    // von_mises (element.rs:189-192)
    const double a = stress[0] - stress[1], b = stress[1] - stress[2], c = stress[2] - stress[0];
    return std::sqrt(0.5 * (a * a + b * b + c * c) + 3.0 * (stress[3] * stress[3] + stress[4] * stress[4] + stress[5] * stress[5]));
}

}  // extern "C"
