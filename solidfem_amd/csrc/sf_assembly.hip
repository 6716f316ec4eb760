// Stiffness assembly on the device.
//
// element_kernel: one wavefront per Tet10 element. Per integration point the wavefront forms
// the Jacobian, cofactors and determinant, the physical shape gradients and the material
// (crates/solidfem-elastic/src/element.rs:25-94), then lane t < 55 accumulates the 3x3 block of
// node pair (a, b), a <= b, of the closed-form B^T D B (element.rs:104-130) with the reference's
// operations in the reference's order. Blocks with a > b are the mirrored (a, b) blocks, which
// the reference writes with the same bits, so only the 55 upper blocks are stored.
//
// gather_kernel: one thread per stored 3x3 block of the free stiffness, adding its element
// contributions in ascending element order from zero, the order in which
// crates/solidfem-elastic/src/assembly.rs:87-130 adds them into CsrMatrix::add's running sum. A
// chunked assembly continues each running sum, so chunks change no bit.
#include <algorithm>
#include <cmath>
#include <functional>
#include <limits>

#include "sf_device.hpp"

namespace sf {
namespace {

using namespace dev;

struct QPoint {
    double r, s, t, w;
};

// crates/solidfem-elastic/src/quadrature.rs:44-77
constexpr double A4 = 0.1381966011250105, B4 = 0.5854101966249685;
constexpr double A14 = 0.3108859192633006, B14 = 1.0 - 3.0 * A14, W14A = 0.11268792571801585;
constexpr double C14 = 0.09273525031089123, D14 = 1.0 - 3.0 * C14, W14B = 0.07349304311636195;
constexpr double E14 = 0.04550370412564965, F14 = 0.5 - E14, W14C = 0.042546020777081466;
__constant__ QPoint RULE4[4] = {{A4, A4, A4, 0.25}, {B4, A4, A4, 0.25}, {A4, B4, A4, 0.25}, {A4, A4, B4, 0.25}};
__constant__ QPoint RULE14[14] = {{A14, A14, A14, W14A}, {B14, A14, A14, W14A}, {A14, B14, A14, W14A}, {A14, A14, B14, W14A},
                                  {C14, C14, C14, W14B}, {D14, C14, C14, W14B}, {C14, D14, C14, W14B}, {C14, C14, D14, W14B},
                                  {E14, E14, F14, W14C}, {E14, F14, E14, W14C}, {F14, E14, E14, W14C}, {E14, F14, F14, W14C},
                                  {F14, E14, F14, W14C}, {F14, F14, E14, W14C}};
__constant__ int8_t EDGE_A[6] = {0, 1, 2, 0, 1, 2};
__constant__ int8_t EDGE_B[6] = {1, 2, 0, 3, 3, 3};
__constant__ double DL[4][3] = {{-1.0, -1.0, -1.0}, {1.0, 0.0, 0.0}, {0.0, 1.0, 0.0}, {0.0, 0.0, 1.0}};

struct FieldArgs {
    int kind;
    double e, nu, core_e, core_nu, depth;
    double lo[3], hi[3];
    int tr0, tr1;
    const double *table;  // SF_MATERIAL_POINTS: (E, nu) per (element, point) of a table_nq-point rule
    int table_nq;
};

// d(shape a)/d(natural axis) at the point (crates/solidfem-elastic/src/shape.rs:46-67)
__device__ double shape_derivative(int a, int axis, const double l[4]) {
    if (a < 4) return (4.0 * l[a] - 1.0) * DL[a][axis];
    const int ea = EDGE_A[a - 4], eb = EDGE_B[a - 4];
    return 4.0 * (DL[ea][axis] * l[eb] + l[ea] * DL[eb][axis]);
}
__device__ double shape_value(int a, const double l[4]) {
    if (a < 4) return l[a] * (2.0 * l[a] - 1.0);
    return 4.0 * l[EDGE_A[a - 4]] * l[EDGE_B[a - 4]];
}

constexpr int EPB = 4;  // elements (wavefronts) per block

template <int NQ>
__global__ __launch_bounds__(256) void element_kernel(size_t first, size_t count, const double *__restrict__ nodes,
                                                      const uint32_t *__restrict__ tets, FieldArgs field,
                                                      double *__restrict__ ke, int32_t *__restrict__ fault) {
    __shared__ double xe[EPB][10][3];
    __shared__ double jac[EPB][9];
    __shared__ double pos[EPB][3];
    __shared__ double cof[EPB][9];
    __shared__ double grad[EPB][10][3];
    __shared__ double mat[EPB][3];  // lambda, mu, measure
    __shared__ int bad[EPB];
    const int w = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    const size_t local = (size_t)blockIdx.x * EPB + w;
    const bool live = local < count;
    const size_t e = first + local;
    if (live && lane < 30) xe[w][lane / 3][lane % 3] = nodes[(size_t)tets[e * 10 + lane / 3] * 3 + lane % 3];
    if (lane == 0) bad[w] = 0;
    // lane -> node pair (a, b), a <= b
    int pa = 0, pb = 0;
    {
        int t = lane < 55 ? lane : 0, a = 0;
        while (t >= 10 - a) {
            t -= 10 - a;
            ++a;
        }
        pa = a;
        pb = a + t;
    }
    double acc[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) acc[i] = 0.0;
    __syncthreads();
    const QPoint *rule = NQ == 14 ? RULE14 : RULE4;
    for (int q = 0; q < NQ; ++q) {
        const QPoint qp = rule[q];
        double l[4];
        l[0] = 1.0 - qp.r - qp.s - qp.t;
        l[1] = qp.r;
        l[2] = qp.s;
        l[3] = qp.t;
        // J[i][k] = sum_a dN_a/dxi_i x_a,k ; position[i] = sum_a N_a x_a,i (element.rs:32-41)
        if (live && lane < 9) {
            const int i = lane / 3, k = lane % 3;
            double s = 0.0;
            for (int a = 0; a < 10; ++a) s += shape_derivative(a, i, l) * xe[w][a][k];
            jac[w][lane] = s;
        } else if (live && lane < 12) {
            const int i = lane - 9;
            double s = 0.0;
            for (int a = 0; a < 10; ++a) s += shape_value(a, l) * xe[w][a][i];
            pos[w][i] = s;
        }
        __syncthreads();
        if (live && lane == 0) {
            const double *j = jac[w];
            // element.rs:42-59
            cof[w][0] = j[4] * j[8] - j[5] * j[7];
            cof[w][1] = j[2] * j[7] - j[1] * j[8];
            cof[w][2] = j[1] * j[5] - j[2] * j[4];
            cof[w][3] = j[5] * j[6] - j[3] * j[8];
            cof[w][4] = j[0] * j[8] - j[2] * j[6];
            cof[w][5] = j[2] * j[3] - j[0] * j[5];
            cof[w][6] = j[3] * j[7] - j[4] * j[6];
            cof[w][7] = j[1] * j[6] - j[0] * j[7];
            cof[w][8] = j[0] * j[4] - j[1] * j[3];
            const double det = j[0] * cof[w][0] + j[1] * cof[w][3] + j[2] * cof[w][6];
            if (isnan(det) || det <= 0.0) {
                bad[w] = 1;
                atomicMin(&fault[0], (int32_t)(e * 64 + (size_t)q * 2));  // element, point, inverted
            }
            // material at the point (material.rs:24-50)
            double E = field.e, nu = field.nu;
            if (field.kind == SF_MATERIAL_POINTS) {
                E = field.table[(e * NQ + q) * 2];
                nu = field.table[(e * NQ + q) * 2 + 1];
            } else if (field.kind == SF_MATERIAL_SHELL_CORE) {
                const int axes[2] = {field.tr0, field.tr1};
                double depth = 1e300;
                for (int s2 = 0; s2 < 2; ++s2) {
                    const int ax = axes[s2];
                    depth = fmin(depth, fmin(pos[w][ax] - field.lo[ax], field.hi[ax] - pos[w][ax]));
                }
                if (!(depth < field.depth)) {
                    E = field.core_e;
                    nu = field.core_nu;
                }
            }
            const bool valid = isfinite(E) && E > 0.0 && nu > -1.0 && nu < 0.5;
            if (!valid) {
                bad[w] = 2;
                atomicMin(&fault[0], (int32_t)(e * 64 + (size_t)q * 2 + 1));  // element, point, material
            }
            mat[w][0] = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
            mat[w][1] = E / (2.0 * (1.0 + nu));
            mat[w][2] = qp.w * det / 6.0;
            jac[w][0] = det;  // reuse: the determinant for the gradients
        }
        __syncthreads();
        // gradients = cof . dN / det (element.rs:66-74)
        if (live && lane < 30) {
            const int a = lane / 3, k = lane % 3;
            const double d0 = shape_derivative(a, 0, l), d1 = shape_derivative(a, 1, l), d2 = shape_derivative(a, 2, l);
            const double sum = cof[w][k * 3 + 0] * d0 + cof[w][k * 3 + 1] * d1 + cof[w][k * 3 + 2] * d2;
            grad[w][a][k] = sum / jac[w][0];
        }
        __syncthreads();
        // element.rs:109-130
        if (live && lane < 55) {
            const double lambda = mat[w][0], mu = mat[w][1], measure = mat[w][2];
            const double ga[3] = {grad[w][pa][0], grad[w][pa][1], grad[w][pa][2]};
            const double gb[3] = {grad[w][pb][0], grad[w][pb][1], grad[w][pb][2]};
            const double dot = ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2];
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    double value = lambda * (ga[i] * gb[k]) + mu * (ga[k] * gb[i]);
                    if (i == k) value += mu * dot;
                    value *= measure;
                    acc[i * 3 + k] += value;
                }
        }
        __syncthreads();
    }
    if (live && lane < 55) {
        double *out = ke + (local * 55 + (size_t)lane) * 9;
#pragma unroll
        for (int i = 0; i < 9; ++i) out[i] = acc[i];
    }
}

__device__ __forceinline__ int pair_index(int a, int b) { return a * 10 - (a * (a - 1)) / 2 + (b - a); }

// Adds the contributions of elements [e0, e1) to every block, in element order.
__global__ void gather_kernel(size_t blocks, const int64_t *__restrict__ cstart, const uint32_t *__restrict__ contrib,
                              uint32_t e0, uint32_t e1, const double *__restrict__ ke, double *__restrict__ values) {
    const size_t at = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (at >= blocks) return;
    double v[9];
    double *out = values + at * 9;
#pragma unroll
    for (int i = 0; i < 9; ++i) v[i] = out[i];
    for (int64_t c = cstart[at]; c < cstart[at + 1]; ++c) {
        const uint32_t code = contrib[c];
        const uint32_t e = code >> 7;
        if (e < e0) continue;
        if (e >= e1) break;
        const int ab = (int)(code & 127u), a = ab / 10, b = ab % 10;
        if (a <= b) {
            const double *src = ke + ((size_t)(e - e0) * 55 + pair_index(a, b)) * 9;
#pragma unroll
            for (int i = 0; i < 9; ++i) v[i] += src[i];
        } else {
            const double *src = ke + ((size_t)(e - e0) * 55 + pair_index(b, a)) * 9;
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int k = 0; k < 3; ++k) v[i * 3 + k] += src[k * 3 + i];
        }
    }
#pragma unroll
    for (int i = 0; i < 9; ++i) out[i] = v[i];
}

// ---- stress recovery (crates/solidfem-elastic/src/results.rs:307-576) --------------------------

// crates/solidfem-elastic/src/shape.rs:9-20
__constant__ double NODE_NATURAL[10][3] = {{0.0, 0.0, 0.0}, {1.0, 0.0, 0.0}, {0.0, 1.0, 0.0}, {0.0, 0.0, 1.0},
                                           {0.5, 0.0, 0.0}, {0.5, 0.5, 0.0}, {0.0, 0.5, 0.0}, {0.0, 0.0, 0.5},
                                           {0.5, 0.0, 0.5}, {0.0, 0.5, 0.5}};

// point_geometry (element.rs:25-80) for one thread: the element's gradients and position at a
// natural point, the same operations in the same order as element_kernel. False when inverted.
__device__ bool thread_geometry(const double xe[10][3], const double nat[3], double grad[10][3], double pos[3],
                                double *det_out = nullptr) {
    double l[4] = {1.0 - nat[0] - nat[1] - nat[2], nat[0], nat[1], nat[2]};
    double j[9];
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k) {
            double s = 0.0;
            for (int a = 0; a < 10; ++a) s += shape_derivative(a, i, l) * xe[a][k];
            j[i * 3 + k] = s;
        }
    for (int i = 0; i < 3; ++i) {
        double s = 0.0;
        for (int a = 0; a < 10; ++a) s += shape_value(a, l) * xe[a][i];
        pos[i] = s;
    }
    double cof[9];
    cof[0] = j[4] * j[8] - j[5] * j[7];
    cof[1] = j[2] * j[7] - j[1] * j[8];
    cof[2] = j[1] * j[5] - j[2] * j[4];
    cof[3] = j[5] * j[6] - j[3] * j[8];
    cof[4] = j[0] * j[8] - j[2] * j[6];
    cof[5] = j[2] * j[3] - j[0] * j[5];
    cof[6] = j[3] * j[7] - j[4] * j[6];
    cof[7] = j[1] * j[6] - j[0] * j[7];
    cof[8] = j[0] * j[4] - j[1] * j[3];
    const double det = j[0] * cof[0] + j[1] * cof[3] + j[2] * cof[6];
    if (det_out) *det_out = det;
    if (isnan(det) || det <= 0.0) return false;
    for (int a = 0; a < 10; ++a)
        for (int k = 0; k < 3; ++k) {
            const double sum = cof[k * 3 + 0] * shape_derivative(a, 0, l) + cof[k * 3 + 1] * shape_derivative(a, 1, l) +
                               cof[k * 3 + 2] * shape_derivative(a, 2, l);
            grad[a][k] = sum / det;
        }
    return true;
}

// MaterialField::at (material.rs:35-50), as element_kernel evaluates it, at point q of element e
// (a caller's field is read from its table of the same rule)
__device__ void material_at(const FieldArgs &field, const double pos[3], size_t e, int q, double &E, double &nu) {
    E = field.e;
    nu = field.nu;
    if (field.kind == SF_MATERIAL_POINTS) {
        E = field.table[(e * (size_t)field.table_nq + (size_t)q) * 2];
        nu = field.table[(e * (size_t)field.table_nq + (size_t)q) * 2 + 1];
    } else if (field.kind == SF_MATERIAL_SHELL_CORE) {
        const int axes[2] = {field.tr0, field.tr1};
        double depth = 1e300;
        for (int s2 = 0; s2 < 2; ++s2) {
            const int ax = axes[s2];
            depth = fmin(depth, fmin(pos[ax] - field.lo[ax], field.hi[ax] - pos[ax]));
        }
        if (!(depth < field.depth)) {
            E = field.core_e;
            nu = field.core_nu;
        }
    }
}

// strain_at, stress_at and von_mises (element.rs:135-192) from the gradients and the element's
// nodal displacements
__device__ double point_stress(const double grad[10][3], const double ue[10][3], double E, double nu, double s[6]) {
    double h[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int a = 0; a < 10; ++a)
        for (int i = 0; i < 3; ++i)
            for (int jj = 0; jj < 3; ++jj) h[i][jj] += ue[a][i] * grad[a][jj];
    const double e0 = h[0][0], e1 = h[1][1], e2 = h[2][2];
    const double e3 = h[0][1] + h[1][0], e4 = h[1][2] + h[2][1], e5 = h[2][0] + h[0][2];
    const double lambda = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu)), mu = E / (2.0 * (1.0 + nu));
    const double trace = e0 + e1 + e2;
    s[0] = lambda * trace + 2.0 * mu * e0;
    s[1] = lambda * trace + 2.0 * mu * e1;
    s[2] = lambda * trace + 2.0 * mu * e2;
    s[3] = mu * e3;
    s[4] = mu * e4;
    s[5] = mu * e5;
    const double a = s[0] - s[1], b = s[1] - s[2], c = s[2] - s[0];
    return sqrt(0.5 * (a * a + b * b + c * c) + 3.0 * (s[3] * s[3] + s[4] * s[4] + s[5] * s[5]));
}

__device__ void load_element(const double *__restrict__ nodes, const uint32_t *__restrict__ tets, const double *__restrict__ u,
                             size_t e, double xe[10][3], double ue[10][3]) {
    for (int a = 0; a < 10; ++a) {
        const size_t n = tets[e * 10 + a];
        for (int i = 0; i < 3; ++i) {
            xe[a][i] = nodes[n * 3 + i];
            ue[a][i] = u[n * 3 + i];
        }
    }
}

// Stress at every integration point (Recovery::advance's point loop): one thread per (element,
// point). `material` keeps each point's (E, nu) for the node values and the patch test of
// uniformity.
template <int NQ>
__global__ __launch_bounds__(256) void recovery_points_kernel(size_t first, size_t count, const double *__restrict__ nodes,
                                                              const uint32_t *__restrict__ tets, const double *__restrict__ u,
                                                              FieldArgs field, double *__restrict__ out, double *__restrict__ material,
                                                              int32_t *__restrict__ fault) {
    const size_t local = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (local >= count * NQ) return;
    const size_t t = first * NQ + local;
    const size_t e = t / NQ;
    const int q = (int)(t - e * NQ);
    double xe[10][3], ue[10][3], grad[10][3], pos[3];
    load_element(nodes, tets, u, e, xe, ue);
    const QPoint qp = (NQ == 14 ? RULE14 : RULE4)[q];
    const double nat[3] = {qp.r, qp.s, qp.t};
    double E, nu;
    if (!thread_geometry(xe, nat, grad, pos)) {
        atomicMin(&fault[0], (int32_t)(e * 64 + (size_t)q * 2));  // element, point, inverted
        return;
    }
    material_at(field, pos, e, q, E, nu);
    if (!(isfinite(E) && E > 0.0 && nu > -1.0 && nu < 0.5)) {
        atomicMin(&fault[0], (int32_t)(e * 64 + (size_t)q * 2 + 1));  // element, point, material
        return;
    }
    double s[6];
    const double vm = point_stress(grad, ue, E, nu, s);
    double *o = out + t * 10;
    for (int i = 0; i < 3; ++i) o[i] = pos[i];
    for (int c = 0; c < 6; ++c) o[3 + c] = s[c];
    o[9] = vm;
    material[t * 2] = E;
    material[t * 2 + 1] = nu;
}

// The element's stress at each of its nodes with the material of its nearest integration point
// (Recovery::advance's node loop): one thread per (element, node).
template <int NQ>
__global__ __launch_bounds__(256) void recovery_nodes_kernel(size_t first, size_t count, const double *__restrict__ nodes,
                                                             const uint32_t *__restrict__ tets, const double *__restrict__ u,
                                                             const double *__restrict__ material, const int8_t *__restrict__ nearest,
                                                             double *__restrict__ out, int32_t *__restrict__ fault) {
    const size_t local = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (local >= count * 10) return;
    const size_t t = first * 10 + local;
    const size_t e = t / 10;
    const int n = (int)(t - e * 10);
    double xe[10][3], ue[10][3], grad[10][3], pos[3];
    load_element(nodes, tets, u, e, xe, ue);
    const double nat[3] = {NODE_NATURAL[n][0], NODE_NATURAL[n][1], NODE_NATURAL[n][2]};
    if (!thread_geometry(xe, nat, grad, pos)) {
        atomicMin(&fault[0], (int32_t)(e * 64 + 28 + (size_t)n * 2));  // element, node, inverted
        return;
    }
    const size_t mp = e * NQ + (size_t)nearest[n];
    double s[6];
    const double vm = point_stress(grad, ue, material[mp * 2], material[mp * 2 + 1], s);
    double *o = out + t * 7;
    for (int c = 0; c < 6; ++c) o[c] = s[c];
    o[6] = vm;
}

// The first largest von Mises among a node's incident element values, in element order (the
// strict comparison of Recovery::advance, from the lowest value).
__global__ void recovery_worst_kernel(size_t nodes, const uint32_t *__restrict__ starts, const uint32_t *__restrict__ at,
                                      const double *__restrict__ node_values, double *__restrict__ worst) {
    const size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nodes) return;
    double best = -1.7976931348623157e308, s[6] = {0, 0, 0, 0, 0, 0};
    for (uint32_t k = starts[n]; k < starts[n + 1]; ++k) {
        const double *v = node_values + (size_t)at[k] * 7;
        if (v[6] > best) {
            best = v[6];
            for (int c = 0; c < 6; ++c) s[c] = v[c];
        }
    }
    for (int c = 0; c < 6; ++c) worst[n * 6 + c] = s[c];
}

// PatchFit::fit (patch.rs:35-119) at each vertex over its incident elements' points in element
// order; a patch of more than one material is left unfitted. fit: flag (1 fitted, 0 flat, -1 more
// than one material), centre 3, scale, coefficients 4 x 6 (29 values).
//
// One wavefront of 64 threads per vertex. The material test and the scale (a maximum, exact in any
// order) are reduced over the wavefront. The normal equations are then formed in chunks of 64
// samples: each lane forms one sample's offsets into LDS, and lane e of the first 40 owns entry e
// (16 of the matrix, 24 of the right-hand side), adding that entry's products over the chunk in
// sample order -- each entry is the same running sum, from zero in sample order, as the serial
// loop forms. Lane 0 solves the 4 x 4 system as the serial loop does.
constexpr int FIT_LANES = 64;

template <int NQ>
__global__ __launch_bounds__(FIT_LANES) void recovery_fit_kernel(size_t first, size_t count, const double *__restrict__ nodes,
                                                                 const uint32_t *__restrict__ starts, const uint32_t *__restrict__ elems,
                                                                 const double *__restrict__ points,
                                                                 const double *__restrict__ material, double *__restrict__ fit) {
    if ((size_t)blockIdx.x >= count) return;
    const size_t v = first + blockIdx.x;
    const int lane = threadIdx.x;
    double *f = fit + v * 29;
    const double c[3] = {nodes[v * 3], nodes[v * 3 + 1], nodes[v * 3 + 2]};
    const uint32_t k0 = starts[v], k1 = starts[v + 1];
    const uint32_t samples = (k1 - k0) * (uint32_t)NQ;
    if (samples == 0) {
        if (lane == 0) f[0] = 0.0;
        return;
    }
    auto point_of = [&](uint32_t sample) { return (size_t)elems[k0 + sample / NQ] * NQ + sample % NQ; };
    // the material test and the scale pass
    const size_t p0 = (size_t)elems[k0] * NQ;
    const long long e0 = __double_as_longlong(material[p0 * 2]), nu0 = __double_as_longlong(material[p0 * 2 + 1]);
    int mixed = 0;
    double scale = 0.0;
    for (uint32_t sample = lane; sample < samples; sample += FIT_LANES) {
        const size_t p = point_of(sample);
        if (__double_as_longlong(material[p * 2]) != e0 || __double_as_longlong(material[p * 2 + 1]) != nu0) mixed = 1;
        for (int a = 0; a < 3; ++a) scale = fmax(scale, fabs(points[p * 10 + a] - c[a]));
    }
    for (int off = FIT_LANES / 2; off > 0; off >>= 1) {
        mixed |= __shfl_xor(mixed, off, FIT_LANES);
        scale = fmax(scale, __shfl_xor(scale, off, FIT_LANES));
    }
    if (mixed || scale == 0.0) {
        if (lane == 0) f[0] = mixed ? -1.0 : 0.0;
        return;
    }
    // the normal equations, entry e of lane e: m[i][j] for e < 16, rhs[i][c] for 16 <= e < 40
    __shared__ double offsets[FIT_LANES][4], stresses[FIT_LANES][6];
    const int ei = lane < 16 ? lane / 4 : (lane - 16) / 6;
    const int ej = lane < 16 ? lane % 4 : (lane - 16) % 6;
    double entry = 0.0;
    for (uint32_t base = 0; base < samples; base += FIT_LANES) {
        const uint32_t sample = base + (uint32_t)lane;
        if (sample < samples) {
            const double *pt = points + point_of(sample) * 10;
            offsets[lane][0] = 1.0;
            for (int a = 0; a < 3; ++a) offsets[lane][1 + a] = (pt[a] - c[a]) / scale;
            for (int k = 0; k < 6; ++k) stresses[lane][k] = pt[3 + k];
        }
        __syncthreads();
        const uint32_t chunk = min((uint32_t)FIT_LANES, samples - base);
        if (lane < 40) {
            // eight samples' factors are read ahead of their additions, which stay in order
            const double *second = lane < 16 ? &offsets[0][ej] : &stresses[0][ej];
            const int stride = lane < 16 ? 4 : 6;
            uint32_t q = 0;
            for (; q + 8 <= chunk; q += 8) {
                double a[8], b[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    a[u] = offsets[q + u][ei];
                    b[u] = second[(q + u) * stride];
                }
#pragma unroll
                for (int u = 0; u < 8; ++u) entry += a[u] * b[u];
            }
            for (; q < chunk; ++q) entry += offsets[q][ei] * second[q * stride];
        }
        __syncthreads();
    }
    __shared__ double m[4][4], rhs[4][6];
    if (lane < 16) m[ei][ej] = entry;
    else if (lane < 40) rhs[ei][ej] = entry;
    __syncthreads();
    if (lane != 0) return;
    for (int column = 0; column < 4; ++column) {
        int pivot = column;
        for (int row = column + 1; row < 4; ++row)
            if (fabs(m[row][column]) > fabs(m[pivot][column])) pivot = row;
        if (fabs(m[pivot][column]) <= 1e-12) {
            f[0] = 0.0;
            return;
        }
        for (int k = 0; k < 4; ++k) {
            const double x = m[column][k];
            m[column][k] = m[pivot][k];
            m[pivot][k] = x;
        }
        for (int k = 0; k < 6; ++k) {
            const double x = rhs[column][k];
            rhs[column][k] = rhs[pivot][k];
            rhs[pivot][k] = x;
        }
        for (int row = column + 1; row < 4; ++row) {
            const double factor = m[row][column] / m[column][column];
            for (int k = column; k < 4; ++k) m[row][k] -= factor * m[column][k];
            for (int k = 0; k < 6; ++k) rhs[row][k] -= factor * rhs[column][k];
        }
    }
    for (int column = 3; column >= 0; --column) {
        double value[6];
        for (int k = 0; k < 6; ++k) value[k] = rhs[column][k];
        for (int j = column + 1; j < 4; ++j)
            for (int k = 0; k < 6; ++k) value[k] -= m[column][j] * rhs[j][k];
        for (int k = 0; k < 6; ++k) rhs[column][k] = value[k] / m[column][column];
    }
    f[0] = 1.0;
    for (int a = 0; a < 3; ++a) f[1 + a] = c[a];
    f[4] = scale;
    for (int i = 0; i < 4; ++i)
        for (int k = 0; k < 6; ++k) f[5 + i * 6 + k] = rhs[i][k];
}

__device__ void fit_at(const double *f, const double p[3], double value[6]) {
    const double t[4] = {1.0, (p[0] - f[1]) / f[4], (p[1] - f[2]) / f[4], (p[2] - f[3]) / f[4]};
    for (int k = 0; k < 6; ++k) value[k] = 0.0;
    for (int i = 0; i < 4; ++i)
        for (int k = 0; k < 6; ++k) value[k] += t[i] * f[5 + i * 6 + k];
}

// Recovery::into_field (results.rs:376-411): a vertex reads its fit, a midside the mean of its
// edge's two fits at its own position (midside_mean, results.rs:544-552); a node without fits
// reads its worst incident value. Then von Mises. out: 7 per node.
__global__ void recovery_nodal_kernel(size_t nodes, size_t vertices, const double *__restrict__ xyz, const uint32_t *__restrict__ ends,
                                      const double *__restrict__ fit, const double *__restrict__ worst, double *__restrict__ out) {
    const size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nodes) return;
    const double p[3] = {xyz[n * 3], xyz[n * 3 + 1], xyz[n * 3 + 2]};
    double s[6];
    if (n < vertices) {
        const double *f = fit + n * 29;
        if (f[0] > 0.0) fit_at(f, p, s);
        else
            for (int c = 0; c < 6; ++c) s[c] = worst[n * 6 + c];
    } else {
        const double *fa = fit + (size_t)ends[(n - vertices) * 2] * 29, *fb = fit + (size_t)ends[(n - vertices) * 2 + 1] * 29;
        if (fa[0] > 0.0 && fb[0] > 0.0) {
            double a[6], b[6];
            fit_at(fa, p, a);
            fit_at(fb, p, b);
            for (int c = 0; c < 6; ++c) s[c] = 0.5 * (a[c] + b[c]);
        } else {
            for (int c = 0; c < 6; ++c) s[c] = worst[n * 6 + c];
        }
    }
    const double a = s[0] - s[1], b = s[1] - s[2], c = s[2] - s[0];
    double *o = out + n * 7;
    for (int k = 0; k < 6; ++k) o[k] = s[k];
    o[6] = sqrt(0.5 * (a * a + b * b + c * c) + 3.0 * (s[3] * s[3] + s[4] * s[4] + s[5] * s[5]));
}

// ---- error estimate (crates/solidfem-elastic/src/error_estimate.rs) ----------------------------

// An element's first refusal is reported as element * 32 + what: 0-3 the corner whose patch has no
// fit, 4 + 2q an inverted point q, 5 + 2q an invalid material at point q -- the order of
// element_integrals (error_estimate.rs:174-205): the fits first, then the points in rule order,
// the geometry before the material. The smallest report is the first refusal in element order.
constexpr unsigned long long NO_ESTIMATE_FAULT = ~0ull;

// recovered_at_nodes (error_estimate.rs:207-231): one thread per (element, node); a corner takes
// its vertex's fit at the corner, a midside the mean of its edge's two fits at the midside
// (midside_mean, results.rs:544-552). recovered: 6 per (element, node) of the slice.
__global__ __launch_bounds__(256) void estimate_nodes_kernel(size_t first, size_t count, const double *__restrict__ nodes,
                                                             const uint32_t *__restrict__ tets, const double *__restrict__ fit,
                                                             double *__restrict__ recovered, unsigned long long *__restrict__ fault) {
    const size_t local = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (local >= count * 10) return;
    const size_t e = first + local / 10;
    const int a = (int)(local % 10);
    const uint32_t *t = tets + e * 10;
    const double p[3] = {nodes[(size_t)t[a] * 3], nodes[(size_t)t[a] * 3 + 1], nodes[(size_t)t[a] * 3 + 2]};
    double s[6];
    if (a < 4) {
        const double *f = fit + (size_t)t[a] * 29;
        if (!(f[0] > 0.0)) {
            atomicMin(fault, (unsigned long long)e * 32 + (unsigned long long)a);
            return;
        }
        fit_at(f, p, s);
    } else {
        const double *fa = fit + (size_t)t[EDGE_A[a - 4]] * 29, *fb = fit + (size_t)t[EDGE_B[a - 4]] * 29;
        if (!(fa[0] > 0.0) || !(fb[0] > 0.0)) return;  // the corner threads report it
        double x[6], y[6];
        fit_at(fa, p, x);
        fit_at(fb, p, y);
        for (int c = 0; c < 6; ++c) s[c] = 0.5 * (x[c] + y[c]);
    }
    double *o = recovered + local * 6;
    for (int c = 0; c < 6; ++c) o[c] = s[c];
}

// compliance_energy (error_estimate.rs:240-248): s : C^-1 : s, one fixed expression
__device__ double compliance_energy(double E, double nu, const double s[6]) {
    const double normal = s[0] * s[0] + s[1] * s[1] + s[2] * s[2];
    const double coupling = s[0] * s[1] + s[1] * s[2] + s[2] * s[0];
    const double shear = s[3] * s[3] + s[4] * s[4] + s[5] * s[5];
    return (normal - 2.0 * nu * coupling + 2.0 * (1.0 + nu) * shear) / E;
}

// element_integrals' terms at each point of the 14-point rule (error_estimate.rs:183-203): one
// thread per (element, point); terms: measure * the difference's energy, measure * the element
// stress's energy. An inverted point leaves its determinant in its first term for the report.
__global__ __launch_bounds__(256) void estimate_points_kernel(size_t first, size_t count, const double *__restrict__ nodes,
                                                              const uint32_t *__restrict__ tets, const double *__restrict__ u,
                                                              const double *__restrict__ recovered, FieldArgs field,
                                                              double *__restrict__ terms, unsigned long long *__restrict__ fault) {
    const size_t local = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (local >= count * 14) return;
    const size_t slot = local / 14;
    const int q = (int)(local - slot * 14);
    const size_t e = first + slot;
    double xe[10][3], ue[10][3], grad[10][3], pos[3], det = 0.0;
    load_element(nodes, tets, u, e, xe, ue);
    const QPoint qp = RULE14[q];
    const double nat[3] = {qp.r, qp.s, qp.t};
    if (!thread_geometry(xe, nat, grad, pos, &det)) {
        terms[local * 2] = det;
        atomicMin(fault, (unsigned long long)e * 32 + 4 + 2 * (unsigned long long)q);
        return;
    }
    double E, nu;
    material_at(field, pos, e, q, E, nu);
    if (!(isfinite(E) && E > 0.0 && nu > -1.0 && nu < 0.5)) {
        atomicMin(fault, (unsigned long long)e * 32 + 5 + 2 * (unsigned long long)q);
        return;
    }
    double stress[6];
    point_stress(grad, ue, E, nu, stress);
    const double l[4] = {1.0 - nat[0] - nat[1] - nat[2], nat[0], nat[1], nat[2]};
    const double *rec = recovered + slot * 60;
    double difference[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int a = 0; a < 10; ++a) {
        const double w = shape_value(a, l);
        for (int c = 0; c < 6; ++c) difference[c] += w * rec[a * 6 + c];
    }
    for (int c = 0; c < 6; ++c) difference[c] -= stress[c];
    const double measure = qp.w * det / 6.0;
    terms[local * 2] = measure * compliance_energy(E, nu, difference);
    terms[local * 2 + 1] = measure * compliance_energy(E, nu, stress);
}

// Each element's two integrals, its point terms added from zero in rule order: one thread per
// element of the slice.
__global__ void estimate_sums_kernel(size_t count, const double *__restrict__ terms, double *__restrict__ sums) {
    const size_t slot = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= count) return;
    double error = 0.0, energy = 0.0;
    for (int q = 0; q < 14; ++q) {
        error += terms[(slot * 14 + q) * 2];
        energy += terms[(slot * 14 + q) * 2 + 1];
    }
    sums[slot * 2] = error;
    sums[slot * 2 + 1] = energy;
}

// The kernels' view of a field; a caller's field (SF_MATERIAL_POINTS) reads `table`, the device copy
// of its values at the points of the `nq`-point rule.
FieldArgs field_args(const MaterialField &f, const double *table = nullptr, int nq = 4) {
    if (f.kind == SF_MATERIAL_POINTS && !table) throw Error(SF_ERR_INTERNAL, "a caller's material field has no device table");
    FieldArgs a{};
    a.table = table;
    a.table_nq = nq;
    a.kind = f.kind;
    a.e = f.e;
    a.nu = f.nu;
    a.core_e = f.core_e;
    a.core_nu = f.core_nu;
    a.depth = f.shell_depth;
    for (int i = 0; i < 3; ++i) {
        a.lo[i] = f.lo[i];
        a.hi[i] = f.hi[i];
    }
    a.tr0 = f.transverse[0];
    a.tr1 = f.transverse[1];
    return a;
}

// Each element's determinant-weighted rotation at its nodes (nodal_rotations, results.rs:262-299):
// one thread per (element, node); out: det_j and det_j * rotation (4 values).
__global__ __launch_bounds__(256) void rotation_nodes_kernel(size_t elements, const double *__restrict__ nodes,
                                                             const uint32_t *__restrict__ tets, const double *__restrict__ u,
                                                             double *__restrict__ out, int32_t *__restrict__ fault) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= elements * 10) return;
    const size_t e = t / 10;
    const int n = (int)(t - e * 10);
    double xe[10][3], ue[10][3], grad[10][3], pos[3], det = 0.0;
    load_element(nodes, tets, u, e, xe, ue);
    const double nat[3] = {NODE_NATURAL[n][0], NODE_NATURAL[n][1], NODE_NATURAL[n][2]};
    if (!thread_geometry(xe, nat, grad, pos, &det)) {
        atomicMin(&fault[0], (int32_t)e);
        return;
    }
    // rotation_at (element.rs:149-156) from displacement_gradient (element.rs:135-146)
    double h[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int a = 0; a < 10; ++a)
        for (int i = 0; i < 3; ++i)
            for (int jj = 0; jj < 3; ++jj) h[i][jj] += ue[a][i] * grad[a][jj];
    const double r[3] = {0.5 * (h[2][1] - h[1][2]), 0.5 * (h[0][2] - h[2][0]), 0.5 * (h[1][0] - h[0][1])};
    double *o = out + t * 4;
    o[0] = det;
    for (int axis = 0; axis < 3; ++axis) o[1 + axis] = det * r[axis];
}

// A node's weighted mean over its incident elements, summed in element order from zero
__global__ void rotation_gather_kernel(size_t nodes, const uint32_t *__restrict__ starts, const uint32_t *__restrict__ at,
                                       const double *__restrict__ values, double *__restrict__ out) {
    const size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nodes) return;
    double sum[3] = {0.0, 0.0, 0.0}, weight = 0.0;
    for (uint32_t k = starts[n]; k < starts[n + 1]; ++k) {
        const double *v = values + (size_t)at[k] * 4;
        for (int axis = 0; axis < 3; ++axis) sum[axis] += v[1 + axis];
        weight += v[0];
    }
    for (int axis = 0; axis < 3; ++axis) out[n * 3 + axis] = sum[axis] / weight;
}

constexpr size_t CHUNK_ELEMENTS = 1u << 20;
constexpr int32_t NO_FAULT = 0x7fffffff;

dim3 grid_of(size_t n) { return dim3((unsigned)((n + 255) / 256)); }

// The integration point nearest each node in natural coordinates, the first on a tie
// (results.rs:555-571).
std::vector<int8_t> nearest_points(int nq) {
    const double r4[4][3] = {{A4, A4, A4}, {B4, A4, A4}, {A4, B4, A4}, {A4, A4, B4}};
    const double r14[14][3] = {{A14, A14, A14}, {B14, A14, A14}, {A14, B14, A14}, {A14, A14, B14}, {C14, C14, C14},
                               {D14, C14, C14}, {C14, D14, C14}, {C14, C14, D14}, {E14, E14, F14}, {E14, F14, E14},
                               {F14, E14, E14}, {E14, F14, F14}, {F14, E14, F14}, {F14, F14, E14}};
    const double natural[10][3] = {{0.0, 0.0, 0.0}, {1.0, 0.0, 0.0}, {0.0, 1.0, 0.0}, {0.0, 0.0, 1.0}, {0.5, 0.0, 0.0},
                                   {0.5, 0.5, 0.0}, {0.0, 0.5, 0.0}, {0.0, 0.0, 0.5}, {0.5, 0.0, 0.5}, {0.0, 0.5, 0.5}};
    std::vector<int8_t> nearest(10);
    for (int n = 0; n < 10; ++n) {
        double best = 1.7976931348623157e308;
        int index = 0;
        for (int q = 0; q < nq; ++q) {
            const double *pt = nq == 14 ? r14[q] : r4[q];
            double d = 0.0;
            for (int a = 0; a < 3; ++a) d += (pt[a] - natural[n][a]) * (pt[a] - natural[n][a]);
            if (d < best) {
                best = d;
                index = q;
            }
        }
        nearest[n] = (int8_t)index;
    }
    return nearest;
}

// A node's incident (element, node) slots in element order, then a vertex's incident elements
// (build_patches, results.rs:414-436) and a midside's edge ends from the first element reaching
// it (into_field, results.rs:386-401).
struct Incidence {
    std::vector<uint32_t> node_starts, node_at, vert_starts, vert_at, ends;
};

Incidence incidence_of(const QuadMesh &mesh, bool with_vertices) {
    const size_t elements = mesh.tets.size(), nodes = mesh.nodes.size(), vertices = mesh.vertex_count;
    Incidence in;
    in.node_starts.assign(nodes + 1, 0);
    for (const auto &t : mesh.tets)
        for (int a = 0; a < 10; ++a) in.node_starts[t[a] + 1] += 1;
    for (size_t n = 0; n < nodes; ++n) in.node_starts[n + 1] += in.node_starts[n];
    in.node_at.resize(in.node_starts[nodes]);
    std::vector<uint32_t> nfill(in.node_starts.begin(), in.node_starts.end() - 1);
    for (size_t e = 0; e < elements; ++e)
        for (int a = 0; a < 10; ++a) in.node_at[nfill[mesh.tets[e][a]]++] = (uint32_t)(e * 10 + a);
    if (!with_vertices) return in;
    in.vert_starts.assign(vertices + 1, 0);
    for (const auto &t : mesh.tets)
        for (int a = 0; a < 4; ++a) in.vert_starts[t[a] + 1] += 1;
    for (size_t v = 0; v < vertices; ++v) in.vert_starts[v + 1] += in.vert_starts[v];
    in.vert_at.resize(in.vert_starts[vertices]);
    std::vector<uint32_t> vfill(in.vert_starts.begin(), in.vert_starts.end() - 1);
    in.ends.assign(2 * (nodes - vertices), 0);
    std::vector<char> seen(nodes - vertices, 0);
    static const int EA[6] = {0, 1, 2, 0, 1, 2}, EB[6] = {1, 2, 0, 3, 3, 3};
    for (size_t e = 0; e < elements; ++e) {
        const auto &t = mesh.tets[e];
        for (int a = 0; a < 4; ++a) in.vert_at[vfill[t[a]]++] = (uint32_t)e;
        for (int slot = 0; slot < 6; ++slot) {
            const size_t m = t[4 + slot] - vertices;
            if (seen[m]) continue;
            seen[m] = 1;
            in.ends[2 * m] = t[EA[slot]];
            in.ends[2 * m + 1] = t[EB[slot]];
        }
    }
    return in;
}

std::vector<double> positions_of(const QuadMesh &mesh) {
    std::vector<double> xyz(mesh.nodes.size() * 3);
    for (size_t n = 0; n < mesh.nodes.size(); ++n)
        for (int a = 0; a < 3; ++a) xyz[3 * n + a] = mesh.nodes[n][a];
    return xyz;
}

std::vector<uint32_t> tets_of(const QuadMesh &mesh) {
    std::vector<uint32_t> tets(mesh.tets.size() * 10);
    for (size_t e = 0; e < mesh.tets.size(); ++e)
        for (int a = 0; a < 10; ++a) tets[10 * e + a] = mesh.tets[e][a];
    return tets;
}

}  // namespace

DeviceRecovery::DeviceRecovery(const Model &model, const double *u, bool converged, hipStream_t s)
    : s_(s),
      elements_(model.mesh.tets.size()),
      nodes_(model.mesh.nodes.size()),
      vertices_(model.mesh.vertex_count),
      nq_(model.quadrature == 14 ? 14 : 4),
      field_(model.field),
      converged_(converged),
      u_(u, u + 3 * model.mesh.nodes.size()) {
    if (elements_ >= (size_t(1) << 25)) throw Error(SF_ERR_INVALID_ARGUMENT, "too many elements for one recovery");
    StageClock clock;
    const Incidence in = incidence_of(model.mesh, true);
    clock.lap("recovery: incidence lists (host)");
    xyz_.upload(positions_of(model.mesh), s);
    u_d_.upload(u_, s);
    tets_.upload(tets_of(model.mesh), s);
    nstart_.upload(in.node_starts, s);
    nat_.upload(in.node_at, s);
    vstart_.upload(in.vert_starts, s);
    vat_.upload(in.vert_at, s);
    ends_.upload(in.ends.empty() ? std::vector<uint32_t>{0, 0} : in.ends, s);
    nearest_.upload(nearest_points(nq_), s);
    if (field_.kind == SF_MATERIAL_POINTS) field_table_.upload(field_.points(nq_), s);
    clock.lap("recovery: uploads");
    points_.resize(elements_ * nq_ * 10);
    material_.resize(elements_ * nq_ * 2);
    nodevals_.resize(elements_ * 10 * 7);
    worst_.resize(nodes_ * 6);
    fit_.resize(std::max<size_t>(vertices_, 1) * 29);
    nodal_.resize(nodes_ * 7);
    fault_.resize(1);
}

void DeviceRecovery::run_elements(size_t count) {
    const int32_t none = NO_FAULT;
    fault_.upload(&none, 1, s_);
    const FieldArgs fa = field_args(field_, field_table_.ptr, nq_);
    if (nq_ == 14) {
        hipLaunchKernelGGL(recovery_points_kernel<14>, grid_of(count * 14), dim3(256), 0, s_, next_, count, xyz_.ptr, tets_.ptr,
                           u_d_.ptr, fa, points_.ptr, material_.ptr, fault_.ptr);
        hipLaunchKernelGGL(recovery_nodes_kernel<14>, grid_of(count * 10), dim3(256), 0, s_, next_, count, xyz_.ptr, tets_.ptr,
                           u_d_.ptr, (const double *)material_.ptr, (const int8_t *)nearest_.ptr, nodevals_.ptr, fault_.ptr);
    } else {
        hipLaunchKernelGGL(recovery_points_kernel<4>, grid_of(count * 4), dim3(256), 0, s_, next_, count, xyz_.ptr, tets_.ptr,
                           u_d_.ptr, fa, points_.ptr, material_.ptr, fault_.ptr);
        hipLaunchKernelGGL(recovery_nodes_kernel<4>, grid_of(count * 10), dim3(256), 0, s_, next_, count, xyz_.ptr, tets_.ptr,
                           u_d_.ptr, (const double *)material_.ptr, (const int8_t *)nearest_.ptr, nodevals_.ptr, fault_.ptr);
    }
    SF_HIP(hipGetLastError());
    int32_t fault = NO_FAULT;
    fault_.download(&fault, 1, s_);
    SF_HIP(hipStreamSynchronize(s_));
    if (fault != NO_FAULT) {
        // the elements before the first faulting one are done; it stays the next to do. The report
        // is element * 64 + what: 2 q (+1 for the material) at integration point q, then 28 + 2 n
        // at node n (the order of Recovery::advance's loops, results.rs:477-536)
        const int64_t e = fault / 64, what = fault % 64;
        next_ = (size_t)e;
        if (what < 28 && what % 2)
            throw Error(SF_ERR_MODEL_INVALID_MATERIAL, "an element has an invalid material at a point", e, what / 2);
        throw Error(SF_ERR_MODEL_INVERTED_JACOBIAN, "an element is inverted at a point", e, what < 28 ? what / 2 : (what - 28) / 2);
    }
    next_ += count;
}

void DeviceRecovery::ensure_worst() {
    if (worst_done_) return;
    hipLaunchKernelGGL(recovery_worst_kernel, grid_of(nodes_), dim3(256), 0, s_, nodes_, (const uint32_t *)nstart_.ptr,
                       (const uint32_t *)nat_.ptr, (const double *)nodevals_.ptr, worst_.ptr);
    SF_HIP(hipGetLastError());
    worst_done_ = true;
}

void DeviceRecovery::run_fits(size_t count) {
    if (nq_ == 14)
        hipLaunchKernelGGL(recovery_fit_kernel<14>, dim3((unsigned)count), dim3(FIT_LANES), 0, s_, next_vertex_, count, (const double *)xyz_.ptr,
                           (const uint32_t *)vstart_.ptr, (const uint32_t *)vat_.ptr, (const double *)points_.ptr,
                           (const double *)material_.ptr, fit_.ptr);
    else
        hipLaunchKernelGGL(recovery_fit_kernel<4>, dim3((unsigned)count), dim3(FIT_LANES), 0, s_, next_vertex_, count, (const double *)xyz_.ptr,
                           (const uint32_t *)vstart_.ptr, (const uint32_t *)vat_.ptr, (const double *)points_.ptr,
                           (const double *)material_.ptr, fit_.ptr);
    SF_HIP(hipGetLastError());
    SF_HIP(hipStreamSynchronize(s_));
    next_vertex_ += count;
}

bool DeviceRecovery::advance(size_t max_work, const std::function<bool()> &cancel) {
    // Recovery::advance (results.rs:477-536): the budget is spent on elements first, and patch
    // fits start only once every element is done; the cancel check is read once per unit.
    size_t budget = max_work, take = 0;
    bool cancelled = false;
    if (!cancel) {  // nothing to poll
        take = std::min(budget, elements_ - next_);
        budget -= take;
    }
    while (cancel && next_ + take < elements_ && budget > 0) {
        budget -= 1;
        if (cancel()) {
            cancelled = true;
            break;
        }
        take += 1;
    }
    if (take) run_elements(take);
    if (cancelled) throw Error(SF_ERR_MODEL_CANCELLED, "recovery cancelled");
    if (next_ == elements_ && next_vertex_ < vertices_ && budget > 0) {
        // the worst incident value at a node reads every element's node values, all done now
        ensure_worst();
        take = 0;
        if (!cancel) {
            take = std::min(budget, vertices_ - next_vertex_);
            budget -= take;
        }
        while (cancel && next_vertex_ + take < vertices_ && budget > 0) {
            budget -= 1;
            if (cancel()) {
                cancelled = true;
                break;
            }
            take += 1;
        }
        if (take) run_fits(take);
        if (cancelled) throw Error(SF_ERR_MODEL_CANCELLED, "recovery cancelled");
    }
    return finished();
}

void DeviceRecovery::field(double *point_values, double *nodal_values) {
    if (!finished()) throw Error(SF_ERR_MODEL_NOT_FINISHED, "the recovery has elements or patch fits left to do",
                                 (int64_t)next_, (int64_t)next_vertex_);
    if (elements_ == 0) return;
    ensure_worst();
    if (!nodal_done_) {
        hipLaunchKernelGGL(recovery_nodal_kernel, grid_of(nodes_), dim3(256), 0, s_, nodes_, vertices_, (const double *)xyz_.ptr,
                           (const uint32_t *)ends_.ptr, (const double *)fit_.ptr, (const double *)worst_.ptr, nodal_.ptr);
        SF_HIP(hipGetLastError());
        nodal_done_ = true;
    }
    StageClock clock;
    SF_HIP(hipStreamSynchronize(s_));
    clock.lap("recovery: nodal field");
    if (point_values) points_.download(point_values, elements_ * nq_ * 10, s_);
    if (nodal_values) nodal_.download(nodal_values, nodes_ * 7, s_);
    SF_HIP(hipStreamSynchronize(s_));
    clock.lap("recovery: downloads");
}

void recover_stress(const Model &model, const double *u, double *point_values, double *nodal_values, hipStream_t s) {
    if (model.mesh.tets.empty()) return;
    DeviceRecovery r(model, u, true, s);
    StageClock clock;
    r.advance(std::numeric_limits<size_t>::max(), nullptr);
    clock.lap("recovery: elements and fits");
    r.field(point_values, nodal_values);
}

DeviceErrorEstimation::DeviceErrorEstimation(const Model &model, DeviceRecovery &recovery, hipStream_t s)
    : recovery_(recovery),
      s_(s),
      elements_(model.mesh.tets.size()),
      nodes_(model.mesh.nodes.size()),
      vertices_(model.mesh.vertex_count),
      field_(model.field) {
    // ErrorEstimation::new (error_estimate.rs:94-103) refuses nothing; a recovery of another mesh
    // is refused by advance, so nothing of this mesh is needed then.
    if (!matches()) return;
    StageClock clock;
    xyz_.upload(positions_of(model.mesh), s);
    tets_.upload(tets_of(model.mesh), s);
    if (field_.kind == SF_MATERIAL_POINTS) field_table_.upload(field_.points(14), s);
    clock.lap("error estimate: mesh upload");
    element_errors_.reserve(elements_);
}

bool DeviceErrorEstimation::matches() const {
    // the counts of error_estimate.rs:128-134: displacements, fits and elements
    return recovery_.nodes() == nodes_ && recovery_.vertices() == vertices_ && recovery_.elements() == elements_;
}

void DeviceErrorEstimation::run(size_t count) {
    constexpr size_t SLICE = size_t(1) << 16;  // elements per launch: 31 MB of recovered values
    if (!recovered_.count) {
        recovered_.resize(std::min(SLICE, elements_) * 60);
        terms_.resize(std::min(SLICE, elements_) * 28);
        sums_.resize(std::min(SLICE, elements_) * 2);
        fault_.resize(1);
    }
    const FieldArgs fa = field_args(field_, field_table_.ptr, 14);
    StageClock clock;
    std::vector<double> sums;
    for (size_t done = 0; done < count;) {
        const size_t n = std::min(SLICE, count - done), first = next_;
        const unsigned long long none = NO_ESTIMATE_FAULT;
        SF_HIP(hipMemcpyAsync(fault_.ptr, &none, sizeof(none), hipMemcpyHostToDevice, s_));
        hipLaunchKernelGGL(estimate_nodes_kernel, grid_of(n * 10), dim3(256), 0, s_, first, n, (const double *)xyz_.ptr,
                           (const uint32_t *)tets_.ptr, recovery_.device_fits(), recovered_.ptr, fault_.ptr);
        hipLaunchKernelGGL(estimate_points_kernel, grid_of(n * 14), dim3(256), 0, s_, first, n, (const double *)xyz_.ptr,
                           (const uint32_t *)tets_.ptr, recovery_.device_displacements(), (const double *)recovered_.ptr, fa,
                           terms_.ptr, fault_.ptr);
        hipLaunchKernelGGL(estimate_sums_kernel, grid_of(n), dim3(256), 0, s_, n, (const double *)terms_.ptr, sums_.ptr);
        SF_HIP(hipGetLastError());
        unsigned long long fault = NO_ESTIMATE_FAULT;
        sums.resize(n * 2);
        fault_.download(&fault, 1, s_);
        sums_.download(sums.data(), n * 2, s_);
        SF_HIP(hipStreamSynchronize(s_));
        // ErrorEstimation::advance's running sums (error_estimate.rs:143-148), in element order,
        // up to the first refusing element
        const size_t good = fault == NO_ESTIMATE_FAULT ? n : (size_t)(fault / 32) - first;
        for (size_t k = 0; k < good; ++k) {
            element_errors_.push_back(std::sqrt(sums[2 * k]));
            error_sum_ += sums[2 * k];
            energy_sum_ += sums[2 * k + 1];
        }
        next_ += good;
        if (fault != NO_ESTIMATE_FAULT) throw refusal(fault, first);
        done += n;
        clock.lap("error estimate: a slice of elements");
    }
}

Error DeviceErrorEstimation::refusal(unsigned long long fault, size_t first) const {
    const size_t e = (size_t)(fault / 32);
    const int what = (int)(fault % 32);
    if (what < 4) {
        // the vertex of the first corner with no fit: a patch of two materials, or a flat one
        uint32_t vertex = 0;
        double flag = 0.0;
        SF_HIP(hipMemcpyAsync(&vertex, tets_.ptr + e * 10 + what, sizeof(vertex), hipMemcpyDeviceToHost, s_));
        SF_HIP(hipStreamSynchronize(s_));
        SF_HIP(hipMemcpyAsync(&flag, recovery_.device_fits() + (size_t)vertex * 29, sizeof(flag), hipMemcpyDeviceToHost, s_));
        SF_HIP(hipStreamSynchronize(s_));
        if (flag < 0.0)
            return Error(SF_ERR_MODEL_MIXED_MATERIAL_PATCH, "the elements at a vertex are of more than one material", vertex);
        return Error(SF_ERR_MODEL_COPLANAR_PATCH, "the elements at a vertex are flat: no smoothed stress was fitted", vertex);
    }
    const int point = (what - 4) / 2;
    if ((what - 4) % 2 == 0) {
        double det = 0.0;
        const size_t slot = e - first;  // the slice's terms start at its first element
        SF_HIP(hipMemcpyAsync(&det, terms_.ptr + (slot * 14 + (size_t)point) * 2, sizeof(det), hipMemcpyDeviceToHost, s_));
        SF_HIP(hipStreamSynchronize(s_));
        return Error(SF_ERR_MODEL_INVERTED_JACOBIAN, "an element is inverted at a point of the 14-point rule", (int64_t)e, point,
                     det);
    }
    return Error(SF_ERR_MODEL_INVALID_MATERIAL, "an element has an invalid material at a point of the 14-point rule", (int64_t)e,
                 point);
}

bool DeviceErrorEstimation::advance(size_t max_elements, const std::function<bool()> &cancel) {
    // ErrorEstimation::advance (error_estimate.rs:121-151)
    if (!matches()) throw Error(SF_ERR_MODEL_RECOVERY_MESH_MISMATCH, "the recovery was begun for another mesh");
    if (!recovery_.finished() || !recovery_.converged())
        throw Error(SF_ERR_MODEL_NOT_FINISHED, "the recovery is unfinished or of a solve that had not converged");
    size_t budget = max_elements, take = 0;
    bool cancelled = false;
    if (!cancel) take = std::min(budget, elements_ - next_);  // nothing to poll
    while (cancel && next_ + take < elements_ && budget > 0) {
        budget -= 1;
        if (cancel()) {
            cancelled = true;
            break;
        }
        take += 1;
    }
    if (take) run(take);
    if (cancelled) throw Error(SF_ERR_MODEL_CANCELLED, "error estimate cancelled");
    return next_ == elements_;
}

void DeviceErrorEstimation::estimate(double *element_errors, double *error_norm_squared, double *solution_norm_squared,
                                     double *relative_error) const {
    // ErrorEstimation::estimate (error_estimate.rs:153-171)
    if (next_ != elements_ || !recovery_.converged())
        throw Error(SF_ERR_MODEL_NOT_FINISHED, "the error estimate has elements left to do", (int64_t)next_);
    if (!std::isfinite(energy_sum_) || !std::isfinite(error_sum_))
        throw Error(SF_ERR_MODEL_NON_FINITE_ENERGY, "the solution's energy is not a finite number");
    if (energy_sum_ <= 0.0) throw Error(SF_ERR_MODEL_ZERO_STRAIN_ENERGY, "the solution stores no strain energy");
    if (element_errors) std::copy(element_errors_.begin(), element_errors_.end(), element_errors);
    if (error_norm_squared) *error_norm_squared = error_sum_;
    if (solution_norm_squared) *solution_norm_squared = energy_sum_;
    if (relative_error) *relative_error = std::sqrt(error_sum_ / (energy_sum_ + error_sum_));
}

void nodal_rotations(const Model &model, const double *u, double *rotations, hipStream_t s) {
    const auto &mesh = model.mesh;
    const size_t elements = mesh.tets.size(), nodes = mesh.nodes.size();
    if (elements == 0) return;
    const Incidence in = incidence_of(mesh, false);
    DeviceBuffer<double> d_xyz, d_u, d_values(elements * 10 * 4), d_out(nodes * 3);
    DeviceBuffer<uint32_t> d_tets, d_starts, d_at;
    DeviceBuffer<int32_t> d_fault;
    d_xyz.upload(positions_of(mesh), s);
    d_u.upload(u, nodes * 3, s);
    d_tets.upload(tets_of(mesh), s);
    d_starts.upload(in.node_starts, s);
    d_at.upload(in.node_at, s);
    const int32_t none = NO_FAULT;
    d_fault.upload(&none, 1, s);
    hipLaunchKernelGGL(rotation_nodes_kernel, grid_of(elements * 10), dim3(256), 0, s, elements, (const double *)d_xyz.ptr,
                       (const uint32_t *)d_tets.ptr, (const double *)d_u.ptr, d_values.ptr, d_fault.ptr);
    SF_HIP(hipGetLastError());
    int32_t fault = NO_FAULT;
    d_fault.download(&fault, 1, s);
    SF_HIP(hipStreamSynchronize(s));
    if (fault != NO_FAULT) throw Error(SF_ERR_MODEL_INVERTED_JACOBIAN, "an element is inverted at a node", fault);
    hipLaunchKernelGGL(rotation_gather_kernel, grid_of(nodes), dim3(256), 0, s, nodes, (const uint32_t *)d_starts.ptr,
                       (const uint32_t *)d_at.ptr, (const double *)d_values.ptr, d_out.ptr);
    SF_HIP(hipGetLastError());
    d_out.download(rotations, nodes * 3, s);
    SF_HIP(hipStreamSynchronize(s));
}

void DeviceAssembly::prepare(const Model &model, const AssemblyPattern &pattern, hipStream_t s) {
    const auto &mesh = model.mesh;
    nodes.upload(positions_of(mesh), s);
    tets.upload(tets_of(mesh), s);
    contribution_starts.upload(pattern.contribution_starts.data(), pattern.contribution_starts.size(), s);
    contributions.upload(pattern.contributions.data(), pattern.contributions.size(), s);
    elements = mesh.tets.size();
    if (elements >= (1u << 25)) throw Error(SF_ERR_INVALID_ARGUMENT, "too many elements for the assembly map");
    field = model.field;
    quadrature = model.quadrature;
    if (field.kind == SF_MATERIAL_POINTS) field_table.upload(field.points(quadrature == 14 ? 14 : 4), s);
    ke.resize(std::min(elements, CHUNK_ELEMENTS) * 55 * 9);
    fault.resize(1);
    const int32_t none = 0x7fffffff;
    fault.upload(&none, 1, s);
    SF_HIP(hipStreamSynchronize(s));
}

void DeviceAssembly::add_elements(size_t first, size_t last, double *values, size_t blocks, hipStream_t s) {
    const FieldArgs fa = field_args(field, field_table.ptr, quadrature == 14 ? 14 : 4);
    for (size_t c0 = first; c0 < last; c0 += CHUNK_ELEMENTS) {
        const size_t c1 = std::min(last, c0 + CHUNK_ELEMENTS);
        const size_t n = c1 - c0;
        const unsigned grid = (unsigned)((n + EPB - 1) / EPB);
        if (quadrature == 14)
            hipLaunchKernelGGL(element_kernel<14>, dim3(grid), dim3(256), 0, s, c0, n, nodes.ptr, tets.ptr, fa, ke.ptr, fault.ptr);
        else
            hipLaunchKernelGGL(element_kernel<4>, dim3(grid), dim3(256), 0, s, c0, n, nodes.ptr, tets.ptr, fa, ke.ptr, fault.ptr);
        SF_HIP(hipGetLastError());
        if (blocks)
            hipLaunchKernelGGL(gather_kernel, dim3((unsigned)((blocks + 255) / 256)), dim3(256), 0, s, blocks, contribution_starts.ptr,
                               contributions.ptr, (uint32_t)c0, (uint32_t)c1, ke.ptr, values);
        SF_HIP(hipGetLastError());
    }
}

void DeviceAssembly::check_fault(hipStream_t s) {
    int32_t f = 0;
    fault.download(&f, 1, s);
    SF_HIP(hipStreamSynchronize(s));
    // the first refusal in element then point order, the geometry before the material at a point
    // (element.rs:98-133): element * 64 + point * 2 + (1 for the material)
    if (f == 0x7fffffff) return;
    if (f % 2) throw Error(SF_ERR_MODEL_INVALID_MATERIAL, "an element has an invalid material at a point", f / 64, (f % 64) / 2);
    throw Error(SF_ERR_MODEL_INVERTED_JACOBIAN, "an element is inverted at a point", f / 64, (f % 64) / 2);
}

void DeviceAssembly::element_matrices(size_t first, size_t count, double *ke900, hipStream_t s) {
    const FieldArgs fa = field_args(field, field_table.ptr, quadrature == 14 ? 14 : 4);
    DeviceBuffer<double> out(std::max<size_t>(count, 1) * 55 * 9);
    const unsigned grid = (unsigned)((count + EPB - 1) / EPB);
    if (quadrature == 14)
        hipLaunchKernelGGL(element_kernel<14>, dim3(grid), dim3(256), 0, s, first, count, nodes.ptr, tets.ptr, fa, out.ptr, fault.ptr);
    else
        hipLaunchKernelGGL(element_kernel<4>, dim3(grid), dim3(256), 0, s, first, count, nodes.ptr, tets.ptr, fa, out.ptr, fault.ptr);
    SF_HIP(hipGetLastError());
    std::vector<double> upper(count * 55 * 9);
    out.download(upper.data(), upper.size(), s);
    SF_HIP(hipStreamSynchronize(s));
    for (size_t e = 0; e < count; ++e) {
        int t = 0;
        for (int a = 0; a < 10; ++a)
            for (int b = a; b < 10; ++b, ++t)
                for (int i = 0; i < 3; ++i)
                    for (int k = 0; k < 3; ++k) {
                        const double v = upper[(e * 55 + t) * 9 + i * 3 + k];
                        ke900[e * 900 + (3 * a + i) * 30 + (3 * b + k)] = v;
                        ke900[e * 900 + (3 * b + k) * 30 + (3 * a + i)] = v;
                    }
    }
}

}  // namespace sf
