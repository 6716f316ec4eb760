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
                atomicMin(&fault[0], (int32_t)e);
            }
            // material at the point (material.rs:24-50)
            double E = field.e, nu = field.nu;
            if (field.kind == SF_MATERIAL_SHELL_CORE) {
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
                atomicMin(&fault[0], (int32_t)e);
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

FieldArgs field_args(const MaterialField &f) {
    FieldArgs a{};
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

constexpr size_t CHUNK_ELEMENTS = 1u << 20;

}  // namespace

void DeviceAssembly::prepare(const Model &model, const AssemblyPattern &pattern, hipStream_t s) {
    const auto &mesh = model.mesh;
    std::vector<double> xyz(mesh.nodes.size() * 3);
    for (size_t n = 0; n < mesh.nodes.size(); ++n)
        for (int a = 0; a < 3; ++a) xyz[3 * n + a] = mesh.nodes[n][a];
    nodes.upload(xyz, s);
    std::vector<uint32_t> t(mesh.tets.size() * 10);
    for (size_t e = 0; e < mesh.tets.size(); ++e)
        for (int a = 0; a < 10; ++a) t[10 * e + a] = mesh.tets[e][a];
    tets.upload(t, s);
    contribution_starts.upload(pattern.contribution_starts, s);
    contributions.upload(pattern.contributions, s);
    elements = mesh.tets.size();
    if (elements >= (1u << 25)) throw Error(SF_ERR_INVALID_ARGUMENT, "too many elements for the assembly map");
    field = model.field;
    quadrature = model.quadrature;
    ke.resize(std::min(elements, CHUNK_ELEMENTS) * 55 * 9);
    fault.resize(1);
    const int32_t none = 0x7fffffff;
    fault.upload(&none, 1, s);
    SF_HIP(hipStreamSynchronize(s));
}

void DeviceAssembly::add_elements(size_t first, size_t last, double *values, size_t blocks, hipStream_t s) {
    const FieldArgs fa = field_args(field);
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
    if (f != 0x7fffffff)
        throw Error(SF_ERR_MODEL_INVERTED_JACOBIAN, "an element has an inverted point or an invalid material", f);
}

void DeviceAssembly::element_matrices(size_t first, size_t count, double *ke900, hipStream_t s) {
    const FieldArgs fa = field_args(field);
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
