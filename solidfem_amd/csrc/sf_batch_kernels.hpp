// Device kernels for several right-hand sides at once (load cases sharing one matrix and one
// multigrid hierarchy).
//
// Vectors of M columns are stored row-interleaved: entry (row i, column m) at v[i * M + m], so
// the M values of a row are contiguous. The matrices are the sliced block matrices of the
// single-column kernels, read with another lane mapping: a group of G lanes (G a power of two,
// G >= M) takes one block row, lane m of the group its column m, and every lane forms the RB
// row sums of its block row for its column. In the sliced layout the RB x CB values of the block
// in slot k of slice row j are contiguous at vals[((start + k) * 64 + j * RB) * CB], so a group
// reads each block once (one broadcast) and each column of x as G consecutive values.
//
// Each row sum runs over the row's blocks in stored order and over a block's columns in order,
// from zero, exactly as the single-column kernels and the reference's block product
// (crates/solidfem-solve/src/multigrid/block.rs:136-154) do: column m of a batched product is
// bit for bit the single-column product of column m. The same holds for every kernel below, so a
// column's solve does not depend on the columns batched with it.
#pragma once

#include "sf_kernels.hpp"

namespace sf {
namespace dev {

// A group's place: the node position (slice * NPW + j) it works on and its column.
struct GroupLane {
    uint32_t pos;  // node position in slice order
    int m;         // column of this lane
};

// With ORDER_GROUP_RUNS in the matrix view's order, each XCD takes one contiguous run of the
// workgroups (as xcd_runs_block): the rows an XCD works on at a time are then neighbours, and so
// are the rows of x they gather, which its L2 can hold for all their columns. Which workgroup
// takes which rows changes no bit.
__device__ __forceinline__ uint32_t group_block(uint32_t order) {
    if (order & ORDER_GROUP_RUNS) return xcd_runs_block();
    if (order & ORDER_GROUP_CHUNKS) return xcd_chunks_block(order >> ORDER_CHUNK_SHIFT);
    return blockIdx.x;
}

__device__ __forceinline__ GroupLane group_lane(int log2g, uint32_t order) {
    const uint32_t t = group_block(order) * blockDim.x + threadIdx.x;
    return {t >> log2g, (int)(t & ((1u << log2g) - 1u))};
}

// CPL consecutive columns of one row-interleaved vector at p (p aligned to CPL values).
template <int CPL, typename V>
__device__ __forceinline__ void load_columns(const V *__restrict__ p, V (&out)[CPL]) {
    if constexpr (CPL == 1) {
        out[0] = p[0];
    } else if constexpr (CPL == 2) {
        using T = typename std::conditional<std::is_same<V, float>::value, float2, double2>::type;
        const T t = *reinterpret_cast<const T *>(p);
        out[0] = t.x;
        out[1] = t.y;
    } else if constexpr (std::is_same<V, float>::value) {
        static_assert(CPL % 4 == 0, "two, four or eight columns per lane");
#pragma unroll
        for (int h = 0; h < CPL / 4; ++h) {
            const float4 t = reinterpret_cast<const float4 *>(p)[h];
            out[4 * h] = t.x;
            out[4 * h + 1] = t.y;
            out[4 * h + 2] = t.z;
            out[4 * h + 3] = t.w;
        }
    } else {
        static_assert(CPL % 4 == 0, "two, four or eight columns per lane");
#pragma unroll
        for (int h = 0; h < CPL / 2; ++h) {
            const double2 t = reinterpret_cast<const double2 *>(p)[h];
            out[2 * h] = t.x;
            out[2 * h + 1] = t.y;
        }
    }
}

// Stores CPL consecutive columns at p (aligned as for load_columns).
template <int CPL, typename V>
__device__ __forceinline__ void store_columns(V *__restrict__ p, const V (&in)[CPL]) {
    if constexpr (CPL == 1) {
        p[0] = in[0];
    } else if constexpr (CPL == 2) {
        using T = typename std::conditional<std::is_same<V, float>::value, float2, double2>::type;
        *reinterpret_cast<T *>(p) = T{in[0], in[1]};
    } else if constexpr (std::is_same<V, float>::value) {
#pragma unroll
        for (int h = 0; h < CPL / 4; ++h) reinterpret_cast<float4 *>(p)[h] = float4{in[4 * h], in[4 * h + 1], in[4 * h + 2], in[4 * h + 3]};
    } else {
#pragma unroll
        for (int h = 0; h < CPL / 2; ++h) reinterpret_cast<double2 *>(p)[h] = double2{in[2 * h], in[2 * h + 1]};
    }
}

// Row sums of one block row for the CPL columns m0 .. m0 + CPL - 1: sum[p][c] = sum over stored
// blocks k and block columns q of A[p][q] x[(col * CB + q) * M + m0 + c], each from zero in the
// reference order. A lane reads a block once for all its columns.
template <int RB, int CB, int BATCH, typename S, typename V, int CPL>
__device__ __forceinline__ void rows_product_columns(const Sliced<S> &a, uint32_t slot0, uint32_t len, int j, int m0, uint32_t M,
                                                     const V *__restrict__ x, V (&sum)[RB][CPL]) {
    constexpr int NPW = Shape<RB>::NPW;
#pragma unroll
    for (int p = 0; p < RB; ++p)
#pragma unroll
        for (int c = 0; c < CPL; ++c) sum[p][c] = (V)0;
    const uint32_t *cp = a.cols + (size_t)slot0 * NPW + j;
    const S *vp = a.vals + ((size_t)slot0 * WAVE + (size_t)j * RB) * CB;
    const uint32_t full = len - len % BATCH;
    // the next batch's block columns are loaded with this batch's values and x entries, so a
    // batch costs one dependent memory round trip instead of two
    uint32_t col[BATCH];
    if (full) {
#pragma unroll
        for (int b = 0; b < BATCH; ++b) col[b] = cp[(size_t)b * NPW];
    }
    for (uint32_t k = 0; k < full; k += BATCH) {
        uint32_t next[BATCH];
        if (k + BATCH < full) {
#pragma unroll
            for (int b = 0; b < BATCH; ++b) next[b] = cp[(size_t)(k + BATCH + b) * NPW];
        }
        V xv[BATCH][CB][CPL];
#pragma unroll
        for (int b = 0; b < BATCH; ++b)
#pragma unroll
            for (int q = 0; q < CB; ++q) load_columns<CPL, V>(x + ((size_t)col[b] * CB + q) * M + m0, xv[b][q]);
        S v[BATCH][RB * CB];
#pragma unroll
        for (int b = 0; b < BATCH; ++b)
#pragma unroll
            for (int e = 0; e < RB * CB; ++e) v[b][e] = vp[(size_t)(k + b) * WAVE * CB + e];
#pragma unroll
        for (int b = 0; b < BATCH; ++b)
#pragma unroll
            for (int p = 0; p < RB; ++p)
#pragma unroll
                for (int q = 0; q < CB; ++q)
#pragma unroll
                    for (int c = 0; c < CPL; ++c) sum[p][c] = sum[p][c] + widen<V>(v[b][p * CB + q]) * xv[b][q][c];
#pragma unroll
        for (int b = 0; b < BATCH; ++b) col[b] = next[b];
    }
    for (uint32_t k = full; k < len; ++k) {
        const uint32_t cl = cp[(size_t)k * NPW];
        V xv[CB][CPL];
#pragma unroll
        for (int q = 0; q < CB; ++q) load_columns<CPL, V>(x + ((size_t)cl * CB + q) * M + m0, xv[q]);
        const S *blk = vp + (size_t)k * WAVE * CB;
#pragma unroll
        for (int p = 0; p < RB; ++p)
#pragma unroll
            for (int q = 0; q < CB; ++q)
#pragma unroll
                for (int c = 0; c < CPL; ++c) sum[p][c] = sum[p][c] + widen<V>(blk[p * CB + q]) * xv[q][c];
    }
}

// The same for one column m.
template <int RB, int CB, int BATCH, typename S, typename V>
__device__ __forceinline__ void rows_product(const Sliced<S> &a, uint32_t slot0, uint32_t len, int j, int m, uint32_t M,
                                             const V *__restrict__ x, V (&sum)[RB]) {
    V s[RB][1];
    rows_product_columns<RB, CB, BATCH, S, V, 1>(a, slot0, len, j, m, M, x, s);
#pragma unroll
    for (int p = 0; p < RB; ++p) sum[p] = s[p][0];
}

// Y = A X (STORE), Y = B - A X (RESIDUAL) or Y = Y + A X (ADD) for M columns: the batched form
// of spmv_kernel (products, residuals, restriction over the stored transpose, prolongation).
template <int RB, int CB, int BATCH, typename S, typename V, int CPL>
__global__ __launch_bounds__(256) void spmm_kernel(Sliced<S> a, int log2g, uint32_t M, int epi, const V *__restrict__ x,
                                                   V *__restrict__ y, const V *__restrict__ b) {
    // CPL > 1: lane m of the group takes the CPL columns from m * CPL (M a multiple of CPL)
    constexpr int NPW = Shape<RB>::NPW;
    const GroupLane g = group_lane(log2g, a.order);
    const uint32_t m0 = (uint32_t)g.m * CPL;
    if (g.pos >= a.slices * (uint32_t)NPW || m0 >= M) return;
    const uint32_t row = a.perm[g.pos];
    if (row == NONE) return;
    const uint32_t slice = g.pos / NPW;
    const int j = (int)(g.pos - slice * NPW);
    V sum[RB][CPL];
    rows_product_columns<RB, CB, BATCH, S, V, CPL>(a, a.slice_start[slice], a.row_len[g.pos], j, (int)m0, M, x, sum);
#pragma unroll
    for (int p = 0; p < RB; ++p) {
        const size_t out = ((size_t)row * RB + p) * M + m0;
        V v[CPL];
        if (epi == EPI_STORE) {
#pragma unroll
            for (int c = 0; c < CPL; ++c) v[c] = sum[p][c];
        } else {
            load_columns<CPL, V>((epi == EPI_RESIDUAL ? b : y) + out, v);
#pragma unroll
            for (int c = 0; c < CPL; ++c) v[c] = epi == EPI_RESIDUAL ? v[c] - sum[p][c] : v[c] + sum[p][c];
        }
        store_columns<CPL, V>(y + out, v);
    }
}

// The same product with one wavefront per block row: the row's position, block columns and block
// values are the same for every lane, so they are read through the scalar unit into registers
// every lane uses, and only the M columns of x are vector loads. Lane m takes column m (lanes
// from M up idle). Row sums in the same order as rows_product.
template <int RB, int CB, int BATCH, typename S, typename V>
__device__ __forceinline__ void node_product(const Sliced<S> &a, uint32_t pos, int m, uint32_t M, const V *__restrict__ x,
                                             uint32_t &row_out, V (&sum)[RB]) {
    constexpr int NPW = Shape<RB>::NPW;
    const uint32_t slice = pos / NPW;
    const int j = (int)(pos - slice * NPW);
    const uint32_t row = a.perm[pos];
    row_out = row;
#pragma unroll
    for (int p = 0; p < RB; ++p) sum[p] = (V)0;
    if (row == NONE) return;
    const uint32_t slot0 = a.slice_start[slice], len = a.row_len[pos];
    const uint32_t *__restrict__ cp = a.cols + (size_t)slot0 * NPW + j;
    const S *__restrict__ vp = a.vals + ((size_t)slot0 * WAVE + (size_t)j * RB) * CB;
    const uint32_t full = len - len % BATCH;
    for (uint32_t k = 0; k < full; k += BATCH) {
        uint32_t c[BATCH];
#pragma unroll
        for (int b = 0; b < BATCH; ++b) c[b] = __builtin_amdgcn_readfirstlane(cp[(size_t)(k + b) * NPW]);
        V xv[BATCH][CB];
#pragma unroll
        for (int b = 0; b < BATCH; ++b)
#pragma unroll
            for (int q = 0; q < CB; ++q) xv[b][q] = (uint32_t)m < M ? x[((size_t)c[b] * CB + q) * M + m] : (V)0;
#pragma unroll
        for (int b = 0; b < BATCH; ++b)
#pragma unroll
            for (int p = 0; p < RB; ++p)
#pragma unroll
                for (int q = 0; q < CB; ++q) sum[p] = sum[p] + widen<V>(vp[(size_t)(k + b) * WAVE * CB + p * CB + q]) * xv[b][q];
    }
    for (uint32_t k = full; k < len; ++k) {
        const uint32_t col = __builtin_amdgcn_readfirstlane(cp[(size_t)k * NPW]);
        V xv[CB];
#pragma unroll
        for (int q = 0; q < CB; ++q) xv[q] = (uint32_t)m < M ? x[((size_t)col * CB + q) * M + m] : (V)0;
#pragma unroll
        for (int p = 0; p < RB; ++p)
#pragma unroll
            for (int q = 0; q < CB; ++q) sum[p] = sum[p] + widen<V>(vp[(size_t)k * WAVE * CB + p * CB + q]) * xv[q];
    }
}

// Position of this wavefront's block row, made uniform for the compiler.
__device__ __forceinline__ uint32_t node_of_wave(uint32_t order) {
    return __builtin_amdgcn_readfirstlane(group_block(order) * (blockDim.x / WAVE) + threadIdx.x / WAVE);
}

// The same product with one workgroup per block row, for matrices of few, long rows (the stored
// transposes of the coarse levels): the workgroup stages a chunk of STAGE_BLOCKS of the row's
// blocks and the x entries they gather into LDS with loads that are all in flight at once, then
// thread (p, m) adds the chunk's terms of row entry p and column m in the stored order -- a chunk
// costs two round trips where a lane walking the row costs one per few blocks. Every sum is the
// lane-group form's: from zero, blocks in order, block columns in order.
constexpr int STAGE_BLOCKS = 32;
template <int RB, int CB, typename S, typename V>
__global__ __launch_bounds__(256) void spmm_staged_kernel(Sliced<S> a, uint32_t M, int epi, const V *__restrict__ x,
                                                          V *__restrict__ y, const V *__restrict__ b) {
    constexpr int NPW = Shape<RB>::NPW;
    const uint32_t pos = blockIdx.x;
    if (pos >= a.slices * (uint32_t)NPW) return;  // the whole workgroup leaves together
    const uint32_t row = a.perm[pos];
    if (row == NONE) return;
    const uint32_t slice = pos / NPW;
    const int j = (int)(pos - slice * NPW);
    const uint32_t slot0 = a.slice_start[slice], len = a.row_len[pos];
    const uint32_t *__restrict__ cp = a.cols + (size_t)slot0 * NPW + j;
    const S *__restrict__ vp = a.vals + ((size_t)slot0 * WAVE + (size_t)j * RB) * CB;
    __shared__ S values[STAGE_BLOCKS][RB * CB];
    __shared__ uint32_t columns[STAGE_BLOCKS];
    extern __shared__ unsigned char staged_x[];  // STAGE_BLOCKS x CB x M entries of x
    V *xs = reinterpret_cast<V *>(staged_x);
    const uint32_t t = threadIdx.x;
    const bool computes = t < (uint32_t)RB * M;
    const uint32_t p = computes ? t / M : 0, m = computes ? t % M : 0;
    V sum = (V)0;
    // a chunk's block columns are loaded during the chunk before it, so its values and its x
    // entries are fetched together: one round trip per chunk
    uint32_t column = t < (uint32_t)STAGE_BLOCKS && t < len ? cp[(size_t)t * NPW] : 0;
    for (uint32_t k0 = 0; k0 < len; k0 += STAGE_BLOCKS) {
        const uint32_t n = min((uint32_t)STAGE_BLOCKS, len - k0);
        if (t < n) columns[t] = column;
        __syncthreads();
        const uint32_t k1 = k0 + STAGE_BLOCKS;
        if (t < (uint32_t)STAGE_BLOCKS && k1 + t < len) column = cp[(size_t)(k1 + t) * NPW];
        for (uint32_t e = t; e < n * RB * CB; e += blockDim.x) {
            const uint32_t k = e / (RB * CB), r = e - k * (RB * CB);
            values[k][r] = vp[(size_t)(k0 + k) * WAVE * CB + r];
        }
        for (uint32_t e = t; e < n * CB * M; e += blockDim.x) {
            const uint32_t mm = e % M, kq = e / M;
            const uint32_t k = kq / CB, q = kq - k * CB;
            xs[e] = x[((size_t)columns[k] * CB + q) * M + mm];
        }
        __syncthreads();
        if (computes)
            for (uint32_t k = 0; k < n; ++k)
#pragma unroll
                for (int q = 0; q < CB; ++q) sum = sum + widen<V>(values[k][p * CB + q]) * xs[(k * CB + q) * M + m];
        __syncthreads();
    }
    if (!computes) return;
    const size_t out = ((size_t)row * RB + p) * M + m;
    V v = sum;
    if (epi == EPI_RESIDUAL) v = b[out] - sum;
    else if (epi == EPI_ADD) v = y[out] + sum;
    y[out] = v;
}

template <int RB, int CB, int BATCH, typename S, typename V>
__global__ __launch_bounds__(256) void spmm_node_kernel(Sliced<S> a, uint32_t M, int epi, const V *__restrict__ x,
                                                        V *__restrict__ y, const V *__restrict__ b) {
    constexpr int NPW = Shape<RB>::NPW;
    const uint32_t pos = node_of_wave(a.order);
    if (pos >= a.slices * (uint32_t)NPW) return;
    const int m = threadIdx.x % WAVE;
    uint32_t row;
    V sum[RB];
    node_product<RB, CB, BATCH, S, V>(a, pos, m, M, x, row, sum);
    if (row == NONE || (uint32_t)m >= M) return;
#pragma unroll
    for (int p = 0; p < RB; ++p) {
        const size_t out = ((size_t)row * RB + p) * M + m;
        if (epi == EPI_STORE) y[out] = sum[p];
        else if (epi == EPI_RESIDUAL) y[out] = b[out] - sum[p];
        else y[out] = y[out] + sum[p];
    }
}

template <int RB, int BATCH, typename S, typename V>
__global__ __launch_bounds__(256) void cheb_step_node_kernel(Sliced<S> a, uint32_t M, const S *__restrict__ dinv,
                                                             const V *__restrict__ d_in, const V *__restrict__ r_in, V keep,
                                                             V step, V *__restrict__ d_out, V *__restrict__ r_out,
                                                             V *__restrict__ x, V *__restrict__ sum_into) {
    constexpr int NPW = Shape<RB>::NPW;
    const uint32_t pos = node_of_wave(a.order);
    if (pos >= a.slices * (uint32_t)NPW) return;
    const int m = threadIdx.x % WAVE;
    uint32_t row;
    V ad[RB];
    node_product<RB, RB, BATCH, S, V>(a, pos, m, M, d_in, row, ad);
    if (row == NONE || (uint32_t)m >= M) return;
    V r_new[RB];
#pragma unroll
    for (int p = 0; p < RB; ++p) r_new[p] = r_in[((size_t)row * RB + p) * M + m] - ad[p];
#pragma unroll
    for (int p = 0; p < RB; ++p) {
        const S *drow = dinv + (size_t)row * RB * RB + (size_t)p * RB;
        V z = (V)0;
#pragma unroll
        for (int q = 0; q < RB; ++q) z = z + widen<V>(drow[q]) * r_new[q];
        const size_t i = ((size_t)row * RB + p) * M + m;
        const V dn = keep * d_in[i] + step * z;
        const V xn = x[i] + dn;
        d_out[i] = dn;
        r_out[i] = r_new[p];
        x[i] = xn;
        if (sum_into) sum_into[i] = sum_into[i] + xn;
    }
}

// The rounding bound terms of CsrMatrix::multiply_rounding_bound (crates/solidfem-solve/src/csr.rs:142-154)
// for M columns: per row and column, ((terms * eps) * sum |a x|)^2.
template <int RB, int CB>
__global__ __launch_bounds__(256) void rounding_bound_batch_kernel(Sliced<double> a, int log2g, uint32_t M,
                                                                   const double *__restrict__ x, double *__restrict__ out) {
    constexpr int NPW = Shape<RB>::NPW;
    const GroupLane g = group_lane(log2g, a.order);
    if (g.pos >= a.slices * (uint32_t)NPW || (uint32_t)g.m >= M) return;
    const uint32_t row = a.perm[g.pos];
    if (row == NONE) return;
    const uint32_t slice = g.pos / NPW;
    const int j = (int)(g.pos - slice * NPW);
    const uint32_t len = a.row_len[g.pos], slot0 = a.slice_start[slice];
    double sum[RB];
#pragma unroll
    for (int p = 0; p < RB; ++p) sum[p] = 0.0;
    for (uint32_t k = 0; k < len; ++k) {
        const uint32_t col = a.cols[(size_t)(slot0 + k) * NPW + j];
        const double *blk = a.vals + ((size_t)(slot0 + k) * WAVE + (size_t)j * RB) * CB;
#pragma unroll
        for (int p = 0; p < RB; ++p)
#pragma unroll
            for (int q = 0; q < CB; ++q) sum[p] = sum[p] + fabs(blk[p * CB + q] * x[((size_t)col * CB + q) * M + g.m]);
    }
#pragma unroll
    for (int p = 0; p < RB; ++p) {
        const double bound = (double)(len * CB) * 2.220446049250313e-16 * sum[p];
        out[((size_t)row * RB + p) * M + g.m] = bound * bound;
    }
}

// The Chebyshev start for M columns (crates/solidfem-solve/src/multigrid/cycle.rs:591-610):
// z = D^-1 b, d = z / theta, x = d. One thread per block row and column.
template <int RB, typename S, typename V>
__global__ __launch_bounds__(256) void cheb_start_batch_kernel(uint32_t rows, uint32_t M, const S *__restrict__ dinv,
                                                               const V *__restrict__ b, V theta, V *__restrict__ d,
                                                               V *__restrict__ x) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)rows * M) return;
    const size_t row = t / M, m = t - row * M;
    V bv[RB];
#pragma unroll
    for (int q = 0; q < RB; ++q) bv[q] = b[(row * RB + q) * M + m];
#pragma unroll
    for (int p = 0; p < RB; ++p) {
        const S *dr = dinv + row * RB * RB + p * RB;
        V z = (V)0;
#pragma unroll
        for (int q = 0; q < RB; ++q) z = z + widen<V>(dr[q]) * bv[q];
        const V dv = z / theta;
        d[(row * RB + p) * M + m] = dv;
        x[(row * RB + p) * M + m] = dv;
    }
}

// One later Chebyshev step for M columns, fused as cheb_step_kernel: ad = A d_in;
// r_out = r_in - ad; z = D^-1 r_out (the lane holds its block row's residuals); d_out = keep d_in
// + step z; x += d_out; with `sum_into`, sum_into += x (cycle.rs:611-650, :534-536, :655-674).
template <int RB, int BATCH, typename S, typename V, int CPL>
__global__ __launch_bounds__(256) void cheb_step_batch_kernel(Sliced<S> a, int log2g, uint32_t M, const S *__restrict__ dinv,
                                                              const V *__restrict__ d_in, const V *__restrict__ r_in, V keep,
                                                              V step, V *__restrict__ d_out, V *__restrict__ r_out,
                                                              V *__restrict__ x, V *__restrict__ sum_into) {
    constexpr int NPW = Shape<RB>::NPW;
    const GroupLane g = group_lane(log2g, a.order);
    const uint32_t m0 = (uint32_t)g.m * CPL;
    if (g.pos >= a.slices * (uint32_t)NPW || m0 >= M) return;
    const uint32_t row = a.perm[g.pos];
    if (row == NONE) return;
    const uint32_t slice = g.pos / NPW;
    const int j = (int)(g.pos - slice * NPW);
    V ad[RB][CPL];
    rows_product_columns<RB, RB, BATCH, S, V, CPL>(a, a.slice_start[slice], a.row_len[g.pos], j, (int)m0, M, d_in, ad);
    // the CPL columns of each row entry move as one vector
    V r_new[RB][CPL];
#pragma unroll
    for (int p = 0; p < RB; ++p) {
        V r[CPL];
        load_columns<CPL, V>(r_in + ((size_t)row * RB + p) * M + m0, r);
#pragma unroll
        for (int c = 0; c < CPL; ++c) r_new[p][c] = r[c] - ad[p][c];
    }
#pragma unroll
    for (int p = 0; p < RB; ++p) {
        const S *drow = dinv + (size_t)row * RB * RB + (size_t)p * RB;
        const size_t i = ((size_t)row * RB + p) * M + m0;
        V dold[CPL], xold[CPL], dn[CPL], xn[CPL];
        load_columns<CPL, V>(d_in + i, dold);
        load_columns<CPL, V>(x + i, xold);
#pragma unroll
        for (int c = 0; c < CPL; ++c) {
            V z = (V)0;
#pragma unroll
            for (int q = 0; q < RB; ++q) z = z + widen<V>(drow[q]) * r_new[q][c];
            dn[c] = keep * dold[c] + step * z;
            xn[c] = xold[c] + dn[c];
        }
        store_columns<CPL, V>(d_out + i, dn);
        store_columns<CPL, V>(r_out + i, r_new[p]);
        store_columns<CPL, V>(x + i, xn);
        if (sum_into) {
            V acc[CPL];
            load_columns<CPL, V>(sum_into + i, acc);
#pragma unroll
            for (int c = 0; c < CPL; ++c) acc[c] = acc[c] + xn[c];
            store_columns<CPL, V>(sum_into + i, acc);
        }
    }
}

// The coarsest solve as a product with the f64 inverse for M columns (coarse_inverse_kernel per
// column): out[i][m] = sum over k ascending of inv(i, k) b[k][m], from zero. A block takes R rows
// (coarse_rows: R * M' = 256 with M' the power of two at or above M, so one sum per thread) and all
// the columns: thread t is row t % R, column t / R. Tiles of 32 k of the inverse's R rows and of
// b pass through two LDS buffers, the next tile's loads in flight while the current one is summed.
constexpr int COARSE_TILE_K = 32;
constexpr int COARSE_STAGE = COARSE_TILE_K * 64 / 256;  // tile entries per thread, R and M at most 64
__host__ __device__ inline uint32_t coarse_rows(uint32_t M) {
    uint32_t p = 1;
    while (p < M) p *= 2;
    return p >= 4 ? 256 / p : 64;
}
template <typename V>
__global__ __launch_bounds__(256) void coarse_inverse_batch_kernel(uint32_t n, uint32_t M, uint32_t R, const double *__restrict__ inv,
                                                                   const V *__restrict__ rhs, V *__restrict__ out) {
    __shared__ double as[2][COARSE_TILE_K * 64];
    __shared__ double bs[2][COARSE_TILE_K * 64];
    const uint32_t t = threadIdx.x, r = t % R, m = t / R;
    const uint32_t i0 = blockIdx.x * R, i = i0 + r;
    double pa[COARSE_STAGE], pb[COARSE_STAGE];
    // tile k0's entries e = t + 256 j: inverse entry (k0 + e / R, i0 + e % R), b entry k0 * M + e
    auto fetch = [&](uint32_t k0) {
        const uint32_t kn = n - k0 < (uint32_t)COARSE_TILE_K ? n - k0 : (uint32_t)COARSE_TILE_K;
#pragma unroll
        for (int j = 0; j < COARSE_STAGE; ++j) {
            const uint32_t e = t + 256 * j, kk = e / R, rr = e % R;
            pa[j] = kk < kn && i0 + rr < n ? inv[(size_t)(k0 + kk) * n + i0 + rr] : 0.0;
            pb[j] = e < kn * M ? (double)rhs[(size_t)k0 * M + e] : 0.0;
        }
    };
    auto put = [&](int buf) {
#pragma unroll
        for (int j = 0; j < COARSE_STAGE; ++j) {
            as[buf][t + 256 * j] = pa[j];
            bs[buf][t + 256 * j] = pb[j];
        }
    };
    fetch(0);
    put(0);
    __syncthreads();
    double sum = 0.0;
    int buf = 0;
    for (uint32_t k0 = 0; k0 < n; k0 += COARSE_TILE_K) {
        const uint32_t kn = n - k0 < (uint32_t)COARSE_TILE_K ? n - k0 : (uint32_t)COARSE_TILE_K;
        const bool more = k0 + COARSE_TILE_K < n;
        if (more) fetch(k0 + COARSE_TILE_K);
        if (m < M && i < n)
            for (uint32_t kk = 0; kk < kn; ++kk) sum = sum + as[buf][kk * R + r] * bs[buf][kk * M + m];
        if (more) put(buf ^ 1);
        __syncthreads();
        buf ^= 1;
    }
    if (m < M && i < n) out[(size_t)i * M + m] = (V)sum;
}

// ---- conjugate gradient, M columns -----------------------------------------------------------

// Batched dot products whose every column is bit for bit the single-column TREE dot
// (dot_partials_kernel then dot_final_kernel): block `blk` forms, for each thread slot t of the
// single kernel and each column, the same grid-stride sum from zero; the sums are staged in LDS
// and reduced by the same pairs as the single kernel's shuffle butterfly (lane l with l + off for
// off = 32 .. 1, lane 0's value), then its four wave sums in order from zero. Columns read
// coalesced: thread u takes column u % M of slots t = u / M, u / M + 256 / M, ..., eight slots at a
// time, so that each grid-stride step has their sixteen loads in flight together.
constexpr int DOT_CHUNK = 8;
template <int MODE>
__global__ __launch_bounds__(DOT_THREADS) void dot_partials_batch_kernel(size_t n, uint32_t M, const double *__restrict__ a,
                                                                          const double *__restrict__ b,
                                                                          double *__restrict__ partials) {
    extern __shared__ double ls[];  // DOT_THREADS x M
    const uint32_t u = threadIdx.x;
    const uint32_t groups = DOT_THREADS / M;
    const size_t stride = (size_t)gridDim.x * DOT_THREADS;
    if (u < groups * M) {
        const uint32_t m = u % M;
        for (uint32_t t0 = u / M; t0 < (uint32_t)DOT_THREADS; t0 += DOT_CHUNK * groups) {
            double v[DOT_CHUNK];
#pragma unroll
            for (int j = 0; j < DOT_CHUNK; ++j) v[j] = 0.0;
            for (size_t base = (size_t)blockIdx.x * DOT_THREADS; base < n; base += stride) {
                double x[DOT_CHUNK], y[DOT_CHUNK];
#pragma unroll
                for (int j = 0; j < DOT_CHUNK; ++j) {
                    const size_t i = base + t0 + (size_t)j * groups;
                    const bool live = t0 + j * groups < (uint32_t)DOT_THREADS && i < n;
                    x[j] = live ? a[i * M + m] : 0.0;
                    y[j] = MODE == 0 && live ? b[i * M + m] : 0.0;
                }
#pragma unroll
                for (int j = 0; j < DOT_CHUNK; ++j)  // each slot's terms in ascending i, as alone
                    if (t0 + j * groups < (uint32_t)DOT_THREADS && base + t0 + (size_t)j * groups < n)
                        v[j] = v[j] + (MODE == 0 ? x[j] * y[j] : x[j]);
            }
#pragma unroll
            for (int j = 0; j < DOT_CHUNK; ++j)
                if (t0 + j * groups < (uint32_t)DOT_THREADS) ls[(t0 + j * groups) * M + m] = v[j];
        }
    }
    __syncthreads();
    // the butterfly of wave_sum, as pairs: s[l] = s[l] + s[l + off] for l < off
    for (uint32_t off = WAVE / 2; off > 0; off >>= 1) {
        const uint32_t work = (DOT_THREADS / WAVE) * off * M;
        for (uint32_t e = u; e < work; e += DOT_THREADS) {
            const uint32_t m = e % M, rest = e / M, l = rest % off, w = rest / off;
            const uint32_t t = w * WAVE + l;
            ls[t * M + m] = ls[t * M + m] + ls[(t + off) * M + m];
        }
        __syncthreads();
    }
    for (uint32_t m = u; m < M; m += DOT_THREADS) {
        double s = 0.0;
        for (int w = 0; w < DOT_THREADS / WAVE; ++w) s = s + ls[(size_t)w * WAVE * M + m];
        partials[(size_t)blockIdx.x * M + m] = s;
    }
}

// What a batched final sum feeds, per column: STORE writes the column's `out` field (`field` is a
// byte offset into PcgScalars); RZ forms rz_next, beta and rz as dot_finish does. Columns whose
// `mask` entry is zero are left as they are.
static __global__ __launch_bounds__(DOT_THREADS) void dot_final_batch_kernel(int parts, uint32_t M, const double *__restrict__ partials,
                                                                             int kind, uint32_t field, PcgScalars *s,
                                                                             const uint32_t *__restrict__ mask) {
    const uint32_t m = blockIdx.x;
    double v = 0.0;
    for (int i = threadIdx.x; i < parts; i += blockDim.x) v = v + partials[(size_t)i * M + m];
    const double total = block_sum(v);
    if (threadIdx.x == 0 && (!mask || mask[m])) {
        PcgScalars *c = s + m;
        if (kind == FINAL_STORE) *(double *)((char *)c + field) = total;
        else dot_finish(FINAL_RZ, total, nullptr, c);
    }
}

// REFERENCE order for M columns: one thread per column sums in index order from zero
// (crates/solidfem-solve/src/pcg.rs:285-291).
template <int MODE>
__global__ void dot_serial_batch_kernel(size_t n, uint32_t M, const double *__restrict__ a, const double *__restrict__ b,
                                        int kind, uint32_t field, PcgScalars *s, const uint32_t *__restrict__ mask) {
    const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= M) return;
    double sum = 0.0;
    for (size_t i = 0; i < n; ++i) sum = sum + (MODE == 0 ? a[i * M + m] * b[i * M + m] : a[i * M + m]);
    if (mask && !mask[m]) return;
    PcgScalars *c = s + m;
    if (kind == FINAL_STORE) *(double *)((char *)c + field) = sum;
    else dot_finish(FINAL_RZ, sum, nullptr, c);
}

// pcg_update_kernel per column; columns whose `active` entry is zero are not touched.
static __global__ void pcg_update_batch_kernel(size_t n, uint32_t M, PcgScalars *__restrict__ s, const uint32_t *__restrict__ active,
                                               const double *__restrict__ p, const double *__restrict__ ap,
                                               double *__restrict__ x, double *__restrict__ r) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n * M) return;
    const uint32_t m = (uint32_t)(e % M);
    if (!active[m]) return;
    const double curvature = s[m].curvature, rz = s[m].rz;
    const bool valid = !(isnan(curvature) || curvature <= 0.0 || !isfinite(rz));
    const double alpha = rz / curvature;
    if (e < M) {
        s[m].valid = valid ? 1 : 0;
        s[m].alpha = alpha;
    }
    if (!valid) return;
    x[e] = x[e] + alpha * p[e];
    r[e] = r[e] + (-alpha) * ap[e];
}

// p = z + beta p per active column.
static __global__ void pcg_direction_batch_kernel(size_t n, uint32_t M, const PcgScalars *__restrict__ s,
                                                  const uint32_t *__restrict__ active, const double *__restrict__ z,
                                                  double *__restrict__ p) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n * M) return;
    const uint32_t m = (uint32_t)(e % M);
    if (active[m]) p[e] = z[e] + s[m].beta * p[e];
}

// dst = src in the columns whose mask entry is set.
template <typename T>
__global__ void copy_columns_kernel(size_t n, uint32_t M, const uint32_t *__restrict__ mask, const T *__restrict__ src,
                                    T *__restrict__ dst) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n * M && mask[e % M]) dst[e] = src[e];
}

// Column m of a row-interleaved array from a plain vector.
template <typename T>
__global__ void set_column_kernel(size_t n, uint32_t M, uint32_t m, const T *__restrict__ src, T *__restrict__ dst) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i * M + m] = src[i];
}
template <typename T>
__global__ void get_column_kernel(size_t n, uint32_t M, uint32_t m, const T *__restrict__ src, T *__restrict__ dst) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i * M + m];
}

// Zeroes the masked columns of a row-interleaved array.
template <typename T>
__global__ void zero_columns_kernel(size_t n, uint32_t M, const uint32_t *__restrict__ mask, T *__restrict__ dst) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n * M && mask[e % M]) dst[e] = (T)0;
}

// Zeroes the scalars of the masked columns.
static __global__ void zero_scalars_kernel(uint32_t M, const uint32_t *__restrict__ mask, PcgScalars *s) {
    const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m < M && mask[m]) s[m] = PcgScalars{};
}

// Column-major (column m at m * n) to row-interleaved, and back.
template <typename T>
__global__ void interleave_kernel(size_t n, uint32_t M, const T *__restrict__ cm, T *__restrict__ il) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n * M) return;
    const size_t i = e / M, m = e - i * M;
    il[e] = cm[m * n + i];
}
template <typename T>
__global__ void deinterleave_kernel(size_t n, uint32_t M, const T *__restrict__ il, T *__restrict__ cm) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n * M) return;
    const size_t m = e / n, i = e - m * n;
    cm[e] = il[i * M + m];
}

// z = r * inverse diagonal for M columns (preconditioner.rs:237-247).
static __global__ void jacobi_batch_kernel(size_t n, uint32_t M, const double *__restrict__ inv, const double *__restrict__ r,
                                           double *__restrict__ z) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n * M) z[e] = r[e] * inv[e / M];
}

}  // namespace dev
}  // namespace sf
