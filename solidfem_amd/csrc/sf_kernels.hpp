// Device kernels of the solve path for gfx950 (CDNA4, wave64).
//
// Sliced block layout ("node slices"). A block sparse matrix of RB x CB blocks is cut into
// slices of NPW = 64 / RB block rows (21 nodes of 3 rows, 10 of 6, 64 scalar rows), one
// wavefront per slice and one lane per scalar row. Within a slice, the k-th block of every row
// is stored side by side: block columns at cols[(start + k) * NPW + j] and values at
// vals[((start + k) * 64 + j * RB + p) * CB + q] for row j, block row p, block column q. A lane's
// CB values of a block are contiguous, one vector load, and a wavefront's loads of one slot cover
// 64 * CB consecutive values; the RB lanes of a row read the same column index (one fetch), and
// every lane sums its own row's blocks in stored order, column by column, from zero: the
// summation order of the reference's block product
// (crates/solidfem-solve/src/multigrid/block.rs:136-154), so the result is the same bits.
// BATCH blocks are loaded before their terms are added, in order, so BATCH loads are in flight
// per lane whatever the latency; the batch depth changes no bit and is chosen per matrix by timing.
// Rows are permuted inside windows so that a slice's rows have similar block counts; perm[]
// maps a slice position back to the block row.
//
// Every kernel is compiled with -ffp-contract=off: no multiply-add is fused, as the reference
// refuses mul_add (clippy.toml:8-9), so the device repeats the reference's roundings.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace sf {
namespace dev {

constexpr int WAVE = 64;
constexpr uint32_t NONE = 0xffffffffu;

template <int RB>
struct Shape {
    static constexpr int NPW = WAVE / RB;  // block rows per slice
    static constexpr int LANES = NPW * RB; // lanes holding a scalar row
};

template <typename T>
struct Sliced {
    const uint32_t *slice_start;  // slices + 1, in block slots per row
    const uint32_t *row_len;      // per slice position: blocks in the row (0 for padding)
    const uint32_t *perm;         // per slice position: block row, NONE for padding
    const uint32_t *cols;         // per block slot and slice row
    const T *vals;                // per block slot, block column and lane
    uint32_t slices;
    uint32_t order;               // workgroup -> slice order: ORDER_* bits (slice_of)
};

// Workgroup-to-slice orders; none changes a bit of any result, only which workgroups run
// together and where. XCD_RUNS gives each XCD one contiguous run of slices; XCD_WINDOWS gives
// each XCD whole sort windows (8 workgroups of 4 slices), so the rows a window permutes share one
// L2; REVERSE walks the slices from the end (alternate passes over a matrix then start where the
// last one ended, in the cache).
enum : uint32_t { ORDER_XCD_RUNS = 1, ORDER_XCD_WINDOWS = 2, ORDER_REVERSE = 4, ORDER_GROUP_RUNS = 8, ORDER_GROUP_CHUNKS = 16 };
// With ORDER_GROUP_CHUNKS, bits 24-31 of an order hold log2 of the chunk (workgroups per XCD).
constexpr uint32_t ORDER_CHUNK_SHIFT = 24;

template <typename V, typename S>
__device__ __forceinline__ V widen(S s) {
    return (V)(double)s;
}

// Workgroups are dispatched to the 8 XCDs in turn (workgroup b to XCD b mod 8). The renumberings
// below are bijections on [0, gridDim.x).
__device__ __forceinline__ uint32_t xcd_runs_block() {
    constexpr uint32_t XCDS = 8;
    const uint32_t nb = gridDim.x, b = blockIdx.x;
    const uint32_t x = b % XCDS, local = b / XCDS, base = nb / XCDS, rem = nb % XCDS;
    return x * base + (x < rem ? x : rem) + local;
}
__device__ __forceinline__ uint32_t xcd_windows_block() {
    // within each group of 64 workgroups, XCD x takes logical workgroups 8x .. 8x + 7
    const uint32_t nb = gridDim.x, b = blockIdx.x;
    if (b >= nb - nb % 64) return b;
    const uint32_t r = b % 64;
    return b - r + (r % 8) * 8 + r / 8;
}
// Within each group of 8 W workgroups, XCD x takes the W consecutive logical workgroups
// x W .. x W + W - 1 (W = 2^lw); the tail past the last whole group keeps its order.
__device__ __forceinline__ uint32_t xcd_chunks_block(uint32_t lw) {
    const uint32_t nb = gridDim.x, b = blockIdx.x, group = 8u << lw;
    if (b >= nb - nb % group) return b;
    const uint32_t r = b & (group - 1u);
    return b - r + (r & 7u) * (1u << lw) + (r >> 3);
}
// The slice a wavefront works on under `order` (4 wavefronts of 64 lanes per workgroup).
__device__ __forceinline__ uint32_t slice_of(uint32_t order) {
    const uint32_t block = (order & ORDER_XCD_RUNS) ? xcd_runs_block() : (order & ORDER_XCD_WINDOWS) ? xcd_windows_block() : blockIdx.x;
    const uint32_t waves = blockDim.x / WAVE;
    uint32_t slice = block * waves + threadIdx.x / WAVE;
    if (order & ORDER_REVERSE) slice = gridDim.x * waves - 1 - slice;
    return slice;
}

// Sum of one lane's row of A x, from zero in stored order: the row product of the reference.
// The column indices of the next batch are loaded with the current batch's values, so each
// batch costs one dependent memory round trip.
template <int RB, int CB, int BATCH, typename S, typename V>
__device__ __forceinline__ V row_product(const Sliced<S> &a, uint32_t slot0, uint32_t len, int j, int lane,
                                         const V *__restrict__ x) {
    constexpr int NPW = Shape<RB>::NPW;
    V sum = (V)0;
    const uint32_t *cp = a.cols + (size_t)slot0 * NPW + j;
    const S *vp = a.vals + ((size_t)slot0 * WAVE + lane) * CB;
    const uint32_t full = len - len % BATCH;
    uint32_t c[BATCH];
    if (full) {
#pragma unroll
        for (int b = 0; b < BATCH; ++b) c[b] = cp[(size_t)b * NPW];
    }
    for (uint32_t k = 0; k < full; k += BATCH) {
        S v[BATCH][CB];
        V xv[BATCH][CB];
        uint32_t cn[BATCH];
#pragma unroll
        for (int b = 0; b < BATCH; ++b)
#pragma unroll
            for (int q = 0; q < CB; ++q) v[b][q] = vp[(size_t)(k + b) * WAVE * CB + q];
#pragma unroll
        for (int b = 0; b < BATCH; ++b)
#pragma unroll
            for (int q = 0; q < CB; ++q) xv[b][q] = x[(size_t)c[b] * CB + q];
        if (k + BATCH < full) {
#pragma unroll
            for (int b = 0; b < BATCH; ++b) cn[b] = cp[(size_t)(k + BATCH + b) * NPW];
        }
#pragma unroll
        for (int b = 0; b < BATCH; ++b)
#pragma unroll
            for (int q = 0; q < CB; ++q) sum = sum + widen<V>(v[b][q]) * xv[b][q];
#pragma unroll
        for (int b = 0; b < BATCH; ++b) c[b] = cn[b];
    }
    for (uint32_t k = full; k < len; ++k) {
        const uint32_t col = cp[(size_t)k * NPW];
#pragma unroll
        for (int q = 0; q < CB; ++q) sum = sum + widen<V>(vp[(size_t)k * WAVE * CB + q]) * x[(size_t)col * CB + q];
    }
    return sum;
}

enum Epilogue { EPI_STORE = 0, EPI_RESIDUAL = 1, EPI_ADD = 2 };

// The Chebyshev start (cycle.rs:591-610) on a vector a kernel has just produced, so that the
// smoothing that follows needs no launch of its own: z = D^-1 y summed over the block row in
// order from zero, d = z / theta, x = d. A null `dinv` means no start.
template <typename S, typename V>
struct Start {
    const S *dinv;
    V theta;
    V *d;
    V *x;
};

// y = A x (STORE), y = b - A x (RESIDUAL) or y = y + A x (ADD): the block products, the
// residual (crates/solidfem-solve/src/multigrid/cycle.rs:548-564), the gather restriction over
// the stored transpose (crates/solidfem-solve/src/gpu/layout.rs:249-291) and the prolongation
// with correction (cycle.rs:512-521); optionally followed by the Chebyshev start on y, the RB
// values of a block row shuffled in from its lanes.
template <int RB, int CB, int BATCH, typename S, typename V, int EPI>
__global__ __launch_bounds__(256) void spmv_kernel(Sliced<S> a, const V *__restrict__ x, V *__restrict__ y,
                                                   const V *__restrict__ b, Start<S, V> start) {
    constexpr int NPW = Shape<RB>::NPW;
    const uint32_t slice = slice_of(a.order);
    if (slice >= a.slices) return;  // whole wave leaves together
    const int lane = threadIdx.x % WAVE;
    const int j = lane / RB, p = lane - j * RB;
    const bool in_wave = j < NPW;
    const uint32_t pos = slice * NPW + (in_wave ? j : 0);
    const uint32_t row = in_wave ? a.perm[pos] : NONE;
    const bool live = row != NONE;
    const size_t out = live ? (size_t)row * RB + p : 0;
    V value = (V)0;
    if (live) {
        const V sum = row_product<RB, CB, BATCH, S, V>(a, a.slice_start[slice], a.row_len[pos], j, lane, x);
        if (EPI == EPI_STORE) value = sum;
        else if (EPI == EPI_RESIDUAL) value = b[out] - sum;
        else value = y[out] + sum;
        y[out] = value;
    }
    if (start.dinv) {  // uniform: every lane takes part in the shuffle
        const S *drow = start.dinv + (live ? (size_t)row * RB * RB + (size_t)p * RB : 0);
        V z = (V)0;
#pragma unroll
        for (int q = 0; q < RB; ++q) {
            const V yq = __shfl(value, (in_wave ? j * RB : 0) + q, WAVE);
            if (live) z = z + widen<V>(drow[q]) * yq;
        }
        if (live) {
            const V dv = z / start.theta;
            start.d[out] = dv;
            start.x[out] = dv;
        }
    }
}

// A block matrix in plain block rows: row r's blocks at [starts[r], starts[r + 1]), their
// row-major RB x CB values contiguous. Used for matrices with few, long rows.
template <typename T>
struct Bsr {
    const uint32_t *starts;  // rows + 1
    const uint32_t *cols;
    const T *vals;
    uint32_t rows;
    uint32_t reverse;  // nonzero: each XCD takes its workgroups' rows in reverse order (xcd_reversed_block)
};

// Workgroup b runs on XCD b mod 8. Each XCD takes the same rows as in dispatch order, in reverse:
// block x + 8 l becomes x + 8 (c - 1 - l), c the XCD's count of blocks. A pass that follows a
// forward pass over the same matrix then starts with the rows that XCD's L2 holds last.
__device__ __forceinline__ uint32_t xcd_reversed_block() {
    const uint32_t nb = gridDim.x, b = blockIdx.x, x = b % 8u, l = b / 8u;
    const uint32_t count = (nb - x + 7u) / 8u;
    return x + 8u * (count - 1u - l);
}

// The longest row among a wavefront's RPW rows (each row's length is uniform over its lanes).
template <int RPW>
__device__ __forceinline__ uint32_t longest_of_wave(uint32_t len) {
#pragma unroll
    for (int o = WAVE / RPW; o < WAVE; o <<= 1) {
        const uint32_t other = (uint32_t)__shfl_xor((int)len, o, WAVE);
        len = other > len ? other : len;
    }
    return len;
}

// Adds `n` terms in order onto `sum`, eight loads ahead.
template <typename V>
__device__ __forceinline__ V add_terms(V sum, const V *t, uint32_t n) {
    uint32_t i = 0;
    for (; i + 8 <= n; i += 8) {
        V v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = t[i + u];
#pragma unroll
        for (int u = 0; u < 8; ++u) sum = sum + v[u];
    }
    for (; i < n; ++i) sum = sum + t[i];
    return sum;
}

// y = A x with an epilogue (as spmv_kernel) for matrices with few, long rows -- the stored
// transposes of the prolongators, whose rows gather every fine row of an aggregate's
// neighbourhood, and the coarse levels. One wavefront (one workgroup) takes RPW block rows,
// 64 / RPW lanes each: the lanes load that many blocks at once and form their terms A[p][q] x[q]
// into LDS; then lane p of the row's group adds the row's terms in stored block order and column
// order, from zero -- the summation of spmv_kernel and of the reference, so the same bits, with
// one memory round trip per 64 / RPW blocks instead of one per few blocks.
template <int RB, int CB, typename S, typename V, int EPI, int RPW>
__global__ __launch_bounds__(64) void long_rows_kernel(Bsr<S> a, const V *__restrict__ x, V *__restrict__ y,
                                                       const V *__restrict__ b, Start<S, V> then) {
    constexpr int LW = WAVE / RPW;
    __shared__ V terms[RPW][RB][LW * CB];
    __shared__ V values[RPW][RB];
    const int lane = threadIdx.x, g = lane / LW, sl = lane % LW;
    const uint32_t row = (a.reverse ? xcd_reversed_block() : blockIdx.x) * RPW + (uint32_t)g;
    const bool live = row < a.rows;
    const uint32_t start = live ? a.starts[row] : 0u, len = live ? a.starts[row + 1] - start : 0u;
    const uint32_t longest = longest_of_wave<RPW>(len);
    V sum = (V)0;
    for (uint32_t k0 = 0; k0 < longest; k0 += LW) {
        const uint32_t k = k0 + sl;
        if (k < len) {
            const uint32_t col = a.cols[start + k];
            const S *blk = a.vals + (size_t)(start + k) * RB * CB;
            V xv[CB];
#pragma unroll
            for (int q = 0; q < CB; ++q) xv[q] = x[(size_t)col * CB + q];
#pragma unroll
            for (int p = 0; p < RB; ++p)
#pragma unroll
                for (int q = 0; q < CB; ++q) terms[g][p][sl * CB + q] = widen<V>(blk[p * CB + q]) * xv[q];
        }
        __syncthreads();
        if (sl < RB && k0 < len) sum = add_terms(sum, terms[g][sl], (len - k0 < (uint32_t)LW ? len - k0 : (uint32_t)LW) * CB);
        __syncthreads();
    }
    const bool owner = live && sl < RB;
    const size_t out = (size_t)row * RB + (owner ? sl : 0);
    if (owner) {
        V value;
        if (EPI == EPI_STORE) value = sum;
        else if (EPI == EPI_RESIDUAL) value = b[out] - sum;
        else value = y[out] + sum;
        y[out] = value;
        values[g][sl] = value;
    }
    if (then.dinv) {  // uniform
        __syncthreads();
        if (owner) {
            const S *drow = then.dinv + (size_t)row * RB * RB + (size_t)sl * RB;
            V z = (V)0;
#pragma unroll
            for (int q = 0; q < RB; ++q) z = z + widen<V>(drow[q]) * values[g][q];
            const V dv = z / then.theta;
            then.d[out] = dv;
            then.x[out] = dv;
        }
    }
}

// The fused Chebyshev step of cheb_step_kernel for levels with few rows, RPW block rows per
// wavefront (one workgroup) as long_rows_kernel: the lanes form the terms of A d_in, lane p of a
// row's group adds the row's terms in stored order, and the RB residuals of the block row meet in
// LDS for the block-Jacobi product (summed over q in order from zero). The same bits as
// cheb_step_kernel.
template <int RB, typename S, typename V, int RPW>
__global__ __launch_bounds__(64) void cheb_rows_kernel(Bsr<S> a, const S *__restrict__ dinv, const V *__restrict__ d_in,
                                                       const V *__restrict__ r_in, V keep, V step, V *__restrict__ d_out,
                                                       V *__restrict__ r_out, V *__restrict__ x, V *__restrict__ sum_into,
                                                       double *__restrict__ widen_into) {
    constexpr int LW = WAVE / RPW;
    __shared__ V terms[RPW][RB][LW * RB];
    __shared__ V r_new[RPW][RB];
    const int lane = threadIdx.x, g = lane / LW, sl = lane % LW;
    const uint32_t row = (a.reverse ? xcd_reversed_block() : blockIdx.x) * RPW + (uint32_t)g;
    const bool live = row < a.rows;
    const uint32_t start = live ? a.starts[row] : 0u, len = live ? a.starts[row + 1] - start : 0u;
    const uint32_t longest = longest_of_wave<RPW>(len);
    V sum = (V)0;
    for (uint32_t k0 = 0; k0 < longest; k0 += LW) {
        const uint32_t k = k0 + sl;
        if (k < len) {
            const uint32_t col = a.cols[start + k];
            const S *blk = a.vals + (size_t)(start + k) * RB * RB;
            V xv[RB];
#pragma unroll
            for (int q = 0; q < RB; ++q) xv[q] = d_in[(size_t)col * RB + q];
#pragma unroll
            for (int p = 0; p < RB; ++p)
#pragma unroll
                for (int q = 0; q < RB; ++q) terms[g][p][sl * RB + q] = widen<V>(blk[p * RB + q]) * xv[q];
        }
        __syncthreads();
        if (sl < RB && k0 < len) sum = add_terms(sum, terms[g][sl], (len - k0 < (uint32_t)LW ? len - k0 : (uint32_t)LW) * RB);
        __syncthreads();
    }
    const bool owner = live && sl < RB;
    const size_t i = (size_t)row * RB + (owner ? sl : 0);
    if (owner) r_new[g][sl] = r_in[i] - sum;
    __syncthreads();
    if (owner) {
        const S *drow = dinv + (size_t)row * RB * RB + (size_t)sl * RB;
        V z = (V)0;
#pragma unroll
        for (int q = 0; q < RB; ++q) z = z + widen<V>(drow[q]) * r_new[g][q];
        const V dn = keep * d_in[i] + step * z;
        const V xn = x[i] + dn;
        d_out[i] = dn;
        r_out[i] = r_new[g][sl];
        x[i] = xn;
        if (sum_into) {
            const V total = sum_into[i] + xn;
            sum_into[i] = total;
            if (widen_into) widen_into[i] = (double)total;
        }
    }
}

// Per scalar row, the rounding bound term of CsrMatrix::multiply_rounding_bound
// (crates/solidfem-solve/src/csr.rs:142-154): (terms * eps) * sum |a x|, squared.
template <int RB, int CB>
__global__ __launch_bounds__(256) void rounding_bound_kernel(Sliced<double> a, const double *__restrict__ x,
                                                             double *__restrict__ out) {
    constexpr int NPW = Shape<RB>::NPW;
    const uint32_t slice = blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
    if (slice >= a.slices) return;
    const int lane = threadIdx.x % WAVE;
    const int j = lane / RB, p = lane - j * RB;
    if (j >= NPW) return;
    const uint32_t pos = slice * NPW + j;
    const uint32_t row = a.perm[pos];
    if (row == NONE) return;
    const uint32_t len = a.row_len[pos], slot0 = a.slice_start[slice];
    double sum = 0.0;
    for (uint32_t k = 0; k < len; ++k) {
        const uint32_t col = a.cols[(size_t)(slot0 + k) * NPW + j];
        for (int q = 0; q < CB; ++q) sum = sum + fabs(a.vals[((size_t)(slot0 + k) * WAVE + lane) * CB + q] * x[(size_t)col * CB + q]);
    }
    const double bound = (double)(len * CB) * 2.220446049250313e-16 * sum;
    out[(size_t)row * RB + p] = bound * bound;
}

// The Chebyshev start (cycle.rs:591-610): z = D^-1 b, d = z / theta, x = d.
template <int RB, typename S, typename V>
__global__ __launch_bounds__(256) void cheb_start_kernel(uint32_t rows, const S *__restrict__ dinv, const V *__restrict__ b,
                                                         V theta, V *__restrict__ d, V *__restrict__ x) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)rows * RB) return;
    const size_t row = i / RB, p = i - row * RB;
    const S *dr = dinv + row * RB * RB + p * RB;
    V z = (V)0;
#pragma unroll
    for (int q = 0; q < RB; ++q) z = z + widen<V>(dr[q]) * b[row * RB + q];
    const V dv = z / theta;
    d[i] = dv;
    x[i] = dv;
}

// One later Chebyshev step, fused (cycle.rs:611-650): ad = A d_in; r_out = r_in - ad;
// z = D^-1 r_out (the RB residuals of a block row shuffled in from its lanes, summed from zero,
// cycle.rs:655-674); d_out = keep d_in + step z; x += d_out. With `sum_into` set, the step is the
// last of a post-smoothing and also adds the smoothed correction to the cycle's x
// (cycle.rs:534-536): sum_into += x.
template <int RB, int BATCH, typename S, typename V>
__global__ __launch_bounds__(256) void cheb_step_kernel(Sliced<S> a, const S *__restrict__ dinv, const V *__restrict__ d_in,
                                                        const V *__restrict__ r_in, V keep, V step, V *__restrict__ d_out,
                                                        V *__restrict__ r_out, V *__restrict__ x, V *__restrict__ sum_into,
                                                        double *__restrict__ widen_into) {
    constexpr int NPW = Shape<RB>::NPW;
    const uint32_t slice = slice_of(a.order);
    if (slice >= a.slices) return;  // whole wave leaves together
    const int lane = threadIdx.x % WAVE;
    const int j = lane / RB, p = lane - j * RB;
    const bool in_wave = j < NPW;
    const uint32_t pos = slice * NPW + (in_wave ? j : 0);
    const uint32_t row = in_wave ? a.perm[pos] : NONE;
    const bool live = row != NONE;
    V ad = (V)0;
    if (live) ad = row_product<RB, RB, BATCH, S, V>(a, a.slice_start[slice], a.row_len[pos], j, lane, d_in);
    const size_t i = live ? (size_t)row * RB + p : 0;
    const V r_new = live ? r_in[i] - ad : (V)0;
    // every lane takes part in the shuffle; padding lanes carry zeros and store nothing
    const S *drow = dinv + (live ? (size_t)row * RB * RB + (size_t)p * RB : 0);
    V z = (V)0;
#pragma unroll
    for (int q = 0; q < RB; ++q) {
        const V rq = __shfl(r_new, (in_wave ? j * RB : 0) + q, WAVE);
        if (live) z = z + widen<V>(drow[q]) * rq;
    }
    if (!live) return;
    const V dn = keep * d_in[i] + step * z;
    const V xn = x[i] + dn;
    d_out[i] = dn;
    r_out[i] = r_new;
    x[i] = xn;
    if (sum_into) {
        const V total = sum_into[i] + xn;
        sum_into[i] = total;
        if (widen_into) widen_into[i] = (double)total;  // the cycle's correction, widened out
    }
}

// x = x + y (cycle.rs:541-545), for a post-smoothing of degree one.
template <typename V>
__global__ void add_kernel(size_t n, V *__restrict__ x, const V *__restrict__ y) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = x[i] + y[i];
}

// The finest level's right-hand side rounded in from f64 (the conversion of cycle.rs:226-243)
// and the Chebyshev start on it (cheb_start_kernel), one thread per block row.
template <int RB, typename S, typename V>
__global__ __launch_bounds__(256) void convert_start_kernel(uint32_t rows, const double *__restrict__ r, V *__restrict__ b,
                                                            const S *__restrict__ dinv, V theta, V *__restrict__ d,
                                                            V *__restrict__ x) {
    const size_t row = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= rows) return;
    V bv[RB];
#pragma unroll
    for (int q = 0; q < RB; ++q) {
        bv[q] = (V)r[row * RB + q];
        b[row * RB + q] = bv[q];
    }
#pragma unroll
    for (int p = 0; p < RB; ++p) {
        const S *dr = dinv + row * RB * RB + p * RB;
        V z = (V)0;
#pragma unroll
        for (int q = 0; q < RB; ++q) z = z + widen<V>(dr[q]) * bv[q];
        const V dv = z / theta;
        d[row * RB + p] = dv;
        x[row * RB + p] = dv;
    }
}

template <typename To, typename From>
__global__ void convert_kernel(size_t n, const From *__restrict__ in, To *__restrict__ out) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (To)in[i];
}

// The coarsest level's direct solve with the f64 Cholesky factor (block.rs:357-373), the
// right-hand side widened from V and the solution rounded to V (cycle.rs:475-486). One
// wavefront; `lt` is the factor transposed (column k of the factor contiguous) and the running
// sums live in LDS. Forward substitution runs right-looking: once x_k is final every later row
// subtracts its term, so each row still subtracts in ascending k, the reference's order. Back
// substitution also subtracts each row's terms in ascending k, which can only start once every
// later x is known: the lanes form the products, then one lane folds them in order.
constexpr uint32_t COARSE_MAX = 3072;
// Several right-hand sides (row-interleaved, `stride` columns) take one workgroup each.
template <typename V>
__global__ __launch_bounds__(64) void coarse_solve_kernel(uint32_t n, const double *__restrict__ lt,
                                                           const V *__restrict__ rhs, V *__restrict__ out, uint32_t stride) {
    __shared__ double xs[COARSE_MAX];
    __shared__ double prod[COARSE_MAX];
    const uint32_t t = threadIdx.x;
    rhs += blockIdx.x;
    out += blockIdx.x;
    for (uint32_t i = t; i < n; i += WAVE) xs[i] = (double)rhs[(size_t)i * stride];
    __syncthreads();
    for (uint32_t k = 0; k < n; ++k) {
        if (t == k % WAVE) xs[k] = xs[k] / lt[(size_t)k * n + k];
        __syncthreads();
        const double xv = xs[k];
        const double *col = lt + (size_t)k * n;
        for (uint32_t i = k + 1 + t; i < n; i += WAVE) xs[i] = xs[i] - col[i] * xv;
        __syncthreads();
    }
    for (uint32_t ii = n; ii-- > 0;) {
        const double *col = lt + (size_t)ii * n;
        for (uint32_t k = ii + 1 + t; k < n; k += WAVE) prod[k] = col[k] * xs[k];
        __syncthreads();
        if (t == 0) {
            double sum = xs[ii];
            for (uint32_t k = ii + 1; k < n; ++k) sum = sum - prod[k];
            xs[ii] = sum / col[ii];
        }
        __syncthreads();
    }
    for (uint32_t i = t; i < n; i += WAVE) out[(size_t)i * stride] = (V)xs[i];
}

// The same solve with the factor held in LDS, packed by columns (column k holds rows k..n-1 at
// offset k n - k (k - 1) / 2), so no step of either substitution waits on global memory. Used
// when the packed factor and two vectors fit in one workgroup's LDS.
__device__ __forceinline__ size_t packed_column(uint32_t n, uint32_t k) { return (size_t)k * n - ((size_t)k * (k - 1)) / 2; }

template <typename V>
__global__ __launch_bounds__(64) void coarse_solve_lds_kernel(uint32_t n, const double *__restrict__ packed,
                                                               const V *__restrict__ rhs, V *__restrict__ out, uint32_t stride) {
    extern __shared__ double smem[];
    const size_t np = (size_t)n * (n + 1) / 2;
    double *l = smem, *xs = smem + np, *prod = xs + n;
    const uint32_t t = threadIdx.x;
    rhs += blockIdx.x;
    out += blockIdx.x;
    for (size_t i = t; i < np; i += WAVE) l[i] = packed[i];
    for (uint32_t i = t; i < n; i += WAVE) xs[i] = (double)rhs[(size_t)i * stride];
    __syncthreads();
    for (uint32_t k = 0; k < n; ++k) {
        const double *col = l + packed_column(n, k) - k;  // col[i] = L[i][k]
        if (t == k % WAVE) xs[k] = xs[k] / col[k];
        __syncthreads();
        const double xv = xs[k];
        for (uint32_t i = k + 1 + t; i < n; i += WAVE) xs[i] = xs[i] - col[i] * xv;
        __syncthreads();
    }
    for (uint32_t ii = n; ii-- > 0;) {
        const double *col = l + packed_column(n, ii) - ii;
        for (uint32_t k = ii + 1 + t; k < n; k += WAVE) prod[k] = col[k] * xs[k];
        __syncthreads();
        if (t == 0) {
            double sum = xs[ii];
            uint32_t k = ii + 1;
            for (; k + 8 <= n; k += 8) {
                double p[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) p[u] = prod[k + u];
#pragma unroll
                for (int u = 0; u < 8; ++u) sum = sum - p[u];
            }
            for (; k < n; ++k) sum = sum - prod[k];
            xs[ii] = sum / col[ii];
        }
        __syncthreads();
    }
    for (uint32_t i = t; i < n; i += WAVE) out[(size_t)i * stride] = (V)xs[i];
}

// The coarsest level's solve as a product with the f64 inverse (stored by columns: inv[k n + i]
// is entry (i, k)), the right-hand side widened from V and staged through LDS, each row summed
// in ascending k from zero: a fixed order, so the same bits on every run.
// The explicit inverse of a coarse factor, L^-T L^-1. First linv = L^-1 by forward substitution
// for all columns at once, one step per row k: row k is finished (divided by L[k][k]), then every
// later row i takes the term L[i][k] linv[k][j] off its entries j <= k. Entry (i, j) thus computes
// ((i == j) - sum over k from j to i - 1 of L[i][k] linv[k][j], subtracted in ascending k) /
// L[i][i]; `lt` holds L by columns (lt[k n + i] = L[i][k]), linv starts as the identity. Then
// entry (i, j) of the inverse = sum over k from max(i, j) of linv[k][i] linv[k][j], added in
// ascending k from zero, stored at inv[j n + i] (the loads of each chunk of 16 terms are issued
// before the terms are taken, in order).
static __global__ void factor_inverse_finish_row_kernel(uint32_t n, uint32_t k, const double *__restrict__ lt, double *__restrict__ linv) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j > k) return;
    linv[(size_t)k * n + j] = linv[(size_t)k * n + j] / lt[(size_t)k * n + k];
}
static __global__ void factor_inverse_update_kernel(uint32_t n, uint32_t k, const double *__restrict__ lt, double *__restrict__ linv) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x, i = k + 1 + blockIdx.y;
    if (j > k || i >= n) return;
    linv[(size_t)i * n + j] -= lt[(size_t)k * n + i] * linv[(size_t)k * n + j];
}
static __global__ __launch_bounds__(256) void factor_inverse_entries_kernel(uint32_t n, const double *__restrict__ linv,
                                                                           double *__restrict__ inv) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)n * n) return;
    const uint32_t i = (uint32_t)(t % n), j = (uint32_t)(t / n);
    constexpr uint32_t CHUNK = 16;
    double v = 0.0;
    uint32_t k = i > j ? i : j;
    for (; k + CHUNK <= n; k += CHUNK) {
        double p[CHUNK], q[CHUNK];
#pragma unroll
        for (uint32_t u = 0; u < CHUNK; ++u) {
            p[u] = linv[(size_t)(k + u) * n + i];
            q[u] = linv[(size_t)(k + u) * n + j];
        }
#pragma unroll
        for (uint32_t u = 0; u < CHUNK; ++u) v += p[u] * q[u];
    }
    for (; k < n; ++k) v += linv[(size_t)k * n + i] * linv[(size_t)k * n + j];
    inv[(size_t)j * n + i] = v;
}

template <typename V>
__global__ __launch_bounds__(256) void coarse_inverse_kernel(uint32_t n, const double *__restrict__ inv,
                                                             const V *__restrict__ rhs, V *__restrict__ out) {
    __shared__ double bs[256];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    double sum = 0.0;
    for (uint32_t k0 = 0; k0 < n; k0 += 256) {
        if (k0 + threadIdx.x < n) bs[threadIdx.x] = (double)rhs[k0 + threadIdx.x];
        __syncthreads();
        const uint32_t kmax = n - k0 < 256u ? n - k0 : 256u;
        if (i < n) {
            // 32 loads in flight (the column is latency-bound), then their terms in ascending k
            uint32_t kk = 0;
            for (; kk + 32 <= kmax; kk += 32) {
                double a[32];
#pragma unroll
                for (int u = 0; u < 32; ++u) a[u] = inv[(size_t)(k0 + kk + u) * n + i];
#pragma unroll
                for (int u = 0; u < 32; ++u) sum = sum + a[u] * bs[kk + u];
            }
            for (; kk + 8 <= kmax; kk += 8) {
                double a[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) a[u] = inv[(size_t)(k0 + kk + u) * n + i];
#pragma unroll
                for (int u = 0; u < 8; ++u) sum = sum + a[u] * bs[kk + u];
            }
            for (; kk < kmax; ++kk) sum = sum + inv[(size_t)(k0 + kk) * n + i] * bs[kk];
        }
        __syncthreads();
    }
    if (i < n) out[i] = (V)sum;
}

// The same product for a coarse level of at most COARSE_SMALL rows (the plate's has 102): one
// workgroup brings the whole inverse (at most 128 KB) and the right-hand side into LDS with every
// load in flight at once, then thread i adds its row's terms in ascending k from zero.
constexpr uint32_t COARSE_SMALL = 128;
constexpr uint32_t COARSE_SMALL_THREADS = 1024;
constexpr size_t coarse_small_lds(uint32_t n) { return ((size_t)n * n + n) * sizeof(double); }
template <typename V>
__global__ __launch_bounds__(COARSE_SMALL_THREADS) void coarse_inverse_small_kernel(uint32_t n, const double *__restrict__ inv,
                                                                                     const V *__restrict__ rhs, V *__restrict__ out) {
    extern __shared__ double ls[];  // the n x n inverse, then the n right-hand-side entries
    constexpr int PER = COARSE_SMALL * COARSE_SMALL / COARSE_SMALL_THREADS;
    const uint32_t t = threadIdx.x, nn = n * n;
    double a[PER];
#pragma unroll
    for (int j = 0; j < PER; ++j) {
        const uint32_t e = t + COARSE_SMALL_THREADS * j;
        a[j] = e < nn ? inv[e] : 0.0;
    }
#pragma unroll
    for (int j = 0; j < PER; ++j) {
        const uint32_t e = t + COARSE_SMALL_THREADS * j;
        if (e < nn) ls[e] = a[j];
    }
    if (t < n) ls[nn + t] = (double)rhs[t];
    __syncthreads();
    if (t < n) {
        double sum = 0.0;
        for (uint32_t k = 0; k < n; ++k) sum = sum + ls[(size_t)k * n + t] * ls[nn + k];
        out[t] = (V)sum;
    }
}

// ---- conjugate gradient vectors -------------------------------------------------------------

// Scalars of a solve, resident on the device.
struct PcgScalars {
    double rz;         // r . z
    double curvature;  // p . A p
    double rr;         // r . r (recursive residual)
    double rz_next;
    double true_rr;    // ||b - A x||^2
    double bound2;     // sum of squared row rounding bounds
    double alpha;
    double beta;
    int32_t valid;     // the curvature test of pcg.rs:219-227 passed
    uint32_t tag;      // host copies only: which readback the copy answers, written last
};

// Copies the scalars to host memory for the host to read without an event: every field first,
// then, after a system-scope fence, the tag the host waits for.
__device__ __forceinline__ void mirror_scalars(const PcgScalars *s, PcgScalars *mirror, uint32_t tag) {
    mirror->rz = s->rz;
    mirror->curvature = s->curvature;
    mirror->rr = s->rr;
    mirror->rz_next = s->rz_next;
    mirror->true_rr = s->true_rr;
    mirror->bound2 = s->bound2;
    mirror->alpha = s->alpha;
    mirror->beta = s->beta;
    mirror->valid = s->valid;
    __threadfence_system();
    __hip_atomic_store(&mirror->tag, tag, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Deterministic dot products. TREE order is a fixed two-stage reduction: DOT_BLOCKS blocks each
// sum a fixed grid-stride pattern through a fixed shuffle tree into one partial, then one block
// sums the partials in the same fixed pattern (a second kernel: no atomics and no fences; a
// one-launch form whose last block finishes the sum measured slower, its per-block release fences
// costing more than the launch they save).
// REFERENCE order is one thread summing in index order from zero, exactly
// crates/solidfem-solve/src/pcg.rs:285-291.
constexpr int DOT_BLOCKS = 512;
constexpr int DOT_THREADS = 256;

__device__ __forceinline__ double wave_sum(double v) {
    // fixed butterfly: the same pairs add in the same order on every run
#pragma unroll
    for (int off = WAVE / 2; off > 0; off >>= 1) v = v + __shfl_xor(v, off, WAVE);
    return v;
}

__device__ __forceinline__ double block_sum(double v) {
    __shared__ double wsum[DOT_THREADS / WAVE];
    v = wave_sum(v);
    if ((threadIdx.x % WAVE) == 0) wsum[threadIdx.x / WAVE] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < DOT_THREADS / WAVE; ++w) s = s + wsum[w];
    return s;  // valid in thread 0
}

// MODE 0: a . b; MODE 1: sum of a.
template <int MODE>
__global__ __launch_bounds__(DOT_THREADS) void dot_partials_kernel(size_t n, const double *__restrict__ a,
                                                                    const double *__restrict__ b, double *__restrict__ partials) {
    double v = 0.0;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) v = v + (MODE == 0 ? a[i] * b[i] : a[i]);
    const double s = block_sum(v);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

// What the final sum feeds: STORE writes it to `out`; RZ makes it the next r.z and forms
// beta = rz_next / rz, rz = rz_next (pcg.rs:258-260).
enum DotFinal { FINAL_STORE = 0, FINAL_RZ = 1 };

__device__ __forceinline__ void dot_finish(int kind, double total, double *out, PcgScalars *s) {
    if (kind == FINAL_STORE) {
        *out = total;
    } else {
        s->rz_next = total;
        s->beta = total / s->rz;
        s->rz = total;
    }
}

// `mirror`, when given, is host memory that receives a copy of the scalars once the sum is stored,
// tagged `tag`: the host waits for the tag instead of an event or a separate copy.
static __global__ __launch_bounds__(DOT_THREADS) void dot_final_kernel(int parts, const double *__restrict__ partials, int kind,
                                                                 double *out, PcgScalars *s, PcgScalars *mirror, uint32_t tag) {
    double v = 0.0;
    for (int i = threadIdx.x; i < parts; i += blockDim.x) v = v + partials[i];
    const double total = block_sum(v);
    if (threadIdx.x == 0) {
        dot_finish(kind, total, out, s);
        if (mirror) mirror_scalars(s, mirror, tag);
    }
}

template <int MODE>
__global__ void dot_serial_kernel(size_t n, const double *__restrict__ a, const double *__restrict__ b, int kind,
                                  double *out, PcgScalars *s, PcgScalars *mirror, uint32_t tag) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double sum = 0.0;
    for (size_t i = 0; i < n; ++i) sum = sum + (MODE == 0 ? a[i] * b[i] : a[i]);
    dot_finish(kind, sum, out, s);
    if (mirror) mirror_scalars(s, mirror, tag);
}

// The update of pcg.rs:219-230: the curvature test, then x += alpha p and r += (-alpha) A p.
static __global__ void pcg_update_kernel(size_t n, PcgScalars *__restrict__ s, const double *__restrict__ p,
                                  const double *__restrict__ ap, double *__restrict__ x, double *__restrict__ r) {
    const double curvature = s->curvature, rz = s->rz;
    const bool valid = !(isnan(curvature) || curvature <= 0.0 || !isfinite(rz));
    const double alpha = rz / curvature;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        s->valid = valid ? 1 : 0;
        s->alpha = alpha;
    }
    if (!valid || i >= n) return;
    x[i] = x[i] + alpha * p[i];
    r[i] = r[i] + (-alpha) * ap[i];
}

// The total of a TREE dot product formed by every block of a consumer kernel of DOT_THREADS
// threads, exactly as dot_final_kernel forms it (the same partials summed in the same order):
// the consumer needs no separate final launch and the bits do not change.
__device__ __forceinline__ double tree_total(const double *__restrict__ partials) {
    double v = 0.0;
    for (int i = threadIdx.x; i < DOT_BLOCKS; i += blockDim.x) v = v + partials[i];
    const double total = block_sum(v);
    __shared__ double shared_total;
    if (threadIdx.x == 0) shared_total = total;
    __syncthreads();
    return shared_total;
}

// The slot of PcgScalars that holds the current r.z in the fused TREE iteration: the direction
// kernel reads it in every block while block 0 stores the next one, so the two alternate.
__device__ __forceinline__ double &rz_in(PcgScalars *s, uint32_t slot) { return slot ? s->rz_next : s->rz; }

// pcg.rs:216-230 with both reductions around it fused, for TREE order, on DOT_BLOCKS blocks of
// DOT_THREADS: p . A p summed from its partials as dot_final_kernel does, the curvature test,
// x += alpha p and r += (-alpha) A p, and the partials of the new r . r over the same grid-stride
// rows as dot_partials_kernel -- the same bits as the separate launches.
static __global__ __launch_bounds__(DOT_THREADS) void pcg_update_fused_kernel(size_t n, PcgScalars *__restrict__ s, uint32_t slot,
                                                                              const double *__restrict__ pap_partials,
                                                                              const double *__restrict__ p,
                                                                              const double *__restrict__ ap, double *__restrict__ x,
                                                                              double *__restrict__ r, double *__restrict__ rr_partials) {
    const double curvature = tree_total(pap_partials);
    const double rz = rz_in(s, slot);
    const bool valid = !(isnan(curvature) || curvature <= 0.0 || !isfinite(rz));
    const double alpha = rz / curvature;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        s->curvature = curvature;
        s->valid = valid ? 1 : 0;
        s->alpha = alpha;
    }
    double v = 0.0;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        double ri = r[i];
        if (valid) {
            x[i] = x[i] + alpha * p[i];
            ri = ri + (-alpha) * ap[i];
            r[i] = ri;
        }
        v = v + ri * ri;
    }
    __syncthreads();  // block_sum's shared words were read by tree_total
    const double part = block_sum(v);
    if (threadIdx.x == 0) rr_partials[blockIdx.x] = part;
}

// pcg.rs:256-263 fused for TREE order, on DOT_BLOCKS blocks of DOT_THREADS: r . z summed from its
// partials as dot_final_kernel does, beta = rz_next / rz, then p = z + beta p. Block 0 stores the
// next r . z in the other slot.
static __global__ __launch_bounds__(DOT_THREADS) void pcg_direction_fused_kernel(size_t n, PcgScalars *__restrict__ s,
                                                                                 uint32_t slot,
                                                                                 const double *__restrict__ rz_partials,
                                                                                 const double *__restrict__ z,
                                                                                 double *__restrict__ p) {
    const double total = tree_total(rz_partials);
    const double beta = total / rz_in(s, slot);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        rz_in(s, slot ^ 1u) = total;
        s->beta = beta;
    }
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = z[i] + beta * p[i];
}

// p = z + beta p (pcg.rs:261-263), beta formed by the r.z reduction.
static __global__ void pcg_direction_kernel(size_t n, const PcgScalars *__restrict__ s, const double *__restrict__ z,
                                            double *__restrict__ p) {
    const double beta = s->beta;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = z[i] + beta * p[i];
}

// Jacobi: z = r * inverse diagonal (preconditioner.rs:237-247).
static __global__ void jacobi_kernel(size_t n, const double *__restrict__ inv, const double *__restrict__ r, double *__restrict__ z) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) z[i] = r[i] * inv[i];
}

// Sliced values from a row-major block array: slot i of the sliced layout takes value
// src[map[i]] rounded to T (the precision conversion of block.rs:26-35).
template <typename T>
__global__ void gather_convert_kernel(size_t n, const uint32_t *__restrict__ block_of_slot, int rb, int cb,
                                      const double *__restrict__ src, T *__restrict__ dst) {
    // dst index = ((slot * 64 + lane) * cb + q), lane = j * rb + p ; map gives the source block
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const size_t q = i % cb;
    const size_t lane = (i / cb) % WAVE;
    const size_t slot_row = i / ((size_t)WAVE * cb);  // the slot
    const int npw = WAVE / rb;
    const int j = (int)(lane / rb), p = (int)(lane % rb);
    if (j >= npw) {
        dst[i] = (T)0;
        return;
    }
    const uint32_t blk = block_of_slot[slot_row * npw + j];
    dst[i] = blk == NONE ? (T)0 : (T)src[(size_t)blk * rb * cb + (size_t)p * cb + q];
}

// A plain copy, 16 bytes per lane per step, grid-stride.
static __global__ void stream_copy_kernel(size_t n16, const uint4 *__restrict__ src, uint4 *__restrict__ dst) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride) dst[i] = src[i];
}

// A plain read, 16 bytes per lane per step, grid-stride; one word per lane is written so the
// loads stay.
static __global__ void stream_read_kernel(size_t n16, const uint4 *__restrict__ src, uint32_t *__restrict__ sink) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t acc = 0;
    for (size_t i = t; i < n16; i += stride) {
        const uint4 v = src[i];
        acc ^= v.x ^ v.y ^ v.z ^ v.w;
    }
    sink[t] = acc;
}

}  // namespace dev
}  // namespace sf
