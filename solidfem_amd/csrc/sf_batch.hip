// Several load cases over one matrix and one hierarchy: the batched block product and the
// lockstep conjugate gradient whose every column repeats the single-column solve's bits.
#include <algorithm>
#include <cmath>
#include <cstddef>

#include "sf_batch_kernels.hpp"
#include "sf_device.hpp"

namespace sf {

using namespace dev;

static inline unsigned grid_for(size_t n, unsigned block) { return (unsigned)((n + block - 1) / block); }

template <typename S>
bool staged_candidate(const DeviceSliced<S> &a, uint32_t M, size_t vector_size) {
    const bool long_rows = a.rows > 0 && a.blocks >= 40 * a.rows;
    const bool forced = forced_variant().load(std::memory_order_relaxed) == STAGED_FORM;
    return (long_rows || forced) && (size_t)a.rb * M <= 256 && (size_t)STAGE_BLOCKS * a.cb * M * vector_size <= 48 * 1024;
}
template bool staged_candidate<double>(const DeviceSliced<double> &, uint32_t, size_t);
template bool staged_candidate<float>(const DeviceSliced<float> &, uint32_t, size_t);

std::vector<int> spmm_candidates(int rb, int cb, size_t value_size, uint32_t M, bool product, bool staged) {
    // Depths 1, 2, 4 of the lane-group kernels, NODE_FORM + depth for one wavefront per block
    // row, and the forms with two or four columns per lane where M allows. A lane holds BATCH
    // blocks of rb * cb values and BATCH * cb * CPL entries of x: keep that within registers.
    const bool large = (size_t)rb * cb * value_size > 144;
    std::vector<int> c = large ? std::vector<int>{1, 2, NODE_FORM + 1, NODE_FORM + 2}
                               : std::vector<int>{1, 2, 4, NODE_FORM + 1, NODE_FORM + 2, NODE_FORM + 4};
    if (!large && product) c.push_back(8);  // products only: the Chebyshev step has no depth 8
    if (M % 2 == 0 && M >= 4) {
        c.push_back(PAIR_FORM + 1);
        c.push_back(PAIR_FORM + 2);
    }
    if (M % 4 == 0 && M >= 8) {
        c.push_back(QUAD_FORM + 1);
        if (!large) c.push_back(QUAD_FORM + 2);
    }
    if (M % 8 == 0 && M >= 16 && !large) {
        c.push_back(OCT_FORM + 1);
        c.push_back(OCT_FORM + 2);
    }
    if (staged && product) c.push_back(STAGED_FORM);
    return c;
}

template <int RB, int CB, int BATCH, int CPL, typename S, typename V>
static void spmm_launch(const DeviceSliced<S> &a, uint32_t M, const V *x, V *y, const V *b, int epi, hipStream_t s, bool node) {
    if (node) {
        const size_t threads = (size_t)a.slices * Shape<RB>::NPW * WAVE;
        hipLaunchKernelGGL((spmm_node_kernel<RB, CB, BATCH, S, V>), dim3(grid_for(threads, 256)), dim3(256), 0, s, a.batch_view(), M, epi,
                           x, y, b);
        return;
    }
    const int lg = columns_log2((M + CPL - 1) / CPL);
    const size_t threads = (size_t)a.slices * Shape<RB>::NPW << lg;
    hipLaunchKernelGGL((spmm_kernel<RB, CB, BATCH, S, V, CPL>), dim3(grid_for(threads, 256)), dim3(256), 0, s, a.batch_view(), lg, M, epi,
                       x, y, b);
}

template <int RB, int CB, int CPL, typename S, typename V>
static void spmm_depth(const DeviceSliced<S> &a, uint32_t M, const V *x, V *y, const V *b, int epi, hipStream_t s, int depth, bool node) {
    switch (depth) {
        case 1: spmm_launch<RB, CB, 1, CPL>(a, M, x, y, b, epi, s, node); break;
        case 2: spmm_launch<RB, CB, 2, CPL>(a, M, x, y, b, epi, s, node); break;
        case 8:
            // eight blocks in flight: for one column per lane (long rows, few of them)
            if constexpr (CPL == 1) {
                spmm_launch<RB, CB, 8, CPL>(a, M, x, y, b, epi, s, node);
                break;
            }
            [[fallthrough]];
        default: spmm_launch<RB, CB, 4, CPL>(a, M, x, y, b, epi, s, node); break;
    }
}

template <int RB, int CB, typename S, typename V>
static void spmm_shape(const DeviceSliced<S> &a, uint32_t M, const V *x, V *y, const V *b, int epi, hipStream_t s, int batch) {
    const int depth = form_depth(batch);
    int cpl = form_columns(batch);
    if (M % (uint32_t)cpl) cpl = 1;
    if (cpl == 8) spmm_depth<RB, CB, 8>(a, M, x, y, b, epi, s, depth, false);
    else if (cpl == 4) spmm_depth<RB, CB, 4>(a, M, x, y, b, epi, s, depth, false);
    else if (cpl == 2) spmm_depth<RB, CB, 2>(a, M, x, y, b, epi, s, depth, false);
    else spmm_depth<RB, CB, 1>(a, M, x, y, b, epi, s, depth, node_form(batch));
}

template <int RB, int CB, typename S, typename V>
static void spmm_staged(const DeviceSliced<S> &a, uint32_t M, const V *x, V *y, const V *b, int epi, hipStream_t s) {
    const size_t lds = (size_t)STAGE_BLOCKS * CB * M * sizeof(V);
    hipLaunchKernelGGL((spmm_staged_kernel<RB, CB, S, V>), dim3((unsigned)((size_t)a.slices * Shape<RB>::NPW)), dim3(256), lds, s,
                       a.batch_view(), M, epi, x, y, b);
}

template <typename S, typename V>
void launch_spmm(const DeviceSliced<S> &a, uint32_t M, const V *x, V *y, const V *b, int epi, hipStream_t s, int batch) {
    if (a.slices == 0 || M == 0) return;
    if (M > MAX_COLUMNS) throw Error(SF_ERR_INVALID_ARGUMENT, "more columns than a batch carries");
    if (batch == 0) {
        int &tuned = a.batch_mm[columns_log2(M)];
        if (!tuned) {
            // a product into scratch vectors of this shape, timed for each depth
            DeviceBuffer<V> tx(std::max<size_t>(a.block_cols * a.cb * M, 1)), ty(std::max<size_t>(a.rows * a.rb * M, 1));
            SF_HIP(hipMemsetAsync(tx.ptr, 0, tx.bytes(), s));
            SF_HIP(hipMemsetAsync(ty.ptr, 0, ty.bytes(), s));
            tuned = tune_batch(s, [&](int depth) { launch_spmm<S, V>(a, M, tx.ptr, ty.ptr, ty.ptr, EPI_STORE, s, depth); },
                               spmm_candidates(a.rb, a.cb, sizeof(S), M, true, staged_candidate(a, M, sizeof(V))), "",
                               "batch product " + std::to_string(columns_log2(M)) + " " + shape_class(a) + "|" + std::to_string(sizeof(V)));
            SF_HIP(hipStreamSynchronize(s));
        }
        batch = tuned;
    }
    if (batch >= STAGED_FORM) {
        // the staged form where it fits, else one column and four blocks in flight per lane
        if (staged_candidate(a, M, sizeof(V))) {
            if (a.rb == 3 && a.cb == 3) spmm_staged<3, 3>(a, M, x, y, b, epi, s);
            else if (a.rb == 6 && a.cb == 6) spmm_staged<6, 6>(a, M, x, y, b, epi, s);
            else if (a.rb == 3 && a.cb == 6) spmm_staged<3, 6>(a, M, x, y, b, epi, s);
            else if (a.rb == 6 && a.cb == 3) spmm_staged<6, 3>(a, M, x, y, b, epi, s);
            else batch = 4;
            if (batch >= STAGED_FORM) {
                SF_HIP(hipGetLastError());
                return;
            }
        } else {
            batch = 4;
        }
    }
    if (a.rb == 3 && a.cb == 3) spmm_shape<3, 3>(a, M, x, y, b, epi, s, batch);
    else if (a.rb == 6 && a.cb == 6) spmm_shape<6, 6>(a, M, x, y, b, epi, s, batch);
    else if (a.rb == 3 && a.cb == 6) spmm_shape<3, 6>(a, M, x, y, b, epi, s, batch);
    else if (a.rb == 6 && a.cb == 3) spmm_shape<6, 3>(a, M, x, y, b, epi, s, batch);
    else if constexpr (std::is_same<S, double>::value && std::is_same<V, double>::value) {
        if (a.rb == 1 && a.cb == 1) spmm_shape<1, 1>(a, M, x, y, b, epi, s, batch);
        else throw Error(SF_ERR_INTERNAL, "unsupported block shape");
    } else {
        throw Error(SF_ERR_INTERNAL, "unsupported block shape");
    }
    SF_HIP(hipGetLastError());
}

template void launch_spmm<double, double>(const DeviceSliced<double> &, uint32_t, const double *, double *, const double *, int,
                                          hipStream_t, int);
template void launch_spmm<float, double>(const DeviceSliced<float> &, uint32_t, const double *, double *, const double *, int,
                                         hipStream_t, int);
template void launch_spmm<float, float>(const DeviceSliced<float> &, uint32_t, const float *, float *, const float *, int,
                                        hipStream_t, int);

// ---- the lockstep conjugate gradient ---------------------------------------------------------

DevicePcgBatch::DevicePcgBatch(std::shared_ptr<Context> ctx, const DeviceSystem &a, DeviceMultigrid *mg, const double *rhs, uint32_t columns,
                               const sf_pcg_options &options)
    : ctx_(std::move(ctx)), a_(a), mg_(mg), n_(a.size), m_(columns), dot_order_(options.dot_order), tolerance_(options.tolerance) {
    if (columns == 0 || columns > MAX_COLUMNS) throw Error(SF_ERR_INVALID_ARGUMENT, "a batch carries 1 to 64 right-hand sides");
    if (mg_ && mg_->dofs() != n_) throw Error(SF_ERR_INVALID_ARGUMENT, "the multigrid does not match the matrix");
    hipStream_t s = ctx_->stream;
    max_iterations_ = options.max_iterations ? options.max_iterations : 10 * (u64)n_ + 1000;
    const size_t total = std::max<size_t>(n_ * m_, 1);
    for (auto *buf : {&b_, &x_, &r_, &z_, &p_, &ap_, &tmp_, &tmp2_, &stage_}) buf->resize(total);
    scalars_.resize(m_);
    partials_.resize((size_t)DOT_BLOCKS * m_);
    active_.resize(m_);
    mask_.resize(m_);
    col_.assign(m_, Column{});
    SF_HIP(hipHostMalloc((void **)&pinned_, sizeof(PcgScalars) * m_, hipHostMallocDefault));
    SF_HIP(hipHostMalloc((void **)&pinned_flags_, sizeof(uint32_t) * m_, hipHostMallocDefault));
    SF_HIP(hipEventCreateWithFlags(&copied_, hipEventDisableTiming));
    static const bool lds_set = [] {
        const int bytes = (int)(DOT_THREADS * MAX_COLUMNS * sizeof(double));
        SF_HIP(hipFuncSetAttribute((const void *)dot_partials_batch_kernel<0>, hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
        SF_HIP(hipFuncSetAttribute((const void *)dot_partials_batch_kernel<1>, hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
        return true;
    }();
    (void)lds_set;
    if (!mg_) jacobi_.upload(jacobi_inverse(ctx_, a_), s);
    restart(rhs);
}

DevicePcgBatch::~DevicePcgBatch() {
    (void)hipStreamSynchronize(ctx_->stream);
    if (pinned_) (void)hipHostFree(pinned_);
    if (pinned_flags_) (void)hipHostFree(pinned_flags_);
    if (copied_) (void)hipEventDestroy(copied_);
}

void DevicePcgBatch::upload_flags(uint32_t *device, const std::vector<uint32_t> &flags) {
    // the pinned staging is reused: nothing may still read it
    SF_HIP(hipStreamSynchronize(ctx_->stream));
    std::copy(flags.begin(), flags.end(), pinned_flags_);
    SF_HIP(hipMemcpyAsync(device, pinned_flags_, sizeof(uint32_t) * m_, hipMemcpyHostToDevice, ctx_->stream));
    SF_HIP(hipStreamSynchronize(ctx_->stream));
}

std::vector<uint32_t> DevicePcgBatch::active_flags() const {
    std::vector<uint32_t> f(m_);
    for (uint32_t m = 0; m < m_; ++m) f[m] = col_[m].status == PcgStatus::Active ? 1u : 0u;
    return f;
}

const PcgScalars *DevicePcgBatch::read_scalars() {
    SF_HIP(hipMemcpyAsync(pinned_, scalars_.ptr, sizeof(PcgScalars) * m_, hipMemcpyDeviceToHost, ctx_->stream));
    SF_HIP(hipStreamSynchronize(ctx_->stream));
    return pinned_;
}

void DevicePcgBatch::dot(const double *a, const double *b, size_t field, int kind, const uint32_t *mask, int mode) {
    hipStream_t s = ctx_->stream;
    if (dot_order_ == SF_DOT_REFERENCE) {
        if (mode == 0)
            hipLaunchKernelGGL((dot_serial_batch_kernel<0>), dim3(1), dim3(WAVE), 0, s, n_, m_, a, b, kind, (uint32_t)field, scalars_.ptr, mask);
        else
            hipLaunchKernelGGL((dot_serial_batch_kernel<1>), dim3(1), dim3(WAVE), 0, s, n_, m_, a, b, kind, (uint32_t)field, scalars_.ptr, mask);
    } else {
        const size_t lds = (size_t)DOT_THREADS * m_ * sizeof(double);
        if (mode == 0)
            hipLaunchKernelGGL((dot_partials_batch_kernel<0>), dim3(DOT_BLOCKS), dim3(DOT_THREADS), lds, s, n_, m_, a, b, partials_.ptr);
        else
            hipLaunchKernelGGL((dot_partials_batch_kernel<1>), dim3(DOT_BLOCKS), dim3(DOT_THREADS), lds, s, n_, m_, a, b, partials_.ptr);
        hipLaunchKernelGGL(dot_final_batch_kernel, dim3(m_), dim3(DOT_THREADS), 0, s, DOT_BLOCKS, m_, (const double *)partials_.ptr,
                           kind, (uint32_t)field, scalars_.ptr, mask);
    }
    SF_HIP(hipGetLastError());
}

void DevicePcgBatch::precondition(const double *r, double *z) {
    if (mg_) {
        mg_->apply_batch(r, z, m_, ctx_->stream);
    } else {
        const size_t total = n_ * m_;
        hipLaunchKernelGGL(jacobi_batch_kernel, dim3(grid_for(total, 256)), dim3(256), 0, ctx_->stream, n_, m_, jacobi_.ptr, r, z);
    }
    SF_HIP(hipGetLastError());
}

void DevicePcgBatch::reset_direction(const uint32_t *mask) {
    // pcg.rs:277-282 for the masked columns: z = M^-1 r (every column's z is recomputed from its
    // own r, the same bits where it was already current), p = z, rz = r . z
    precondition(r_.ptr, z_.ptr);
    const size_t total = n_ * m_;
    hipLaunchKernelGGL((copy_columns_kernel<double>), dim3(grid_for(total, 256)), dim3(256), 0, ctx_->stream, n_, m_, mask,
                       (const double *)z_.ptr, p_.ptr);
    dot(r_.ptr, z_.ptr, offsetof(PcgScalars, rz), FINAL_STORE, mask);
}

void DevicePcgBatch::restart(const double *rhs) {
    // pcg.rs:135-164 per column: r = b, x = 0, then the first direction
    hipStream_t s = ctx_->stream;
    SF_HIP(hipStreamSynchronize(s));
    const size_t total = n_ * m_;
    if (rhs) {
        stage_.upload(rhs, total, s);
        hipLaunchKernelGGL((interleave_kernel<double>), dim3(grid_for(total, 256)), dim3(256), 0, s, n_, m_, (const double *)stage_.ptr, b_.ptr);
    }
    SF_HIP(hipMemcpyAsync(r_.ptr, b_.ptr, total * 8, hipMemcpyDeviceToDevice, s));
    SF_HIP(hipMemsetAsync(x_.ptr, 0, total * 8, s));
    SF_HIP(hipMemsetAsync(z_.ptr, 0, total * 8, s));
    SF_HIP(hipMemsetAsync(p_.ptr, 0, total * 8, s));
    SF_HIP(hipMemsetAsync(scalars_.ptr, 0, sizeof(PcgScalars) * m_, s));
    dot(b_.ptr, b_.ptr, offsetof(PcgScalars, rr), FINAL_STORE, nullptr);
    const PcgScalars *h = read_scalars();
    std::vector<uint32_t> live(m_, 0);
    for (uint32_t m = 0; m < m_; ++m) {
        Column &c = col_[m];
        c = Column{};
        c.rhs_norm = std::sqrt(h[m].rr);
        if (c.rhs_norm == 0.0) {
            c.status = PcgStatus::Converged;
        } else {
            c.relative = 1.0;
            live[m] = 1;
        }
    }
    upload_flags(active_.ptr, live);
    upload_flags(mask_.ptr, live);
    reset_direction(mask_.ptr);
    SF_HIP(hipStreamSynchronize(s));
}

void DevicePcgBatch::iterate() {
    // DevicePcg::iterate for every running column; a column's branch is its own.
    hipStream_t s = ctx_->stream;
    const size_t total = n_ * m_;
    const unsigned g = grid_for(total, 256);
    const int token = Profiler::get().begin(4, s, (double)a_.sliced.algorithmic_matrix_bytes() + 2.0 * total * 8);
    launch_spmm<double, double>(a_.sliced, m_, p_.ptr, ap_.ptr, nullptr, EPI_STORE, s);
    Profiler::get().end(4, token, s);
    dot(p_.ptr, ap_.ptr, offsetof(PcgScalars, curvature), FINAL_STORE, nullptr);
    hipLaunchKernelGGL(pcg_update_batch_kernel, dim3(g), dim3(256), 0, s, n_, m_, scalars_.ptr, (const uint32_t *)active_.ptr,
                       (const double *)p_.ptr, (const double *)ap_.ptr, x_.ptr, r_.ptr);
    dot(r_.ptr, r_.ptr, offsetof(PcgScalars, rr), FINAL_STORE, nullptr);
    SF_HIP(hipMemcpyAsync(pinned_, scalars_.ptr, sizeof(PcgScalars) * m_, hipMemcpyDeviceToHost, s));
    SF_HIP(hipEventRecord(copied_, s));
    precondition(r_.ptr, z_.ptr);
    dot(r_.ptr, z_.ptr, 0, FINAL_RZ, active_.ptr);
    hipLaunchKernelGGL(pcg_direction_batch_kernel, dim3(g), dim3(256), 0, s, n_, m_, (const PcgScalars *)scalars_.ptr,
                       (const uint32_t *)active_.ptr, (const double *)z_.ptr, p_.ptr);
    SF_HIP(hipGetLastError());
    SF_HIP(hipEventSynchronize(copied_));
    std::vector<PcgScalars> h(pinned_, pinned_ + m_);
    bool changed = false;
    std::vector<uint32_t> recheck(m_, 0);
    bool any_recheck = false;
    for (uint32_t m = 0; m < m_; ++m) {
        Column &c = col_[m];
        if (c.status != PcgStatus::Active) continue;
        if (!h[m].valid) {
            c.status = PcgStatus::NotPositiveDefinite;
            c.failure = Error(SF_ERR_PCG_NOT_POSITIVE_DEFINITE, "the system is not positive definite", (int64_t)c.iterations, 0,
                              h[m].curvature);
            changed = true;
            continue;
        }
        c.iterations += 1;
        c.relative = std::sqrt(h[m].rr) / c.rhs_norm;
        if (c.relative <= tolerance_) {
            recheck[m] = 1;
            any_recheck = true;
        }
    }
    if (any_recheck) {
        upload_flags(mask_.ptr, recheck);
        launch_spmm<double, double>(a_.sliced, m_, x_.ptr, tmp_.ptr, b_.ptr, EPI_RESIDUAL, s);
        dot(tmp_.ptr, tmp_.ptr, offsetof(PcgScalars, true_rr), FINAL_STORE, mask_.ptr);
        const int lg = columns_log2(m_);
        const size_t threads = (size_t)a_.sliced.slices * (a_.rb == 3 ? Shape<3>::NPW : Shape<1>::NPW) << lg;
        if (a_.rb == 3)
            hipLaunchKernelGGL((rounding_bound_batch_kernel<3, 3>), dim3(grid_for(threads, 256)), dim3(256), 0, s, a_.sliced.batch_view(), lg, m_,
                               (const double *)x_.ptr, tmp2_.ptr);
        else
            hipLaunchKernelGGL((rounding_bound_batch_kernel<1, 1>), dim3(grid_for(threads, 256)), dim3(256), 0, s, a_.sliced.batch_view(), lg, m_,
                               (const double *)x_.ptr, tmp2_.ptr);
        dot(tmp2_.ptr, tmp2_.ptr, offsetof(PcgScalars, bound2), FINAL_STORE, mask_.ptr, 1);
        const PcgScalars *t = read_scalars();
        std::vector<uint32_t> restart_cols(m_, 0);
        bool any_restart = false;
        for (uint32_t m = 0; m < m_; ++m) {
            if (!recheck[m]) continue;
            Column &c = col_[m];
            c.relative = std::sqrt(t[m].true_rr) / c.rhs_norm;
            const double floor = std::sqrt(t[m].bound2) / c.rhs_norm;
            c.floor = c.relative > tolerance_ && c.relative <= floor;
            if (c.relative <= tolerance_ || c.floor) {
                c.status = PcgStatus::Converged;
                changed = true;
            } else {
                restart_cols[m] = 1;
                any_restart = true;
            }
        }
        if (any_restart) {
            // rounding made the recursive residual over-optimistic: restart from the true one
            upload_flags(mask_.ptr, restart_cols);
            hipLaunchKernelGGL((copy_columns_kernel<double>), dim3(g), dim3(256), 0, s, n_, m_, (const uint32_t *)mask_.ptr,
                               (const double *)tmp_.ptr, r_.ptr);
            reset_direction(mask_.ptr);
        }
    }
    for (uint32_t m = 0; m < m_; ++m) {
        Column &c = col_[m];
        if (c.status != PcgStatus::Active || c.iterations < max_iterations_) continue;
        c.status = PcgStatus::NotConverged;
        c.failure = Error(SF_ERR_PCG_NOT_CONVERGED, "the solver did not converge", (int64_t)c.iterations, 0, c.relative);
        changed = true;
    }
    if (changed) upload_flags(active_.ptr, active_flags());
}

uint32_t DevicePcgBatch::step(u64 max, const Cancel &cancel) {
    auto running = [&] {
        uint32_t k = 0;
        for (const Column &c : col_) k += c.status == PcgStatus::Active ? 1 : 0;
        return k;
    };
    for (u64 i = 0; i < max && running() > 0; ++i) {
        if (cancel()) {
            SF_HIP(hipStreamSynchronize(ctx_->stream));
            for (Column &c : col_)
                if (c.status == PcgStatus::Active) c.status = PcgStatus::Cancelled;
            upload_flags(active_.ptr, active_flags());
            break;
        }
        iterate();
    }
    SF_HIP(hipStreamSynchronize(ctx_->stream));
    return running();
}

void DevicePcgBatch::estimate(double *out) const {
    const size_t total = n_ * m_;
    hipLaunchKernelGGL((deinterleave_kernel<double>), dim3(grid_for(total, 256)), dim3(256), 0, ctx_->stream, n_, m_,
                       (const double *)x_.ptr, stage_.ptr);
    SF_HIP(hipGetLastError());
    stage_.download(out, total, ctx_->stream);
    SF_HIP(hipStreamSynchronize(ctx_->stream));
}

void DevicePcgBatch::estimate_column(uint32_t column, double *out) const {
    if (column >= m_) throw Error(SF_ERR_INVALID_ARGUMENT, "no such column");
    hipLaunchKernelGGL((get_column_kernel<double>), dim3(grid_for(n_, 256)), dim3(256), 0, ctx_->stream, n_, m_, column,
                       (const double *)x_.ptr, stage_.ptr);
    SF_HIP(hipGetLastError());
    stage_.download(out, n_, ctx_->stream);
    SF_HIP(hipStreamSynchronize(ctx_->stream));
}

void DevicePcgBatch::replace_columns(const std::vector<uint32_t> &columns, const std::vector<const double *> &rhs) {
    // restart() for the listed columns only. Every other running column has z = M^-1 r from its
    // last iteration, so the preconditioner application of reset_direction recomputes the same
    // bits there, and p and r.z change only in the listed columns.
    hipStream_t s = ctx_->stream;
    SF_HIP(hipStreamSynchronize(s));
    if (columns.size() != rhs.size()) throw Error(SF_ERR_INVALID_ARGUMENT, "one right-hand side per column");
    if (columns.empty()) return;
    std::vector<uint32_t> listed(m_, 0);
    for (size_t k = 0; k < columns.size(); ++k) {
        const uint32_t m = columns[k];
        if (m >= m_ || listed[m]) throw Error(SF_ERR_INVALID_ARGUMENT, "no such column, or a column listed twice");
        listed[m] = 1;
        SF_HIP(hipMemcpyAsync(stage_.ptr, rhs[k], n_ * sizeof(double), hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL((set_column_kernel<double>), dim3(grid_for(n_, 256)), dim3(256), 0, s, n_, m_, m,
                           (const double *)stage_.ptr, b_.ptr);
        SF_HIP(hipStreamSynchronize(s));  // the staging is reused for the next column
    }
    const size_t total = n_ * m_;
    const unsigned g = grid_for(total, 256);
    upload_flags(mask_.ptr, listed);
    hipLaunchKernelGGL((copy_columns_kernel<double>), dim3(g), dim3(256), 0, s, n_, m_, (const uint32_t *)mask_.ptr,
                       (const double *)b_.ptr, r_.ptr);
    for (double *v : {x_.ptr, z_.ptr, p_.ptr})
        hipLaunchKernelGGL((zero_columns_kernel<double>), dim3(g), dim3(256), 0, s, n_, m_, (const uint32_t *)mask_.ptr, v);
    hipLaunchKernelGGL(zero_scalars_kernel, dim3(1), dim3(WAVE), 0, s, m_, (const uint32_t *)mask_.ptr, scalars_.ptr);
    dot(b_.ptr, b_.ptr, offsetof(PcgScalars, rr), FINAL_STORE, mask_.ptr);
    const PcgScalars *h = read_scalars();
    std::vector<uint32_t> live(m_, 0);
    bool any = false;
    for (uint32_t m = 0; m < m_; ++m) {
        if (!listed[m]) continue;
        Column &c = col_[m];
        c = Column{};
        c.rhs_norm = std::sqrt(h[m].rr);
        if (c.rhs_norm == 0.0) {
            c.status = PcgStatus::Converged;
        } else {
            c.relative = 1.0;
            live[m] = 1;
            any = true;
        }
    }
    upload_flags(active_.ptr, active_flags());
    if (any) {
        upload_flags(mask_.ptr, live);
        reset_direction(mask_.ptr);
    }
    SF_HIP(hipStreamSynchronize(s));
}

size_t DevicePcgBatch::device_bytes() const {
    size_t bytes = 0;
    for (const auto *buf : {&b_, &x_, &r_, &z_, &p_, &ap_, &tmp_, &tmp2_, &jacobi_, &stage_, &partials_}) bytes += buf->bytes();
    return bytes + scalars_.bytes() + active_.bytes() + mask_.bytes();
}

void DevicePcgBatch::progress(sf_column_progress *out) const {
    for (uint32_t m = 0; m < m_; ++m) {
        const Column &c = col_[m];
        sf_column_progress p{};
        p.iteration = c.iterations;
        p.relative_residual = c.relative;
        p.at_rounding_floor = c.floor ? 1 : 0;
        switch (c.status) {
            case PcgStatus::Active: p.outcome = SF_RUNNING; break;
            case PcgStatus::Converged: p.outcome = SF_CONVERGED; break;
            case PcgStatus::Cancelled: p.outcome = SF_CANCELLED; break;
            default:
                p.outcome = SF_FAILED;
                p.error = c.failure.status;
                p.error_iteration = (u64)c.failure.a;
                p.error_value = c.failure.value;
                break;
        }
        out[m] = p;
    }
}

}  // namespace sf
