// The allocation-free form of the rounding bound (CsrMatrix::multiply_rounding_bound,
// crates/solidfem-solve/src/csr.rs:142-154) for callers that cannot allocate device memory: code
// under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc. Its scratch (the
// per-row terms and the reduction's partials) lives in one file-scope device array per GPU,
// looked up once per device when a context is created; every call on a GPU shares that array.
#include <mutex>

#include "sf_handles.hpp"
#include "sf_kernels.hpp"

namespace {

// 16 MiB per GPU, held until the process ends: the reduction's partials, its sum, then one term
// per scalar row.
constexpr size_t SCRATCH_DOUBLES = (size_t)1 << 21;
constexpr size_t PARTIALS_AT = 0;
constexpr size_t SUM_AT = 512;
constexpr size_t ROWS_AT = 1024;
static_assert(sf::dev::DOT_BLOCKS <= (int)SUM_AT, "the partials overlap the sum");

static __device__ __align__(256) double scratch[SCRATCH_DOUBLES];

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_lock;
double *g_scratch[MAX_DEVICES] = {};

__global__ void root_kernel(const double *__restrict__ sum, double *__restrict__ bound) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *bound = sqrt(*sum);
}

}  // namespace

namespace sf {

extern const size_t NOALLOC_MAX_ROWS = SCRATCH_DOUBLES - ROWS_AT;

void register_scratch(int device) {
    // with `device` current: its own copy of the array, at its own address
    std::lock_guard<std::mutex> g(g_scratch_lock);
    if (device < 0 || device >= MAX_DEVICES) throw Error(SF_ERR_INVALID_ARGUMENT, "device index beyond the scratch table");
    if (g_scratch[device]) return;
    void *ptr = nullptr;
    SF_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) throw Error(SF_ERR_HIP, "the scratch array has no device address");
    g_scratch[device] = static_cast<double *>(ptr);
}

void rounding_bound_noalloc(const sf_csr &c, const double *x, int32_t dot_order, double *bound, hipStream_t s) {
    if (!c.ctx) throw Error(SF_ERR_INVALID_ARGUMENT, "the matrix is not on a device");
    if (dot_order != SF_DOT_TREE && dot_order != SF_DOT_REFERENCE) throw Error(SF_ERR_INVALID_ARGUMENT, "unknown dot order");
    const size_t n = c.system.size;
    if (n > NOALLOC_MAX_ROWS)
        throw Error(SF_ERR_INVALID_ARGUMENT, "the matrix has more rows than the allocation-free scratch holds", (int64_t)n,
                    (int64_t)NOALLOC_MAX_ROWS);
    if (!bound || (n && !x)) throw Error(SF_ERR_INVALID_ARGUMENT, "a device pointer is missing");
    double *base = nullptr;
    {
        std::lock_guard<std::mutex> g(g_scratch_lock);
        base = g_scratch[c.ctx->device];
    }
    if (!base) throw Error(SF_ERR_INTERNAL, "no scratch address for this device");
    double *partials = base + PARTIALS_AT, *sum = base + SUM_AT, *rows = base + ROWS_AT;
    // every scratch entry read below is written first by this call: nothing needs clearing
    if (c.system.sliced.slices) {
        const dim3 grid((unsigned)((c.system.sliced.slices + 3) / 4)), block(256);
        if (c.system.rb == 3)
            hipLaunchKernelGGL((dev::rounding_bound_kernel<3, 3>), grid, block, 0, s, c.system.sliced.view(), x, rows);
        else
            hipLaunchKernelGGL((dev::rounding_bound_kernel<1, 1>), grid, block, 0, s, c.system.sliced.view(), x, rows);
    }
    if (dot_order == SF_DOT_REFERENCE) {
        hipLaunchKernelGGL((dev::dot_serial_kernel<1>), dim3(1), dim3(64), 0, s, n, (const double *)rows, (const double *)rows,
                           (int)dev::FINAL_STORE, sum, (dev::PcgScalars *)nullptr, (dev::PcgScalars *)nullptr, 0u);
    } else {
        hipLaunchKernelGGL((dev::dot_partials_kernel<1>), dim3(dev::DOT_BLOCKS), dim3(dev::DOT_THREADS), 0, s, n, (const double *)rows,
                           (const double *)rows, partials);
        hipLaunchKernelGGL(dev::dot_final_kernel, dim3(1), dim3(dev::DOT_THREADS), 0, s, dev::DOT_BLOCKS, (const double *)partials,
                           (int)dev::FINAL_STORE, sum, (dev::PcgScalars *)nullptr, (dev::PcgScalars *)nullptr, 0u);
    }
    hipLaunchKernelGGL(root_kernel, dim3(1), dim3(64), 0, s, (const double *)sum, bound);
    SF_HIP(hipGetLastError());
}

}  // namespace sf
