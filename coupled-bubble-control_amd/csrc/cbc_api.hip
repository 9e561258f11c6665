// cbc_api.hip -- the C ABI of libcbc.so (declared in include/cbc.h).
//
// Replaces the reference's backend layer: the kernel loader
// (backends/cupy_cuda/kernel_loader.py:9-99, models/ckm1d.py:214-253) becomes
// cbc_plan_create, the host launch loop (rkck45_coupled_cupy_solver.py:29-80)
// becomes cbc_solve.  No exception crosses this boundary.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>

#include "../../include/cbc.h"
#include "cbc_kernels.h"
#include "cbc_linalg.h"
#include "cbc_math.h"

static_assert(cbc::kServeReset == CBC_SERVE_RESET && cbc::kServeUseControl == CBC_SERVE_USE_CONTROL,
              "item option bits of the ABI and the ring");

namespace {

thread_local std::string g_last_error;

int32_t fail(int32_t code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

int32_t hip_fail(hipError_t e, const char* what)
{
    return fail(CBC_ERR_HIP, "%s: %s (%d)", what, hipGetErrorString(e), (int)e);
}

int next_pow2(int v)
{
    int t = 1;
    while (t < v) t <<= 1;
    return t;
}

// Clears n 32-bit words; a kernel rather than hipMemsetAsync so that launch
// sequences stay valid under stream capture.
__global__ void zero_words(uint32_t* p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

hipError_t zero_async(uint32_t* p, int n, hipStream_t st)
{
    void* args[] = {&p, &n};
    return hipLaunchKernel(reinterpret_cast<const void*>(&zero_words), dim3((n + 63) / 64), dim3(64), args, 0, st);
}

// Device evaluation of the elementary functions for cbc_selftest_math.
__global__ void selftest_math(int fn, const double* x, const double* y, double* out, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double a = x[i];
    const double b = y ? y[i] : 0.0;
    double u = 0.0, v = 0.0;
    switch (fn) {
    case CBC_MATH_SINCOS: cbc::sincos_fast(a, &u, &v); break;
    case CBC_MATH_EXP: u = cbc::exp_fast(a); break;
    case CBC_MATH_LOG: u = cbc::log_fast(a); break;
    case CBC_MATH_POW: u = cbc::pow_pos(a, b); break;
    case CBC_MATH_RCP: u = cbc::rcp_fast(a); break;
    case CBC_MATH_SINCOS_KERN: cbc::sincos_kern(a, &u, &v); break;
    case CBC_MATH_SINCOS_TINY: cbc::sincos_tiny(a, &u, &v); break;
    case CBC_MATH_ROTATE: {
        double sa, ca, sb, cm1;
        cbc::sincos_fast(a, &sa, &ca);
        cbc::sincos_tiny(b, &sb, &cm1);
        cbc::rotate(sa, ca, sb, cm1, &u, &v);
        break;
    }
    case CBC_MATH_POW1P: u = cbc::pow1p_small(a, b); break;
    default: break;
    }
    out[2 * (size_t)i] = u;
    out[2 * (size_t)i + 1] = v;
}

// The in-register coupling solve (lu_solve) for cbc_selftest_solve, one system
// per thread.
template <int n>
__global__ void __launch_bounds__(64) selftest_solve(const double* a, double* x, int count)
{
    const int s = blockIdx.x * 64 + threadIdx.x;
    if (s >= count) return;
    const double* src = a + (size_t)s * n * (n + 1);
    double m[n][n + 1], v[n];
#pragma unroll
    for (int r = 0; r < n; ++r)
#pragma unroll
        for (int j = 0; j <= n; ++j) m[r][j] = src[r * (n + 1) + j];
    cbc::lu_solve<n>(m, v);
#pragma unroll
    for (int j = 0; j < n; ++j) x[(size_t)s * n + j] = v[j];
}

}  // namespace

struct cbc_plan {
    cbc_model_spec model;
    cbc_exec_spec exec;
    cbc_solver_spec solver;
    int32_t mapping;
    int32_t block;
    uint32_t* d_active;      // device counter of systems left ACTIVE
    uint32_t* h_active;      // pinned host mirror
    int device;
};

namespace cbc {
// Error hook for the library's other translation units (cbc_env.hip): records
// the message that cbc_last_error() returns on this thread.
int32_t report_error(int32_t code, const char* msg)
{
    g_last_error = msg;
    return code;
}
}  // namespace cbc

extern "C" {

int32_t cbc_abi_version(void) { return CBC_ABI_VERSION; }

const char* cbc_last_error(void) { return g_last_error.c_str(); }

int32_t cbc_plan_create(const cbc_model_spec* model, const cbc_exec_spec* exec, const cbc_solver_spec* solver,
                        cbc_plan** out)
{
    try {
        if (!model || !exec || !solver || !out) return fail(CBC_ERR_INVALID, "cbc_plan_create: null argument");
        *out = nullptr;
        if (model->sd != 4) return fail(CBC_ERR_INVALID, "sd must be 4 (R, X, U, V), got %d", model->sd);
        if (model->nup != 13) return fail(CBC_ERR_INVALID, "nup must be 13, got %d", model->nup);
        if (model->nsp != 5) return fail(CBC_ERR_INVALID, "nsp must be 5, got %d", model->nsp);
        if (model->ncm != 3) return fail(CBC_ERR_INVALID, "ncm must be 3, got %d", model->ncm);
        if (model->nk < 1) return fail(CBC_ERR_INVALID, "nk must be >= 1, got %d", model->nk);
        if (model->field < CBC_FIELD_CONST || model->field > CBC_FIELD_SW_A)
            return fail(CBC_ERR_INVALID, "unknown acoustic field %d", model->field);
        if (exec->ns < 1) return fail(CBC_ERR_INVALID, "ns must be >= 1, got %d", exec->ns);
        if (exec->ups < 1) return fail(CBC_ERR_INVALID, "ups must be >= 1, got %d", exec->ups);
        if (exec->ndo < 0) return fail(CBC_ERR_INVALID, "ndo must be >= 0, got %d", exec->ndo);
        if (exec->tile != next_pow2(exec->ups))
            return fail(CBC_ERR_INVALID, "tile must be next_pow2(ups)=%d, got %d", next_pow2(exec->ups), exec->tile);
        if (!(solver->atol > 0.0) || !(solver->rtol >= 0.0))
            return fail(CBC_ERR_INVALID, "atol must be > 0 and rtol >= 0");
        if (!(solver->min_step > 0.0) || !(solver->max_step >= solver->min_step))
            return fail(CBC_ERR_INVALID, "need 0 < min_step <= max_step");
        if (solver->linsolve != CBC_LIN_DIRECT)
            return fail(CBC_ERR_UNSUPPORTED, "only the direct coupling solve (CBC_LIN_DIRECT) is built");
        if (solver->event_mode != CBC_EVENT_BRACKET && solver->event_mode != CBC_EVENT_EXACT)
            return fail(CBC_ERR_INVALID, "unknown event mode %d", solver->event_mode);

        // AUTO: N = 1 in one lane.  N = 2 in lane pairs (the group mapping, bubble
        // data and coupling rows passed by DPP, 256 VGPRs) while the pairs' waves
        // fit one per SIMD: an attempt ~15 % faster than one system per lane
        // there (4 096 / 16 384 systems: 9.9 / 10.0 vs 11.5 / 11.7 us); with more
        // systems one per lane (395 VGPRs, one wave per SIMD), since two pair
        // waves sharing a SIMD take longer (65 536 systems: 14.6 vs 12.9 us).
        // N >= 3 one lane per bubble (group mapping).
        int mapping = exec->mapping;
        if (mapping == CBC_MAP_AUTO) {
            if (exec->ups == 1) {
                mapping = CBC_MAP_LANE;
            } else if (exec->ups == 2) {
                int dev = 0, cus = 0;
                hipError_t e = hipGetDevice(&dev);
                if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
                const int64_t pair_waves = ((int64_t)exec->ns * 2 + 63) / 64;
                mapping = (e == hipSuccess && cus > 0 && pair_waves <= 4 * (int64_t)cus) ? CBC_MAP_GROUP : CBC_MAP_LANE;
            } else {
                mapping = CBC_MAP_GROUP;
            }
        }
        if (mapping == CBC_MAP_GROUP && exec->ups == 1) mapping = CBC_MAP_LANE;
        int block = 64;
        if (mapping == CBC_MAP_LANE) {
            if (!cbc::lane_supported(exec->ups, model->nk, model->field))
                return fail(CBC_ERR_UNSUPPORTED, "lane kernel not instantiated for ups=%d nk=%d field=%d",
                            exec->ups, model->nk, model->field);
            block = exec->block_size > 0 ? exec->block_size : 64;
            if (block > 256 || block % 64 != 0)
                return fail(CBC_ERR_INVALID, "block_size must be a multiple of 64 in [64, 256], got %d", block);
        } else if (mapping == CBC_MAP_GROUP) {
            if (!cbc::group_supported(exec->ups, model->nk, model->field))
                return fail(CBC_ERR_UNSUPPORTED, "group kernel not instantiated for ups=%d nk=%d field=%d",
                            exec->ups, model->nk, model->field);
            if (exec->block_size > 0 && exec->block_size != 64)
                return fail(CBC_ERR_INVALID, "the group mapping runs one wave (64 threads) per block");
        } else {
            return fail(CBC_ERR_INVALID, "unknown mapping %d", mapping);
        }

        cbc_plan* p = new (std::nothrow) cbc_plan();
        if (!p) return fail(CBC_ERR_NOMEM, "cbc_plan_create: out of host memory");
        p->model = *model;
        p->exec = *exec;
        p->solver = *solver;
        p->mapping = mapping;
        p->block = block;
        hipError_t e = hipGetDevice(&p->device);
        if (e != hipSuccess) {
            delete p;
            return hip_fail(e, "hipGetDevice");
        }
        e = hipMalloc(&p->d_active, 16);
        if (e != hipSuccess) {
            delete p;
            return hip_fail(e, "hipMalloc(active counter)");
        }
        e = hipHostMalloc(&p->h_active, 16, hipHostMallocDefault);
        if (e != hipSuccess) {
            (void)hipFree(p->d_active);
            delete p;
            return hip_fail(e, "hipHostMalloc(active counter)");
        }
        *out = p;
        return CBC_OK;
    } catch (...) {
        return fail(CBC_ERR_INVALID, "cbc_plan_create: unexpected exception");
    }
}

int32_t cbc_plan_query(const cbc_plan* plan, cbc_plan_info* out)
{
    if (!plan || !out) return fail(CBC_ERR_INVALID, "cbc_plan_query: null argument");
    cbc::KernelAttr at{};
    const bool lane = plan->mapping == CBC_MAP_LANE;
    hipError_t e = lane ? cbc::lane_attr(plan->exec.ups, plan->model.nk, plan->model.field, &at)
                        : cbc::group_attr(plan->exec.ups, plan->model.nk, plan->model.field, &at);
    if (e != hipSuccess) return hip_fail(e, "hipFuncGetAttributes");
    const int spb = lane ? plan->block : cbc::group_systems_per_block(plan->exec.ups);
    out->mapping = plan->mapping;
    out->block_size = plan->block;
    out->grid_size = (plan->exec.ns + spb - 1) / spb;
    out->systems_per_block = spb;
    out->lds_bytes = at.lds_bytes;
    out->num_regs = at.num_regs;
    out->local_bytes = at.local_bytes;
    out->max_threads = at.max_threads;
    return CBC_OK;
}

void cbc_plan_destroy(cbc_plan* plan)
{
    if (!plan) return;
    if (plan->d_active) (void)hipFree(plan->d_active);
    if (plan->h_active) (void)hipHostFree(plan->h_active);
    delete plan;
}

static hipError_t launch_kernel(const cbc_plan* p, const cbc::LaunchArgs& a, hipStream_t st)
{
    return p->mapping == CBC_MAP_LANE ? cbc::lane_launch(a, p->block, st) : cbc::group_launch(a, st);
}

static int32_t fill_args(const cbc_plan* p, const cbc_buffers* b, cbc::LaunchArgs& a)
{
    if (!b) return fail(CBC_ERR_INVALID, "null buffer set");
    if (!b->actual_time || !b->time_end || !b->time_step || !b->actual_state || !b->unit_params ||
        !b->system_params || !b->control_params || !b->coupling_matrices || !b->actual_event || !b->status_flags)
        return fail(CBC_ERR_INVALID, "a required buffer pointer is null");
    if (p->exec.ndo > 0 && (!b->dense_index || !b->dense_time || !b->dense_state || !b->time_begin))
        return fail(CBC_ERR_INVALID, "ndo > 0 needs dense_index, dense_time, dense_state and time_begin");
    std::memset(&a, 0, sizeof(a));
    a.ns = p->exec.ns;
    a.tile = p->exec.tile;
    a.ups = p->exec.ups;
    a.ndo = p->exec.ndo;
    a.nk = p->model.nk;
    a.field = p->model.field;
    a.atol = p->solver.atol;
    a.rtol = p->solver.rtol;
    a.min_step = p->solver.min_step;
    a.max_step = p->solver.max_step;
    a.etol = p->solver.etol;
    a.col_factor = 1.0 + p->solver.col_threshold;
    a.event_exact = p->solver.event_mode == CBC_EVENT_EXACT;
    a.t = b->actual_time;
    a.t_end = b->time_end;
    a.t_begin = b->time_begin;
    a.h = b->time_step;
    a.state = b->actual_state;
    a.up = b->unit_params;
    a.sp = b->system_params;
    a.cp = b->control_params;
    a.cm = b->coupling_matrices;
    a.dense_index = b->dense_index;
    a.dense_time = b->dense_time;
    a.dense_state = b->dense_state;
    a.event = b->actual_event;
    a.flags = b->status_flags;
    a.total_steps = b->total_steps;
    a.rejected_steps = b->rejected_steps;
    a.period = 0.0;
    a.queue = nullptr;
    a.queue_len = 0;
    a.refill = 1;
    a.queue_head = nullptr;
    return CBC_OK;
}

int32_t cbc_launch(cbc_plan* plan, const cbc_buffers* buf, int32_t kernel_steps, void* stream)
{
    if (!plan) return fail(CBC_ERR_INVALID, "cbc_launch: null plan");
    if (kernel_steps < 1) return fail(CBC_ERR_INVALID, "kernel_steps must be >= 1, got %d", kernel_steps);
    cbc::LaunchArgs a;
    int32_t rc = fill_args(plan, buf, a);
    if (rc != CBC_OK) return rc;
    a.kernel_steps = kernel_steps;
    a.active_count = nullptr;
    hipError_t e = launch_kernel(plan, a, (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "kernel launch");
    return CBC_OK;
}

int32_t cbc_solve(cbc_plan* plan, const cbc_buffers* buf, int32_t kernel_steps, int64_t max_steps, void* stream,
                  double* launch_times, int32_t max_times, int32_t* n_launches)
{
    if (!plan) return fail(CBC_ERR_INVALID, "cbc_solve: null plan");
    if (kernel_steps < 1) return fail(CBC_ERR_INVALID, "kernel_steps must be >= 1, got %d", kernel_steps);
    if (max_steps < 1) return fail(CBC_ERR_INVALID, "max_steps must be >= 1");
    cbc::LaunchArgs a;
    int32_t rc = fill_args(plan, buf, a);
    if (rc != CBC_OK) return rc;
    a.kernel_steps = kernel_steps;
    a.active_count = plan->d_active;
    hipStream_t st = (hipStream_t)stream;
    const int64_t max_launches = (max_steps + kernel_steps - 1) / kernel_steps;
    int32_t launches = 0;
    for (int64_t l = 0; l < max_launches; ++l) {
        auto t0 = std::chrono::steady_clock::now();
        a.resume = l > 0;
        hipError_t e = zero_async(plan->d_active, 1, st);
        if (e != hipSuccess) return hip_fail(e, "zero active counter");
        e = launch_kernel(plan, a, st);
        if (e != hipSuccess) return hip_fail(e, "kernel launch");
        e = hipMemcpyAsync(plan->h_active, plan->d_active, sizeof(uint32_t), hipMemcpyDeviceToHost, st);
        if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync(active counter)");
        e = hipStreamSynchronize(st);
        if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
        auto t1 = std::chrono::steady_clock::now();
        if (launch_times && launches < max_times)
            launch_times[launches] = std::chrono::duration<double>(t1 - t0).count();
        ++launches;
        if (*plan->h_active == 0) break;
    }
    if (n_launches) *n_launches = launches;
    return CBC_OK;
}

int32_t cbc_solve_periodic(cbc_plan* plan, const cbc_buffers* buf, const cbc_period_spec* period, int64_t max_steps,
                           void* stream)
{
    if (!plan) return fail(CBC_ERR_INVALID, "cbc_solve_periodic: null plan");
    if (!period) return fail(CBC_ERR_INVALID, "cbc_solve_periodic: null period spec");
    if (!(period->period > 0.0)) return fail(CBC_ERR_INVALID, "period must be > 0");
    if (period->n_windows < 0 || (period->n_windows > 0 && (!period->max_r || !period->mean_x)))
        return fail(CBC_ERR_INVALID, "n_windows > 0 needs max_r and mean_x");
    if (max_steps < 1) return fail(CBC_ERR_INVALID, "max_steps must be >= 1");
    if (plan->exec.ndo > 0) return fail(CBC_ERR_INVALID, "period windows and dense output are exclusive");
    cbc::LaunchArgs a;
    int32_t rc = fill_args(plan, buf, a);
    if (rc != CBC_OK) return rc;
    a.kernel_steps = max_steps;
    a.period = period->period;
    a.first_window = period->first_window;
    a.n_windows = period->n_windows;
    a.max_r = period->max_r;
    a.mean_x = period->mean_x;
    a.resume = period->resume != 0;
    a.active_count = nullptr;
    hipError_t e = launch_kernel(plan, a, (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "kernel launch (periodic)");
    return CBC_OK;
}

static int32_t queued_args(cbc_plan* plan, const cbc_buffers* buf, const int32_t* order, int32_t order_len,
                           int32_t& waves, int32_t refill, int32_t kernel_steps, cbc::LaunchArgs& a)
{
    if (!plan) return fail(CBC_ERR_INVALID, "queued solve: null plan");
    if (!order || order_len < 1) return fail(CBC_ERR_INVALID, "order must hold at least one system index");
    if (order_len > plan->exec.ns) return fail(CBC_ERR_INVALID, "order_len %d exceeds ns %d", order_len, plan->exec.ns);
    if (kernel_steps < 1) return fail(CBC_ERR_INVALID, "kernel_steps must be >= 1, got %d", kernel_steps);
    if (refill < 1 || refill > 64) return fail(CBC_ERR_INVALID, "refill must be in [1, 64], got %d", refill);
    if (waves < 0) return fail(CBC_ERR_INVALID, "waves must be >= 0");
    int32_t rc = fill_args(plan, buf, a);
    if (rc != CBC_OK) return rc;
    const bool lane = plan->mapping == CBC_MAP_LANE;
    if (waves == 0) {
        int dev = 0, cus = 0, per_cu = 0;
        hipError_t e = hipGetDevice(&dev);
        if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        if (e != hipSuccess) return hip_fail(e, "device query");
        cbc::KernelAttr at{};
        e = lane ? cbc::lane_attr(plan->exec.ups, plan->model.nk, plan->model.field, &at)
                 : cbc::group_attr(plan->exec.ups, plan->model.nk, plan->model.field, &at);
        if (e != hipSuccess) return hip_fail(e, "hipFuncGetAttributes");
        // waves resident per SIMD from the register budget (512 per lane on gfx950)
        const int regs = at.num_regs > 0 ? at.num_regs : 512;
        per_cu = 4 * (512 / regs > 0 ? 512 / regs : 1);
        waves = cus * per_cu;
    }
    a.kernel_steps = kernel_steps;
    a.queue = order;
    a.queue_len = order_len;
    a.refill = refill;
    a.queue_head = plan->d_active + 1;
    a.active_count = plan->d_active;
    return CBC_OK;
}

static hipError_t queued_launch(cbc_plan* plan, const cbc::LaunchArgs& a, int32_t waves, hipStream_t st)
{
    hipError_t e = zero_async(plan->d_active, 2, st);
    if (e != hipSuccess) return e;
    return plan->mapping == CBC_MAP_LANE ? cbc::lane_launch_persistent(a, plan->block, waves, st)
                                         : cbc::group_launch_persistent(a, waves, st);
}

int32_t cbc_launch_queued(cbc_plan* plan, const cbc_buffers* buf, const int32_t* order, int32_t order_len,
                          int32_t waves, int32_t refill, int32_t kernel_steps, void* stream)
{
    cbc::LaunchArgs a;
    int32_t rc = queued_args(plan, buf, order, order_len, waves, refill, kernel_steps, a);
    if (rc != CBC_OK) return rc;
    hipError_t e = queued_launch(plan, a, waves, (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "persistent kernel launch");
    return CBC_OK;
}

// Bubble counts of a fused launch: distinct, 1..kFusedMaxBins of them.
static int32_t fused_check_list(const int32_t* ups, int32_t n_bins)
{
    if (!ups) return fail(CBC_ERR_INVALID, "fused launch: null bubble-count list");
    if (n_bins < 1 || n_bins > cbc::kFusedMaxBins)
        return fail(CBC_ERR_INVALID, "fused launch: n_bins = %d not in [1, %d]", n_bins, cbc::kFusedMaxBins);
    for (int b = 0; b < n_bins; ++b)
        for (int c = 0; c < b; ++c)
            if (ups[b] == ups[c]) return fail(CBC_ERR_INVALID, "fused launch: bubble count %d in two bins", ups[b]);
    return CBC_OK;
}

int32_t cbc_fused_supported(const int32_t* ups, int32_t n_bins, int32_t nk, int32_t field)
{
    int32_t rc = fused_check_list(ups, n_bins);
    if (rc != CBC_OK) return rc;
    int u[cbc::kFusedMaxBins], slot[cbc::kFusedMaxBins];
    for (int b = 0; b < n_bins; ++b) u[b] = ups[b];
    return cbc::fused_pick(u, n_bins, nk, field, slot) != nullptr ? 1 : 0;
}

int32_t cbc_launch_fused(cbc_plan* const* plans, const cbc_buffers* const* bufs, const int32_t* const* orders,
                         const int32_t* order_lens, const int32_t* refills, int32_t n_bins, int32_t waves,
                         int32_t kernel_steps, void* stream)
{
    if (!plans || !bufs || !orders || !order_lens || !refills)
        return fail(CBC_ERR_INVALID, "cbc_launch_fused: null argument array");
    if (n_bins < 1 || n_bins > cbc::kFusedMaxBins)
        return fail(CBC_ERR_INVALID, "cbc_launch_fused: n_bins = %d not in [1, %d]", n_bins, cbc::kFusedMaxBins);
    if (waves < 0) return fail(CBC_ERR_INVALID, "waves must be >= 0");
    int32_t ups[cbc::kFusedMaxBins];
    for (int b = 0; b < n_bins; ++b) {
        const cbc_plan* p = plans[b];
        if (!p || !bufs[b]) return fail(CBC_ERR_INVALID, "cbc_launch_fused: null plan or buffers of bin %d", b);
        if (p->mapping != CBC_MAP_GROUP)
            return fail(CBC_ERR_UNSUPPORTED, "cbc_launch_fused: bin %d (N = %d) does not use the group mapping", b,
                        p->exec.ups);
        if (p->model.nk != plans[0]->model.nk || p->model.field != plans[0]->model.field)
            return fail(CBC_ERR_INVALID, "cbc_launch_fused: bins differ in harmonic count or field type");
        if (p->device != plans[0]->device) return fail(CBC_ERR_INVALID, "cbc_launch_fused: plans on different devices");
        ups[b] = p->exec.ups;
    }
    int32_t rc = fused_check_list(ups, n_bins);
    if (rc != CBC_OK) return rc;
    int u[cbc::kFusedMaxBins], slot[cbc::kFusedMaxBins];
    for (int b = 0; b < n_bins; ++b) u[b] = ups[b];
    const void* fn = cbc::fused_pick(u, n_bins, plans[0]->model.nk, plans[0]->model.field, slot);
    if (!fn) return fail(CBC_ERR_UNSUPPORTED, "cbc_launch_fused: no fused kernel covers these bubble counts");
    cbc::FusedArgs f;
    std::memset(&f, 0, sizeof(f));
    for (int b = 0; b < n_bins; ++b) {
        int32_t w = 1;  // the grid is the fused kernel's, not the bin's
        rc = queued_args(plans[b], bufs[b], orders[b], order_lens[b], w, refills[b], kernel_steps, f.bin[slot[b]]);
        if (rc != CBC_OK) return rc;
    }
    if (waves == 0) {
        int dev = 0, cus = 0;
        hipError_t e = hipGetDevice(&dev);
        if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        if (e != hipSuccess) return hip_fail(e, "device query");
        cbc::KernelAttr at{};
        e = cbc::fused_attr(fn, &at);
        if (e != hipSuccess) return hip_fail(e, "hipFuncGetAttributes");
        const int regs = at.num_regs > 0 ? at.num_regs : 512;
        waves = cus * 4 * (512 / regs > 0 ? 512 / regs : 1);
    }
    hipStream_t st = (hipStream_t)stream;
    for (int b = 0; b < n_bins; ++b) {
        hipError_t e = zero_async(plans[b]->d_active, 2, st);
        if (e != hipSuccess) return hip_fail(e, "zero queue head");
    }
    hipError_t e = cbc::fused_launch(fn, f, waves, st);
    if (e != hipSuccess) return hip_fail(e, "fused persistent kernel launch");
    return CBC_OK;
}

int32_t cbc_selftest_math(int32_t fn, const double* x, const double* y, double* out, int32_t n, void* stream)
{
    if (fn < CBC_MATH_SINCOS || fn > CBC_MATH_POW1P) return fail(CBC_ERR_INVALID, "unknown function code %d", fn);
    if (n < 0 || (n > 0 && (!x || !out))) return fail(CBC_ERR_INVALID, "cbc_selftest_math: bad arrays");
    if ((fn == CBC_MATH_POW || fn == CBC_MATH_ROTATE || fn == CBC_MATH_POW1P) && n > 0 && !y)
        return fail(CBC_ERR_INVALID, "function %d needs the second argument array", fn);
    if (n == 0) return CBC_OK;
    int f = fn;
    void* args[] = {&f, &x, &y, &out, &n};
    hipError_t e = hipLaunchKernel(reinterpret_cast<const void*>(&selftest_math), dim3((n + 255) / 256), dim3(256),
                                   args, 0, (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "selftest kernel launch");
    return CBC_OK;
}

int32_t cbc_selftest_solve(int32_t n, const double* a, double* x, int32_t count, void* stream)
{
    if (n != 2 && n != 4 && n != 6 && n != 8)
        return fail(CBC_ERR_INVALID, "cbc_selftest_solve: n = %d not in {2, 4, 6, 8}", n);
    if (count < 0 || (count > 0 && (!a || !x))) return fail(CBC_ERR_INVALID, "cbc_selftest_solve: bad arrays");
    if (count == 0) return CBC_OK;
    const void* k = n == 2 ? reinterpret_cast<const void*>(&selftest_solve<2>)
                  : n == 4 ? reinterpret_cast<const void*>(&selftest_solve<4>)
                  : n == 6 ? reinterpret_cast<const void*>(&selftest_solve<6>)
                           : reinterpret_cast<const void*>(&selftest_solve<8>);
    void* args[] = {&a, &x, &count};
    hipError_t e = hipLaunchKernel(k, dim3((count + 63) / 64), dim3(64), args, 0, (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "selftest kernel launch");
    return CBC_OK;
}

int32_t cbc_plan_active_count(cbc_plan* plan, void* stream, int32_t* out)
{
    if (!plan || !out) return fail(CBC_ERR_INVALID, "cbc_plan_active_count: null argument");
    hipStream_t st = (hipStream_t)stream;
    hipError_t e = hipMemcpyAsync(plan->h_active, plan->d_active, sizeof(uint32_t), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(e, "active count readback");
    *out = (int32_t)*plan->h_active;
    return CBC_OK;
}

int32_t cbc_solve_queued(cbc_plan* plan, const cbc_buffers* buf, const int32_t* order, int32_t order_len,
                         int32_t waves, int32_t refill, int32_t kernel_steps, int64_t max_steps, void* stream,
                         int32_t* n_launches)
{
    if (max_steps < 1) return fail(CBC_ERR_INVALID, "max_steps must be >= 1");
    cbc::LaunchArgs a;
    int32_t rc = queued_args(plan, buf, order, order_len, waves, refill, kernel_steps, a);
    if (rc != CBC_OK) return rc;
    hipStream_t st = (hipStream_t)stream;
    const int64_t max_launches = (max_steps + kernel_steps - 1) / kernel_steps;
    int32_t launches = 0;
    for (int64_t l = 0; l < max_launches; ++l) {
        a.resume = l > 0;
        hipError_t e = queued_launch(plan, a, waves, st);
        if (e != hipSuccess) return hip_fail(e, "persistent kernel launch");
        e = hipMemcpyAsync(plan->h_active, plan->d_active, sizeof(uint32_t), hipMemcpyDeviceToHost, st);
        if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync(active counter)");
        e = hipStreamSynchronize(st);
        if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
        ++launches;
        if (*plan->h_active == 0) break;
    }
    if (n_launches) *n_launches = launches;
    return CBC_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// serving: a persistent kernel fed from a host-mapped work ring
// ---------------------------------------------------------------------------
struct cbc_serve {
    cbc_plan* plan;
    cbc::LaunchArgs args;
    hipStream_t stream;
    bool own_stream;             // created here
    hipEvent_t ended;            // recorded after each launch
    bool launched;
    int32_t waves;
    int32_t capacity;
    int32_t nk;
    cbc::RingHost rh;            // ring slots, per-system sequence numbers (cbc_ring.h)
    cbc::ServeItem* ring;        // host-mapped, capacity entries
    uint32_t* ctl;               // host-mapped: [0] tail, [1] stop
    uint32_t* done_seq;          // host-mapped (NS)
    double* out_state;           // host-mapped (4, NS, TILE) or null
    double* out_misc;            // host-mapped (NS, 4)
    double* stage_state;         // host-mapped staging mirrors for episode resets
    double* stage_up;
    double* stage_cm;
    uint32_t* last_seq;          // host (NS): seq of the last item submitted per system
    int64_t launches;
};

namespace {

bool serve_running(cbc_serve* s)
{
    if (!s->launched) return false;
    return hipEventQuery(s->ended) == hipErrorNotReady;
}

void serve_retire(cbc_serve* s) { cbc::ring_retire(s->rh); }

int32_t serve_launch(cbc_serve* s)
{
    cbc_plan* p = s->plan;
    hipError_t e = p->mapping == CBC_MAP_LANE ? cbc::lane_launch_persistent(s->args, p->block, s->waves, s->stream)
                                              : cbc::group_launch_persistent(s->args, s->waves, s->stream);
    if (e != hipSuccess) return hip_fail(e, "serving kernel launch");
    e = hipEventRecord(s->ended, s->stream);
    if (e != hipSuccess) return hip_fail(e, "hipEventRecord");
    s->launched = true;
    ++s->launches;
    return CBC_OK;
}

void serve_free(cbc_serve* s)
{
    if (s->ended) (void)hipEventDestroy(s->ended);
    if (s->own_stream && s->stream) (void)hipStreamDestroy(s->stream);
    if (s->args.queue_head) (void)hipFree(s->args.queue_head);
    if (s->ring) (void)hipHostFree(s->ring);
    if (s->ctl) (void)hipHostFree(s->ctl);
    if (s->done_seq) (void)hipHostFree(s->done_seq);
    if (s->out_state) (void)hipHostFree(s->out_state);
    if (s->out_misc) (void)hipHostFree(s->out_misc);
    if (s->stage_state) (void)hipHostFree(s->stage_state);
    if (s->stage_up) (void)hipHostFree(s->stage_up);
    if (s->stage_cm) (void)hipHostFree(s->stage_cm);
    delete[] s->last_seq;
    delete s;
}

}  // namespace

extern "C" {

int32_t cbc_serve_create(cbc_plan* plan, const cbc_buffers* buf, const cbc_serve_spec* spec, void* stream,
                         cbc_serve** out)
{
    try {
        if (!plan || !spec || !out) return fail(CBC_ERR_INVALID, "cbc_serve_create: null argument");
        *out = nullptr;
        const int32_t cap = spec->capacity;
        if (cap < 64 || cap > (1 << 24) || (cap & (cap - 1)) != 0)
            return fail(CBC_ERR_INVALID, "capacity must be a power of two in [64, 2^24], got %d", cap);
        if (spec->refill < 1 || spec->refill > 64)
            return fail(CBC_ERR_INVALID, "refill must be in [1, 64], got %d", spec->refill);
        if (spec->poll_interval < 1 || spec->poll_interval > (1 << 20))
            return fail(CBC_ERR_INVALID, "poll_interval must be in [1, 2^20], got %d", spec->poll_interval);
        if (spec->max_attempts < 1) return fail(CBC_ERR_INVALID, "max_attempts must be >= 1");
        if (spec->idle_us < 1 || spec->idle_us > 10000000)
            return fail(CBC_ERR_INVALID, "idle_us must be in [1, 1e7], got %d", spec->idle_us);
        if (spec->waves < 0) return fail(CBC_ERR_INVALID, "waves must be >= 0");
        if (plan->model.nk > 2) return fail(CBC_ERR_UNSUPPORTED, "serving items carry at most 2 drive amplitudes");
        if (plan->exec.ndo > 0) return fail(CBC_ERR_UNSUPPORTED, "serving does not record dense output");

        cbc_serve* s = new (std::nothrow) cbc_serve();
        if (!s) return fail(CBC_ERR_NOMEM, "cbc_serve_create: out of host memory");
        int32_t rc = fill_args(plan, buf, s->args);
        if (rc != CBC_OK) {
            delete s;
            return rc;
        }
        s->plan = plan;
        s->stream = (hipStream_t)stream;
        if (!s->stream) {
            hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
            if (e != hipSuccess) {
                s->stream = nullptr;
                serve_free(s);
                return hip_fail(e, "serving stream creation");
            }
            s->own_stream = true;
        }
        s->capacity = cap;
        s->nk = plan->model.nk;
        const size_t ns = (size_t)plan->exec.ns;
        s->last_seq = new (std::nothrow) uint32_t[ns]();
        if (!s->last_seq) {
            serve_free(s);
            return fail(CBC_ERR_NOMEM, "cbc_serve_create: out of host memory");
        }
        const unsigned hf = hipHostMallocCoherent | hipHostMallocMapped;
        hipError_t e = hipHostMalloc((void**)&s->ring, sizeof(cbc::ServeItem) * (size_t)cap, hf);
        if (e == hipSuccess) e = hipHostMalloc((void**)&s->ctl, 64, hf);
        if (e == hipSuccess) e = hipHostMalloc((void**)&s->done_seq, sizeof(uint32_t) * ns, hf);
        if (e == hipSuccess) e = hipHostMalloc((void**)&s->out_misc, sizeof(double) * 4 * ns, hf);
        const size_t nst = ns * (size_t)plan->exec.tile;
        const size_t n_cm = 3 * (size_t)plan->exec.ups * nst;
        if (e == hipSuccess && spec->copy_state) e = hipHostMalloc((void**)&s->out_state, sizeof(double) * 4 * nst, hf);
        if (e == hipSuccess) e = hipHostMalloc((void**)&s->stage_state, sizeof(double) * 4 * nst, hf);
        if (e == hipSuccess) e = hipHostMalloc((void**)&s->stage_up, sizeof(double) * 13 * nst, hf);
        if (e == hipSuccess) e = hipHostMalloc((void**)&s->stage_cm, sizeof(double) * n_cm, hf);
        if (e == hipSuccess) e = hipMalloc((void**)&s->args.queue_head, 16);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&s->ended, hipEventDisableTiming);
        if (e != hipSuccess) {
            serve_free(s);
            return hip_fail(e, "cbc_serve_create allocation");
        }
        std::memset(s->ring, 0, sizeof(cbc::ServeItem) * (size_t)cap);
        s->rh.ring = s->ring;
        s->rh.capacity = (uint32_t)cap;
        s->rh.ns = plan->exec.ns;
        s->rh.nk = plan->model.nk;
        s->rh.done_seq = s->done_seq;
        s->rh.last_seq = s->last_seq;
        s->rh.tail = 0;
        s->rh.retired = 0;
        std::memset(s->ctl, 0, 64);
        std::memset(s->done_seq, 0, sizeof(uint32_t) * ns);
        std::memset(s->out_misc, 0, sizeof(double) * 4 * ns);
        if (s->out_state) std::memset(s->out_state, 0, sizeof(double) * 4 * nst);
        std::memset(s->stage_state, 0, sizeof(double) * 4 * nst);
        std::memset(s->stage_up, 0, sizeof(double) * 13 * nst);
        std::memset(s->stage_cm, 0, sizeof(double) * n_cm);
        e = hipMemsetAsync(s->args.queue_head, 0, 16, s->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
        if (e != hipSuccess) {
            serve_free(s);
            return hip_fail(e, "cbc_serve_create counter reset");
        }

        int dev = 0, cus = 0, clk_khz = 0;
        e = hipGetDevice(&dev);
        if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        if (e == hipSuccess) e = hipDeviceGetAttribute(&clk_khz, hipDeviceAttributeWallClockRate, dev);
        if (e != hipSuccess) {
            serve_free(s);
            return hip_fail(e, "device query");
        }
        const bool lane = plan->mapping == CBC_MAP_LANE;
        cbc::KernelAttr at{};
        e = lane ? cbc::lane_attr(plan->exec.ups, plan->model.nk, plan->model.field, &at)
                 : cbc::group_attr(plan->exec.ups, plan->model.nk, plan->model.field, &at);
        if (e != hipSuccess) {
            serve_free(s);
            return hip_fail(e, "hipFuncGetAttributes");
        }
        int waves = spec->waves;
        if (waves == 0) {
            // enough lanes for every ring item in flight at once, at most one
            // wave per SIMD slot the register budget allows
            const int per_wave = lane ? 64 : cbc::group_systems_per_block(plan->exec.ups);
            const int regs = at.num_regs > 0 ? at.num_regs : 512;
            const int64_t slots = (int64_t)cus * 4 * (512 / regs > 0 ? 512 / regs : 1);
            const int64_t want = ((int64_t)std::min<size_t>((size_t)cap, ns) + per_wave - 1) / per_wave;
            waves = (int)std::max<int64_t>(1, std::min<int64_t>(want, slots));
        }
        s->waves = waves;

        // device-side addresses of the host-mapped buffers
        void *d_ring = nullptr, *d_ctl = nullptr, *d_done = nullptr, *d_misc = nullptr, *d_state = nullptr;
        void *d_sst = nullptr, *d_sup = nullptr, *d_scm = nullptr;
        e = hipHostGetDevicePointer(&d_ring, s->ring, 0);
        if (e == hipSuccess) e = hipHostGetDevicePointer(&d_ctl, s->ctl, 0);
        if (e == hipSuccess) e = hipHostGetDevicePointer(&d_done, s->done_seq, 0);
        if (e == hipSuccess) e = hipHostGetDevicePointer(&d_misc, s->out_misc, 0);
        if (e == hipSuccess && s->out_state) e = hipHostGetDevicePointer(&d_state, s->out_state, 0);
        if (e == hipSuccess) e = hipHostGetDevicePointer(&d_sst, s->stage_state, 0);
        if (e == hipSuccess) e = hipHostGetDevicePointer(&d_sup, s->stage_up, 0);
        if (e == hipSuccess) e = hipHostGetDevicePointer(&d_scm, s->stage_cm, 0);
        if (e != hipSuccess) {
            serve_free(s);
            return hip_fail(e, "hipHostGetDevicePointer");
        }
        cbc::LaunchArgs& a = s->args;
        a.kernel_steps = spec->max_attempts;
        a.refill = spec->refill;
        a.ring = static_cast<const cbc::ServeItem*>(d_ring);
        a.ring_tail = static_cast<const uint32_t*>(d_ctl);
        a.ring_stop = static_cast<const uint32_t*>(d_ctl) + 1;
        a.ring_mask = (uint32_t)(cap - 1);
        a.poll_interval = spec->poll_interval;
        a.idle_ticks = (uint64_t)spec->idle_us * (uint64_t)(clk_khz > 0 ? clk_khz : 100000) / 1000ull;
        a.done_seq = static_cast<uint32_t*>(d_done);
        a.out_state = static_cast<double*>(d_state);
        a.out_misc = static_cast<double*>(d_misc);
        a.stage_state = static_cast<const double*>(d_sst);
        a.stage_up = static_cast<const double*>(d_sup);
        a.stage_cm = static_cast<const double*>(d_scm);
        a.active_count = nullptr;
        *out = s;
        return CBC_OK;
    } catch (...) {
        return fail(CBC_ERR_INVALID, "cbc_serve_create: unexpected exception");
    }
}

int32_t cbc_serve_get_view(const cbc_serve* s, cbc_serve_view* out)
{
    if (!s || !out) return fail(CBC_ERR_INVALID, "cbc_serve_get_view: null argument");
    out->done_seq = s->done_seq;
    out->out_state = s->out_state;
    out->out_misc = s->out_misc;
    out->stage_state = s->stage_state;
    out->stage_up = s->stage_up;
    out->stage_cm = s->stage_cm;
    out->capacity = s->capacity;
    out->waves = s->waves;
    return CBC_OK;
}

int32_t cbc_serve_submit(cbc_serve* s, const int32_t* sys, const double* pa, int32_t n, double t_end,
                         const double* h0, const int32_t* opts, uint32_t* seq_out)
{
    if (!s) return fail(CBC_ERR_INVALID, "cbc_serve_submit: null handle");
    if (n < 0 || (n > 0 && (!sys || !pa))) return fail(CBC_ERR_INVALID, "cbc_serve_submit: bad item arrays");
    if (!(t_end > 0.0) || !(t_end < 1e300)) return fail(CBC_ERR_INVALID, "t_end must be positive and finite");
    if (n == 0) return CBC_OK;
    int32_t bad = 0;
    switch (cbc::ring_submit(s->rh, sys, pa, n, t_end, h0, opts, seq_out, &bad)) {
    case cbc::kRingOk: break;
    case cbc::kRingFull:
        return fail(CBC_ERR_BUSY, "serving ring full: %u of %d slots in use, %d requested", s->rh.tail - s->rh.retired,
                    s->capacity, n);
    case cbc::kRingBadSystem: return fail(CBC_ERR_INVALID, "system index %d out of [0, %d)", bad, s->rh.ns);
    case cbc::kRingBusySystem: return fail(CBC_ERR_INVALID, "system %d already has an unfinished item", bad);
    case cbc::kRingDuplicate: return fail(CBC_ERR_INVALID, "system %d appears twice in one submission", bad);
    case cbc::kRingBadOptions: return fail(CBC_ERR_INVALID, "unknown item option bits 0x%x", (unsigned)bad);
    }
    __atomic_store_n(&s->ctl[0], s->rh.tail, __ATOMIC_RELEASE);
    if (!serve_running(s)) return serve_launch(s);
    return CBC_OK;
}

int32_t cbc_serve_poll(cbc_serve* s, int32_t* pending)
{
    if (!s) return fail(CBC_ERR_INVALID, "cbc_serve_poll: null handle");
    serve_retire(s);
    const uint32_t left = s->rh.tail - s->rh.retired;
    if (pending) *pending = (int32_t)left;
    if (left > 0 && !serve_running(s)) {
        // the kernel has ended (idle timeout) with items it never claimed;
        // retire again first: items it finished are visible once it has ended
        serve_retire(s);
        if (s->rh.tail != s->rh.retired) return serve_launch(s);
    }
    return CBC_OK;
}

int32_t cbc_serve_stats(const cbc_serve* s, int64_t* launches, uint32_t* published, uint32_t* retired)
{
    if (!s) return fail(CBC_ERR_INVALID, "cbc_serve_stats: null handle");
    if (launches) *launches = s->launches;
    if (published) *published = s->rh.tail;
    if (retired) *retired = s->rh.retired;
    return CBC_OK;
}

int32_t cbc_serve_destroy(cbc_serve* s)
{
    if (!s) return CBC_OK;
    __atomic_store_n(&s->ctl[1], 1u, __ATOMIC_RELEASE);
    hipError_t e = hipSuccess;
    if (s->launched) e = hipEventSynchronize(s->ended);
    serve_free(s);
    if (e != hipSuccess) return hip_fail(e, "cbc_serve_destroy: waiting for the serving kernel");
    return CBC_OK;
}

}  // extern "C"
