// cbc_api.hip -- the C ABI of libcbc.so (declared in include/cbc.h).
//
// Replaces the reference's backend layer: the kernel loader
// (backends/cupy_cuda/kernel_loader.py:9-99, models/ckm1d.py:214-253) becomes
// cbc_plan_create, the host launch loop (rkck45_coupled_cupy_solver.py:29-80)
// becomes cbc_solve.  No exception crosses this boundary.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>

#include "../../include/cbc.h"
#include "cbc_kernels.h"

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

        // AUTO: whole system in one lane while it fits in VGPRs (N <= 2), else one
        // lane per bubble with the coupling solve staged in LDS
        int mapping = exec->mapping;
        if (mapping == CBC_MAP_AUTO) mapping = (exec->ups <= 2) ? CBC_MAP_LANE : CBC_MAP_GROUP;
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
