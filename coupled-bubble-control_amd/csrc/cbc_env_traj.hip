// cbc_env_traj.hip -- per-episode trajectory collection on the device
// (TrajectorBuffer, coupledbubble_control/envs/common/trajectory_buffer.py:70-166).
//
// The reference copies observations, actions, rewards and the whole dense-output
// buffers to the host after every step and appends them to per-env Python
// lists, then zeroes the dense buffers; at an episode's end it merges the lists
// (Trajectory.merge_list, :38-56) and writes them.  Here every env's running
// episode stays in a device arena:
//
//   traj_append  after each step: the step's observation, action and reward
//                rows, and its dense points -- the first point dropped for every
//                step after the first and the times offset by k * step_length,
//                as merge_list concatenates them -- converted to float32 (the
//                dtype the writer stores); then dense_index := 0.
//   traj_gather  at episode ends: the finished envs' arenas packed back to back
//                in done-list order for one device-to-host copy; their counters
//                cleared.
//
// Compiled with -ffp-contract=off: the time offset is t + (k * step_length),
// two rounded operations as in NumPy.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>

#include "../../include/cbc.h"

namespace cbc {
int32_t report_error(int32_t code, const char* msg);
}

namespace {

constexpr int kAppendThreads = 256;   // 4 waves of 8 envs x 8 points
constexpr int kEnvsPerWave = 8;
constexpr int kGatherThreads = 256;   // one workgroup per finished env

int32_t fail(int32_t code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    return cbc::report_error(code, buf);
}

int32_t hip_fail(hipError_t e, const char* what)
{
    return fail(CBC_ERR_HIP, "%s: %s (%d)", what, hipGetErrorString(e), (int)e);
}

struct GatherArgs {
    cbc_traj_spec s;
    cbc_traj_buffers b;
    const int64_t* ids;
    const int64_t* step_off;    // (n,) first step row of env i in the packed output
    const int64_t* point_off;   // (n,) first point of env i
    cbc_traj_packed out;
};

// A wave covers 8 envs x 8 points: lane = point * 8 + env.  For one point the
// 8 lanes of consecutive envs read 64 contiguous bytes of the (NDO, NS) float64
// dense buffers; the 8 lanes of one env write 8 consecutive arena points.
__global__ __launch_bounds__(kAppendThreads) void traj_append(cbc_traj_spec s, cbc_traj_buffers b)
{
    const int lane = threadIdx.x & 63;
    const int sub = lane & (kEnvsPerWave - 1);          // env within the wave
    const int pt = lane >> 3;                            // point within a chunk of 8
    const int e = (blockIdx.x * (kAppendThreads / 64) + (threadIdx.x >> 6)) * kEnvsPerWave + sub;
    if (e >= s.num_envs) return;
    const size_t NE = (size_t)s.num_envs, N = (size_t)s.num_bubbles, TILE = (size_t)s.tile;
    const int k = b.steps[e];
    const int64_t p0 = b.points[e];
    const int npts = min(b.dense_index[e], s.ndo);
    const int first = k == 0 ? 0 : 1;
    const int add = npts > first ? npts - first : 0;
    const bool fits = k < s.max_steps && p0 + add <= s.max_points;
    if (fits) {
        const size_t row = (size_t)e * s.max_steps + k;
        for (int c = pt; c < s.obs_dim; c += 8) b.obs[row * s.obs_dim + c] = b.observations[(size_t)e * s.obs_dim + c];
        for (int c = pt; c < s.num_actions; c += 8)
            b.act[row * s.num_actions + c] = b.actions[(size_t)e * s.num_actions + c];
        if (pt == 0) b.rew[row] = b.rewards[e];
        const double toff = (double)k * s.step_length;
        float* tdst = b.time + (size_t)e * s.max_points + p0;
        for (int q = pt; q < add; q += 8) {
            const double t = b.dense_time[(size_t)(first + q) * NE + e];
            tdst[q] = (float)(k == 0 ? t : t + toff);
        }
        for (int j = 0; j < s.num_saved; ++j) {
            float* d = b.states + (((size_t)e * s.num_saved + j) * s.max_points + p0) * N;
            for (int q = pt; q < add; q += 8) {
                const double* src = b.dense_state + (((size_t)(first + q) * 4 + s.saved_rows[j]) * NE + e) * TILE;
                for (size_t i = 0; i < N; ++i) d[(size_t)q * N + i] = (float)src[i];
            }
        }
    }
    if (pt == 0) {   // the counters were read above by every lane of this env (same wave)
        if (fits) {
            b.steps[e] = k + 1;
            b.points[e] = p0 + add;
        } else {
            b.overflow[0] = 1;
        }
        b.dense_index[e] = 0;
    }
}

__global__ __launch_bounds__(kGatherThreads) void traj_gather(GatherArgs a)
{
    const cbc_traj_spec& s = a.s;
    const cbc_traj_buffers& b = a.b;
    const int i = blockIdx.x;
    const int64_t e64 = a.ids[i];
    if (e64 < 0 || e64 >= s.num_envs) return;
    const size_t e = (size_t)e64;
    const int T = b.steps[e];
    const int64_t P = b.points[e];
    const size_t so = (size_t)a.step_off[i], po = (size_t)a.point_off[i];
    const size_t N = (size_t)s.num_bubbles;
    for (size_t x = threadIdx.x; x < (size_t)T * s.obs_dim; x += kGatherThreads)
        a.out.obs[so * s.obs_dim + x] = b.obs[e * s.max_steps * s.obs_dim + x];
    for (size_t x = threadIdx.x; x < (size_t)T * s.num_actions; x += kGatherThreads)
        a.out.act[so * s.num_actions + x] = b.act[e * s.max_steps * s.num_actions + x];
    for (size_t x = threadIdx.x; x < (size_t)T; x += kGatherThreads) a.out.rew[so + x] = b.rew[e * s.max_steps + x];
    for (size_t x = threadIdx.x; x < (size_t)P; x += kGatherThreads)
        a.out.time[po + x] = b.time[e * s.max_points + x];
    // states: saved row j of env i occupies P * N floats after the previous rows of all envs
    for (int j = 0; j < s.num_saved; ++j) {
        const float* src = b.states + (e * s.num_saved + j) * s.max_points * N;
        float* dst = a.out.states + ((size_t)j * a.out.total_points + po) * N;
        for (size_t x = threadIdx.x; x < (size_t)P * N; x += kGatherThreads) dst[x] = src[x];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        b.steps[e] = 0;
        b.points[e] = 0;
    }
}

int32_t check(const cbc_traj_spec* s, const cbc_traj_buffers* b)
{
    if (!s || !b) return fail(CBC_ERR_INVALID, "traj: null spec or buffers");
    if (s->num_envs < 1 || s->num_bubbles < 1 || s->tile < s->num_bubbles || s->ndo < 1)
        return fail(CBC_ERR_INVALID, "traj: bad sizes (num_envs %d, num_bubbles %d, tile %d, ndo %d)", s->num_envs,
                    s->num_bubbles, s->tile, s->ndo);
    if (s->num_saved < 0 || s->num_saved > 4) return fail(CBC_ERR_INVALID, "traj: num_saved must be in [0, 4]");
    for (int j = 0; j < s->num_saved; ++j)
        if (s->saved_rows[j] < 0 || s->saved_rows[j] > 3)
            return fail(CBC_ERR_INVALID, "traj: saved row %d out of [0, 3]", s->saved_rows[j]);
    if (s->obs_dim < 0 || s->num_actions < 0 || s->max_steps < 1 || s->max_points < 1)
        return fail(CBC_ERR_INVALID, "traj: bad arena sizes");
    if (!b->steps || !b->points || !b->obs || !b->act || !b->rew || !b->time || (s->num_saved && !b->states))
        return fail(CBC_ERR_INVALID, "traj: null arena buffer");
    return CBC_OK;
}

}  // namespace

extern "C" {

int32_t cbc_traj_append(const cbc_traj_spec* spec, const cbc_traj_buffers* buf, void* stream)
{
    int32_t rc = check(spec, buf);
    if (rc != CBC_OK) return rc;
    if (!buf->observations || !buf->actions || !buf->rewards || !buf->dense_index || !buf->dense_time ||
        !buf->dense_state || !buf->overflow)
        return fail(CBC_ERR_INVALID, "traj_append: null step buffer");
    cbc_traj_spec s = *spec;
    cbc_traj_buffers b = *buf;
    void* args[] = {&s, &b};
    hipError_t e = hipLaunchKernel(reinterpret_cast<const void*>(&traj_append),
                                   dim3((s.num_envs + 4 * kEnvsPerWave - 1) / (4 * kEnvsPerWave)), dim3(kAppendThreads),
                                   args, 0, (hipStream_t)stream);
    return e == hipSuccess ? CBC_OK : hip_fail(e, "traj_append launch");
}

int32_t cbc_traj_gather(const cbc_traj_spec* spec, const cbc_traj_buffers* buf, const int64_t* ids, int32_t n,
                        const int64_t* step_off, const int64_t* point_off, const cbc_traj_packed* out, void* stream)
{
    int32_t rc = check(spec, buf);
    if (rc != CBC_OK) return rc;
    if (n < 0 || n > spec->num_envs) return fail(CBC_ERR_INVALID, "traj_gather: n = %d out of range", n);
    if (n == 0) return CBC_OK;
    if (!ids || !step_off || !point_off || !out || !out->obs || !out->act || !out->rew || !out->time ||
        (spec->num_saved && !out->states))
        return fail(CBC_ERR_INVALID, "traj_gather: null argument");
    GatherArgs a{*spec, *buf, ids, step_off, point_off, *out};
    void* args[] = {&a};
    hipError_t e = hipLaunchKernel(reinterpret_cast<const void*>(&traj_gather), dim3(n), dim3(kGatherThreads), args,
                                   0, (hipStream_t)stream);
    return e == hipSuccess ? CBC_OK : hip_fail(e, "traj_gather launch");
}

}  // extern "C"
