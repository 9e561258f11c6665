// cbc_env.hip -- the vector environment's step around the solver (PosNBC1D,
// coupledbubble_control/envs/bubble_1D_pos_control.py:503-853).
//
// The reference runs set_action, observe, check_termination, get_rewards,
// advance_time, final_termination and reset_envs as some forty torch kernels
// and three host synchronisations per step.  Here they are four kernels, one
// thread per environment:
//
//   env_set_action  actions -> control_params rows                     (:732-749)
//   env_post        observe + check_termination + get_rewards
//                   + advance_time + done counts and offsets             (:756-842)
//   env_compact     the ordered done list from the per-workgroup ids
//   env_reset       final outputs of the done envs + reset_envs with the
//                   scene arithmetic of SceneGenerator                   (:112-249, 583-614)
//
// Every access is a few contiguous words per environment (state rows,
// observation rows), so the kernels are launch- and latency-bound; they exist
// to remove launches and host round trips, not FLOPs.
//
// The float32 arithmetic of observations and rewards is the reference's
// operation by operation (elementwise torch kernels round every operation):
// this file is compiled with -ffp-contract=off so that no product and sum are
// fused.  Reductions over the bubbles run in index order; torch.mean is taken
// as sum * (1/N), the form of its reduction kernel.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <mutex>

#include "../../include/cbc.h"

namespace cbc {
int32_t report_error(int32_t code, const char* msg);
}

namespace {

constexpr int kBlock = 256;                   // threads per workgroup (4 waves)
constexpr int kWaves = kBlock / 64;
constexpr int kMaxEnvs = 1 << 20;             // 4096 workgroups: 16 per scanning thread
constexpr int kMaxStack = 64;
// workspace words: [0] ticket, [1..4] counts (done, failures, successes,
// time-outs), [8 .. 8 + 4 nb) per-workgroup counts, [8 + 4 nb .. 8 + 5 nb)
// per-workgroup offsets of the done list, then nb * kBlock int32 local ids
constexpr int kWsCounts = 1;
constexpr int kWsBlocks = 8;

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

int n_blocks(int ne) { return (ne + kBlock - 1) / kBlock; }

struct PostArgs {
    cbc_env_spec s;
    cbc_env_buffers b;
    int64_t observed;
};

struct ResetArgs {
    cbc_env_spec s;
    cbc_env_buffers b;
    cbc_env_scene sc;
    const int64_t* ids;
    int32_t n;
    int64_t observed;
    float* final_obs;
    float* final_return;
    int64_t* final_length;
};

// ---------------------------------------------------------------------------
// device helpers

// torch.pow(x, e) for a float tensor and a Python scalar: the exponents that
// ATen evaluates by dedicated formulas, powf otherwise.
__device__ __forceinline__ float pow_scalar(float x, float e)
{
    if (e == 0.5f) return sqrtf(x);
    if (e == 1.0f) return x;
    if (e == 2.0f) return x * x;
    if (e == 3.0f) return x * x * x;
    if (e == -0.5f) return 1.0f / sqrtf(x);
    if (e == -1.0f) return 1.0f / x;
    if (e == -2.0f) return 1.0f / (x * x);
    return powf(x, e);
}

__device__ __forceinline__ const float* part_row(const cbc_env_spec& s, const float* hist, int p, int slot, int e)
{
    const size_t NE = (size_t)s.num_envs, N = (size_t)s.num_bubbles;
    return hist + s.part_hist[p] + ((size_t)slot * NE + (size_t)e) * N;
}

// Observation row of env e (ObservationBufferContainer.get_stacked then
// Box.scale, buffers.py:75-83, 131-135; spaces.py:116-117): parts in order,
// each bubble-major with its ring oldest first.  `count` is the number of
// observe() calls so far: a ring of depth st starts at slot count % st.
__device__ void build_observation(const cbc_env_spec& s, const cbc_env_buffers& b, int e, int64_t count)
{
    const int N = s.num_bubbles;
    float* o = b.observations + (size_t)e * (size_t)s.obs_dim;
    int col = 0;
    for (int p = 0; p < s.num_parts; ++p) {
        const int st = s.part_stack[p];
        const int first = (int)(count % st);
        for (int i = 0; i < N; ++i) {
            for (int h = 0; h < st; ++h) {
                int slot = first + h;
                if (slot >= st) slot -= st;
                const float v = part_row(s, b.history, p, slot, e)[i];
                o[col] = (v - b.obs_bias[col]) * b.obs_rscale[col];
                ++col;
            }
        }
    }
}

// SceneGenerator positions of scene i for one side (bubble_1D_pos_control.py:139-249).
__device__ void scene_positions(const cbc_env_scene& sc, const cbc_scene_side& sd, int i, int N, double* pos)
{
    if (sd.method == CBC_SCENE_RANDOM) {
        // _random_uniform_position_1D (:183-242): extras ~ U(0, room), scaled to fit
        const int G = N - 1;
        double ex[CBC_ENV_MAX_BUBBLES];
        double s = 0.0;
        for (int g = 0; g < G; ++g) {
            ex[g] = sd.u_gaps[(size_t)i * G + g] * sc.room_per_gap;
            s = s + ex[g];
        }
        const double denom = s < 1e-12 ? 1e-12 : s;
        double scale = sc.extra_gap_limit / denom;
        scale = scale > 1.0 ? 1.0 : scale;
        double gsum = 0.0;
        for (int g = 0; g < G; ++g) {
            ex[g] = ex[g] * scale;                 // extras = extras_raw * scale
            ex[g] = ex[g] + sc.min_distance;       // gaps = extras + min_distance
            gsum = gsum + ex[g];
        }
        const double remaining = sc.available - gsum;
        const double off = sd.random_align ? sc.lo + sd.u_align[i] * remaining : sc.lo + 0.5 * remaining;
        pos[0] = off;
        double c = 0.0;
        for (int g = 0; g < G; ++g) {
            c = c + ex[g];
            pos[g + 1] = off + c;
        }
    } else if (sd.method == CBC_SCENE_EQUIDISTANT) {
        // _equidistant_position_1D (:139-181): base + offset
        const double off = sd.random_align ? sc.lo + sd.u_align[i] * (sd.max_start - sc.lo) : sd.offset;
        for (int b = 0; b < N; ++b) pos[b] = sd.base[b] + off;
    } else {
        // _expand_positions_1D (:244-249)
        for (int b = 0; b < N; ++b) pos[b] = sd.base[b];
    }
}

__device__ __forceinline__ uint32_t popc64(uint64_t m) { return (uint32_t)__popcll(m); }

// ---------------------------------------------------------------------------
// kernels

// envs a launch covers: [env_begin, env_begin + env_count), all when env_count is 0
__host__ __device__ __forceinline__ int range_count(const cbc_env_spec& s)
{
    return s.env_count > 0 ? s.env_count : s.num_envs;
}

__global__ __launch_bounds__(kBlock) void env_set_action(cbc_env_spec s, const float* actions, double* cp)
{
    const int k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= range_count(s)) return;
    const int e = s.env_begin + k;
    const size_t NE = (size_t)s.num_envs;
    const float* a = actions + (size_t)e * s.num_actions;
    for (int c = 0; c < s.num_actions; ++c) cp[(size_t)s.action_row[c] * NE + e] = (double)a[c] * s.action_scale[c];
}

__global__ __launch_bounds__(kBlock) void env_post(PostArgs a)
{
    const cbc_env_spec& s = a.s;
    const cbc_env_buffers& b = a.b;
    const int k = blockIdx.x * kBlock + threadIdx.x;
    const bool live = k < range_count(s);
    const int e = s.env_begin + k;
    bool f_fail = false, f_succ = false, f_tout = false;
    if (live) {
        const int N = s.num_bubbles;
        const size_t NE = (size_t)s.num_envs;
        // observe (:756-769): the STATE rows into their rings as float32
        for (int p = 0; p < s.num_parts; ++p) {
            if (s.part_kind[p] != CBC_ENV_PART_STATE) continue;
            const double* row = b.actual_state + ((size_t)s.part_row[p] * NE + (size_t)e) * (size_t)s.tile;
            float* dst = b.history + s.part_hist[p] +
                         ((size_t)(a.observed % s.part_stack[p]) * NE + (size_t)e) * (size_t)N;
            for (int i = 0; i < N; ++i) dst[i] = (float)row[i];
        }
        build_observation(s, b, e, a.observed + 1);

        // check_termination (:772-789)
        f_fail = b.status_flags[e] != CBC_SUCCESS;
        b.status_flags[e] = CBC_ACTIVE;
        const float* x = part_row(s, b.history, s.x_part, (int)(a.observed % s.part_stack[s.x_part]), e);
        const float* xt = part_row(s, b.history, s.target_part, 0, e);
        bool all_in = true;
        float sum_d = 0.0f;
        for (int i = 0; i < N; ++i) {
            const float td = fabsf(xt[i] - x[i]);
            all_in = all_in && (td <= s.position_tolerance);
            // get_rewards (:792-800): masked, normalised, clamped, powered
            float m = td < s.position_tolerance ? 0.0f : td;
            float nd = m * s.r_dnorm;
            nd = nd < 0.0f ? 0.0f : nd;
            sum_d = sum_d + pow_scalar(nd, s.reward_exp);
        }
        f_succ = s.apply_termination && all_in;
        const float distance = sum_d * (1.0f / (float)N);
        // actuation penalty (:802-804)
        const float* act = b.actions + (size_t)e * s.num_actions;
        float sum_a = 0.0f;
        for (int c = 0; c < s.num_pa; ++c) {
            const float v = act[s.pa_col[c]];
            sum_a = sum_a + v * v;
        }
        const float intensity = sqrtf(sum_a);
        // weighted reward (:806-813)
        float r = (-s.w_distance) * distance;
        r = r - s.w_intensity * intensity;
        r = r + s.negative_terminal_reward * (f_fail ? 1.0f : 0.0f);
        r = r + s.positive_terminal_reward * (f_succ ? 1.0f : 0.0f);
        b.rewards[e] = r;
        b.total_rewards[e] = b.total_rewards[e] + r;

        // advance_time (:816-827)
        b.actual_time[e] = 0.0;
        if (s.next_time_step > 0.0) b.time_step[e] = s.next_time_step;
        b.actual_event[e] = -1.0;
        const int64_t steps = b.algo_steps[e] + 1;
        b.algo_steps[e] = steps;
        f_tout = steps >= s.episode_length;
        b.failure[e] = f_fail;
        b.success[e] = f_succ;
        b.terminated[e] = f_fail || f_succ;
        b.time_outs[e] = f_tout;
    }

    // final_termination's done list (:832): local ids per workgroup; the last
    // workgroup to finish turns the counts into offsets
    const bool done = live && (f_fail || f_succ || f_tout);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t m_done = __ballot(done);
    const uint64_t m_fail = __ballot(live && f_fail);
    const uint64_t m_succ = __ballot(live && f_succ);
    const uint64_t m_tout = __ballot(live && f_tout);
    __shared__ uint32_t wcnt[4][kWaves];
    __shared__ uint32_t last;
    if (lane == 0) {
        wcnt[0][wave] = popc64(m_done);
        wcnt[1][wave] = popc64(m_fail);
        wcnt[2][wave] = popc64(m_succ);
        wcnt[3][wave] = popc64(m_tout);
    }
    __syncthreads();
    uint32_t* ws = b.workspace;
    const int nb = gridDim.x;
    int32_t* local = reinterpret_cast<int32_t*>(ws + kWsBlocks + 5 * nb);
    if (done) {
        uint32_t before = 0;
        for (int w = 0; w < wave; ++w) before += wcnt[0][w];
        const uint64_t below = lane ? (m_done << (64 - lane)) : 0ull;   // lanes < lane
        local[(size_t)blockIdx.x * kBlock + before + popc64(below)] = e;
    }
    if (threadIdx.x < 4) {
        uint32_t c = 0;
        for (int w = 0; w < kWaves; ++w) c += wcnt[threadIdx.x][w];
        ws[kWsBlocks + 4 * blockIdx.x + threadIdx.x] = c;
    }
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) last = atomicAdd(ws, 1u) == (uint32_t)(nb - 1);
    __syncthreads();
    if (!last) return;
    __threadfence();

    // the last workgroup: exclusive scan of the per-workgroup done counts into
    // per-workgroup offsets (env_compact then moves every workgroup's ids)
    __shared__ uint32_t part[kBlock];
    __shared__ uint32_t tot[4];
    const int q = (nb + kBlock - 1) / kBlock;
    const int b0 = threadIdx.x * q;
    const int b1 = min(nb, b0 + q);
    uint32_t sums[4] = {0u, 0u, 0u, 0u};
    for (int bb = b0; bb < b1; ++bb)
        for (int j = 0; j < 4; ++j) sums[j] += ws[kWsBlocks + 4 * bb + j];
    part[threadIdx.x] = sums[0];
    if (threadIdx.x < 4) tot[threadIdx.x] = 0u;
    __syncthreads();
    for (int j = 1; j < 4; ++j) atomicAdd(&tot[j], sums[j]);
    // inclusive Hillis--Steele scan of the per-thread done sums
    for (int d = 1; d < kBlock; d <<= 1) {
        const uint32_t v = threadIdx.x >= d ? part[threadIdx.x - d] : 0u;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t run = part[threadIdx.x] - sums[0];
    uint32_t* offs = ws + kWsBlocks + 4 * nb;
    for (int bb = b0; bb < b1; ++bb) {
        offs[bb] = run;
        run += ws[kWsBlocks + 4 * bb];
    }
    if (threadIdx.x == 0) {
        ws[kWsCounts + 0] = part[kBlock - 1];
        ws[kWsCounts + 1] = tot[1];
        ws[kWsCounts + 2] = tot[2];
        ws[kWsCounts + 3] = tot[3];
        ws[0] = 0u;   // ticket for the next launch (stream-ordered after this one)
    }
}

// Second half of the done list: workgroup b of env_post wrote its ids in
// ascending order at local[b * kBlock ..); they move to done_ids at the
// workgroup's offset.  One workgroup per env_post workgroup, no serial chain.
__global__ __launch_bounds__(kBlock) void env_compact(const uint32_t* ws, int nb, int64_t* done_ids)
{
    const int bb = blockIdx.x;
    const uint32_t c = ws[kWsBlocks + 4 * bb];
    if ((uint32_t)threadIdx.x >= c) return;
    const uint32_t o = ws[kWsBlocks + 4 * nb + bb];
    const int32_t* local = reinterpret_cast<const int32_t*>(ws + kWsBlocks + 5 * nb);
    done_ids[o + threadIdx.x] = (int64_t)local[(size_t)bb * kBlock + threadIdx.x];
}

__global__ __launch_bounds__(kBlock) void env_reset(ResetArgs a)
{
    const cbc_env_spec& s = a.s;
    const cbc_env_buffers& b = a.b;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= a.n) return;
    const int64_t e64 = a.ids[i];
    if (e64 < 0 || e64 >= s.num_envs) return;
    const int e = (int)e64;
    const int N = s.num_bubbles;
    const size_t NE = (size_t)s.num_envs;
    // final_termination (:833-840): taken before the reset
    if (a.final_obs) {
        const float* o = b.observations + (size_t)e * s.obs_dim;
        float* f = a.final_obs + (size_t)i * s.obs_dim;
        for (int c = 0; c < s.obs_dim; ++c) f[c] = o[c];
    }
    if (a.final_return) a.final_return[i] = b.total_rewards[e];
    if (a.final_length) a.final_length[i] = b.algo_steps[e];

    double pos[CBC_ENV_MAX_BUBBLES], tgt[CBC_ENV_MAX_BUBBLES];
    scene_positions(a.sc, a.sc.bubbles, i, N, pos);
    scene_positions(a.sc, a.sc.targets, i, N, tgt);
    // reset_envs (:596-613): R = 1, X = positions, U = V = 0; rings and targets
    const size_t T = (size_t)s.tile;
    double* st = b.actual_state;
    for (int k = 0; k < N; ++k) {
        st[(0 * NE + e) * T + k] = 1.0;
        st[(1 * NE + e) * T + k] = pos[k];
        st[(2 * NE + e) * T + k] = 0.0;
        st[(3 * NE + e) * T + k] = 0.0;
    }
    const int sx = s.part_stack[s.x_part];
    for (int slot = 0; slot < sx; ++slot) {
        float* d = b.history + s.part_hist[s.x_part] + ((size_t)slot * NE + e) * (size_t)N;
        for (int k = 0; k < N; ++k) d[k] = (float)pos[k];
    }
    float* t = b.history + s.part_hist[s.target_part] + (size_t)e * (size_t)N;
    for (int k = 0; k < N; ++k) t[k] = (float)tgt[k];
    b.algo_steps[e] = 0;
    b.total_rewards[e] = 0.0f;
    build_observation(s, b, e, a.observed);
}

__global__ __launch_bounds__(kBlock) void env_scenes(cbc_env_scene sc, int n, int N, double* pos_out, double* tgt_out)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    double v[CBC_ENV_MAX_BUBBLES];
    if (pos_out) {
        scene_positions(sc, sc.bubbles, i, N, v);
        for (int k = 0; k < N; ++k) pos_out[(size_t)i * N + k] = v[k];
    }
    if (tgt_out) {
        scene_positions(sc, sc.targets, i, N, v);
        for (int k = 0; k < N; ++k) tgt_out[(size_t)i * N + k] = v[k];
    }
}

// ---------------------------------------------------------------------------
// host-side checks

int32_t check_spec(const cbc_env_spec* s)
{
    if (!s) return fail(CBC_ERR_INVALID, "env: null spec");
    if (s->num_envs < 1 || s->num_envs > kMaxEnvs)
        return fail(CBC_ERR_INVALID, "env: num_envs must be in [1, %d], got %d", kMaxEnvs, s->num_envs);
    if (s->env_begin < 0 || s->env_count < 0 || (int64_t)s->env_begin + s->env_count > s->num_envs ||
        (s->env_count == 0 && s->env_begin != 0))
        return fail(CBC_ERR_INVALID, "env: range [%d, %d + %d) outside the %d envs (env_count 0 needs env_begin 0)",
                    s->env_begin, s->env_begin, s->env_count, s->num_envs);
    if (s->num_bubbles < 1 || s->num_bubbles > CBC_ENV_MAX_BUBBLES)
        return fail(CBC_ERR_INVALID, "env: num_bubbles must be in [1, %d], got %d", CBC_ENV_MAX_BUBBLES,
                    s->num_bubbles);
    if (s->tile < s->num_bubbles) return fail(CBC_ERR_INVALID, "env: tile %d < num_bubbles %d", s->tile, s->num_bubbles);
    if (s->num_actions < 0 || s->num_actions > CBC_ENV_MAX_ACTIONS)
        return fail(CBC_ERR_INVALID, "env: num_actions must be in [0, %d]", CBC_ENV_MAX_ACTIONS);
    for (int c = 0; c < s->num_actions; ++c)
        if (s->action_row[c] < 0 || s->action_row[c] >= s->num_control_rows)
            return fail(CBC_ERR_INVALID, "env: action %d targets control row %d of %d", c, s->action_row[c],
                        s->num_control_rows);
    if (s->num_pa < 0 || s->num_pa > s->num_actions) return fail(CBC_ERR_INVALID, "env: num_pa out of range");
    for (int c = 0; c < s->num_pa; ++c)
        if (s->pa_col[c] < 0 || s->pa_col[c] >= s->num_actions)
            return fail(CBC_ERR_INVALID, "env: pa_col[%d] = %d out of range", c, s->pa_col[c]);
    if (s->num_parts < 1 || s->num_parts > CBC_ENV_MAX_PARTS)
        return fail(CBC_ERR_INVALID, "env: num_parts must be in [1, %d]", CBC_ENV_MAX_PARTS);
    const int64_t NE = s->num_envs, N = s->num_bubbles;
    int64_t dim = 0;
    for (int p = 0; p < s->num_parts; ++p) {
        const int k = s->part_kind[p], st = s->part_stack[p];
        if (k != CBC_ENV_PART_STATE && k != CBC_ENV_PART_TARGET && k != CBC_ENV_PART_FIXED)
            return fail(CBC_ERR_INVALID, "env: part %d has unknown kind %d", p, k);
        if (st < 1 || st > kMaxStack || (k == CBC_ENV_PART_TARGET && st != 1))
            return fail(CBC_ERR_INVALID, "env: part %d has stack %d", p, st);
        if (k == CBC_ENV_PART_STATE && (s->part_row[p] < 0 || s->part_row[p] > 3))
            return fail(CBC_ERR_INVALID, "env: part %d reads state row %d", p, s->part_row[p]);
        const int64_t span = (int64_t)st * NE * N;
        if (s->part_hist[p] < 0 || s->part_hist[p] + span > s->history_floats)
            return fail(CBC_ERR_INVALID, "env: part %d [%d, +%lld) exceeds history of %lld floats", p,
                        s->part_hist[p], (long long)span, (long long)s->history_floats);
        dim += (int64_t)st * N;
    }
    if (dim != s->obs_dim) return fail(CBC_ERR_INVALID, "env: obs_dim %d != parts' %lld", s->obs_dim, (long long)dim);
    if (s->x_part < 0 || s->x_part >= s->num_parts || s->part_kind[s->x_part] != CBC_ENV_PART_STATE ||
        s->part_row[s->x_part] != 1)
        return fail(CBC_ERR_INVALID, "env: x_part must be the STATE part of row 1 (X)");
    if (s->target_part < 0 || s->target_part >= s->num_parts || s->part_kind[s->target_part] != CBC_ENV_PART_TARGET)
        return fail(CBC_ERR_INVALID, "env: target_part must be a TARGET part (XT)");
    if (s->episode_length < 0) return fail(CBC_ERR_INVALID, "env: episode_length must be >= 0");
    return CBC_OK;
}

int32_t check_buffers(const cbc_env_buffers* b, bool post, bool need_workspace = true)
{
    if (!b) return fail(CBC_ERR_INVALID, "env: null buffers");
    if (!b->actual_state || !b->history || !b->observations || !b->obs_bias || !b->obs_rscale ||
        !b->total_rewards || !b->algo_steps)
        return fail(CBC_ERR_INVALID, "env: null state, history, observation or counter buffer");
    if (post && (!b->actual_time || !b->time_step || !b->actual_event || !b->status_flags || !b->actions ||
                 !b->rewards || !b->failure || !b->success || !b->terminated || !b->time_outs || !b->done_ids ||
                 (need_workspace && !b->workspace)))
        return fail(CBC_ERR_INVALID, "env: null buffer for the post-step kernel");
    return CBC_OK;
}

int32_t check_side(const cbc_scene_side& sd, int n, int N, const char* name)
{
    if (sd.method != CBC_SCENE_RANDOM && sd.method != CBC_SCENE_EQUIDISTANT && sd.method != CBC_SCENE_FIXED)
        return fail(CBC_ERR_INVALID, "env: %s scene method %d unknown", name, sd.method);
    if (n > 0 && sd.method == CBC_SCENE_RANDOM && N > 1 && !sd.u_gaps)
        return fail(CBC_ERR_INVALID, "env: %s random scenes need u_gaps", name);
    if (n > 0 && sd.method != CBC_SCENE_FIXED && sd.random_align && !sd.u_align)
        return fail(CBC_ERR_INVALID, "env: %s random alignment needs u_align", name);
    return CBC_OK;
}

// ---- the workspace of cbc_env_post_static -------------------------------------
// A file-scope device array sized for CBC_ENV_STATIC_MAX_ENVS envs (4.1 MiB), so the
// call needs no workspace argument and allocates nothing.  In host code `scratch`
// names a host-side stand-in, not device memory: its device address on each GPU comes
// from hipGetSymbolAddress, once per device, and only that address is ever used.
constexpr int kStaticBlocks = (CBC_ENV_STATIC_MAX_ENVS + kBlock - 1) / kBlock;
constexpr int64_t kStaticWords = (int64_t)kWsBlocks + 5 * (int64_t)kStaticBlocks + (int64_t)kStaticBlocks * kBlock;
static_assert(CBC_ENV_STATIC_MAX_ENVS <= kMaxEnvs, "the static workspace serves at most the kernels' env limit");
static __device__ __align__(256) uint32_t scratch[kStaticWords];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint32_t* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device (looked up on the first call there).
int32_t static_workspace(uint32_t** out)
{
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) return fail(CBC_ERR_UNSUPPORTED, "env_post_static: device %d beyond %d", dev, kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(env scratch)");
        if (!p) return fail(CBC_ERR_HIP, "hipGetSymbolAddress(env scratch) returned null");
        g_scratch_ptr[dev] = static_cast<uint32_t*>(p);
    }
    *out = g_scratch_ptr[dev];
    return CBC_OK;
}

__global__ void zero_ticket(uint32_t* ws) { ws[0] = 0u; }

// cbc_env_post's launches and counts read-back on workspace `ws`.
int32_t env_post_run(const cbc_env_spec* spec, cbc_env_buffers b, uint32_t* ws, int64_t observed, hipStream_t st,
                     cbc_env_counts* counts)
{
    b.workspace = ws;
    PostArgs a{*spec, b, observed};
    void* args[] = {&a};
    hipError_t e = hipLaunchKernel(reinterpret_cast<const void*>(&env_post), dim3(n_blocks(range_count(*spec))),
                                   dim3(kBlock), args, 0, st);
    if (e != hipSuccess) return hip_fail(e, "env_post launch");
    {
        const uint32_t* cws = ws;
        int nb = n_blocks(range_count(*spec));
        int64_t* ids = b.done_ids;
        void* cargs[] = {&cws, &nb, &ids};
        e = hipLaunchKernel(reinterpret_cast<const void*>(&env_compact), dim3(nb), dim3(kBlock), cargs, 0, st);
        if (e != hipSuccess) return hip_fail(e, "env_compact launch");
    }
    if (!counts) return CBC_OK;
    uint32_t h[4];
    e = hipMemcpyAsync(h, ws + kWsCounts, sizeof(h), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(e, "env_post counts readback");
    counts->done = (int32_t)h[0];
    counts->failures = (int32_t)h[1];
    counts->successes = (int32_t)h[2];
    counts->time_outs = (int32_t)h[3];
    return CBC_OK;
}

}  // namespace

extern "C" {

int64_t cbc_env_workspace_words(int32_t num_envs)
{
    if (num_envs < 1) return 0;
    return (int64_t)kWsBlocks + 5 * (int64_t)n_blocks(num_envs) + (int64_t)n_blocks(num_envs) * kBlock;
}

int32_t cbc_env_set_action(const cbc_env_spec* spec, const cbc_env_buffers* buf, void* stream)
{
    int32_t rc = check_spec(spec);
    if (rc != CBC_OK) return rc;
    if (!buf || !buf->actions || !buf->control_params)
        return fail(CBC_ERR_INVALID, "env_set_action: null actions or control_params");
    if (spec->num_actions == 0) return CBC_OK;
    cbc_env_spec s = *spec;
    const float* act = buf->actions;
    double* cp = buf->control_params;
    void* args[] = {&s, &act, &cp};
    hipError_t e = hipLaunchKernel(reinterpret_cast<const void*>(&env_set_action), dim3(n_blocks(range_count(s))),
                                   dim3(kBlock), args, 0, (hipStream_t)stream);
    return e == hipSuccess ? CBC_OK : hip_fail(e, "env_set_action launch");
}

int32_t cbc_env_post(const cbc_env_spec* spec, const cbc_env_buffers* buf, int64_t observed, void* stream,
                     cbc_env_counts* counts)
{
    int32_t rc = check_spec(spec);
    if (rc != CBC_OK) return rc;
    rc = check_buffers(buf, true);
    if (rc != CBC_OK) return rc;
    if (observed < 0) return fail(CBC_ERR_INVALID, "env_post: observed must be >= 0");
    if (spec->num_pa > 0 && !buf->actions) return fail(CBC_ERR_INVALID, "env_post: null actions");
    return env_post_run(spec, *buf, buf->workspace, observed, (hipStream_t)stream, counts);
}

int32_t cbc_env_post_static(const cbc_env_spec* spec, const cbc_env_buffers* buf, int64_t observed, void* stream,
                            cbc_env_counts* counts)
{
    int32_t rc = check_spec(spec);
    if (rc != CBC_OK) return rc;
    if (spec->num_envs > CBC_ENV_STATIC_MAX_ENVS)
        return fail(CBC_ERR_INVALID, "env_post_static: %d envs exceed the static workspace's %d", spec->num_envs,
                    CBC_ENV_STATIC_MAX_ENVS);
    rc = check_buffers(buf, true, false);
    if (rc != CBC_OK) return rc;
    if (observed < 0) return fail(CBC_ERR_INVALID, "env_post_static: observed must be >= 0");
    if (spec->num_pa > 0 && !buf->actions) return fail(CBC_ERR_INVALID, "env_post_static: null actions");
    uint32_t* ws = nullptr;
    rc = static_workspace(&ws);
    if (rc != CBC_OK) return rc;
    hipStream_t st = (hipStream_t)stream;
    // the done-list ticket must start at zero (env_post's last workgroup leaves it zero,
    // but a call on this array may have been cut short); a kernel, not a memset, so the
    // call stays valid under stream capture
    void* zargs[] = {&ws};
    hipError_t e = hipLaunchKernel(reinterpret_cast<const void*>(&zero_ticket), dim3(1), dim3(1), zargs, 0, st);
    if (e != hipSuccess) return hip_fail(e, "env_post_static ticket reset");
    return env_post_run(spec, *buf, ws, observed, st, counts);
}

int32_t cbc_env_reset(const cbc_env_spec* spec, const cbc_env_buffers* buf, const cbc_env_scene* scene,
                      const int64_t* ids, int32_t n, int64_t observed, float* final_observation,
                      float* episode_return, int64_t* episode_length, void* stream)
{
    int32_t rc = check_spec(spec);
    if (rc != CBC_OK) return rc;
    rc = check_buffers(buf, false);
    if (rc != CBC_OK) return rc;
    if (!scene) return fail(CBC_ERR_INVALID, "env_reset: null scene");
    if (n < 0 || n > spec->num_envs) return fail(CBC_ERR_INVALID, "env_reset: n = %d out of [0, %d]", n, spec->num_envs);
    if (observed < 0) return fail(CBC_ERR_INVALID, "env_reset: observed must be >= 0");
    if (n == 0) return CBC_OK;
    if (!ids) return fail(CBC_ERR_INVALID, "env_reset: null ids");
    rc = check_side(scene->bubbles, n, spec->num_bubbles, "bubble");
    if (rc == CBC_OK) rc = check_side(scene->targets, n, spec->num_bubbles, "target");
    if (rc != CBC_OK) return rc;
    ResetArgs a{*spec, *buf, *scene, ids, n, observed, final_observation, episode_return, episode_length};
    void* args[] = {&a};
    hipError_t e = hipLaunchKernel(reinterpret_cast<const void*>(&env_reset), dim3(n_blocks(n)), dim3(kBlock), args,
                                   0, (hipStream_t)stream);
    return e == hipSuccess ? CBC_OK : hip_fail(e, "env_reset launch");
}

int32_t cbc_env_scene_positions(const cbc_env_scene* scene, int32_t n, int32_t num_bubbles, double* positions,
                                double* targets, void* stream)
{
    if (!scene) return fail(CBC_ERR_INVALID, "env_scene_positions: null scene");
    if (num_bubbles < 1 || num_bubbles > CBC_ENV_MAX_BUBBLES)
        return fail(CBC_ERR_INVALID, "env_scene_positions: num_bubbles must be in [1, %d]", CBC_ENV_MAX_BUBBLES);
    if (n < 0 || n > kMaxEnvs) return fail(CBC_ERR_INVALID, "env_scene_positions: n = %d out of range", n);
    if (n == 0 || (!positions && !targets)) return CBC_OK;
    int32_t rc = positions ? check_side(scene->bubbles, n, num_bubbles, "bubble") : CBC_OK;
    if (rc == CBC_OK && targets) rc = check_side(scene->targets, n, num_bubbles, "target");
    if (rc != CBC_OK) return rc;
    cbc_env_scene sc = *scene;
    int N = num_bubbles;
    void* args[] = {&sc, &n, &N, &positions, &targets};
    hipError_t e = hipLaunchKernel(reinterpret_cast<const void*>(&env_scenes), dim3(n_blocks(n)), dim3(kBlock), args,
                                   0, (hipStream_t)stream);
    return e == hipSuccess ? CBC_OK : hip_fail(e, "env_scene_positions launch");
}

}  // extern "C"
