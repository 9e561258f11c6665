/*
 * cbc.h -- C ABI of libcbc.so, the MI355X (gfx950) ensemble solver for the
 * coupled Keller--Miksis 1-D bubble model (CKM1D).
 *
 * Every entry point takes plain pointers and sizes, returns 0 on success or a
 * negative CBC_ERR_* code, and never lets an exception cross the boundary.
 * cbc_last_error() gives a readable message for the calling thread's last
 * failure.  Device buffers are owned by the caller (any HIP allocation, e.g.
 * torch tensors); a plan owns only its code selection and a small workspace.
 * Work is stream-ordered on the hipStream_t passed as `stream` (NULL = the
 * legacy default stream).  One host thread per plan.
 *
 * Buffer layouts are the reference's (coupledbubble_control/models/model.py:56-106),
 * all FP64 unless stated:
 *   actual_time, time_end, time_begin, time_step, actual_event   (NS,)
 *   status_flags                                                 (NS,)   int32
 *   actual_state                                                 (4, NS, TILE)
 *   unit_params                                                  (13, NS, TILE)
 *   system_params                                                (5, NS)
 *   control_params                                               (4*NK, NS)
 *   coupling_matrices                                            (3, UPS, NS, TILE)
 *   dense_index (NS,) int32, dense_time (NDO, NS), dense_state (NDO, 4, NS, TILE)
 * TILE = next power of two >= UPS (models/specs.py:47).
 */
#ifndef CBC_H
#define CBC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CBC_ABI_VERSION 5

/* return codes */
#define CBC_OK 0
#define CBC_ERR_INVALID (-1)      /* bad argument, shape or spec */
#define CBC_ERR_HIP (-2)          /* a HIP runtime call failed */
#define CBC_ERR_UNSUPPORTED (-3)  /* configuration has no kernel */
#define CBC_ERR_NOMEM (-4)        /* host or device allocation failed */
#define CBC_ERR_BUSY (-5)         /* serving queue full / admission refused */

/* per-system status flags; rkck45_coupled_numba_stepper.py:27-33 */
#define CBC_ACTIVE 0
#define CBC_SUCCESS 1
#define CBC_EVENT_TERMINAL 2
#define CBC_MAX_STEPS_REACHED (-1)
#define CBC_STEP_SIZE_LIMIT (-2)
#define CBC_CONVERGENCE_FAILURE (-3)
#define CBC_SINGULARITY_OR_NAN (-4)

/* acoustic field (ModelSpec.ac, models/specs.py:16) */
#define CBC_FIELD_CONST 0
#define CBC_FIELD_SW_N 1
#define CBC_FIELD_SW_A 2

/* system-to-hardware mapping */
#define CBC_MAP_AUTO 0   /* N = 1 lane; N = 2 group while its waves fit one per SIMD, else lane; N >= 3 group */
#define CBC_MAP_LANE 1   /* one system per lane, whole state in VGPRs (N = 1..4) */
#define CBC_MAP_GROUP 2  /* TILE lanes per system (16 for N = 5..8, a wave for N = 9..16), coupling solve across the group (N = 2..16) */

/* coupling solve */
#define CBC_LIN_DIRECT 0      /* LU with partial pivoting */

/* ModelSpec, models/specs.py:5-23 (name and ac are replaced by `field`) */
typedef struct cbc_model_spec {
    int32_t sd, nup, nsp, ncm, nk, ne, ncf, nct;
    int32_t field;
} cbc_model_spec;

/* ExecutionSpec, models/specs.py:26-52, plus the mapping request */
typedef struct cbc_exec_spec {
    int32_t ns, spb, ups, ndo, tile;
    int32_t mapping;       /* CBC_MAP_* */
    int32_t block_size;    /* threads per block, 0 = plan default */
} cbc_exec_spec;

/* Collision-event localisation.
 * BRACKET: the reference GPU / Numba rule (rkck45_coupled_numba_stepper.py:500-515,
 *   .cu:608-641): a step across a sign change of the gap is rejected and retried
 *   with h * 0.99 |g_old| / (|g_old| + |g_new|) until |g| < etol.
 * EXACT: the root of the gap as the reference's CPU path finds it with
 *   solve_ivp events (CKM1D_reference.py:702-714, 848-856): the rejected step is
 *   retried with h * theta, theta the root of the gap along the step's cubic
 *   Hermite interpolant of R and x (slopes U and V, the state's own derivatives),
 *   and a step that crosses with |g_new| < etol is accepted; with etol ~1e-12
 *   the run ends at the root to that tolerance.  Built for both mappings and
 *   every bubble count (for 8 and 16 bubbles per system in kernels of its own). */
#define CBC_EVENT_BRACKET 0
#define CBC_EVENT_EXACT 1

/* SolverSpec, models/specs.py:55-71, plus the collision threshold of
 * ckm1d_def.py:534, the solve selector and the event mode */
typedef struct cbc_solver_spec {
    double atol, rtol, min_step, max_step, etol;
    double lin_atol, lin_rtol, lin_relaxation;
    int32_t lin_max_iter;
    int32_t linsolve;       /* CBC_LIN_* */
    double col_threshold;
    int32_t event_mode;     /* CBC_EVENT_* */
} cbc_solver_spec;

/* The 14 kernel buffers of rkck45_coupled_cuda_sover.cu:206-226 plus the two
 * step counters of models/buffers.py:57-63 (int32, may be NULL). */
typedef struct cbc_buffers {
    double *actual_time;
    const double *time_end;
    const double *time_begin;
    double *time_step;
    double *actual_state;
    const double *unit_params;
    const double *system_params;
    const double *control_params;
    const double *coupling_matrices;
    int32_t *dense_index;
    double *dense_time;
    double *dense_state;
    double *actual_event;
    int32_t *status_flags;
    int32_t *total_steps;
    int32_t *rejected_steps;
} cbc_buffers;

/* Per-window statistics for parameter studies: windows of `period` tau;
 * windows with index < first_window are skipped; outputs are
 * (n_windows, NS, TILE): maximum R/R0 over the accepted states of the window
 * (including its first state) and the trapezoidal time-mean of x/lambda. */
typedef struct cbc_period_spec {
    double period;
    int32_t first_window;
    int32_t n_windows;
    double *max_r;
    double *mean_x;
    int32_t resume;       /* nonzero: a continuation launch of the same integration (a
                             budget-limited call repeated); systems whose flag is negative
                             (a failure stop) are left untouched */
} cbc_period_spec;

typedef struct cbc_plan_info {
    int32_t mapping;          /* CBC_MAP_LANE or CBC_MAP_GROUP */
    int32_t block_size;
    int32_t grid_size;
    int32_t systems_per_block;
    int32_t lds_bytes;
    int32_t num_regs;         /* from hipFuncGetAttributes */
    int32_t local_bytes;      /* private (spill) bytes per lane */
    int32_t max_threads;
} cbc_plan_info;

typedef struct cbc_plan cbc_plan;

/* ABI version of the loaded library (CBC_ABI_VERSION). */
int32_t cbc_abi_version(void);

/* Message of the calling thread's last failed call ("" if none). */
const char *cbc_last_error(void);

/* Selects and sizes the kernel for a model/execution/solver spec; replaces
 * load_coupled_kernel (backends/cupy_cuda/kernel_loader.py:9-99) and the
 * Numba make_model/make_kernel pair (models/ckm1d.py:214-232).  Allocates a
 * 16-byte device workspace (the active-system counter of cbc_solve and the
 * work-queue head of cbc_solve_queued). */
int32_t cbc_plan_create(const cbc_model_spec *model, const cbc_exec_spec *exec,
                        const cbc_solver_spec *solver, cbc_plan **out);

/* Reports the kernel chosen by a plan. */
int32_t cbc_plan_query(const cbc_plan *plan, cbc_plan_info *out);

void cbc_plan_destroy(cbc_plan *plan);

/* One launch of up to kernel_steps step attempts per system; the device
 * entry rkck45_coupled_solver (rkck45_coupled_cuda_sover.cu:206-846).
 * Asynchronous on `stream`. */
int32_t cbc_launch(cbc_plan *plan, const cbc_buffers *buf, int32_t kernel_steps, void *stream);

/* The host launch loop of rkck45_coupled_cupy_solver.py:29-80: up to
 * ceil(max_steps/kernel_steps) launches, stopping once no system is ACTIVE.
 * The stop test reads one device counter instead of every flag.  When
 * launch_times is not NULL it receives the wall seconds of each launch (at
 * most max_times entries).  *n_launches receives the launch count.  Returns
 * after the stream is synchronised. */
int32_t cbc_solve(cbc_plan *plan, const cbc_buffers *buf, int32_t kernel_steps, int64_t max_steps,
                  void *stream, double *launch_times, int32_t max_times, int32_t *n_launches);

/* One launch that integrates every system to time_end in windows of
 * period->period tau, recording window statistics (no counterpart in the
 * reference; its per-period outputs are what the parameter-study scripts
 * derive from dense output, scripts/verify_CKM1D_paramstudy.py:268-330).
 * max_steps bounds the attempts per system.  Asynchronous on `stream`. */
int32_t cbc_solve_periodic(cbc_plan *plan, const cbc_buffers *buf, const cbc_period_spec *period,
                           int64_t max_steps, void *stream);

/* Persistent launches over a device work queue (no counterpart in the
 * reference, whose blocks own fixed groups of systems, rkck45_coupled_cuda_sover.cu:228-239):
 * lanes (lane mapping) or lane groups (group mapping) take system indices from
 * `order` (device int32[order_len], a permutation of a subset of 0..ns-1, e.g.
 * longest horizon first) and refill as their systems stop, so uneven horizons
 * and step counts do not leave lanes idle.  `waves` is the persistent grid in
 * wavefronts (0 = every SIMD filled at the kernel's register occupancy);
 * `refill` (1..64) is how many lanes (groups) of a wave must be free before the
 * wave takes new work.  Each launch gives every system at most kernel_steps
 * attempts; launches repeat while systems remain ACTIVE, up to
 * ceil(max_steps/kernel_steps) launches.  Returns after the stream is
 * synchronised; *n_launches receives the launch count. */
int32_t cbc_solve_queued(cbc_plan *plan, const cbc_buffers *buf, const int32_t *order, int32_t order_len,
                         int32_t waves, int32_t refill, int32_t kernel_steps, int64_t max_steps,
                         void *stream, int32_t *n_launches);

/* One asynchronous persistent launch of the queued solve (no relaunch, no
 * synchronisation): the building block for running several plans (e.g. one
 * per bubble count) concurrently on separate streams and for stream capture
 * into a HIP graph (the launch sequence contains no memset node).  The number
 * of systems left ACTIVE is read with cbc_plan_active_count. */
int32_t cbc_launch_queued(cbc_plan *plan, const cbc_buffers *buf, const int32_t *order, int32_t order_len,
                          int32_t waves, int32_t refill, int32_t kernel_steps, void *stream);

/* The queued launches of n_bins plans with different bubble counts fused into
 * ONE persistent kernel (no counterpart in the reference, which compiles one
 * kernel per bubble count, backends/cupy_cuda/kernel_loader.py:49, and so can
 * integrate only one N per launch).  Bin b is plans[b] over bufs[b], with work
 * queue orders[b][0..order_lens[b]) and refill threshold refills[b] as in
 * cbc_launch_queued.  Every wave of the grid drains the bins' queues in turn,
 * largest bubble count first; each bin's systems step exactly as in its own
 * queued launch (bit-identical results).  All plans must use the group mapping,
 * the same harmonic count (1 or 2) and field type, and distinct bubble counts
 * that one fused kernel covers (cbc_fused_supported: the lists {8, 4, 3, 2} and
 * {16, 2}); otherwise CBC_ERR_UNSUPPORTED or CBC_ERR_INVALID, with nothing
 * launched.  `waves` = 0: one wave per SIMD.  Asynchronous and capturable; each
 * plan's count of systems left ACTIVE is read with cbc_plan_active_count. */
int32_t cbc_launch_fused(cbc_plan *const *plans, const cbc_buffers *const *bufs, const int32_t *const *orders,
                         const int32_t *order_lens, const int32_t *refills, int32_t n_bins, int32_t waves,
                         int32_t kernel_steps, void *stream);

/* 1 if a fused kernel covers the distinct bubble counts ups[0..n_bins) at
 * harmonic count nk and field type `field` (CBC_FIELD_*), else 0; CBC_ERR_INVALID
 * for a malformed list. */
int32_t cbc_fused_supported(const int32_t *ups, int32_t n_bins, int32_t nk, int32_t field);

/* Synchronises `stream` and returns the number of systems the plan's last
 * launch left ACTIVE. */
int32_t cbc_plan_active_count(cbc_plan *plan, void *stream, int32_t *out);

/* ---- serving ------------------------------------------------------------
 * The vector environment's step (set_action, step_physics, advance_time of
 * envs/bubble_1D_pos_control.py:731-753, 815-826) as a stream of work items
 * served by one persistent kernel per plan, with every system's state and last
 * accepted step size resident in the device buffers between items.  The host
 * appends items to a ring in coherent pinned host memory; waves with free lanes
 * (lane groups) claim them as they arrive, so steps of different requests
 * overlap on the device instead of waiting for one launch per batch.  A
 * finished item's end state is stored in the device buffers and copied to
 * host-mapped outputs, then its sequence number is published in done_seq.
 * Idle waves exit after idle_us without work (or at cbc_serve_destroy); the
 * next submit or poll relaunches the kernel.  One host thread per handle. */
typedef struct cbc_serve_spec {
    int32_t capacity;       /* ring slots: items queued or running at once (power of two, >= 64) */
    int32_t waves;          /* persistent grid in waves (0 = lanes for min(capacity, ns) systems, at most one wave per SIMD slot) */
    int32_t refill;         /* free lanes (groups) of a wave that trigger a claim, 1..64 */
    int32_t poll_interval;  /* attempts between ring polls of a wave that has live systems
                               (doubling after each poll that claims nothing, up to 8x);
                               1 .. 2^20 */
    int64_t max_attempts;   /* per item; a system reaching it finishes with CBC_ACTIVE */
    int32_t idle_us;        /* an idle wave exits after this many microseconds without work */
    int32_t copy_state;     /* nonzero: finished items also copy their state to out_state */
} cbc_serve_spec;

typedef struct cbc_serve_view {
    const uint32_t *done_seq;  /* (NS): sequence number of the last finished item per system */
    const double *out_state;   /* (4, NS, TILE): end state of that item, or NULL */
    const double *out_misc;    /* (NS, 4): tau reached, step size, status flag, attempts */
    double *stage_state;       /* (4, NS, TILE) staging rows of new episodes (host writes) */
    double *stage_up;          /* (13, NS, TILE) */
    double *stage_cm;          /* (3, UPS, NS, TILE) */
    int32_t capacity;
    int32_t waves;
} cbc_serve_view;

/* item option: the system starts a new episode (reset_envs,
 * envs/bubble_1D_pos_control.py:583-614, with new radii): its state, unit
 * parameter and coupling rows are read from the staging mirrors, which the
 * host filled before submitting, and stored into the device buffers
 * (actual_state, unit_params, coupling_matrices are then written). */
#define CBC_SERVE_RESET 1
/* item option: the drive amplitudes come from control_params (as the environment's
 * set_action wrote them, phases included) instead of the item's pa */
#define CBC_SERVE_USE_CONTROL 2

typedef struct cbc_serve cbc_serve;

/* Creates a serving handle over the plan's device buffers (no counterpart in
 * the reference, whose environment blocks in model.solve every step,
 * envs/bubble_1D_pos_control.py:751-753).  Allocates the ring and the outputs
 * (pinned host memory) and a 16-byte device counter; launches nothing.
 * stream == NULL: the handle creates its own non-blocking stream.  Work on
 * another stream that the runtime maps to the same hardware queue waits
 * until the serving kernel idles out, so the serving loop keeps other device
 * work (episode resets included) out of the way while serving. */
int32_t cbc_serve_create(cbc_plan *plan, const cbc_buffers *buf, const cbc_serve_spec *spec, void *stream,
                         cbc_serve **out);

/* Host pointers of the outputs (valid until cbc_serve_destroy). */
int32_t cbc_serve_get_view(const cbc_serve *serve, cbc_serve_view *out);

/* Publishes n items (one env step each): system sys[k] integrates tau in
 * [0, t_end] from its resident state with drive amplitudes pa[k], pa[n + k]
 * (NK x n, Pa), as set_action + advance_time + step_physics do; h0[k] > 0
 * replaces the resident step size (the reference's advance_time resets it to
 * 1e-6), h0[k] <= 0 or h0 == NULL keeps it; opts[k] (may be NULL) holds
 * CBC_SERVE_* bits.  Each system must have no unfinished item and appear
 * once.  seq_out[k] (may be NULL) receives the item's sequence number; the
 * item is finished when done_seq[sys[k]] == seq_out[k].  Returns
 * CBC_ERR_BUSY, publishing nothing, when fewer than n ring slots are free
 * (admission).  Launches the serving kernel if none is running. */
int32_t cbc_serve_submit(cbc_serve *serve, const int32_t *sys, const double *pa, int32_t n, double t_end,
                         const double *h0, const int32_t *opts, uint32_t *seq_out);

/* Frees the ring slots of finished items and relaunches the kernel if it
 * ended with items unclaimed; *pending (may be NULL) receives the number of
 * published items not yet known finished.  Call it while waiting. */
int32_t cbc_serve_poll(cbc_serve *serve, int32_t *pending);

/* Kernel launches so far, items published, ring slots freed. */
int32_t cbc_serve_stats(const cbc_serve *serve, int64_t *launches, uint32_t *published, uint32_t *retired);

/* Stops the serving kernel once its live items finish, waits for it and frees
 * the handle; the device buffers keep every finished item's state. */
int32_t cbc_serve_destroy(cbc_serve *serve);

/* ---- vector-environment step ----------------------------------------------
 * The device side of PosNBC1D.step / reset_envs
 * (coupledbubble_control/envs/bubble_1D_pos_control.py:503-538, 583-614,
 * 731-853) as three kernels over the solver buffers plus the environment's own
 * buffers:
 *   cbc_env_set_action  set_action (:732-749)
 *   cbc_env_post        observe, check_termination, get_rewards, advance_time
 *                       and the done list of final_termination (:756-842)
 *   cbc_env_reset       final_observation / episode_return / episode_length of
 *                       the done envs, then reset_envs (:583-614) with scenes
 *                       from SceneGenerator (:112-249) computed on the device
 *                       from the generator's uniform draws.
 * Observations, rewards and targets are float32 as in the reference
 * (envs/common/buffers.py, spaces.py); state rows are FP64.
 *
 * Observation parts follow SpaceSchema's components in order
 * (envs/common/specs.py:35-44): a STATE part is a row of actual_state kept as
 * a ring of `stack` observations (RingBuffer, buffers.py:36-83; stack 1 is the
 * PersistentBuffer), a TARGET part holds the episode's target positions, a
 * FIXED part is never written (a component that observe() skips; it may be
 * stacked and then stays an all-zero ring).  history
 * holds the parts at part_hist offsets (floats): STATE (stack, NE, N), others
 * (NE, N).  Ring slots advance once per cbc_env_post; `observed` is the number
 * of cbc_env_post calls made before the call, so slot = observed % stack.
 * The observation row is the concatenation of the parts, each bubble-major
 * with its history oldest first (RingBuffer.get_stacked), scaled per column
 * as Box.scale: (x - bias) * rscale. */
#define CBC_ENV_MAX_ACTIONS 16
#define CBC_ENV_MAX_PARTS 8
#define CBC_ENV_MAX_BUBBLES 16
#define CBC_ENV_PART_STATE 0
#define CBC_ENV_PART_TARGET 1
#define CBC_ENV_PART_FIXED 2

typedef struct cbc_env_spec {
    int32_t num_envs;            /* NE = NS of the solver buffers */
    int32_t num_bubbles;         /* N = UPS */
    int32_t tile;                /* state row stride (TILE) */
    int32_t num_control_rows;    /* 4 * NK */
    int32_t num_actions;         /* action columns */
    int32_t action_row[CBC_ENV_MAX_ACTIONS];   /* control_params row of each column (ActionMapingManager._dev_cols) */
    double action_scale[CBC_ENV_MAX_ACTIONS];  /* its scale (1e5 for PA, pi for PS) */
    int32_t num_pa;              /* action columns of the PA component (actuation penalty) */
    int32_t pa_col[CBC_ENV_MAX_ACTIONS];
    int32_t num_parts;
    int32_t part_kind[CBC_ENV_MAX_PARTS];
    int32_t part_row[CBC_ENV_MAX_PARTS];       /* state row of a STATE part */
    int32_t part_stack[CBC_ENV_MAX_PARTS];     /* ring depth (1 for TARGET) */
    int32_t part_hist[CBC_ENV_MAX_PARTS];      /* float offset of the part in history */
    int32_t obs_dim;             /* sum over parts of N * stack */
    int64_t history_floats;      /* floats in history (bounds of the part offsets) */
    int32_t x_part;              /* the STATE part of the positions (X, row 1) */
    int32_t target_part;         /* the TARGET part (XT) */
    int32_t apply_termination;   /* success = every |XT - X| <= position_tolerance */
    int64_t episode_length;
    float position_tolerance;
    float r_dnorm;               /* 1 / dnorm */
    float reward_exp;
    float w_distance, w_intensity;                          /* reward weights */
    float positive_terminal_reward, negative_terminal_reward;
    double next_time_step;       /* time_step after each step (1e-6); <= 0 keeps the last accepted step */
    int32_t env_begin;           /* cbc_env_set_action / cbc_env_post act on envs [env_begin, env_begin + */
    int32_t env_count;           /* env_count) only (0: every env); the buffers keep all NE rows */
} cbc_env_spec;

typedef struct cbc_env_buffers {
    /* solver buffers (cbc_buffers layouts) */
    double *actual_state;
    double *actual_time;
    double *time_step;
    double *actual_event;
    int32_t *status_flags;
    double *control_params;
    /* environment buffers, (NE,) unless stated */
    const float *actions;        /* (NE, num_actions) */
    float *history;              /* observation parts, see above */
    float *observations;         /* (NE, obs_dim) */
    const float *obs_bias;       /* (obs_dim,) */
    const float *obs_rscale;     /* (obs_dim,) */
    float *rewards;
    float *total_rewards;
    int64_t *algo_steps;
    uint8_t *failure;            /* bool: status != SUCCESS */
    uint8_t *success;
    uint8_t *terminated;         /* failure | success */
    uint8_t *time_outs;          /* algo_steps >= episode_length */
    int64_t *done_ids;           /* (NE,) ascending ids of terminated or timed-out envs */
    uint32_t *workspace;         /* cbc_env_workspace_words(NE) words, zeroed once before the first call */
} cbc_env_buffers;

/* Scene of one side (bubbles or targets), SceneGenerator.generate_new_scenes:
 * RANDOM: _random_uniform_position_1D (:183-242), draws (n, N-1) uniforms plus
 *   (n, 1) for random alignment; EQUIDISTANT: _equidistant_position_1D
 *   (:139-181), base positions `base` plus an offset (random alignment: one
 *   uniform per scene, offset = lo + u (max_start - lo)); FIXED: `base`. */
#define CBC_SCENE_RANDOM 0
#define CBC_SCENE_EQUIDISTANT 1
#define CBC_SCENE_FIXED 2
typedef struct cbc_scene_side {
    int32_t method;
    int32_t random_align;        /* alignment == "random" */
    double base[CBC_ENV_MAX_BUBBLES];
    double offset;               /* EQUIDISTANT, centred or no distance */
    double max_start;            /* EQUIDISTANT, random alignment */
    const double *u_gaps;        /* RANDOM: (n, N-1) draws */
    const double *u_align;       /* random alignment: (n,) draws */
} cbc_scene_side;

typedef struct cbc_env_scene {
    double lo;                   /* domain_min + margin */
    double available;            /* hi - lo */
    double min_distance;
    double room_per_gap;         /* max_distance - min_distance */
    double extra_gap_limit;      /* available - (N-1) min_distance */
    cbc_scene_side bubbles;
    cbc_scene_side targets;
} cbc_env_scene;

/* Counts of the last cbc_env_post (host values). */
typedef struct cbc_env_counts {
    int32_t done;
    int32_t failures;
    int32_t successes;
    int32_t time_outs;
} cbc_env_counts;

/* Words of cbc_env_buffers.workspace for NE environments. */
int64_t cbc_env_workspace_words(int32_t num_envs);

/* set_action (envs/bubble_1D_pos_control.py:732-749): control_params[action_row[c]][e]
 * = (double)actions[e][c] * action_scale[c].  Asynchronous. */
int32_t cbc_env_set_action(const cbc_env_spec *spec, const cbc_env_buffers *buf, void *stream);

/* observe (:756-769), check_termination (:772-789), get_rewards (:792-813),
 * advance_time (:816-827) and the done list of final_termination (:830-842),
 * one thread per env, in one launch; its last workgroup to finish turns the
 * per-workgroup done counts into offsets, and a second launch moves each
 * workgroup's ids into the ordered done_ids.  When `counts` is not NULL the call waits
 * for the stream and returns them (the reference's host syncs of the same
 * step); otherwise it is asynchronous. */
int32_t cbc_env_post(const cbc_env_spec *spec, const cbc_env_buffers *buf, int64_t observed, void *stream,
                     cbc_env_counts *counts);

/* No counterpart in the reference: cbc_env_post without the workspace argument, for
 * callers that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Same kernels, arguments and results
 * as cbc_env_post; buf->workspace is ignored (may be NULL).  The workspace is a
 * file-scope device array of the library sized for CBC_ENV_STATIC_MAX_ENVS envs
 * (4.1 MiB), which every GPU that runs this call holds until the process ends; the call
 * itself allocates nothing.  All calls on one GPU share that array, so calls that
 * overlap on one GPU (other streams, other threads) race and must not be issued.
 * num_envs above CBC_ENV_STATIC_MAX_ENVS returns CBC_ERR_INVALID. */
#define CBC_ENV_STATIC_MAX_ENVS (1 << 20)
int32_t cbc_env_post_static(const cbc_env_spec *spec, const cbc_env_buffers *buf, int64_t observed, void *stream,
                            cbc_env_counts *counts);

/* final_termination's per-done-env outputs and reset_envs (:583-614) for the n
 * envs ids[0..n): final_observation (n, obs_dim), episode_return (n,),
 * episode_length (n,) (each may be NULL) are taken before the reset; then new
 * positions and targets from `scene`, R = 1, X = positions, U = V = 0 in
 * actual_state, the X ring filled with the positions and the targets stored,
 * algo_steps = 0, total_rewards = 0, and the observation rows rebuilt.
 * `observed` counts the cbc_env_post calls made so far.  Asynchronous. */
int32_t cbc_env_reset(const cbc_env_spec *spec, const cbc_env_buffers *buf, const cbc_env_scene *scene,
                      const int64_t *ids, int32_t n, int64_t observed, float *final_observation,
                      float *episode_return, int64_t *episode_length, void *stream);

/* SceneGenerator.generate_new_scenes (:112-137) alone: positions (n, N) and
 * targets (n, N) of n scenes (either may be NULL), by the arithmetic of
 * cbc_env_reset.  Asynchronous. */
int32_t cbc_env_scene_positions(const cbc_env_scene *scene, int32_t n, int32_t num_bubbles, double *positions,
                                double *targets, void *stream);

/* ---- trajectory collection -----------------------------------------------
 * TrajectorBuffer (coupledbubble_control/envs/common/trajectory_buffer.py:70-166)
 * on the device: every env's running episode is kept in an arena of float32
 * rows (the dtype TrajectoryZarrWriter stores, :196-226):
 *   obs (NE, max_steps, obs_dim), act (NE, max_steps, num_actions),
 *   rew (NE, max_steps), time (NE, max_points),
 *   states (NE, num_saved, max_points, N), steps (NE,) int32, points (NE,) int64.
 * Points of step k > 0 skip the step's first dense point and carry times
 * t + k * step_length (Trajectory.merge_list, :38-56). */
typedef struct cbc_traj_spec {
    int32_t num_envs;
    int32_t num_bubbles;
    int32_t tile;
    int32_t ndo;                 /* dense points per solve (NDO) */
    int32_t num_saved;           /* state rows recorded (saved_states) */
    int32_t saved_rows[4];
    int32_t obs_dim;
    int32_t num_actions;
    int32_t max_steps;           /* arena rows per env (episode_length) */
    int64_t max_points;          /* arena points per env */
    double step_length;          /* time_step_length */
} cbc_traj_spec;

typedef struct cbc_traj_buffers {
    /* one step's inputs (device) */
    const float *observations;   /* (NE, obs_dim) */
    const float *actions;        /* (NE, num_actions) */
    const float *rewards;        /* (NE,) */
    int32_t *dense_index;        /* the solver's dense buffers; dense_index is cleared */
    const double *dense_time;
    const double *dense_state;
    /* the arena */
    float *obs;
    float *act;
    float *rew;
    float *time;
    float *states;
    int32_t *steps;
    int64_t *points;
    int32_t *overflow;           /* set to 1 when an append did not fit (the step is not recorded) */
} cbc_traj_buffers;

/* Packed trajectories of n envs, back to back in done-list order: obs / act /
 * rew rows (sum of steps), time (total_points), states (num_saved, total_points, N). */
typedef struct cbc_traj_packed {
    float *obs;
    float *act;
    float *rew;
    float *time;
    float *states;
    int64_t total_points;
} cbc_traj_packed;

/* TrajectorBuffer.step (:95-127) for every env; one wave per 8 envs.  Asynchronous. */
int32_t cbc_traj_append(const cbc_traj_spec *spec, const cbc_traj_buffers *buf, void *stream);

/* The arenas of envs ids[0..n) into `out` (step_off / point_off: device int64
 * prefix offsets of each env's rows and points, from the steps / points
 * counters), then their counters cleared -- the data TrajectorBuffer.end_episode
 * (:129-155) merges and writes.  One workgroup per env.  Asynchronous. */
int32_t cbc_traj_gather(const cbc_traj_spec *spec, const cbc_traj_buffers *buf, const int64_t *ids, int32_t n,
                        const int64_t *step_off, const int64_t *point_off, const cbc_traj_packed *out,
                        void *stream);

/* ---- diagnostics ---------------------------------------------------------
 * Evaluates the kernels' elementary functions on n device inputs (x, and y for
 * two-argument functions) into out[2 n] (value pairs (u, v) per input):
 *   CBC_MATH_SINCOS       (sin x, cos x)
 *   CBC_MATH_EXP / LOG / RCP   (e^x / log x / 1/x, 0)
 *   CBC_MATH_POW          (x^y for x > 0, 0)
 *   CBC_MATH_SINCOS_KERN  (sin x, cos x - 1) for |x| <= pi/4
 *   CBC_MATH_SINCOS_TINY  (sin x, cos x - 1) for |x| <= 1/20
 *   CBC_MATH_ROTATE       (sin(x + y), cos(x + y)) by the addition theorem, |y| <= 1/20
 *   CBC_MATH_POW1P        ((1 + x)^y, 0) for |x| <= 1/100, |y| <= 6
 * No counterpart in the reference (which calls the CUDA/Numba libm); it lets
 * tests pin the device arithmetic against a host library.  Asynchronous. */
#define CBC_MATH_SINCOS 0
#define CBC_MATH_EXP 1
#define CBC_MATH_LOG 2
#define CBC_MATH_POW 3
#define CBC_MATH_RCP 4
#define CBC_MATH_SINCOS_KERN 5
#define CBC_MATH_SINCOS_TINY 6
#define CBC_MATH_ROTATE 7
#define CBC_MATH_POW1P 8
int32_t cbc_selftest_math(int32_t fn, const double *x, const double *y, double *out, int32_t n, void *stream);

/* The kernels' in-register coupling solve (LU with partial pivoting by conditional row
 * exchanges) on `count` augmented systems a (count x n x (n + 1), row major), n in
 * {2, 4, 6, 8} (the 2N unknowns for N = 1..4); solutions into x (count x n).  No
 * counterpart in the reference (which calls a library solve); it lets tests drive every
 * exchange path of the solve on chosen matrices.  Asynchronous. */
int32_t cbc_selftest_solve(int32_t n, const double *a, double *x, int32_t count, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* CBC_H */
