/*
 * inundation_hip.h - C ABI of the MI355X shallow-water step backend.
 *
 * libinundation_hip.so replaces the reference's GPU-resident step pipeline,
 * i.e. the Python backend class SWESolverGpuResidentBatch
 * (inundation/vulkan/gpu_resident_batch.py:140-762) and the GLSL kernels it
 * dispatches through Kompute (inundation/vulkan/shaders/resident.py,
 * flux.py, cfl.py). The reference has no C ABI of its own: its boundary is the
 * backend-class protocol that SWEWorkflow drives (workflow.py:783-826). Each
 * entry point below names the reference method or kernel it stands in for.
 *
 * Conventions
 *   - Every function returns 0 on success or a negative INU_E* code;
 *     inu_last_error() gives the message of the last failure on a context.
 *   - Plain pointers and sizes only. Host arrays are copied; the caller keeps
 *     ownership. Device memory belongs to the context and is freed by
 *     inu_destroy().
 *   - One HIP stream per context; calls on one context are not thread-safe.
 *     Calls that return data synchronise that stream.
 *   - Arrays are in the solver's cell order (Hilbert) and in the edge order
 *     sorted by the block of the left cell (reference blocks.py:28-30).
 */
#ifndef INUNDATION_HIP_H
#define INUNDATION_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define INU_ABI_VERSION 2

/* Error codes. INU_EINVAL maps to ValueError on the Python side, the others to
 * RuntimeError (reference raises: gpu_resident_batch.py:641-655, 720-724). */
#define INU_OK 0
#define INU_EINVAL -1 /* bad argument */
#define INU_EHIP -2   /* HIP runtime error */
#define INU_ESTATE -3 /* call order (no mesh yet, ...) */
#define INU_ERANGE -5 /* mesh exceeds a size limit of the kernels */
#define INU_ENODEV -6 /* no HIP device */

/* inu_set_params flags: the solver attributes the workflow sets after
 * construction (reference gpu_resident_batch.py:143-158, workflow.py:816-826). */
/* Fixed-order CSR gather (bit-reproducible). Off: each edge's words are added
 * into its cells' sums as threads arrive (the reference's atomic variant,
 * resident.py:110-180), results vary in the last bits from run to run; not
 * with open boundaries or an adjoint recording. */
#define INU_FLAG_DETERMINISTIC 1u
#define INU_FLAG_COMPENSATED 2u   /* depth as a compensated float32 pair */
#define INU_FLAG_ACTIVE_SET 4u    /* skip blocks beyond the two-ring of water */
#define INU_FLAG_TRACK_CLAMP 8u   /* accumulate the positivity-clamp mass */
#define INU_FLAG_EXACT_ORDER 16u  /* the reference's operation order: one IEEE division or
                                     square root per occurrence, no fused multiply-add;
                                     without it the kernels share reciprocals and use FMA */

/* Fields of inu_download / inu_upload. */
#define INU_FIELD_H 0
#define INU_FIELD_HU 1
#define INU_FIELD_HV 2
#define INU_FIELD_H_LO 3
#define INU_FIELD_H1 4
#define INU_FIELD_HU1 5
#define INU_FIELD_HV1 6
#define INU_FIELD_H1_LO 7
#define INU_FIELD_CLAMP 8
#define INU_FIELD_BLOCK_WET 9     /* uint32 per 256-cell block */
#define INU_FIELD_BLOCK_ACTIVE 10 /* uint32 per block */
#define INU_FIELD_BLOCK_FORCED 11 /* uint32 per block */

/* Open boundary conditions (reference tuning.py:64-68, the OPEN_* constants of
 * shaders/flux.py:194-202). */
#define INU_OPEN_WALL 0u
#define INU_OPEN_FREE 1u   /* zero-gradient outflow */
#define INU_OPEN_NORMAL 2u /* outflow at the Manning velocity; value sqrt(S) / n */
#define INU_OPEN_STAGE 3u  /* a level beyond the edge; value the depth over the cell's bed */
#define INU_OPEN_INFLOW 4u /* value the discharge per unit width */

/* The 48-byte resident control block: the integer clock and step planning
 * (reference RESIDENT_CTRL_DTYPE, gpu_resident_batch.py:65-77; the CTRL block
 * of STEP_CONTROL_GLSL, resident.py:823-833). */
typedef struct inu_ctrl {
  uint64_t now_ticks;
  uint64_t stop_ticks;
  uint64_t steps;
  uint32_t dt_ticks;
  uint32_t dt_max_ticks;
  uint32_t dry_dt_ticks;
  float last_raw;
  float cfl_safety;
  uint32_t pad;
} inu_ctrl;

/* One mesh, as the arguments of inu_set_mesh (for inu_set_batch). */
typedef struct inu_mesh {
  uint32_t n_cells;
  uint32_t n_edges;
  const int32_t* edge_cell_l;
  const int32_t* edge_cell_r;
  const float* nx;
  const float* ny;
  const float* len;
  const float* edge_dz;
  const float* cell_area;
  const uint32_t* cell_edge_ptr;
  const uint32_t* cell_edge_idx;
} inu_mesh;

typedef struct inu_ctx inu_ctx;

/* ABI version of the loaded library (INU_ABI_VERSION). No reference counterpart. */
int inu_version(void);

/* Context on HIP device `device`; replaces kp.Manager() in SWESolver.__init__
 * (reference solver.py:186). */
int inu_create(int device, inu_ctx** out);
void inu_destroy(inu_ctx* ctx);
const char* inu_last_error(const inu_ctx* ctx);

/* Mesh upload; replaces _build_tensors (reference solver.py:219-296 and
 * gpu_resident_batch.py:160-282). Edges must be sorted by the block of their
 * left cell; edge_dz holds 2E interleaved (dzL, dzR) words (solver.py:233);
 * cell_edge_idx holds (edge << 1) | is_right in that sorted edge numbering
 * (blocks.py:33-44). edge_cell_r is a cell, -1 (a wall) or -2 - k (open edge
 * k, k in 0..n_open-1, each once, on an edge without a bed step: reference
 * solver.py:239-243). Open edges need a context of one scenario; they start
 * closed (INU_OPEN_WALL), in one stretch. Resets the state, parameters and
 * sources. */
int inu_set_mesh(inu_ctx* ctx, uint32_t n_cells, uint32_t n_edges, const int32_t* edge_cell_l,
                 const int32_t* edge_cell_r, const float* edge_nx, const float* edge_ny,
                 const float* edge_len, const float* edge_dz, const float* cell_area,
                 const uint32_t* cell_edge_ptr, const uint32_t* cell_edge_idx);

/* Batched scenarios (an extension with no reference counterpart:
 * the reference runs one scenario per solver). n_scenarios scenarios, scenario
 * s on mesh scenario_mesh[s] of the n_meshes meshes, each with its own state,
 * sources, control block and CFL dt, all advanced by the same launches.
 * Every cell-indexed array of the other calls then holds the scenarios' cells
 * back to back (scenario 0's cells first; inu_scenario_info gives the
 * ranges); block-indexed arrays hold the scenarios' blocks back to back;
 * inu_ctrl_write/read and inu_dt_words take one entry per scenario. A
 * scenario's results are bit-identical to running it alone. inu_set_mesh is
 * the one-scenario case. */
int inu_set_batch(inu_ctx* ctx, uint32_t n_meshes, const inu_mesh* meshes, uint32_t n_scenarios,
                  const uint32_t* scenario_mesh);

/* Number of scenarios (>= 1 once a mesh is set), or a negative INU_E* code. */
int inu_scenario_count(inu_ctx* ctx);

/* out4 = {first cell, cells, first block, blocks} of scenario s. */
int inu_scenario_info(inu_ctx* ctx, uint32_t s, uint32_t* out4);

/* out6 = {N, E, slots, blocks, max edges per block, stage grid size}, summed
 * over the scenarios. No reference counterpart (diagnostics). */
int inu_mesh_info(inu_ctx* ctx, uint32_t* out6);

/* g, dry_tol, CFL number and INU_FLAG_* (reference constructor arguments,
 * solver.py:149-172, and the solver attributes above). */
int inu_set_params(inu_ctx* ctx, float g, float dry_tol, float cfl, uint32_t flags);

/* Depth, momenta and the compensated low word (NULL = zeros); replaces
 * reset() (solver.py:390-421, gpu_resident_batch.py:436-445). */
int inu_set_state(inu_ctx* ctx, const float* h, const float* hu, const float* hv, const float* h_lo);

/* Per-cell Manning roughness (binding 16, solver.py:252). */
int inu_set_manning(inu_ctx* ctx, const float* n);

/* Constant per-run sources as (cell, dh/dt) pairs; zero rates are dropped;
 * replaces _set_sources (gpu_resident_batch.py:529-560). */
int inu_set_sources(inu_ctx* ctx, uint32_t n, const uint32_t* cell, const float* rate);

/* New rates for cells already in the source set (a zero rate turns a source
 * off until a later call turns it on again); the cell set, its block lists
 * and the recorded graphs stay. Replaces the per-phase source re-upload of
 * the reference's run loop (gpu_resident_batch.py:529-560, 666-670) when only
 * the rates change between phases. */
int inu_set_source_rates(inu_ctx* ctx, uint32_t n, const uint32_t* cell, const float* rate);

/* Control block up/down, one per scenario; replaces
 * _sync_ctrl_to_device/_read_ctrl (gpu_resident_batch.py:483-487). */
int inu_ctrl_write(inu_ctx* ctx, const inu_ctrl* ctrl);
int inu_ctrl_read(inu_ctx* ctx, inu_ctrl* ctrl);

/* The dt words, four per scenario: [0] this step's dt bits, [1] next-step CFL
 * bound bits, [2], [3] zero (binding 17, resident.py:6-11). */
int inu_dt_words(inu_ctx* ctx, uint32_t* out4);

/* Start of a run: block marks, forced dilation, CFL bound of the current
 * state, step_control(START); replaces _record_prologue
 * (gpu_resident_batch.py:567-591). */
int inu_prologue(inu_ctx* ctx);

/* Work the active set let through, summed over the real steps since the last
 * reset: out4 = {active cells, active owned edges, active CSR slots, steps}
 * (a block's cells and the edges whose left cell lies in it, as the
 * algorithmic-bytes rule counts them). reset != 0 zeroes the sums. No
 * reference counterpart (measurement). */
int inu_activity(inu_ctx* ctx, uint64_t* out4, int reset);

/* New output window: step_control(RETARGET); replaces _record_retarget
 * (gpu_resident_batch.py:593-599). */
int inu_retarget(inu_ctx* ctx);

/* Run n recorded steps and wait; use_graph replays cached hipGraphs of at
 * most 100 steps each (recorded on first use), back to back. Replaces
 * steps_seq.eval() of _record_steps (gpu_resident_batch.py:601-607, 717). */
int inu_steps(inu_ctx* ctx, uint32_t n, int use_graph);

/* n steps launched one by one with HIP events around the two stage kernels:
 * ms_out3 = {predictor kernel ms, corrector kernel ms, whole span ms}. No
 * reference counterpart (measurement). */
int inu_steps_timed(inu_ctx* ctx, uint32_t n, float* ms_out3);

/* Field readback/upload (INU_FIELD_*); replaces download_h/hu/hv,
 * download_h_residual, download_block_flags, download_clamp_mass
 * (solver.py:473-497, gpu_resident_batch.py:469-481). */
int inu_download(inu_ctx* ctx, int which, void* out);
int inu_upload(inu_ctx* ctx, int which, const void* in);

/* Arithmetic probe on n (a, b) pairs: out7 per pair = {two-sum error of a + b,
 * a / b, sqrt|a|, a^(4/3), film drag of depth a over dt b, a^(-1/3), film drag
 * in the fast arithmetic (0 outside its domain a >= dry_tol, b >= 0)}. Generalises the EFT probe (reference
 * EFT_PROBE_GLSL, resident.py:889-907, and _check_error_free_arithmetic,
 * solver.py:449-467). */
int inu_probe(inu_ctx* ctx, uint32_t n, const float* pairs, float* out7, float g, float dry_tol);

/* Test probe of the fast arithmetic's reciprocal: every float bit pattern u
 * in [lo, hi) and its negative, compared bitwise with the IEEE division 1 / x
 * (NaNs equal); *mismatches gets the count that differ and *first the smallest
 * such pattern (0xffffffff if none). hi <= 0x80000000. No reference
 * counterpart: it pins the claim that the reciprocal is correctly rounded on
 * [2^-126, 2^126] (lo = 0x00800000, hi = 0x7e800001). */
int inu_probe_rcp(inu_ctx* ctx, uint32_t lo, uint32_t hi, uint64_t* mismatches, uint32_t* first);

/* Test probe of the open boundary formulas (the reference tests them the same
 * way: tests/unit/test_gpu_open_ghost.py:46-68 runs open_ghost and
 * boundary_mass_flux, flux.py:224-281, on the device). Case k is (h, hu, hv,
 * nx, ny, fh) of the cell inside an open edge of condition kinds[k] and value
 * values[k]; out4 gets the ghost (h, hu, hv) and the edge's mass flux for an
 * HLLC mass flux fh. */
int inu_probe_open(inu_ctx* ctx, uint32_t n, const float* cases6, const uint32_t* kinds, const float* values,
                   float* out4, float g, float dry_tol);

/* Test probe of the integer clock: step_control (mode 0 START, 1 STEP, 2
 * RETARGET) once on a hand-set control block and dt words, in place (dt4[0]
 * the prologue bound bits, dt4[1] the next-step bound bits; on return this
 * step's dt bits, the sentinel 1e30, 0, 0). The reference tests its kernel
 * the same way: tests/unit/test_gpu_resident_clock.py:46-104 runs
 * step_control (resident.py:813-885) on a hand-built control block. Uses
 * device memory of its own; the context's mesh and clock are not touched. */
int inu_probe_clock(inu_ctx* ctx, int mode, inu_ctrl* ctrl, uint32_t* dt4);

/* A plain device copy of n_bytes (a positive multiple of 16), launched
 * `reps` times after one untimed launch; *ms_out = mean duration of one copy.
 * Measurement support with no counterpart in the reference: a kernel of known
 * traffic (n_bytes read, n_bytes written) that HBM byte counters are scaled
 * against, and the attainable copy bandwidth. Allocates and frees its own
 * buffers. */
int inu_stream_copy(inu_ctx* ctx, uint64_t n_bytes, uint32_t reps, float* ms_out);

/* Attainable HBM rate, for the roofline's context (measurement, no reference
 * counterpart): kind 0 copies n_bytes (read and written), 1 reads them (plus
 * one word per thread written), 2 writes them, 3 is hipMemcpyAsync device to
 * device. Kinds 0-2 move `unroll` (1, 2, 4, 8) 16-byte words per thread per
 * trip, loads before stores, with nontemporal hints, on wg_per_cu 256-thread
 * workgroups per CU. `reps` launches after one untimed; *ms_out = mean
 * milliseconds of one. Allocates and frees its own buffers. */
int inu_stream_bw(inu_ctx* ctx, uint64_t n_bytes, uint32_t reps, int kind, uint32_t unroll, uint32_t wg_per_cu,
                  float* ms_out);

/* Autotuner of the stage kernels' tiles (batching extension; no
 * reference counterpart). Times `reps` launches of the predictor at each tile
 * level on the current state (after inu_prologue; the state is left as it
 * was) and keeps the fastest for both stages: out4 = {chosen level,
 * microseconds per launch at levels 0, 1, 2}. A level is the number of edge
 * records and CSR slots each thread loads before the first barrier: (2, 4),
 * (3, 6), (4, 8). Results do not depend on the level. */
int inu_autotune(inu_ctx* ctx, uint32_t reps, uint32_t* out4);

/* Sets the tile level (0, 1 or 2) directly; inu_set_mesh/inu_set_batch reset it to 0. */
int inu_set_tile(inu_ctx* ctx, uint32_t tile);

/* inu_set_source_rates without device allocation. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * hipGraph stream capture, tight serving loops, harnesses that forbid
 * hipMalloc). It stages the cell ids and rates through a fixed 512 KiB device
 * array that each GPU holds for the life of the process, so it takes at most
 * 65536 rates per call (INU_EINVAL beyond). Calls that overlap on one GPU,
 * from any context or thread, race on that array. Refused (INU_ESTATE) during
 * an adjoint recording, whose rate log allocates, and (INU_EINVAL) for a
 * negative rate while every source is non-negative (the step graphs would be
 * recorded again); inu_set_source_rates takes both. Same results as
 * inu_set_source_rates otherwise. */
int inu_set_source_rates_noalloc(inu_ctx* ctx, uint32_t n, const uint32_t* cell, const float* rate);

/* Open boundary stretches: open edge k belongs to stretch s when
 * stretch_ptr[s] <= k < stretch_ptr[s + 1] (n_stretches + 1 words, from 0 to
 * the open edge count; reference SWESolver's open_edges, solver.py:140-165,
 * one edge array per stretch, numbered in stretch order, and
 * gpu_resident_batch.py:250-263). Clears the volume tally. */
int inu_set_open_stretches(inu_ctx* ctx, uint32_t n_stretches, const uint32_t* stretch_ptr);

/* Condition (INU_OPEN_*) and value of every open edge, before a run or a
 * phase; replaces set_open_conditions (reference solver.py:427-438,
 * gpu_resident_batch.py:447-458). A block whose open edge can bring water in
 * (inflow or stage with value > 0) counts as wet while these hold; call
 * inu_prologue (or inu_retarget) before stepping so the lists see it. */
int inu_set_open_conditions(inu_ctx* ctx, uint32_t n_open, const uint32_t* kinds, const float* values);

/* Volume [m^3] out of and into the mesh through each stretch since the last
 * inu_set_state or inu_set_open_stretches: out2 = (out, in) per stretch, for
 * n_stretches stretches (the count last set; reference download_open_volumes,
 * gpu_resident_batch.py:460-467; the tally of resident.py:613-658). */
int inu_open_volumes(inu_ctx* ctx, uint32_t n_stretches, double* out2);

/* Bed elevation per cell: recomputes every edge's bed step with the face on
 * the higher bed (reference mesh/geometry.py:241-254, edge_bed_steps, FP64
 * then FP32; the left cell on a tie, walls none). For differentiating and
 * perturbing the bed without a new mesh. */
int inu_set_bed(inu_ctx* ctx, const float* z);

/* ---- Adjoint (an extension: the reference has no gradient path). Reverse
 * mode of recorded steps, for d loss / d (Manning n, bed z) of a loss on
 * sampled gauge depths. One scenario; source rates must be non-negative;
 * gauge cells distinct. The forward is recorded at step 0 after inu_prologue
 * with inu_adj_steps: CFL-driven steps as inu_steps, plus each step's dt and
 * active-set lists (listed blocks, idle source blocks), gauge depths h + h_lo
 * after every `every` steps and the state every `seg` steps;
 * inu_set_source_rates may change rates between recorded steps (they are
 * replayed). inu_adj_backward replays each `seg`-step segment from its
 * checkpoint with the recorded dt, rates and lists (device memory: seg x 56
 * bytes per cell for the segment's tape) and runs it in reverse over the
 * listed blocks, following the forward's branches; blocks the forward skipped
 * keep their state, so their gradient passes through. Replays and reverse
 * segments run as hipGraphs cached with the recording. After a replay or
 * backward the context's state is that of the replay and must be set again
 * before further steps. */
int inu_adj_begin(inu_ctx* ctx, uint32_t max_steps, uint32_t seg, uint32_t n_gauges, const uint32_t* gauges,
                  uint32_t every);
int inu_adj_steps(inu_ctx* ctx, uint32_t n);
/* out = gauge depths, (max_steps / every) rows of n_gauges; row r after step (r + 1) every. */
int inu_adj_series(inu_ctx* ctx, float* out);
/* seeds = d loss / d series (same shape as inu_adj_series). Out: grad_n and
 * grad_z per cell; grad_state (NULL, or 3 N words) = d loss / d (h, hu, hv)
 * of the recording's initial state. */
int inu_adj_backward(inu_ctx* ctx, const float* seeds, float* grad_n, float* grad_z, float* grad_state);
/* The recorded steps again from the recording's start, with the recorded dt
 * and source rates and the current Manning n and bed (for finite
 * differences); out as inu_adj_series. Leaves the replay's state. */
int inu_adj_replay(inu_ctx* ctx, float* series_out);
/* The recorded dt of each step (float bits; max_steps words, 0 past the last). */
int inu_adj_dt_log(inu_ctx* ctx, uint32_t* out);
/* Counters of the last inu_adj_backward: listed block-steps reversed, idle
 * source block-steps passed through, and edges whose gradient reached a cell
 * outside the listed blocks (film on the active set's boundary; those few
 * additions are atomic, so their order, and only theirs, may vary). */
int inu_adj_stats(inu_ctx* ctx, unsigned long long* out3);
/* Face depth below which a side's velocity and celerity are held in the
 * reverse pass (its depth still enters the fluxes): u = q / h and sqrt(g h)
 * have slopes of order 1 / h in a thin layer, which the reverse pass would
 * compound at every wetting front. 0 holds none; negative restores the
 * default, 100 x dry_tol. Also used by inu_probe_vjp. */
int inu_adj_set_thin(inu_ctx* ctx, float depth);
/* Checkpointing of the next recording (call before inu_adj_begin). slots = 0
 * (the default): a checkpoint every `seg` steps, written by the recording,
 * and each segment replayed once by the backward pass. slots in [1, 64]:
 * binomial checkpointing (Griewank's revolve) over the recording's segments
 * with `slots` checkpoints of 16 B per cell: the recording keeps only its
 * start, and the backward pass follows the optimal offline schedule,
 * advancing segments without tape to store checkpoints and replaying each
 * segment with tape once, just before reversing it. Device memory for the
 * checkpoints: slots x 16 B per cell instead of (steps / seg + 1) x 16 B;
 * the gradient is bitwise the same. At most 2048 segments. */
int inu_adj_set_slots(inu_ctx* ctx, uint32_t slots);
/* Steps the last inu_adj_backward replayed: out2 = (without tape, with tape). */
int inu_adj_replayed(inu_ctx* ctx, unsigned long long* out2);
/* Host only: the fewest segments the binomial schedule advances without tape
 * to reverse n_segments segments with `slots` checkpoints (slot 0 holding the
 * start); the backward pass also replays each segment once with tape. */
int inu_adj_schedule_cost(uint32_t n_segments, uint32_t slots, unsigned long long* out);
int inu_adj_end(inu_ctx* ctx);

/* Edge-word VJP probe: per row in17 = {hL, huL, hvL, hR, huR, hvR, dzL, dzR,
 * nx, ny, len, wall, dloss/dword 0..4}; out8 = d loss / d (hL, huL, hvL, hR,
 * huR, hvR, dzL, dzR). Test support with no reference counterpart. */
int inu_probe_vjp(inu_ctx* ctx, uint32_t n, const float* in17, float* out8, float g, float dry_tol);

#ifdef __cplusplus
}
#endif

#endif /* INUNDATION_HIP_H */
