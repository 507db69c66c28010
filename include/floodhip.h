/*
 * floodhip -- C ABI of the MI355X per-timestep solver for the unstructured
 * urban flood model (surface HLL shallow water + coupled 1-D drainage).
 *
 * One context = one scenario, or a batch of independent scenarios ("members",
 * each with its own mesh, network, forcing, clock, dt and ledger), resident on
 * one GPU and bound to one HIP stream.  The library owns every device allocation; all
 * pointers passed in are HOST pointers borrowed for the duration of the call
 * (copied to the device), all pointers passed out are host buffers the caller
 * owns.  Contexts are not thread-safe.  Arrays are float32 / int32 with the
 * meaning of the reference's NamedTuple fields; NULL marks an optional array
 * as absent where stated.
 *
 * Reference interface replaced (paths relative to
 * HF_CUPY_JAX_COMPLETED_PROJECT/reproduction/source/src/):
 *   hybrid_flood/operational/parity_runner.py:198-234  _advance_to_time_while
 *   hybrid_flood/operational/parity_runner.py:118-158  _perform_numerical_step
 *   hybrid_flood/jax_solver/v1_parity.py:430-469       advance_parity_step
 *   hybrid_flood/jax_solver/v1_parity.py:153-158       stable_timestep
 *   hybrid_flood/operational/parity_inputs.py:130-220  load_parity_inputs (device upload)
 *   gurugram_flood/solver/engine.py:219-344            run_simulation hot loop
 */
#ifndef FLOODHIP_H
#define FLOODHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes (mapped by the Python layer to the reference's exception types,
 * hybrid_flood/operational/errors.py:4-25 and triangular_backend.py:486-489) */
#define FH_OK 0
#define FH_ERR_VALUE 1          /* ValueError: bad shape / argument            */
#define FH_ERR_SIMULATION 2     /* SimulationError: non-finite state            */
#define FH_ERR_BUDGET 3         /* SimulationError: max_steps before target     */
#define FH_ERR_UNAVAILABLE 4    /* BackendUnavailableError: no usable GPU       */
#define FH_ERR_RUNTIME 5        /* HIP runtime failure                          */

#define FH_LEDGER_TERMS 13      /* parity_runner.py:38-51 MassLedger field order */

typedef struct fh_ctx fh_ctx;

/* Solver constants.  Defaults (fh_default_params) are the reference's:
 * v1_parity.py:14-15 (g, min depth), :245 (max speed 5), :275 (orifice
 * 0.62*0.04), :328 (slot wave speed 30); parity_runner.py:124-135 (CFL 0.85,
 * re-evaluate every 32 steps, x0.8 reuse, dt cap 1 s, floor 1e-3 s). */
typedef struct {
    double gravity;
    double min_depth;
    double cfl;
    double reuse_factor;
    double max_dt;
    double dt_floor;
    double max_speed;
    double inlet_orifice_q;     /* 0.62 * 0.04 */
    double slot_wave_speed;     /* 30 m/s      */
    double fixed_dt;            /* > 0: every step uses exactly this dt (no CFL) */
    int32_t cfl_interval;       /* 32 */
    int32_t compensated;        /* 1: Kahan-compensated cell update (carried per cell) */
    int32_t precision;          /* 32 (float state) or 64 (double state: FP64 oracle runs) */
    int32_t use_graph;          /* 1: replay captured hipGraphs of steps_per_launch steps */
    int32_t steps_per_launch;   /* steps per graph / per host check (default 32) */
    int32_t block_size;         /* threads per block for the edge and cell kernels (64..256; default 64, one wavefront) */
    int32_t xcd_remap;          /* 1: map blocks so each XCD works a contiguous slice */
    int32_t device;             /* HIP device ordinal */
    int32_t edges_per_thread;   /* 1, 2 or 4 edges per thread in the flux kernel */
    int32_t pack_order;         /* batch member order: 0 as given, 1 largest first, 2 large/small alternating */
    int32_t cell_order;         /* device layout of cells and edges: 0 as given, 1 breadth-first (Cuthill-McKee)
                                   locality order; results are returned in the caller's order either way */
    int32_t source_classes;     /* 1: store the per-cell source tuples (Manning n^2, rain multiplier, Horton
                                   f0, fc, k) as a 16-bit class per cell plus a class table when there are
                                   at most 65,535 distinct tuples; 0: one array each.  Same values. */
    int32_t surface_mode;       /* 0: full shallow-water step (v1_parity.py:161-214); 1: local-inertial step
                                   (triangular_backend.py:1181-1291) with the gravity-wave CFL timestep
                                   (engine.py:220-225).  Sources, clamp, drainage and ledger are shared. */
    int32_t share_mesh;         /* 1: a batch whose members all have the same mesh stores its edge tables and
                                   cell->edge table once and runs the members' copies of each tile back to
                                   back (same results); 0: one copy per member */
} fh_params;

typedef struct {
    double time_s;              /* FP64 simulated time                           */
    double cached_dt_s;         /* last evaluated CFL dt                         */
    double last_dt_s;           /* dt of the last executed step                  */
    int64_t global_step;        /* steps executed since the clock was set        */
    int64_t interval_steps;     /* steps executed by the last advance call        */
    int64_t cfl_evaluations;
} fh_clock;

void fh_default_params(fh_params* p);
const char* fh_version(void);
int fh_device_count(void);

/* Context lifetime.  fh_create returns NULL on failure (message via fh_global_error). */
fh_ctx* fh_create(const fh_params* params);
/* Batched ensemble: `members` independent scenarios advanced together, every
 * stage of a step one launch over all members (no counterpart in the
 * reference, which has no batched mode; fh_create is fh_create_batch(p, 1)).
 * Upload, state, clock and ledger calls apply to the member chosen with
 * fh_select_member (default 0); fh_advance_to and fh_step advance every member,
 * each with its own CFL dt. */
fh_ctx* fh_create_batch(const fh_params* params, int32_t members);
int fh_select_member(fh_ctx* ctx, int32_t member);
int32_t fh_member_count(const fh_ctx* ctx);
void fh_destroy(fh_ctx* ctx);
const char* fh_last_error(const fh_ctx* ctx);
const char* fh_global_error(void);

/* SurfaceTopology (v1_parity.py:23-34).  active may be NULL (all active). */
int fh_upload_surface(fh_ctx* ctx, int64_t n_cells, int64_t n_edges,
                      const float* bed, const float* manning_n, const float* cell_area,
                      const float* characteristic_length,
                      const int32_t* edge_left, const int32_t* edge_right,
                      const float* edge_nx, const float* edge_ny, const float* edge_length,
                      const int32_t* boundary_code, const uint8_t* active);

/* CouplingTopology (v1_parity.py:37-49).  n_entries = surcharge table rows (may be 0). */
int fh_upload_coupling(fh_ctx* ctx, int64_t n_nodes, int64_t n_inlets, int64_t n_entries,
                       const int32_t* inlet_cell, const int32_t* inlet_node,
                       const int32_t* node_surface_cell, const float* node_invert,
                       const float* node_storage_area, const float* node_max_depth,
                       const float* inlet_capture_capacity, const float* inlet_surcharge_capacity,
                       const int32_t* surcharge_cell, const int32_t* surcharge_node,
                       const float* surcharge_weight, int32_t surcharge_weight_normalized);

/* PipeTopology (v1_parity.py:52-69); outfall_* arrays have n_nodes entries. */
int fh_upload_pipe(fh_ctx* ctx, int64_t n_links,
                   const int32_t* link_from, const int32_t* link_to, const float* link_length,
                   const float* link_manning, const float* link_capacity, const float* link_invert,
                   const float* link_diameter, const float* link_minor_loss,
                   const int32_t* link_geometry_code, const float* link_rect_width,
                   const float* link_rect_height, const float* outfall_area,
                   const float* outfall_coefficient, const float* outfall_tailwater,
                   const float* outfall_capacity);

/* RechargeTopology (v1_parity.py:72-77). */
int fh_upload_recharge(fh_ctx* ctx, int64_t n_rows, const int32_t* source_cell,
                       const float* storage_capacity, const float* capture_capacity,
                       const float* recharge_capacity, const float* stop_depth);

/* ParityForcing (parity_inputs.py:36-46).  external_q is [n_times][n_receivers] row-major. */
int fh_set_forcing(fh_ctx* ctx, int64_t n_intervals, const float* rain_start_s,
                   const float* rain_end_s, const float* rain_rate_mps,
                   const float* rainfall_multiplier, const float* horton_f0_mps,
                   const float* horton_fc_mps, const float* horton_k_per_s,
                   int64_t n_times, const float* external_times_s,
                   int64_t n_receivers, const int32_t* external_receiver_cell,
                   const float* external_q_m3ps);

/* Stage values of the code-3 (Dirichlet stage) boundary edges of the selected
 * member: TriangularSurfaceForcing.boundary_stage_m
 * (gurugram_flood/kernels/triangular_backend.py:93-101), applied as the V1
 * ghost of triangular_backend.py:1294-1327 / triangular_swe_cuda_kernels.py:86-93:
 * depth max(stage - z_left, 0), zero momentum, the left cell's bed.
 * n_values == 1: one stage for every edge; n_values == the n_edges given to
 * fh_upload_surface: one per edge in the caller's order; 0 or NULL: none, and
 * code 3 copies the left state as v1_parity.py's step does.  The local-inertial
 * step treats code 3 as a wall either way.  A stage edge whose own (left) cell
 * is inactive lets the stage flow in against a zero state; that inflow counts as
 * boundary volume and is removed with the inactive cell's water, as in the V1
 * step (v1_parity.py:170-214). */
int fh_set_boundary_stage(fh_ctx* ctx, int64_t n_values, const float* stage_m);

/* ParityState (v1_parity.py:80-87).  NULL leaves a field unchanged (set) or unread (get). */
int fh_set_state(fh_ctx* ctx, const float* h, const float* hu, const float* hv,
                 const float* wet_time_s, const float* node_volume, const float* link_flow,
                 const float* recharge_storage);
int fh_get_state(fh_ctx* ctx, float* h, float* hu, float* hv, float* wet_time_s,
                 float* node_volume, float* link_flow, float* recharge_storage);
/* FP64 views of the cell state (meaningful when precision == 64). */
int fh_get_state_f64(fh_ctx* ctx, double* h, double* hu, double* hv);
/* FP64 views of the node volumes and link flows (meaningful when precision == 64);
 * NULL leaves a field unread. */
int fh_get_drainage_f64(fh_ctx* ctx, double* node_volume, double* link_flow);

/* Clock and ledger (parity_runner.py:54-61 AdvanceCarry).  The ledger is 13
 * FP64 terms in MassLedger order, kept on the device. */
int fh_set_clock(fh_ctx* ctx, double time_s, double cached_dt_s, int64_t global_step);
int fh_get_clock(fh_ctx* ctx, fh_clock* out);
int fh_set_ledger(fh_ctx* ctx, const double* ledger13);
int fh_get_ledger(fh_ctx* ctx, double* ledger13);

/* _advance_to_time_while: step while time < target_s and fewer than max_steps
 * steps were taken in this call; dt = max(min(cfl_or_reuse, max_dt, final_s - t), floor).
 * Blocks until done.  Returns FH_ERR_BUDGET when the budget ran out first and
 * FH_ERR_SIMULATION when depth or node volume became non-finite. */
int fh_advance_to(fh_ctx* ctx, double target_s, double final_s, int64_t max_steps,
                  fh_clock* clock_out, double* ledger_out);

/* advance_parity_step with an explicit dt and rain rate (v1_parity.py:430-469):
 * one step, external inflow q given per receiver (may be NULL = 0).  diag_out gets
 * the 13 step diagnostics (StepDiagnostics order).  The clock advances by dt. */
int fh_step(fh_ctx* ctx, double dt, double rain_rate_mps, const float* external_q_m3ps,
            double* diag_out);

/* stable_timestep (v1_parity.py:153-158) on the resident state. */
int fh_stable_timestep(fh_ctx* ctx, double* dt_out);

/* Allocation-free stable_timestep.  No counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The CFL minimum of
 * fh_stable_timestep (v1_parity.py:153-158: min over cells of
 * cfl * Lc / (|u| + sqrt(g max(h, 1e-5))), non-finite values skipped, +inf for
 * no cells), on caller-owned DEVICE arrays h, hu, hv and characteristic_length
 * (float32, n_cells each); the result (a float32 value, as double) is written to
 * the DEVICE address dt_out.  Enqueued on `stream` (a hipStream_t, NULL = the
 * null stream) without synchronizing; calls no allocating routine.  Its block
 * minima go to a fixed 8 KiB device array per GPU (1,024 doubles, held until the
 * process ends), written before they are read in every call, so no state
 * carries between calls; calls that overlap on one GPU (other streams or host
 * threads) race on that array.  The first call on a device looks the array's
 * address up: make it outside a graph capture.  n_cells must be at most
 * FH_NOALLOC_MAX_CELLS; larger values return FH_ERR_VALUE (message via
 * fh_global_error). */
#define FH_NOALLOC_MAX_CELLS 2147483647LL
int fh_stable_timestep_noalloc(const float* h, const float* hu, const float* hv,
                               const float* characteristic_length, int64_t n_cells, double cfl,
                               double* dt_out, void* stream);

/* Sewer-only dry-weather flow, step_triangular_pipe_dry_weather_network
 * (gurugram_flood/kernels/triangular_backend.py:833-952): node inflow (m^3/s per
 * node in the caller's order, NULL = none), the link update and outfalls of the
 * coupled step, the overflow above storage capacity reported and capped when
 * cap_node_overflow != 0.  No surface exchange; the clock and the run ledger are
 * unchanged.  outfall_tailwater_head_m (finite, per node, NULL = the pipe
 * table's) overrides the tailwater.  diag_out gets 5 values: dry-weather inflow,
 * sum of |link transfer|, outfall discharge, overflow, final node volume (m^3).
 * Single-scenario contexts only. */
int fh_dry_weather_step(fh_ctx* ctx, double dt, const float* node_inflow_m3ps,
                        const float* outfall_tailwater_head_m, int32_t cap_node_overflow, double* diag_out);

/* Measurement: runs `steps` steps of the plain step sequence (no graph), timing
 * each kernel with HIP events on the context stream.  names/ms/bytes are
 * arrays of length cap; returns the number of kernels reported in *count.
 * bytes = algorithmic bytes per launch of that kernel. */
int fh_profile_kernels(fh_ctx* ctx, int32_t steps, int32_t cap, int32_t* count,
                       char* names /* cap x 32 chars */, double* avg_ms, double* bytes,
                       int64_t* launches);

/* Size of the resident device footprint in bytes. */
int64_t fh_device_bytes(const fh_ctx* ctx);

/* Verification helper (no counterpart in the reference): checks on device 0
 * the division the float32 kernels use (a shared, Newton-refined reciprocal and
 * one remainder correction) against IEEE division.  counts[6]: reciprocal
 * mismatches and reciprocals tested (every significand at nine exponents from
 * 2^-100 to 2^99), quotient mismatches and quotients tested over random pairs
 * in the exact domain (a = +-0, or 2^-100 <= |a|, |a / b| <= 2^100, with
 * 2^-100 <= b <= 2^100), and the same two for dividends below 2^-100 (outside
 * the domain: at most one ulp off). */
int fh_division_selftest(int64_t* counts);

/* Verification helper (no counterpart in the reference): checks on device 0
 * the power x^p, p = (float)(4/3), that the float32 friction terms use (FP64
 * Newton step on a hardware estimate) against the correctly rounded value over
 * every float x in [2^-40, 2^20).  counts[5]: its mismatches, the values
 * tested, the device library powf's mismatches on the same values, and the
 * largest ulp distance of each. */
int fh_pow_selftest(int64_t* counts);

/* Measurement helper (no counterpart in the reference): times `reps` launches
 * of a streaming copy of n_floats floats (n_floats % 4 == 0) on device 0, one
 * 16-byte element per thread, with ordinary (nontemporal == 0) or non-temporal
 * loads and stores, and returns the average milliseconds per launch.  Used as
 * the reference point for HBM byte counters and achievable bandwidth. */
int fh_copy_probe(int64_t n_floats, int32_t reps, int32_t nontemporal, double* avg_ms);

#ifdef __cplusplus
}
#endif
#endif /* FLOODHIP_H */
