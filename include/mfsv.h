/*
 * mfsv.h -- C ABI of libmfsv_hip.so, the MI355X (gfx950) implementation of
 * mfcm's explicit 2-D Saint-Venant time step on the Cartesian mesh.
 *
 * Plain C: no C++ types, no exceptions and no torch types cross this
 * boundary.  Every entry point returns MFSV_OK (0) or a negative error code;
 * mfsv_last_error(ctx) returns the message of the last failure on a context.
 * A context owns one HIP stream and its device allocations and is externally
 * synchronised (one host thread at a time per context).  Several contexts may
 * share a device; the batch API advances many catchments per launch.
 *
 * The reference has no C ABI: its boundary is C++ templates and virtuals.  The
 * comment on each entry point names the reference interface it replaces
 * (file:line relative to the reference's src/).
 */
#ifndef MFSV_H
#define MFSV_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MFSV_OK 0
#define MFSV_ERR_ARG (-1)         /* bad argument (null, size, range)       */
#define MFSV_ERR_HIP (-2)         /* a HIP runtime call failed              */
#define MFSV_ERR_STATE (-3)       /* call not valid in the current state    */
#define MFSV_ERR_NOMEM (-4)       /* device allocation failed / over budget */
#define MFSV_ERR_UNSUPPORTED (-5) /* option not available in this build     */

/* Value precision of a context.  Time, the controller and all mass/volume
 * accumulation are FP64 in both. */
#define MFSV_F64 0 /* plain method: the reference's arithmetic in FP64 */
#define MFSV_F32 1 /* precision form: FP32 depth-based arithmetic      */

/* Extension flags (0 = the reference's semantics). */
#define MFSV_E1_RAIN_GUARD 0x01   /* discharge dry guard h + dh*dt <= 0     */
#define MFSV_E2_WEST_OUTFLOW 0x02 /* transmissive outflow on the x = 0 edge */
#define MFSV_E4_STORM 0x08        /* analytic drifting Gaussian storm        */
#define MFSV_E7_Y_ORIENT 0x40     /* y slope taken south->north             */
#define MFSV_E8_DRY_VELOCITY 0x80 /* completed stage states: u = v = 0
                                     where h <= h_dry                        */
#define MFSV_E9_VELOCITY_CAP 0x100 /* completed stage states: |u|, |v| <=
                                      u_max                                 */

/* Cell fields (Field/DataArray containers of the reference, one value per
 * cell, row-major i = y*nx + x as Mesh/Cartesian2DMesh.hpp:308-309). */
#define MFSV_FIELD_H 0         /* depth h of the current state U0         */
#define MFSV_FIELD_U 1         /* x velocity u                            */
#define MFSV_FIELD_V 2         /* y velocity v                            */
#define MFSV_FIELD_ZB 3        /* bed elevation (NaN = coded out)         */
#define MFSV_FIELD_MANNING_N 4 /* per-cell Manning n (shallow = deep)     */
#define MFSV_FIELD_QBDY0 5     /* discharge endpoint at step start        */
#define MFSV_FIELD_QBDY1 6     /* discharge endpoint at step end          */
#define MFSV_FIELD_HBDY0 7     /* head endpoint at step start (NaN = off) */
#define MFSV_FIELD_HBDY1 8     /* head endpoint at step end (NaN = off)   */
#define MFSV_FIELD_DHDT 9      /* k of the last mfsv_dudt call, h part    */
#define MFSV_FIELD_DUDT 10
#define MFSV_FIELD_DVDT 11
#define MFSV_FIELD_BRANCH_X 12 /* x-face branch codes, (nx+1)*ny values   */
#define MFSV_FIELD_BRANCH_Y 13 /* y-face branch codes, nx*(ny+1) values   */
#define MFSV_FIELD_ELOSS_FX 14 /* energy-loss coefficient along x (1/m)  */
#define MFSV_FIELD_ELOSS_FY 15 /* energy-loss coefficient along y (1/m)  */
#define MFSV_FIELD_INFIL_RATE 16 /* infiltration rate (m/s)              */
#define MFSV_FIELD_INFIL_CAP 17  /* remaining infiltration capacity (m);
                                    state: every stage consumes it       */
#define MFSV_FIELD_SBDY0 18    /* stage boundary: water level at step start
                                  (NaN = off)                            */
#define MFSV_FIELD_SBDY1 19    /* stage boundary: water level at step end  */
#define MFSV_FIELD_EDDY_MU 20  /* eddy viscosity coefficient (constant
                                  model)                                 */
/* Per-cell Manning parameters (ManningRoughnessSourceTerm.hpp:129-140, the
 * fields n_shallow, n_deep, d_shallow, d_deep).  When all three are uploaded
 * MANNING_N is n_shallow (it is required then). */
#define MFSV_FIELD_N_DEEP 21
#define MFSV_FIELD_D_SHALLOW 22
#define MFSV_FIELD_D_DEEP 23
/* Output fields filled by mfsv_diagnose (download only).  Face fields hold
 * (nx+1)*ny x-faces then nx*(ny+1) y-faces in the reference's face order
 * (Mesh/Cartesian2DMesh.hpp:150-164): the fluxes of Fluxes.hpp:81-97. */
#define MFSV_FIELD_FLUX_H 24
#define MFSV_FIELD_FLUX_U 25
#define MFSV_FIELD_FLUX_V 26
#define MFSV_FIELD_FLUX_Z 27
#define MFSV_FIELD_MANNINGS_N 28    /* n(h) of the Manning term           */
#define MFSV_FIELD_FRICTION_SLOPE 29 /* S_f (0 where h <= 1e-6)            */
#define MFSV_FIELD_DZDX_BED 30      /* Minmod3 bed slopes (Constants.cpp)  */
#define MFSV_FIELD_DZDY_BED 31
/* z_bed + h of U0 at the last mfsv_update_stage_field (Solver.hpp:90-98). */
#define MFSV_FIELD_STAGE 32
#define MFSV_FIELD_COUNT 33

typedef struct mfsv_ctx mfsv_ctx;

/* Mesh geometry as Cartesian2DMesh (Mesh/Cartesian2DMesh.cpp:25-61) with
 * zero rotation: dx, dy are the cell sizes, origin the lower-left corner. */
typedef struct mfsv_grid {
  int64_t nx, ny;
  double dx, dy;
  double origin_x, origin_y;
} mfsv_grid;

typedef struct mfsv_params {
  int32_t flags;    /* MFSV_E* bits                                     */
  double cn_target; /* controller target; 1.0 in RungeKutta.cpp:154      */
  /* Manning source (ManningRoughnessSourceTerm.cpp:49-109); a per-cell n
   * uploaded as MFSV_FIELD_MANNING_N replaces n_shallow and n_deep. */
  int32_t manning;
  double n_shallow, n_deep, d_shallow, d_deep;
  /* Discharge boundary blocks (DischargeBoundarySourceTerm): 0 = none,
   * 1 = global selection with scalar per-step endpoints (mfsv_set_forcing
   * or the forcing table), 2 = per-cell endpoint fields QBDY0/QBDY1. */
  int32_t discharge;
  /* Head boundary (HeadBoundarySourceTerm): 0 = none, 1 = the columns
   * [head_x0, head_x1) with scalar endpoints, 2 = per-cell HBDY0/HBDY1. */
  int32_t head;
  int64_t head_x0, head_x1;
  /* E4 storm: rate = peak*exp(-r^2/(2 sigma^2)) around (cx0+vx t, cy0+vy t). */
  double storm_peak, storm_sigma, storm_cx0, storm_cy0, storm_vx, storm_vy;
  double h_dry; /* E8 dry depth (m)          */
  double u_max; /* E9 velocity cap (m/s)     */
  /* Further source terms and the stage boundary; each 0 = off.  They run
   * after Manning in the order energy loss, infiltration, eddy viscosity,
   * and the stage boundary after the head boundary.
   * energy_loss 1 (EnergyLossSourceTerm.cpp:59-88): per-cell ELOSS_FX/FY.
   * infiltration 1 (InfiltrationSourceTerm.cpp:42-64): per-cell INFIL_RATE
   *   and INFIL_CAP (capacity consumed at every stage).
   * stage 1 (StageBoundarySourceTerm.cpp:43-68): per-cell SBDY0/SBDY1.
   * eddy (EddyViscositySourceTerm.cpp:44-98, .hpp:216-256): 1 = per-cell
   *   EDDY_MU, 2 = Smagorinsky with mu_const and mu_smag. */
  int32_t energy_loss;
  int32_t infiltration;
  int32_t stage;
  int32_t eddy;
  double mu_const, mu_smag;
} mfsv_params;

typedef struct mfsv_stats {
  double time;         /* global time of U0                                */
  double t_local;      /* time within the current outer step               */
  double dt;           /* next timestep                                    */
  double cn;           /* control number of the last timestep              */
  double volume;       /* FP64 sum of h*dx*dy of U0                        */
  double source_volume;/* FP64 integral of the net boundary + source rate  */
  double volume0;      /* volume when the run started                      */
  int64_t step;        /* outer steps completed                            */
  int64_t nsteps;      /* outer steps in the run                           */
  int64_t accepted;    /* accepted timesteps                               */
  int64_t rejected;    /* rejected (repeated) timesteps                    */
  int32_t done;        /* 1 once every outer step has completed            */
  int32_t current;     /* which state buffer holds U0                      */
} mfsv_stats;

/* ---- lifetime (replaces SaintVenantSolver's constructor, Solver.cpp:36-71,
 * and the SYCL queue selection of sycl.cpp:22-126) ---------------------- */
const char* mfsv_version(void);
int mfsv_device_count(int* count);
/* Each field must stay below 2 GiB (fewer than 2^28 cells in FP64, 2^29 in
 * FP32): larger grids return MFSV_ERR_ARG. */
int mfsv_create(const mfsv_grid* grid, int precision, int device, mfsv_ctx** out);
int mfsv_destroy(mfsv_ctx* ctx);
const char* mfsv_last_error(const mfsv_ctx* ctx);
int mfsv_synchronize(mfsv_ctx* ctx);
/* Bytes of device memory the context holds. */
int mfsv_device_bytes(const mfsv_ctx* ctx, int64_t* bytes);

/* ---- configuration ------------------------------------------------------ */
int mfsv_set_params(mfsv_ctx* ctx, const mfsv_params* p);
/* Butcher table, (nstages+1)^2 doubles row-major laid out as the rows of
 * RungeKuttaCoefficientSet (TemporalScheme/RungeKutta.hpp:59-123):
 * row i < n = {c_i, a_i1..}, row n = {0, b_1..b_n}.  1 <= nstages <= 4. */
int mfsv_set_scheme(mfsv_ctx* ctx, int nstages, const double* tableau);
/* Work decomposition of the stage kernel: interior rows per wavefront tile
 * (1..64; 16 by default).  Results do not depend on it; the autotuner picks
 * it per grid shape.  No reference counterpart (SYCL chose its own ranges). */
int mfsv_set_tile_rows(mfsv_ctx* ctx, int rows);
int mfsv_get_tile_rows(const mfsv_ctx* ctx, int* rows);
/* TimeParameters (Config/TimeParameters.cpp:22-58), all in seconds. */
int mfsv_set_time(mfsv_ctx* ctx, double start, double end, double step_duration,
                  double initial_dt);
/* Start of an outer step at global time time_now (start_new_step,
 * Solver.cpp:119-129): the scalar discharge/head endpoints the boundaries
 * computed for [time_now, time_now + step) (NaN head = inactive). */
int mfsv_start_new_step(mfsv_ctx* ctx, double time_now, double q0, double q1, double hb0,
                        double hb1);
/* Per-outer-step endpoint table (4 doubles per step: q0, q1, hb0, hb1) that
 * the device controller loads at each step start. */
int mfsv_set_forcing_table(mfsv_ctx* ctx, int64_t nsteps, const double* table);

/* ---- fields (replaces DataArray<T>::move_to_device/move_to_host,
 * DataArray/DataArray.cpp:84-116).  Host arrays are FP64 in both
 * precisions; an F32 context rounds on upload. ---------------------------- */
int mfsv_upload(mfsv_ctx* ctx, int field, const double* host, int64_t count);
int mfsv_download(mfsv_ctx* ctx, int field, double* host, int64_t count);

/* ---- the time step ------------------------------------------------------- */
/* Temporal derivative of the current state U0 at stage time t_sub with
 * timestep dt into the DHDT/DUDT/DVDT fields: the slope passes, face fluxes,
 * cell pass, Manning, discharge and head terms of
 * SaintVenantSolver::update_dUdt (SaintVenant/Solver.cpp:77-114).  With
 * want_branches != 0 also fills BRANCH_X/BRANCH_Y (FluxKernel.cpp:291-293). */
int mfsv_dudt(mfsv_ctx* ctx, double t_sub, double dt, int want_branches);
/* Output-only fields of the reference's solver (the fields that
 * MultiFieldOutputFile and the check files write, Solver.hpp:124-161):
 * one derivative evaluation without side effects that fills the selected
 * output fields.  They depend on the state, the bed and the Manning
 * parameters only, so boundary and further source-term fields need not be
 * set (the infiltration capacity is not consumed).
 * source 0: the current state U0 at stage time t_sub with timestep dt.
 * source 1: the input of the last RK stage of the last timestep -- the state
 *   the reference's last update_dUdt saw, whose fluxes, Manning n and
 *   friction slope it holds at output time (t_sub and dt are then ignored).
 * what: MFSV_DIAG_* bits. */
#define MFSV_DIAG_FLUXES 1     /* FLUX_H, FLUX_U, FLUX_V, FLUX_Z           */
#define MFSV_DIAG_MANNING 2    /* MANNINGS_N, FRICTION_SLOPE (Manning on)   */
#define MFSV_DIAG_BED_SLOPES 4 /* DZDX_BED, DZDY_BED                        */
int mfsv_diagnose(mfsv_ctx* ctx, int source, double t_sub, double dt, int what);
/* end_of_step (Solver.hpp:90-98): STAGE = z_bed + h of U0. */
int mfsv_update_stage_field(mfsv_ctx* ctx);
/* SaintVenantSolver::control_number(state_no, dt) (State.cpp:199-229):
 * max over cells of dt*((|u| + c)/dx + (|v| + c)/dy), c = sqrt(9.81 max(h,0)),
 * of state 0 (U0), 1..nstages-1 (the intermediate RK rows of the last
 * timestep) or nstages (the last timestep's result, until it is accepted). */
int mfsv_control_number(mfsv_ctx* ctx, int state, double dt, double* cn);
/* Allocation-free control number.  No counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The same value as
 * mfsv_control_number, bit for bit, of a state the caller holds in device
 * memory: h, u, v are device arrays of nx*ny values (double for MFSV_F64,
 * float for MFSV_F32) laid out as the context fields; dx, dy are the cell
 * sizes.  The result dt * max_i((|u| + c)/dx + (|v| + c)/dy) is written to
 * cn_out, a device pointer to one double, by work queued on `stream` (a
 * hipStream_t; NULL = the null stream); the call does not synchronise.
 * Allocates nothing: the block maxima go to a fixed array of 4096 doubles
 * (32 KiB) in the library's device memory, one per GPU, held until the
 * process ends.  Grids of more than MFSV_NOALLOC_MAX_CELLS cells return
 * MFSV_ERR_ARG, as do bad sizes, spacings, dt or null pointers (before any
 * device call).  Calls on one GPU share that array: calls that overlap on
 * one GPU (concurrent streams or host threads) race and must be ordered by
 * the caller. */
#define MFSV_NOALLOC_MAX_CELLS ((int64_t)67108864) /* 4096 * 256 * 64 */
int mfsv_control_number_noalloc(int precision, int device, int64_t nx, int64_t ny, double dx,
                                double dy, const void* h, const void* u, const void* v,
                                double dt, double* cn_out, void* stream);
/* SpatialDerivativeOperator<T, Mesh, Cell, Minmod3>::apply<Axis>
 * (SpatialDerivative.hpp:48-53, 104-122; Minmod3.hpp:34-82) on an FP64 cell
 * field of the context's mesh: axis 0 = X (west, east), axis 1 = Y with the
 * reference's orientation (north as the left neighbour).  in and out hold
 * nx*ny values each. */
int mfsv_spatial_derivative(mfsv_ctx* ctx, int axis, const double* in, double* out,
                            int64_t count);
/* One whole timestep from U0 at local time t_local with timestep dt: every
 * RK stage (RungeKutta.cpp:143-153, update_k + update_y) fused per stage, and
 * the control number of the result (State.cpp:199-229) returned in *cn.
 * Synchronises the stream.  The result is held until mfsv_accept. */
int mfsv_timestep(mfsv_ctx* ctx, double t_local, double dt, double* cn);
/* accept_timestep (RungeKutta.hpp:428-431): the last result becomes U0. */
int mfsv_accept(mfsv_ctx* ctx);
/* FP64 volume sum(h)*dx*dy of U0 (Output/Measure.hpp:137-153). */
int mfsv_volume(mfsv_ctx* ctx, double* volume);

/* ---- device-controlled run (TemporalScheme::solve/step,
 * TemporalScheme.hpp:112-149 and :244-276, with the controller of
 * RungeKutta.cpp:154-167 evaluated on the device) ------------------------- */
/* Reset to outer step 0 with dt = initial dt (forcing table required when a
 * scalar discharge or head boundary is configured). */
int mfsv_run_begin(mfsv_ctx* ctx);
/* Attempt up to max_timesteps timesteps (fewer if the run finishes).
 * graph: 0 = launch each kernel; 1 = replay hipGraph-captured batches of
 * timesteps; 2 = launch each kernel and bracket every timestep's stage
 * kernels with HIP events (read with mfsv_last_stage_ms). */
int mfsv_run(mfsv_ctx* ctx, int64_t max_timesteps, int graph);
int mfsv_get_stats(mfsv_ctx* ctx, mfsv_stats* out);
/* The inner loop of one outer step (TemporalScheme::step,
 * TemporalScheme.hpp:112-149) on the device: after mfsv_run_begin and the
 * step's mfsv_start_new_step, attempt timesteps until the current outer step
 * is complete, so the host can prepare the next step's boundary fields
 * between steps.  No-op when every outer step is complete. */
int mfsv_run_step(mfsv_ctx* ctx, int graph);

/* Timing of the last mfsv_run: milliseconds measured with HIP events on the
 * context's stream around the launches of that call. */
int mfsv_last_run_ms(mfsv_ctx* ctx, double* ms);
/* Stage-kernel time of the last mfsv_run with graph == 2: the summed HIP
 * event intervals around the stage launches, and the number of launches. */
int mfsv_last_stage_ms(mfsv_ctx* ctx, double* ms, int64_t* launches);

/* ---- batches: many catchments advanced per launch.  No reference
 * counterpart (the reference runs one mesh per process, mfcm.cpp:38-49); the
 * batch drives the same per-catchment TemporalScheme loop as mfsv_run for
 * every member, with one stage launch per RK stage over the
 * concatenated tile tables and one controller launch for all members. ---- */
typedef struct mfsv_batch mfsv_batch;
int mfsv_batch_create(int precision, int device, mfsv_batch** out);
int mfsv_batch_destroy(mfsv_batch* batch);
const char* mfsv_batch_last_error(const mfsv_batch* batch);
/* The member must have the batch's precision, device and RK scheme and must
 * have been started with mfsv_run_begin.  The batch does not own it; its
 * fields and statistics stay readable through the member's own calls. */
int mfsv_batch_add(mfsv_batch* batch, mfsv_ctx* member);
int mfsv_batch_remove(mfsv_batch* batch, mfsv_ctx* member);
/* members: added and not removed; active: members not yet finished. */
int mfsv_batch_size(const mfsv_batch* batch, int* members, int* active);
/* Advance every active member by up to max_timesteps attempted timesteps
 * (graph != 0: replay hipGraph-captured chunks).  Finished members leave the
 * tile table; their results stay in their contexts. */
int mfsv_batch_run(mfsv_batch* batch, int64_t max_timesteps, int graph);
int mfsv_batch_last_run_ms(const mfsv_batch* batch, double* ms);

#ifdef __cplusplus
}
#endif
#endif /* MFSV_H */
