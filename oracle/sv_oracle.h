/*
 * sv_oracle.h -- CPU oracle for the explicit 2-D Saint-Venant time step of
 * mfcm on its Cartesian mesh.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  It is never linked into, called by, or used as a fallback for the
 * product library (mfcm_amd/libmfsv_hip.so).
 *
 * It restates, in plain C (FP64 values, FP64 time), the reference's kernel
 * decomposition and operation order: Minmod3 slope passes -> face pass ->
 * cell pass -> source terms -> boundary terms -> Runge-Kutta field
 * arithmetic -> control number -> controller.  Each function in sv_oracle.c
 * cites the reference file:line it follows (paths relative to the reference's
 * src/ directory).
 *
 * Pinning (what checks each function; the tests live in tests/):
 *   or_minmod3           closed-form KATs incl. one-sided edges, ties, NaN
 *                        neighbours (test_oracle_kat.py)
 *   or_face              hand-computed KATs for branches 1, 2.25, 2.75, 3, 4,
 *                        edge faces and coded-out faces; flat-lake invariant
 *   or_td (cell pass)    lake-at-rest => dU/dt == 0 exactly; telescoping sum
 *   or_manning           closed form incl. overshoot cap; the shallow/deep
 *                        blend n = mix(n_sh, n_dp, smoothstep) with distinct
 *                        values below, inside and above [d_sh, d_dp]
 *   E2 outflow ghost     uniform westward flow passes h u at x = 0 (depth
 *                        steady there; a closed wall piles up -h u / dx);
 *                        flow away from the edge sees a ghost of zero normal
 *                        velocity (Rusanov average h u / 2)
 *   E4 storm             peak at the drifted centre, exp(-1/2) at one sigma,
 *                        the separable Gaussian at random points, and its
 *                        entry into the discharge term at each cell centre
 *   or_discharge/head    closed form incl. the dry guard
 *   energy loss          closed form incl. the overshoot cap and dry cut
 *   infiltration         closed form: rate, depth and capacity limits, the
 *                        capacity consumed once per application
 *   eddy viscosity       quadratic fields: exact second derivatives; the
 *                        Smagorinsky coefficient of linear fields
 *   stage boundary       closed form: target = level - bed (head invariant)
 *   RK + controller      tableau KATs; Heun-2 on a linear ODE; controller
 *                        table over cn; landing rules: the exact timestep
 *                        sequence into a step end with and without the
 *                        40/60 split
 *   volume               exact for constant fields
 * Parity unpinned (no reference test or golden vector exists and the
 * behaviour is implementation-defined in SYCL 1.2.1):
 *   - sycl::min with a NaN operand inside Minmod3 (only reachable next to
 *     coded-out, NaN-bed cells); restated as (b < a) ? b : a;
 *   - the last bits of sycl::pow in the Manning term: here x^(4/3) comes
 *     from or_pow43 (IEEE-only Newton iteration, within 7 ulp, the same
 *     bits as the HIP kernels);
 *   - smoothstep with d_shallow == d_deep (0/0); configs keep d_sh < d_dp;
 *   - sycl::pow(a, 2) in the Smagorinsky coefficient: restated as a * a;
 *   - the order of source terms and boundaries, which the reference takes
 *     from its config file: restated in one fixed order (Manning, energy
 *     loss, infiltration, eddy viscosity; discharge, head, stage).
 */
#ifndef MFSV_SV_ORACLE_H
#define MFSV_SV_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Extension flags.  All OFF = the reference's semantics ("parity mode"). */
#define OR_E1_RAIN_GUARD   0x01 /* discharge dry guard uses h + dh*dt <= 0 */
#define OR_E2_WEST_OUTFLOW 0x02 /* transmissive ghost on the x = 0 edge faces */
#define OR_E4_STORM        0x08 /* analytic drifting Gaussian storm rainfall */
#define OR_E7_Y_ORIENT     0x40 /* y slope taken south->north (physical sign) */
#define OR_E8_DRY_VELOCITY 0x80 /* no velocity in completed stage states with h <= h_dry */
#define OR_E9_VELOCITY_CAP 0x100 /* |u|, |v| <= u_max in completed stage states */

/* Face branch codes written to the branch face field (FluxKernel.cpp:206-293). */
#define OR_BRANCH_WET     1.0
#define OR_BRANCH_DRY_R   2.25
#define OR_BRANCH_DRY_L   2.75
#define OR_BRANCH_LEFT    3.0
#define OR_BRANCH_RIGHT   4.0

/* Cell-level data one side of a face needs (values and limited slopes). */
typedef struct or_cellv {
  double h, u, v, zb;
  double dzbdx, dzbdy;
  double dhdx, dudx, dvdx;
  double dhdy, dudy, dvdy;
} or_cellv;

/* Drifting Gaussian storm (extension E4): rate in m/s at cell centre (x, y)
 * and global time t is peak * exp(-dx^2/(2 sigma^2)) * exp(-dy^2/(2 sigma^2))
 * with (dx, dy) the offset from (cx0 + vx t, cy0 + vy t); exp from
 * or_exp_neg. */
typedef struct or_storm {
  double peak, sigma, cx0, cy0, vx, vy;
} or_storm;

typedef struct or_sim or_sim;

/* --- single-kernel restatements (pure functions, used by KAT tests) ----- */
double or_minmod3(double Ul, double dxl, double Uc, double dxr, double Ur);
/* One face.  edge in {-1, 0, 1}, dir 0 = x-face, 1 = y-face; L/R may alias.
 * out = {hflux, uflux, vflux, zflux, branch}. */
void or_face(const or_cellv* L, const or_cellv* R, int edge, int dir,
             double dx, double out[5]);
void or_slope_x(int64_t nx, int64_t ny, double dx, const double* f, double* out);
void or_slope_y(int64_t nx, int64_t ny, double dy, const double* f, double* out,
                int flags);
double or_storm_rate(const or_storm* s, double x, double y, double t);
double or_pow43(double x);
double or_exp_neg(double x);

/* --- whole-solver restatement ------------------------------------------ */
/* tableau: (nstages+1) x (nstages+1) row-major, rows as the reference's a_
 * (row i: c_i then a_{i,1..i}; last row: 0 then b_1..b_n), zero padded. */
or_sim* or_sim_create(int64_t nx, int64_t ny, double dx, double dy,
                      int nstages, const double* tableau, int flags,
                      double cn_target);
void or_sim_destroy(or_sim* S);
void or_sim_set_bed(or_sim* S, const double* zb);
void or_sim_set_state(or_sim* S, const double* h, const double* u,
                      const double* v);
void or_sim_get_state(const or_sim* S, int idx, double* h, double* u, double* v);
void or_sim_set_manning(or_sim* S, const double* n_sh, const double* n_dp,
                        const double* d_sh, const double* d_dp);
int or_sim_add_discharge(or_sim* S, int n, const double* t, const double* v,
                         const uint8_t* sel);
int or_sim_add_head(or_sim* S, int n, const double* t, const double* v,
                    const uint8_t* sel);
void or_sim_set_storm(or_sim* S, const or_storm* st, double origin_x,
                      double origin_y);
/* Stage boundary blocks (water-level time series on a selection). */
int or_sim_add_stage(or_sim* S, int n, const double* t, const double* v,
                     const uint8_t* sel);
/* Energy loss with per-cell coefficients fx, fy (per metre). */
void or_sim_set_energy_loss(or_sim* S, const double* fx, const double* fy);
/* Infiltration: per-cell rate (m/s) and initial capacity (m); the capacity
 * is state (array 26). */
void or_sim_set_infiltration(or_sim* S, const double* rate, const double* capacity);
/* Eddy viscosity: a per-cell constant mu, or (mu == NULL) Smagorinsky with
 * mu = mu_const + mu_smag * dx dy * sqrt(...). */
void or_sim_set_eddy_viscosity(or_sim* S, const double* mu, double mu_const,
                               double mu_smag);
void or_sim_set_time(or_sim* S, double start, double end, double step_duration,
                     double initial_dt);
void or_sim_set_dry_depth(or_sim* S, double h_dry);
void or_sim_set_velocity_cap(or_sim* S, double u_max);
int64_t or_sim_num_steps(const or_sim* S);

void or_sim_start_new_step(or_sim* S, double time_now);
void or_sim_update_dudt(or_sim* S, int state_no, double t_sub, double dt);
void or_sim_update_y(or_sim* S, int step, double dt);
double or_sim_control_number(or_sim* S, int state_no, double dt);
/* Returns 1 when the timestep must be repeated; *new_dt gets the next dt. */
int or_sim_do_timestep(or_sim* S, double t_local, double dt, double* new_dt);
void or_sim_accept(or_sim* S);
void or_sim_step(or_sim* S);               /* one outer step */
void or_sim_solve(or_sim* S);              /* all outer steps */
void or_sim_run_steps(or_sim* S, int64_t first, int64_t count);
double or_sim_dt(const or_sim* S);
void or_sim_set_dt(or_sim* S, double dt);
/* stats[0] accepted, [1] rejected timesteps. */
void or_sim_stats(const or_sim* S, int64_t* stats);
double or_sim_volume(const or_sim* S, int state_no);
/* Timestep trace: records (t_local, dt, cn, repeat) per attempted timestep
 * into a caller buffer of cap entries (4 doubles each); returns the count. */
void or_sim_set_trace(or_sim* S, double* buf, int64_t cap);
int64_t or_sim_trace_len(const or_sim* S);
/* Internal arrays for white-box tests.  which: 0 h, 1 u, 2 v (state idx);
 * 3..5 dUdt h,u,v (idx); 6..8 dUdx h,u,v; 9..11 dUdy h,u,v; 12 zb; 13 dzbdx;
 * 14 dzbdy; 15..19 faces hflux,uflux,vflux,zflux,branch; 20 qbdy0; 21 qbdy1;
 * 22 hbdy0; 23 hbdy1; 24 mannings_n; 25 friction_slope; 26 infiltration
 * capacity; 27 eddy mu; 28..31 d2u/dx2, d2u/dy2, d2v/dx2, d2v/dy2; 32, 33
 * stage endpoints. */
double* or_sim_array(or_sim* S, int which, int idx);

#ifdef __cplusplus
}
#endif
#endif
