// sv_stage.hpp -- the fused Runge-Kutta stage kernel for one catchment on
// gfx950.
//
// One stage kernel launch performs, for every cell, what the reference does
// with ~15 SYCL kernels per stage (SaintVenantSolver::update_dUdt,
// Solver.cpp:77-114, plus RungeKuttaSolver::update_y, RungeKutta.hpp:371-405):
//   Minmod3 slopes of h, u, v, zb along x and y          (State.cpp:180-192)
//   face fluxes of every x- and y-face                   (FluxKernel.cpp:66-295)
//   cell pass: flux divergence, pressure, bed slope     (TemporalDerivativeKernel.cpp:47-99)
//   Manning friction                                     (ManningRoughnessSourceTerm.cpp:49-91)
//   discharge (rain) and head boundary terms             (DischargeBoundarySourceTerm.cpp:43-67,
//                                                         HeadBoundarySourceTerm.cpp:43-66)
//   RK accumulation y_i += k_s * (a_{i,s} dt) for every later row i
// and, in the last stage, the control-number maximum (State.cpp:199-229) and
// FP64 volume / source-rate partial sums.
//
// Work decomposition: one wavefront owns a strip of 64 columns (60 interior
// + 2 halo columns each side) and marches down `rows` rows plus a 2-row halo.
// The live rows are kept in registers (one row prefetched ahead);
// x-neighbours move between lanes with cross-lane shuffles; each face and
// each face difference is evaluated once per wave.  Every row load and store
// is a contiguous 64-element segment of a structure-of-arrays field.
#pragma once
#include "sv_math.hpp"

namespace mfsv {

constexpr int kWave = 64;
constexpr int kHalo = 2;
constexpr int kStrip = kWave - 2 * kHalo;  // interior columns per wave
constexpr int kWavesPerBlock = 4;
constexpr int kMaxStages = 4;

// Per-catchment device control block: the state of TypedTemporalScheme::step
// (TemporalScheme.hpp:112-149) held on the device.
struct Ctrl {
  double t_local;       // time within the current outer step
  double dt;            // timestep of the next attempt
  double step_dur;
  double start_time;
  double cn;            // control number of the last attempt (raw)
  double cn_target;
  double volume;        // FP64 volume of U0
  double srcvol;        // FP64 integral of sum(k_h)*area over accepted steps
  double q0, q1, hb0, hb1;  // current outer step's scalar forcing
  double gstep;         // global time at the start of the current outer step
  int64_t step;         // outer steps completed
  int64_t nsteps;
  int64_t n_acc, n_rej;
  unsigned long long cn_bits;  // atomic max of non-negative doubles
  int32_t cur;          // state buffer holding U0
  int32_t done;
  int32_t repeat;
  int32_t pad;
};

// Drifting Gaussian storm (extension E4): peak, sigma, centre at t = 0 and
// drift velocity.
struct Storm {
  double peak, sigma, cx0, cy0, vx, vy;
};

template <typename T>
struct StageArgs {
  int32_t nx, ny;           // cells per row / rows (nx*ny < 2^31)
  int32_t nrb, ntiles;      // row blocks per strip, tiles (= waves)
  int32_t rows;             // interior rows per tile
  int32_t stage, nstages;
  int32_t last;             // last stage: control number + volume
  int32_t flags;
  int32_t nanbed;           // the bed has coded-out (NaN) cells
  int32_t manning, discharge, head;
  int32_t hx0, hx1;
  T dx, dy, rdx, rdy, hdx, hdy;  // spacing, reciprocal, half spacing
  double dxd, dyd;                // FP64 spacing for the storm
  double ox, oy;
  double c_s;               // c coefficient of this stage
  double rstep;             // 1 / outer step duration
  double acoef[kMaxStages]; // a_{i, s+1} for the rows this stage feeds
  // inputs
  const T* in_h; const T* in_u; const T* in_v;   // stage input; null = U0
  T* Uh[2]; T* Uu[2]; T* Uv[2];
  const T* zb;
  const T* nfield;
  const T* q0f; const T* q1f; const T* hb0f; const T* hb1f;
  T n_sh, n_dp, d_sh, d_dp;
  T h_dry, u_max;           // E8 dry depth, E9 velocity cap
  Storm storm;
  // accumulators: dst[j]; null = the new-state buffer U[cur^1]
  T* acc_h[kMaxStages]; T* acc_u[kMaxStages]; T* acc_v[kMaxStages];
  T* k_h; T* k_u; T* k_v;
  float* br_x; float* br_y;
  Ctrl* ctrl;
  double* part_vol;         // [ntiles]
  double* part_rate;        // [ntiles], this stage's slice
};

template <typename T> struct Row { T h, u, v, z; };
template <typename T> struct FaceUV { T h, u, v, z; };

__device__ __forceinline__ double shfl_up1(double v) { return __shfl_up(v, 1, kWave); }
__device__ __forceinline__ float shfl_up1(float v) { return __shfl_up(v, 1, kWave); }
__device__ __forceinline__ double shfl_dn1(double v) { return __shfl_down(v, 1, kWave); }
__device__ __forceinline__ float shfl_dn1(float v) { return __shfl_down(v, 1, kWave); }

template <typename T>
__device__ __forceinline__ Row<T> shfl_up_row(const Row<T>& r) {
  return {shfl_up1(r.h), shfl_up1(r.u), shfl_up1(r.v), shfl_up1(r.z)};
}
template <typename T>
__device__ __forceinline__ Row<T> shfl_dn_row(const Row<T>& r) {
  return {shfl_dn1(r.h), shfl_dn1(r.u), shfl_dn1(r.v), shfl_dn1(r.z)};
}
template <typename T>
__device__ __forceinline__ FaceUV<T> shfl_dn_face(const FaceUV<T>& r) {
  return {shfl_dn1(r.h), shfl_dn1(r.u), shfl_dn1(r.v), shfl_dn1(r.z)};
}

// Workgroups are dispatched round-robin over the 8 XCDs; remap so that runs
// of consecutive tiles (vertically adjacent tiles sharing halo rows) land on
// one XCD and its L2.
__device__ __forceinline__ int xcd_block(int b, int nb) {
  const int nb8 = (nb / 8) * 8;
  if (b >= nb8) return b;
  return (b % 8) * (nb8 / 8) + b / 8;
}

// One separable factor of the E4 storm, exp(-(a - c)^2 / (2 sigma^2)) for
// coordinate a and centre c (the oracle's storm_axis).
template <typename T>
__device__ __forceinline__ T storm_axis(const Storm& s, double a, double c) {
  if constexpr (sizeof(T) == 8) {
    const double d = a - c;
    return exp_neg(-(d * d) / (2.0 * s.sigma * s.sigma));
  } else {
    const float d = float(a - c), sg = float(s.sigma);
    return __expf(-(d * d) / (2.0f * sg * sg));
  }
}

// Division by a per-cell value: correctly rounded in FP64 (the reference's
// a / b), reciprocal-based in the FP32 form.
__device__ __forceinline__ double vdiv(double a, double b) { return a / b; }
__device__ __forceinline__ float vdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }

// NACC: accumulator rows this stage writes (0 with KOUT: write k instead).
// QUICK: divide by dx, dy with one multiplication by the reciprocal (exact
// for power-of-two spacings; the FP32 form always does).
template <typename T, int NACC, bool KOUT, bool QUICK>
__device__ __forceinline__ void stage_tile(const StageArgs<T>& A, const int tile, const int lane) {
  constexpr bool kRel = sizeof(T) == 4;
  Ctrl* __restrict__ C = A.ctrl;
  if (C->done) return;

  const int nx = A.nx, ny = A.ny;
  const int strip = tile / A.nrb, rb = tile - strip * A.nrb;
  const int x = strip * kStrip - kHalo + lane;
  const int y0 = rb * A.rows;
  const int y1 = (y0 + A.rows < ny) ? y0 + A.rows : ny;
  const bool xin = x >= 0 && x < nx;
  const bool own = lane >= kHalo && lane < kWave - kHalo && xin;
  const bool hasW = x > 0, hasE = x < nx - 1;
  const bool e7 = (A.flags & 0x40) != 0;
  const bool nanbed = A.nanbed != 0;

  const double dt = C->dt;
  const double tsub = C->t_local + A.c_s * dt;
  const int cur = C->cur;
  // select, not index: a run-time index into the argument block would copy
  // it to private memory
  T* const Uh0 = cur ? A.Uh[1] : A.Uh[0], * const Uh1 = cur ? A.Uh[0] : A.Uh[1];
  T* const Uu0 = cur ? A.Uu[1] : A.Uu[0], * const Uu1 = cur ? A.Uu[0] : A.Uu[1];
  T* const Uv0 = cur ? A.Uv[1] : A.Uv[0], * const Uv1 = cur ? A.Uv[0] : A.Uv[1];
  const T* __restrict__ xh = A.in_h ? A.in_h : Uh0;
  const T* __restrict__ xu = A.in_u ? A.in_u : Uu0;
  const T* __restrict__ xv = A.in_v ? A.in_v : Uv0;
  const T* __restrict__ zb = A.zb;
  const T dx = A.dx, dy = A.dy, rdx = A.rdx, rdy = A.rdy, hdx = A.hdx, hdy = A.hdy;
  const T G = T(kG);

  // RK coefficients of this stage: (T)(a * dt) as State::operator*(ValueType).
  T sc[NACC > 0 ? NACC : 1];
  T* dh_[NACC > 0 ? NACC : 1];
  T* du_[NACC > 0 ? NACC : 1];
  T* dv_[NACC > 0 ? NACC : 1];
#pragma unroll
  for (int j = 0; j < NACC; ++j) {
    sc[j] = T(A.acoef[j] * dt);
    dh_[j] = A.acc_h[j] ? A.acc_h[j] : Uh1;
    du_[j] = A.acc_u[j] ? A.acc_u[j] : Uu1;
    dv_[j] = A.acc_v[j] ? A.acc_v[j] : Uv1;
  }

  // Uniform discharge / head terms (global selection, scalar endpoints).
  const double step_len = C->step_dur;
  const double q0s = C->q0, q1s = C->q1, hb0s = C->hb0, hb1s = C->hb1;
  const double gstep = C->gstep;
  const bool storm = (A.flags & 0x08) != 0;
  // E4 storm factors: this lane's column at the step's two endpoint times,
  // and (in lane k) row y0 + k, broadcast to all lanes row by row.
  T ex0 = T(0), ex1 = T(0), eyA0 = T(0), eyA1 = T(0);
  if (storm) {
    const Storm& st = A.storm;
    const double t1 = gstep + step_len;
    const double xc = A.ox + (double(x) + 0.5) * A.dxd;
    ex0 = storm_axis<T>(st, xc, st.cx0 + st.vx * gstep);
    ex1 = storm_axis<T>(st, xc, st.cx0 + st.vx * t1);
    const double yc = A.oy + (double(y0 + lane) + 0.5) * A.dyd;
    eyA0 = storm_axis<T>(st, yc, st.cy0 + st.vy * gstep);
    eyA1 = storm_axis<T>(st, yc, st.cy0 + st.vy * t1);
  }

  auto load = [&](int r) {
    Row<T> R;
    if (xin && r >= 0 && r < ny) {
      const int i = r * nx + x;
      R.h = xh[i]; R.u = xu[i]; R.v = xv[i]; R.z = zb[i];
    } else {
      R.h = T(0); R.u = T(0); R.v = T(0); R.z = T(0);
    }
    return R;
  };

  Row<T> R0 = load(y0 - 2), R1 = load(y0 - 1), R2 = load(y0);
  // Per-cell inputs of the cell pass (Manning n, accumulator bases of a
  // later stage), loaded one row before the row that uses them.
  constexpr int NB = NACC > 0 ? NACC : 1;
  struct Extra {
    T n;
    T bh[NB], bu[NB], bv[NB];
  };
  const bool base_ld = A.stage > 0;
  auto load_extra = [&](int r) {
    Extra e;
    e.n = A.n_sh;
#pragma unroll
    for (int j = 0; j < NB; ++j) { e.bh[j] = T(0); e.bu[j] = T(0); e.bv[j] = T(0); }
    if (own && r >= y0 && r < y1) {
      const int i = r * nx + x;
      if (A.nfield) e.n = A.nfield[i];
      if (base_ld) {
#pragma unroll
        for (int j = 0; j < NACC; ++j) { e.bh[j] = dh_[j][i]; e.bu[j] = du_[j][i]; e.bv[j] = dv_[j][i]; }
      }
    }
    return e;
  };
  Extra X0 = load_extra(y0 - 1);
  // south quotients (U_c - U_{c-1})/dy of the current row, carried down
  // (first row: rows y0-2 and y0-1; unused when y0-1 has no south neighbour)
  T qsh = dv<QUICK>(R1.h - R0.h, dy, rdy), qsu = dv<QUICK>(R1.u - R0.u, dy, rdy);
  T qsv = dv<QUICK>(R1.v - R0.v, dy, rdy), qsz = dv<QUICK>(R1.z - R0.z, dy, rdy);
  // slopes of row c-1 that the cell pass and the y-face still need
  T p_hx = T(0), p_zx = T(0), p_hy = T(0), p_uy = T(0), p_vy = T(0), p_zy = T(0);
  FaceUV<T> Fs0 = {}, Xw0 = {};
  double vol = 0.0, rate = 0.0, cnmax = 0.0;

#pragma unroll 3
  for (int c = y0 - 1; c <= y1; ++c) {
    const Row<T> Rn = load(c + 2);  // prefetch
    const Extra X1 = load_extra(c);
    const bool rowin = c >= 0 && c < ny;
    const Row<T> W = shfl_up_row(R1), E = shfl_dn_row(R1);
    // ---- slopes of row c (State.cpp:187-192, SpatialDerivative.hpp:104-122)
    T shx = T(0), sux = T(0), svx = T(0), szx = T(0);
    T shy = T(0), suy = T(0), svy = T(0), szy = T(0);
    T qnh = T(0), qnu = T(0), qnv = T(0), qnz = T(0);
    if (rowin) {
      const bool hasN = c + 1 < ny, hasS = c > 0;
      if (!nanbed) {
        const T qwh = dv<QUICK>(R1.h - W.h, dx, rdx), qwu = dv<QUICK>(R1.u - W.u, dx, rdx);
        const T qwv = dv<QUICK>(R1.v - W.v, dx, rdx), qwz = dv<QUICK>(R1.z - W.z, dx, rdx);
        const T qeh = shfl_dn1(qwh), qeu = shfl_dn1(qwu), qev = shfl_dn1(qwv), qez = shfl_dn1(qwz);
        shx = minmod_q(qwh, qeh, dv<QUICK>(E.h - W.h, dx, rdx), hasW, hasE);
        sux = minmod_q(qwu, qeu, dv<QUICK>(E.u - W.u, dx, rdx), hasW, hasE);
        svx = minmod_q(qwv, qev, dv<QUICK>(E.v - W.v, dx, rdx), hasW, hasE);
        szx = minmod_q(qwz, qez, dv<QUICK>(E.z - W.z, dx, rdx), hasW, hasE);
        qnh = dv<QUICK>(R2.h - R1.h, dy, rdy); qnu = dv<QUICK>(R2.u - R1.u, dy, rdy);
        qnv = dv<QUICK>(R2.v - R1.v, dy, rdy); qnz = dv<QUICK>(R2.z - R1.z, dy, rdy);
        // south -> north slope; the reference's north -> south slope is its
        // exact negation (Minmod3 commutes with negation)
        const T yh = minmod_q(qsh, qnh, dv<QUICK>(R2.h - R0.h, dy, rdy), hasS, hasN);
        const T yu = minmod_q(qsu, qnu, dv<QUICK>(R2.u - R0.u, dy, rdy), hasS, hasN);
        const T yv = minmod_q(qsv, qnv, dv<QUICK>(R2.v - R0.v, dy, rdy), hasS, hasN);
        const T yz = minmod_q(qsz, qnz, dv<QUICK>(R2.z - R0.z, dy, rdy), hasS, hasN);
        shy = e7 ? yh : -yh; suy = e7 ? yu : -yu; svy = e7 ? yv : -yv; szy = e7 ? yz : -yz;
      } else {
        // general Minmod3 with the reference's NaN rules
        const Row<T> Wc = hasW ? W : R1, Ec = hasE ? E : R1;
        shx = minmod3(Wc.h, hasW, R1.h, hasE, Ec.h, dx, rdx);
        sux = minmod3(Wc.u, hasW, R1.u, hasE, Ec.u, dx, rdx);
        svx = minmod3(Wc.v, hasW, R1.v, hasE, Ec.v, dx, rdx);
        szx = minmod3(Wc.z, hasW, R1.z, hasE, Ec.z, dx, rdx);
        const Row<T> N = hasN ? R2 : R1, S = hasS ? R0 : R1;
        if (e7) {
          shy = minmod3(S.h, hasS, R1.h, hasN, N.h, dy, rdy);
          suy = minmod3(S.u, hasS, R1.u, hasN, N.u, dy, rdy);
          svy = minmod3(S.v, hasS, R1.v, hasN, N.v, dy, rdy);
          szy = minmod3(S.z, hasS, R1.z, hasN, N.z, dy, rdy);
        } else {
          shy = minmod3(N.h, hasN, R1.h, hasS, S.h, dy, rdy);
          suy = minmod3(N.u, hasN, R1.u, hasS, S.u, dy, rdy);
          svy = minmod3(N.v, hasN, R1.v, hasS, S.v, dy, rdy);
          szy = minmod3(N.z, hasN, R1.z, hasS, S.z, dy, rdy);
        }
      }
    }
    // ---- west x-face of row c (face id c*(nx+1) + x)
    const T Whx = shfl_up1(shx), Wux = shfl_up1(sux), Wvx = shfl_up1(svx), Wzx = shfl_up1(szx);
    FaceUV<T> Xw1 = {};
    if (c >= y0 && c < y1) {
      Side<T> R = {R1.h, R1.u, R1.v, R1.z, shx, sux, svx, szx};
      Side<T> L = {W.h, W.u, W.v, W.z, Whx, Wux, Wvx, Wzx};
      float br;
      Flux<T> F;
      if (x == 0 && (A.flags & 0x02)) {  // E2: zero-gradient ghost that only drains
        L = {R1.h, fmin_t(R1.u, T(0)), R1.v, R1.z, T(0), T(0), T(0), T(0)};
        F = face_flux<T, kRel>(L, R, 0, hdx, dx, rdx, &br);
      } else if (!nanbed) {
        if (x == 0) L = wall_ghost(R);
        else if (x == nx) R = wall_ghost(L);
        F = face_fast<T, kRel, QUICK>(L, R, hdx, dx, rdx, &br);
      } else {
        int edge = 0;
        if (x == 0) { L = R; edge = -1; }
        else if (x == nx) { R = L; edge = 1; }
        F = face_flux<T, kRel>(L, R, edge, hdx, dx, rdx, &br);
      }
      Xw1 = {F.h, F.n, F.t, F.z};
      if (A.br_x && x >= 0 && x <= nx && lane >= kHalo && lane <= kWave - kHalo)
        A.br_x[c * (nx + 1) + x] = br;
    }
    // ---- south y-face of row c (face id (nx+1)ny + c*nx + x)
    FaceUV<T> Fs1 = {};
    if (c >= y0) {
      Side<T> L, R;
      float br;
      Flux<T> F;
      if (c == ny) {  // north edge of the top row
        L = {R0.h, R0.v, R0.u, R0.z, p_hy, p_vy, p_uy, p_zy};
        if (!nanbed) F = face_fast<T, kRel, QUICK>(L, wall_ghost(L), hdy, dy, rdy, &br);
        else F = face_flux<T, kRel>(L, L, 1, hdy, dy, rdy, &br);
      } else {
        R = {R1.h, R1.v, R1.u, R1.z, shy, svy, suy, szy};
        if (c == 0) {
          if (!nanbed) F = face_fast<T, kRel, QUICK>(wall_ghost(R), R, hdy, dy, rdy, &br);
          else F = face_flux<T, kRel>(R, R, -1, hdy, dy, rdy, &br);
        } else {
          L = {R0.h, R0.v, R0.u, R0.z, p_hy, p_vy, p_uy, p_zy};
          if (!nanbed) F = face_fast<T, kRel, QUICK>(L, R, hdy, dy, rdy, &br);
          else F = face_flux<T, kRel>(L, R, 0, hdy, dy, rdy, &br);
        }
      }
      Fs1 = {F.h, F.t, F.n, F.z};
      if (A.br_y && own) A.br_y[c * nx + x] = br;
    }
    // ---- cell pass + sources + RK for row r = c - 1
    const FaceUV<T> Xe0 = shfl_dn_face(Xw0);
    const int r = c - 1;
    T ey0 = T(0), ey1 = T(0);
    if (storm) {  // uniform: every lane takes part in the broadcast
      const int k = r - y0;
      if (k >= 0 && k < kWave) {
        ey0 = __shfl(eyA0, k, kWave);
        ey1 = __shfl(eyA1, k, kWave);
      } else {
        const Storm& st = A.storm;
        const double yc = A.oy + (double(r) + 0.5) * A.dyd;
        ey0 = storm_axis<T>(st, yc, st.cy0 + st.vy * gstep);
        ey1 = storm_axis<T>(st, yc, st.cy0 + st.vy * (gstep + step_len));
      }
    }
    if (r >= y0 && r < y1 && own) {
      const int i = r * nx + x;
      const T h = R0.h, u = R0.u, v = R0.v;
      // TemporalDerivativeKernel.cpp:57-90
      T dhdt = dv<QUICK>(Xw0.h - Xe0.h, dx, rdx) + dv<QUICK>(Fs0.h - Fs1.h, dy, rdy);
      T dudt = dv<QUICK>(Xw0.u - Xe0.u, dx, rdx) + dv<QUICK>(Fs0.u - Fs1.u, dy, rdy);
      T dvdt = dv<QUICK>(Xw0.v - Xe0.v, dx, rdx) + dv<QUICK>(Fs0.v - Fs1.v, dy, rdy);
      dudt += p_hx * T(-kG);
      dvdt += p_hy * T(-kG);
      T dzdx = p_zx;
      const T hx_ = dv<QUICK>(h, dx, rdx);
      if (fabs_t(dzdx) > hx_) dzdx = sy_sign(dzdx) * hx_;
      T dzdy = p_zy;
      const T hy_ = dv<QUICK>(h, dy, rdy);
      if (fabs_t(dzdy) > hy_) dzdy = sy_sign(dzdy) * hy_;
      const T dudt_bed = T(-kG) * dzdx;
      const T dvdt_bed = T(-kG) * dzdy;
      dudt += dv<QUICK>(Xw0.z - Xe0.z, dx, rdx);
      dvdt += dv<QUICK>(Fs0.z - Fs1.z, dy, rdy);
      dudt += dudt_bed;
      dvdt += dvdt_bed;
      // ManningRoughnessSourceTerm.cpp:55-90
      if (A.manning) {
        T nsh = A.n_sh, ndp = A.n_dp;
        if (A.nfield) { nsh = X0.n; ndp = nsh; }
        T n = nsh;  // mix(a, a, t) == a
        if (nsh != ndp) {
          T t = vdiv(h - A.d_sh, A.d_dp - A.d_sh);
          t = fmin_t(fmax_t(t, T(0)), T(1));
          n = nsh + (ndp - nsh) * (t * t * (T(3) - T(2) * t));
        }
        if (h > T(1e-6)) {
          const T inv_h = vdiv(h, h * h + T(1e-3));
          const T Sf = n * n * fsqrt_t(u * u + v * v) * pow43(inv_h);
          T dum = T(-kG) * Sf * u;
          T dvm = T(-kG) * Sf * v;
          const double ts = dt;
          if (fabs_t(T(dum * ts)) > fabs_t(u) && sy_sign(T(dum * ts)) != sy_sign(u)) dum = T(-u / ts);
          if (fabs_t(T(dvm * ts)) > fabs_t(v) && sy_sign(T(dvm * ts)) != sy_sign(v)) dvm = T(-v / ts);
          dudt += dum;
          dvdt += dvm;
        }
      }
      // DischargeBoundarySourceTerm.cpp:43-67 (E1 flips the guard)
      if (A.discharge || storm) {
        T tq0, tq1;
        if (A.discharge == 2) { tq0 = A.q0f[i]; tq1 = A.q1f[i]; }
        else if (A.discharge == 1) { tq0 = T(q0s); tq1 = T(q1s); }
        else { tq0 = T(0); tq1 = T(0); }
        if (storm) {
          tq0 += T(A.storm.peak) * ex0 * ey0;
          tq1 += T(A.storm.peak) * ex1 * ey1;
        }
        if (tq0 != T(0) || tq1 != T(0)) {
          T dh;
          if constexpr (sizeof(T) == 8) {
            const double dqdt = div_c(tq1 - tq0, step_len, A.rstep);
            const double qnow = tq0 + tsub * dqdt;
            const double qnext = qnow + dt * dqdt;
            dh = 0.5 * (qnow + qnext);
          } else {
            const float dqdt = (tq1 - tq0) * float(A.rstep);
            const float qnow = __builtin_fmaf(float(tsub), dqdt, tq0);
            const float qnext = __builtin_fmaf(float(dt), dqdt, qnow);
            dh = 0.5f * (qnow + qnext);
          }
          const bool dry = (A.flags & 0x01) ? (h + dh * dt <= 0.0) : (h - dh * dt <= 0.0);
          if (dry) dh = T(-h / dt);
          dhdt += dh;
        }
      }
      // HeadBoundarySourceTerm.cpp:43-66 (overwrites dh)
      if (A.head) {
        T h0, h1;
        if (A.head == 2) { h0 = A.hb0f[i]; h1 = A.hb1f[i]; }
        else if (x >= A.hx0 && x < A.hx1) { h0 = T(hb0s); h1 = T(hb1s); }
        else { h0 = Num<T>::nan(); h1 = h0; }
        if (!(h0 != h0) && !(h1 != h1)) {
          const T dhb = T((h1 - h0) / step_len);
          const T now = T(h0 + tsub * dhb);
          const T next = T(now + dt * dhb);
          T target = T(0.5) * (now + next);
          if (target <= T(0)) target = T(0);
          dhdt = T((target - h) / dt);
        }
      }
      rate += double(dhdt);
      if (KOUT) {
        A.k_h[i] = dhdt; A.k_u[i] = dudt; A.k_v[i] = dvdt;
      } else {
        // RungeKutta.hpp:377-388: y_i = y_i + k_s * (a dt)
        T oh = h, ou = u, ov = v;
#pragma unroll
        for (int j = 0; j < NACC; ++j) {
          T bh = h, bu = u, bv = v;
          if (base_ld) { bh = X0.bh[j]; bu = X0.bu[j]; bv = X0.bv[j]; }
          T nh = bh + dhdt * sc[j];
          T nu = bu + dudt * sc[j];
          T nv = bv + dvdt * sc[j];
          if (j == 0) {  // row s+1 is complete after this stage
            if ((A.flags & 0x80) && nh <= A.h_dry) { nu = T(0); nv = T(0); }  // E8
            if (A.flags & 0x100) {                                           // E9
              nu = fmin_t(fmax_t(nu, -A.u_max), A.u_max);
              nv = fmin_t(fmax_t(nv, -A.u_max), A.u_max);
            }
            oh = nh; ou = nu; ov = nv;
          }
          dh_[j][i] = nh; du_[j][i] = nu; dv_[j][i] = nv;
        }
        if (A.last) {
          // State.cpp:211-222 on the new state
          const T hh = fmax_t(oh, T(0));
          const T cc = fsqrt_t(G * hh);
          const double cn = dt * double(dv<QUICK>(fabs_t(ou) + cc, dx, rdx) + dv<QUICK>(fabs_t(ov) + cc, dy, rdy));
          if (cn > cnmax) cnmax = cn;
          vol += double(oh);
        }
      }
    }
    R0 = R1; R1 = R2; R2 = Rn;
    X0 = X1;
    qsh = qnh; qsu = qnu; qsv = qnv; qsz = qnz;
    p_hx = shx; p_zx = szx; p_hy = shy; p_uy = suy; p_vy = svy; p_zy = szy;
    Fs0 = Fs1;
    Xw0 = Xw1;
  }

  // wave reductions (fixed butterfly order -> deterministic per tile)
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    rate += __shfl_xor(rate, off, kWave);
    if (A.last) {
      vol += __shfl_xor(vol, off, kWave);
      const double o = __shfl_xor(cnmax, off, kWave);
      cnmax = (o > cnmax) ? o : cnmax;
    }
  }
  if (lane == 0) {
    A.part_rate[tile] = rate;
    if (A.last) {
      A.part_vol[tile] = vol;
      atomicMax(&C->cn_bits, (unsigned long long)__double_as_longlong(cnmax));
    }
  }
}

// One catchment: one wavefront per tile.
template <typename T, int NACC, bool KOUT, bool QUICK>
__global__ __launch_bounds__(kWave * kWavesPerBlock) void k_stage(const StageArgs<T> A) {
  const int lane = threadIdx.x & (kWave - 1);
  const int tile = xcd_block(blockIdx.x, gridDim.x) * kWavesPerBlock + (threadIdx.x >> 6);
  if (tile >= A.ntiles) return;
  stage_tile<T, NACC, KOUT, QUICK>(A, tile, lane);
}

// Many catchments in one launch: a persistent grid walks a tile table of
// (member, tile) pairs; member m's arguments are margs[m].  Each tile runs
// exactly the code of the single-catchment kernel, so a catchment's result
// does not depend on which other catchments share the launch.  The table
// length is read from device memory so that captured graphs survive changes
// of membership.
template <typename T, int NACC, bool QUICK>
__global__ __launch_bounds__(kWave * kWavesPerBlock) void k_stage_batch(
    const StageArgs<T>* __restrict__ margs, const int2* __restrict__ tiles,
    const int* __restrict__ ntiles_p) {
  const int lane = threadIdx.x & (kWave - 1);
  const int nwaves = gridDim.x * kWavesPerBlock;
  const int ntiles = *ntiles_p;
  for (int k = xcd_block(blockIdx.x, gridDim.x) * kWavesPerBlock + (threadIdx.x >> 6); k < ntiles;
       k += nwaves) {
    const int m = __builtin_amdgcn_readfirstlane(tiles[k].x);
    const int t = __builtin_amdgcn_readfirstlane(tiles[k].y);
    stage_tile<T, NACC, false, QUICK>(margs[m], t, lane);
  }
}

}  // namespace mfsv
