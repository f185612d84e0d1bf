// sv_stage.hpp -- the fused Runge-Kutta stage kernel and the device
// controller for one catchment on gfx950.
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
// The three live rows are kept in registers; x-neighbours move between lanes
// with cross-lane shuffles; each face is evaluated exactly once per wave.
// Every row load and store is a contiguous 64-element segment of a
// structure-of-arrays field.
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

template <typename T>
struct StageArgs {
  int64_t nx, ny;
  int64_t nrb, ntiles;      // row blocks per strip, tiles (= waves)
  int32_t rows;             // interior rows per tile
  int32_t stage, nstages;
  int32_t nacc;             // accumulator rows updated by this stage
  int32_t last;             // last stage: control number + volume
  int32_t kout;             // write k to k_* instead of accumulating
  int32_t flags;
  int32_t manning, discharge, head;
  int64_t hx0, hx1;
  T dx, dy, rdx, rdy, hdx, hdy;  // spacing, reciprocal, half spacing
  double dxd, dyd, area;          // FP64 spacing for the storm / volume
  double ox, oy;
  double c_s;               // c coefficient of this stage
  double acoef[kMaxStages]; // a_{i, s+1} for the rows this stage feeds
  // inputs
  const T* in_h; const T* in_u; const T* in_v;   // stage input; null = U0
  T* Uh[2]; T* Uu[2]; T* Uv[2];
  const T* zb;
  const T* nfield;
  const T* q0f; const T* q1f; const T* hb0f; const T* hb1f;
  T n_sh, n_dp, d_sh, d_dp;
  T h_dry, u_max;           // E8 dry depth, E9 velocity cap
  double storm[6];          // peak, sigma, cx0, cy0, vx, vy
  // accumulators: dst[j]; null = the new-state buffer U[cur^1]
  T* acc_h[kMaxStages]; T* acc_u[kMaxStages]; T* acc_v[kMaxStages];
  T* k_h; T* k_u; T* k_v;
  float* br_x; float* br_y;
  Ctrl* ctrl;
  double* part_vol;         // [ntiles]
  double* part_rate;        // [ntiles], this stage's slice
};

template <typename T> struct Row { T h, u, v, z; };
template <typename T> struct Slopes { T hx, ux, vx, zx, hy, uy, vy, zy; };
template <typename T> struct FaceUV { T h, u, v, z; };

__device__ __forceinline__ double shfl_up1(double v) { return __shfl_up(v, 1, kWave); }
__device__ __forceinline__ float shfl_up1(float v) { return __shfl_up(v, 1, kWave); }
__device__ __forceinline__ double shfl_dn1(double v) { return __shfl_down(v, 1, kWave); }
__device__ __forceinline__ float shfl_dn1(float v) { return __shfl_down(v, 1, kWave); }

// Workgroups are dispatched round-robin over the 8 XCDs; remap so that runs
// of consecutive tiles (vertically adjacent strips sharing halo rows) land on
// one XCD and its L2.
__device__ __forceinline__ int64_t xcd_block(int64_t b, int64_t nb) {
  const int64_t nb8 = (nb / 8) * 8;
  if (b >= nb8) return b;
  return (b % 8) * (nb8 / 8) + b / 8;
}

template <typename T>
__device__ __forceinline__ double storm_rate(const double* s, double x, double y, double t) {
  const double cx = s[2] + s[4] * t, cy = s[3] + s[5] * t;
  const double ddx = x - cx, ddy = y - cy;
  return s[0] * exp(-(ddx * ddx + ddy * ddy) / (2.0 * s[1] * s[1]));
}
template <>
__device__ __forceinline__ double storm_rate<float>(const double* s, double x, double y, double t) {
  const float cx = float(s[2] + s[4] * t), cy = float(s[3] + s[5] * t);
  const float ddx = float(x) - cx, ddy = float(y) - cy;
  const float sg = float(s[1]);
  return double(float(s[0]) * __expf(-(ddx * ddx + ddy * ddy) / (2.0f * sg * sg)));
}

template <typename T>
__global__ __launch_bounds__(kWave * kWavesPerBlock) void k_stage(const StageArgs<T> A) {
  constexpr bool kRel = sizeof(T) == 4;
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t tile = xcd_block(blockIdx.x, gridDim.x) * kWavesPerBlock + (threadIdx.x >> 6);
  if (tile >= A.ntiles) return;
  Ctrl* __restrict__ C = A.ctrl;
  if (C->done) return;

  const int64_t nx = A.nx, ny = A.ny;
  const int64_t strip = tile / A.nrb, rb = tile % A.nrb;
  const int64_t x = strip * kStrip - kHalo + lane;
  const int64_t y0 = rb * A.rows;
  const int64_t y1 = (y0 + A.rows < ny) ? y0 + A.rows : ny;
  const bool xin = x >= 0 && x < nx;
  const bool own = lane >= kHalo && lane < kWave - kHalo && xin;
  const bool hasW = x > 0, hasE = x < nx - 1;

  const double dt = C->dt;
  const double tsub = C->t_local + A.c_s * dt;
  const int cur = C->cur;
  const T* __restrict__ xh = A.in_h ? A.in_h : A.Uh[cur];
  const T* __restrict__ xu = A.in_u ? A.in_u : A.Uu[cur];
  const T* __restrict__ xv = A.in_v ? A.in_v : A.Uv[cur];
  const T* __restrict__ zb = A.zb;
  const T dx = A.dx, dy = A.dy, rdx = A.rdx, rdy = A.rdy, hdx = A.hdx, hdy = A.hdy;
  const T G = T(kG);

  // RK coefficients of this stage: (T)(a * dt) as State::operator*(ValueType).
  T sc[kMaxStages];
#pragma unroll
  for (int j = 0; j < kMaxStages; ++j) sc[j] = T(A.acoef[j] * dt);

  // Uniform discharge / head terms (global selection, scalar endpoints).
  const double step_len = C->step_dur;
  double q0s = C->q0, q1s = C->q1, hb0s = C->hb0, hb1s = C->hb1;
  const double gstep = C->gstep;

  auto load = [&](int64_t r) {
    Row<T> R;
    if (xin && r >= 0 && r < ny) {
      const int64_t i = r * nx + x;
      R.h = xh[i]; R.u = xu[i]; R.v = xv[i]; R.z = zb[i];
    } else {
      R.h = T(0); R.u = T(0); R.v = T(0); R.z = T(0);
    }
    return R;
  };

  Row<T> R0 = load(y0 - 2), R1 = load(y0 - 1), R2;
  Slopes<T> S0 = {}, S1 = {};
  FaceUV<T> Fs0 = {}, Fs1 = {}, Xw0 = {}, Xw1 = {};
  double vol = 0.0, rate = 0.0;
  double cnmax = 0.0;

  for (int64_t c = y0 - 1; c <= y1; ++c) {
    R2 = load(c + 1);
    const bool rowin = c >= 0 && c < ny;
    // ---- slopes of row c (State.cpp:187-192, SpatialDerivative.hpp:104-122)
    Row<T> W, E;
    W.h = shfl_up1(R1.h); W.u = shfl_up1(R1.u); W.v = shfl_up1(R1.v); W.z = shfl_up1(R1.z);
    E.h = shfl_dn1(R1.h); E.u = shfl_dn1(R1.u); E.v = shfl_dn1(R1.v); E.z = shfl_dn1(R1.z);
    if (rowin) {
      const T wh = hasW ? W.h : R1.h, wu = hasW ? W.u : R1.u, wv = hasW ? W.v : R1.v, wz = hasW ? W.z : R1.z;
      const T eh = hasE ? E.h : R1.h, eu = hasE ? E.u : R1.u, ev = hasE ? E.v : R1.v, ez = hasE ? E.z : R1.z;
      S1.hx = minmod3(wh, hasW, R1.h, hasE, eh, dx, rdx);
      S1.ux = minmod3(wu, hasW, R1.u, hasE, eu, dx, rdx);
      S1.vx = minmod3(wv, hasW, R1.v, hasE, ev, dx, rdx);
      S1.zx = minmod3(wz, hasW, R1.z, hasE, ez, dx, rdx);
      const bool hasN = c + 1 < ny, hasS = c - 1 >= 0;
      const Row<T> N = hasN ? R2 : R1;
      const Row<T> S = hasS ? R0 : R1;
      if (A.flags & 0x40) {  // E7: south as the left neighbour
        S1.hy = minmod3(S.h, hasS, R1.h, hasN, N.h, dy, rdy);
        S1.uy = minmod3(S.u, hasS, R1.u, hasN, N.u, dy, rdy);
        S1.vy = minmod3(S.v, hasS, R1.v, hasN, N.v, dy, rdy);
        S1.zy = minmod3(S.z, hasS, R1.z, hasN, N.z, dy, rdy);
      } else {               // reference: north as the left neighbour
        S1.hy = minmod3(N.h, hasN, R1.h, hasS, S.h, dy, rdy);
        S1.uy = minmod3(N.u, hasN, R1.u, hasS, S.u, dy, rdy);
        S1.vy = minmod3(N.v, hasN, R1.v, hasS, S.v, dy, rdy);
        S1.zy = minmod3(N.z, hasN, R1.z, hasS, S.z, dy, rdy);
      }
    }
    // ---- west x-face of row c (face id c*(nx+1) + x)
    const T Whx = shfl_up1(S1.hx), Wux = shfl_up1(S1.ux), Wvx = shfl_up1(S1.vx), Wzx = shfl_up1(S1.zx);
    if (c >= y0 && c < y1) {
      Side<T> L, R;
      R = {R1.h, R1.u, R1.v, R1.z, S1.hx, S1.ux, S1.vx, S1.zx};
      int edge = 0;
      if (x == 0) {
        if (A.flags & 0x02) {  // E2: zero-gradient ghost that only drains
          L = {R1.h, fmin_t(R1.u, T(0)), R1.v, R1.z, T(0), T(0), T(0), T(0)};
        } else {
          L = R;
          edge = -1;
        }
      } else {
        L = {W.h, W.u, W.v, W.z, Whx, Wux, Wvx, Wzx};
        if (x == nx) { R = L; edge = 1; }
      }
      float br;
      const Flux<T> F = face_flux<T, kRel>(L, R, edge, hdx, dx, rdx, &br);
      Xw1 = {F.h, F.n, F.t, F.z};
      if (A.br_x && x >= 0 && x <= nx && lane >= kHalo && lane <= kWave - kHalo)
        A.br_x[c * (nx + 1) + x] = br;
    }
    // ---- south y-face of row c (face id (nx+1)ny + c*nx + x)
    if (c >= y0) {
      Side<T> L, R;
      int edge = 0;
      if (c == ny) {          // north edge of the top row
        L = {R0.h, R0.v, R0.u, R0.z, S0.hy, S0.vy, S0.uy, S0.zy};
        R = L;
        edge = 1;
      } else {
        R = {R1.h, R1.v, R1.u, R1.z, S1.hy, S1.vy, S1.uy, S1.zy};
        if (c == 0) { L = R; edge = -1; }
        else L = {R0.h, R0.v, R0.u, R0.z, S0.hy, S0.vy, S0.uy, S0.zy};
      }
      float br;
      const Flux<T> F = face_flux<T, kRel>(L, R, edge, hdy, dy, rdy, &br);
      Fs1 = {F.h, F.t, F.n, F.z};
      if (A.br_y && own) A.br_y[(nx + 1) * ny + c * nx + x - (nx + 1) * ny] = br;
    }
    // ---- cell pass + sources + RK for row r = c - 1
    const FaceUV<T> Xe0 = {shfl_dn1(Xw0.h), shfl_dn1(Xw0.u), shfl_dn1(Xw0.v), shfl_dn1(Xw0.z)};
    const int64_t r = c - 1;
    if (r >= y0 && r < y1 && own) {
      const int64_t i = r * nx + x;
      const T h = R0.h, u = R0.u, v = R0.v;
      // TemporalDerivativeKernel.cpp:57-90
      T dhdt = div_c(Xw0.h - Xe0.h, dx, rdx) + div_c(Fs0.h - Fs1.h, dy, rdy);
      T dudt = div_c(Xw0.u - Xe0.u, dx, rdx) + div_c(Fs0.u - Fs1.u, dy, rdy);
      T dvdt = div_c(Xw0.v - Xe0.v, dx, rdx) + div_c(Fs0.v - Fs1.v, dy, rdy);
      dudt += S0.hx * T(-kG);
      dvdt += S0.hy * T(-kG);
      T dzdx = S0.zx;
      const T hx_ = div_c(h, dx, rdx);
      if (fabs_t(dzdx) > hx_) dzdx = sy_sign(dzdx) * hx_;
      T dzdy = S0.zy;
      const T hy_ = div_c(h, dy, rdy);
      if (fabs_t(dzdy) > hy_) dzdy = sy_sign(dzdy) * hy_;
      const T dudt_bed = T(-kG) * dzdx;
      const T dvdt_bed = T(-kG) * dzdy;
      dudt += div_c(Xw0.z - Xe0.z, dx, rdx);
      dvdt += div_c(Fs0.z - Fs1.z, dy, rdy);
      dudt += dudt_bed;
      dvdt += dvdt_bed;
      // ManningRoughnessSourceTerm.cpp:55-90
      if (A.manning) {
        T nsh = A.n_sh, ndp = A.n_dp;
        if (A.nfield) { nsh = A.nfield[i]; ndp = nsh; }
        T n = nsh;
        if (nsh != ndp) {
          T t = (h - A.d_sh) / (A.d_dp - A.d_sh);
          t = fmin_t(fmax_t(t, T(0)), T(1));
          n = nsh + (ndp - nsh) * (t * t * (T(3) - T(2) * t));
        }
        if (h > T(1e-6)) {
          const T inv_h = h / (h * h + T(1e-3));
          const T Sf = n * n * sqrt_t(u * u + v * v) * pow43(inv_h);
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
      if (A.discharge || (A.flags & 0x08)) {
        double q0, q1;
        if (A.discharge == 2) { q0 = A.q0f[i]; q1 = A.q1f[i]; }
        else if (A.discharge == 1) { q0 = q0s; q1 = q1s; }
        else { q0 = 0.0; q1 = 0.0; }
        if (A.flags & 0x08) {
          const double cxp = A.ox + (double(x) + 0.5) * A.dxd;
          const double cyp = A.oy + (double(r) + 0.5) * A.dyd;
          q0 += storm_rate<T>(A.storm, cxp, cyp, gstep);
          q1 += storm_rate<T>(A.storm, cxp, cyp, gstep + step_len);
        }
        const T tq0 = T(q0), tq1 = T(q1);
        if (tq0 != T(0) || tq1 != T(0)) {
          const T dqdt = T((tq1 - tq0) / step_len);
          const T qnow = T(tq0 + tsub * dqdt);
          const T qnext = T(qnow + dt * dqdt);
          T dh = T(0.5) * (qnow + qnext);
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
      if (A.kout) {
        A.k_h[i] = dhdt; A.k_u[i] = dudt; A.k_v[i] = dvdt;
      } else {
        // RungeKutta.hpp:377-388: y_i = y_i + k_s * (a dt)
        T oh = h, ou = u, ov = v;
#pragma unroll
        for (int j = 0; j < kMaxStages; ++j) {
          if (j < A.nacc) {
            T* dh_ = A.acc_h[j] ? A.acc_h[j] : A.Uh[cur ^ 1];
            T* du_ = A.acc_u[j] ? A.acc_u[j] : A.Uu[cur ^ 1];
            T* dv_ = A.acc_v[j] ? A.acc_v[j] : A.Uv[cur ^ 1];
            T bh = h, bu = u, bv = v;
            if (A.stage > 0) { bh = dh_[i]; bu = du_[i]; bv = dv_[i]; }
            oh = bh + dhdt * sc[j];
            ou = bu + dudt * sc[j];
            ov = bv + dvdt * sc[j];
            if (j == 0) {  // row s+1 is complete after this stage
              if ((A.flags & 0x80) && oh <= A.h_dry) { ou = T(0); ov = T(0); }  // E8
              if (A.flags & 0x100) {                                           // E9
                ou = fmin_t(fmax_t(ou, -A.u_max), A.u_max);
                ov = fmin_t(fmax_t(ov, -A.u_max), A.u_max);
              }
            }
            dh_[i] = oh; du_[i] = ou; dv_[i] = ov;
          }
        }
        if (A.last) {
          // State.cpp:211-222 on the new state
          const T hh = fmax_t(oh, T(0));
          const T cc = sqrt_t(G * hh);
          const double cn = dt * double(div_c(fabs_t(ou) + cc, dx, rdx) + div_c(fabs_t(ov) + cc, dy, rdy));
          if (cn > cnmax) cnmax = cn;
          vol += double(oh);
        }
      }
    }
    R0 = R1; R1 = R2;
    S0 = S1;
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

struct CtrlArgs {
  Ctrl* ctrl;
  const double* part_vol;   // [ntiles]
  const double* part_rate;  // [nstages][ntiles]
  int64_t ntiles;
  int32_t nstages;
  int32_t mode;             // 0 = device controller, 1 = host-driven (record only)
  double b[kMaxStages];
  double area;
  const double* forcing;    // [nforcing][4] or null
  int64_t nforcing;
};

// Fixed-order block sum: thread t sums tiles t, t+256, ... then a tree.
__device__ __forceinline__ double block_sum(const double* p, int64_t n, double* sh) {
  double s = 0.0;
  for (int64_t k = threadIdx.x; k < n; k += blockDim.x) s += p[k];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

// Controller (RungeKutta.cpp:154-167) + accept (RungeKutta.hpp:428-431) +
// landing rules and outer-step transitions (TemporalScheme.hpp:112-149).
__global__ __launch_bounds__(256) void k_ctrl(const CtrlArgs A) {
  __shared__ double sh[256];
  Ctrl* C = A.ctrl;
  if (C->done) return;
  const double vol = block_sum(A.part_vol, A.ntiles, sh) * A.area;
  double rate = 0.0;
  for (int s = 0; s < A.nstages; ++s)
    rate += A.b[s] * block_sum(A.part_rate + (int64_t)s * A.ntiles, A.ntiles, sh);
  if (threadIdx.x != 0) return;
  const double cn_raw = __longlong_as_double((long long)C->cn_bits);
  C->cn_bits = 0ull;
  C->cn = cn_raw;
  if (A.mode == 1) { C->volume = vol; return; }
  const double cn = (C->cn_target == 1.0) ? cn_raw : cn_raw / C->cn_target;
  const double dt = C->dt;
  bool repeat;
  double ndt;
  if (cn > 1.0) {
    repeat = true;
    ndt = (cn > 5.0) ? dt / 5.0 : dt / (cn * 1.1);
  } else if (cn > 0.9) {
    repeat = false; ndt = dt;
  } else {
    repeat = false; ndt = dt / cn;
  }
  C->repeat = repeat ? 1 : 0;
  if (repeat) {
    C->n_rej++;
  } else {
    C->n_acc++;
    C->t_local += dt;
    C->cur ^= 1;
    C->volume = vol;
    C->srcvol += dt * rate * A.area;
  }
  C->dt = ndt;
  const double sd = C->step_dur, tl = C->t_local;
  if (tl >= sd) {
    C->step++;
    C->t_local = 0.0;
    // TemporalScheme.hpp:254: time_now = start_time + i * step_duration
    C->gstep = C->start_time + double(C->step) * sd;
    if (C->step >= C->nsteps) {
      C->done = 1;
    } else if (A.forcing && C->step < A.nforcing) {
      const double* f = A.forcing + 4 * C->step;
      C->q0 = f[0]; C->q1 = f[1]; C->hb0 = f[2]; C->hb1 = f[3];
    }
  } else if (tl + ndt > sd) {
    C->dt = sd - tl;
  } else if (tl + 2.0 * ndt > sd) {
    C->dt = (sd - tl) * 0.4;
  }
}

// Host-driven mode: set the attempt's t_local / dt (and the buffer flip of an
// accept) without a memcpy node, so the same launches can be graph-captured.
__global__ void k_set_ctrl(Ctrl* C, double t_local, double dt, int flip, int set_time) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (flip) C->cur ^= 1;
  if (set_time) { C->t_local = t_local; C->dt = dt; }
  C->cn_bits = 0ull;
}

}  // namespace mfsv
