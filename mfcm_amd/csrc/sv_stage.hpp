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
// x-neighbours move between lanes with DPP shifts; each face and each face
// difference is evaluated once per wave.  Every row load and store is a
// contiguous 64-element segment of a structure-of-arrays field, addressed by
// a uniform base and a 32-bit offset.  Sources are evaluated branch-free with
// selects.
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
  unsigned long long cn_bits;  // unused (the control number comes from tile partials)
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
  T* Uhl[2];                // FP32 form: low-order part of h of each state
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
  double* part_vol;         // [2 * ntiles]: volume, then control-number maxima
  double* part_rate;        // [ntiles], this stage's slice
  // further source terms (0 = off) and their per-cell fields
  int32_t eloss, infil, stageb, eddy;
  T mu_const, mu_smag, area;
  const T* fx; const T* fy;
  const T* irate; T* icap;
  const T* sb0; const T* sb1;
  const T* mu;
  // per-cell Manning parameters (ManningRoughnessSourceTerm.hpp:129-140):
  // with mfull, nfield is n_shallow and these give n_deep, d_shallow, d_deep
  int32_t mfull;
  const T* ndp_f; const T* dsh_f; const T* ddp_f;
  // output fields of a derivative evaluation (KOUT only; null = not wanted):
  // face fluxes h, u, v, z ((nx+1)*ny x-faces, then nx*(ny+1) y-faces), the
  // Manning n and friction slope, the bed slopes.  diag != 0 leaves the
  // infiltration capacity untouched.
  int32_t diag;
  // the stage input is a convex-form y1 stored before E8/E9: apply them to
  // the velocities as the rows are loaded (the completed state they stand for)
  int32_t adj_in;
  T* dg_fh; T* dg_fu; T* dg_fv; T* dg_fz;
  T* dg_nh; T* dg_sf; T* dg_zx; T* dg_zy;
};

template <typename T> struct Row { T h, u, v, z; };
template <typename T> struct FaceUV { T h, u, v, z; };

// ---- cross-lane moves: DPP row shifts across the whole wave (wave_shr:1 /
// wave_shl:1).  from_w(v) in lane l is v of lane l-1, from_e(v) is v of lane
// l+1; the lane without a source reads 0 (bound_ctrl).  Full-rate VALU moves,
// no LDS round trip.
__device__ __forceinline__ float from_w(float v) {
  return __builtin_amdgcn_update_dpp(0.0f, v, 0x138, 0xf, 0xf, true);
}
__device__ __forceinline__ double from_w(double v) {
  return __builtin_amdgcn_update_dpp(0.0, v, 0x138, 0xf, 0xf, true);
}
__device__ __forceinline__ float from_e(float v) {
  return __builtin_amdgcn_update_dpp(0.0f, v, 0x130, 0xf, 0xf, true);
}
__device__ __forceinline__ double from_e(double v) {
  return __builtin_amdgcn_update_dpp(0.0, v, 0x130, 0xf, 0xf, true);
}
template <typename T>
__device__ __forceinline__ Row<T> from_w(const Row<T>& r) {
  return {from_w(r.h), from_w(r.u), from_w(r.v), from_w(r.z)};
}
template <typename T>
__device__ __forceinline__ Row<T> from_e(const Row<T>& r) {
  return {from_e(r.h), from_e(r.u), from_e(r.v), from_e(r.z)};
}
template <typename T>
__device__ __forceinline__ FaceUV<T> from_e(const FaceUV<T>& r) {
  return {from_e(r.h), from_e(r.u), from_e(r.v), from_e(r.z)};
}
// Lane k's value, broadcast (k uniform).
__device__ __forceinline__ float lane_value(float v, int k) {
  return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), k));
}
__device__ __forceinline__ double lane_value(double v, int k) {
  const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, k);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), k);
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// ---- field access: a uniform base pointer plus a 32-bit byte offset per
// lane (the saddr form of global loads: one shared offset register for all
// fields of a row).  Lanes and rows outside the grid read a clamped cell,
// whose value never reaches an owned cell; stores are masked to owned cells.
template <typename T>
__device__ __forceinline__ T ld(const T* __restrict__ p, uint32_t off) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(p) + off);
}
template <typename T>
__device__ __forceinline__ void st(T* __restrict__ p, uint32_t off, T v) {
  *reinterpret_cast<T*>(reinterpret_cast<char*>(p) + off) = v;
}

template <typename T>
__device__ __forceinline__ T sel(bool c, T a, T b) { return c ? a : b; }
template <typename T>
__device__ __forceinline__ Side<T> sel(bool c, const Side<T>& a, const Side<T>& b) {
  return {c ? a.h : b.h, c ? a.un : b.un, c ? a.ut : b.ut, c ? a.z : b.z,
          c ? a.sh : b.sh, c ? a.sun : b.sun, c ? a.sut : b.sut, c ? a.sz : b.sz};
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
    // accurate expf: once per column / row of a tile, and its bias would
    // show in the FP64 rain volume
    const float d = float(a - c), sg = float(s.sigma);
    return expf(-(d * d) / (2.0f * sg * sg));
  }
}

// Division by a per-cell value: correctly rounded in FP64 (the reference's
// a / b), reciprocal-based in the FP32 form.
__device__ __forceinline__ double vdiv(double a, double b) { return a / b; }
__device__ __forceinline__ float vdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }

// Division by the (uniform) timestep: a / dt correctly rounded in FP64 from
// rdt = RN(1/dt) (div_c), reciprocal multiply in the FP32 form.
__device__ __forceinline__ double div_dt(double a, double dt, double rdt) { return div_c(a, dt, rdt); }
__device__ __forceinline__ float div_dt(float a, float, float rdt) { return a * rdt; }

// E2 ghost of the x = 0 edge face: zero-gradient depth and tangential
// velocity, the normal velocity only where it drains out of the domain.
template <typename T>
__device__ __forceinline__ Side<T> outflow_ghost(const Side<T>& c) {
  Side<T> g;
  g.h = c.h; g.un = fmin_t(c.un, T(0)); g.ut = c.ut; g.z = c.z;
  g.sh = T(0); g.sun = T(0); g.sut = T(0); g.sz = T(0);
  return g;
}

// NACC: accumulator rows this stage writes (0 with KOUT: write k instead).
// QUICK: the difference form, for power-of-two spacing in FP64 and for any
// spacing in the FP32 form without coded-out cells.  Slopes are then carried
// as per-cell differences m = d * slope: for d = 2^k the Minmod3 of the
// quotients is exactly the Minmod3 of the differences scaled by 1/d, the face
// projection (d/2) * slope is exactly m/2, and -g * slope is exactly
// m * (-g/d), so every value is the reference's bit for bit (other spacings
// in FP32 multiply by the once-rounded reciprocal).  Otherwise (and with
// coded-out cells, NANBED) slopes are exact quotients.
// EXT: the further source terms (energy loss, infiltration, eddy viscosity)
// and the stage boundary are compiled in (each still switched at run time).
// BASE: a later stage (s > 0), whose accumulator rows are read back as the
// bases of this stage's updates; the first stage updates from its input.
// BND: per-cell discharge / head endpoint fields (discharge or head mode 2)
// may be set; without it the kernel neither loads nor holds their pointers
// (the FP32 form loads them with the row, FP64 where they are used).
// CONVEX (FP32 form, Heun-2 only, NACC == 1): the stages in the convex form
// y1 = y0 + k0 dt, y2 = (y0 + y1)/2 + k1 dt/2 (the same method; FP32 is not
// held to the reference's summation order): the first stage writes y1 (its
// depth rounded once from y0's high part, the low term of (y0 + y1)/2 that
// this leaves out into the new state's low-order buffer), the second reads
// y0, y1 and that low word and writes y2 with a compensated depth, so no
// partial row y0 + k0 dt/2 is stored and read back and the FP64 volume of
// hi + lo still tracks the accumulated rates.
template <typename T, int NACC, bool KOUT, bool QUICK, bool NANBED, bool EXT, bool BASE, bool BND,
          bool CONVEX = false>
__device__ __forceinline__ void stage_tile(const StageArgs<T>& A, const int tile, const int lane) {
  static_assert(!CONVEX || (sizeof(T) == 4 && NACC == 1 && !KOUT), "convex form: FP32 Heun-2");
  static_assert(!(QUICK && NANBED), "coded-out cells take the quotient path");
  constexpr bool kRel = sizeof(T) == 4;
  constexpr uint32_t S = sizeof(T);
  constexpr int NB = NACC > 0 ? NACC : 1;
  Ctrl* __restrict__ C = A.ctrl;
  // a finished device-controlled run skips its remaining captured stages (a
  // derivative evaluation, KOUT, is always an explicit request): the tile
  // then marches no row and stores nothing.  No branch: the control block's
  // fields are loaded together, not one round trip before a branch and
  // another after it.
  const bool finished = !KOUT && C->done != 0;

  const int nx = A.nx, ny = A.ny;
  // tile order.  FP64: row block by row block (horizontal neighbours come
  // next; with the blocks of an XCD consecutive, the strips that share a halo
  // column read it through one L2 at the same time).  FP32: groups of 4
  // adjacent strips marched down together, row block by row block within a
  // group: the 4 waves of a workgroup (one CU) are horizontal neighbours at
  // the same rows and share the cache lines their strips overlap in, and
  // consecutive workgroups (one XCD) follow down the same columns.  On the
  // 2048^2 storm catchment, per FP32 stage launch: strip by strip 47.5 us
  // (24-row tiles) / 48.2 us (18), row block by row block 49.5 / -, groups of
  // 4 strips 49.3 / 46.6.
  int strip, rb;
  {
    const int nstrips = A.ntiles / A.nrb;
    if constexpr (sizeof(T) == 8) {
      rb = tile / nstrips;
      strip = tile - rb * nstrips;
    } else {
      constexpr int kGroup = kWavesPerBlock;
      const int gsz = kGroup * A.nrb;
      const int g = tile / gsz, rem = tile - g * gsz;
      const int gw = (g + 1) * kGroup <= nstrips ? kGroup : nstrips - g * kGroup;  // last group
      rb = rem / gw;
      strip = g * kGroup + (rem - rb * gw);
    }
  }
  const int x = strip * kStrip - kHalo + lane;
  const int y0 = rb * A.rows;
  const int y1 = finished ? y0 - 3 : ((y0 + A.rows < ny) ? y0 + A.rows : ny);
  const bool xin = x >= 0 && x < nx;
  const bool own = lane >= kHalo && lane < kWave - kHalo && xin;
  const bool hasW = x > 0, hasE = x < nx - 1;
  // every lane of the strip has both x neighbours
  const bool xint = strip * kStrip - kHalo > 0 && strip * kStrip - kHalo + kWave - 1 < nx - 1;
  const bool e7 = (A.flags & 0x40) != 0;
  const bool e2 = (A.flags & 0x02) != 0;
  // does this strip hold the x = 0 or the x = nx face?
  const bool edge_strip = strip == 0 || (strip * kStrip - kHalo <= nx && nx < strip * kStrip - kHalo + kWave);

  const double dt = C->dt;
  const double tsub = C->t_local + A.c_s * dt;
  const int cur = C->cur;
  // select, not index: a run-time index into the argument block would copy
  // it to private memory
  T* const Uh1 = cur ? A.Uh[0] : A.Uh[1];
  T* const Uu1 = cur ? A.Uu[0] : A.Uu[1];
  T* const Uv1 = cur ? A.Uv[0] : A.Uv[1];
  const T* const xh = A.in_h ? A.in_h : (cur ? A.Uh[1] : A.Uh[0]);
  const T* const xu = A.in_u ? A.in_u : (cur ? A.Uu[1] : A.Uu[0]);
  const T* const xv = A.in_v ? A.in_v : (cur ? A.Uv[1] : A.Uv[0]);
  const T* __restrict__ zb = A.zb;
  const bool has_n = A.nfield != nullptr;
  // FP32: without a roughness field the row's n load reads the bed instead
  // (a valid address; the value is not used): no branch around the load, so
  // its wait does not also wait for the row prefetch
  const T* __restrict__ nf = has_n ? A.nfield : A.zb;
  const T dx = A.dx, dy = A.dy, rdx = A.rdx, rdy = A.rdy;
  const T hdx = QUICK ? T(0.5) : A.hdx, hdy = QUICK ? T(0.5) : A.hdy;
  const T gx = QUICK ? T(-kG) * rdx : T(-kG), gy = QUICK ? T(-kG) * rdy : T(-kG);
  // FP32 difference form: the y slopes stay south -> north (the reference's
  // run north -> south unless E7, their exact negation) and the sign rides on
  // the y half spacing and the y pressure and bed factors instead (exact: a
  // product changes sign with either factor)
  constexpr bool kFoldY = sizeof(T) == 4 && QUICK && !NANBED && !EXT && !KOUT;
  // ... and, under the same conditions, the slopes are carried halved and
  // the face fluxes doubled (exact power-of-two scalings: the results are
  // those of the unscaled form): the half spacing of the reconstructions
  // becomes 1, the pressure and bed factors double, the flux divergence
  // takes half the reciprocal spacing; this drops a doubling per slope and
  // the halvings of the Rusanov average per face
  constexpr bool kScaled = kFoldY;
  const T hdx_s = kScaled ? T(1) : hdx;
  const T hdy_s = (kFoldY && !e7) ? -(kScaled ? T(1) : hdy) : (kScaled ? T(1) : hdy);
  const T gx_s = kScaled ? T(2) * gx : gx;
  const T gy_s = ((kFoldY && !e7) ? -gy : gy) * (kScaled ? T(2) : T(1));
  const T rdx_f = kScaled ? T(2) * rdx : rdx, rdy_f = kScaled ? T(2) * rdy : rdy;
  const T rdx_c = kScaled ? T(0.5) * rdx : rdx, rdy_c = kScaled ? T(0.5) * rdy : rdy;
  auto mm_c = [&](T ql, T qr, T qc) -> T {
    if constexpr (kScaled) return minmod_ch(ql, qr, qc);
    else return minmod_c(ql, qr, qc);
  };
  // the edge cases of a slope whose two-sided value mm_c gave
  auto mm_e = [&](T in, T ql, T qr, bool hl, bool hr) -> T {
    return minmod_edge<kScaled>(in, ql, qr, hl, hr);
  };
  const T G = T(kG);

  // RK coefficients of this stage: (T)(a * dt) as State::operator*(ValueType).
  T sc[NB];
  T* oh_[NB];
  T* ou_[NB];
  T* ov_[NB];
#pragma unroll
  for (int j = 0; j < NB; ++j) {
    sc[j] = T(0);
    oh_[j] = Uh1; ou_[j] = Uu1; ov_[j] = Uv1;
  }
#pragma unroll
  for (int j = 0; j < NACC; ++j) {
    sc[j] = T(A.acoef[j] * dt);
    oh_[j] = A.acc_h[j] ? A.acc_h[j] : Uh1;
    ou_[j] = A.acc_u[j] ? A.acc_u[j] : Uu1;
    ov_[j] = A.acc_v[j] ? A.acc_v[j] : Uv1;
  }

  // byte offsets of row r in this lane's column (clamped to the grid)
  const int xc_ = x < 0 ? 0 : (x < nx ? x : nx - 1);
  const uint32_t lane_off = uint32_t(xc_) * S;
  const uint32_t row_bytes = uint32_t(nx) * S;
  auto roff = [&](int r) -> uint32_t {
    const int rc = r < 0 ? 0 : (r < ny ? r : ny - 1);
    return lane_off + uint32_t(rc) * row_bytes;
  };

  // Uniform discharge / head terms (global selection, scalar endpoints).
  const double step_len = C->step_dur;
  const double q0s = C->q0, q1s = C->q1, hb0s = C->hb0, hb1s = C->hb1;
  const double gstep = C->gstep;
  const bool storm = (A.flags & 0x08) != 0;
  const T dtT = T(dt);
  const T rdt = sizeof(T) == 8 ? T(1.0 / dt) : T(1.0f / float(dt));

  // E8 / E9 on a completed state's velocities (no velocity where h <= h_dry,
  // components within +-u_max)
  auto complete = [&](T hh, T& uu, T& vv) {
    if (A.flags & 0x80) {  // E8
      const bool dryc = hh <= A.h_dry;
      uu = dryc ? T(0) : uu;
      vv = dryc ? T(0) : vv;
    }
    if (A.flags & 0x100) {  // E9
      uu = clamp_t(uu, -A.u_max, A.u_max);
      vv = clamp_t(vv, -A.u_max, A.u_max);
    }
  };
  // rawu, rawv: the row's velocities as stored, before the completion
  // (the convex second stage combines y1's stored velocities; the compiler
  // drops them elsewhere)
  auto load = [&](int r, T& rawu, T& rawv) {
    const uint32_t o = roff(r);
    Row<T> w = {ld(xh, o), ld(xu, o), ld(xv, o), ld(zb, o)};
    rawu = w.u;
    rawv = w.v;
    // always in the convex second stage; a derivative evaluation of that y1
    // says so at run time
    if constexpr (CONVEX && BASE) complete(w.h, w.u, w.v);
    else if constexpr (sizeof(T) == 4 && KOUT) { if (A.adj_in) complete(w.h, w.u, w.v); }
    return w;
  };
  // a difference along x / y as the slope quantity (see QUICK)
  auto qx = [&](T d) -> T {
    if constexpr (QUICK) return d;
    else return div_c(d, dx, rdx);
  };
  auto qy = [&](T d) -> T {
    if constexpr (QUICK) return d;
    else return div_c(d, dy, rdy);
  };

  T r0u, r0v, r1u, r1v, r2u, r2v;  // stored velocities of rows R0, R1, R2
  Row<T> R0 = load(y0 - 2, r0u, r0v), R1 = load(y0 - 1, r1u, r1v), R2 = load(y0, r2u, r2v);
  // Per-cell inputs of the cell pass (Manning n, accumulator bases of a
  // later stage).
  struct Extra {
    T n;
    T bh[NB], bu[NB], bv[NB];
    T lo;  // FP32 form: low-order depth of the final row's base
    // FP32 form: per-cell boundary endpoints (discharge, head)
    T q0, q1, hb0, hb1;
    // FP32 form with the further source terms: their per-cell fields
    T ndp, dsh, ddp, fx, fy, irate, icap, mu, sb0, sb1;
  };
  // low-order depth read by this stage: U0's (first stage) or what the
  // previous stage left in the new state's low-order buffer (the final row's
  // partial; in the convex form, the low term of (y0 + y1)/2)
  T* const hl_in = (sizeof(T) == 4 && !KOUT)
                       ? ((A.stage == 0) ? (cur ? A.Uhl[1] : A.Uhl[0])
                                         : (cur ? A.Uhl[0] : A.Uhl[1]))
                       : nullptr;
  constexpr bool kLoIn = sizeof(T) == 4 && !KOUT;
  // convex second stage: the bases are U0 (the new state is written to the
  // other buffer)
  const T* const u0h = cur ? A.Uh[1] : A.Uh[0];
  const T* const u0u = cur ? A.Uu[1] : A.Uu[0];
  const T* const u0v = cur ? A.Uv[1] : A.Uv[0];
  // FP32: the per-cell boundary endpoints are loaded with the other cell-pass
  // inputs, ahead of the row prefetch and without a branch, so that no wait
  // in the cell pass also waits for the prefetch.  A field that is off reads
  // the words the row's low-order depth load has just fetched (the bed where
  // there is none); its value is not used.  FP64 loads them where they are
  // used.
  constexpr bool kPreBnd = sizeof(T) == 4 && BND;
  constexpr bool kBnd = BND;  // per-cell endpoint fields possible
  const T* const fb = kLoIn ? hl_in : zb;
  const T* const q0p = A.discharge == 2 ? A.q0f : fb;
  const T* const q1p = A.discharge == 2 ? A.q1f : fb;
  const T* const hb0p = A.head == 2 ? A.hb0f : fb;
  const T* const hb1p = A.head == 2 ? A.hb1f : fb;
  // the further source terms' fields, likewise (FP32 outside derivative
  // evaluations)
  constexpr bool kPreExt = sizeof(T) == 4 && EXT && !KOUT;
  const T* ndpp = fb; const T* dshp = fb; const T* ddpp = fb; const T* fxp = fb; const T* fyp = fb;
  const T* iratep = fb; const T* icapp = fb; const T* mup = fb; const T* sb0p = fb; const T* sb1p = fb;
  if constexpr (kPreExt) {
    if (A.mfull) { ndpp = A.ndp_f; dshp = A.dsh_f; ddpp = A.ddp_f; }
    if (A.eloss) { fxp = A.fx; fyp = A.fy; }
    if (A.infil) { iratep = A.irate; icapp = A.icap; }
    if (A.eddy == 1) mup = A.mu;
    if (A.stageb) { sb0p = A.sb0; sb1p = A.sb1; }
  }
  constexpr bool base_ld = BASE;
  auto load_extra = [&](int r) {
    Extra e;
    const uint32_t o = roff(r);
    if constexpr (sizeof(T) == 4) e.n = ld(nf, o);
    else e.n = has_n ? ld(nf, o) : A.n_sh;
    e.lo = T(0);
    if constexpr (kLoIn) e.lo = ld(hl_in, o);
    if constexpr (kPreBnd) {
      e.q0 = ld(q0p, o); e.q1 = ld(q1p, o); e.hb0 = ld(hb0p, o); e.hb1 = ld(hb1p, o);
    } else {
      e.q0 = e.q1 = e.hb0 = e.hb1 = T(0);
    }
    if constexpr (kPreExt) {
      e.ndp = ld(ndpp, o); e.dsh = ld(dshp, o); e.ddp = ld(ddpp, o);
      e.fx = ld(fxp, o); e.fy = ld(fyp, o); e.irate = ld(iratep, o); e.icap = ld(icapp, o);
      e.mu = ld(mup, o); e.sb0 = ld(sb0p, o); e.sb1 = ld(sb1p, o);
    } else {
      e.ndp = e.dsh = e.ddp = e.fx = e.fy = e.irate = e.icap = e.mu = e.sb0 = e.sb1 = T(0);
    }
#pragma unroll
    for (int j = 0; j < NB; ++j) { e.bh[j] = T(0); e.bu[j] = T(0); e.bv[j] = T(0); }
    if constexpr (base_ld) {
#pragma unroll
      for (int j = 0; j < NACC; ++j) {
        if constexpr (CONVEX) {
          e.bh[j] = ld(u0h, o); e.bu[j] = ld(u0u, o); e.bv[j] = ld(u0v, o);
        } else {
          e.bh[j] = ld(oh_[j], o); e.bu[j] = ld(ou_[j], o); e.bv[j] = ld(ov_[j], o);
        }
      }
    }
    return e;
  };
  Extra Xn;
  // the cell-pass inputs of the rows above the tile are never used: those
  // iterations load row y0's (needed a row or two later, so the loads are not
  // wasted on HBM); likewise the last iteration's prefetch reloads row y1 + 1
  if constexpr (sizeof(T) == 8) Xn = load_extra(y0);
  // (the storm factors are computed while the first rows' loads are in
  // flight: the loads above come first in the program)
  // E4 storm factors: this lane's column at the step's two endpoint times
  // (registers), and the factors of rows y0 .. y0+63 (computed by lane k for
  // row y0 + k, kept in this wave's LDS slot and read back row by row as a
  // broadcast, so they hold no registers across the march).
  // The FP32 form has registers to spare and keeps the row factors in lane
  // registers, read with v_readlane.
  constexpr bool kStormLds = sizeof(T) == 8;
  __shared__ T storm_rows[kWavesPerBlock][2][kWave];
  const int wv = (int)(threadIdx.x >> 6);
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
    if constexpr (kStormLds) {
      storm_rows[wv][0][lane] = eyA0;
      storm_rows[wv][1][lane] = eyA1;
      // the wave's own writes, visible to all its lanes before the reads
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
  }
  // south slope quantities of the current row (rows y0-2 and y0-1 first)
  T qsh = qy(R1.h - R0.h), qsu = qy(R1.u - R0.u), qsv = qy(R1.v - R0.v), qsz = qy(R1.z - R0.z);
  // slopes of row c-1 that the cell pass and the y-face still need
  T p_hx = T(0), p_zx = T(0), p_hy = T(0), p_uy = T(0), p_vy = T(0), p_zy = T(0);
  // EXT: x slopes of u, v of row c-1 and y slopes of u, v of row c-2 (the
  // second derivatives of the eddy-viscosity term)
  T p_ux = T(0), p_vx = T(0), pp_uy = T(0), pp_vy = T(0);
  FaceUV<T> Fs0 = {}, Xw0 = {};
  double vol = 0.0, rate = 0.0;
  T qmax = T(0);

  // One row of the march: slopes of row c, the west x-faces of row c, the
  // south y-faces of row c and the cell pass of row c - 1.
  auto row_step = [&](const int c) __attribute__((always_inline)) {
    // inputs of row c-1's cell pass: FP32 loads them here (consumed after the
    // slopes and faces; fewer live registers) and before the row prefetch,
    // so that waiting for them leaves the prefetch in flight; FP64 one row
    // ahead (its occupancy is lower, so memory latency needs the longer lead)
    Extra X0;
    if constexpr (sizeof(T) == 4) X0 = load_extra(c - 1 < y0 ? y0 : c - 1);
    T rnu, rnv;
    const Row<T> Rn = load(c + 2 > y1 + 1 ? y1 + 1 : c + 2, rnu, rnv);  // prefetch
    if constexpr (sizeof(T) == 8) {
      X0 = Xn;
      Xn = load_extra(c < y0 ? y0 : c);
    }
    const Row<T> W = from_w(R1), E = from_e(R1);
    const bool hasN = c + 1 < ny, hasS = c > 0;
    // ---- slopes of row c (State.cpp:187-192, SpatialDerivative.hpp:104-122)
    T shx, sux, svx, szx, shy, suy, svy, szy;
    T qnh = T(0), qnu = T(0), qnv = T(0), qnz = T(0);
    if constexpr (!NANBED) {
      const T qwh = qx(R1.h - W.h), qwu = qx(R1.u - W.u), qwv = qx(R1.v - W.v), qwz = qx(R1.z - W.z);
      // east differences: the west difference one lane east (a single DPP
      // move in FP32); the FP64 difference form subtracts the east values it
      // already holds instead (the same operands: bit-identical, and no
      // NaN-quieting pass over the shifted words before the minimum)
      T qeh, qeu, qev, qez;
      if constexpr (sizeof(T) == 8 && QUICK) {
        qeh = E.h - R1.h; qeu = E.u - R1.u; qev = E.v - R1.v; qez = E.z - R1.z;
      } else {
        qeh = from_e(qwh); qeu = from_e(qwu); qev = from_e(qwv); qez = from_e(qwz);
      }
      // central differences, once for both forms below (the east shift then
      // feeds a single subtraction)
      const T qch = qx(E.h - W.h), qcu = qx(E.u - W.u), qcv = qx(E.v - W.v), qcz = qx(E.z - W.z);
      // two-sided everywhere, then the lanes on an x edge where the strip
      // has one (a wave-uniform branch that overwrites the values in place,
      // so that the common path carries no copies)
      shx = mm_c(qwh, qeh, qch);
      sux = mm_c(qwu, qeu, qcu);
      svx = mm_c(qwv, qev, qcv);
      szx = mm_c(qwz, qez, qcz);
      if (!xint) {
        shx = mm_e(shx, qwh, qeh, hasW, hasE);
        sux = mm_e(sux, qwu, qeu, hasW, hasE);
        svx = mm_e(svx, qwv, qev, hasW, hasE);
        szx = mm_e(szx, qwz, qez, hasW, hasE);
      }
      qnh = qy(R2.h - R1.h); qnu = qy(R2.u - R1.u); qnv = qy(R2.v - R1.v); qnz = qy(R2.z - R1.z);
      // south -> north slope; the reference's north -> south slope is its
      // exact negation (Minmod3 commutes with negation)
      // south differences: the previous row's north differences (carried;
      // recomputed, bit-identical, in the FP64 difference form, which keeps
      // eight fewer registers live across the row)
      if constexpr (QUICK && sizeof(T) == 8) {
        qsh = R1.h - R0.h; qsu = R1.u - R0.u; qsv = R1.v - R0.v; qsz = R1.z - R0.z;
      }
      T yh = mm_c(qsh, qnh, qy(R2.h - R0.h));
      T yu = mm_c(qsu, qnu, qy(R2.u - R0.u));
      T yv = mm_c(qsv, qnv, qy(R2.v - R0.v));
      T yz = mm_c(qsz, qnz, qy(R2.z - R0.z));
      if (!(hasS && hasN)) {  // wave-uniform: the grid's first or last row
        yh = mm_e(yh, qsh, qnh, hasS, hasN);
        yu = mm_e(yu, qsu, qnu, hasS, hasN);
        yv = mm_e(yv, qsv, qnv, hasS, hasN);
        yz = mm_e(yz, qsz, qnz, hasS, hasN);
      }
      if constexpr (kFoldY) {
        shy = yh; suy = yu; svy = yv; szy = yz;
      } else {
        shy = e7 ? yh : -yh; suy = e7 ? yu : -yu; svy = e7 ? yv : -yv; szy = e7 ? yz : -yz;
      }
    } else {
      // general Minmod3 with the reference's NaN rules
      const Row<T> Wc = hasW ? W : R1, Ec = hasE ? E : R1;
      shx = minmod3(Wc.h, hasW, R1.h, hasE, Ec.h, dx, rdx);
      sux = minmod3(Wc.u, hasW, R1.u, hasE, Ec.u, dx, rdx);
      svx = minmod3(Wc.v, hasW, R1.v, hasE, Ec.v, dx, rdx);
      szx = minmod3(Wc.z, hasW, R1.z, hasE, Ec.z, dx, rdx);
      const Row<T> N = hasN ? R2 : R1, Sr = hasS ? R0 : R1;
      if (e7) {
        shy = minmod3(Sr.h, hasS, R1.h, hasN, N.h, dy, rdy);
        suy = minmod3(Sr.u, hasS, R1.u, hasN, N.u, dy, rdy);
        svy = minmod3(Sr.v, hasS, R1.v, hasN, N.v, dy, rdy);
        szy = minmod3(Sr.z, hasS, R1.z, hasN, N.z, dy, rdy);
      } else {
        shy = minmod3(N.h, hasN, R1.h, hasS, Sr.h, dy, rdy);
        suy = minmod3(N.u, hasN, R1.u, hasS, Sr.u, dy, rdy);
        svy = minmod3(N.v, hasN, R1.v, hasS, Sr.v, dy, rdy);
        szy = minmod3(N.z, hasN, R1.z, hasS, Sr.z, dy, rdy);
      }
    }
    // ---- west x-face of row c (face id c*(nx+1) + x).  In the rows outside
    // the tile's faces it is left undefined (no instruction): the cell pass
    // reads only faces of rows y0 .. y1-1.
    FaceUV<T> Xw1;
    if (c >= y0 && c < y1) {
      const Side<T> R = {R1.h, R1.u, R1.v, R1.z, shx, sux, svx, szx};
      const Side<T> L = {W.h, W.u, W.v, W.z, from_w(shx), from_w(sux), from_w(svx), from_w(szx)};
      float br;
      Flux<T> F;
      if constexpr (!NANBED && sizeof(T) == 4) {
        F = face_fast<T, kRel, QUICK, kScaled>(L, R, hdx_s, dx, rdx_f, &br);
        // wave-uniform: the strip holds the x = 0 or the x = nx face, whose
        // lanes take a ghost side; the face is evaluated again (the same
        // values elsewhere) and overwrites the first, so that the common
        // path carries no copies of the sides
        if (edge_strip) {
          const Side<T> Le = sel(x == 0, e2 ? outflow_ghost(R) : wall_ghost(R), L);
          const Side<T> Re = sel(x == nx, wall_ghost(L), R);
          F = face_fast<T, kRel, QUICK, kScaled>(Le, Re, hdx_s, dx, rdx_f, &br);
        }
        Xw1 = {F.h, F.n, F.t, F.z};
      } else if constexpr (!NANBED) {
        // FP64: the ghost sides selected before a single evaluation (a
        // second evaluation costs this kernel registers it does not have)
        Side<T> Lf = L, Rf = R;
        if (edge_strip) {
          Lf = sel(x == 0, e2 ? outflow_ghost(R) : wall_ghost(R), L);
          Rf = sel(x == nx, wall_ghost(L), R);
        }
        F = face_fast<T, kRel, QUICK, kScaled>(Lf, Rf, hdx_s, dx, rdx_f, &br);
        Xw1 = {F.h, F.n, F.t, F.z};
      } else {
        // FluxKernel.cpp:66-295 term for term (coded-out cells)
        const CellV<T> Rc = {R1.h, R1.u, R1.v, R1.z, szx, szy, shx, sux, svx, shy, suy, svy};
        CellV<T> Lc = {W.h, W.u, W.v, W.z, from_w(szx), from_w(szy), from_w(shx), from_w(sux),
                       from_w(svx), from_w(shy), from_w(suy), from_w(svy)};
        FluxXY<T> G2;
        if (x == 0 && e2) {
          // E2 ghost: the cell's depth, bed and tangential velocity, the
          // normal velocity only outwards, no slopes
          const CellV<T> Gc = {R1.h, fmin_t(R1.u, T(0)), R1.v, R1.z, T(0), T(0), T(0), T(0),
                               T(0), T(0), T(0), T(0)};
          G2 = face_general<T, kRel>(Gc, Rc, 0, 0, dx, rdx, &br);
        } else {
          CellV<T> Rf = Rc;
          int edge = 0;
          if (x == 0) { Lc = Rc; edge = -1; }
          else if (x == nx) { Rf = Lc; edge = 1; }
          G2 = face_general<T, kRel>(Lc, Rf, edge, 0, dx, rdx, &br);
        }
        Xw1 = {G2.h, G2.u, G2.v, G2.z};
      }
      if constexpr (KOUT) {
        if (x >= 0 && x <= nx && lane >= kHalo && lane <= kWave - kHalo) {
          const int64_t fi = int64_t(c) * (nx + 1) + x;
          if (A.br_x) A.br_x[fi] = br;
          if (A.dg_fh) {
            A.dg_fh[fi] = Xw1.h; A.dg_fu[fi] = Xw1.u; A.dg_fv[fi] = Xw1.v; A.dg_fz[fi] = Xw1.z;
          }
        }
      }
    }
    // ---- south y-face of row c (face id (nx+1)ny + c*nx + x); undefined in
    // the rows outside y0 .. y1, which no cell pass reads
    FaceUV<T> Fs1;
    if (c >= y0 && c <= y1) {
      float br;
      Flux<T> F;
      if constexpr (!NANBED) {
        if (c == ny) {  // north edge of the top row
          const Side<T> L = {R0.h, R0.v, R0.u, R0.z, p_hy, p_vy, p_uy, p_zy};
          F = face_fast<T, kRel, QUICK, kScaled>(L, wall_ghost(L), hdy_s, dy, rdy_f, &br);
        } else {
          const Side<T> R = {R1.h, R1.v, R1.u, R1.z, shy, svy, suy, szy};
          if (c == 0) {
            F = face_fast<T, kRel, QUICK, kScaled>(wall_ghost(R), R, hdy_s, dy, rdy_f, &br);
          } else {
            const Side<T> L = {R0.h, R0.v, R0.u, R0.z, p_hy, p_vy, p_uy, p_zy};
            F = face_fast<T, kRel, QUICK, kScaled>(L, R, hdy_s, dy, rdy_f, &br);
          }
        }
        Fs1 = {F.h, F.t, F.n, F.z};
      } else {
        // FluxKernel.cpp:66-295 term for term (coded-out cells); the cell
        // below is row c-1 with its slopes of the previous iteration
        const CellV<T> Lc = {R0.h, R0.u, R0.v, R0.z, p_zx, p_zy, p_hx, p_ux, p_vx, p_hy, p_uy, p_vy};
        FluxXY<T> G2;
        if (c == ny) {
          G2 = face_general<T, kRel>(Lc, Lc, 1, 1, dy, rdy, &br);
        } else {
          const CellV<T> Rc = {R1.h, R1.u, R1.v, R1.z, szx, szy, shx, sux, svx, shy, suy, svy};
          if (c == 0) G2 = face_general<T, kRel>(Rc, Rc, -1, 1, dy, rdy, &br);
          else G2 = face_general<T, kRel>(Lc, Rc, 0, 1, dy, rdy, &br);
        }
        Fs1 = {G2.h, G2.u, G2.v, G2.z};
      }
      if constexpr (KOUT) {
        if (own) {
          if (A.br_y) A.br_y[c * nx + x] = br;
          if (A.dg_fh) {
            const int64_t fi = int64_t(nx + 1) * ny + int64_t(c) * nx + x;
            A.dg_fh[fi] = Fs1.h; A.dg_fu[fi] = Fs1.u; A.dg_fv[fi] = Fs1.v; A.dg_fz[fi] = Fs1.z;
          }
        }
      }
    }
    // ---- cell pass + sources + RK for row r = c - 1
    const FaceUV<T> Xe0 = from_e(Xw0);
    const int r = c - 1;
    if (r >= y0 && r < y1) {
      T ey0 = T(0), ey1 = T(0);
      if (storm) {
        const int k = r - y0;
        // a tile has at most kWave rows (mfsv_set_tile_rows): row r's
        // factors are always in the wave's table
        if constexpr (kStormLds) {
          ey0 = storm_rows[wv][0][k];
          ey1 = storm_rows[wv][1][k];
        } else {
          ey0 = lane_value(eyA0, k);
          ey1 = lane_value(eyA1, k);
        }
      }
      const uint32_t wo = lane_off + uint32_t(r) * row_bytes;
      const T h = R0.h, u = R0.u, v = R0.v;
      // TemporalDerivativeKernel.cpp:57-90
      T dhdt = dv2<QUICK>(Xw0.h - Xe0.h, dx, rdx_c, Fs0.h - Fs1.h, dy, rdy_c);
      T dudt = dv2<QUICK>(Xw0.u - Xe0.u, dx, rdx_c, Fs0.u - Fs1.u, dy, rdy_c);
      T dvdt = dv2<QUICK>(Xw0.v - Xe0.v, dx, rdx_c, Fs0.v - Fs1.v, dy, rdy_c);
      dudt = mad_t(p_hx, gx_s, dudt);
      dvdt = mad_t(p_hy, gy_s, dvdt);
      // bed slope limited to |dz/dx| <= h/dx
      T dudt_bed, dvdt_bed;
      if constexpr (QUICK) {
        T zx, zy;
        if constexpr (sizeof(T) == 4) {
          // half slopes are limited to half the depth
          const T hl = kScaled ? T(0.5) * h : h;
          const T hp = fmax_t(hl, T(0));
          zx = bed_limit(p_zx, hl, hp);
          zy = bed_limit(p_zy, hl, hp);
        } else {
          // sycl::sign, not copysign: a negative depth passes the test with a
          // zero slope, whose sign is zero
          zx = (fabs_t(p_zx) > h) ? sy_sign(p_zx) * h : p_zx;
          zy = (fabs_t(p_zy) > h) ? sy_sign(p_zy) * h : p_zy;
        }
        dudt_bed = gx_s * zx;
        dvdt_bed = gy_s * zy;
      } else {
        T dzdx = p_zx;
        const T hx_ = div_c(h, dx, rdx);
        if (fabs_t(dzdx) > hx_) dzdx = sy_sign(dzdx) * hx_;
        T dzdy = p_zy;
        const T hy_ = div_c(h, dy, rdy);
        if (fabs_t(dzdy) > hy_) dzdy = sy_sign(dzdy) * hy_;
        dudt_bed = T(-kG) * dzdx;
        dvdt_bed = T(-kG) * dzdy;
      }
      if constexpr (QUICK) {
        dudt = mad_t(Xw0.z - Xe0.z, rdx_c, dudt);
        dvdt = mad_t(Fs0.z - Fs1.z, rdy_c, dvdt);
      } else {
        dudt += div_c(Xw0.z - Xe0.z, dx, rdx);
        dvdt += div_c(Fs0.z - Fs1.z, dy, rdy);
      }
      dudt += dudt_bed;
      dvdt += dvdt_bed;
      // ManningRoughnessSourceTerm.cpp:55-90
      if (A.manning) {
        T nsh = A.n_sh, ndp = A.n_dp, dsh = A.d_sh, ddp = A.d_dp;
        if (has_n) { nsh = X0.n; ndp = nsh; }
        if constexpr (EXT) {
          if (A.mfull) {
            if constexpr (kPreExt) { ndp = X0.ndp; dsh = X0.dsh; ddp = X0.ddp; }
            else { ndp = ld(A.ndp_f, wo); dsh = ld(A.dsh_f, wo); ddp = ld(A.ddp_f, wo); }
          }
        }
        T n = nsh;  // mix(a, a, t) == a
        // one roughness field without per-cell deep n: n_deep == n_shallow
        // in every cell, the mix is the identity (a wave-uniform skip)
        bool may_mix = !has_n;
        if constexpr (EXT) may_mix = may_mix || A.mfull;
        if (may_mix) {
          T t = vdiv(h - dsh, ddp - dsh);
          t = clamp_t(t, T(0), T(1));
          n = (nsh != ndp) ? mad_t(ndp - nsh, t * t * (T(3) - T(2) * t), nsh) : nsh;
        }
        // the friction acts on wet cells only: a wave without one skips it
        const bool wet = h > T(1e-6);
        T Sf = T(0);
        if (wet) {
          const T inv_h = vdiv(h, mad_t(h, h, T(1e-3)));
          Sf = n * n * fsqrt_t(mad_t(u, u, v * v)) * pow43(inv_h);
          T dum = T(-kG) * Sf * u;
          T dvm = T(-kG) * Sf * v;
          // overshoot cap with the full timestep (ManningRoughnessSourceTerm.cpp:74-81):
          // pu = ((-g Sf) u) dt with Sf >= 0 (a product of n^2, a square root
          // and a power; NaN fails every test), so where |pu| > |u| both are
          // non-zero and pu's sign is the opposite of u's: the reference's
          // sign test always passes there and is not repeated
          const T pu = dum * dtT, pv = dvm * dtT;
          if (fabs_t(pu) > fabs_t(u)) dum = div_dt(-u, dtT, rdt);
          if (fabs_t(pv) > fabs_t(v)) dvm = div_dt(-v, dtT, rdt);
          dudt = dudt + dum;
          dvdt = dvdt + dvm;
        }
        if constexpr (KOUT) {
          if (A.dg_nh && own) {
            st(A.dg_nh, wo, n);
            st(A.dg_sf, wo, Sf);  // 0 where dry
          }
        }
      }
      if constexpr (KOUT) {
        if (A.dg_zx && own) {  // Constants.cpp:27-44: Minmod3 slopes of the bed
          // the difference form holds dz (exact multiple of the slope for
          // power-of-two spacing), the quotient form the slope itself
          st(A.dg_zx, wo, QUICK ? p_zx * rdx : p_zx);
          st(A.dg_zy, wo, QUICK ? p_zy * rdy : p_zy);
        }
      }
      if constexpr (EXT) {
        // overshoot cap shared by the velocity sources: a change that would
        // reverse the velocity within the timestep becomes -vel/dt
        auto vcap = [&](T d, T vel) -> T {
          const T p = d * dtT;
          return (fabs_t(p) > fabs_t(vel) && sy_sign(p) != sy_sign(vel)) ? div_dt(-vel, dtT, rdt) : d;
        };
        const bool wet = h > T(1e-6);
        // EnergyLossSourceTerm.cpp:59-88
        if (A.eloss) {
          const T hyp = fsqrt_t(u * u + v * v);
          const T fxv = kPreExt ? X0.fx : ld(A.fx, wo), fyv = kPreExt ? X0.fy : ld(A.fy, wo);
          const T due = vcap(T(-0.5) * u * fxv * hyp, u);
          const T dve = vcap(T(-0.5) * v * fyv * hyp, v);
          dudt = wet ? dudt + due : dudt;
          dvdt = wet ? dvdt + dve : dvdt;
        }
        // InfiltrationSourceTerm.cpp:42-64 (the capacity is consumed here)
        if (A.infil) {
          T di = (kPreExt ? X0.irate : ld(A.irate, wo)) * dtT;
          di = (di > h) ? h : di;
          const T capv = kPreExt ? X0.icap : ld(A.icap, wo);
          di = (di > capv) ? capv : di;
          dhdt -= div_dt(di, dtT, rdt);
          if (own && !(KOUT && A.diag)) st(A.icap, wo, T(capv - di));
        }
        // EddyViscositySourceTerm.cpp:44-98: Minmod3 of the first-derivative
        // fields (x of du/dx, y of du/dy), Smagorinsky or per-cell mu
        if (A.eddy) {
          const T Wu = from_w(p_ux), Eu = from_e(p_ux), Wv = from_w(p_vx), Ev = from_e(p_vx);
          const bool hS = r > 0, hN = r + 1 < ny;
          T d2ux, d2vx, d2uy, d2vy;
          if constexpr (!NANBED) {
            d2ux = minmod_q(qx(p_ux - Wu), qx(Eu - p_ux), qx(Eu - Wu), hasW, hasE);
            d2vx = minmod_q(qx(p_vx - Wv), qx(Ev - p_vx), qx(Ev - Wv), hasW, hasE);
            const T yu = minmod_q(qy(p_uy - pp_uy), qy(suy - p_uy), qy(suy - pp_uy), hS, hN);
            const T yv = minmod_q(qy(p_vy - pp_vy), qy(svy - p_vy), qy(svy - pp_vy), hS, hN);
            d2uy = e7 ? yu : -yu;
            d2vy = e7 ? yv : -yv;
            if constexpr (QUICK) {
              d2ux = d2ux * rdx * rdx; d2vx = d2vx * rdx * rdx;
              d2uy = d2uy * rdy * rdy; d2vy = d2vy * rdy * rdy;
            }
          } else {
            d2ux = minmod3(hasW ? Wu : p_ux, hasW, p_ux, hasE, hasE ? Eu : p_ux, dx, rdx);
            d2vx = minmod3(hasW ? Wv : p_vx, hasW, p_vx, hasE, hasE ? Ev : p_vx, dx, rdx);
            const T Nu = hN ? suy : p_uy, Su = hS ? pp_uy : p_uy;
            const T Nv = hN ? svy : p_vy, Sv = hS ? pp_vy : p_vy;
            if (e7) {
              d2uy = minmod3(Su, hS, p_uy, hN, Nu, dy, rdy);
              d2vy = minmod3(Sv, hS, p_vy, hN, Nv, dy, rdy);
            } else {
              d2uy = minmod3(Nu, hN, p_uy, hS, Su, dy, rdy);
              d2vy = minmod3(Nv, hN, p_vy, hS, Sv, dy, rdy);
            }
          }
          T mu;
          if (A.eddy == 2) {
            T dudx_ = p_ux, dvdy_ = p_vy, dudy_ = p_uy, dvdx_ = p_vx;
            if constexpr (QUICK) { dudx_ *= rdx; dvdx_ *= rdx; dudy_ *= rdy; dvdy_ *= rdy; }
            const T t = fabs_t(dudy_) + fabs_t(dvdx_);
            mu = A.mu_const + A.mu_smag * A.area *
                 fsqrt_t(dudx_ * dudx_ + dvdy_ * dvdy_ + T(0.5) * (t * t));
          } else {
            mu = kPreExt ? X0.mu : ld(A.mu, wo);
          }
          const T due = vcap(-mu * (d2ux + d2uy), u);
          const T dve = vcap(-mu * (d2vx + d2vy), v);
          dudt = wet ? dudt + due : dudt;
          dvdt = wet ? dvdt + dve : dvdt;
        }
      }
      // DischargeBoundarySourceTerm.cpp:43-67 (E1 flips the guard)
      if (A.discharge || storm) {
        T tq0, tq1;
        if (kBnd && A.discharge == 2) {
          if constexpr (kPreBnd) { tq0 = X0.q0; tq1 = X0.q1; }
          else { tq0 = ld(A.q0f, wo); tq1 = ld(A.q1f, wo); }
        }
        else if (A.discharge == 1) { tq0 = T(q0s); tq1 = T(q1s); }
        else { tq0 = T(0); tq1 = T(0); }
        if (storm) {
          tq0 += T(A.storm.peak) * ex0 * ey0;
          tq1 += T(A.storm.peak) * ex1 * ey1;
        }
        T dh;
        if constexpr (sizeof(T) == 8) {
          const double dqdt = div_c(tq1 - tq0, step_len, A.rstep);
          const double qnow = tq0 + tsub * dqdt;
          const double qnext = qnow + dt * dqdt;
          dh = 0.5 * (qnow + qnext);
        } else {
          const float dqdt = (tq1 - tq0) * float(A.rstep);
          const float qnow = __builtin_fmaf(float(tsub), dqdt, tq0);
          const float qnext = __builtin_fmaf(dtT, dqdt, qnow);
          dh = 0.5f * (qnow + qnext);
        }
        const bool dry = (A.flags & 0x01) ? (h + dh * dtT <= T(0)) : (h - dh * dtT <= T(0));
        dh = dry ? div_dt(-h, dtT, rdt) : dh;
        dhdt = (tq0 != T(0) || tq1 != T(0)) ? dhdt + dh : dhdt;
      }
      // HeadBoundarySourceTerm.cpp:43-66 (overwrites dh)
      if (A.head) {
        T h0, h1;
        if (kBnd && A.head == 2) {
          if constexpr (kPreBnd) { h0 = X0.hb0; h1 = X0.hb1; }
          else { h0 = ld(A.hb0f, wo); h1 = ld(A.hb1f, wo); }
        }
        else {
          const bool in = x >= A.hx0 && x < A.hx1;
          h0 = in ? T(hb0s) : Num<T>::nan();
          h1 = in ? T(hb1s) : Num<T>::nan();
        }
        T target;
        if constexpr (sizeof(T) == 8) {
          const double dhb = div_c(h1 - h0, step_len, A.rstep);
          const double now = h0 + tsub * dhb;
          const double next = now + dt * dhb;
          target = 0.5 * (now + next);
        } else {
          const float dhb = (h1 - h0) * float(A.rstep);
          const float now = __builtin_fmaf(float(tsub), dhb, h0);
          const float next = __builtin_fmaf(dtT, dhb, now);
          target = 0.5f * (now + next);
        }
        target = (target <= T(0)) ? T(0) : target;
        dhdt = (!(h0 != h0) && !(h1 != h1)) ? div_dt(target - h, dtT, rdt) : dhdt;
      }
      // StageBoundarySourceTerm.cpp:43-68: the head boundary on level - bed
      if constexpr (EXT) {
        if (A.stageb) {
          const T h0 = (kPreExt ? X0.sb0 : ld(A.sb0, wo)) - R0.z;
          const T h1 = (kPreExt ? X0.sb1 : ld(A.sb1, wo)) - R0.z;
          T target;
          if constexpr (sizeof(T) == 8) {
            const double dhb = div_c(h1 - h0, step_len, A.rstep);
            const double now = h0 + tsub * dhb;
            const double next = now + dt * dhb;
            target = 0.5 * (now + next);
          } else {
            const float dhb = (h1 - h0) * float(A.rstep);
            const float now = __builtin_fmaf(float(tsub), dhb, h0);
            const float next = __builtin_fmaf(dtT, dhb, now);
            target = 0.5f * (now + next);
          }
          target = (target <= T(0)) ? T(0) : target;
          dhdt = (!(h0 != h0) && !(h1 != h1)) ? div_dt(target - h, dtT, rdt) : dhdt;
        }
      }
      rate += double(dhdt);  // lanes that own no cell are dropped after the march
      if constexpr (KOUT) {
        if (own) {
          st(A.k_h, wo, dhdt);
          st(A.k_u, wo, dudt);
          st(A.k_v, wo, dvdt);
        }
      } else {
        // RungeKutta.hpp:377-388: y_i = y_i + k_s * (a dt)
        T oh = h, ou = u, ov = v, ol = T(0);
        // FP32 form: the final row's depth is kept as hi + lo; its increment
        // (with the product's rounding error) and the running low part are
        // added to hi with an error-free TwoSum, so the FP64 volume sum of
        // hi + lo tracks the accumulated rates without FP32 rounding drift
        const T lo_in = X0.lo;
#pragma unroll
        for (int j = 0; j < NACC; ++j) {
          T bh = h, bu = u, bv = v;
          if constexpr (base_ld) { bh = X0.bh[j]; bu = X0.bu[j]; bv = X0.bv[j]; }
          T nh, nlo = T(0);
          T nu, nv;
          if constexpr (CONVEX) {
            if constexpr (BASE) {
              // y2 = (y0 + y1)/2 + k1 dt/2: the halves of the high parts of
              // y0 and y1 summed error-free, the low term of (y0 + y1)/2
              // that the first stage left (the input low word), the
              // increment and its rounding error added, then TwoSum into
              // hi + lo
              const T b0 = T(0.5) * bh, b1 = T(0.5) * h;
              const T s2 = b0 + b1, sb = s2 - b0;
              const T sl = (b0 - (s2 - sb)) + (b1 - sb);
              const T p = dhdt * sc[j];
              const T pe = fma_t(dhdt, sc[j], -p);
              const T inc = ((pe + lo_in) + sl) + p;
              nh = s2 + inc;
              const T bb = nh - s2;
              nlo = (s2 - (nh - bb)) + (inc - bb);
              // y1's velocities as stored: y2 = y0 + k0 dt/2 + k1 dt/2
              // whatever E8/E9 made of y1 as a stage input
              nu = mad_t(dudt, sc[j], T(0.5) * (bu + r0u));
              nv = mad_t(dvdt, sc[j], T(0.5) * (bv + r0v));
            } else {
              // y1 = hi(y0) + k0 dt rounded once, as the stencil of the
              // second stage reads it; what that rounding and y0's low part
              // leave out (y1 - y0 - k0 dt exactly, up to the rounding of
              // this small sum) goes to the new state's low-order buffer as
              // lo(y0) + (rounding)/2, the low term of (y0 + y1)/2
              const T p = dhdt * sc[j];
              const T pe = fma_t(dhdt, sc[j], -p);
              nh = bh + p;
              const T bb = nh - bh;
              const T e = (bh - (nh - bb)) + (p - bb);  // hi(y0) + p == nh + e
              nlo = fma_t(T(0.5), e + pe, lo_in);
              nu = mad_t(dudt, sc[j], bu);
              nv = mad_t(dvdt, sc[j], bv);
            }
          } else if (sizeof(T) == 4 && j == NACC - 1) {
            const T p = dhdt * sc[j];
            const T pe = fma_t(dhdt, sc[j], -p);   // exact: dhdt*sc == p + pe
            const T b = (pe + lo_in) + p;
            nh = bh + b;                              // TwoSum(bh, b)
            const T bb = nh - bh;
            nlo = (bh - (nh - bb)) + (b - bb);
          } else {
            nh = mad_t(dhdt, sc[j], bh);
          }
          if constexpr (!CONVEX) {
            nu = mad_t(dudt, sc[j], bu);
            nv = mad_t(dvdt, sc[j], bv);
          }
          if (j == 0) {  // row s+1 is complete after this stage
            // (the convex form stores y1 as it is; its readers complete it)
            if constexpr (!(CONVEX && !BASE)) complete(nh, nu, nv);
            oh = nh; ou = nu; ov = nv; ol = nlo;
          }
          if (own) {
            st(oh_[j], wo, nh);
            st(ou_[j], wo, nu);
            st(ov_[j], wo, nv);
            if constexpr (sizeof(T) == 4) {
              if (j == NACC - 1) st(cur ? A.Uhl[0] : A.Uhl[1], wo, nlo);
            }
          }
        }
        if (A.last) {
          // State.cpp:211-222 on the new state; max_i dt*q_i == dt*max_i q_i
          const T cc = fsqrt_t(G * max0(oh));
          const T q = dv2<QUICK>(fabs_t(ou) + cc, dx, rdx, fabs_t(ov) + cc, dy, rdy);
          qmax = fmax_t(qmax, q);  // q >= 0; a NaN q is skipped, as by q > qmax
          if constexpr (sizeof(T) == 4) vol += double(oh) + double(ol);
          else vol += oh;  // ol == 0: the sum is the same
        }
      }
    }
    R0 = R1; R1 = R2; R2 = Rn;
    r0u = r1u; r1u = r2u; r2u = rnu;
    r0v = r1v; r1v = r2v; r2v = rnv;
    if constexpr (!(QUICK && sizeof(T) == 8)) { qsh = qnh; qsu = qnu; qsv = qnv; qsz = qnz; }
    if constexpr (EXT) { pp_uy = p_uy; pp_vy = p_vy; }
    if constexpr (EXT || NANBED) { p_ux = sux; p_vx = svx; }
    p_hx = shx; p_zx = szx; p_hy = shy; p_uy = suy; p_vy = svy; p_zy = szy;
    Fs0 = Fs1;
    Xw0 = Xw1;
  };
  // FP32: two rows per trip (the loop holds cross-lane operations, which the
  // compiler will not unroll with a remainder), so that the row registers are
  // renamed instead of copied.  An odd row count runs one extra step at
  // c = y1 + 1, which computes no face or cell of the tile and stores nothing
  // (its loads are clamped into the grid).  FP64 keeps one row per trip: its
  // register budget at 3 waves per SIMD has no room for the second row.
  // Issue priority falls as the wave gets through its tile.  A SIMD's
  // arbiter otherwise favours its oldest wave among equals: the waves of a
  // launch that fits the chip in one round start together but finish in
  // the order they were dispatched (on the 2048^2 storm catchment the FP32
  // first stage's waves took 29.5, 31.9, 35.1 and 39.1 us by dispatch
  // quartile), and the rest of the launch ran on fewer and fewer waves per
  // SIMD.  A wave past half of its rows yields to those behind it (priority
  // 3 -> 2 at 1/2, 1 at 3/4, 0 at 15/16), so a SIMD's waves finish nearer
  // together: FP32 stages 46.4 -> 41.7 us, FP64 at 24-row tiles 114 -> 105
  // us.  Results do not depend on it.
  const int nsteps = y1 - y0 + 2;
  const int q1 = nsteps / 2, q2 = (3 * nsteps) / 4, q3 = (15 * nsteps) / 16;
  __builtin_amdgcn_s_setprio(3);
  auto yield_by_progress = [&](const int c) {
    const int k = c - (y0 - 1);
    if (k >= q3) __builtin_amdgcn_s_setprio(0);
    else if (k >= q2) __builtin_amdgcn_s_setprio(1);
    else if (k >= q1) __builtin_amdgcn_s_setprio(2);
  };
  if constexpr (sizeof(T) == 4) {
    for (int c = y0 - 1; c <= y1; c += 2) {
      yield_by_progress(c);
      row_step(c);
      row_step(c + 1);
    }
  } else {
    for (int c = y0 - 1; c <= y1; ++c) {
      yield_by_progress(c);
      row_step(c);
    }
  }
  __builtin_amdgcn_s_setprio(0);

  // the lanes that own a cell, then wave reductions (fixed butterfly order ->
  // deterministic per tile)
  if (!own) { rate = 0.0; vol = 0.0; qmax = T(0); }
  double cnmax = dt * double(qmax);
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
      A.part_vol[A.ntiles + tile] = cnmax;  // reduced by the controller: no same-address atomics
    }
  }
}

// One catchment: one wavefront per tile.
// Occupancy: the FP64 difference-form kernel is held to 3 waves per SIMD
// (<= 168 registers, at most a few bytes of scratch); the other FP64
// variants keep 2 (they would spill), FP32 runs at 4.
template <typename T, bool QUICK, bool NANBED, bool EXT>
constexpr int stage_waves() {
  return sizeof(T) == 4 ? 4 : ((QUICK && !NANBED && !EXT) ? 3 : 2);
}
template <typename T, int NACC, bool KOUT, bool QUICK, bool NANBED, bool EXT, bool BASE, bool BND,
          bool CONVEX>
__global__ __launch_bounds__(kWave * kWavesPerBlock)
__attribute__((amdgpu_waves_per_eu(stage_waves<T, QUICK, NANBED, EXT>()))) void k_stage(const StageArgs<T> A) {
  const int lane = threadIdx.x & (kWave - 1);
  // wave-uniform: everything derived from the tile index stays scalar
  const int tile = __builtin_amdgcn_readfirstlane(xcd_block(blockIdx.x, gridDim.x) * kWavesPerBlock +
                                                  (int)(threadIdx.x >> 6));
  if (tile >= A.ntiles) return;
  stage_tile<T, NACC, KOUT, QUICK, NANBED, EXT, BASE, BND, CONVEX>(A, tile, lane);
}

// Many catchments in one launch: the grid walks a tile table of (member,
// tile) pairs, a wave per entry when the grid was sized for the table, more
// with the stride loop when the table has grown since; member m's arguments
// are margs[m].  Each tile runs
// exactly the code of the single-catchment kernel for its member (with
// GENERAL, the coded-out or further-sources variant where the member needs
// it), so a catchment's result
// does not depend on which other catchments share the launch.  The table
// length is read from device memory, so a captured launch is correct for any
// table (the stride loop covers one longer than its grid).
template <typename T, int NACC, bool QUICK, bool GENERAL, bool BASE, bool CONVEX>
__global__ __launch_bounds__(kWave * kWavesPerBlock) void k_stage_batch(
    const StageArgs<T>* __restrict__ margs, const int2* __restrict__ tiles,
    const int* __restrict__ ntiles_p) {
  const int lane = threadIdx.x & (kWave - 1);
  const int nwaves = gridDim.x * kWavesPerBlock;
  const int ntiles = *ntiles_p;
  const int k0 = __builtin_amdgcn_readfirstlane(xcd_block(blockIdx.x, gridDim.x) * kWavesPerBlock +
                                                (int)(threadIdx.x >> 6));
  // the member arguments are read through the constant address space: loads
  // that no store of the kernel can alias, which the compiler may repeat
  // where it needs a value instead of keeping it (as it does with a kernel's
  // own arguments) rather than reloading and waiting after the row stores
#if defined(__HIP_DEVICE_COMPILE__)
  using ConstArgs = const __attribute__((address_space(4))) StageArgs<T>;
#else  // the host pass of a kernel body has no GPU address spaces
  using ConstArgs = const StageArgs<T>;
#endif
  ConstArgs* const cargs = (ConstArgs*)margs;
  for (int k = k0; k < ntiles; k += nwaves) {
    const int m = __builtin_amdgcn_readfirstlane(tiles[k].x);
    const int t = __builtin_amdgcn_readfirstlane(tiles[k].y);
    const StageArgs<T> a = cargs[m];
    if constexpr (GENERAL) {
      const bool ext = (a.eloss | a.infil | a.stageb | a.eddy) != 0;
      if (a.nanbed) {
        stage_tile<T, NACC, false, false, true, true, BASE, true, CONVEX>(a, t, lane);
        continue;
      }
      if (ext) {
        stage_tile<T, NACC, false, QUICK, false, true, BASE, true, CONVEX>(a, t, lane);
        continue;
      }
    }
    stage_tile<T, NACC, false, QUICK, false, false, BASE, GENERAL, CONVEX>(a, t, lane);
  }
}

}  // namespace mfsv
