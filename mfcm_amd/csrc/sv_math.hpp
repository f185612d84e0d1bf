// sv_math.hpp -- device arithmetic of the Saint-Venant stage (gfx950).
//
// Restates the per-point arithmetic of mfcm's Cartesian-mesh solver:
//   Minmod3                SpatialDerivative/Minmod3.hpp:34-82
//   face flux (4 branches) SaintVenant/FluxKernel.cpp:66-295
// For T = double the operation order is the reference's, term by term, so
// the FP64 path reproduces it to the last bit (the file is compiled with
// -ffp-contract=off; the only fused operations are the explicit fma()s of
// div_c, which return the correctly rounded quotient).  For T = float the
// same routines form the precision form, whose wet/dry tests compare depths
// with bed steps instead of summed surface elevations.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mfsv {

constexpr double kG = 9.81;

template <typename T> struct Num;
template <> struct Num<double> {
  static __device__ __forceinline__ double nan() { return __builtin_nan(""); }
};
template <> struct Num<float> {
  static __device__ __forceinline__ float nan() { return __builtin_nanf(""); }
};

__device__ __forceinline__ double fma_t(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float fma_t(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double sqrt_t(double a) { return __builtin_sqrt(a); }
__device__ __forceinline__ float sqrt_t(float a) { return __builtin_sqrtf(a); }
__device__ __forceinline__ double fabs_t(double a) { return __builtin_fabs(a); }
__device__ __forceinline__ float fabs_t(float a) { return __builtin_fabsf(a); }
__device__ __forceinline__ double fmax_t(double a, double b) { return __builtin_fmax(a, b); }
__device__ __forceinline__ float fmax_t(float a, float b) { return __builtin_fmaxf(a, b); }
__device__ __forceinline__ double fmin_t(double a, double b) { return __builtin_fmin(a, b); }
__device__ __forceinline__ float fmin_t(float a, float b) { return __builtin_fminf(a, b); }
__device__ __forceinline__ bool isnan_t(double a) { return __builtin_isnan(a); }
__device__ __forceinline__ bool isnan_t(float a) { return __builtin_isnan(a); }

// a / b for a divisor b known on the host with rb = 1/b rounded to nearest:
// q = a*rb, then one residual step with exact fma.  For binary64 this is the
// correctly rounded quotient (Markstein's correction with a correctly
// rounded reciprocal; brute-force checked on 3.5e8 operand pairs), so it
// equals the reference's a / b wherever neither overflows nor underflows.
template <typename T>
__device__ __forceinline__ T div_c(T a, T b, T rb) {
  T q = a * rb;
  T e = fma_t(-q, b, a);
  return fma_t(e, rb, q);
}

// x^(4/3) for the Manning term (sycl::pow(inv_h, 4/3),
// ManningRoughnessSourceTerm.cpp:66-67) from IEEE operations only, so the
// CPU oracle computes the same bits: five Newton steps for r = x^(-1/3) from
// an exponent-bit estimate, then (x r)^2.  Within 7 ulp of x^(4/3) for the
// Manning range x in (1e-3, 16).
__device__ __forceinline__ double pow43(double x) {
  unsigned long long i = (unsigned long long)__double_as_longlong(x);
  i = 0x5540000000000000ULL - i / 3ull;
  double r = __longlong_as_double((long long)i);
  const double third = 1.0 / 3.0;
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    const double r3 = r * r * r;
    const double t = 4.0 - x * r3;
    r = r * t * third;
  }
  const double c = x * r;
  return c * c;
}
// exp(x) for x <= 0 from IEEE operations only, the same expression as the
// CPU oracle's or_exp_neg (Cody-Waite reduction by ln 2, degree-12 Taylor
// polynomial, exact 2^k scaling): within 2.3 ulp, identical bits on both.
__device__ __forceinline__ double exp_neg(double x) {
  if (x < -708.0) return 0.0;
  const double ln2_hi = 6.93147180369123816490e-01, ln2_lo = 1.90821492927058770002e-10;
  const double log2e = 1.44269504088896338700e+00;
  const double k = floor(x * log2e + 0.5);
  const double r = (x - k * ln2_hi) - k * ln2_lo;
  double p = 1.0 / 479001600.0;
  p = p * r + 1.0 / 39916800.0;
  p = p * r + 1.0 / 3628800.0;
  p = p * r + 1.0 / 362880.0;
  p = p * r + 1.0 / 40320.0;
  p = p * r + 1.0 / 5040.0;
  p = p * r + 1.0 / 720.0;
  p = p * r + 1.0 / 120.0;
  p = p * r + 1.0 / 24.0;
  p = p * r + 1.0 / 6.0;
  p = p * r + 0.5;
  p = p * r + 1.0;
  p = p * r + 1.0;
  return ldexp(p, (int)k);
}
__device__ __forceinline__ float exp_neg(float x) { return __expf(x); }

// FP32 form: hardware log2/exp2.
__device__ __forceinline__ float pow43(float x) {
  return __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(x) * (4.0f / 3.0f));
}

// sycl::sign: 1 for x > 0, -1 for x < 0, the signed zero itself, 0 for NaN.
template <typename T>
__device__ __forceinline__ T sy_sign(T x) {
  return x > T(0) ? T(1) : (x < T(0) ? T(-1) : (x != x ? T(0) : x));
}
// min(max(x, lo), hi) for lo <= hi: one median instruction in FP32.
__device__ __forceinline__ float clamp_t(float x, float lo, float hi) {
  return __builtin_amdgcn_fmed3f(x, lo, hi);
}
__device__ __forceinline__ double clamp_t(double x, double lo, double hi) {
  return __builtin_fmin(__builtin_fmax(x, lo), hi);
}
// max(x, 0): one median instruction in FP32 (fmax of a value that the
// compiler cannot prove canonical costs a second instruction that quiets a
// signalling NaN first; the upper bound is the largest finite float, since
// the compiler turns a median with +inf back into that fmax).  Equal to
// max(x, 0) for every finite x; a NaN gives 0 in both forms.
__device__ __forceinline__ float max0(float x) {
  return __builtin_amdgcn_fmed3f(x, 0.0f, 3.40282347e+38f);
}
__device__ __forceinline__ double max0(double x) { return __builtin_fmax(x, 0.0); }
// Sign bits differ.
__device__ __forceinline__ bool signs_differ(float a, float b) {
  return (__float_as_int(a) ^ __float_as_int(b)) < 0;
}
__device__ __forceinline__ bool signs_differ(double a, double b) {
  return (__double2hiint(a) ^ __double2hiint(b)) < 0;
}
// The bed slope limit of TemporalDerivativeKernel.cpp:72-80 in FP32,
// (|x| > h) ? sycl::sign(x) * h : x, without the sign function: for x != 0
// the product is h with x's sign bit applied, and with hp = max(h, 0) the
// test passes exactly where the reference's does for x != 0 (a negative
// depth passes it for every non-zero slope); a zero x is returned as it is
// where the reference returns a zero of either sign.
__device__ __forceinline__ float bed_limit(float x, float h, float hp) {
  const float s = __int_as_float(__float_as_int(h) ^ (__float_as_int(x) & int(0x80000000u)));
  return (fabs_t(x) > hp) ? s : x;
}
// sycl::min on floating point: y < x ? y : x.
template <typename T>
__device__ __forceinline__ T sy_min(T x, T y) { return (y < x) ? y : x; }

// Minmod3 with theta = 2 on a uniform axis.  A missing neighbour is the
// cell itself at distance 0 (Cartesian2DMesh.hpp:477-520): pass its value as
// Uc with has = false.  d/rd: the axis spacing and its reciprocal.
template <typename T>
__device__ __forceinline__ T minmod3(T Ul, bool hasl, T Uc, bool hasr, T Ur, T d, T rd) {
  T lhs = hasl ? div_c(T(2) * (Uc - Ul), d, rd) : Num<T>::nan();
  T rhs = hasr ? div_c(T(2) * (Ur - Uc), d, rd) : Num<T>::nan();
  T central = T(0);
  if (hasl || hasr) {
    T q = div_c(Ur - Ul, d, rd);  // (Ur-Ul)/(2d) == 0.5*((Ur-Ul)/d) exactly
    central = (hasl && hasr) ? T(0.5) * q : q;
  }
  T outer;
  if (isnan_t(lhs)) outer = rhs;
  else if (isnan_t(rhs)) outer = lhs;
  else outer = T(0.5) * (sy_sign(lhs) + sy_sign(rhs)) * sy_min(fabs_t(lhs), fabs_t(rhs));
  if (isnan_t(outer)) return central;
  return T(0.5) * (sy_sign(outer) + sy_sign(central)) * sy_min(fabs_t(outer), fabs_t(central));
}

// One side of a face: cell values and the limited slopes along the face
// normal.  un/ut: velocity normal/tangential to the face (u/v on x-faces,
// v/u on y-faces); s*: the matching slopes.
template <typename T> struct Side {
  T h, un, ut, z;
  T sh, sun, sut, sz;
};

// Face flux in normal/tangential form: h, normal momentum, tangential
// momentum, wall force ("z-flux").
template <typename T> struct Flux {
  T h, n, t, z;
};

// A cell as one side of a face sees it in the general form: values and all
// eight limited slopes (along x and along y).
template <typename T> struct CellV {
  T h, u, v, z;
  T zx, zy, hx, ux, vx, hy, uy, vy;
};

// A face flux in x/y components: h, u momentum, v momentum, wall force.
template <typename T> struct FluxXY {
  T h, u, v, z;
};

// FluxKernel.cpp:66-295 for one face, term for term, as the catchments with
// coded-out (NaN-bed) cells need it: the projections keep the tangential
// slopes multiplied by the zero direction factor and the normal speed keeps
// the tangential velocity times zero, because next to coded-out cells a
// slope (or, after a timestep, a velocity) can be NaN and then poisons the
// reference's face where a finite value would only add a signed zero.
// dir 0: x-face (d = dx), 1: y-face (d = dy); rd = 1/d.
// DEPTHREL: wet tests as h_m > max(zb_p - zb_m, 0) (the FP32 form).
template <typename T, bool DEPTHREL>
__device__ __forceinline__ FluxXY<T> face_general(CellV<T> L, CellV<T> R, int edge, int dir, T d,
                                                  T rd, float* branch) {
  FluxXY<T> F;
  const T xdir = T(1) - T(dir), ydir = T(dir);              // :76-77
  T zb_L = L.z, zb_R = R.z;                                 // :80-81
  if (zb_L != zb_L) {                                       // :85-95
    L = R;
    if (edge == 1) goto zero;
    edge = -1;
  }
  if (zb_R != zb_R) {                                       // :96-107
    if (edge == -1) goto zero;
    R = L;
    edge = 1;
  }
  {
    const T mL = edge < 0 ? T(0) : T(1), mR = edge > 0 ? T(0) : T(1);
    const T mLx = (edge < 0 && dir == 0) ? T(0) : T(1), mRx = (edge > 0 && dir == 0) ? T(0) : T(1);
    const T mLy = (edge < 0 && dir == 1) ? T(0) : T(1), mRy = (edge > 0 && dir == 1) ? T(0) : T(1);
    const T h_L = L.h * mL, h_R = R.h * mR;                 // :111-112
    const T u_L = L.u * mLx, u_R = R.u * mRx;               // :116-117
    const T v_L = L.v * mLy, v_R = R.v * mRy;               // :121-122
    const T hx_L = L.hx * mL, hx_R = R.hx * mR;             // :125-128
    const T hy_L = L.hy * mL, hy_R = R.hy * mR;
    const T ux_L = L.ux * mLx, ux_R = R.ux * mRx;           // :132-135
    const T uy_L = L.uy * mLx, uy_R = R.uy * mRx;
    const T vx_L = L.vx * mLy, vx_R = R.vx * mRy;           // :139-142
    const T vy_L = L.vy * mLy, vy_R = R.vy * mRy;
    const T zx_L = L.zx * mL, zx_R = R.zx * mR;             // :145-148
    const T zy_L = L.zy * mL, zy_R = R.zy * mR;
    if (edge < 0) zb_L = zb_R + h_R * T(2);                 // :152-157
    if (edge > 0) zb_R = zb_L + h_L * T(2);
    // :161-187, evaluated left to right ((0.5 d) s) dir
    const T hd = T(0.5) * d, mhd = T(-0.5) * d;
    const T zb_m = zb_L + hd * zx_L * xdir + hd * zy_L * ydir;
    T h_m = h_L + hd * hx_L * xdir + hd * hy_L * ydir;
    const T u_m = u_L + hd * ux_L * xdir + hd * uy_L * ydir;
    const T v_m = v_L + hd * vx_L * xdir + hd * vy_L * ydir;
    const T zb_p = zb_R + mhd * zx_R * xdir + mhd * zy_R * ydir;
    T h_p = h_R + mhd * hx_R * xdir + mhd * hy_R * ydir;
    const T u_p = u_R + mhd * ux_R * xdir + mhd * uy_R * ydir;
    const T v_p = v_R + mhd * vx_R * xdir + mhd * vy_R * ydir;
    // :191-195: the wet tests and their dry counterparts, all on the
    // unclamped projections; a NaN operand fails both, which gives branch 4
    bool wm, wp, dm, dp;
    if (DEPTHREL) {
      const T dz = zb_p - zb_m;
      const T lm = fmax_t(dz, T(0)), lp = fmax_t(-dz, T(0));
      wm = h_m > lm; wp = h_p > lp;
      dm = h_m <= lm; dp = h_p <= lp;
    } else {
      const T zb_f = fmax_t(zb_m, zb_p);
      const T y_m = zb_m + h_m, y_p = zb_p + h_p;
      wm = y_m > zb_f; wp = y_p > zb_f;
      dm = y_m <= zb_f; dp = y_p <= zb_f;
    }
    h_m = fmax_t(h_m, T(0));                                // :199-200
    h_p = fmax_t(h_p, T(0));
    const T G = T(kG);
    const T c_m = sqrt_t(G * h_m), c_p = sqrt_t(G * h_p);   // :203-204
    if (wm && wp) {                                         // :208-234
      const T spd_m = u_m * xdir + v_m * ydir;
      const T spd_p = u_p * xdir + v_p * ydir;
      const T hf_m = h_m * spd_m, hf_p = h_p * spd_p;
      const T uf_m = u_m * ((T(1) - T(0.5) * xdir) * spd_m) + G * h_m * xdir;
      const T uf_p = u_p * ((T(1) - T(0.5) * xdir) * spd_p) - G * h_p * xdir;
      const T vf_m = v_m * ((T(1) - T(0.5) * ydir) * spd_m) + G * h_m * ydir;
      const T vf_p = v_p * ((T(1) - T(0.5) * ydir) * spd_p) - G * h_p * ydir;
      const T a = fmax_t(fabs_t(spd_p + sy_sign(spd_p) * c_p), fabs_t(spd_m + sy_sign(spd_m) * c_m));
      F.h = T(0.5) * (hf_p + hf_m) - T(0.5) * a * (h_p - h_m);
      F.u = T(0.5) * (uf_p + uf_m) - T(0.5) * a * (u_p - u_m);
      F.v = T(0.5) * (vf_p + vf_m) - T(0.5) * a * (v_p - v_m);
      F.z = (zb_m - zb_p) * G;
      *branch = 1.0f;
    } else if (dm && dp) {                                  // :235-254
      F.h = T(0);
      if (zb_p > zb_m) {
        F.u = T(0.5) * (G * h_m * xdir);
        F.v = T(0.5) * (G * h_m * ydir);
        F.z = -h_m * T(0.5) * G;
        *branch = 2.25f;
      } else {
        F.u = T(-0.5) * (G * h_p * xdir);
        F.v = T(-0.5) * (G * h_p * ydir);
        F.z = h_p * T(0.5) * G;
        *branch = 2.75f;
      }
    } else if (wm) {                                        // :255-271
      const T spd_m = u_m * xdir + v_m * ydir;
      const T hf_m = h_m * spd_m;
      const T uf_m = u_m * ((T(1) - T(0.5) * xdir) * spd_m) + G * h_m * xdir;
      const T vf_m = v_m * ((T(1) - T(0.5) * ydir) * spd_m) + G * h_m * ydir;
      const T a = fabs_t(spd_m + sy_sign(spd_m) * c_m);
      F.h = T(0.5) * hf_m - T(0.5) * a * (-h_m);
      F.u = T(0.5) * uf_m - T(0.5) * a * (-u_m);
      F.v = T(0.5) * vf_m - T(0.5) * a * (-v_m);
      F.z = div_c(h_p, d, rd);
      *branch = 3.0f;
    } else {                                                // :272-289
      const T spd_p = u_p * xdir + v_p * ydir;
      const T hf_p = h_p * spd_p;
      const T uf_p = u_p * ((T(1) - T(0.5) * xdir) * spd_p) + G * h_p * xdir;
      const T vf_p = v_p * ((T(1) - T(0.5) * ydir) * spd_p) + G * h_p * ydir;
      const T a = fabs_t(spd_p + sy_sign(spd_p) * c_p);
      F.h = T(0.5) * hf_p - T(0.5) * a * h_p;
      F.u = T(0.5) * uf_p - T(0.5) * a * u_p;
      F.v = T(0.5) * vf_p - T(0.5) * a * v_p;
      F.z = div_c(h_m, d, rd);
      *branch = 4.0f;
    }
    return F;
  }
zero:                                                       // :47-60
  F.h = T(0); F.u = T(0); F.v = T(0); F.z = T(0);
  *branch = 0.0f;
  return F;
}

// ---------------------------------------------------------------------------
// Fast forms used when the catchment has no coded-out (NaN-bed) cell.  They
// return the same values as minmod3 / face_flux (signed zeros aside) with far
// fewer instructions.

__device__ __forceinline__ double copysign_t(double m, double s) { return __builtin_copysign(m, s); }
__device__ __forceinline__ float copysign_t(float m, float s) { return __builtin_copysignf(m, s); }

// a*b + c: unfused in FP64 (the reference rounds the product), fused in the
// FP32 form.
__device__ __forceinline__ double mad_t(double a, double b, double c) { return a * b + c; }
__device__ __forceinline__ float mad_t(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// Square root: correctly rounded in FP64, the hardware v_sqrt_f32 in FP32.
__device__ __forceinline__ double fsqrt_t(double a) { return __builtin_sqrt(a); }
__device__ __forceinline__ float fsqrt_t(float a) { return __builtin_amdgcn_sqrtf(a); }

// Minmod3 from the three quotients ql = (Uc-Ul)/d, qr = (Ur-Uc)/d and
// qc = (Ur-Ul)/d (no NaN among them).  With theta = 2 the reference's
// lhs = 2 ql, rhs = 2 qr, central = qc/2 exactly; minmod of those is
// sign * min(2|ql|, 2|qr|, |qc|/2) when ql and qr share a sign (then qc
// does too), else 0.  One-sided at an edge: the reference reduces to the
// remaining quotient.
//
// The two-argument minmod is the median of a, b and 0 (the operand nearer
// zero when they share a sign, else zero), so Minmod3 is
// med3(2 med3(ql, qr, 0), qc/2, 0): the same operand is selected and the
// scalings are exact; only the sign of a zero result may differ, which no
// later operation distinguishes (a zero slope enters sums and products only).
__device__ __forceinline__ float minmod2(float a, float b) {
  return __builtin_amdgcn_fmed3f(a, b, 0.0f);
}
__device__ __forceinline__ double minmod2(double a, double b) {
  return __builtin_fmax(__builtin_fmin(a, b), __builtin_fmin(__builtin_fmax(a, b), 0.0));
}
template <typename T>
__device__ __forceinline__ T minmod_c(T ql, T qr, T qc) {
  return minmod2(T(2) * minmod2(ql, qr), T(0.5) * qc);
}
// The one-sided and isolated cases given the two-sided value `in`: the
// remaining quotient at an edge, zero without neighbours.  HALF: the halved
// form of minmod_ch (the remaining quotient halved, exactly).
template <bool HALF, typename T>
__device__ __forceinline__ T minmod_edge(T in, T ql, T qr, bool hasl, bool hasr) {
  const T l = HALF ? T(0.5) * ql : ql, r = HALF ? T(0.5) * qr : qr;
  return hasl ? (hasr ? in : l) : (hasr ? r : T(0));
}
template <typename T>
__device__ __forceinline__ T minmod_q(T ql, T qr, T qc, bool hasl, bool hasr) {
  return minmod_edge<false>(minmod_c(ql, qr, qc), ql, qr, hasl, hasr);
}
// Half of minmod_c / minmod_q, exactly: med3(med3(ql, qr, 0), qc/4, 0)
// selects the same operand at half scale.
template <typename T>
__device__ __forceinline__ T minmod_ch(T ql, T qr, T qc) {
  return minmod2(minmod2(ql, qr), T(0.25) * qc);
}
template <typename T>
__device__ __forceinline__ T minmod_qh(T ql, T qr, T qc, bool hasl, bool hasr) {
  return minmod_edge<true>(minmod_ch(ql, qr, qc), ql, qr, hasl, hasr);
}

// The Rusanov speed |un + sign(un) c| of FluxKernel.cpp:225-226 with c >= 0:
// |un| + c when un != 0 (the sum of two same-signed values rounds
// symmetrically), and 0 when un is a zero (sycl::sign(0) == 0).  NaN in,
// NaN out.
template <typename T>
__device__ __forceinline__ T rus_speed(T un, T c) {
  return (un == T(0)) ? T(0) : fabs_t(un) + c;
}

// a / b with rb = 1/b: exact division (div_c) unless QUICK, where a * rb is
// used (exact for power-of-two b; the FP32 form uses it always).
template <bool QUICK, typename T>
__device__ __forceinline__ T dv(T a, T b, T rb) {
  if constexpr (QUICK) return a * rb;
  else return div_c(a, b, rb);
}

// a / d + b / e: in the difference form a * rd + b * re with the first
// product fused in FP32 (FP64 rounds both products, as the reference does)
template <bool QUICK, typename T>
__device__ __forceinline__ T dv2(T a, T d, T rd, T b, T e, T re) {
  if constexpr (QUICK) return mad_t(a, rd, b * re);
  else return div_c(a, d, rd) + div_c(b, e, re);
}

// Ghost side of a closed edge face (FluxKernel.cpp:109-157 with the other
// side's cell): no depth, no normal velocity, tangential velocity and its
// normal slope kept, bed raised to the cell's water level plus its depth.
template <typename T>
__device__ __forceinline__ Side<T> wall_ghost(const Side<T>& c) {
  Side<T> g;
  g.h = T(0); g.un = T(0); g.ut = c.ut; g.z = c.z + c.h * T(2);
  g.sh = T(0); g.sun = T(0); g.sut = c.sut; g.sz = T(0);
  return g;
}

// FluxKernel.cpp:159-289 for a face whose sides are given explicitly (wall
// faces pass a wall_ghost): branches 1, 3 and 4 as one Rusanov form over the
// wet sides (a dry side contributes zero, and a side that is wet alone takes
// +g h on the plus side as branch 4 does), branch 2 selected after.
// X2: every component doubled (the halvings of the Rusanov average and of
// the wall force dropped, 2 g for g, and rd passed doubled), exactly twice
// the X2 = false result: the caller halves once per cell instead.
template <typename T, bool DEPTHREL, bool QUICK, bool X2 = false>
__device__ __forceinline__ Flux<T> face_fast(const Side<T>& L, const Side<T>& R, T hd, T d, T rd,
                                             float* branch) {
  const T zm = mad_t(hd, L.sz, L.z);
  T hm = mad_t(hd, L.sh, L.h);
  const T unm = mad_t(hd, L.sun, L.un);
  const T utm = mad_t(hd, L.sut, L.ut);
  const T zp = mad_t(-hd, R.sz, R.z);
  T hp = mad_t(-hd, R.sh, R.h);
  const T unp = mad_t(-hd, R.sun, R.un);
  const T utp = mad_t(-hd, R.sut, R.ut);
  bool wm, wp;
  if (DEPTHREL) {
    const T dz = zp - zm;
    wm = hm > fmax_t(dz, T(0));
    wp = hp > fmax_t(-dz, T(0));
  } else {
    const T zf = fmax_t(zm, zp);
    wm = zm + hm > zf;
    wp = zp + hp > zf;
  }
  const T G = T(kG);
  Flux<T> F;
  if (wm || wp) {
    // a dry side contributes zero: depth, velocities and with them g h and
    // the celerity are selected to zero (a wet side's depth is > 0, so the
    // reference's clamp to >= 0 changes nothing there)
    const T hme = wm ? hm : T(0), unme = wm ? unm : T(0), utme = wm ? utm : T(0);
    const T hpe = wp ? hp : T(0), unpe = wp ? unp : T(0), utpe = wp ? utp : T(0);
    const T gm = G * hme, gpe = G * hpe;
    const T cm = fsqrt_t(gm), cp = fsqrt_t(gpe);
    const T gp = wm ? -gpe : gpe;
    // mad_t: the FP64 form rounds the product (the reference's operation
    // order), the FP32 form fuses it
    const T nfm = mad_t(unme, T(0.5) * unme, gm);
    const T nfp = mad_t(unpe, T(0.5) * unpe, gp);
    const T a = fmax_t(rus_speed(unpe, cp), rus_speed(unme, cm));
    // one side dry: the wall force of the dry side's depth, clamped to >= 0
    const T hdry = max0(wm ? hp : hm);
    if constexpr (X2) {
      F.h = mad_t(-a, hpe - hme, mad_t(hpe, unpe, hme * unme));
      F.n = mad_t(-a, unpe - unme, nfp + nfm);
      F.t = mad_t(-a, utpe - utme, mad_t(utpe, unpe, utme * unme));
      F.z = (wm && wp) ? (zm - zp) * (T(2) * G) : dv<QUICK>(hdry, d, rd);
    } else {
      const T ha = T(0.5) * a;
      F.h = mad_t(-ha, hpe - hme, T(0.5) * mad_t(hpe, unpe, hme * unme));
      F.n = mad_t(-ha, unpe - unme, T(0.5) * (nfp + nfm));
      F.t = mad_t(-ha, utpe - utme, T(0.5) * mad_t(utpe, unpe, utme * unme));
      F.z = (wm && wp) ? (zm - zp) * G : dv<QUICK>(hdry, d, rd);
    }
    *branch = (wm && wp) ? 1.0f : (wm ? 3.0f : 4.0f);
  } else {
    hm = max0(hm);                           // FluxKernel.cpp:199-204
    hp = max0(hp);
    const T ghm = G * hm, ghp = G * hp;
    const bool r = zp > zm;
    F.h = T(0);
    F.t = T(0);
    if constexpr (X2) {
      F.n = r ? ghm : -ghp;
      F.z = r ? -ghm : ghp;                  // -hm G == -(G hm): the same product
    } else {
      F.n = r ? T(0.5) * ghm : T(-0.5) * ghp;
      F.z = r ? -hm * T(0.5) * G : hp * T(0.5) * G;
    }
    *branch = r ? 2.25f : 2.75f;
  }
  return F;
}

}  // namespace mfsv
