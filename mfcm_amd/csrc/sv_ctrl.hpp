// sv_ctrl.hpp -- device-side controller of the Runge-Kutta timestep loop.
#pragma once
#include "sv_stage.hpp"

namespace mfsv {

struct CtrlArgs {
  Ctrl* ctrl;
  const double* part_vol;   // [2 * ntiles]: volume, then control-number maxima
  const double* part_rate;  // [nstages][ntiles]
  int64_t ntiles;
  int32_t nstages;
  int32_t mode;             // 0 = device controller, 1 = host-driven (record only)
  double b[kMaxStages];
  double area;
  const double* forcing;    // [nforcing][4] or null
  int64_t nforcing;
};

constexpr int kCtrlThreads = 1024;
constexpr int kCtrlWaves = kCtrlThreads / kWave;

constexpr int kCtrlVals = 2 + kMaxStages;  // volume, control number, stage rates

// Volume and per-stage source-rate sums and the control-number maximum of the
// tile partials, in one pass: thread t takes tiles t, t + 1024, ...
// (independent accumulators), then a butterfly inside each wave and a
// fixed-order pass over the waves.  Deterministic for a given tile count.
// out[0] = volume partial sum, out[1] = max control number, out[2 + s] =
// stage s rate sum.
__device__ __forceinline__ void tile_sums(const CtrlArgs& A, double* sh, double* out) {
  double acc[kCtrlVals];
#pragma unroll
  for (int q = 0; q < kCtrlVals; ++q) acc[q] = 0.0;
  for (int64_t k = threadIdx.x; k < A.ntiles; k += blockDim.x) {
    acc[0] += A.part_vol[k];
    const double cn = A.part_vol[A.ntiles + k];
    acc[1] = (cn > acc[1]) ? cn : acc[1];
#pragma unroll
    for (int s = 0; s < kMaxStages; ++s)
      if (s < A.nstages) acc[2 + s] += A.part_rate[(int64_t)s * A.ntiles + k];
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
    for (int q = 0; q < kCtrlVals; ++q) {
      const double o = __shfl_xor(acc[q], off, kWave);
      acc[q] = (q == 1) ? ((o > acc[q]) ? o : acc[q]) : acc[q] + o;
    }
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & (kWave - 1)) == 0)
#pragma unroll
    for (int q = 0; q < kCtrlVals; ++q) sh[w * kCtrlVals + q] = acc[q];
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nw = (int)(blockDim.x >> 6);
#pragma unroll
    for (int q = 0; q < kCtrlVals; ++q) {
      double t = 0.0;
      for (int i = 0; i < nw; ++i) {
        const double v = sh[i * kCtrlVals + q];
        t = (q == 1) ? ((v > t) ? v : t) : t + v;
      }
      out[q] = t;
    }
  }
}

// Controller (RungeKutta.cpp:154-167) + accept (RungeKutta.hpp:428-431) +
// landing rules and outer-step transitions (TemporalScheme.hpp:112-149).
__device__ __forceinline__ void ctrl_body(const CtrlArgs& A, double* sh) {
  Ctrl* C = A.ctrl;
  if (C->done) return;
  double sums[kCtrlVals];
  tile_sums(A, sh, sums);
  if (threadIdx.x != 0) return;
  const double vol = sums[0] * A.area;
  double rate = 0.0;
#pragma unroll
  for (int s = 0; s < kMaxStages; ++s)
    if (s < A.nstages) rate += A.b[s] * sums[2 + s];
  const double cn_raw = sums[1];
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

__global__ __launch_bounds__(kCtrlThreads) void k_ctrl(const CtrlArgs A) {
  __shared__ double sh[kCtrlWaves * kCtrlVals];
  ctrl_body(A, sh);
}

// One block per member of a batch; the member count is read from device
// memory so that captured graphs survive changes of membership.
__global__ __launch_bounds__(kCtrlThreads) void k_ctrl_batch(const CtrlArgs* __restrict__ args,
                                                             const int* __restrict__ nmembers_p) {
  __shared__ double sh[kCtrlWaves * kCtrlVals];
  if ((int)blockIdx.x >= *nmembers_p) return;
  ctrl_body(args[blockIdx.x], sh);
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
