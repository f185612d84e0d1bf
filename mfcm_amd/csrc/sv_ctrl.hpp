// sv_ctrl.hpp -- device-side controller of the Runge-Kutta timestep loop.
#pragma once
#include "sv_stage.hpp"

namespace mfsv {

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
__device__ __forceinline__ void ctrl_body(const CtrlArgs& A, double* sh) {
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

__global__ __launch_bounds__(256) void k_ctrl(const CtrlArgs A) {
  __shared__ double sh[256];
  ctrl_body(A, sh);
}

// One block per member of a batch; the member count is read from device
// memory so that captured graphs survive changes of membership.
__global__ __launch_bounds__(256) void k_ctrl_batch(const CtrlArgs* __restrict__ args,
                                                    const int* __restrict__ nmembers_p) {
  __shared__ double sh[256];
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
