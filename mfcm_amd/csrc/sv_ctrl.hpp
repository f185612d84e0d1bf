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
// the control block in 8-byte words, read by the first threads of the block
constexpr int kCtrlWords = int(sizeof(Ctrl) / 8);
static_assert(sizeof(Ctrl) % 8 == 0 && kCtrlWords <= kWave, "control block: whole words, one wave");

// Volume and per-stage source-rate sums and the control-number maximum of the
// tile partials, in one pass: thread t takes tiles t, t + 1024, ... in that
// order, issuing the loads of four tiles before summing them (the pass is
// latency-bound), then a butterfly inside each wave and a fixed-order pass
// over the waves.  Deterministic for a given tile count.
// out[0] = volume partial sum, out[1] = max control number, out[2 + s] =
// stage s rate sum.  The first kCtrlWords threads also copy the control block
// into ctrl_sh: their loads go out with the partials' (one memory round trip
// for the whole kernel, instead of a wait for the block before the partials'
// loads are issued), and the copy is visible to every thread on return.
__device__ __forceinline__ void tile_sums(const CtrlArgs& A, double* sh, double* out,
                                          unsigned long long* ctrl_sh) {
  constexpr int kU = 4;
  const bool cword = threadIdx.x < kCtrlWords;
  unsigned long long cw = 0ull;
  if (cword) cw = reinterpret_cast<const unsigned long long*>(A.ctrl)[threadIdx.x];
  double acc[kCtrlVals];
#pragma unroll
  for (int q = 0; q < kCtrlVals; ++q) acc[q] = 0.0;
  const int64_t nt = A.ntiles;
  for (int64_t k0 = threadIdx.x; k0 < nt; k0 += int64_t(kU) * blockDim.x) {
    double v[kU][kCtrlVals];
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const int64_t k = k0 + int64_t(u) * blockDim.x;
      const bool in = k < nt;
      v[u][0] = in ? A.part_vol[k] : 0.0;
      v[u][1] = in ? A.part_vol[nt + k] : 0.0;
#pragma unroll
      for (int s = 0; s < kMaxStages; ++s)
        v[u][2 + s] = (in && s < A.nstages) ? A.part_rate[int64_t(s) * nt + k] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      acc[0] += v[u][0];
      acc[1] = (v[u][1] > acc[1]) ? v[u][1] : acc[1];
#pragma unroll
      for (int s = 0; s < kMaxStages; ++s) acc[2 + s] += v[u][2 + s];
    }
  }
  // every value takes part (the rates of stages beyond nstages are zeros), so
  // that each level's cross-lane reads are all issued before one wait
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
  if (cword) ctrl_sh[threadIdx.x] = cw;
  __syncthreads();
  // thread q < kCtrlVals folds value q over the waves in wave order, all
  // its LDS reads issued before the first addition
  if (threadIdx.x < kCtrlVals) {
    const int q = threadIdx.x;
    double v[kCtrlWaves];
#pragma unroll
    for (int i = 0; i < kCtrlWaves; ++i) v[i] = (i * kWave < (int)blockDim.x) ? sh[i * kCtrlVals + q] : 0.0;
    double t = 0.0;
#pragma unroll
    for (int i = 0; i < kCtrlWaves; ++i) t = (q == 1) ? ((v[i] > t) ? v[i] : t) : t + v[i];
    sh[kCtrlWaves * kCtrlVals + q] = t;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < kCtrlVals; ++q) out[q] = sh[kCtrlWaves * kCtrlVals + q];
  }
}

// Controller (RungeKutta.cpp:154-167) + accept (RungeKutta.hpp:428-431) +
// landing rules and outer-step transitions (TemporalScheme.hpp:112-149).
__device__ __forceinline__ void ctrl_body(const CtrlArgs& A, double* sh, unsigned long long* ctrl_sh) {
  Ctrl* C = A.ctrl;
  // the control block arrives with the partial sums; thread 0 works on a
  // copy of it and writes it back once
  double sums[kCtrlVals];
  tile_sums(A, sh, sums, ctrl_sh);
  if (threadIdx.x != 0) return;
  Ctrl L;
  __builtin_memcpy(&L, ctrl_sh, sizeof(Ctrl));
  if (L.done) return;
  const double vol = sums[0] * A.area;
  double rate = 0.0;
#pragma unroll
  for (int s = 0; s < kMaxStages; ++s)
    if (s < A.nstages) rate += A.b[s] * sums[2 + s];
  const double cn_raw = sums[1];
  L.cn_bits = 0ull;
  L.cn = cn_raw;
  if (A.mode == 1) {
    L.volume = vol;
    *C = L;
    return;
  }
  const double cn = (L.cn_target == 1.0) ? cn_raw : cn_raw / L.cn_target;
  const double dt = L.dt;
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
  L.repeat = repeat ? 1 : 0;
  if (repeat) {
    L.n_rej++;
  } else {
    L.n_acc++;
    L.t_local += dt;
    L.cur ^= 1;
    L.volume = vol;
    L.srcvol += dt * rate * A.area;
  }
  L.dt = ndt;
  const double sd = L.step_dur, tl = L.t_local;
  if (tl >= sd) {
    L.step++;
    L.t_local = 0.0;
    // TemporalScheme.hpp:254: time_now = start_time + i * step_duration
    L.gstep = L.start_time + double(L.step) * sd;
    if (L.step >= L.nsteps) {
      L.done = 1;
    } else if (A.forcing && L.step < A.nforcing) {
      const double* f = A.forcing + 4 * L.step;
      L.q0 = f[0]; L.q1 = f[1]; L.hb0 = f[2]; L.hb1 = f[3];
    }
  } else if (tl + ndt > sd) {
    L.dt = sd - tl;
  } else if (tl + 2.0 * ndt > sd) {
    L.dt = (sd - tl) * 0.4;
  }
  *C = L;
}

__global__ __launch_bounds__(kCtrlThreads) void k_ctrl(const CtrlArgs A) {
  __shared__ double sh[(kCtrlWaves + 1) * kCtrlVals];
  __shared__ unsigned long long ctrl_sh[kCtrlWords];
  ctrl_body(A, sh, ctrl_sh);
}

// One block per member of a batch; the member count is read from device
// memory so that captured graphs survive changes of membership.
__global__ __launch_bounds__(kCtrlThreads) void k_ctrl_batch(const CtrlArgs* __restrict__ args,
                                                             const int* __restrict__ nmembers_p) {
  __shared__ double sh[(kCtrlWaves + 1) * kCtrlVals];
  __shared__ unsigned long long ctrl_sh[kCtrlWords];
  if ((int)blockIdx.x >= *nmembers_p) return;
  ctrl_body(args[blockIdx.x], sh, ctrl_sh);
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
