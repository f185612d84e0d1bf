// mfsv.hip -- libmfsv_hip.so: the C ABI declared in include/mfsv.h.
//
// A context is one catchment on one device: HIP stream, structure-of-arrays
// device fields, the device control block and the per-tile partial sums.
// All launches go to the context's stream; the only host synchronisations
// are the ones an entry point documents (mfsv_timestep returns the control
// number, mfsv_run checks the finished flag once per batch of timesteps).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/mfsv.h"
#include "sv_ctrl.hpp"

using namespace mfsv;

namespace {

constexpr int kVolBlocks = 256;

struct Buf {
  void* p = nullptr;
  size_t bytes = 0;
};

}  // namespace

struct mfsv_ctx {
  int dev = 0;
  hipStream_t stream = nullptr;
  int prec = MFSV_F64;
  size_t vsz = 8;
  int64_t nx = 0, ny = 0, nc = 0;
  double dx = 1, dy = 1, ox = 0, oy = 0;
  double zoff = 0.0;  // bed offset removed by the F32 form
  mfsv_params P{};
  int nstages = 2;
  double tab[kMaxStages + 1][kMaxStages + 1] = {};
  double start = 0, end = 0, step_dur = 60, dt0 = 1;
  // device memory
  Buf U[2][3];
  Buf Ul[2];               // FP32 form: low-order part of h of each state
  Buf acc[kMaxStages][3];  // accumulators for RK rows 1..nstages-1
  Buf zb, nfield, q0f, q1f, hb0f, hb1f, kf[3], brx, bry;
  Buf fxf, fyf, iratef, icapf, sb0f, sb1f, muf;  // further sources and the stage boundary
  Buf ndpf, dshf, ddpf;    // per-cell Manning n_deep, d_shallow, d_deep
  Buf dgf[4], dgnh, dgsf, dgzx, dgzy, stagef;  // output-only fields
  Buf ctrl, part_vol, part_rate, forcing, volp;
  int64_t nforcing = 0;
  Ctrl* hctrl = nullptr;  // pinned mirror
  int rows = 16;
  int64_t nstrips = 0, nrb = 0, ntiles = 0, nblocks = 0;
  int cur = 0;
  bool nfield_set = false;
  bool nanbed = false;  // the bed has coded-out (NaN) cells
  bool begun = false;   // mfsv_run_begin has initialised the control block
  bool ran = false;     // a timestep has run (its stage inputs are held)
  mfsv_batch* batch = nullptr;  // the batch this context is a member of
  uint64_t gen = 0;     // bumped by every configuration change
  hipGraphExec_t gexec = nullptr;
  hipGraph_t graph = nullptr;
  int64_t graph_steps = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::vector<hipEvent_t> sev;  // per-timestep stage brackets (timed runs)
  double last_ms = 0.0;
  double stage_ms = 0.0;
  int64_t stage_launches = 0;
  double volume0 = 0.0;
  std::string err;
  int64_t bytes = 0;
};

namespace {

int fail(mfsv_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

#define HIPCHK(c, expr)                                                          \
  do {                                                                          \
    hipError_t e_ = (expr);                                                     \
    if (e_ != hipSuccess)                                                       \
      return fail((c), MFSV_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

int alloc(mfsv_ctx* c, Buf& b, size_t bytes) {
  if (b.p && b.bytes == bytes) return MFSV_OK;
  if (b.p) {
    (void)hipFree(b.p);
    c->bytes -= (int64_t)b.bytes;
    b.p = nullptr;
  }
  if (hipMalloc(&b.p, bytes) != hipSuccess) {
    b.p = nullptr;
    return fail(c, MFSV_ERR_NOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
  }
  b.bytes = bytes;
  c->bytes += (int64_t)bytes;
  return MFSV_OK;
}

void release(Buf& b) {
  if (b.p) (void)hipFree(b.p);
  b.p = nullptr;
  b.bytes = 0;
}

// Every configuration change (parameters, scheme, tile height, bed,
// forcing table) passes through here: the captured graph is stale, and so
// are the launch tables of a batch holding this context (it compares gen).
void drop_graph(mfsv_ctx* c) {
  ++c->gen;
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  c->gexec = nullptr;
  c->graph = nullptr;
  c->graph_steps = 0;
}

void geometry(mfsv_ctx* c) {
  c->nstrips = (c->nx + kStrip - 1) / kStrip;
  c->nrb = (c->ny + c->rows - 1) / c->rows;
  c->ntiles = c->nstrips * c->nrb;
  c->nblocks = (c->ntiles + kWavesPerBlock - 1) / kWavesPerBlock;
}

int ensure_partials(mfsv_ctx* c) {
  int rc;
  if ((rc = alloc(c, c->part_vol, sizeof(double) * 2 * c->ntiles))) return rc;
  if ((rc = alloc(c, c->part_rate, sizeof(double) * c->ntiles * kMaxStages))) return rc;
  return MFSV_OK;
}

// Upload a host FP64 array as the context's value type.
int put(mfsv_ctx* c, Buf& b, const double* host, int64_t n, double offset) {
  int rc = alloc(c, b, c->vsz * n);
  if (rc) return rc;
  if (c->prec == MFSV_F64 && offset == 0.0) {
    HIPCHK(c, hipMemcpyAsync(b.p, host, 8 * n, hipMemcpyHostToDevice, c->stream));
  } else if (c->prec == MFSV_F64) {
    std::vector<double> t(n);
    for (int64_t i = 0; i < n; ++i) t[i] = host[i] - offset;
    HIPCHK(c, hipMemcpyAsync(b.p, t.data(), 8 * n, hipMemcpyHostToDevice, c->stream));
  } else {
    std::vector<float> t(n);
    for (int64_t i = 0; i < n; ++i) t[i] = (float)(host[i] - offset);
    HIPCHK(c, hipMemcpyAsync(b.p, t.data(), 4 * n, hipMemcpyHostToDevice, c->stream));
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return MFSV_OK;
}

int get(mfsv_ctx* c, const Buf& b, double* host, int64_t n, double offset, size_t esz) {
  if (!b.p) return fail(c, MFSV_ERR_STATE, "field not allocated");
  if (esz == 8) {
    HIPCHK(c, hipMemcpyAsync(host, b.p, 8 * n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (offset != 0.0)
      for (int64_t i = 0; i < n; ++i) host[i] += offset;
  } else {
    std::vector<float> t(n);
    HIPCHK(c, hipMemcpyAsync(t.data(), b.p, 4 * n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int64_t i = 0; i < n; ++i) host[i] = (double)t[i] + offset;
  }
  return MFSV_OK;
}

// FP32 form: a depth is held as hi + lo (two floats, |lo| <= ulp(hi)/2);
// upload splits the FP64 value, download returns hi + lo.
int put_split(mfsv_ctx* c, Buf& hi, Buf& lo, const double* host, int64_t n) {
  int rc;
  if ((rc = alloc(c, hi, 4 * n)) || (rc = alloc(c, lo, 4 * n))) return rc;
  std::vector<float> th(n), tl(n);
  for (int64_t i = 0; i < n; ++i) {
    th[i] = (float)host[i];
    tl[i] = (float)(host[i] - (double)th[i]);
  }
  HIPCHK(c, hipMemcpyAsync(hi.p, th.data(), 4 * n, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(lo.p, tl.data(), 4 * n, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return MFSV_OK;
}

int get_split(mfsv_ctx* c, const Buf& hi, const Buf& lo, double* host, int64_t n) {
  if (!hi.p || !lo.p) return fail(c, MFSV_ERR_STATE, "field not allocated");
  std::vector<float> th(n), tl(n);
  HIPCHK(c, hipMemcpyAsync(th.data(), hi.p, 4 * n, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(tl.data(), lo.p, 4 * n, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int64_t i = 0; i < n; ++i) host[i] = (double)th[i] + (double)tl[i];
  return MFSV_OK;
}

template <typename T>
T* tp(const Buf& b) { return static_cast<T*>(b.p); }

// The FP32 form advances Heun-2 in its convex form (a first stage that
// writes y1 only; see stage_tile's CONVEX).
bool convex_rk(const mfsv_ctx* c) {
  return c->prec == MFSV_F32 && c->nstages == 2 && c->tab[1][0] == 1.0 && c->tab[1][1] == 1.0 &&
         c->tab[2][0] == 0.0 && c->tab[2][1] == 0.5 && c->tab[2][2] == 0.5;
}

template <typename T>
StageArgs<T> stage_args(mfsv_ctx* c, int s, bool kout, bool branches, int* nacc) {
  StageArgs<T> A{};
  A.nx = (int32_t)c->nx; A.ny = (int32_t)c->ny;
  A.nrb = (int32_t)c->nrb; A.ntiles = (int32_t)c->ntiles; A.rows = c->rows;
  A.stage = s; A.nstages = c->nstages;
  A.last = (!kout && s == c->nstages - 1) ? 1 : 0;
  A.flags = c->P.flags;
  A.nanbed = c->nanbed ? 1 : 0;
  A.manning = c->P.manning;
  A.discharge = c->P.discharge;
  A.head = c->P.head;
  A.hx0 = (int32_t)std::max<int64_t>(0, std::min<int64_t>(c->P.head_x0, c->nx));
  A.hx1 = (int32_t)std::max<int64_t>(0, std::min<int64_t>(c->P.head_x1, c->nx));
  A.dx = T(c->dx); A.dy = T(c->dy);
  A.rdx = T(1.0 / c->dx); A.rdy = T(1.0 / c->dy);
  A.hdx = T(0.5 * c->dx); A.hdy = T(0.5 * c->dy);
  A.dxd = c->dx; A.dyd = c->dy;
  A.ox = c->ox; A.oy = c->oy;
  A.c_s = c->tab[s][0];
  A.rstep = 1.0 / c->step_dur;
  A.Uh[0] = tp<T>(c->U[0][0]); A.Uu[0] = tp<T>(c->U[0][1]); A.Uv[0] = tp<T>(c->U[0][2]);
  A.Uh[1] = tp<T>(c->U[1][0]); A.Uu[1] = tp<T>(c->U[1][1]); A.Uv[1] = tp<T>(c->U[1][2]);
  A.Uhl[0] = tp<T>(c->Ul[0]); A.Uhl[1] = tp<T>(c->Ul[1]);
  if (s > 0) {  // stage input: accumulator row s
    A.in_h = tp<T>(c->acc[s][0]); A.in_u = tp<T>(c->acc[s][1]); A.in_v = tp<T>(c->acc[s][2]);
  }
  A.zb = tp<T>(c->zb);
  A.nfield = c->nfield_set ? tp<T>(c->nfield) : nullptr;
  A.q0f = tp<T>(c->q0f); A.q1f = tp<T>(c->q1f);
  A.hb0f = tp<T>(c->hb0f); A.hb1f = tp<T>(c->hb1f);
  A.eloss = c->P.energy_loss; A.infil = c->P.infiltration;
  A.stageb = c->P.stage; A.eddy = c->P.eddy;
  A.mu_const = T(c->P.mu_const); A.mu_smag = T(c->P.mu_smag);
  A.area = T(c->dx * c->dy);
  A.fx = tp<T>(c->fxf); A.fy = tp<T>(c->fyf);
  A.irate = tp<T>(c->iratef); A.icap = tp<T>(c->icapf);
  A.sb0 = tp<T>(c->sb0f); A.sb1 = tp<T>(c->sb1f);
  A.mu = tp<T>(c->muf);
  A.mfull = (c->ndpf.p && c->dshf.p && c->ddpf.p) ? 1 : 0;
  A.ndp_f = tp<T>(c->ndpf); A.dsh_f = tp<T>(c->dshf); A.ddp_f = tp<T>(c->ddpf);
  A.n_sh = T(c->P.n_shallow); A.n_dp = T(c->P.n_deep);
  A.d_sh = T(c->P.d_shallow); A.d_dp = T(c->P.d_deep);
  A.h_dry = T(c->P.h_dry); A.u_max = T(c->P.u_max);
  A.storm = {c->P.storm_peak, c->P.storm_sigma, c->P.storm_cx0, c->P.storm_cy0,
             c->P.storm_vx, c->P.storm_vy};
  *nacc = 0;
  if (kout) {
    A.k_h = tp<T>(c->kf[0]); A.k_u = tp<T>(c->kf[1]); A.k_v = tp<T>(c->kf[2]);
  } else {
    // rows i = s+1 .. n receive k_s * a(i, s+1)  (RungeKutta.hpp:377-388);
    // the convex form's first stage feeds row 1 only, which the second
    // stage completes (E8/E9) as it reads it
    *nacc = (convex_rk(c) && s == 0) ? 1 : c->nstages - s;
    A.adj_in = (convex_rk(c) && s == 1) ? 1 : 0;
    for (int j = 0; j < *nacc; ++j) {
      const int i = s + 1 + j;
      A.acoef[j] = c->tab[i][s + 1];
      if (i < c->nstages) {
        A.acc_h[j] = tp<T>(c->acc[i][0]); A.acc_u[j] = tp<T>(c->acc[i][1]);
        A.acc_v[j] = tp<T>(c->acc[i][2]);
      }  // row n: null -> the new-state buffer
    }
  }
  if (branches) { A.br_x = tp<float>(c->brx); A.br_y = tp<float>(c->bry); }
  A.ctrl = tp<Ctrl>(c->ctrl);
  A.part_vol = tp<double>(c->part_vol);
  A.part_rate = tp<double>(c->part_rate) + (int64_t)s * c->ntiles;
  return A;
}

template <typename T, bool QUICK, bool NANBED, bool EXT, bool BND>
void launch_q(mfsv_ctx* c, const StageArgs<T>& A, int nacc, bool kout) {
  dim3 grid((unsigned)c->nblocks), block(kWave * kWavesPerBlock);
  const bool base = A.stage > 0;
#define MFSV_LAUNCH(N, K, B) \
  hipLaunchKernelGGL((k_stage<T, N, K, QUICK, NANBED, EXT, B, BND, false>), grid, block, 0, \
                     c->stream, A)
  if (kout) { MFSV_LAUNCH(0, true, false); return; }
  if constexpr (sizeof(T) == 4) {
    if (convex_rk(c)) {  // one row per stage
      if (base)
        hipLaunchKernelGGL((k_stage<T, 1, false, QUICK, NANBED, EXT, true, BND, true>), grid,
                           block, 0, c->stream, A);
      else
        hipLaunchKernelGGL((k_stage<T, 1, false, QUICK, NANBED, EXT, false, BND, true>), grid,
                           block, 0, c->stream, A);
      return;
    }
  }
  switch (nacc) {
    case 1: if (base) MFSV_LAUNCH(1, false, true); else MFSV_LAUNCH(1, false, false); break;
    case 2: if (base) MFSV_LAUNCH(2, false, true); else MFSV_LAUNCH(2, false, false); break;
    case 3: if (base) MFSV_LAUNCH(3, false, true); else MFSV_LAUNCH(3, false, false); break;
    default: MFSV_LAUNCH(4, false, false); break;  // four rows: only the first of four stages
  }
#undef MFSV_LAUNCH
}

// Further source terms or the stage boundary switched on?
bool has_ext(const mfsv_ctx* c) {
  return c->P.energy_loss || c->P.infiltration || c->P.stage || c->P.eddy ||
         (c->P.manning && c->ndpf.p && c->dshf.p && c->ddpf.p);
}

// Per-cell discharge or head endpoint fields in use (the stage kernels
// load them only when they are).
bool has_bnd(const mfsv_ctx* c) { return c->P.discharge == 2 || c->P.head == 2; }

template <typename T, bool QUICK, bool NANBED, bool EXT>
void launch_b(mfsv_ctx* c, const StageArgs<T>& A, int nacc, bool kout) {
  if (has_bnd(c)) launch_q<T, QUICK, NANBED, EXT, true>(c, A, nacc, kout);
  else launch_q<T, QUICK, NANBED, EXT, false>(c, A, nacc, kout);
}

template <typename T, bool QUICK, bool NANBED>
void launch_x(mfsv_ctx* c, const StageArgs<T>& A, int nacc, bool kout) {
  if (has_ext(c)) launch_b<T, QUICK, NANBED, true>(c, A, nacc, kout);
  else launch_b<T, QUICK, NANBED, false>(c, A, nacc, kout);
}

// True when x is a power of two: x/d == x * (1/d) exactly.
bool pow2(double d) {
  int e;
  return std::frexp(d, &e) == 0.5;
}

// Coded-out cells take the exact-quotient path with the reference's NaN
// rules; otherwise power-of-two spacing (and the FP32 form) takes the
// difference form, any other spacing exact quotients.
template <typename T>
void launch_typed(mfsv_ctx* c, int s, bool kout, bool branches) {
  int nacc = 0;
  const StageArgs<T> A = stage_args<T>(c, s, kout, branches, &nacc);
  if (c->nanbed) {
    launch_x<T, false, true>(c, A, nacc, kout);
  } else if constexpr (sizeof(T) == 4) {
    launch_x<T, true, false>(c, A, nacc, kout);
  } else {
    if (pow2(c->dx) && pow2(c->dy)) launch_x<T, true, false>(c, A, nacc, kout);
    else launch_x<T, false, false>(c, A, nacc, kout);
  }
}

// One side-effect-free derivative evaluation of stage s's input (or of the
// previous U0 when prev_u0) that fills the selected output fields.
template <typename T>
void launch_diag_typed(mfsv_ctx* c, int s, bool prev_u0, int what) {
  int nacc = 0;
  StageArgs<T> A = stage_args<T>(c, s, true, false, &nacc);
  if (prev_u0) {
    const int o = c->cur ^ 1;
    A.in_h = tp<T>(c->U[o][0]); A.in_u = tp<T>(c->U[o][1]); A.in_v = tp<T>(c->U[o][2]);
  }
  A.diag = 1;
  // the convex form's y1 (the last stage's input) is stored before E8/E9
  A.adj_in = (!prev_u0 && s == 1 && convex_rk(c)) ? 1 : 0;
  // the output fields depend on the state, the bed and the Manning
  // parameters only: no boundary or further source field is read
  A.discharge = 0; A.head = 0; A.stageb = 0;
  A.eloss = 0; A.infil = 0; A.eddy = 0;
  if (what & MFSV_DIAG_FLUXES) {
    A.dg_fh = tp<T>(c->dgf[0]); A.dg_fu = tp<T>(c->dgf[1]);
    A.dg_fv = tp<T>(c->dgf[2]); A.dg_fz = tp<T>(c->dgf[3]);
  }
  if ((what & MFSV_DIAG_MANNING) && c->P.manning) { A.dg_nh = tp<T>(c->dgnh); A.dg_sf = tp<T>(c->dgsf); }
  if (what & MFSV_DIAG_BED_SLOPES) { A.dg_zx = tp<T>(c->dgzx); A.dg_zy = tp<T>(c->dgzy); }
  if (c->nanbed) {
    launch_x<T, false, true>(c, A, 0, true);
  } else if constexpr (sizeof(T) == 4) {
    launch_x<T, true, false>(c, A, 0, true);
  } else {
    if (pow2(c->dx) && pow2(c->dy)) launch_x<T, true, false>(c, A, 0, true);
    else launch_x<T, false, false>(c, A, 0, true);
  }
}

template <typename T>
__global__ void k_stage_field(const T* __restrict__ zb, const T* __restrict__ h, int64_t n,
                              T* __restrict__ out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = zb[i] + h[i];
}

int launch_stage(mfsv_ctx* c, int s, bool kout, bool branches) {
  if (c->prec == MFSV_F64) launch_typed<double>(c, s, kout, branches);
  else launch_typed<float>(c, s, kout, branches);
  HIPCHK(c, hipGetLastError());
  return MFSV_OK;
}

CtrlArgs ctrl_args(const mfsv_ctx* c, int mode) {
  CtrlArgs A{};
  A.ctrl = tp<Ctrl>(c->ctrl);
  A.part_vol = tp<double>(c->part_vol);
  A.part_rate = tp<double>(c->part_rate);
  A.ntiles = c->ntiles;
  A.nstages = c->nstages;
  A.mode = mode;
  for (int s = 0; s < c->nstages; ++s) A.b[s] = c->tab[c->nstages][s + 1];
  A.area = c->dx * c->dy;
  A.forcing = c->nforcing ? tp<double>(c->forcing) : nullptr;
  A.nforcing = c->nforcing;
  return A;
}

int launch_ctrl(mfsv_ctx* c, int mode) {
  const CtrlArgs A = ctrl_args(c, mode);
  hipLaunchKernelGGL(k_ctrl, dim3(1), dim3(kCtrlThreads), 0, c->stream, A);
  HIPCHK(c, hipGetLastError());
  return MFSV_OK;
}

int launch_timestep(mfsv_ctx* c, int mode) {
  int rc;
  c->ran = true;
  for (int s = 0; s < c->nstages; ++s)
    if ((rc = launch_stage(c, s, false, false))) return rc;
  return launch_ctrl(c, mode);
}

int read_ctrl(mfsv_ctx* c) {
  HIPCHK(c, hipMemcpyAsync(c->hctrl, c->ctrl.p, sizeof(Ctrl), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return MFSV_OK;
}

template <typename T>
__global__ void k_volume(const T* __restrict__ h, const T* __restrict__ hl, int64_t n, double* part) {
  __shared__ double sh[256];
  const int64_t per = (n + gridDim.x - 1) / gridDim.x;
  const int64_t b0 = blockIdx.x * per;
  const int64_t b1 = (b0 + per < n) ? b0 + per : n;
  double s = 0.0;
  for (int64_t i = b0 + threadIdx.x; i < b1; i += blockDim.x)
    s += hl ? (double)h[i] + (double)hl[i] : (double)h[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ void k_sum(const double* part, int n, double scale, double* out) {
  __shared__ double sh[256];
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) s += part[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = sh[0] * scale;
}

int compute_volume(mfsv_ctx* c, int which, double* out) {
  int rc;
  if ((rc = alloc(c, c->volp, sizeof(double) * (kVolBlocks + 1)))) return rc;
  double* part = tp<double>(c->volp);
  if (c->prec == MFSV_F64)
    hipLaunchKernelGGL(k_volume<double>, dim3(kVolBlocks), dim3(256), 0, c->stream,
                       tp<double>(c->U[which][0]), (const double*)nullptr, c->nc, part);
  else
    hipLaunchKernelGGL(k_volume<float>, dim3(kVolBlocks), dim3(256), 0, c->stream,
                       tp<float>(c->U[which][0]), tp<float>(c->Ul[which]), c->nc, part);
  hipLaunchKernelGGL(k_sum, dim3(1), dim3(256), 0, c->stream, part, kVolBlocks, c->dx * c->dy,
                     part + kVolBlocks);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(out, part + kVolBlocks, 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return MFSV_OK;
}

// State.cpp:211-222: per-block maxima of (|u| + c)/dx + (|v| + c)/dy with
// c = sqrt(9.81 max(h, 0)) (quotients correctly rounded); the caller takes
// the maximum of the block maxima and multiplies by dt (max_i dt q_i ==
// dt max_i q_i for dt > 0).
// adj != 0: the state is a convex-form y1 stored before E8/E9 (flags,
// h_dry, u_max), which apply to its velocities first.
template <typename T>
__global__ void k_cn(const T* __restrict__ h, const T* __restrict__ u, const T* __restrict__ v,
                     int64_t n, T dx, T rdx, T dy, T rdy, double* part, int adj = 0,
                     int flags = 0, T h_dry = T(0), T u_max = T(0)) {
  __shared__ double sh[256];
  double m = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    T ui = u[i], vi = v[i];
    if (adj) {
      if ((flags & 0x80) && h[i] <= h_dry) { ui = T(0); vi = T(0); }
      if (flags & 0x100) { ui = clamp_t(ui, -u_max, u_max); vi = clamp_t(vi, -u_max, u_max); }
    }
    const T c = fsqrt_t(T(kG) * fmax_t(h[i], T(0)));
    const T q = div_c(fabs_t(ui) + c, dx, rdx) + div_c(fabs_t(vi) + c, dy, rdy);
    m = (double(q) > m) ? double(q) : m;
  }
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// ---- allocation-free control number ------------------------------------
// The block maxima of k_cn live in a file-scope device array instead of a
// context's buffer: one copy per GPU, held until the process ends.  Host code
// reaches it only through the address hipGetSymbolAddress returns for the
// current device (the array's name is a host-side stand-in there).
constexpr int kNoAllocThreads = 256;
constexpr int64_t kNoAllocCellsPerBlock = int64_t(kNoAllocThreads) * 64;
constexpr int kNoAllocParts = 4096;
constexpr int kNoAllocDevices = 64;
static_assert(int64_t(kNoAllocParts) * kNoAllocCellsPerBlock == MFSV_NOALLOC_MAX_CELLS,
              "the header's cell limit is what the scratch array holds");
static __device__ __align__(256) double scratch[kNoAllocParts];

// dt times the maximum of n block maxima (each >= 0: k_cn skips NaN), as the
// host does for mfsv_control_number (max is exact, so the grouping of the
// cells into blocks does not change the result).
__global__ void k_cn_final(const double* __restrict__ part, int n, double dt, double* __restrict__ out) {
  __shared__ double sh[kNoAllocThreads];
  double m = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) m = part[i] > m ? part[i] : m;
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int w = kNoAllocThreads / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = dt * sh[0];
}

// Device address of `scratch` on each device, looked up once per device with
// that device current.
std::mutex g_scratch_mu;
double* g_scratch[kNoAllocDevices] = {};

int scratch_on(int device, double** out) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[device]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return MFSV_ERR_HIP;
    g_scratch[device] = static_cast<double*>(p);
  }
  *out = g_scratch[device];
  return MFSV_OK;
}

// SpatialDerivativeOperationKernel with Minmod3 (SpatialDerivative.hpp:104-122,
// Minmod3.hpp:34-82) on a FP64 cell field: axis 0 takes (west, east) as
// (left, right), axis 1 the reference's (north, south).  A missing neighbour
// makes the difference one-sided.
__global__ void k_slope(const double* __restrict__ f, int64_t nx, int64_t ny, int axis, double d,
                        double rd, double* __restrict__ out) {
  const int64_t n = nx * ny;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t x = i % nx, y = i / nx;
    bool hl, hr;
    int64_t il, ir;
    if (axis == 0) {
      hl = x > 0; hr = x < nx - 1; il = i - 1; ir = i + 1;
    } else {
      hl = y < ny - 1; hr = y > 0; il = i + nx; ir = i - nx;
    }
    const double uc = f[i];
    const double ul = hl ? f[il] : uc, ur = hr ? f[ir] : uc;
    out[i] = minmod3(ul, hl, uc, hr, ur, d, rd);
  }
}

int set_ctrl(mfsv_ctx* c, double t_local, double dt, int flip, int set_time) {
  hipLaunchKernelGGL(k_set_ctrl, dim3(1), dim3(64), 0, c->stream, tp<Ctrl>(c->ctrl), t_local, dt,
                     flip, set_time);
  HIPCHK(c, hipGetLastError());
  return MFSV_OK;
}

bool valid_field(int f) { return f >= 0 && f < MFSV_FIELD_COUNT; }

}  // namespace

extern "C" {

const char* mfsv_version(void) { return "mfsv 0.1 (gfx950)"; }

int mfsv_device_count(int* count) {
  if (!count) return MFSV_ERR_ARG;
  if (hipGetDeviceCount(count) != hipSuccess) {
    *count = 0;
    return MFSV_ERR_HIP;
  }
  return MFSV_OK;
}

int mfsv_create(const mfsv_grid* g, int precision, int device, mfsv_ctx** out) {
  if (!g || !out || g->nx < 1 || g->ny < 1 || !(g->dx > 0) || !(g->dy > 0)) return MFSV_ERR_ARG;
  if (g->nx > (1 << 30) || g->ny > (1 << 30) ||
      g->nx * g->ny * (precision == MFSV_F64 ? 8 : 4) >= (int64_t(1) << 31))
    return MFSV_ERR_ARG;  // 32-bit byte offsets into each field in the kernels
  if (precision != MFSV_F64 && precision != MFSV_F32) return MFSV_ERR_ARG;
  *out = nullptr;
  mfsv_ctx* c = new (std::nothrow) mfsv_ctx();
  if (!c) return MFSV_ERR_NOMEM;
  c->dev = device;
  if (hipSetDevice(device) != hipSuccess) {
    delete c;
    return MFSV_ERR_HIP;
  }
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return MFSV_ERR_HIP;
  }
  c->prec = precision;
  c->vsz = precision == MFSV_F64 ? 8 : 4;
  c->nx = g->nx; c->ny = g->ny; c->nc = g->nx * g->ny;
  c->dx = g->dx; c->dy = g->dy; c->ox = g->origin_x; c->oy = g->origin_y;
  c->P.cn_target = 1.0;
  c->P.n_shallow = 0.03; c->P.n_deep = 0.3; c->P.d_shallow = 0.1; c->P.d_deep = 0.3;
  // heun2 by default (RungeKutta.cpp:69-72)
  c->nstages = 2;
  c->tab[1][0] = 1.0; c->tab[1][1] = 1.0;
  c->tab[2][1] = 0.5; c->tab[2][2] = 0.5;
  geometry(c);
  int rc = 0;
  const size_t fb = c->vsz * c->nc;
  for (int k = 0; k < 2 && !rc; ++k)
    for (int q = 0; q < 3 && !rc; ++q) rc = alloc(c, c->U[k][q], fb);
  for (int k = 0; k < 2 && !rc && precision == MFSV_F32; ++k) rc = alloc(c, c->Ul[k], fb);
  for (int i = 1; i < c->nstages && !rc; ++i)
    for (int q = 0; q < 3 && !rc; ++q) rc = alloc(c, c->acc[i][q], fb);
  if (!rc) rc = alloc(c, c->zb, fb);
  if (!rc) rc = alloc(c, c->ctrl, sizeof(Ctrl));
  if (!rc) rc = ensure_partials(c);
  if (!rc && hipHostMalloc((void**)&c->hctrl, sizeof(Ctrl)) != hipSuccess) rc = MFSV_ERR_HIP;
  if (!rc && (hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess))
    rc = MFSV_ERR_HIP;
  if (!rc) {
    for (int k = 0; k < 2; ++k)
      for (int q = 0; q < 3; ++q)
        if (hipMemsetAsync(c->U[k][q].p, 0, fb, c->stream) != hipSuccess) rc = MFSV_ERR_HIP;
    for (int k = 0; k < 2; ++k)
      if (c->Ul[k].p && hipMemsetAsync(c->Ul[k].p, 0, fb, c->stream) != hipSuccess) rc = MFSV_ERR_HIP;
    if (hipMemsetAsync(c->zb.p, 0, fb, c->stream) != hipSuccess) rc = MFSV_ERR_HIP;
    Ctrl h{};
    h.step_dur = 60.0; h.dt = 1.0; h.cn_target = 1.0; h.nsteps = 0;
    h.hb0 = NAN; h.hb1 = NAN;
    *c->hctrl = h;
    if (hipMemcpyAsync(c->ctrl.p, c->hctrl, sizeof(Ctrl), hipMemcpyHostToDevice, c->stream) != hipSuccess)
      rc = MFSV_ERR_HIP;
    if (hipStreamSynchronize(c->stream) != hipSuccess) rc = MFSV_ERR_HIP;
  }
  if (rc) {
    mfsv_destroy(c);
    return rc;
  }
  *out = c;
  return MFSV_OK;
}

int mfsv_destroy(mfsv_ctx* c) {
  if (!c) return MFSV_ERR_ARG;
  if (c->batch) (void)mfsv_batch_remove(c->batch, c);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  drop_graph(c);
  for (auto& a : c->U) for (auto& b : a) release(b);
  for (auto& b : c->Ul) release(b);
  for (auto& a : c->acc) for (auto& b : a) release(b);
  for (Buf* b : {&c->fxf, &c->fyf, &c->iratef, &c->icapf, &c->sb0f, &c->sb1f, &c->muf,
                 &c->ndpf, &c->dshf, &c->ddpf, &c->dgf[0], &c->dgf[1], &c->dgf[2], &c->dgf[3],
                 &c->dgnh, &c->dgsf, &c->dgzx, &c->dgzy, &c->stagef})
    release(*b);
  for (Buf* b : {&c->zb, &c->nfield, &c->q0f, &c->q1f, &c->hb0f, &c->hb1f, &c->kf[0], &c->kf[1],
                 &c->kf[2], &c->brx, &c->bry, &c->ctrl, &c->part_vol, &c->part_rate, &c->forcing,
                 &c->volp})
    release(*b);
  if (c->hctrl) (void)hipHostFree(c->hctrl);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  for (auto e : c->sev)
    if (e) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return MFSV_OK;
}

const char* mfsv_last_error(const mfsv_ctx* c) { return c ? c->err.c_str() : "null context"; }

int mfsv_synchronize(mfsv_ctx* c) {
  if (!c) return MFSV_ERR_ARG;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return MFSV_OK;
}

int mfsv_device_bytes(const mfsv_ctx* c, int64_t* bytes) {
  if (!c || !bytes) return MFSV_ERR_ARG;
  *bytes = c->bytes;
  return MFSV_OK;
}

int mfsv_set_params(mfsv_ctx* c, const mfsv_params* p) {
  if (!c || !p) return MFSV_ERR_ARG;
  if (p->discharge < 0 || p->discharge > 2 || p->head < 0 || p->head > 2)
    return fail(c, MFSV_ERR_ARG, "discharge/head mode out of range");
  if (p->energy_loss < 0 || p->energy_loss > 1 || p->infiltration < 0 || p->infiltration > 1 ||
      p->stage < 0 || p->stage > 1 || p->eddy < 0 || p->eddy > 2)
    return fail(c, MFSV_ERR_ARG, "source-term mode out of range");
  if (!(p->cn_target > 0)) return fail(c, MFSV_ERR_ARG, "cn_target must be positive");
  if ((p->flags & MFSV_E4_STORM) && !(p->storm_sigma > 0))
    return fail(c, MFSV_ERR_ARG, "storm sigma must be positive");
  c->P = *p;
  drop_graph(c);
  if (hipSetDevice(c->dev) != hipSuccess) return fail(c, MFSV_ERR_HIP, "hipSetDevice");
  // cn_target lives in the control block
  if (read_ctrl(c)) return MFSV_ERR_HIP;
  c->hctrl->cn_target = p->cn_target;
  HIPCHK(c, hipMemcpyAsync(c->ctrl.p, c->hctrl, sizeof(Ctrl), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return MFSV_OK;
}

int mfsv_set_scheme(mfsv_ctx* c, int nstages, const double* tab) {
  if (!c || !tab || nstages < 1 || nstages > kMaxStages) return MFSV_ERR_ARG;
  for (int i = 0; i <= kMaxStages; ++i)
    for (int j = 0; j <= kMaxStages; ++j)
      c->tab[i][j] = (i <= nstages && j <= nstages) ? tab[i * (nstages + 1) + j] : 0.0;
  c->nstages = nstages;
  drop_graph(c);
  int rc = 0;
  const size_t fb = c->vsz * c->nc;
  for (int i = 1; i < nstages && !rc; ++i)
    for (int q = 0; q < 3 && !rc; ++q) rc = alloc(c, c->acc[i][q], fb);
  for (int i = nstages; i < kMaxStages; ++i)
    for (int q = 0; q < 3; ++q) {
      if (c->acc[i][q].p) { c->bytes -= (int64_t)c->acc[i][q].bytes; release(c->acc[i][q]); }
    }
  return rc;
}

int mfsv_set_tile_rows(mfsv_ctx* c, int rows) {
  if (!c) return MFSV_ERR_ARG;
  // the storm's row factors of a tile come from a table of one row per lane
  if (rows < 1 || rows > kWave) return fail(c, MFSV_ERR_ARG, "tile rows must be in 1..64");
  if (rows == c->rows) return MFSV_OK;
  drop_graph(c);
  c->rows = rows;
  geometry(c);
  return ensure_partials(c);
}

int mfsv_get_tile_rows(const mfsv_ctx* c, int* rows) {
  if (!c || !rows) return MFSV_ERR_ARG;
  *rows = c->rows;
  return MFSV_OK;
}

int mfsv_set_time(mfsv_ctx* c, double start, double end, double step_duration, double initial_dt) {
  if (!c || !(step_duration > 0) || !(initial_dt > 0) || !(end >= start)) return MFSV_ERR_ARG;
  c->start = start; c->end = end; c->step_dur = step_duration; c->dt0 = initial_dt;
  return MFSV_OK;
}

int mfsv_start_new_step(mfsv_ctx* c, double time_now, double q0, double q1, double hb0, double hb1) {
  if (!c) return MFSV_ERR_ARG;
  if (c->batch) return fail(c, MFSV_ERR_STATE, "the context is a batch member: remove it first");
  if (read_ctrl(c)) return MFSV_ERR_HIP;
  Ctrl& h = *c->hctrl;
  h.q0 = q0; h.q1 = q1; h.hb0 = hb0; h.hb1 = hb1;
  h.gstep = time_now;
  h.step_dur = c->step_dur;
  h.start_time = c->start;
  h.done = 0;
  HIPCHK(c, hipMemcpyAsync(c->ctrl.p, c->hctrl, sizeof(Ctrl), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return MFSV_OK;
}

int mfsv_set_forcing_table(mfsv_ctx* c, int64_t nsteps, const double* table) {
  if (!c || nsteps < 0 || (nsteps && !table)) return MFSV_ERR_ARG;
  drop_graph(c);
  if (nsteps == 0) {
    c->nforcing = 0;
    return MFSV_OK;
  }
  int rc = alloc(c, c->forcing, sizeof(double) * 4 * nsteps);
  if (rc) return rc;
  HIPCHK(c, hipMemcpyAsync(c->forcing.p, table, sizeof(double) * 4 * nsteps, hipMemcpyHostToDevice,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->nforcing = nsteps;
  return MFSV_OK;
}

int mfsv_upload(mfsv_ctx* c, int field, const double* host, int64_t n) {
  if (!c || !host || !valid_field(field)) return MFSV_ERR_ARG;
  if (n != c->nc) return fail(c, MFSV_ERR_ARG, "count must equal nx*ny");
  if (hipSetDevice(c->dev) != hipSuccess) return fail(c, MFSV_ERR_HIP, "hipSetDevice");
  if (field <= MFSV_FIELD_V) {  // the device controller (or a batch) may have flipped U0
    int rc = read_ctrl(c);
    if (rc) return rc;
    c->cur = c->hctrl->cur;
  }
  switch (field) {
    case MFSV_FIELD_H:
      if (c->prec == MFSV_F32) return put_split(c, c->U[c->cur][0], c->Ul[c->cur], host, n);
      return put(c, c->U[c->cur][0], host, n, 0.0);
    case MFSV_FIELD_U: case MFSV_FIELD_V:
      return put(c, c->U[c->cur][field], host, n, 0.0);
    case MFSV_FIELD_ZB: {
      double off = 0.0;
      if (c->prec == MFSV_F32) {  // depth-based form: bed relative to its minimum
        double m = INFINITY;
        for (int64_t i = 0; i < n; ++i)
          if (host[i] < m) m = host[i];
        off = std::isfinite(m) ? m : 0.0;
      }
      c->zoff = off;
      c->nanbed = false;
      for (int64_t i = 0; i < n && !c->nanbed; ++i) c->nanbed = std::isnan(host[i]);
      drop_graph(c);
      return put(c, c->zb, host, n, off);
    }
    case MFSV_FIELD_MANNING_N: {
      int rc = put(c, c->nfield, host, n, 0.0);
      if (!rc) c->nfield_set = true;
      return rc;
    }
    case MFSV_FIELD_QBDY0: return put(c, c->q0f, host, n, 0.0);
    case MFSV_FIELD_QBDY1: return put(c, c->q1f, host, n, 0.0);
    case MFSV_FIELD_HBDY0: return put(c, c->hb0f, host, n, 0.0);
    case MFSV_FIELD_HBDY1: return put(c, c->hb1f, host, n, 0.0);
    case MFSV_FIELD_ELOSS_FX: return put(c, c->fxf, host, n, 0.0);
    case MFSV_FIELD_ELOSS_FY: return put(c, c->fyf, host, n, 0.0);
    case MFSV_FIELD_INFIL_RATE: return put(c, c->iratef, host, n, 0.0);
    case MFSV_FIELD_INFIL_CAP: return put(c, c->icapf, host, n, 0.0);
    // water levels, relative to the bed offset of the FP32 form
    case MFSV_FIELD_SBDY0: return put(c, c->sb0f, host, n, c->zoff);
    case MFSV_FIELD_SBDY1: return put(c, c->sb1f, host, n, c->zoff);
    case MFSV_FIELD_EDDY_MU: return put(c, c->muf, host, n, 0.0);
    case MFSV_FIELD_N_DEEP: drop_graph(c); return put(c, c->ndpf, host, n, 0.0);
    case MFSV_FIELD_D_SHALLOW: drop_graph(c); return put(c, c->dshf, host, n, 0.0);
    case MFSV_FIELD_D_DEEP: drop_graph(c); return put(c, c->ddpf, host, n, 0.0);
    default: return fail(c, MFSV_ERR_ARG, "field is not uploadable");
  }
}

int mfsv_download(mfsv_ctx* c, int field, double* host, int64_t n) {
  if (!c || !host || !valid_field(field)) return MFSV_ERR_ARG;
  if (hipSetDevice(c->dev) != hipSuccess) return fail(c, MFSV_ERR_HIP, "hipSetDevice");
  const int64_t nbx = (c->nx + 1) * c->ny, nby = c->nx * (c->ny + 1);
  if (field == MFSV_FIELD_BRANCH_X || field == MFSV_FIELD_BRANCH_Y) {
    const int64_t want = field == MFSV_FIELD_BRANCH_X ? nbx : nby;
    if (n != want) return fail(c, MFSV_ERR_ARG, "branch field count mismatch");
    const Buf& b = field == MFSV_FIELD_BRANCH_X ? c->brx : c->bry;
    return get(c, b, host, n, 0.0, 4);
  }
  if (field >= MFSV_FIELD_FLUX_H && field <= MFSV_FIELD_FLUX_Z) {
    if (n != nbx + nby) return fail(c, MFSV_ERR_ARG, "face field count must be (nx+1)*ny + nx*(ny+1)");
    return get(c, c->dgf[field - MFSV_FIELD_FLUX_H], host, n, 0.0, c->vsz);
  }
  if (n != c->nc) return fail(c, MFSV_ERR_ARG, "count must equal nx*ny");
  if (field <= MFSV_FIELD_V) {
    int rc = read_ctrl(c);
    if (rc) return rc;
    c->cur = c->hctrl->cur;
  }
  switch (field) {
    case MFSV_FIELD_H:
      if (c->prec == MFSV_F32) return get_split(c, c->U[c->cur][0], c->Ul[c->cur], host, n);
      return get(c, c->U[c->cur][0], host, n, 0.0, c->vsz);
    case MFSV_FIELD_U: case MFSV_FIELD_V:
      return get(c, c->U[c->cur][field], host, n, 0.0, c->vsz);
    case MFSV_FIELD_ZB: return get(c, c->zb, host, n, c->zoff, c->vsz);
    case MFSV_FIELD_MANNING_N: return get(c, c->nfield, host, n, 0.0, c->vsz);
    case MFSV_FIELD_QBDY0: return get(c, c->q0f, host, n, 0.0, c->vsz);
    case MFSV_FIELD_QBDY1: return get(c, c->q1f, host, n, 0.0, c->vsz);
    case MFSV_FIELD_HBDY0: return get(c, c->hb0f, host, n, 0.0, c->vsz);
    case MFSV_FIELD_HBDY1: return get(c, c->hb1f, host, n, 0.0, c->vsz);
    case MFSV_FIELD_ELOSS_FX: return get(c, c->fxf, host, n, 0.0, c->vsz);
    case MFSV_FIELD_ELOSS_FY: return get(c, c->fyf, host, n, 0.0, c->vsz);
    case MFSV_FIELD_INFIL_RATE: return get(c, c->iratef, host, n, 0.0, c->vsz);
    case MFSV_FIELD_INFIL_CAP: return get(c, c->icapf, host, n, 0.0, c->vsz);
    case MFSV_FIELD_SBDY0: return get(c, c->sb0f, host, n, c->zoff, c->vsz);
    case MFSV_FIELD_SBDY1: return get(c, c->sb1f, host, n, c->zoff, c->vsz);
    case MFSV_FIELD_EDDY_MU: return get(c, c->muf, host, n, 0.0, c->vsz);
    case MFSV_FIELD_DHDT: case MFSV_FIELD_DUDT: case MFSV_FIELD_DVDT:
      return get(c, c->kf[field - MFSV_FIELD_DHDT], host, n, 0.0, c->vsz);
    case MFSV_FIELD_N_DEEP: return get(c, c->ndpf, host, n, 0.0, c->vsz);
    case MFSV_FIELD_D_SHALLOW: return get(c, c->dshf, host, n, 0.0, c->vsz);
    case MFSV_FIELD_D_DEEP: return get(c, c->ddpf, host, n, 0.0, c->vsz);
    case MFSV_FIELD_MANNINGS_N: return get(c, c->dgnh, host, n, 0.0, c->vsz);
    case MFSV_FIELD_FRICTION_SLOPE: return get(c, c->dgsf, host, n, 0.0, c->vsz);
    case MFSV_FIELD_DZDX_BED: return get(c, c->dgzx, host, n, 0.0, c->vsz);
    case MFSV_FIELD_DZDY_BED: return get(c, c->dgzy, host, n, 0.0, c->vsz);
    case MFSV_FIELD_STAGE: return get(c, c->stagef, host, n, c->zoff, c->vsz);
    default: return MFSV_ERR_ARG;
  }
}

static int check_inputs(mfsv_ctx* c) {
  if (c->P.discharge == 2 && (!c->q0f.p || !c->q1f.p))
    return fail(c, MFSV_ERR_STATE, "per-cell discharge needs QBDY0 and QBDY1");
  if (c->P.energy_loss && (!c->fxf.p || !c->fyf.p))
    return fail(c, MFSV_ERR_STATE, "energy loss needs ELOSS_FX and ELOSS_FY");
  if (c->P.infiltration && (!c->iratef.p || !c->icapf.p))
    return fail(c, MFSV_ERR_STATE, "infiltration needs INFIL_RATE and INFIL_CAP");
  if (c->P.stage && (!c->sb0f.p || !c->sb1f.p))
    return fail(c, MFSV_ERR_STATE, "the stage boundary needs SBDY0 and SBDY1");
  if (c->P.eddy == 1 && !c->muf.p)
    return fail(c, MFSV_ERR_STATE, "constant eddy viscosity needs EDDY_MU");
  if (c->P.head == 2 && (!c->hb0f.p || !c->hb1f.p))
    return fail(c, MFSV_ERR_STATE, "per-cell head needs HBDY0 and HBDY1");
  const int mset = (c->ndpf.p ? 1 : 0) + (c->dshf.p ? 1 : 0) + (c->ddpf.p ? 1 : 0);
  if (mset != 0 && mset != 3)
    return fail(c, MFSV_ERR_STATE, "per-cell Manning needs N_DEEP, D_SHALLOW and D_DEEP together");
  if (mset == 3 && !c->nfield_set)
    return fail(c, MFSV_ERR_STATE, "per-cell Manning needs MANNING_N (the shallow n)");
  return MFSV_OK;
}

int mfsv_dudt(mfsv_ctx* c, double t_sub, double dt, int want_branches) {
  if (!c || !(dt > 0)) return MFSV_ERR_ARG;
  if (c->batch) return fail(c, MFSV_ERR_STATE, "the context is a batch member: remove it first");
  if (hipSetDevice(c->dev) != hipSuccess) return fail(c, MFSV_ERR_HIP, "hipSetDevice");
  int rc;
  if ((rc = check_inputs(c))) return rc;
  for (int q = 0; q < 3; ++q)
    if ((rc = alloc(c, c->kf[q], c->vsz * c->nc))) return rc;
  if (want_branches) {
    if ((rc = alloc(c, c->brx, 4 * (c->nx + 1) * c->ny))) return rc;
    if ((rc = alloc(c, c->bry, 4 * c->nx * (c->ny + 1)))) return rc;
  }
  // stage 0 has c = 0, so t_local = t_sub gives the stage time t_sub
  if ((rc = set_ctrl(c, t_sub, dt, 0, 1))) return rc;
  if ((rc = launch_stage(c, 0, true, want_branches != 0))) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return MFSV_OK;
}

int mfsv_diagnose(mfsv_ctx* c, int source, double t_sub, double dt, int what) {
  if (!c || (source != 0 && source != 1) || what < 0 || what > 7) return MFSV_ERR_ARG;
  if (source == 0 && !(dt > 0)) return MFSV_ERR_ARG;
  if (c->batch) return fail(c, MFSV_ERR_STATE, "the context is a batch member: remove it first");
  if (hipSetDevice(c->dev) != hipSuccess) return fail(c, MFSV_ERR_HIP, "hipSetDevice");
  if (source == 1 && !c->ran)
    return fail(c, MFSV_ERR_STATE, "no timestep has run: there is no last stage input");
  if (!c->zb.p || !c->U[0][0].p) return fail(c, MFSV_ERR_STATE, "bed and state must be uploaded");
  const int mset = (c->ndpf.p ? 1 : 0) + (c->dshf.p ? 1 : 0) + (c->ddpf.p ? 1 : 0);
  if (mset != 0 && (mset != 3 || !c->nfield_set))
    return fail(c, MFSV_ERR_STATE, "per-cell Manning needs MANNING_N, N_DEEP, D_SHALLOW and D_DEEP");
  int rc;
  if ((rc = read_ctrl(c))) return rc;
  c->cur = c->hctrl->cur;
  const double keep_t = c->hctrl->t_local, keep_dt = c->hctrl->dt;
  const int64_t nf = (c->nx + 1) * c->ny + c->nx * (c->ny + 1);
  for (int q = 0; q < 3; ++q)
    if ((rc = alloc(c, c->kf[q], c->vsz * c->nc))) return rc;
  if (what & MFSV_DIAG_FLUXES)
    for (int q = 0; q < 4; ++q)
      if ((rc = alloc(c, c->dgf[q], c->vsz * nf))) return rc;
  if ((what & MFSV_DIAG_MANNING) && c->P.manning)
    if ((rc = alloc(c, c->dgnh, c->vsz * c->nc)) || (rc = alloc(c, c->dgsf, c->vsz * c->nc))) return rc;
  if (what & MFSV_DIAG_BED_SLOPES)
    if ((rc = alloc(c, c->dgzx, c->vsz * c->nc)) || (rc = alloc(c, c->dgzy, c->vsz * c->nc))) return rc;
  // source 1: the last stage read accumulator row n-1 (row 0 = the U0 that
  // the accept replaced); its fluxes, n and S_f depend on the state only
  const int s = source == 1 ? c->nstages - 1 : 0;
  const bool prev = source == 1 && s == 0;
  if ((rc = set_ctrl(c, source == 1 ? 0.0 : t_sub, source == 1 ? 1.0 : dt, 0, 1))) return rc;
  if (c->prec == MFSV_F64) launch_diag_typed<double>(c, s, prev, what);
  else launch_diag_typed<float>(c, s, prev, what);
  HIPCHK(c, hipGetLastError());
  // leave the control block's time as it was (a device-controlled run may
  // continue after this call)
  if ((rc = set_ctrl(c, keep_t, keep_dt, 0, 1))) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return MFSV_OK;
}

int mfsv_control_number(mfsv_ctx* c, int state, double dt, double* cn) {
  if (!c || !cn || !(dt > 0) || state < 0 || state > c->nstages) return MFSV_ERR_ARG;
  if (hipSetDevice(c->dev) != hipSuccess) return fail(c, MFSV_ERR_HIP, "hipSetDevice");
  if (!c->U[0][0].p) return fail(c, MFSV_ERR_STATE, "the state must be uploaded");
  int rc;
  if ((rc = read_ctrl(c))) return rc;
  c->cur = c->hctrl->cur;
  const Buf* f;  // U_0, the rows U_1..U_{n-1}, U_n (the result buffer)
  if (state == 0) f = c->U[c->cur];
  else if (state == c->nstages) f = c->U[c->cur ^ 1];
  else f = c->acc[state];
  if (!f[0].p) return fail(c, MFSV_ERR_STATE, "that state holds no values");
  if ((rc = alloc(c, c->volp, sizeof(double) * (kVolBlocks + 1)))) return rc;
  double* part = tp<double>(c->volp);
  if (c->prec == MFSV_F64)
    hipLaunchKernelGGL(k_cn<double>, dim3(kVolBlocks), dim3(256), 0, c->stream, tp<double>(f[0]),
                       tp<double>(f[1]), tp<double>(f[2]), c->nc, c->dx, 1.0 / c->dx, c->dy,
                       1.0 / c->dy, part);
  else
    hipLaunchKernelGGL(k_cn<float>, dim3(kVolBlocks), dim3(256), 0, c->stream, tp<float>(f[0]),
                       tp<float>(f[1]), tp<float>(f[2]), c->nc, float(c->dx), float(1.0 / c->dx),
                       float(c->dy), float(1.0 / c->dy), part,
                       (state > 0 && state < c->nstages && convex_rk(c)) ? 1 : 0, c->P.flags,
                       float(c->P.h_dry), float(c->P.u_max));
  HIPCHK(c, hipGetLastError());
  std::vector<double> hp(kVolBlocks);
  HIPCHK(c, hipMemcpyAsync(hp.data(), part, sizeof(double) * kVolBlocks, hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  double m = 0.0;
  for (double x : hp) m = x > m ? x : m;
  *cn = dt * m;
  return MFSV_OK;
}

int mfsv_control_number_noalloc(int precision, int device, int64_t nx, int64_t ny, double dx,
                                double dy, const void* h, const void* u, const void* v, double dt,
                                double* cn_out, void* stream) {
  if (precision != MFSV_F64 && precision != MFSV_F32) return MFSV_ERR_ARG;
  if (device < 0 || device >= kNoAllocDevices || nx < 1 || ny < 1) return MFSV_ERR_ARG;
  if (nx > MFSV_NOALLOC_MAX_CELLS || ny > MFSV_NOALLOC_MAX_CELLS / nx) return MFSV_ERR_ARG;
  if (!(dx > 0) || !(dy > 0) || !(dt > 0) || !h || !u || !v || !cn_out) return MFSV_ERR_ARG;
  if (hipSetDevice(device) != hipSuccess) return MFSV_ERR_HIP;
  double* part = nullptr;
  int rc = scratch_on(device, &part);
  if (rc) return rc;
  const int64_t n = nx * ny;
  const int blocks = (int)((n + kNoAllocCellsPerBlock - 1) / kNoAllocCellsPerBlock);
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (precision == MFSV_F64)
    hipLaunchKernelGGL(k_cn<double>, dim3(blocks), dim3(kNoAllocThreads), 0, s,
                       static_cast<const double*>(h), static_cast<const double*>(u),
                       static_cast<const double*>(v), n, dx, 1.0 / dx, dy, 1.0 / dy, part);
  else
    hipLaunchKernelGGL(k_cn<float>, dim3(blocks), dim3(kNoAllocThreads), 0, s,
                       static_cast<const float*>(h), static_cast<const float*>(u),
                       static_cast<const float*>(v), n, float(dx), float(1.0 / dx), float(dy),
                       float(1.0 / dy), part);
  hipLaunchKernelGGL(k_cn_final, dim3(1), dim3(kNoAllocThreads), 0, s, (const double*)part, blocks,
                     dt, cn_out);
  return hipGetLastError() == hipSuccess ? MFSV_OK : MFSV_ERR_HIP;
}

int mfsv_spatial_derivative(mfsv_ctx* c, int axis, const double* in, double* out, int64_t count) {
  if (!c || !in || !out || (axis != 0 && axis != 1)) return MFSV_ERR_ARG;
  if (count != c->nc) return fail(c, MFSV_ERR_ARG, "count must equal nx*ny");
  if (hipSetDevice(c->dev) != hipSuccess) return fail(c, MFSV_ERR_HIP, "hipSetDevice");
  Buf a, b;
  int rc;
  if ((rc = alloc(c, a, 8 * c->nc)) || (rc = alloc(c, b, 8 * c->nc))) {
    release(a); release(b);
    return rc;
  }
  const double d = axis == 0 ? c->dx : c->dy;
  const unsigned blocks = (unsigned)std::min<int64_t>((c->nc + 255) / 256, 4096);
  hipError_t e = hipMemcpyAsync(a.p, in, 8 * c->nc, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_slope, dim3(blocks), dim3(256), 0, c->stream, tp<double>(a), c->nx, c->ny,
                       axis, d, 1.0 / d, tp<double>(b));
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(out, b.p, 8 * c->nc, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  c->bytes -= (int64_t)(a.bytes + b.bytes);
  release(a);
  release(b);
  if (e != hipSuccess) return fail(c, MFSV_ERR_HIP, std::string("spatial derivative: ") + hipGetErrorString(e));
  return MFSV_OK;
}

int mfsv_update_stage_field(mfsv_ctx* c) {
  if (!c) return MFSV_ERR_ARG;
  if (c->batch) return fail(c, MFSV_ERR_STATE, "the context is a batch member: remove it first");
  if (hipSetDevice(c->dev) != hipSuccess) return fail(c, MFSV_ERR_HIP, "hipSetDevice");
  if (!c->zb.p || !c->U[0][0].p) return fail(c, MFSV_ERR_STATE, "bed and state must be uploaded");
  int rc;
  if ((rc = read_ctrl(c))) return rc;
  c->cur = c->hctrl->cur;
  if ((rc = alloc(c, c->stagef, c->vsz * c->nc))) return rc;
  const unsigned blocks = (unsigned)std::min<int64_t>((c->nc + 255) / 256, 4096);
  if (c->prec == MFSV_F64)
    hipLaunchKernelGGL(k_stage_field<double>, dim3(blocks), dim3(256), 0, c->stream,
                       tp<double>(c->zb), tp<double>(c->U[c->cur][0]), c->nc, tp<double>(c->stagef));
  else
    hipLaunchKernelGGL(k_stage_field<float>, dim3(blocks), dim3(256), 0, c->stream,
                       tp<float>(c->zb), tp<float>(c->U[c->cur][0]), c->nc, tp<float>(c->stagef));
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return MFSV_OK;
}

int mfsv_timestep(mfsv_ctx* c, double t_local, double dt, double* cn) {
  if (!c || !(dt > 0) || !cn) return MFSV_ERR_ARG;
  if (c->batch) return fail(c, MFSV_ERR_STATE, "the context is a batch member: remove it first");
  if (hipSetDevice(c->dev) != hipSuccess) return fail(c, MFSV_ERR_HIP, "hipSetDevice");
  int rc;
  if ((rc = check_inputs(c))) return rc;
  if ((rc = set_ctrl(c, t_local, dt, 0, 1))) return rc;
  if ((rc = launch_timestep(c, 1))) return rc;
  if ((rc = read_ctrl(c))) return rc;
  *cn = c->hctrl->cn;
  return MFSV_OK;
}

int mfsv_accept(mfsv_ctx* c) {
  if (!c) return MFSV_ERR_ARG;
  if (c->batch) return fail(c, MFSV_ERR_STATE, "the context is a batch member: remove it first");
  int rc = set_ctrl(c, 0, 0, 1, 0);
  if (rc) return rc;
  c->cur ^= 1;
  return MFSV_OK;
}

int mfsv_volume(mfsv_ctx* c, double* v) {
  if (!c || !v) return MFSV_ERR_ARG;
  if (hipSetDevice(c->dev) != hipSuccess) return fail(c, MFSV_ERR_HIP, "hipSetDevice");
  return compute_volume(c, c->cur, v);
}

// TimeParameters.cpp:40-58
static int64_t num_steps(const mfsv_ctx* c) {
  const double eps = 1e-6;
  const double dur = c->end - c->start;
  const double frac = dur / c->step_dur;
  const int64_t n = (int64_t)frac;
  if (c->step_dur * n >= dur - eps && c->step_dur * n <= dur + eps) return n;
  return n + 1;
}

int mfsv_run_begin(mfsv_ctx* c) {
  if (!c) return MFSV_ERR_ARG;
  if (c->batch) return fail(c, MFSV_ERR_STATE, "the context is a batch member: remove it first");
  if (hipSetDevice(c->dev) != hipSuccess) return fail(c, MFSV_ERR_HIP, "hipSetDevice");
  int rc;
  if ((rc = check_inputs(c))) return rc;
  const int64_t ns = num_steps(c);
  if ((c->P.discharge == 1 || c->P.head == 1) && c->nforcing < ns)
    return fail(c, MFSV_ERR_STATE, "scalar discharge/head needs a forcing table row per outer step");
  if ((rc = compute_volume(c, c->cur, &c->volume0))) return rc;
  if ((rc = read_ctrl(c))) return rc;
  Ctrl& h = *c->hctrl;
  h.t_local = 0.0;
  h.dt = c->dt0;
  h.step_dur = c->step_dur;
  h.start_time = c->start;
  h.gstep = c->start;
  h.cn = 0.0;
  h.cn_target = c->P.cn_target;
  h.volume = c->volume0;
  h.srcvol = 0.0;
  h.step = 0;
  h.nsteps = ns;
  h.n_acc = 0; h.n_rej = 0;
  h.cn_bits = 0;
  h.cur = c->cur;
  h.done = ns <= 0 ? 1 : 0;
  h.repeat = 0;
  c->begun = true;
  if (c->nforcing) {
    std::vector<double> f0(4);
    HIPCHK(c, hipMemcpy(f0.data(), c->forcing.p, 32, hipMemcpyDeviceToHost));
    h.q0 = f0[0]; h.q1 = f0[1]; h.hb0 = f0[2]; h.hb1 = f0[3];
  } else {
    h.q0 = 0; h.q1 = 0; h.hb0 = NAN; h.hb1 = NAN;
  }
  HIPCHK(c, hipMemcpyAsync(c->ctrl.p, c->hctrl, sizeof(Ctrl), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return MFSV_OK;
}

static int build_graph(mfsv_ctx* c, int64_t steps) {
  drop_graph(c);
  HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  int rc = 0;
  for (int64_t k = 0; k < steps && !rc; ++k) rc = launch_timestep(c, 0);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(c->stream, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return fail(c, MFSV_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  c->graph = g;
  HIPCHK(c, hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0));
  c->graph_steps = steps;
  return MFSV_OK;
}

int mfsv_run(mfsv_ctx* c, int64_t max_timesteps, int use_graph) {
  if (!c || max_timesteps < 0) return MFSV_ERR_ARG;
  if (c->batch) return fail(c, MFSV_ERR_STATE, "the context is a batch member: remove it first");
  if (hipSetDevice(c->dev) != hipSuccess) return fail(c, MFSV_ERR_HIP, "hipSetDevice");
  int rc;
  const int64_t batch = 32;
  const bool timed = use_graph == 2;
  if (use_graph == 1 && c->graph_steps != batch)
    if ((rc = build_graph(c, batch))) return rc;
  if (timed && c->sev.empty()) {
    c->sev.resize(2 * batch, nullptr);
    for (auto& e : c->sev) HIPCHK(c, hipEventCreate(&e));
  }
  c->stage_ms = 0.0;
  c->stage_launches = 0;
  c->ran = c->ran || max_timesteps > 0;
  HIPCHK(c, hipEventRecord(c->ev0, c->stream));
  int64_t left = max_timesteps;
  while (left > 0) {
    const int64_t n = left < batch ? left : batch;
    if (use_graph == 1 && n == batch) {
      HIPCHK(c, hipGraphLaunch(c->gexec, c->stream));
    } else {
      for (int64_t k = 0; k < n; ++k) {
        if (timed) HIPCHK(c, hipEventRecord(c->sev[2 * k], c->stream));
        for (int s = 0; s < c->nstages; ++s)
          if ((rc = launch_stage(c, s, false, false))) return rc;
        if (timed) HIPCHK(c, hipEventRecord(c->sev[2 * k + 1], c->stream));
        if ((rc = launch_ctrl(c, 0))) return rc;
      }
    }
    left -= n;
    // the done flag decides whether another chunk follows; after the last
    // chunk the control block is read once below, with the closing event
    if (left > 0 || timed)
      if ((rc = read_ctrl(c))) return rc;
    if (timed) {
      for (int64_t k = 0; k < n; ++k) {
        float m = 0.f;
        HIPCHK(c, hipEventElapsedTime(&m, c->sev[2 * k], c->sev[2 * k + 1]));
        c->stage_ms += m;
      }
      c->stage_launches += n * c->nstages;
    }
    if (c->hctrl->done) break;
  }
  // the control block's copy goes in the stream ahead of the closing event,
  // so one wait covers both
  HIPCHK(c, hipMemcpyAsync(c->hctrl, c->ctrl.p, sizeof(Ctrl), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipEventRecord(c->ev1, c->stream));
  HIPCHK(c, hipEventSynchronize(c->ev1));
  float ms = 0.f;
  HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
  c->last_ms = ms;
  c->cur = c->hctrl->cur;
  return MFSV_OK;
}

int mfsv_run_step(mfsv_ctx* c, int use_graph) {
  if (!c || use_graph < 0 || use_graph > 2) return MFSV_ERR_ARG;
  if (c->batch) return fail(c, MFSV_ERR_STATE, "the context is a batch member: remove it first");
  if (!c->begun) return fail(c, MFSV_ERR_STATE, "mfsv_run_begin must come first");
  if (hipSetDevice(c->dev) != hipSuccess) return fail(c, MFSV_ERR_HIP, "hipSetDevice");
  int rc;
  if ((rc = read_ctrl(c))) return rc;
  const int64_t total = c->hctrl->nsteps;
  if (c->hctrl->step >= total) return MFSV_OK;
  // the device controller finishes the run at the end of this outer step
  c->hctrl->nsteps = c->hctrl->step + 1;
  c->hctrl->done = 0;
  HIPCHK(c, hipMemcpyAsync(c->ctrl.p, c->hctrl, sizeof(Ctrl), hipMemcpyHostToDevice, c->stream));
  rc = mfsv_run(c, int64_t(1) << 50, use_graph);
  const int rc2 = read_ctrl(c);
  if (rc) return rc;
  if (rc2) return rc2;
  c->hctrl->nsteps = total;
  c->hctrl->done = c->hctrl->step >= total ? 1 : 0;
  HIPCHK(c, hipMemcpyAsync(c->ctrl.p, c->hctrl, sizeof(Ctrl), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return MFSV_OK;
}

int mfsv_get_stats(mfsv_ctx* c, mfsv_stats* s) {
  if (!c || !s) return MFSV_ERR_ARG;
  int rc = read_ctrl(c);
  if (rc) return rc;
  const Ctrl& h = *c->hctrl;
  s->time = h.gstep + h.t_local;
  s->t_local = h.t_local;
  s->dt = h.dt;
  s->cn = h.cn;
  s->volume = h.volume;
  s->source_volume = h.srcvol;
  s->volume0 = c->volume0;
  s->step = h.step;
  s->nsteps = h.nsteps;
  s->accepted = h.n_acc;
  s->rejected = h.n_rej;
  s->done = h.done;
  s->current = h.cur;
  c->cur = h.cur;
  return MFSV_OK;
}

int mfsv_last_run_ms(mfsv_ctx* c, double* ms) {
  if (!c || !ms) return MFSV_ERR_ARG;
  *ms = c->last_ms;
  return MFSV_OK;
}

int mfsv_last_stage_ms(mfsv_ctx* c, double* ms, int64_t* launches) {
  if (!c || !ms || !launches) return MFSV_ERR_ARG;
  *ms = c->stage_ms;
  *launches = c->stage_launches;
  return MFSV_OK;
}

}  // extern "C"

#include "mfsv_batch.hpp"
