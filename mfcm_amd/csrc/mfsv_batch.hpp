// mfsv_batch.hpp -- the batch API of include/mfsv.h (many catchments per
// launch).  Part of the mfsv.hip translation unit: it uses that file's
// context internals.
//
// A batch references member contexts; each member keeps its own fields,
// control block and tile height.  One timestep of the batch is one
// k_stage_batch launch per RK stage over the concatenated tile table of all
// unfinished members, then one k_ctrl_batch launch (a block per member).
// Members that finish are dropped from the table at the next check.

struct mfsv_batch {
  int dev = 0;
  hipStream_t stream = nullptr;
  int prec = MFSV_F64;
  int nstages = 0;
  double tab[kMaxStages + 1][kMaxStages + 1] = {};
  std::vector<mfsv_ctx*> members;  // in insertion order
  std::vector<mfsv_ctx*> active;   // members in the current tables
  std::vector<uint64_t> gens;      // active members' configuration generations
  Buf margs[kMaxStages];           // StageArgs<T>[cap] per stage
  Buf cargs;                       // CtrlArgs[cap]
  Buf tiles;                       // int2[tile_cap]
  Buf counts;                      // int[2]: tiles, members
  Buf done;                        // int[cap]
  int cap = 0, tile_cap = 0;
  int* hdone = nullptr;            // pinned
  int ntiles = 0;
  bool dirty = true;
  bool quick = true;
  bool any_nanbed = false;         // a member needs the general stage path
  bool convex = false;             // FP32 Heun-2: the convex form (convex_rk)
  int grid_tiles = 0;              // table length the stage launches' grid was sized for
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  int64_t graph_steps = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double last_ms = 0.0;
  int64_t bytes = 0;
  std::string err;
};

namespace {

int bfail(mfsv_batch* b, int code, const std::string& msg) {
  if (b) b->err = msg;
  return code;
}

#define BHIPCHK(b, expr)                                                           \
  do {                                                                            \
    hipError_t e_ = (expr);                                                       \
    if (e_ != hipSuccess)                                                         \
      return bfail((b), MFSV_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

int balloc(mfsv_batch* b, Buf& buf, size_t bytes) {
  if (buf.p && buf.bytes >= bytes) return MFSV_OK;
  if (buf.p) {
    (void)hipFree(buf.p);
    b->bytes -= (int64_t)buf.bytes;
    buf.p = nullptr;
  }
  if (hipMalloc(&buf.p, bytes) != hipSuccess) {
    buf.p = nullptr;
    buf.bytes = 0;
    return bfail(b, MFSV_ERR_NOMEM, "hipMalloc of batch tables failed");
  }
  buf.bytes = bytes;
  b->bytes += (int64_t)bytes;
  return MFSV_OK;
}

void bdrop_graph(mfsv_batch* b) {
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->gexec = nullptr;
  b->graph = nullptr;
  b->graph_steps = 0;
}

__global__ void k_gather_done(const CtrlArgs* __restrict__ args, const int* __restrict__ n_p,
                              int* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < *n_p) out[i] = args[i].ctrl->done;
}

template <typename T>
int upload_tables(mfsv_batch* b) {
  const int n = (int)b->active.size();
  int rc;
  if (n > b->cap) {
    const int cap = std::max(n, 2 * b->cap);
    for (int s = 0; s < b->nstages; ++s)
      if ((rc = balloc(b, b->margs[s], sizeof(StageArgs<T>) * cap))) return rc;
    if ((rc = balloc(b, b->cargs, sizeof(CtrlArgs) * cap))) return rc;
    if ((rc = balloc(b, b->done, sizeof(int) * cap))) return rc;
    if (b->hdone) (void)hipHostFree(b->hdone);
    b->hdone = nullptr;
    if (hipHostMalloc((void**)&b->hdone, sizeof(int) * cap) != hipSuccess)
      return bfail(b, MFSV_ERR_HIP, "hipHostMalloc");
    b->cap = cap;
    bdrop_graph(b);  // table addresses changed
  }
  int nt = 0;
  for (auto* c : b->active) nt += (int)c->ntiles;
  if (nt > b->tile_cap) {
    const int tc = std::max(nt, 2 * b->tile_cap);
    if ((rc = balloc(b, b->tiles, sizeof(int2) * tc))) return rc;
    b->tile_cap = tc;
    bdrop_graph(b);
  }
  if ((rc = balloc(b, b->counts, sizeof(int) * 2))) return rc;
  std::vector<int2> tiles;
  tiles.reserve(nt);
  bool quick = true;  // exact division stays exact for power-of-two spacing
  if (sizeof(T) == 8)
    for (auto* c : b->active) quick = quick && pow2(c->dx) && pow2(c->dy);
  // a member needs the general path: coded-out cells, further sources, or
  // per-cell discharge / head endpoint fields
  bool any_nanbed = false;
  for (auto* c : b->active)
    any_nanbed = any_nanbed || c->nanbed || has_ext(c) || has_bnd(c);
  if (quick != b->quick || any_nanbed != b->any_nanbed) bdrop_graph(b);
  b->quick = quick;
  b->any_nanbed = any_nanbed;
  // the members share the scheme and precision: the first one decides
  const bool cvx = !b->active.empty() && convex_rk(b->active[0]);
  if (cvx != b->convex) bdrop_graph(b);
  b->convex = cvx;
  for (int s = 0; s < b->nstages; ++s) {
    std::vector<StageArgs<T>> a(n);
    for (int j = 0; j < n; ++j) {
      int nacc = 0;
      a[j] = stage_args<T>(b->active[j], s, false, false, &nacc);
    }
    if (n) BHIPCHK(b, hipMemcpyAsync(b->margs[s].p, a.data(), sizeof(StageArgs<T>) * n,
                                     hipMemcpyHostToDevice, b->stream));
  }
  std::vector<CtrlArgs> ca(n);
  for (int j = 0; j < n; ++j) {
    ca[j] = ctrl_args(b->active[j], 0);
    for (int t = 0; t < (int)b->active[j]->ntiles; ++t) tiles.push_back(make_int2(j, t));
  }
  if (n) {
    BHIPCHK(b, hipMemcpyAsync(b->cargs.p, ca.data(), sizeof(CtrlArgs) * n, hipMemcpyHostToDevice,
                              b->stream));
    BHIPCHK(b, hipMemcpyAsync(b->tiles.p, tiles.data(), sizeof(int2) * nt, hipMemcpyHostToDevice,
                              b->stream));
  }
  const int cnt[2] = {nt, n};
  BHIPCHK(b, hipMemcpyAsync(b->counts.p, cnt, sizeof(cnt), hipMemcpyHostToDevice, b->stream));
  BHIPCHK(b, hipStreamSynchronize(b->stream));
  b->ntiles = nt;
  // the stage launches' grid follows the table: a captured timestep with a
  // grid for another table length is captured again
  if (nt != b->grid_tiles) {
    bdrop_graph(b);
    b->grid_tiles = nt;
  }
  b->gens.clear();
  for (auto* c : b->active) b->gens.push_back(c->gen);
  b->dirty = false;
  return MFSV_OK;
}

template <typename T, bool QUICK, bool GENERAL>
void launch_batch_stage(mfsv_batch* b, int s) {
  const int* cnt = static_cast<const int*>(b->counts.p);
  const StageArgs<T>* m = static_cast<const StageArgs<T>*>(b->margs[s].p);
  const int2* t = static_cast<const int2*>(b->tiles.p);
  // a wave per table entry (the hardware hands the next entry to whichever
  // SIMD frees up first; the tiles of uneven members do not leave a static
  // round-robin's tail); a table that grew after a capture is still covered
  // by the kernel's stride loop
  dim3 grid((unsigned)std::max(1, (b->grid_tiles + kWavesPerBlock - 1) / kWavesPerBlock)),
      block(kWave * kWavesPerBlock);
  const bool base = s > 0;
  if constexpr (sizeof(T) == 4) {
    if (b->convex) {  // FP32 Heun-2 in the convex form: one row per stage
      if (base)
        hipLaunchKernelGGL((k_stage_batch<T, 1, QUICK, GENERAL, true, true>), grid, block, 0,
                           b->stream, m, t, cnt);
      else
        hipLaunchKernelGGL((k_stage_batch<T, 1, QUICK, GENERAL, false, true>), grid, block, 0,
                           b->stream, m, t, cnt);
      return;
    }
  }
#define MFSV_BLAUNCH(N, B) \
  hipLaunchKernelGGL((k_stage_batch<T, N, QUICK, GENERAL, B, false>), grid, block, 0, b->stream, m, \
                     t, cnt)
  switch (b->nstages - s) {
    case 1: if (base) MFSV_BLAUNCH(1, true); else MFSV_BLAUNCH(1, false); break;
    case 2: if (base) MFSV_BLAUNCH(2, true); else MFSV_BLAUNCH(2, false); break;
    case 3: if (base) MFSV_BLAUNCH(3, true); else MFSV_BLAUNCH(3, false); break;
    default: MFSV_BLAUNCH(4, false); break;
  }
#undef MFSV_BLAUNCH
}

template <typename T, bool QUICK>
void launch_batch_stage_n(mfsv_batch* b, int s) {
  if (b->any_nanbed) launch_batch_stage<T, QUICK, true>(b, s);
  else launch_batch_stage<T, QUICK, false>(b, s);
}

int launch_batch_timestep(mfsv_batch* b) {
  for (int s = 0; s < b->nstages; ++s) {
    if (b->prec == MFSV_F64) {
      if (b->quick) launch_batch_stage_n<double, true>(b, s);
      else launch_batch_stage_n<double, false>(b, s);
    } else {
      launch_batch_stage_n<float, true>(b, s);
    }
  }
  hipLaunchKernelGGL(k_ctrl_batch, dim3((unsigned)std::max(b->cap, 1)), dim3(kCtrlThreads), 0,
                     b->stream, static_cast<const CtrlArgs*>(b->cargs.p),
                     static_cast<const int*>(b->counts.p) + 1);
  BHIPCHK(b, hipGetLastError());
  return MFSV_OK;
}

int refresh_tables(mfsv_batch* b) {
  return b->prec == MFSV_F64 ? upload_tables<double>(b) : upload_tables<float>(b);
}

}  // namespace

extern "C" {

int mfsv_batch_create(int precision, int device, mfsv_batch** out) {
  if (!out || (precision != MFSV_F64 && precision != MFSV_F32)) return MFSV_ERR_ARG;
  *out = nullptr;
  mfsv_batch* b = new (std::nothrow) mfsv_batch();
  if (!b) return MFSV_ERR_NOMEM;
  b->dev = device;
  b->prec = precision;
  hipDeviceProp_t prop;
  if (hipSetDevice(device) != hipSuccess || hipGetDeviceProperties(&prop, device) != hipSuccess ||
      hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess) {
    mfsv_batch_destroy(b);
    return MFSV_ERR_HIP;
  }
  *out = b;
  return MFSV_OK;
}

int mfsv_batch_destroy(mfsv_batch* b) {
  if (!b) return MFSV_ERR_ARG;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  for (auto* c : b->members) c->batch = nullptr;
  bdrop_graph(b);
  for (auto& m : b->margs) release(m);
  for (Buf* p : {&b->cargs, &b->tiles, &b->counts, &b->done}) release(*p);
  if (b->hdone) (void)hipHostFree(b->hdone);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return MFSV_OK;
}

const char* mfsv_batch_last_error(const mfsv_batch* b) { return b ? b->err.c_str() : "null batch"; }

int mfsv_batch_add(mfsv_batch* b, mfsv_ctx* c) {
  if (!b || !c) return MFSV_ERR_ARG;
  if (c->prec != b->prec) return bfail(b, MFSV_ERR_ARG, "member precision differs from the batch");
  if (c->dev != b->dev) return bfail(b, MFSV_ERR_ARG, "member lives on another device");
  if (b->members.empty()) {
    b->nstages = c->nstages;
    std::memcpy(b->tab, c->tab, sizeof(b->tab));
  } else if (c->nstages != b->nstages || std::memcmp(b->tab, c->tab, sizeof(b->tab)) != 0) {
    return bfail(b, MFSV_ERR_ARG, "member Runge-Kutta scheme differs from the batch");
  }
  if (c->batch == b) return bfail(b, MFSV_ERR_STATE, "member already in the batch");
  if (c->batch) return bfail(b, MFSV_ERR_STATE, "member of another batch");
  if (!c->begun) return bfail(b, MFSV_ERR_STATE, "call mfsv_run_begin on the member first");
  int rc = check_inputs(c);
  if (rc) return bfail(b, rc, c->err);
  if (hipSetDevice(b->dev) != hipSuccess) return bfail(b, MFSV_ERR_HIP, "hipSetDevice");
  BHIPCHK(b, hipStreamSynchronize(c->stream));
  b->members.push_back(c);
  b->active.push_back(c);
  c->batch = b;
  b->dirty = true;
  return MFSV_OK;
}

int mfsv_batch_remove(mfsv_batch* b, mfsv_ctx* c) {
  if (!b || !c) return MFSV_ERR_ARG;
  auto it = std::find(b->members.begin(), b->members.end(), c);
  if (it == b->members.end()) return bfail(b, MFSV_ERR_STATE, "not a member");
  BHIPCHK(b, hipStreamSynchronize(b->stream));
  b->members.erase(it);
  c->batch = nullptr;
  auto ja = std::find(b->active.begin(), b->active.end(), c);
  if (ja != b->active.end()) b->active.erase(ja);
  b->dirty = true;
  return MFSV_OK;
}

int mfsv_batch_size(const mfsv_batch* b, int* members, int* active) {
  if (!b || !members || !active) return MFSV_ERR_ARG;
  *members = (int)b->members.size();
  *active = (int)b->active.size();
  return MFSV_OK;
}

int mfsv_batch_run(mfsv_batch* b, int64_t max_timesteps, int use_graph) {
  if (!b || max_timesteps < 0) return MFSV_ERR_ARG;
  if (hipSetDevice(b->dev) != hipSuccess) return bfail(b, MFSV_ERR_HIP, "hipSetDevice");
  int rc;
  const int64_t chunk = 16;
  // a member reconfigured since the tables were built (tile height,
  // parameters, bed, forcing) invalidates them
  if (!b->dirty) {
    if (b->gens.size() != b->active.size()) b->dirty = true;
    for (size_t j = 0; j < b->active.size() && !b->dirty; ++j)
      b->dirty = b->active[j]->gen != b->gens[j];
  }
  for (auto* c : b->active)
    if (c->nstages != b->nstages || std::memcmp(b->tab, c->tab, sizeof(b->tab)) != 0)
      return bfail(b, MFSV_ERR_STATE, "a member's Runge-Kutta scheme changed after it joined");
  if (max_timesteps > 0)
    for (auto* c : b->active) c->ran = true;
  BHIPCHK(b, hipEventRecord(b->ev0, b->stream));
  int64_t left = max_timesteps;
  while (left > 0) {
    if (b->dirty && (rc = refresh_tables(b))) return rc;
    if (b->active.empty()) break;
    const int64_t n = left < chunk ? left : chunk;
    if (use_graph && n == chunk) {
      if (b->graph_steps != chunk) {
        bdrop_graph(b);
        BHIPCHK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        rc = 0;
        for (int64_t k = 0; k < chunk && !rc; ++k) rc = launch_batch_timestep(b);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(b->stream, &g);
        if (rc || e != hipSuccess) {
          if (g) (void)hipGraphDestroy(g);
          return rc ? rc : bfail(b, MFSV_ERR_HIP, "graph capture failed");
        }
        b->graph = g;
        BHIPCHK(b, hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0));
        b->graph_steps = chunk;
      }
      BHIPCHK(b, hipGraphLaunch(b->gexec, b->stream));
    } else {
      for (int64_t k = 0; k < n; ++k)
        if ((rc = launch_batch_timestep(b))) return rc;
    }
    left -= n;
    // which members finished?
    const int na = (int)b->active.size();
    hipLaunchKernelGGL(k_gather_done, dim3((na + 255) / 256), dim3(256), 0, b->stream,
                       static_cast<const CtrlArgs*>(b->cargs.p),
                       static_cast<const int*>(b->counts.p) + 1, static_cast<int*>(b->done.p));
    BHIPCHK(b, hipMemcpyAsync(b->hdone, b->done.p, sizeof(int) * na, hipMemcpyDeviceToHost, b->stream));
    BHIPCHK(b, hipStreamSynchronize(b->stream));
    std::vector<mfsv_ctx*> still;
    for (int j = 0; j < na; ++j)
      if (!b->hdone[j]) still.push_back(b->active[j]);
    if ((int)still.size() != na) {
      b->active.swap(still);
      b->dirty = true;
    }
  }
  BHIPCHK(b, hipEventRecord(b->ev1, b->stream));
  BHIPCHK(b, hipEventSynchronize(b->ev1));
  float ms = 0.f;
  BHIPCHK(b, hipEventElapsedTime(&ms, b->ev0, b->ev1));
  b->last_ms = ms;
  return MFSV_OK;
}

int mfsv_batch_last_run_ms(const mfsv_batch* b, double* ms) {
  if (!b || !ms) return MFSV_ERR_ARG;
  *ms = b->last_ms;
  return MFSV_OK;
}

}  // extern "C"
