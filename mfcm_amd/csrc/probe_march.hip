// probe_march.hip -- the memory floor of the stage kernels' access pattern.
//
// A measurement aid, not part of the library: kernels that move exactly the
// bytes of one fused RK stage launch (sv_stage.hpp) in the same pattern --
// one wavefront per tile of 64 lanes (60 owned columns and a 2-column halo
// on each side), the row march with a 2-row halo above and below and the
// next row loaded a row ahead, the per-cell inputs of the cell pass at the
// owned rows, the stores masked to the owned cells, the same blocks-to-XCD
// remapping and tile order -- but with no arithmetic beyond a running sum.
// Their launch time is what the stage kernels would take if their
// arithmetic were free; a plain device copy and a read-only stream give the
// device's streaming limits beside it.
//
//   build: hipcc -O3 --offload-arch=gfx950 -o probe_march probe_march.hip
//   run:   ./probe_march [nx ny [rows_fp32 rows_fp64]]
//          (default 2048 2048, tile heights 18 and 24 in both precisions)
// Each result line ends with "key value" pairs for scripts: copy_GBs,
// read_GBs, f32_s1_us / f32_s2_us / f64_s1_us / f64_s2_us (per rows).
//
// Every access stays inside the fields: rows and columns are clamped into
// the grid, stores go to owned cells inside the grid only.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <vector>

#define CK(x)                                                                        \
  do {                                                                               \
    hipError_t e_ = (x);                                                             \
    if (e_ != hipSuccess) {                                                          \
      std::fprintf(stderr, "HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__); \
      std::exit(1);                                                                  \
    }                                                                                \
  } while (0)

constexpr int kMaxFields = 16;
struct Fields {
  void* f[kMaxFields];
};

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// the stage kernel's remapping: consecutive blocks on one XCD
__device__ __forceinline__ int xcd_block(int b, int nb) {
  const int nb8 = (nb / 8) * 8;
  if (b >= nb8) return b;
  return (b % 8) * (nb8 / 8) + b / 8;
}

// NR row fields (read rows y0-2 .. y1+1), NX per-cell fields read at the
// owned rows, NW fields written at the owned cells.  Tile order as the stage
// kernels number them: ROWMAJOR (FP64) row block by row block, else (FP32)
// groups of 4 adjacent strips marched down together.
template <typename T, int NR, int NX, int NW, bool ROWMAJOR>
__global__ __launch_bounds__(256) void k_march(Fields F, int nx, int ny, int rows, int nrb,
                                               int ntiles) {
  const int lane = threadIdx.x & 63;
  const int tile = xcd_block(blockIdx.x, gridDim.x) * 4 + (int)(threadIdx.x >> 6);
  if (tile >= ntiles) return;
  const int nstrips = ntiles / nrb;
  int strip, rb;
  if (ROWMAJOR) {
    rb = tile / nstrips;
    strip = tile - rb * nstrips;
  } else {
    const int gsz = 4 * nrb, g = tile / gsz, rem = tile - g * gsz;
    const int gw = (g + 1) * 4 <= nstrips ? 4 : nstrips - g * 4;
    rb = rem / gw;
    strip = g * 4 + (rem - rb * gw);
  }
  const int x = strip * 60 - 2 + lane;
  const int xc = clampi(x, 0, nx - 1);
  const bool own = lane >= 2 && lane < 62 && x < nx;
  const int y0 = rb * rows, y1 = min(y0 + rows, ny);
  T acc[NR], cur[NR], nxt[NR];
  for (int k = 0; k < NR; ++k) {
    acc[k] = T(0);
    nxt[k] = static_cast<const T*>(F.f[k])[int64_t(clampi(y0 - 2, 0, ny - 1)) * nx + xc];
  }
  for (int c = y0 - 2; c <= y1 + 1; ++c) {
    for (int k = 0; k < NR; ++k) cur[k] = nxt[k];
    const int cn = clampi(min(c + 1, y1 + 1), 0, ny - 1);
    for (int k = 0; k < NR; ++k) nxt[k] = static_cast<const T*>(F.f[k])[int64_t(cn) * nx + xc];
    T s = T(0);
    for (int k = 0; k < NR; ++k) {
      acc[k] = acc[k] * T(0.5) + cur[k];
      s += acc[k];
    }
    const int r = c - 1;
    if (r >= y0 && r < y1) {
      const int64_t i = int64_t(r) * nx + xc;
      for (int k = 0; k < NX; ++k) s += static_cast<const T*>(F.f[NR + k])[i];
      if (own)
        for (int k = 0; k < NW; ++k) static_cast<T*>(F.f[NR + NX + k])[i] = s + T(k);
    }
  }
}

__global__ void k_copy(const float4* __restrict__ a, float4* __restrict__ b, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    b[i] = a[i];
}
__global__ void k_read(const float4* __restrict__ a, float* out, size_t n) {
  float s = 0.f;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const float4 v = a[i];
    s += v.x + v.y + v.z + v.w;
  }
  if (s == -1.0f) out[0] = s;  // keeps the loads
}

template <typename K, typename... Args>
double time_us(K kernel, dim3 grid, dim3 block, int iters, Args... args) {
  hipEvent_t a, b;
  CK(hipEventCreate(&a));
  CK(hipEventCreate(&b));
  for (int i = 0; i < 3; ++i) hipLaunchKernelGGL(kernel, grid, block, 0, 0, args...);
  CK(hipGetLastError());
  CK(hipEventRecord(a));
  for (int i = 0; i < iters; ++i) hipLaunchKernelGGL(kernel, grid, block, 0, 0, args...);
  CK(hipEventRecord(b));
  CK(hipEventSynchronize(b));
  float ms = 0.f;
  CK(hipEventElapsedTime(&ms, a, b));
  CK(hipEventDestroy(a));
  CK(hipEventDestroy(b));
  return 1000.0 * ms / iters;
}

template <typename T, int NR, int NX, int NW, bool ROWMAJOR>
void march(const Fields& F, int nx, int ny, int rows, const char* name, const char* key) {
  const int nstrips = (nx + 59) / 60, nrb = (ny + rows - 1) / rows, ntiles = nstrips * nrb;
  const double us = time_us(k_march<T, NR, NX, NW, ROWMAJOR>, dim3((ntiles + 3) / 4), dim3(256), 50, F,
                            nx, ny, rows, nrb, ntiles);
  const double bytes = double(nx) * ny * sizeof(T) * (NR + NX + NW);
  std::printf("%-44s rows %2d  %8.2f us  %6.1f MB of fields  %6.0f GB/s  | %s rows %d us %.3f\n", name,
              rows, us, bytes / 1e6, bytes / us / 1e3, key, rows, us);
}

int main(int argc, char** argv) {
  const int nx = argc > 2 ? std::atoi(argv[1]) : 2048, ny = argc > 2 ? std::atoi(argv[2]) : 2048;
  if (nx < 1 || ny < 1 || int64_t(nx) * ny * 8 >= (int64_t(1) << 31)) {
    std::fprintf(stderr, "grid %d x %d out of range\n", nx, ny);
    return 2;
  }
  std::vector<int> rows32 = {18, 24}, rows64 = {18, 24};
  if (argc > 4) {
    rows32 = {std::atoi(argv[3])};
    rows64 = {std::atoi(argv[4])};
  }
  for (int r : rows32)
    if (r < 1 || r > 64) { std::fprintf(stderr, "tile rows %d out of range\n", r); return 2; }
  for (int r : rows64)
    if (r < 1 || r > 64) { std::fprintf(stderr, "tile rows %d out of range\n", r); return 2; }
  Fields F;
  for (int k = 0; k < kMaxFields; ++k) {
    CK(hipMalloc(&F.f[k], size_t(nx) * ny * 8));
    CK(hipMemset(F.f[k], 0, size_t(nx) * ny * 8));
  }
  {
    const size_t n = (size_t(256) << 20) / 16;  // 256 MiB
    float4 *a, *b;
    float* o;
    CK(hipMalloc(&a, n * 16));
    CK(hipMalloc(&b, n * 16));
    CK(hipMalloc(&o, 64));
    CK(hipMemset(a, 0, n * 16));
    const double cu = time_us(k_copy, dim3(8192), dim3(256), 20, (const float4*)a, b, n);
    const double ru = time_us(k_read, dim3(8192), dim3(256), 20, (const float4*)a, o, n);
    std::printf("device copy 256 MiB (read + write)           %8.2f us  %6.0f GB/s  | copy_GBs %.1f\n", cu,
                2.0 * n * 16 / cu / 1e3, 2.0 * n * 16 / cu / 1e3);
    std::printf("read-only stream 256 MiB                     %8.2f us  %6.0f GB/s  | read_GBs %.1f\n", ru,
                1.0 * n * 16 / ru / 1e3, 1.0 * n * 16 / ru / 1e3);
    CK(hipFree(a));
    CK(hipFree(b));
    CK(hipFree(o));
  }
  std::printf("grid %d x %d\n", nx, ny);
  // FP32 convex Heun-2: stage 1 reads y0 (h, u, v) and the bed as row fields,
  // n and y0's low-order depth at the owned rows, writes y1 and a low word;
  // stage 2 reads y1 and the bed as rows, n, the low word and y0 at the
  // owned rows, writes y2 and its low word.
  for (int rows : rows32) {
    march<float, 4, 2, 4, false>(F, nx, ny, rows, "FP32 stage 1 pattern (4 rows, 2 cell, 4 out)", "f32_s1");
    march<float, 4, 5, 4, false>(F, nx, ny, rows, "FP32 stage 2 pattern (4 rows, 5 cell, 4 out)", "f32_s2");
  }
  // FP64 reference summation order: stage 1 reads y0 and the bed as rows and
  // n at the owned rows, writes y1 and the partial row; stage 2 reads y1 and
  // the bed as rows, n and the partial at the owned rows, writes y2.
  for (int rows : rows64) {
    march<double, 4, 1, 6, true>(F, nx, ny, rows, "FP64 stage 1 pattern (4 rows, 1 cell, 6 out)", "f64_s1");
    march<double, 4, 4, 3, true>(F, nx, ny, rows, "FP64 stage 2 pattern (4 rows, 4 cell, 3 out)", "f64_s2");
  }
  for (int k = 0; k < kMaxFields; ++k) CK(hipFree(F.f[k]));
  return 0;
}
