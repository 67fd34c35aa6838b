// Internal definitions shared by the HIP sources of libemc_hip.so.
//
// Design notes (gfx950 / MI355X):
//  * one HIP stream per emc_ctx, default priority;
//  * FP32 volumes, FP64 only for Euler matrices and reduction accumulators;
//  * reductions are deterministic: a fixed grid writes per-block partials,
//    one block folds them in a fixed order (no floating-point atomics), so a
//    repeated call returns bit-identical sums and argmax ties resolve the same
//    way every time;
//  * wavefronts are 64 lanes; block reductions shuffle within a wavefront
//    and go through LDS across wavefronts.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <map>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/emc_hip.h"

namespace emc {

constexpr int kWave = 64;
constexpr int kRedBlocks = 512;     // fixed grid of every reduction
constexpr int kRedThreads = 256;    // 4 wavefronts
constexpr int kMaxSym = 60;

struct Error : public std::runtime_error {
  emc_status status;
  Error(emc_status s, const std::string& m) : std::runtime_error(m), status(s) {}
};

[[noreturn]] void fail(emc_status s, const char* fmt, ...);
void set_last_error(const char* msg);

#define EMC_HIP(x)                                                                   \
  do {                                                                               \
    hipError_t e_ = (x);                                                             \
    if (e_ != hipSuccess)                                                            \
      ::emc::fail(EMC_ERR_HIP, "%s:%d %s: %s", __FILE__, __LINE__, #x,               \
                  hipGetErrorString(e_));                                            \
  } while (0)

#define EMC_REQUIRE(cond, ...)                                                       \
  do {                                                                               \
    if (!(cond)) ::emc::fail(EMC_ERR_INVALID, __VA_ARGS__);                          \
  } while (0)

// Run `f` and convert exceptions into an emc_status for the C ABI.
template <class F>
emc_status guard(F&& f) {
  try {
    f();
    return EMC_OK;
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.status;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return EMC_ERR_INVALID;
  }
}

// Device scratch owned by a context, grown on demand, reused between calls.
struct Scratch {
  void* ptr = nullptr;
  size_t bytes = 0;
};

}  // namespace emc

struct emc_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  // named scratch slots (grown on demand, freed with the context)
  std::map<std::string, emc::Scratch> scratch;
  // small host-side facts about cached scratch contents (e.g. pass counts)
  std::map<std::string, int> meta;
  // internal FFT plans keyed by (nx, ny, nz, batch)
  std::map<std::tuple<int, int, int, int>, emc_fft_plan*> plans;
  // caching allocator behind emc_malloc / emc_free: freed blocks are kept
  // by size and reused in the order of this context's stream
  std::multimap<size_t, void*> cached;  // size -> free block
  std::map<void*, size_t> live;         // block handed out -> size
  size_t cached_bytes = 0;
  double* h_pinned = nullptr;  // small pinned host buffer for results
  size_t h_pinned_bytes = 0;
  // pinned staging ring for host-to-device uploads that must not wait for
  // the stream: each upload's region is held until the event recorded after
  // its copy has passed
  unsigned char* h_ring = nullptr;
  size_t ring_bytes = 0, ring_head = 0;
  struct StageRegion {
    size_t beg, end;
    hipEvent_t ev;
  };
  std::vector<StageRegion> ring_busy;
  std::vector<hipEvent_t> ring_events;  // free events (timing disabled)
  // optional per-kernel timing with HIP events on the context stream
  bool timing = false;
  struct Pending {
    std::string name;
    hipEvent_t a, b;
  };
  std::vector<Pending> pending;
  std::vector<hipEvent_t> event_pool;
  std::map<std::string, std::pair<double, int64_t>> ktimes;  // name -> (ms, launches)
  // captured launch sequences replayed as HIP graphs, keyed by name; `key`
  // holds every argument the capture baked in (pointers included), so a
  // changed argument means a new capture
  struct Graph {
    hipGraphExec_t exec = nullptr;
    std::vector<unsigned char> key;
  };
  std::map<std::string, Graph> graphs;
};

namespace emc {

// Brackets one kernel launch with HIP events when the context's timing is on.
struct KTimer {
  emc_ctx* c;
  emc_ctx::Pending p;
  bool on;
  KTimer(emc_ctx* ctx, const char* name);
  ~KTimer();
};

}  // namespace emc

namespace emc {

// Make `ctx`'s device current for the calling thread.
void use_device(emc_ctx* ctx);
// Launch-shape parameter `name` for box size n (emc_tune_set), else `dflt`.
int tune_value(emc_ctx* ctx, const char* name, int n, int dflt);
// Scratch buffer of at least `bytes` under `name` (contents undefined).
void* scratch(emc_ctx* ctx, const char* name, size_t bytes);
// Pinned host staging of at least `bytes`.
void* pinned(emc_ctx* ctx, size_t bytes);
// Internal cached plan (owned by the context).
emc_fft_plan* cached_plan(emc_ctx* ctx, int nx, int ny, int nz, int batch);

inline int ceil_div(int64_t a, int64_t b) { return (int)((a + b - 1) / b); }

// Host-side R2C/C2R/C2C on an internal plan.
void fft_r2c(emc_fft_plan* p, const float* in, void* out);
void fft_c2r(emc_fft_plan* p, void* in, float* out);
void fft_c2c(emc_fft_plan* p, const void* in, void* out, int direction);

// ---------------------------------------------------------------- reductions

// Launch helpers implemented in reduce.hip.
// sum and sum of squares of a real array (FP64), out[0]=sum, out[1]=sumsq.
void reduce_sum_sumsq(emc_ctx* ctx, const float* d, int64_t n, double* d_out2);
// minimum of a real array read with a stride (stride 2 reads real parts of
// complex data), out[0] = min as double.
// Copy `bytes` from host memory to d_dst on the context's stream through the
// pinned staging ring: returns once `src` may be reused, without waiting for
// the work already queued (uploads above a quarter of the ring wait for the
// stream instead).
void upload_async(emc_ctx* ctx, void* d_dst, const void* src, size_t bytes);

void reduce_min(emc_ctx* ctx, const float* d, int64_t n, int stride, double* d_out1);
// minima of `batch` volumes of n values (stride between values, bstride floats
// between volumes) into d_out[0 .. batch)
void reduce_min_batched(emc_ctx* ctx, const float* d, int64_t n, int stride, int64_t bstride, int batch,
                        double* d_out);

// ---------------------------------------------------------- device helpers

__device__ inline double wave_sum(double v) {
  for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_down(v, o, kWave);
  return v;
}

__device__ inline float wave_min(float v) {
  for (int o = kWave / 2; o > 0; o >>= 1) v = fminf(v, __shfl_down(v, o, kWave));
  return v;
}

__device__ inline unsigned long long wave_max_u64(unsigned long long v) {
  for (int o = kWave / 2; o > 0; o >>= 1) {
    unsigned long long w = __shfl_down(v, o, kWave);
    v = w > v ? w : v;
  }
  return v;
}

// Block-wide sum of K doubles; result valid in thread 0. blockDim.x must be a
// multiple of 64 and at most 1024.
template <int K>
__device__ inline void block_sum(double (&v)[K]) {
  __shared__ double lds[K][16];
  const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
  const int nw = blockDim.x / kWave;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    double s = wave_sum(v[k]);
    if (lane == 0) lds[k][wid] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      double s = 0.0;
      for (int w = 0; w < nw; ++w) s += lds[k][w];
      v[k] = s;
    }
  }
  __syncthreads();
}

__device__ inline unsigned long long block_max_u64(unsigned long long v) {
  __shared__ unsigned long long lds[16];
  const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
  const int nw = blockDim.x / kWave;
  v = wave_max_u64(v);
  if (lane == 0) lds[wid] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < nw; ++w) v = lds[w] > v ? lds[w] : v;
  }
  __syncthreads();
  return v;
}

__device__ inline float block_min(float v) {
  __shared__ float lds[16];
  const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
  const int nw = blockDim.x / kWave;
  v = wave_min(v);
  if (lane == 0) lds[wid] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < nw; ++w) v = fminf(v, lds[w]);
  }
  __syncthreads();
  return v;
}

// Fold kRedBlocks-style partials of K doubles (layout part[K*b + k]) in a
// fixed order into out[0..K-1]. One block of kRedThreads per reduction;
// block i folds part + i*K*nb into out + i*K.
template <int K>
__global__ void __launch_bounds__(kRedThreads) fold_partials_kernel(const double* __restrict__ part,
                                                                    int nb,
                                                                    double* __restrict__ out) {
  part += (int64_t)blockIdx.x * K * nb;
  out += (int64_t)blockIdx.x * K;
  double v[K];
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x)
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] += part[(int64_t)K * b + k];
  block_sum<K>(v);
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < K; ++k) out[k] = v[k];
}

// Orderable key for "max value, first index on ties": larger key = larger
// value, and among equal values the smaller index. NaN sorts lowest (MATLAB
// max ignores NaN).
__device__ inline unsigned long long argmax_key(float v, uint32_t idx) {
  uint32_t u = __float_as_uint(v);
  if (v != v) u = 0u;  // NaN -> lowest
  else u = (u & 0x80000000u) ? ~u : (u | 0x80000000u);
  return ((unsigned long long)u << 32) | (unsigned long long)(0xFFFFFFFFu - idx);
}

__device__ __host__ inline float key_value(unsigned long long key) {
  uint32_t u = (uint32_t)(key >> 32);
  if (u == 0u) return __builtin_nanf("");
  u = (u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u;
  float f;
  __builtin_memcpy(&f, &u, 4);
  return f;
}

__device__ __host__ inline uint32_t key_index(unsigned long long key) {
  return 0xFFFFFFFFu - (uint32_t)(key & 0xFFFFFFFFull);
}

// (x, y, z) of linear index i in an nx x ny x * grid (x fastest). The GPU has
// no integer divider: a 64-bit division is a long emulated sequence, a
// 32-bit one a short reciprocal-based one, so indices below 2^32 (every
// volume here) take the 32-bit path.
__device__ inline void split3(int64_t i, int nx, int ny, int& x, int& y, int& z) {
  if ((uint64_t)i <= 0xFFFFFFFFull) {
    const uint32_t u = (uint32_t)i, q = u / (uint32_t)nx, zz = q / (uint32_t)ny;
    x = (int)(u - q * (uint32_t)nx);
    y = (int)(q - zz * (uint32_t)ny);
    z = (int)zz;
  } else {
    const int64_t q = i / nx;
    x = (int)(i - q * nx);
    z = (int)(q / ny);
    y = (int)(q - (int64_t)z * ny);
  }
}

// Trilinear sample at index coordinates (x,y,z), zero outside [0,n-1]
// (texture border addressing of the reference's mexXform3d).
__device__ inline float sample_trilinear_zero(const float* __restrict__ v, int nx, int ny, int nz,
                                              float x, float y, float z) {
  const float xf = floorf(x), yf = floorf(y), zf = floorf(z);
  const float fx = x - xf, fy = y - yf, fz = z - zf;
  // guard the int conversion far outside the volume
  if (xf < -2.0f || yf < -2.0f || zf < -2.0f || xf > (float)nx || yf > (float)ny || zf > (float)nz)
    return 0.0f;
  const int x0 = (int)xf, y0 = (int)yf, z0 = (int)zf;
  float c[2][2][2];
#pragma unroll
  for (int dz = 0; dz < 2; ++dz)
#pragma unroll
    for (int dy = 0; dy < 2; ++dy)
#pragma unroll
      for (int dx = 0; dx < 2; ++dx) {
        const int xi = x0 + dx, yi = y0 + dy, zi = z0 + dz;
        const bool ok = (unsigned)xi < (unsigned)nx && (unsigned)yi < (unsigned)ny &&
                        (unsigned)zi < (unsigned)nz;
        c[dz][dy][dx] = ok ? v[((int64_t)zi * ny + yi) * nx + xi] : 0.0f;
      }
  const float c00 = c[0][0][0] + fx * (c[0][0][1] - c[0][0][0]);
  const float c01 = c[0][1][0] + fx * (c[0][1][1] - c[0][1][0]);
  const float c10 = c[1][0][0] + fx * (c[1][0][1] - c[1][0][0]);
  const float c11 = c[1][1][0] + fx * (c[1][1][1] - c[1][1][0]);
  const float c0 = c00 + fy * (c01 - c00);
  const float c1 = c10 + fy * (c11 - c10);
  return c0 + fz * (c1 - c0);
}

// Index of the centred (fftshifted) position of FFT-order index j on an axis
// of length n: fftshift moves j to (j + n/2) mod n.
__device__ __host__ inline int shifted_index(int j, int n) {
  int s = j + n / 2;
  return s >= n ? s - n : s;
}

// FFT-order index of -j on an axis of length n.
__device__ __host__ inline int mirror_index(int j, int n) { return j == 0 ? 0 : n - j; }

}  // namespace emc
