// Context, memory, error reporting and the host-side FP64 geometry
// (BH_defineMatrix, eulerSearch) of libemc_hip.so.
#include <cmath>
#include <cstring>
#include <iterator>

#include "emc_internal.hpp"

namespace emc {

static thread_local std::string g_last_error;

void set_last_error(const char* msg) { g_last_error = msg ? msg : ""; }

void fail(emc_status s, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  throw Error(s, buf);
}

void use_device(emc_ctx* ctx) {
  EMC_REQUIRE(ctx != nullptr, "null context");
  EMC_HIP(hipSetDevice(ctx->device));
}

void* scratch(emc_ctx* ctx, const char* name, size_t bytes) {
  Scratch& s = ctx->scratch[name];
  if (s.bytes < bytes) {
    if (s.ptr) {
      EMC_HIP(hipStreamSynchronize(ctx->stream));
      EMC_HIP(hipFree(s.ptr));
      s.ptr = nullptr;
      s.bytes = 0;
    }
    size_t want = bytes < 256 ? 256 : bytes;
    if (hipMalloc(&s.ptr, want) != hipSuccess) {
      (void)hipGetLastError();
      fail(EMC_ERR_ALLOC, "scratch '%s': cannot allocate %zu bytes", name, want);
    }
    s.bytes = want;
  }
  return s.ptr;
}

void* pinned(emc_ctx* ctx, size_t bytes) {
  if (ctx->h_pinned_bytes < bytes) {
    if (ctx->h_pinned) EMC_HIP(hipHostFree(ctx->h_pinned));
    ctx->h_pinned = nullptr;
    size_t want = bytes < 4096 ? 4096 : bytes;
    EMC_HIP(hipHostMalloc((void**)&ctx->h_pinned, want, hipHostMallocDefault));
    ctx->h_pinned_bytes = want;
  }
  return ctx->h_pinned;
}

// ------------------------------------------------------------- staging

constexpr size_t kRingBytes = size_t(32) << 20;

void upload_async(emc_ctx* ctx, void* d_dst, const void* src, size_t bytes) {
  if (bytes == 0) return;
  const size_t need = (bytes + 255) & ~size_t(255);
  if (need > kRingBytes / 4) {  // large: the copy waits for the stream
    EMC_HIP(hipMemcpyAsync(d_dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    EMC_HIP(hipStreamSynchronize(ctx->stream));
    return;
  }
  if (!ctx->h_ring) {
    EMC_HIP(hipHostMalloc((void**)&ctx->h_ring, kRingBytes, hipHostMallocDefault));
    ctx->ring_bytes = kRingBytes;
    ctx->ring_head = 0;
  }
  auto& busy = ctx->ring_busy;
  // retire finished uploads, oldest first (they are kept in upload order)
  size_t done = 0;
  while (done < busy.size() && hipEventQuery(busy[done].ev) == hipSuccess) {
    ctx->ring_events.push_back(busy[done].ev);
    ++done;
  }
  busy.erase(busy.begin(), busy.begin() + (std::ptrdiff_t)done);
  if (ctx->ring_head + need > ctx->ring_bytes) ctx->ring_head = 0;
  const size_t beg = ctx->ring_head, end = beg + need;
  // wait for (and retire) every earlier upload whose region this one reuses
  for (size_t i = 0; i < busy.size();) {
    if (busy[i].beg < end && beg < busy[i].end) {
      EMC_HIP(hipEventSynchronize(busy[i].ev));
      ctx->ring_events.push_back(busy[i].ev);
      busy.erase(busy.begin() + (std::ptrdiff_t)i);
    } else {
      ++i;
    }
  }
  std::memcpy(ctx->h_ring + beg, src, bytes);
  EMC_HIP(hipMemcpyAsync(d_dst, ctx->h_ring + beg, bytes, hipMemcpyHostToDevice, ctx->stream));
  hipEvent_t ev;
  if (!ctx->ring_events.empty()) {
    ev = ctx->ring_events.back();
    ctx->ring_events.pop_back();
  } else {
    EMC_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  }
  EMC_HIP(hipEventRecord(ev, ctx->stream));
  busy.push_back({beg, end, ev});
  ctx->ring_head = end;
}

// ------------------------------------------------------------- timing

static hipEvent_t take_event(emc_ctx* c) {
  if (!c->event_pool.empty()) {
    hipEvent_t e = c->event_pool.back();
    c->event_pool.pop_back();
    return e;
  }
  hipEvent_t e;
  EMC_HIP(hipEventCreate(&e));
  return e;
}

KTimer::KTimer(emc_ctx* ctx, const char* name) : c(ctx), on(ctx && ctx->timing) {
  if (!on) return;
  p.name = name;
  p.a = take_event(c);
  p.b = take_event(c);
  EMC_HIP(hipEventRecord(p.a, c->stream));
}

KTimer::~KTimer() {
  if (!on) return;
  if (hipEventRecord(p.b, c->stream) == hipSuccess) c->pending.push_back(p);
}

static void fold_timing(emc_ctx* c) {
  EMC_HIP(hipStreamSynchronize(c->stream));
  for (auto& p : c->pending) {
    float ms = 0.0f;
    EMC_HIP(hipEventElapsedTime(&ms, p.a, p.b));
    auto& t = c->ktimes[p.name];
    t.first += ms;
    t.second += 1;
    c->event_pool.push_back(p.a);
    c->event_pool.push_back(p.b);
  }
  c->pending.clear();
}

// ------------------------------------------------------------- geometry

static void rz(double a, double m[3][3]) {
  double c = std::cos(a), s = std::sin(a);
  double r[3][3] = {{c, -s, 0}, {s, c, 0}, {0, 0, 1}};
  std::memcpy(m, r, sizeof(r));
}
static void rx(double a, double m[3][3]) {
  double c = std::cos(a), s = std::sin(a);
  double r[3][3] = {{1, 0, 0}, {0, c, -s}, {0, s, c}};
  std::memcpy(m, r, sizeof(r));
}
static void ry(double a, double m[3][3]) {
  double c = std::cos(a), s = std::sin(a);
  double r[3][3] = {{c, 0, s}, {0, 1, 0}, {-s, 0, c}};
  std::memcpy(m, r, sizeof(r));
}
static void mul3(const double a[3][3], const double b[3][3], double o[3][3]) {
  double t[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) t[i][j] = a[i][0] * b[0][j] + a[i][1] * b[1][j] + a[i][2] * b[2][j];
  std::memcpy(o, t, sizeof(t));
}

void define_matrix(const double ang_deg[3], int convention, int direction, double out_colmajor[9]) {
  double a[3];
  for (int i = 0; i < 3; ++i) a[i] = ang_deg[i] * (M_PI / 180.0);
  if (convention == 1) a[1] = a[2] = 0.0;  // TILT takes one angle
  switch (direction) {
    case 0:  // forward
    case 3:  // invVector
      for (double& v : a) v = -v;
      break;
    case 1:  // inv
    case 2:  // forwardVector
      if (convention != 1) std::swap(a[0], a[2]);
      break;
    default:
      fail(EMC_ERR_INVALID, "direction must be 0..3, got %d", direction);
  }
  double m[3][3], p[3][3], q[3][3];
  switch (convention) {
    case 0:  // Bah: Rz(a3) Rx(a2) Rz(a1)
      rz(a[2], m); rx(a[1], p); mul3(m, p, m); rz(a[0], q); mul3(m, q, m);
      break;
    case 1: {  // TILT about Y
      double c = std::cos(a[0]), s = std::sin(a[0]);
      double r[3][3] = {{c, 0, s}, {0, 1, 0}, {-s, 0, c}};
      std::memcpy(m, r, sizeof(r));
      break;
    }
    case 2:  // SPIDER: Rz Ry Rz
    case 4:  // IMOD (as written in the reference): Rz Ry Rz
      rz(a[2], m); ry(a[1], p); mul3(m, p, m); rz(a[0], q); mul3(m, q, m);
      break;
    case 3:  // Helical: Rz Rx Ry
      rz(a[2], m); rx(a[1], p); mul3(m, p, m); ry(a[0], q); mul3(m, q, m);
      break;
    default:
      fail(EMC_ERR_INVALID, "convention must be 0..4, got %d", convention);
  }
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r) out_colmajor[c * 3 + r] = m[r][c];
}

// MATLAB a:step:b
static std::vector<double> colon(double a, double step, double b) {
  std::vector<double> v;
  if (step == 0 || (step > 0 && a > b) || (step < 0 && a < b)) return v;
  int64_t n = (int64_t)std::floor((b - a) / step + 1e-10) + 1;
  v.reserve(n);
  for (int64_t i = 0; i < n; ++i) v.push_back(a + step * (double)i);
  return v;
}

void euler_search(const char* sym, double theta_max, double theta_step, double psi_max,
                  double psi_step, std::vector<double>& out) {
  EMC_REQUIRE(sym && sym[0], "empty symmetry");
  bool bipolar = theta_max < 0;
  theta_max = std::fabs(theta_max);
  double psi_full = 2.0 * psi_max, phi_max = 360.0;
  switch (sym[0]) {
    case 'C': {
      int n = std::atoi(sym + 1);
      EMC_REQUIRE(n > 0, "Cyclic symmetry requires an int specifying CX");
      psi_full = std::min(psi_full, 360.0 / n);
      theta_max = std::min(180.0, theta_max);
      break;
    }
    case 'D': {
      int n = std::atoi(sym + 1);
      EMC_REQUIRE(n > 0, "D symmetry requires an int specifying DX");
      psi_full = 360.0 / n;
      theta_max = std::min(theta_max, 90.0);
      break;
    }
    case 'O':
      psi_full = std::min(psi_full, 90.0);
      theta_max = std::min(theta_max, 54.7);
      phi_max = 90.0;
      break;
    case 'I':
      psi_full = std::min(psi_full, 180.0);
      theta_max = 31.7;
      phi_max = 180.0;
      break;
    default:
      fail(EMC_ERR_INVALID, "symmetry symbol (%s) not recognized", sym);
  }
  std::vector<double> psi = colon(-psi_full / 2.0, psi_step, psi_full / 2.0);
  std::vector<double> theta = colon(0.0, theta_step, theta_max);
  if (theta.empty()) theta.push_back(0.0);
  if (bipolar) {
    size_t n = theta.size();
    for (size_t i = 0; i < n; ++i) theta.push_back(180.0 - theta[n - 1 - i]);
  }
  if (psi.empty()) psi.push_back(0.0);
  out.clear();
  for (double th : theta) {
    double phi_step;
    if (th == 0.0 || th == 180.0) {
      phi_step = phi_max;
    } else {
      phi_step = std::fabs(theta_step / std::sin(th * M_PI / 180.0));
      if (phi_step > phi_max) phi_step = phi_max;
      else phi_step = phi_max / std::floor(phi_max / phi_step + 0.5);
    }
    for (double ph : colon(0.0, phi_step, phi_max - 1.0))
      for (double ps : psi) {
        out.push_back(ph);
        out.push_back(th);
        out.push_back(ps);
      }
  }
}

constexpr size_t kCacheLimit = size_t(16) << 30;  // bytes of free blocks kept per context

// Block size class: 512 B granules below 1 MiB, 2 MiB granules above.
size_t block_size(size_t bytes) {
  const size_t g = bytes < (size_t(1) << 20) ? 512 : (size_t(2) << 20);
  return (std::max<size_t>(bytes, 1) + g - 1) / g * g;
}

// Return every cached block to HIP (after the stream drains).
void release_cache(emc_ctx* ctx) {
  if (ctx->cached.empty()) return;
  (void)hipStreamSynchronize(ctx->stream);
  for (auto& kv : ctx->cached) (void)hipFree(kv.second);
  ctx->cached.clear();
  ctx->cached_bytes = 0;
}

}  // namespace emc

using namespace emc;

namespace {
struct TuneRange {
  const char* name;
  int lo, hi, dflt;
};
constexpr TuneRange kTune[] = {{"search_tile", 2, 128, 16},
                               {"grad_blocks", 256, 65536, 4096},
                               {"search_passes", 1, 4, 2},
                               {"refine_graph", 0, 1, 1},
                               {"search_zsplit", 1, 4, 1},
                               {"tsearch_pruned", 0, 2, 2},
                               {"grad_kernel", 0, 1, 1},
                               {"grad_brick", 16, 32, 32},
                               {"align_grid_batch", 1, 64, 16},
                               {"grad_waves", 4, 5, 4}};

const TuneRange* find_tune(const char* name) {
  for (const auto& t : kTune)
    if (name && std::strcmp(name, t.name) == 0) return &t;
  return nullptr;
}
}  // namespace

int emc::tune_value(emc_ctx* ctx, const char* name, int n, int dflt) {
  auto it = ctx->meta.find(std::string("tune:") + name + ":" + std::to_string(n));
  if (it != ctx->meta.end()) return it->second;
  it = ctx->meta.find(std::string("tune:") + name + ":0");
  return it != ctx->meta.end() ? it->second : dflt;
}

extern "C" {

const char* emc_last_error(void) { return g_last_error.c_str(); }

const char* emc_version(void) { return "emclarity_amd 0.1 (gfx950, rocFFT)"; }

emc_status emc_ctx_create(int device, emc_ctx** out) {
  return guard([&] {
    EMC_REQUIRE(out != nullptr, "null output pointer");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
      (void)hipGetLastError();
      fail(EMC_ERR_NODEVICE, "no HIP device available");
    }
    EMC_REQUIRE(device >= 0 && device < n, "device %d out of range (%d devices)", device, n);
    emc_ctx* c = new emc_ctx();
    c->device = device;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
      delete c;
      fail(EMC_ERR_HIP, "context creation: %s", hipGetErrorString(e));
    }
    *out = c;
  });
}

emc_status emc_fft_plan_destroy(emc_fft_plan* plan);

emc_status emc_ctx_destroy(emc_ctx* ctx) {
  return guard([&] {
    if (!ctx) return;
    use_device(ctx);
    (void)hipStreamSynchronize(ctx->stream);
    for (auto& kv : ctx->plans) emc_fft_plan_destroy(kv.second);
    for (auto& p : ctx->pending) {
      (void)hipEventDestroy(p.a);
      (void)hipEventDestroy(p.b);
    }
    for (auto e : ctx->event_pool) (void)hipEventDestroy(e);
    for (auto& r : ctx->ring_busy) (void)hipEventDestroy(r.ev);
    for (auto e : ctx->ring_events) (void)hipEventDestroy(e);
    if (ctx->h_ring) (void)hipHostFree(ctx->h_ring);
    for (auto& kv : ctx->graphs)
      if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
    for (auto& kv : ctx->scratch)
      if (kv.second.ptr) (void)hipFree(kv.second.ptr);
    release_cache(ctx);
    if (ctx->h_pinned) (void)hipHostFree(ctx->h_pinned);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
  });
}

emc_status emc_ctx_wait(emc_ctx* waiter, emc_ctx* signaller) {
  return guard([&] {
    EMC_REQUIRE(waiter && signaller, "null context");
    EMC_REQUIRE(waiter->device == signaller->device, "contexts on different devices");
    use_device(signaller);
    hipEvent_t e;
    EMC_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    EMC_HIP(hipEventRecord(e, signaller->stream));
    EMC_HIP(hipStreamWaitEvent(waiter->stream, e, 0));
    EMC_HIP(hipEventDestroy(e));  // released once the recorded work completes
  });
}

emc_status emc_ctx_synchronize(emc_ctx* ctx) {
  return guard([&] {
    use_device(ctx);
    EMC_HIP(hipStreamSynchronize(ctx->stream));
  });
}

void* emc_ctx_stream(emc_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

emc_status emc_tune_set(emc_ctx* ctx, const char* name, int box, int value) {
  return guard([&] {
    EMC_REQUIRE(ctx, "null context");
    const TuneRange* t = find_tune(name);
    EMC_REQUIRE(t, "unknown launch parameter '%s'", name ? name : "(null)");
    EMC_REQUIRE(box >= 0 && box <= 4096, "box size %d out of range", box);
    EMC_REQUIRE(value >= t->lo && value <= t->hi, "%s must be %d..%d", t->name, t->lo, t->hi);
    ctx->meta[std::string("tune:") + t->name + ":" + std::to_string(box)] = value;
  });
}

emc_status emc_tune_get(emc_ctx* ctx, const char* name, int box, int* value) {
  return guard([&] {
    EMC_REQUIRE(ctx && value, "null pointer");
    const TuneRange* t = find_tune(name);
    EMC_REQUIRE(t, "unknown launch parameter '%s'", name ? name : "(null)");
    *value = tune_value(ctx, t->name, box, t->dflt);
  });
}

emc_status emc_timing_enable(emc_ctx* ctx, int on) {
  return guard([&] {
    use_device(ctx);
    fold_timing(ctx);
    ctx->timing = on != 0;
  });
}

emc_status emc_timing_reset(emc_ctx* ctx) {
  return guard([&] {
    use_device(ctx);
    fold_timing(ctx);
    ctx->ktimes.clear();
  });
}

emc_status emc_timing_read(emc_ctx* ctx, const char* name, double* ms_total, int64_t* launches) {
  return guard([&] {
    use_device(ctx);
    EMC_REQUIRE(name && ms_total && launches, "null pointer");
    fold_timing(ctx);
    auto it = ctx->ktimes.find(name);
    *ms_total = it == ctx->ktimes.end() ? 0.0 : it->second.first;
    *launches = it == ctx->ktimes.end() ? 0 : it->second.second;
  });
}


emc_status emc_malloc(emc_ctx* ctx, size_t bytes, void** d_ptr) {
  return guard([&] {
    use_device(ctx);
    EMC_REQUIRE(d_ptr != nullptr, "null output pointer");
    const size_t sz = block_size(bytes);
    auto range = ctx->cached.equal_range(sz);
    if (range.first != range.second) {  // most recently freed block of this size
      auto it = std::prev(range.second);
      *d_ptr = it->second;
      ctx->cached.erase(it);
      ctx->cached_bytes -= sz;
    } else {
      void* p = nullptr;
      if (hipMalloc(&p, sz) != hipSuccess) {
        (void)hipGetLastError();
        release_cache(ctx);
        if (hipMalloc(&p, sz) != hipSuccess) {
          (void)hipGetLastError();
          fail(EMC_ERR_ALLOC, "cannot allocate %zu bytes", bytes);
        }
      }
      *d_ptr = p;
    }
    ctx->live[*d_ptr] = sz;
  });
}

emc_status emc_free(emc_ctx* ctx, void* d_ptr) {
  return guard([&] {
    use_device(ctx);
    if (!d_ptr) return;
    auto it = ctx->live.find(d_ptr);
    if (it == ctx->live.end()) {  // not from emc_malloc: plain HIP free
      for (auto& kv : ctx->cached)
        if (kv.second == d_ptr) fail(EMC_ERR_INVALID, "emc_free: block %p was already freed", d_ptr);
      EMC_HIP(hipStreamSynchronize(ctx->stream));
      EMC_HIP(hipFree(d_ptr));
      return;
    }
    ctx->cached.emplace(it->second, d_ptr);
    ctx->cached_bytes += it->second;
    ctx->live.erase(it);
    if (ctx->cached_bytes > kCacheLimit) release_cache(ctx);
  });
}

emc_status emc_memcpy_h2d(emc_ctx* ctx, void* d_dst, const void* src, size_t bytes) {
  return guard([&] {
    use_device(ctx);
    EMC_HIP(hipMemcpyAsync(d_dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    EMC_HIP(hipStreamSynchronize(ctx->stream));
  });
}

emc_status emc_memcpy_h2d_async(emc_ctx* ctx, void* d_dst, const void* src, size_t bytes) {
  return guard([&] {
    use_device(ctx);
    EMC_REQUIRE(d_dst && (src || bytes == 0), "null pointer");
    upload_async(ctx, d_dst, src, bytes);
  });
}

emc_status emc_memcpy_d2h(emc_ctx* ctx, void* dst, const void* d_src, size_t bytes) {
  return guard([&] {
    use_device(ctx);
    EMC_HIP(hipMemcpyAsync(dst, d_src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    EMC_HIP(hipStreamSynchronize(ctx->stream));
  });
}

emc_status emc_memcpy_d2d(emc_ctx* ctx, void* d_dst, const void* d_src, size_t bytes) {
  return guard([&] {
    use_device(ctx);
    EMC_HIP(hipMemcpyAsync(d_dst, d_src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
  });
}

emc_status emc_memset_zero(emc_ctx* ctx, void* d_dst, size_t bytes) {
  return guard([&] {
    use_device(ctx);
    EMC_HIP(hipMemsetAsync(d_dst, 0, bytes, ctx->stream));
  });
}

emc_status emc_define_matrix(const double angles_deg[3], int convention, int direction,
                             double R_colmajor[9]) {
  return guard([&] {
    EMC_REQUIRE(angles_deg && R_colmajor, "null pointer");
    define_matrix(angles_deg, convention, direction, R_colmajor);
  });
}

emc_status emc_euler_search(const char* symmetry, double theta_max, double theta_step,
                            double psi_max, double psi_step, double* angles_out,
                            int64_t capacity, int64_t* n_out) {
  return guard([&] {
    EMC_REQUIRE(n_out != nullptr, "null count pointer");
    std::vector<double> v;
    euler_search(symmetry, theta_max, theta_step, psi_max, psi_step, v);
    int64_t n = (int64_t)(v.size() / 3);
    *n_out = n;
    if (angles_out && capacity > 0) {
      int64_t m = n < capacity ? n : capacity;
      std::memcpy(angles_out, v.data(), sizeof(double) * 3 * m);
    }
  });
}

}  // extern "C"
