// rocFFT plans behind the mexFFT-style handle (mexFiles/mexFFT.cu:178-223).
//
// One emc_fft_plan holds out-of-place FP32 R2C and C2R plans for `batch`
// contiguous volumes of dims {nx,ny,nz} (x fastest; rocFFT lengths are given
// fastest-first, so no dimension reversal is needed, unlike cuFFT's
// row-major dims at mexFFT.cu:136-146), and a C2C pair created on first use.
// Each plan owns its work buffer; executions run on the context's stream.
#include <rocfft/rocfft.h>

#include <mutex>

#include "ops.hpp"

struct emc_fft_plan {
  emc_ctx* ctx = nullptr;
  int nx = 0, ny = 0, nz = 0, batch = 1;
  rocfft_plan r2c = nullptr, c2r = nullptr, c2c_fwd = nullptr, c2c_inv = nullptr;
  rocfft_execution_info info = nullptr;
  void* work = nullptr;
  size_t work_bytes = 0;
};

namespace emc {

static void rocfft_check(rocfft_status s, const char* what) {
  if (s != rocfft_status_success) fail(EMC_ERR_FFT, "rocFFT %s failed (status %d)", what, (int)s);
}

static void rocfft_init_once() {
  static std::once_flag flag;
  static rocfft_status st = rocfft_status_success;
  std::call_once(flag, [] { st = rocfft_setup(); });
  rocfft_check(st, "setup");
}

static rocfft_plan make_plan(emc_fft_plan* p, rocfft_transform_type type) {
  size_t lengths[3] = {(size_t)p->nx, (size_t)p->ny, (size_t)p->nz};
  rocfft_plan plan = nullptr;
  rocfft_check(rocfft_plan_create(&plan, rocfft_placement_notinplace, type, rocfft_precision_single,
                                  3, lengths, (size_t)p->batch, nullptr),
               "plan_create");
  size_t wb = 0;
  rocfft_check(rocfft_plan_get_work_buffer_size(plan, &wb), "work size");
  if (wb > p->work_bytes) {
    if (p->work) EMC_HIP(hipFree(p->work));
    p->work = nullptr;
    EMC_HIP(hipMalloc(&p->work, wb));
    p->work_bytes = wb;
    rocfft_check(rocfft_execution_info_set_work_buffer(p->info, p->work, p->work_bytes),
                 "set work buffer");
  }
  return plan;
}

static emc_fft_plan* create_plan(emc_ctx* ctx, int nx, int ny, int nz, int batch) {
  use_device(ctx);
  rocfft_init_once();
  EMC_REQUIRE(nx > 0 && ny > 0 && nz > 0 && batch > 0, "bad FFT dims %d %d %d batch %d", nx, ny,
              nz, batch);
  emc_fft_plan* p = new emc_fft_plan();
  p->ctx = ctx;
  p->nx = nx;
  p->ny = ny;
  p->nz = nz;
  p->batch = batch;
  try {
    rocfft_check(rocfft_execution_info_create(&p->info), "execution_info_create");
    rocfft_check(rocfft_execution_info_set_stream(p->info, ctx->stream), "set stream");
    p->r2c = make_plan(p, rocfft_transform_type_real_forward);
    p->c2r = make_plan(p, rocfft_transform_type_real_inverse);
  } catch (...) {
    if (p->r2c) rocfft_plan_destroy(p->r2c);
    if (p->info) rocfft_execution_info_destroy(p->info);
    if (p->work) (void)hipFree(p->work);
    delete p;
    throw;
  }
  return p;
}

emc_fft_plan* cached_plan(emc_ctx* ctx, int nx, int ny, int nz, int batch) {
  auto key = std::make_tuple(nx, ny, nz, batch);
  auto it = ctx->plans.find(key);
  if (it != ctx->plans.end()) return it->second;
  emc_fft_plan* p = create_plan(ctx, nx, ny, nz, batch);
  ctx->plans[key] = p;
  return p;
}

void fft_plan_dims(const emc_fft_plan* p, int dims[3]) {
  dims[0] = p->nx;
  dims[1] = p->ny;
  dims[2] = p->nz;
}

void fft_r2c(emc_fft_plan* p, const float* in, void* out) {
  void* ib[1] = {const_cast<float*>(in)};
  void* ob[1] = {out};
  rocfft_check(rocfft_execute(p->r2c, ib, ob, p->info), "execute r2c");
}

void fft_c2r(emc_fft_plan* p, void* in, float* out) {
  void* ib[1] = {in};
  void* ob[1] = {out};
  rocfft_check(rocfft_execute(p->c2r, ib, ob, p->info), "execute c2r");
}

void fft_c2c(emc_fft_plan* p, const void* in, void* out, int direction) {
  rocfft_plan* slot = direction < 0 ? &p->c2c_fwd : &p->c2c_inv;
  if (!*slot)
    *slot = make_plan(p, direction < 0 ? rocfft_transform_type_complex_forward
                                       : rocfft_transform_type_complex_inverse);
  void* ib[1] = {const_cast<void*>(in)};
  void* ob[1] = {out};
  rocfft_check(rocfft_execute(*slot, ib, ob, p->info), "execute c2c");
}

}  // namespace emc

using namespace emc;

extern "C" {

emc_status emc_fft_plan_create(emc_ctx* ctx, const int dims[3], int batch, emc_fft_plan** plan) {
  return guard([&] {
    EMC_REQUIRE(dims && plan, "null pointer");
    *plan = create_plan(ctx, dims[0], dims[1], dims[2], batch);
  });
}

emc_status emc_fft_plan_destroy(emc_fft_plan* p) {
  return guard([&] {
    if (!p) return;
    use_device(p->ctx);
    (void)hipStreamSynchronize(p->ctx->stream);
    if (p->r2c) rocfft_plan_destroy(p->r2c);
    if (p->c2r) rocfft_plan_destroy(p->c2r);
    if (p->c2c_fwd) rocfft_plan_destroy(p->c2c_fwd);
    if (p->c2c_inv) rocfft_plan_destroy(p->c2c_inv);
    if (p->info) rocfft_execution_info_destroy(p->info);
    if (p->work) (void)hipFree(p->work);
    delete p;
  });
}

emc_status emc_fft_r2c(emc_fft_plan* plan, const float* d_in, void* d_out_c64) {
  return guard([&] {
    EMC_REQUIRE(plan && d_in && d_out_c64, "null pointer");
    use_device(plan->ctx);
    fft_r2c(plan, d_in, d_out_c64);
  });
}

emc_status emc_fft_c2r(emc_fft_plan* plan, void* d_in_c64, float* d_out) {
  return guard([&] {
    EMC_REQUIRE(plan && d_in_c64 && d_out, "null pointer");
    use_device(plan->ctx);
    fft_c2r(plan, d_in_c64, d_out);
  });
}

emc_status emc_fft_c2c(emc_fft_plan* plan, const void* d_in_c64, void* d_out_c64, int direction) {
  return guard([&] {
    EMC_REQUIRE(plan && d_in_c64 && d_out_c64, "null pointer");
    EMC_REQUIRE(direction == -1 || direction == 1, "direction must be -1 or +1");
    use_device(plan->ctx);
    fft_c2c(plan, d_in_c64, d_out_c64, direction);
  });
}

}  // extern "C"
