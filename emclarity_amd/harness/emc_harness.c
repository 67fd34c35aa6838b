/*
 * Plain C harness for libemc_hip.so: drives the C ABI without MATLAB or
 * Python. Builds a seeded 64^3 Gaussian-blob reference, makes a particle by
 * rotating and shifting it with emc_xform3d, and recovers the shift with
 * the translational cross-correlation (emc_xcf_translational) and the best
 * orientation of a small grid with emc_align_grid, printing timings.
 *
 * Build: gcc -O2 -I include emclarity_amd/harness/emc_harness.c \
 *            -L emclarity_amd/lib -lemc_hip -Wl,-rpath,$PWD/emclarity_amd/lib -lm
 * Exit status 0 when the recovered integer shift is within 0.1 px and the grid
 * search returns the true orientation.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "emc_hip.h"

#define CHECK(call)                                                              \
  do {                                                                           \
    emc_status s_ = (call);                                                      \
    if (s_ != EMC_OK) {                                                          \
      fprintf(stderr, "%s failed (%d): %s\n", #call, (int)s_, emc_last_error()); \
      exit(1);                                                                   \
    }                                                                            \
  } while (0)

static unsigned long long g_state = 88172645463325252ull;
static double urand(void) {
  g_state ^= g_state << 13;
  g_state ^= g_state >> 7;
  g_state ^= g_state << 17;
  return (double)(g_state >> 11) / 9007199254740992.0;
}

static double now_s(void) {
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return t.tv_sec + 1e-9 * t.tv_nsec;
}

int main(void) {
  const int n = 64, nv = n * n * n, nh = (n / 2 + 1) * n * n;
  const int dims[3] = {n, n, n};
  emc_ctx* ctx = NULL;
  CHECK(emc_ctx_create(0, &ctx));
  printf("%s\n", emc_version());

  /* 30 Gaussian blobs, sigma 2-4 px, inside radius 18 */
  float* h = (float*)calloc(nv, sizeof(float));
  for (int b = 0; b < 30; ++b) {
    double c[3], r2;
    do {
      for (int k = 0; k < 3; ++k) c[k] = (urand() * 2 - 1) * 18.0;
      r2 = c[0] * c[0] + c[1] * c[1] + c[2] * c[2];
    } while (r2 > 18.0 * 18.0);
    const double s = 2.0 + 2.0 * urand(), a = 0.5 + 0.5 * urand();
    for (int z = 0; z < n; ++z)
      for (int y = 0; y < n; ++y)
        for (int x = 0; x < n; ++x) {
          const double dx = x - n / 2 - c[0], dy = y - n / 2 - c[1], dz = z - n / 2 - c[2];
          h[(z * n + y) * n + x] += (float)(a * exp(-(dx * dx + dy * dy + dz * dz) / (2 * s * s)));
        }
  }
  float *d_ref, *d_part, *d_ones, *d_mask, *d_bp;
  void *d_fa, *d_fb, *d_full_a, *d_full_b, *d_work, *d_ft1;
  CHECK(emc_malloc(ctx, nv * 4, (void**)&d_ref));
  CHECK(emc_malloc(ctx, nv * 4, (void**)&d_part));
  CHECK(emc_malloc(ctx, nv * 4, (void**)&d_ones));
  CHECK(emc_malloc(ctx, nv * 4, (void**)&d_mask));
  CHECK(emc_malloc(ctx, nv * 4, (void**)&d_bp));
  CHECK(emc_malloc(ctx, (size_t)nh * 8, &d_fa));
  CHECK(emc_malloc(ctx, (size_t)nh * 8, &d_fb));
  CHECK(emc_malloc(ctx, (size_t)nh * 8, &d_ft1));
  CHECK(emc_malloc(ctx, (size_t)nv * 8, &d_full_a));
  CHECK(emc_malloc(ctx, (size_t)nv * 8, &d_full_b));
  CHECK(emc_malloc(ctx, (size_t)nv * 16, &d_work));
  CHECK(emc_memcpy_h2d(ctx, d_ref, h, nv * 4));

  /* particle: reference rotated by the grid orientation (30, 20, 40) and
   * shifted by (1, -2, 1): RotMat = BH_defineMatrix([phi,theta,psi-phi],'Bah','inv'),
   * particle(x) = ref(RotMat^T (x - t)) = xform fwd with R = RotMat^T ... */
  const double truth[3] = {30.0, 20.0, 40.0};
  const double ang[3] = {truth[0], truth[1], truth[2] - truth[0]};
  double Rm[9];
  CHECK(emc_define_matrix(ang, 0, 1, Rm));
  float RT[9];
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r) RT[c * 3 + r] = (float)Rm[r * 3 + c]; /* transpose, col-major */
  const float t[3] = {1.0f, -2.0f, 1.0f};
  double t0 = now_s();
  CHECK(emc_xform3d(ctx, d_ref, dims, RT, t, 1, d_part));
  CHECK(emc_ctx_synchronize(ctx));
  printf("xform3d 64^3: %.3f ms (first call)\n", 1e3 * (now_s() - t0));

  /* translational xcf between the particle and the reference rotated back */
  float* ones = (float*)malloc(nv * 4);
  for (int i = 0; i < nv; ++i) ones[i] = 1.0f;
  CHECK(emc_memcpy_h2d(ctx, d_ones, ones, nv * 4));
  emc_fft_plan* plan;
  CHECK(emc_fft_plan_create(ctx, dims, 1, &plan));
  float* rot = d_mask; /* scratch: reference rotated like the particle, unshifted */
  const float zero[3] = {0, 0, 0};
  CHECK(emc_xform3d(ctx, d_ref, dims, RT, zero, 1, rot));
  /* full complex spectra via R2C + host Hermitian expansion would be long;
   * use the C2C plan on complex copies instead */
  float* hc = (float*)malloc((size_t)nv * 8);
  float* hr = (float*)malloc((size_t)nv * 4);
  CHECK(emc_memcpy_d2h(ctx, hr, d_part, nv * 4));
  for (int i = 0; i < nv; ++i) hc[2 * i] = hr[i], hc[2 * i + 1] = 0.0f;
  CHECK(emc_memcpy_h2d(ctx, d_work, hc, (size_t)nv * 8));
  CHECK(emc_fft_c2c(plan, d_work, d_full_a, -1));
  CHECK(emc_memcpy_d2h(ctx, hr, rot, nv * 4));
  for (int i = 0; i < nv; ++i) hc[2 * i] = hr[i], hc[2 * i + 1] = 0.0f;
  CHECK(emc_memcpy_h2d(ctx, d_work, hc, (size_t)nv * 8));
  CHECK(emc_fft_c2c(plan, d_work, d_full_b, -1));
  /* conj(B) on the host */
  CHECK(emc_memcpy_d2h(ctx, hc, d_full_b, (size_t)nv * 8));
  for (int i = 0; i < nv; ++i) hc[2 * i + 1] = -hc[2 * i + 1];
  CHECK(emc_memcpy_h2d(ctx, d_full_b, hc, (size_t)nv * 8));
  const int com[3] = {3, 3, 3};
  double peak[3], mx;
  t0 = now_s();
  CHECK(emc_xcf_translational(ctx, plan, d_full_a, d_full_b, dims, NULL, com, d_work, peak, &mx));
  printf("xcf_translational: peak (%.3f, %.3f, %.3f) expected (1, -2, 1), %.3f ms\n", peak[0],
         peak[1], peak[2], 1e3 * (now_s() - t0));
  int ok = fabs(peak[0] - 1.0) < 0.1 && fabs(peak[1] + 2.0) < 0.1 && fabs(peak[2] - 1.0) < 0.1;

  /* plain grid search over 5 orientations, the truth among them */
  CHECK(emc_bandpass3d(ctx, dims, 0, 1e-3, 40.0, 8.0, 2.5, d_bp));
  CHECK(emc_ref_ft1(ctx, plan, d_ref, NULL, dims, (const int[3]){0, 0, 0}, (const int[3]){0, 0, 0}, d_ft1));
  float* sph = (float*)malloc(nv * 4);
  for (int z = 0; z < n; ++z)
    for (int y = 0; y < n; ++y)
      for (int x = 0; x < n; ++x) {
        const int dx = x - n / 2, dy = y - n / 2, dz = z - n / 2;
        sph[(z * n + y) * n + x] = (dx * dx + dy * dy + dz * dz <= 12 * 12) ? 1.0f : 0.0f;
      }
  CHECK(emc_memcpy_h2d(ctx, d_mask, sph, nv * 4));
  const double grid[15] = {0, 0, 0, 30, 20, 40, 60, 20, 40, 30, 40, 40, 30, 20, 10};
  const double I9[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, z3[3] = {0, 0, 0};
  double ccc[5], est[15];
  t0 = now_s();
  CHECK(emc_align_grid(ctx, n, d_part, d_ref, NULL, NULL, d_ft1, NULL, d_mask, d_bp, d_ones, I9, z3, com,
                       grid, 5, 3, ccc, est));
  int best = 0;
  for (int i = 1; i < 5; ++i)
    if (ccc[i] > ccc[best]) best = i;
  printf("align_grid (5 orientations): %.3f ms, best %d (ccc %.4f), shift (%.2f, %.2f, %.2f)\n",
         1e3 * (now_s() - t0), best, ccc[best], est[3 * best], est[3 * best + 1], est[3 * best + 2]);
  ok = ok && best == 1;

  emc_fft_plan_destroy(plan);
  emc_free(ctx, d_ref);
  emc_free(ctx, d_part);
  emc_free(ctx, d_ones);
  emc_free(ctx, d_mask);
  emc_free(ctx, d_bp);
  emc_free(ctx, d_fa);
  emc_free(ctx, d_fb);
  emc_free(ctx, d_ft1);
  emc_free(ctx, d_full_a);
  emc_free(ctx, d_full_b);
  emc_free(ctx, d_work);
  emc_ctx_destroy(ctx);
  free(h);
  free(ones);
  free(hc);
  free(hr);
  free(sph);
  printf(ok ? "harness OK\n" : "harness FAILED\n");
  return ok ? 0 : 1;
}
