/*
 * emc_cpu.c -- FP32 CPU restatement of the plain alignment loop body,
 * alignment/BH_alignRaw3d_v2.m:1182-1402 (the two-stage orientation body)
 * for the configuration the alignment workload uses: window = calc size
 * (no padding), no reference weight (refWgtROT = 1, refWGT = 1), C1.
 *
 * TEST INFRASTRUCTURE: this file is part of the oracle. It is the timed CPU
 * baseline of the benchmark and is checked against the FP64 numpy oracle
 * (oracle/emc_oracle.py: align_orientation) by the tests; the library never
 * calls it.
 *
 * Per orientation R (rotation of the particle's current pose):
 *   stage 1  rotate the reference (fwd, R^T, shift est0) and the peak mask,
 *            band-limit-centre-normalise the reference (no bandpass),
 *            xcf_translational(particle_FT .* 1, conj(ref_FT) .* wedge)
 *            -> shift est = peak + est0            (:1287-1344)
 *   stage 2  rotate the particle (inv, R, est) and its wedge, x volMask,
 *            band-limit-centre-normalise with the bandpass,
 *            xcf_rotational(.., ref_FT1, wedge, 1) -> CCC   (:1199-1256, :1375-1387)
 * The particle's stage-1 spectrum and the wedge in FFT order do not depend on
 * R and are computed once per particle (the reference recomputes the
 * former per orientation). Orientations run in parallel (OpenMP), each
 * thread with its own buffers.
 *
 * The 3-D FFT is a mixed-radix (4, 2, 3, 5, generic) Cooley-Tukey transform
 * applied along x, y and z; unnormalised forward, inverse scaled by 1/N as
 * numpy's ifftn.
 */
#include <complex.h>
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

typedef float complex cf;

/* ------------------------------------------------------------------ FFT */

typedef struct {
  int n;
  int nfac;
  int fac[64]; /* (radix, remaining length) pairs */
  cf* tw_fwd;  /* e^{-2 pi i k / n} */
  cf* tw_inv;  /* e^{+2 pi i k / n} */
} fft1d;

static void fft1d_init(fft1d* p, int n) {
  p->n = n;
  p->nfac = 0;
  int m = n;
  const int pref[4] = {4, 2, 3, 5};
  while (m > 1) {
    int r = 0;
    for (int i = 0; i < 4 && !r; ++i)
      if (m % pref[i] == 0) r = pref[i];
    if (!r) {
      r = 7;
      while (m % r) r += 2;
    }
    m /= r;
    p->fac[2 * p->nfac] = r;
    p->fac[2 * p->nfac + 1] = m;
    p->nfac++;
  }
  p->tw_fwd = malloc(sizeof(cf) * (size_t)n);
  p->tw_inv = malloc(sizeof(cf) * (size_t)n);
  for (int k = 0; k < n; ++k) {
    const double a = 2.0 * M_PI * (double)k / (double)n;
    p->tw_fwd[k] = (float)cos(a) - I * (float)sin(a);
    p->tw_inv[k] = (float)cos(a) + I * (float)sin(a);
  }
}

static void fft1d_free(fft1d* p) {
  free(p->tw_fwd);
  free(p->tw_inv);
}

/* out[0 .. r*m) <- DFT of in[0], in[s], in[2s], ... (s = fstride), by
 * recursion over the factor list, then radix-r butterflies with twiddles. */
static void fft_work(cf* out, const cf* in, int fstride, const int* fac, const cf* tw, int n, int inv) {
  const int r = fac[0], m = fac[1];
  if (m == 1) {
    for (int k = 0; k < r; ++k) out[k] = in[(size_t)k * fstride];
  } else {
    for (int k = 0; k < r; ++k) fft_work(out + k * m, in + (size_t)k * fstride, fstride * r, fac + 2, tw, n, inv);
  }
  if (r == 2) {
    for (int u = 0; u < m; ++u) {
      const cf a = out[u], b = out[u + m] * tw[u * fstride];
      out[u] = a + b;
      out[u + m] = a - b;
    }
  } else if (r == 4) {
    for (int u = 0; u < m; ++u) {
      const cf a0 = out[u];
      const cf a1 = out[u + m] * tw[u * fstride];
      const cf a2 = out[u + 2 * m] * tw[2 * u * fstride];
      const cf a3 = out[u + 3 * m] * tw[3 * u * fstride];
      const cf s02 = a0 + a2, d02 = a0 - a2, s13 = a1 + a3;
      /* -i (forward) or +i (inverse) times (a1 - a3) */
      const cf d13 = inv ? I * (a1 - a3) : -I * (a1 - a3);
      out[u] = s02 + s13;
      out[u + m] = d02 + d13;
      out[u + 2 * m] = s02 - s13;
      out[u + 3 * m] = d02 - d13;
    }
  } else {
    cf sc[64];
    for (int u = 0; u < m; ++u) {
      for (int q = 0; q < r; ++q) sc[q] = out[u + q * m] * tw[(size_t)q * u * fstride % n];
      for (int q1 = 0; q1 < r; ++q1) {
        cf acc = 0;
        for (int q = 0; q < r; ++q) acc += sc[q] * tw[(size_t)q * q1 * m * fstride % n];
        out[u + q1 * m] = acc;
      }
    }
  }
}

typedef struct {
  int nx, ny, nz;
  fft1d fx, fy, fz;
} fft3d;

static void fft3d_init(fft3d* p, int nx, int ny, int nz) {
  p->nx = nx;
  p->ny = ny;
  p->nz = nz;
  fft1d_init(&p->fx, nx);
  fft1d_init(&p->fy, ny);
  fft1d_init(&p->fz, nz);
}

static void fft3d_free(fft3d* p) {
  fft1d_free(&p->fx);
  fft1d_free(&p->fy);
  fft1d_free(&p->fz);
}

/* in place; line buffers a, b of max(nx, ny, nz) */
static void fft3d_run(const fft3d* p, cf* v, int inv, cf* a, cf* b) {
  const int nx = p->nx, ny = p->ny, nz = p->nz;
  const cf* twx = inv ? p->fx.tw_inv : p->fx.tw_fwd;
  const cf* twy = inv ? p->fy.tw_inv : p->fy.tw_fwd;
  const cf* twz = inv ? p->fz.tw_inv : p->fz.tw_fwd;
  for (int zy = 0; zy < ny * nz; ++zy) {
    cf* line = v + (size_t)zy * nx;
    memcpy(a, line, sizeof(cf) * nx);
    fft_work(line, a, 1, p->fx.fac, twx, nx, inv);
  }
  for (int z = 0; z < nz; ++z)
    for (int x = 0; x < nx; ++x) {
      cf* base = v + (size_t)z * nx * ny + x;
      for (int y = 0; y < ny; ++y) a[y] = base[(size_t)y * nx];
      fft_work(b, a, 1, p->fy.fac, twy, ny, inv);
      for (int y = 0; y < ny; ++y) base[(size_t)y * nx] = b[y];
    }
  for (int yx = 0; yx < nx * ny; ++yx) {
    cf* base = v + yx;
    for (int z = 0; z < nz; ++z) a[z] = base[(size_t)z * nx * ny];
    fft_work(b, a, 1, p->fz.fac, twz, nz, inv);
    for (int z = 0; z < nz; ++z) base[(size_t)z * nx * ny] = b[z];
  }
  if (inv) {
    const float s = 1.0f / ((float)nx * ny * nz);
    for (size_t i = 0; i < (size_t)nx * ny * nz; ++i) v[i] *= s;
  }
}

/* ------------------------------------------------------------- volumes */

static float trilinear_zero(const float* v, int nx, int ny, int nz, float x, float y, float z) {
  const float xf = floorf(x), yf = floorf(y), zf = floorf(z);
  const float fx = x - xf, fy = y - yf, fz = z - zf;
  const int x0 = (int)xf, y0 = (int)yf, z0 = (int)zf;
  float acc = 0.0f;
  for (int dz = 0; dz < 2; ++dz) {
    const int zi = z0 + dz;
    if (zi < 0 || zi >= nz) continue;
    const float wz = dz ? fz : 1.0f - fz;
    for (int dy = 0; dy < 2; ++dy) {
      const int yi = y0 + dy;
      if (yi < 0 || yi >= ny) continue;
      const float wy = dy ? fy : 1.0f - fy;
      for (int dx = 0; dx < 2; ++dx) {
        const int xi = x0 + dx;
        if (xi < 0 || xi >= nx) continue;
        const float wx = dx ? fx : 1.0f - fx;
        acc += wx * wy * wz * v[((size_t)zi * ny + yi) * nx + xi];
      }
    }
  }
  return acc;
}

/* mexXform3d semantics: fwd p = M (u - t), inv p = M u + t, u = index - n/2,
 * M row-major; sampled at p + n/2, zero outside */
static void xform(const float* in, float* out, int n, const double* M, const double* t, int fwd) {
  const float c = (float)(n / 2);
  float m[9];
  for (int i = 0; i < 9; ++i) m[i] = (float)M[i];
  for (int z = 0; z < n; ++z)
    for (int y = 0; y < n; ++y)
      for (int x = 0; x < n; ++x) {
        float u = (float)x - c, v = (float)y - c, w = (float)z - c;
        float px, py, pz;
        if (fwd) {
          u -= (float)t[0];
          v -= (float)t[1];
          w -= (float)t[2];
          px = m[0] * u + m[1] * v + m[2] * w;
          py = m[3] * u + m[4] * v + m[5] * w;
          pz = m[6] * u + m[7] * v + m[8] * w;
        } else {
          px = m[0] * u + m[1] * v + m[2] * w + (float)t[0];
          py = m[3] * u + m[4] * v + m[5] * w + (float)t[1];
          pz = m[6] * u + m[7] * v + m[8] * w + (float)t[2];
        }
        out[((size_t)z * n + y) * n + x] = trilinear_zero(in, n, n, n, px + c, py + c, pz + c);
      }
}

/* BH_bandLimitCenterNormalize without padding (masking/BH_bandLimitCenterNormalize.m:48-61):
 * fftn, F(0) = 0, x bandpass (NULL = none), F / (sqrt(sum |F|^2) / numel) */
static void blcn(const float* img, const float* bp, int n, const fft3d* f, cf* F, cf* a, cf* b) {
  const size_t N = (size_t)n * n * n;
  for (size_t i = 0; i < N; ++i) F[i] = img[i];
  fft3d_run(f, F, 0, a, b);
  F[0] = 0;
  double s = 0.0;
  for (size_t i = 0; i < N; ++i) {
    if (bp) F[i] *= bp[i];
    s += (double)crealf(F[i]) * crealf(F[i]) + (double)cimagf(F[i]) * cimagf(F[i]);
  }
  const float k = (float)((double)N / sqrt(s));
  for (size_t i = 0; i < N; ++i) F[i] *= k;
}

static double mround(double x) { return x >= 0 ? floor(x + 0.5) : ceil(x - 0.5); }

/* BH_multi_xcf_Translational (statistics/BH_multi_xcf_Translational.m:31-103),
 * peak mask > 0.95 kept, COM box of half-width com (all three nonzero) */
static int xcf_translational(cf* AB, const float* mask, int n, const int com[3], const fft3d* f, float* M, cf* a,
                             cf* b, double out[3]) {
  const size_t N = (size_t)n * n * n;
  fft3d_run(f, AB, 1, a, b);
  const int h = n / 2;
  for (int z = 0; z < n; ++z)
    for (int y = 0; y < n; ++y)
      for (int x = 0; x < n; ++x) {
        const int sz = (z + h) % n, sy = (y + h) % n, sx = (x + h) % n;  /* fftshift */
        M[((size_t)sz * n + sy) * n + sx] = crealf(AB[((size_t)z * n + y) * n + x]);
      }
  float mn = M[0];
  for (size_t i = 1; i < N; ++i) mn = M[i] < mn ? M[i] : mn;
  size_t best = 0;
  float mx = -INFINITY;
  for (size_t i = 0; i < N; ++i) {
    if (mask && mask[i] < 0.95f) M[i] = mn;
    if (M[i] > mx) {
      mx = M[i];
      best = i;
    }
  }
  double peak[3] = {(double)(best % n) + 1, (double)((best / n) % n) + 1, (double)(best / ((size_t)n * n)) + 1};
  if (!isnan(mx) && mx != 0.0f) {
    for (int icom = 1; icom <= 2; ++icom) {
      double pc[3];
      for (int k = 0; k < 3; ++k) pc[k] = icom == 1 ? peak[k] : mround(peak[k]);
      int lo[3], hi[3];
      for (int k = 0; k < 3; ++k) {
        lo[k] = (int)(pc[k] - com[k]);
        hi[k] = (int)(pc[k] + com[k]);
        if (lo[k] < 1 || hi[k] > n) return -1;
      }
      double s = 0, cx = 0, cy = 0, cz = 0;
      for (int z = lo[2]; z <= hi[2]; ++z)
        for (int y = lo[1]; y <= hi[1]; ++y)
          for (int x = lo[0]; x <= hi[0]; ++x) {
            const double v = (double)M[((size_t)(z - 1) * n + (y - 1)) * n + (x - 1)] - mn;
            s += v;
            cx += v * (x - pc[0]);
            cy += v * (y - pc[1]);
            cz += v * (z - pc[2]);
          }
      peak[0] = pc[0] + cx / s;
      peak[1] = pc[1] + cy / s;
      peak[2] = pc[2] + cz / s;
    }
  } else {
    peak[0] = peak[1] = peak[2] = 0;
  }
  for (int k = 0; k < 3; ++k) {
    out[k] = peak[k] - ceil((n + 1) / 2.0);
    if (isnan(out[k])) out[k] = 0;
  }
  return 0;
}

/* ------------------------------------------------------------ the loop */

/* All volumes n^3, x fastest. particle, ref, vol_mask, peak_mask (centred),
 * bandpass (FFT order), wedge (centred). A: the particle's 3x3 (row-major).
 * grid: n_orient (phi, theta, psi); R = A * defineMatrix([phi, theta,
 * psi - phi], 'Bah', 'inv') is formed by the caller: mats holds n_orient
 * row-major 3x3. est0: initial shift. Outputs ccc[n_orient], est[3*n_orient].
 * Returns 0, or -1 if a COM box fell outside a map. */
int emc_cpu_align_grid(int n, const float* particle, const float* ref, const float* vol_mask,
                       const float* peak_mask, const float* bandpass, const float* wedge, const double* mats,
                       int n_orient, const double* est0, int threads, double* ccc, double* est) {
  const size_t N = (size_t)n * n * n;
  const int com[3] = {3, 3, 3};
  fft3d f;
  fft3d_init(&f, n, n, n);
  /* per particle: stage-1 particle spectrum (with bandpass), ref_FT1 =
   * conj(blcn(ref .* volMask)), wedge in FFT order */
  cf* part_ft = malloc(sizeof(cf) * N);
  cf* ref_ft1 = malloc(sizeof(cf) * N);
  float* wedge_ifft = malloc(sizeof(float) * N);
  float* tmp = malloc(sizeof(float) * N);
  cf* la = malloc(sizeof(cf) * (size_t)n);
  cf* lb = malloc(sizeof(cf) * (size_t)n);
  blcn(particle, bandpass, n, &f, part_ft, la, lb);
  for (size_t i = 0; i < N; ++i) tmp[i] = ref[i] * vol_mask[i];
  blcn(tmp, NULL, n, &f, ref_ft1, la, lb);
  for (size_t i = 0; i < N; ++i) ref_ft1[i] = conjf(ref_ft1[i]);
  const int h = (n + 1) / 2; /* ifftshift: centred index (k + n/2) -> FFT index k */
  for (int z = 0; z < n; ++z)
    for (int y = 0; y < n; ++y)
      for (int x = 0; x < n; ++x)
        wedge_ifft[((size_t)z * n + y) * n + x] =
            wedge[((size_t)((z + n - h) % n) * n + (y + n - h) % n) * n + (x + n - h) % n];
  free(tmp);
  free(la);
  free(lb);
  int status = 0;
  if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel
  {
    float* rref = malloc(sizeof(float) * N);
    float* rmask = malloc(sizeof(float) * N);
    float* rwedge = malloc(sizeof(float) * N);
    float* M = malloc(sizeof(float) * N);
    cf* F = malloc(sizeof(cf) * N);
    cf* G = malloc(sizeof(cf) * N);
    cf* a = malloc(sizeof(cf) * (size_t)n);
    cf* b = malloc(sizeof(cf) * (size_t)n);
#pragma omp for schedule(dynamic, 1)
    for (int o = 0; o < n_orient; ++o) {
      const double* R = mats + 9 * (size_t)o;
      double RT[9];
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) RT[r * 3 + c] = R[c * 3 + r];
      /* stage 1 */
      xform(ref, rref, n, RT, est0, 1);
      xform(peak_mask, rmask, n, RT, (const double[3]){0, 0, 0}, 1);
      blcn(rref, NULL, n, &f, F, a, b);
      for (size_t i = 0; i < N; ++i) G[i] = part_ft[i] * (conjf(F[i]) * wedge_ifft[i]);
      double pk[3];
      if (xcf_translational(G, rmask, n, com, &f, M, a, b, pk) != 0) {
#pragma omp atomic write
        status = -1;
        pk[0] = pk[1] = pk[2] = 0;
      }
      double e[3] = {pk[0] + est0[0], pk[1] + est0[1], pk[2] + est0[2]};
      /* stage 2: rotate the particle and its wedge */
      xform(particle, rref, n, R, e, 0);
      xform(wedge, rwedge, n, R, (const double[3]){0, 0, 0}, 0);
      for (size_t i = 0; i < N; ++i) rref[i] *= vol_mask[i];
      blcn(rref, bandpass, n, &f, F, a, b);
      double num = 0, dr = 0, dp = 0;
      for (int z = 0; z < n; ++z)
        for (int y = 0; y < n; ++y)
          for (int x = 0; x < n; ++x) {
            const size_t i = ((size_t)z * n + y) * n + x;
            const float w = rwedge[((size_t)((z + n - h) % n) * n + (y + n - h) % n) * n + (x + n - h) % n];
            const cf p = F[i] * w;   /* P .* Wp, Wr = 1 */
            const cf rr = ref_ft1[i];
            num += crealf(p * rr);
            dr += (double)crealf(rr * w) * crealf(rr * w) + (double)cimagf(rr * w) * cimagf(rr * w);
            dp += (double)crealf(F[i]) * crealf(F[i]) + (double)cimagf(F[i]) * cimagf(F[i]);
          }
      ccc[o] = num / sqrt(dr * dp);
      est[3 * (size_t)o] = e[0];
      est[3 * (size_t)o + 1] = e[1];
      est[3 * (size_t)o + 2] = e[2];
    }
    free(rref);
    free(rmask);
    free(rwedge);
    free(M);
    free(F);
    free(G);
    free(a);
    free(b);
  }
  free(part_ft);
  free(ref_ft1);
  free(wedge_ifft);
  fft3d_free(&f);
  return status;
}

int emc_cpu_max_threads(void) { return omp_get_max_threads(); }
