// Batched alignment of many subtomograms x many orientations in one launch,
// and the hand-written adjoint (score gradient w.r.t. 3 Euler angles and 3
// shifts) used for gradient refinement.
//
// This path has no counterpart in the reference: it is the batched form of
// the rotate-weight-correlate chain that alignment/BH_alignRaw3d_v2.m runs
// one orientation and one particle at a time (:1119-1419). The reference
// rotates in real space (mexXform3d), then FFTs, multiplies and inverse-FFTs
// the full volume for every orientation. Here the reference is rotated in
// Fourier space instead: its spectrum, oversampled 2x by zero padding, is
// computed once per box size and stays resident in L2 / MALL; for each
// (particle, orientation) one workgroup
//   1. samples the rotated reference spectrum F(2 R^T k) trilinearly for
//      every frequency k inside the bandpass support (half grid),
//   2. multiplies by the particle's prepared spectrum
//      Q = P_centred * bandpass^2 * wedge, accumulates the normalisation
//      sum |F|^2 * wedge^2 * bandpass^2 in FP64,
//   3. runs a pruned inverse DFT that evaluates the correlation only on the
//      (2r+1)^3 cube of candidate shifts (z, then y, then the Hermitian x
//      sum), entirely in registers and LDS,
//   4. takes the peak inside the sphere |d| <= r with parabolic sub-pixel
//      refinement, and writes the normalised CC and the shift.
// No FFT and no per-orientation volume: a unit writes a few bytes of
// results. What it does read is the reference table, by trilinear gathers
// (a 2x2x2 footprint per frequency) that the units running together on an
// XCD share only in part; the traffic past L2 from those gathers, not the
// FP32 arithmetic, is what bounds the kernel. The reference table can be
// complex64 or fp16 parts (half the gather bytes), and the frequencies are
// visited in compact (kx, ky) patches so that nearby orientations read
// nearby parts of the table at the same time.
//
// CC definition (Parseval form of the normalised real-space CC of the
// band-passed particle and the wedge-weighted rotated reference):
//   a(k) = P(k) b(k),   c_o(k) = F(2 R_o^T k) S(k) b(k)
//   CC_o(s) = sum_k Re[a(k) conj(c_o(k)) e^{2 pi i k.s/n}]
//             / sqrt(sum_k |a|^2 * sum_k |c_o|^2)
// with b the bandpass, S the particle's SF3D wedge symmetrised over k/-k so
// that the product is Hermitian and the half grid is exact.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <tuple>

#include "ops.hpp"

namespace emc {

void define_matrix(const double ang_deg[3], int convention, int direction, double out_colmajor[9]);

constexpr int kBatchThreads = 256;
constexpr int kMaxN = 128;

// --------------------------------------------------------------- preparation

// Reference of edge n, zero-padded to 2n with its centre n/2 moved to the
// origin (ifftshift of the centred padding).
__global__ void pad_origin_kernel(const float* __restrict__ ref, float* __restrict__ out, int n) {
  const int m = 2 * n;
  const int X = blockIdx.x * blockDim.x + threadIdx.x;
  const int Y = blockIdx.y * blockDim.y + threadIdx.y;
  const int Z = blockIdx.z;
  if (X >= m || Y >= m) return;
  // signed offset of this output voxel from the origin
  const int ux = X < n ? X : X - m, uy = Y < n ? Y : Y - m, uz = Z < n ? Z : Z - m;
  const int x = ux + n / 2, y = uy + n / 2, z = uz + n / 2;
  float v = 0.0f;
  if ((unsigned)x < (unsigned)n && (unsigned)y < (unsigned)n && (unsigned)z < (unsigned)n)
    v = ref[((int64_t)z * n + y) * n + x];
  out[((int64_t)Z * m + Y) * m + X] = v;
}

// Particles of one preparation chunk (kernel-argument pointer table): up to
// 64 (1 KiB of kernel arguments); prep_chunk picks the chunk per box size.
constexpr int kPrepChunk = 64;

// Particles per preparation chunk: 16 at 128^3 and above, 32 up to 96^3, 64
// up to 64^3, so a chunk's batched transform is ~32 MB of input or more.
inline int prep_chunk(int n) { return n > 96 ? 16 : (n > 64 ? 32 : 64); }
struct PrepPtrs {
  const float* part[kPrepChunk];
  const float* wedge[kPrepChunk];  // centred SF3D or null (= 1)
};

// Each chunk volume into a contiguous batch for the R2C, times the centred
// real-space mask when one is given (the volMask the reference multiplies
// the rotated particle by, BH_alignRaw3d_v2.m:1375; a mask symmetric about
// the centre commutes with the rotation). shift: the ifftshift (centre
// n/2 -> origin) as well, for odd n; for even n the shift is the factor
// (-1)^(kx + ky + kz) of the spectrum, which prep_particle_kernel applies,
// so the volume is copied in place order (or, unmasked and already
// contiguous, not copied at all). blockIdx.z = z + n * particle.
__global__ void ifftshift_many_kernel(PrepPtrs pp, const float* __restrict__ mask, float* __restrict__ out, int n,
                                      int shift) {
  const int X = blockIdx.x * blockDim.x + threadIdx.x;
  const int Y = blockIdx.y * blockDim.y + threadIdx.y;
  const int Z = blockIdx.z % n, b = blockIdx.z / n;
  if (X >= n || Y >= n) return;
  const int x = shift ? shifted_index(X, n) : X, y = shift ? shifted_index(Y, n) : Y;
  const int z = shift ? shifted_index(Z, n) : Z;
  const int64_t c = ((int64_t)z * n + y) * n + x;
  const float v = pp.part[b][c];
  out[(((int64_t)b * n + Z) * n + Y) * n + X] = mask ? v * mask[c] : v;
}

// Q, N in the fused kernel's layout [kz][kx][ky] (ky fastest, coalesced for
// threads running along ky); per-block partials of A = sum_full |P b|^2.
// flip: P is the spectrum of the unshifted volume (even n): the ifftshift's
// factor e^{2 pi i k (n/2) / n} = (-1)^(kx + ky + kz) is applied here.
// blockIdx.y = particle of the chunk; spectra, Q, N and partials are
// stacked per particle. The spectrum ([kz][ky][kx]) is read and Q, N are
// written through a 16 x 16 (kx, ky) tile in LDS, so both sides are
// coalesced.
constexpr int kPrepTile = 16;

__global__ void __launch_bounds__(kRedThreads) prep_particle_kernel(
    const float2* __restrict__ P, const float* __restrict__ bp, PrepPtrs pp, int n, int flip,
    float2* __restrict__ Q, float* __restrict__ N, double* __restrict__ part) {
  static_assert(kPrepTile * kPrepTile == kRedThreads, "one tile per workgroup");
  __shared__ float2 sq[kPrepTile][kPrepTile + 1];
  __shared__ float sn[kPrepTile][kPrepTile + 1];
  const int cx = n / 2 + 1;
  const int64_t nh = (int64_t)cx * n * n;
  const int b = blockIdx.y;
  const float* __restrict__ wedge_c = pp.wedge[b];
  P += b * nh;
  Q += b * nh;
  N += b * nh;
  part += (int64_t)b * gridDim.x;
  const int tx = threadIdx.x % kPrepTile, ty = threadIdx.x / kPrepTile;
  const int nxt = (cx + kPrepTile - 1) / kPrepTile, nyt = (n + kPrepTile - 1) / kPrepTile;
  const int ntiles = n * nyt * nxt;
  double v[1] = {0.0};
  for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {  // uniform per workgroup
    const int xt = tile % nxt, r = tile / nxt;
    const int yt = r % nyt, z = r / nyt;
    const int x = xt * kPrepTile + tx, y = yt * kPrepTile + ty;
    if (x < cx && y < n) {
      const int64_t i = ((int64_t)z * n + y) * cx + x;
      const float bb = bp[((int64_t)z * n + y) * n + x];
      const int64_t cp = ((int64_t)shifted_index(z, n) * n + shifted_index(y, n)) * n + shifted_index(x, n);
      const int mx = mirror_index(x, n), my = mirror_index(y, n), mz = mirror_index(z, n);
      const int64_t cm = ((int64_t)shifted_index(mz, n) * n + shifted_index(my, n)) * n + shifted_index(mx, n);
      const float S = wedge_c ? 0.5f * (wedge_c[cp] + wedge_c[cm]) : 1.0f;
      float2 p = P[i];
      if (flip && ((x + y + z) & 1)) p = make_float2(-p.x, -p.y);
      if (i == 0) p = make_float2(0.0f, 0.0f);  // mean removed
      const float w = bb * bb * S;
      sq[ty][tx] = make_float2(p.x * w, p.y * w);
      sn[ty][tx] = (i == 0) ? 0.0f : S * S * bb * bb;
      const double m = (x == 0 || (n % 2 == 0 && x == n / 2)) ? 1.0 : 2.0;
      v[0] += m * ((double)p.x * p.x + (double)p.y * p.y) * (double)bb * bb;
    }
    __syncthreads();
    // transposed write: this thread stores (kx = xt*16 + ty, ky = yt*16 + tx)
    const int xo = xt * kPrepTile + ty, yo = yt * kPrepTile + tx;
    if (xo < cx && yo < n) {
      const int64_t o = ((int64_t)z * cx + xo) * n + yo;
      Q[o] = sq[tx][ty];
      N[o] = sn[tx][ty];
    }
    __syncthreads();
  }
  block_sum<1>(v);
  if (threadIdx.x == 0) part[blockIdx.x] = v[0];
}

// ---------------------------------------------------------- rotation matrices

// R = A * BH_defineMatrix([phi, theta, psi - phi], 'Bah', 'inv') in FP64 on
// the device (alignment/BH_alignRaw3d_v2.m:1175-1176), stored as the FP32
// sampling matrix M = 2 R^T (row-major) per (particle, orientation).
__device__ inline void rz_d(double a, double m[3][3]) {
  double s, c;
  sincos(a, &s, &c);
  m[0][0] = c; m[0][1] = -s; m[0][2] = 0;
  m[1][0] = s; m[1][1] = c;  m[1][2] = 0;
  m[2][0] = 0; m[2][1] = 0;  m[2][2] = 1;
}
__device__ inline void rx_d(double a, double m[3][3]) {
  double s, c;
  sincos(a, &s, &c);
  m[0][0] = 1; m[0][1] = 0; m[0][2] = 0;
  m[1][0] = 0; m[1][1] = c; m[1][2] = -s;
  m[2][0] = 0; m[2][1] = s; m[2][2] = c;
}
__device__ inline void mm_d(const double a[3][3], const double b[3][3], double o[3][3]) {
  double t[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) t[i][j] = a[i][0] * b[0][j] + a[i][1] * b[1][j] + a[i][2] * b[2][j];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) o[i][j] = t[i][j];
}

// Local matrix of (phi, theta, psi): inv reverses [phi, theta, psi-phi]
// -> Rz(phi) Rx(theta) Rz(psi - phi).
__device__ inline void local_mat(double phi, double theta, double psi, double L[3][3]) {
  const double d2r = 3.14159265358979323846 / 180.0;
  double a[3][3], b[3][3], c[3][3];
  rz_d(phi * d2r, a);
  rx_d(theta * d2r, b);
  rz_d((psi - phi) * d2r, c);
  mm_d(a, b, L);
  mm_d(L, c, L);
}

// Best orientation of each particle (the first maximum of its scores, as
// numpy's argmax; NaN lowest) and the refinement's start pose R0 = A *
// local(pts[best]) (row-major, FP64), s0 = the shift found with it.
__global__ void __launch_bounds__(kRedThreads) pick_kernel(const float* __restrict__ score,
                                                          const float* __restrict__ shift,
                                                          const double* __restrict__ start_cm,
                                                          const double* __restrict__ pts, int n_orient,
                                                          double* __restrict__ R0, double* __restrict__ s0,
                                                          int* __restrict__ best, float* __restrict__ best_score) {
  const int p = blockIdx.x;
  const float* sc = score + (int64_t)p * n_orient;
  unsigned long long key = 0ull;
  for (int o = threadIdx.x; o < n_orient; o += blockDim.x) {
    const unsigned long long k = argmax_key(sc[o], (uint32_t)o);
    key = k > key ? k : key;
  }
  key = block_max_u64(key);  // valid in thread 0
  if (threadIdx.x != 0) return;
  int o = (int)key_index(key);
  if (o < 0 || o >= n_orient) o = 0;
  double A[3][3], L[3][3], R[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) A[r][c] = start_cm[9 * p + c * 3 + r];
  local_mat(pts[3 * o], pts[3 * o + 1], pts[3 * o + 2], L);
  mm_d(A, L, R);
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) R0[9 * p + 3 * r + c] = R[r][c];
  for (int k = 0; k < 3; ++k) s0[3 * p + k] = (double)shift[((int64_t)p * n_orient + o) * 3 + k];
  if (best) best[p] = o;
  if (best_score) best_score[p] = sc[o];
}

__global__ void compose_kernel(const double* __restrict__ start_cm, const double* __restrict__ pts,
                               int n_part, int n_orient, float* __restrict__ M) {
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= n_part * n_orient) return;
  const int p = u / n_orient, o = u % n_orient;
  double A[3][3], L[3][3], R[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) A[r][c] = start_cm[9 * p + c * 3 + r];
  local_mat(pts[3 * o], pts[3 * o + 1], pts[3 * o + 2], L);
  mm_d(A, L, R);
  float* m = M + 9 * (int64_t)u;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) m[3 * r + c] = (float)(2.0 * R[c][r]);  // 2 R^T
}

// ------------------------------------------------------------- sampling

// Reference table formats: complex64 texels, or complex with fp16 parts
// (half the bytes per gather). The fp16 table holds F * s with s = 2^(14-e),
// max |Re|,|Im| of F < 2^e, so every texel is below 2^14; the normalised
// correlation and its gradient do not depend on s.
typedef _Float16 half2_t __attribute__((ext_vector_type(2)));

// Table entry (x, y, z) holds the four texels (x, y, z), (x, y+1, z),
// (x, y, z+1), (x, y+1, z+1); entries x and x+1 are adjacent, so a trilinear
// footprint (x0..x0+1, y0..y0+1, z0..z0+1) is one contiguous octet: 64 bytes
// (complex64) or 32 bytes (fp16) inside one 128-byte cache line for all but
// the octets that start in a line's last quarter. Octet order: x0 (y0z0,
// y1z0, y0z1, y1z1), then x1 likewise.
struct TabC64 {
  using T = float2;
  struct alignas(16) Oct {
    float2 t[8];
  };
  __device__ static void store(T* dst, float2 v, float) { *dst = v; }
  __device__ static Oct load(const T* p) { return *reinterpret_cast<const Oct*>(p); }
  __device__ static float2 texel(const Oct& o, int i) { return o.t[i]; }
};

struct TabF16 {
  using T = half2_t;
  struct alignas(16) Oct {  // 32 bytes, two loads
    half2_t t[8];
  };
  __device__ static void store(T* dst, float2 v, float s) {
    *dst = half2_t{(_Float16)(v.x * s), (_Float16)(v.y * s)};
  }
  __device__ static Oct load(const T* p) { return *reinterpret_cast<const Oct*>(p); }
  __device__ static float2 texel(const Oct& o, int i) { return make_float2((float)o.t[i].x, (float)o.t[i].y); }
};

// Scale of the fp16 table: per-block maxima of max(|Re|, |Im|) (as
// order-preserving bits of non-negative floats), then s = 2^(14 - e).
__global__ void __launch_bounds__(kRedThreads) f2_maxabs_kernel(const float2* __restrict__ raw, int64_t count,
                                                                unsigned long long* __restrict__ part) {
  unsigned long long m = 0ull;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (int64_t)gridDim.x * blockDim.x) {
    const float2 v = raw[i];
    const float a = fmaxf(fabsf(v.x), fabsf(v.y));
    const unsigned long long u = (a == a) ? (unsigned long long)__float_as_uint(a) : 0ull;
    m = u > m ? u : m;
  }
  m = block_max_u64(m);
  if (threadIdx.x == 0) part[blockIdx.x] = m;
}

__global__ void __launch_bounds__(kRedThreads) f2_scale_kernel(const unsigned long long* __restrict__ part,
                                                               int nb, float* __restrict__ scale) {
  unsigned long long m = 0ull;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) m = part[i] > m ? part[i] : m;
  m = block_max_u64(m);
  if (threadIdx.x == 0) {
    const float a = __uint_as_float((uint32_t)m);
    int e = 0;
    if (a > 0.0f && a <= 3.0e38f) (void)frexpf(a, &e);  // a = f 2^e, f in [0.5, 1)
    scale[0] = (a > 0.0f && a <= 3.0e38f) ? ldexpf(1.0f, 14 - e) : 1.0f;
  }
}

// Texel (tx, ty, tz) of the 2n-padded half spectrum for tx in [0, n+1],
// ty, tz in [-n, n+1]: indices repeat with period 2n, tx = n + 1 =
// -(n - 1) mod 2n comes from the Hermitian mirror (n - 1, -ty, -tz).
__device__ inline float2 raw_texel(const float2* __restrict__ raw, int n, int tx, int ty, int tz) {
  const int m = 2 * n;
  bool cj = false;
  if (tx > n) {
    tx = m - tx;
    ty = -ty;
    tz = -tz;
    cj = true;
  }
  ty = ty < 0 ? ty + m : (ty >= m ? ty - m : ty);
  tz = tz < 0 ? tz + m : (tz >= m ? tz - m : tz);
  const float2 v = raw[((int64_t)tz * m + ty) * (n + 1) + tx];
  return cj ? make_float2(v.x, -v.y) : v;
}

// Centred reference table F2c: entries (x, y, z) for x in [0, n+1], y and z
// in [-n, n], entry index [(z + n)(2n+1) + (y + n)](n+2) + x, each the four
// texels (x, y, z), (x, y+1, z), (x, y, z+1), (x, y+1, z+1). A folded sample
// with |j| <= n reads its footprint, entries x0 and x0+1, without wrapping.
template <class Tab>
__global__ void f2_centre_kernel(const float2* __restrict__ raw, int n, const float* __restrict__ scale,
                                 typename Tab::T* __restrict__ F2c) {
  const int X = n + 2, Y = 2 * n + 1;
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int yi = blockIdx.y * blockDim.y + threadIdx.y;
  const int zi = blockIdx.z;
  if (x >= X || yi >= Y) return;
  const float s = scale ? scale[0] : 1.0f;
  typename Tab::T* e = F2c + 4 * (((int64_t)zi * Y + yi) * X + x);
  const int y = yi - n, z = zi - n;
  Tab::store(e, raw_texel(raw, n, x, y, z), s);
  Tab::store(e + 1, raw_texel(raw, n, x, y + 1, z), s);
  Tab::store(e + 2, raw_texel(raw, n, x, y, z + 1), s);
  Tab::store(e + 3, raw_texel(raw, n, x, y + 1, z + 1), s);
}

struct Tri {
  float2 f;
  float2 g[3];  // d f / d j (per index unit)
};

// Trilinear sample of the centred table at j (index units of the padded
// grid), split into fetch (the 2x2x2 footprint, one octet) and
// interpolation so a caller can keep several fetches in flight. The
// trilinear interpolant of a Hermitian table is Hermitian, I(-j) = conj I(j),
// so a sample with jx < 0 is folded as a whole. The clamp keeps every read
// inside the table for any j; samples with |j| < n are never clamped.
template <class Tab>
struct Foot {
  typename Tab::Oct t;  // the 2x2x2 footprint
  float fx, fy, fz;
  bool cj;
};

template <class Tab>
__device__ inline Foot<Tab> f2_fetch(const typename Tab::T* __restrict__ F2c, int n, float jx, float jy,
                                     float jz) {
  Foot<Tab> q;
  q.cj = jx < 0.0f;
  if (q.cj) {
    jx = -jx;
    jy = -jy;
    jz = -jz;
  }
  const float fn = (float)n;
  jx = fminf(jx, fn);
  jy = fminf(fmaxf(jy, -fn), fn);
  jz = fminf(fmaxf(jz, -fn), fn);
  const float xf = floorf(jx), yf = floorf(jy), zf = floorf(jz);
  q.fx = jx - xf;
  q.fy = jy - yf;
  q.fz = jz - zf;
  const int X = n + 2, Y = 2 * n + 1;
  q.t = Tab::load(F2c + 4 * ((((int)zf + n) * Y + ((int)yf + n)) * X + (int)xf));
  return q;
}

template <bool GRAD, class Tab>
__device__ inline Tri f2_interp(const Foot<Tab>& q) {
  // t[z][y][x]; octet order x0 (y0z0, y1z0, y0z1, y1z1), x1 likewise
  const float2 t000 = Tab::texel(q.t, 0), t010 = Tab::texel(q.t, 1), t100 = Tab::texel(q.t, 2),
               t110 = Tab::texel(q.t, 3), t001 = Tab::texel(q.t, 4), t011 = Tab::texel(q.t, 5),
               t101 = Tab::texel(q.t, 6), t111 = Tab::texel(q.t, 7);
  const float fx = q.fx, fy = q.fy, fz = q.fz;
  auto lerp2 = [](float2 a, float2 b, float w) {
    return make_float2(a.x + w * (b.x - a.x), a.y + w * (b.y - a.y));
  };
  const float2 c00 = lerp2(t000, t001, fx), c01 = lerp2(t010, t011, fx);
  const float2 c10 = lerp2(t100, t101, fx), c11 = lerp2(t110, t111, fx);
  const float2 c0 = lerp2(c00, c01, fy), c1 = lerp2(c10, c11, fy);
  Tri t;
  t.f = lerp2(c0, c1, fz);
  if (q.cj) t.f.y = -t.f.y;
  if (GRAD) {
    t.g[2] = make_float2(c1.x - c0.x, c1.y - c0.y);
    const float2 dy0 = make_float2(c01.x - c00.x, c01.y - c00.y);
    const float2 dy1 = make_float2(c11.x - c10.x, c11.y - c10.y);
    t.g[1] = lerp2(dy0, dy1, fz);
    const float2 d00 = make_float2(t001.x - t000.x, t001.y - t000.y);
    const float2 d01 = make_float2(t011.x - t010.x, t011.y - t010.y);
    const float2 d10 = make_float2(t101.x - t100.x, t101.y - t100.y);
    const float2 d11 = make_float2(t111.x - t110.x, t111.y - t110.y);
    t.g[0] = lerp2(lerp2(d00, d01, fy), lerp2(d10, d11, fy), fz);
    // d/dj of conj(I(-j)) = -conj(I'(-j))
    if (q.cj)
#pragma unroll
      for (int i = 0; i < 3; ++i) t.g[i].x = -t.g[i].x;
  }
  return t;
}

template <bool GRAD, class Tab>
__device__ inline Tri f2_sample(const typename Tab::T* __restrict__ F2c, int n, float jx, float jy, float jz) {
  return f2_interp<GRAD, Tab>(f2_fetch<Tab>(F2c, n, jx, jy, jz));
}

__device__ inline int signed_freq(int i, int n) { return i <= n / 2 ? i : i - n; }

// ------------------------------------------------------------ fused search

// One frequency of a column: X = Q conj(f) into the z-pruned sums
// acc[d] += X e^{2 pi i kz d / n} (twiddle row e), and the norm term.
template <int D>
__device__ inline void accumulate_z(float2 (&acc)[D], float& dpart, float2 qv, float w, float2 f,
                                    const float2* __restrict__ e) {
  const float xr = qv.x * f.x + qv.y * f.y, xi = qv.y * f.x - qv.x * f.y;
  dpart += (f.x * f.x + f.y * f.y) * w;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    acc[d].x = fmaf(xr, e[d].x, fmaf(-xi, e[d].y, acc[d].x));
    acc[d].y = fmaf(xr, e[d].y, fmaf(xi, e[d].x, acc[d].y));
  }
}

// One pass of the column schedule for every (particle, orientation) unit:
// the z-pruned DFT of the pass's support columns (registers), then the
// y-pruned DFT per plane piece and the Hermitian x expansion (LDS), added to
// the unit's shift-cube map and normalisation sum in global memory. One
// launch per pass keeps all resident units of an XCD on the same pass, so
// their reads of the reference table (nearby orientations of one particle
// on nearby patches) share that XCD's L2; the map round trip per pass is
// 4 D^3 bytes each way.
template <int RR, class Tab>
__global__ void __launch_bounds__(kBatchThreads) __attribute__((amdgpu_waves_per_eu(RR <= 6 ? 4 : 3))) align_pass_kernel(
    const typename Tab::T* __restrict__ F2, int n, const float2* __restrict__ Q, const float* __restrict__ Nw,
    const float* __restrict__ Mall, int n_orient, int unit0, int n_units, int units_per_xcd,
    const int* __restrict__ sched, int g0, int npass, const float2* __restrict__ twg, float* __restrict__ maps,
    double* __restrict__ dsums) {
  constexpr int D = 2 * RR + 1;
  __shared__ float2 tw[kMaxN][D];              // e^{2 pi i k d / n}
  __shared__ float2 T1[kBatchThreads][D];      // z-pruned column results
  __shared__ float Ms[9];
  __shared__ uint8_t kyslot[kBatchThreads];    // FFT-order ky of each column slot (< 128)
  __shared__ double red[kBatchThreads / kWave];

  // Passes g0 .. g0 + npass - 1 of the schedule in one launch, pass-major
  // (the blocks of a pass are dispatched together). Pass g accumulates into
  // map buffer g % npass, so no two blocks of a launch update one map.
  // XCD-aware unit: workgroups are dealt round-robin over the 8 XCDs;
  // consecutive units (the orientations of one particle) land on one XCD.
  // lu indexes this launch's units (maps, dsums), unit all units.
  const int per_pass = 8 * units_per_xcd;
  const int g = g0 + (int)blockIdx.x / per_pass;
  const int slot = (int)blockIdx.x % per_pass;
  const int lu = (slot % 8) * units_per_xcd + slot / 8;
  if (lu >= n_units || g >= sched[0]) return;  // uniform per block
  const int buf = g % npass;
  maps += (int64_t)buf * n_units * (D * D * D);
  dsums += (int64_t)buf * n_units;
  const int unit = unit0 + lu;
  const int t = threadIdx.x;
  const int p = unit / n_orient;
  const int cx = n / 2 + 1;
  const int64_t nh = (int64_t)cx * n * n;
  const float2* Qp = Q + (int64_t)p * nh;
  const float* Np = Nw + (int64_t)p * nh;
  for (int i = t; i < n * D; i += kBatchThreads) tw[i / D][i % D] = twg[i];
  if (t < 9) Ms[t] = Mall[9 * (int64_t)unit + t];

  // Column schedule (support_schedule): per pass at most kBatchThreads
  // support columns (kx, ky, [lo, hi]), [lo, hi] a segment of the column's
  // kz range [-h, h] inside the support sphere; a piece is a run of slots of
  // one kx inside a pass.
  const int ngroups = sched[0], npieces = sched[1];
  const int4* groups = reinterpret_cast<const int4*>(sched + 4);
  const int4* pieces = groups + ngroups;
  const int* cols = reinterpret_cast<const int*>(pieces + npieces);
  const int4 G = groups[g];  // (first column, column count, first piece, piece count)
  const bool active = t < G.y;
  const int packed = active ? cols[G.x + t] : (128 << 8);  // idle slot: k = (0, 0), empty kz range
  const int kx = packed & 255, ky = ((packed >> 8) & 255) - 128;
  const int lo = active ? ((packed >> 16) & 127) - 64 : 1;   // the column's kz segment [lo, hi]
  const int hi = active ? ((packed >> 23) & 127) - 64 : 0;
  const int yy = ky < 0 ? ky + n : ky;
  kyslot[t] = (uint8_t)yy;
  __syncthreads();
  const float m00 = Ms[0], m01 = Ms[1], m02 = Ms[2], m10 = Ms[3], m11 = Ms[4], m12 = Ms[5],
              m20 = Ms[6], m21 = Ms[7], m22 = Ms[8];
  int lw = lo, hw = hi;  // kz bounds uniform over the wavefront (twiddle reads stay broadcast)
  for (int o = kWave / 2; o > 0; o >>= 1) {
    lw = min(lw, __shfl_xor(lw, o, kWave));
    hw = max(hw, __shfl_xor(hw, o, kWave));
  }
  float2 acc[D];
#pragma unroll
  for (int d = 0; d < D; ++d) acc[d] = make_float2(0.0f, 0.0f);
  float dpart = 0.0f;
  const float fkx = (float)kx, fky = (float)ky;
  const float bx = m00 * fkx + m01 * fky, by = m10 * fkx + m11 * fky, bz = m20 * fkx + m21 * fky;
  // two kz per iteration: both samples' loads are issued in one block
  // before either is consumed (accumulation order stays kz, kz + 1). A lane
  // outside its segment also fetches (kz >= lw >= -(n/2 - 1), kz + 1 <=
  // hw + 1 <= n/2: its Q index is valid and the table read is clamped); it
  // is not summed.
  for (int kz = lw; kz <= hw; kz += 2) {
    const int kz1 = kz + 1;
    const bool pa = kz >= lo && kz <= hi, pb = kz1 >= lo && kz1 <= hi;
    if (pa || pb) {
      const int za = kz < 0 ? kz + n : kz, zb = kz1 < 0 ? kz1 + n : kz1;
      const int oa = (za * cx + kx) * n + yy, ob = (zb * cx + kx) * n + yy;
      const float2 qa = Qp[oa], qb = Qp[ob];
      const float wa = Np[oa], wb = Np[ob];
      const float fa = (float)kz, fb = (float)kz1;
      const Foot<Tab> ta = f2_fetch<Tab>(F2, n, bx + m02 * fa, by + m12 * fa, bz + m22 * fa);
      const Foot<Tab> tb = f2_fetch<Tab>(F2, n, bx + m02 * fb, by + m12 * fb, bz + m22 * fb);
      if (pa) accumulate_z<D>(acc, dpart, qa, wa, f2_interp<false, Tab>(ta).f, tw[za]);
      if (pb) accumulate_z<D>(acc, dpart, qb, wb, f2_interp<false, Tab>(tb).f, tw[zb]);
    }
  }
  const double dsum = active ? ((kx == 0 || (n % 2 == 0 && kx == n / 2)) ? 1.0 : 2.0) * (double)dpart : 0.0;
#pragma unroll
  for (int d = 0; d < D; ++d) T1[t][d] = acc[d];
  {
    const double v = wave_sum(dsum);
    if (t % kWave == 0) red[t / kWave] = v;
  }
  __syncthreads();
  // y-pruned sums over each plane piece and the Hermitian x expansion; each
  // thread owns (dy, dz) and its row of D map entries (at most kOut per
  // thread), kept in registers until every thread is done with T1
  constexpr int kOut = (D * D + kBatchThreads - 1) / kBatchThreads;
  float vals[kOut][D];
#pragma unroll
  for (int q = 0; q < kOut; ++q) {
    const int o = t + q * kBatchThreads;
#pragma unroll
    for (int d = 0; d < D; ++d) vals[q][d] = 0.0f;
    if (o >= D * D) continue;
    const int dy = o / D, dz = o % D;
    for (int pc = G.z; pc < G.z + G.w; ++pc) {
      const int4 pz = pieces[pc];  // (kx, first slot, end slot, 0)
      const int kxx = pz.x, s0 = pz.y, s1 = pz.z;
      float sr = 0.0f, si = 0.0f;
      for (int sl = s0; sl < s1; ++sl) {
        const float2 a = T1[sl][dz];
        const float2 e = tw[kyslot[sl]][dy];
        sr = fmaf(a.x, e.x, fmaf(-a.y, e.y, sr));
        si = fmaf(a.x, e.y, fmaf(a.y, e.x, si));
      }
      const float mk = (kxx == 0 || (n % 2 == 0 && kxx == n / 2)) ? 1.0f : 2.0f;
#pragma unroll
      for (int dx = 0; dx < D; ++dx) {
        const float2 e = tw[kxx][dx];
        vals[q][dx] = fmaf(mk * sr, e.x, fmaf(-mk * si, e.y, vals[q][dx]));
      }
    }
  }
  // The unit's D^3 map goes through LDS (T1's storage, D^3 floats <= 256 D
  // float2) so its read-modify-write in global memory is coalesced: lane i
  // of a wave touches float i of a run of 64.
  static_assert(D * D * D <= 2 * kBatchThreads * D, "map staging must fit in T1");
  __syncthreads();
  float* stage = &T1[0][0].x;
#pragma unroll
  for (int q = 0; q < kOut; ++q) {
    const int o = t + q * kBatchThreads;
    if (o >= D * D) continue;
    const int dy = o / D, dz = o % D;
#pragma unroll
    for (int dx = 0; dx < D; ++dx) stage[(dz * D + dy) * D + dx] = vals[q][dx];
  }
  __syncthreads();
  float* mp = maps + (int64_t)lu * (D * D * D);
  for (int i = t; i < D * D * D; i += kBatchThreads) mp[i] = (g < npass ? 0.0f : mp[i]) + stage[i];
  if (t == 0) {
    double v = 0.0;
    for (int w = 0; w < kBatchThreads / kWave; ++w) v += red[w];
    dsums[lu] = (g < npass ? 0.0 : dsums[lu]) + v;
  }
}

// Peak of each unit's shift-cube map inside the sphere |d| <= RR (first
// index on ties, z, y, x order), parabolic sub-pixel refinement per axis,
// normalised score.
template <int RR>
__global__ void __launch_bounds__(kBatchThreads) align_final_kernel(const float* __restrict__ maps,
                                                                   const double* __restrict__ dsums, int nbuf,
                                                                   int n_units, const double* __restrict__ A,
                                                                   int n_orient, int unit0,
                                                                   float* __restrict__ score,
                                                                   float* __restrict__ shift) {
  constexpr int D = 2 * RR + 1;
  __shared__ float map[D * D * D];
  const int lu = blockIdx.x, unit = unit0 + lu, t = threadIdx.x;
  // the pass buffers added in a fixed order
  for (int i = t; i < D * D * D; i += kBatchThreads) {
    float v = maps[(int64_t)lu * (D * D * D) + i];
    for (int b = 1; b < nbuf; ++b) v += maps[((int64_t)b * n_units + lu) * (D * D * D) + i];
    map[i] = v;
  }
  __syncthreads();
  unsigned long long key = 0ull;
  for (int i = t; i < D * D * D; i += kBatchThreads) {
    const int dx = i % D - RR, dy = (i / D) % D - RR, dz = i / (D * D) - RR;
    if (dx * dx + dy * dy + dz * dz > RR * RR) continue;
    const unsigned long long k = argmax_key(map[i], (uint32_t)i);
    key = k > key ? k : key;
  }
  key = block_max_u64(key);  // contains a barrier; valid in thread 0
  if (t == 0) {
    const int best = (int)key_index(key);
    const float bv = map[best];
    const int bx = best % D, by = (best / D) % D, bz = best / (D * D);
    auto para = [&](int i0, int stride, int coord) {
      if (coord <= 0 || coord >= D - 1) return 0.0f;
      const float a = map[i0 - stride], b = map[i0], c = map[i0 + stride];
      const float den = a - 2.0f * b + c;
      if (!(den < 0.0f)) return 0.0f;
      float d = 0.5f * (a - c) / den;
      return fminf(0.5f, fmaxf(-0.5f, d));
    };
    const float sx = (float)(bx - RR) + para(best, 1, bx);
    const float sy = (float)(by - RR) + para(best, D, by);
    const float sz = (float)(bz - RR) + para(best, D * D, bz);
    double ds = dsums[lu];
    for (int b = 1; b < nbuf; ++b) ds += dsums[(int64_t)b * n_units + lu];
    const double norm = sqrt(A[unit / n_orient] * ds);
    score[unit] = norm > 0.0 ? (float)((double)bv / norm) : 0.0f;
    shift[3 * (int64_t)unit + 0] = sx;
    shift[3 * (int64_t)unit + 1] = sy;
    shift[3 * (int64_t)unit + 2] = sz;
  }
}

// e^{2 pi i k d / n} for k in [0, n), d in [-RR, RR] (FP64 sincospi).
__global__ void twiddle_kernel(int n, int RR, float2* __restrict__ tw) {
  const int D = 2 * RR + 1;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * D) return;
  const int k = i / D, d = i % D - RR;
  double s, c;
  sincospi(2.0 * (double)k * (double)d / (double)n, &s, &c);
  tw[i] = make_float2((float)c, (float)s);
}

// ------------------------------------------------------------ adjoint

// Score and gradient at one pose per particle:
//   num(s, R) = sum_k Re[Q(k) conj(f(k)) e^{2 pi i k.s/n}],  f = F(M k), M = 2 R^T
//   D(R) = sum_k |f|^2 N(k),  score = num / sqrt(A D)
// d num / d s_i = -sum_k Im[Q conj(f) e] 2 pi k_i / n
// d num / d a   =  sum_k Re[Q conj(grad f . dM_a k) e]
// d D / d a     =  2 sum_k Re[conj(f) (grad f . dM_a k)] N
// The derivative of the trilinear interpolant is taken inside each cell
// (piecewise constant across cells). The support frequencies come from a
// compact list (index into the Q/N layout, packed signed k); each particle's
// list is split over gridDim.x workgroups that write FP64 partial sums,
// folded by adjoint_fold_kernel into [score, d/dphi, d/dtheta, d/dpsi
// (per radian), d/dsx, d/dsy, d/dsz].
template <class Tab>
__global__ void __launch_bounds__(kBatchThreads) align_eval_grad_kernel(
    const typename Tab::T* __restrict__ F2, int n, const float2* __restrict__ Q, const float* __restrict__ Nw,
    const float* __restrict__ MdM, const float* __restrict__ shifts, const int2* __restrict__ supp,
    int n_supp, double* __restrict__ part) {
  // blockIdx.x = particle, blockIdx.y = chunk of the support list (4 x 4 x
  // 4 frequency cubes: the lanes of a wave sample neighbouring frequencies)
  const int p = blockIdx.x, chunk = blockIdx.y, nchunk = gridDim.y;
  const int cx = n / 2 + 1;
  const int64_t nh = (int64_t)cx * n * n;
  const float2* Qp = Q + (int64_t)p * nh;
  const float* Np = Nw + (int64_t)p * nh;
  __shared__ float Ms[36];
  if (threadIdx.x < 36) Ms[threadIdx.x] = MdM[36 * (int64_t)p + threadIdx.x];
  __syncthreads();
  const float s0 = shifts[3 * p], s1 = shifts[3 * p + 1], s2 = shifts[3 * p + 2];
  const float inv_n2 = 2.0f / (float)n, tp = 6.283185307179586f / (float)n;
  float v[kGradTerms];
#pragma unroll
  for (int i = 0; i < kGradTerms; ++i) v[i] = 0.0f;
  const int per = (n_supp + nchunk - 1) / nchunk;
  const int i0 = chunk * per, i1 = min(n_supp, i0 + per);
  for (int i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
    const int2 e = supp[i];
    const float fk[3] = {(float)((e.y & 255) - 128), (float)(((e.y >> 8) & 255) - 128),
                         (float)(((e.y >> 16) & 255) - 128)};
    float j[3];
#pragma unroll
    for (int r = 0; r < 3; ++r) j[r] = Ms[3 * r] * fk[0] + Ms[3 * r + 1] * fk[1] + Ms[3 * r + 2] * fk[2];
    const Tri tr = f2_sample<true, Tab>(F2, n, j[0], j[1], j[2]);
    const float2 qv = Qp[e.x];
    const float w = Np[e.x];
    const float m = (e.y >> 24) ? 2.0f : 1.0f;   // Hermitian multiplicity of the x plane
    float sn, cs;
    sincospif((fk[0] * s0 + fk[1] * s1 + fk[2] * s2) * inv_n2, &sn, &cs);
    // X = Q conj(f); Y = X e
    const float xr = qv.x * tr.f.x + qv.y * tr.f.y;
    const float xi = qv.y * tr.f.x - qv.x * tr.f.y;
    const float yr = xr * cs - xi * sn, yi = xr * sn + xi * cs;
    v[0] += m * yr;
    v[1] -= m * yi * tp * fk[0];
    v[2] -= m * yi * tp * fk[1];
    v[3] -= m * yi * tp * fk[2];
    v[4] += m * (tr.f.x * tr.f.x + tr.f.y * tr.f.y) * w;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const float* dM = Ms + 9 * (a + 1);
      float dj[3];
#pragma unroll
      for (int r = 0; r < 3; ++r) dj[r] = dM[3 * r] * fk[0] + dM[3 * r + 1] * fk[1] + dM[3 * r + 2] * fk[2];
      const float dfr = tr.g[0].x * dj[0] + tr.g[1].x * dj[1] + tr.g[2].x * dj[2];
      const float dfi = tr.g[0].y * dj[0] + tr.g[1].y * dj[1] + tr.g[2].y * dj[2];
      // Re[Q conj(df) e]
      const float zr = qv.x * dfr + qv.y * dfi;
      const float zi = qv.y * dfr - qv.x * dfi;
      v[5 + a] += m * (zr * cs - zi * sn);
      v[8 + a] += m * 2.0f * (tr.f.x * dfr + tr.f.y * dfi) * w;
    }
  }
  double d[kGradTerms];
#pragma unroll
  for (int i = 0; i < kGradTerms; ++i) d[i] = (double)v[i];
  block_sum<kGradTerms>(d);
  if (threadIdx.x == 0) {
    double* o = part + ((int64_t)p * nchunk + chunk) * kGradTerms;
#pragma unroll
    for (int i = 0; i < kGradTerms; ++i) o[i] = d[i];
  }
}

// The same score and gradient, brick-major. The samples j = M k of every
// particle cover the same ball of the table whatever its rotation, so the
// table is cut into bricks of S^3 positions (the half ball jx >= 0, |j| <=
// 2 r; brick list from brick_host) and one workgroup takes one (brick,
// particle): it enumerates the integer k of the full support ball with M k
// inside its brick (S b <= j < S (b + 1) per axis: the partition of the
// samples) and sums their terms. The full ball with the jx >= 0 half
// weighted 2 (1 on jx = 0) equals the half-grid sum with its Hermitian
// multiplicities: every term is the real part of a Hermitian-symmetric
// product, and k, -k have j, -j. Q and N of k with kx < 0 are conj(Q(-k)),
// N(-k).
// Workgroups are dealt to the XCDs round-robin; XCD x walks bricks x, x + 8,
// x + 16, ..., each for every particle in turn, so the table region of one
// brick (S^3 entries, 1.1 MB complex64 at S = 32) stays in that XCD's L2
// while all particles read it.
// Enumeration: the brick is the rotated cube k = M^T j / 4 (M = 2 R^T); its
// axis-aligned bounding box in k is cut into (kz, kx) rows, each with the ky
// interval of the cube and the ball (widened by 1e-3: the exact membership
// test on the sample's own j decides), the row's phase e^{2 pi i (kx sx +
// kz sz) / n} and an exclusive scan of the row lengths in LDS. Each wave
// takes a contiguous quarter of the items, 64 consecutive items per step
// (consecutive lanes along ky: Q and N read coalesced), each lane finding
// its row by a short forward scan from the step's first row.
// GEN: the derivative matrices are those of right-multiplied rotation
// increments, dM_a = G_a^T M (the refinement's), so dj_a = G_a^T j is a
// cross product with the axis and the three dM are not read.
template <int S>
struct BrickRows {
  static constexpr int kMax = (S * 7 / 8 + 3) * (S * 7 / 8 + 3);  // rows of one pass
};

// WPE: the waves per SIMD the register allocation aims for (4: up to 128
// VGPRs; 5: at most 102, the brick kernel fits 96 without spills); the
// kernel waits on dependent gathers, so more resident waves hide more of it.
template <int S, class Tab, bool GEN, int WPE>
__global__ void __launch_bounds__(kBatchThreads) __attribute__((amdgpu_waves_per_eu(WPE, 8)))
align_eval_brick_kernel(
    const typename Tab::T* __restrict__ F2, int n, const float2* __restrict__ Q, const float* __restrict__ Nw,
    const float* __restrict__ MdM, const float* __restrict__ shifts, const int4* __restrict__ bricks, int nb,
    int n_part, int brick_slots, double r2, double* __restrict__ part) {
  constexpr int kRows = BrickRows<S>::kMax;
  constexpr int kNW = kBatchThreads / kWave;
  __shared__ float Ms[36];
  __shared__ int row_key[kRows];    // kz + 128 | (kx + 128) << 8 | (lo + 128) << 16
  __shared__ int row_off[kRows + 1];
  __shared__ float2 row_ph[kRows];  // e^{2 pi i (kx sx + kz sz) / n}
  __shared__ float2 ph_y[128];      // e^{2 pi i ky sy / n}, ky + 64
  __shared__ int wsum[kNW];
  // XCD-aware (brick, particle): block L goes to XCD L % 8
  const int L = blockIdx.x, xcd = L % 8, q = L / 8;
  const int p = q % n_part, slot = q / n_part;
  const int b = slot * 8 + xcd;
  if (slot >= brick_slots || b >= nb) return;  // uniform per block
  const int t = threadIdx.x, lane = t % kWave, wave = t / kWave;
  if (t < (GEN ? 9 : 36)) Ms[t] = MdM[36 * (int64_t)p + t];
  const float s0 = shifts[3 * p], s1 = shifts[3 * p + 1], s2 = shifts[3 * p + 2];
  const float inv_n2 = 2.0f / (float)n, tp = 6.283185307179586f / (float)n;
  if (t < 128) {
    float sn, cs;
    sincospif((float)(t - 64) * s1 * inv_n2, &sn, &cs);
    ph_y[t] = make_float2(cs, sn);
  }
  __syncthreads();
  const int4 bi = bricks[b];
  const float j0x = (float)(bi.x * S), j0y = (float)(bi.y * S), j0z = (float)(bi.z * S);
  const float j1x = j0x + (float)S, j1y = j0y + (float)S, j1z = j0z + (float)S;
  const float m00 = Ms[0], m01 = Ms[1], m02 = Ms[2], m10 = Ms[3], m11 = Ms[4], m12 = Ms[5], m20 = Ms[6],
              m21 = Ms[7], m22 = Ms[8];
  // bounding box of the rotated brick in k: k_i = sum_a M_ai j_a / 4
  const float hS = 0.5f * (float)S;
  const float jcx = j0x + hS, jcy = j0y + hS, jcz = j0z + hS;
  const float kcx = 0.25f * (m00 * jcx + m10 * jcy + m20 * jcz);
  const float kcy = 0.25f * (m01 * jcx + m11 * jcy + m21 * jcz);
  const float kcz = 0.25f * (m02 * jcx + m12 * jcy + m22 * jcz);
  const float hkx = 0.25f * hS * (fabsf(m00) + fabsf(m10) + fabsf(m20));
  const float hky = 0.25f * hS * (fabsf(m01) + fabsf(m11) + fabsf(m21));
  const float hkz = 0.25f * hS * (fabsf(m02) + fabsf(m12) + fabsf(m22));
  const int ir = (int)sqrt(r2);
  const int ir2 = (int)floor(r2);  // k2 <= r2 for integer k2
  const int kx0 = max((int)ceilf(kcx - hkx - 1e-3f), -ir), kx1 = min((int)floorf(kcx + hkx + 1e-3f), ir);
  const int ky0 = max((int)ceilf(kcy - hky - 1e-3f), -ir), ky1 = min((int)floorf(kcy + hky + 1e-3f), ir);
  const int kz0 = max((int)ceilf(kcz - hkz - 1e-3f), -ir), kz1 = min((int)floorf(kcz + hkz + 1e-3f), ir);
  const int nxr = kx1 - kx0 + 1, nzr = kz1 - kz0 + 1;
  const int nrows = (nxr > 0 && nzr > 0 && ky1 >= ky0) ? nxr * nzr : 0;
  // ky interval of a row from the three slabs a0 <= c + m ky < a0 + S
  const bool fx = fabsf(m01) > 1e-6f, fy = fabsf(m11) > 1e-6f, fz = fabsf(m21) > 1e-6f;
  const float ix = fx ? 1.0f / m01 : 0.0f, iy = fy ? 1.0f / m11 : 0.0f, iz = fz ? 1.0f / m21 : 0.0f;
  const int cx = n / 2 + 1;
  const int64_t nh = (int64_t)cx * n * n;
  const float2* Qp = Q + (int64_t)p * nh;
  const float* Np = Nw + (int64_t)p * nh;
  float v[kGradTerms];
#pragma unroll
  for (int i = 0; i < kGradTerms; ++i) v[i] = 0.0f;
  for (int r0 = 0; r0 < nrows; r0 += kRows) {  // uniform
    const int rn = min(kRows, nrows - r0);
    constexpr int kPer = (kRows + kBatchThreads - 1) / kBatchThreads;
    int len[kPer];
    int local = 0;
#pragma unroll
    for (int c = 0; c < kPer; ++c) {
      const int r = t * kPer + c;
      len[c] = 0;
      if (r >= rn) continue;
      const int rr_ = r0 + r;
      const int kz = kz0 + rr_ / nxr, kx = kx0 + rr_ % nxr;
      const int rem = ir2 - (kx * kx + kz * kz);
      int l = 0, ilo = 0;
      if (rem >= 0) {
        const float sr = sqrtf((float)rem);
        float lo = fmaxf((float)ky0, -sr - 1e-3f), hi = fminf((float)ky1, sr + 1e-3f);
        const float fkx = (float)kx, fkz = (float)kz;
        auto clip = [&](bool f, float inv, float c0, float a0, float a1) {
          if (f) {
            float ya = (a0 - c0) * inv, yb = (a1 - c0) * inv;
            if (ya > yb) { const float tmp = ya; ya = yb; yb = tmp; }
            lo = fmaxf(lo, ya - 1e-3f);
            hi = fminf(hi, yb + 1e-3f);
          } else if (c0 < a0 - 1e-3f || c0 > a1 + 1e-3f) {
            hi = lo - 1.0f;
          }
        };
        clip(fx, ix, m00 * fkx + m02 * fkz, j0x, j1x);
        clip(fy, iy, m10 * fkx + m12 * fkz, j0y, j1y);
        clip(fz, iz, m20 * fkx + m22 * fkz, j0z, j1z);
        ilo = (int)ceilf(lo);
        const int ihi = (int)floorf(hi);
        l = ihi >= ilo ? ihi - ilo + 1 : 0;
        if (l > 0) {
          float sn, cs;
          sincospif((fkx * s0 + fkz * s2) * inv_n2, &sn, &cs);
          row_ph[r] = make_float2(cs, sn);
        }
      }
      row_key[r] = (kz + 128) | ((kx + 128) << 8) | ((ilo + 128) << 16);
      len[c] = l;
      local += l;
    }
    // exclusive scan of the row lengths (rows t*kPer .. t*kPer + kPer - 1 per thread)
    int incl = local;
    for (int o = 1; o < kWave; o <<= 1) {
      const int y = __shfl_up(incl, o, kWave);
      if (lane >= o) incl += y;
    }
    if (lane == kWave - 1) wsum[wave] = incl;
    __syncthreads();
    int base = 0;
    for (int w = 0; w < wave; ++w) base += wsum[w];
    int run = base + incl - local;
#pragma unroll
    for (int c = 0; c < kPer; ++c) {
      const int r = t * kPer + c;
      if (r < rn) row_off[r] = run;
      run += len[c];
    }
    int total = 0;
    for (int w = 0; w < kNW; ++w) total += wsum[w];
    if (t == 0) row_off[rn] = total;
    __syncthreads();
    // each wave a contiguous quarter of the items; its first row by bisection
    const int per = (total + kNW - 1) / kNW;
    const int ibeg = min(total, wave * per), iend = min(total, ibeg + per);
    int rw = 0;
    if (ibeg < iend) {
      int a = 0, bnd = rn;
      while (bnd - a > 1) {
        const int mid = (a + bnd) >> 1;
        if (row_off[mid] <= ibeg) a = mid; else bnd = mid;
      }
      rw = a;
    }
    for (int i0 = ibeg; i0 < iend; i0 += kWave) {  // uniform over the wave
      const int i = i0 + lane;
      int r = rw;
      const int ic = min(i, iend - 1);
      while (row_off[r + 1] <= ic) ++r;
      rw = __shfl(r, kWave - 1, kWave);
      if (i >= iend) continue;
      const int key = row_key[r];
      const int kz = (key & 255) - 128, kx = ((key >> 8) & 255) - 128;
      const int ky = ((key >> 16) & 255) - 128 + (i - row_off[r]);
      const int k2 = kx * kx + ky * ky + kz * kz;
      if (k2 > ir2 || k2 == 0) continue;
      const float fk0 = (float)kx, fk1 = (float)ky, fk2 = (float)kz;
      const float jx = m00 * fk0 + m01 * fk1 + m02 * fk2;
      const float jy = m10 * fk0 + m11 * fk1 + m12 * fk2;
      const float jz = m20 * fk0 + m21 * fk1 + m22 * fk2;
      if (!(jx >= j0x && jx < j1x && jy >= j0y && jy < j1y && jz >= j0z && jz < j1z)) continue;
      const float m = jx > 0.0f ? 2.0f : 1.0f;
      // Q, N at k (the stored half grid holds kx >= 0; conj(Q(-k)) otherwise);
      // |k| <= r < n/2 keeps the index inside the particle's arrays
      const bool neg = kx < 0;
      const int ax = neg ? -kx : kx, ay = neg ? -ky : ky, az = neg ? -kz : kz;
      const int64_t idx = ((int64_t)(az < 0 ? az + n : az) * cx + ax) * n + (ay < 0 ? ay + n : ay);
      float2 qv = Qp[idx];
      if (neg) qv.y = -qv.y;
      const float w = Np[idx];
      const Tri tr = f2_sample<true, Tab>(F2, n, jx, jy, jz);
      const float2 pr = row_ph[r], py = ph_y[ky + 64];
      const float cs = pr.x * py.x - pr.y * py.y, sn = pr.x * py.y + pr.y * py.x;
      const float xr = qv.x * tr.f.x + qv.y * tr.f.y;
      const float xi = qv.y * tr.f.x - qv.x * tr.f.y;
      const float yr = xr * cs - xi * sn, yi = xr * sn + xi * cs;
      v[0] += m * yr;
      const float myi = m * yi * tp;
      v[1] -= myi * fk0;
      v[2] -= myi * fk1;
      v[3] -= myi * fk2;
      v[4] += m * (tr.f.x * tr.f.x + tr.f.y * tr.f.y) * w;
#pragma unroll
      for (int aa = 0; aa < 3; ++aa) {
        float dfr, dfi;
        if (GEN) {
          // dj_x = (0, jz, -jy), dj_y = (-jz, 0, jx), dj_z = (jy, -jx, 0)
          if (aa == 0) {
            dfr = tr.g[1].x * jz - tr.g[2].x * jy;
            dfi = tr.g[1].y * jz - tr.g[2].y * jy;
          } else if (aa == 1) {
            dfr = tr.g[2].x * jx - tr.g[0].x * jz;
            dfi = tr.g[2].y * jx - tr.g[0].y * jz;
          } else {
            dfr = tr.g[0].x * jy - tr.g[1].x * jx;
            dfi = tr.g[0].y * jy - tr.g[1].y * jx;
          }
        } else {
          const float* dM = Ms + 9 * (aa + 1);
          const float dj0 = dM[0] * fk0 + dM[1] * fk1 + dM[2] * fk2;
          const float dj1 = dM[3] * fk0 + dM[4] * fk1 + dM[5] * fk2;
          const float dj2 = dM[6] * fk0 + dM[7] * fk1 + dM[8] * fk2;
          dfr = tr.g[0].x * dj0 + tr.g[1].x * dj1 + tr.g[2].x * dj2;
          dfi = tr.g[0].y * dj0 + tr.g[1].y * dj1 + tr.g[2].y * dj2;
        }
        const float zr = qv.x * dfr + qv.y * dfi;
        const float zi = qv.y * dfr - qv.x * dfi;
        v[5 + aa] += m * (zr * cs - zi * sn);
        v[8 + aa] += m * 2.0f * (tr.f.x * dfr + tr.f.y * dfi) * w;
      }
    }
    __syncthreads();  // rows of the next pass reuse the LDS arrays
  }
  double d[kGradTerms];
#pragma unroll
  for (int i = 0; i < kGradTerms; ++i) d[i] = (double)v[i];
  block_sum<kGradTerms>(d);
  if (t == 0) {
    double* o = part + ((int64_t)p * nb + b) * kGradTerms;
#pragma unroll
    for (int i = 0; i < kGradTerms; ++i) o[i] = d[i];
  }
}

__global__ void __launch_bounds__(kRedThreads) adjoint_fold_kernel(const double* __restrict__ part, int g,
                                                                   const double* __restrict__ A,
                                                                   double* __restrict__ out) {
  const int p = blockIdx.x;
  double v[kGradTerms];
#pragma unroll
  for (int i = 0; i < kGradTerms; ++i) v[i] = 0.0;
  for (int b = threadIdx.x; b < g; b += blockDim.x)
#pragma unroll
    for (int i = 0; i < kGradTerms; ++i) v[i] += part[((int64_t)p * g + b) * kGradTerms + i];
  block_sum<kGradTerms>(v);
  if (threadIdx.x == 0) {
    const double num = v[0], Dn = v[4];
    const double norm = sqrt(A[p] * Dn);
    double* o = out + 7 * (int64_t)p;
    o[0] = num / norm;
    for (int a = 0; a < 3; ++a) o[1 + a] = v[5 + a] / norm - num * v[8 + a] / (2.0 * Dn * norm);
    for (int i = 0; i < 3; ++i) o[4 + i] = v[1 + i] / norm;
  }
}

// Compact support list of the half grid for (n, radius): x = index in the
// [kz][kx][ky] layout of Q/N, y = packed signed (kx, ky, kz) + 128 in bytes
// 0..2 and the Hermitian multiplicity flag (x plane not self-mirrored) in
// byte 3. DC excluded; padded to at least 32 entries with DC entries (Q = N
// = 0 there, so they add nothing).
std::vector<int2> support_host(int n, double radius) {
  const int cx = n / 2 + 1;
  const double r2 = radius * radius;
  std::vector<int2> v;
  for (int zz = 0; zz < n; ++zz)
    for (int kx = 0; kx < cx; ++kx)
      for (int yy = 0; yy < n; ++yy) {
        const int ky = yy <= n / 2 ? yy : yy - n, kz = zz <= n / 2 ? zz : zz - n;
        if ((double)(kx * kx + ky * ky + kz * kz) > r2 || (kx == 0 && ky == 0 && kz == 0)) continue;
        const int herm = (kx == 0 || (n % 2 == 0 && kx == n / 2)) ? 0 : 1;
        int2 e;
        e.x = (zz * cx + kx) * n + yy;
        e.y = (kx + 128) | ((ky + 128) << 8) | ((kz + 128) << 16) | (herm << 24);
        v.push_back(e);
      }
  // 4 x 4 x 4 cubes of (kx, ky, kz), one wavefront each where the sphere
  // fills them, so a wavefront's rotated samples fall in a compact region
  // of the reference table
  auto cube_key = [](const int2& e) {
    const int kx = e.y & 255, ky = (e.y >> 8) & 255, kz = (e.y >> 16) & 255;  // offset by 128
    return ((((kz >> 2) * 64 + (kx >> 2)) * 64 + (ky >> 2)) << 6) | ((kz & 3) << 4) | ((kx & 3) << 2) | (ky & 3);
  };
  std::stable_sort(v.begin(), v.end(), [&](const int2& a, const int2& b) { return cube_key(a) < cube_key(b); });
  while (v.size() < 32) v.push_back(int2{0, 128 | (128 << 8) | (128 << 16)});
  return v;
}

// The support list cached per (n, radius) in the context.
const int2* support_list(emc_ctx* ctx, int n, double radius, int* count) {
  char key[64];
  snprintf(key, sizeof(key), "supp_%d_%.6f", n, radius);
  auto it = ctx->scratch.find(key);
  if (it != ctx->scratch.end() && it->second.ptr) {
    *count = (int)(it->second.bytes / sizeof(int2));  // allocated with the exact size (>= 32 entries)
    return (const int2*)it->second.ptr;
  }
  const std::vector<int2> v = support_host(n, radius);
  *count = (int)v.size();
  int2* d = (int2*)scratch(ctx, key, sizeof(int2) * v.size());
  EMC_HIP(hipMemcpyAsync(d, v.data(), sizeof(int2) * v.size(), hipMemcpyHostToDevice, ctx->stream));
  EMC_HIP(hipStreamSynchronize(ctx->stream));
  return d;
}

template <int RR, class Tab>
void launch_fused(emc_ctx* ctx, const void* F2, int n, const float2* Q, const float* N,
                  const double* A, const float* M, int n_orient, int n_units, const int* sched, int ngroups,
                  float* score, float* shift) {
  constexpr int D = 2 * RR + 1;
  char key[32];
  snprintf(key, sizeof(key), "bat_tw_%d_%d", n, RR);
  auto it = ctx->scratch.find(key);
  const bool fresh = it == ctx->scratch.end() || !it->second.ptr;
  float2* tw = (float2*)scratch(ctx, key, sizeof(float2) * n * D);
  if (fresh) {
    hipLaunchKernelGGL(twiddle_kernel, dim3(ceil_div(n * D, 256)), dim3(256), 0, ctx->stream, n, RR, tw);
    EMC_HIP(hipGetLastError());
  }
  // npass passes per launch ("search_passes"): the last, partly filled
  // round of workgroups of each launch is shared by npass passes; each pass
  // of a launch has its own map buffer
  const int npass = std::max(1, std::min(tune_value(ctx, "search_passes", n, 2), ngroups));
  // units in chunks so the per-unit maps stay bounded (<= 1.3 GB per buffer at r = 8)
  constexpr int kChunkUnits = 65536;
  const int cu = std::min(n_units, kChunkUnits);
  float* maps = (float*)scratch(ctx, "bat_maps", sizeof(float) * D * D * D * (size_t)cu * npass);
  double* dsums = (double*)scratch(ctx, "bat_dsums", sizeof(double) * (size_t)cu * npass);
  for (int u0 = 0; u0 < n_units; u0 += kChunkUnits) {
    const int nu = std::min(kChunkUnits, n_units - u0);
    const int per = ceil_div(nu, 8);
    for (int g = 0; g < ngroups; g += npass) {
      const int np = std::min(npass, ngroups - g);
      KTimer kt(ctx, "align_fused");
      hipLaunchKernelGGL((align_pass_kernel<RR, Tab>), dim3(8 * per * np), dim3(kBatchThreads), 0, ctx->stream,
                         (const typename Tab::T*)F2, n, Q, N, M, n_orient, u0, nu, per, sched, g, npass,
                         (const float2*)tw, maps, dsums);
      EMC_HIP(hipGetLastError());
    }
    KTimer kt(ctx, "align_final");
    hipLaunchKernelGGL(align_final_kernel<RR>, dim3(nu), dim3(kBatchThreads), 0, ctx->stream,
                       (const float*)maps, (const double*)dsums, npass, nu, A, n_orient, u0, score, shift);
    EMC_HIP(hipGetLastError());
  }
}

template <class Tab>
void launch_fused_r(emc_ctx* ctx, int peak_radius, const void* F2, int n, const float2* Q, const float* N,
                    const double* A, const float* M, int n_orient, int n_units, const int* sched, int ngroups,
                    float* score, float* shift) {
  switch (peak_radius) {
    case 4: launch_fused<4, Tab>(ctx, F2, n, Q, N, A, M, n_orient, n_units, sched, ngroups, score, shift); break;
    case 6: launch_fused<6, Tab>(ctx, F2, n, Q, N, A, M, n_orient, n_units, sched, ngroups, score, shift); break;
    default: launch_fused<8, Tab>(ctx, F2, n, Q, N, A, M, n_orient, n_units, sched, ngroups, score, shift); break;
  }
}

// Column schedule of the fused search for (n, radius), cached in the
// context: [ngroups, npieces, 0, 0, groups (int4 x ngroups: first column,
// column count, first piece, piece count), pieces (int4: kx, first slot, end
// slot, 0), columns]. A column is a (kx, ky) with kx^2 + ky^2 <= r^2, packed
// kx | (ky + 128) << 8 | h << 16, h = max |kz| with the frequency inside the
// sphere (the same float test as the kernel). Columns are tiled in (kx, ky)
// patches of "search_tile" x (256 / "search_tile") columns (16 x 16 by
// default), consecutive patches merged while they fit one
// workgroup: the units running together on an XCD (nearby orientations)
// then read a compact region of the reference table per pass.
const int* support_schedule(emc_ctx* ctx, int n, float kmax2, int* ngroups) {
  // patches of tx (kx) x ty (ky) columns, tx * ty = one workgroup's columns;
  // each column's kz range [-h, h] cut into zs consecutive segments, one
  // pass per (patch, segment): a pass then reads 1/zs of the patch's table
  // region (the z-DFT sums and the norm are additive over segments)
  const int tx = tune_value(ctx, "search_tile", n, 16), ty = kBatchThreads / tx;
  const int zs = tune_value(ctx, "search_zsplit", n, 1);
  char key[96];
  snprintf(key, sizeof(key), "sched_%d_%.6f_t%d_z%d", n, (double)kmax2, tx, zs);
  auto it = ctx->scratch.find(key);
  auto mt = ctx->meta.find(key);
  if (it != ctx->scratch.end() && it->second.ptr && mt != ctx->meta.end()) {
    *ngroups = mt->second;
    return (const int*)it->second.ptr;
  }
  // column entry: kx | (ky + 128) << 8 | (lo + 64) << 16 | (hi + 64) << 23,
  // [lo, hi] the column's kz segment (|kz| <= n/2 - 1 <= 63)
  std::map<std::tuple<int, int, int>, std::vector<int>> tiles;  // ordered by (kx tile, ky tile, segment)
  for (int kx = 0; kx <= n / 2; ++kx) {
    if ((float)(kx * kx) > kmax2) break;
    for (int ky = -(n / 2 - 1); ky <= n / 2 - 1; ++ky) {
      if ((float)(kx * kx + ky * ky) > kmax2) continue;
      int h = 0;
      while (h + 1 <= n / 2 - 1 && (float)(kx * kx + ky * ky + (h + 1) * (h + 1)) <= kmax2) ++h;
      const int len = 2 * h + 1;
      for (int sg = 0; sg < zs; ++sg) {
        const int lo = -h + sg * len / zs, hi = -h + (sg + 1) * len / zs - 1;
        if (lo > hi) continue;
        tiles[{kx / tx, (ky + 128) / ty, sg}].push_back(kx | ((ky + 128) << 8) | ((lo + 64) << 16) |
                                                        ((hi + 64) << 23));
      }
    }
  }
  std::vector<int> cols, groups, pieces;
  std::vector<int> cur;
  auto flush = [&]() {
    if (cur.empty()) return;
    std::sort(cur.begin(), cur.end(), [](int a, int b) {
      const int ka = a & 255, kb = b & 255;
      return ka != kb ? ka < kb : ((a >> 8) & 255) < ((b >> 8) & 255);
    });
    const int c0 = (int)cols.size(), p0 = (int)pieces.size() / 4;
    for (size_t i = 0; i < cur.size(); ++i) {
      if (i == 0 || (cur[i] & 255) != (cur[i - 1] & 255)) {
        if (i > 0) pieces[pieces.size() - 2] = (int)i;
        pieces.insert(pieces.end(), {cur[i] & 255, (int)i, (int)cur.size(), 0});
      }
      cols.push_back(cur[i]);
    }
    groups.insert(groups.end(), {c0, (int)cur.size(), p0, (int)pieces.size() / 4 - p0});
    cur.clear();
  };
  for (auto& kv : tiles) {
    if (cur.size() + kv.second.size() > (size_t)kBatchThreads) flush();
    cur.insert(cur.end(), kv.second.begin(), kv.second.end());
  }
  flush();
  for (size_t i = 1; i < groups.size(); i += 4)
    EMC_REQUIRE(groups[i] <= kBatchThreads, "support tile wider than one workgroup");
  const int ng = (int)groups.size() / 4, np = (int)pieces.size() / 4;
  std::vector<int> buf = {ng, np, 0, 0};
  buf.insert(buf.end(), groups.begin(), groups.end());
  buf.insert(buf.end(), pieces.begin(), pieces.end());
  buf.insert(buf.end(), cols.begin(), cols.end());
  int* d = (int*)scratch(ctx, key, sizeof(int) * buf.size());
  EMC_HIP(hipMemcpyAsync(d, buf.data(), sizeof(int) * buf.size(), hipMemcpyHostToDevice, ctx->stream));
  EMC_HIP(hipStreamSynchronize(ctx->stream));
  ctx->meta[key] = ng;
  *ngroups = ng;
  return d;
}

// Workgroups per particle of an adjoint evaluation: every particle's support
// split over enough workgroups to fill the device ("grad_blocks" over the
// whole batch, 4096 = ~16 per CU by default), at least 256 frequencies each.
int eval_grad_blocks(emc_ctx* ctx, int n, int n_part, int ns) {
  const int target = tune_value(ctx, "grad_blocks", n, 4096);
  return std::max(1, std::min((ns + 255) / 256, (target + n_part - 1) / n_part));
}

// Bricks of side S (table index units) covering the half ball jx >= 0,
// |j| <= 2 radius the samples j = M k of the support reach: int4 {bx, by,
// bz, 0} brick indices, z-major.
std::vector<int4> brick_host(int n, double radius, int S) {
  (void)n;
  const double R = 2.0 * radius + 1e-3;
  const int nbr = (int)std::ceil(R / S);
  auto gap = [&](int b) {  // distance from 0 to the interval [b S, b S + S]
    const double lo = (double)b * S, hi = lo + S;
    return lo > 0.0 ? lo : (hi < 0.0 ? -hi : 0.0);
  };
  std::vector<int4> v;
  for (int bz = -nbr; bz < nbr; ++bz)
    for (int by = -nbr; by < nbr; ++by)
      for (int bx = 0; bx < nbr; ++bx) {
        const double dx = gap(bx), dy = gap(by), dz = gap(bz);
        if (dx * dx + dy * dy + dz * dz <= R * R) v.push_back(int4{bx, by, bz, 0});
      }
  return v;
}

// The brick list cached per (radius, S) in the context.
static const int4* brick_list(emc_ctx* ctx, int n, double radius, int S, int* count) {
  char key[64];
  snprintf(key, sizeof(key), "bricks_%d_%.6f_%d", n, radius, S);
  auto it = ctx->scratch.find(key);
  auto mt = ctx->meta.find(key);
  if (it != ctx->scratch.end() && it->second.ptr && mt != ctx->meta.end()) {
    *count = mt->second;
    return (const int4*)it->second.ptr;
  }
  const std::vector<int4> v = brick_host(n, radius, S);
  EMC_REQUIRE(!v.empty(), "empty brick list");
  int4* d = (int4*)scratch(ctx, key, sizeof(int4) * v.size());
  EMC_HIP(hipMemcpyAsync(d, v.data(), sizeof(int4) * v.size(), hipMemcpyHostToDevice, ctx->stream));
  EMC_HIP(hipStreamSynchronize(ctx->stream));
  ctx->meta[key] = (int)v.size();
  *count = (int)v.size();
  return d;
}

// Which adjoint kernel ("grad_kernel": 1 brick-major, 0 support-list order)
// and the brick side ("grad_brick": 16 or 32).
void eval_grad_kind(emc_ctx* ctx, int n, int* kind, int* S, int* waves) {
  *kind = tune_value(ctx, "grad_kernel", n, 1) ? 1 : 0;
  *S = tune_value(ctx, "grad_brick", n, 32) == 16 ? 16 : 32;
  // 5 waves per SIMD only for the 32-brick kernel with the refinement's
  // generator derivatives (the 16-brick one, and general derivative
  // matrices, keep 4)
  *waves = (*S == 32 && tune_value(ctx, "grad_waves", n, 4) >= 5) ? 5 : 4;
}

// Score + gradient launch shared by emc_align_eval_grad and the refinement:
// the support (or brick) list and the partial-sum buffer are set up first
// (they may allocate or copy), the launches alone can then be captured into
// a graph.
EvalSetup eval_grad_setup(emc_ctx* ctx, int n, int n_part, double support_radius) {
  EvalSetup e{};
  eval_grad_kind(ctx, n, &e.kind, &e.S, &e.waves);
  e.r2 = support_radius * support_radius;
  if (e.kind == 1) {
    e.list = brick_list(ctx, n, support_radius, e.S, &e.ns);
    e.g = e.ns;
  } else {
    e.list = support_list(ctx, n, support_radius, &e.ns);
    EMC_REQUIRE(e.ns > 0, "empty support");
    e.g = eval_grad_blocks(ctx, n, n_part, e.ns);
  }
  e.part = (double*)scratch(ctx, "grad_part", sizeof(double) * kGradTerms * (size_t)e.g * n_part);
  return e;
}

template <class Tab, bool GEN>
static void launch_brick(emc_ctx* ctx, const EvalSetup& e, int n, int n_part, const void* d_F2, const void* d_Q,
                         const float* d_N, const float* d_MdM, const float* d_shift) {
  using T = typename Tab::T;
  const int slots = ceil_div(e.ns, 8);
  const dim3 grid((unsigned)(8 * slots * n_part));
  if (e.S == 16)
    hipLaunchKernelGGL((align_eval_brick_kernel<16, Tab, GEN, 4>), grid, dim3(kBatchThreads), 0, ctx->stream,
                       (const T*)d_F2, n, (const float2*)d_Q, d_N, d_MdM, d_shift, (const int4*)e.list, e.ns, n_part,
                       slots, e.r2, e.part);
  else if (GEN && e.waves == 5)  // general dM (GEN false) needs more than 102 VGPRs: it keeps 4 waves
    hipLaunchKernelGGL((align_eval_brick_kernel<32, Tab, GEN, GEN ? 5 : 4>), grid, dim3(kBatchThreads), 0,
                       ctx->stream, (const T*)d_F2, n, (const float2*)d_Q, d_N, d_MdM, d_shift, (const int4*)e.list,
                       e.ns, n_part, slots, e.r2, e.part);
  else
    hipLaunchKernelGGL((align_eval_brick_kernel<32, Tab, GEN, 4>), grid, dim3(kBatchThreads), 0, ctx->stream,
                       (const T*)d_F2, n, (const float2*)d_Q, d_N, d_MdM, d_shift, (const int4*)e.list, e.ns, n_part,
                       slots, e.r2, e.part);
}

template <class Tab>
static void launch_eval_kernel(emc_ctx* ctx, const EvalSetup& e, int n, int n_part, const void* d_F2,
                               const void* d_Q, const float* d_N, const float* d_MdM, const float* d_shift,
                               bool gen) {
  using T = typename Tab::T;
  if (e.kind == 1) {
    if (gen)
      launch_brick<Tab, true>(ctx, e, n, n_part, d_F2, d_Q, d_N, d_MdM, d_shift);
    else
      launch_brick<Tab, false>(ctx, e, n, n_part, d_F2, d_Q, d_N, d_MdM, d_shift);
  } else {
    hipLaunchKernelGGL(align_eval_grad_kernel<Tab>, dim3(n_part, e.g), dim3(kBatchThreads), 0, ctx->stream,
                       (const T*)d_F2, n, (const float2*)d_Q, d_N, d_MdM, d_shift, (const int2*)e.list, e.ns, e.part);
  }
}

void eval_grad_launch(emc_ctx* ctx, const EvalSetup& e, int n, int n_part, int table_format, const void* d_F2,
                      const void* d_Q, const float* d_N, const double* d_A, const float* d_MdM, const float* d_shift,
                      double* d_out, bool timed, bool gen) {
  {
    KTimer kt(timed ? ctx : nullptr, "align_eval_grad");
    if (table_format == EMC_TABLE_C64)
      launch_eval_kernel<TabC64>(ctx, e, n, n_part, d_F2, d_Q, d_N, d_MdM, d_shift, gen);
    else
      launch_eval_kernel<TabF16>(ctx, e, n, n_part, d_F2, d_Q, d_N, d_MdM, d_shift, gen);
  }
  hipLaunchKernelGGL(adjoint_fold_kernel, dim3(n_part), dim3(kRedThreads), 0, ctx->stream, (const double*)e.part,
                     e.g, d_A, d_out);
  EMC_HIP(hipGetLastError());
}

void eval_grad_impl(emc_ctx* ctx, int n, int n_part, int table_format, const void* d_F2, const void* d_Q,
                    const float* d_N, const double* d_A, const float* d_MdM, const float* d_shift,
                    double support_radius, double* d_out) {
  const EvalSetup e = eval_grad_setup(ctx, n, n_part, support_radius);
  eval_grad_launch(ctx, e, n, n_part, table_format, d_F2, d_Q, d_N, d_A, d_MdM, d_shift, d_out, true, false);
}

// ------------------------------------------------------------ refinement
//
// Batched quasi-Newton ascent of the CC over (rotation increment w, shift s)
// per particle, all on the device (FP64, one thread per particle): the
// candidate pose is R exp([w]x), s + ds with (w, ds) = t H g; the step is
// taken when the CC rises by at least 1e-4 t (H g).g (Armijo), then H gets
// the BFGS update of the inverse Hessian of -CC and t doubles (capped at 1),
// else t shrinks by 0.3; a non-ascent direction resets H to its start
// diag(step_angle/|g_w|, step_shift/|g_s|). Derivatives are w.r.t.
// right-multiplied rotations, dR/dw_a = R G_a (no gimbal lock).

struct RefineState {
  double *R, *s, *f, *g, *H, *H0, *t, *Rt, *st, *step, *dir;  // per particle: 9,3,1,6,36,36,1,9,3,6,6
  float *MdM, *sf;                                             // eval inputs: 36 + 3 floats
  double* out;                                                 // eval outputs: 7
  static size_t bytes(int P) { return (size_t)P * (sizeof(double) * (9 + 3 + 1 + 6 + 36 + 36 + 1 + 9 + 3 + 6 + 6 + 7) +
                                                   sizeof(float) * (36 + 3)) + 256; }
  void bind(char* b, int P) {
    double* d = (double*)b;
    R = d; d += 9 * P;
    s = d; d += 3 * P;
    f = d; d += P;
    g = d; d += 6 * P;
    H = d; d += 36 * P;
    H0 = d; d += 36 * P;
    t = d; d += P;
    Rt = d; d += 9 * P;
    st = d; d += 3 * P;
    step = d; d += 6 * P;
    dir = d; d += 6 * P;
    out = d; d += 7 * P;
    float* q = (float*)d;
    MdM = q; q += 36 * P;
    sf = q;
  }
};

// exp([w]x) (Rodrigues), row-major
__device__ inline void rodrigues_d(const double w[3], double E[9]) {
  const double th = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
  double k[3] = {0.0, 0.0, 0.0};
  if (th > 1e-15)
    for (int i = 0; i < 3; ++i) k[i] = w[i] / th;
  const double K[9] = {0.0, -k[2], k[1], k[2], 0.0, -k[0], -k[1], k[0], 0.0};
  const double sn = sin(th), c = 1.0 - cos(th);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double kk = 0.0;
      for (int l = 0; l < 3; ++l) kk += K[3 * i + l] * K[3 * l + j];
      E[3 * i + j] = (i == j ? 1.0 : 0.0) + sn * K[3 * i + j] + c * kk;
    }
}

// Eval inputs of pose (R, s): M = 2 R^T and dM_a = 2 (R G_a)^T, G_a the
// generators of rotations about x, y, z; FP32 as the adjoint kernel reads.
__device__ inline void refine_eval_inputs(const double* R, const double* s, float* MdM, float* sf) {
  // R G_x = [0, R02, -R01; ...], R G_y = [-R02, 0, R00; ...], R G_z = [R01, -R00, 0; ...]
  double D[4][9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) D[0][3 * i + j] = R[3 * i + j];
  for (int i = 0; i < 3; ++i) {
    const double r0 = R[3 * i], r1 = R[3 * i + 1], r2 = R[3 * i + 2];
    D[1][3 * i] = 0.0, D[1][3 * i + 1] = r2, D[1][3 * i + 2] = -r1;
    D[2][3 * i] = -r2, D[2][3 * i + 1] = 0.0, D[2][3 * i + 2] = r0;
    D[3][3 * i] = r1, D[3][3 * i + 1] = -r0, D[3][3 * i + 2] = 0.0;
  }
  for (int m = 0; m < 4; ++m)
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) MdM[9 * m + 3 * i + j] = (float)(2.0 * D[m][3 * j + i]);
  for (int i = 0; i < 3; ++i) sf[i] = (float)s[i];
}

// Next candidate: dir = H g (H reset to H0 if that is not an ascent
// direction), step = t dir, Rt = R exp([step_w]x), st = s + step_s.
__device__ inline void refine_candidate(RefineState& S, int p) {
  double* H = S.H + 36 * p;
  const double* g = S.g + 6 * p;
  double d[6], dg = 0.0;
  for (int i = 0; i < 6; ++i) {
    d[i] = 0.0;
    for (int j = 0; j < 6; ++j) d[i] += H[6 * i + j] * g[j];
    dg += d[i] * g[i];
  }
  if (dg <= 0.0) {
    const double* H0 = S.H0 + 36 * p;
    for (int i = 0; i < 36; ++i) H[i] = H0[i];
    for (int i = 0; i < 6; ++i) {
      d[i] = 0.0;
      for (int j = 0; j < 6; ++j) d[i] += H0[6 * i + j] * g[j];
    }
  }
  const double t = S.t[p];
  double w[3], E[9];
  for (int i = 0; i < 6; ++i) {
    S.dir[6 * p + i] = d[i];
    S.step[6 * p + i] = t * d[i];
  }
  for (int i = 0; i < 3; ++i) w[i] = t * d[i];
  rodrigues_d(w, E);
  const double* R = S.R + 9 * p;
  double* Rt = S.Rt + 9 * p;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      Rt[3 * i + j] = R[3 * i] * E[j] + R[3 * i + 1] * E[3 + j] + R[3 * i + 2] * E[6 + j];
  for (int i = 0; i < 3; ++i) S.st[3 * p + i] = S.s[3 * p + i] + t * d[3 + i];
  refine_eval_inputs(Rt, S.st + 3 * p, S.MdM + 36 * p, S.sf + 3 * p);
}

__global__ void refine_begin_kernel(RefineState S, const double* __restrict__ R0, const double* __restrict__ s0,
                                    int P) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  for (int i = 0; i < 9; ++i) S.R[9 * p + i] = R0[9 * p + i];
  for (int i = 0; i < 3; ++i) S.s[3 * p + i] = s0[3 * p + i];
  refine_eval_inputs(S.R + 9 * p, S.s + 3 * p, S.MdM + 36 * p, S.sf + 3 * p);
}

__global__ void refine_first_kernel(RefineState S, int P, double step_angle, double step_shift) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  const double* o = S.out + 7 * p;
  S.f[p] = o[0];
  for (int i = 0; i < 6; ++i) S.g[6 * p + i] = o[1 + i];
  const double* g = S.g + 6 * p;
  const double gw = sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]) + 1e-12;
  const double gs = sqrt(g[3] * g[3] + g[4] * g[4] + g[5] * g[5]) + 1e-12;
  for (int i = 0; i < 36; ++i) S.H[36 * p + i] = 0.0;
  for (int i = 0; i < 6; ++i) S.H[36 * p + 7 * i] = i < 3 ? step_angle / gw : step_shift / gs;
  for (int i = 0; i < 36; ++i) S.H0[36 * p + i] = S.H[36 * p + i];
  S.t[p] = 1.0;
  refine_candidate(S, p);
}

__global__ void refine_step_kernel(RefineState S, int P) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  const double* o = S.out + 7 * p;
  const double ft = o[0];
  const double* gt = o + 1;
  double* g = S.g + 6 * p;
  const double* step = S.step + 6 * p;
  const double* dir = S.dir + 6 * p;
  const double t = S.t[p];
  double dg = 0.0;
  for (int i = 0; i < 6; ++i) dg += dir[i] * g[i];
  const bool ok = ft > S.f[p] + 1e-4 * t * dg;
  if (ok) {
    // BFGS update of the inverse Hessian of -CC: y = (-gt) - (-g)
    double y[6], sy = 0.0;
    for (int i = 0; i < 6; ++i) {
      y[i] = g[i] - gt[i];
      sy += step[i] * y[i];
    }
    if (sy > 1e-18) {
      // H' = V H V^T + rho s s^T with V = I - rho s y^T; for symmetric H
      // this is H - rho (s (Hy)^T + (Hy) s^T) + (rho^2 y^T H y + rho) s s^T
      const double rho = 1.0 / sy;
      double* H = S.H + 36 * p;
      double hy[6], yhy = 0.0;
#pragma unroll
      for (int i = 0; i < 6; ++i) {
        double a = 0.0;
#pragma unroll
        for (int j = 0; j < 6; ++j) a += H[6 * i + j] * y[j];
        hy[i] = a;
        yhy += y[i] * a;
      }
      const double c = rho * rho * yhy + rho;
#pragma unroll
      for (int i = 0; i < 6; ++i)
#pragma unroll
        for (int j = 0; j < 6; ++j)
          H[6 * i + j] = H[6 * i + j] - rho * (step[i] * hy[j] + hy[i] * step[j]) + c * step[i] * step[j];
    }
    for (int i = 0; i < 9; ++i) S.R[9 * p + i] = S.Rt[9 * p + i];
    for (int i = 0; i < 3; ++i) S.s[3 * p + i] = S.st[3 * p + i];
    S.f[p] = ft;
    for (int i = 0; i < 6; ++i) g[i] = gt[i];
    S.t[p] = fmin(1.0, 2.0 * t);
  } else {
    S.t[p] = 0.3 * t;
  }
  refine_candidate(S, p);
}

__global__ void refine_end_kernel(RefineState S, int P, double* __restrict__ R, double* __restrict__ s,
                                  double* __restrict__ cc) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  for (int i = 0; i < 9; ++i) R[9 * p + i] = S.R[9 * p + i];
  for (int i = 0; i < 3; ++i) s[3 * p + i] = S.s[3 * p + i];
  cc[p] = S.f[p];
}

}  // namespace emc

using namespace emc;

extern "C" {

emc_status emc_batch_reference_spectrum(emc_ctx* ctx, int n, const float* d_ref, int table_format,
                                        void* d_F2) {
  return guard([&] {
    use_device(ctx);
    EMC_REQUIRE(n >= 8 && n <= kMaxN, "box size must be 8..%d", kMaxN);
    EMC_REQUIRE(d_ref && d_F2, "null pointer");
    EMC_REQUIRE(table_format == EMC_TABLE_C64 || table_format == EMC_TABLE_F16, "unknown table format %d",
                table_format);
    const int m = 2 * n;
    float* pad = (float*)scratch(ctx, "bat_pad", sizeof(float) * (size_t)m * m * m);
    float2* raw = (float2*)scratch(ctx, "bat_F2raw", sizeof(float2) * (size_t)(n + 1) * m * m);
    hipLaunchKernelGGL(pad_origin_kernel, dim3(ceil_div(m, 64), ceil_div(m, 4), m), dim3(64, 4, 1), 0,
                       ctx->stream, d_ref, pad, n);
    EMC_HIP(hipGetLastError());
    fft_r2c(cached_plan(ctx, m, m, m, 1), pad, raw);
    const dim3 grid(ceil_div(n + 2, 64), ceil_div(m + 1, 4), m + 1), block(64, 4, 1);
    if (table_format == EMC_TABLE_C64) {
      hipLaunchKernelGGL(f2_centre_kernel<TabC64>, grid, block, 0, ctx->stream, (const float2*)raw, n,
                         (const float*)nullptr, (float2*)d_F2);
    } else {
      const int nb = 256;
      auto* part = (unsigned long long*)scratch(ctx, "bat_f2max", sizeof(unsigned long long) * nb + 256);
      float* sc = (float*)(part + nb);
      hipLaunchKernelGGL(f2_maxabs_kernel, dim3(nb), dim3(kRedThreads), 0, ctx->stream, (const float2*)raw,
                         (int64_t)(n + 1) * m * m, part);
      hipLaunchKernelGGL(f2_scale_kernel, dim3(1), dim3(kRedThreads), 0, ctx->stream,
                         (const unsigned long long*)part, nb, sc);
      hipLaunchKernelGGL(f2_centre_kernel<TabF16>, grid, block, 0, ctx->stream, (const float2*)raw, n,
                         (const float*)sc, (half2_t*)d_F2);
    }
    EMC_HIP(hipGetLastError());
  });
}

emc_status emc_batch_prepare_particles(emc_ctx* ctx, int n, int n_part, const float* const* h_particles,
                                       const float* d_bandpass, const float* const* h_wedges, void* d_Q,
                                       float* d_N, double* d_A) {
  return emc_batch_prepare_particles_masked(ctx, n, n_part, h_particles, nullptr, d_bandpass, h_wedges, d_Q, d_N,
                                            d_A);
}

emc_status emc_batch_prepare_particles_masked(emc_ctx* ctx, int n, int n_part, const float* const* h_particles,
                                              const float* d_mask, const float* d_bandpass,
                                              const float* const* h_wedges, void* d_Q, float* d_N, double* d_A) {
  return guard([&] {
    use_device(ctx);
    EMC_REQUIRE(n >= 8 && n <= kMaxN, "box size must be 8..%d", kMaxN);
    EMC_REQUIRE(n_part >= 0, "n_part must be >= 0");
    if (n_part == 0) return;
    EMC_REQUIRE(h_particles && d_bandpass && d_Q && d_N && d_A, "null pointer");
    for (int i = 0; i < n_part; ++i) EMC_REQUIRE(h_particles[i], "null particle pointer %d", i);
    const int64_t nv = (int64_t)n * n * n, nh = (int64_t)(n / 2 + 1) * n * n;
    const bool even = n % 2 == 0;
    const int chunk = prep_chunk(n);
    float* sh = (float*)scratch(ctx, "bat_shift", sizeof(float) * nv * chunk);
    float2* P = (float2*)scratch(ctx, "bat_P", sizeof(float2) * nh * chunk);
    const int nb = 256;
    double* part = (double*)scratch(ctx, "bat_part", sizeof(double) * nb * chunk);
    // chunks of `chunk`, then halves down to 1 particle: a handful of batched
    // FFT plans per size
    for (int i0 = 0; i0 < n_part;) {
      int B = chunk;
      while (B > n_part - i0) B /= 2;
      PrepPtrs pp{};
      for (int b = 0; b < B; ++b) {
        pp.part[b] = h_particles[i0 + b];
        pp.wedge[b] = h_wedges ? h_wedges[i0 + b] : nullptr;
      }
      // even n: the shift goes into the spectrum's sign; unmasked particles
      // already stacked at a stride of one volume go to the R2C as they are
      bool direct = even && !d_mask;
      for (int b = 1; b < B && direct; ++b) direct = pp.part[b] == pp.part[0] + b * nv;
      const float* src = sh;
      if (direct) {
        src = pp.part[0];
      } else {
        hipLaunchKernelGGL(ifftshift_many_kernel, dim3(ceil_div(n, 64), ceil_div(n, 4), n * B), dim3(64, 4, 1),
                           0, ctx->stream, pp, d_mask, sh, n, even ? 0 : 1);
        EMC_HIP(hipGetLastError());
      }
      fft_r2c(cached_plan(ctx, n, n, n, B), src, P);
      {
        KTimer kt(ctx, "prep_particle");
        hipLaunchKernelGGL(prep_particle_kernel, dim3(nb, B), dim3(kRedThreads), 0, ctx->stream,
                           (const float2*)P, d_bandpass, pp, n, even ? 1 : 0, (float2*)d_Q + i0 * nh, d_N + i0 * nh,
                           part);
      }
      hipLaunchKernelGGL(fold_partials_kernel<1>, dim3(B), dim3(kRedThreads), 0, ctx->stream,
                         (const double*)part, nb, d_A + i0);
      EMC_HIP(hipGetLastError());
      i0 += B;
    }
  });
}

emc_status emc_batch_prepare_particle(emc_ctx* ctx, int n, const float* d_particle,
                                      const float* d_bandpass, const float* d_wedge, void* d_Q,
                                      float* d_N, double* d_A) {
  return emc_batch_prepare_particles(ctx, n, 1, &d_particle, d_bandpass, &d_wedge, d_Q, d_N, d_A);
}

emc_status emc_align_batched(emc_ctx* ctx, int n, int n_part, int table_format, const void* d_F2, const void* d_Q,
                             const float* d_N, const double* d_A, const double* d_start_cm,
                             const double* angles_pts, int n_orient, int peak_radius,
                             double support_radius, float* d_score, float* d_shift) {
  return guard([&] {
    use_device(ctx);
    EMC_REQUIRE(n >= 8 && n <= kMaxN, "box size must be 8..%d", kMaxN);
    EMC_REQUIRE(n_part > 0 && n_orient > 0, "empty batch");
    EMC_REQUIRE(d_F2 && d_Q && d_N && d_A && d_start_cm && angles_pts && d_score && d_shift,
                "null pointer");
    EMC_REQUIRE(support_radius > 0 && support_radius < 0.5 * n, "support radius must be in (0, n/2)");
    EMC_REQUIRE(peak_radius == 4 || peak_radius == 6 || peak_radius == 8, "peak radius must be 4, 6 or 8");
    EMC_REQUIRE(table_format == EMC_TABLE_C64 || table_format == EMC_TABLE_F16, "unknown table format %d",
                table_format);
    const int64_t units64 = (int64_t)n_part * n_orient;
    EMC_REQUIRE(units64 < ((int64_t)1 << 31) / 8, "too many units in one call");
    const int units = (int)units64;
    double* d_pts = (double*)scratch(ctx, "bat_pts", sizeof(double) * 3 * n_orient);
    float* d_M = (float*)scratch(ctx, "bat_M", sizeof(float) * 9 * (size_t)units);
    upload_async(ctx, d_pts, angles_pts, sizeof(double) * 3 * n_orient);
    hipLaunchKernelGGL(compose_kernel, dim3(ceil_div(units, 256)), dim3(256), 0, ctx->stream, d_start_cm,
                       (const double*)d_pts, n_part, n_orient, d_M);
    EMC_HIP(hipGetLastError());
    const float kmax2 = (float)(support_radius * support_radius);
    const float2* Q = (const float2*)d_Q;
    int ngroups = 0;
    const int* sched = support_schedule(ctx, n, kmax2, &ngroups);
    if (table_format == EMC_TABLE_C64)
      launch_fused_r<TabC64>(ctx, peak_radius, d_F2, n, Q, d_N, d_A, d_M, n_orient, units, sched, ngroups, d_score,
                             d_shift);
    else
      launch_fused_r<TabF16>(ctx, peak_radius, d_F2, n, Q, d_N, d_A, d_M, n_orient, units, sched, ngroups, d_score,
                             d_shift);
  });
}

emc_status emc_align_eval_grad(emc_ctx* ctx, int n, int n_part, int table_format, const void* d_F2,
                               const void* d_Q,
                               const float* d_N, const double* d_A, const float* d_MdM,
                               const float* d_shift, double support_radius, double* d_out) {
  return guard([&] {
    use_device(ctx);
    EMC_REQUIRE(n >= 8 && n <= kMaxN, "box size must be 8..%d", kMaxN);
    EMC_REQUIRE(n_part > 0, "empty batch");
    EMC_REQUIRE(d_F2 && d_Q && d_N && d_A && d_MdM && d_shift && d_out, "null pointer");
    EMC_REQUIRE(table_format == EMC_TABLE_C64 || table_format == EMC_TABLE_F16, "unknown table format %d",
                table_format);
    EMC_REQUIRE(support_radius > 0 && support_radius < 0.5 * n, "support radius must be in (0, n/2)");
    eval_grad_impl(ctx, n, n_part, table_format, d_F2, d_Q, d_N, d_A, d_MdM, d_shift, support_radius, d_out);
  });
}

emc_status emc_align_pick(emc_ctx* ctx, int n_part, int n_orient, const float* d_score, const float* d_shift,
                          const double* d_start_cm, const double* d_angles_pts, double* d_R0, double* d_s0,
                          int32_t* d_best, float* d_best_score) {
  return guard([&] {
    use_device(ctx);
    EMC_REQUIRE(n_part > 0 && n_orient > 0, "empty batch");
    EMC_REQUIRE(d_score && d_shift && d_start_cm && d_angles_pts && d_R0 && d_s0, "null pointer");
    hipLaunchKernelGGL(pick_kernel, dim3(n_part), dim3(kRedThreads), 0, ctx->stream, d_score, d_shift, d_start_cm,
                       d_angles_pts, n_orient, d_R0, d_s0, (int*)d_best, d_best_score);
    EMC_HIP(hipGetLastError());
  });
}

emc_status emc_align_refine(emc_ctx* ctx, int n, int n_part, int table_format, const void* d_F2, const void* d_Q,
                            const float* d_N, const double* d_A, double* d_R, double* d_s, int iters,
                            double step_angle, double step_shift, double support_radius, double* d_cc) {
  return guard([&] {
    use_device(ctx);
    EMC_REQUIRE(n >= 8 && n <= kMaxN, "box size must be 8..%d", kMaxN);
    EMC_REQUIRE(n_part > 0, "empty batch");
    EMC_REQUIRE(iters >= 0, "iters must be >= 0");
    EMC_REQUIRE(step_angle > 0 && step_shift > 0, "initial steps must be positive");
    EMC_REQUIRE(d_F2 && d_Q && d_N && d_A && d_R && d_s && d_cc, "null pointer");
    EMC_REQUIRE(table_format == EMC_TABLE_C64 || table_format == EMC_TABLE_F16, "unknown table format %d",
                table_format);
    EMC_REQUIRE(support_radius > 0 && support_radius < 0.5 * n, "support radius must be in (0, n/2)");
    RefineState st;
    char* base = (char*)scratch(ctx, "refine_state", RefineState::bytes(n_part));
    st.bind(base, n_part);
    const EvalSetup e = eval_grad_setup(ctx, n, n_part, support_radius);
    const dim3 grid(ceil_div(n_part, 64)), block(64);
    // kernels only (no allocation, copy, memset or host wait): capturable
    auto body = [&](bool timed) {
      hipLaunchKernelGGL(refine_begin_kernel, grid, block, 0, ctx->stream, st, (const double*)d_R,
                         (const double*)d_s, n_part);
      eval_grad_launch(ctx, e, n, n_part, table_format, d_F2, d_Q, d_N, d_A, st.MdM, st.sf, st.out, timed, true);
      hipLaunchKernelGGL(refine_first_kernel, grid, block, 0, ctx->stream, st, n_part, step_angle, step_shift);
      for (int it = 0; it < iters; ++it) {
        eval_grad_launch(ctx, e, n, n_part, table_format, d_F2, d_Q, d_N, d_A, st.MdM, st.sf, st.out, timed, true);
        hipLaunchKernelGGL(refine_step_kernel, grid, block, 0, ctx->stream, st, n_part);
      }
      hipLaunchKernelGGL(refine_end_kernel, grid, block, 0, ctx->stream, st, n_part, d_R, d_s, d_cc);
      EMC_HIP(hipGetLastError());
    };
    // Replay the whole iteration chain as one HIP graph, captured again
    // whenever an argument it bakes in changes; the "refine_graph"
    // launch-shape value 0 launches eagerly instead. Per-kernel timing
    // (host-side event bookkeeping) brackets the evaluations of eager
    // launches only.
    if (tune_value(ctx, "refine_graph", n, 1) == 0) {
      body(true);
      return;
    }
    const int ki[9] = {n, n_part, table_format, iters, e.ns, e.g, e.kind, e.S, e.waves};
    const double kd[3] = {step_angle, step_shift, support_radius};
    const void* kp[12] = {d_F2, d_Q, d_N, d_A, d_R, d_s, d_cc, base, e.list, e.part, nullptr, nullptr};
    std::vector<unsigned char> kb(sizeof(ki) + sizeof(kd) + sizeof(kp));
    std::memcpy(kb.data(), ki, sizeof(ki));
    std::memcpy(kb.data() + sizeof(ki), kd, sizeof(kd));
    std::memcpy(kb.data() + sizeof(ki) + sizeof(kd), kp, sizeof(kp));
    // one graph per support list (a coarse-to-fine refinement alternates two)
    auto& gr = ctx->graphs["refine:" + std::to_string(e.kind) + ":" + std::to_string(e.ns) + ":" +
                           std::to_string(support_radius)];
    if (!gr.exec || gr.key != kb) {
      hipGraph_t graph = nullptr;
      EMC_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
      try {
        body(false);
      } catch (...) {
        (void)hipStreamEndCapture(ctx->stream, &graph);
        if (graph) (void)hipGraphDestroy(graph);
        throw;
      }
      EMC_HIP(hipStreamEndCapture(ctx->stream, &graph));
      bool updated = false;
      if (gr.exec) {
        hipGraphNode_t err_node = nullptr;
        hipGraphExecUpdateResult res;
        updated = hipGraphExecUpdate(gr.exec, graph, &err_node, &res) == hipSuccess &&
                  res == hipGraphExecUpdateSuccess;
        if (!updated) {
          (void)hipGetLastError();
          (void)hipGraphExecDestroy(gr.exec);
          gr.exec = nullptr;
        }
      }
      if (!updated) {
        const hipError_t ie = hipGraphInstantiate(&gr.exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ie != hipSuccess) {
          gr.exec = nullptr;
          gr.key.clear();
          EMC_HIP(ie);
        }
      } else {
        (void)hipGraphDestroy(graph);
      }
      gr.key = kb;
    }
    EMC_HIP(hipGraphLaunch(gr.exec, ctx->stream));
  });
}

}  // extern "C"
