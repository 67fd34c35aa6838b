// MRC files and batched subtomogram extraction.
//
// Host side: the 1024-byte header of @MRCImage/private/readHeader.m:43-186 and
// private/writeHeader.m, the (sub)volume read of @MRCImage/getVolume.m:40-298,
// SAVE_IMG.m's header defaults and statistics. Device side: the tomogram is
// uploaded once in its stored type (int8 / int16 / float32 / uint16) through
// two pinned staging buffers, byte order fixed on the device, and every
// particle's window is cut from it in one launch: BH_isWindowValid
// (logicals/BH_isWindowValid.m:42-65) on the host per particle, then the crop
// + 'singleTaper' pad of BH_padZeros3d (masking/BH_padZeros3d.m:189-229) per
// voxel ('single' without the taper, as the averaging uses). The cut is
// HBM-bound: mode bytes read + 4 B written per window voxel.
#include <cerrno>
#include <cmath>
#include <cstring>
#include <memory>

#include "emc_internal.hpp"

namespace emc {
namespace {

constexpr int kHeaderBytes = 1024;
constexpr int32_t kImodStamp = 1146047817;  // "IMOD" little endian (private/defaultIMODStamp.m)

struct File {
  FILE* f = nullptr;
  explicit File(const char* path, const char* mode) : f(std::fopen(path, mode)) {
    if (!f) fail(EMC_ERR_IO, "cannot open %s: %s", path, std::strerror(errno));
  }
  ~File() {
    if (f) std::fclose(f);
  }
  void seek(int64_t off, const char* path) {
    if (fseeko(f, (off_t)off, SEEK_SET) != 0) fail(EMC_ERR_IO, "seek failed in %s", path);
  }
  void read(void* dst, size_t bytes, const char* path) {
    if (bytes && std::fread(dst, 1, bytes, f) != bytes)
      fail(EMC_ERR_IO, "short read in %s (expected %zu bytes)", path, bytes);
  }
  void write(const void* src, size_t bytes, const char* path) {
    if (bytes && std::fwrite(src, 1, bytes, f) != bytes) fail(EMC_ERR_IO, "short write to %s", path);
  }
};

template <class T>
T get(const unsigned char* b, int off, bool swap) {
  unsigned char t[sizeof(T)];
  std::memcpy(t, b + off, sizeof(T));
  if (swap)
    for (size_t i = 0; i < sizeof(T) / 2; ++i) std::swap(t[i], t[sizeof(T) - 1 - i]);
  T v;
  std::memcpy(&v, t, sizeof(T));
  return v;
}

template <class T>
void put(unsigned char* b, int off, T v) {
  std::memcpy(b + off, &v, sizeof(T));
}

int mode_bytes(int mode) {
  switch (mode) {
    case 0: return 1;
    case 1: return 2;
    case 2: return 4;
    case 6: return 2;
    default: fail(EMC_ERR_INVALID, "unsupported MRC mode %d (0, 1, 2 and 6 are read)", mode);
  }
}

// Field offsets follow the read order of readHeader.m:43-172.
void parse_header(const unsigned char* b, bool sw, emc_mrc_header* h) {
  std::memset(h, 0, sizeof(*h));
  h->nx = get<int32_t>(b, 0, sw);
  h->ny = get<int32_t>(b, 4, sw);
  h->nz = get<int32_t>(b, 8, sw);
  h->mode = get<int32_t>(b, 12, sw);
  h->nxstart = get<int32_t>(b, 16, sw);
  h->nystart = get<int32_t>(b, 20, sw);
  h->nzstart = get<int32_t>(b, 24, sw);
  h->mx = get<int32_t>(b, 28, sw);
  h->my = get<int32_t>(b, 32, sw);
  h->mz = get<int32_t>(b, 36, sw);
  for (int k = 0; k < 3; ++k) {
    h->cell[k] = get<float>(b, 40 + 4 * k, sw);
    h->cell_angle[k] = get<float>(b, 52 + 4 * k, sw);
  }
  h->mapc = get<int32_t>(b, 64, sw);
  h->mapr = get<int32_t>(b, 68, sw);
  h->maps = get<int32_t>(b, 72, sw);
  h->dmin = get<float>(b, 76, sw);
  h->dmax = get<float>(b, 80, sw);
  h->dmean = get<float>(b, 84, sw);
  h->space_group = get<int16_t>(b, 88, sw);
  h->n_symmetry_bytes = get<int16_t>(b, 90, sw);
  h->n_bytes_extended = get<int32_t>(b, 92, sw);
  h->creator_id = get<int16_t>(b, 96, sw);
  h->n_bytes_per_section = get<int16_t>(b, 128, sw);
  h->serialem_type = get<int16_t>(b, 130, sw);
  h->imod_stamp = get<int32_t>(b, 152, sw);
  h->imod_flags = get<int32_t>(b, 156, sw);
  h->idtype = get<int16_t>(b, 160, sw);
  h->lens = get<int16_t>(b, 162, sw);
  h->nd1 = get<int16_t>(b, 164, sw);
  h->nd2 = get<int16_t>(b, 166, sw);
  h->vd1 = get<int16_t>(b, 168, sw);
  h->vd2 = get<int16_t>(b, 170, sw);
  for (int k = 0; k < 6; ++k) h->tilt_angles[k] = get<float>(b, 172 + 4 * k, sw);
  for (int k = 0; k < 3; ++k) h->origin[k] = get<float>(b, 196 + 4 * k, sw);
  std::memcpy(h->map, b + 208, 4);
  std::memcpy(h->machine_stamp, b + 212, 4);
  h->rms = get<float>(b, 216, sw);
  h->n_labels = get<int32_t>(b, 220, sw);
  const int nl = h->n_labels < 0 ? 0 : (h->n_labels > 10 ? 10 : h->n_labels);
  for (int l = 0; l < nl; ++l) std::memcpy(h->labels[l], b + 224 + 80 * l, 80);
  h->swapped = sw ? 1 : 0;
  h->data_offset = kHeaderBytes + (int64_t)(h->n_bytes_extended > 0 ? h->n_bytes_extended : 0);
}

void read_header_file(File& f, const char* path, emc_mrc_header* h) {
  unsigned char b[kHeaderBytes];
  f.seek(0, path);
  f.read(b, kHeaderBytes, path);
  // readHeader.m:54-78: the other byte order when a size reads negative or
  // all three exceed 65535
  const int32_t nx = get<int32_t>(b, 0, false), ny = get<int32_t>(b, 4, false), nz = get<int32_t>(b, 8, false);
  const bool sw = nx < 0 || ny < 0 || nz < 0 || (nx > 65535 && ny > 65535 && nz > 65535);
  parse_header(b, sw, h);
  if (h->nx <= 0 || h->ny <= 0 || h->nz <= 0) fail(EMC_ERR_IO, "%s: bad MRC sizes", path);
}

template <int MODE>
struct ModeT;
template <>
struct ModeT<0> {
  using T = int8_t;
};
template <>
struct ModeT<1> {
  using T = int16_t;
};
template <>
struct ModeT<2> {
  using T = float;
};
template <>
struct ModeT<6> {
  using T = uint16_t;
};

template <class T>
void convert_host(const unsigned char* src, int64_t n, bool sw, float* dst) {
  for (int64_t i = 0; i < n; ++i) dst[i] = (float)get<T>(src, (int)(i * sizeof(T)), sw);
}

void convert_host_mode(int mode, const unsigned char* src, int64_t n, bool sw, float* dst) {
  switch (mode) {
    case 0: convert_host<int8_t>(src, n, sw, dst); break;
    case 1: convert_host<int16_t>(src, n, sw, dst); break;
    case 2: convert_host<float>(src, n, sw, dst); break;
    case 6: convert_host<uint16_t>(src, n, sw, dst); break;
    default: mode_bytes(mode);
  }
}

// MATLAB std(X) along the first non-singleton axis of d (x fastest), N-1
// normalisation; all axes singleton -> one 0. d is updated to the result.
template <class T>
std::vector<double> std_first_axis(const T* v, int64_t d[3]) {
  int a = 0;
  while (a < 3 && d[a] == 1) ++a;
  if (a == 3) return std::vector<double>(1, 0.0);
  const int64_t len = d[a];
  const int64_t outer = d[0] * d[1] * d[2] / len;  // the axes before a are singleton
  std::vector<double> out((size_t)outer);
  for (int64_t o = 0; o < outer; ++o) {
    const T* r = v + o * len;
    double s = 0.0, s2 = 0.0;
    for (int64_t k = 0; k < len; ++k) s += (double)r[k];
    const double mu = s / (double)len;
    for (int64_t k = 0; k < len; ++k) s2 += ((double)r[k] - mu) * ((double)r[k] - mu);
    out[(size_t)o] = std::sqrt(s2 / (double)(len - 1));
  }
  d[a] = 1;
  return out;
}

// ------------------------------------------------------------ device side

template <int B>
__global__ void __launch_bounds__(256) byteswap_kernel(unsigned char* p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  unsigned char* q = p + i * B;
  unsigned char t[B];
#pragma unroll
  for (int k = 0; k < B; ++k) t[k] = q[k];
#pragma unroll
  for (int k = 0; k < B; ++k) q[k] = t[B - 1 - k];
}

template <int MODE>
__device__ __forceinline__ float load_mode(const void* p, int64_t i) {
  return (float)static_cast<const typename ModeT<MODE>::T*>(p)[i];
}

template <int MODE>
__global__ void __launch_bounds__(256) to_float_kernel(const void* __restrict__ in, int64_t n,
                                                       float* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = load_mode<MODE>(in, i);
}

// One particle's window: where the kept block starts in the tomogram
// (0-based), where it starts in the window (padLOW) and its extent; valid 0
// zero-fills the window.
struct Window {
  int src[3];
  int pad[3];
  int ext[3];
  int valid;
};

// BH_padZeros3d 'singleTaper' weights along one axis of the kept block (extent
// n): the low 7 voxels take flip(taper), the high 7 taper, taper(i) = 0.5 +
// 0.5 cos(i pi / 7); both apply where they overlap. With a zero fill value
// the six sequential edge passes of :189-229 reduce to the product of the
// three axis weights.
__device__ __forceinline__ float edge_weight(int i, int n) {
  float w = 1.0f;
  if (i < 7) w *= 0.5f + 0.5f * cosf((float)(6 - i) * 3.14159265358979f / 7.0f);
  if (i >= n - 7) w *= 0.5f + 0.5f * cosf((float)(i - (n - 7)) * 3.14159265358979f / 7.0f);
  return w;
}

// Exact quotient u / d for u < 2^24 and 1 <= d <= 2^16 from m = ceil(2^40 / d):
// u m / 2^40 = u / d + u (m - 2^40 / d) / 2^40 lies in [u / d, u / d + u / 2^40)
// and u d < 2^40, so the floor is floor(u / d).
__device__ __forceinline__ uint32_t div_magic(uint32_t u, uint64_t m) { return (uint32_t)(((uint64_t)u * m) >> 40); }

// grid: x over quads of window voxels (4 consecutive x of one row), y over
// particles. Stores are float4 (vec4) when the window's x size is a multiple
// of 4 and the output is 16-byte aligned. The quad index splits into (x
// quad, row, plane) by multiply-shift division (mq = ceil(2^40 / qx), my =
// ceil(2^40 / wy); mq = 0: plain division). For 2-byte modes a quad's four
// source values come from two aligned 8-byte loads (elements a0 .. a0 + 7,
// a0 = the first element rounded down to a multiple of 4) when those lie
// inside the tomogram's `total` elements, else one load per voxel.
template <int MODE>
__global__ void __launch_bounds__(256) extract_kernel(const void* __restrict__ tomo, int64_t tnx, int64_t tny,
                                                      int64_t total, const Window* __restrict__ wins, int wx, int wy,
                                                      int wz, uint64_t mq, uint64_t my, int taper, int vec4,
                                                      float* __restrict__ out) {
  using T = typename ModeT<MODE>::T;
  const int p = blockIdx.y;
  const int qx = (wx + 3) >> 2;
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (int64_t)qx * wy * wz) return;
  int X0, Y, Z;
  if (mq) {  // qx wy wz < 2^24, wy <= 2^16
    const uint32_t r = div_magic((uint32_t)q, mq), zz = div_magic(r, my);
    X0 = 4 * (int)((uint32_t)q - r * (uint32_t)qx);
    Y = (int)(r - zz * (uint32_t)wy);
    Z = (int)zz;
  } else {
    split3(q, qx, wy, X0, Y, Z);
    X0 *= 4;
  }
  const Window w = wins[p];
  const int y = Y - w.pad[1], z = Z - w.pad[2];
  const bool row_in = w.valid && (unsigned)y < (unsigned)w.ext[1] && (unsigned)z < (unsigned)w.ext[2];
  float v[4];
  const float wyz = row_in ? (taper ? edge_weight(y, w.ext[1]) * edge_weight(z, w.ext[2]) : 1.0f) : 0.0f;
  const int64_t row = ((int64_t)(w.src[2] + z) * tny + (w.src[1] + y)) * tnx + w.src[0];
  float e[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  bool wide = false;
  if constexpr (sizeof(T) == 2) {
    const int64_t base = row + X0 - w.pad[0];  // source element of the quad's first voxel
    const int64_t a0 = base & ~(int64_t)3;
    if (row_in && a0 >= 0 && a0 + 8 <= total) {
      const uint64_t* src = reinterpret_cast<const uint64_t*>(static_cast<const T*>(tomo) + a0);
      const uint64_t lo = src[0], hi = src[1];
      const int sh = 16 * (int)(base - a0);
      const uint64_t bits = sh ? ((lo >> sh) | (hi << (64 - sh))) : lo;
#pragma unroll
      for (int k = 0; k < 4; ++k) e[k] = (float)(T)(uint16_t)(bits >> (16 * k));
      wide = true;
    }
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int x = X0 + k - w.pad[0];
    v[k] = 0.0f;
    if (row_in && (unsigned)x < (unsigned)w.ext[0]) {
      const float a = wide ? e[k] : load_mode<MODE>(tomo, row + x);
      v[k] = taper ? a * (edge_weight(x, w.ext[0]) * wyz) : a;
    }
  }
  float* o = out + ((int64_t)p * wz + Z) * ((int64_t)wy * wx) + (int64_t)Y * wx + X0;
  if (vec4) {
    *reinterpret_cast<float4*>(o) = make_float4(v[0], v[1], v[2], v[3]);
  } else {
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if (X0 + k < wx) o[k] = v[k];
  }
}

// logicals/BH_isWindowValid.m:42-65 for one particle, per axis; returns the
// 1-based inclusive index range, the pads and shiftVAL, and whether the
// kept area is large enough for the mask ('noUse' otherwise).
bool window_valid(const int vol[3], const int win[3], const double mask_radius[3], const double c[3],
                  int64_t ind[2][3], int64_t pad[2][3], double shift[3]) {
  const double mr = std::max(mask_radius[0], std::max(mask_radius[1], mask_radius[2]));
  const double min_size = (mr + 6.0) * 2.0;
  bool ok = true;
  for (int k = 0; k < 3; ++k) {
    const int64_t low_c = (int64_t)std::ceil((win[k] - 1) / 2.0);
    const int64_t top_c = low_c - ((win[k] - 1) % 2);
    const double fc = std::floor(c[k]);
    const int64_t wc = (int64_t)fc;
    shift[k] = c[k] - fc;
    const int64_t LOW = wc - low_c;
    const int64_t TOP = vol[k] - (wc + top_c);
    pad[0][k] = LOW <= 0 ? (-LOW + 1) : 0;
    pad[1][k] = TOP <= 0 ? -TOP : 0;
    ind[0][k] = LOW + pad[0][k];
    ind[1][k] = wc + top_c - pad[1][k];
    const double avail = (double)(win[k] - pad[0][k] - pad[1][k]);
    if (avail - min_size < -2) ok = false;
  }
  return ok;
}

}  // namespace
}  // namespace emc

using namespace emc;

extern "C" {

emc_status emc_mrc_read_header(const char* path, emc_mrc_header* hdr) {
  return guard([&] {
    EMC_REQUIRE(path && hdr, "null pointer");
    File f(path, "rb");
    read_header_file(f, path, hdr);
  });
}

emc_status emc_mrc_read_volume(const char* path, const int range[6], float* h_out) {
  return guard([&] {
    EMC_REQUIRE(path && h_out, "null pointer");
    File f(path, "rb");
    emc_mrc_header h;
    read_header_file(f, path, &h);
    const int B = mode_bytes(h.mode);
    const int full[6] = {1, h.nx, 1, h.ny, 1, h.nz};
    const int* r = range ? range : full;
    const int n3[3] = {h.nx, h.ny, h.nz};
    for (int k = 0; k < 3; ++k)
      if (r[2 * k] < 1 || r[2 * k + 1] > n3[k] || r[2 * k] > r[2 * k + 1])
        fail(EMC_ERR_INVALID, "index range [%d %d] outside 1..%d (axis %d)", r[2 * k], r[2 * k + 1], n3[k], k);
    const int64_t rx = r[1] - r[0] + 1, ry = r[3] - r[2] + 1;
    // one read of the section's rows r[2]..r[3] at full width, then the x crop
    std::vector<unsigned char> buf((size_t)(ry * h.nx * B));
    float* o = h_out;
    for (int s = r[4]; s <= r[5]; ++s) {
      const int64_t off = h.data_offset + (((int64_t)(s - 1) * h.ny + (r[2] - 1)) * h.nx) * B;
      f.seek(off, path);
      f.read(buf.data(), buf.size(), path);
      for (int64_t j = 0; j < ry; ++j, o += rx)
        convert_host_mode(h.mode, buf.data() + (j * h.nx + (r[0] - 1)) * B, rx, h.swapped != 0, o);
    }
  });
}

emc_status emc_mrc_upload(emc_ctx* ctx, const char* path, int k0, int k1, void* d_raw) {
  return guard([&] {
    use_device(ctx);
    EMC_REQUIRE(path && d_raw, "null pointer");
    File f(path, "rb");
    emc_mrc_header h;
    read_header_file(f, path, &h);
    const int B = mode_bytes(h.mode);
    EMC_REQUIRE(k0 >= 1 && k1 <= h.nz && k0 <= k1, "section range [%d %d] outside 1..%d", k0, k1, h.nz);
    const int64_t section = (int64_t)h.nx * h.ny * B;
    const int64_t total = section * (k1 - k0 + 1);
    const int64_t chunk = std::max<int64_t>(section, std::min<int64_t>(total, (int64_t)64 << 20) / section * section);
    unsigned char* stage[2] = {nullptr, nullptr};
    hipEvent_t done[2] = {nullptr, nullptr};
    auto release = [&] {
      for (int b = 0; b < 2; ++b) {
        if (done[b]) (void)hipEventSynchronize(done[b]), (void)hipEventDestroy(done[b]);
        if (stage[b]) (void)hipHostFree(stage[b]);
      }
    };
    try {
      for (int b = 0; b < 2; ++b) {
        EMC_HIP(hipHostMalloc((void**)&stage[b], (size_t)chunk, hipHostMallocDefault));
        EMC_HIP(hipEventCreateWithFlags(&done[b], hipEventDisableTiming));
      }
      f.seek(h.data_offset + section * (k0 - 1), path);
      unsigned char* dst = static_cast<unsigned char*>(d_raw);
      int b = 0;
      for (int64_t off = 0; off < total; off += chunk, b ^= 1) {
        const int64_t bytes = std::min(chunk, total - off);
        EMC_HIP(hipEventSynchronize(done[b]));  // the copy that last used this buffer
        f.read(stage[b], (size_t)bytes, path);
        EMC_HIP(hipMemcpyAsync(dst + off, stage[b], (size_t)bytes, hipMemcpyHostToDevice, ctx->stream));
        EMC_HIP(hipEventRecord(done[b], ctx->stream));
      }
      if (h.swapped && B > 1) {
        const int64_t n = total / B;
        const int grid = ceil_div(n, 256);
        if (B == 2)
          hipLaunchKernelGGL(byteswap_kernel<2>, dim3(grid), dim3(256), 0, ctx->stream, dst, n);
        else
          hipLaunchKernelGGL(byteswap_kernel<4>, dim3(grid), dim3(256), 0, ctx->stream, dst, n);
        EMC_HIP(hipGetLastError());
      }
      EMC_HIP(hipStreamSynchronize(ctx->stream));
    } catch (...) {
      release();
      throw;
    }
    release();
  });
}

emc_status emc_mrc_to_float(emc_ctx* ctx, const void* d_raw, int mode, int64_t count, float* d_out) {
  return guard([&] {
    use_device(ctx);
    mode_bytes(mode);
    EMC_REQUIRE(count >= 0, "negative count");
    if (count == 0) return;
    EMC_REQUIRE(d_raw && d_out, "null pointer");
    const int grid = ceil_div(count, 256);
    switch (mode) {
      case 0: hipLaunchKernelGGL(to_float_kernel<0>, dim3(grid), dim3(256), 0, ctx->stream, d_raw, count, d_out); break;
      case 1: hipLaunchKernelGGL(to_float_kernel<1>, dim3(grid), dim3(256), 0, ctx->stream, d_raw, count, d_out); break;
      case 2: hipLaunchKernelGGL(to_float_kernel<2>, dim3(grid), dim3(256), 0, ctx->stream, d_raw, count, d_out); break;
      default: hipLaunchKernelGGL(to_float_kernel<6>, dim3(grid), dim3(256), 0, ctx->stream, d_raw, count, d_out); break;
    }
    EMC_HIP(hipGetLastError());
  });
}

emc_status emc_mrc_write(const char* path, const float* h_data, const int dims[3], const double* pixel,
                         const double* origin, int n_origin) {
  return guard([&] {
    EMC_REQUIRE(path && h_data && dims, "null pointer");
    EMC_REQUIRE(dims[0] > 0 && dims[1] > 0 && dims[2] > 0, "bad dims");
    EMC_REQUIRE(origin == nullptr || n_origin == 1 || n_origin == 3, "origin takes 1 or 3 values");
    const int64_t nx = dims[0], ny = dims[1], nz = dims[2];
    const int64_t n = nx * ny * nz;
    // min, max, mean and the nested std(std(std(vol))) of SAVE_IMG.m (FP64)
    double mn = h_data[0], mx = h_data[0], sum = 0.0;
    for (int64_t i = 0; i < n; ++i) {
      mn = std::min(mn, (double)h_data[i]);
      mx = std::max(mx, (double)h_data[i]);
      sum += h_data[i];
    }
    // std(std(std(vol))): each MATLAB std without a dimension reduces the
    // first non-singleton axis (N-1 normalisation); a scalar gives 0, so a
    // volume with a singleton axis ends at 0
    int64_t d[3] = {nx, ny, nz};
    std::vector<double> cur = std_first_axis(h_data, d);
    for (int step = 1; step < 3; ++step) cur = std_first_axis(cur.data(), d);
    const double rms = cur[0];

    unsigned char b[kHeaderBytes];
    std::memset(b, 0, sizeof(b));
    put<int32_t>(b, 0, (int32_t)nx);
    put<int32_t>(b, 4, (int32_t)ny);
    put<int32_t>(b, 8, (int32_t)nz);
    put<int32_t>(b, 12, 2);
    put<int32_t>(b, 28, (int32_t)nx);
    put<int32_t>(b, 32, (int32_t)ny);
    put<int32_t>(b, 36, (int32_t)nz);
    for (int k = 0; k < 3; ++k) {
      double o = 0.0;
      if (origin) o = n_origin == 1 ? -0.5 * dims[k] : origin[k];  // -cellDimension/2 before scaling
      put<float>(b, 196 + 4 * k, (float)o);
      put<float>(b, 40 + 4 * k, (float)(dims[k] * (pixel ? pixel[k] : 1.0)));
      put<float>(b, 52 + 4 * k, 90.0f);
    }
    put<int32_t>(b, 64, 1);
    put<int32_t>(b, 68, 2);
    put<int32_t>(b, 72, 3);
    put<float>(b, 76, (float)mn);
    put<float>(b, 80, (float)mx);
    put<float>(b, 84, (float)(sum / n));
    put<int16_t>(b, 88, 1);
    put<int32_t>(b, 152, kImodStamp);
    std::memcpy(b + 208, "MAP ", 4);
    const unsigned char stamp[4] = {68, 65, 0, 0};
    std::memcpy(b + 212, stamp, 4);
    put<float>(b, 216, (float)rms);
    put<int32_t>(b, 220, 1);  // writeHeader.m writes nLabels + 1 when nLabels < 10
    File f(path, "wb");
    f.write(b, kHeaderBytes, path);
    f.write(h_data, (size_t)n * sizeof(float), path);
  });
}

emc_status emc_extract_subtomos(emc_ctx* ctx, const void* d_tomo, int mode, const int tomo_dims[3], int n,
                                const double* h_centres, const int window[3], const double mask_radius[3],
                                int taper, float* d_out, int32_t* h_valid, double* h_shifts) {
  return guard([&] {
    use_device(ctx);
    mode_bytes(mode);
    EMC_REQUIRE(tomo_dims && window && mask_radius, "null pointer");
    EMC_REQUIRE(n >= 0, "negative particle count");
    for (int k = 0; k < 3; ++k) {
      EMC_REQUIRE(tomo_dims[k] > 0 && window[k] > 0, "bad tomogram or window size");
      EMC_REQUIRE(std::isfinite(mask_radius[k]), "mask radius not finite");
    }
    if (n == 0) return;
    EMC_REQUIRE(d_tomo && h_centres && d_out && h_valid && h_shifts, "null pointer");
    const int64_t wvox = (int64_t)window[0] * window[1] * window[2];
    EMC_REQUIRE(wvox <= ((int64_t)1 << 31), "window too large");
    std::vector<Window> wins((size_t)n);
    for (int p = 0; p < n; ++p) {
      const double* c = h_centres + 3 * (int64_t)p;
      EMC_REQUIRE(std::isfinite(c[0]) && std::isfinite(c[1]) && std::isfinite(c[2]),
                  "particle %d: centre not finite", p);
      int64_t ind[2][3], pad[2][3];
      const bool ok = window_valid(tomo_dims, window, mask_radius, c, ind, pad, h_shifts + 3 * (int64_t)p);
      Window& w = wins[(size_t)p];
      w.valid = ok ? 1 : 0;
      for (int k = 0; k < 3; ++k) {
        const int64_t ext = window[k] - pad[0][k] - pad[1][k];
        if (ok) {
          // the kept block lies inside the tomogram whenever it is non-empty
          if (!(ind[0][k] >= 1 && ind[1][k] <= tomo_dims[k] && ext == ind[1][k] - ind[0][k] + 1 && ext > 0))
            fail(EMC_ERR_RANGE, "particle %d: window outside the tomogram on axis %d", p, k);
          w.src[k] = (int)(ind[0][k] - 1);
          w.pad[k] = (int)pad[0][k];
          w.ext[k] = (int)ext;
        } else {
          w.src[k] = 0;
          w.pad[k] = 0;
          w.ext[k] = 0;
        }
      }
      h_valid[p] = w.valid;
    }
    const size_t bytes = sizeof(Window) * (size_t)n;
    Window* d_wins = static_cast<Window*>(scratch(ctx, "extract_windows", bytes));
    // the window table through the staging ring: the cut stays queued
    // behind the work already on the stream
    upload_async(ctx, d_wins, wins.data(), bytes);
    const int64_t quads = (int64_t)((window[0] + 3) / 4) * window[1] * window[2];
    const int64_t tnx = tomo_dims[0], tny = tomo_dims[1], total = tnx * tny * (int64_t)tomo_dims[2];
    const int vec4 = (window[0] % 4 == 0) && ((reinterpret_cast<uintptr_t>(d_out) & 15) == 0);
    // the kernel's multiply-shift division needs quads < 2^24 and divisors
    // <= 2^16; larger windows divide the plain way (mq = 0)
    const bool magic = quads < ((int64_t)1 << 24) && window[1] <= (1 << 16);
    const int qxw = (window[0] + 3) / 4;
    const uint64_t mq = magic ? (((uint64_t)1 << 40) + (uint64_t)qxw - 1) / (uint64_t)qxw : 0;
    const uint64_t my = magic ? (((uint64_t)1 << 40) + (uint64_t)window[1] - 1) / (uint64_t)window[1] : 0;
    KTimer t(ctx, "extract_subtomos");
    // grid.y carries the particles: at most 65535 per launch
    for (int p0 = 0; p0 < n; p0 += 65535) {
      const int np = std::min(65535, n - p0);
      const dim3 grid(ceil_div(quads, 256), (unsigned)np);
      const Window* w = d_wins + p0;
      float* o = d_out + (int64_t)p0 * wvox;
      switch (mode) {
        case 0: hipLaunchKernelGGL(extract_kernel<0>, grid, dim3(256), 0, ctx->stream, d_tomo, tnx, tny, total, w, window[0], window[1], window[2], mq, my, taper, vec4, o); break;
        case 1: hipLaunchKernelGGL(extract_kernel<1>, grid, dim3(256), 0, ctx->stream, d_tomo, tnx, tny, total, w, window[0], window[1], window[2], mq, my, taper, vec4, o); break;
        case 2: hipLaunchKernelGGL(extract_kernel<2>, grid, dim3(256), 0, ctx->stream, d_tomo, tnx, tny, total, w, window[0], window[1], window[2], mq, my, taper, vec4, o); break;
        default: hipLaunchKernelGGL(extract_kernel<6>, grid, dim3(256), 0, ctx->stream, d_tomo, tnx, tny, total, w, window[0], window[1], window[2], mq, my, taper, vec4, o); break;
      }
      EMC_HIP(hipGetLastError());
    }
  });
}

}  // extern "C"
