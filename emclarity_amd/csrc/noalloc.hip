// The allocation-free variant of emc_align_eval_grad, for callers that cannot
// allocate device memory (code under HIP graph capture, tight serving loops,
// harnesses that forbid hipMalloc). It runs the same kernels
// (the adjoint kernel the context selects + adjoint_fold_kernel, batched.hip)
// with the same list order and workgroup split, so its results equal
// emc_align_eval_grad's bit for bit; the brick (or support) list and the
// per-workgroup partial sums live in `scratch`, a file-scope device array of kScratchBytes
// that every GPU the library runs on holds until the process ends.
//
// Every call on one GPU shares that GPU's array and leaves its contents
// behind (the list of the last (n, radius, kernel) is kept and reused by the
// next call with the same values; the partial sums are written in full before
// they are read, so nothing needs clearing). Calls that overlap on one GPU
// (two host threads, or two contexts' streams) race on the array.
#include <cstring>
#include <mutex>

#include "ops.hpp"

namespace {

constexpr size_t kScratchBytes = size_t(16) << 20;  // 16 MiB per GPU
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Per device: the array's device address (hipGetSymbolAddress with the
// device current; the host-side name `scratch` is not a device address) and
// the list the array holds: (n, radius, kernel kind, brick side) and its
// entry count.
struct DeviceState {
  void* ptr = nullptr;
  int n = 0, kind = -1, S = 0;
  double radius = -1.0;
  int ns = 0;
};
std::mutex g_mutex;
DeviceState g_state[kMaxDevices];

DeviceState& device_state(int device) {
  EMC_REQUIRE(device >= 0 && device < kMaxDevices, "device %d beyond the %d the scratch table holds", device,
              kMaxDevices);
  DeviceState& s = g_state[device];
  if (!s.ptr) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    EMC_REQUIRE(e == hipSuccess && p, "hipGetSymbolAddress(scratch) on device %d: %s", device,
                hipGetErrorString(e));
    s.ptr = p;
  }
  return s;
}

size_t round256(size_t b) { return (b + 255) & ~size_t(255); }

}  // namespace

extern "C" {

emc_status emc_align_eval_grad_noalloc(emc_ctx* ctx, int n, int n_part, int table_format, const void* d_F2,
                                       const void* d_Q, const float* d_N, const double* d_A, const float* d_MdM,
                                       const float* d_shift, double support_radius, double* d_out) {
  return emc::guard([&] {
    emc::use_device(ctx);  // the context's device is current
    EMC_REQUIRE(n >= 8 && n <= 128, "box size must be 8..128");
    EMC_REQUIRE(n_part > 0, "empty batch");
    EMC_REQUIRE(d_F2 && d_Q && d_N && d_A && d_MdM && d_shift && d_out, "null pointer");
    EMC_REQUIRE(table_format == EMC_TABLE_C64 || table_format == EMC_TABLE_F16, "unknown table format %d",
                table_format);
    EMC_REQUIRE(support_radius > 0 && support_radius < 0.5 * n, "support radius must be in (0, n/2)");
    std::lock_guard<std::mutex> lock(g_mutex);
    DeviceState& st = device_state(ctx->device);
    unsigned char* base = (unsigned char*)st.ptr;
    emc::EvalSetup e{};
    emc::eval_grad_kind(ctx, n, &e.kind, &e.S, &e.waves);
    e.r2 = support_radius * support_radius;
    const bool same = st.n == n && st.radius == support_radius && st.kind == e.kind && st.S == e.S;
    // the list in host memory only: bricks (kind 1) or support frequencies
    std::vector<int4> bricks;
    std::vector<int2> supp;
    size_t entry = e.kind == 1 ? sizeof(int4) : sizeof(int2);
    int ns = st.ns;
    if (!same) {
      if (e.kind == 1) {
        bricks = emc::brick_host(n, support_radius, e.S);
        ns = (int)bricks.size();
      } else {
        supp = emc::support_host(n, support_radius);
        ns = (int)supp.size();
      }
    }
    e.ns = ns;
    e.g = e.kind == 1 ? ns : emc::eval_grad_blocks(ctx, n, n_part, ns);
    const size_t list_bytes = round256(entry * (size_t)ns);
    const size_t part_bytes = sizeof(double) * emc::kGradTerms * (size_t)e.g * (size_t)n_part;
    EMC_REQUIRE(list_bytes + part_bytes <= kScratchBytes,
                "list (%zu bytes) and partial sums (%zu bytes) exceed the %zu-byte scratch of the "
                "allocation-free entry; use emc_align_eval_grad",
                list_bytes, part_bytes, kScratchBytes);
    e.list = base;
    e.part = (double*)(base + list_bytes);
    if (!same) {
      st.n = 0;  // the array holds no valid list until the copy is done
      const void* src = e.kind == 1 ? (const void*)bricks.data() : (const void*)supp.data();
      EMC_HIP(hipMemcpyAsync(base, src, entry * (size_t)ns, hipMemcpyHostToDevice, ctx->stream));
      EMC_HIP(hipStreamSynchronize(ctx->stream));  // the host list goes away at return
      st.n = n;
      st.radius = support_radius;
      st.kind = e.kind;
      st.S = e.S;
      st.ns = ns;
    }
    emc::eval_grad_launch(ctx, e, n, n_part, table_format, d_F2, d_Q, d_N, d_A, d_MdM, d_shift, d_out, false,
                          false);
  });
}

}  // extern "C"
