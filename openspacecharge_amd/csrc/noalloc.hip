// interpolate_field without device allocation, for callers that cannot
// allocate (code under HIP graph capture, tight serving loops, harnesses that
// forbid hipMalloc).  No upstream counterpart.
//
// Its kernels are those of osc_interpolate_field's packed-record gather; the
// node records (64 bytes per mesh node) and the clamped count live in
// `scratch`, a file-scope device array of fixed size that each GPU the
// library runs on holds until the process ends.  Host code never passes
// `scratch` itself to a kernel (on the host the name is only a stand-in):
// osc_ctx_create looks up the array's device address once per device with
// hipGetSymbolAddress, and the entry point passes that address.  Every call
// on one GPU shares the array: calls that overlap on one GPU (on different
// streams or host threads) race, and must be serialised by the caller.
#include <mutex>

#include "osc_internal.h"

#define OSC_API extern "C" __attribute__((visibility("default")))

namespace osc {
namespace {

// the clamped count at byte 0, the node records from byte kRecOff
constexpr int64_t kRecOff = 256;
constexpr int64_t kMaxNodes = OSC_NOALLOC_MAX_NODES;
constexpr size_t kScratchBytes = size_t(kRecOff + 64 * kMaxNodes);
constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};

void* scratch_ptr(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_mu);
  return g_ptr[device];
}

}  // namespace

int noalloc_setup(int device, std::string* err) {
  if (device < 0 || device >= kMaxDevices) {
    if (err) *err = "no-allocation gather: device index out of range";
    return OSC_ERR_ARG;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_ptr[device]) return OSC_OK;
  DeviceGuard g(device);
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    (void)hipGetLastError();
    if (err) *err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
    return OSC_ERR_HIP;
  }
  g_ptr[device] = p;
  return OSC_OK;
}

}  // namespace osc

using namespace osc;

OSC_API int osc_interpolate_field_noalloc(osc_ctx* c, const osc_mesh* m, const double* efield,
                                          const double* bfield, int64_t n, const double* x,
                                          const double* y, const double* z, double* E, double* B,
                                          int8_t* clamped, int64_t* n_clamped) {
  if (!c) return OSC_ERR_ARG;
  auto fail = [&](int code, const char* msg) {
    c->last_error = msg;
    return code;
  };
  try {
    DeviceGuard g(c->device);
    if (!m || n < 0 || (n > 0 && (!x || !y || !z)))
      return fail(OSC_ERR_ARG, "interpolate_field_noalloc: bad arguments");
    InterpArgs a{};
    int64_t nt = 1;
    for (int d = 0; d < 3; ++d) {
      a.nlo[d] = m->nlo[d];
      a.nhi[d] = m->nhi[d];
      a.nn[d] = m->nhi[d] - m->nlo[d] + 1;
      a.mn[d] = m->min[d];
      a.delta[d] = m->delta[d];
      if (a.nn[d] < 2) return fail(OSC_ERR_ARG, "interpolation needs at least 2 nodes per axis");
      nt *= a.nn[d];
    }
    if (nt > kMaxNodes)
      return fail(OSC_ERR_ARG,
                  "interpolate_field_noalloc: the mesh has more nodes than the fixed scratch "
                  "holds (OSC_NOALLOC_MAX_NODES = 1048576); use osc_interpolate_field");
    const bool wantE = efield && E, wantB = bfield && B;
    const void* dev_args[] = {wantE ? efield : nullptr, wantB ? bfield : nullptr,
                              n ? x : nullptr, n ? y : nullptr, n ? z : nullptr,
                              wantE && n ? E : nullptr, wantB && n ? B : nullptr,
                              n ? clamped : nullptr};
    for (const void* p : dev_args)
      if (p && !is_device_ptr(p))
        return fail(OSC_ERR_ARG,
                    "interpolate_field_noalloc: every array must be device memory (host arrays "
                    "would need staging buffers)");
    char* base = static_cast<char*>(scratch_ptr(c->device));
    if (!base) return fail(OSC_ERR_HIP, "interpolate_field_noalloc: scratch not set up");
    uint32_t* cnt = reinterpret_cast<uint32_t*>(base);
    interpolate_packed(c, a, wantE ? efield : nullptr, wantB ? bfield : nullptr, n, x, y, z,
                       wantE ? E : nullptr, wantB ? B : nullptr, clamped,
                       reinterpret_cast<double*>(base + kRecOff), cnt);
    if (n_clamped) {
      uint32_t* h = reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(c->mesh_host) + kNoallocHost);
      OSC_HIP(hipMemcpyAsync(h, cnt, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
      OSC_HIP(hipStreamSynchronize(c->stream));
      *n_clamped = *h;
    } else if (!c->async) {
      OSC_HIP(hipStreamSynchronize(c->stream));
    }
    c->last_error.clear();
    return OSC_OK;
  } catch (const Error& e) {
    c->last_error = e.msg;
    return e.code;
  }
}
