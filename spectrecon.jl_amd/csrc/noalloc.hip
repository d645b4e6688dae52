// noalloc.hip -- sr_fft_conv_noalloc: sr_fft_conv (fft_conv! / fft_conv_adj!,
// src/fft_convolve.jl:33-130) for callers that cannot allocate device memory.
//
// sr_fft_conv allocates its tap tables, view id, chunk partials and (for
// host arrays) staging buffers on every call.  This variant keeps them in
// one file-scope device array, `scratch`, sized for the largest plane and
// PSF the kernels take (sr::kMaxPlane floats per plane buffer, sr::kMaxTaps
// taps per side): 102,848 floats (411,392 bytes) on each GPU the library runs
// on, held until the process ends.  Host code addresses it only through the per-device
// pointer that hipGetSymbolAddress returns (the array's name is a host-side
// stand-in).  Every call on a GPU shares that GPU's array and rewrites what
// it reads (tap tables in full, zero padding included; partials before they
// are summed), so calls that overlap on one GPU race; calls return after
// their work finished.  No hipMalloc / hipFree, no host heap.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <mutex>

#include "../../include/spectrecon_amd.h"
#include "kernels.h"
#include "taps.h"

namespace {

constexpr int kT = sr::kMaxTaps;
constexpr long kP = sr::kMaxPlane;
// scratch layout in floats (every region 256-byte aligned): the view id (an
// int, 0), taps u | w | cum (4 kT) | h (kT x kT), then three plane buffers:
// chunk partials, staged input, staged output
constexpr long kOffVid = 0;
constexpr long kOffU = 64;
constexpr long kOffW = kOffU + kT;
constexpr long kOffC = kOffW + kT;
constexpr long kOffH = kOffC + 4 * kT;
constexpr long kTabFloats = kOffH + (long)kT * kT;  // the table block copied per call
constexpr long kOffPart = (kTabFloats + 63) / 64 * 64;
constexpr long kOffIn = kOffPart + kP;
constexpr long kOffOut = kOffIn + kP;
constexpr long kScratchFloats = kOffOut + kP;
static_assert(kScratchFloats == 102848, "scratch size as documented");

__device__ __align__(256) float scratch[kScratchFloats];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
float* g_ptr[kMaxDevices] = {};  // scratch's address on each device, set once

int fail(int code, const char* msg) { return sr::record_error(code, msg); }

// the current device's copy of `scratch` (looked up once per device)
int scratch_ptr(float** out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    (void)hipGetLastError();
    return fail(SR_ERR_HIP, "no current HIP device");
  }
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      (void)hipGetLastError();
      char b[160];
      std::snprintf(b, sizeof b, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
      return fail(SR_ERR_HIP, b);
    }
    g_ptr[dev] = static_cast<float*>(p);
  }
  *out = g_ptr[dev];
  return SR_OK;
}

#define NA_TRY(expr)                                    \
  do {                                                  \
    const hipError_t _e = (expr);                       \
    if (_e != hipSuccess) {                             \
      (void)hipGetLastError();                          \
      char _b[200];                                     \
      std::snprintf(_b, sizeof _b, "%s failed: %s", #expr, hipGetErrorString(_e)); \
      return fail(SR_ERR_HIP, _b);                      \
    }                                                   \
  } while (0)

}  // namespace

extern "C" int sr_fft_conv_noalloc(float* out, const float* img, int32_t nx, int32_t nz, const float* ker, int32_t px,
                                   int32_t pz, int32_t adj, int32_t mem) {
  if (!out || !img || !ker) return fail(SR_ERR_ARG, "null argument");
  if (nx <= 0 || nz <= 0 || px <= 0 || pz <= 0) return fail(SR_ERR_ARG, "size");
  if (px % 2 == 0 || pz % 2 == 0) return fail(SR_ERR_ARG, "non-odd size psfs");
  if (px > kT || pz > kT) {
    char b[128];
    std::snprintf(b, sizeof b, "kernel %dx%d exceeds the %d-tap scratch tables", px, pz, kT);
    return fail(SR_ERR_UNSUPPORTED, b);
  }
  const long P = (long)nx * nz;
  if (P > kP) {
    char b[128];
    std::snprintf(b, sizeof b, "plane %dx%d exceeds the %ld-float scratch planes", nx, nz, kP);
    return fail(SR_ERR_UNSUPPORTED, b);
  }
  float* base = nullptr;
  int rc;
  if ((rc = scratch_ptr(&base))) return rc;
  const bool host = (mem & 0xff) == SR_MEM_HOST;
  // the kernel on the host (it is factorised there), then the tap tables
  float hk[kT * kT];
  if (host) std::memcpy(hk, ker, sizeof(float) * px * pz);
  else NA_TRY(hipMemcpy(hk, ker, sizeof(float) * px * pz, hipMemcpyDeviceToHost));
  double u[kT], w[kT];
  sr::Geo g{nx, 1, nz, 0, 0, px, pz};
  const bool sep = sr::psf_sep_supported(g) && sr::factor_rank1(hk, px, pz, u, w);
  const int nt = sr::round_taps(std::max(px, pz));
  static_assert(sizeof(int) == sizeof(float), "view id stored in a float slot");
  float tab[kTabFloats];
  std::memset(tab, 0, sizeof tab);
  long ncopy;  // floats of the table block this call uses
  if (sep) {
    for (int a = 0; a < px; ++a) tab[kOffU + a] = (float)u[a];
    for (int b = 0; b < pz; ++b) tab[kOffW + b] = (float)w[b];
    // (cum_sums takes [plane][nt] rows: u and w rows of length nt, contiguous per table)
    float hu[kT] = {}, hw[kT] = {};
    std::memcpy(hu, tab + kOffU, sizeof(float) * nt);
    std::memcpy(hw, tab + kOffW, sizeof(float) * nt);
    sr::cum_sums(hu, hw, 1, nt, tab + kOffC);
    ncopy = kOffC + 4 * nt;
  } else {
    for (int b = 0; b < pz; ++b)
      for (int a = 0; a < px; ++a) tab[kOffH + (long)b * nt + a] = hk[a + px * b];
    ncopy = kOffH + (long)pz * nt;
  }
  NA_TRY(hipMemcpy(base, tab, sizeof(float) * ncopy, hipMemcpyHostToDevice));
  sr::Taps t{};
  t.ntx = t.ntz = nt;
  t.vstride = 0;
  t.u = base + kOffU;
  t.w = base + kOffW;
  t.cum = base + kOffC;
  t.h = base + kOffH;
  const int* vid = reinterpret_cast<const int*>(base + kOffVid);
  const float* din = img;
  float* dout = out;
  if (host) {
    NA_TRY(hipMemcpy(base + kOffIn, img, sizeof(float) * P, hipMemcpyHostToDevice));
    din = base + kOffIn;
    dout = base + kOffOut;
  }
  if (adj) {
    NA_TRY(sr::launch_psf_adj(sep, din, t, vid, 1, dout, nullptr, 0, 1, g, nullptr));
  } else {
    int nchunk = 0;
    NA_TRY(sr::launch_psf_fwd(sep, din, t, vid, 1, base + kOffPart, 1, &nchunk, g, nullptr));
    NA_TRY(sr::launch_view_reduce(base + kOffPart, nchunk, vid, 1, dout, nullptr, nullptr, sr::EPI_VIEWS, g,
                                  nullptr));
  }
  NA_TRY(hipStreamSynchronize(nullptr));
  if (host) NA_TRY(hipMemcpy(out, dout, sizeof(float) * P, hipMemcpyDeviceToHost));
  return SR_OK;
}
