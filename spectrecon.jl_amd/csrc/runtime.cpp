// runtime.cpp -- device-resident SPECT plans and the C-ABI (spectrecon_amd.h).
//
// A plan owns: the mu-map, the per-view rotation parameters, the per-view
// attenuation factors (the reference recomputes them for every view of every
// call, src/project.jl:86-113; they depend only on mu, the angle and dy, so
// they are built once here), the PSF taps, and the batch work buffers.  All
// work of a plan runs on its own HIP stream; calls on one plan must not
// overlap (the reference's plans are not re-entrant either).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/spectrecon_amd.h"
#include "kernels.h"
#include "taps.h"

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIP_TRY(expr)                                                                   \
  do {                                                                                  \
    hipError_t _e = (expr);                                                             \
    if (_e != hipSuccess)                                                               \
      return fail(SR_ERR_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e), __FILE__, __LINE__); \
  } while (0)

constexpr double kPi = 3.14159265358979323846;

long jmod(long x, long y) {
  long r = x % y;
  return r < 0 ? r + y : r;
}

// Plan-time host loops over independent items (PSF planes, views), split
// over up to 16 host threads.
template <class F>
void host_parallel(long n, F&& f) {
  const long hw = std::max(1L, (long)std::thread::hardware_concurrency());
  const long nth = std::min({16L, hw, n / 16});
  if (nth <= 1) {
    for (long i = 0; i < n; ++i) f(i);
    return;
  }
  std::vector<std::thread> th;
  for (long w = 0; w < nth; ++w)
    th.emplace_back([&, w] {
      for (long i = w; i < n; i += nth) f(i);
    });
  for (auto& t : th) t.join();
}

// Plan-creation phase times, printed to stderr when SR_PLAN_TIMING is set
// (a diagnostic: the phases run the same either way).
struct PhaseClock {
  bool on = std::getenv("SR_PLAN_TIMING") != nullptr;
  std::chrono::steady_clock::time_point t = std::chrono::steady_clock::now();
  std::string line;
  void mark(const char* phase) {
    if (!on) return;
    const auto now = std::chrono::steady_clock::now();
    char b[96];
    std::snprintf(b, sizeof b, " %s %.2f", phase, std::chrono::duration<double, std::milli>(now - t).count());
    line += b;
    t = now;
  }
  ~PhaseClock() {
    if (on && !line.empty()) std::fprintf(stderr, "sr_plan_create ms:%s\n", line.c_str());
  }
};

// src/plan-rotate.jl:56
int rot_padsize(int nx) { return (int)std::ceil(1.0 + nx * std::sqrt(2.0) / 2.0 - nx / 2.0); }

// A device buffer that frees itself.
struct DBuf {
  void* p = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  ~DBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  int ensure(size_t bytes) {
    if (bytes <= n) return SR_OK;
    release();
    if (bytes == 0) return SR_OK;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
      p = nullptr;
      (void)hipGetLastError();
      return fail(SR_ERR_ALLOC, "hipMalloc of %zu bytes failed: %s", bytes, hipGetErrorString(e));
    }
    n = bytes;
    return SR_OK;
  }
  template <class T>
  T* as() const { return static_cast<T*>(p); }
};

// Per-view rotation parameters from the angle (host, Float64 where the
// reference computes in Float64).
//  :one -- src/rotatez.jl:249-265 and linearinterp! (:21-51), per-row tables
//          (b, C_odd, C_even) for the x pass (tan(θm/2)) and the y pass (-sin θm)
//  :two -- the affine map of src/rotatez.jl:348-353, coefficients rounded to
//          Float32 as AffineTransform2D{Float32} holds them.
void rot_params(double theta, int interp, int Np, sr::RotView& rv, int* tb, float* tw) {
  rv = sr::RotView{};
  if (interp == SR_INTERP_TWO) {
    rv.kind = (std::fmod(theta, 2 * kPi) == 0.0) ? 0 : 2;
    const double c = (1.0 + Np) / 2.0;
    const double phi = 2 * kPi - theta;
    const double cs = std::cos(phi), sn = std::sin(phi);
    const float cT = (float)c;
    rv.xx = (float)(cs * 1.0 - sn * 0.0);
    rv.xy = (float)(cs * 0.0 - sn * 1.0);
    const float xr = (float)(cs * (double)(-cT) - sn * (double)(-cT));
    rv.yx = (float)(cs * 0.0 + sn * 1.0);
    rv.yy = (float)(cs * 1.0 + sn * 0.0);
    const float yr = (float)(cs * (double)(-cT) + sn * (double)(-cT));
    rv.x0 = (float)((double)xr + c);
    rv.y0 = (float)((double)yr + c);
    return;
  }
  if (std::fmod(theta, 2 * kPi) == 0.0) {
    rv.kind = 0;
    return;
  }
  const long m = jmod((long)std::floor(0.5 + theta / (kPi / 2)), 4);
  rv.m = (int)m;
  const double rtol = std::sqrt((double)std::numeric_limits<float>::epsilon());
  const double q = m * (kPi / 2);
  if (theta == q || std::fabs(theta - q) <= rtol * std::max(std::fabs(theta), std::fabs(q))) {
    rv.kind = 1;
    return;
  }
  rv.kind = 2;
  const double mod_theta = theta - m * (kPi / 2);
  const double coef[2] = {std::tan(mod_theta / 2), -std::sin(mod_theta)};
  const double c = (Np + 1) / 2.0;
  for (int pass = 0; pass < 2; ++pass)
    for (int i = 1; i <= Np; ++i) {
      const double s = (i - c) * coef[pass] + 1;
      const double e = (i - c) * coef[pass] + Np;
      const double dec = std::ceil(s) - s;
      int b;
      if (e <= (double)Np) b = (int)std::ceil(e) - Np - 1;  // J = max(1, ceil(e) - Np + k [-1])
      else b = (int)std::floor(s) - 1;                     // J = min(Np, floor(s) + k [-1])
      tb[pass * Np + (i - 1)] = b;
      tw[pass * 2 * Np + 2 * (i - 1)] = (float)dec;
      tw[pass * 2 * Np + 2 * (i - 1) + 1] = (float)(1.0 - dec);
    }
}

// the tap tables as vectors (taps.h works on caller-provided arrays)
std::vector<float> cum_sums(const std::vector<float>& hu, const std::vector<float>& hw, int nplane, int nt) {
  std::vector<float> c((size_t)nplane * 4 * nt);
  sr::cum_sums(hu.data(), hw.data(), nplane, nt, c.data());
  return c;
}

bool factor_rank1(const float* h, int px, int pz, std::vector<double>& u, std::vector<double>& w) {
  u.assign(px, 0.0);
  w.assign(pz, 0.0);
  return sr::factor_rank1(h, px, pz, u.data(), w.data());
}

using sr::round_taps;

}  // namespace

// the calling thread's error message, for the other translation units
int sr::record_error(int code, const char* msg) {
  g_err = msg;
  return code;
}

// ===========================================================================
// Plan
// ===========================================================================
struct sr_plan_s {
  sr_plan_desc d{};
  sr::Geo g{};
  long V = 0, P = 0;  // voxels, pixels per view
  hipStream_t stream = nullptr;
  int separable = 0;
  int nt = 8;
  int vstride = 0;
  int max_batch = 32;
  int ncu = 256;  // compute units of the plan's device
  bool mfma = false;  // separable PSF passes on the matrix cores
  // z-fastest work layout: work, attenuation and partial images are
  // [.][y][x][z]; the :two rotations are z-vector gathers and the rotation
  // adjoint reads plan-time per-target source lists (zent)
  bool zfast = false;
  size_t bytes = 0;

  DBuf mumap, atten, taps_u, taps_w, taps_h, taps_c, rotv, tab_b, tab_w;
  // PSF by FFT (x-fastest plans whose padded planes fit the FFT kernels): the
  // kernels' spectra [plane][L2][L1/2 + 1] and the twiddles
  DBuf fspec, ftw;
  sr::FftPlan fft{};
  bool fft_ok = false;       // the FFT kernels take this plan
  bool fft_default = false;  // the built-in PSF method is the FFT
  DBuf work, fpart, ipart, vids, ones;
  DBuf st_img, st_views, st_aux1, st_aux2, st_x0;  // staging for host-pointer calls and EM temporaries
  DBuf em_ratio, em_back, em_asum;
  DBuf flag;  // one int: device-side argument checks
  DBuf zimg, zent;  // z-fastest image copy, per-target rotation lists
  DBuf zh;          // OS-EM: the PSF adjoint of a ones view (one PSF stack for all views)
  // OS-EM with one PSF stack on the fused path: H .* atten[v] for every view
  // (the normalisers' sources, [view][y][x][z]), valid while atten_h_fam is
  // the PSF adjoint family H was computed with (-1: not built)
  DBuf atten_h;
  int atten_h_fam = -1;
  DBuf gview, gwork2, gpart, gloss;  // sr_grad: residual views, G .* atten, loss partial sums
  int zk = 0;        // slots per target of zent
  std::vector<int> h_vids, h_order;

  // per-kernel event timing (sr_plan_set_timing)
  int timing = 0;  // bit mask of the kernel classes bracketed by events
  struct Pending {
    int cls;
    hipEvent_t a, b;
  };
  std::vector<Pending> pending;
  std::vector<hipEvent_t> event_pool;
  double t_ms[sr::K_NCLASS] = {};
  long t_n[sr::K_NCLASS] = {};

  // HIP graphs of the EM kernel sequence (sr_plan_set_graph), keyed by every
  // device address and count the sequence bakes in; most recent last
  int graph = 0;
  struct GraphEntry {
    std::vector<uintptr_t> key;
    hipGraphExec_t exec;
  };
  std::vector<GraphEntry> graphs;
  std::vector<std::vector<uintptr_t>> graph_seen;  // keys run once eagerly, not yet captured
  long graph_captures = 0, graph_replays = 0;

  sr::Taps taps() const {
    sr::Taps t;
    t.u = taps_u.as<float>();
    t.w = taps_w.as<float>();
    t.h = taps_h.as<float>();
    t.cum = taps_c.as<float>();
    t.ntx = nt;
    t.ntz = nt;
    t.vstride = vstride;
    return t;
  }
  long atten_vstride() const { return V; }
  // planes are split into chunks so that a launch fills the device: one
  // block per CU for the matrix-core PSF kernels (their LDS takes a CU),
  // about 512 blocks otherwise
  // kernel choices (sr_plan_set_choice / sr_plan_autotune), SR_CHOICE_* order:
  // PSF forward family, its chunk count in percent of the heuristic one, the
  // same two for the PSF adjoint, view slices of the rotation-adjoint gather,
  // the PSF method, where OS-EM sums its normalisers
  int32_t choice[SR_NCHOICE] = {0, 100, 0, 100, 0, 0, 0};
  // the PSF through the plane FFTs (else the direct / separable taps)
  bool use_fft() const {
    return fft_ok && (choice[SR_CHOICE_PSF_METHOD] == 2 || (choice[SR_CHOICE_PSF_METHOD] == 0 && fft_default));
  }
  // plane chunks per view for the PSF kernels: their blocks take a compute
  // unit each (LDS), so a launch of c chunks x nb views runs in
  // ceil(c nb / ncu) rounds of about (planes per chunk + 2) plane times (the
  // 2: a block's prologue and pipeline fill); the cheapest c, fewest on ties
  int chunks_base(int nb) const {
    int best = 1;
    long best_cost = -1;
    for (int c = 1; c <= d.ny; ++c) {
      const int yc = (d.ny + c - 1) / c;
      if ((d.ny + yc - 1) / yc != c) continue;  // same chunks as a smaller c
      const long rounds = ((long)c * std::max(1, nb) + ncu - 1) / ncu;
      const long cost = rounds * (yc + 2);
      if (best_cost < 0 || cost < best_cost) {
        best_cost = cost;
        best = c;
      }
    }
    return best;
  }
  int chunks_scaled(int nb, int pct) const {
    return std::min(d.ny, std::max(1, (int)(((long)chunks_base(nb) * pct + 50) / 100)));
  }
  int nchunk_fwd(int nb) const {
    if (mfma || use_fft()) return chunks_scaled(nb, choice[SR_CHOICE_FWD_CHUNKS]);
    const int target = std::max(1, (int)std::ceil(512.0 / std::max(1, nb)));
    return std::min(d.ny, target);
  }
  // PSF adjoint: one block per CU (its LDS holds the padded view and the
  // intermediate plane), the planes of a view split over ceil(ncu/nb) chunks
  int nchunk_adj(int nb) const { return chunks_scaled(nb, choice[SR_CHOICE_ADJ_CHUNKS]); }
  int fam_fwd() const { return choice[SR_CHOICE_FWD_FAMILY]; }
  int fam_adj() const { return choice[SR_CHOICE_ADJ_FAMILY]; }
  int ell_vs() const { return choice[SR_CHOICE_GATHER_SLICES]; }
  // OS-EM normalisers (one shared PSF stack, fused loop): each in its
  // subset's first-iteration gather (built-in), else all subsets in one
  // gather launch before the iterations
  bool em_norm_upfront() const { return choice[SR_CHOICE_EM_NORM] == 2; }
  int groups_adj(int nb, int* vpg) const {
    // views per rotation-adjoint group: enough blocks (nz * ngroups) to fill
    // the device; the :one kernel takes one view per block
    if (d.interpmeth == SR_INTERP_ONE) {
      *vpg = 1;
      return nb;
    }
    if (zfast) {  // one group: the gather kernel sums all views of the batch in its blocks
      *vpg = nb;
      return 1;
    }
    // time ~ (rounds of blocks over the CUs) x (views per block); each group
    // also writes and re-reads one partial image (about 1/16 of a view's time)
    const long slots = (long)ncu * sr::rot2_adj_blocks_per_cu(g);
    long best = -1;
    int bv = nb;
    for (int v = 1; v <= nb; ++v) {
      const long ng = (nb + v - 1) / v;
      const long cost = 16 * v * (((long)d.nz * ng + slots - 1) / slots) + ng;
      if (best < 0 || cost < best) {
        best = cost;
        bv = v;
      }
    }
    *vpg = bv;
    return (nb + bv - 1) / bv;
  }
};

namespace {

hipEvent_t pool_event(sr_plan p) {
  if (!p->event_pool.empty()) {
    hipEvent_t e = p->event_pool.back();
    p->event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) {
    (void)hipGetLastError();
    return nullptr;
  }
  return e;
}

// Brackets one kernel launch with events on the plan's stream when timing is on.
struct KTimer {
  sr_plan p;
  int cls;
  hipEvent_t a = nullptr;
  KTimer(sr_plan p_, int cls_) : p(p_), cls(cls_) {
    if ((p->timing & SR_TIME_CLASS(cls)) && (a = pool_event(p))) (void)hipEventRecord(a, p->stream);
  }
  ~KTimer() {
    if (!a) return;
    hipEvent_t b = pool_event(p);
    if (!b) {
      p->event_pool.push_back(a);
      return;
    }
    (void)hipEventRecord(b, p->stream);
    p->pending.push_back({cls, a, b});
  }
};

// fold completed event pairs into the per-class totals (after a stream sync)
void resolve_timing(sr_plan p) {
  for (auto& e : p->pending) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, e.a, e.b) == hipSuccess) {
      p->t_ms[e.cls] += ms;
      p->t_n[e.cls] += 1;
    } else {
      (void)hipGetLastError();
    }
    p->event_pool.push_back(e.a);
    p->event_pool.push_back(e.b);
  }
  p->pending.clear();
}

__attribute__((format(printf, 2, 3))) void appendf(std::string& s, const char* fmt, ...) {
  char b[256];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(b, sizeof b, fmt, ap);
  va_end(ap);
  s += b;
}

void drop_graphs(sr_plan p) {
  for (auto& e : p->graphs) (void)hipGraphExecDestroy(e.exec);
  p->graphs.clear();
  p->graph_seen.clear();
}

int ensure_batch_buffers(sr_plan p, int nb) {
  int rc;
  if ((rc = p->work.ensure(sizeof(float) * p->V * nb))) return rc;
  if ((rc = p->fpart.ensure(sizeof(float) * p->P * (size_t)nb * p->d.ny))) return rc;
  int vpg;
  const int ng = p->groups_adj(nb, &vpg);
  if ((rc = p->ipart.ensure(sizeof(float) * p->V * (size_t)ng))) return rc;
  return SR_OK;
}

int upload_vids(sr_plan p, const int32_t* index, int32_t nindex) {
  int rc;
  if ((rc = p->vids.ensure(sizeof(int) * std::max(1, nindex)))) return rc;
  p->h_vids.resize(nindex);
  for (int i = 0; i < nindex; ++i) {
    const int v = index ? index[i] : i;
    if (v < 0 || v >= p->d.nview) return fail(SR_ERR_ARG, "view index %d out of range 0..%d", v, p->d.nview - 1);
    p->h_vids[i] = v;
  }
  HIP_TRY(hipMemcpyAsync(p->vids.p, p->h_vids.data(), sizeof(int) * nindex, hipMemcpyHostToDevice, p->stream));
  return SR_OK;
}

// The PSF forward / adjoint of a batch on the plan's PSF method (x-fastest
// plans may take the plane FFTs; z-fastest ones fuse the rotation)
hipError_t psf_fwd(sr_plan p, const float* src, const int* vids, int nb, int nchunk, int* used, int family) {
  if (!p->zfast && p->use_fft())
    return sr::launch_psf_fwd_fft(src, p->fspec.as<float>(), p->vstride, vids, nb, p->fpart.as<float>(), nchunk, used,
                                  p->g, p->fft, p->ftw.as<float2>(), p->stream);
  return sr::launch_psf_fwd(p->separable, src, p->taps(), vids, nb, p->fpart.as<float>(), nchunk, used, p->g,
                            p->stream, p->zfast, p->atten.as<float>(), p->atten_vstride(), p->rotv.as<sr::RotView>(),
                            family);
}
hipError_t psf_adj(sr_plan p, const float* views, const int* vids, int nb, float* work, const float* atten, int nchunk,
                   int family) {
  if (!p->zfast && p->use_fft())
    return sr::launch_psf_adj_fft(views, p->fspec.as<float>(), p->vstride, vids, nb, work, atten, p->atten_vstride(),
                                  nchunk, p->g, p->fft, p->ftw.as<float2>(), p->stream);
  return sr::launch_psf_adj(p->separable, views, p->taps(), vids, nb, work, atten, atten ? p->atten_vstride() : 0,
                            nchunk, p->g, p->stream, p->zfast, family);
}

// views[ii] = A_{vids[ii]} image (or the ML-EM ratio y/(A x + b) when yobs)
// (zimg_ready: the plan's z-fastest image copy already holds `image`)
int run_project(sr_plan p, float* views, const float* image, const int* d_vids, int n, const float* yobs,
                const float* bkg, bool zimg_ready = false) {
  if (p->zfast && n > 0 && !zimg_ready) {
    KTimer kt(p, sr::K_ROT_FWD);
    HIP_TRY(sr::launch_to_zfast(image, p->zimg.as<float>(), p->g, p->stream));
  }
  for (int s0 = 0; s0 < n; s0 += p->max_batch) {
    const int nb = std::min(p->max_batch, n - s0);
    int rc;
    if ((rc = ensure_batch_buffers(p, nb))) return rc;
    if (!p->zfast) {  // (z-fastest: the PSF forward rotates while it stages)
      KTimer kt(p, sr::K_ROT_FWD);
      HIP_TRY(sr::launch_rot_fwd(p->d.interpmeth, image, p->work.as<float>(), p->atten.as<float>(),
                                 p->atten_vstride(), p->rotv.as<sr::RotView>(), d_vids + s0, nb, p->tab_b.as<int>(),
                                 p->tab_w.as<float>(), p->g, p->stream));
    }
    int nchunk = 0;
    {
      KTimer kt(p, sr::K_PSF_FWD);
      HIP_TRY(psf_fwd(p, p->zfast ? p->zimg.as<float>() : p->work.as<float>(), d_vids + s0, nb, p->nchunk_fwd(nb),
                      &nchunk, p->fam_fwd()));
    }
    {
      KTimer kt(p, sr::K_VIEW_REDUCE);
      HIP_TRY(sr::launch_view_reduce(p->fpart.as<float>(), nchunk, d_vids + s0, nb, views + (long)s0 * p->P, yobs,
                                     bkg, yobs ? sr::EPI_RATIO : sr::EPI_VIEWS, p->g, p->stream));
    }
  }
  return SR_OK;
}

// image = sum_ii A_{vids[ii]}' views[ii]  (mode IMG_WRITE), or with IMG_EM:
// image .*= (A' views) ./ asum.  *zimg_written (optional): set when the
// z-fastest image copy was written with the new image too
int run_backproject(sr_plan p, float* image, const float* views, const int* d_vids, int n, int mode,
                    const float* asum, bool* zimg_written = nullptr) {
  if (zimg_written) *zimg_written = false;
  const bool single = n <= p->max_batch;
  float* target = image;
  if (mode == sr::IMG_EM && !single) {
    int rc;
    if ((rc = p->em_back.ensure(sizeof(float) * p->V))) return rc;
    target = p->em_back.as<float>();
  }
  if (n == 0) {
    if (mode == sr::IMG_WRITE) HIP_TRY(sr::launch_fill(image, 0.f, p->V, p->stream));
    return SR_OK;
  }
  for (int s0 = 0; s0 < n; s0 += p->max_batch) {
    const int nb = std::min(p->max_batch, n - s0);
    int rc;
    if ((rc = ensure_batch_buffers(p, nb))) return rc;
    {
      KTimer kt(p, sr::K_PSF_ADJ);
      HIP_TRY(psf_adj(p, views + (long)s0 * p->P, d_vids + s0, nb, p->work.as<float>(), p->atten.as<float>(),
                      p->nchunk_adj(nb), p->fam_adj()));
    }
    int vpg;
    const int ng = p->groups_adj(nb, &vpg);
    int m = (s0 == 0) ? sr::IMG_WRITE : sr::IMG_ACCUM;
    if (mode == sr::IMG_EM && single) m = sr::IMG_EM;
    if (p->zfast) {
      {
        KTimer kt(p, sr::K_ROT_ADJ);
        HIP_TRY(sr::launch_rot2_adj_z(p->work.as<float>(), p->ipart.as<float>(), p->zent.as<int2>(), p->zk,
                                      d_vids + s0, nb, p->g, p->stream, nullptr, 0, nullptr, false, p->ell_vs()));
      }
      KTimer kt(p, sr::K_IMG_REDUCE);
      float* zout = (m == sr::IMG_EM && zimg_written) ? p->zimg.as<float>() : nullptr;
      HIP_TRY(sr::launch_img_reduce_z(p->ipart.as<float>(), ng, target, asum, m, p->g, p->stream, zout));
      if (zout) *zimg_written = true;
      continue;
    }
    {
      KTimer kt(p, sr::K_ROT_ADJ);
      HIP_TRY(sr::launch_rot_adj(p->d.interpmeth, p->work.as<float>(), p->ipart.as<float>(),
                                 p->rotv.as<sr::RotView>(), d_vids + s0, nb, vpg, p->tab_b.as<int>(),
                                 p->tab_w.as<float>(), p->g, p->stream));
    }
    KTimer kt(p, sr::K_IMG_REDUCE);
    HIP_TRY(sr::launch_img_reduce(p->ipart.as<float>(), ng, target, asum, p->V, m, p->stream));
  }
  if (mode == sr::IMG_EM && !single) {
    KTimer kt(p, sr::K_IMG_REDUCE);
    HIP_TRY(sr::launch_img_reduce(target, 1, image, asum, p->V, sr::IMG_EM, p->stream));
  }
  return SR_OK;
}

int finish(sr_plan p, int32_t mem) {
  if (mem & SR_ASYNC) return SR_OK;
  HIP_TRY(hipStreamSynchronize(p->stream));
  resolve_timing(p);
  return SR_OK;
}

// Host-pointer staging: copies `bytes` from host to a device buffer of at
// least max(bytes, min_bytes) bytes (an empty view list still gets a buffer
// the kernels may address, and nothing is read from the caller's array).
int stage_in(sr_plan p, DBuf& buf, const void* host, size_t bytes, size_t min_bytes = 0) {
  int rc;
  if ((rc = buf.ensure(std::max(bytes, min_bytes)))) return rc;
  if (bytes) HIP_TRY(hipMemcpyAsync(buf.p, host, bytes, hipMemcpyHostToDevice, p->stream));
  return SR_OK;
}

// The attenuation factors of every view from the plan's mu-map (device):
// each view's rotated mu volume, then the running sums along y and exp, in
// batches of 32 views on the plan's stream.
int build_atten(sr_plan p) {
  p->atten_h_fam = -1;  // (H .* atten[v] follows the factors)
  const sr_plan_desc& d = p->d;
  std::vector<int> all(d.nview);
  for (int v = 0; v < d.nview; ++v) all[v] = v;
  int rc;
  if ((rc = p->vids.ensure(sizeof(int) * d.nview))) return rc;
  HIP_TRY(hipMemcpyAsync(p->vids.p, all.data(), sizeof(int) * d.nview, hipMemcpyHostToDevice, p->stream));
  // (the copy reads `all` on the host: it lands before the stream moves on)
  HIP_TRY(hipStreamSynchronize(p->stream));
  const int chunk = 32;
  for (int v0 = 0; v0 < d.nview; v0 += chunk) {
    const int nb = std::min(chunk, d.nview - v0);
    float* dst = p->atten.as<float>() + p->V * v0;
    {
      KTimer kt(p, sr::K_ROT_FWD);
      if (p->zfast) {  // [v][y][x][z]
        if (v0 == 0) HIP_TRY(sr::launch_to_zfast(p->mumap.as<float>(), p->zimg.as<float>(), p->g, p->stream));
        HIP_TRY(sr::launch_rot2_fwd_z(p->zimg.as<float>(), dst, nullptr, 0, p->rotv.as<sr::RotView>(),
                                      p->vids.as<int>() + v0, nb, p->g, p->stream));
      } else {
        HIP_TRY(sr::launch_rot_fwd(d.interpmeth, p->mumap.as<float>(), dst, nullptr, 0, p->rotv.as<sr::RotView>(),
                                   p->vids.as<int>() + v0, nb, p->tab_b.as<int>(), p->tab_w.as<float>(), p->g,
                                   p->stream));
      }
    }
    KTimer kt(p, sr::K_ATTEN);
    HIP_TRY(sr::launch_atten_scan(dst, nb, p->g, d.dy, p->zfast, p->stream));
  }
  return SR_OK;
}

int check_device() {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
    (void)hipGetLastError();
    return fail(SR_ERR_HIP, "no HIP device is available");
  }
  return SR_OK;
}

int check_background(const float* b, size_t n) {
  for (size_t i = 0; i < n; ++i)
    if (!(b[i] > 0.f)) return fail(SR_ERR_ARG, "need background > 0");
  return SR_OK;
}

}  // namespace

extern "C" {

const char* sr_last_error(void) { return g_err.c_str(); }
int sr_version(void) { return 10000; }

int sr_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return n;
}

int sr_malloc(void** dptr, size_t bytes) {
  if (!dptr) return fail(SR_ERR_ARG, "null pointer");
  hipError_t e = hipMalloc(dptr, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return fail(SR_ERR_ALLOC, "hipMalloc of %zu bytes failed: %s", bytes, hipGetErrorString(e));
  }
  return SR_OK;
}
int sr_free(void* dptr) {
  HIP_TRY(hipFree(dptr));
  return SR_OK;
}
int sr_memcpy_h2d(void* dst, const void* src, size_t bytes) {
  HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return SR_OK;
}
int sr_memcpy_d2h(void* dst, const void* src, size_t bytes) {
  HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return SR_OK;
}
int sr_memcpy_d2d(void* dst, const void* src, size_t bytes) {
  if (bytes == 0) return SR_OK;
  if (!dst || !src) return fail(SR_ERR_ARG, "null argument");
  const bool al = ((reinterpret_cast<uintptr_t>(dst) | reinterpret_cast<uintptr_t>(src)) & 15) == 0;
  const size_t n16 = al ? bytes / 16 : 0;
  if (n16) HIP_TRY(sr::launch_copy16(dst, src, (long)n16, nullptr));
  if (bytes > 16 * n16)
    HIP_TRY(hipMemcpy(static_cast<char*>(dst) + 16 * n16, static_cast<const char*>(src) + 16 * n16, bytes - 16 * n16,
                      hipMemcpyDeviceToDevice));
  HIP_TRY(hipDeviceSynchronize());
  return SR_OK;
}
int sr_memset(void* dptr, int value, size_t bytes) {
  HIP_TRY(hipMemset(dptr, value, bytes));
  return SR_OK;
}
int sr_device_synchronize(void) {
  HIP_TRY(hipDeviceSynchronize());
  return SR_OK;
}

// src/spectplan.jl:58-137
int sr_plan_create(const sr_plan_desc* desc, const float* mumap, const float* psfs, const float* viewangle,
                   sr_plan* out) {
  if (!desc || !mumap || !psfs || !viewangle || !out) return fail(SR_ERR_ARG, "null argument");
  *out = nullptr;
  const sr_plan_desc& d = *desc;
  // checks in the reference's order and with its messages (src/spectplan.jl:76-93)
  if (d.nx != d.ny) return fail(SR_ERR_ARG, "nx != ny");
  if (d.nx % 2 != 0 || d.ny % 2 != 0) return fail(SR_ERR_ARG, "nx odd");
  if (d.px % 2 == 0 || d.pz % 2 == 0) return fail(SR_ERR_ARG, "non-odd size psfs");
  if (d.nx <= 0 || d.nz <= 0 || d.px <= 0 || d.pz <= 0 || d.nview <= 0) return fail(SR_ERR_ARG, "size");
  PhaseClock clk;
  const long pxz = (long)d.px * d.pz;
  const long nplane = (long)d.ny * d.nview;
  {  // every plane point-symmetric (src/spectplan.jl:83), planes checked in parallel
    std::atomic<bool> asym{false};
    host_parallel(nplane, [&](long pl) {
      const float* h = psfs + pl * pxz;
      for (long k = 0; k < pxz / 2 + 1; ++k)
        if (!(h[k] == h[pxz - 1 - k])) {
          asym.store(true, std::memory_order_relaxed);
          return;
        }
    });
    if (asym.load()) return fail(SR_ERR_ARG, "asym. psf");
  }
  clk.mark("checks");
  if (d.interpmeth != SR_INTERP_ONE && d.interpmeth != SR_INTERP_TWO) return fail(SR_ERR_ARG, "bad interpmeth");
  if (d.mode != SR_MODE_FAST && d.mode != SR_MODE_MEM) return fail(SR_ERR_ARG, "bad mode");
  if (d.px > sr::kMaxTaps || d.pz > sr::kMaxTaps)
    return fail(SR_ERR_UNSUPPORTED, "PSF size %dx%d exceeds the %d-tap kernels", d.px, d.pz, sr::kMaxTaps);
  if ((long)d.nx * d.nz > sr::kMaxPlane) return fail(SR_ERR_UNSUPPORTED, "plane %dx%d too large", d.nx, d.nz);
  int rc;
  if ((rc = check_device())) return rc;

  std::unique_ptr<sr_plan_s> guard(new sr_plan_s());
  sr_plan p = guard.get();
  p->d = d;
  const int pad = rot_padsize(d.nx);
  p->g = sr::Geo{d.nx, d.ny, d.nz, pad, d.nx + 2 * pad, d.px, d.pz};
  if (d.interpmeth == SR_INTERP_ONE && p->g.Np > sr::kMaxNpOne)
    return fail(SR_ERR_UNSUPPORTED, ":one rotation supports nx up to 140 (padded %d > %d)", d.nx + 2 * pad,
                sr::kMaxNpOne);
  p->V = (long)d.nx * d.ny * d.nz;
  p->P = (long)d.nx * d.nz;
  p->max_batch = std::max(1, std::min(d.nview, 64));
  {
    int dev = 0, n = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && n > 0)
      p->ncu = n;
    (void)hipGetLastError();
  }
  auto bail = [](int code) { return code; };
  if (hipStreamCreate(&p->stream) != hipSuccess) {
    (void)hipGetLastError();
    return bail(fail(SR_ERR_HIP, "hipStreamCreate failed"));
  }

  // --- PSF taps: dedupe views, choose separable vs direct ---
  std::atomic<bool> differs{false};
  host_parallel(d.nview - 1, [&](long v) {
    if (!differs.load(std::memory_order_relaxed) &&
        std::memcmp(psfs, psfs + (v + 1) * d.ny * pxz, sizeof(float) * d.ny * pxz) != 0)
      differs.store(true, std::memory_order_relaxed);
  });
  const bool shared = !differs.load();
  clk.mark("dedupe");
  p->vstride = shared ? 0 : d.ny;
  const int nstore = shared ? d.ny : d.ny * d.nview;
  p->nt = round_taps(std::max(d.px, d.pz));
  const bool mfma_ok = sr::psf_mfma_supported(p->g, p->nt);
  bool sep = d.conv != SR_CONV_DIRECT && (mfma_ok || sr::psf_sep_supported(p->g));
  std::vector<double> u, w;
  if (sep)
    for (int pl = 0; pl < nstore && sep; ++pl) sep = factor_rank1(psfs + (long)pl * pxz, d.px, d.pz, u, w);
  if (d.conv == SR_CONV_SEPARABLE && !sep) return bail(fail(SR_ERR_ARG, "PSF planes are not rank 1"));
  if (d.conv < SR_CONV_AUTO || d.conv > SR_CONV_FFT) return bail(fail(SR_ERR_ARG, "bad conv"));
  p->separable = sep ? 1 : 0;
  p->mfma = sep && mfma_ok;
  // the matrix-core kernels take any table length: the exact one keeps their
  // LDS tap tables and wave buffers smallest (the VALU kernels are
  // instantiated per multiple of 8)
  if (p->mfma) p->nt = std::max(d.px, d.pz);
  if (sep) {
    std::vector<float> hu((size_t)nstore * p->nt, 0.f), hw((size_t)nstore * p->nt, 0.f);
    for (int pl = 0; pl < nstore; ++pl) {
      factor_rank1(psfs + (long)pl * pxz, d.px, d.pz, u, w);
      for (int a = 0; a < d.px; ++a) hu[(size_t)pl * p->nt + a] = (float)u[a];
      for (int b = 0; b < d.pz; ++b) hw[(size_t)pl * p->nt + b] = (float)w[b];
    }
    const std::vector<float> hc = cum_sums(hu, hw, nstore, p->nt);
    if ((rc = p->taps_u.ensure(sizeof(float) * hu.size())) || (rc = p->taps_w.ensure(sizeof(float) * hw.size())) ||
        (rc = p->taps_c.ensure(sizeof(float) * hc.size())))
      return bail(rc);
    HIP_TRY(hipMemcpy(p->taps_u.p, hu.data(), sizeof(float) * hu.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(p->taps_w.p, hw.data(), sizeof(float) * hw.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(p->taps_c.p, hc.data(), sizeof(float) * hc.size(), hipMemcpyHostToDevice));
  } else {
    std::vector<float> hh((size_t)nstore * d.pz * p->nt, 0.f);
    for (int pl = 0; pl < nstore; ++pl)
      for (int b = 0; b < d.pz; ++b)
        for (int a = 0; a < d.px; ++a)
          hh[((size_t)pl * d.pz + b) * p->nt + a] = psfs[(long)pl * pxz + a + (long)d.px * b];
    if ((rc = p->taps_h.ensure(sizeof(float) * hh.size()))) return bail(rc);
    HIP_TRY(hipMemcpy(p->taps_h.p, hh.data(), sizeof(float) * hh.size(), hipMemcpyHostToDevice));
  }

  clk.mark("taps");
  // --- rotation parameters ---
  const int Np = p->g.Np;
  std::vector<sr::RotView> rv(d.nview);
  std::vector<int> tb((size_t)d.nview * 2 * Np, 0);
  std::vector<float> tw((size_t)d.nview * 4 * Np, 0.f);
  for (int v = 0; v < d.nview; ++v)
    rot_params((double)viewangle[v], d.interpmeth, Np, rv[v], tb.data() + (size_t)v * 2 * Np,
               tw.data() + (size_t)v * 4 * Np);
  if ((rc = p->rotv.ensure(sizeof(sr::RotView) * d.nview)) || (rc = p->tab_b.ensure(sizeof(int) * tb.size())) ||
      (rc = p->tab_w.ensure(sizeof(float) * tw.size())))
    return bail(rc);
  HIP_TRY(hipMemcpy(p->rotv.p, rv.data(), sizeof(sr::RotView) * d.nview, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(p->tab_b.p, tb.data(), sizeof(int) * tb.size(), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(p->tab_w.p, tw.data(), sizeof(float) * tw.size(), hipMemcpyHostToDevice));

  clk.mark("rotation");
  // --- work layout: z-fastest for :two with the matrix-core PSF passes
  // (SR_LAYOUT=zyx in the environment keeps the x-fastest layout) ---
  {
    const char* lay = std::getenv("SR_LAYOUT");
    const bool force_zyx = lay && std::strcmp(lay, "zyx") == 0;
    p->zfast = !force_zyx && p->mfma && d.interpmeth == SR_INTERP_TWO && d.conv != SR_CONV_FFT &&
               sr::zfast_supported(p->g);
  }
  if (p->zfast) {
    const char* lists = std::getenv("SR_LISTS");
    const long NN = (long)d.nx * d.nx;
    if (lists && std::strcmp(lists, "host") == 0) {
      // (the host builder: SR_LISTS=host, for A/B tests of the device one)
      std::vector<int2> ent;
      p->zfast = sr::build_rot2_lists(rv.data(), d.nview, p->g, &p->zk, ent);
      if (p->zfast) {
        if ((rc = p->zent.ensure(sizeof(int2) * ent.size())) || (rc = p->zimg.ensure(sizeof(float) * p->V)))
          return bail(rc);
        HIP_TRY(hipMemcpy(p->zent.p, ent.data(), sizeof(int2) * ent.size(), hipMemcpyHostToDevice));
      }
    } else {
      // on the device: the longest list (K, at least 4), then the slots
      const int nblk = sr::rot2_list_blocks(p->g);
      std::vector<int> bmax((size_t)nblk * d.nview);
      DBuf dmax;
      if ((rc = dmax.ensure(sizeof(int) * bmax.size()))) return bail(rc);
      HIP_TRY(sr::launch_rot2_list_count(p->rotv.as<sr::RotView>(), d.nview, p->g, dmax.as<int>(), p->stream));
      HIP_TRY(hipMemcpyAsync(bmax.data(), dmax.p, sizeof(int) * bmax.size(), hipMemcpyDeviceToHost, p->stream));
      HIP_TRY(hipStreamSynchronize(p->stream));
      p->zk = std::max(4, *std::max_element(bmax.begin(), bmax.end()));
      p->zfast = p->zk <= sr::kMaxEll;
      if (p->zfast) {
        if ((rc = p->zent.ensure(sizeof(int2) * d.nview * p->zk * NN)) ||
            (rc = p->zimg.ensure(sizeof(float) * p->V)))
          return bail(rc);
        HIP_TRY(sr::launch_rot2_list_fill(p->rotv.as<sr::RotView>(), d.nview, p->g, p->zk, p->zent.as<int2>(),
                                          p->stream));
      }
    }
  }

  clk.mark("lists");
  // --- PSF spectra for the plane FFTs (x-fastest plans whose planes fit):
  // from the PSF planes on the device, in double, once ---
  if (!p->zfast && sr::psf_fft_plan(p->g, &p->fft)) {
    const long nf = (long)p->fft.L2 * (p->fft.L1 / 2 + 1);
    DBuf dpsf;
    std::vector<float2> tw;
    sr::psf_fft_twiddles(p->fft, tw);
    if ((rc = dpsf.ensure(sizeof(float) * pxz * nstore)) || (rc = p->fspec.ensure(sizeof(float) * nf * nstore)) ||
        (rc = p->ftw.ensure(sizeof(float2) * tw.size())))
      return bail(rc);
    HIP_TRY(hipMemcpy(dpsf.p, psfs, sizeof(float) * pxz * nstore, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(p->ftw.p, tw.data(), sizeof(float2) * tw.size(), hipMemcpyHostToDevice));
    HIP_TRY(sr::launch_psf_spectrum(dpsf.as<float>(), nstore, p->g, p->fft, p->fspec.as<float>(), p->stream));
    HIP_TRY(hipStreamSynchronize(p->stream));
    p->fft_ok = true;
    // built in: the FFT for general PSFs of 11 x 11 taps and more (their direct
    // 2-D cost grows with px pz, the FFT's with log L); sr_plan_autotune
    // times both
    p->fft_default = d.conv == SR_CONV_FFT || (d.conv == SR_CONV_AUTO && !p->separable && pxz >= 121);
  } else if (d.conv == SR_CONV_FFT) {
    return bail(fail(SR_ERR_UNSUPPORTED, "the PSF FFT kernels do not take %dx%d planes with %dx%d PSFs", d.nx, d.nz,
                     d.px, d.pz));
  }

  clk.mark("spectra");
  // --- mu-map and attenuation factors for every view ---
  if ((rc = p->mumap.ensure(sizeof(float) * p->V))) return bail(rc);
  HIP_TRY(hipMemcpy(p->mumap.p, mumap, sizeof(float) * p->V, hipMemcpyHostToDevice));
  if ((rc = p->atten.ensure(sizeof(float) * p->V * d.nview))) return bail(rc);
  clk.mark("alloc");
  if ((rc = build_atten(p))) return bail(rc);
  HIP_TRY(hipStreamSynchronize(p->stream));
  clk.mark("atten");
  p->bytes = p->mumap.n + p->atten.n + p->taps_u.n + p->taps_w.n + p->taps_h.n + p->taps_c.n + p->rotv.n + p->tab_b.n +
             p->tab_w.n + p->zimg.n + p->zent.n + p->fspec.n + p->ftw.n;  // (zh, atten_h: the OS-EM calls')
  *out = guard.release();
  return SR_OK;
}

int sr_plan_destroy(sr_plan p) {
  if (!p) return SR_OK;
  if (p->stream) {
    (void)hipStreamSynchronize(p->stream);
    resolve_timing(p);
    for (hipEvent_t e : p->event_pool) (void)hipEventDestroy(e);
    for (auto& e : p->graphs) (void)hipGraphExecDestroy(e.exec);
    (void)hipStreamDestroy(p->stream);
  }
  delete p;
  return SR_OK;
}

// A new mu-map for an existing plan: the attenuation factors of every view
// are rebuilt in place (SPECTplan's mu-dependent state, src/spectplan.jl:71,
// src/project.jl:77-116); the rotation tables, PSF taps and lists stay.
int sr_plan_set_mumap(sr_plan p, const float* mumap, int32_t mem) {
  if (!p || !mumap) return fail(SR_ERR_ARG, "null argument");
  const bool host = (mem & 0xff) == SR_MEM_HOST;
  HIP_TRY(hipMemcpyAsync(p->mumap.p, mumap, sizeof(float) * p->V,
                         host ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice, p->stream));
  if (host) HIP_TRY(hipStreamSynchronize(p->stream));  // (the caller may reuse its array at once)
  int rc;
  if ((rc = build_atten(p))) return rc;
  return finish(p, host ? (mem & ~SR_ASYNC) : mem);
}

int sr_plan_set_timing(sr_plan p, int32_t enable) {
  if (!p) return fail(SR_ERR_ARG, "null plan");
  HIP_TRY(hipStreamSynchronize(p->stream));
  resolve_timing(p);
  p->timing = enable == 1 ? SR_TIME_ALL : (enable & SR_TIME_ALL);
  for (int k = 0; k < sr::K_NCLASS; ++k) {
    p->t_ms[k] = 0;
    p->t_n[k] = 0;
  }
  return SR_OK;
}

int sr_plan_set_graph(sr_plan p, int32_t enable) {
  if (!p) return fail(SR_ERR_ARG, "null plan");
  HIP_TRY(hipStreamSynchronize(p->stream));
  p->graph = enable ? 1 : 0;
  if (!p->graph) drop_graphs(p);
  return SR_OK;
}

int sr_plan_graph_stats(sr_plan p, int64_t* captures, int64_t* replays) {
  if (!p) return fail(SR_ERR_ARG, "null plan");
  if (captures) *captures = p->graph_captures;
  if (replays) *replays = p->graph_replays;
  return SR_OK;
}

int sr_plan_set_choice(sr_plan p, const int32_t* c) {
  if (!p || !c) return fail(SR_ERR_ARG, "null argument");
  if (c[SR_CHOICE_FWD_FAMILY] < 0 || c[SR_CHOICE_FWD_FAMILY] > 5)
    return fail(SR_ERR_ARG, "PSF forward family must be 0..5");
  if (c[SR_CHOICE_ADJ_FAMILY] < 0 || c[SR_CHOICE_ADJ_FAMILY] > 2)
    return fail(SR_ERR_ARG, "PSF adjoint family must be 0..2");
  for (int k : {SR_CHOICE_FWD_CHUNKS, SR_CHOICE_ADJ_CHUNKS})
    if (c[k] < 10 || c[k] > 1000) return fail(SR_ERR_ARG, "chunk percent must be in 10..1000");
  const int vs = c[SR_CHOICE_GATHER_SLICES];
  if (vs != 0 && vs != 2 && vs != 4 && vs != 8) return fail(SR_ERR_ARG, "gather slices must be 0, 2, 4 or 8");
  if (c[SR_CHOICE_PSF_METHOD] < 0 || c[SR_CHOICE_PSF_METHOD] > 2) return fail(SR_ERR_ARG, "PSF method must be 0..2");
  if (c[SR_CHOICE_EM_NORM] < 0 || c[SR_CHOICE_EM_NORM] > 2) return fail(SR_ERR_ARG, "EM normaliser choice must be 0..2");
  if (c[SR_CHOICE_PSF_METHOD] == 2 && !p->fft_ok)
    return fail(SR_ERR_UNSUPPORTED, "the plan has no PSF spectra (z-fastest, or planes the FFT kernels do not take)");
  HIP_TRY(hipStreamSynchronize(p->stream));
  drop_graphs(p);
  std::copy(c, c + SR_NCHOICE, p->choice);
  return SR_OK;
}

int sr_plan_get_choice(sr_plan p, int32_t* c) {
  if (!p || !c) return fail(SR_ERR_ARG, "null argument");
  std::copy(p->choice, p->choice + SR_NCHOICE, c);
  return SR_OK;
}

// Times each choice's candidates on the kernel it steers, one choice at a
// time (the others at their current values), over batches of nsub views
// spread over the orbit; inputs are the plan's own buffers filled with ones.
// A candidate replaces the built-in choice only when it is faster by more
// than the timing noise (kTuneMargin), so that equal kernels do not flip
// between runs.
constexpr double kTuneMargin = 0.97;
int sr_plan_autotune(sr_plan p, int32_t nsub, char* log, size_t log_bytes) {
  if (!p) return fail(SR_ERR_ARG, "null plan");
  HIP_TRY(hipStreamSynchronize(p->stream));
  drop_graphs(p);
  const int32_t defaults[SR_NCHOICE] = {0, 100, 0, 100, 0, 0, 0};
  std::string out;
  if (nsub <= 0) {
    std::copy(defaults, defaults + SR_NCHOICE, p->choice);
    out = "{\"reset\": true}";
  } else {
    const int nb = std::min({(int)nsub, p->d.nview, p->max_batch});
    int rc;
    if ((rc = ensure_batch_buffers(p, nb)) || (rc = p->em_ratio.ensure(sizeof(float) * p->P * nb)) ||
        (rc = p->st_img.ensure(sizeof(float) * p->V)) || (rc = p->vids.ensure(sizeof(int) * nb)))
      return rc;
    std::vector<int> vid(nb);
    for (int k = 0; k < nb; ++k) vid[k] = (int)((long)k * p->d.nview / nb);
    HIP_TRY(hipMemcpyAsync(p->vids.p, vid.data(), sizeof(int) * nb, hipMemcpyHostToDevice, p->stream));
    const int* dv = p->vids.as<int>();
    float* img = p->zfast ? p->zimg.as<float>() : p->st_img.as<float>();
    HIP_TRY(sr::launch_fill(img, 1.f, p->V, p->stream));
    HIP_TRY(sr::launch_fill(p->em_ratio.as<float>(), 1.f, p->P * nb, p->stream));
    HIP_TRY(sr::launch_fill(p->work.as<float>(), 1.f, p->V * nb, p->stream));
    // the two timing events, destroyed on every return path
    struct EventPair {
      hipEvent_t a = nullptr, b = nullptr;
      ~EventPair() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
      }
    } ev;
    HIP_TRY(hipEventCreate(&ev.a));
    HIP_TRY(hipEventCreate(&ev.b));
    hipEvent_t e0 = ev.a, e1 = ev.b;
    const sr::Taps t = p->taps();
    // microseconds per launch: best of 5 rounds of 4 launches, after one
    // untimed launch; negative when the candidate fails to launch
    auto time_us = [&](auto&& launch) -> double {
      if (launch() != hipSuccess) {
        (void)hipGetLastError();
        return -1;
      }
      double best = 1e30;
      for (int r = 0; r < 5; ++r) {
        (void)hipEventRecord(e0, p->stream);
        for (int k = 0; k < 4; ++k) (void)launch();
        (void)hipEventRecord(e1, p->stream);
        float ms = 0;
        if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
          (void)hipGetLastError();
          return -1;
        }
        best = std::min(best, 250.0 * ms);
      }
      return best;
    };
    const bool psf_choices = p->separable && p->mfma;
    // the device ramps up over the first milliseconds of work: about 10 ms of
    // the built-in PSF passes, untimed, so that the first candidate timed
    // (the built-in one) is not the one that pays for it
    for (int k = 0; k < 80; ++k) {
      int used = 0;
      HIP_TRY(psf_fwd(p, p->zfast ? img : p->work.as<float>(), dv, nb, p->nchunk_fwd(nb), &used, p->fam_fwd()));
    }
    appendf(out, "{\"nsub\": %d, \"views_per_launch\": %d, \"layout\": \"%s\"", (int)nsub, nb,
         p->zfast ? "z-fastest" : "x-fastest");
    // PSF method (plans with PSF spectra): the direct taps against the plane
    // FFTs, forward (with its view reduction) plus adjoint, each at the
    // built-in chunk counts
    if (p->fft_ok) {
      appendf(out, ", \"psf_method\": {\"candidates\": [");
      double best = -1, def_total = -1;
      int bm = 0;
      for (int m : {1, 2}) {
        p->choice[SR_CHOICE_PSF_METHOD] = m;
        const double fu = time_us([&]() {
          int used = 0;
          hipError_t e = psf_fwd(p, p->work.as<float>(), dv, nb, p->nchunk_fwd(nb), &used, p->fam_fwd());
          if (e == hipSuccess)
            e = sr::launch_view_reduce(p->fpart.as<float>(), used, dv, nb, p->em_ratio.as<float>(), nullptr, nullptr,
                                       sr::EPI_VIEWS, p->g, p->stream);
          return e;
        });
        const double au = time_us([&]() {
          return psf_adj(p, p->em_ratio.as<float>(), dv, nb, p->work.as<float>(), p->atten.as<float>(),
                         p->nchunk_adj(nb), p->fam_adj());
        });
        const double total = (fu > 0 && au > 0) ? fu + au : -1;
        const bool is_default = (m == 2) == p->fft_default;
        appendf(out, "%s{\"method\": \"%s\", \"code\": %d, \"fwd_us\": %.2f, \"adj_us\": %.2f}", m == 1 ? "" : ", ",
                m == 2 ? "fft" : p->separable ? "separable" : "direct", m, fu, au);
        if (is_default) def_total = total;
        if (total > 0 && (best < 0 || total < best)) {
          best = total;
          bm = m;
        }
      }
      // the built-in method stays unless the other is clearly faster
      const int dm = p->fft_default ? 2 : 1;
      if (def_total > 0 && best > kTuneMargin * def_total) {
        bm = dm;
        best = def_total;
      }
      p->choice[SR_CHOICE_PSF_METHOD] = bm == dm ? 0 : bm;
      appendf(out, "], \"choice\": {\"method\": \"%s\", \"code\": %d, \"us\": %.2f}}",
              bm == 2 ? "fft" : p->separable ? "separable" : "direct", p->choice[SR_CHOICE_PSF_METHOD], best);
    }
    // PSF forward (with the view reduction its chunk count sets) and adjoint:
    // family x chunk percent
    for (int dir = 0; dir < 2 && psf_choices && !p->use_fft(); ++dir) {
      const int fk = dir == 0 ? SR_CHOICE_FWD_FAMILY : SR_CHOICE_ADJ_FAMILY;
      const int ck = dir == 0 ? SR_CHOICE_FWD_CHUNKS : SR_CHOICE_ADJ_CHUNKS;
      // families that fit this plan (forward: producer/consumer waves with 4
      // or 8 producers, column-owned, tile; adjoint: column-owned, tile)
      // (the built-in one first: PSF_AUTO)
      std::vector<int> fams;
      const int auto_npw = p->zfast ? sr::psf_fwd_auto_npw(p->g, t.ntx) : 0;
      if (dir == 0) {
        fams.push_back(sr::PSF_AUTO);
        if (auto_npw != 8 && p->zfast && sr::psf_fwd_ws_ok(p->g, t.ntx, 8)) fams.push_back(sr::PSF_WS8);
        if (auto_npw != 4 && p->zfast && sr::psf_fwd_ws_ok(p->g, t.ntx, 4)) fams.push_back(sr::PSF_WS4);
        if (p->zfast && sr::psf_fwd_ws2_ok(p->g, t.ntx)) fams.push_back(sr::PSF_WS2);
        if (auto_npw != 0 && p->zfast && sr::psf_fwd_col_ok(p->g, t.ntx)) fams.push_back(sr::PSF_COL);
        if (p->zfast) fams.push_back(sr::PSF_TILE);
      } else {
        fams.push_back(sr::PSF_AUTO);
        if (p->zfast && sr::psf_adj_col_ok(p->g, t.ntx)) fams.push_back(sr::PSF_TILE);
      }
      auto fam_name = [&](int f) {
        if (f == sr::PSF_AUTO)
          return dir == 1 ? (p->zfast && sr::psf_adj_col_ok(p->g, t.ntx) ? "column" : "tile")
                 : auto_npw == 8 ? "overlap8" : auto_npw == 4 ? "overlap" : p->zfast && sr::psf_fwd_col_ok(p->g, t.ntx) ? "column" : "tile";
        return f == sr::PSF_TILE ? "tile" : f == sr::PSF_COL ? "column" : f == sr::PSF_WS8 ? "overlap8"
               : f == sr::PSF_WS2 ? "pairs" : "overlap";
      };
      appendf(out, ", \"%s\": {\"candidates\": [", dir == 0 ? "psf_fwd" : "psf_adj");
      double best = -1, def_us = -1;
      int bf = defaults[fk], bc = defaults[ck], n = 0;
      for (int fam : fams) {
        int last_chunks = -1;
        for (int pct : {50, 100, 200, 400}) {
          const int chunks = p->chunks_scaled(nb, pct);
          if (chunks == last_chunks) continue;
          last_chunks = chunks;
          double us;
          if (dir == 0)
            us = time_us([&]() {
              int used = 0;
              hipError_t e = sr::launch_psf_fwd(p->separable, p->zfast ? img : p->work.as<float>(), t, dv, nb,
                                                p->fpart.as<float>(), chunks, &used, p->g, p->stream, p->zfast,
                                                p->atten.as<float>(), p->atten_vstride(), p->rotv.as<sr::RotView>(),
                                                fam);
              if (e == hipSuccess)
                e = sr::launch_view_reduce(p->fpart.as<float>(), used, dv, nb, p->em_ratio.as<float>(), nullptr,
                                           nullptr, sr::EPI_VIEWS, p->g, p->stream);
              return e;
            });
          else
            us = time_us([&]() {
              return sr::launch_psf_adj(p->separable, p->em_ratio.as<float>(), t, dv, nb, p->work.as<float>(),
                                        p->atten.as<float>(), p->atten_vstride(), chunks, p->g, p->stream, p->zfast,
                                        fam);
            });
          appendf(out, "%s{\"family\": \"%s\", \"code\": %d, \"chunks\": %d, \"us\": %.2f}", n++ ? ", " : "",
                  fam_name(fam), fam, chunks, us);
          if (fam == defaults[fk] && pct == defaults[ck]) def_us = us;
          if (us > 0 && (best < 0 || us < best)) {
            best = us;
            bf = fam;
            bc = pct;
          }
        }
      }
      // the built-in choice stays unless a candidate is clearly faster
      if (def_us > 0 && best > kTuneMargin * def_us) {
        bf = defaults[fk];
        bc = defaults[ck];
        best = def_us;
      }
      p->choice[fk] = bf;
      p->choice[ck] = bc;
      appendf(out, "], \"choice\": {\"family\": \"%s\", \"code\": %d, \"chunk_percent\": %d, \"us\": %.2f}}",
              fam_name(bf), bf, bc, best);
    }
    // rotation-adjoint gather (z-fastest): view slices per block
    if (p->zfast) {
      appendf(out, ", \"gather\": {\"candidates\": [");
      double best = -1, def_us = -1;
      int bv = 0, n = 0;
      for (int vs : {0, 2, 4, 8}) {  // (0: the built-in count for this batch)
        const double us = time_us([&]() {
          return sr::launch_rot2_adj_z(p->work.as<float>(), p->ipart.as<float>(), p->zent.as<int2>(), p->zk, dv, nb,
                                       p->g, p->stream, nullptr, 0, nullptr, false, vs);
        });
        appendf(out, "%s{\"slices\": %d, \"us\": %.2f}", n++ ? ", " : "", vs, us);
        if (vs == 0) {
          def_us = us;
          continue;
        }
        if (us > 0 && (best < 0 || us < best)) {
          best = us;
          bv = vs;
        }
      }
      if (def_us > 0 && best > kTuneMargin * def_us) {
        bv = 0;
        best = def_us;
      }
      p->choice[SR_CHOICE_GATHER_SLICES] = bv;
      appendf(out, "], \"choice\": {\"slices\": %d, \"us\": %.2f}}", bv, best);
    }
    out += "}";
    HIP_TRY(hipStreamSynchronize(p->stream));
  }
  if (log && log_bytes) {
    const size_t n = std::min(out.size(), log_bytes - 1);
    std::memcpy(log, out.data(), n);
    log[n] = 0;
  }
  return SR_OK;
}

int sr_plan_get_timing(sr_plan p, int32_t kernel, double* total_ms, int64_t* launches) {
  if (!p) return fail(SR_ERR_ARG, "null plan");
  if (kernel < 0 || kernel >= sr::K_NCLASS) return fail(SR_ERR_ARG, "kernel class out of range");
  HIP_TRY(hipStreamSynchronize(p->stream));
  resolve_timing(p);
  if (total_ms) *total_ms = p->t_ms[kernel];
  if (launches) *launches = p->t_n[kernel];
  return SR_OK;
}

int sr_plan_info(sr_plan p, size_t* device_bytes, int32_t* conv_method, int32_t* max_batch) {
  if (!p) return fail(SR_ERR_ARG, "null plan");
  if (device_bytes)
    *device_bytes = p->bytes + p->work.n + p->fpart.n + p->ipart.n + p->st_img.n + p->st_views.n + p->st_aux1.n +
                    p->st_aux2.n + p->st_x0.n + p->em_ratio.n + p->em_back.n + p->em_asum.n + p->ones.n +
                    p->zh.n + p->atten_h.n;
  if (conv_method) *conv_method = p->use_fft() ? SR_CONV_FFT : p->separable ? SR_CONV_SEPARABLE : SR_CONV_DIRECT;
  if (max_batch) *max_batch = p->max_batch;
  return SR_OK;
}

int sr_plan_layout(sr_plan p, int32_t* layout) {
  if (!p || !layout) return fail(SR_ERR_ARG, "null argument");
  *layout = p->zfast ? SR_LAYOUT_Z_FASTEST : SR_LAYOUT_X_FASTEST;
  return SR_OK;
}

int sr_plan_set_batch(sr_plan p, int32_t max_batch) {
  if (!p) return fail(SR_ERR_ARG, "null plan");
  if (max_batch < 1) return fail(SR_ERR_ARG, "max_batch must be >= 1");
  p->max_batch = max_batch;
  return SR_OK;
}

int sr_plan_synchronize(sr_plan p) {
  if (!p) return fail(SR_ERR_ARG, "null plan");
  HIP_TRY(hipStreamSynchronize(p->stream));
  return SR_OK;
}

// src/project.jl:143-163
int sr_project(sr_plan p, float* views, const float* image, const int32_t* index, int32_t nindex, int32_t mem) {
  if (!p || !views || !image) return fail(SR_ERR_ARG, "null argument");
  if (!index && nindex != p->d.nview) return fail(SR_ERR_DIM, "nindex must equal nview when index is NULL");
  int rc;
  if ((rc = upload_vids(p, index, nindex))) return rc;
  const bool host = (mem & 0xff) == SR_MEM_HOST;
  const float* d_img = image;
  float* d_views = views;
  if (host) {
    if ((rc = stage_in(p, p->st_img, image, sizeof(float) * p->V))) return rc;
    if ((rc = p->st_views.ensure(sizeof(float) * p->P * std::max(1, nindex)))) return rc;
    d_img = p->st_img.as<float>();
    d_views = p->st_views.as<float>();
  }
  if ((rc = run_project(p, d_views, d_img, p->vids.as<int>(), nindex, nullptr, nullptr))) return rc;
  if (host) {
    HIP_TRY(hipMemcpyAsync(views, d_views, sizeof(float) * p->P * nindex, hipMemcpyDeviceToHost, p->stream));
    HIP_TRY(hipStreamSynchronize(p->stream));
    resolve_timing(p);
    return SR_OK;
  }
  return finish(p, mem);
}

// src/backproject.jl:132-162
int sr_backproject(sr_plan p, float* image, const float* views, const int32_t* index, int32_t nindex, int32_t mem) {
  if (!p || !views || !image) return fail(SR_ERR_ARG, "null argument");
  if (!index && nindex != p->d.nview) return fail(SR_ERR_DIM, "nindex must equal nview when index is NULL");
  int rc;
  if ((rc = upload_vids(p, index, nindex))) return rc;
  const bool host = (mem & 0xff) == SR_MEM_HOST;
  float* d_img = image;
  const float* d_views = views;
  if (host) {
    if ((rc = stage_in(p, p->st_views, views, sizeof(float) * p->P * nindex, sizeof(float) * p->P))) return rc;
    if ((rc = p->st_img.ensure(sizeof(float) * p->V))) return rc;
    d_img = p->st_img.as<float>();
    d_views = p->st_views.as<float>();
  }
  if ((rc = run_backproject(p, d_img, d_views, p->vids.as<int>(), nindex, sr::IMG_WRITE, nullptr))) return rc;
  if (host) {
    HIP_TRY(hipMemcpyAsync(image, d_img, sizeof(float) * p->V, hipMemcpyDeviceToHost, p->stream));
    HIP_TRY(hipStreamSynchronize(p->stream));
    resolve_timing(p);
    return SR_OK;
  }
  return finish(p, mem);
}

namespace {
int run_mu_vjp(sr_plan p, float* d_gmu, const float* r, const int* d_vids, int nindex);
}  // namespace

int sr_grad(sr_plan p, double* loss, float* grad_x, float* grad_mu, const float* image, const float* y,
            const int32_t* index, int32_t nindex, int32_t mem) {
  if (!p || !loss || !grad_x || !grad_mu || !image || !y) return fail(SR_ERR_ARG, "null argument");
  if (!index && nindex != p->d.nview) return fail(SR_ERR_DIM, "nindex must equal nview when index is NULL");
  if (!p->zfast)
    return fail(SR_ERR_UNSUPPORTED,
                "the mu-gradient needs a z-fastest plan (:two, rank-1 PSFs on planes of multiples of 16)");
  int rc;
  if ((rc = upload_vids(p, index, nindex))) return rc;
  const bool host = (mem & 0xff) == SR_MEM_HOST;
  const long n = std::max(1, nindex), PN = p->P * n;
  const float *d_img = image, *d_y = y;
  float *d_gx = grad_x, *d_gmu = grad_mu;
  if (host) {
    if ((rc = stage_in(p, p->st_x0, image, sizeof(float) * p->V)) ||
        (rc = stage_in(p, p->st_aux1, y, sizeof(float) * p->P * nindex, sizeof(float) * PN)) ||
        (rc = p->st_img.ensure(sizeof(float) * p->V)) ||
        (rc = p->st_aux2.ensure(sizeof(float) * p->V)))
      return rc;
    d_img = p->st_x0.as<float>();
    d_y = p->st_aux1.as<float>();
    d_gx = p->st_img.as<float>();
    d_gmu = p->st_aux2.as<float>();
  }
  if ((rc = p->gview.ensure(sizeof(float) * PN)) || (rc = p->gpart.ensure(sizeof(double) * ((PN + 255) / 256))) ||
      (rc = p->gloss.ensure(sizeof(double))) ||
      (rc = p->gwork2.ensure(sizeof(float) * p->V * std::min<long>(n, p->max_batch))))
    return rc;
  const int* d_vids = p->vids.as<int>();
  float* r = p->gview.as<float>();
  // r = 2 (A x - y), L = sum (A x - y)^2; the projection leaves x's z-fastest copy in zimg
  if ((rc = run_project(p, r, d_img, d_vids, nindex, nullptr, nullptr))) return rc;
  HIP_TRY(sr::launch_resid_loss(r, d_y, r, p->P * nindex, p->gpart.as<double>(), p->gloss.as<double>(), p->stream));
  if (nindex == 0) {
    HIP_TRY(sr::launch_fill(d_gx, 0.f, p->V, p->stream));
    HIP_TRY(sr::launch_fill(d_gmu, 0.f, p->V, p->stream));
  }
  // both gradients from one PSF adjoint per batch: G = conv^T(r) (no
  // attenuation); the scan forms E = G .* (R(x) .* a) for dL/dmu, rotating x
  // on the fly, and leaves G .* a, the back-projector's work volume, for
  // dL/dx = A' r
  const sr::Taps t = p->taps();
  for (int s0 = 0; s0 < nindex; s0 += p->max_batch) {
    const int nb = std::min(p->max_batch, nindex - s0);
    if ((rc = ensure_batch_buffers(p, nb))) return rc;
    const int m = s0 == 0 ? sr::IMG_WRITE : sr::IMG_ACCUM;
    {
      KTimer kt(p, sr::K_PSF_ADJ);
      HIP_TRY(sr::launch_psf_adj(p->separable, r + (long)s0 * p->P, t, d_vids + s0, nb, p->work.as<float>(), nullptr,
                                 0, p->nchunk_adj(nb), p->g, p->stream, true, p->fam_adj()));
    }
    {  // (the rotation of x formed inside the scan, from zimg)
      KTimer kt(p, sr::K_ROT_FWD);
      HIP_TRY(sr::launch_mu_grad_scan_rot(p->work.as<float>(), p->gwork2.as<float>(), p->zimg.as<float>(),
                                          p->atten.as<float>(), p->atten_vstride(), p->rotv.as<sr::RotView>(),
                                          d_vids + s0, nb, p->g, p->d.dy, p->stream));
    }
    for (int k = 0; k < 2; ++k) {  // dL/dx from G .* a, then dL/dmu from the scan
      {
        KTimer kt(p, sr::K_ROT_ADJ);
        HIP_TRY(sr::launch_rot2_adj_z(k ? p->work.as<float>() : p->gwork2.as<float>(), p->ipart.as<float>(),
                                      p->zent.as<int2>(), p->zk, d_vids + s0, nb, p->g, p->stream, nullptr, 0,
                                      nullptr, false, p->ell_vs()));
      }
      KTimer kt(p, sr::K_IMG_REDUCE);
      HIP_TRY(sr::launch_img_reduce_z(p->ipart.as<float>(), 1, k ? d_gmu : d_gx, nullptr, m, p->g, p->stream));
    }
  }
  // (no view: the residual kernel writes no loss, and the sum over no views is 0)
  if (nindex > 0) HIP_TRY(hipMemcpyAsync(loss, p->gloss.p, sizeof(double), hipMemcpyDeviceToHost, p->stream));
  else *loss = 0.0;
  if (host) {
    HIP_TRY(hipMemcpyAsync(grad_x, d_gx, sizeof(float) * p->V, hipMemcpyDeviceToHost, p->stream));
    HIP_TRY(hipMemcpyAsync(grad_mu, d_gmu, sizeof(float) * p->V, hipMemcpyDeviceToHost, p->stream));
  }
  // the loss lands in host memory: always wait
  HIP_TRY(hipStreamSynchronize(p->stream));
  resolve_timing(p);
  return SR_OK;
}

// src/project.jl:103-116 differentiated in mu: grad_mu = sum over the indexed
// views of the mu-cotangent of <dy, A(x, mu)>
int sr_project_vjp_mu(sr_plan p, float* grad_mu, const float* image, const float* dy, const int32_t* index,
                      int32_t nindex, int32_t mem) {
  if (!p || !grad_mu || !image || !dy) return fail(SR_ERR_ARG, "null argument");
  if (!index && nindex != p->d.nview) return fail(SR_ERR_DIM, "nindex must equal nview when index is NULL");
  if (!p->zfast)
    return fail(SR_ERR_UNSUPPORTED,
                "the mu-gradient needs a z-fastest plan (:two, rank-1 PSFs on planes of multiples of 16)");
  int rc;
  if ((rc = upload_vids(p, index, nindex))) return rc;
  const bool host = (mem & 0xff) == SR_MEM_HOST;
  const long PN = p->P * std::max(1, nindex);
  const float *d_img = image, *d_dy = dy;
  float* d_gmu = grad_mu;
  if (host) {
    if ((rc = stage_in(p, p->st_x0, image, sizeof(float) * p->V)) ||
        (rc = stage_in(p, p->st_aux1, dy, sizeof(float) * p->P * nindex, sizeof(float) * PN)) ||
        (rc = p->st_aux2.ensure(sizeof(float) * p->V)))
      return rc;
    d_img = p->st_x0.as<float>();
    d_dy = p->st_aux1.as<float>();
    d_gmu = p->st_aux2.as<float>();
  }
  {  // the image's z-fastest copy, which the rotation of x reads
    KTimer kt(p, sr::K_ROT_FWD);
    HIP_TRY(sr::launch_to_zfast(d_img, p->zimg.as<float>(), p->g, p->stream));
  }
  if ((rc = run_mu_vjp(p, d_gmu, d_dy, p->vids.as<int>(), nindex))) return rc;
  if (host) HIP_TRY(hipMemcpyAsync(grad_mu, d_gmu, sizeof(float) * p->V, hipMemcpyDeviceToHost, p->stream));
  return finish(p, mem);
}

namespace {
// dL/dmu = sum_v R_v' (-dy (reverse cumulative sum along y of conv'(r_v) .* R_v x .* atten_v, halved at j)) for
// the view cotangents r (the z-fastest image copy holds x)
int run_mu_vjp(sr_plan p, float* d_gmu, const float* r, const int* d_vids, int nindex) {
  int rc;
  const sr::Taps t = p->taps();
  if (nindex == 0) HIP_TRY(sr::launch_fill(d_gmu, 0.f, p->V, p->stream));
  for (int s0 = 0; s0 < nindex; s0 += p->max_batch) {
    const int nb = std::min(p->max_batch, nindex - s0);
    if ((rc = ensure_batch_buffers(p, nb))) return rc;
    {
      KTimer kt(p, sr::K_PSF_ADJ);
      HIP_TRY(sr::launch_psf_adj(p->separable, r + (long)s0 * p->P, t, d_vids + s0, nb, p->work.as<float>(), nullptr,
                                 0, p->nchunk_adj(nb), p->g, p->stream, true, p->fam_adj()));
    }
    {  // (the rotation of x formed inside the scan, from zimg)
      KTimer kt(p, sr::K_ROT_FWD);
      HIP_TRY(sr::launch_mu_grad_scan_rot(p->work.as<float>(), nullptr, p->zimg.as<float>(), p->atten.as<float>(),
                                          p->atten_vstride(), p->rotv.as<sr::RotView>(), d_vids + s0, nb, p->g,
                                          p->d.dy, p->stream));
    }
    {
      KTimer kt(p, sr::K_ROT_ADJ);
      HIP_TRY(sr::launch_rot2_adj_z(p->work.as<float>(), p->ipart.as<float>(), p->zent.as<int2>(), p->zk,
                                    d_vids + s0, nb, p->g, p->stream, nullptr, 0, nullptr, false, p->ell_vs()));
    }
    KTimer kt(p, sr::K_IMG_REDUCE);
    HIP_TRY(sr::launch_img_reduce_z(p->ipart.as<float>(), 1, d_gmu, nullptr, s0 == 0 ? sr::IMG_WRITE : sr::IMG_ACCUM,
                                    p->g, p->stream));
  }
  return SR_OK;
}
}  // namespace

namespace {

// One EM call after its arguments are staged and its buffers sized: only
// kernel launches on the plan's stream, so that the sequence can be captured
// into a HIP graph and replayed.
struct EmCall {
  float* out;
  const float *x0, *y, *b;
  const int* vids;                 // subset view lists, block after block
  std::vector<int> cnt, off;       // views per block, offset of each block in vids
  int nblocks, niter;
  bool fused, shared;
};

// where a fused EM call with one PSF stack for all views sums its subset
// normalisers: all in one launch before the iterations, or each in its
// subset's first-iteration gather
bool em_norm_upfront(sr_plan p, const EmCall& c) { return c.shared && c.niter > 0 && p->em_norm_upfront(); }
bool em_norm_in_step(sr_plan p, const EmCall& c, int it) { return c.shared && it == 0 && !p->em_norm_upfront(); }

// The EM kernel sequence of a z-fastest plan whose subsets fit one launch
// batch (c.fused) in three parts, so that the subset steps of many plans can
// also run as shared launches (sr_osem_batch): the normalisers and the
// image's z-fastest copy; one subset step; the result in the caller's layout.
int em_fused_begin(sr_plan p, const EmCall& c) {
  int rc;
  // (one PSF stack: H and H .* atten[v] are plan state, built by em_setup)
  if (!c.shared) {
    HIP_TRY(sr::launch_fill(p->ones.as<float>(), 1.f, p->P * c.cnt[0], p->stream));
    for (int nb = 0; nb < c.nblocks; ++nb) {
      float* as = p->em_asum.as<float>() + p->V * nb;
      if ((rc = run_backproject(p, p->em_back.as<float>(), p->ones.as<float>(), c.vids + c.off[nb], c.cnt[nb],
                                sr::IMG_WRITE, nullptr)))
        return rc;
      HIP_TRY(sr::launch_to_zfast(p->em_back.as<float>(), as, p->g, p->stream));
      HIP_TRY(sr::launch_zero_to_inf(as, p->V, p->stream));
    }
  }
  {
    KTimer kt(p, sr::K_ROT_FWD);
    HIP_TRY(sr::launch_to_zfast(c.x0, p->zimg.as<float>(), p->g, p->stream));
  }
  // every subset's normaliser from one gather launch (grid row k: block k,
  // the lists laid out block after block, the first `rem` one view longer)
  if (em_norm_upfront(p, c)) {
    const int base = c.cnt[c.nblocks - 1];
    int rem = 0;
    while (rem < c.nblocks && c.cnt[rem] == base + 1) ++rem;
    for (int k = 0; k < c.nblocks; ++k)
      if (c.cnt[k] != base + (k < rem ? 1 : 0) || c.off[k] != k * base + std::min(k, rem))
        return fail(SR_ERR_ARG, "internal: subset lists are not strided blocks");
    KTimer kt(p, sr::K_ROT_ADJ);
    HIP_TRY(sr::launch_rot2_adj_z_norms(p->zh.as<float>(), p->em_asum.as<float>(), p->zent.as<int2>(), p->zk,
                                        c.vids, base, rem, c.nblocks, p->g, p->stream, p->atten.as<float>(),
                                        p->atten_vstride(), p->ell_vs()));
  }
  return SR_OK;
}

// subset nb of iteration it: project (ratio y ./ (A x + b)), PSF adjoint,
// and the gather that applies the update to the z-fastest image in place
int em_fused_step(sr_plan p, const EmCall& c, int it, int nb) {
  const int* sv = c.vids + c.off[nb];
  const int ns = c.cnt[nb];
  int rc;
  if ((rc = run_project(p, p->em_ratio.as<float>(), nullptr, sv, ns, c.y, c.b, true))) return rc;
  {
    KTimer kt(p, sr::K_PSF_ADJ);
    HIP_TRY(sr::launch_psf_adj(p->separable, p->em_ratio.as<float>(), p->taps(), sv, ns, p->work.as<float>(),
                               p->atten.as<float>(), p->atten_vstride(), p->nchunk_adj(ns), p->g, p->stream, true,
                               p->fam_adj()));
  }
  KTimer kt(p, sr::K_ROT_ADJ);
  float* as = p->em_asum.as<float>() + p->V * nb;
  if (em_norm_in_step(p, c, it))  // this subset's normaliser, summed in the same pass
    HIP_TRY(sr::launch_rot2_adj_z_em_norm(p->work.as<float>(), p->zimg.as<float>(), p->zent.as<int2>(), p->zk, sv,
                                          ns, p->g, p->stream, nullptr, p->atten_h.as<float>(),
                                          p->atten_vstride(), as, p->ell_vs()));
  else
    HIP_TRY(sr::launch_rot2_adj_z(p->work.as<float>(), p->zimg.as<float>(), p->zent.as<int2>(), p->zk, sv, ns,
                                  p->g, p->stream, nullptr, 0, as, false, p->ell_vs()));
  return SR_OK;
}

int em_fused_end(sr_plan p, const EmCall& c) {
  KTimer kt(p, sr::K_IMG_REDUCE);
  HIP_TRY(sr::launch_img_reduce_z(p->zimg.as<float>(), 1, c.out, nullptr, sr::IMG_WRITE, p->g, p->stream));
  return SR_OK;
}

int em_body(sr_plan p, const EmCall& c) {
  int rc;
  if (c.fused) {
    if ((rc = em_fused_begin(p, c))) return rc;
    for (int it = 0; it < c.niter; ++it)
      for (int nb = 0; nb < c.nblocks; ++nb)
        if ((rc = em_fused_step(p, c, it, nb))) return rc;
    return em_fused_end(p, c);
  }
  const int nsub = c.cnt[0];
  HIP_TRY(sr::launch_fill(p->ones.as<float>(), 1.f, p->P * nsub, p->stream));
  if (c.shared) {
    const sr::Taps t = p->taps();
    KTimer kt(p, sr::K_PSF_ADJ);
    HIP_TRY(sr::launch_psf_adj(p->separable, p->ones.as<float>(), t, c.vids, 1, p->zh.as<float>(), nullptr, 0,
                               p->nchunk_adj(1), p->g, p->stream, true, p->fam_adj()));
    for (int nb = 0; nb < c.nblocks; ++nb) {
      float* as = p->em_asum.as<float>() + p->V * nb;
      {
        KTimer kt(p, sr::K_ROT_ADJ);
        HIP_TRY(sr::launch_rot2_adj_z(p->zh.as<float>(), p->ipart.as<float>(), p->zent.as<int2>(), p->zk,
                                      c.vids + c.off[nb], c.cnt[nb], p->g, p->stream, p->atten.as<float>(),
                                      p->atten_vstride(), nullptr, false, p->ell_vs()));
      }
      {
        KTimer kt(p, sr::K_IMG_REDUCE);
        HIP_TRY(sr::launch_img_reduce_z(p->ipart.as<float>(), 1, as, nullptr, sr::IMG_WRITE, p->g, p->stream));
      }
      HIP_TRY(sr::launch_zero_to_inf(as, p->V, p->stream));
    }
  } else {
    for (int nb = 0; nb < c.nblocks; ++nb) {
      float* as = p->em_asum.as<float>() + p->V * nb;
      if ((rc = run_backproject(p, as, p->ones.as<float>(), c.vids + c.off[nb], c.cnt[nb], sr::IMG_WRITE, nullptr)))
        return rc;
      HIP_TRY(sr::launch_zero_to_inf(as, p->V, p->stream));
    }
  }
  // out = x0 (a copy kernel: the sequence holds kernels only)
  HIP_TRY(sr::launch_img_reduce(c.x0, 1, c.out, nullptr, p->V, sr::IMG_WRITE, p->stream));
  bool zready = false;  // the EM update also refreshed the z-fastest image copy
  for (int it = 0; it < c.niter; ++it)
    for (int nb = 0; nb < c.nblocks; ++nb) {
      const int* sv = c.vids + c.off[nb];
      if ((rc = run_project(p, p->em_ratio.as<float>(), c.out, sv, c.cnt[nb], c.y, c.b, zready))) return rc;
      if ((rc = run_backproject(p, c.out, p->em_ratio.as<float>(), sv, c.cnt[nb], sr::IMG_EM,
                                p->em_asum.as<float>() + p->V * nb, &zready)))
        return rc;
    }
  return SR_OK;
}

constexpr size_t kMaxGraphs = 8;  // instantiated EM graphs kept per plan

// em_body through a HIP graph: the first call with a given key runs eagerly
// (kernel attributes are set on first launch), the second captures and
// replays, later ones replay.  A failed capture runs the sequence eagerly.
int em_graph(sr_plan p, const EmCall& c) {
  std::vector<uintptr_t> key = {
      (uintptr_t)c.out, (uintptr_t)c.x0, (uintptr_t)c.y, (uintptr_t)c.b, (uintptr_t)c.vids,
      (uintptr_t)c.nblocks, (uintptr_t)c.niter, (uintptr_t)p->max_batch, (uintptr_t)p->work.p,
      (uintptr_t)p->fpart.p, (uintptr_t)p->ipart.p, (uintptr_t)p->ones.p, (uintptr_t)p->em_asum.p,
      (uintptr_t)p->em_ratio.p, (uintptr_t)p->em_back.p, (uintptr_t)p->zimg.p, (uintptr_t)p->zh.p,
      (uintptr_t)p->atten_h.p};
  for (auto& e : p->graphs)
    if (e.key == key) {
      HIP_TRY(hipGraphLaunch(e.exec, p->stream));
      ++p->graph_replays;
      return SR_OK;
    }
  auto seen = std::find(p->graph_seen.begin(), p->graph_seen.end(), key);
  if (seen == p->graph_seen.end()) {
    p->graph_seen.push_back(key);
    if (p->graph_seen.size() > kMaxGraphs) p->graph_seen.erase(p->graph_seen.begin());
    return em_body(p, c);
  }
  p->graph_seen.erase(seen);
  HIP_TRY(hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal));
  const int rc = em_body(p, c);
  hipGraph_t g = nullptr;
  const hipError_t ec = hipStreamEndCapture(p->stream, &g);
  hipGraphExec_t x = nullptr;
  const hipError_t ei = (rc == SR_OK && ec == hipSuccess && g) ? hipGraphInstantiate(&x, g, nullptr, nullptr, 0)
                                                               : hipErrorUnknown;
  if (g) (void)hipGraphDestroy(g);
  if (ei != hipSuccess) {
    (void)hipGetLastError();
    return em_body(p, c);
  }
  if (p->graphs.size() >= kMaxGraphs) {
    (void)hipGraphExecDestroy(p->graphs.front().exec);
    p->graphs.erase(p->graphs.begin());
  }
  p->graphs.push_back({std::move(key), x});
  ++p->graph_captures;
  HIP_TRY(hipGraphLaunch(x, p->stream));
  return SR_OK;
}

// shared body of mlem!/osem!: blocks of views nb:nblocks:nview
// (src/ml-os-em.jl:21-49 with nblocks == 1; :97-131 otherwise)
// H = the PSF adjoint of a ones view (one PSF stack for all views: the same
// for every view) into zh, and H .* atten[v] for every view into atten_h,
// rounded as the two-sum gather rounds them: the sources of the OS-EM
// normalisers A_b' 1 = sum over b of rot2^T(H .* atten[v]).  Plan state,
// rebuilt when the attenuation factors or the PSF adjoint family change.
// vids: any device view id (the taps are those of every view).
int ensure_atten_h(sr_plan p, const int* vids) {
  if (p->atten_h_fam == p->fam_adj()) return SR_OK;
  int rc;
  if ((rc = p->zh.ensure(sizeof(float) * p->V)) || (rc = p->ones.ensure(sizeof(float) * p->P)) ||
      (rc = p->atten_h.ensure(sizeof(float) * p->V * p->d.nview)))
    return rc;
  HIP_TRY(sr::launch_fill(p->ones.as<float>(), 1.f, p->P, p->stream));
  {
    KTimer kt(p, sr::K_PSF_ADJ);
    HIP_TRY(sr::launch_psf_adj(p->separable, p->ones.as<float>(), p->taps(), vids, 1, p->zh.as<float>(), nullptr, 0,
                               p->nchunk_adj(1), p->g, p->stream, true, p->fam_adj()));
  }
  HIP_TRY(sr::launch_mul_volumes(p->zh.as<float>(), p->atten.as<float>(), p->atten_h.as<float>(), p->V, p->d.nview,
                                 p->stream));
  p->atten_h_fam = p->fam_adj();
  return SR_OK;
}

// checks, host staging, the subset view lists and every buffer of one EM
// call, leaving *c ready for its kernel sequence
int em_setup(sr_plan p, EmCall& c, float* out, const float* x0, const float* y, const float* b, int nblocks,
             int niter, int32_t mem) {
  const int nview = p->d.nview;
  const long M = p->P * nview;
  const bool host = (mem & 0xff) == SR_MEM_HOST;
  int rc;
  // background > 0 (src/ml-os-em.jl:30, :106), checked before any update:
  // on the host for host arrays, else by a device scan that returns one flag
  // (which waits for the stream: SR_NO_CHECK, for callers that checked b,
  // leaves an SR_ASYNC call free of host waits)
  if (host) {
    if ((rc = check_background(b, (size_t)M))) return rc;
  } else if (!(mem & SR_NO_CHECK)) {
    if ((rc = p->flag.ensure(sizeof(int)))) return rc;
    HIP_TRY(sr::launch_flag_nonpositive(b, M, p->flag.as<int>(), p->stream));
    int bad = 0;
    HIP_TRY(hipMemcpyAsync(&bad, p->flag.p, sizeof(int), hipMemcpyDeviceToHost, p->stream));
    HIP_TRY(hipStreamSynchronize(p->stream));
    if (bad) return fail(SR_ERR_ARG, "need background > 0");
  }
  c.x0 = x0;
  c.y = y;
  c.b = b;
  c.out = out;
  if (host) {
    if ((rc = stage_in(p, p->st_x0, x0, sizeof(float) * p->V)) || (rc = stage_in(p, p->st_aux1, y, sizeof(float) * M)) ||
        (rc = stage_in(p, p->st_aux2, b, sizeof(float) * M)) || (rc = p->st_img.ensure(sizeof(float) * p->V)))
      return rc;
    c.x0 = p->st_x0.as<float>();
    c.y = p->st_aux1.as<float>();
    c.b = p->st_aux2.as<float>();
    c.out = p->st_img.as<float>();
  }
  // subset view lists, uploaded once (from a plan member: the copy may still
  // read it after an SR_ASYNC call returns)
  std::vector<int>& order = p->h_order;
  order.clear();
  for (int nb = 0; nb < nblocks; ++nb)
    for (int v = nb; v < nview; v += nblocks) order.push_back(v);
  // block nb: views nb, nb + nblocks, ... (src/ml-os-em.jl:76); counts
  // differ by at most one when nblocks does not divide nview
  c.cnt.resize(nblocks);
  c.off.resize(nblocks);
  for (int nb = 0, o = 0; nb < nblocks; ++nb) {
    c.cnt[nb] = (nview - nb + nblocks - 1) / nblocks;
    c.off[nb] = o;
    o += c.cnt[nb];
  }
  c.nblocks = nblocks;
  c.niter = niter;
  const int nsub = c.cnt[0];  // the largest block
  if ((rc = p->vids.ensure(sizeof(int) * nview))) return rc;
  HIP_TRY(hipMemcpyAsync(p->vids.p, order.data(), sizeof(int) * nview, hipMemcpyHostToDevice, p->stream));
  c.vids = p->vids.as<int>();
  // z-fastest plans whose subsets fit one launch batch run the whole loop on
  // the z-fastest image copy: the rotation-adjoint gather applies the EM
  // update in place, the normalisers are kept z-fastest too, and the caller's
  // layout is written once at the end
  c.fused = p->zfast && nsub <= p->max_batch;
  // one PSF for all views: the PSF adjoint of a ones view is the same volume
  // H for every view, so A_v' 1 = rot2^T(H .* atten[v]); H is computed once.
  // On the fused path each normaliser is then summed by the first
  // iteration's gather of its subset, in the same pass as the back-projection
  c.shared = p->zfast && p->vstride == 0 && p->separable;
  // every buffer the kernel sequence touches is sized here, before it runs
  // (asum per block = A_b' 1 with zeros -> Inf)
  if ((rc = p->ones.ensure(sizeof(float) * p->P * nsub)) || (rc = p->em_asum.ensure(sizeof(float) * p->V * nblocks)) ||
      (rc = p->em_ratio.ensure(sizeof(float) * p->P * nsub)) ||
      (rc = ensure_batch_buffers(p, std::min(nsub, p->max_batch))) ||
      (c.shared && (rc = p->zh.ensure(sizeof(float) * p->V))) ||
      (((c.fused && !c.shared) || (!c.fused && nsub > p->max_batch)) &&
       (rc = p->em_back.ensure(sizeof(float) * p->V))))
    return rc;
  if (c.fused && c.shared && (rc = ensure_atten_h(p, c.vids))) return rc;
  return SR_OK;
}

int run_em(sr_plan p, float* out, const float* x0, const float* y, const float* b, int nblocks, int niter,
           int32_t mem) {
  const bool host = (mem & 0xff) == SR_MEM_HOST;
  EmCall c;
  int rc;
  if ((rc = em_setup(p, c, out, x0, y, b, nblocks, niter, mem))) return rc;
  if ((rc = (p->graph && !p->timing) ? em_graph(p, c) : em_body(p, c))) return rc;
  if (host) {
    HIP_TRY(hipMemcpyAsync(out, c.out, sizeof(float) * p->V, hipMemcpyDeviceToHost, p->stream));
    HIP_TRY(hipStreamSynchronize(p->stream));
    resolve_timing(p);
    return SR_OK;
  }
  return finish(p, mem);
}

// ---------------------------------------------------------------------------
// Shared launches across plans (sr_osem_batch)
// ---------------------------------------------------------------------------
// The EM steps of many instances, each on its own plan, as one launch
// sequence on a batch stream: at global step s every instance with more
// than s steps runs subset s % nblocks of iteration s / nblocks, and the
// forwards, view reductions, adjoints and gathers of the step that use one
// kernel instantiation run as one grouped launch (one grid over all their
// blocks; descriptors in device memory).  Each block computes what it
// computes in the instance's own launch, so results equal the instance's
// own call bit for bit.

extern "C++" {  // (templates and C++ return types inside the C-ABI file)
// a plan's own launches (the begin and end parts of its EM call) queued on
// the batch stream: the plan's stream is swapped for the guard's lifetime
struct StreamSwap {
  sr_plan p;
  hipStream_t old;
  StreamSwap(sr_plan q, hipStream_t s) : p(q), old(q->stream) { p->stream = s; }
  ~StreamSwap() { p->stream = old; }
};

enum { G_FWD = 0, G_REDUCE = 1, G_ADJ = 2, G_ELL = 3 };
// one grouped launch: class, kernel key, byte offset of its descriptors in
// the batch's descriptor image, their count, grid blocks, dynamic LDS
struct GLaunch {
  int cls, key;
  size_t off;
  int n, blocks;
  size_t lds;
};
// the instances of one lane (the batch's launches run in lanes, one stream
// each: instances whose kernels share instantiations share a lane), their
// grouped launches, descriptors in the batch's descriptor image
struct Lane {
  std::vector<sr_plan> ps;
  std::vector<const EmCall*> cs;
  std::vector<GLaunch> seq;
};

template <class D>
void put_group(std::vector<unsigned char>& desc, std::vector<GLaunch>& seq, int cls, int key, std::vector<D>& ds,
               size_t lds) {
  if (ds.empty()) return;
  int blk = 0;
  for (auto& d : ds) {
    d.h.blk0 = blk;
    blk += d.h.nblk;
  }
  const size_t off = (desc.size() + 255) & ~(size_t)255;
  desc.resize(off + sizeof(D) * ds.size());
  std::memcpy(desc.data() + off, ds.data(), sizeof(D) * ds.size());
  seq.push_back({cls, key, off, (int)ds.size(), blk, lds});
}

// the four descriptors of instance (p, c)'s step (it, nb); false when one of
// its launches has no grouped form
struct StepDesc {
  sr::PsfFwdGrp f;
  sr::ViewReduceGrp r;
  sr::PsfAdjGrp a;
  sr::EllGrp e;
  int kf, kr, ka, ke;
  size_t lf, la;
};
bool step_desc(sr_plan p, const EmCall& c, int it, int nb, StepDesc& d) {
  const int* sv = c.vids + c.off[nb];
  const int ns = c.cnt[nb];
  const sr::Taps t = p->taps();
  int nchunk = 0;
  d.lf = d.la = 0;
  d.kf = sr::psf_fwd_grp_desc(p->zimg.as<float>(), t, sv, ns, p->fpart.as<float>(), p->nchunk_fwd(ns), &nchunk, p->g,
                              p->atten.as<float>(), p->atten_vstride(), p->rotv.as<sr::RotView>(), p->fam_fwd(), &d.f,
                              &d.lf);
  d.kr = sr::view_reduce_grp_desc(p->fpart.as<float>(), nchunk, sv, ns, p->em_ratio.as<float>(), c.y, c.b,
                                  sr::EPI_RATIO, p->g, &d.r);
  d.ka = sr::psf_adj_grp_desc(p->em_ratio.as<float>(), t, sv, ns, p->work.as<float>(), p->atten.as<float>(),
                              p->atten_vstride(), p->nchunk_adj(ns), p->g, p->fam_adj(), &d.a, &d.la);
  float* as = p->em_asum.as<float>() + p->V * nb;
  d.ke = em_norm_in_step(p, c, it)
             ? sr::ell_grp_desc(3, p->work.as<float>(), p->zimg.as<float>(), p->zent.as<int2>(), p->zk, sv, ns, p->g,
                                nullptr, p->atten_h.as<float>(), p->atten_vstride(), as, p->ell_vs(), &d.e)
             : sr::ell_grp_desc(0, p->work.as<float>(), p->zimg.as<float>(), p->zent.as<int2>(), p->zk, sv, ns, p->g,
                                nullptr, nullptr, 0, as, p->ell_vs(), &d.e);
  return d.kf && d.kr && d.ka && d.ke;
}

// whether an instance's EM call can join the shared launches: the fused
// z-fastest loop with one PSF stack and the kernels that have grouped forms
bool batch_eligible(sr_plan p, const EmCall& c) {
  if (!(c.fused && c.shared && p->zfast && p->separable && p->mfma) || p->timing) return false;
  StepDesc d;
  for (int it = 0; it < std::min(c.niter, 2); ++it)
    for (int nb = 0; nb < c.nblocks; ++nb)
      if (!step_desc(p, c, it, nb, d)) return false;
  return true;
}


// plane chunks per view for a grouped PSF launch over views of planes of up
// to ny_max planes (one block per CU): the cheapest c of ceil(c nviews / ncu)
// rounds of (planes per chunk + 2) plane times, fewest on ties -- the plan's
// own rule (chunks_base) applied to the whole group, so that instances that
// share a launch take fewer, longer chunks than each would alone
int group_chunks(long nviews, int ny_max, int ncu) {
  int best = 1;
  long best_cost = -1;
  for (int c = 1; c <= ny_max; ++c) {
    const int yc = (ny_max + c - 1) / c;
    if ((ny_max + yc - 1) / yc != c) continue;
    const long cost = ((long)c * std::max(1L, nviews) + ncu - 1) / ncu * (yc + 2);
    if (best_cost < 0 || cost < best_cost) {
      best = c;
      best_cost = cost;
    }
  }
  return best;
}

int batch_build(Lane& ln, std::vector<unsigned char>& desc) {
  const std::vector<sr_plan>& ps = ln.ps;
  const std::vector<const EmCall*>& cs = ln.cs;
  int steps = 0;
  for (const EmCall* c : cs) steps = std::max(steps, c->niter * c->nblocks);
  struct Member {
    size_t i;
    int it, nb, ns;
    const int* sv;
    StepDesc d;
    int nchunk;  // the forward's chunks, for the view reduction
  };
  // groups: kernel key -> members (in instance order)
  using Groups = std::vector<std::pair<int, std::vector<Member*>>>;
  auto slot = [](Groups& g, int key) -> std::vector<Member*>& {
    for (auto& e : g)
      if (e.first == key) return e.second;
    g.push_back({key, {}});
    return g.back().second;
  };
  for (int st = 0; st < steps; ++st) {
    std::vector<Member> mem;
    mem.reserve(ps.size());
    for (size_t i = 0; i < ps.size(); ++i) {
      const EmCall& c = *cs[i];
      if (st >= c.niter * c.nblocks) continue;
      Member m{i, st / c.nblocks, st % c.nblocks, 0, nullptr, {}, 0};
      m.sv = c.vids + c.off[m.nb];
      m.ns = c.cnt[m.nb];
      if (!step_desc(ps[i], c, m.it, m.nb, m.d))
        return fail(SR_ERR_UNSUPPORTED, "internal: instance %zu has no shared-launch form", i);
      mem.push_back(m);
    }
    Groups gf, ga, ge;
    for (Member& m : mem) {
      slot(gf, m.d.kf).push_back(&m);
      slot(ga, m.d.ka).push_back(&m);
      // gathers grouped without their view-slice count, which the group sets
      // (unless a plan pins it)
      const int vs = ps[m.i]->ell_vs();
      slot(ge, vs ? m.d.ke : (m.d.ke & ~(15 << 8))).push_back(&m);
    }
    // PSF forward: the group's chunk count
    for (auto& [key, ms] : gf) {
      long nv = 0;
      int ny = 1;
      for (Member* m : ms) nv += m->ns, ny = std::max(ny, ps[m->i]->d.ny);
      const int cg = group_chunks(nv, ny, ps[ms[0]->i]->ncu);
      std::vector<sr::PsfFwdGrp> ds;
      size_t lds = 0;
      for (Member* m : ms) {
        const sr_plan p = ps[m->i];
        sr::PsfFwdGrp d;
        size_t l = 0;
        const int k = sr::psf_fwd_grp_desc(p->zimg.as<float>(), p->taps(), m->sv, m->ns, p->fpart.as<float>(),
                                           std::min(cg, p->d.ny), &m->nchunk, p->g, p->atten.as<float>(),
                                           p->atten_vstride(), p->rotv.as<sr::RotView>(), p->fam_fwd(), &d, &l);
        if (k != key) return fail(SR_ERR_UNSUPPORTED, "internal: forward key changed with the chunk count");
        ds.push_back(d);
        lds = std::max(lds, l);
      }
      put_group(desc, ln.seq, G_FWD, key, ds, lds);
    }
    // view reductions over each forward's chunks, with the EM ratio
    {
      std::vector<sr::ViewReduceGrp> ds;
      for (Member& m : mem) {
        const sr_plan p = ps[m.i];
        const EmCall& c = *cs[m.i];
        sr::ViewReduceGrp d;
        sr::view_reduce_grp_desc(p->fpart.as<float>(), m.nchunk, m.sv, m.ns, p->em_ratio.as<float>(), c.y, c.b,
                                 sr::EPI_RATIO, p->g, &d);
        ds.push_back(d);
      }
      put_group(desc, ln.seq, G_REDUCE, 1, ds, 0);
    }
    // PSF adjoint: the group's chunk count (planes are independent here: any
    // count gives the same result)
    for (auto& [key, ms] : ga) {
      long nv = 0;
      int ny = 1;
      for (Member* m : ms) nv += m->ns, ny = std::max(ny, ps[m->i]->d.ny);
      const int cg = group_chunks(nv, ny, ps[ms[0]->i]->ncu);
      std::vector<sr::PsfAdjGrp> ds;
      size_t lds = 0;
      for (Member* m : ms) {
        const sr_plan p = ps[m->i];
        sr::PsfAdjGrp d;
        size_t l = 0;
        const int k = sr::psf_adj_grp_desc(p->em_ratio.as<float>(), p->taps(), m->sv, m->ns, p->work.as<float>(),
                                           p->atten.as<float>(), p->atten_vstride(), std::min(cg, p->d.ny), p->g,
                                           p->fam_adj(), &d, &l);
        if (k != key) return fail(SR_ERR_UNSUPPORTED, "internal: adjoint key changed with the chunk count");
        ds.push_back(d);
        lds = std::max(lds, l);
      }
      put_group(desc, ln.seq, G_ADJ, key, ds, lds);
    }
    // rotation-adjoint gathers: view slices from the group's wave count (the
    // fewest of 2, 4, 8 giving 4096 waves, at most the shortest view list)
    for (auto& [gkey, ms] : ge) {
      int vs = (gkey >> 8) & 15;
      if (!vs) {
        long blocks = 0;
        int nmin = 1 << 30;
        for (Member* m : ms) blocks += m->d.e.h.nblk, nmin = std::min(nmin, m->ns);
        vs = 2;
        while (vs < 8 && 2 * vs <= nmin && blocks * vs < 4096) vs *= 2;
      }
      std::vector<sr::EllGrp> ds;
      int key = 0;
      for (Member* m : ms) {
        const sr_plan p = ps[m->i];
        const EmCall& c = *cs[m->i];
        float* as = p->em_asum.as<float>() + p->V * m->nb;
        sr::EllGrp d;
        const int k =
            em_norm_in_step(p, c, m->it)
                ? sr::ell_grp_desc(3, p->work.as<float>(), p->zimg.as<float>(), p->zent.as<int2>(), p->zk, m->sv, m->ns,
                                   p->g, nullptr, p->atten_h.as<float>(), p->atten_vstride(), as, vs, &d)
                : sr::ell_grp_desc(0, p->work.as<float>(), p->zimg.as<float>(), p->zent.as<int2>(), p->zk, m->sv, m->ns,
                                   p->g, nullptr, nullptr, 0, as, vs, &d);
        if (key && k != key) return fail(SR_ERR_UNSUPPORTED, "internal: gather keys differ within a group");
        key = k;
        ds.push_back(d);
      }
      put_group(desc, ln.seq, G_ELL, key, ds, 0);
    }
  }
  return SR_OK;
}

// the whole batch sequence on stream s: every instance's begin part, the
// grouped steps (descriptors at ddesc), every instance's end part
int batch_run(const Lane& ln, const unsigned char* ddesc, hipStream_t s) {
  const std::vector<sr_plan>& ps = ln.ps;
  const std::vector<const EmCall*>& cs = ln.cs;
  int rc;
  for (size_t i = 0; i < ps.size(); ++i) {
    StreamSwap sw(ps[i], s);
    if ((rc = em_fused_begin(ps[i], *cs[i]))) return rc;
  }
  for (const GLaunch& g : ln.seq) {
    const void* d = ddesc + g.off;
    switch (g.cls) {
      case G_FWD:
        HIP_TRY(sr::launch_psf_fwd_grp(g.key, static_cast<const sr::PsfFwdGrp*>(d), g.n, g.blocks, g.lds, s));
        break;
      case G_REDUCE:
        HIP_TRY(sr::launch_view_reduce_grp(static_cast<const sr::ViewReduceGrp*>(d), g.n, g.blocks, s));
        break;
      case G_ADJ:
        HIP_TRY(sr::launch_psf_adj_grp(g.key, static_cast<const sr::PsfAdjGrp*>(d), g.n, g.blocks, g.lds, s));
        break;
      default:
        HIP_TRY(sr::launch_ell_grp(g.key, static_cast<const sr::EllGrp*>(d), g.n, g.blocks, s));
    }
  }
  for (size_t i = 0; i < ps.size(); ++i) {
    StreamSwap sw(ps[i], s);
    if ((rc = em_fused_end(ps[i], *cs[i]))) return rc;
  }
  return SR_OK;
}

// per device: the batch stream, the descriptor buffer of eager runs, and
// graphs of whole batch sequences keyed by everything they bake in (the
// descriptor image and every instance's begin/end arguments); the first
// call with a key runs eagerly, the second captures, later ones replay
constexpr int kLanes = 8;
// lanes a batch uses: HIP's default of 4 hardware queues per process, so
// that every lane's stream has a queue of its own
constexpr int kLaneQueues = 4;
struct BatchCtx {
  std::mutex mu;
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr, fork = nullptr;
  hipStream_t lane[kLanes] = {};
  hipEvent_t join[kLanes] = {};
  DBuf desc;
  struct Graph {
    std::vector<uintptr_t> key;
    std::unique_ptr<DBuf> desc;
    hipGraphExec_t exec = nullptr;
  };
  std::vector<Graph> graphs;
  std::vector<std::vector<uintptr_t>> seen;
  long instances = 0, captures = 0, replays = 0, launches = 0, lanes_last = 0;
};
constexpr size_t kMaxBatchGraphs = 4;

BatchCtx* batch_ctx() {
  static BatchCtx ctx[64];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) {
    (void)hipGetLastError();
    dev = 0;
  }
  return &ctx[dev];
}

std::vector<uintptr_t> batch_key(const std::vector<sr_plan>& ps, const std::vector<const EmCall*>& cs) {
  // every device address, count and choice the sequence bakes in (the
  // descriptors are derived from these)
  std::vector<uintptr_t> key;
  key.push_back(ps.size());
  for (size_t i = 0; i < ps.size(); ++i) {
    const sr_plan p = ps[i];
    const EmCall& c = *cs[i];
    for (uintptr_t v :
         {(uintptr_t)p, (uintptr_t)c.out, (uintptr_t)c.x0, (uintptr_t)c.y, (uintptr_t)c.b, (uintptr_t)c.vids,
          (uintptr_t)c.nblocks, (uintptr_t)c.niter, (uintptr_t)p->max_batch, (uintptr_t)p->work.p,
          (uintptr_t)p->fpart.p, (uintptr_t)p->ipart.p, (uintptr_t)p->ones.p, (uintptr_t)p->em_asum.p,
          (uintptr_t)p->em_ratio.p, (uintptr_t)p->em_back.p, (uintptr_t)p->zimg.p, (uintptr_t)p->zh.p,
          (uintptr_t)p->zent.p, (uintptr_t)p->zk, (uintptr_t)p->atten.p, (uintptr_t)p->atten_h.p, (uintptr_t)p->rotv.p,
          (uintptr_t)p->taps_u.p, (uintptr_t)p->taps_w.p, (uintptr_t)p->taps_c.p})
      key.push_back(v);
    for (int k = 0; k < SR_NCHOICE; ++k) key.push_back((uintptr_t)(uint32_t)p->choice[k]);
  }
  return key;
}

// runs the grouped instances on the batch stream (graph or eager) and makes
// every grouped plan's stream wait for it
// runs the grouped instances (graph or eager): lanes fork from the batch
// stream, each on a stream of its own, and join it again; every grouped
// plan's stream then waits for the batch stream
int batch_exec(BatchCtx& bx, const std::vector<sr_plan>& ps, const std::vector<const EmCall*>& cs, bool graph) {
  if (!bx.stream) {
    HIP_TRY(hipStreamCreateWithFlags(&bx.stream, hipStreamNonBlocking));
    HIP_TRY(hipEventCreateWithFlags(&bx.done, hipEventDisableTiming));
    HIP_TRY(hipEventCreateWithFlags(&bx.fork, hipEventDisableTiming));
    for (int l = 0; l < kLanes; ++l) {
      HIP_TRY(hipStreamCreateWithFlags(&bx.lane[l], hipStreamNonBlocking));
      HIP_TRY(hipEventCreateWithFlags(&bx.join[l], hipEventDisableTiming));
    }
  }
  // lanes: the instances ordered by z tiles per column (which fix their PSF
  // kernels' instantiations), then split into up to kLaneQueues contiguous
  // runs of about equal work (voxels x views x iterations), one stream each
  std::vector<size_t> order(ps.size());
  for (size_t i = 0; i < order.size(); ++i) order[i] = i;
  auto work_of = [&](size_t i) { return (double)ps[i]->V * ps[i]->d.nview * std::max(1, cs[i]->niter); };
  std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) {
    const int ta = ps[a]->g.nz / 16, tb = ps[b]->g.nz / 16;
    return ta != tb ? ta < tb : work_of(a) > work_of(b);
  });
  double total = 0;
  for (size_t i : order) total += work_of(i);
  const int nl = (int)std::min<size_t>(kLaneQueues, ps.size());
  std::vector<Lane> lanes(nl);
  double acc = 0;
  for (size_t i : order) {
    const int l = std::min(nl - 1, (int)(acc / total * nl));
    acc += work_of(i);
    lanes[l].ps.push_back(ps[i]);
    lanes[l].cs.push_back(cs[i]);
  }
  lanes.erase(std::remove_if(lanes.begin(), lanes.end(), [](const Lane& ln) { return ln.ps.empty(); }), lanes.end());
  std::vector<unsigned char> desc;
  int rc;
  for (Lane& ln : lanes)
    if ((rc = batch_build(ln, desc))) return rc;
  // the plans' queued work (their setup copies) and any earlier batch land first
  for (sr_plan p : ps) HIP_TRY(hipStreamSynchronize(p->stream));
  HIP_TRY(hipStreamSynchronize(bx.stream));
  auto run = [&](const unsigned char* dd) -> int {
    HIP_TRY(hipEventRecord(bx.fork, bx.stream));
    for (size_t l = 0; l < lanes.size(); ++l) {
      HIP_TRY(hipStreamWaitEvent(bx.lane[l], bx.fork, 0));
      int r;
      if ((r = batch_run(lanes[l], dd, bx.lane[l]))) return r;
      HIP_TRY(hipEventRecord(bx.join[l], bx.lane[l]));
      HIP_TRY(hipStreamWaitEvent(bx.stream, bx.join[l], 0));
    }
    return SR_OK;
  };
  bool ran = false;
  if (graph) {
    std::vector<uintptr_t> key = batch_key(ps, cs);
    for (auto& e : bx.graphs)
      if (e.key == key) {
        HIP_TRY(hipGraphLaunch(e.exec, bx.stream));
        ++bx.replays;
        ran = true;
        break;
      }
    if (!ran) {
      auto seen = std::find(bx.seen.begin(), bx.seen.end(), key);
      if (seen == bx.seen.end()) {
        bx.seen.push_back(key);
        if (bx.seen.size() > kMaxBatchGraphs) bx.seen.erase(bx.seen.begin());
      } else {
        bx.seen.erase(seen);
        auto buf = std::make_unique<DBuf>();
        if ((rc = buf->ensure(std::max<size_t>(desc.size(), 256)))) return rc;
        HIP_TRY(hipMemcpy(buf->p, desc.data(), desc.size(), hipMemcpyHostToDevice));
        HIP_TRY(hipStreamBeginCapture(bx.stream, hipStreamCaptureModeThreadLocal));
        rc = run(buf->as<unsigned char>());
        hipGraph_t g = nullptr;
        const hipError_t ec = hipStreamEndCapture(bx.stream, &g);
        hipGraphExec_t x = nullptr;
        const hipError_t ei = (rc == SR_OK && ec == hipSuccess && g) ? hipGraphInstantiate(&x, g, nullptr, nullptr, 0)
                                                                     : hipErrorUnknown;
        if (g) (void)hipGraphDestroy(g);
        if (ei == hipSuccess) {
          if (bx.graphs.size() >= kMaxBatchGraphs) {
            (void)hipGraphExecDestroy(bx.graphs.front().exec);
            bx.graphs.erase(bx.graphs.begin());
          }
          bx.graphs.push_back({std::move(key), std::move(buf), x});
          ++bx.captures;
          HIP_TRY(hipGraphLaunch(x, bx.stream));
          ran = true;
        } else {
          (void)hipGetLastError();
        }
      }
    }
  }
  if (!ran) {
    if ((rc = bx.desc.ensure(std::max<size_t>(desc.size(), 256)))) return rc;
    HIP_TRY(hipMemcpy(bx.desc.p, desc.data(), desc.size(), hipMemcpyHostToDevice));
    if ((rc = run(bx.desc.as<unsigned char>()))) return rc;
  }
  for (const Lane& ln : lanes) bx.launches += (long)ln.seq.size();
  bx.instances += (long)ps.size();
  bx.lanes_last = (long)lanes.size();
  HIP_TRY(hipEventRecord(bx.done, bx.stream));
  for (sr_plan p : ps) HIP_TRY(hipStreamWaitEvent(p->stream, bx.done, 0));
  return SR_OK;
}
}  // extern "C++"

}  // namespace

int sr_mlem(sr_plan p, float* out, const float* x0, const float* y, const float* b, int32_t niter, int32_t mem) {
  if (!p || !out || !x0 || !y || !b) return fail(SR_ERR_ARG, "null argument");
  if (niter < 0) return fail(SR_ERR_ARG, "niter < 0");
  return run_em(p, out, x0, y, b, 1, niter, mem);
}

int sr_osem(sr_plan p, float* out, const float* x0, const float* y, const float* b, int32_t nblocks, int32_t niter,
            int32_t mem) {
  if (!p || !out || !x0 || !y || !b) return fail(SR_ERR_ARG, "null argument");
  if (nblocks < 1) return fail(SR_ERR_ARG, "nblocks < 1");
  if (p->d.nview % nblocks != 0) return fail(SR_ERR_ARG, "nview must be divisible by nblocks!");
  if (niter < 0) return fail(SR_ERR_ARG, "niter < 0");
  return run_em(p, out, x0, y, b, nblocks, niter, mem);
}

int sr_osem_strided(sr_plan p, float* out, const float* x0, const float* y, const float* b, int32_t nblocks,
                    int32_t niter, int32_t mem) {
  if (!p || !out || !x0 || !y || !b) return fail(SR_ERR_ARG, "null argument");
  if (nblocks < 1 || nblocks > p->d.nview) return fail(SR_ERR_ARG, "nblocks must be in 1..nview");
  if (niter < 0) return fail(SR_ERR_ARG, "niter < 0");
  return run_em(p, out, x0, y, b, nblocks, niter, mem);
}

int sr_plan_memcpy_async(sr_plan p, void* dst, const void* src, size_t bytes, int32_t kind) {
  if (!p || (!dst && bytes) || (!src && bytes)) return fail(SR_ERR_ARG, "null argument");
  if (kind != SR_COPY_H2D && kind != SR_COPY_D2H && kind != SR_COPY_D2D) return fail(SR_ERR_ARG, "bad copy kind");
  if (bytes == 0) return SR_OK;
  const hipMemcpyKind k = kind == SR_COPY_H2D   ? hipMemcpyHostToDevice
                          : kind == SR_COPY_D2H ? hipMemcpyDeviceToHost
                                                : hipMemcpyDeviceToDevice;
  HIP_TRY(hipMemcpyAsync(dst, src, bytes, k, p->stream));
  return SR_OK;
}

int sr_host_alloc(void** ptr, size_t bytes) {
  if (!ptr) return fail(SR_ERR_ARG, "null pointer");
  *ptr = nullptr;
  const hipError_t e = hipHostMalloc(ptr, std::max<size_t>(bytes, 1), hipHostMallocDefault);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    *ptr = nullptr;
    return fail(SR_ERR_ALLOC, "hipHostMalloc of %zu bytes failed: %s", bytes, hipGetErrorString(e));
  }
  return SR_OK;
}

int sr_host_free(void* ptr) {
  if (ptr) HIP_TRY(hipHostFree(ptr));
  return SR_OK;
}

// Many reconstructions, each on its own plan (and stream), from one call:
// all are checked first (the reference's messages; nothing runs when one
// fails), then each one's EM sequence is queued on its plan's stream (a HIP
// graph replay where the plan has graphs on), so the device runs them
// concurrently; returns when all finished, or at once with SR_ASYNC.
int sr_osem_batch(int32_t n, const sr_plan* plans, float* const* outs, const float* const* x0s,
                  const float* const* ys, const float* const* bs, const int32_t* nblocks, const int32_t* niters,
                  int32_t mem) {
  if (n < 0 || (n > 0 && (!plans || !outs || !x0s || !ys || !bs || !nblocks || !niters)))
    return fail(SR_ERR_ARG, "null argument");
  if ((mem & 0xff) != SR_MEM_DEVICE) return fail(SR_ERR_ARG, "sr_osem_batch takes device arrays (SR_MEM_DEVICE)");
  for (int i = 0; i < n; ++i) {
    sr_plan p = plans[i];
    if (!p || !outs[i] || !x0s[i] || !ys[i] || !bs[i]) return fail(SR_ERR_ARG, "null argument (instance %d)", i);
    if (nblocks[i] < 1 || nblocks[i] > p->d.nview) return fail(SR_ERR_ARG, "nblocks must be in 1..nview (instance %d)", i);
    if (niters[i] < 0) return fail(SR_ERR_ARG, "niter < 0 (instance %d)", i);
    for (int k = 0; k < i; ++k)
      if (plans[k] == p) return fail(SR_ERR_ARG, "instances %d and %d share a plan (a plan runs one call at a time)", k, i);
  }
  if (!(mem & SR_NO_CHECK)) {
    // every background scanned on its plan's stream, the flags read once
    for (int i = 0; i < n; ++i) {
      sr_plan p = plans[i];
      int rc;
      if ((rc = p->flag.ensure(sizeof(int)))) return rc;
      HIP_TRY(sr::launch_flag_nonpositive(bs[i], p->P * p->d.nview, p->flag.as<int>(), p->stream));
    }
    for (int i = 0; i < n; ++i) {
      int bad = 0;
      HIP_TRY(hipMemcpyAsync(&bad, plans[i]->flag.p, sizeof(int), hipMemcpyDeviceToHost, plans[i]->stream));
      HIP_TRY(hipStreamSynchronize(plans[i]->stream));
      if (bad) return fail(SR_ERR_ARG, "need background > 0");
    }
  }
  // every instance set up (subset lists, buffers); those whose EM call has
  // grouped forms run as shared launches on the batch stream, the others
  // on their plans' streams, concurrently
  std::vector<EmCall> calls(n);
  std::vector<sr_plan> gp;
  std::vector<const EmCall*> gc;
  bool graph = true;
  for (int i = 0; i < n; ++i) {
    sr_plan p = plans[i];
    int rc;
    if ((rc = em_setup(p, calls[i], outs[i], x0s[i], ys[i], bs[i], nblocks[i], niters[i],
                       SR_MEM_DEVICE | SR_NO_CHECK)))
      return rc;
    if (!(mem & SR_BATCH_STREAMS) && batch_eligible(p, calls[i])) {
      gp.push_back(p);
      gc.push_back(&calls[i]);
      graph = graph && p->graph;
      continue;
    }
    if ((rc = (p->graph && !p->timing) ? em_graph(p, calls[i]) : em_body(p, calls[i]))) return rc;
  }
  if (!gp.empty()) {
    BatchCtx* bx = batch_ctx();
    std::lock_guard<std::mutex> lk(bx->mu);
    int rc;
    if ((rc = batch_exec(*bx, gp, gc, graph))) return rc;
  }
  if (mem & SR_ASYNC) return SR_OK;
  for (int i = 0; i < n; ++i) {
    HIP_TRY(hipStreamSynchronize(plans[i]->stream));
    resolve_timing(plans[i]);
  }
  return SR_OK;
}

int sr_batch_stats(int64_t* instances, int64_t* launches, int64_t* captures, int64_t* replays) {
  BatchCtx* bx = batch_ctx();
  std::lock_guard<std::mutex> lk(bx->mu);
  if (instances) *instances = bx->instances;
  if (launches) *launches = bx->launches;
  if (captures) *captures = bx->captures;
  if (replays) *replays = bx->replays;
  return SR_OK;
}

int sr_plan_query(sr_plan p, int32_t* done) {
  if (!p || !done) return fail(SR_ERR_ARG, "null argument");
  const hipError_t e = hipStreamQuery(p->stream);
  if (e == hipSuccess) {
    *done = 1;
    resolve_timing(p);
    return SR_OK;
  }
  if (e == hipErrorNotReady) {
    (void)hipGetLastError();
    *done = 0;
    return SR_OK;
  }
  return fail(SR_ERR_HIP, "hipStreamQuery: %s", hipGetErrorString(e));
}

// ---------------------------------------------------------------------------
// Unit operators
// ---------------------------------------------------------------------------

}  // extern "C"

namespace {

// Device buffers of the unit operators (sr_imrotate, sr_fft_conv,
// sr_fft_conv3), kept per device and grown as needed, so repeated calls
// allocate nothing after the first of a size; calls are serialised per
// device (the buffers are shared) and run on the null stream.
struct UnitWs {
  std::mutex mu;
  DBuf rv, tb, tw, vids, u, w, c, h, in, out, part;
  int nvids = 0;  // vids holds 0, 1, ..., nvids - 1
  // what the tap and rotation tables hold now (a call with the same kernels
  // and geometry, or the same rotation, uploads nothing)
  std::vector<float> ker;
  int knx = -1, knz = -1, kpx = -1, kpz = -1, kplanes = -1, ksep = 0;
  sr::Taps ktaps{};
  float theta = 0.f;
  int rn = -1, rinterp = -1;
};
UnitWs* unit_ws() {
  static UnitWs ws[64];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) {
    (void)hipGetLastError();
    dev = 0;
  }
  return &ws[dev];
}

// 0, 1, ..., n-1 into the workspace's view ids
int unit_vids(UnitWs* ws, int n) {
  n = std::max(1, n);
  if (n <= ws->nvids) return SR_OK;
  int rc;
  ws->nvids = 0;
  if ((rc = ws->vids.ensure(sizeof(int) * n))) return rc;
  std::vector<int> v(n);
  for (int i = 0; i < n; ++i) v[i] = i;
  HIP_TRY(hipMemcpy(ws->vids.p, v.data(), sizeof(int) * n, hipMemcpyHostToDevice));
  ws->nvids = n;
  return SR_OK;
}

// The tap tables of nplanes PSF planes ker[plane][pz][px] (host) into the
// workspace: separable (u, w and the tail/head sums) when every plane is rank
// 1 and the geometry takes the separable kernels, else the 2-D table.
int unit_taps(UnitWs* ws, const float* ker, int nplanes, int px, int pz, const sr::Geo& g, sr::Taps* t, int* sep) {
  const long pxz = (long)px * pz;
  if (ws->kplanes == nplanes && ws->kpx == px && ws->kpz == pz && ws->knx == g.nx && ws->knz == g.nz &&
      std::memcmp(ws->ker.data(), ker, sizeof(float) * pxz * nplanes) == 0) {
    *t = ws->ktaps;
    *sep = ws->ksep;
    return SR_OK;
  }
  ws->kplanes = -1;  // (valid again only once the new tables are up)
  const int nt = round_taps(std::max(px, pz));
  std::vector<double> u, w;
  bool s = sr::psf_sep_supported(g);
  for (int pl = 0; pl < nplanes && s; ++pl) s = factor_rank1(ker + pl * pxz, px, pz, u, w);
  *t = sr::Taps{};
  t->ntx = t->ntz = nt;
  t->vstride = 1;  // plane of "view" v, row 0: v
  int rc;
  if (s) {
    std::vector<float> hu((size_t)nplanes * nt, 0.f), hw((size_t)nplanes * nt, 0.f);
    for (int pl = 0; pl < nplanes; ++pl) {
      factor_rank1(ker + pl * pxz, px, pz, u, w);
      for (int a = 0; a < px; ++a) hu[(size_t)pl * nt + a] = (float)u[a];
      for (int b = 0; b < pz; ++b) hw[(size_t)pl * nt + b] = (float)w[b];
    }
    const std::vector<float> hc = cum_sums(hu, hw, nplanes, nt);
    if ((rc = ws->u.ensure(sizeof(float) * hu.size())) || (rc = ws->w.ensure(sizeof(float) * hw.size())) ||
        (rc = ws->c.ensure(sizeof(float) * hc.size())))
      return rc;
    HIP_TRY(hipMemcpy(ws->u.p, hu.data(), sizeof(float) * hu.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(ws->w.p, hw.data(), sizeof(float) * hw.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(ws->c.p, hc.data(), sizeof(float) * hc.size(), hipMemcpyHostToDevice));
    t->u = ws->u.as<float>();
    t->w = ws->w.as<float>();
    t->cum = ws->c.as<float>();
  } else {
    std::vector<float> hh((size_t)nplanes * pz * nt, 0.f);
    for (int pl = 0; pl < nplanes; ++pl)
      for (int b = 0; b < pz; ++b)
        for (int a = 0; a < px; ++a) hh[((size_t)pl * pz + b) * nt + a] = ker[pl * pxz + a + (long)px * b];
    if ((rc = ws->h.ensure(sizeof(float) * hh.size()))) return rc;
    HIP_TRY(hipMemcpy(ws->h.p, hh.data(), sizeof(float) * hh.size(), hipMemcpyHostToDevice));
    t->h = ws->h.as<float>();
  }
  *sep = s ? 1 : 0;
  ws->ker.assign(ker, ker + pxz * nplanes);
  ws->knx = g.nx;
  ws->knz = g.nz;
  ws->kpx = px;
  ws->kpz = pz;
  ws->ksep = *sep;
  ws->ktaps = *t;
  ws->kplanes = nplanes;
  return SR_OK;
}

int unit_conv_checks(const float* out, const float* img, const float* ker, int nx, int nz, int px, int pz) {
  if (!out || !img || !ker) return fail(SR_ERR_ARG, "null argument");
  if (nx <= 0 || nz <= 0 || px <= 0 || pz <= 0) return fail(SR_ERR_ARG, "size");
  if (px % 2 == 0 || pz % 2 == 0) return fail(SR_ERR_ARG, "non-odd size psfs");
  if (px > sr::kMaxTaps || pz > sr::kMaxTaps) return fail(SR_ERR_UNSUPPORTED, "kernel larger than %d taps", sr::kMaxTaps);
  if ((long)nx * nz > sr::kMaxPlane) return fail(SR_ERR_UNSUPPORTED, "plane too large");
  return check_device();
}

}  // namespace

extern "C" {

// src/rotatez.jl:240-416 and the 3-D forms :472-495
int sr_imrotate(float* out, const float* img, int32_t n, int32_t nz, float theta, int32_t interpmeth, int32_t adj,
                int32_t mem) {
  if (!out || !img) return fail(SR_ERR_ARG, "null argument");
  if (n <= 0 || nz <= 0) return fail(SR_ERR_ARG, "size");
  if (interpmeth != SR_INTERP_ONE && interpmeth != SR_INTERP_TWO) return fail(SR_ERR_ARG, "bad interpmeth");
  int rc;
  if ((rc = check_device())) return rc;
  const int pad = rot_padsize(n);
  sr::Geo g{n, n, nz, pad, n + 2 * pad, 1, 1};
  if (interpmeth == SR_INTERP_ONE && g.Np > sr::kMaxNpOne)
    return fail(SR_ERR_UNSUPPORTED, ":one rotation supports n up to 140");
  const long V = (long)n * n * nz;
  UnitWs* ws = unit_ws();
  std::lock_guard<std::mutex> lock(ws->mu);
  if ((rc = unit_vids(ws, 1))) return rc;
  if (!(ws->rn == n && ws->rinterp == interpmeth && std::memcmp(&ws->theta, &theta, sizeof theta) == 0)) {
    ws->rn = -1;
    sr::RotView rv;
    std::vector<int> tb(2 * g.Np);
    std::vector<float> tw(4 * g.Np);
    rot_params((double)theta, interpmeth, g.Np, rv, tb.data(), tw.data());
    if ((rc = ws->rv.ensure(sizeof rv)) || (rc = ws->tb.ensure(sizeof(int) * tb.size())) ||
        (rc = ws->tw.ensure(sizeof(float) * tw.size())))
      return rc;
    HIP_TRY(hipMemcpy(ws->rv.p, &rv, sizeof rv, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(ws->tb.p, tb.data(), sizeof(int) * tb.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(ws->tw.p, tw.data(), sizeof(float) * tw.size(), hipMemcpyHostToDevice));
    ws->rn = n;
    ws->rinterp = interpmeth;
    ws->theta = theta;
  }
  const bool host = (mem & 0xff) == SR_MEM_HOST;
  const float* din = img;
  float* dout = out;
  if (host) {
    if ((rc = ws->in.ensure(sizeof(float) * V)) || (rc = ws->out.ensure(sizeof(float) * V))) return rc;
    HIP_TRY(hipMemcpy(ws->in.p, img, sizeof(float) * V, hipMemcpyHostToDevice));
    din = ws->in.as<float>();
    dout = ws->out.as<float>();
  }
  if (adj)
    HIP_TRY(sr::launch_rot_adj(interpmeth, din, dout, ws->rv.as<sr::RotView>(), ws->vids.as<int>(), 1, 1,
                               ws->tb.as<int>(), ws->tw.as<float>(), g, nullptr));
  else
    HIP_TRY(sr::launch_rot_fwd(interpmeth, din, dout, nullptr, 0, ws->rv.as<sr::RotView>(), ws->vids.as<int>(), 1,
                               ws->tb.as<int>(), ws->tw.as<float>(), g, nullptr));
  HIP_TRY(hipStreamSynchronize(nullptr));
  if (host) HIP_TRY(hipMemcpy(out, dout, sizeof(float) * V, hipMemcpyDeviceToHost));
  return SR_OK;
}

// src/fft_convolve.jl:33-130 (fft_conv!, fft_conv_adj!), one plane
int sr_fft_conv(float* out, const float* img, int32_t nx, int32_t nz, const float* ker, int32_t px, int32_t pz,
                int32_t adj, int32_t mem) {
  int rc;
  if ((rc = unit_conv_checks(out, img, ker, nx, nz, px, pz))) return rc;
  return sr_fft_conv3(out, img, nx, 1, nz, ker, px, pz, adj ? SR_CONV3_ADJ : SR_CONV3_FWD, mem);
}

// src/fft_convolve.jl:160-239: the 3-D forms over the planes iy of [nx, ny, nz]
// arrays with kernels ker[:, :, iy]; every plane in one launch
int sr_fft_conv3(float* out, const float* img, int32_t nx, int32_t ny, int32_t nz, const float* ker, int32_t px,
                 int32_t pz, int32_t form, int32_t mem) {
  int rc;
  if ((rc = unit_conv_checks(out, img, ker, nx, nz, px, pz))) return rc;
  if (ny <= 0) return fail(SR_ERR_ARG, "size");
  if (form != SR_CONV3_FWD && form != SR_CONV3_ADJ && form != SR_CONV3_ADJ2) return fail(SR_ERR_ARG, "bad form");
  const bool host = (mem & 0xff) == SR_MEM_HOST;
  const long P = (long)nx * nz, V = P * ny, nin = form == SR_CONV3_ADJ2 ? P : V;
  const long pxz = (long)px * pz;
  std::vector<float> hk((size_t)pxz * ny);
  if (host) std::memcpy(hk.data(), ker, sizeof(float) * hk.size());
  else HIP_TRY(hipMemcpy(hk.data(), ker, sizeof(float) * hk.size(), hipMemcpyDeviceToHost));
  UnitWs* ws = unit_ws();
  std::lock_guard<std::mutex> lock(ws->mu);
  // the planes as a batch of one-row "views": plane iy is view iy with PSF
  // plane iy (Geo ny = 1, taps vstride 1), [view][z][x] contiguous
  sr::Geo gp{nx, 1, nz, 0, 0, px, pz};
  sr::Taps t;
  int sep = 0;
  if ((rc = unit_taps(ws, hk.data(), ny, px, pz, gp, &t, &sep)) || (rc = unit_vids(ws, ny)) ||
      (rc = ws->in.ensure(sizeof(float) * V)) || (rc = ws->out.ensure(sizeof(float) * V)) ||
      (rc = ws->part.ensure(sizeof(float) * V)))
    return rc;
  float* din = ws->in.as<float>();
  float* dout = ws->out.as<float>();
  const float* src = img;
  if (host) {
    HIP_TRY(hipMemcpy(din, img, sizeof(float) * nin, hipMemcpyHostToDevice));
    src = din;
  }
  const int* vids = ws->vids.as<int>();
  if (form == SR_CONV3_ADJ2) {
    // one 2-D image, every plane's adjoint: the PSF adjoint of one view over
    // ny planes writes [z][y][x], the caller's [nx, ny, nz] itself
    sr::Geo g{nx, ny, nz, 0, 0, px, pz};
    sr::Taps t3 = t;
    t3.vstride = 0;
    float* dst = host ? dout : out;
    HIP_TRY(sr::launch_psf_adj(sep, src, t3, vids, 1, dst, nullptr, 0, ny, g, nullptr));  // a block per plane
  } else {
    // the caller's [x][y][z] -> planes [y][z][x] (one plane: the same layout,
    // read and written in place unless the output aliases the input)
    float* res = host ? din : out;  // (the staged input is free after the permutation)
    const bool direct = ny == 1 && (host || src != res);
    const float* planes = src;
    if (!direct) {
      HIP_TRY(sr::launch_permute_rows(src, ws->part.as<float>(), nx, ny, nz, 0, nullptr));
      planes = ws->part.as<float>();
    }
    float* res_planes = direct && !host ? res : dout;
    if (form == SR_CONV3_FWD) {
      int used = 0;
      HIP_TRY(sr::launch_psf_fwd(sep, planes, t, vids, ny, res_planes, 1, &used, gp, nullptr));
    } else {
      HIP_TRY(sr::launch_psf_adj(sep, planes, t, vids, ny, res_planes, nullptr, 0, 1, gp, nullptr));
    }
    if (!direct) HIP_TRY(sr::launch_permute_rows(res_planes, res, nx, ny, nz, 1, nullptr));
    dout = direct ? res_planes : res;
  }
  HIP_TRY(hipStreamSynchronize(nullptr));
  if (host) HIP_TRY(hipMemcpy(out, dout, sizeof(float) * V, hipMemcpyDeviceToHost));
  return SR_OK;
}

int sr_imrotate_adj(float* out, const float* img, int32_t n, int32_t nz, float theta, int32_t interpmeth,
                    int32_t mem) {
  return sr_imrotate(out, img, n, nz, theta, interpmeth, 1, mem);
}

int sr_fft_conv_adj(float* out, const float* img, int32_t nx, int32_t nz, const float* ker, int32_t px, int32_t pz,
                    int32_t mem) {
  return sr_fft_conv(out, img, nx, nz, ker, px, pz, 1, mem);
}

}  // extern "C"
