// kernels.h -- launch interface of the HIP kernels (kernels.hip) used by the
// plan runtime (runtime.cpp).  Data is FP32, column-major as in the
// reference: a volume [x,y,z] has x fastest; the per-batch work volume is
// [b][z][y][x] (same as [x,y,z,b]).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <vector>

namespace sr {

// Per-view rotation parameters.
//  :two -- kind 0: forward copy (theta == 0 shortcut, src/rotatez.jl:341-344),
//          kind 2: bilinear through the affine map (xx..y0); the adjoint
//          always uses the map (src/rotatez.jl:390-416 has no shortcut).
//  :one -- kind 0: copy, 1: quadrant turn m only, 2: quadrant turn then the
//          three shears whose per-row tables live in the plan's tab arrays.
struct RotView {
  int kind;
  int m;
  float xx, xy, x0, yx, yy, y0;
};

struct Geo {
  int nx, ny, nz;  // image (nx == ny)
  int pad, Np;     // rotation pad and padded size, src/plan-rotate.jl:56
  int px, pz;      // PSF size
};

// PSF taps for a plan, padded with zeros to the template tap counts.
//  separable: u [planes][ntx], w [planes][ntz]; plane = v*vstride + y
//  direct:    h [planes][pz][ntx]
struct Taps {
  const float* u;
  const float* w;
  const float* h;
  // separable only: per plane, tail sums sum_{a>=k} and head sums sum_{a<=k}
  // of u and w (the replicate-pad fold of the adjoint), [planes][4][ntx]
  // ordered (u tail, u head, w tail, w head)
  const float* cum;
  int ntx, ntz;
  int vstride;  // 0 when every view shares one PSF stack, else ny
};

enum FwdEpilogue { EPI_VIEWS = 0, EPI_RATIO = 1 };
enum ImgMode { IMG_WRITE = 0, IMG_ACCUM = 1, IMG_EM = 2 };

// ---- rotation -----------------------------------------------------------
// out[b] (batch work volume) = rot_{vids[b]}(img) .* atten[vids[b]]
// (atten may be null).  img is one volume (shared by all b).
hipError_t launch_rot_fwd(int interp, const float* img, float* out, const float* atten, long atten_vstride,
                          const RotView* rv, const int* vids, int nb, const int* tab_b, const float* tab_w,
                          const Geo& g, hipStream_t s);
// part[grp] = sum over b in group of rot_{vids[b]}^T(work[b]); groups of vpg
// views; returns the number of groups through *ngroups.
hipError_t launch_rot_adj(int interp, const float* work, float* part, const RotView* rv, const int* vids, int nb,
                          int vpg, const int* tab_b, const float* tab_w, const Geo& g, hipStream_t s);
// image-reduce: out = combine(sum_g part[g]) per ImgMode
hipError_t launch_img_reduce(const float* part, int ngroups, float* out, const float* asum, long n, int mode,
                             hipStream_t s);

// attenuation factors, in place over the rotated mu volumes of the batch:
// a = exp(-dy * (sum_{j<=y} mu_r[j] - mu_r[y]/2)) from running sums along y
// (O(ny) per column, a wave-level segmented scan when the batch has few
// columns), over [b][y][x][z] volumes (zfast) or [b][z][y][x]
hipError_t launch_atten_scan(float* mur, int nb, const Geo& g, float dy, bool zfast, hipStream_t s);

// ---- PSF ----------------------------------------------------------------
// per-chunk partial views of the batch from the work volume (rotated and
// attenuated already); returns the chunk count actually used through
// *nchunk_used.  zfast: `work` is the z-fastest image [y][x][z] and the :two
// rotation (rv) and attenuation ([v][y][x][z]) are applied while staging
// family (z-fastest matrix-core path): PSF_AUTO takes the built-in kernel
// (forward: producer/consumer waves, psf_fwd_auto_npw producers, else
// column-owned; adjoint: column-owned), PSF_TILE the tile kernels, PSF_COL
// the column-owned ones, PSF_WS8 / PSF_WS4 (forward) producer/consumer with
// 8 / 4 producer waves, PSF_WS2 (forward) 4 producer waves and 4 consumer
// waves of two columns each
constexpr int PSF_AUTO = 0, PSF_TILE = 1, PSF_COL = 2, PSF_WS8 = 3, PSF_WS4 = 4, PSF_WS2 = 5;
bool psf_fwd_col_ok(const Geo& g, int ntx);
bool psf_fwd_ws2_ok(const Geo& g, int ntx);
bool psf_fwd_ws_ok(const Geo& g, int ntx, int npw);
int psf_fwd_auto_npw(const Geo& g, int ntx);
bool psf_adj_col_ok(const Geo& g, int ntx);
hipError_t launch_psf_fwd(int separable, const float* work, const Taps& t, const int* vids, int nb, float* part,
                          int nchunk, int* nchunk_used, const Geo& g, hipStream_t s, bool zfast = false,
                          const float* atten = nullptr, long atten_vstride = 0, const RotView* rv = nullptr,
                          int family = PSF_AUTO);
// views (or the ML-EM ratio y/(view + b) when epi == EPI_RATIO) = sum of partials
hipError_t launch_view_reduce(const float* part, int nchunk, const int* vids, int nb, float* out, const float* yobs,
                              const float* bkg, int epi, const Geo& g, hipStream_t s);
// work[b][:,y,:] = atten[v][:,y,:] .* conv_adj(views[b], psf[v,y])
hipError_t launch_psf_adj(int separable, const float* views, const Taps& t, const int* vids, int nb, float* work,
                          const float* atten, long atten_vstride, int nchunk, const Geo& g, hipStream_t s, bool zfast = false,
                          int family = PSF_AUTO);

// ---- PSF by FFT (x-fastest layout; fft_psf.hip) ---------------------------
// Padded plane sizes and LDS of the FFT kernels for a geometry.
struct FftPlan {
  int L1, L2;      // Power2(nx + px - 1), Power2(nz + pz - 1)
  int up, left;    // pads before x and z (src/helper.jl:9-12)
  int NPS;         // x-buffer stride per position (row pairs)
  int boff, poff;  // z-buffer and (adjoint) z-pad sums, float2 offsets after the x-buffer
  size_t lds_fwd, lds_adj;
};
// false when the padded plane or its LDS buffers exceed what the kernels take
bool psf_fft_plan(const Geo& g, FftPlan* f);
// W_L1^m then W_L2^m (e^{-2 pi i m / L}), rounded from double
void psf_fft_twiddles(const FftPlan& f, std::vector<float2>& tw);
// kernel spectra of nplanes planes (psfs [plane][pz][px]) into spec
// [plane][L2][L1/2 + 1], scaled by 1/(L1 L2), in the kernels' digit order
hipError_t launch_psf_spectrum(const float* psfs, int nplanes, const Geo& g, const FftPlan& f, float* spec,
                               hipStream_t s);
// as launch_psf_fwd / launch_psf_adj (x-fastest work volumes), through the
// plane FFTs; spectra of plane v * vstride + y
hipError_t launch_psf_fwd_fft(const float* work, const float* spec, int vstride, const int* vids, int nb, float* part,
                              int nchunk, int* nchunk_used, const Geo& g, const FftPlan& f, const float2* tw,
                              hipStream_t s);
hipError_t launch_psf_adj_fft(const float* views, const float* spec, int vstride, const int* vids, int nb,
                              float* work, const float* atten, long atten_vstride, int nchunk, const Geo& g,
                              const FftPlan& f, const float2* tw, hipStream_t s);

// ---- elementwise --------------------------------------------------------
hipError_t launch_fill(float* p, float v, long n, hipStream_t s);
hipError_t launch_zero_to_inf(float* p, long n, hipStream_t s);
// rows of nx floats: [z][y][x] -> [y][z][x] (mode 0) or back (mode 1)
hipError_t launch_permute_rows(const float* src, float* dst, int nx, int ny, int nz, int mode, hipStream_t s);
// dst[i] = src[i] for n 16-byte vectors (grid-stride streaming copy)
hipError_t launch_copy16(void* dst, const void* src, long n, hipStream_t s);
// z-fastest work layout ([y][x][z] volumes): the rotations as z-vector gathers
bool zfast_supported(const Geo& g);
// per-target (source pixel, weight) lists of the :two rotation of every view,
// padded to *K slots ([view][K][nx*nx]); false when a list exceeds kMaxEll
constexpr int kMaxEll = 8;
bool build_rot2_lists(const RotView* rv, int nview, const Geo& g, int* K, std::vector<int2>& ent);
// the same lists built on the device: the largest list per 256-target block
// (rot2_list_blocks(g) blocks per view) into blockmax, then, with K their
// maximum (at least 4), the slots into ent ([view][K][nx*nx])
int rot2_list_blocks(const Geo& g);
hipError_t launch_rot2_list_count(const RotView* rv, int nview, const Geo& g, int* blockmax, hipStream_t s);
hipError_t launch_rot2_list_fill(const RotView* rv, int nview, const Geo& g, int K, int2* ent, hipStream_t s);
// OS-EM update of the z-fastest image x with the subset normaliser computed
// in the same pass: asum = sum rot2^T(hvol .* atten[v]) (zeros as Inf, written
// to asum), x .*= (sum rot2^T(work[b])) ./ asum; hvol null: atten holds the
// products hvol .* atten[v] already (launch_mul_volumes)
hipError_t launch_rot2_adj_z_em_norm(const float* work, float* x, const int2* ent, int K, const int* vids, int nb,
                                     const Geo& g, hipStream_t s, const float* hvol, const float* atten,
                                     long atten_vstride, float* asum, int vs = 0);
// every OS-EM normaliser of a reconstruction in one launch: asum[k] =
// sum over view list k of rot2^T(hvol .* atten[v]) (zeros as Inf), for
// nsubsets consecutive lists in vids of nb views, the first nb_rem of them
// nb + 1 (the strided blocks laid out block after block); asum [k][y][x][z]
hipError_t launch_rot2_adj_z_norms(const float* hvol, float* asum, const int2* ent, int K, const int* vids, int nb,
                                   int nb_rem, int nsubsets, const Geo& g, hipStream_t s, const float* atten,
                                   long atten_vstride, int vs = 0);
// out[v] = h .* atten[v] for nview volumes of V floats (V % 4 == 0)
hipError_t launch_mul_volumes(const float* h, const float* atten, float* out, long V, int nview, hipStream_t s);
// img [z][y][x] -> out [y][x][z]
hipError_t launch_to_zfast(const float* img, float* out, const Geo& g, hipStream_t s);
// out[b] = atten[v] .* rot2(img_t) per view, all [y][x][z]; atten may be null
hipError_t launch_rot2_fwd_z(const float* img_t, float* out, const float* atten, long atten_vstride,
                             const RotView* rv, const int* vids, int nb, const Geo& g, hipStream_t s);
// part [y][x][z] = sum over the batch's views of rot2^T(work[b]) through the
// K-slot per-target lists (vs: views slices per block, 2, 4 or 8; 0 picks
// by the batch size); with atten: of rot2^T(work .* atten[v]), work one
// volume shared by the views; with em_asum: part is the z-fastest image x,
// updated in place as x .*= sum ./ em_asum (z-fastest); zero_inf: zero sums
// are written as Inf (OS-EM normalisers)
hipError_t launch_rot2_adj_z(const float* work, float* part, const int2* ent, int K, const int* vids, int nb,
                             const Geo& g, hipStream_t s, const float* atten = nullptr, long atten_vstride = 0,
                             const float* em_asum = nullptr, bool zero_inf = false, int vs = 0);
// out [z][y][x] = combine(sum_g part[g] [y][x][z]) per ImgMode; zout
// (optional) receives the new out as [y][x][z] too
hipError_t launch_img_reduce_z(const float* part, int ngroups, float* out, const float* asum, int mode, const Geo& g,
                               hipStream_t s, float* zout = nullptr);
// r = 2 (p - y) over n entries; *out = sum (p - y)^2 in double (partial: one
// double per 256 entries)
hipError_t launch_resid_loss(const float* p, const float* y, float* r, long n, double* partial, double* out,
                             hipStream_t s);
// mu-gradient in the rotated frame, in place over g ([b][y][x][z]):
// g[j] = -dy (sum_{y>j} g_y w_y + g_j w_j / 2) with w_y = R(x)_y a_y formed
// on the fly (R(x) the :two rotation of the z-fastest image img_t, as
// launch_rot2_fwd_z; a = atten[vids[b]]); ga (optional) receives g .* a
hipError_t launch_mu_grad_scan_rot(float* g, float* ga, const float* img_t, const float* atten, long atten_vstride,
                                   const RotView* rv, const int* vids, int nb, const Geo& geo, float dy,
                                   hipStream_t s);
// *flag = 0, then 1 if some p[k] is not > 0
hipError_t launch_flag_nonpositive(const float* p, long n, int* flag, hipStream_t s);

// ---- grouped launches -----------------------------------------------------
// One grid over the blocks of several launches of one kernel, each with its
// own plan's geometry, tables and buffers (the OS-EM steps of many instances
// in shared launches).  Descriptor i owns grid blocks [blk0, blk0 + nblk);
// its local block l runs as block (l % gx, l / gx) of its own launch.
// Descriptors of one grouped launch are sorted by blk0, blk0 of the first is
// 0, and each range starts where the previous one ends.
struct GrpHdr {
  int blk0, nblk, gx, pad;
};
struct PsfFwdGrp {  // k_psf_fwd_ws
  GrpHdr h;
  const float* img_t;
  Taps t;
  const int* vids;
  int Yc;
  float* part;
  Geo g;
  const float* atten;
  long atten_vstride;
  const RotView* rv;
};
struct ViewReduceGrp {  // k_view_reduce
  GrpHdr h;
  const float* part;
  int nchunk, P;
  const int* vids;
  float* out;
  const float* yobs;
  const float* bkg;
  int epi;
};
struct PsfAdjGrp {  // k_psf_adj_col
  GrpHdr h;
  const float* views;
  Taps t;
  const int* vids;
  int Yc;
  float* work;
  const float* atten;
  long atten_vstride;
  Geo g;
  int TBP;
};
struct EllGrp {  // k_rot2_adj_ell, one subset (gx = 1; blocks = 8 C, XCD-ordered)
  GrpHdr h;
  const float* work;
  float* part;
  const int2* ent;
  const int* vids;
  int nb, C;
  Geo g;
  const float* hvol;
  const float* atten;
  long atten_vstride;
  float* em_asum;
  int zero_inf;
};
// The kernel that launch_psf_fwd / launch_view_reduce / launch_psf_adj /
// launch_rot2_adj_z(_em_norm) would run for the same arguments, as a
// descriptor (its blk0 left 0) and a key naming the kernel instantiation
// (0: that launch has no grouped form); *lds: its dynamic LDS bytes.
// Launches of one key may share a grouped launch.
int psf_fwd_grp_desc(const float* work, const Taps& t, const int* vids, int nb, float* part, int nchunk,
                     int* nchunk_used, const Geo& g, const float* atten, long atten_vstride, const RotView* rv,
                     int family, PsfFwdGrp* d, size_t* lds);
int view_reduce_grp_desc(const float* part, int nchunk, const int* vids, int nb, float* out, const float* yobs,
                         const float* bkg, int epi, const Geo& g, ViewReduceGrp* d);
int psf_adj_grp_desc(const float* views, const Taps& t, const int* vids, int nb, float* work, const float* atten,
                     long atten_vstride, int nchunk, const Geo& g, int family, PsfAdjGrp* d, size_t* lds);
// mode 0: with em_asum, the OS-EM update of part (the z-fastest image) in
// place; mode 2: the update with the normaliser summed in the same pass
int ell_grp_desc(int mode, const float* work, float* part, const int2* ent, int K, const int* vids, int nb,
                 const Geo& g, const float* hvol, const float* atten, long atten_vstride, float* em_asum, int vs,
                 EllGrp* d);
// one grouped launch of key `key` over n descriptors (device memory) that
// cover `blocks` grid blocks; lds: the largest of the descriptors' LDS
hipError_t launch_psf_fwd_grp(int key, const PsfFwdGrp* d, int n, int blocks, size_t lds, hipStream_t s);
hipError_t launch_view_reduce_grp(const ViewReduceGrp* d, int n, int blocks, hipStream_t s);
hipError_t launch_psf_adj_grp(int key, const PsfAdjGrp* d, int n, int blocks, size_t lds, hipStream_t s);
hipError_t launch_ell_grp(int key, const EllGrp* d, int n, int blocks, hipStream_t s);

// psf kernel limits: the fast separable / direct kernels handle tap counts
// up to this many (padded to a multiple of 8)
constexpr int kMaxTaps = 64;
// LDS per workgroup on gfx950
constexpr size_t kMaxLds = 160 * 1024;
// largest padded rotation size the :one LDS kernel handles
constexpr int kMaxNpOne = 202;
// largest plane (nx * nz) the PSF kernels handle (one strip of <= 32 outputs
// per thread and phase, <= 1024 threads)
constexpr long kMaxPlane = 32768;

// whether the separable PSF kernels handle this geometry (nx % 4 == 0 and a
// plane small enough for one strip per thread); otherwise the direct kernel
bool psf_sep_supported(const Geo& g);
// the separable PSF passes run on the matrix cores for this geometry / tap count
bool psf_mfma_supported(const Geo& g, int nt);
// blocks of the :two rotation adjoint that one CU holds at once (LDS-bound)
int rot2_adj_blocks_per_cu(const Geo& g);

// sets the calling thread's sr_last_error() message and returns `code`
int record_error(int code, const char* msg);

// kernel classes for per-kernel event timing
enum KernelClass {
  K_ROT_FWD = 0, K_PSF_FWD, K_VIEW_REDUCE, K_PSF_ADJ, K_ROT_ADJ, K_IMG_REDUCE, K_ATTEN, K_OTHER, K_NCLASS
};

}  // namespace sr
