/*
 * spectrecon_amd.h -- C-ABI of the MI355X SPECT projector / reconstructor.
 *
 * A drop-in for the hot path of SPECTrecon.jl (SPECTplan -> project! /
 * backproject! -> mlem / osem).  Every entry point takes plain pointers and
 * sizes, FP32 data in the reference's column-major (Julia) layout:
 *   image, mumap : [nx, ny, nz]      (x fastest)
 *   views        : [nx, nz, nview]
 *   psfs         : [px, pz, ny, nview]
 * View indices are 0-based int32.  `mem` says where the array pointers live:
 * SR_MEM_HOST (the library stages them through device memory and returns
 * when the results are back on the host) or SR_MEM_DEVICE (pointers from
 * sr_malloc; the call runs on the plan's stream and returns after it
 * finished unless SR_ASYNC is or-ed into `mem`).
 *
 * Every function returns SR_OK (0) or an error code; sr_last_error() gives
 * the calling thread's last message.  The messages of argument checks equal
 * the strings the reference throws, so a host wrapper can rethrow them.
 */
#ifndef SPECTRECON_AMD_H
#define SPECTRECON_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SR_OK 0
#define SR_ERR_ARG 1   /* reference: throw(String) from an argument check   */
#define SR_ERR_DIM 2   /* reference: DimensionMismatch                      */
#define SR_ERR_HIP 3   /* HIP runtime or kernel launch failure              */
#define SR_ERR_ALLOC 4 /* device memory could not be allocated              */
#define SR_ERR_UNSUPPORTED 5 /* valid for the reference, not built here     */

#define SR_MEM_HOST 0
#define SR_MEM_DEVICE 1
#define SR_ASYNC 0x100
/* mlem / osem with device arrays: the caller has checked background > 0, so
 * the call skips its device scan of b (which waits for the stream) and an
 * SR_ASYNC call returns without any host wait */
#define SR_NO_CHECK 0x200
/* sr_osem_batch: every instance on its own plan's stream, no shared launches */
#define SR_BATCH_STREAMS 0x400

#define SR_INTERP_ONE 1 /* :one, 3-pass 1-D linear interpolation          */
#define SR_INTERP_TWO 2 /* :two, single-pass bilinear (reference default) */

#define SR_MODE_FAST 0 /* :fast */
#define SR_MODE_MEM 1  /* :mem  */

#define SR_CONV_AUTO 0     /* pick per plan (separable when every PSF is rank 1) */
#define SR_CONV_DIRECT 1   /* direct 2-D clamped convolution                     */
#define SR_CONV_SEPARABLE 2
#define SR_CONV_FFT 3      /* plane FFTs as src/fft_convolve.jl (x-fastest layout)   */

typedef struct sr_plan_s* sr_plan;

/* Plan geometry and options: SPECTplan(mumap, psfs, dy; T, viewangle,
 * interpmeth, nthread, mode), src/spectplan.jl:58-67.  T is Float32 here;
 * nthread is a Julia-side check and has no counterpart. */
typedef struct sr_plan_desc {
  int32_t nx, ny, nz; /* image size; nx == ny and even (src/spectplan.jl:76-77) */
  int32_t px, pz;     /* PSF size, both odd (src/spectplan.jl:82)               */
  int32_t nview;      /* size(psfs, 4)                                          */
  float dy;           /* voxel size used in exp(-mu*dy) (src/project.jl:111)    */
  int32_t interpmeth; /* SR_INTERP_ONE / SR_INTERP_TWO (src/spectplan.jl:87)    */
  int32_t mode;       /* SR_MODE_FAST / SR_MODE_MEM (src/spectplan.jl:93)       */
  int32_t conv;       /* SR_CONV_*: PSF method choice (no reference counterpart) */
} sr_plan_desc;

/* ---- errors, devices, memory --------------------------------------------- */
const char* sr_last_error(void);
int sr_version(void);
/* number of visible GPUs (0 when there is none); never fails */
int sr_device_count(void);
int sr_malloc(void** dptr, size_t bytes);
int sr_free(void* dptr);
int sr_memcpy_h2d(void* dst, const void* src, size_t bytes);
int sr_memcpy_d2h(void* dst, const void* src, size_t bytes);
/* device to device through the library's own streaming copy kernel (16-byte
 * vectors when both pointers are 16-byte aligned): also the plain copy of
 * known size that calibrates the HBM byte counters of a profile */
int sr_memcpy_d2d(void* dst, const void* src, size_t bytes);
int sr_memset(void* dptr, int value, size_t bytes);
int sr_device_synchronize(void);

/* ---- SPECTplan ------------------------------------------------------------
 * Replaces SPECTplan(...) (src/spectplan.jl:58-137) with device-resident
 * state: mu-map, per-view rotation tables, per-view attenuation factors
 * exp(-dy*(sum_{j<=y} mu_r[j] - mu_r[y]/2)) and PSF taps.  `viewangle` has
 * nview radians (the reference default is (0:nview-1)/nview*T(2pi),
 * src/spectplan.jl:63, computed by the caller).  `mumap`, `psfs` and
 * `viewangle` are host pointers. */
int sr_plan_create(const sr_plan_desc* desc, const float* mumap, const float* psfs, const float* viewangle,
                   sr_plan* plan);
int sr_plan_destroy(sr_plan plan);
/* Replaces the plan's mu-map (SPECTplan's mumap, src/spectplan.jl:71) and
 * rebuilds every view's attenuation factors in place, for a geometry, PSF
 * and angles that stay (a serving loop whose requests bring their own
 * mu-maps keeps one plan per geometry).  The factors are
 * exp(-dy (S_y - mu_r[y]/2)) with S_y the running sum of the rotated mu-map
 * along y (src/project.jl:103-113 sums each y anew: the same values up to
 * rounding).  `mumap` [nx,ny,nz] is a host (SR_MEM_HOST) or device pointer. */
int sr_plan_set_mumap(sr_plan plan, const float* mumap, int32_t mem);
/* resident device bytes and the PSF method the plan runs (SR_CONV_*; with
 * SR_CONV_AUTO, x-fastest plans whose PSFs are not rank 1 take the plane FFTs
 * when their padded planes fit the FFT kernels and the PSF has at least
 * 121 taps, else the direct 2-D kernels; the autotuner may switch) */
int sr_plan_info(sr_plan plan, size_t* device_bytes, int32_t* conv_method, int32_t* max_batch);
/* Device layout of the plan's work volumes, attenuation factors and partial
 * images (no reference counterpart; the caller's arrays keep the reference
 * layout either way): SR_LAYOUT_X_FASTEST [.][z][y][x], or
 * SR_LAYOUT_Z_FASTEST [.][y][x][z], which :two plans with matrix-core PSF
 * passes take (rotations as z-vector gathers, the rotation adjoint through
 * plan-time per-target source lists).  The environment variable
 * SR_LAYOUT=zyx at plan creation keeps the x-fastest layout. */
#define SR_LAYOUT_X_FASTEST 0
#define SR_LAYOUT_Z_FASTEST 1
int sr_plan_layout(sr_plan plan, int32_t* layout);
/* cap on views processed per launch batch (memory/parallelism trade-off) */
int sr_plan_set_batch(sr_plan plan, int32_t max_batch);
/* wait for all work queued on the plan's stream */
int sr_plan_synchronize(sr_plan plan);
/* Per-kernel timing with HIP events on the plan's stream (no reference
 * counterpart; for measurement).  While enabled, every kernel launch of the
 * chosen classes is bracketed by events; sr_plan_get_timing returns the
 * summed device time and launch count of one kernel class since timing was
 * (re)enabled: 0 rotate, 1 PSF forward, 2 view reduce, 3 PSF adjoint,
 * 4 rotate adjoint, 5 image reduce / EM update, 6 attenuation factors
 * (sr_plan_set_mumap).  `enable`: 0 off, 1 all
 * classes, otherwise an or of SR_TIME_CLASS(SR_K_*) bits (bracketing only
 * the classes of interest keeps the event overhead off the others). */
#define SR_TIME_CLASS(k) (2 << (k))
#define SR_TIME_ALL 0xfe
#define SR_K_ROT_FWD 0
#define SR_K_PSF_FWD 1
#define SR_K_VIEW_REDUCE 2
#define SR_K_PSF_ADJ 3
#define SR_K_ROT_ADJ 4
#define SR_K_IMG_REDUCE 5
#define SR_K_ATTEN 6
int sr_plan_set_timing(sr_plan plan, int32_t enable);
int sr_plan_get_timing(sr_plan plan, int32_t kernel, double* total_ms, int64_t* launches);

/* ---- project! / backproject! ----------------------------------------------
 * sr_project replaces project!(views, image, plan; index) (src/project.jl:143-163):
 *   views[:,:,ii] = A_{index[ii]} image, ii = 0..nindex-1.
 * sr_backproject replaces backproject!(image, views, plan; index)
 * (src/backproject.jl:132-162): image = sum_ii A_{index[ii]}' views[:,:,ii]
 * (image is overwritten, as the reference zeroes it).
 * index == NULL means all views 0..nview-1 (nindex must then be nview). */
int sr_project(sr_plan plan, float* views, const float* image, const int32_t* index, int32_t nindex, int32_t mem);
int sr_backproject(sr_plan plan, float* image, const float* views, const int32_t* index, int32_t nindex,
                   int32_t mem);

/* ---- mlem! / osem! ---------------------------------------------------------
 * sr_mlem replaces mlem!(out, x0, ynoisy, background, A; niter) with A the
 * plan's operator (src/ml-os-em.jl:21-49); sr_osem replaces
 * osem!(out, x0, ynoisy, background, Ablock(plan, nblocks); niter)
 * (src/ml-os-em.jl:70-131).  y, b: [nx,nz,nview]; b must be > 0 and
 * nview % nblocks == 0, with the reference's messages otherwise. */
int sr_mlem(sr_plan plan, float* out, const float* x0, const float* y, const float* b, int32_t niter, int32_t mem);
int sr_osem(sr_plan plan, float* out, const float* x0, const float* y, const float* b, int32_t nblocks,
            int32_t niter, int32_t mem);

/* Batched / served use (no reference counterpart): OS-EM over the strided
 * subsets nb:nblocks:nview for any 1 <= nblocks <= nview (block sizes differ
 * by at most one when nblocks does not divide nview; the reference's Ablock,
 * src/ml-os-em.jl:73, throws then, as sr_osem does).  With nblocks dividing
 * nview it equals sr_osem. */
int sr_osem_strided(sr_plan plan, float* out, const float* x0, const float* y, const float* b, int32_t nblocks,
                    int32_t niter, int32_t mem);
/* *done = 1 when all work queued on the plan's stream has finished (after an
 * SR_ASYNC call), else 0; never blocks.  Several plans (each with its own
 * stream) run their queued work concurrently on the device. */
int sr_plan_query(sr_plan plan, int32_t* done);
/* An asynchronous copy on the plan's stream, ordered with the plan's work
 * (no reference counterpart): for serving loops that stage request arrays
 * through pinned host memory (sr_host_alloc) while other plans compute. */
#define SR_COPY_H2D 1
#define SR_COPY_D2H 2
#define SR_COPY_D2D 3
int sr_plan_memcpy_async(sr_plan plan, void* dst, const void* src, size_t bytes, int32_t kind);
/* page-locked host memory (hipHostMalloc), for asynchronous copies */
int sr_host_alloc(void** ptr, size_t bytes);
int sr_host_free(void* ptr);
/* Batched OS-EM (no reference counterpart): n reconstructions, instance i on
 * plans[i] (distinct plans) with device arrays outs[i], x0s[i], ys[i], bs[i]
 * ([nx,nz,nview] of that plan) over the strided subsets of sr_osem_strided
 * with nblocks[i] blocks and niters[i] iterations.  All instances are checked
 * first (the reference's messages; nothing runs if one fails; b > 0 unless
 * SR_NO_CHECK), then every instance's kernel sequence is queued on its plan's
 * stream (replayed from the plan's HIP graph when sr_plan_set_graph is on),
 * so the device runs the instances concurrently.  Instances on z-fastest
 * plans with one separable PSF stack whose subsets fit one launch batch run
 * as shared launches instead: one launch sequence on a batch stream where,
 * at step s, subset s % nblocks[i] of iteration s / nblocks[i] of every
 * instance with more than s steps runs, and the PSF forwards, view
 * reductions, PSF adjoints and rotation-adjoint gathers of the step that use
 * one kernel instantiation share one grid (per-block descriptors of each
 * instance's launch), captured into a HIP graph and replayed when every
 * such plan has sr_plan_set_graph on.  Each instance's result equals its
 * own sr_osem_strided call bit for bit.  SR_BATCH_STREAMS runs every
 * instance on its own stream.  Returns after all finished, or at once with
 * SR_ASYNC (then sr_plan_query / sr_plan_synchronize per plan).  `mem` must
 * be SR_MEM_DEVICE (| SR_ASYNC | SR_NO_CHECK | SR_BATCH_STREAMS). */
int sr_osem_batch(int32_t n, const sr_plan* plans, float* const* outs, const float* const* x0s,
                  const float* const* ys, const float* const* bs, const int32_t* nblocks, const int32_t* niters,
                  int32_t mem);
/* Counters of sr_osem_batch's shared launches on the current device:
 * instances run as shared launches, grouped launches issued, batch graphs
 * captured and replayed (any pointer may be NULL). */
int sr_batch_stats(int64_t* instances, int64_t* launches, int64_t* captures, int64_t* replays);
/* HIP-graph replay of the EM kernel sequence (no reference counterpart; for
 * callers that run many small reconstructions, whose time is in launches).
 * With enable != 0, sr_mlem / sr_osem / sr_osem_strided run their first call
 * for a given set of buffers (caller pointers, plan buffers, nblocks, niter)
 * eagerly, capture the sequence on the second and replay it after that;
 * argument checks, host staging and result copies stay outside the graph,
 * and results equal the eager ones bit for bit.  Up to 8 graphs are kept per
 * plan; enable = 0 frees them.  Calls with timing on run eagerly.
 * sr_plan_graph_stats reports how many graphs were captured and replayed. */
int sr_plan_set_graph(sr_plan plan, int32_t enable);
int sr_plan_graph_stats(sr_plan plan, int64_t* captures, int64_t* replays);

/* Kernel choices of a plan (no reference counterpart; the reference has one
 * CPU code path).  choice[SR_CHOICE_*]:
 *   FWD_FAMILY: 0 the built-in matrix-core PSF forward (z-fastest plans:
 *     producer/consumer waves, the gathers overlapping the matrix products;
 *     8 producer waves for nz <= 80, else 4), 1 the tile kernel, 2 the
 *     column-owned kernel, 3 / 4 producer/consumer with 8 / 4 producer waves,
 *     5 producer/consumer with consumer waves of two columns whose x and z
 *     passes share one K loop;
 *   ADJ_FAMILY: 0 (or 2) the column-owned adjoint where it fits, 1 the tile
 *     kernel;
 *   FWD_CHUNKS / ADJ_CHUNKS: plane chunks per view, in percent of the built-in
 *     count (one block per compute unit over the batch), 10..1000;
 *   GATHER_SLICES: view slices per block of the rotation-adjoint gather
 *     (2, 4, 8; 0 picks by the launch's wave count and the batch size);
 *   PSF_METHOD: 0 the plan's built-in method, 1 the direct taps (separable
 *     when every PSF plane is rank 1, else the 2-D clamped convolution), 2
 *     the plane FFTs (plans with PSF spectra: x-fastest layout and padded
 *     planes the FFT kernels take; SR_ERR_UNSUPPORTED otherwise);
 *   EM_NORM: where OS-EM on a z-fastest plan with one PSF stack for all views
 *     sums its subset normalisers A_b' 1: 0 built-in (1), 1 in the gather of
 *     each subset's first iteration (the same pass as its back-projection),
 *     2 all subsets in one gather launch before the first iteration.
 * Defaults {0, 100, 0, 100, 0, 0, 0}.  Every choice computes the same operator;
 * the gather's slice count changes the order of its sums (last-bit
 * differences).  sr_plan_autotune times the candidates of each choice on
 * batches of nsub views (the OS-EM subset size) on the plan's stream, keeps
 * the fastest, and writes a JSON record of every candidate's time to `log`
 * (NUL-terminated, truncated to log_bytes; may be NULL); nsub <= 0 restores
 * the defaults.  A candidate replaces a built-in choice only when it is at
 * least 3% faster.  Changing the choices drops the plan's EM graphs. */
#define SR_CHOICE_FWD_FAMILY 0
#define SR_CHOICE_FWD_CHUNKS 1
#define SR_CHOICE_ADJ_FAMILY 2
#define SR_CHOICE_ADJ_CHUNKS 3
#define SR_CHOICE_GATHER_SLICES 4
#define SR_CHOICE_PSF_METHOD 5
#define SR_CHOICE_EM_NORM 6
#define SR_NCHOICE 7
int sr_plan_set_choice(sr_plan plan, const int32_t* choice);
int sr_plan_get_choice(sr_plan plan, int32_t* choice);
int sr_plan_autotune(sr_plan plan, int32_t nsub, char* log, size_t log_bytes);

/* ---- gradients ----------------------------------------------------------------
 * No reference counterpart (the reference differentiates only through the
 * linear operators, docs/lit/examples/6-dl.jl:145-156): for
 * L = sum over the indexed views of ||A(x, mu) - y||^2 with mu the plan's
 * mu-map, *loss = L, grad_x = dL/dx = A'(2 (A x - y)) and grad_mu = dL/dmu
 * through the rotated exponential cumulative sum of the attenuation.  y holds
 * the indexed views in order ([nx,nz,nindex]).  grad_x and grad_mu equal
 * sr_backproject and sr_project_vjp_mu of the residual r = 2 (A x - y) bit
 * for bit (both come from one PSF adjoint of r per launch batch).  Needs a
 * z-fastest plan (sr_plan_layout); SR_ERR_UNSUPPORTED otherwise.  `loss` is a
 * host pointer; the call returns after the stream finished. */
int sr_grad(sr_plan plan, double* loss, float* grad_x, float* grad_mu, const float* image, const float* y,
            const int32_t* index, int32_t nindex, int32_t mem);
/* The mu vector-Jacobian product of the projector (the suggested
 * sr_project_vjp_mu of the boundary): grad_mu = d<dy, A(x, mu)>/dmu for view
 * cotangents dy ([nx,nz,nindex], the indexed views in order), through the
 * attenuation of src/project.jl:103-116.  It is sr_grad's dL/dmu with the
 * residual 2 (A x - y) replaced by dy, without the projection and the loss:
 * what a ChainRulesCore rrule for a mu-dependent project calls (the image
 * cotangent is backproject(dy)).  Needs a z-fastest plan. */
int sr_project_vjp_mu(sr_plan plan, float* grad_mu, const float* image, const float* dy, const int32_t* index,
                      int32_t nindex, int32_t mem);

/* ---- unit operators ----------------------------------------------------------
 * The unit operators keep their device buffers and their tap and rotation
 * tables per device between calls (a repeated size, kernel or angle allocates
 * and uploads nothing), serialise on a per-device lock, run on the null
 * stream and return after the device finished.
 * sr_imrotate replaces imrotate!/imrotate_adj! on a stack of nz square n x n
 * slices (src/rotatez.jl:240-416, 3-D forms :472-495). */
int sr_imrotate(float* out, const float* img, int32_t n, int32_t nz, float theta, int32_t interpmeth, int32_t adj,
                int32_t mem);
/* sr_fft_conv replaces fft_conv!/fft_conv_adj! (src/fft_convolve.jl:33-130):
 * replicate-padded convolution of an nx x nz image with a px x pz kernel. */
int sr_fft_conv(float* out, const float* img, int32_t nx, int32_t nz, const float* ker, int32_t px, int32_t pz,
                int32_t adj, int32_t mem);
/* sr_fft_conv3 replaces the 3-D forms of src/fft_convolve.jl:160-239 on
 * [nx, ny, nz] arrays with kernels ker [px, pz, ny], every plane in one launch:
 *   SR_CONV3_FWD:  out[:, iy, :] = fft_conv(img[:, iy, :], ker[:, :, iy])
 *   SR_CONV3_ADJ:  out[:, iy, :] = fft_conv_adj(img[:, iy, :], ker[:, :, iy])
 *   SR_CONV3_ADJ2: out[:, iy, :] = fft_conv_adj(img2, ker[:, :, iy]) for one
 *                  [nx, nz] image img2 (fft_conv_adj2!, :217-239).
 * The same kernels as sr_fft_conv; the arrays are [nx, ny, nz] column-major
 * like the reference's (ker: [px, pz, ny]). */
#define SR_CONV3_FWD 0
#define SR_CONV3_ADJ 1
#define SR_CONV3_ADJ2 2
int sr_fft_conv3(float* out, const float* img, int32_t nx, int32_t ny, int32_t nz, const float* ker, int32_t px,
                 int32_t pz, int32_t form, int32_t mem);
/* The adjoints under their own names: imrotate_adj! (src/rotatez.jl:390-416,
 * :497-558) and fft_conv_adj! (src/fft_convolve.jl:80-130); the same as
 * sr_imrotate / sr_fft_conv with adj = 1. */
int sr_imrotate_adj(float* out, const float* img, int32_t n, int32_t nz, float theta, int32_t interpmeth,
                    int32_t mem);
int sr_fft_conv_adj(float* out, const float* img, int32_t nx, int32_t nz, const float* ker, int32_t px, int32_t pz,
                    int32_t mem);

/* sr_fft_conv for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * No counterpart in the reference.  Same arguments and results as
 * sr_fft_conv, but every buffer beyond the caller's arrays (tap tables, chunk
 * partials, staging of host arrays) lives in one fixed device array of
 * 411,392 bytes per GPU, held until the process ends: it calls no
 * hipMalloc / hipFree and allocates no host memory.  Limits: nx * nz <= 32768
 * and px, pz <= 64 (SR_ERR_UNSUPPORTED beyond).  Calls on one GPU share that
 * array, so calls that overlap on one GPU race; each call returns after its
 * work finished. */
int sr_fft_conv_noalloc(float* out, const float* img, int32_t nx, int32_t nz, const float* ker, int32_t px,
                        int32_t pz, int32_t adj, int32_t mem);

/* ---- input generators (host only) --------------------------------------------
 * sr_psf_gauss replaces psf_gauss(; ny, px, pz, fwhm_x, fwhm_z)
 * (src/psf-gauss.jl:24-55); psf is [px, pz, ny]. */
int sr_psf_gauss(float* psf, int32_t ny, int32_t px, int32_t pz, const double* fwhm_x, const double* fwhm_z);
/* Seeded synthetic inputs (no reference counterpart; SplitMix64): a sum of
 * 6 random ellipsoids with intensities U(0.5,4), and the attenuation map of
 * an outer 0.15 ellipsoid with two 0.05 inner ellipsoids. */
int sr_phantom(float* image, int32_t nx, int32_t ny, int32_t nz, uint64_t seed);
int sr_mumap(float* mumap, int32_t nx, int32_t ny, int32_t nz);
/* out[i] ~ Poisson(mean[i]), seeded */
int sr_poisson(float* out, const float* mean, size_t n, uint64_t seed);

#ifdef __cplusplus
}
#endif
#endif /* SPECTRECON_AMD_H */
