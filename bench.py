#!/usr/bin/env python3
"""Benchmark: OS-EM reconstructions/s at 128x128x80, 128 views, 37x37 PSF.

One step = one OS-EM reconstruction (8 subsets x 4 iterations, i.e. 5
back-projections and 4 projections of 128 views, plus the EM updates) on
device-resident inputs.  Prints one JSON line with the metric, per-step
latency percentiles, proj+backproj milliseconds, the roofline of the
dominant kernel (HIP-event timed inside the timed region) and a CPU baseline
(the oracle's Float32 restatement of the reference, threaded over views).
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "spectrecon.jl_amd"))

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3

NX, NZ, NVIEW, PSIZE = 128, 80, 128, 37
NBLOCKS, NITER = 8, 4
DY = 0.48  # cm: mu is in 1/cm and the reference multiplies mu*dy directly


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def workload():
    import spectrecon_amd as S

    image = S.phantom(NX, NX, NZ, seed=0)
    mumap = S.mumap_phantom(NX, NX, NZ)
    psf = S.psf_gauss(ny=NX, px=PSIZE, fwhm_start=1.5, fwhm_end=6.0)
    psfs = np.asfortranarray(np.repeat(psf[:, :, :, None], NVIEW, 3))
    return image, mumap, psfs


def algorithmic(cls, nb, g):
    """(bytes, flops) of one launch of a kernel class over nb views, for the
    z-fastest :two path, counted one way everywhere (the bench line and the
    profiles): per view, the projection reads the attenuation volume and the
    image and writes the view (2 V + P floats), the back-projection reads the
    view and the attenuation volume and writes the attenuated volume the
    rotation adjoint gathers (P + 2 V); the rotation adjoint with the EM update
    reads the nb volumes, x and the normaliser and writes x; the ratio reads
    A x, y and b and writes y / (A x + b); a rotation or an attenuation scan
    reads and writes one volume per view.  The PSF's flops are the separable
    clamped convolution's, 2 passes x p MACs per voxel."""
    nx, ny, nz, p = g
    V, P = nx * ny * nz, nx * nz
    sep_flops = 2 * 2 * p * nx * nz * ny * nb  # separable clamped conv: 2 passes x p MACs per voxel
    if cls == "psf_fwd":  # read attenuation + image, write the view
        return 4 * (2 * V + P) * nb, sep_flops
    if cls == "psf_adj":  # read view and attenuation, write the attenuated volume
        return 4 * (P + 2 * V) * nb, sep_flops
    if cls == "rot_fwd":  # the image to its z-fastest copy
        return 4 * 2 * V, 0
    if cls == "rot_adj":  # read the volumes, x and the normaliser; write x (the EM update is fused)
        return 4 * (V * nb + 3 * V), 0
    if cls == "view_reduce":  # the ratio y / (A x + b): read A x, y, b, write the ratio
        return 4 * 4 * P * nb, 0
    if cls == "img_reduce":  # the z-fastest image to the caller's layout
        return 4 * 2 * V, 0
    if cls == "atten":  # running sums along y: read the rotated mu volume, write the factors
        return 4 * 2 * V * nb, 0
    return 0, 0


def cpu_baseline(image, mumap, psfs, x0, y, b, xhat):
    """The CPU leg, with the oracle as the checker.

    - Baseline: one whole OS-EM reconstruction (8 subsets x 4 iterations) by
      the oracle's Float32 restatement of the reference (an FFT per plane, the
      O(ny^2) attenuation, views chunked over the host cores like spawner),
      timed on this host.
    - Errors of the GPU's reconstruction `xhat` (from the same x0, y, b):
      relative L2 against that Float32 reference-path result and against the
      FP64 oracle's OS-EM."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    import spectrecon_amd as S

    cores = min(os.cpu_count() or 1, NVIEW // NBLOCKS)
    ang = S.default_viewangle(NVIEW).astype(np.float64)
    P32 = O.Problem(mumap, psfs, DY, ang, "two", np.float32, "fft", nthread=cores)
    t0 = time.perf_counter()
    x32 = P32.osem(x0, y, b, NBLOCKS, NITER)
    t_recon = time.perf_counter() - t0
    log(f"cpu baseline: one OS-EM reconstruction in {t_recon:.1f} s on {cores} cores")
    P64 = O.Problem(mumap, psfs, DY, ang, "two", np.float64, "fft", nthread=cores)
    t0 = time.perf_counter()
    x64 = P64.osem(x0.astype(np.float64), y.astype(np.float64), b.astype(np.float64), NBLOCKS, NITER)
    log(f"fp64 oracle: {time.perf_counter() - t0:.1f} s")

    def rel(a, r):
        a, r = np.asarray(a, np.float64), np.asarray(r, np.float64)
        return float(np.linalg.norm(a - r) / np.linalg.norm(r))

    errors = {"gpu_vs_fp32_reference_path_rel_l2": rel(xhat, x32), "gpu_vs_fp64_oracle_rel_l2": rel(xhat, x64),
              "fp32_reference_path_vs_fp64_oracle_rel_l2": rel(x32, x64),
              "gpu_vs_fp64_oracle_max_rel": float(np.abs(np.asarray(xhat, np.float64) - x64).max()
                                                  / np.abs(x64).max()),
              "note": "the timed reconstruction's output; the FP32 reference path is the oracle's literal "
                      "restatement (FFT per plane), the FP64 oracle the same algorithm in double"}
    base = {"value": 1.0 / t_recon, "unit": "recons/s", "cores": cores, "kind": "port",
            "sample": (f"one OS-EM reconstruction (8 subsets x 4 iterations) at {NX}x{NX}x{NZ}, {NVIEW} views, "
                       f"{PSIZE}x{PSIZE} PSF by the oracle's Float32 restatement, {t_recon:.1f} s"),
            "calibration": calibration(O, S, cores)}
    return base, errors


def calibration(O, S, cores):
    """The reference's published timing problem (time/project.jl:17-43): 64x64x40,
    60 views, random mu-map, random 19x19 PSFs symmetrised by reversal and
    transposition (not separable), dy = 4.7952, mode :fast.  project! was
    published at 289.804 ms (:two) and 378.348 ms (:one) on an unstated CPU
    (time/project.jl:60-61, :56-57).  Here: the oracle's Float32 restatement
    on this host's cores (median of 3) and the library on the GPU (median of
    5, device arrays), with seeded inputs of the same construction."""
    rng = np.random.default_rng(0)
    nx, nz, nview, p = 64, 40, 60, 19
    mumap = np.asfortranarray(rng.random((nx, nx, nz), dtype=np.float32))
    psfs = rng.random((p, p, nx, nview), dtype=np.float32)
    psfs = psfs + psfs[::-1, ::-1]
    psfs = psfs + psfs.transpose(1, 0, 2, 3)
    psfs = np.asfortranarray(psfs / psfs.sum(axis=(0, 1), keepdims=True))
    x = np.asfortranarray(rng.random((nx, nx, nz), dtype=np.float32))
    dy = 4.7952
    ang = S.default_viewangle(nview).astype(np.float64)
    out = {"problem": "64x64x40, 60 views, 19x19 non-separable PSF (time/project.jl)",
           "reference_ms": {"two": 289.804, "one": 378.348}, "cores": cores}
    for interp in ("two", "one"):
        P = O.Problem(mumap, psfs, dy, ang, interp, np.float32, "fft", nthread=min(cores, nview))
        ts = []
        for _ in range(3):
            t0 = time.perf_counter()
            P.project(x)
            ts.append(time.perf_counter() - t0)
        out.setdefault("cpu_port_ms", {})[interp] = 1e3 * float(np.median(ts))
        plan = S.SPECTplan(mumap, psfs, dy, interpmeth=interp)
        d_x, d_v = S.DeviceArray.from_host(x), S.DeviceArray((nx, nz, nview))
        lib = S.lib()
        from spectrecon_amd import _lib as L
        L.check(lib.sr_project(plan.handle, d_v.ptr, d_x.ptr, None, nview, L.SR_MEM_DEVICE))
        tg = []
        for _ in range(5):
            t0 = time.perf_counter()
            L.check(lib.sr_project(plan.handle, d_v.ptr, d_x.ptr, None, nview, L.SR_MEM_DEVICE))
            tg.append(time.perf_counter() - t0)
        out.setdefault("gpu_ms", {})[interp] = 1e3 * float(np.median(tg))
        plan.close()
    return out


def stream_workload(kind, nreq, rate, seed=7, autotune=True, small_batch=0):
    """The batched and served patterns (one JSON line each; not the headline).

    batched: 32 seeded instances over nx in {64,96,128}, nz in {32,64,80},
      views in {60,120,128}, square PSF side in {15,25,37}, subsets in {4,8}
      (strided, uneven where the subsets do not divide the views), OS-EM x 4
      iterations, device-resident inputs on resident plans: all from one
      sr_osem_batch call (the subset steps of all instances as shared
      launches, one HIP graph replay), against the same call with every
      instance on its own stream, one call per instance, and the plain
      drop-in call per instance; every batched result must equal its
      one-at-a-time result.
    served: `nreq` requests with Poisson arrivals at `rate` per second, 70%
      128x128x80 / 128 views / 37x37 PSF and 30% 64x64x32 / 60 views / 15x15
      PSF, OS-EM 8 x 4 (strided subsets: 60 views over 8 blocks)."""
    from spectrecon_amd import serve as V

    rng = np.random.default_rng(seed)
    plans = {}
    reqs = []
    inst = []
    if kind == "batched":
        # the seeded uneven draw, or (small_batch) `small_batch` instances of
        # the served stream's small configuration, 64x64x32 / 60 views / 15x15
        # PSF, 8 subsets: launches that one instance alone does not fill
        draw = (V.batched_instances(seed) if not small_batch else
                [(V.Config(64, 32, 60, 15, DY, 1.0, 4.0), 8)] * small_batch)
        for k, (cfg, nblocks) in enumerate(draw):
            inst.append((cfg, nblocks))
            reqs.append(V.synthetic_request(cfg, seed=1 + k, nblocks=nblocks, niter=4, rid=k, plans=plans))
    else:
        big = V.Config(NX, NZ, NVIEW, PSIZE, DY, 1.5, 6.0)
        small = V.Config(64, 32, 60, 15, DY, 1.0, 4.0)
        t = 0.0
        for k in range(nreq):
            t += float(rng.exponential(1.0 / rate))
            cfg = big if rng.random() < 0.7 else small
            reqs.append(V.synthetic_request(cfg, seed=1000 + k, nblocks=8, niter=4, arrival=t, rid=k, plans=plans))
            if (k + 1) % 500 == 0:
                log(f"served: {k + 1} requests prepared")
    for p in plans.values():
        p.close()
    log(f"{kind}: {len(reqs)} requests prepared")
    nsubs = {}
    for r in reqs:
        nsubs.setdefault(r.config, set()).add(-(-r.config.nview // r.nblocks))
    out = {"workload": kind, "requests": len(reqs), "data": "synthetic (seeded phantoms, Poisson counts)"}
    if kind == "batched":
        # one resident plan per instance (built and, per configuration and
        # subset size, autotuned first: the batched pattern keeps them), the
        # inputs uploaded; then all instances from one sr_osem_batch call
        # (shared launches per step and kernel instantiation, one graph
        # replay after two calls), against every instance on its own stream
        # (concurrent, each plan replaying its own graph) and one call per
        # instance, each waited for
        backend = V.DeviceBackend(autotune=autotune)
        t0 = time.perf_counter()
        bplans = [backend.make_plan(cfg) for cfg, _ in inst]
        for plan, r in zip(bplans, reqs):
            if autotune:
                nsub = -(-r.config.nview // r.nblocks)
                if (r.config, nsub) not in backend.tuned:
                    backend.tune(plan, r.config, nsub)
                plan.set_choice(**backend.tuned[(r.config, nsub)])
        prep_s = time.perf_counter() - t0
        batch = V.Batch(bplans, reqs)
        for _ in range(2):  # (eager first call, graph capture on the second)
            batch.run()
            batch.run(streams=True)
        t_one = float(np.median([batch.run(sequential=True, streams=True) for _ in range(3)]))
        alone = batch.results()
        t_streams = float(np.median([batch.run(streams=True) for _ in range(5)]))
        same_streams = all(np.array_equal(a, b) for a, b in zip(alone, batch.results()))
        s0 = V.batch_stats()
        # the plain path: the drop-in call per request (host arrays, staged by
        # the library), one request after the other
        import spectrecon_amd as S
        t0 = time.perf_counter()
        plain = [S.osem_strided(r.x0, r.y, r.b, p, r.nblocks, r.niter) for p, r in zip(bplans, reqs)]
        t_plain = time.perf_counter() - t0
        t_all = float(np.median([batch.run() for _ in range(5)]))
        s1 = V.batch_stats()
        res = batch.results()
        same = same_streams and all(np.array_equal(a, b) for a, b in zip(alone, plain))
        rel_shared = max(float(np.linalg.norm(a - b) / np.linalg.norm(a)) for a, b in zip(alone, res))
        caps = [p.graph_stats() for p in bplans]
        batch.free()
        for p in bplans:
            backend.close_plan(p)
        out.update({"metric": (f"batched uneven OS-EM recons/s ({len(reqs)} instances, one sr_osem_batch call)"
                               if not small_batch else
                               f"batched OS-EM recons/s ({len(reqs)} instances of 64x64x32 / 60 views / 15x15 / "
                               "8 subsets, one sr_osem_batch call)"),
                    "unit": "recons/s", "value": len(reqs) / t_all,
                    "plain_one_at_a_time": len(reqs) / t_plain, "speedup_vs_plain": t_plain / t_all,
                    "resident_one_at_a_time": len(reqs) / t_one, "speedup_vs_resident": t_one / t_all,
                    "streams": len(reqs) / t_streams, "speedup_vs_streams": t_streams / t_all,
                    "results_equal_alone": bool(same),
                    "shared_vs_alone_max_rel_l2": rel_shared,
                    "results_note": "streams and plain results equal the one-at-a-time results bit for bit; the "
                                    "shared launches' forwards take the group's plane chunks, so their view sums "
                                    "round differently (max relative L2 difference reported)",
                    "batch_ms": 1e3 * t_all, "plain_ms": 1e3 * t_plain,
                    "resident_sequential_ms": 1e3 * t_one, "streams_ms": 1e3 * t_streams, "prepare_s": prep_s,
                    "shared_launches": {"instances_per_run": (s1["instances"] - s0["instances"]) // 5,
                                        "grouped_launches_per_run": (s1["launches"] - s0["launches"]) // 5,
                                        "graph_replays": s1["replays"] - s0["replays"]},
                    "baselines": "plain: one drop-in call per instance with host arrays (sr_osem_strided, "
                                 "SR_MEM_HOST), each returning before the next; resident: one sr_osem_batch call "
                                 "per instance on its plan's stream (SR_BATCH_STREAMS) on the same device-resident "
                                 "inputs with graph replays, each waited for; streams: one sr_osem_batch call with "
                                 "every instance on its own stream, concurrently",
                    "graphs": {"captured": sum(c[0] for c in caps), "replays": sum(c[1] for c in caps)},
                    "instances": [{"nx": c.nx, "nz": c.nz, "nview": c.nview, "psf": c.psf, "nblocks": nb}
                                  for c, nb in inst],
                    "autotune": backend.tune_log})
    else:
        srv = V.Server(hbm_budget=200e9, max_active=8, backend=V.DeviceBackend(autotune=autotune))
        try:
            srv.prepare({r.config for r in reqs}, per_config=2, nsubs=nsubs)
            st = srv.run(reqs, progress=lambda i: log(f"served: {i} requests submitted"))
            # the plain path on the same requests: one drop-in call per request
            # (host arrays, staged by the library), one after the other with no
            # arrival gaps, on the server's resident plans with the same choices
            import spectrecon_amd as S
            t0 = time.perf_counter()
            for r in reqs:
                plan = srv.pools[r.config][0]
                nsub = -(-r.config.nview // r.nblocks)
                if autotune and (r.config, nsub) in srv.backend.tuned and \
                        plan.choice() != srv.backend.tuned[(r.config, nsub)]:
                    plan.set_choice(**srv.backend.tuned[(r.config, nsub)])
                S.osem_strided(r.x0, r.y, r.b, plan, r.nblocks, r.niter)
            t_plain = time.perf_counter() - t0
            log(f"served: plain path {len(reqs) / t_plain:.1f} recons/s")
        finally:
            srv.close()
        st["plain_one_at_a_time_recons_per_s"] = len(reqs) / t_plain
        st["speedup_vs_plain"] = st["recons_per_s"] / st["plain_one_at_a_time_recons_per_s"]
        out.update({"metric": f"served OS-EM stream, Poisson arrivals {rate:g}/s", "unit": "recons/s",
                    "value": st["recons_per_s"], "latency_ms": st["latency_ms"], "rejected": st["rejected"],
                    "max_queued": st["max_queued"],
                    "plain_one_at_a_time": st["plain_one_at_a_time_recons_per_s"],
                    "speedup_vs_plain": st["speedup_vs_plain"],
                    "note": "value: completed requests per second of the stream (the arrival rate bounds it "
                            "unless the stream saturates the device); plain: the same requests through the "
                            "drop-in host-array call, one after the other with no arrival gaps",
                    "stats": st})
    print(json.dumps(out), flush=True)


def gradient_workload(steps):
    """L = ||A(x, mu) - y||^2 and its gradients w.r.t. x and mu at 128x128x80,
    128 views, 37x37 PSF, for a batch of 8 seeded phantoms (device arrays)."""
    import ctypes as C

    import spectrecon_amd as S
    from spectrecon_amd import _lib as L

    image, mumap, psfs = workload()
    plan = S.SPECTplan(mumap, psfs, DY)
    lib = S.lib()
    ys = [S.DeviceArray.from_host(0.9 * S.project(S.phantom(NX, NX, NZ, seed=50 + k), plan)) for k in range(8)]
    xs = [S.DeviceArray.from_host(S.phantom(NX, NX, NZ, seed=k)) for k in range(8)]
    gx, gmu = S.DeviceArray((NX, NX, NZ)), S.DeviceArray((NX, NX, NZ))
    loss = C.c_double()

    def batch():
        for x, y in zip(xs, ys):
            L.check(lib.sr_grad(plan.handle, C.byref(loss), gx.ptr, gmu.ptr, x.ptr, y.ptr, None, NVIEW,
                                L.SR_MEM_DEVICE))
    batch()
    t0 = time.perf_counter()
    for _ in range(steps):
        batch()
    dt = (time.perf_counter() - t0) / steps
    print(json.dumps({"workload": "gradient", "metric": "gradients (x and mu) per second, batch of 8 phantoms",
                      "unit": "gradients/s", "value": 8 / dt, "ms_per_batch": 1e3 * dt, "steps": steps,
                      "config": {"nx": NX, "nz": NZ, "nview": NVIEW, "psf": PSIZE, "batch": 8},
                      "data": "synthetic (seeded phantoms; y = 0.9 A x' of other phantoms)"}), flush=True)


def psf_workload(reps=5):
    """The PSF methods on general (non-separable) PSFs, device arrays:
    projection and back-projection of all views with the direct 2-D taps and
    with the plane FFTs, and the FFT kernels' FP32 rate counted as
    5 N log2(N) flops per padded plane (N = L1 L2) and direction (the
    convention of the reference's FFT path: forward and inverse transform per
    plane, src/fft_convolve.jl:16-25; the spectral product is not counted).

    - clinical: 128x128x80, 128 views, one random point-symmetric 37x37 PSF per
      depth (seeded), the phantom mu-map, dy = 0.48;
    - timing: the reference's timing problem (time/project.jl: 64x64x40, 60
      views, random 19x19 PSFs symmetrised, dy = 4.7952), on the method the
      autotuner picks."""
    import spectrecon_amd as S
    from spectrecon_amd import _lib as L

    lib = S.lib()
    out = {"workload": "psf", "unit": "ms", "note": "device-resident project!/backproject! of all views; "
           "fft_TFLOPs counts 5 N log2 N per padded plane and direction"}

    def timed(plan, dv, dx, nview):
        def proj():
            L.check(lib.sr_project(plan.handle, dv.ptr, dx.ptr, None, nview, L.SR_MEM_DEVICE))

        def back():
            L.check(lib.sr_backproject(plan.handle, dx.ptr, dv.ptr, None, nview, L.SR_MEM_DEVICE))
        res = {}
        for name, fn in (("project_ms", proj), ("backproject_ms", back)):
            fn()
            ts = []
            for _ in range(reps):
                t0 = time.perf_counter()
                fn()
                ts.append(time.perf_counter() - t0)
            res[name] = 1e3 * float(np.median(ts))
        plan.set_timing(True, classes=["psf_fwd", "psf_adj"])
        proj()
        back()
        tm = plan.timing()
        plan.set_timing(False)
        for cls in ("psf_fwd", "psf_adj"):
            res[cls + "_launch_ms"] = tm[cls][0] / max(1, tm[cls][1])
            res[cls + "_launches"] = tm[cls][1]
        return res

    def fft_rate(res, nx, nz, p, ny, nview):
        L1 = 1 << math.ceil(math.log2(nx + p - 1))
        L2 = 1 << math.ceil(math.log2(nz + p - 1))
        N = L1 * L2
        fl = 5 * N * math.log2(N) * ny * nview  # all planes of all views, one direction
        for cls, key in (("psf_fwd", "fwd"), ("psf_adj", "adj")):
            t = res[cls + "_launch_ms"] * res[cls + "_launches"] / 1e3
            res[f"fft_{key}_TFLOPs"] = fl / t / 1e12
            res[f"fft_{key}_frac"] = fl / t / 1e12 / FP32_PEAK_TFLOPS
        res["padded_plane"] = [L1, L2]

    rng = np.random.default_rng(0)
    nx, nz, nview, p = NX, NZ, NVIEW, PSIZE
    mumap = S.mumap_phantom(nx, nx, nz)
    k = rng.random((p, p, nx)).astype(np.float32)
    k = k + k[::-1, ::-1]
    k = k / k.sum(axis=(0, 1), keepdims=True)
    k = 0.5 * (k + k[::-1, ::-1])
    psfs = np.asfortranarray(np.repeat(k[:, :, :, None], nview, 3).astype(np.float32))
    plan = S.SPECTplan(mumap, psfs, DY)
    dx = S.DeviceArray.from_host(S.phantom(nx, nx, nz, seed=0))
    dv = S.DeviceArray((nx, nz, nview))
    clinical = {"problem": f"{nx}x{nx}x{nz}, {nview} views, {p}x{p} non-separable PSF", "info": plan.info()}
    for m, name in ((1, "direct"), (2, "fft")):
        plan.set_choice(psf_method=m)
        log(f"psf: clinical, {name}")
        clinical[name] = timed(plan, dv, dx, nview)
    fft_rate(clinical["fft"], nx, nz, p, nx, nview)
    plan.close()
    out["clinical"] = clinical

    nx, nz, nview, p = 64, 40, 60, 19
    mumap = np.asfortranarray(rng.random((nx, nx, nz), dtype=np.float32))
    psfs = rng.random((p, p, nx, nview), dtype=np.float32)
    psfs = psfs + psfs[::-1, ::-1]
    psfs = psfs + psfs.transpose(1, 0, 2, 3)
    psfs = np.asfortranarray(psfs / psfs.sum(axis=(0, 1), keepdims=True))
    plan = S.SPECTplan(mumap, psfs, 4.7952)
    rec = plan.autotune(nview)
    dx = S.DeviceArray.from_host(np.asfortranarray(rng.random((nx, nx, nz), dtype=np.float32)))
    dv = S.DeviceArray((nx, nz, nview))
    log("psf: timing problem")
    tp = {"problem": "64x64x40, 60 views, 19x19 random PSFs (time/project.jl)", "autotune": rec.get("psf_method"),
          "info": plan.info(), "published_ms": {"project": 289.804, "backproject": 208.163}}
    tp.update(timed(plan, dv, dx, nview))
    if plan.info()["conv"] == "fft":
        fft_rate(tp, nx, nz, p, nx, nview)
    plan.close()
    out["timing_problem"] = tp
    print(json.dumps(out), flush=True)


def units_workload(reps=20):
    """Each of the reference's published timings (time/*.jl; unstated CPU)
    beside the library on the GPU for the same operation and sizes (device
    arrays, median of `reps` calls; every call returns after its stream
    finished, so small operations measure launch latency)."""
    import spectrecon_amd as S
    from spectrecon_amd import _lib as L

    lib = S.lib()
    rng = np.random.default_rng(0)

    def med(fn):
        fn()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        return 1e3 * float(np.median(ts))

    rows = []
    # imrotate! of one 100x100 slice, theta = 3 pi / 16 (time/rotate2.jl:9-40)
    sl = S.DeviceArray.from_host(np.asfortranarray(rng.random((100, 100, 1), dtype=np.float32)))
    so = S.DeviceArray((100, 100, 1))
    th = float(np.float32(3 * np.pi / 16))
    for interp, name, pub in ((L.SR_INTERP_TWO, "two", (0.162171, 0.168518)), (L.SR_INTERP_ONE, "one", (0.533316, 0.517367))):
        for adj in (0, 1):
            rows.append({"op": f"imrotate{'_adj' if adj else ''}! 100x100 slice, :{name}", "published_ms": pub[adj],
                         "gpu_ms": med(lambda: L.check(lib.sr_imrotate(so.ptr, sl.ptr, 100, 1, th, interp, adj,
                                                                        L.SR_MEM_DEVICE)))})
    # 3-D rotation of 128x128x100, theta = pi / 6 (time/rotate3.jl:9-24, :two)
    v3 = S.DeviceArray.from_host(np.asfortranarray(rng.standard_normal((128, 128, 100)).astype(np.float32)))
    o3 = S.DeviceArray((128, 128, 100))
    th3 = float(np.float32(np.pi / 6))
    rows.append({"op": "imrotate! 128x128x100, :two, threaded", "published_ms": 5.9,
                 "gpu_ms": med(lambda: L.check(lib.sr_imrotate(o3.ptr, v3.ptr, 128, 100, th3, L.SR_INTERP_TWO, 0,
                                                                L.SR_MEM_DEVICE)))})
    # fft_conv! / fft_conv_adj! of a 200x64 plane with a 5x5 box kernel (time/fft_convolve.jl:8-24)
    img = np.zeros((200, 64), np.float32)
    img[19:150, 19:40] = rng.random((131, 21))
    di, do = S.DeviceArray.from_host(np.asfortranarray(img)), S.DeviceArray((200, 64))
    dk = S.DeviceArray.from_host(np.asfortranarray(np.full((5, 5), 1 / 25, np.float32)))
    for adj, pub in ((0, 0.565631), (1, 0.561371)):
        rows.append({"op": f"fft_conv{'_adj' if adj else ''}! 200x64, 5x5", "published_ms": pub,
                     "gpu_ms": med(lambda: L.check(lib.sr_fft_conv(do.ptr, di.ptr, 200, 64, dk.ptr, 5, 5, adj,
                                                                    L.SR_MEM_DEVICE)))})
    # the 3-D fft_conv! / fft_conv_adj! forms (src/fft_convolve.jl:160-209) over 128 planes of 128x80 with
    # 37x37 Gaussian kernels: no published timing; every plane in one call and launch
    nx3, ny3, nz3, p3 = 128, 128, 80, 37
    img3 = S.DeviceArray.from_host(np.asfortranarray(rng.random((nx3, ny3, nz3), dtype=np.float32)))
    out3 = S.DeviceArray((nx3, ny3, nz3))
    ker3 = S.DeviceArray.from_host(np.asfortranarray(S.psf_gauss(ny=ny3, px=p3, pz=p3, fwhm_start=1.5,
                                                                  fwhm_end=6.0)))
    for form, name in ((L.SR_CONV3_FWD, "fft_conv!"), (L.SR_CONV3_ADJ, "fft_conv_adj!")):
        rows.append({"op": f"3-D {name} 128x128x80, 37x37 per plane", "published_ms": None,
                     "gpu_ms": med(lambda: L.check(lib.sr_fft_conv3(out3.ptr, img3.ptr, nx3, ny3, nz3, ker3.ptr, p3,
                                                                     p3, form, L.SR_MEM_DEVICE)))})
    # SPECTplan, project!, backproject! on the timing problem (time/SPECTplan.jl, time/project.jl,
    # time/backproject.jl): 64x64x40, 60 views, random 19x19 PSFs symmetrised, dy = 4.7952
    nx, nz, nview, p = 64, 40, 60, 19
    mumap = np.asfortranarray(rng.random((nx, nx, nz), dtype=np.float32))
    psfs = rng.random((p, p, nx, nview), dtype=np.float32)
    psfs = psfs + psfs[::-1, ::-1]
    psfs = psfs + psfs.transpose(1, 0, 2, 3)
    psfs = np.asfortranarray(psfs / psfs.sum(axis=(0, 1), keepdims=True))
    dx = S.DeviceArray.from_host(np.asfortranarray(rng.random((nx, nx, nz), dtype=np.float32)))
    dv = S.DeviceArray.from_host(np.asfortranarray(rng.random((nx, nz, nview), dtype=np.float32)))
    # the reference times a construction after a first one (@btime, which reports the minimum); here the
    # first construction in the process is reported beside the median of later ones
    t0 = time.perf_counter()
    S.SPECTplan(mumap, psfs, 4.7952).close()
    first_ms = 1e3 * (time.perf_counter() - t0)
    rows.append({"op": "SPECTplan 64x64x40, 60 views, 19x19", "published_ms": 13.585,
                 "gpu_ms": med(lambda: S.SPECTplan(mumap, psfs, 4.7952).close()), "first_ms": first_ms,
                 "note": "host inputs, kernel spectra included; gpu_ms the median of later constructions, "
                         "first_ms the process's first of this shape"})
    pubs = {("two", "fast"): (289.804, 208.163), ("one", "fast"): (378.348, 343.419),
            ("two", "mem"): (618.369, 246.847), ("one", "mem"): (793.918, 379.044)}
    for (interp, mode), (pp, pb) in pubs.items():
        plan = S.SPECTplan(mumap, psfs, 4.7952, interpmeth=interp, mode=mode)
        rows.append({"op": f"project! 64x64x40, 60 views, :{interp}, :{mode}", "published_ms": pp,
                     "gpu_ms": med(lambda: L.check(lib.sr_project(plan.handle, dv.ptr, dx.ptr, None, nview,
                                                                   L.SR_MEM_DEVICE)))})
        rows.append({"op": f"backproject! 64x64x40, 60 views, :{interp}, :{mode}", "published_ms": pb,
                     "gpu_ms": med(lambda: L.check(lib.sr_backproject(plan.handle, dx.ptr, dv.ptr, None, nview,
                                                                       L.SR_MEM_DEVICE)))})
        plan.close()
    for r in rows:
        r["speedup"] = r["published_ms"] / r["gpu_ms"] if r["published_ms"] else None
    print(json.dumps({"workload": "units", "note": "published: the reference's own timings on an unstated CPU "
                      "(time/*.jl); gpu: this library on one MI355X", "rows": rows}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", choices=["osem", "batched", "served", "gradient", "units", "psf"], default="osem",
                    help="osem: the headline OS-EM line (default); batched / served: the concurrent patterns; "
                         "gradient: the differentiable projector; units: the reference's published timings; "
                         "psf: direct vs FFT PSF methods on general PSFs")
    ap.add_argument("--requests", type=int, default=100, help="served: number of requests")
    ap.add_argument("--small-batch", type=int, default=0,
                    help="batched: this many instances of 64x64x32 / 60 views / 15x15 / 8 subsets instead of the "
                         "seeded uneven draw")
    ap.add_argument("--rate", type=float, default=20.0, help="served: Poisson arrival rate per second")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--latency-samples", type=int, default=200,
                    help="reconstructions after the timed region for the latency percentiles")
    ap.add_argument("--no-autotune", action="store_true", help="keep the built-in kernel choices")
    ap.add_argument("--choices", default=None,
                    help="pin the kernel choices (six comma-separated SR_CHOICE_* values) instead of autotuning")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="after the benchmark, copy this many MiB 3 times with the library's copy kernel "
                         "(the known-size reference for the FETCH_SIZE / WRITE_SIZE counters)")
    ap.add_argument("--traffic-from", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="no benchmark: turn the counter_collection.csv files of a FETCH_SIZE and a WRITE_SIZE "
                         "rocprofv3 pass (run with --copy-mib) into profiles/traffic.json")
    args = ap.parse_args()
    if args.traffic_from:
        from spectrecon_amd import profiling

        if args.copy_mib <= 0:
            raise SystemExit("--traffic-from needs the --copy-mib of the profiled run")
        t = profiling.write_traffic(*args.traffic_from, args.copy_mib << 20,
                                    os.path.join(ROOT, "profiles", "traffic.json"))
        print(json.dumps(t, indent=1))
        return
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {args.gpus}); the multi-GPU "
              "sharding over views is described in DESIGN.md but not built", file=sys.stderr)
        sys.exit(2)

    import spectrecon_amd as S
    from spectrecon_amd import _lib as L

    if args.workload == "gradient":
        return gradient_workload(max(1, args.steps))
    if args.workload == "units":
        return units_workload()
    if args.workload == "psf":
        return psf_workload()
    if args.workload != "osem":
        return stream_workload(args.workload, args.requests, args.rate, autotune=not args.no_autotune,
                               small_batch=args.small_batch)
    if S.device_count() == 0:
        raise SystemExit("no HIP device")
    lib = S.lib()
    image, mumap, psfs = workload()
    # the process's first plan also pays the HIP runtime's one-time start-up
    # (context, code objects): a 16x16x16 one-view plan takes it first
    t0 = time.perf_counter()
    S.SPECTplan(np.zeros((16, 16, 16), np.float32, order="F"),
                np.asfortranarray(S.psf_gauss(ny=16, px=3, fwhm_start=1.0, fwhm_end=1.0)[:, :, :, None]), DY).close()
    init_ms = 1e3 * (time.perf_counter() - t0)
    t0 = time.perf_counter()
    plan = S.SPECTplan(mumap, psfs, DY)
    plan_ms = 1e3 * (time.perf_counter() - t0)
    info = plan.info()
    log(f"plan: {plan_ms:.1f} ms, {info}")
    # kernel choices for 16-view subsets, timed on this device (part of
    # setting the plan up, outside the timed region)
    tune_ms, tune = None, None
    if args.choices:
        plan.set_choice(**dict(zip(plan.CHOICES, (int(c) for c in args.choices.split(",")))))
    elif not args.no_autotune:
        t0 = time.perf_counter()
        tune = plan.autotune(NVIEW // NBLOCKS)
        tune_ms = 1e3 * (time.perf_counter() - t0)
        log(f"autotune: {tune_ms:.1f} ms, {plan.choice()}")

    # measurements: y ~ Poisson(s*A x + b), s = 5e6 / sum(A x), b = 0.1 mean(s A x)
    ytrue = S.project(image, plan)
    s = 5e6 / float(ytrue.sum())
    bval = 0.1 * float((s * ytrue).mean())
    b = np.full_like(ytrue, bval)
    y = S.poisson(s * ytrue + b, seed=0)
    d_y = S.DeviceArray.from_host(y)
    d_b = S.DeviceArray.from_host(b)
    d_x0 = S.DeviceArray.from_host(np.ones(image.shape, np.float32))
    d_out = S.DeviceArray(image.shape)

    def recon():
        L.check(lib.sr_osem(plan.handle, d_out.ptr, d_x0.ptr, d_y.ptr, d_b.ptr, NBLOCKS, NITER, L.SR_MEM_DEVICE))

    # the EM kernel sequence replays from a HIP graph (captured on the second
    # call with these buffers; calls with event timing on run eagerly)
    plan.set_graph(True)
    for w in range(max(2, args.warmup)):
        recon()
        log(f"warmup {w + 1}/{max(2, args.warmup)} done")
    # per-class kernel breakdown from one reconstruction with every launch
    # bracketed by events (outside the timed region: the events cost time)
    plan.set_timing(True)
    recon()
    breakdown = plan.timing()
    dom = max(breakdown, key=lambda k: breakdown[k][0])
    log(f"breakdown: {breakdown}")
    # timed region: the dominant kernel's launches of the first timed step
    # carry events (its average launch time for the roofline, measured inside
    # the timed region); the other steps run without events (events around
    # 32 launches cost about 0.25 ms per reconstruction) and replay the graph
    plan.set_timing(True, classes=[dom])
    L.check(lib.sr_device_synchronize())
    lat = []
    timing = None
    t_start = time.perf_counter()
    for k in range(args.steps):
        if k == 1:
            timing = plan.timing()
            plan.set_timing(False)
        t1 = time.perf_counter()
        recon()  # returns after the plan's stream finished
        lat.append(time.perf_counter() - t1)
        log(f"step {k + 1}/{args.steps}: {1e3 * lat[-1]:.2f} ms")
    L.check(lib.sr_device_synchronize())
    t_total = time.perf_counter() - t_start
    if timing is None:
        timing = plan.timing()
    plan.set_timing(False)
    xhat = d_out.to_host()
    if not np.all(np.isfinite(xhat)):
        raise SystemExit("non-finite reconstruction")
    # latency percentiles from more samples than the timed steps (p99 of 20
    # samples is their maximum): the same reconstruction, outside the timed region
    lat_n = max(args.latency_samples, args.steps)
    lat_s = []
    for _ in range(lat_n):
        t1 = time.perf_counter()
        recon()
        lat_s.append(time.perf_counter() - t1)
    log(f"latency: {lat_n} more reconstructions")

    # proj + backproj of all 128 views, device resident, timed separately
    d_img = S.DeviceArray.from_host(image)
    d_views = S.DeviceArray((NX, NZ, NVIEW))
    reps = 5
    L.check(lib.sr_project(plan.handle, d_views.ptr, d_img.ptr, None, NVIEW, L.SR_MEM_DEVICE))
    L.check(lib.sr_backproject(plan.handle, d_img.ptr, d_views.ptr, None, NVIEW, L.SR_MEM_DEVICE))
    t1 = time.perf_counter()
    for _ in range(reps):
        L.check(lib.sr_project(plan.handle, d_views.ptr, d_img.ptr, None, NVIEW, L.SR_MEM_DEVICE))
    proj_ms = 1e3 * (time.perf_counter() - t1) / reps
    t1 = time.perf_counter()
    for _ in range(reps):
        L.check(lib.sr_backproject(plan.handle, d_img.ptr, d_views.ptr, None, NVIEW, L.SR_MEM_DEVICE))
    back_ms = 1e3 * (time.perf_counter() - t1) / reps

    # PCIe-inclusive: the same reconstruction from host arrays (staged in and out by the library)
    hx0, hout = np.ones(image.shape, np.float32, order="F"), np.empty(image.shape, np.float32, order="F")
    hy, hb = np.asfortranarray(y), np.asfortranarray(b)
    host_ms = []
    for _ in range(7):  # (the first two calls run eagerly and capture the graph)
        t1 = time.perf_counter()
        L.check(lib.sr_osem(plan.handle, hout.ctypes.data, hx0.ctypes.data, hy.ctypes.data, hb.ctypes.data,
                            NBLOCKS, NITER, L.SR_MEM_HOST))
        host_ms.append(1e3 * (time.perf_counter() - t1))
    host_ms = host_ms[2:]

    # roofline of the dominant kernel over the timed region
    nsub = NVIEW // NBLOCKS
    g = (NX, NX, NZ, PSIZE)
    tot_ms, nl = timing[dom]
    # launches in OS-EM are one subset (nsub views) each
    by, fl = algorithmic(dom, nsub, g)
    avg_s = tot_ms / 1e3 / max(nl, 1)
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath):
        traffic = json.load(open(tpath)).get("kernels", {}).get(dom, {}).get("traffic_bytes")
    if fl > 0:
        roof = {"bound": "mfma", "achieved": fl / avg_s / 1e12, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s"}
    else:
        roof = {"bound": "hbm", "achieved": by / avg_s / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s"}
    roof["frac"] = roof["achieved"] / roof["peak"]
    roof["traffic"] = traffic
    roof["kernel"] = dom
    roof["avg_launch_ms"] = 1e3 * avg_s
    # per reconstruction, from the breakdown run
    kernels = {k: {"total_ms": round(v[0], 4), "launches": v[1],
                   "avg_ms": round(v[0] / max(v[1], 1), 5)} for k, v in breakdown.items()}

    # HBM (or FP32) fraction of every kernel class of a reconstruction, from
    # the event-timed one, plus the attenuation scan (plan-time work, timed
    # over a rebuild of the plan's factors: 4 launches of 32 views)
    plan.set_timing(True, classes=["atten"])
    plan.set_mumap(mumap)
    at_ms, at_n = plan.timing()["atten"]
    plan.set_timing(False)
    per_kernel = {}
    for cls, (ms, n) in list(breakdown.items()) + [("atten", (at_ms, at_n))]:
        if n == 0:
            continue
        nbv = 32 if cls == "atten" else nsub
        byk, flk = algorithmic(cls, nbv, g)
        avg = ms / 1e3 / n
        ent = {"avg_launch_ms": 1e3 * avg, "launches": n, "views_per_launch": nbv if cls != "img_reduce" else None,
               "algorithmic_bytes": byk, "hbm_GBs": byk / avg / 1e9, "hbm_frac": byk / avg / 1e9 / HBM_PEAK_GBS}
        if flk:
            ent.update({"fp32_TFLOPs": flk / avg / 1e12, "fp32_frac": flk / avg / 1e12 / FP32_PEAK_TFLOPS})
        per_kernel[cls] = ent

    cpu = errors = None
    if not args.no_cpu_baseline:
        log("cpu baseline and oracle errors ...")
        cpu, errors = cpu_baseline(image, mumap, psfs, np.ones(image.shape, np.float32), y, b, xhat)

    if args.copy_mib > 0:
        n = args.copy_mib << 20
        src, dst = S.DeviceArray((n // 4,)), S.DeviceArray((n // 4,))
        for _ in range(3):
            L.check(lib.sr_memcpy_d2d(dst.ptr, src.ptr, n))
        log(f"copied {args.copy_mib} MiB x 3 (counter calibration)")

    lat_ms = sorted(1e3 * t for t in lat)
    value = args.steps / t_total
    out = {
        "metric": "OS-EM recons/sec + p50/p99 latency (128²×80, 128 views); proj+backproj ms",
        "value": value,
        "unit": "recons/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": max(2, args.warmup),
        "ms_per_step": 1e3 * t_total / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded ellipsoid phantom, mu-map, Gaussian PSF, Poisson counts)",
        "config": {"workload": "OS-EM 8 subsets x 4 iterations, 128x128x80 image, 128 views over 360 deg, "
                               "37x37 depth-varying Gaussian PSF (FWHM 1.5->6 px), dy=0.48",
                   "nx": NX, "ny": NX, "nz": NZ, "nview": NVIEW, "psf": PSIZE, "nblocks": NBLOCKS,
                   "niter": NITER, "interpmeth": "two", "conv": info["conv"]},
        "latency_ms": {"p50": float(np.percentile(1e3 * np.array(lat_s), 50)),
                       "p99": float(np.percentile(1e3 * np.array(lat_s), 99)),
                       "min": 1e3 * min(lat_s), "max": 1e3 * max(lat_s), "samples": lat_n,
                       "note": "per-reconstruction latency (each returns after its stream finished), sampled "
                               "after the timed region"},
        "latency_timed_steps_ms": {"p50": float(np.percentile(lat_ms, 50)), "min": lat_ms[0], "max": lat_ms[-1]},
        "proj_ms": proj_ms,
        "backproj_ms": back_ms,
        "proj_backproj_ms": proj_ms + back_ms,
        "plan_ms": plan_ms,
        "runtime_init_ms": init_ms,
        "autotune_ms": tune_ms,
        "kernel_choices": plan.choice(),
        "autotune": tune,
        "osem_host_ms": float(np.median(host_ms)),
        "em_graph": dict(zip(("captures", "replays"), plan.graph_stats())),
        "kernels": kernels,
        "kernel_roofline": per_kernel,
        "errors": errors,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
