"""Benchmark: ptdirect on the 1.8M-triangle sphere-field scene at 1024x1024.

One step = `--spp-per-step` samples per pixel over the whole image (paths of
at most 32 segments, Russian roulette from depth 4), rendered into a batch
film that is folded into per-pixel moments, so the RMSE of the estimate is
measured from the batch variance.  The defaults (16 steps of 64 spp) render
the full 1024-spp image of this workload.  Prints one JSON line: samples/s,
rays/s, the estimated RMSE at 1024 spp, the time each engine and precision
needs to reach the RMSE of the FP64 estimator at 1024 spp against a
16,384-spp FP64 reference (the metric's fixed-RMSE part), the roofline of
the dominant kernel with its measured limiter, and the CPU FP64 oracle timed
on the host cores on a bounded sample of the same workload.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "samples/sec at fixed RMSE vs FP64 ref (ptdirect, 1M tri, 1024²); adjoint evals/s"
HBM_PEAK_GBS = 8000.0
TRI_BYTES = 48      # one leaf triangle record (3 x float4)
RAY_IO_BYTES = 48   # 32 B ray in + 16 B hit out
SHQ_BYTES = 96      # shadow-queue record of the persistent kernel: 48 B written, 48 B read back
FILM_BYTES = 48     # FP64 RGB read-modify-write of one contribution


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def _profile_value(key):
    """A `key: value` line of the committed rocprofv3 summary of the default
    workload (profiles/persistent_ptdirect_sphere_field_1024.md), or None."""
    path = os.path.join(ROOT, "profiles", "persistent_ptdirect_sphere_field_1024.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(key + ":"):
            return float(line.split(":", 1)[1])
    return None


def is_default_workload(a, engine, precision):
    import nanogi_amd as ngi
    return (engine == ngi.ENGINE_PERSISTENT and precision == ngi.PRECISION_FP32_FP64ACC and a.bvh == 8 and
            a.width == 1024 and a.height == 1024 and a.spp_per_step == 64 and a.detail == 8 and a.leaf == 0 and a.ct == 0)


def measured_traffic(a, engine, precision):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE x 2.0 +
    WRITE_SIZE x 1.0, calibrated on a 1 GiB copy) for the default workload;
    None for other configurations."""
    if not is_default_workload(a, engine, precision):
        return None
    v = _profile_value("hbm_bytes_per_launch")
    return int(v) if v is not None else None


def host_cpu_info():
    """CPU model and counts of the host (/proc/cpuinfo and the scheduler)."""
    model, sockets, cores = None, set(), set()
    phys = None
    try:
        for line in open("/proc/cpuinfo"):
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k == "model name" and model is None:
                model = v
            elif k == "physical id":
                phys = v
                sockets.add(v)
            elif k == "core id":
                cores.add((phys, v))
    except OSError:
        pass
    return {"model": model, "logical_cpus": os.cpu_count(), "physical_cores": len(cores) or None,
            "sockets": len(sockets) or None, "cpus_usable_by_this_process": len(os.sched_getaffinity(0))}


def pct(xs, q):
    import numpy as np
    return float(np.percentile(np.asarray(xs, dtype=np.float64), q))


def bench_batch(a):
    """Many small scenes in one persistent launch against one launch per scene.

    256 seeded Cornell-style rooms with boxes and icospheres (10-5,000
    triangles, log-uniform), square images of 64-512 pixels a side and
    16-512 spp (log-uniform), ptdirect, at most 16 segments, Russian roulette
    from depth 4.  Completion time of a scene = time from the start of the
    frame until its last path ended (device clock for the batch; host clock
    after each launch's synchronisation for the per-scene baseline)."""
    import numpy as np

    import nanogi_amd as ngi
    rng = np.random.default_rng(2024)
    lu = lambda lo, hi: float(np.exp(rng.uniform(np.log(lo), np.log(hi))))
    scenes, params = [], []
    t0 = time.perf_counter()
    for i in range(a.batch_scenes):
        tris = int(round(lu(10, 5000)))
        side = int(round(lu(64, 512)))
        spp = int(round(lu(16, 512)))
        s = ngi.Scene.generate(ngi.GEN_BATCH_ROOM, 1000 + i, 1.0, max(tris, 12))
        s.build(8)
        scenes.append(s)
        p = ngi.default_params(renderer="ptdirect", width=side, height=side, sampling=ngi.SAMPLING_PER_PIXEL, spp=spp,
                               max_num_vertices=17, rr_depth=4, rr_prob=0.5, seed=77, engine=ngi.ENGINE_PERSISTENT)
        p.stream_id = i
        params.append(p)
    gen_s = time.perf_counter() - t0
    total = sum(p.width * p.height * p.spp for p in params)
    tris = [sc.num_triangles for sc in scenes]
    log(f"batch: {len(scenes)} scenes, {total} paths, {min(tris)}-{max(tris)} triangles, generated in {gen_s:.2f} s")

    t0 = time.perf_counter()
    b = ngi.Batch(scenes, params)
    create_s = time.perf_counter() - t0
    cfg = b.config()
    log(f"batch: autotuned block {cfg['block_size']}, {cfg['blocks_per_cu']} blocks/CU in {cfg['tune_ms'] / 1e3:.2f} s")
    b.render(0)   # warm-up frame
    b.sync()
    frames = 2
    t0 = time.perf_counter()
    for f in range(1, frames + 1):
        b.render(f)
    b.sync()
    batch_s = (time.perf_counter() - t0) / frames
    ms, st = b.times()   # completion times of the last frame
    batch_rate = total / batch_s
    log(f"batch: {batch_rate / 1e6:.1f} M samples/s, frame {batch_s * 1e3:.1f} ms")

    # Baseline: one persistent launch per scene, in the same order.
    sessions = [ngi.Session(sc, p) for sc, p in zip(scenes, params)]
    for se, p in zip(sessions[:4], params[:4]):
        se.render(p.spp)
        se.sync()
    for se in sessions:
        se.reset()
    done = []
    t0 = time.perf_counter()
    for se, p in zip(sessions, params):
        se.render(p.spp)
        se.sync()
        done.append((time.perf_counter() - t0) * 1e3)
    single_s = done[-1] / 1e3
    for se in sessions:
        se.close()
    b.close()
    log(f"one launch per scene: {total / single_s / 1e6:.1f} M samples/s, {single_s * 1e3:.1f} ms")
    return {
        "workload": f"{len(scenes)} Cornell-style rooms, {min(tris)}-{max(tris)} triangles, 64-512 px square, 16-512 spp "
                    f"(log-uniform, seed 2024), ptdirect, max 16 segments, RR from depth 4",
        "paths": total, "samples_per_s": round(batch_rate, 1), "frame_ms": round(batch_s * 1e3, 3),
        "rays_per_s": round((st["extension_rays"] + st["shadow_rays"]) / batch_s, 1),
        "completion_ms": {"p50": round(pct(ms, 50), 3), "p99": round(pct(ms, 99), 3), "max": round(float(np.max(ms)), 3)},
        "one_launch_per_scene": {"samples_per_s": round(total / single_s, 1), "total_ms": round(single_s * 1e3, 3),
                                 "completion_ms": {"p50": round(pct(done, 50), 3), "p99": round(pct(done, 99), 3)}},
        "speedup": round(single_s / batch_s, 3),
        "launch": {"graph": bool(cfg["use_graph"]), "block_size": cfg["block_size"], "blocks_per_cu": cfg["blocks_per_cu"],
                   "autotune_ms": round(cfg["tune_ms"], 1), "create_ms": round(create_s * 1e3, 1)},
    }


def bench_adjoint(a):
    """Inverse rendering of the ~77k-triangle nave: 6 albedos x 3 + 4 light
    radiances x 3 from a 512x512 target (ptdirect, 64 spp forward + 64 spp
    replayed adjoint with a decorrelated seed, L2 loss, Adam)."""
    import numpy as np

    import nanogi_amd as ngi
    W = H = 512
    scene = ngi.Scene.generate(ngi.GEN_NAVE, 1, 1.0)
    scene.build(8)
    pr = scene.desc()["primitives"]
    is_r = (pr["type"] & ngi.PRIM_D) > 0
    is_l = (pr["type"] & ngi.PRIM_L) > 0
    R_true = np.where(is_r[:, None], pr["D_R"], 0.0)
    L_true = pr["Le"].copy()

    def params(seed, spp=64):
        return ngi.default_params(renderer="ptdirect", width=W, height=H, sampling=ngi.SAMPLING_PER_PIXEL, spp=spp,
                                  max_num_vertices=17, rr_depth=4, rr_prob=0.5, seed=seed, engine=ngi.ENGINE_PERSISTENT)

    target, _ = ngi.render(scene, params(999, 256))
    R = np.where(is_r[:, None], np.full(3, 0.5), 0.0)
    L = np.where(is_l[:, None], np.full(3, 10.0), 0.0)

    def loss_grad(R, L, seed, adj_seed):
        scene.set_params(R, L)
        img, _ = ngi.render(scene, params(seed))
        d = img - target
        g = 2.0 * d / d.size
        gR, gL, _ = ngi.render_adjoint(scene, params(adj_seed), g)
        return float(np.mean(d * d)), gR, gL

    # Gradient check at the start point: adjoint on the forward's own paths
    # against central differences with common random numbers (h = 1e-3).
    scene.set_params(R, L)
    img0, _ = ngi.render(scene, params(5))
    g0 = 2.0 * (img0 - target) / img0.size
    gR0, gL0, _ = ngi.render_adjoint(scene, params(5), g0)
    fd_err = []
    checks = [("R", int(np.nonzero(is_r)[0][k]), c) for k, c in ((0, 0), (2, 1), (4, 2))] + \
             [("L", int(np.nonzero(is_l)[0][k]), c) for k, c in ((0, 0), (3, 2))]
    for which, m, c in checks:
        vals = []
        for sgn in (1, -1):
            R2, L2 = R.copy(), L.copy()
            (R2 if which == "R" else L2)[m, c] += sgn * 1e-3
            scene.set_params(R2, L2)
            vals.append(float(np.sum(g0 * ngi.render(scene, params(5))[0])))
        fd = (vals[0] - vals[1]) / 2e-3
        an = (gR0 if which == "R" else gL0)[m, c]
        fd_err.append(abs(an - fd) / max(abs(fd), 1e-12))

    # Cost of one forward and one adjoint pass (synchronous, same workload).
    scene.set_params(R, L)
    reps = 5
    t0 = time.perf_counter()
    for k in range(reps):
        ngi.render(scene, params(100 + k))
    fwd_s = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for k in range(reps):
        ngi.render_adjoint(scene, params(200 + k), g0)
    adj_s = (time.perf_counter() - t0) / reps

    # Adam over albedos and radiances.
    mR, vR, mL, vL = (np.zeros_like(R) for _ in range(4))
    lrR, lrL, b1, b2, eps = 0.01, 0.2, 0.9, 0.999, 1e-8
    losses = []
    t0 = time.perf_counter()
    for it in range(a.adjoint_steps):
        loss, gR, gL = loss_grad(R, L, 10_000 + 2 * it, 10_001 + 2 * it)
        losses.append(loss)
        gR = np.where(is_r[:, None], gR, 0.0)
        gL = np.where(is_l[:, None], gL, 0.0)
        for x, g, m, v, lr in ((R, gR, mR, vR, lrR), (L, gL, mL, vL, lrL)):
            m *= b1
            m += (1 - b1) * g
            v *= b2
            v += (1 - b2) * g * g
            x -= lr * (m / (1 - b1 ** (it + 1))) / (np.sqrt(v / (1 - b2 ** (it + 1))) + eps)
        np.clip(R, 0.0, 1.0, out=R)
        np.clip(L, 0.0, None, out=L)
        if (it + 1) % 50 == 0:
            log(f"adjoint: Adam step {it + 1}/{a.adjoint_steps}, loss {loss:.4g}")
    opt_s = time.perf_counter() - t0
    scene.set_params(R_true, L_true)
    rel = lambda x, y, m: float(np.linalg.norm((x - y)[m]) / np.linalg.norm(y[m]))
    return {
        "workload": f"nave ({scene.num_triangles} triangles), 512x512, ptdirect 64 spp forward + 64 spp adjoint "
                    f"(decorrelated seeds), max 16 segments, RR from depth 4; 6x3 albedo + 4x3 radiance parameters",
        "evals_per_s": round(a.adjoint_steps / opt_s, 3) if a.adjoint_steps else None,
        "forward_ms": round(fwd_s * 1e3, 3), "adjoint_ms": round(adj_s * 1e3, 3),
        "adjoint_over_forward": round(adj_s / fwd_s, 3),
        "fd_rel_err_max": float(max(fd_err)), "fd_checks": len(fd_err),
        "adam_steps": a.adjoint_steps,
        "loss_first": losses[0] if losses else None, "loss_last": losses[-1] if losses else None,
        "param_rel_err": {"albedo": rel(R, R_true, is_r), "radiance": rel(L, L_true, is_l)},
        "param_rel_err_start": {"albedo": rel(np.where(is_r[:, None], np.full(3, 0.5), 0.0), R_true, is_r),
                                "radiance": rel(np.where(is_l[:, None], np.full(3, 10.0), 0.0), L_true, is_l)},
    }


def rmse_study(a, scene, params, ngi):
    """RMSE of each estimator variant at `--rmse-spp` spp against an FP64
    reference at `--rmse-ref-spp` spp (FP64 shading, persistent engine), a
    calibrated z-test against it, and the time each variant needs to reach
    the FP64 estimator's RMSE (time scales with spp, error with 1/sqrt(spp)).

    RMSE_target = the RMSE, net of the reference's own noise, of the FP64
    estimator (persistent engine, FP64 shading) at `--rmse-spp` spp.  The
    plain megakernel is the one-path-per-lane kernel with FP64 shading, on
    the binary BVH and on the wide one."""
    import numpy as np
    W, H = a.width, a.height
    batch = 16   # spp per batch: the variants' variance comes from rmse_spp / 16 batch means

    def run(engine, precision, spp, seed, width=8):
        if scene.bvh_info()["width"] != width:
            scene.build(width)
        se = ngi.Session(scene, params(engine=engine, precision=precision, seed=seed, spp=batch))
        se.render_batch(batch)
        se.sync()
        se.reset()
        se.sync()
        t0 = time.perf_counter()
        for k in range(spp // batch):
            se.render_batch(batch)
            if (k + 1) % 64 == 0:
                log(f"rmse study: {(k + 1) * batch}/{spp} spp")
        se.sync()
        dt = time.perf_counter() - t0
        film, var = se.film(variance=True)
        se.close()
        return film, var, dt

    log(f"rmse study: FP64 reference at {a.rmse_ref_spp} spp")
    ref, ref_var, ref_s = run(ngi.ENGINE_PERSISTENT, ngi.PRECISION_FP64_ALL, a.rmse_ref_spp, 4242)
    variants = [("persistent_fp32_fp64acc", ngi.ENGINE_PERSISTENT, ngi.PRECISION_FP32_FP64ACC, 8),
                ("persistent_fp32_all", ngi.ENGINE_PERSISTENT, ngi.PRECISION_FP32_ALL, 8),
                ("wavefront_fp32_fp64acc", ngi.ENGINE_WAVEFRONT, ngi.PRECISION_FP32_FP64ACC, 8),
                ("wavefront_fp32_all", ngi.ENGINE_WAVEFRONT, ngi.PRECISION_FP32_ALL, 8),
                ("megakernel_fp32_fp64acc", ngi.ENGINE_MEGAKERNEL, ngi.PRECISION_FP32_FP64ACC, 8),
                ("persistent_fp64", ngi.ENGINE_PERSISTENT, ngi.PRECISION_FP64_ALL, 8),
                ("megakernel_fp64", ngi.ENGINE_MEGAKERNEL, ngi.PRECISION_FP64_ALL, 8),
                ("megakernel_bvh2_fp64", ngi.ENGINE_MEGAKERNEL, ngi.PRECISION_FP64_ALL, 2)]
    out = {"reference": {"spp": a.rmse_ref_spp, "precision": "fp64 shading, fp32 traversal", "seconds": round(ref_s, 2),
                         "rmse_noise": float(np.sqrt(np.mean(ref_var)))}, "spp": a.rmse_spp, "variants": {}}
    bs = 8
    blk = lambda x: x[:H // bs * bs, :W // bs * bs].reshape(H // bs, bs, W // bs, bs, 3).mean(axis=(1, 3))
    for name, eng, prec, width in variants:
        film, var, dt = run(eng, prec, a.rmse_spp, 17, width)
        err = film - ref
        rmse = float(np.sqrt(np.mean(err * err)))
        # Remove the reference's own noise from the measured error.
        rmse_vs_truth = float(np.sqrt(max(np.mean(err * err) - np.mean(ref_var), 0.0)))
        z = err / np.sqrt(var + ref_var + 1e-300)
        ok = (var + ref_var) > 0
        # Calibrated tests.  Block means over 8x8 pixels (single pixels are
        # heavy-tailed: a few paths reach the small lights); pixel errors are
        # independent, so block variances are mean variances / 64.  Under
        # the null the block z-scores are close to N(0, 1) and sum(z^2) is
        # chi-square with n degrees of freedom (reported as (chi2 - n) /
        # sqrt(2 n) standard deviations).  A bias shared by the image shows
        # in the z-score of the image mean, sum(err) / sqrt(sum(variances)),
        # which does not suffer from the noise of per-block variance
        # estimates.
        zb = (blk(err) / np.sqrt((blk(var) + blk(ref_var)) / (bs * bs) + 1e-300)).ravel()
        nb = zb.size
        chi2 = float(np.sum(zb * zb))
        z_img = float(np.sum(err) / np.sqrt(np.sum(var + ref_var) + 1e-300))
        from scipy import stats as _st
        dof = max(a.rmse_spp // batch - 1, 1)
        out["variants"][name] = {"rmse": rmse, "rmse_minus_ref_noise": rmse_vs_truth,
                                 "rmse_from_variance": float(np.sqrt(np.mean(var))),
                                 "seconds": round(dt, 3), "samples_per_s": round(a.rmse_spp * W * H / dt, 1),
                                 "bvh_width": width,
                                 "z_gt3_frac": float(np.mean(np.abs(z[ok]) > 3.0)),
                                 "z_gt3_expected_t": float(2 * _st.t.sf(3.0, dof)), "batches": a.rmse_spp // batch,
                                 "block_z": {"blocks": nb, "chi2_per_dof": round(chi2 / nb, 4),
                                             "chi2_sigma": round((chi2 - nb) / math.sqrt(2 * nb), 2),
                                             "image_mean_z": round(z_img, 3),
                                             "gt3_frac": float(np.mean(np.abs(zb) > 3.0)), "gt3_expected": 0.0027},
                                 "mean_bias": float(np.mean(err)), "mean_rel_bias": float(np.mean(err) / np.mean(ref))}
        log(f"rmse study: {name}: rmse {rmse:.5g} in {dt:.2f} s, block chi2/dof {chi2 / nb:.3f}")
    target = out["variants"]["persistent_fp64"]["rmse_minus_ref_noise"]
    out["rmse_target"] = target
    for name, v in out["variants"].items():
        e = max(v["rmse_minus_ref_noise"], 1e-30)
        v["time_to_target_s"] = round(v["seconds"] * (e / target) ** 2, 3)
    for base_name in ("megakernel_fp64", "megakernel_bvh2_fp64"):
        base = out["variants"][base_name]["time_to_target_s"]
        for name, v in out["variants"].items():
            v["speedup_vs_" + base_name] = round(base / v["time_to_target_s"], 3)
    if scene.bvh_info()["width"] != a.bvh:
        scene.build(a.bvh)
    return out


def bench_renderers(a):
    """Every renderer on the 36-triangle Cornell box at 256x256, 64 spp, at
    most 8 segments, no Russian roulette (the light tracers shoot the same
    number of light paths): GPU samples/s on the persistent engine, and the
    FP64 CPU oracle on a bounded sample of the same paths."""
    import numpy as np

    import nanogi_amd as ngi
    import oracle
    scene = ngi.Scene.generate(ngi.GEN_CORNELL, 1, 1.0)
    scene.build(8)
    W = H = 256
    out = {"workload": "Cornell box (36 triangles), 256x256, 64 spp, max 8 segments, no Russian roulette"}
    ol = oracle.LoadedScene(scene)
    threads = max(1, min(a.cpu_threads, len(os.sched_getaffinity(0))))
    for r in ("pt", "ptdirect", "lt", "ltdirect", "bdpt"):
        p = ngi.default_params(renderer=r, width=W, height=H, sampling=ngi.SAMPLING_PER_PIXEL, spp=64, max_num_vertices=9,
                               rr_depth=-1, seed=1, engine=ngi.ENGINE_PERSISTENT)
        ngi.render(scene, p)   # warm-up (code load, allocations)
        t0 = time.perf_counter()
        reps = 3
        for _ in range(reps):
            _, st = ngi.render(scene, p)
        dt = (time.perf_counter() - t0) / reps
        entry = {"gpu_samples_per_s": round(W * H * 64 / dt, 1), "gpu_rays_per_s": round((st["extension_rays"] + st["shadow_rays"]) / dt, 1)}
        if a.cpu_seconds > 0:
            n = 65536
            t0 = time.perf_counter()
            ol.render_paths(p, 0, n, threads)
            cdt = time.perf_counter() - t0
            entry["cpu_samples_per_s"] = round(n / cdt, 1)
            entry["cpu_threads"] = threads
        out[r] = entry
        log(f"{r}: {entry}")
    return out


def bench_nave(a):
    """The nave interior (about 77k triangles) with ptdirect at 1024x1024,
    `--nave-spp` samples per pixel, at most 16 segments, roulette from depth
    4: one render per engine (persistent, the plain megakernel, wavefront),
    plus the FP64-shading megakernel and bdpt on 16-spp samples and the
    light tracer ltdirect on 64 light paths per pixel."""
    import nanogi_amd as ngi
    scene = ngi.Scene.generate(ngi.GEN_NAVE, 1, 1.0)
    scene.build(8)
    W = H = 1024
    spp = a.nave_spp
    out = {"workload": f"ptdirect (and bdpt at 16 spp, ltdirect at 64 light paths per pixel), nave ({scene.num_triangles} triangles), "
                       f"{W}x{H}, {spp} spp, max 16 segments, "
                       f"RR from depth 4", "engines": {}}
    variants = [("persistent", "ptdirect", ngi.ENGINE_PERSISTENT, ngi.PRECISION_FP32_FP64ACC, spp),
                ("megakernel", "ptdirect", ngi.ENGINE_MEGAKERNEL, ngi.PRECISION_FP32_FP64ACC, spp),
                ("wavefront", "ptdirect", ngi.ENGINE_WAVEFRONT, ngi.PRECISION_FP32_FP64ACC, spp),
                ("megakernel_fp64", "ptdirect", ngi.ENGINE_MEGAKERNEL, ngi.PRECISION_FP64_ALL, min(spp, 16)),
                ("bdpt", "bdpt", ngi.ENGINE_PERSISTENT, ngi.PRECISION_FP32_FP64ACC, min(spp, 16)),
                ("ltdirect", "ltdirect", ngi.ENGINE_PERSISTENT, ngi.PRECISION_FP32_FP64ACC, min(spp, 64))]
    for name, renderer, eng, prec, n in variants:
        p = ngi.default_params(renderer=renderer, width=W, height=H, sampling=ngi.SAMPLING_PER_PIXEL, spp=n,
                               max_num_vertices=17, rr_depth=4, seed=1, engine=eng, precision=prec)
        s = ngi.Session(scene, p)
        s.render(1)
        s.sync()
        s.reset()
        t0 = time.perf_counter()
        s.render(n)
        s.sync()
        dt = time.perf_counter() - t0
        st = s.stats()
        s.close()
        out["engines"][name] = {"spp": n, "seconds": round(dt, 3), "samples_per_s": round(W * H * n / dt, 1),
                                "rays_per_s": round((st["extension_rays"] + st["shadow_rays"]) / dt, 1)}
        log(f"nave {name}: {out['engines'][name]}")
    return out


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=16)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--spp-per-step", type=int, default=64)
    ap.add_argument("--width", type=int, default=1024)
    ap.add_argument("--height", type=int, default=1024)
    ap.add_argument("--engine", choices=["wavefront", "megakernel", "persistent"], default="persistent")
    ap.add_argument("--compare-steps", type=int, default=4,
                    help="steps for the other engines reported beside the main one (0 = skip)")
    ap.add_argument("--precision", choices=["fp32_fp64acc", "fp32_all", "fp64_all"], default="fp32_fp64acc")
    ap.add_argument("--bvh", type=int, choices=[2, 8], default=8)
    ap.add_argument("--detail", type=int, default=8, help="icosphere subdivision levels")
    ap.add_argument("--leaf", type=int, default=0, help="BVH maximum leaf size (0 = builder default)")
    ap.add_argument("--ct", type=float, default=0.0, help="BVH SAH node-visit cost (0 = builder default)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="CPU oracle budget (0 = skip)")
    ap.add_argument("--cpu-threads", type=int, default=16,
                    help="CPU oracle threads: at most the CPUs this process may run on (the GPU box grants 16 per GPU)")
    ap.add_argument("--no-counters", action="store_true")
    ap.add_argument("--paths-in-flight", type=int, default=0,
                    help="wavefront queue capacity or persistent lanes (0 = auto)")
    ap.add_argument("--batch-scenes", type=int, default=256,
                    help="scenes of the batched many-scene workload reported beside the main line (0 = skip)")
    ap.add_argument("--adjoint-steps", type=int, default=200,
                    help="Adam steps of the inverse-rendering workload reported beside the main line (0 = skip)")
    ap.add_argument("--autotune-paths", type=int, default=1 << 24,
                    help="paths per candidate of the single-scene launch tuner on this workload (0 = skip)")
    ap.add_argument("--nave-spp", type=int, default=256,
                    help="spp of the nave (ptdirect, 1024^2) engine comparison (0 = skip)")
    ap.add_argument("--renderers", type=int, default=1, help="time every renderer on the Cornell box (0 = skip)")
    ap.add_argument("--rmse-ref-spp", type=int, default=16384,
                    help="render an FP64 reference at this spp and report each variant's RMSE against it (0 = skip)")
    ap.add_argument("--rmse-spp", type=int, default=1024, help="spp of each variant in the RMSE study")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only copy this many bytes device-to-device (calibrates rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    return ap.parse_args()


def main():
    a = parse()
    if a.copy_probe:
        import torch
        n = a.copy_probe // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": n * 4}))
        return 0
    if a.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {a.gpus} is not supported (multi-GPU sharding is a design "
              f"note, not part of this build)", file=sys.stderr)
        return 2
    import numpy as np

    import nanogi_amd as ngi
    if ngi.device_count() <= 0:
        print("no HIP device", file=sys.stderr)
        return 1
    W, H = a.width, a.height
    npix = W * H
    t0 = time.perf_counter()
    scene = ngi.Scene.generate(ngi.GEN_SPHERE_FIELD, 1, W / H, a.detail)
    gen_s = time.perf_counter() - t0
    info = scene.build(a.bvh, max_leaf_size=a.leaf, traversal_cost=a.ct)
    ntris = scene.num_triangles
    log(f"scene: {ntris} triangles generated in {gen_s:.2f} s; BVH{a.bvh} {info['num_nodes']} nodes, "
        f"depth {info['max_depth']}, built in {info['build_ms'] / 1e3:.2f} s")
    engine = {"wavefront": ngi.ENGINE_WAVEFRONT, "megakernel": ngi.ENGINE_MEGAKERNEL, "persistent": ngi.ENGINE_PERSISTENT}[a.engine]
    precision = {"fp32_fp64acc": ngi.PRECISION_FP32_FP64ACC, "fp32_all": ngi.PRECISION_FP32_ALL,
                 "fp64_all": ngi.PRECISION_FP64_ALL}[a.precision]

    def params(**kw):
        base = dict(renderer="ptdirect", width=W, height=H, sampling=ngi.SAMPLING_PER_PIXEL, spp=a.spp_per_step,
                    max_num_vertices=33, rr_depth=4, rr_prob=0.5, seed=1, precision=precision, engine=engine,
                    paths_in_flight=a.paths_in_flight)
        base.update(kw)
        return ngi.default_params(**base)

    # Algorithmic traversal work per ray, counted by the traversal itself in
    # a separate (untimed) one-spp pass over the same image.
    counts = None
    if not a.no_counters:
        cs = ngi.Session(scene, params(collect_counters=1, seed=7))
        cs.render(1)
        st = cs.stats()
        cs.close()
        counts = {
            "ext_nodes_per_ray": st["node_visits"] / max(1, st["extension_rays"]),
            "ext_tris_per_ray": st["tri_tests"] / max(1, st["extension_rays"]),
            "shadow_nodes_per_ray": st["shadow_node_visits"] / max(1, st["shadow_rays"]),
            "shadow_tris_per_ray": st["shadow_tri_tests"] / max(1, st["shadow_rays"]),
            "ext_rays_per_path": st["extension_rays"] / max(1, st["samples"]),
            "shadow_rays_per_path": st["shadow_rays"] / max(1, st["samples"]),
        }
        log(f"counters: {json.dumps({k: round(v, 3) for k, v in counts.items()})}")

    sess = ngi.Session(scene, params(time_kernels=1))
    for _ in range(a.warmup):
        sess.render_batch(a.spp_per_step)
    sess.sync()
    sess.reset()
    sess.sync()
    t0 = time.perf_counter()
    for k in range(a.steps):
        sess.render_batch(a.spp_per_step)
        if k == 0 or (k + 1) % 8 == 0:
            log(f"step {k + 1}/{a.steps} enqueued")
    sess.sync()
    elapsed = time.perf_counter() - t0
    st = sess.stats()
    # The batch variance needs two batches (steps).
    var = sess.film(variance=True)[1] if a.steps >= 2 else None
    samples = a.steps * a.spp_per_step * npix
    value = samples / elapsed
    rays = st["extension_rays"] + st["shadow_rays"]
    spp_done = a.steps * a.spp_per_step
    rmse_now = float(np.sqrt(np.mean(var))) if var is not None else None
    rmse_1024 = rmse_now * math.sqrt(spp_done / 1024.0) if rmse_now is not None else None
    log(f"timed: {samples} paths in {elapsed:.3f} s -> {value / 1e6:.2f} M samples/s, {rays / elapsed / 1e9:.3f} G rays/s")

    # Roofline of the dominant kernel.  The unit is one ray query; the
    # algorithmic bytes follow the traversal's own node-visit and
    # triangle-test counts (128 B per wide node, 48 B per triangle), plus
    # the memory the kernel's design moves per query.
    node_bytes = info["node_bytes"]
    roof = None
    if counts is not None:
        er, sr = counts["ext_rays_per_path"], counts["shadow_rays_per_path"]
        trav_per_path = (er * (node_bytes * counts["ext_nodes_per_ray"] + TRI_BYTES * counts["ext_tris_per_ray"]) +
                         sr * (node_bytes * counts["shadow_nodes_per_ray"] + TRI_BYTES * counts["shadow_tris_per_ray"]))
        if engine in (ngi.ENGINE_PERSISTENT, ngi.ENGINE_MEGAKERNEL):
            # Path and ray state live in registers: no ray I/O or path-state
            # bytes.  The persistent FP32 ptdirect kernel queues its shadow
            # rays in HBM (one 48-B record written, read once); every traced
            # shadow ray may add one FP64 RGB contribution (an upper bound:
            # occluded ones add nothing).
            shq = SHQ_BYTES * sr if (engine == ngi.ENGINE_PERSISTENT and precision != ngi.PRECISION_FP64_ALL) else 0.0
            per_path = trav_per_path + shq + FILM_BYTES * sr
            if engine == ngi.ENGINE_PERSISTENT:
                avg_ms = st["ms_shade"] / a.steps          # one launch per step
                paths_per_launch = a.spp_per_step * npix
                kernel = "persistent path kernel (path regeneration, resumable traversals)"
            else:
                mk_launches = max(1, math.ceil(samples / max(1, (1 << 22) // npix * npix)))
                avg_ms = st["ms_shade"] / mk_launches
                paths_per_launch = samples / mk_launches
                kernel = "megakernel (one path per lane)"
            units = paths_per_launch * (er + sr)
            bytes_per_launch = per_path * paths_per_launch
            excluded = "ray in/out and path state (registers in this kernel)"
        else:
            # Wavefront extension stage: rays in and hits out through HBM.
            avg_ms = st["ms_extend"] / max(1, st["extend_launches"])
            units = st["extend_rays_timed"] / max(1, st["extend_launches"])
            bytes_per_launch = units * (RAY_IO_BYTES + node_bytes * counts["ext_nodes_per_ray"] +
                                        TRI_BYTES * counts["ext_tris_per_ray"])
            kernel = "extend (closest-hit traversal)"
            excluded = "none"
        achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
        traffic = measured_traffic(a, engine, precision)
        roof = {"bound": "hbm", "kernel": kernel, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "avg_launch_ms": round(avg_ms, 4), "algorithmic_bytes_per_launch": round(bytes_per_launch),
                "unit_of_work": "ray query (extension + shadow)", "units_per_launch": round(units),
                "bytes_per_unit": round(bytes_per_launch / max(units, 1.0), 2), "excluded_bytes": excluded}
        if traffic is not None:
            roof["traffic_frac"] = round(traffic / (avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
        if is_default_workload(a, engine, precision):
            # What limits the kernel (rocprofv3 --pmc on this workload): the
            # share of wave cycles spent waiting on dependencies and the
            # share of lanes active per VALU instruction.
            wait, valu = _profile_value("sq_wait_any_frac"), _profile_value("valu_utilization")
            if wait is not None and valu is not None:
                roof["limiter"] = "latency of dependent traversal loads and SIMD divergence, not HBM bandwidth"
                roof["sq_wait_any_frac"] = wait
                roof["valu_utilization"] = valu

    # The other engines on the same workload (short runs, same seeds).
    engines = {}
    if a.compare_steps > 0:
        variants = [("megakernel_bvh8_fp32", ngi.ENGINE_MEGAKERNEL, 8, ngi.PRECISION_FP32_FP64ACC),
                    ("wavefront_bvh8_fp32", ngi.ENGINE_WAVEFRONT, 8, ngi.PRECISION_FP32_FP64ACC),
                    ("persistent_bvh8_fp32", ngi.ENGINE_PERSISTENT, 8, ngi.PRECISION_FP32_FP64ACC),
                    ("persistent_bvh8_fp64", ngi.ENGINE_PERSISTENT, 8, ngi.PRECISION_FP64_ALL),
                    ("megakernel_bvh8_fp64", ngi.ENGINE_MEGAKERNEL, 8, ngi.PRECISION_FP64_ALL),
                    ("megakernel_bvh2_fp64", ngi.ENGINE_MEGAKERNEL, 2, ngi.PRECISION_FP64_ALL)]
        for name, eng, width, prec in variants:
            if eng == engine and width == a.bvh and prec == precision:
                engines[name] = round(value, 1)
                continue
            if scene.bvh_info()["width"] != width:
                scene.build(width)
            vs = ngi.Session(scene, params(engine=eng, precision=prec))
            vs.render(a.spp_per_step)
            vs.sync()
            t1 = time.perf_counter()
            for _ in range(a.compare_steps):
                vs.render(a.spp_per_step)
            vs.sync()
            engines[name] = round(a.compare_steps * a.spp_per_step * npix / (time.perf_counter() - t1), 1)
            vs.close()
            log(f"{name}: {engines[name] / 1e6:.1f} M samples/s")
        if scene.bvh_info()["width"] != a.bvh:
            scene.build(a.bvh)

    # Single-scene launch tuner on the same workload (reported; the timed
    # run above keeps the library defaults).
    tune = None
    if a.autotune_paths > 0:
        names = {ngi.ENGINE_MEGAKERNEL: "megakernel", ngi.ENGINE_WAVEFRONT: "wavefront", ngi.ENGINE_PERSISTENT: "persistent"}
        t1 = time.perf_counter()
        tuned, trials = ngi.autotune(scene, params(), a.autotune_paths)
        tune = {"tune_paths": a.autotune_paths, "tune_s": round(time.perf_counter() - t1, 3),
                "chosen": {"engine": names[tuned.engine], "paths_in_flight": tuned.paths_in_flight},
                "trials": [{"engine": names[t["engine"]], "paths_in_flight": t["paths_in_flight"],
                            "ms": round(t["ms"], 3), "samples_per_s": round(t["samples_per_s"], 1)} for t in trials]}
        log(f"autotune: {tune['chosen']} in {tune['tune_s']} s")

    # CPU baseline: the FP64 oracle on the host cores, bounded sample of paths.
    cpu = None
    if a.cpu_seconds > 0:
        try:
            import oracle
            threads = max(1, min(a.cpu_threads, len(os.sched_getaffinity(0))))
            ol = oracle.LoadedScene(scene)
            p = params()

            def timed(first, count, nthreads):
                t1 = time.perf_counter()
                _, cc = ol.render_paths(p, first, count, nthreads)
                return time.perf_counter() - t1, cc

            # Calibrate on growing prefixes until one takes >= 1 s.
            n, dt = 4096, 0.0
            while True:
                dt, _ = timed(0, n, threads)
                if dt >= 1.0 or n >= npix:
                    break
                n = min(npix, n * 4)
            n2 = int(min(max(n, n * a.cpu_seconds / max(dt, 1e-3)), 64 * npix))
            dt, cc = timed(n, n2, threads)
            # One thread on a sample of about a third of the budget.
            n1 = int(max(4096, min(n2, n2 / threads * a.cpu_seconds / 3.0 / max(dt, 1e-3))))
            dt1, cc1 = timed(n + n2, n1, 1)
            cpu = {"value": round(n2 / dt, 1), "unit": "samples/s", "cores": threads, "kind": "port",
                   "sample": f"{n2} path samples (path ids {n}..{n + n2 - 1}; pixel = id mod W*H, sample = id div W*H) "
                             f"of the same ptdirect workload; FP64 CPU restatement of the reference estimator with its own "
                             f"median-split BVH (not the reference binary, which needs Embree 2 and TBB), {threads} threads",
                   "rays_per_s": round(float(cc.sum()) / dt, 1),
                   "one_thread": {"value": round(n1 / dt1, 1), "rays_per_s": round(float(cc1.sum()) / dt1, 1),
                                  "sample": f"{n1} path samples (path ids {n + n2}..{n + n2 + n1 - 1})"},
                   "host": host_cpu_info()}
            log(f"cpu oracle: {n2} paths in {dt:.2f} s on {threads} threads; {n1} paths in {dt1:.2f} s on 1 thread")
        except Exception as e:  # the baseline is reported, never required
            log(f"cpu baseline skipped: {e}")

    study = None
    if a.rmse_ref_spp > 0:
        study = rmse_study(a, scene, params, ngi)
    renderers = bench_renderers(a) if a.renderers else None
    nave = bench_nave(a) if a.nave_spp > 0 else None
    batch = None
    if a.batch_scenes > 0:
        batch = bench_batch(a)
    adjoint = None
    if a.adjoint_steps > 0:
        adjoint = bench_adjoint(a)

    # The metric's fixed-RMSE part: time for the timed engine and precision
    # to reach RMSE_target, against the plain FP64 megakernels.
    fixed_rmse = None
    if study is not None:
        key = {(ngi.ENGINE_PERSISTENT, ngi.PRECISION_FP32_FP64ACC): "persistent_fp32_fp64acc",
               (ngi.ENGINE_PERSISTENT, ngi.PRECISION_FP32_ALL): "persistent_fp32_all",
               (ngi.ENGINE_PERSISTENT, ngi.PRECISION_FP64_ALL): "persistent_fp64",
               (ngi.ENGINE_WAVEFRONT, ngi.PRECISION_FP32_FP64ACC): "wavefront_fp32_fp64acc",
               (ngi.ENGINE_WAVEFRONT, ngi.PRECISION_FP32_ALL): "wavefront_fp32_all",
               (ngi.ENGINE_MEGAKERNEL, ngi.PRECISION_FP32_FP64ACC): "megakernel_fp32_fp64acc",
               (ngi.ENGINE_MEGAKERNEL, ngi.PRECISION_FP64_ALL): "megakernel_fp64"}.get((engine, precision))
        if key is not None:
            v = study["variants"][key]
            # At the timed run's rate (64 spp per launch), which the study's
            # 16-spp batches do not reach.
            t_main = v["time_to_target_s"] * v["samples_per_s"] / value
            fixed_rmse = {"variant": key, "rmse_target": study["rmse_target"],
                          "rmse_target_definition": f"RMSE of the FP64 estimator (persistent engine, FP64 shading) at "
                                                    f"{a.rmse_spp} spp against a {a.rmse_ref_spp}-spp FP64 reference, "
                                                    f"net of the reference's own noise",
                          "time_to_target_s": round(t_main, 3), "time_to_target_s_in_study": v["time_to_target_s"],
                          "speedup_vs_plain_megakernel_bvh2_fp64": round(study["variants"]["megakernel_bvh2_fp64"]["time_to_target_s"] / t_main, 3),
                          "speedup_vs_plain_megakernel_bvh8_fp64": round(study["variants"]["megakernel_fp64"]["time_to_target_s"] / t_main, 3),
                          "block_z": v["block_z"]}

    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "samples/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(elapsed * 1e3 / a.steps, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64" if precision == ngi.PRECISION_FP64_ALL else "f32",
        "precision": {"traversal": "f32", "shading": "f64" if precision == ngi.PRECISION_FP64_ALL else "f32",
                      "film": "f32" if precision == ngi.PRECISION_FP32_ALL else "f64"},
        "data": "synthetic (seeded procedural scene)",
        "config": {"workload": f"ptdirect, displaced icosphere ({a.detail} subdivisions) on 512^2 heightfield, "
                               f"{ntris} triangles, {W}x{H}, max 32 segments, RR from depth 4",
                   "engine": a.engine, "precision": a.precision, "bvh_width": a.bvh, "spp_per_step": a.spp_per_step,
                   "triangles": ntris},
        "rays_per_s": round(rays / elapsed, 1),
        "fixed_rmse": fixed_rmse,
        "rmse_estimate": {"spp": spp_done, "rmse": rmse_now, "rmse_at_1024spp": rmse_1024,
                          "time_to_1024spp_s": round(1024 * npix / value, 3)},
        "roofline": roof,
        "engines_samples_per_s": engines,
        "autotune": tune,
        "cpu_baseline": cpu,
        "counters": counts,
        "bvh_build_s": round(info["build_ms"] / 1e3, 3),
        "batch": batch,
        "adjoint": adjoint,
        "rmse_study": study,
        "renderers": renderers,
        "nave": nave,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
