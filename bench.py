#!/usr/bin/env python3
"""Render-path benchmark: Msamples/s of the MI355X wavefront path tracer.

Workload (default): 8 displaced icospheres (subdivision 7, 327,680 triangles
each) on a 2-triangle ground plane + a 2-triangle area light = 2,621,444
triangles; 1024x1024 pixels, 64 spp, depth 16, hero wavelengths (360..830 nm),
fp32.  One step = one full render of that image (67,108,864 camera samples)
with the scene and its BVH already resident in HBM.

Besides value/ms_per_step the JSON line carries:
  roofline     : the extension-ray traversal kernel (the dominant kernel):
                 algorithmic bytes (32 B ray + 16 B hit + 128 B per node
                 visited + 48 B per triangle tested, counted by the kernel's
                 own counters on the identical rays) / its HIP-event time.
  cpu_baseline : the CPU oracle (the reference integrator restated, 60-bin
                 spectra) on a bounded slice of the same image, host threads.
  cornell      : the small Cornell workload (256^2 x 16 spp, depth 8): GPU
                 Msamples/s, the oracle's Msamples/s and their ratio.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "filianore-core_amd"))

METRIC = "Msamples/s + relMSE at equal time (guided vs plain), 1 MI355X vs host CPU"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def traffic_from_profile():
    """HBM bytes per k_trace launch measured with rocprofv3 PMC counters
    (FETCH_SIZE + WRITE_SIZE, scaled by a copy-kernel calibration), recorded
    in profiles/traffic_k_trace_spheres.json; None until measured."""
    path = os.path.join(ROOT, "profiles", "traffic_k_trace_spheres.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("hbm_bytes_per_launch")


# The GPU box allots this many host CPUs to a process on its one GPU (worker
# pools are sized to it whatever the machine's total core count is).
HOST_CPU_SHARE = 16


def host_info():
    """Host CPUs: online count, the process's affinity mask, the CPU model,
    and the threads the CPU baseline uses (the affinity mask, at most the
    box's per-GPU share)."""
    online = os.cpu_count() or 1
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = online
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": online, "affinity": affinity, "model": model,
            "threads": max(1, min(HOST_CPU_SHARE, affinity))}


def host_threads():
    return host_info()["threads"]


def issue_from_profile():
    """What bounds k_trace besides bytes: VALU issue and lane utilisation,
    L2 hit rate, from rocprofv3 SQ/TCC counters recorded under profiles/;
    None until measured."""
    path = os.path.join(ROOT, "profiles", "issue_k_trace_spheres.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def cpu_baseline(gen, spp, depth, budget_s, threads):
    """Oracle (CPU restatement) on horizontal slices of the same image until
    about budget_s seconds of CPU work; returns Msamples/s."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    sc = oracle.Scene(gen)
    H = gen.camera.height
    rows = 8
    done_samples, secs, y, sweep = 0, 0.0, H // 2 - 64, 0
    while secs < budget_s:
        if y >= H:
            y, sweep = 0, sweep + 1
        _, _, st = sc.render(spp, depth, seed=17 + y + 7919 * sweep, threads=threads, rows=(y, min(H, y + rows)))
        done_samples += st.samples
        secs += st.seconds
        y += rows
        rows = min(64, rows * 2)
    sc.close()
    return done_samples / secs / 1e6, done_samples, secs


MFMA_BF16_PEAK_TFLOPS = 2500.0
GUIDE_FWD_FLOP = 2 * 18240          # algorithmic FLOP per network query (53*64 + 3*64*64 + 64*40 MAC)
GUIDE_TRAIN_FLOP = 3 * GUIDE_FWD_FLOP


# CIE XYZ -> linear sRGB (D65), the matrix of xyz_to_rgb
# (color/spectrumoperations.cpp:58-66)
XYZ_TO_SRGB = ((3.240479, -1.537150, -0.498535),
               (-0.969256, 1.875991, 0.041556),
               (0.055648, -0.204043, 1.057311))


def to_srgb(xyz):
    import numpy as np
    return xyz @ np.asarray(XYZ_TO_SRGB, np.float64).T


def intersection_section(flgpu, ctx, gen, sc, cpu_seconds, threads):
    """The 4M-ray intersection workload on the spheres scene: 2,097,152
    coherent primary rays (2048 x 1024 pixel centres) and 2,097,152
    cosine-diffuse secondaries from their hits. GPU Mrays/s with the rays
    resident (fl_intersect_device), the CPU BVH's Mrays/s on a bounded slice
    (host threads), hit IDs compared bit for bit on that slice, and the
    near-ties (best two distinct triangles within 1e-6 relative t) counted
    on it."""
    import numpy as np
    from raysets import primary_rays, secondary_rays
    prim = primary_rays(gen.camera, 2048, 1024)
    pt, ptri, _ = sc.intersect(prim)
    sec = secondary_rays(gen.triangles(), prim, pt, ptri, prim.shape[0])
    rays = np.concatenate([prim, sec])
    out = {"rays": int(rays.shape[0])}
    for name, sub_rays in (("all", rays), ("primary", prim), ("secondary", sec)):
        t, tri, uv, ms = sc.intersect_resident(sub_rays, reps=5)
        out["gpu_mrays_s_" + name] = round(sub_rays.shape[0] / ms / 1e3, 1)
        if name == "all":
            g_t, g_tri, g_uv = t, tri, uv
    if cpu_seconds > 0:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        ors = oracle.Scene(gen)
        # an interleaved slice of both halves, grown until the time budget
        n, secs, done, mism, ties, checked = 1 << 14, 0.0, 0, 0, 0, 0
        rng = np.random.default_rng(3)
        while secs < cpu_seconds and done < rays.shape[0]:
            idx = np.sort(rng.choice(rays.shape[0], min(n, rays.shape[0] - done), replace=False))
            ts = time.perf_counter()
            ot, otri, ouv = ors.intersect(rays[idx], threads=threads)
            secs += time.perf_counter() - ts
            done += idx.size
            mism += int(np.sum(otri != g_tri[idx]))
            both = (otri == g_tri[idx]) & (otri != flgpu.FL_NO_HIT)
            mism += int(np.sum(ot[both].view(np.uint32) != g_t[idx][both].view(np.uint32)))
            t2, _ = ors.intersect2(rays[idx], threads=threads)
            ok = np.isfinite(t2[:, 1]) & (t2[:, 0] > 0)
            gap = np.abs(np.where(ok, t2[:, 1], 0.0) - np.where(ok, t2[:, 0], 0.0))
            ties += int(np.sum(ok & (gap <= 1e-6 * t2[:, 0])))
            checked += idx.size
            n *= 2
        ors.close()
        out.update({"cpu_mrays_s": round(done / secs / 1e6, 2), "cpu_threads": threads,
                    "cpu_rays_timed": done, "checked_rays": checked, "hit_mismatches": mism,
                    "near_ties": ties,
                    "gpu_over_cpu": round(out["gpu_mrays_s_all"] / (done / secs / 1e6), 1)})
    return out


def batch_section(flgpu, ctx, n_scenes, tune):
    """64 scenes (log-uniform 1e3..2e6 triangles, 128^2..1024^2, 16 spp, depth
    12) in one fl_render_batch, against the same renders one by one."""
    t0 = time.time()
    gens = [flgpu.GenSceneData(flgpu.FL_GEN_BATCH, scene_index=k) for k in range(n_scenes)]
    t_gen = time.time() - t0
    scenes = [ctx.scene(g) for g in gens]
    try:
        tris = sum(g.n_triangles for g in gens)
        samples = sum(g.camera.width * g.camera.height * g.spp for g in gens)
        p = flgpu.default_params(spp=16, depth=12, seed=1, flags=flgpu.FL_RENDER_VNDF_PDF | flgpu.FL_RENDER_GRAPH)
        entries = [(sc, None, g.spp) for sc, g in zip(scenes, gens)]
        tuned = None
        if tune:
            n = len(entries)
            items = (flgpu.BatchItem * n)()
            for k, (sc, _, spp) in enumerate(entries):
                items[k].scene, items[k].camera, items[k].spp = sc.h, sc.gen.camera, spp
            r = flgpu.TuneResult()
            flgpu._check(ctx.L.fl_autotune_batch(ctx.h, items, n, flgpu.C.byref(p), flgpu.C.byref(r)), ctx.h,
                         "fl_autotune_batch")
            r.apply(p)
            tuned = {"max_paths": r.max_paths, "stack_depth": r.stack_depth, "trav_waves": r.trav_waves,
                     "trav_budget": r.trav_budget, "refill_lanes": r.refill_lanes, "shade_waves": r.shade_waves,
                     "sort_shading": r.sort_shading, "node_width": r.node_width, "trav_prefetch": r.trav_prefetch,
                     "ms_default": round(r.ms_default, 2), "ms_tuned": round(r.ms_best, 2), "probes": r.n_tried}
        ctx.render_batch(entries, p, want_film=False)
        ts = time.perf_counter()
        _, st = ctx.render_batch(entries, p, want_film=False)
        t_batch = time.perf_counter() - ts
        t_seq = 0.0
        for sc, g in zip(scenes, gens):
            q = flgpu.default_params(spp=g.spp, depth=12, seed=1, flags=p.flags)
            sc.render(q, want_film=False)
            ts = time.perf_counter()
            sc.render(q, want_film=False)
            t_seq += time.perf_counter() - ts
    finally:
        for sc in scenes:
            sc.close()
    return {"scenes": n_scenes, "triangles": tris, "samples": samples, "gen_s": round(t_gen, 1),
            "ms_batch": round(t_batch * 1e3, 2), "msamples_s": round(samples / t_batch / 1e6, 1),
            "ms_one_by_one": round(t_seq * 1e3, 2), "batch_speedup": round(t_seq / t_batch, 2),
            "passes": st.passes, "autotune": tuned}


NO_RR = 1 << 20   # rr_start beyond any depth: no Russian roulette


def guided_section(flgpu, ctx, config, res, ref_spp, spp=64):
    """Plain vs guided at equal wall time on a room lit only through a 5%
    ceiling slot. Each method runs with and without Russian roulette (the
    reference rule after bounce 3); the better setting of each (lower relMSE x
    time) is compared. Guided renders train the proposal online on their
    first 15% of samples, inside the timed render. relMSE is against a plain
    reference of ref_spp samples per pixel, rendered in 4096-spp parts, with
    the reference's own noise (its per-pixel variance of the mean, from the
    film's sums of squares) subtracted so that the figure estimates the error
    against the converged image."""
    import numpy as np
    gen = flgpu.GenSceneData(config, width=res, height=res)
    base = flgpu.FL_RENDER_VNDF_PDF | flgpu.FL_RENDER_GRAPH
    out = {"scene": {flgpu.FL_GEN_GUIDED: "attic slot", flgpu.FL_GEN_GUIDED_SKYLIGHT: "skylight strip"}[config],
           "resolution": res, "triangles": gen.n_triangles, "depth": gen.depth, "reference_spp": ref_spp}
    with ctx.scene(gen) as sc:
        ts = time.perf_counter()
        halves = [None, None]
        parts = max(2, ref_spp // 4096)
        for k in range(parts):
            p = flgpu.default_params(spp=ref_spp // parts, depth=gen.depth, seed=1000 + k, flags=base)
            s_, _, _ = sc.render(p)
            h = k * 2 // parts
            halves[h] = s_.astype(np.float64) if halves[h] is None else halves[h] + s_
            log("  reference part %d/%d (%.0f s)" % (k + 1, parts, time.perf_counter() - ts))
        # relMSE in linear sRGB against the mean of two independent halves;
        # the reference's own noise, the variance of that mean, is estimated
        # per pixel and channel without bias by (A - B)^2 / 4 and subtracted
        A = to_srgb(halves[0] / (ref_spp // 2))
        B = to_srgb(halves[1] / (ref_spp // 2))
        ref = 0.5 * (A + B)
        ref_var = 0.25 * (A - B) ** 2
        den = ref * ref + 0.01
        out["reference_s"] = round(time.perf_counter() - ts, 1)
        out["reference_noise_relmse"] = round(float(np.mean(ref_var / den)), 5)

        def relmse(img_xyz):
            img = to_srgb(img_xyz.astype(np.float64))
            return float(np.mean(((img - ref) ** 2 - ref_var) / den))

        def timed(p):
            t0 = time.perf_counter()
            img, _, st = sc.render(p)
            return time.perf_counter() - t0, img, st

        plain = {}
        for name, rr in (("rr3", 3), ("no_rr", NO_RR)):
            p = flgpu.default_params(spp=spp, depth=gen.depth, seed=7, flags=base)
            p.rr_start = rr
            sc.render(p, want_film=False)
            t, img, _ = timed(p)
            e = relmse(img / spp)
            plain[name] = {"spp": spp, "s": round(t, 4), "relmse": round(e, 5), "relmse_x_s": round(e * t, 5)}
        guided = {}
        mfma = None
        for name, rr in (("no_rr", NO_RR), ("rr3", 3)):
            pts = []
            for g_spp in ((8, 16, 32, spp) if name == "no_rr" else (spp,)):
                with ctx.guide(seed=1) as g:
                    p = flgpu.default_params(spp=g_spp, depth=gen.depth, seed=7,
                                             flags=base | flgpu.FL_RENDER_GUIDED | flgpu.FL_RENDER_TRAIN_GUIDE)
                    p.rr_start = rr
                    p.guide = g.h
                    t, img, st = timed(p)
                    e = relmse(img / g_spp)
                    pts.append({"spp": g_spp, "s": round(t, 4), "relmse": round(e, 5), "relmse_x_s": round(e * t, 5),
                                "train_ms": round(st.ms_guide, 1), "passes": st.passes})
                    if name == "no_rr" and g_spp == spp:
                        steps = g.stats()[0]
                        tp = flgpu.default_params(spp=spp, depth=gen.depth, seed=8,
                                                  flags=flgpu.FL_RENDER_VNDF_PDF | flgpu.FL_RENDER_GUIDED |
                                                  flgpu.FL_RENDER_TIMING)
                        tp.rr_start = rr
                        tp.guide = g.h
                        _, _, tst = sc.render(tp, want_film=False)
                        # the network runs inside the guided shading kernel:
                        # its FLOP over that kernel's whole time (shading
                        # included) is a lower bound on the matrix cores' rate
                        infer_tf = tst.guide_queries * GUIDE_FWD_FLOP / (tst.ms_shade * 1e-3) / 1e12
                        train_tf = steps * 16384 * GUIDE_TRAIN_FLOP / (st.ms_guide * 1e-3) / 1e12
                        mfma = {"infer_queries": int(tst.guide_queries), "shade_ms_with_infer": round(tst.ms_shade, 2),
                                "infer_tflops_over_shade": round(infer_tf, 2),
                                "infer_mfma_util_lower_bound": round(infer_tf / MFMA_BF16_PEAK_TFLOPS, 5),
                                "render_ms": round(tst.ms_total, 1), "train_steps": steps,
                                "train_ms": round(st.ms_guide, 1), "train_tflops": round(train_tf, 2),
                                "train_mfma_util": round(train_tf / MFMA_BF16_PEAK_TFLOPS, 5)}
            guided[name] = pts
    out["plain"] = plain
    out["guided"] = guided
    # the better setting of each method at the config's spp, by relMSE x time
    best_plain = min(plain, key=lambda k: plain[k]["relmse_x_s"])
    best_guided = min(guided, key=lambda k: guided[k][-1]["relmse_x_s"])
    target_t, target_e = plain[best_plain]["s"], plain[best_plain]["relmse"]
    # wall time the guided render needs to reach that relMSE: log-log
    # interpolation between guided points of the chosen setting
    need = None
    pts = sorted((q["s"], q["relmse"]) for q in guided[best_guided])
    for (t0, e0), (t1, e1) in zip(pts, pts[1:]):
        if e0 >= target_e >= e1 and e0 > 0 and e1 > 0 and e0 != e1:
            f = (np.log(target_e) - np.log(e0)) / (np.log(e1) - np.log(e0))
            need = float(np.exp(np.log(t0) + f * (np.log(t1) - np.log(t0))))
            break
    if need is None and pts and pts[0][1] <= target_e:
        need = pts[0][0]   # the cheapest guided point already beats it: an upper bound
    out["plain_setting"] = best_plain
    out["guided_setting"] = best_guided
    out["equal_spp_relmse_ratio"] = round(guided[best_guided][-1]["relmse"] / target_e, 4)
    out["time_ratio_to_plain_relmse"] = round(need / target_t, 3) if need else None
    out["mfma"] = mfma
    return out


def nee_one_section(flgpu, sc, gen, spp, ref_spp=2048):
    """Direct lighting as the reference does it, one sample per light
    (uniform_sample_all_lights, integrators/integrator.cpp:10-23), against one
    uniformly chosen light per vertex (FL_RENDER_NEE_ONE), on the same scene
    at equal spp: render time, relMSE (linear sRGB, against an all-lights
    reference of ref_spp with its own noise subtracted) and relMSE x time.
    The headline figure always uses the reference's estimator."""
    import numpy as np
    base = flgpu.FL_RENDER_VNDF_PDF | flgpu.FL_RENDER_GRAPH
    halves = []
    for k in range(2):
        p = flgpu.default_params(spp=ref_spp // 2, depth=gen.depth, mode=flgpu.FL_SPECTRAL_HERO, seed=2000 + k,
                                 flags=base)
        s_, _, _ = sc.render(p)
        halves.append(to_srgb(s_.astype(np.float64) / (ref_spp // 2)))
    ref = 0.5 * (halves[0] + halves[1])
    ref_var = 0.25 * (halves[0] - halves[1]) ** 2
    den = ref * ref + 0.01
    out = {"reference_spp": ref_spp, "spp": spp}
    for name, flags in (("all_lights", base), ("one_light", base | flgpu.FL_RENDER_NEE_ONE)):
        p = flgpu.default_params(spp=spp, depth=gen.depth, mode=flgpu.FL_SPECTRAL_HERO, seed=7, flags=flags)
        img, _, st = sc.render(p)
        e = float(np.mean(((to_srgb(img.astype(np.float64) / spp) - ref) ** 2 - ref_var) / den))
        ts = []
        for _ in range(3):
            t0 = time.perf_counter()
            sc.render(p, want_film=False)
            ts.append(time.perf_counter() - t0)
        t = sum(ts) / len(ts)
        out[name] = {"s": round(t, 4), "msamples_s": round(img.shape[0] * img.shape[1] * spp / t / 1e6, 1),
                     "shadow_rays": int(st.rays_shadow), "relmse": round(e, 6), "relmse_x_s": round(e * t, 7)}
    out["relmse_x_s_ratio"] = round(out["one_light"]["relmse_x_s"] / out["all_lights"]["relmse_x_s"], 3)
    return out


def terrain_section(flgpu, ctx, spp, res_div):
    """The full terrain config: 4096 x 2048 fBm heightfield (16.8 M
    triangles) + 200 Cauchy prisms + 24 GGX metals, 8 cameras at
    3840x2160 / res_div rendered as one guided batch (trained online)."""
    t0 = time.time()
    gen = flgpu.GenSceneData(flgpu.FL_GEN_TERRAIN)
    W, H = 3840 // res_div, 2160 // res_div
    cams = []
    for k in range(8):
        c = flgpu.GenSceneData(flgpu.FL_GEN_TERRAIN, grid_x=2, grid_y=2, camera_index=k, width=W, height=H).camera
        cams.append(c)
    t_gen = time.time() - t0
    out = {"triangles": gen.n_triangles, "cameras": 8, "width": W, "height": H, "spp": spp, "gen_s": round(t_gen, 1)}
    with ctx.scene(gen) as sc, ctx.guide(seed=1) as g:
        out["bvh_build_ms"] = round(sc.info().build_ms, 1)
        p = flgpu.default_params(spp=spp, depth=gen.depth, seed=1,
                                 flags=flgpu.FL_RENDER_VNDF_PDF | flgpu.FL_RENDER_GRAPH | flgpu.FL_RENDER_GUIDED |
                                 flgpu.FL_RENDER_TRAIN_GUIDE)
        p.guide = g.h
        ts = time.perf_counter()
        _, st = ctx.render_batch([(sc, c, spp) for c in cams], p, want_film=False)
        dt = time.perf_counter() - ts
        steps, loss, div = g.stats()
    samples = 8 * W * H * spp
    out.update({"s": round(dt, 2), "samples": samples, "msamples_s": round(samples / dt / 1e6, 1),
                "rays": int(st.rays_extension + st.rays_shadow), "passes": st.passes, "guide_steps": steps,
                "guide_train_ms": round(st.ms_guide, 1), "guide_diverged": div, "nan_samples": int(st.nan_samples)})
    return out


def profile_guided(flgpu, spp, config):
    """A short program for the profiler: on the guided room at 1024^2, one
    plain render (no roulette) and two guided renders (fresh network, trained
    online) of spp samples per pixel, with their times on stderr. No
    reference image and no JSON line."""
    gen = flgpu.GenSceneData(config, width=1024, height=1024)
    base = flgpu.FL_RENDER_VNDF_PDF | flgpu.FL_RENDER_GRAPH
    with flgpu.Context(0) as ctx, ctx.scene(gen) as sc:
        p = flgpu.default_params(spp=spp, depth=gen.depth, seed=7, flags=base)
        p.rr_start = NO_RR
        sc.render(p, want_film=False)
        ts = time.perf_counter()
        sc.render(p, want_film=False)
        log("plain %d spp: %.1f ms" % (spp, (time.perf_counter() - ts) * 1e3))
        for rep in range(2):
            with ctx.guide(seed=1) as g:
                q = flgpu.default_params(spp=spp, depth=gen.depth, seed=7,
                                         flags=base | flgpu.FL_RENDER_GUIDED | flgpu.FL_RENDER_TRAIN_GUIDE)
                q.rr_start = NO_RR
                q.guide = g.h
                ts = time.perf_counter()
                _, _, st = sc.render(q, want_film=False)
                log("guided %d spp: %.1f ms (training %.1f ms, %d passes)" % (
                    spp, (time.perf_counter() - ts) * 1e3, st.ms_guide, st.passes))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="spheres", choices=["spheres", "cornell", "guided"])
    ap.add_argument("--spp", type=int, default=0)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-cornell", action="store_true", help="skip the small Cornell section")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-scene batch section")
    ap.add_argument("--no-guided", action="store_true", help="skip the guided (learned proposal) section")
    ap.add_argument("--batch-scenes", type=int, default=64)
    ap.add_argument("--guided-res", type=int, default=1024)
    ap.add_argument("--guided-ref-spp", type=int, default=16384)
    ap.add_argument("--no-guided-skylight", action="store_true",
                    help="skip the guided comparison on the skylight reading of the room")
    ap.add_argument("--no-intersect", action="store_true", help="skip the 4M-ray intersection section")
    ap.add_argument("--no-nee-one", action="store_true",
                    help="skip the one-light-per-vertex direct lighting comparison")
    ap.add_argument("--terrain", action="store_true", help="also render the full terrain config (8 x 4K x 256 spp)")
    ap.add_argument("--terrain-spp", type=int, default=256)
    ap.add_argument("--terrain-res-div", type=int, default=1)
    ap.add_argument("--profile-guided", type=int, default=0, metavar="SPP",
                    help="only render the guided room plain and guided at SPP (a short program to profile)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run a plain 1 GiB -> 1 GiB copy kernel (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: measures one GPU only (a render would shard across GPUs by pixels and samples)",
              file=sys.stderr)
        sys.exit(2)

    import flgpu

    if args.profile_guided:
        profile_guided(flgpu, args.profile_guided, flgpu.FL_GEN_GUIDED)
        return
    cfg = {"spheres": flgpu.FL_GEN_SPHERES, "cornell": flgpu.FL_GEN_CORNELL, "guided": flgpu.FL_GEN_GUIDED}[args.workload]
    t0 = time.time()
    gen = flgpu.GenSceneData(cfg)
    spp = args.spp or gen.spp
    W, H = gen.camera.width, gen.camera.height
    log("scene: %d triangles, %dx%d, %d spp, depth %d (gen %.1fs)" % (gen.n_triangles, W, H, spp, gen.depth,
                                                                      time.time() - t0))
    out = {}
    with flgpu.Context(0) as ctx:
        with ctx.scene(gen) as sc:
            info = sc.info()
            log("BVH: %d nodes, depth %d, build %.1f ms" % (info.n_nodes, info.depth, info.build_ms))
            params = flgpu.default_params(spp=spp, depth=gen.depth, mode=flgpu.FL_SPECTRAL_HERO, seed=1,
                                          flags=flgpu.FL_RENDER_VNDF_PDF | flgpu.FL_RENDER_GRAPH)
            for i in range(args.warmup):
                _, _, st = sc.render(params, want_film=False)
                log("warmup %d: %.1f ms" % (i, st.ms_total))
            times = []
            t_start = time.perf_counter()
            for i in range(args.steps):
                ts = time.perf_counter()
                _, _, st = sc.render(params, want_film=False)
                times.append(time.perf_counter() - ts)
                log("step %d: %.1f ms (device %.1f ms), %d ext + %d shadow rays" % (
                    i, times[-1] * 1e3, st.ms_total, st.rays_extension, st.rays_shadow))
            wall = time.perf_counter() - t_start
            samples = W * H * spp
            ms_per_step = wall / args.steps * 1e3
            value = samples * args.steps / wall / 1e6
            rays = st.rays_extension + st.rays_shadow
            # timing + counter passes over the identical rays (same seed)
            tp = flgpu.default_params(spp=spp, depth=gen.depth, mode=flgpu.FL_SPECTRAL_HERO, seed=1,
                                      flags=flgpu.FL_RENDER_VNDF_PDF | flgpu.FL_RENDER_TIMING)
            _, _, tst = sc.render(tp, want_film=False)
            cp = flgpu.default_params(spp=spp, depth=gen.depth, mode=flgpu.FL_SPECTRAL_HERO, seed=1,
                                      flags=flgpu.FL_RENDER_VNDF_PDF | flgpu.FL_RENDER_COUNTERS)
            _, _, cst = sc.render(cp, want_film=False)
            # k_trace: per bounce, the previous bounce's shadow rays + this bounce's extension rays
            n_rays = cst.rays_extension + cst.rays_shadow
            trace_bytes = n_rays * (32 + 16) + cst.nodes_visited * 128 + cst.tris_tested * 48
            launches = max(1, tst.launches_extend)
            achieved = trace_bytes / (tst.ms_extend * 1e-3) / 1e9
            roofline = {
                "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_from_profile(),
                "kernel": "k_trace (BVH traversal: closest-hit extension + any-hit shadow rays)",
                "bytes_per_launch": int(trace_bytes / launches),
                "avg_launch_ms": round(tst.ms_extend / launches, 4),
                "nodes_per_ray": round(cst.nodes_visited / max(1, n_rays), 2),
                "tris_per_ray": round(cst.tris_tested / max(1, n_rays), 2),
                # by ray type: any-hit shadow rays and closest-hit extension rays
                "shadow_rays": int(cst.rays_shadow), "extension_rays": int(cst.rays_extension),
                "shadow_nodes_per_ray": round(cst.shadow_nodes_visited / max(1, cst.rays_shadow), 2),
                "shadow_tris_per_ray": round(cst.shadow_tris_tested / max(1, cst.rays_shadow), 2),
                "extension_nodes_per_ray": round((cst.nodes_visited - cst.shadow_nodes_visited) /
                                                 max(1, cst.rays_extension), 2),
                "extension_tris_per_ray": round((cst.tris_tested - cst.shadow_tris_tested) /
                                                max(1, cst.rays_extension), 2),
                "ms_trace": round(tst.ms_extend, 2), "ms_shade": round(tst.ms_shade, 2),
                "launches": launches,
                "achieved_definition": "algorithmic bytes (32 B ray + 16 B hit + 128 B per node + 48 B per "
                                       "triangle visited) per launch / average launch time; they include the "
                                       "reads L2 and the Infinity Cache serve, so frac can pass 1",
            }
            if roofline["traffic"]:
                # the measured HBM traffic per launch (rocprof FETCH_SIZE + WRITE_SIZE,
                # calibrated) over this run's average launch time
                hbm = roofline["traffic"] / (tst.ms_extend / launches * 1e-3) / 1e9
                roofline["hbm_achieved"] = round(hbm, 1)
                roofline["hbm_frac"] = round(hbm / HBM_PEAK_GBS, 4)
            issue = issue_from_profile()
            if issue:
                # the byte roofline is the one the workload is specified
                # against; what binds the kernel is instruction issue under
                # divergence, with L2 misses served from the Infinity Cache
                roofline["binding"] = issue.get("binding")
                roofline["issue"] = issue
            if not args.no_intersect:
                log("intersection section: 4M rays")
                out["intersection"] = intersection_section(flgpu, ctx, gen, sc, 0.0 if args.no_cpu else
                                                           args.cpu_seconds / 2, host_threads())
                log("intersection: %s" % out["intersection"])
            if not args.no_nee_one:
                log("direct lighting section: all lights vs one light per vertex")
                out["nee_one"] = nee_one_section(flgpu, sc, gen, spp)
                log("nee_one: %s" % out["nee_one"])
            out["mrays_per_s"] = round(rays * args.steps / wall / 1e6, 1)
            out["rays_per_sample"] = round(rays / samples, 3)
            out["bvh_build_ms"] = round(info.build_ms, 2)
            out["nan_samples"] = int(st.nan_samples)

        if not args.no_batch:
            log("batch section: %d scenes" % args.batch_scenes)
            out["batch"] = batch_section(flgpu, ctx, args.batch_scenes, tune=True)
            log("batch: %s" % out["batch"])
        if not args.no_guided:
            log("guided section: %d^2, reference %d spp" % (args.guided_res, args.guided_ref_spp))
            out["guided"] = guided_section(flgpu, ctx, flgpu.FL_GEN_GUIDED, args.guided_res, args.guided_ref_spp)
            log("guided: %s" % out["guided"])
            if not args.no_guided_skylight:
                out["guided_skylight"] = guided_section(flgpu, ctx, flgpu.FL_GEN_GUIDED_SKYLIGHT, args.guided_res,
                                                        args.guided_ref_spp)
                log("guided skylight: %s" % out["guided_skylight"])

        if args.terrain:
            log("terrain section")
            out["terrain"] = terrain_section(flgpu, ctx, args.terrain_spp, args.terrain_res_div)
            log("terrain: %s" % out["terrain"])

        if args.calibrate_copy:
            # plain streaming kernel (k_copy_probe), 1 GiB read + 1 GiB written
            # per launch: the reference point for FETCH_SIZE / WRITE_SIZE
            out["copy_probe_gbs"] = round(ctx.probe_copy(1 << 30, 3), 1)

        # the small Cornell workload (the CPU-runnable one): GPU rate
        cornell = None
        if not args.no_cornell:
            cgen = flgpu.GenSceneData(flgpu.FL_GEN_CORNELL)
            with ctx.scene(cgen) as csc:
                cparams = flgpu.default_params(spp=cgen.spp, depth=cgen.depth, mode=flgpu.FL_SPECTRAL_HERO, seed=1,
                                               flags=flgpu.FL_RENDER_GRAPH)
                csc.render(cparams, want_film=False)
                ct = []
                for _ in range(5):
                    ts = time.perf_counter()
                    csc.render(cparams, want_film=False)
                    ct.append(time.perf_counter() - ts)
                c_samples = cgen.camera.width * cgen.camera.height * cgen.spp
                cornell = {"gpu_msamples_s": round(c_samples * len(ct) / sum(ct) / 1e6, 2)}

    host = host_info()
    threads = host["threads"]
    cpu = None
    if not args.no_cpu:
        v, ns, secs = cpu_baseline(gen, 1, gen.depth, args.cpu_seconds, threads)
        cpu = {"value": round(v, 4), "unit": "Msamples/s", "cores": threads, "kind": "port",
               "host_nproc": host["nproc"], "host_affinity": host["affinity"], "host_model": host["model"],
               "per_thread": round(v / threads, 4),
               "sample": "%d samples (1 spp, rows of the same %dx%d image, depth %d), %.1f s" % (ns, W, H, gen.depth,
                                                                                             secs)}
        if cornell is not None:
            cv, cns, csecs = cpu_baseline(cgen, cgen.spp, cgen.depth, args.cpu_seconds / 3, threads)
            # the ratio against every online host core, assuming the CPU
            # rate scaled linearly from the threads used (a bound in the
            # CPU's favour)
            cornell.update({"cpu_msamples_s": round(cv, 4), "cpu_cores": threads,
                            "speedup": round(cornell["gpu_msamples_s"] / cv, 1),
                            "cpu_per_thread": round(cv / threads, 4),
                            "speedup_vs_all_host_cores_linear": round(
                                cornell["gpu_msamples_s"] / (cv / threads * host["nproc"]), 1),
                            "cpu_sample": "%d samples of the 256x256x16 image" % cns})
        out["speedup_vs_cpu"] = round(value / v, 1)

    res = {
        "metric": METRIC, "value": round(value, 2), "unit": "Msamples/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 2), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "%s: %d tris, %dx%d, %d spp, depth %d, hero 360-830nm" % (
            args.workload, gen.n_triangles, W, H, spp, gen.depth), "samples_per_step": samples},
        "roofline": roofline, "cpu_baseline": cpu, "cornell": cornell,
    }
    res.update(out)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
