#!/usr/bin/env python3
"""Benchmark of the imaging major cycle's grid+degrid hot path on one MI355X.

A step is one pass of the major cycle's visibility-domain work over the whole
synthetic data set, inputs resident in HBM:
    degrid the model grid (kernel-weighted, w-projection) and subtract from the
    measured visibilities (std_degridding + subtract_from_measurements_finegrid)
    -> grid the residual visibilities (std_gridding + merge).
Workload (default): 256 antennas log-radial to 5 km, 512 time steps
(16,711,680 visibilities), 4096^2 grid, 32 w-planes with half support 4-16,
oversampling 8.

Prints ONE JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "simsdp-generic-imaging-pipeline_amd"))

METRIC = "Mvis/s grid+degrid; major-cycle s to fixed residual RMS, learned vs Hogbom"
HBM_PEAK_GBS = 8000.0
VALU_PEAK_TFLOPS = 78.6   # FP32 FMA without packed math: 256 CUs x 64 lanes x 2 flops x 2.4 GHz


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def measured_traffic(phase):
    """HBM bytes per bench step of `phase` (its launches together), from the
    FETCH_SIZE / WRITE_SIZE counter summary committed under profiles/ for
    this workload (calibrated against a plain copy); None when absent."""
    path = os.path.join(REPO, "profiles", "degrid_grid_config2_hbm_traffic.md")
    try:
        text = open(path).read()
        block = text.split("```json", 1)[1].split("```", 1)[0]
        return int(json.loads(block)[phase])
    except (OSError, IndexError, KeyError, ValueError):
        return None


def copy_bandwidth(hip, nbytes=1 << 30, reps=5):
    """HBM bandwidth of a plain 16-byte device copy (read + write bytes per
    second), the measured ceiling beside the 8 TB/s specification."""
    a, b = hip.DeviceBuffer(nbytes), hip.DeviceBuffer(nbytes)
    a.zero()
    ev = [hip.Event() for _ in range(2)]
    for _ in range(2):
        hip.check(hip.load().sdp_device_copy(b.ptr, a.ptr, nbytes, None))
    ev[0].record()
    for _ in range(reps):
        hip.check(hip.load().sdp_device_copy(b.ptr, a.ptr, nbytes, None))
    ev[1].record()
    ev[1].sync()
    ms = ev[0].elapsed_ms(ev[1])
    a.close()
    b.close()
    return round(2.0 * nbytes * reps / (ms * 1e-3) / 1e9, 1)


def cpu_baseline(d, seconds):
    """Oracle std_gridding + std_degridding (the reference's CPU actors,
    restated) on a random sample of the same visibilities: on one core, and
    on `threads` cores the way the reference parallelises inside a node
    (splitmerge.c: visibility slices gridded into per-slice grids, merged in
    slice order, then degridded per slice).  The multi-core figure is the
    baseline's value."""
    import numpy as np
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import sdp_hip
    import sdp_oracle

    cfg = sdp_hip.make_config(d.grid_size, d.cell_size, oversampling=d.oversampling,
                              num_kernels=d.supports.shape[0], max_w=d.max_w)
    rng = np.random.default_rng(5)
    grid = np.zeros((d.grid_size ** 2, 2), np.float32)

    def run(n):
        idx = np.sort(rng.choice(d.n_vis, n, replace=False))
        uvw, vis = d.uvw[idx], d.vis[idx]
        t0 = time.perf_counter()
        g = sdp_oracle.std_gridding(d.grid_size, d.oversampling, d.kernels, d.supports, uvw, vis, cfg)
        t1 = time.perf_counter()
        sdp_oracle.std_degridding(d.grid_size, d.oversampling, d.degrid_kernels, d.supports, g, uvw, cfg)
        t2 = time.perf_counter()
        return t1 - t0, t2 - t1

    tg, td = run(2000)
    per_vis = (tg + td) / 2000
    n = int(max(2000, min(d.n_vis, seconds / max(per_vis, 1e-9))))
    tg, td = run(n)
    del grid
    single = {"value": round(n / (tg + td) / 1e6, 4), "cores": 1,
              "sample": f"{n} of {d.n_vis} visibilities (uniform random), {d.grid_size}^2 grid: oracle std_gridding "
                        f"{tg:.2f} s + std_degridding {td:.2f} s, -O3 C, 1 thread"}
    # the same sample size per thread, within the host's share of cores
    threads = max(1, min(os.cpu_count() or 1, int(os.environ.get("OMP_NUM_THREADS") or 0) or 64, 64))
    m = int(min(d.n_vis, n * threads))
    idx = np.sort(rng.choice(d.n_vis, m, replace=False))
    _, _, (sg, sd) = sdp_oracle.sliced_grid_degrid(threads, d.grid_size, d.oversampling, d.kernels, d.degrid_kernels,
                                                   d.supports, d.uvw[idx], d.vis[idx], cfg)
    return {"value": round(m / (sg + sd) / 1e6, 4), "unit": "Mvis/s", "cores": threads, "kind": "port",
            "sample": f"{m} of {d.n_vis} visibilities (uniform random), {d.grid_size}^2 grid, {threads} threads "
                      f"(splitmerge.c slices): gridding and merge {sg:.2f} s + degridding {sd:.2f} s, -O3 C",
            "single_core": single}


def major_cycle(hip, sdp_synth, d, d_uvw, d_meas, d_gk, d_dk, tile, chunk, target_frac=0.1):
    """Full major cycles on the device (sdp_pipeline.h): predict, degrid,
    subtract, grid, inverse FFT, Hogbom; replayed from a captured graph
    after the first.  Reports the residual RMS per cycle (Jy) and the time
    to bring it to target_frac of the dirty image's RMS."""
    d_tap = hip.DeviceBuffer.from_array(sdp_synth.taper_padded(d.grid_size))
    # stage breakdown from a profiled (eager) pipeline: second cycle
    prof = hip.Pipeline(d.grid_size, d.oversampling, d.supports, d.uv_scale, d.w_scale, d.cell_size, d_uvw, d_meas,
                        d.n_vis, d_gk, d_dk, d_tap, d.n_minor, loop_gain=0.1, flags=hip.PIPE_PROFILE,
                        tile=tile, chunk=chunk)
    prof.cycle()
    prof.cycle()
    stage_ms = {k: round(v, 3) for k, v in prof.stage_ms().items()}
    prof.close()
    log(f"major-cycle stages (ms): {stage_ms}")
    t = time.perf_counter()
    pipe = hip.Pipeline(d.grid_size, d.oversampling, d.supports, d.uv_scale, d.w_scale, d.cell_size, d_uvw, d_meas,
                        d.n_vis, d_gk, d_dk, d_tap, d.n_minor, loop_gain=0.1, flags=hip.PIPE_GRAPH,
                        tile=tile, chunk=chunk)
    setup_s = time.perf_counter() - t
    cycles = []
    target = None
    s_to_target = None
    elapsed_ms = 0.0
    for c in range(d.n_major):
        before, after, ms = pipe.cycle()
        if target is None:
            target = target_frac * before
        elapsed_ms += ms
        if s_to_target is None and before <= target:
            s_to_target = elapsed_ms * 1e-3   # this cycle's residual certifies the target
        cycles.append({"rms_residual": float(f"{before:.6g}"), "rms_after_clean": float(f"{after:.6g}"),
                       "ms": round(ms, 3)})
        log(f"major cycle {c}: residual RMS {before:.4g} Jy -> {after:.4g} after CLEAN, {ms:.2f} ms")
    n_src = pipe.sources().shape[0]
    pipe.close()
    G = d.grid_size
    hog_ms = stage_ms["clean"]
    # Hogbom per iteration: scan 4 G^2 bytes + PSF window read and residual
    # window read-modify-write 12 A_win (A_win = G^2 for a central peak)
    hog_bytes = (4 + 12) * G * G * d.n_minor
    return {"minor": "hogbom", "n_minor": d.n_minor, "gain": 0.1, "setup_s": round(setup_s, 3),
            "stage_ms": stage_ms, "hogbom_us_per_iter": round(hog_ms * 1e3 / max(d.n_minor, 1), 2),
            "hogbom_gbs": round(hog_bytes / (hog_ms * 1e-3) / 1e9, 1),
            "cycles": cycles, "s_total": round(elapsed_ms * 1e-3, 4),
            "target_rms": float(f"{target:.6g}"), "target": f"{target_frac} x dirty-image RMS",
            "s_to_target": None if s_to_target is None else round(s_to_target, 4),
            "components_last_cycle": n_src}


LEARNED_N_MINOR = 50   # Hogbom iterations per major cycle on the learned path


def learned_vs_hogbom(hip, nets, seed=5):
    """Major-cycle time to a fixed residual RMS, learned proposal + Hogbom
    against plain Hogbom, on the learned-proposal workload: config-1 array
    (516,096 visibilities, 1024^2 grid, OS 16, no w-term) observing 50
    power-law point sources and 20 elliptical Gaussians (FWHM 2-30 px) in a
    2 x 2 degree box, noise 0.01 Jy.  The plain path runs 5 major cycles x
    1000 Hogbom iterations (gain 0.1); its final residual RMS is the target.
    Each network of `nets` (dicts: levels, base channels, training pairs,
    steps and optionally head_scale and lr; BF16 MFMA, FP32 accumulation)
    is trained here by the library's kernels
    on pairs the GPU makes from the pipeline's own PSF (make_pool_device);
    the training time is reported, not counted in the major-cycle time.
    The learned path runs LEARNED_N_MINOR Hogbom iterations per major cycle
    after the proposals (also timed at 300, 100 and 25 iterations)."""
    import numpy as np
    import sdp_learned as L

    w = L.learned_workload(seed=seed)
    G, T, margin = w.ds.grid_size, 256, 32

    def run(pipe, target, max_cycles):
        hip.synchronize()
        t0 = time.perf_counter()
        ms, rows = L.cycles_to_target(pipe, target, max_cycles)
        hip.synchronize()
        return ms, rows, (time.perf_counter() - t0) * 1e3

    def fmt(rows):
        return [[float(f"{a:.5g}"), float(f"{b:.5g}"), round(ms, 3)] for a, b, ms in rows]

    plain = w.pipeline(1000, flags=hip.PIPE_GRAPH)
    psf = plain.image(hip.IMAGE_PSF)
    _, hog_rows, hog_wall = run(plain, 0.0, 5)
    plain.close()
    target = hog_rows[-1][1]
    hog_ms = sum(r[2] for r in hog_rows)
    log(f"learned workload: plain Hogbom 5 x 1000 -> RMS {target:.4g} Jy in {hog_ms:.1f} ms")
    p300 = w.pipeline(300, flags=hip.PIPE_GRAPH)
    h300_ms, h300_rows, _ = run(p300, target, 40)
    p300.close()
    out = {
        "workload": f"config-1 array ({w.ds.n_vis} vis), {G}^2 grid, OS 16, no w-term; sky of 50 power-law points "
                    f"+ 20 elliptical Gaussians FWHM 2-30 px in a 2 deg box (seed {seed}), noise 0.01 Jy",
        "target_rms": float(f"{target:.6g}"), "target": "residual RMS after 5 x 1000 Hogbom iterations",
        "hogbom": {"n_minor": 1000, "gain": 0.1, "ms": round(hog_ms, 3), "wall_ms": round(hog_wall, 3),
                   "cycles": fmt(hog_rows)},
        "hogbom_n300": {"n_minor": 300, "ms": None if h300_ms is None else round(h300_ms, 3),
                        "n_cycles": len(h300_rows)},
        "networks": [],
    }
    for spec in nets:
        levels, base, n_pairs, steps = spec["levels"], spec["base"], spec["pairs"], spec["steps"]
        head, lr = spec.get("head_scale", 1.0), spec.get("lr", 3e-4)
        t = time.perf_counter()
        pool = L.make_pool_device(psf, n_pairs, seed=1, T=T, crops_per_sky=8, with_dirty=True,
                                  crop_range=(G // 2 - T - 64, G // 2 + 64), n_points=(20, 80), n_gauss=(0, 20),
                                  box=w.box)
        pool_s = time.perf_counter() - t
        net = hip.Cnn(levels, base, T, T, 16, precision=hip.CNN_BF16)
        L.init_params(net, 0, head_scale=head)
        t = time.perf_counter()
        losses = L.train(net, pool, steps, 16, lr=lr, lr_final=lr / 10, data_weight=1.0, psf=L.psf_patch(psf, T))
        hip.synchronize()
        train_s = time.perf_counter() - t
        for b in (pool.d_x, pool.d_t, pool.d_d):
            b.close()
        log(f"network {levels}x{base}: {n_pairs} pairs in {pool_s:.1f} s, {steps} steps in {train_s:.1f} s, "
            f"loss {losses[0]:.3g} -> {np.mean(losses[-50:]):.3g}")
        # Hogbom iterations per learned cycle: the proposals remove most of
        # the flux, so the minor cycle after them is short; the other counts
        # are timed too and reported beside it
        sweep = {}
        for n_minor in (LEARNED_N_MINOR, 300, 100, 25):
            lp = w.pipeline(n_minor)
            lp.set_learned(net, T, margin, gain=1.0, threshold=0.05, floor=0.0, max_cycles=3,
                           scale=hip.LEARNED_GLOBAL_SCALE)
            ms_n, rows_n, wall_n = run(lp, target, 14)
            lp.close()
            if n_minor == LEARNED_N_MINOR and n_minor not in sweep:
                l_ms, l_rows, l_wall = ms_n, rows_n, wall_n
            sweep[n_minor] = {"ms": None if ms_n is None else round(ms_n, 3), "n_cycles": len(rows_n),
                              "ratio": None if ms_n is None else round(ms_n / hog_ms, 4)}
        net.close()
        log(f"learned ({levels}x{base}) + Hogbom: {'not reached' if l_ms is None else f'{l_ms:.1f} ms'} "
            f"in {len(l_rows)} cycles")
        out["networks"].append({
            "net": {"levels": levels, "base": base, "params": int(net_params(hip, levels, base)),
                    "precision": "bf16 MFMA, fp32 acc"},
            "train": {"pairs": n_pairs, "steps": steps, "batch": 16, "lr": lr, "head_scale": head,
                      "s": round(train_s, 2),
                      "pool_s": round(pool_s, 2), "loss_first": float(f"{losses[0]:.4g}"),
                      "loss_last50": float(f"{np.mean(losses[-50:]):.4g}")},
            "learned": {"n_minor": LEARNED_N_MINOR, "learned_cycles": 3, "gain": 1.0, "threshold": 0.05,
                        "patch": T, "margin": margin, "scale": "global",
                        "ms": None if l_ms is None else round(l_ms, 3), "wall_ms": round(l_wall, 3),
                        "cycles": fmt(l_rows)},
            "n_minor_sweep": {str(k): v for k, v in sorted(sweep.items(), reverse=True)},
            "ratio": None if l_ms is None else round(l_ms / hog_ms, 4),
        })
    # the configured network (the last) gives the headline ratio
    out["ratio"] = out["networks"][-1]["ratio"] if out["networks"] else None
    return out


def dft_precision_variant(hip, d, n_src=4096):
    """DFT predict of n_src power-law sources onto the array's 16.7M
    visibilities: the array-factorised matrix-core form (BF16 and FP16
    operands with the hi + lo split, and without it) against the
    per-visibility FP64-phase kernel; errors against the FP64 restatement
    over the first instant's 32,640 baselines, relative to sum |g|."""
    import numpy as np
    import sdp_synth

    enu, ha = d.extra["enu"], d.extra["ha_hours"]
    ant = sdp_synth.antenna_uvw(enu, -30.0, -30.0, ha, 150e6)
    src = sdp_synth.config3_sources(n_src)
    n_time, n_ant = ant.shape[0], ant.shape[1]
    n_pairs = n_ant * (n_ant - 1) // 2
    N = n_time * n_pairs
    pairs = sdp_synth.receiver_pairs(n_ant)
    ref = sdp_synth.dft_f64((ant[:1, pairs[:, 1]] - ant[:1, pairs[:, 0]]).reshape(-1, 3), src)
    l, m, f = (src[:, k].astype(np.float64) for k in range(3))
    gsum = float(np.abs(f / (1 - 0.5 * (l * l + m * m))).sum())
    d_ant, d_src = hip.DeviceBuffer.from_array(ant), hip.DeviceBuffer.from_array(src)
    d_out = hip.DeviceBuffer(8 * N)
    work = hip.DeviceBuffer(hip.dft_array_work_bytes(n_src))
    ev = [hip.Event(), hip.Event()]
    nt = (n_ant + 127) // 128
    out = {"workload": f"{n_src} sources x {N} vis ({n_ant} antennas x {n_time} instants)",
           "error_ref": "FP64 restatement, first instant, max |V - V64| / sum|g|"}
    for name, flags in (("bf16_split", hip.DFT_ARRAY_BF16), ("fp16_split", hip.DFT_ARRAY_FP16),
                        ("bf16", hip.DFT_ARRAY_BF16 | hip.DFT_ARRAY_NO_SPLIT),
                        ("fp16", hip.DFT_ARRAY_FP16 | hip.DFT_ARRAY_NO_SPLIT)):
        hip.dft_predict_array(n_time, n_ant, d_ant, n_src, d_src, d_out, flags, work)
        ev[0].record()
        for _ in range(3):
            hip.dft_predict_array(n_time, n_ant, d_ant, n_src, d_src, d_out, flags, work)
        ev[1].record()
        ev[1].sync()
        ms = ev[0].elapsed_ms(ev[1]) / 3
        got = d_out.download((n_pairs, 2), np.float32)
        products = 1 if flags & hip.DFT_ARRAY_NO_SPLIT else 3
        mfma_flops = n_time * nt * (nt + 1) // 2 * 128 * 128 * ((n_src + 15) // 16 * 32) * 2 * 2 * products
        alg_flops = 8.0 * N * n_src   # one complex multiply-add per visibility-source pair
        out[name] = {"ms": round(ms, 3), "gpairs_per_s": round(N * n_src / ms / 1e6, 1),
                     "max_err": float(f"{np.abs(got - ref).max() / gsum:.3g}"),
                     "alg_tflops": round(alg_flops / ms / 1e9, 1),
                     "mfma_frac": round(alg_flops / ms / 1e9 / 2500.0, 4),
                     "mfma_issued_tflops": round(mfma_flops / ms / 1e9, 1),
                     "mfma_issued_frac": round(mfma_flops / ms / 1e9 / 2500.0, 4)}
        log(f"array DFT {name}: {ms:.3f} ms, error {out[name]['max_err']}")
    uvw = (ant[:, pairs[:, 1]] - ant[:, pairs[:, 0]]).reshape(-1, 3).astype(np.float32)
    d_uvw = hip.DeviceBuffer.from_array(uvw)
    ev[0].record()
    hip.dft_predict(N, n_src, d_src, d_uvw, d_out, hip.DFT_FP64_PHASE)
    ev[1].record()
    ev[1].sync()
    ms = ev[0].elapsed_ms(ev[1])
    got = d_out.download((n_pairs, 2), np.float32)
    out["per_vis_fp64_phase"] = {"ms": round(ms, 3), "gpairs_per_s": round(N * n_src / ms / 1e6, 1),
                                 "max_err": float(f"{np.abs(got - ref).max() / gsum:.3g}")}
    return out


def batched_channels(hip, n_channels=64, passes=2):
    """The batched multi-channel workload: config-1 array with 8192 time
    steps and sky (spectral index -0.7) at 64 frequencies on 100-200 MHz,
    flagged fraction U[0, 0.6] per channel (6.6-16.5M visibilities each),
    64 grids of 2048^2, 16 w-planes.  One plan covers all channels, so a
    residual pass (batched predict FFT, degrid + subtract + grid, merge,
    batched inverse FFT) takes the launches of one channel; it is captured
    once and replayed as one graph; the grid and degrid stage times come
    from a profiled (eager) pass.  Tile and work-item size from the
    autotuner on the first channel."""
    import numpy as np
    import sdp_synth

    t = time.perf_counter()
    m = sdp_synth.config4(n_channels=n_channels)
    G = m.grid_size
    base = hip.DeviceBuffer.from_array(m.base_uvw)
    uvws, viss = [], []
    for c in range(n_channels):
        n = int(m.n_vis[c])
        u = hip.DeviceBuffer(12 * n)
        hip.axpby(3 * n, float(m.freqs[c] / sdp_synth.C_LIGHT), base, 0.0, base, u)
        src = hip.DeviceBuffer.from_array(m.channel_sources(c))
        v = hip.DeviceBuffer(8 * n)
        hip.dft_predict(n, src.nbytes // 12, src, u, v, hip.DFT_FP64_PHASE)
        uvws.append(u)
        viss.append(v)
    hip.synchronize()
    gen_s = time.perf_counter() - t
    common = [hip.DeviceBuffer.from_array(a) for a in (m.kernels, m.degrid_kernels, sdp_synth.taper_padded(G))]
    t = time.perf_counter()
    tile, chunk, tune_ms = hip.autotune(G, m.oversampling, m.supports, m.uv_scale, m.w_scale, uvws[0],
                                        int(m.n_vis[0]), common[0], common[1])
    tune_s = time.perf_counter() - t
    t = time.perf_counter()
    b = hip.Batch(G, m.oversampling, m.supports, m.uv_scale, m.w_scale, uvws, viss, m.n_vis, *common, tile=tile,
                  chunk=chunk, flags=hip.BATCH_GRAPH)
    setup_s = time.perf_counter() - t
    total = int(m.n_vis.sum())
    info = b.plan_info(0)                                       # the plan over all channels
    ms = [b.residuals()[0] for _ in range(passes + 1)]          # eager, capture + replay, replay
    b.close()
    prof = hip.Batch(G, m.oversampling, m.supports, m.uv_scale, m.w_scale, uvws, viss, m.n_vis, *common, tile=tile,
                     chunk=chunk, flags=hip.BATCH_PROFILE)
    _, stage = prof.residuals()
    prof.close()
    st = dict(zip(hip.STAGES, stage))
    gd_ms = st["degrid_subtract_grid"] + st["grid_merge"]
    alg = total * 40 + n_channels * 16 * G * G
    out = {"workload": f"{n_channels} channels {m.freqs[0] / 1e6:.0f}-{m.freqs[-1] / 1e6:.0f} MHz, {total} vis "
                       f"({int(m.n_vis.min())}-{int(m.n_vis.max())} per channel), {G}^2 grids, "
                       f"{m.supports.shape[0]} w-planes h {m.supports[:, 0].min()}-{m.supports[:, 0].max()}, "
                       f"OS {m.oversampling}",
           "n_vis": total, "graph_pass_ms": round(min(ms[1:]), 3), "eager_pass_ms": round(ms[0], 3),
           "stage_ms": {k: round(v, 3) for k, v in st.items()},
           "mvis_per_s_grid_degrid": round(total / (gd_ms * 1e-3) / 1e6, 1),
           "mvis_per_s_pass": round(total / (min(ms[1:]) * 1e-3) / 1e6, 1),
           "hbm_frac_grid_degrid": round(alg / (gd_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 5),
           "taps": info["taps"], "work_items": info["chunks"], "plan_scratch_bytes": info["scratch_bytes"],
           "gtaps_per_s": round(2 * info["taps"] / (gd_ms * 1e-3) / 1e9, 1),
           "autotune": {"tile": tile, "chunk_taps": chunk, "s": round(tune_s, 2)},
           "generate_s": round(gen_s, 2), "setup_s": round(setup_s, 2)}
    for u in uvws + viss:
        u.close()
    base.close()
    return out


def net_params(hip, levels, base):
    return sum(9 * ci * co + co for _, ci, co in hip.cnn_layer_shapes(levels, base, 2, 1))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="config2", choices=["config1", "config2"])
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-vs-Hogbom major-cycle comparison")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched 64-channel workload")
    ap.add_argument("--learned-steps", type=int, default=2000, help="training steps of the 3x16 proposal network")
    ap.add_argument("--learned-steps-4x32", type=int, default=3000,
                    help="training steps of the configured 4-level, 32-channel proposal network")
    ap.add_argument("--learned-pairs", type=int, default=49152, help="training pairs of the 4x32 network")
    ap.add_argument("--accumulate", choices=["blocked", "compensated"], default="blocked",
                    help="gridder accumulation inside a work item: float (blocked) or TwoSum; "
                         "work items are merged TwoSum-compensated either way")
    ap.add_argument("--tile", type=int, default=0, help="uv tile of the plan (with --chunk: no autotuning)")
    ap.add_argument("--chunk", type=int, default=0, help="taps per work item of the plan (with --tile)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain 1 GiB device copy 5 times (HBM counter calibration)")
    ap.add_argument("--cnn-only", action="store_true",
                    help="only run the 4-level, 32-channel network's BF16 forward pass on 16 patches of 256^2 "
                         "10 times (kernel traces of the learned path's convolutions)")
    ap.add_argument("--fft-only", action="store_true",
                    help="only run the major cycle's 4096^2 inverse FFT (shift / real / taper) 5 times "
                         "(counter runs on the FFT kernels)")
    ap.add_argument("--dft-only", action="store_true",
                    help="only run the array DFT predict of 4096 sources onto the 16.7M-visibility array "
                         "(BF16 and FP16 hi + lo split), 5 times each (counter runs on the matrix-core kernel)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import sdp_hip as hip
    import sdp_synth

    if args.copy_probe:
        n = 1 << 30
        a, b = hip.DeviceBuffer(n), hip.DeviceBuffer(n)
        a.zero()
        for _ in range(5):
            hip.check(hip.load().sdp_device_copy(b.ptr, a.ptr, n, None))
        hip.synchronize()
        log(f"copied {n} bytes x 5")
        return
    if args.cnn_only:
        import numpy as np
        import sdp_learned as L
        net = hip.Cnn(4, 32, 256, 256, 16, precision=hip.CNN_BF16)
        L.init_params(net, 1)
        x = np.random.default_rng(0).standard_normal((16, 256, 256, 2)).astype(np.float32)
        d_x, d_y = hip.DeviceBuffer.from_array(x), hip.DeviceBuffer(4 * 16 * 256 * 256)
        for _ in range(10):
            net.forward(d_x, d_y, 16)
        hip.synchronize()
        net.close()
        log("4x32 forward x 10 on 16 x 256^2")
        return
    if args.fft_only:
        G = 4096
        d_grid = hip.DeviceBuffer(8 * G * G)
        d_grid.zero()
        d_dirty = hip.DeviceBuffer(4 * G * G)
        d_work = hip.DeviceBuffer(hip.fft_work_bytes(G))
        d_tap = hip.DeviceBuffer.from_array(sdp_synth.taper_padded(G))
        flags = hip.FFT_SHIFT_IN | hip.FFT_SHIFT_OUT | hip.FFT_REAL_OUT | hip.FFT_CORRECT_OUT
        for _ in range(5):
            hip.fft2d(d_grid, d_dirty, G, hip.FFT_INVERSE, flags, d_tap, d_work)
        hip.synchronize()
        log(f"{G}^2 inverse FFT x 5")
        return

    if args.dft_only:
        enu = sdp_synth.log_radial_layout(256, 50.0, 5000.0, seed=11)
        ant = sdp_synth.antenna_uvw(enu, -30.0, -30.0, np.linspace(-4.0, 4.0, 512), 150e6)
        src = sdp_synth.config3_sources(4096)
        n_time, n_ant, n_src = ant.shape[0], ant.shape[1], src.shape[0]
        d_ant, d_src = hip.DeviceBuffer.from_array(ant), hip.DeviceBuffer.from_array(src)
        d_out = hip.DeviceBuffer(8 * n_time * (n_ant * (n_ant - 1) // 2))
        work = hip.DeviceBuffer(hip.dft_array_work_bytes(n_src))
        for flags in (hip.DFT_ARRAY_BF16, hip.DFT_ARRAY_FP16):
            for _ in range(5):
                hip.dft_predict_array(n_time, n_ant, d_ant, n_src, d_src, d_out, flags, work)
        hip.synchronize()
        log(f"array DFT {n_src} sources x {n_ant} antennas x {n_time} instants, BF16 and FP16 split, x 5 each")
        return

    t = time.perf_counter()
    d = sdp_synth.config2() if args.config == "config2" else sdp_synth.config1(with_vis=False)
    log(f"dataset {d.name}: {d.n_vis} vis, G={d.grid_size}, planes={d.supports.shape[0]} "
        f"h={d.supports[:, 0].min()}..{d.supports[:, 0].max()}, OS={d.oversampling} ({time.perf_counter() - t:.1f} s)")
    G, N = d.grid_size, d.n_vis
    hip.check(hip.load().sdp_init())

    d_uvw = hip.DeviceBuffer.from_array(d.uvw)
    d_gk = hip.DeviceBuffer.from_array(d.kernels)
    d_dk = hip.DeviceBuffer.from_array(d.degrid_kernels)
    d_src = hip.DeviceBuffer.from_array(d.sources)
    d_meas = hip.DeviceBuffer(8 * N)
    ev = [hip.Event() for _ in range(6)]

    # measured visibilities: FP64-phase DFT predict of the sky (also timed)
    ev[0].record()
    hip.dft_predict(N, d.sources.shape[0], d_src, d_uvw, d_meas, hip.DFT_FP64_PHASE)
    ev[1].record()
    ev[1].sync()
    dft_ms = ev[0].elapsed_ms(ev[1])
    pairs = N * d.sources.shape[0]
    log(f"DFT predict {d.sources.shape[0]} src x {N} vis: {dft_ms:.1f} ms ({pairs / dft_ms / 1e6:.1f} G pairs/s)")
    d.vis = d_meas.download((N, 2), np.float32)

    t = time.perf_counter()
    acc_mode = hip.ACC_COMPENSATED if args.accumulate == "compensated" else hip.ACC_BLOCKED
    if args.tile > 0 and args.chunk > 0:   # a fixed plan shape (profiling runs)
        tile, chunk, tune_ms = args.tile, args.chunk, 0.0
    else:
        tile, chunk, tune_ms = hip.autotune(G, d.oversampling, d.supports, d.uv_scale, d.w_scale, d_uvw, N, d_gk,
                                            d_dk, accumulate=acc_mode)
    tune_s = time.perf_counter() - t
    log(f"autotune {tune_s:.1f} s: tile {tile}, {chunk} taps per work item ({tune_ms:.2f} ms per pass)")
    t = time.perf_counter()
    plan = hip.GridPlan(G, d.oversampling, d.supports, d.uv_scale, d.w_scale, d_uvw, N, tile=tile, chunk=chunk,
                        accumulate=acc_mode)
    hip.synchronize()
    plan_s = time.perf_counter() - t
    info = plan.info()
    log(f"plan {plan_s * 1e3:.0f} ms: {info}")

    # model grid = FFT of a model image holding the sky's point sources
    img = np.zeros((G, G), np.float32)
    px = np.clip(np.round(d.sources[:, 0] / d.cell_size).astype(int) + G // 2, 0, G - 1)
    py = np.clip(np.round(d.sources[:, 1] / d.cell_size).astype(int) + G // 2, 0, G - 1)
    np.add.at(img, (py, px), d.sources[:, 2])
    d_img = hip.DeviceBuffer.from_array(img)
    d_model = hip.DeviceBuffer(8 * G * G)
    d_work = hip.DeviceBuffer(hip.fft_work_bytes(G))
    hip.fft2d(d_img, d_model, G, hip.FFT_FORWARD, hip.FFT_REAL_IN | hip.FFT_SHIFT_IN | hip.FFT_SHIFT_OUT, None,
              d_work)

    d_meas_p = hip.DeviceBuffer(8 * N)
    plan.to_plan_order(d_meas, d_meas_p)
    d_delta = hip.DeviceBuffer(8 * N)
    d_grid = hip.DeviceBuffer(8 * G * G)
    info["classes"] = plan.class_launches()

    # one step: the major cycle's fused degrid + subtract + grid pass (into
    # the work items' sub-grids), then their merge into the residual grid
    def run_steps(pl, steps, warmup):
        for _ in range(warmup):
            pl.degrid_grid_accumulate(d_dk, d_gk, d_model, d_meas_p, d_residual_vis=d_delta, order=hip.ORDER_PLAN)
            pl.merge(d_grid)
        hip.synchronize()
        acc = {"degrid_subtract_grid": 0.0, "grid_merge": 0.0}
        t0 = time.perf_counter()
        for _ in range(steps):
            ev[0].record()
            pl.degrid_grid_accumulate(d_dk, d_gk, d_model, d_meas_p, d_residual_vis=d_delta, order=hip.ORDER_PLAN)
            ev[1].record()
            pl.merge(d_grid)
            ev[2].record()
            ev[2].sync()
            acc["degrid_subtract_grid"] += ev[0].elapsed_ms(ev[1])
            acc["grid_merge"] += ev[1].elapsed_ms(ev[2])
        hip.synchronize()
        return (time.perf_counter() - t0) * 1e3 / steps, {k: v / steps for k, v in acc.items()}

    ms_per_step, k_ms = run_steps(plan, args.steps, args.warmup)
    log(f"step {ms_per_step:.3f} ms (events: {k_ms})")

    # algorithmic bytes per step of each phase
    alg = {
        # model grid read, uvw read, measured visibility read, residual visibility written
        "degrid_subtract_grid": 8 * G * G + N * (12 + 8 + 8),
        "grid_merge": 8 * G * G,                   # residual grid written once
    }
    dom = max(k_ms, key=k_ms.get)
    traffic = measured_traffic(dom) if args.config == "config2" else None
    achieved = alg[dom] / (k_ms[dom] * 1e-3) / 1e9
    b_gd = N * 40 + 16 * G * G
    gd_ms = sum(k_ms.values())
    taps = plan.total_taps()
    # the fused pass's arithmetic: a complex multiply-add per tap for the
    # prediction and another for the gridding, 8 flops each, against the
    # non-packed FP32 FMA rate (256 CUs x 64 lanes x 2 flops x 2.4 GHz)
    valu_tflops = 16.0 * taps / (k_ms["degrid_subtract_grid"] * 1e-3) / 1e12

    extras = {}
    if not args.no_extras:
        # FFT: inverse with the shift / real / taper epilogue, as in the major cycle
        d_tap = hip.DeviceBuffer.from_array(sdp_synth.taper_padded(G))
        d_dirty = hip.DeviceBuffer(4 * G * G)
        flags = hip.FFT_SHIFT_IN | hip.FFT_SHIFT_OUT | hip.FFT_REAL_OUT | hip.FFT_CORRECT_OUT
        hip.fft2d(d_grid, d_dirty, G, hip.FFT_INVERSE, flags, d_tap, d_work)
        ev[0].record()
        for _ in range(5):
            hip.fft2d(d_grid, d_dirty, G, hip.FFT_INVERSE, flags, d_tap, d_work)
        ev[1].record()
        ev[1].sync()
        fft_ms = ev[0].elapsed_ms(ev[1]) / 5
        extras["fft_ms"] = round(fft_ms, 4)
        extras["fft_gflops"] = round(5 * G * G * np.log2(G * G) / (fft_ms * 1e-3) / 1e9, 1)
        extras["dft_predict_gpairs_per_s"] = round(pairs / dft_ms / 1e6, 2)
        extras["plan_ms"] = round(plan_s * 1e3, 1)
        extras["grid_degrid_hbm_frac"] = round(b_gd / (gd_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 5)
        extras["kernel_ms"] = {k: round(v, 4) for k, v in k_ms.items()}
        # the same step as separate degrid + subtract, grid accumulate and merge passes
        sep = {}
        for rep in range(3):
            ev[0].record()
            plan.degrid(d_dk, d_model, d_delta, d_measured=d_meas_p, order=hip.ORDER_PLAN)
            ev[1].record()
            plan.accumulate(d_gk, d_delta, order=hip.ORDER_PLAN)
            ev[2].record()
            plan.merge(d_grid)
            ev[3].record()
            ev[3].sync()
            if rep:
                for k, (a, b) in (("degrid_subtract", (0, 1)), ("grid_accumulate", (1, 2)), ("grid_merge", (2, 3))):
                    sep[k] = sep.get(k, 0.0) + ev[a].elapsed_ms(ev[b]) / 2
        extras["separate_passes_ms"] = {k: round(v, 4) for k, v in sep.items()}
        # the same step with TwoSum-compensated sums inside the work items
        t = time.perf_counter()
        cplan = hip.GridPlan(G, d.oversampling, d.supports, d.uv_scale, d.w_scale, d_uvw, N, tile=tile, chunk=chunk,
                             accumulate=hip.ACC_COMPENSATED)
        c_ms, c_k = run_steps(cplan, 3, 1)
        cplan.close()
        extras["compensated_step"] = {"ms_per_step": round(c_ms, 4), "value": round(N / (c_ms * 1e-3) / 1e6, 2),
                                      "kernel_ms": {k: round(v, 4) for k, v in c_k.items()},
                                      "s": round(time.perf_counter() - t, 2)}
        extras["autotune"] = {"tile": tile, "chunk_taps": chunk, "pass_ms": round(tune_ms, 3), "s": round(tune_s, 2)}
        extras["taps_per_step"] = taps
        extras["gtaps_per_s"] = round(2 * taps / (k_ms["degrid_subtract_grid"] * 1e-3) / 1e9, 1)
        extras["plan"] = info
        extras["major_cycle"] = major_cycle(hip, sdp_synth, d, d_uvw, d_meas, d_gk, d_dk, tile, chunk)
        if "enu" in d.extra:
            log("DFT precision variant ...")
            extras["dft_precision_variant"] = dft_precision_variant(hip, d)
        if not args.no_batch:
            log("batched multi-channel ...")
            extras["batched_channels"] = batched_channels(hip)
        if not args.no_learned:
            log("learned vs Hogbom ...")
            extras["learned_vs_hogbom"] = learned_vs_hogbom(
                hip, [dict(levels=3, base=16, pairs=3072, steps=args.learned_steps, head_scale=0.0),
                      dict(levels=4, base=32, pairs=args.learned_pairs, steps=args.learned_steps_4x32,
                           head_scale=0.0)])

    copy_gbs = copy_bandwidth(hip)
    out = {
        "metric": METRIC,
        "value": round(N / (ms_per_step * 1e-3) / 1e6, 2),
        "unit": "Mvis/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{d.name}: degrid(w-proj, weighted)+subtract+grid (one fused pass, then the merge) of {N} vis on a {G}^2 grid, "
                               f"{d.supports.shape[0]} w-planes h {d.supports[:, 0].min()}-{d.supports[:, 0].max()}, "
                               f"OS {d.oversampling}",
                   "n_vis": N, "grid": G, "w_planes": int(d.supports.shape[0]), "oversampling": d.oversampling,
                   "accumulate": args.accumulate},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic, "traffic_per": "step",
                     "alg_bytes_per_step": alg[dom], "ms_per_step": round(k_ms[dom], 4),
                     "copy_gbs": copy_gbs, "frac_of_copy": round(achieved / copy_gbs, 5) if copy_gbs else None,
                     "valu": {"kernel": "degrid_subtract_grid", "achieved": round(valu_tflops, 2), "peak": VALU_PEAK_TFLOPS,
                              "unit": "TFLOP/s", "frac": round(valu_tflops / VALU_PEAK_TFLOPS, 4),
                              "flops_per_tap": 8, "taps_per_step": 2 * taps,
                              "note": "prediction and gridding taps, 8 flops each, against non-packed FP32 FMA"}},
    }
    out.update(extras)
    if not args.no_cpu:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(d, args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
