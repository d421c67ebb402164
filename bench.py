"""Benchmark of the 2D forward single shot: 2048 x 2048 nodes, 8th-order Laplacian, seeded
smoothed random velocity (1500-4500 m/s) plus three layers, dx = 10 m, 40-node quadratic sponge,
Ricker 10 Hz source, 512 surface receivers, fp32.

A bench "step" is one leapfrog time step of that shot (one launch of the step kernel over the
whole grid, source injection and receiver sampling fused).  The default times the full 4000-step
shot.  Prints one JSON line with Gcell-updates/s, the HBM roofline of the step kernel, the
FWI gradient part of the metric (16 shots of 1024^2, order 8, 3000 steps: s/shot batched forward +
adjoint against plain per-shot torch autograd, `fwi_gradient`) and the CPU oracle timed on the
host cores on a bounded sample of the same workload.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
# Bound on the relative L2 error of the fp32 shot (traces, final field) against the same shot in
# fp64: fp32 rounding of kc and the weights accumulates to 3.3e-4 (traces) / 1.5e-3 (field) over
# 4200 steps; a time-reversed or otherwise wrong continuation reads 1e-1 and more.
FP64_BOUND = 1e-2


def workload(n=2048, nsteps=4000, dtype=None):
    import numpy as np
    import torch

    from neuralmech_amd import media, solver

    dtype = dtype or torch.float32
    h = (10.0, 10.0)
    v = media.forward_2d_model(n, seed=2)
    dt = 0.9 * solver.stable_dt(h, float(v.max()), 8)
    g = solver.WaveGrid((n, n), h, dt, 8, dtype)
    beta = [media.sponge_profile(n, 40, 0.05) for _ in range(2)]
    med = solver.Medium(g, velocity=torch.as_tensor(v), beta_profiles=beta)
    src = solver.PointSet(g, [(50, n // 2)])
    rec = solver.PointSet(g, [(45, int(j)) for j in np.round(np.linspace(1, n - 2, 512)).astype(int)])
    sig = media.ricker(np.arange(nsteps) * dt, 10.0) / (h[0] * h[1])
    return dict(grid=g, medium=med, src=src, rec=rec, sig=sig, v=v, dt=dt, h=h, beta=beta, n=n, nsteps=nsteps)


def cpu_baseline(w, budget_s=15.0):
    """C oracle (scalar restatement, OpenMP over rows) on the same model, bounded number of steps."""
    import numpy as np

    from oracle import c_oracle

    threads = host_threads()
    kc = ((w["v"] * w["dt"]) ** 2).astype(np.float32)
    rec = [(45, int(j)) for j in np.round(np.linspace(1, w["n"] - 2, 512)).astype(int)]
    kw = dict(beta_profiles=w["beta"], src_nodes=[(50, w["n"] // 2)], rec_nodes=rec, nthreads=threads)
    sig = w["sig"][:, None].astype(np.float32)
    c_oracle.forward(kc, 8, w["h"], 2, src_sig=sig[:2], **kw)  # thread start-up
    t0 = time.perf_counter()
    c_oracle.forward(kc, 8, w["h"], 20, src_sig=sig[:20], **kw)
    probe = time.perf_counter() - t0
    # about budget_s of host work, at most the whole shot
    steps = int(max(20, min(len(sig), budget_s / max(probe / 20, 1e-5))))
    t0 = time.perf_counter()
    traces, _, _ = c_oracle.forward(kc, 8, w["h"], steps, src_sig=sig[:steps], **kw)
    el = time.perf_counter() - t0
    cells = (w["n"] - 2) ** 2 * steps
    return ({"value": cells / el / 1e9, "unit": "Gcell-updates/s", "cores": threads, "kind": "port",
             "sample": f"C oracle, {steps} of {w['nsteps']} steps of the same 2048^2 order-8 shot "
                       f"(sponge, source, 512 receivers), {threads} OpenMP threads, {el:.1f} s",
             "host": host_info(), "dtype": "f32"}, traces)


def host_threads() -> int:
    """Host threads for the CPU baseline: the CPUs this process may run on, at most OMP_NUM_THREADS
    when that is set (the share of a machine's cores given to this job)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    omp = os.environ.get("OMP_NUM_THREADS", "")
    return max(1, min(n, int(omp))) if omp.isdigit() and int(omp) > 0 else max(1, n)


def host_info() -> dict:
    """The host the CPU baseline ran on: CPU model, os.cpu_count(), usable CPUs, torch threads."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None
    try:
        import torch

        tthreads = torch.get_num_threads()
    except ImportError:
        tthreads = None
    return {"cpu_model": model, "os_cpu_count": os.cpu_count(), "usable_cpus": usable, "torch_threads": tthreads}


def fwi_main(args):
    """Batched forward + adjoint FWI gradient (16 shots, 1024^2, order 8, 3000 steps) against the
    plain per-shot torch autograd form; s/shot for both, gradient error against the fp64 run."""
    print(json.dumps(fwi_result(args)))


def fwi_result(args) -> dict:
    """The FWI gradient measurement of fwi_main as a dict."""
    import torch

    from neuralmech_amd import autograd, fwi, solver

    nsteps = args.fwi_steps
    w = fwi.setup(nsteps=nsteps)
    tuned = None
    if not args.no_tune:  # launch shape of the 16-shot grid, timed once on scratch fields (untimed)
        tuned = w["grid"].tune(solver.Medium(w["grid"], velocity=torch.as_tensor(w["start"]),
                                             beta_profiles=w["beta"]), steps=20)
    obs = fwi.observed(w)
    torch.cuda.synchronize()
    def plans():
        return [{"seg": p.seg, "checkpoints": p.nckpt, "GB": round(p.bytes / 1e9, 1)}
                for p in autograd._PLAN_CACHE.values()]

    # the configuration's sqrt(n) checkpointing (value), then every forward level kept in device
    # memory (no recompute; same gradient bits), each after an untimed warm-up that creates its plan
    fwi.batched_gradient(w, obs)
    (J, grad), t_b = fwi.timed(fwi.batched_gradient, w, obs)
    plan_info = plans()
    fwi.batched_gradient(w, obs, seg=solver.SEG_FASTEST)
    (_, grad_k), t_k = fwi.timed(fwi.batched_gradient, w, obs, seg=solver.SEG_FASTEST)
    kept = {"plan": plans(), "bit_identical": bool(torch.equal(grad_k, grad))}
    del grad_k
    autograd.clear_plans()
    w64 = fwi.setup(nsteps=nsteps, dtype=torch.float64)
    (J64, grad64), _ = fwi.timed(fwi.batched_gradient, w64, obs.double())
    err64 = float((grad - grad64).norm() / grad64.norm())
    autograd.clear_plans()  # the plain leg's autograd tape needs the memory
    plain_shots = max(0, args.plain_shots)  # 0: no plain leg (e.g. for counter runs)
    t_p = 0.0
    for s in range(plain_shots):
        _, dt_s = fwi.timed(fwi.plain_shot_gradient, w, obs[s], s)
        t_p += dt_s
    s_b = t_b / w["nshots"]
    s_p = t_p / plain_shots if plain_shots else None
    cells = (w["n"] - 2) ** 2 * nsteps
    out = {
        "metric": "FWI gradient s/shot (batched forward+adjoint)",
        "value": round(s_b, 5),
        "unit": "s/shot",
        "n_gpus": 1,
        "steps": nsteps,
        "warmup": 1,
        "ms_per_step": round(t_b / nsteps * 1e3, 5),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded layered model +-10 %, smoothed start, Ricker shots)",
        "config": {"workload": "2D FWI gradient, 16 shots batched: 1024x1024 nodes, order 8, 201 receivers, "
                               f"{nsteps} steps, sqrt(n) checkpointing", "shots": w["nshots"],
                   "launch": {"in_use": list(w["grid"].launch()), "tuned": tuned},
                   "plan": plan_info},
        "plain_s_per_shot": None if s_p is None else round(s_p, 4),
        "speedup_vs_plain": None if s_p is None else round(s_p / s_b, 2),
        "kept_levels": dict(kept, value=round(t_k / w["nshots"], 5), unit="s/shot",
                            speedup_vs_plain=None if s_p is None else round(s_p * w["nshots"] / t_k, 2),
                            note="every forward level kept in device memory instead of sqrt(n) "
                                 "checkpoints and recompute (WFD_SEG_FASTEST)"),
        "plain_sample": f"{plain_shots} shot(s) of torch-on-GPU slice-form step + autograd",
        "grad_rel_l2_vs_fp64": err64,
        "misfit": J,
        "gcell_updates_per_s_gradient": round(3 * cells * w["nshots"] / t_b / 1e9, 2),
    }
    autograd.clear_plans()
    return out


def uneven_main(args):
    """Uneven batch: 64 shots of different grid sizes (512^2-1536^2), step counts (1500-5000) and
    orders (4/8), every shot's forward + adjoint gradient in one persistent launch replayed from a
    hipGraph; against the same shots run one by one through the single-shot gradient plan."""
    import torch

    from neuralmech_amd import uneven

    t0 = time.perf_counter()
    shots = uneven.setup(nshots=args.uneven_shots, step_scale=args.uneven_scale)
    t_setup = time.perf_counter() - t0
    cs = uneven.cell_steps(shots)
    t0 = time.perf_counter()
    b = uneven.batched(shots, tune_steps=0 if args.no_tune else 10, max_active_cells=args.max_active)
    t_tune = time.perf_counter() - t0
    info = b.info()
    classes = {}
    for k in range(len(shots)):
        c, v, r, seg, _, _ = b.shot_info(k)
        classes[str(c)] = [v, r]
    stream = torch.cuda.current_stream()
    for _ in range(max(1, args.warmup if args.warmup < 10 else 1)):  # graph capture + replay
        b.run(graph=True)
    b.status()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    runs = max(1, args.uneven_runs)
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(runs):
        b.run(graph=True)
    e1.record(stream)
    b.status()
    t_run = e0.elapsed_time(e1) / 1e3 / runs
    t0 = time.perf_counter()
    grads = b.gradients()
    torch.cuda.synchronize()
    t_fin = time.perf_counter() - t0
    t_batch = t_run + t_fin
    # one launch per phase, same tiles (the persistent kernel's barrier against kernel boundaries)
    t_phases = None
    if not args.no_phases:
        bp = uneven.batched(shots, tune_steps=0, max_active_cells=args.max_active, mode="phases")
        for c, (v, r) in classes.items():
            bp.set_config(v, r, int(c))
        bp.run(graph=True)
        bp.status()
        torch.cuda.synchronize()
        e0.record(stream)
        bp.run(graph=True)
        e1.record(stream)
        bp.status()
        t_phases = e0.elapsed_time(e1) / 1e3
        del bp
    del b
    t_seq, diff = None, None
    if not args.no_sequential:
        seq, t_seq = uneven.sequential(shots)
        diff = max(float((a[1] - s_).norm() / s_.norm()) for a, s_ in zip(grads, seq))
        del seq
    nshots = len(shots)
    plain = None
    if args.uneven_plain > 0:  # plain per-shot torch autograd on a sample, extrapolated by cell-steps
        from neuralmech_amd import fwi

        torch.cuda.empty_cache()
        pick = uneven.plain_sample(shots, args.uneven_plain)
        t_p, cs_p, errs = 0.0, 0, []
        for k in pick:
            (_, gp), dt_k = fwi.timed(uneven.plain_shot_gradient, shots[k])
            t_p += dt_k
            cs_p += uneven.cell_steps([shots[k]])
            errs.append(float((gp.double() - grads[k][1]).norm() / grads[k][1].norm()))
            del gp
            torch.cuda.empty_cache()
        t_plain_all = t_p / cs_p * cs
        plain = {"plain_s_per_shot": round(t_plain_all / nshots, 4),
                 "speedup_vs_plain": round(t_plain_all / t_batch, 2),
                 "plain_sample": f"shots {pick} ({cs_p / 1e9:.1f} G cell-steps, {t_p:.1f} s) of torch-on-GPU "
                                 "slice-form step + autograd, extrapolated to all shots by cell-steps",
                 "plain_grad_rel_l2_vs_batch": errs}
    gbs = 64.0 * cs / t_run / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "uneven_batch_traffic.json")
    if os.path.exists(tpath) and nshots == 64 and args.uneven_scale == 1.0:
        with open(tpath) as f:
            traffic = json.load(f)["bytes_per_run"]
    out = {
        "metric": "FWI gradient s/shot (uneven batch, one persistent launch)",
        "value": round(t_batch / nshots, 5),
        "unit": "s/shot",
        "n_gpus": 1,
        "steps": runs,
        "warmup": 1,
        "ms_per_step": round(t_run * 1e3, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded layered models +-10 %, smoothed start models, Ricker shots)",
        "config": {"workload": f"uneven batch: {nshots} shots, n ~ U{{512..1536}}^2, steps ~ U{{1500..5000}}"
                               f" x {args.uneven_scale}, order 4/8, 201 receivers, sqrt(n) checkpointing, rng(4)",
                   "shots": nshots, "cell_steps": cs, "phases": info["phases"], "blocks": info["blocks"],
                   "store_gb": round(info["bytes"] / 1e9, 2), "tiles_per_class": classes},
        "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "note": "64 algorithmic B per cell-step (forward 16 + recompute 20 + adjoint 28); "
                             "traffic = FETCH/WRITE_SIZE bytes of one persistent launch"},
        "gcell_updates_per_s": round(cs / t_run / 1e9, 2),
        "batch_run_s": round(t_run, 4),
        "finalize_s": round(t_fin, 4),
        "per_phase_launches_s": None if t_phases is None else round(t_phases, 4),
        "sequential_s_per_shot": None if t_seq is None else round(t_seq / nshots, 5),
        "speedup_vs_sequential": None if t_seq is None else round(t_seq / t_batch, 2),
        "grad_rel_l2_vs_sequential": diff,
        **(plain or {}),
        "setup_s": round(t_setup, 1),
        "tune_s": round(t_tune, 1),
    }
    print(json.dumps(out))


def roofline_main(args):
    """The reference's roofline harness (2D_wave_roofline.py) over the reference-layout kernels:
    copy ceiling, fp32 FMA peak, Gflop/s at 18 flop per dof on 1600 x 3200 for 2000 steps."""
    from neuralmech_amd import harness

    r = harness.roofline(steps=args.steps if args.steps != 4000 else 2000)
    padded = next(k for k in r["kernels"] if k["kernel"] == "padded")
    out = {"metric": "Gflop/s at 18 flop/dof, padded_step drop-in (1600x3200, 2000 steps)",
           "value": round(padded["gflops"], 1), "unit": "Gflop/s", "n_gpus": 1, "steps": r["steps"], "warmup": 20,
           "ms_per_step": round(padded["s_per_step"] * 1e3, 5), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32", "data": "synthetic (uniform random fields, as the reference)",
           "config": {"workload": "reference roofline harness, 1600x3200 nodes, order 2, constant c"},
           "bandwidth_GBs": round(r["bandwidth_GBs"], 1), "peak_Gflops": round(r["peak_Gflops"], 1),
           "kernels": [{k: (round(v, 6) if isinstance(v, float) else v) for k, v in x.items()} for x in r["kernels"]]}
    if args.csv_dir:
        harness.save_csv(os.path.join(args.csv_dir, "wave_roofline.csv"), kernel=[k["kernel"] for k in r["kernels"]],
                         intensity=[k["intensity"] for k in r["kernels"]], gflops=[k["gflops"] for k in r["kernels"]])
        harness.save_csv(os.path.join(args.csv_dir, "wave_roofline_ceilings.csv"), bandwidth=[r["bandwidth_GBs"]],
                         peak=[r["peak_Gflops"]])
    print(json.dumps(out))


def scaling_main(args):
    """The reference's grid-size sweep (2D_wave_scaling.py): dofs against s/step for the padded
    (fp32) and mixed (fp16) drop-ins, 20-step graphs."""
    from neuralmech_amd import harness

    res = {}
    for version, kind in enumerate(("padded", "mixed")):
        dofs, t = harness.scaling(kind, max_resolution=args.scaling_max, steps=args.scaling_steps)
        res[kind] = (dofs, t)
        if args.csv_dir:
            harness.save_csv(os.path.join(args.csv_dir, f"wave_scaling_{version}.csv"), x=dofs, y=t)
    d, t = res["padded"]
    thr = [a / b for a, b in zip(d, t)]
    best = max(range(len(thr)), key=lambda i: thr[i])
    out = {"metric": "peak dofs/s over grid sizes, padded_step drop-in", "value": round(thr[best] / 1e9, 3),
           "unit": "Gdofs/s", "n_gpus": 1, "steps": args.scaling_steps, "warmup": 20,
           "ms_per_step": round(t[best] * 1e3, 5), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32", "data": "synthetic (uniform random fields)",
           "config": {"workload": f"reference scaling sweep, square grids up to {max(d) ** 0.5 + 2:.0f}^2",
                      "peak_at_dofs": d[best]},
           "mixed_peak_Gdofs": round(max(a / b for a, b in zip(*res["mixed"])) / 1e9, 3),
           "largest_dofs": max(d)}
    print(json.dumps(out))


def copy_main(args):
    """Device-to-device copy ceiling (the reference measures the same with a 1 GiB copy,
    2D_wave_roofline.py:46-57), at 1 GiB (HBM) and 64 MiB (fits the Infinity Cache); also the
    calibration run for the FETCH_SIZE / WRITE_SIZE counters."""
    import torch

    out = {}
    for label, nbytes in (("1GiB", 1 << 30), ("64MiB", 64 << 20)):
        a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        reps = 20 if nbytes >= (1 << 30) else 400
        b.copy_(a)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            b.copy_(a)
        e1.record()
        torch.cuda.synchronize()
        s = e0.elapsed_time(e1) / 1e3
        out[label] = {"GB_per_s": round(2 * nbytes * reps / s / 1e9, 1), "bytes_read_per_copy": nbytes,
                      "bytes_written_per_copy": nbytes, "copies": reps + 1}
        del a, b
    # hand-written streaming kernels (wfd_stream_bench): copy, the stencil's 3-read / 1-write mix
    # (plain and with the nontemporal hint), read-only; 1 GiB per array (HBM) and 32 MiB (cache)
    import ctypes

    from neuralmech_amd import _lib

    lib = _lib.load()
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    kernels = {}
    for label, nbytes in (("1GiB", 1 << 30), ("32MiB", 32 << 20)):
        n = nbytes // 4
        arr = [torch.ones(n, dtype=torch.float32, device="cuda") for _ in range(4)]
        reps = 10 if nbytes >= (1 << 30) else 200
        row = {}
        for mode, name, moved in ((0, "copy", 2), (1, "read3_write1", 4), (2, "read3_write1_nontemporal", 4),
                                  (3, "read", 1)):
            def launch():
                _lib.check(lib.wfd_stream_bench(arr[0].data_ptr(), arr[1].data_ptr(), arr[2].data_ptr(),
                                                arr[3].data_ptr(), n, mode, st))
            launch()
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                launch()
            e1.record()
            torch.cuda.synchronize()
            row[name] = round(moved * nbytes * reps / (e0.elapsed_time(e1) / 1e3) / 1e9, 1)
        if not torch.equal(arr[3][: n // 4096], torch.full((n // 4096,), 4096.0, device="cuda")):
            raise RuntimeError("wfd_stream_bench: wrong block sums")
        kernels[label] = row
        del arr
    print(json.dumps({"metric": "device copy bandwidth", "unit": "GB/s", "results": out,
                      "stream_kernels": kernels}))


def volume_main(args):
    """3D forward (and, with --grad, forward+adjoint) of 512^3 nodes, order 8, sponge 32, Ricker
    8 Hz at (40, 256, 256), 128 x 128 receivers on plane 36."""
    import numpy as np
    import torch

    from neuralmech_amd import media, solver

    n, nsteps = args.n3d, args.steps3d
    h = (10.0, 10.0, 10.0)
    v = media.volume_3d_model(n, seed=5, device="cuda")
    dt = 0.9 * solver.stable_dt(h, float(v.max()), 8)
    g = solver.WaveGrid((n, n, n), h, dt, 8, torch.float32)
    beta = [media.sponge_profile(n, 32, 0.05) for _ in range(3)]
    med = solver.Medium(g, velocity=v, beta_profiles=beta)
    del v
    torch.cuda.empty_cache()
    src = solver.PointSet(g, [(40, n // 2, n // 2)])
    idx = 64 + 3 * np.arange(128)
    idx = idx[idx < n - 1]
    rec = solver.PointSet(g, [(36, int(i), int(j)) for i in idx for j in idx])
    sig = media.ricker(np.arange(nsteps) * dt, 8.0) / (h[0] * h[1] * h[2])
    tuned = None
    if args.launch:
        g.set_launch(*[int(x) for x in args.launch.split(",")])
    elif not args.no_tune and not args.two_step:
        tuned = g.tune(med, steps=6)
    # staged outside the timing; --two-step P: two steps per launch, P planes per block
    run = solver.ForwardRun(g, med, nsteps, src=src, sig=sig, rec=rec, two_step=args.two_step)
    run.advance(4)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    run.advance(nsteps - 4)
    e1.record()
    torch.cuda.synchronize()
    traces = run.traces[:, 4:]
    ms = e0.elapsed_time(e1)
    steps = nsteps - 4
    cells = g.interior_cells
    ms_step = ms / steps
    bpc = 10 if args.two_step else 16  # u_prev, u_cur, kc read and u_next written once per (two) step(s)
    achieved = bpc * cells / (ms_step / 1e3) / 1e9
    out = {
        "metric": "Gcell-updates/s (3D forward, fp32)", "value": round(cells * steps / (ms / 1e3) / 1e9, 3),
        "unit": "Gcell-updates/s", "n_gpus": 1, "steps": steps, "warmup": 4, "ms_per_step": round(ms_step, 4),
        "higher_is_better": True, "dtype": "f32", "data": "synthetic (seeded smoothed random velocity)",
        "config": {"workload": f"3D forward {n}^3 nodes, order 8 (25-point), 32-node sponge, Ricker 8 Hz, "
                               f"{rec.count} receivers",
                   "launch": ({"two_step_planes": args.two_step} if args.two_step
                              else {"rows": g.launch()[1], "tuned": tuned})},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "bytes_per_cell": bpc},
        "finite": bool(torch.isfinite(traces).all()),
    }
    del run, traces
    if args.grad:
        plan = solver.GradientPlan(g, nsteps)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        tr = plan.forward(med, src, sig, rec)
        gk = plan.backward(med, src, rec, tr)
        torch.cuda.synchronize()
        tg = time.perf_counter() - t0
        out["gradient"] = {"seconds": round(tg, 3), "checkpoint_spacing": plan.seg, "checkpoints": plan.nckpt,
                           "plan_GB": round(plan.bytes / 1e9, 2), "finite": bool(torch.isfinite(gk).all()),
                           "bytes_per_cell_step": 64,
                           "achieved_GB_per_s": round(64 * cells * nsteps / tg / 1e9, 1)}
    print(json.dumps(out))


def acoustic_main(args):
    """Variable mass / stiffness model at the analog-RNN driver's size
    (projects/4_specialized_nns/analog_rnn_fixture.py:22-45: 3000 x 1500 region, 10 m sponge on
    every edge, air and polymer foam, fp32, order 2): forward steps with the reconstruction strip
    recorded, then the reconstruction adjoint over the same steps."""
    import numpy as np
    import torch

    from neuralmech_amd import acoustic as ac
    from neuralmech_amd import boundary, media
    from neuralmech_amd.solver import PointSet

    res, lengths = (3000, 1500), (200.0, 100.0)
    dx = tuple(lengths[d] / (res[d] - 3) for d in range(2))
    Nx, width, origin, region = boundary.pad_for_sponge(res, dx, 10.0)
    rho1, rho2, k1, k2 = 1.204, 30.0, 1.419e5, 9.72e5
    cmax = max(np.sqrt(k1 / rho1), np.sqrt(k2 / rho2))
    from neuralmech_amd import wave

    dt = 0.99 * wave.stable_dt(dx, cmax, 2)
    nsteps = args.acoustic_steps
    sim = ac.AcousticWave(Nx, dx, nsteps, dt, precision="float32", rho1=rho1, rho2=rho2, kappa1=k1, kappa2=k2)
    prof = boundary.sponge(sim, None, width, 0.05)
    rng = np.random.default_rng(0)
    gamma = np.zeros(Nx)
    blk = (rng.random(((res[0] // 2) // 25, res[1] // 25)) > 0.5).repeat(25, 0).repeat(25, 1)
    x0, y0 = region[0].start + res[0] // 4, region[1].start
    gamma[x0:x0 + blk.shape[0], y0:y0 + blk.shape[1]] = blk
    grid = sim.grid()
    m, s = sim.material(torch.as_tensor(gamma))
    med = ac.AcMedium(grid, m, s, "neumann", prof)
    src = PointSet(grid, [(region[0].start + 150, region[1].start + 750)])
    rec = PointSet(grid, [(region[0].stop - 150, region[1].start + j) for j in (375, 750, 1125)])
    sn = med.strip_nodes()
    strip = PointSet(grid, [tuple(int(c) for c in p) for p in sn])
    sig = torch.as_tensor(media.ricker(np.arange(nsteps) * dt, 50.0)[None, :, None] / (dx[0] * dx[1]))
    cells = (Nx[0] - 2) * (Nx[1] - 2)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    ac.forward(grid, med, 20, src, sig[:, :20], rec, strip)  # warm-up
    torch.cuda.synchronize()
    ev[0].record()
    tr, store, state = ac.forward(grid, med, nsteps, src, sig, rec, strip)
    ev[1].record()
    dJ = 2.0 * tr.double()
    ev[2].record()
    G, mu_src = ac.backward(grid, med, nsteps, state, src, sig, rec, dJ, strip, store)
    ev[3].record()
    torch.cuda.synchronize()
    tf = ev[0].elapsed_time(ev[1]) / 1e3
    tb = ev[2].elapsed_time(ev[3]) / 1e3
    und = int(med.undamped.sum())
    fwd_bytes = 24.0 * cells
    bwd_bytes = 60.0 * und + 32.0 * (cells - und)  # imaging only where nothing damps
    out = {
        "metric": "Gcell-updates/s, variable mass/stiffness forward and reconstruction adjoint",
        "value": round(cells * nsteps / tf / 1e9, 2), "unit": "Gcell-updates/s", "n_gpus": 1, "steps": nsteps,
        "warmup": 20, "ms_per_step": round(tf / nsteps * 1e3, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded binary air/foam design, Ricker source)",
        "config": {"workload": f"AcousticWave {Nx[0]}x{Nx[1]} nodes (3000x1500 region + {width}-node sponge), "
                               "order 2, no-flux walls, air/foam", "strip_nodes": int(len(sn)),
                   "strip_MB_per_1000_steps": round(1001 * len(sn) * 4 / 1e6, 1)},
        "roofline": {"bound": "hbm", "achieved": round(fwd_bytes * nsteps / tf / 1e9, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(fwd_bytes * nsteps / tf / 1e9 / HBM_PEAK_GBS, 4),
                     "traffic": None, "note": "forward: 24 algorithmic B per cell-step"},
        "backward": {"gcell_updates_per_s": round(cells * nsteps / tb / 1e9, 2),
                     "ms_per_step": round(tb / nsteps * 1e3, 4),
                     "achieved_GB_per_s": round(bwd_bytes * nsteps / tb / 1e9, 1),
                     "frac": round(bwd_bytes * nsteps / tb / 1e9 / HBM_PEAK_GBS, 4),
                     "note": "60 B per undamped cell-step (7 fields read, 2 written, 3 fp32 partial sums "
                             "read and written), 32 B per damped cell-step",
                     "finite": bool(torch.isfinite(G).all())},
    }
    print(json.dumps(out))


def plate_main(args):
    """The plate FWI of projects/18_fwi (plate.py:19-47, 2D_fwi.py:21-43): 512 x 512 nodes, 32
    elements on the top edge, a void indicator measured at order 8, every fourth element firing;
    one iteration = the density-indicator misfit gradient of all 8 shots at order 4 (batched
    forward recording the reconstruction strip + reconstruction adjoint) and an Adam step, as the
    driver's loop; reports the reconstruction's relative L2 error and void F1 score."""
    import math

    import numpy as np
    import torch

    from neuralmech_amd import geometry, signals, utils, wave
    from neuralmech_amd.scalar import ScalarWave

    Nx, L, T, safety = (512, 512), (1.0, 1.0), 2.5, 0.9
    dx = tuple(l / (n - 3) for l, n in zip(L, Nx))
    elements = utils.line((0.1, L[1]), (0.9, L[1]), 32)

    def simulation(order):
        N = math.ceil(T / (safety * wave.stable_dt(dx, 1.0, order))) + 1
        sim = ScalarWave(Nx, dx, N, T / (N - 1), (4, 64), precision="float32", space_order=order, wavespeed=1.0,
                         density=1.0)
        t = np.linspace(0, T, N)
        return sim, t, utils.shots(sim, elements, signals.sineburst(t, 1.0, 20.0, 2)), utils.Sensors(sim, elements)

    sim_obs, t_obs, src_obs, sens_obs = simulation(8)
    sim, t, src, sens = simulation(4)
    coords = wave.grid_coords(Nx, dx)
    void = np.zeros(Nx, bool)
    for k, y in enumerate((0.15, 0.25, 0.35, 0.45)):  # four voids stacked along the array axis
        void |= geometry.circle(coords, (0.5, y), 0.02 + 0.01 * k)
    truth = np.where(void, 1e-4, 1.0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    observed = utils.measure(sim_obs, src_obs[::4], truth, sens_obs)
    torch.cuda.synchronize()
    t_meas = time.perf_counter() - t0
    observed = [utils.resample(r, t_obs[1], t[1], sim.N) for r in observed]
    from neuralmech_amd import evals, optimization

    gamma = torch.ones(Nx, dtype=torch.float64, device="cuda")
    utils.misfit_gradient(sim, src[::4][:1], gamma, sens, observed[:1])  # warm-up
    iters = max(1, args.plate_iters)
    costs = []
    optimizer = optimization.Adam(lr=0.1)  # 2D_fwi.py:35-41
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(iters):
        J, g = utils.misfit_gradient(sim, src[::4], gamma, sens, observed)
        costs.append(J)
        gamma = optimizer.step(gamma, g).clamp(1e-3, 1.0)
    torch.cuda.synchronize()
    t_it = (time.perf_counter() - t0) / iters
    inner = utils.interior_slice(sim)
    tr = torch.as_tensor(truth, device="cuda")
    cells = (Nx[0] - 2) * (Nx[1] - 2) * 8 * sim.N
    out = {
        "metric": "plate FWI ms per iteration (density-indicator misfit gradient, 8 shots)",
        "value": round(t_it * 1e3, 2), "unit": "ms/iteration", "n_gpus": 1, "steps": iters, "warmup": 1,
        "ms_per_step": round(t_it / sim.N * 1e3, 4), "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (four circular voids, gamma 1e-4; sine-burst array)",
        "config": {"workload": f"plate FWI: {Nx[0]}x{Nx[1]} nodes, order 4 inversion / order 8 measurement, "
                               f"32 elements, 8 shots x {sim.N} steps", "steps_per_shot": sim.N,
                   "measurement_order8_s": round(t_meas, 3)},
        "gcell_updates_per_s_forward_plus_adjoint": round(2 * cells / t_it / 1e9, 2),
        "misfit_first_last": [round(costs[0], 6), round(costs[-1], 6)],
        "l2_error": round(evals.l2_error(gamma[inner], tr[inner]), 4),
        "f1_voids": round(evals.f1_score(gamma[inner], tr[inner], threshold=0.5), 4),
    }
    print(json.dumps(out))


def tfm_main(args):
    """The TFM imager of projects/18_fwi (2D_tfm.py, 2D_tfm_scaling.py): the plate's full matrix
    capture (32 elements, order 8, voids minus no voids) from this solver, its analytic signal,
    then the delay-and-sum image at sides 32 .. --tfm-max (powers of two), the reciprocal kernel
    form; the array form and the NumPy restatement of it on the host (up to 256^2, as the
    reference's cpu run) beside it."""
    import math

    import numpy as np
    import torch

    from neuralmech_amd import geometry, signals, tfm, utils, wave
    from neuralmech_amd.scalar import ScalarWave
    from oracle import tfm_ref

    Nx, L, T, safety, E = (512, 512), (1.0, 1.0), 2.5, 0.9, 32
    dx = tuple(l / (n - 3) for l, n in zip(L, Nx))
    elements = utils.line((0.1, L[1]), (0.9, L[1]), E)
    N = math.ceil(T / (safety * wave.stable_dt(dx, 1.0, 8))) + 1
    sim = ScalarWave(Nx, dx, N, T / (N - 1), (4, 64), precision="float32", space_order=8, wavespeed=1.0, density=1.0)
    t = np.linspace(0, T, N)
    sources = utils.shots(sim, elements, signals.sineburst(t, 1.0, 20.0, 2))
    sensors = utils.Sensors(sim, elements)
    coords = wave.grid_coords(Nx, dx)
    void = np.zeros(Nx, bool)
    for k, yv in enumerate((0.15, 0.25, 0.35, 0.45)):
        void |= geometry.circle(coords, (0.5, yv), 0.02 + 0.01 * k)
    truth = np.where(void, 1e-4, 1.0)
    torch.cuda.synchronize()
    t0_ = time.perf_counter()
    scattered = torch.stack(utils.measure(sim, sources, truth, sensors))
    scattered -= torch.stack(utils.measure(sim, sources, np.ones(Nx), sensors))
    torch.cuda.synchronize()
    t_fmc = time.perf_counter() - t0_
    h = tfm.analytic_signal(scattered.double(), axis=1).transpose(1, 2).contiguous().to(torch.complex64)
    t0 = 2 / (2 * 20.0)
    dt = t[1]
    sizes = [2 ** k for k in range(5, int(math.log2(args.tfm_max)) + 1)]
    rows = []
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    for n in sizes:
        x, y = torch.meshgrid(torch.linspace(0, L[0], n, device="cuda"), torch.linspace(0, L[1], n, device="cuda"),
                              indexing="ij")
        tfm.tfm_kernel(h, elements, t0, dt, 1.0, x, y)  # warm-up
        times = []
        for _ in range(3):
            e0.record()
            tfm.tfm_kernel(h, elements, t0, dt, 1.0, x, y)
            e1.record()
            torch.cuda.synchronize()
            times.append(e0.elapsed_time(e1) / 1e3)
        rows.append({"pixels": n * n, "kernel_s": min(times)})
    n_big = sizes[-1]
    pairs = E * (E + 1) // 2
    best = rows[-1]["kernel_s"]
    # host baseline: the reference's array form in NumPy at 256^2 (its cpu run stops there); the
    # GPU array form on the same pixels is compared with it
    xs, ys = np.meshgrid(np.linspace(0, 1, 256), np.linspace(0, 1, 256), indexing="ij")
    hc = h.cpu().numpy()
    t0_ = time.perf_counter()
    ref = tfm_ref.tfm_all_pairs(hc, np.asarray(elements), t0, dt, 1.0, xs, ys)
    t_cpu = time.perf_counter() - t0_
    img = tfm.tfm(h, elements, t0, dt, 1.0, torch.as_tensor(xs, dtype=torch.float32, device="cuda"),
                  torch.as_tensor(ys, dtype=torch.float32, device="cuda")).cpu().numpy()
    gpu_rel_diff = float(np.abs(img - ref).max() / np.abs(ref).max())
    out = {
        "metric": "TFM image time (32-element full matrix capture, reciprocal delay and sum)",
        "value": round(best * 1e3, 3), "unit": "ms/image", "n_gpus": 1, "steps": 3, "warmup": 1,
        "ms_per_step": round(best * 1e3, 3), "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (plate with four voids, this solver's order-8 full matrix capture)",
        "config": {"workload": f"TFM {n_big}^2 pixels, {E} elements, {h.shape[2]} samples", "fmc_s": round(t_fmc, 3)},
        "gpair_samples_per_s": round(n_big * n_big * pairs / best / 1e9, 2),
        "scaling_rows": rows,
        "cpu_baseline": {"value": round(t_cpu * 1e3, 1), "unit": "ms/image", "cores": 1, "kind": "port",
                         "sample": "NumPy array form of tfm.py:39-53 (all pairs), 256^2 pixels, float64",
                         "gpu_array_form_max_rel_diff": gpu_rel_diff, "host": host_info()},
    }
    if args.csv_dir:
        from neuralmech_amd import harness

        harness.save_csv(os.path.join(args.csv_dir, "tfm_scaling_kernel.csv"), x=[r["pixels"] for r in rows],
                         y=[r["kernel_s"] for r in rows])
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--graph", type=int, default=0, help="replay a hipGraph of this many steps (even)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-fwi", action="store_true", help="default workload: skip the FWI gradient part")
    ap.add_argument("--no-tune", action="store_true", help="use the default launch shape")
    ap.add_argument("--workload", choices=["forward", "fwi", "copy", "3d", "uneven", "roofline", "scaling",
                                           "acoustic", "plate", "tfm"],
                    default="forward")
    ap.add_argument("--launch", default="", help="forward / 3d: force VARIANT,ROWS instead of tuning")
    ap.add_argument("--two-step", type=int, default=0,
                    help="forward: two steps per launch, ROWS output rows per wave (2D) or planes per block (3D)")
    ap.add_argument("--n3d", type=int, default=512)
    ap.add_argument("--steps3d", type=int, default=2000)
    ap.add_argument("--grad", action="store_true", help="3D: also time the checkpointed gradient")
    ap.add_argument("--fwi-steps", type=int, default=3000)
    ap.add_argument("--plain-shots", type=int, default=1)
    ap.add_argument("--uneven-shots", type=int, default=64)
    ap.add_argument("--uneven-scale", type=float, default=1.0, help="scale the drawn step counts")
    ap.add_argument("--uneven-runs", type=int, default=2)
    ap.add_argument("--max-active", type=int, default=0, help="uneven: cells in flight (0 = default)")
    ap.add_argument("--no-phases", action="store_true", help="uneven: skip the per-phase launch timing")
    ap.add_argument("--no-sequential", action="store_true", help="uneven: skip the shot-by-shot baseline")
    ap.add_argument("--uneven-plain", type=int, default=3,
                    help="uneven: shots timed with plain per-shot torch autograd (0 = skip)")
    ap.add_argument("--csv-dir", default="", help="roofline / scaling: write the reference's CSV files here")
    ap.add_argument("--scaling-max", type=int, default=100000, help="scaling: largest grid side")
    ap.add_argument("--scaling-steps", type=int, default=400, help="scaling: steps timed per grid")
    ap.add_argument("--acoustic-steps", type=int, default=1000, help="acoustic: forward / adjoint steps timed")
    ap.add_argument("--plate-iters", type=int, default=5, help="plate: gradient iterations timed")
    ap.add_argument("--tfm-max", type=int, default=4096, help="tfm: largest image side")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one MI355X: the shot runs on a single GPU and this build has no "
              "multi-GPU path; run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.workload == "fwi":
        fwi_main(args)
        return
    if args.workload == "copy":
        copy_main(args)
        return
    if args.workload == "3d":
        volume_main(args)
        return
    if args.workload == "uneven":
        uneven_main(args)
        return
    if args.workload == "roofline":
        roofline_main(args)
        return
    if args.workload == "scaling":
        scaling_main(args)
        return
    if args.workload == "acoustic":
        acoustic_main(args)
        return
    if args.workload == "plate":
        plate_main(args)
        return
    if args.workload == "tfm":
        tfm_main(args)
        return

    import numpy as np
    import torch

    from neuralmech_amd import solver

    nsteps = args.warmup + args.steps
    w = workload(nsteps=max(nsteps, 4000))
    g, med, src, rec = w["grid"], w["medium"], w["src"], w["rec"]
    tuned = None
    if args.launch:
        g.set_launch(*[int(x) for x in args.launch.split(",")])
    elif not args.no_tune:
        tuned = g.tune(med, steps=40)
    # signal, trace buffer and descriptors staged once; the warm-up steps advance the shot and
    # the timed steps continue it from the state the warm-up left (levels and step index)
    run = solver.ForwardRun(g, med, nsteps, src=src, sig=w["sig"], rec=rec, graph_chunk=args.graph,
                            two_step=args.two_step)
    run.advance(args.warmup)
    run.prepare()  # graph capture, if any, outside the timing
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    t_start, t_end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t_start.record(stream)
    run.advance(args.steps)
    t_end.record(stream)
    torch.cuda.synchronize()
    ms = t_start.elapsed_time(t_end)
    traces = run.traces[:, args.warmup:nsteps]
    if run.state.step != nsteps or not torch.isfinite(traces).all():
        raise RuntimeError("the timed run did not end at the expected step with finite traces")
    cells = g.interior_cells
    ms_step = ms / args.steps
    value = cells * args.steps / (ms / 1e3) / 1e9
    bytes_per_launch = 16 * cells  # u_prev r, u_cur r, kc r, u_next w (fp32)
    achieved = bytes_per_launch / (ms_step / 1e3) / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "step2d_order8_2048_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("bytes_per_launch")
    working_set = 4 * g.elems * 4  # u_prev, u_cur, u_next, kc (fp32)
    out = {
        "metric": "Gcell-updates/s & HBM roofline %; FWI grad s/shot batched+adjoint vs plain",
        "value": round(value, 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded smoothed random velocity + layers, Ricker source)",
        "config": {"workload": "2D forward single shot: 2048x2048 nodes, order 8, dx 10 m, 40-node sponge, "
                               "Ricker 10 Hz, 512 receivers, 4000 steps",
                   "grid": [w["n"], w["n"]], "order": 8, "receivers": 512, "graph_chunk": args.graph,
                   "launch": {"variant": g.launch()[0], "rows": g.launch()[1], "tuned": tuned}},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": (("step2d_xl_kernel" if g.launch()[0] & 1 else "step2d_kernel")
                                + f"<float, K=4, VARC, SPONGE, rows {g.launch()[1]}"
                                + (", streaming>" if g.launch()[0] & 2 else ">")), "bytes_per_cell": 16,
                     "residency": f"working set {working_set / 2**20:.0f} MiB (u_prev, u_cur, u_next, kc) fits the "
                                  "256 MB Infinity Cache: traffic is fabric traffic (L2 misses), mostly served "
                                  "by the cache, not HBM; the clean HBM-resident case is the 3D workload"},
    }
    # error of the timed fp32 run against the same shot run in fp64 on the GPU (outside the timing)
    w64 = workload(nsteps=max(nsteps, 4000), dtype=torch.float64)
    tr64, st64 = solver.run(w64["grid"], w64["medium"], nsteps, src=w64["src"], sig=w64["sig"], rec=w64["rec"])
    torch.cuda.synchronize()
    t64 = tr64[:, args.warmup:nsteps].double()
    f32 = g.nodes(run.state.cur)[0].double()
    f64 = w64["grid"].nodes(st64.cur)[0].double()
    err_tr = float(((traces.double() - t64).norm() / t64.norm()).item())
    err_f = float(((f32 - f64).norm() / f64.norm()).item())
    out["fp64_gpu_oracle"] = {
        "traces_rel_l2": err_tr, "field_rel_l2": err_f, "bound": FP64_BOUND,
        "run": "the same shot in float64 on the GPU (float64 instantiation of the same kernels)"}
    if not args.no_fwi:  # the metric's second part: the FWI gradient, batched + adjoint vs plain autograd
        del w64, tr64, st64
        torch.cuda.empty_cache()
        f = fwi_result(args)
        out["fwi_gradient"] = {k: f[k] for k in ("value", "unit", "plain_s_per_shot", "speedup_vs_plain",
                                                 "grad_rel_l2_vs_fp64", "gcell_updates_per_s_gradient",
                                                 "plain_sample", "kept_levels")}
        out["fwi_gradient"]["workload"] = f["config"]["workload"]
        out["fwi_gradient"]["launch"] = f["config"]["launch"]
    parity_ok = err_tr <= FP64_BOUND and err_f <= FP64_BOUND
    if not args.no_cpu:
        cb, oracle_traces = cpu_baseline(w)
        rows = min(oracle_traces.shape[0], nsteps)
        same = bool(np.array_equal(run.traces[0, :rows].cpu().numpy(), oracle_traces[:rows]))
        cb["parity_vs_c_oracle"] = {"trace_rows": rows, "bit_exact": same}
        parity_ok = parity_ok and same
        out["cpu_baseline"] = cb
    print(json.dumps(out))
    if not parity_ok:
        print("bench: the timed fp32 run fails its parity checks (fp64 GPU run bound "
              f"{FP64_BOUND}, C oracle bit-exactness); the measurement is invalid", file=sys.stderr)
        sys.exit(1)


if __name__ == "__main__":
    main()
