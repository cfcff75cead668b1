"""Benchmark: MLUPS of the FUSED FP32 step on one MI355X, its HBM roofline
fraction, FP16-storage MLUPS, and the reference's CPU build timed on the
host cores beside it.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

One "step" is one lattice update of the whole 4096x4096 lid-driven cavity
(Re = 10^4: svisc = 0.04096, u0 = 0.1), the single-solve bandwidth
configuration (20,000 steps by default).  Populations are resident in HBM
before the timed region; the timed region is bgk_run(K) bracketed by
stream synchronisation.  Prints one JSON line.
"""
import argparse
import json
import os
import re
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "MLUPS (% HBM roofline) fp32 vs fp16-store; batched-sweep p50/p99 latency"
BYTES_PER_LUP = {"single": 72, "fp16": 36, "bf16": 36, "double": 144, "mixed2": 72}


def gpu_solve(n, prec, bc, steps, warmup, svisc, u0, tblock=0):
    """FP16/BF16 run with the configuration that meets the accuracy bar:
    scale 1024, deviation form, stochastic rounding.  tblock: FUSED steps
    per kernel (0: the library's choice, 1, 2)."""
    from bgk2d_gpu_amd import Solver
    narrow = prec in ("fp16", "bf16")
    s = Solver(n, n, bc=bc, impl="fused", prec=prec, svisc=svisc, u0=u0,
               scale=1024.0 if narrow else 1.0, amplitude=0.01, form=1 if narrow else 0,
               rounding=1 if narrow else 0, tblock=tblock)
    if warmup:
        s.run(warmup)
    s.sync()
    t0 = time.perf_counter()
    s.run(steps)
    s.sync()
    wall = time.perf_counter() - t0
    kernel_ms = s.last_run_ms()  # HIP events on the context stream, per step
    ok = s.check(0.5)
    diag = s.diagno()
    s.close()
    return dict(wall=wall, kernel_ms=kernel_ms, ok=ok, diag=diag.tolist())


def fp16_leg(steps):
    """FP16 shifted storage against FP32: Taylor-Green 8192^2 (tau = 0.8,
    u0 = 0.01), each with one and with two steps per pass; the ratio of the
    best of each"""
    n3 = 8192
    fp = {}
    for prec in ("fp16", "single"):
        for tb in (1, 2):
            r = gpu_solve(n3, prec, "periodic", steps, 50, 0.1, 0.0, tblock=tb)
            fp[(prec, tb)] = (n3 * n3 * steps / r["wall"] / 1e6, r)
    ml16 = max(fp[("fp16", 1)][0], fp[("fp16", 2)][0])
    ml32 = max(fp[("single", 1)][0], fp[("single", 2)][0])
    return {"workload": "Taylor-Green 8192x8192 periodic, tau=0.8, u0=0.01, "
                        f"{steps} steps, scale=1024, deviation form, stochastic rounding",
            "mlups_fp16": round(ml16, 1), "mlups_fp32": round(ml32, 1),
            "ratio": round(ml16 / ml32, 3),
            "one_step_per_pass": {
                "mlups_fp16": round(fp[("fp16", 1)][0], 1),
                "mlups_fp32": round(fp[("single", 1)][0], 1),
                "ratio": round(fp[("fp16", 1)][0] / fp[("single", 1)][0], 3),
                "hbm_frac_fp16": round(36 * n3 * n3 / (fp[("fp16", 1)][1]["kernel_ms"] / 1e3)
                                       / 1e9 / HBM_PEAK_GBS, 4),
                "kernel_ms_per_step_fp16": round(fp[("fp16", 1)][1]["kernel_ms"], 5)},
            "two_steps_per_pass": {
                "mlups_fp16": round(fp[("fp16", 2)][0], 1),
                "mlups_fp32": round(fp[("single", 2)][0], 1),
                "ratio": round(fp[("fp16", 2)][0] / fp[("single", 2)][0], 3),
                "kernel_ms_per_step_fp16": round(fp[("fp16", 2)][1]["kernel_ms"], 5)}}


def host_info():
    """nproc (this process's CPU set), the CPU model and the number of
    physical cores (lscpu)"""
    info = {"nproc": len(os.sched_getaffinity(0))}
    try:
        out = subprocess.run(["lscpu"], stdout=subprocess.PIPE, text=True, timeout=30).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                info["model"] = line.split(":", 1)[1].strip()
        cores = subprocess.run(["lscpu", "-p=Core,Socket"], stdout=subprocess.PIPE, text=True,
                               timeout=30).stdout
        info["physical_cores"] = len({l for l in cores.splitlines() if l and not l.startswith("#")})
    except (OSError, subprocess.SubprocessError):
        pass
    return info


def run_reference(exe, n, steps, threads=None):
    """one timed run of a reference build (oracle/_ref): its own loop MLUPS
    (HARNESS line, SRC/finalize.F90's definition) and an external wall clock
    around the whole process"""
    env = dict(os.environ)
    if threads:
        env["OMP_NUM_THREADS"] = str(threads)
    t0 = time.perf_counter()
    r = subprocess.run([exe, str(n), str(n), "0.04096", "0.1", str(steps), "16", "-"],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, cwd="/tmp",
                       timeout=600, env=env)
    wall = time.perf_counter() - t0
    m = re.search(r"loop_s=\s*([0-9.]+)\s+mlups=\s*([0-9.]+)", r.stdout)
    if not m:
        return None
    return {"mlups": float(m.group(2)), "loop_s": float(m.group(1)), "process_wall_s": round(wall, 3),
            "steps": steps, "threads": threads or 1}


def cpu_baseline():
    """The reference's own FUSED SINGLE LDC sources (oracle/_ref, flang -O3),
    on a bounded sample of the same 4096^2 workload: the serial build on one
    core, and the reference's two multicore forms on the host's physical
    cores (at most 16, the CPU share of one GPU): its do-concurrent loops
    mapped to OpenMP (the MULTICORE build) and its OpenMP-offload loops
    built without a device (target regions on the host)."""
    n = 4096
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_SINGLE_FUSED_LDC")
    host = host_info()
    if os.path.exists(exe):
        serial = run_reference(exe, n, 150)  # about 10 s of one core
        if serial:
            out = {"value": serial["mlups"], "unit": "MLUPS", "cores": 1, "kind": "reference",
                   "sample": f"4096x4096 LDC Re=1e4, {serial['steps']} steps, serial flang -O3 build "
                             "of the reference's FUSED SINGLE sources (oracle/_ref)",
                   "serial": serial, "host": host}
            threads = min(16, host.get("physical_cores") or host["nproc"], host["nproc"])
            mc = {}
            # flang's builds of both run 100-250 times slower than its serial
            # build on these cores (2 and 1.1 MLUPS on 16 EPYC cores against
            # 274 serial), so their samples are 2 and 1 steps
            for name, tag, steps in (("do_concurrent_openmp", "refmc", 2), ("openmp_target_on_host", "refomp", 1)):
                e = os.path.join(ROOT, "oracle", "_ref", f"{tag}_SINGLE_FUSED_LDC")
                if os.path.exists(e):
                    mc[name] = run_reference(e, n, steps, threads)
            out["multicore"] = mc
            best = max([v for v in mc.values() if v] + [serial], key=lambda v: v["mlups"])
            out["best"] = {"mlups": best["mlups"], "threads": best["threads"]}
            return out
    # fallback: the oracle (C restatement), one core
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    o = pyoracle.Oracle(n, n, bc="driven", impl="fused", prec="single", svisc=0.04096, u0=0.1)
    steps = 20
    t0 = time.perf_counter()
    o.step(steps)
    dt = time.perf_counter() - t0
    return {"value": n * n * steps / dt / 1e6, "unit": "MLUPS", "cores": 1, "kind": "port",
            "sample": f"4096x4096 LDC Re=1e4, {steps} steps, oracle C restatement", "host": host}


CX = (1, 1, 1, 0, -1, -1, -1, 0, 0)
CY = (-1, 1, 0, 1, -1, 1, 0, -1, 0)


def velocity(solver, scale):
    """u, v of the interior from the downloaded populations (float64)"""
    import numpy as np
    raw = solver.raw()[:, 1:-1, 1:-1]
    if solver.prec == "fp16":
        f = raw.view(np.float16).astype(np.float32).astype(np.float64) / scale
    elif solver.prec == "bf16":
        f = (raw.astype(np.uint32) << 16).view(np.float32).astype(np.float64) / scale
    else:
        f = raw.astype(np.float64)
    del raw
    rho = f.sum(axis=0) + 1.0
    u = sum(c * f[k] for k, c in enumerate(CX) if c) / rho
    v = sum(c * f[k] for k, c in enumerate(CY) if c) / rho
    return u, v


def precision_study(n=8192, steps=5000, amp=0.01, svisc=0.1, init_v=0):
    """Taylor-Green n^2, tau = 0.8, u0 = 0.01: velocity relative L2 error of
    every storage mode against the FP64 run, and against the analytic
    field u(t) = u(0) exp(-2 nu k^2 t)."""
    import numpy as np
    from bgk2d_gpu_amd import Solver
    res = {"workload": f"Taylor-Green {n}x{n}, tau=0.8 (svisc={svisc}), u0={amp}, {steps} steps, "
                       + ("rho=1 start" if init_v == 0 else "pressure-consistent density start"),
           "metric": "relative L2 = ||u-u_ref||_2/||u_ref||_2 over all nodes, both components"}
    s = Solver(n, n, bc="periodic", impl="fused", prec="double", svisc=svisc, amplitude=amp,
               init_v=init_v)
    s.run(steps)
    s.sync()
    ur, vr = velocity(s, 1.0)
    s.close()
    norm = np.sqrt((ur * ur + vr * vr).sum())
    x = (np.arange(1, n + 1) - 0.5) / n
    k = 2 * np.pi
    decay = np.exp(-svisc * 2 * (k / n) ** 2 * steps)  # nu |k|^2 t, |k|^2 = 2 (2 pi/n)^2
    ua = amp * np.sin(k * x)[None, :] * np.cos(k * x)[:, None] * decay
    va = -amp * np.cos(k * x)[None, :] * np.sin(k * x)[:, None] * decay
    anorm = np.sqrt((ua * ua + va * va).sum())
    res["double_vs_analytic"] = float(np.sqrt(((ur - ua) ** 2 + (vr - va) ** 2).sum()) / anorm)
    for prec, scale, form, rnd in (("single", 1.0, 0, 0), ("single", 1.0, 1, 0),
                                   ("fp16", 1.0, 0, 0), ("fp16", 1024.0, 1, 0),
                                   ("fp16", 1024.0, 1, 1), ("bf16", 1024.0, 1, 1)):
        s = Solver(n, n, bc="periodic", impl="fused", prec=prec, svisc=svisc, amplitude=amp,
                   scale=scale, form=form, init_v=init_v, rounding=rnd)
        s.run(steps)
        s.sync()
        u, v = velocity(s, scale)
        s.close()
        err = float(np.sqrt(((u - ur) ** 2 + (v - vr) ** 2).sum()) / norm)
        tag = f"{prec}_scale{int(scale)}_form{form}_{'sr' if rnd else 'rne'}"
        res[tag + "_vs_double"] = err
        erra = float(np.sqrt(((u - ua) ** 2 + (v - va) ** 2).sum()) / anorm)
        res[tag + "_vs_analytic"] = erra
        print(f"# {tag}: rel L2 {err:.3e} vs double, {erra:.3e} vs analytic",
              file=sys.stderr, flush=True)
        del u, v
    return res


def kernel_ceilings(steps=400):
    """MLUPS of bgk_run alone on a 4096^2 lattice of each serving kind and
    storage kind (the kernels a sequential solve of that size uses): the
    rate a perfectly packed batch could reach"""
    from bgk2d_gpu_amd import serve
    out = {}
    for kind in serve.KINDS:
        for prec in ("single", "fp16"):
            r = serve.Request(0, kind, 4096, 1000.0, steps, prec)
            lx, ly, bc, svisc, u0, amp = r.geometry()
            g = gpu_solve_xy(lx, ly, prec, bc, steps, 40, svisc, u0, amp)
            out[f"{kind}/{prec}"] = round(lx * ly * steps / g / 1e6, 1)
    return out


def gpu_solve_xy(lx, ly, prec, bc, steps, warmup, svisc, u0, amp):
    """wall seconds of bgk_run(steps) on an lx x ly lattice (serving settings)"""
    from bgk2d_gpu_amd import Solver
    narrow = prec in ("fp16", "bf16")
    s = Solver(lx, ly, bc=bc, impl="fused", prec=prec, svisc=svisc, u0=u0, amplitude=amp,
               scale=1024.0 if narrow else 1.0, form=1 if narrow else 0, rounding=1 if narrow else 0)
    s.run(warmup)
    s.sync()
    t0 = time.perf_counter()
    s.run(steps)
    s.sync()
    wall = time.perf_counter() - t0
    s.close()
    return wall


def serving_leg(n=48, seed=20261203, side=(128, 4096), steps=(1000, 50000), chunk=100,
                budget=200e9, fill=1 << 23):
    """The serving configuration: a seeded Poisson stream (200 req/s) of
    cavity / Poiseuille / Taylor-Green solves, sides log-uniform in
    128..4096, Re log-uniform in 100..1e4, steps uniform in 1k..50k, FP32
    or FP16 storage.  Served by continuous batching with FIFO admission,
    each chunk advancing the admitted requests with the least remaining
    work whose lattices hold at least `fill` nodes (2^23: the p50 below the
    sequential one at most of the all-member throughput); by the same loop
    advancing every admitted request each chunk (the most throughput); and
    by sequential FIFO solves with the same kernels.  Reports p50/p99
    time-to-solution, aggregate MLUPS and their ratios, and the batched
    aggregate as a fraction of the stream's kernel ceiling (its lattice
    updates at the bgk_run rate of each request's kind at 4096^2)."""
    from bgk2d_gpu_amd import serve
    reqs = serve.make_stream(n, seed=seed, side=side, steps=steps)
    tiles = {p: serve.autotune(p, 1024) for p in ("single", "fp16")}
    ceil = kernel_ceilings()
    bo, bs = serve.serve_batched(reqs, budget_bytes=budget, chunk=chunk, check_every=1000,
                                 tiles=tiles, fill_nodes=fill)
    ao, as_ = serve.serve_batched(reqs, budget_bytes=budget, chunk=chunk, check_every=1000,
                                  tiles=tiles)
    so, ss = serve.serve_sequential(reqs, check_every=1000)
    b, q = serve.summarize(reqs, bo, bs), serve.summarize(reqs, so, ss)
    a = serve.summarize(reqs, ao, as_)
    # the stream's kernel ceiling: its updates at each kind's bgk_run rate
    ok = [r for r in reqs if so[r.rid].status == "ok"]
    t_ceil = sum(r.lups / (ceil[f"{r.kind}/{r.prec}"] * 1e6) for r in ok)
    ceiling = sum(r.lups for r in ok) / t_ceil / 1e6 if t_ceil else None
    return {"stream": f"{n} requests, seed {seed}, sides log-uniform {side[0]}..{side[1]}, "
                      f"Re log-uniform 100..1e4, steps uniform {steps[0]}..{steps[1]}, "
                      "FP32/FP16 50/50, Poisson 200 req/s",
            "lattice_updates": int(sum(r.lups for r in reqs)),
            "batched": dict(b, policy=f"shortest remaining work first, >= {fill} nodes per chunk"),
            "batched_all_members": a, "sequential": q,
            "aggregate_mlups_ratio": round(b["mlups"] / q["mlups"], 3),
            "all_members_mlups_ratio": round(a["mlups"] / q["mlups"], 3),
            "batched_p50_over_sequential_p50": round(b["p50_s"] / q["p50_s"], 3),
            "all_members_p50_over_sequential_p50": round(a["p50_s"] / q["p50_s"], 3),
            "kernel_ceilings_4096": ceil,
            "stream_kernel_ceiling_mlups": round(ceiling, 1) if ceiling else None,
            "batched_over_ceiling": round(b["mlups"] / ceiling, 3) if ceiling else None,
            "all_members_over_ceiling": round(a["mlups"] / ceiling, 3) if ceiling else None,
            "tiles": tiles,
            "same_outcomes": all(bo[r.rid].status == so[r.rid].status == ao[r.rid].status for r in reqs),
            "same_diagnostics": all(bo[r.rid].diag == so[r.rid].diag == ao[r.rid].diag
                                    for r in reqs if so[r.rid].status == "ok")}


def c4_leg(steps=2000):
    """Channel past a cylinder: 8192 x 2048 OPEN channel (parabolic inflow
    of peak 0.05, zero-gradient outflow, free-slip walls), cylinder of
    diameter 128 centred at x = 1024 and one node above mid-height,
    Re = u D / nu = 200 (nu = 0.032, tau = 0.596); FP32 and FP16 storage
    (scale 1024, deviation form, stochastic rounding); MLUPS over `steps`
    steps and the drag and lift coefficients Fx,y / (u^2 R) at the end.  The
    channel starts at rest and the inflow front needs about 20,000 steps to
    reach the cylinder, so after a short run these coefficients are those of
    the start-up pressure waves passing it (negative drag included): they
    show FP32 and FP16 agreeing, not the shedding values."""
    import numpy as np
    from bgk2d_gpu_amd import Solver
    l, m, D, u = 8192, 2048, 128, 0.05
    nu = u * D / 200.0
    xc, yc, r = 1024, m // 2 + 1, D // 2
    jj, ii = np.mgrid[1:m + 1, 1:l + 1]
    mask = ((ii - xc) ** 2 + (jj - yc) ** 2 <= r * r).astype(np.int32)
    del ii, jj
    box = (xc - r - 3, xc + r + 3, yc - r - 3, yc + r + 3)
    out = {"workload": f"OPEN channel {l}x{m}, cylinder D={D} at ({xc},{yc}), u={u}, Re=200, "
                       f"{steps} steps", "solid_nodes": int(mask.sum())}
    for prec in ("single", "fp16"):
        narrow = prec == "fp16"
        s = Solver(l, m, bc="open", prec=prec, svisc=nu, u0=u, scale=1024.0 if narrow else 1.0,
                   form=1 if narrow else 0, rounding=1 if narrow else 0)
        s.set_obstacle(mask)
        s.run(50)
        s.sync()
        t0 = time.perf_counter()
        s.run(steps)
        s.sync()
        wall = time.perf_counter() - t0
        fx, fy = s.draglift(box)
        out[prec] = {"mlups": round(l * m * steps / wall / 1e6, 1),
                     "kernel_ms_per_step": round(s.last_run_ms(), 5),
                     "cd_startup": round(fx / (u * u * r), 4), "cl_startup": round(fy / (u * u * r), 4),
                     "healthy": bool(s.check(0.5))}
        s.close()
    return out


def c4_shedding(steps=400000, every=256):
    """The C4 channel (geometry of c4_leg) run long enough to shed: from rest,
    the inflow front reaches the cylinder after about 20,000 steps, the wake
    turns unsteady, and the last half of the run is the measurement window.
    Sampled every `every` steps: the drag and lift coefficients
    Fx,y / (u^2 R) and the cross-stream velocity 4 D behind the cylinder.
    Strouhal number St = D / (u T) from the mean period T between upward
    zero crossings of that velocity (linearly interpolated); mean drag, lift
    amplitude.  FP32 and FP16 storage (serving settings)."""
    import numpy as np
    from bgk2d_gpu_amd import Solver
    l, m, D, u = 8192, 2048, 128, 0.05
    nu = u * D / 200.0
    xc, yc, r = 1024, m // 2 + 1, D // 2
    jj, ii = np.mgrid[1:m + 1, 1:l + 1]
    mask = ((ii - xc) ** 2 + (jj - yc) ** 2 <= r * r).astype(np.int32)
    del ii, jj
    box = (xc - r - 3, xc + r + 3, yc - r - 3, yc + r + 3)
    out = {"workload": f"OPEN channel {l}x{m}, cylinder D={D} at ({xc},{yc}), u={u}, Re=200, from rest, "
                       f"{steps} steps, sampled every {every}; statistics over the last half",
           "convective_times": round(steps * u / D, 1)}
    for prec in ("single", "fp16"):
        narrow = prec == "fp16"
        s = Solver(l, m, bc="open", prec=prec, svisc=nu, u0=u, scale=1024.0 if narrow else 1.0,
                   form=1 if narrow else 0, rounding=1 if narrow else 0)
        s.set_obstacle(mask)
        cd, cl, v = [], [], []
        t0 = time.perf_counter()
        for k in range(steps // every):
            s.run(every)
            fx, fy = s.draglift(box)
            cd.append(fx / (u * u * r))
            cl.append(fy / (u * u * r))
            v.append(s.probe(xc + 4 * D, yc)[1])
            if (k + 1) % 200 == 0:
                print(f"# c4 {prec}: step {(k + 1) * every}, Cd {cd[-1]:.4f}, Cl {cl[-1]:.4f}",
                      file=sys.stderr, flush=True)
        wall = time.perf_counter() - t0
        ok = bool(s.check(0.5))
        s.close()
        h = len(v) // 2
        y = np.array(v[h:]) - np.mean(v[h:])
        up = [i + y[i] / (y[i] - y[i + 1]) for i in range(len(y) - 1) if y[i] < 0 <= y[i + 1]]
        period = (up[-1] - up[0]) / (len(up) - 1) * every if len(up) >= 3 else None
        out[prec] = {"mlups_with_sampling": round(l * m * steps / wall / 1e6, 1), "healthy": ok,
                     "strouhal": round(D / (u * period), 4) if period else None,
                     "periods_in_window": len(up) - 1,
                     "cd_mean": round(float(np.mean(cd[h:])), 4),
                     "cl_amplitude": round(float((np.max(cl[h:]) - np.min(cl[h:])) / 2), 4),
                     "v_amplitude_over_u": round(float(np.abs(y).max() / u), 4)}
    a, b = out["single"], out["fp16"]
    if a["strouhal"] and b["strouhal"]:
        out["fp16_vs_fp32"] = {"strouhal_rel": round(abs(b["strouhal"] / a["strouhal"] - 1), 5),
                               "cd_mean_rel": round(abs(b["cd_mean"] / a["cd_mean"] - 1), 5)}
    return out


def copy_calib():
    """Plain device copies of a known size, for scaling the FETCH_SIZE /
    WRITE_SIZE counters to bytes."""
    import torch
    a = torch.empty(1 << 28, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    a.fill_(1.0)
    torch.cuda.synchronize()
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calib_bytes_each_way": a.numel() * 4, "copies": 5}))


def copy_rate():
    """Measured device-copy rate on this box (GB/s, read + write bytes): a
    plain 1 GiB copy, best of 10, timed with events on torch's stream (the
    stream the copy runs on).  Run in a process of its own (--copy-rate):
    torch's HIP runtime and the library's do not share a process."""
    import torch
    a = torch.empty(1 << 28, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    a.fill_(1.0)
    b.copy_(a)
    torch.cuda.synchronize()
    best = None
    for _ in range(10):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        b.copy_(a)
        e1.record()
        e1.synchronize()
        ms = e0.elapsed_time(e1)
        best = ms if best is None or ms < best else best
    nbytes = 2 * a.numel() * 4
    del a, b
    torch.cuda.empty_cache()
    return round(nbytes / (best / 1e3) / 1e9, 1)


def copy_rate_child():
    """copy_rate() in a child process, before this one touches the GPU"""
    try:
        r = subprocess.run([sys.executable, os.path.abspath(__file__), "--copy-rate"], stdout=subprocess.PIPE,
                           stderr=subprocess.DEVNULL, text=True, timeout=300)
        return json.loads(r.stdout.strip().splitlines()[-1])["copy_gbs"]
    except (OSError, subprocess.SubprocessError, ValueError, KeyError, IndexError):
        return None


def profiled_traffic(name):
    """HBM bytes per launch of a kernel measured with FETCH_SIZE and
    WRITE_SIZE (calibrated on a 1 GiB copy), from its committed summary"""
    path = os.path.join(ROOT, "profiles", name)
    if os.path.exists(path):
        m = re.search(r"traffic_bytes_per_launch:\s*([0-9]+)", open(path).read())
        if m:
            return int(m.group(1))
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-only", action="store_true", help="only the CPU baseline leg")
    ap.add_argument("--fp16-steps", type=int, default=2000)
    ap.add_argument("--no-fp16", action="store_true")
    ap.add_argument("--fp16-only", action="store_true",
                    help="only the FP16/FP32 Taylor-Green 8192^2 leg (one and two steps per pass)")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-only", action="store_true")
    ap.add_argument("--serve-n", type=int, default=48)
    ap.add_argument("--serve-full", action="store_true", help="the serving leg only (same as --serve-only)")
    ap.add_argument("--no-c4", action="store_true")
    ap.add_argument("--c4-only", action="store_true")
    ap.add_argument("--c4-steps", type=int, default=2000)
    ap.add_argument("--c4-shedding", action="store_true",
                    help="only the C4 channel run long enough to shed (Strouhal number, mean drag)")
    ap.add_argument("--c4-shedding-steps", type=int, default=400000)
    ap.add_argument("--precision-study", action="store_true",
                    help="only run the FP32/FP16/BF16 vs FP64 Taylor-Green error study")
    ap.add_argument("--study-n", type=int, default=8192)
    ap.add_argument("--study-steps", type=int, default=5000)
    ap.add_argument("--study-init", type=int, default=0)
    ap.add_argument("--copy-calib", action="store_true",
                    help="only run 5 torch copies of 1 GiB (HBM counter calibration)")
    ap.add_argument("--copy-rate", action="store_true", help="only the measured device-copy rate")
    ap.add_argument("--no-copy-rate", action="store_true",
                    help="skip the device-copy child process (profiled runs: it is a second GPU process)")
    args = ap.parse_args()
    if args.copy_rate:
        print(json.dumps({"copy_gbs": copy_rate()}))
        return
    if args.copy_calib:
        copy_calib()
        return
    if args.precision_study:
        print(json.dumps(precision_study(args.study_n, args.study_steps, init_v=args.study_init)))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(a large solve would shard by y-slabs, serving by request, one GPU each)", file=sys.stderr)
        sys.exit(2)
    if args.serve_only:
        print(json.dumps(serving_leg(args.serve_n)))
        return
    if args.serve_full:
        print(json.dumps(serving_leg(args.serve_n)))
        return
    if args.c4_only:
        print(json.dumps(c4_leg(args.c4_steps)))
        return
    if args.c4_shedding:
        print(json.dumps(c4_shedding(args.c4_shedding_steps)))
        return
    if args.fp16_only:
        print(json.dumps(fp16_leg(args.fp16_steps)))
        return
    if args.cpu_only:
        print(json.dumps(cpu_baseline()))
        return

    copy_gbs = None if args.no_copy_rate else copy_rate_child()
    n = 4096
    # the FP32 solve as bgk_run does it: two FUSED steps per pass over HBM
    r32 = gpu_solve(n, "single", "driven", args.steps, args.warmup, 0.04096, 0.1, tblock=2)
    mlups = n * n * args.steps / r32["wall"] / 1e6
    # one launch of k_fused_tb = two steps; its algorithmic HBM bytes are
    # 18*S per node (each population read once and written once per pass)
    launch_s = 2 * r32["kernel_ms"] / 1e3
    algo = BYTES_PER_LUP["single"] * n * n
    achieved = algo / launch_s / 1e9
    traffic = profiled_traffic("k_fused_tb_fp32_ldc4096_hbm_counters.md")
    # the same solve with one step per pass (k_fused), for comparison
    one_steps = max(2, args.steps // 4)
    r1 = gpu_solve(n, "single", "driven", one_steps, args.warmup, 0.04096, 0.1, tblock=1)
    ml1 = n * n * one_steps / r1["wall"] / 1e6
    one = {"kernel": "k_fused<float,float,DRIVEN> (one launch = one step)", "steps": one_steps,
           "mlups": round(ml1, 1),
           "achieved_gbs": round(algo / (r1["kernel_ms"] / 1e3) / 1e9, 1),
           "frac": round(algo / (r1["kernel_ms"] / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
           "traffic_bytes_per_launch": profiled_traffic("k_fused_fp32_ldc4096_hbm_counters.md")}

    if not args.no_fp16:
        fp16 = fp16_leg(args.fp16_steps)

    out = {
        "metric": METRIC,
        "value": round(mlups, 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r32["wall"] / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "lid-driven cavity 4096x4096, FUSED, FP32 storage+compute, "
                               "Re=1e4 (svisc=0.04096, u0=0.1), boundary folded into the pull",
                   "global_batch": 1, "lattice": [n, n], "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": (round(traffic / launch_s / 1e9, 1) if traffic else None),
                     "algorithmic_bytes_per_launch": algo,
                     "lattice_updates_per_launch": 2 * n * n,
                     "traffic_bytes_per_launch": traffic,
                     "kernel": "k_fused_tb<float,float,DRIVEN> (one launch = two steps)",
                     "kernel_ms_per_launch": round(launch_s * 1e3, 5),
                     "traffic_source": "profiles/k_fused_tb_fp32_ldc4096_hbm_counters.md"},
        "mlups_over_one_step_roofline": round(mlups * 1e6 * BYTES_PER_LUP["single"] / 1e9
                                              / HBM_PEAK_GBS, 4),
        "measured_copy_gbs": copy_gbs,
        "one_step_per_pass": one,
        "kernel_ms_per_step": r32["kernel_ms"],
        "healthy": bool(r32["ok"] and r1["ok"]),
    }
    if not args.no_fp16:
        out["fp16_store"] = fp16
    if not args.no_c4:
        out["cylinder_channel"] = c4_leg(args.c4_steps)
    if not args.no_serve:
        out["serving"] = serving_leg(args.serve_n)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
