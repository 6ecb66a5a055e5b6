#!/usr/bin/env python3
"""Throughput of the full FP64 time step on one MI355X.

Workload: the 256x256x512 ACDI rising bubble (density ratio 1000, viscosity
ratio 100, Re=35, Eo=10, periodic x/y, no-slip z; one sphere r=0.25 at
(0.5,0.5,0.5) in a 1x1x2 box), synthetic initial state from the reference's
own initial-condition rules. One "step" is one reference AB2 time step
(src/main.f90:414-562): phase-field transport, phi/normals/curvature,
momentum, projection with the FFT Poisson solve, gamma, and the dt/divergence
checks every 10 steps. Fields stay resident in HBM; the timed region holds
only device work and the checks' 8-byte read-backs.

Prints one JSON line: metric Mcell-steps/s (cells x steps / seconds / 1e6)
over the timed region, where the steps replay their captured HIP graphs with
no per-launch events; the dominant kernel's roofline from a second,
instrumented pass over as many steps (direct launches, HIP events around
every launch on the case's stream); and the CPU oracle timed on this host.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for sub in ("cans-fizzy_amd", "oracle", ""):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E

# algorithmic bytes per cell of one launch: distinct fields read + written once, 8 B each
# (the minimal dataflow of each routine; halo'd-array kernels count halo cells)
ALG_BYTES = {
    # fused momentum: reads u,v,w,psi,psio,flux x3,history x3,p,po,kappa (14 fields), writes u,v,w,history x3
    "mom_fused": 160,
    # fused phase-field step: reads u,v,w,normals x3,phi,psio,history (9), writes psi,history,flux x3 (5)
    "acdi_step": 112,
    # fillps evaluated in the load of the forward x transform: reads u,v,w, writes the spectral array
    "fillps_fft_x_fwd": 32,
    # correc + updatep + gamma speeds over the halo-inclusive grid: reads u,v,w,dp,p, writes u,v,w,p
    "correc_step": 72,
    "fft_x_fwd": 16, "fft_y_fwd": 16, "fft_y_bwd": 16, "fft_x_bwd": 16, "thomas": 16, "thomas_periodic": 16,
    "mom_ad": 80, "acdi_transport": 96, "youngs": 40, "mom_oth": 56, "rk_mom_ab2": 112, "rk_mom_oth": 80,
    "rk_psi_oop": 32, "acdi_phi": 16, "clip": 16, "fillps": 32, "correc": 56, "updatep_oop": 24,
    "red_gamma": 24, "red_div": 24, "red_dt": 24, "vof_flux": 88, "vof_div": 32, "vof_phi": 16, "forcing": 32,
}
HALO_KERNELS = {"acdi_phi", "clip", "updatep_oop", "vof_phi", "correc", "correc_step"}
# minimal per-step dataflow of the whole step (stage model S1-S9): 61 passes of 8 B
STEP_ALG_BYTES_PER_CELL = 488


WORKLOAD_NAMES = {"c2": "acdi_rising_bubble_256x256x512", "c0": "acdi_rising_bubble_64x64x128",
                  "c3": "vof_droplets_forced_256x256x256"}


KERNEL_SYMBOLS = {"mom_fused": ("k_mom_lds", "k_mom_fused"),
                  "correc_step": ("k_correc_lds", "k_correc_step_mz", "k_correc_step"),
                  "acdi_step": ("k_acdi_sf", "k_acdi_step", "k_acdi_lds"),
                  "youngs": ("k_youngs_mz", "k_youngs")}


def profiled_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` from the committed counter summary
    (FETCH_SIZE + WRITE_SIZE, copy-kernel calibrated), or None."""
    path = os.path.join(ROOT, "profiles", f"step_kernels_{WORKLOAD_NAMES[workload]}.txt")
    if not os.path.exists(path):
        return None
    # the kernels a launch name stands for (variants share one name)
    symbols = KERNEL_SYMBOLS.get(kernel, ("k_" + kernel,))
    for line in open(path):
        cols = line.split()
        # kernel names carry their template arguments (the FP64 instances here)
        if len(cols) >= 6 and cols[0].split("<")[0] in symbols and "float" not in cols[0] and cols[1].isdigit():
            try:
                return round((float(cols[4]) + float(cols[5])) * 1e6)
            except ValueError:
                continue  # not a table row
    return None


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def prof_report(F, ctx):
    buf = C.create_string_buffer(1 << 16)
    F.call("fz_prof_report", ctx, buf, len(buf))
    rows = {}
    for line in buf.value.decode().splitlines():
        name, calls, ms = line.split()
        rows[name] = (int(calls), float(ms))
    return rows


def dc_count(dc):
    return dc.field_ptr("u")[1]


def copy_bandwidth(F, ctx, nbytes, reps=10):
    """HBM rate of the library's plain copy kernel (k_copy) over nbytes: the
    achievable streaming reference, and the calibration kernel of the
    FETCH_SIZE/WRITE_SIZE counter runs."""
    src, dst = C.c_void_p(), C.c_void_p()
    F.call("fz_malloc", ctx, nbytes, C.byref(src))
    F.call("fz_malloc", ctx, nbytes, C.byref(dst))
    try:
        F.call("fz_copy", ctx, dst, src, nbytes)
        F.call("fz_ctx_sync", ctx)
        t0 = time.perf_counter()
        for _ in range(reps):
            F.call("fz_copy", ctx, dst, src, nbytes)
        F.call("fz_ctx_sync", ctx)
        t1 = time.perf_counter()
    finally:
        F.call("fz_free", ctx, src)
        F.call("fz_free", ctx, dst)
    return 2 * nbytes * reps / (t1 - t0) / 1e9


def cpu_baseline(spec, budget_s=15.0, max_steps=20):
    """The oracle (plain C, OpenMP over k-planes) on this host's cores: AB2
    steps of the same case after the first (is_first) step, until about
    budget_s seconds of stepping or max_steps steps."""
    import oracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    os.environ.setdefault("OMP_NUM_THREADS", str(threads))
    t0 = time.perf_counter()
    oc = O.Case(spec)
    oc.init_fields()
    oc.step(1)  # the is_first step (history copies), not timed
    t1 = time.perf_counter()
    steps = 0
    while steps < max_steps and time.perf_counter() - t1 < budget_s:
        oc.step(1)
        steps += 1
    dt = time.perf_counter() - t1
    n = spec["ng"]
    cells = n[0] * n[1] * n[2]
    log(f"cpu oracle: set-up + first step {t1 - t0:.1f} s, {steps} timed steps {dt:.2f} s on {threads} threads")
    return {"value": round(cells * steps / dt / 1e6, 3), "unit": "Mcell-steps/s", "cores": threads, "kind": "port",
            "sample": f"AB2 steps 2-{steps + 1} of the same {n[0]}x{n[1]}x{n[2]} case by the C oracle "
                      f"(OpenMP, {threads} threads), {dt:.1f} s"}


def cpu_baseline_poisson(n, f, budget_s=12.0, max_solves=5):
    """The oracle's FP64 solve (plain C, OpenMP) of the same 512^3 right-hand
    side on this host's cores: whole solves until about budget_s seconds or
    max_solves solves (the plan's set-up not timed)."""
    import numpy as np

    import oracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    dzc, dzf, zc, zf = O.initgrid(1, n[2], 0.0, 1.0)
    dli = [n[0] / 1.0, n[1] / 1.0]
    sol = O.Solver(n, dli + [float(n[2])], 1.0 / dzc, 1.0 / dzf, list("PPPPNN"))
    p = np.empty_like(f)
    solves, t = 0, 0.0
    while solves < max_solves and t < budget_s:
        p[...] = f
        t0 = time.perf_counter()
        sol.solve(p)
        t += time.perf_counter() - t0
        solves += 1
    cells = n[0] * n[1] * n[2]
    log(f"cpu oracle: {solves} solves of {n[0]}x{n[1]}x{n[2]} in {t:.2f} s on {threads} threads")
    return {"value": round(cells * solves / t / 1e6, 2), "unit": "Mcell-solves/s", "cores": threads, "kind": "port",
            "sample": f"{solves} FP64 solves of the same {n[0]}x{n[1]}x{n[2]} right-hand side by the C oracle "
                      f"(OpenMP, {threads} threads), {t:.1f} s"}


def cpu_baseline_serving(reqs, budget_s=15.0, max_steps=20):
    """The oracle on this host's cores over the serving mix's requests in
    arrival order: each request's case initialised and stepped (up to
    max_steps of its steps, after its first), until about budget_s seconds of
    stepping; Mcell-steps/s over the timed steps."""
    import numpy as np

    import oracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    t, cell_steps, used = 0.0, 0, 0
    for _, spec, nsteps in reqs:
        if t >= budget_s:
            break
        oc = O.Case(spec)
        oc.init_fields()
        oc.step(1)
        k = min(max_steps, max(nsteps - 1, 0))
        t0 = time.perf_counter()
        oc.step(k)
        t += time.perf_counter() - t0
        cell_steps += int(np.prod(spec["ng"])) * k
        used += 1
        del oc
    log(f"cpu oracle: {used} requests, {cell_steps / 1e6:.1f} Mcell-steps in {t:.2f} s on {threads} threads")
    return {"value": round(cell_steps / t / 1e6, 3), "unit": "Mcell-steps/s", "cores": threads, "kind": "port",
            "sample": f"the first {used} requests of the mix in arrival order, steps 2-{max_steps + 1} of each "
                      f"(FP64), by the C oracle (OpenMP, {threads} threads), {t:.1f} s"}


def main_serving(args):
    """The serving mix: Poisson arrivals (mean 50 ms) of 2-D/3-D rising bubbles of
    uneven grids and density ratios, 200 steps each, admitted by HBM footprint
    and stepped together; p50/p99 per-step latency, aggregate throughput, and
    the same requests run one after another for comparison."""
    import ctypes as C

    import numpy as np

    import fizzy_amd as F
    from fizzy_amd import serve

    reqs = serve.serving_requests(args.requests, mean_interarrival=args.arrival_ms / 1e3, nsteps=args.case_steps,
                                  precision=args.precision, refine=args.refine,
                                  grids=[g for g in serve.GRIDS if g[2]] if args.grids == "2d" else None)
    os.environ["FZ_SERVE_GROUP"] = "1" if args.group else "0"  # read when the server is made
    cells = [int(np.prod(spec["ng"])) * n for _, spec, n in reqs]
    f, t = C.c_size_t(), C.c_size_t()
    F.call("fz_device_memory", C.byref(f), C.byref(t))
    srv = serve.Server(int(0.5 * f.value))
    try:
        wall, ids = serve.run_stream(srv, reqs)
        lat = srv.step_times(0)
    finally:
        srv.close()
    log(f"serving: {len(reqs)} requests, {sum(cells) / 1e6:.1f} Mcell-steps in {wall:.2f} s, "
        f"p50 {np.percentile(lat, 50):.3f} ms, p99 {np.percentile(lat, 99):.3f} ms")
    t0 = time.perf_counter()
    for _, spec, n in reqs:  # sequential: one case at a time, same work (graphs as in the server)
        dc = F.DeviceCase(spec)
        try:
            dc.init()
            F.call("fz_case_set_graphs", dc._h, 1)
            dc.step(n)
            dc.sync()
        finally:
            dc.close()
    seq = time.perf_counter() - t0
    thr, seq_thr = sum(cells) / wall / 1e6, sum(cells) / seq / 1e6
    log(f"sequential: {seq:.2f} s ({seq_thr:.1f} Mcell-steps/s); served {thr:.1f} Mcell-steps/s")
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline_serving(reqs)
        cpu["cpu_model"] = cpu_model()
    out = {
        "metric": "ensemble p50/p99 step latency",
        "value": round(float(np.percentile(lat, 99)), 4), "unit": "ms (p99 per-case step)",
        "p50_ms": round(float(np.percentile(lat, 50)), 4), "p99_ms": round(float(np.percentile(lat, 99)), 4),
        "throughput_Mcell_steps_s": round(thr, 2), "sequential_Mcell_steps_s": round(seq_thr, 2),
        "speedup_vs_sequential": round(thr / seq_thr, 3), "wall_s": round(wall, 3),
        "last_arrival_s": round(max(r[0] for r in reqs), 3), "n_gpus": 1,
        "higher_is_better": False, "vs_baseline": None, "dtype": "f64" if args.precision == "fp64" else "mixed",
        "data": "synthetic",
        "config": {"workload": "serving_mix", "requests": len(reqs), "steps_per_case": args.case_steps,
                   "batched_steps": bool(args.group), "grid_set": args.grids,
                   "refinements": args.refine if args.precision == "mixed" else None,
                   "arrivals": f"Poisson, mean {args.arrival_ms:g} ms", "grids": "64^2, 128^2, 256^2 (2-D)" + ("" if args.grids == "2d" else ", 64^2x128, 128^2x256"),
                   "density_ratio": "uniform [2, 1000]"},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def batched_poisson(F, np, k=16, n=64, rounds=20):
    """k same-shaped n^3 solves: one fz_solver per context against one
    fz_solver_batched call for all of them."""
    import ctypes as C
    ctxs = [F.Context([n, n, n], [1.0, 1.0, 1.0], "PPPPNN") for _ in range(k)]
    try:
        rng = np.random.default_rng(5)
        fields = []
        for c in ctxs:
            F.call("fz_initsolver", c.h, None, None, None, None, None, None, None, None, None)
            fields.append(c.halo(rng.uniform(-1.0, 1.0, (n + 2, n + 2, n + 2))))
        cv = (C.c_void_p * k)(*[c.h.value for c in ctxs])
        pv = (C.c_void_p * k)(*[f.ptr.value for f in fields])
        t = {}
        for mode in ("single", "batched"):
            for rep in range(rounds + 2):
                if rep == 2:
                    for c in ctxs:
                        c.sync()
                    t0 = time.perf_counter()
                if mode == "single":
                    for c, f in zip(ctxs, fields):
                        F.call("fz_solver", c.h, f.ptr)
                else:
                    F.call("fz_solver_batched", cv, pv, k)
            for c in ctxs:
                c.sync()
            t[mode] = (time.perf_counter() - t0) / rounds * 1e3
        log(f"{k} solves of {n}^3: one per context {t['single']:.3f} ms, batched {t['batched']:.3f} ms")
        return {"solves": k, "grid": [n, n, n], "ms_one_per_context": round(t["single"], 4),
                "ms_batched": round(t["batched"], 4), "speedup": round(t["single"] / t["batched"], 3)}
    finally:
        for c in ctxs:
            c.close()


def cpu_baseline_one_core(workload):
    """The same oracle timing on one thread (a separate process, since the
    OpenMP thread count is fixed at start-up), over one AB2 step."""
    import subprocess
    env = dict(os.environ, OMP_NUM_THREADS="1")
    try:
        r = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-baseline-only", "--workload", workload],
                           env=env, capture_output=True, text=True, timeout=300)
        return json.loads(r.stdout.strip().splitlines()[-1])
    except Exception as e:  # reported, not fatal: the GPU figures stand without it
        log(f"one-core baseline not measured: {e}")
        return None


def main_poisson(args):
    """The standalone Poisson solve: 512^3, periodic x/y, Neumann z, a
    right-hand side seeded uniform in [-1, 1] with its (volume-weighted)
    mean removed; the FP64 solver against FP32 transforms with FP64-residual
    refinement to a 1e-12 relative residual, each timed over the solves."""
    import ctypes as C

    import numpy as np

    import fizzy_amd as F

    n = [512, 512, 512]
    nsolve = args.steps
    cells = n[0] * n[1] * n[2]
    ctx = F.Context(n, [1.0, 1.0, 1.0], "PPPPNN")
    try:
        F.call("fz_initsolver", ctx.h, None, None, None, None, None, None, None, None, None)
        rng = np.random.default_rng(20261202)
        f = np.zeros((n[0] + 2, n[1] + 2, n[2] + 2), order="F")
        fi = f[1:-1, 1:-1, 1:-1]
        fi[...] = rng.uniform(-1.0, 1.0, n)
        wz = ctx.dzf[1:-1]
        fi -= (fi * wz).sum() / (wz.sum() * n[0] * n[1])
        b = ctx.halo(f)
        x = ctx.halo()
        res = {}
        for mode in ("fp64", "mixed"):
            its, rels = [], []
            for rep in range(args.warmup + nsolve):
                if rep == args.warmup:
                    ctx.sync()
                    t0 = time.perf_counter()
                F.call("fz_copy", ctx.h, x.ptr, b.ptr, C.c_size_t(x.nbytes))
                if mode == "fp64":
                    F.call("fz_solver", ctx.h, x.ptr)
                else:
                    it, rr = C.c_int(), C.c_double()
                    F.call("fz_solver_mixed", ctx.h, x.ptr, 1e-12, 10, C.byref(it), C.byref(rr))
                    its.append(it.value)
                    rels.append(rr.value)
            ctx.sync()
            t1 = time.perf_counter()
            # the residual of the last solution, by the reference's chk_poisson
            zero = (C.c_double * 6)()
            F.call("fz_boundp", ctx.h, b"PPPPNN", zero, x.ptr)
            dm = C.c_double()
            F.call("fz_chk_poisson", ctx.h, b.ptr, x.ptr, C.byref(dm))
            ms = 1e3 * (t1 - t0) / nsolve
            res[mode] = {"ms_per_solve": round(ms, 3), "Mcell_solves_s": round(cells / ms / 1e3, 1),
                         "relres": dm.value / np.abs(fi).max()}
            if its:
                res[mode]["refinements"] = sorted(set(its))
            log(f"{mode}: {ms:.3f} ms/solve (incl. the rhs copy), relres {res[mode]['relres']:.2e}"
                + (f", refinements {sorted(set(its))}" if its else ""))
    finally:
        ctx.close()
    res["batched"] = batched_poisson(F, np)
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline_poisson(n, f)
        cpu["cpu_model"] = cpu_model()
    out = {
        "metric": "standalone Poisson solve, FP64 vs FP32 + FP64-residual refinement",
        "value": res["fp64"]["Mcell_solves_s"], "unit": "Mcell-solves/s (FP64)", "n_gpus": 1, "steps": nsolve,
        "warmup": args.warmup, "higher_is_better": True, "vs_baseline": None, "dtype": "f64",
        "data": "synthetic: uniform [-1,1] right-hand side, seeded, volume-weighted mean removed",
        "config": {"workload": "poisson_512x512x512_PPPPNN", "grid": n}, "fp64": res["fp64"],
        "mixed": res["mixed"], "batched": res["batched"], "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def time_steps(F, dc, steps, graphs, prof):
    """Wall time of `steps` steps of a case (stream synchronised on both sides):
    replayed from captured graphs or launched directly, with per-launch HIP
    events (prof) or without; returns (seconds, killed, per-launch rows)."""
    F.call("fz_case_set_graphs", dc._h, 1 if graphs else 0)
    F.call("fz_prof_enable", dc.ctx, 1 if prof else 0)
    prof_report(F, dc.ctx)  # clears earlier records
    dc.sync()
    t0 = time.perf_counter()
    killed = dc.step(steps)
    dc.sync()
    t1 = time.perf_counter()
    F.call("fz_prof_enable", dc.ctx, 0)
    rows = prof_report(F, dc.ctx) if prof else {}
    return t1 - t0, killed, rows


def log_rows(rows, width=22):
    total_ms = sum(m for _, m in rows.values())
    for name, (calls, m) in sorted(rows.items(), key=lambda kv: -kv[1][1])[:12]:
        log(f"  {name:{width}s} {calls:5d} launches {m / calls * 1e3:9.1f} us/launch {100 * m / total_ms:5.1f}%")


def run_case_timed(F, spec, args, label):
    """Initialise, autotune and warm up a case (its step graphs are captured
    in the warm-up), then: the timed region, `steps` steps replayed from the
    graphs with no per-launch events (the headline); and an instrumented
    pass over as many further steps, launched directly with HIP events
    around every launch (the per-kernel table and the kernel roofline)."""
    dc = F.DeviceCase(spec)
    try:
        dc.init()
        tuning = None
        if not args.no_autotune:
            tuning = dc.autotune(3)
            log(f"autotune ({label}): {tuning}")
        graphs = not args.no_graphs
        F.call("fz_case_set_graphs", dc._h, 1 if graphs else 0)
        if args.warmup:
            dc.step(args.warmup)
        elapsed, killed, _ = time_steps(F, dc, args.steps, graphs, prof=False)
        sc = dc.scalars()
        inst, ikilled, rows = (None, 0, {})
        if not args.no_prof:
            inst, ikilled, rows = time_steps(F, dc, args.steps, graphs=False, prof=True)
        copy_gbs = None
        if label == "fp64":
            copy_gbs = copy_bandwidth(F, dc.ctx, 8 * dc_count(dc))
        return dict(elapsed=elapsed, killed=killed or ikilled, divmax=float(sc[7]), rows=rows, inst=inst,
                    tuning=tuning, copy_gbs=copy_gbs)
    finally:
        dc.close()


def run_mixed(F, spec, args, cells, fp64_s_per_step, refine=None):
    """The same workload in the mixed-precision mode, timed the same way as
    the FP64 run. refine -1: the FP32 phase field with the plain step's FP64
    projection."""
    refine = args.refine if refine is None else refine
    what = "FP64 projection" if refine < 0 else f"{refine} refinements"
    r = run_case_timed(F, dict(spec, precision="mixed", refine=refine), args, "mixed")
    s_per_step = r["elapsed"] / args.steps
    value = cells / s_per_step / 1e6
    log(f"mixed ({what}): {1e3 * s_per_step:.3f} ms/step, {value:.1f} Mcell-steps/s, divmax {r['divmax']:.2e}, "
        f"killed {r['killed']}")
    if r["rows"]:
        log(f"  instrumented pass: {1e3 * r['inst'] / args.steps:.3f} ms/step")
        log_rows(r["rows"])
    return {"value": round(value, 2), "ms_per_step": round(1e3 * s_per_step, 4),
            "speedup_vs_fp64": round(fp64_s_per_step / s_per_step, 3), "refinements": refine,
            "divmax": r["divmax"],
            "dtype": ("f32 phase field, normals, curvature, fluxes and AB2 histories; f64 velocity, pressure, "
                      "momentum arithmetic and the whole projection" if refine < 0 else
                      "f32 phase field, normals, curvature, fluxes, AB2 histories and transforms; "
                      "f64 velocity, pressure, momentum arithmetic, pivots, residuals and sums")}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", choices=["c2", "c0", "c1", "c3", "c4"], default="c2",
                    help="c2: 256x256x512 bubble (default); c0: 64x64x128; c1: the standalone 512^3 Poisson "
                         "solve; c3: THINC/QQ droplets in forced turbulence, 256^3; c4: the serving mix")
    ap.add_argument("--requests", type=int, default=24, help="c4: number of requests")
    ap.add_argument("--case-steps", type=int, default=200, help="c4: steps per request")
    ap.add_argument("--arrival-ms", type=float, default=50.0,
                    help="c4: mean gap between arrivals (0: every request waiting at the start)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-prof", action="store_true", help="skip the instrumented pass (per-launch events)")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision run of the same workload")
    ap.add_argument("--refine", type=int, default=1, help="mixed precision: FP64-residual refinements per solve")
    ap.add_argument("--no-graphs", action="store_true", help="time the steps launched directly, not from graphs")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch parameters")
    ap.add_argument("--grids", choices=["mix", "2d"], default="mix",
                    help="c4: the five grids of the serving mix, or its three 2-D grids only")
    ap.add_argument("--group", action="store_true",
                    help="c4: batched steps of same-shaped small cases (fz_case_batch) in the server")
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64", help="c4: precision of the cases")
    ap.add_argument("--cpu-baseline-only", action="store_true", help=argparse.SUPPRESS)
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one MI355X: this build runs the hot path on a single GPU "
              f"(asked for --gpus {args.gpus}, WORLD_SIZE {world}); nothing measured.", file=sys.stderr)
        sys.exit(2)

    if args.cpu_baseline_only:  # the one-core oracle timing, run as a child process
        from fizzy_amd import cases
        spec = {"c2": cases.config2, "c0": cases.config0, "c3": cases.config3}[args.workload]()
        print(json.dumps(cpu_baseline(spec, budget_s=10.0, max_steps=1)), flush=True)
        return
    if args.workload == "c4":
        return main_serving(args)
    if args.workload == "c1":
        return main_poisson(args)

    import fizzy_amd as F
    from fizzy_amd import cases

    spec = {"c2": cases.config2, "c0": cases.config0, "c3": cases.config3}[args.workload]()
    n = spec["ng"]
    cells = n[0] * n[1] * n[2]
    log(f"workload {args.workload}: {n[0]}x{n[1]}x{n[2]} = {cells} cells, warmup {args.warmup}, steps {args.steps}")
    r = run_case_timed(F, spec, args, "fp64")
    elapsed, rows = r["elapsed"], r["rows"]
    ms_per_step = 1e3 * elapsed / args.steps
    value = cells * args.steps / elapsed / 1e6
    log(f"{ms_per_step:.3f} ms/step ({'graphs' if not args.no_graphs else 'direct launches'}, no per-launch events), "
        f"{value:.1f} Mcell-steps/s, divmax {r['divmax']:.2e}, killed {r['killed']}")

    roof = None
    if rows:
        inst_ms = 1e3 * r["inst"] / args.steps
        log(f"instrumented pass (direct launches, HIP events around each): {inst_ms:.3f} ms/step")
        log_rows(rows, 18)
        total_ms = sum(ms for _, ms in rows.values())
        ranked = sorted(rows.items(), key=lambda kv: -kv[1][1])
        dom = next(name for name, _ in ranked if name in ALG_BYTES)
        calls, ms = rows[dom]
        units = (n[0] + 2) * (n[1] + 2) * (n[2] + 2) if dom in HALO_KERNELS else cells
        bytes_per_launch = ALG_BYTES[dom] * units
        avg_s = ms / calls / 1e3
        achieved = bytes_per_launch / avg_s / 1e9
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": profiled_traffic(dom, args.workload),
                "kernel": dom, "symbols": list(KERNEL_SYMBOLS.get(dom, ("k_" + dom,))),
                "avg_launch_us": round(avg_s * 1e6, 2), "alg_bytes_per_launch": bytes_per_launch,
                "kernel_time_share": round(ms / total_ms, 4), "instrumented_ms_per_step": round(inst_ms, 4)}
    step_bw = STEP_ALG_BYTES_PER_CELL * cells / (elapsed / args.steps) / 1e9
    log(f"plain copy kernel over one field ({8 * (n[0] + 2) * (n[1] + 2) * (n[2] + 2) / 1e6:.0f} MB): "
        f"{r['copy_gbs']:.0f} GB/s")

    mixed = None
    if not args.no_mixed:
        mixed = run_mixed(F, spec, args, cells, elapsed / args.steps)
        # the same FP32 phase field with the plain step's FP64 projection
        mixed["fp64_projection"] = run_mixed(F, spec, args, cells, elapsed / args.steps, refine=-1)

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(spec)
        cpu["one_core"] = cpu_baseline_one_core(args.workload)
        cpu["cpu_model"] = cpu_model()
    out = {
        "metric": "Mcell-steps/s per RK3 step, FP64 vs mixed; ensemble p50/p99 step latency",
        "value": round(value, 2),
        "unit": "Mcell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": ("synthetic: seeded droplets (SplitMix64 centres), fluid at rest, Taylor-Green forcing"
                 if spec.get("vof") else
                 "synthetic: the reference's rising-bubble initial condition (sphere, fluid at rest)"),
        "config": {"workload": WORKLOAD_NAMES[args.workload], "interface": "THINC/QQ" if spec.get("vof") else "ACDI",
                   "grid": n, "density_ratio": spec["rho12"][0] / spec["rho12"][1],
                   "viscosity_ratio": spec["mu12"][0] / spec["mu12"][1], "precision": "fp64",
                   "step": "AB2 (one reference time step)",
                   "launch": "captured step graphs" if not args.no_graphs else "direct launches",
                   "launch_parameters": r["tuning"]},
        "roofline": roof,
        "step_roofline": {"alg_bytes_per_cell": STEP_ALG_BYTES_PER_CELL, "achieved_GBs": round(step_bw, 1),
                          "frac": round(step_bw / HBM_PEAK_GBS, 4)},
        "cpu_baseline": cpu,
        "hbm_copy_GBs": round(r["copy_gbs"], 1),
        "mixed": mixed,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
