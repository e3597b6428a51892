#!/usr/bin/env python3
"""Benchmark: 3LPT initial conditions at 512^3 on one MI355X.

Workload (the mixed-precision 3LPT configuration): simple-cubic 512^3 lattice,
box 1000 Mpc/h, z_start = 24, MUSIC1 white noise, Eisenstein-Hu P(k) with
Omega_m = 0.3, Omega_b = 0.045, h = 0.7, n_s = 0.96, sigma_8 = 0.81.  FFTs and
Orszag convolutions (768^3 padded) run in FP32; amplitude table, growth
factors and the position/velocity assembly run in FP64.

A step is one whole realisation on the device: noise fill -> r2c -> phi(1) ->
phi(2), phi(3a), phi(3b), A(3) -> assembly of the six growth-scaled
displacement and velocity components -> 6 c2r transforms writing the FP64
positions and velocities, all resident in HBM.  value = particles/s = 512^3 /
(time per step).

Also reported on the same line:
  roofline      achieved HBM GB/s of the dominant kernel (algorithmic bytes
                per launch / average launch time from HIP events on the
                library's stream) against 8 TB/s; traffic = its calibrated
                rocprofv3 FETCH_SIZE + WRITE_SIZE bytes per launch from the
                summary committed under profiles/ for this workload
  fft_gbps      algorithmic bytes / time summed over every FFT pass
  fp64          particles/s of the same workload with FP64 bulk
  accuracy      per output field, FP32 bulk vs FP64 bulk: RMS and max error
                relative to the field RMS, and the P(k) ratio deviation
  cpu_baseline  the FP64 CPU oracle (OpenMP) on a 256^3 3LPT sample (box 500)
  accuracy_vs_oracle  per output field of that 256^3 sample, FP32-bulk and
                FP64 GPU paths vs the FP64 CPU oracle: RMS and max error
                relative to the oracle field's RMS

--config batched reports the 48 uneven 2LPT realisations (16 each at 96^3,
128^3, 192^3) run sequentially, as per-generator graphs on streams, and as
one captured graph holding all 48.

--config eulerian times the semiclassical (PPT) Eulerian fields of a 2LPT
512^3 FP32 realisation (generic_out_eulerian: density and velocities),
replayed as a graph, with the PPT passes' kernel times and the FP64 oracle's
PPT step on a 64^3 sample as the CPU baseline.

--config panphasia times RNG_panphasia::Fill_Grid of a Panphasia-HO
descriptor at 512^3 (a 256^3 coefficient cube, 2.4e7 octree cells), with the
oracle's field at 64^3 as the CPU baseline.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "3LPT IC particles/s at 512^3, FP32 vs FP64 on GPU; 3D FFT achieved HBM GB/s"
HBM_PEAK_GBPS = 8000.0
FFT_KERNELS = ("fft_", "conv_p", "lpt_")


def log(msg):
    sys.stderr.write(msg + "\n")
    sys.stderr.flush()


def run_gpu(M, ctx, n, precision, steps, warmup, fused, profile_step=True, generator="MUSIC1", plt=False,
            boxlen=1000.0):
    params = dict(GridRes=n, BoxLength=boxlen, zstart=24.0, LPTorder=3, precision=precision, generator=generator,
                  seed=12345, fused=bool(fused), PLT=bool(plt))
    ic = M.ICGenerator(ctx, **params)
    for w in range(warmup):
        ic.run()
        ctx.synchronize()
        log(f"[bench] {precision} warmup {w + 1}/{warmup} done")
    ctx.synchronize()
    ctx.record(0)
    t0 = time.perf_counter()
    for s in range(steps):
        ic.run()
        if s + 1 < steps:
            ctx.synchronize()
            log(f"[bench] {precision} step {s + 1}/{steps} done")
    ctx.record(1)
    ctx.synchronize()
    wall = time.perf_counter() - t0
    ev_ms = ctx.elapsed_ms(0, 1)
    log(f"[bench] {precision} {steps} steps: wall {wall:.4f} s, events {ev_ms / 1e3:.4f} s")
    stats = None
    if profile_step:
        ctx.profile_reset()
        ctx.profile(True)
        ic.run()
        ctx.synchronize()
        stats = ctx.profile_stats()
        ctx.profile(False)
        log(f"[bench] {precision} profiled step done")
        calibration_copy(M, ctx, n, precision)
    return wall, ev_ms, stats, ic


def calibration_copy(M, ctx, n, precision):
    """one device-to-device copy of an n^3 grid (n^2 (n+2) reals read and
    written, a blit kernel of known traffic), after the timed steps: the
    reference dispatch for the calibration of the rocprofv3 FETCH_SIZE and
    WRITE_SIZE counters"""
    a = M.Grid_FFT(ctx, n, 1.0, precision)
    b = M.Grid_FFT(ctx, n, 1.0, precision)
    b.copy_from(a)
    ctx.synchronize()
    del a, b


def roofline(stats):
    dom = max(stats.items(), key=lambda kv: kv[1]["ms"])
    name, s = dom
    avg_ms = s["ms"] / s["launches"]
    bytes_per = s["bytes"] / s["launches"]
    achieved = bytes_per / (avg_ms * 1e-3) / 1e9
    fft = [v for k, v in stats.items() if k.startswith(FFT_KERNELS)]
    fft_bytes = sum(v["bytes"] for v in fft)
    fft_ms = sum(v["ms"] for v in fft)
    return {
        "bound": "hbm", "kernel": name, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": None, "launches": s["launches"],
        "avg_ms": round(avg_ms, 4), "bytes_per_launch": bytes_per,
    }, (fft_bytes / (fft_ms * 1e-3) / 1e9 if fft_ms > 0 else None)


def profiled_traffic(kernel, workload):
    """calibrated HBM bytes per launch of `kernel` on `workload`, from the
    rocprofv3 FETCH_SIZE/WRITE_SIZE summaries committed under profiles/"""
    d = os.path.join(ROOT, "profiles")
    if not os.path.isdir(d):
        return None
    for f in sorted(os.listdir(d)):
        if not f.endswith((".md", ".txt")):
            continue
        for line in open(os.path.join(d, f)):
            w = line.split()
            if len(w) == 4 and w[0] == "traffic" and w[1] == kernel and w[2] == workload:
                return float(w[3])
    return None


def accuracy(M, ctx, ic32, ic64, n, boxlen):
    """per output field: RMS and max |FP32 - FP64| relative to the FP64 field's
    RMS (sigma_Psi for displacements), and the largest deviation of the P(k)
    ratio FP32/FP64 over 0.01 < k < k_Nyquist (P(k) from Compute_PowerSpectrum
    of each field)"""
    import numpy as np
    names = ["psi_x", "psi_y", "psi_z", "v_x", "v_y", "v_z"]
    q1 = np.arange(n, dtype=np.float64) / n
    kny = np.pi * n / boxlen
    out = {}
    for w in range(6):
        a = ic32.particles(w)
        b = ic64.particles(w)
        if w < 3:  # positions q + Psi: compare the displacements Psi
            sh = [1, 1, 1]
            sh[w] = n
            a -= q1.reshape(sh)
            b -= q1.reshape(sh)
        sig = float(np.sqrt(np.mean(b * b)))
        d = a - b
        rms = float(np.sqrt(np.mean(d * d))) / sig
        mx = float(np.max(np.abs(d))) / sig
        del d
        pk = []
        for f in (a, b):
            g = M.Grid_FFT(ctx, n, boxlen, "double")
            buf = np.zeros((n, n, n + 2))
            buf[:, :, :n] = f
            g.upload(buf)
            del buf
            g.FourierTransformForward()
            k, P, _, _ = g.Compute_PowerSpectrum()
            pk.append(P)
            g.reset()
        sel = (k > 0.01) & (k < kny) & (pk[1] > 0)
        dev = float(np.max(np.abs(pk[0][sel] / pk[1][sel] - 1)))
        out[names[w]] = {"rms_rel": float(f"{rms:.3e}"), "max_rel": float(f"{mx:.3e}"),
                         "pk_ratio_dev": float(f"{dev:.3e}")}
        del a, b
        log(f"[bench] accuracy {names[w]}: rms {rms:.2e} max {mx:.2e} P(k) {dev:.2e}")
    return out


def cpu_baseline(n=256):
    """FP64 CPU oracle (OpenMP) on an n^3 3LPT sample of the same workload;
    also returns the oracle's six fields for the accuracy check"""
    from oracle import oracle as O
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from test_gpu_lpt import oracle_pipeline
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    cos = O.Cosmology()
    t0 = time.perf_counter()
    fields, _, _ = oracle_pipeline(O, n, 1000.0 * n / 512, 24.0, 3, 12345, cos)
    dt = time.perf_counter() - t0
    return {"value": round(n ** 3 / dt, 1), "unit": "particles/s", "cores": cores, "kind": "port",
            "sample": f"3LPT {n}^3 sc, box {1000.0 * n / 512:g} Mpc/h, MUSIC1, FP64 oracle (OpenMP), {dt:.2f} s"}, fields


def accuracy_vs_oracle(M, ctx, n, fields):
    """per field RMS and max |GPU - oracle| relative to the oracle field's RMS
    (sigma_Psi for displacements) for the FP32-bulk and FP64-bulk GPU paths,
    same seed and configuration as the oracle sample"""
    import numpy as np
    names = ["psi_x", "psi_y", "psi_z", "v_x", "v_y", "v_z"]
    q1 = np.arange(n, dtype=np.float64) / n
    out = {}
    for prec in ("float", "double"):
        ic = M.ICGenerator(ctx, n, 1000.0 * n / 512, 24.0, LPTorder=3, precision=prec, generator="MUSIC1", seed=12345)
        ic.run()
        res = {}
        for w in range(6):
            a = ic.particles(w)
            b = fields[w].copy()
            if w < 3:
                sh = [1, 1, 1]
                sh[w] = n
                a -= q1.reshape(sh)
                b -= q1.reshape(sh)
            sig = float(np.sqrt(np.mean(b * b)))
            d = a - b
            res[names[w]] = {"rms_rel": float(f"{np.sqrt(np.mean(d * d)) / sig:.3e}"),
                             "max_rel": float(f"{np.max(np.abs(d)) / sig:.3e}")}
        out["fp32_bulk" if prec == "float" else "fp64"] = res
        del ic
        log(f"[bench] {prec} vs oracle at {n}^3: max rms {max(v['rms_rel'] for v in res.values()):.2e}")
    return out


def run_eulerian(M, n=512, steps=3, warmup=1):
    """2LPT FP32 realisations with the semiclassical Eulerian fields"""
    import numpy as np
    ctx = M.Context(0)
    ic = M.ICGenerator(ctx, n, 1000.0 * n / 512, 24.0, LPTorder=2, precision="float", generator="MUSIC1",
                       seed=12345, field_output=True, generic_out_eulerian=True)
    ic.capture()
    for _ in range(warmup):
        ic.launch()
    ctx.synchronize()
    ctx.record(0)
    for s in range(steps):
        ic.launch()
        ctx.synchronize()
        log(f"[bench] eulerian step {s + 1}/{steps} done")
    ctx.record(1)
    ctx.synchronize()
    ms = ctx.elapsed_ms(0, 1) / steps
    ctx.profile_reset()
    ctx.profile(True)
    ic.run()
    ctx.synchronize()
    stats = ctx.profile_stats()
    ctx.profile(False)
    ppt = {k: v for k, v in stats.items() if k.startswith("ppt")}
    ppt_ms = sum(v["ms"] for v in ppt.values())
    roof, _ = roofline(ppt)
    delta = ic.density()
    # CPU baseline: the oracle's PPT step (transforms, drift, density,
    # velocities) on a 64^3 sample of growth-scaled potentials
    from oracle import oracle as O
    m = 64
    rng = np.random.default_rng(1)
    phi = np.zeros((m, m, m + 2))
    phi[:, :, :m] = rng.standard_normal((m, m, m))
    t0 = time.perf_counter()
    O.ppt_fields(phi, phi.copy(), 100.0, 0.04, 1.0)
    dt = time.perf_counter() - t0
    return {"metric": "semiclassical Eulerian fields: realisations/s", "value": round(1e3 / ms, 3),
            "unit": "realisations/s", "ms_per_step": round(ms, 3), "steps": steps, "warmup": warmup,
            "higher_is_better": True, "dtype": "f32", "data": "synthetic",
            "config": {"workload": f"2LPT {n}^3 FP32 MUSIC1, generic_out_eulerian (PPT density + velocities)",
                       "n": n, "hbar": ic.hbar(), "delta_std": float(np.std(delta))},
            "ppt_ms": round(ppt_ms, 3), "roofline": roof,
            "cpu_baseline": {"value": round(m ** 3 / dt, 1), "unit": "cells/s (PPT step only)", "cores":
                             int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)), "kind": "port",
                             "sample": f"oracle PPT fields at {m}^3, {dt:.2f} s"},
            "gpu_ppt_cells_per_s": round(n ** 3 / (ppt_ms * 1e-3), 1),
            "kernels": {k: {"launches": v["launches"], "ms": round(v["ms"], 3),
                            "GBps": round(v["bytes"] / (v["ms"] * 1e-3) / 1e9, 1) if v["ms"] > 0 else None}
                        for k, v in sorted(stats.items(), key=lambda kv: -kv[1]["ms"])}}


PAN_DESCRIPTOR = "[Panph6,L20,(424060,82570,148256),S1,KK0,CH-999,Auriga_100_vol2]"


def run_panphasia(M, n=512, steps=3, warmup=1):
    """RNG_panphasia::Fill_Grid of a Panphasia-HO descriptor"""
    ctx = M.Context(0)
    g = M.Grid_FFT(ctx, n, 1000.0, "float")
    rng = M.RNG_panphasia(PAN_DESCRIPTOR)
    for _ in range(warmup):
        rng.Fill_Grid(g)
    ctx.synchronize()
    t0 = time.perf_counter()
    for s in range(steps):
        rng.Fill_Grid(g)
        ctx.synchronize()
        log(f"[bench] panphasia fill {s + 1}/{steps} done")
    ms = (time.perf_counter() - t0) * 1e3 / steps
    info = M.panphasia_info(PAN_DESCRIPTOR, n)
    from oracle import oracle as O
    m = 64
    mi = M.panphasia_info(PAN_DESCRIPTOR, m)
    t1 = time.perf_counter()
    O.panphasia_noise((mi["x"], mi["y"], mi["z"]), mi["level"], mi["size"], mi["kk"], m)
    dt = time.perf_counter() - t1
    return {"metric": "Panphasia-HO white noise: grid points/s", "value": round(n ** 3 / (ms * 1e-3), 1),
            "unit": "points/s", "ms_per_step": round(ms, 3), "steps": steps, "warmup": warmup,
            "higher_is_better": True, "dtype": "f64 cell solve, f32 transforms", "data": "synthetic",
            "config": {"workload": f"Fill_Grid {n}^3, descriptor {PAN_DESCRIPTOR}", "n": n,
                       "coefficient_cube": info["np"], "fourier_grid": info["N0"], "relative_level": info["rel_level"]},
            "cpu_baseline": {"value": round(m ** 3 / dt, 1), "unit": "points/s", "kind": "port",
                             "cores": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)),
                             "sample": f"oracle Panphasia-HO field at {m}^3 ({mi['np']}^3 cube), {dt:.2f} s"}}


BATCH_CLASSES = [(96, 150.0), (128, 200.0), (192, 300.0)]  # (grid, box Mpc/h), 16 realisations each


def run_batched(M, reps=16, streams=4, seed0=1000):
    """The batched uneven 2LPT configuration: 16 realisations each at 96^3,
    128^3 and 192^3 (boxes 150/200/300 Mpc/h), NGENIC noise with seeds
    seed0 + r, FP32 bulk with FP64 assembly.  Timed (a) sequentially, eager
    runs on one stream, and (b) as captured HIP graphs replayed on `streams`
    streams per size class, with the NGENIC seed tables of the next round
    computed on host worker threads while the device runs."""
    kw = dict(zstart=24.0, LPTorder=2, precision="float", generator="NGENIC")
    npart = reps * sum(n ** 3 for n, _ in BATCH_CLASSES)
    # (a) sequential
    # generators are addressed by index only, so that none outlives its context
    ctx = M.Context(0)
    gens = [M.ICGenerator(ctx, n, L, seed=seed0, **kw) for n, L in BATCH_CLASSES]
    for i in range(len(gens)):
        gens[i].run()
    ctx.synchronize()
    t0 = time.perf_counter()
    for i in range(len(gens)):
        for r in range(reps):
            gens[i].set_seed(seed0 + r)
            gens[i].run()
    ctx.synchronize()
    seq_s = time.perf_counter() - t0
    gens.clear()
    log(f"[bench] batched: sequential {seq_s * 1e3:.1f} ms")
    # (b) graphs on streams: generator q = ci * streams + k on context q
    nq = streams * len(BATCH_CLASSES)
    ctxs = [M.Context(0) for _ in range(nq)]
    ggens = [M.ICGenerator(ctxs[q], BATCH_CLASSES[q // streams][0], BATCH_CLASSES[q // streams][1], seed=seed0,
                           **kw) for q in range(nq)]
    for q in range(nq):
        ggens[q].capture()
    for q in range(nq):
        ctxs[q].synchronize()
    rounds = reps // streams
    t0 = time.perf_counter()
    for rnd in range(rounds):
        for q in range(nq):
            ggens[q].set_seed(seed0 + rnd * streams + q % streams)
        for q in range(nq):
            ggens[q].launch()
    for q in range(nq):
        ctxs[q].synchronize()
    graph_s = time.perf_counter() - t0
    log(f"[bench] batched: graphs on {nq} streams {graph_s * 1e3:.1f} ms")
    ggens.clear()
    for q in range(nq):
        ctxs[q].close()
    # (c) all 48 realisations, one generator each, captured into ONE graph
    # whose branches run on `bstreams` streams; seeds staged per generator
    import numpy as np
    bctx = M.Context(0)
    order = [ci for r in range(reps) for ci in range(len(BATCH_CLASSES))]  # interleave the sizes
    bgens = [M.ICGenerator(bctx, BATCH_CLASSES[ci][0], BATCH_CLASSES[ci][1], seed=seed0, **kw) for ci in order]
    bstreams = 16
    batch = M.Batch(bgens, nstreams=bstreams)
    seeds = [seed0 + r for r in range(len(bgens))]
    for g, sd in zip(bgens, seeds):
        g.set_seed(sd)
    batch.launch()  # warm-up
    batch.synchronize()
    for g, sd in zip(bgens, seeds):
        g.set_seed(sd)
    t0 = time.perf_counter()
    batch.launch()
    batch.synchronize()
    one_s = time.perf_counter() - t0
    log(f"[bench] batched: one graph, {len(bgens)} realisations on {bstreams} streams {one_s * 1e3:.1f} ms")
    # check: one realisation of each size equals an eager FP32 run with its
    # seed bit for bit; its RMS distance to the eager FP64-bulk run
    check = []
    for i in range(len(BATCH_CLASSES)):
        n, L = BATCH_CLASSES[order[i]]
        got = [bgens[i].particles(w) for w in range(6)]
        eq = True
        for prec in ("float", "double"):
            e = M.ICGenerator(bctx, n, L, seed=seeds[i], **dict(kw, precision=prec))
            e.run()
            ref = [e.particles(w) for w in range(6)]
            if prec == "float":
                eq = all(np.array_equal(a, b) for a, b in zip(got, ref))
            else:
                q1 = np.arange(n) / n
                errs = []
                for w in range(6):
                    a, b = got[w], ref[w]
                    if w < 3:
                        sh = [1, 1, 1]
                        sh[w] = n
                        a, b = a - q1.reshape(sh), b - q1.reshape(sh)
                    errs.append(float(np.sqrt(np.mean((a - b) ** 2)) / np.sqrt(np.mean(b ** 2))))
            del e
        check.append({"n": n, "seed": seeds[i], "equals_eager_fp32": bool(eq),
                      "max_rms_rel_vs_eager_fp64": float(f"{max(errs):.3e}")})
        log(f"[bench] batched check {n}^3: equal to eager {eq}, vs FP64 {max(errs):.2e}")
    batch.close()
    del g, got, batch
    bgens.clear()
    bctx.close()
    return {
        "metric": "batched uneven 2LPT particles/s (16 x 96^3, 16 x 128^3, 16 x 192^3), one captured graph",
        "value": round(npart / one_s, 1), "unit": "particles/s", "n_gpus": 1, "higher_is_better": True,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "48 realisations, 2LPT sc, NGENIC seeds s0+r, FP32 bulk + FP64 assembly",
                   "realisations": 3 * reps, "particles": npart, "graph_branches": bstreams},
        "one_graph_ms": round(one_s * 1e3, 3),
        "sequential_ms": round(seq_s * 1e3, 3), "graph_streams_ms": round(graph_s * 1e3, 3),
        "graph_streams": nq, "speedup_vs_sequential": round(seq_s / one_s, 3),
        "sequential_particles_per_s": round(npart / seq_s, 1),
        "graph_streams_particles_per_s": round(npart / graph_s, 1),
        "check": check,
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--fused", type=int, default=1, help="1: fused LPT source engine, 0: reference convolve sequence")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64-bulk comparison leg")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--no-accuracy", action="store_true", help="skip the per-field FP32 vs FP64 error report")
    ap.add_argument("--cpu-n", type=int, default=256)
    ap.add_argument("--generator", default="MUSIC1", choices=["MUSIC1", "NGENIC", "PANPHASIA_THREEFRY"],
                    help="white-noise plugin (the 1024^3 configuration uses PANPHASIA_THREEFRY)")
    ap.add_argument("--plt", action="store_true", help="PLT-corrected assembly (the 1024^3 configuration)")
    ap.add_argument("--box", type=float, default=None, help="box length in Mpc/h (default 1000 * n / 512)")
    ap.add_argument("--config", default="c3", choices=["c3", "batched", "eulerian", "panphasia"],
                    help="c3: the 512^3 3LPT headline line; batched: the 48 uneven 2LPT realisations; "
                         "eulerian: semiclassical fields at 512^3; panphasia: Panphasia-HO noise at 512^3")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {args.gpus}); "
              "the path is single-device by design", file=sys.stderr)
        sys.exit(2)

    import monofonic_amd as M
    if args.config == "batched":
        print(json.dumps(run_batched(M)))
        return
    if args.config == "eulerian":
        print(json.dumps(run_eulerian(M, steps=args.steps, warmup=args.warmup)))
        return
    if args.config == "panphasia":
        print(json.dumps(run_panphasia(M, steps=args.steps, warmup=args.warmup)))
        return
    ctx = M.Context(0)
    n = args.n
    box = args.box if args.box else 1000.0 * n / 512
    extra = dict(generator=args.generator, plt=args.plt, boxlen=box)
    wall, ev_ms, stats, ic32 = run_gpu(M, ctx, n, "float", args.steps, args.warmup, args.fused, **extra)
    ms_per_step = ev_ms / args.steps
    value = n ** 3 / (ms_per_step * 1e-3)
    roof, fft_gbps = roofline(stats)
    workload_id = f"3lpt-{n}-f32" + ("" if args.fused else "-seq") + ("-plt" if args.plt else "") + \
        ("" if args.generator == "MUSIC1" else "-" + args.generator.lower())
    tr = profiled_traffic(roof["kernel"], workload_id)
    roof["traffic"] = round(tr) if tr else None
    lpt_path = "stored" if "lpt_zC" in stats else ("streamed" if "lpt_zC1" in stats else None)

    fp64 = None
    acc = None
    if not args.no_fp64:
        w64, e64, st64, ic64 = run_gpu(M, ctx, n, "double", max(1, args.steps // 2), 1, args.fused, **extra)
        ms64 = e64 / max(1, args.steps // 2)
        _, fft64 = roofline(st64)
        fp64 = {"value": round(n ** 3 / (ms64 * 1e-3), 1), "unit": "particles/s", "ms_per_step": round(ms64, 3),
                "fft_gbps": round(fft64, 1) if fft64 else None, "fp32_speedup": round(ms64 / ms_per_step, 3)}
        if not args.no_accuracy:
            fields = accuracy(M, ctx, ic32, ic64, n, box)
            acc = {"reference": "FP64-bulk GPU path, same seed (it matches the FP64 CPU oracle to 1e-11 rms at "
                                "32^3-64^3 in the parity tests)",
                   "normalisation": "rms_rel, max_rel: |FP32 - FP64| / RMS of the FP64 field (sigma_Psi for Psi)",
                   "fields": fields,
                   "max_rms_rel": max(v["rms_rel"] for v in fields.values()),
                   "max_pk_ratio_dev": max(v["pk_ratio_dev"] for v in fields.values())}
        del ic64
    del ic32

    cpu = None
    acc_oracle = None
    if not args.no_cpu:
        log("[bench] CPU oracle baseline ...")
        cpu, ofields = cpu_baseline(args.cpu_n)
        log(f"[bench] CPU oracle baseline done: {cpu['value']} particles/s")
        acc_oracle = {"n": args.cpu_n, "box": 1000.0 * args.cpu_n / 512,
                      "normalisation": "|GPU - FP64 CPU oracle| / RMS of the oracle field",
                      "fields": accuracy_vs_oracle(M, ctx, args.cpu_n, ofields)}
        del ofields

    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "particles/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"3LPT sc {n}^3, box {box:g} Mpc/h, z_start 24, {args.generator} noise, "
                               f"Eisenstein-Hu, FP32 bulk ({3 * n // 2}^3 padded) + FP64 tables/assembly"
                               + (", PLT corrections" if args.plt else ""),
                   "n": n, "lpt_order": 3, "padded": 3 * n // 2, "fused_lpt": bool(args.fused),
                   "lpt_path": lpt_path, "generator": args.generator, "plt": bool(args.plt),
                   "column_tiles": ctx.autotune_report().strip().splitlines(), "wall_s": round(wall, 4)},
        "roofline": roof, "fft_gbps": round(fft_gbps, 1) if fft_gbps else None,
        "fft_frac": round(fft_gbps / HBM_PEAK_GBPS, 4) if fft_gbps else None,
        "fp64": fp64, "accuracy": acc, "accuracy_vs_oracle": acc_oracle, "cpu_baseline": cpu,
        "kernels": {k: {"launches": v["launches"], "ms": round(v["ms"], 3),
                        "GBps": round(v["bytes"] / (v["ms"] * 1e-3) / 1e9, 1) if v["ms"] > 0 else None}
                    for k, v in sorted(stats.items(), key=lambda kv: -kv[1]["ms"])},
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
