"""Benchmark: simulated days per wall-day of the GFS spectral dycore step.

Workload (default T254L64): 768x384 Gauss grid, triangular truncation 254,
64 hybrid levels, dry Held-Suarez forcing, IMEX RK3 with dt = 600 s, from a
resting isothermal atmosphere (300 K, 1000 hPa) plus seeded random vorticity
(rms 1e-6 s^-1 in total wavenumbers 1-21). One "step" is one call of
run.f90 advance: three getdyntend passes with semi-implicit solves, implicit
damping, the inverse-transform pass, Held-Suarez physics and the dry-mass
fixer, all on the GPU, replayed from a HIP graph.

value = mixed-precision (FP32 Legendre/FFT/columns, FP64 tables and global
sums) simulated days per wall-day; the FP64 GPU path is timed beside it and
the spectral error of the mixed path after --error-steps steps (default one
model day) against the FP64 path is reported. roofline = the dominant kernel
of the step, timed with HIP events on the context stream; cpu_baseline = the
FP64 CPU oracle timed on one step of the same workload on the host cores.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "gfs-dycore_amd"))

import gfsdc  # noqa: E402
import levels  # noqa: E402

PEAK = {"fp32_mfma": 157.3e12, "fp64_mfma": 78.6e12, "hbm": 8.0e12}


def make(cfgname, precision):
    c = levels.CONFIGS[cfgname]
    ak, bk = levels.levels_for(cfgname)
    ctx = gfsdc.Context(c["nlon"], c["nlat"], c["ntrunc"], precision, nlev=c["nlev"], ak=ak, bk=bk,
                        dt=c["dt"], physics=gfsdc.PHYS_HS)
    ctx.init_rest_state(300.0, 1.0e5, 1.0e-6, 1, 21, seed=0)
    return ctx


def device_sync(ctx):
    ctx.sync()
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except Exception:
        pass


def time_steps(ctx, steps, warmup):
    ctx.advance(warmup, use_graph=True)
    device_sync(ctx)
    t0 = time.perf_counter()
    ctx.advance(steps, use_graph=True)
    device_sync(ctx)
    return (time.perf_counter() - t0) / steps


def kernel_table(ctx, nprof):
    """mean per-launch time (ms) and launches per step, by kernel name"""
    acc = {}
    for _ in range(nprof):
        for name, ms in ctx.profile_step():
            a = acc.setdefault(name, [0.0, 0])
            a[0] += ms
            a[1] += 1
    return {k: (v[0] / v[1], v[1] / nprof) for k, v in acc.items()}


def algorithmic_work(name, ctx, esz):
    """(kind, work per launch) of one launch of a step kernel: flops for the
    Legendre GEMMs (2 nlat ndimspec per scalar field), bytes otherwise
    (each input / output array read or written once)."""
    nlat, nlon, T, nlm = ctx.nlat, ctx.nlon, ctx.ntrunc, ctx.nlm
    nl = ctx.nlev
    nsyn, nan, nhs = 7 * nl + 3, 4 * nl + 1, 3 * nl + 1
    ng = nlat * nlon
    four = (T + 1) * nlat * 2 * esz  # Fourier coefficients of one field, both hemispheres
    if name == "leg_synth":
        return "mfma", nsyn * 2.0 * nlat * nlm
    if name == "leg_anal":
        return "mfma", None  # batch size differs between dynamics and physics passes
    if name == "fft_synth":
        return "hbm", nsyn * (four + ng * esz)
    if name == "columns":
        return "hbm", ((7 * nl + 3) + (4 * nl + 1) + (2 * nl + 1) + 2) * ng * esz
    return "hbm", None


def spectral_error(a, b):
    sa, sb = a.get_state(), b.get_state()
    return {k: float(np.linalg.norm(sa[k] - sb[k]) / np.linalg.norm(sb[k]))
            for k in ("vrt", "div", "tv", "lnps")}


def cpu_baseline(cfgname):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import cases
    import oracle as O
    c = levels.CONFIGS[cfgname]
    ak, bk = levels.levels_for(cfgname)
    m = O.Model(c["nlon"], c["nlat"], c["ntrunc"], c["nlev"], ak, bk, c["dt"])
    vrt, div, tv, lnps = cases.rest_state(m.sht.degree, m.sht.order, c["nlev"], seed=0)
    m.set_state(vrt[0], div[0], tv[0], lnps[0])
    t0 = time.perf_counter()
    m.advance(1)
    t = time.perf_counter() - t0
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": c["dt"] / t, "unit": "simulated days/wall-day", "cores": cores, "kind": "port",
            "sample": f"1 step (advance: 3 getdyntend + HS physics + mass fixer) of the FP64 C "
                      f"oracle at {cfgname}, OpenMP over fields, {t:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="T254L64", choices=sorted(levels.CONFIGS))
    ap.add_argument("--error-steps", type=int, default=144, help="steps for the mixed-vs-FP64 error (0: skip)")
    ap.add_argument("--prof-steps", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU: the step is timed on a single MI355X; "
              "multi-GPU sharding is described in DESIGN.md but not run", file=sys.stderr)
        sys.exit(2)
    cfg = levels.CONFIGS[args.config]
    dt = cfg["dt"]

    c32 = make(args.config, gfsdc.FP32)
    t32 = time_steps(c32, args.steps, args.warmup)
    out = {"fp32_ms_per_step": t32 * 1e3}
    if not args.no_fp64:
        c64 = make(args.config, gfsdc.FP64)
        t64 = time_steps(c64, args.steps, args.warmup)
        out["fp64_ms_per_step"] = t64 * 1e3
        out["fp64_days_per_day"] = dt / t64
        out["speedup_mixed_vs_fp64"] = t64 / t32
        del c64

    # per-kernel timing (HIP events on the context stream) of the mixed path
    tab = kernel_table(c32, args.prof_steps)
    per_step = {k: v[0] * v[1] for k, v in tab.items()}
    dom = max(per_step, key=per_step.get)
    kind, work = algorithmic_work(dom, c32, 4)
    avg_s = tab[dom][0] * 1e-3
    roof = {"kernel": dom, "bound": kind, "avg_launch_ms": tab[dom][0], "traffic": None}
    if work is not None:
        if kind == "mfma":
            roof.update(achieved=work / avg_s / 1e12, peak=PEAK["fp32_mfma"] / 1e12, unit="TFLOP/s")
        else:
            roof.update(achieved=work / avg_s / 1e9, peak=PEAK["hbm"] / 1e9, unit="GB/s")
        roof["frac"] = roof["achieved"] / roof["peak"]
    leg = tab.get("leg_synth")
    legendre = None
    if leg:
        _, f = algorithmic_work("leg_synth", c32, 4)
        legendre = {"kernel": "leg_synth", "avg_launch_ms": leg[0], "achieved_tflops": f / (leg[0] * 1e-3) / 1e12,
                    "peak_tflops": PEAK["fp32_mfma"] / 1e12,
                    "frac": f / (leg[0] * 1e-3) / PEAK["fp32_mfma"]}

    err = None
    if args.error_steps > 0 and not args.no_fp64:
        a, b = make(args.config, gfsdc.FP32), make(args.config, gfsdc.FP64)
        a.advance(args.error_steps, use_graph=True)
        b.advance(args.error_steps, use_graph=True)
        err = spectral_error(a, b)
        del a, b

    cpu = None if args.no_cpu_baseline else cpu_baseline(args.config)

    line = {
        "metric": "simulated days/wall-day at T254L64, mixed vs FP64 GPU, at stated error",
        "value": dt / t32,
        "unit": "simulated days/wall-day",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t32 * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (resting 300 K isothermal state + seeded random vorticity)",
        "config": {"workload": f"{args.config} dry Held-Suarez dycore step (IMEX RK3, dt={dt:g} s)",
                   "grid": f"{cfg['nlon']}x{cfg['nlat']}", "ntrunc": cfg["ntrunc"], "nlev": cfg["nlev"],
                   "members": 1, "hip_graph": True},
        **out,
        "error_vs_fp64_after_steps": args.error_steps if err else 0,
        "spectral_rel_error_vs_fp64": err,
        "roofline": roof,
        "legendre_gemm": legendre,
        "kernels_ms_per_step": {k: round(v, 4) for k, v in sorted(per_step.items(), key=lambda x: -x[1])},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
