"""Throughput benchmark of the MI355X time-stepping core.

Workload: 512^3 Yee grid, one Lorentz pole (deps 1.1, Om = 2 omega_L) + Kerr (chi3 2e-22 m^2/V^2),
16-cell CPML on every face, 20 fs / 60-cell-waist short pulse at 5e9 V/m, FP32-bulk storage
(E, B and CPML auxiliaries float32; PLRC accumulators and the Kerr Newton solve float64).
A step is one full Yee step (E side then B side) of the whole grid; value = cell-updates/s
= Nx*Ny*Nz*steps / time, with all state resident in HBM.  The float64 fused path and the
float64 reference-structured path are timed beside it, and the CPU oracle (OpenMP) is timed
on a bounded sample of the same workload.  The headline form is the faster of the two
FP32-bulk kernel structures (two fused kernels per step, or the single-pass sweep).  The
64-run batched parameter sweep (see batch_sweep), the adjoint-gradient workload (see
adjoint_gradient), the focused-pulse workload (see focused_pulse) and the FP32-bulk field
error against float64 after 1000 steps (see accuracy_fp32_bulk) are measured beside it.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 512] [--no-cpu] [--no-fp64] [--no-batch]
                [--no-adjoint] [--no-accuracy] [--no-focused]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "cell-updates/s at 512³ PLRC+Kerr, FP32-bulk vs FP64, % HBM roofline"
HBM_PEAK_GBS = 8000.0


def kernel_bytes(desc, cpml_d, n_poles):
    """Algorithmic HBM bytes per launch of the fused E and B kernels: every state word the
    kernel owns is read once and written once, neighbour re-reads are free, static tables are
    ignored (B is read-only in the E kernel and E read-only in the B kernel)."""
    nx, ny, nz = desc.nx, desc.ny, desc.nz
    cells = nx * ny * nz
    wf = 4 if desc.precision == 1 else 8      # E, B, CPML psi word
    e_kernel = cells * (2 * 3 * wf + 3 * wf + 2 * 6 * 8 * n_poles)
    b_kernel = cells * (2 * 3 * wf + 3 * wf)
    # CPML face-layer cells (interior part), two live auxiliaries read + written per side
    layer = 2 * cpml_d * (nx * ny + ny * nz + nx * nz)
    e_kernel += layer * 2 * 2 * wf
    b_kernel += layer * 2 * 2 * wf
    return e_kernel, b_kernel


def run_gpu(desc, steps, warmup):
    import torch

    from mpulse_amd import abi
    s = abi.Solver(desc)
    s.step(warmup)
    s.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    s.step(steps)
    s.sync()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    wall_ms = (t1 - t0) * 1e3 / steps
    # per-kernel durations with HIP events on the solver's own stream
    prof = s.time_steps(min(steps, 50))
    s.close()
    torch.cuda.empty_cache()
    return wall_ms, prof


def cpu_baseline(n=512, n_sample_steps=3):
    """Oracle (C, OpenMP over x planes) on the same 512^3 workload for a few steps."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import Oracle

    from mpulse_amd import configs
    d, _ = configs.config_roofline(precision=0, n=n)
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count()
    o = Oracle(d, omp=True)
    o.step(1)  # first touch of the 34 GB state
    t0 = time.perf_counter()
    o.step(n_sample_steps)
    dt = time.perf_counter() - t0
    o.close()
    cells = d.nx * d.ny * d.nz
    return {"value": cells * n_sample_steps / dt, "unit": "cell-updates/s", "cores": threads, "kind": "port",
            "sample": f"oracle (C restatement, float64, OpenMP {threads} threads) on the full {n}^3 workload, "
                      f"{n_sample_steps} steps after 1 untimed step, {dt:.1f} s"}


def batch_sweep(steps=500, nruns=64, autotune_steps=3):
    """Uneven-size parameter sweep: nruns independent cubes (48^3..160^3, one pole + Kerr,
    10-cell CPML, 500 steps each) stepped together by the batched persistent kernels replayed
    from captured graphs with autotuned tiles, in float64 and in FP32-bulk, against the same
    runs launched one after another on the float64 plain (reference-structured) path and on
    the fused path of the same precision.  Times cover the steps only (set-up excluded), wall
    clock around each set."""
    import torch

    from mpulse_amd import abi, configs

    def sequential(descs, path):
        total = 0.0
        for d in descs:
            d.path = path
            s = abi.Solver(d)
            s.sync()
            t0 = time.perf_counter()
            s.step(steps)
            s.sync()
            total += time.perf_counter() - t0
            s.close()
        torch.cuda.empty_cache()
        return total

    def batched(descs):
        b = abi.Batch(descs)
        tune = b.autotune(autotune_steps)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        b.step(steps)
        b.sync()
        t = time.perf_counter() - t0
        b.close()
        torch.cuda.empty_cache()
        return t, tune

    descs64 = [d for d, _ in configs.config_sweep(nruns=nruns, steps=steps, precision=abi.PREC_FP64)]
    cells = sum(d.nx * d.ny * d.nz for d in descs64)
    out = {"runs": nruns, "steps": steps, "cells_per_step": cells, "sizes": [d.nx for d in descs64]}
    t_plain = sequential(descs64, abi.PATH_REFERENCE)
    out["sequential_plain"] = {"precision": "fp64", "s": t_plain, "cell_updates_per_s": cells * steps / t_plain}
    print(f"# sequential_plain (float64): {t_plain:.3f} s", file=sys.stderr, flush=True)
    for name, prec in (("fp64", abi.PREC_FP64), ("fp32_bulk", abi.PREC_FP32_BULK)):
        descs = [d for d, _ in configs.config_sweep(nruns=nruns, steps=steps, precision=prec)]
        t_b, tune = batched(descs)
        t_f = sequential(descs, abi.PATH_FUSED)
        out[name] = {"batched": {"s": t_b, "cell_updates_per_s": cells * steps / t_b, "tuning": tune},
                     "sequential_fused": {"s": t_f, "cell_updates_per_s": cells * steps / t_f},
                     "speedup_vs_sequential_plain": t_plain / t_b,
                     "speedup_vs_sequential_fused": t_f / t_b}
        print(f"# batch {name}: {t_b:.3f} s for {nruns} runs x {steps} steps, tuning {tune}; "
              f"{t_plain / t_b:.2f}x sequential float64 plain, {t_f / t_b:.2f}x sequential {name} fused "
              f"({t_f:.3f} s)", file=sys.stderr, flush=True)
    return out


def focused_pulse():
    """The focused-pulse self-focusing workload: 384^2 x 768, two Lorentz poles, Kerr from n2,
    focused pulse (NA 0.1, 30 fs, 5e16 W/m^2), 16-cell CPML, FP32-bulk on the fused path, its
    full 2000 steps timed after 10 untimed ones; peak |E| on the focal plane at the end."""
    import numpy as np
    import torch

    from mpulse_amd import abi, configs
    d, meta = configs.config_focus(precision=abi.PREC_FP32_BULK)
    steps = meta["steps"]
    cells = d.nx * d.ny * d.nz
    with abi.Solver(d) as s:
        s.step(10)
        s.sync()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        s.step(steps)
        s.sync()
        t = time.perf_counter() - t0
        prof = s.time_steps(20)
        kz = d.nz // 2
        ex = s.get("Ex")[:, :, kz]
    torch.cuda.empty_cache()
    out = {"grid": [d.nx, d.ny, d.nz], "poles": 2, "steps": steps, "s": t, "ms_per_step": t / steps * 1e3,
           "cell_updates_per_s": cells * steps / t, "kernels_ms": {k: v[0] for k, v in prof["kernels"].items()},
           "peak_abs_Ex_focal_plane": float(np.max(np.abs(ex)))}
    print(f"# focused pulse {d.nx}^2 x {d.nz}, 2 poles, FP32-bulk: {t:.2f} s for {steps} steps, "
          f"{out['ms_per_step']:.3f} ms/step, {out['cell_updates_per_s'] / 1e9:.2f} G cell-updates/s",
          file=sys.stderr, flush=True)
    return out


def accuracy_fp32_bulk(n=512, steps=1000):
    """Relative L2 error of the FP32-bulk fields against the float64 fields after `steps` steps
    of the same workload (both on the fused path): E and B as vector fields and per component."""
    import numpy as np
    import torch

    from mpulse_amd import abi, configs
    d64, _ = configs.config_roofline(precision=abi.PREC_FP64, n=n)
    d32, _ = configs.config_roofline(precision=abi.PREC_FP32_BULK, n=n)
    out, num, den = {}, {"E": 0.0, "B": 0.0}, {"E": 0.0, "B": 0.0}
    with abi.Solver(d64) as a, abi.Solver(d32) as b:
        a.step(steps)
        b.step(steps)
        for f in ("Ex", "Ey", "Ez", "Bx", "By", "Bz"):
            x, y = b.get(f), a.get(f)
            dd, yy = float(np.sum((x - y) ** 2)), float(np.sum(y ** 2))
            out[f] = (dd / yy) ** 0.5
            num[f[0]] += dd
            den[f[0]] += yy
            del x, y
    torch.cuda.empty_cache()
    for g in ("E", "B"):
        out[g] = (num[g] / den[g]) ** 0.5
    print(f"# fp32-bulk vs float64 after {steps} steps at {n}^3: rel-L2 E {out['E']:.2e}, B {out['B']:.2e}",
          file=sys.stderr, flush=True)
    return {"grid": [n, n, n], "steps": steps, "rel_l2": out, "bar": 1e-5}


def adjoint_gradient(fd=True):
    """dJ/dEps and dJ/dchi of the plane objective on the 128^2 x 256 float64 workload
    (per-cell permittivity, Kerr, 10-cell CPML, 2000 steps, checkpoints every 50 steps), timed
    next to the forward run alone, and checked against central finite differences (h = 1e-4)
    at 16 seeded cells and for chi."""
    import numpy as np
    import torch

    from mpulse_amd import abi, configs
    d, full, meta = configs.config_adjoint()
    k, w, n, K = meta["k_obj"], meta["weight"], meta["steps"], meta["checkpoint_every"]

    def run(desc, eps_full, grad):
        s = abi.Solver(desc)
        s.set("Eps", eps_full)
        s.sync()
        t0 = time.perf_counter()
        res = s.adjoint_grad(k, w, n, checkpoint_every=K, grad=grad)
        s.sync()
        dt_ = time.perf_counter() - t0
        s.close()
        torch.cuda.empty_cache()
        return res, dt_

    (J0, _, _), t_fwd = run(d, full, False)
    (J, g, gchi), t_adj = run(d, full, True)
    cells = d.nx * d.ny * d.nz
    out = {"grid": [d.nx, d.ny, d.nz], "steps": n, "checkpoint_every": K, "J": J,
           "forward_s": t_fwd, "forward_and_adjoint_s": t_adj, "adjoint_over_forward": t_adj / t_fwd,
           "forward_cell_updates_per_s": cells * n / t_fwd, "dJ_dchi": gchi}
    print(f"# adjoint: J {J:.6e}, forward {t_fwd:.2f} s, forward+adjoint {t_adj:.2f} s", file=sys.stderr,
          flush=True)
    if fd:
        h = meta["fd_h"]
        errs = []
        gmax = float(np.abs(g).max())
        for c in meta["fd_cells"]:
            fp, fm = full.copy(), full.copy()
            fp[c] += h
            fm[c] -= h
            fd_ = (run(d, fp, False)[0][0] - run(d, fm, False)[0][0]) / (2 * h)
            errs.append({"cell": list(c), "adjoint": float(g[c]), "fd": fd_,
                         "rel": abs(g[c] - fd_) / max(abs(fd_), 1e-300)})
        dp, _, _ = configs.config_adjoint()
        dm, _, _ = configs.config_adjoint()
        hc = 1e-4 * d.chi
        dp.chi, dm.chi = d.chi + hc, d.chi - hc
        fdc = (run(dp, full, False)[0][0] - run(dm, full, False)[0][0]) / (2 * hc)
        out["fd_check"] = {"h": h, "cells": errs, "max_rel_eps": max(e["rel"] for e in errs),
                           "max_abs_grad": gmax, "chi_fd": fdc, "chi_rel": abs(gchi - fdc) / abs(fdc)}
        print(f"# adjoint vs finite differences: max rel (eps) {out['fd_check']['max_rel_eps']:.2e}, "
              f"chi rel {out['fd_check']['chi_rel']:.2e}", file=sys.stderr, flush=True)
    return out


def load_traffic(name):
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


FORMS = {
    # name: (precision, path, dominant kernel class)
    "fp32_bulk_fused": (1, 0, "fused_E"),
    "fp32_bulk_sweep": (1, 2, "sweep"),
    "fp64_fused": (0, 0, "fused_E"),
    "fp64_sweep": (0, 2, "sweep"),
    "fp64_reference": (0, 1, "stepD_ref"),
}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--forms", default="", help="comma list of " + ", ".join(FORMS))
    ap.add_argument("--copy-probe", action="store_true", help="also time a 4 GiB device copy")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-run batched sweep")
    ap.add_argument("--batch-only", action="store_true", help="only the 64-run batched sweep")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint-gradient workload")
    ap.add_argument("--adjoint-only", action="store_true", help="only the adjoint-gradient workload")
    ap.add_argument("--no-focused", action="store_true", help="skip the focused-pulse workload")
    ap.add_argument("--no-accuracy", action="store_true",
                    help="skip the 1000-step FP32-bulk against float64 field comparison")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(the sharding design is described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import torch

    from mpulse_amd import abi, configs
    torch.cuda.set_device(0)
    if args.batch_only:
        print(json.dumps({"batch_sweep": batch_sweep()}))
        return
    if args.adjoint_only:
        print(json.dumps({"adjoint_gradient": adjoint_gradient()}))
        return
    n = args.n
    if args.forms:
        names = [f for f in args.forms.split(",") if f in FORMS]
    elif args.no_fp64:
        names = ["fp32_bulk_fused", "fp32_bulk_sweep"]
    else:
        names = list(FORMS)
    cells = n ** 3
    results = {}
    for name in names:
        prec, path, _ = FORMS[name]
        d, meta = configs.config_roofline(precision=prec, path=path, n=n)
        wall_ms, prof = run_gpu(d, args.steps, args.warmup)
        results[name] = {"ms_per_step": wall_ms, "cell_updates_per_s": cells / (wall_ms * 1e-3),
                         "kernels_ms": {k: v[0] for k, v in prof["kernels"].items()}}
        print(f"# {name}: {wall_ms:.3f} ms/step, {cells / (wall_ms * 1e-3) / 1e9:.2f} G cell-updates/s, "
              f"kernels {prof['kernels']}", file=sys.stderr, flush=True)

    fp32 = [f for f in ("fp32_bulk_sweep", "fp32_bulk_fused") if f in results]
    if not fp32:
        raise SystemExit("the headline needs an fp32_bulk form")
    head = max(fp32, key=lambda f: results[f]["cell_updates_per_s"])
    main_res = results[head]
    prec, path, kclass = FORMS[head]
    d, meta = configs.config_roofline(precision=prec, path=path, n=n)
    eb, bb = kernel_bytes(d, meta["d_cpml"], 1)
    wf = 4 if prec == 1 else 8
    # the single-pass kernel reads E and B once for both halves of the step
    step_bytes = eb + bb - 2 * 3 * wf * cells
    k_bytes = step_bytes if path == 2 else eb
    k_ms = main_res["kernels_ms"][kclass]
    achieved = k_bytes / (k_ms * 1e-3) / 1e9
    traffic = load_traffic(f"{kclass}_fp32_{n}_traffic.json")
    alg_step = step_bytes if path == 2 else eb + bb
    out = {
        "metric": METRIC,
        "value": main_res["cell_updates_per_s"],
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": main_res["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32-bulk (f32 E/B/CPML storage, f64 PLRC psi and Kerr solve, f64 arithmetic)",
        "data": "synthetic (zero initial fields, analytic short-pulse source)",
        "config": {"workload": f"{n}^3 PLRC+Kerr, 1 Lorentz pole, 16-cell CPML, 20 fs short pulse, FP32-bulk",
                   "grid": [n, n, n], "cpml": 16, "poles": 1, "kerr": True, "path": head,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": kclass, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic["bytes_per_launch"] if traffic else None,
                     "algorithmic_bytes_per_launch": k_bytes, "mean_launch_ms": k_ms},
        "step_roofline": {"algorithmic_bytes_per_step": alg_step,
                          "achieved_GBs": alg_step / (main_res["ms_per_step"] * 1e-3) / 1e9,
                          "frac": alg_step / (main_res["ms_per_step"] * 1e-3) / 1e9 / HBM_PEAK_GBS},
        "forms": results,
    }
    if args.copy_probe:
        nb = 4 << 30
        ms = abi.copy_probe(nb, 20)
        out["copy_probe"] = {"bytes_moved": 2 * nb, "ms": ms, "GBs": 2 * nb / (ms * 1e-3) / 1e9}
    for ref in ("fp64_reference", "fp64_fused", "fp64_sweep"):
        if ref in results:
            out[f"fp32_bulk_over_{ref}"] = main_res["cell_updates_per_s"] / results[ref]["cell_updates_per_s"]
    if not args.no_accuracy:
        out["accuracy_fp32_bulk_vs_fp64"] = accuracy_fp32_bulk(n)
    if not args.no_focused:
        out["focused_pulse"] = focused_pulse()
    if not args.no_batch:
        out["batch_sweep"] = batch_sweep()
    if not args.no_adjoint:
        out["adjoint_gradient"] = adjoint_gradient()
    if not args.no_cpu:
        try:
            out["cpu_baseline"] = cpu_baseline(n)
        except Exception as exc:  # the baseline is reported, never required
            out["cpu_baseline"] = {"value": None, "error": str(exc)}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
