"""Throughput benchmark of the MI355X time-stepping core.

Workload: 512^3 Yee grid, one Lorentz pole (deps 1.1, Om = 2 omega_L) + Kerr (chi3 2e-22 m^2/V^2),
16-cell CPML on every face, 20 fs / 60-cell-waist short pulse at 5e9 V/m.  A step is one full
Yee step (E side then B side) of the whole grid; value = cell-updates/s = Nx*Ny*Nz*steps / time
with all state resident in HBM.  The timed steps start after a pre-roll of 880 steps, when the
pulse has entered the grid (field everywhere near the source plane, Kerr solves iterating).

Headline form: FP32-res -- E, B and the CPML auxiliaries as float32 words each backed by a
10-bit residual slot, the PLRC accumulators and all arithmetic (the Kerr solve included) in
float64 -- the float32-storage form that meets the per-field accuracy bar against float64
(plain float32 words do not: their cross-polarised components sit at the float32 storage
floor).  The faster of its fused and single-pass structures is the headline line.  FP32-bulk (plain float32), the
float64 fused, single-pass and reference-structured forms are timed beside it; the CPU oracle
(OpenMP) is timed on a bounded sample of the same workload, and on one core on a 192^3 sample.
The 64-run batched parameter sweep (see batch_sweep), the adjoint-gradient workload (see
adjoint_gradient), the focused-pulse workload (see focused_pulse) and the field error of the
float32 forms against float64 after 1000 steps (see accuracy) are measured beside it.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--preroll P] [--n 512] [--forms a,b] [--no-cpu]
                [--no-batch] [--no-adjoint] [--no-accuracy] [--no-focused]
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


def step_bytes(prec, n, cpml_d, n_poles):
    """Algorithmic HBM bytes of one step: every mutable state word read once and written once,
    neighbour re-reads free, static tables ignored.  Per interior cell 2 (6 w_f + r + 6 P 8),
    w_f = 8 (float64) or 4 (float32 forms), r = 8 for the two residual words of FP32-res (E, B);
    per CPML face-layer cell (2 d (NxNy + NyNz + NxNz) of them) 8 w_c, plus 16 for the residual
    words of the E- and H-side auxiliaries in FP32-res.  Returns (total, E side, B side): the
    E side is E, the accumulators and the E auxiliaries, the B side B and the H auxiliaries."""
    cells = n ** 3
    wf = 8 if prec == 0 else 4
    r = 4 if prec == 2 else 0  # one residual word per cell and field group
    layer = 2 * cpml_d * 3 * n * n
    e_side = cells * (2 * (3 * wf + r) + 2 * 6 * 8 * n_poles) + layer * (4 * wf + 2 * r)
    b_side = cells * 2 * (3 * wf + r) + layer * (4 * wf + 2 * r)
    return e_side + b_side, e_side, b_side


def two_kernel_bytes(prec, n, cpml_d, n_poles):
    """HBM bytes the two fused kernels move if every word they touch crosses HBM once per
    launch: the E kernel also reads B, the B kernel also reads E (the split's extra bytes)."""
    _, e_side, b_side = step_bytes(prec, n, cpml_d, n_poles)
    wf = 8 if prec == 0 else 4
    r = 4 if prec == 2 else 0
    extra = n ** 3 * (3 * wf + r)
    return e_side + extra, b_side + extra


def run_gpu(desc, steps, warmup, preroll):
    """Time `steps` steps after `warmup` from the start (zero-field window) and again after
    `preroll` steps (pulse window); per-kernel durations from HIP events on the solver's own
    stream over the pulse window."""
    import torch

    from mpulse_amd import abi
    s = abi.Solver(desc)

    def window():
        s.sync()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        s.step(steps)
        s.sync()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3 / steps

    s.step(warmup)
    early_ms = window()
    s.step(max(0, preroll - s.time))
    s.step(warmup)
    wall_ms = window()
    prof = s.time_steps(min(steps, 50))
    s.close()
    torch.cuda.empty_cache()
    return wall_ms, early_ms, prof


def cpu_baseline(n=512, n_sample_steps=20, n_single=192, single_steps=10):
    """The oracle (C restatement, float64, OpenMP over x planes) on the same 512^3 workload for
    n_sample_steps steps after one untimed step, and on one core on the same workload at
    n_single^3 (single_steps steps after one untimed step)."""
    import psutil
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import Oracle

    from mpulse_amd import configs
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count()
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next(l.split(":", 1)[1].strip() for l in f if l.startswith("model name"))
    except (OSError, StopIteration):
        pass

    def rate(size, steps, omp):
        d, _ = configs.config_roofline(precision=0, n=size)
        o = Oracle(d, omp=omp)
        o.step(1)  # first touch of the state
        t0 = time.perf_counter()
        o.step(steps)
        dt = time.perf_counter() - t0
        o.close()
        return size ** 3 * steps / dt, dt

    v, dt = rate(n, n_sample_steps, True)
    v1, dt1 = rate(n_single, single_steps, False)
    return {"value": v, "unit": "cell-updates/s", "cores": threads, "kind": "port",
            "sample": f"oracle (C restatement, float64, OpenMP, {threads} threads) on the full {n}^3 workload, "
                      f"{n_sample_steps} steps after 1 untimed step, {dt:.1f} s; no MPI launcher on the host",
            "single_core": {"value": v1, "unit": "cell-updates/s",
                            "sample": f"serial oracle, same workload at {n_single}^3, {single_steps} steps after "
                                      f"1 untimed step, {dt1:.1f} s"},
            "host": {"cpu": model, "physical_cores": psutil.cpu_count(logical=False),
                     "logical_cpus": psutil.cpu_count(), "threads_used": threads}}


def batch_sweep(steps=500, nruns=64, autotune_steps=10):
    """Uneven-size parameter sweep: nruns independent cubes (48^3..160^3, one pole + Kerr,
    10-cell CPML, 500 steps each) stepped together by the batched kernels replayed from
    captured graphs with autotuned tiles, in float64 and in FP32-bulk, against the same runs
    launched one after another on the float64 plain (reference-structured) path and on the
    path of the batch's tuned structure (fused, sweep or pass) in the same precision.  Every
    batched run's E and B after the last step are compared bit for bit with that run on its
    own (SHA-256 of the six fields).  Times cover the steps only (set-up excluded), wall clock
    around each set."""
    import hashlib

    import torch

    from mpulse_amd import abi, configs

    def digest(view):
        h = hashlib.sha256()
        for name in ("Ex", "Ey", "Ez", "Bx", "By", "Bz"):
            h.update(view.get(name).tobytes())
        return h.hexdigest()

    def sequential(descs, path, digests=None):
        total = 0.0
        for d in descs:
            d.path = path
            s = abi.Solver(d)
            s.sync()
            t0 = time.perf_counter()
            s.step(steps)
            s.sync()
            total += time.perf_counter() - t0
            if digests is not None:
                digests.append(digest(s))
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
        digests = [digest(b.run(r)) for r in range(len(descs))]
        b.close()
        torch.cuda.empty_cache()
        return t, tune, digests

    descs64 = [d for d, _ in configs.config_sweep(nruns=nruns, steps=steps, precision=abi.PREC_FP64)]
    cells = sum(d.nx * d.ny * d.nz for d in descs64)
    out = {"runs": nruns, "steps": steps, "cells_per_step": cells, "sizes": [d.nx for d in descs64]}
    t_plain = sequential(descs64, abi.PATH_REFERENCE)
    out["sequential_plain"] = {"precision": "fp64", "s": t_plain, "cell_updates_per_s": cells * steps / t_plain}
    print(f"# sequential_plain (float64): {t_plain:.3f} s", file=sys.stderr, flush=True)
    for name, prec in (("fp64", abi.PREC_FP64), ("fp32_bulk", abi.PREC_FP32_BULK)):
        descs = [d for d, _ in configs.config_sweep(nruns=nruns, steps=steps, precision=prec)]
        t_b, tune, got = batched(descs)
        same_path = {"fused": abi.PATH_FUSED, "sweep": abi.PATH_SWEEP, "pass": abi.PATH_PASS}[tune["structure"]]
        want = []
        t_f = sequential(descs, same_path, want)
        identical = [g == w for g, w in zip(got, want)]
        t_fused = t_f if same_path == abi.PATH_FUSED else sequential(descs, abi.PATH_FUSED)
        out[name] = {"batched": {"s": t_b, "cell_updates_per_s": cells * steps / t_b, "tuning": tune},
                     "sequential_same_path": {"path": tune["structure"], "s": t_f,
                                              "cell_updates_per_s": cells * steps / t_f},
                     "speedup_vs_sequential_plain": t_plain / t_b,
                     "speedup_vs_sequential_same_path": t_f / t_b,
                     "sequential_fused": {"s": t_fused, "cell_updates_per_s": cells * steps / t_fused},
                     "speedup_vs_sequential_fused": t_fused / t_b,
                     "runs_bit_identical_to_sequential": int(sum(identical)),
                     "all_runs_bit_identical": all(identical)}
        print(f"# batch {name}: {t_b:.3f} s for {nruns} runs x {steps} steps, tuning {tune}; "
              f"{t_plain / t_b:.2f}x sequential float64 plain, {t_f / t_b:.2f}x sequential {name} "
              f"{tune['structure']} ({t_f:.3f} s), {t_fused / t_b:.2f}x sequential {name} fused ({t_fused:.3f} s); "
              f"{sum(identical)}/{nruns} runs bit-identical to their "
              f"sequential runs", file=sys.stderr, flush=True)
    return out


def focused_pulse():
    """The focused-pulse self-focusing workload: 384^2 x 768, two Lorentz poles, Kerr from n2,
    focused pulse (NA 0.1, 30 fs, 5e16 W/m^2), 16-cell CPML, its full 2000 steps timed after 10
    untimed ones: FP32-res on the pass and fused paths, FP32-bulk on the fused path, and a float64
    run of the same workload whose fields after the 2000 steps each float32 run is compared with,
    per component (relative L2, bar 1e-5), with the peak |Ex| on the focal plane."""
    import numpy as np
    import torch

    from mpulse_amd import abi, configs
    fields = ("Ex", "Ey", "Ez", "Bx", "By", "Bz")

    def run(prec, path, keep):
        d, meta = configs.config_focus(precision=prec, path=path)
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
            got = {f: s.get(f) for f in fields} if keep else None
        torch.cuda.empty_cache()
        r = {"grid": [d.nx, d.ny, d.nz], "poles": 2, "steps": steps, "s": t, "ms_per_step": t / steps * 1e3,
             "cell_updates_per_s": cells * steps / t, "kernels_ms": {k: v[0] for k, v in prof["kernels"].items()}}
        return r, got

    out = {}
    ref_r, ref = run(abi.PREC_FP64, abi.PATH_FUSED, True)
    out["fp64_fused"] = ref_r
    nz = ref_r["grid"][2]
    ref_r["peak_abs_Ex_focal_plane"] = float(np.max(np.abs(ref["Ex"][:, :, nz // 2])))
    for name, prec, path in (("fp32_res_pass", abi.PREC_FP32_RES, abi.PATH_PASS),
                             ("fp32_res_fused", abi.PREC_FP32_RES, abi.PATH_FUSED),
                             ("fp32_bulk_fused", abi.PREC_FP32_BULK, abi.PATH_FUSED)):
        r, got = run(prec, path, True)
        r["peak_abs_Ex_focal_plane"] = float(np.max(np.abs(got["Ex"][:, :, nz // 2])))
        rel = {f: float(np.linalg.norm(got[f] - ref[f]) / np.linalg.norm(ref[f])) for f in fields}
        del got
        r["rel_l2_vs_fp64"] = rel
        r["bar_met"] = {f: rel[f] <= 1e-5 for f in fields}
        out[name] = r
        print(f"# focused pulse {r['grid'][0]}^2 x {r['grid'][2]}, 2 poles, {name}: {r['s']:.2f} s for {r['steps']} "
              f"steps, {r['ms_per_step']:.3f} ms/step, {r['cell_updates_per_s'] / 1e9:.2f} G cell-updates/s; vs "
              f"float64: " + ", ".join(f"{f} {rel[f]:.2e}" for f in fields), file=sys.stderr, flush=True)
    print(f"# focused pulse float64 fused: {ref_r['ms_per_step']:.3f} ms/step; peak |Ex| on the focal plane "
          f"{ref_r['peak_abs_Ex_focal_plane']:.4f}", file=sys.stderr, flush=True)
    return out


def accuracy(n=512, steps=1000):
    """Relative L2 error of the float32 forms' fields against the float64 fields after `steps`
    steps of the same workload (fused path), per component and for E and B as vector fields,
    with the per-field bar (1e-5) met or not."""
    import numpy as np
    import torch

    from mpulse_amd import abi, configs
    fields = ("Ex", "Ey", "Ez", "Bx", "By", "Bz")
    d64, _ = configs.config_roofline(precision=abi.PREC_FP64, n=n)
    with abi.Solver(d64) as a:
        a.step(steps)
        ref = {f: a.get(f) for f in fields}
    torch.cuda.empty_cache()
    res = {"grid": [n, n, n], "steps": steps, "bar": 1e-5}
    for name, prec in (("fp32_res", abi.PREC_FP32_RES), ("fp32_bulk", abi.PREC_FP32_BULK)):
        d, _ = configs.config_roofline(precision=prec, n=n)
        out, num, den = {}, {"E": 0.0, "B": 0.0}, {"E": 0.0, "B": 0.0}
        with abi.Solver(d) as b:
            b.step(steps)
            for f in fields:
                x, y = b.get(f), ref[f]
                dd, yy = float(np.sum((x - y) ** 2)), float(np.sum(y ** 2))
                out[f] = (dd / yy) ** 0.5
                num[f[0]] += dd
                den[f[0]] += yy
                del x
        torch.cuda.empty_cache()
        for g in ("E", "B"):
            out[g] = (num[g] / den[g]) ** 0.5
        res[name] = {"rel_l2": out, "bar_met": {f: out[f] <= 1e-5 for f in fields},
                     "all_fields_met": all(out[f] <= 1e-5 for f in fields)}
        print(f"# {name} vs float64 after {steps} steps at {n}^3: " +
              ", ".join(f"{f} {out[f]:.2e}" for f in fields), file=sys.stderr, flush=True)
    return res


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
        fc = {"h": h, "cells": errs, "max_rel_eps": max(e["rel"] for e in errs), "max_abs_grad": gmax,
              "chi_fd": fdc, "chi_rel": abs(gchi - fdc) / abs(fdc), "bar": 1e-4}
        fc["pass"] = fc["max_rel_eps"] <= 1e-4 and fc["chi_rel"] <= 1e-4
        out["fd_check"] = fc
        print(f"# adjoint vs finite differences: max rel (eps) {fc['max_rel_eps']:.2e}, chi rel "
              f"{fc['chi_rel']:.2e}: {'pass' if fc['pass'] else 'FAIL'} (bar 1e-4)", file=sys.stderr, flush=True)
    return out


def load_traffic(name):
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


FORMS = {
    # name: (precision, path, dominant kernel class)
    "fp32_res_fused": (2, 0, "fused_E"),
    "fp32_res_pass": (2, 3, "pass"),
    "fp32_bulk_fused": (1, 0, "fused_E"),
    "fp32_bulk_pass": (1, 3, "pass"),
    "fp32_bulk_sweep": (1, 2, "sweep"),
    "fp64_fused": (0, 0, "fused_E"),
    "fp64_pass": (0, 3, "pass"),
    "fp64_sweep": (0, 2, "sweep"),
    "fp64_reference": (0, 1, "stepD_ref"),
}
DEFAULT_FORMS = ("fp32_res_fused", "fp32_res_pass", "fp32_bulk_fused", "fp32_bulk_pass", "fp64_fused",
                 "fp64_pass", "fp64_reference")
PREC_NAME = {0: "fp64", 1: "fp32-bulk", 2: "fp32-res"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--preroll", type=int, default=880,
                    help="steps run before the timed window (the pulse reaches the grid near step 880)")
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--forms", default="", help="comma list of " + ", ".join(FORMS))
    ap.add_argument("--copy-probe", action="store_true", help="also time a 4 GiB device copy")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-run batched sweep")
    ap.add_argument("--batch-only", action="store_true", help="only the 64-run batched sweep")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint-gradient workload")
    ap.add_argument("--adjoint-only", action="store_true", help="only the adjoint-gradient workload")
    ap.add_argument("--no-focused", action="store_true", help="skip the focused-pulse workload")
    ap.add_argument("--no-accuracy", action="store_true",
                    help="skip the 1000-step comparison of the float32 forms with float64")
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
    names = [f for f in args.forms.split(",") if f in FORMS] if args.forms else list(DEFAULT_FORMS)
    cells = n ** 3
    d_cpml = 16
    results = {}
    for name in names:
        prec, path, kclass = FORMS[name]
        d, meta = configs.config_roofline(precision=prec, path=path, n=n)
        wall_ms, early_ms, prof = run_gpu(d, args.steps, args.warmup, args.preroll)
        kern = {k: v[0] for k, v in prof["kernels"].items()}
        alg, e_side, b_side = step_bytes(prec, n, d_cpml, 1)
        kb = e_side if kclass == "fused_E" else alg  # the dominant kernel's share
        r = {"ms_per_step": wall_ms, "cell_updates_per_s": cells / (wall_ms * 1e-3),
             "zero_field_ms_per_step": early_ms, "kernels_ms": kern,
             "algorithmic_bytes_per_step": alg, "step_frac": alg / (wall_ms * 1e-3) / 1e9 / HBM_PEAK_GBS}
        if kclass in kern:
            r["dominant"] = {"kernel": kclass, "algorithmic_bytes": kb, "mean_launch_ms": kern[kclass],
                             "frac": kb / (kern[kclass] * 1e-3) / 1e9 / HBM_PEAK_GBS}
        results[name] = r
        print(f"# {name}: {wall_ms:.3f} ms/step ({early_ms:.3f} from zero fields), "
              f"{cells / (wall_ms * 1e-3) / 1e9:.2f} G cell-updates/s, step frac {r['step_frac']:.3f}, "
              f"kernels {prof['kernels']}", file=sys.stderr, flush=True)

    heads = [f for f in ("fp32_res_fused", "fp32_res_pass") if f in results]
    if not heads:
        heads = [f for f in results if FORMS[f][0] != 0]
    if not heads:
        raise SystemExit("the headline needs a float32 form")
    head = max(heads, key=lambda f: results[f]["cell_updates_per_s"])
    main_res = results[head]
    prec, path, kclass = FORMS[head]
    alg, e_side, b_side = step_bytes(prec, n, d_cpml, 1)
    dom = main_res["dominant"]
    achieved = dom["algorithmic_bytes"] / (dom["mean_launch_ms"] * 1e-3) / 1e9
    traffic = load_traffic(f"{kclass}_{PREC_NAME[prec]}_{n}_traffic.json")
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
        "dtype": ("fp32-res (f32 E/B/CPML words + 10-bit residual slots, f64 PLRC psi and Kerr solve, "
                  "f64 arithmetic)" if prec == 2
                  else "fp32-bulk (f32 E/B/CPML, f64 psi and Kerr, f64 arithmetic)"),
        "data": "synthetic (zero initial fields, analytic short-pulse source; timed after a pre-roll of "
                f"{args.preroll} steps with the pulse in the grid)",
        "config": {"workload": f"{n}^3 PLRC+Kerr, 1 Lorentz pole, 16-cell CPML, 20 fs short pulse, "
                               f"{PREC_NAME[prec]}", "grid": [n, n, n], "cpml": d_cpml, "poles": 1, "kerr": True,
                   "form": head, "preroll_steps": args.preroll, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": kclass, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic["bytes_per_launch"] if traffic else None,
                     "algorithmic_bytes_per_launch": dom["algorithmic_bytes"],
                     "algorithmic_share": "E side of the step's algorithmic bytes (E, accumulators, E auxiliaries)"
                                          if kclass == "fused_E" else "the whole step",
                     "mean_launch_ms": dom["mean_launch_ms"]},
        "step_roofline": {"algorithmic_bytes_per_step": alg,
                          "bytes_per_cell_update": alg / cells,
                          "achieved_GBs": alg / (main_res["ms_per_step"] * 1e-3) / 1e9,
                          "frac": main_res["step_frac"], "bar": 0.70, "bar_met": main_res["step_frac"] >= 0.70},
        "forms": results,
    }
    if path == 0:
        ek, bk = two_kernel_bytes(prec, n, d_cpml, 1)
        out["step_roofline"]["traffic_model_two_kernels"] = {
            "bytes_per_step": ek + bk, "over_algorithmic": (ek + bk) / alg,
            "note": "the fused E kernel also reads B and the fused B kernel also reads E"}
        tb = load_traffic(f"fused_B_{PREC_NAME[prec]}_{n}_traffic.json")
        if traffic and tb:
            out["step_roofline"]["measured_traffic_per_step"] = traffic["bytes_per_launch"] + tb["bytes_per_launch"]
            out["step_roofline"]["measured_over_algorithmic"] = (traffic["bytes_per_launch"] +
                                                                 tb["bytes_per_launch"]) / alg
    if args.copy_probe:
        nb = 4 << 30
        ms = abi.copy_probe(nb, 20)
        out["copy_probe"] = {"bytes_moved": 2 * nb, "ms": ms, "GBs": 2 * nb / (ms * 1e-3) / 1e9}
    for ref in ("fp64_reference", "fp64_fused"):
        if ref in results:
            out[f"value_over_{ref}"] = main_res["cell_updates_per_s"] / results[ref]["cell_updates_per_s"]
    if not args.no_accuracy:
        out["accuracy_vs_fp64"] = accuracy(n)
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
