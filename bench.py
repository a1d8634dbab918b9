"""Benchmark of the rfsyn FDTD core on one MI355X.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c1|...]

One step is one full Yee step (H update, sources, E update with CPML and
materials, probes) over the whole grid.  W untimed steps run first; then
K steps are timed with the wall clock around fdtd_sim_run (which
synchronises its stream), with the H and E kernels each bracketed by HIP
events on the simulation stream.  Inputs are resident in HBM before the
timed region starts.  Prints one JSON line.

Roofline: the dominant kernel's algorithmic bytes per launch (every field
value it must read or write once, psi memories only in CPML slab cells,
one material tag per node) divided by its average event-timed duration,
against the 8 TB/s HBM3E peak.

CPU baseline: the rfsyn reference compiled from its own sources
(oracle/_ref, kind "reference"; the CPU oracle, kind "port", where that
library is absent) on one pinned core, on a bounded sample of the same
scene driven through the same low-level calls the reference's own
validators use.
"""
import argparse
import ctypes as C
import json
import math
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

import rfsyn_amd  # noqa: E402
from rfsyn_amd import capi  # noqa: E402
from rfsyn_amd.driver import NativeRun, c1_cpml_params, c1_freqs, c1_probes, modgauss_params  # noqa: E402

HBM_PEAK_GBS = 8000.0
METRIC = "Mcell-updates/s + %HBM roofline, FP32 vs FP64 at <=1e-4 probe-DFT error"


def slab_counts(n, npml_axes):
    """Cells (i<nx, j<ny, k<nz) whose coordinate along each axis lies in a
    CPML slab, for a CPML of depth npml on that axis (0 = none)."""
    out = []
    for a in range(3):
        d = npml_axes[a]
        if not d:
            out.append(0)
            continue
        # half-node profile nonzero on [0, d) and [n-d, n); integer-node on
        # [0, d) and (n-d, n]: the union restricted to [0, n) is d + d cells
        along = min(2 * d, n[a])
        other = n[(a + 1) % 3] * n[(a + 2) % 3]
        out.append(along * other)
    return out


def kernel_bytes(n, s, npml_axes, tag_bytes):
    cells = n[0] * n[1] * n[2]
    slabs = sum(slab_counts(n, npml_axes))
    h = s * (9 * cells + 4 * slabs)
    e = s * (9 * cells + 4 * slabs) + tag_bytes * cells
    return h, e


# ------------------------------------------------------------ workloads
def build_c1(lib, precision, steps_cap):
    """64^3 vacuum, dx = 1 mm, S = 0.99, 10-cell CPML on all faces (m=3,
    R0=1e-8, kappa_max=1, alpha_max=0.05), Ez modulated-Gaussian dipole at
    the centre (30 GHz, 20 GHz -20 dB bandwidth, delay 4 sigma), 8 seeded Ez
    probes with 64-frequency DFTs over 20-40 GHz."""
    n = 64
    run = NativeRun(lib, (n, n, n), 1e-3, 0.99, cpml=c1_cpml_params(10), precision=precision,
                    compensated=1 if precision == capi.FDTD_FP32 else 0)
    run.add_point_source(capi.FIELD_EZ, (n // 2,) * 3, "source_waveform_modulated_gaussian",
                         modgauss_params(30e9, 20e9))
    freqs = c1_freqs()
    pids = [run.add_probe(capi.FIELD_EZ, p, steps_cap, freqs) for p in c1_probes(n, (n // 2,) * 3, 8)]
    info = {"n": (n, n, n), "npml": (10, 10, 10), "tag_bytes": 1,
            "config": {"workload": "c1: 64^3 vacuum dipole, 10-cell CPML, 8 probes x 64-freq DFT",
                       "grid": [n, n, n], "dx_m": 1e-3, "steps_total": 2000}}
    return run, pids, info


def cpu_c1_sample(lib, steps):
    """The C1 step loop through the reference's low-level calls
    (validate_probe_dft.c:66-82 stepping): H with CPML, soft source at
    t = n dt, E with CPML and (vacuum) materials, 8 DFT probes at
    t = (n+1) dt.  Returns seconds for `steps` steps."""
    n = 64
    g = lib.fdtd_grid_create(n, n, n, 1e-3, 0.99)
    pml = lib.cpml_create(g, c1_cpml_params(10))
    mat = lib.materials_create(g)
    p = modgauss_params(30e9, 20e9)
    wf = capi.waveform_ptr(lib, "source_waveform_modulated_gaussian")
    src = lib.source_make_point(32, 32, 32, capi.FIELD_EZ, capi.SRC_SOFT, 1.0, wf,
                                C.cast(C.pointer(p), C.c_void_p))
    freqs = np.ascontiguousarray(c1_freqs())
    fp = freqs.ctypes.data_as(C.POINTER(C.c_double))
    probes = []
    for q in c1_probes(n, (32, 32, 32), 8):
        pr = capi.ProbeDFT()
        assert lib.probe_dft_create(C.byref(pr), q[0], q[1], q[2], capi.FIELD_EZ, fp, len(freqs)) == 0
        probes.append(pr)
    dt = g.contents.dt
    t0 = time.perf_counter()
    for s in range(steps):
        lib.fdtd_update_h_cpml(g, pml)
        lib.source_apply(C.byref(src), g, s * dt)
        lib.fdtd_update_e_cpml_materials(g, pml, mat)
        for pr in probes:
            lib.probe_dft_record(C.byref(pr), g, (s + 1) * dt, dt)
    el = time.perf_counter() - t0
    for pr in probes:
        lib.probe_dft_destroy(C.byref(pr))
    lib.materials_destroy(mat)
    lib.cpml_destroy(pml)
    lib.fdtd_grid_destroy(g)
    return el, n ** 3


WORKLOADS = {"c1": (build_c1, cpu_c1_sample, 2000, capi.FDTD_FP64)}


def cpu_baseline(workload, budget_s=15.0):
    ref_so = os.path.join(REPO, "oracle", "_ref", "libref_fdtd.so")
    port_so = os.path.join(REPO, "oracle", "liboracle_fdtd.so")
    kind, path = ("reference", ref_so) if os.path.exists(ref_so) else ("port", port_so)
    if not os.path.exists(path):
        return None
    lib = capi.bind(C.CDLL(path))
    sample = WORKLOADS[workload][1]
    old = os.sched_getaffinity(0)
    cpu = min(old)
    os.sched_setaffinity(0, {cpu})
    try:
        t, cells = sample(lib, 5)
        steps = max(5, min(WORKLOADS[workload][2], int(budget_s / max(t / 5, 1e-9))))
        t, cells = sample(lib, steps)
    finally:
        os.sched_setaffinity(0, old)
    return {"value": cells * steps / t / 1e6, "unit": "Mcell-updates/s", "cores": 1, "kind": kind,
            "sample": f"{workload} scene, {steps} steps on 1 pinned core (host has {os.cpu_count()} cpus)"}


def timed(run, steps):
    t0 = time.perf_counter()
    run.run(steps)
    return time.perf_counter() - t0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--workload", default="c1", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for {args.gpus}); "
              "multi-GPU sharding is designed but not built", file=sys.stderr)
        sys.exit(2)

    lib = rfsyn_amd.load_library()
    build, _, steps_total, headline_prec = WORKLOADS[args.workload]
    total = args.warmup + args.steps
    res = {}
    for prec in (capi.FDTD_FP64, capi.FDTD_FP32):
        run, pids, info = build(lib, prec, total)
        run.run(args.warmup)
        run.set_profiling(True)
        wall = timed(run, args.steps)
        h_ms, e_ms, nk = run.kernel_times()
        run.set_profiling(False)
        dft = np.stack([run.dft(p) for p in pids])
        run.close()
        res[prec] = dict(wall=wall, h_ms=h_ms / nk, e_ms=e_ms / nk, dft=dft, info=info)
        print(f"[bench] {args.workload} {'fp32' if prec else 'fp64'}: {wall / args.steps * 1e3:.4f} ms/step, "
              f"H {h_ms / nk * 1e3:.1f} us, E {e_ms / nk * 1e3:.1f} us", file=sys.stderr, flush=True)

    info = res[headline_prec]["info"]
    n = info["n"]
    cells = n[0] * n[1] * n[2]
    r = res[headline_prec]
    value = cells * args.steps / r["wall"] / 1e6
    s = 8 if headline_prec == capi.FDTD_FP64 else 4
    hb, eb = kernel_bytes(n, s, info["npml"], info["tag_bytes"])
    if r["e_ms"] >= r["h_ms"]:
        dom, dur, byt = "update_e", r["e_ms"], eb
    else:
        dom, dur, byt = "update_h", r["h_ms"], hb
    achieved = byt / (dur * 1e-3) / 1e9
    d64, d32 = res[capi.FDTD_FP64]["dft"], res[capi.FDTD_FP32]["dft"]
    peak = np.max(np.abs(d64), axis=1, keepdims=True)
    err = float(np.max(np.abs(d32 - d64) / peak))
    other = res[capi.FDTD_FP32 if headline_prec == capi.FDTD_FP64 else capi.FDTD_FP64]
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["wall"] / args.steps * 1e3, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if headline_prec == capi.FDTD_FP64 else "f32",
        "data": "synthetic",
        "config": info["config"],
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "bytes_per_launch": byt, "avg_launch_us": round(dur * 1e3, 3)},
        "other_precision": {"dtype": "f32" if headline_prec == capi.FDTD_FP64 else "f64",
                            "value": round(cells * args.steps / other["wall"] / 1e6, 3)},
        "fp32_vs_fp64_probe_dft_err": err,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.workload)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
