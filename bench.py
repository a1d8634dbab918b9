"""Benchmark of the rfsyn FDTD core on one MI355X.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c3|c2|c1|c4]

One step is one full Yee step (H update, sources, E update with CPML,
materials and dispersive ADE, probes and plane monitors) over the whole
grid.  The default workload is the 512x512x256 periodic dispersive stack
(c3); c4 steps 96 seeded scenes together (fdtd_sim_run_many).  Each
handle first runs the tiling autotuner (fdtd_sim_autotune, before the
first step; --no-autotune keeps the default tiling).  W untimed steps run
first; then K steps are timed with the wall
clock around fdtd_sim_run (which synchronises its stream), with the step's
kernels bracketed by HIP events on the simulation stream.  Inputs are
resident in HBM before the timed region starts.  The run then continues to
the workload's full step count in both precisions so the FP32 probe DFTs
can be compared with the FP64 ones.  Prints one JSON line.

Roofline: the dominant kernel's algorithmic bytes per launch divided by its
average event-timed duration, against the 8 TB/s HBM3E peak.  The bytes
follow the two-sweep accounting of a Yee step (18 values per cell, a
material tag, 8 psi values per cell and CPML axis, 6 per Drude cell, 9 per
Lorentz pole and cell, 6 per tensor cell) over the cells the launch
updates; on the fused step the dominant kernel is the leapfrog launch over
its box, whose dispersive state moves in the split ADE pass (reported
beside it with its own bytes).  "step" gives the survey accounting for the
whole grid against the wall time per step.  "traffic" is the HBM bytes per
leapfrog launch measured with rocprofv3 counters (FETCH_SIZE, WRITE_SIZE)
and scaled by what the same counters read for a copy of known size; it is
read from the committed profile summary of the same workload and dtype
(profiles/), or null when there is none.

CPU baseline: c1 runs the rfsyn reference compiled from its own sources
(oracle/_ref, kind "reference") on one pinned core over the same
low-level calls its validators use; c2 (3x3 tensors) and c3 (Drude/Lorentz
ADE, plane sources) need physics the reference does not have, so they time
the CPU oracle (kind "port") on a bounded sample of steps of the same scene.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

import rfsyn_amd  # noqa: E402
from rfsyn_amd import capi  # noqa: E402
from rfsyn_amd.driver import (NativeRun, build_c3, c1_cpml_params, c1_freqs, c1_probes,  # noqa: E402
                              modgauss_params)

HBM_PEAK_GBS = 8000.0
METRIC = "Mcell-updates/s + %HBM roofline, FP32 vs FP64 at <=1e-4 probe-DFT error"


def slab_cells(n, npml_axes):
    """Cells (i<nx, j<ny, k<nz) whose coordinate along each axis lies in a
    CPML slab (depth d on both faces of that axis; 0 = no CPML)."""
    out = 0
    for a in range(3):
        d = npml_axes[a]
        if d:
            out += min(2 * d, n[a]) * n[(a + 1) % 3] * n[(a + 2) % 3]
    return out


def kernel_bytes(n, s, npml_axes, tag_bytes, ade_values=0):
    """Algorithmic HBM bytes of one H and one E launch.

    H: read Ex Ey Ez Hx Hy Hz, write Hx Hy Hz (9 values per cell) plus, per
    CPML axis a cell lies in, the two psi memories of that axis read and
    written (4 values).  E: the same plus one material tag per node and the
    ADE state values read and written (ade_values per launch)."""
    cells = n[0] * n[1] * n[2]
    slabs = slab_cells(n, npml_axes)
    h = s * (9 * cells + 4 * slabs)
    e = s * (9 * cells + 4 * slabs + ade_values) + tag_bytes * cells
    return h, e


def box_cells(box, n):
    """Cells (coordinate < n on every axis) inside a node box [(lo, hi)] * 3."""
    out = 1
    for a, (lo, hi) in enumerate(box):
        out *= max(0, min(hi, n[a]) - lo)
    return out


def box_slab_cells(box, n, npml_axes):
    """Sum over axes of the cells of the box whose coordinate along that
    axis lies in its CPML slab (depth d on both faces; 0 = none)."""
    out = 0
    for a in range(3):
        d = npml_axes[a]
        if not d:
            continue
        lo, hi = box[a][0], min(box[a][1], n[a])
        inside = max(0, min(hi, d) - lo) + max(0, hi - max(lo, n[a] - d))
        others = 1
        for b in range(3):
            if b != a:
                others *= max(0, min(box[b][1], n[b]) - box[b][0])
        out += inside * others
    return out


def step_bytes(cells, slab, s, tag_bytes, drude=0, lorentz_pole_cells=0, tensor=0):
    """Algorithmic bytes of one Yee step over `cells` cells in the two-sweep
    accounting: 18 values per cell (each component read and written in its
    own sweep and read once as curl source in the other), a material tag,
    8 psi values per cell and CPML axis, 6 per Drude cell (J read and
    written, 3 components), 9 per Lorentz pole and cell, 6 per tensor cell."""
    return (s * (18 * cells + 8 * slab + 6 * drude + 9 * lorentz_pole_cells + 6 * tensor)
            + tag_bytes * cells)


# ------------------------------------------------------------ workloads
class C1:
    """64^3 vacuum, dx = 1 mm, S = 0.99, 10-cell CPML on all faces (m=3,
    R0=1e-8, kappa_max=1, alpha_max=0.05), Ez modulated-Gaussian dipole at
    the centre (30 GHz, 20 GHz -20 dB bandwidth, delay 4 sigma), 8 seeded
    Ez probes with 64-frequency DFTs over 20-40 GHz, 2000 steps."""
    steps_total = 2000
    headline = capi.FDTD_FP64
    n = (64, 64, 64)
    npml = (10, 10, 10)
    tag_bytes = 1
    ade_values = 0
    config = {"workload": "c1: 64^3 vacuum dipole, 10-cell CPML, 8 Ez probes x 64-freq DFT, 2000 steps",
              "grid": [64, 64, 64], "dx_m": 1e-3, "courant": 0.99}

    def build(self, lib, prec, td_cap):
        n = self.n[0]
        run = NativeRun(lib, self.n, 1e-3, 0.99, cpml=c1_cpml_params(10), precision=prec,
                        compensated=1 if prec == capi.FDTD_FP32 else 0)
        run.add_point_source(capi.FIELD_EZ, (n // 2,) * 3, "source_waveform_modulated_gaussian",
                             modgauss_params(30e9, 20e9))
        freqs = c1_freqs()
        pids = [run.add_probe(capi.FIELD_EZ, p, td_cap, freqs) for p in c1_probes(n, (n // 2,) * 3, 8)]
        return run, pids, []

    def cpu_sample(self, lib, steps):
        """The C1 step loop through the reference's low-level calls
        (validate_probe_dft.c:66-82 stepping)."""
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
        return el, n ** 3, "reference"


class C3:
    """512x512x256, dx = 0.5 mm, periodic x/y, 16-cell CPML on +-z, Ex soft
    plane source (40 GHz, 30 GHz bandwidth), seeded Drude grating / 2-pole
    Lorentz / eps=4 stack, two full-plane DFT monitors at 128 frequencies,
    8 Ex probes with 64-frequency DFTs, 6000 steps; FP32 fields, FP64 DFT."""
    steps_total = 6000
    headline = capi.FDTD_FP32
    n = (512, 512, 256)
    npml = (0, 0, 16)
    tag_bytes = 1
    mon_freqs = 128

    def __init__(self):
        from rfsyn_amd.driver import c3_layers
        lay = c3_layers()
        nxy = self.n[0] * self.n[1]
        per, width = lay["period"], max(1, int(round(lay["duty"] * lay["period"])))
        drude_cells = nxy // per * width * (lay["grating"][1] - lay["grating"][0])
        lorentz_cells = nxy * (lay["lorentz"][1] - lay["lorentz"][0])
        # Drude: J per component read+write (6 values); 2 Lorentz poles:
        # J and P per pole per component read+write (24 values)
        self.ade_values = 6 * drude_cells + 24 * lorentz_cells
        self.drude_cells, self.lorentz_pole_cells = drude_cells, 2 * lorentz_cells
        self.lorentz_cells = lorentz_cells
        # ADE pass box: the ADE state box (grating to the end of the Lorentz
        # layer, all x and y)
        self.ade_box_cells = nxy * (lay["lorentz"][1] - lay["grating"][0])
        self.config = {"workload": "c3: 512x512x256 periodic Drude/Lorentz stack, 16-cell z CPML, "
                                   "2 plane monitors x 128 freqs, 6000 steps",
                       "grid": list(self.n), "dx_m": 0.5e-3, "courant": 0.99, "layers": lay}

    def ade_pass_bytes(self, s):
        """Algorithmic bytes of the split ADE pass: a tag per node of its
        box; per dispersive node E^n and the plain part read and E^{n+1}
        written (3 components); J read and written (Drude, 3 components);
        J and P read and written per Lorentz pole (3 components)."""
        return (self.tag_bytes * self.ade_box_cells + s * 9 * (self.drude_cells + self.lorentz_cells)
                + s * 6 * self.drude_cells + s * 12 * self.lorentz_pole_cells)

    def build(self, lib, prec, td_cap):
        run, pids, mons, _ = build_c3(lib, prec, td_cap, self.n)
        return run, pids, mons

    def cpu_sample(self, lib, steps):
        run, pids, mons = self.build(lib, capi.FDTD_FP64, 0)
        run.run(1)    # first touch of the CPU arrays, untimed
        t0 = time.perf_counter()
        run.run(steps)
        el = time.perf_counter() - t0
        run.close()
        return el, self.n[0] * self.n[1] * self.n[2], "port"


class C2:
    """256^3, dx = 1 mm, 12-cell CPML (C1 grading), 40 seeded spheres
    (radius 4-24 cells, eps_r 1-12) and an eps_r = 6 slab tilted 23 degrees
    about x, with 3x3 subpixel inverse-permittivity tensors on interface
    edges; Ez dipole at the centre, 16 Ez probes with 64-frequency DFTs,
    3000 steps.  FP32 headline, FP64 alongside."""
    steps_total = 3000
    headline = capi.FDTD_FP32
    n = (256, 256, 256)
    npml = (12, 12, 12)

    def __init__(self):
        self.tag_bytes = 2
        self.ade_values = 0
        self.config = {"workload": "c2: 256^3 40 spheres + 23-degree tilted slab, 3x3 subpixel tensors, "
                                   "12-cell CPML, 16 Ez probes x 64-freq DFT, 3000 steps",
                       "grid": list(self.n), "dx_m": 1e-3, "courant": 0.99}

    def build(self, lib, prec, td_cap, record=True):
        """The scene painted by lib (on the device for the product library);
        with record, its counts and paint time go into the config (the last
        device build, FP64, paints with the runtime already initialised)."""
        from rfsyn_amd.driver import build_c2
        stats = {}
        run, pids = build_c2(lib, prec, td_cap, stats_out=stats)
        if record:
            self.config["scene"] = stats
        return run, pids, []

    def cpu_sample(self, lib, steps):
        run, pids, _ = self.build(lib, capi.FDTD_FP64, 0, record=False)
        run.run(1)
        t0 = time.perf_counter()
        run.run(steps)
        el = time.perf_counter() - t0
        run.close()
        return el, self.n[0] * self.n[1] * self.n[2], "port"


WORKLOADS = {"c1": C1, "c2": C2, "c3": C3, "c4": None, "c5": None, "c5train": None}


def sur_gemm_flops(n, grid=64, chans=(3, 16, 32, 64, 128), width=512, queries=8, outputs=128, coords=78):
    """(forward, training step) FLOPs of the surrogate's matrix products
    for a batch of n scenes, 2 M N K per product as the network issues
    them (rfsyn_amd/csrc/surrogate.c: forward, backward)."""
    fwd = bwd = 0
    S = [grid >> l for l in range(len(chans))]
    for l in range(len(chans) - 1):
        rows, K, N = n * S[l + 1] ** 3, 27 * chans[l], chans[l + 1]
        fwd += 2 * rows * K * N
        bwd += 2 * K * N * rows + 2 * rows * N + (2 * rows * K * N if l > 0 else 0)
    F = S[-1] ** 3 * chans[-1]
    q, W, O = n * queries, width, outputs
    fwd += 2 * n * F * W + 2 * q * coords * W + 2 * q * W * W + 2 * q * W * O
    bwd += (2 * W * O * q + 2 * q * O + 2 * q * O * W + 2 * W * W * q + 2 * q * W + 2 * q * W * W
            + 2 * coords * W * q + 2 * q * W + 2 * F * W * n + 2 * n * F * W)
    return fwd, fwd + bwd


def bench_c5train(args, lib):
    """The C5 network's Adam step alone at its training batch (256 scenes of
    the dispersive stack distribution as 64^3 maps, 8 queries, random
    labels; no FDTD): W untimed then K timed steps, each ending with the
    loss read back.  Roofline: the step's product FLOPs against the BF16
    MFMA peak."""
    from rfsyn_amd import surrogate_pipeline as sp
    bt = args.c5_batch
    rng = np.random.default_rng(3)
    scs = [sp.scene(i) for i in range(bt)]
    maps = np.stack([sp.scene_map(sc) for sc in scs])
    coords = np.stack([sp.scene_coords(sc) for sc in scs])
    labels = rng.standard_normal((bt, sp.N_PROBES, 2 * sp.N_FREQS)).astype(np.float32)
    net = sp.make_net(lib, batch=bt)
    for _ in range(args.warmup):
        net.train_step(maps, coords, labels, args.c5_lr)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        net.train_step(maps, coords, labels, args.c5_lr)
    wall = time.perf_counter() - t0
    net.close()
    flops = sur_gemm_flops(bt)[1]
    tf = flops * args.steps / wall / 1e12
    print(json.dumps({
        "metric": "surrogate training step throughput (BF16 products, FP32 accumulation)",
        "value": round(tf, 3), "unit": "TFLOP/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "bf16", "data": "synthetic",
        "config": {"workload": "c5 network training step: 64^3 maps, 4 stride-2 3x3x3 convs 16-32-64-128, "
                               "MLP width 512, 8 queries x 128 outputs", "batch": bt},
        "roofline": {"bound": "mfma", "kernel": "training step (all products)", "achieved": round(tf, 3),
                     "peak": 2500.0, "unit": "TFLOP/s", "frac": round(tf / 2500.0, 5), "traffic": None,
                     "flops_per_step": flops}}))


def bench_c5(args, lib):
    """Learned surrogate plus FDTD judge on 64^3 scenes of the periodic
    dispersive stack distribution: label train and test scenes with the
    batched GPU FDTD, train the BF16 network, then infer the test scenes at
    batch 1024 with a 5% FDTD spot check and FDTD fallback above 2e-2
    relative L2.  Reports scenes/s of both paths and the accuracy."""
    from rfsyn_amd import surrogate_pipeline as sp
    steps = args.c5_fdtd_steps
    n_tr, n_te = args.c5_train, args.c5_test
    log = lambda m: print(f"[c5] {m}", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    mtr, ctr, ltr, fdtd_tr = sp.label_scenes(lib, list(range(n_tr)), steps, batch=args.c5_label_batch, log=log)
    mte, cte, lte, fdtd_te = sp.label_scenes(lib, list(range(n_tr, n_tr + n_te)), steps, batch=args.c5_label_batch,
                                             log=log)
    label_wall = time.perf_counter() - t0
    log(f"labelled {n_tr + n_te} scenes: {label_wall:.1f} s wall, {fdtd_tr + fdtd_te:.1f} s of FDTD")
    # the plain single-scene solver on the same scenes: one handle at a time
    n_one = 16
    _, _, _, fdtd_one = sp.label_scenes(lib, list(range(n_one)), steps, batch=1)
    log(f"single-scene FDTD: {n_one} scenes in {fdtd_one:.2f} s")
    scale = float(np.std(ltr))
    net = sp.make_net(lib, batch=1024)
    t1 = time.perf_counter()
    bt = min(args.c5_batch, n_tr)
    losses = sp.train(net, mtr, ctr, ltr / scale, args.c5_epochs, lr=args.c5_lr, batch=bt,
                      log=log, schedule="cos")
    train_s = time.perf_counter() - t1
    n_steps_tr = args.c5_epochs * max(1, n_tr // bt)
    step_flops = sur_gemm_flops(bt)[1]
    # accuracy on the held-out scenes, and the predictor without the network
    pred = sp.predict(net, mte, cte, 1024) * scale
    err = sp.rel_l2(pred, lte)
    base = sp.rel_l2(np.broadcast_to(ltr.mean(axis=0), lte.shape), lte)
    # inference with the FDTD judge
    out, inf_t, fallbacks, spot = sp.spot_checked_inference(lib, net, list(range(n_tr, n_tr + n_te)), mte, cte,
                                                            steps, 1024, scale=scale)
    net.close()
    # GPU time of each path (the FDTD's host-side scene set-up is left out
    # of both, as in the labelling rate)
    fdtd_rate = (n_tr + n_te) / (fdtd_tr + fdtd_te)
    sur_rate = n_te / (inf_t["network"] + inf_t["fdtd"])
    res = {
        "metric": "scenes/s (surrogate with FDTD spot checks vs batched FDTD), relative L2 vs FDTD",
        "value": round(sur_rate, 3), "unit": "scenes/s", "n_gpus": 1, "steps": steps, "warmup": 0,
        "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "bf16",
        "data": "synthetic",
        "config": {"workload": f"c5: 64^3 periodic dispersive stacks, {n_tr} train / {n_te} held-out scenes, "
                               f"8 probes x 64-freq transfer functions, FDTD {steps} steps",
                   "epochs": args.c5_epochs, "lr": args.c5_lr, "lr_schedule": "cosine", "batch_train": bt,
                   "batch_infer": 1024},
        "roofline": {"bound": "mfma", "kernel": "training step (all matrix products, BF16 operands, FP32 accumulation)",
                     "achieved": round(step_flops * n_steps_tr / train_s / 1e12, 3), "peak": 2500.0,
                     "unit": "TFLOP/s", "frac": round(step_flops * n_steps_tr / train_s / 1e12 / 2500.0, 5),
                     "traffic": None, "flops_per_step": step_flops,
                     "ms_per_train_step": round(train_s / n_steps_tr * 1e3, 3),
                     "note": "wall time of whole Adam steps (products, im2col, GELU, loss, Adam, host uploads); "
                             "the products' own share is in the rocprof summary"},
        "fdtd_scenes_per_s": round(fdtd_rate, 3),
        "fdtd_single_scene_per_s": round(n_one / fdtd_one, 3),
        "batching_gain": round(fdtd_rate / (n_one / fdtd_one), 3),
        "speedup_vs_fdtd": round(sur_rate / fdtd_rate, 3),
        "network_only_scenes_per_s": round(n_te / max(inf_t["network"], 1e-9), 1),
        "held_out_rel_l2_median": float(np.median(err)), "held_out_rel_l2_p90": float(np.percentile(err, 90)),
        "mean_predictor_rel_l2_median": float(np.median(base)),
        "spot_checks": len(spot), "spot_check_rel_l2_max": float(np.max(spot)) if spot else None,
        "fallback_batches": fallbacks, "train_seconds": round(train_s, 2), "train_loss_last": losses[-1],
        "inference_seconds": {k: round(v, 3) for k, v in inf_t.items()},
        "label_wall_seconds": round(label_wall, 2),
    }
    print(json.dumps(res))


def bench_c4(args, lib):
    """96 seeded scenes (sizes U{48..192} per axis, 10-cell CPML, 1-12
    lossy boxes, one dipole, 4 probes x 64-frequency DFT), FP32, advanced
    together by a packed batch (fdtd_batch_*: one leapfrog launch per
    tiling variant, one launch each for all shells' H and E, one for all
    probe gathers per step, a block of steps captured as one graph);
    Mcell-updates/s over all scenes.  The same handles are then stepped
    with one stream per handle (the earlier batching) for comparison."""
    from rfsyn_amd.driver import Batch, build_c4_scene, c4_scene, run_many
    specs = [c4_scene(q) for q in range(96)]
    cells = sum(sp["n"][0] * sp["n"][1] * sp["n"][2] for sp in specs)
    runs = [build_c4_scene(lib, sp, capi.FDTD_FP32, 0) for sp in specs]
    handles = [r for r, _ in runs]
    if not args.no_autotune:   # per scene tiling; the scenes stay on the fused step (batched)
        t_tune = time.perf_counter()
        for r in handles:
            assert lib.fdtd_sim_autotune(r.sim, 2, None) == 0, lib.fdtd_last_error()
        print(f"[c4] autotuned 96 scenes in {time.perf_counter() - t_tune:.1f} s", file=sys.stderr, flush=True)
    batch = Batch(lib, handles)
    batch.run(args.warmup)
    t0 = time.perf_counter()
    batch.run(args.steps)
    wall = time.perf_counter() - t0
    launches = batch.launches()
    batch.close()
    wall_streams = one = None
    if not args.c4_packed_only:
        os.environ["RFSYN_PACKED"] = "0"
        run_many(lib, handles, args.warmup)
        t1 = time.perf_counter()
        run_many(lib, handles, args.steps)
        wall_streams = time.perf_counter() - t1
        os.environ.pop("RFSYN_PACKED")
        one = time.perf_counter()
        for r in handles[:8]:   # the same 8 scenes one after another, for the batching gain
            r.run(args.steps)
        one = time.perf_counter() - one
    cells8 = sum(sp["n"][0] * sp["n"][1] * sp["n"][2] for sp in specs[:8])
    slab = sum(slab_cells(sp["n"], (10, 10, 10)) for sp in specs)
    byt = step_bytes(cells, slab, 4, 1)
    for r in handles:
        r.close()
    value = cells * args.steps / wall / 1e6
    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "Mcell-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(wall / args.steps * 1e3, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "c4: 96 seeded scenes, n U{48..192} per axis, 10-cell CPML, 1-12 lossy boxes, "
                               "dipole, 4 probes x 64-freq DFT, all scenes stepped together",
                   "total_cells": cells},
        "roofline": {"bound": "hbm", "kernel": "step (all scenes)", "achieved": round(byt / (wall / args.steps) / 1e9, 2),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(byt / (wall / args.steps) / 1e9 / HBM_PEAK_GBS, 4), "traffic": None,
                     "bytes_per_launch": byt},
        "packed_launches_per_step": {"leapfrog": launches[0], "shell": launches[1], "ade_pass": launches[2],
                                     "probes": 1},
        "streams_per_scene_mcell_s": round(cells * args.steps / wall_streams / 1e6, 3) if wall_streams else None,
        "sequential_8_scenes_mcell_s": round(cells8 * args.steps / one / 1e6, 3) if one else None,
        "autotuned": not args.no_autotune,
    }
    print(json.dumps(out))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(wl, budget_s=15.0):
    ref_so = os.path.join(REPO, "oracle", "_ref", "libref_fdtd.so")
    port_so = os.path.join(REPO, "oracle", "liboracle_fdtd.so")
    old = os.sched_getaffinity(0)
    cpu = min(old)
    os.sched_setaffinity(0, {cpu})
    try:
        path = ref_so if isinstance(wl, C1) and os.path.exists(ref_so) else port_so
        if not os.path.exists(path):
            return None
        lib = capi.bind(C.CDLL(path))
        if isinstance(wl, (C2, C3)):
            steps = 2 if isinstance(wl, C3) else 4   # seconds per step on one core
        else:
            t, cells, kind = wl.cpu_sample(lib, 5)
            steps = max(5, min(wl.steps_total, int(budget_s / max(t / 5, 1e-9))))
        t, cells, kind = wl.cpu_sample(lib, steps)
    finally:
        os.sched_setaffinity(0, old)
    if path == port_so:
        kind = "port"
    return {"value": round(cells * steps / t / 1e6, 3), "unit": "Mcell-updates/s", "cores": 1, "kind": kind,
            "cpu_model": cpu_model(),
            "sample": f"{wl.config['workload'].split(':')[0]} scene, {steps} steps on 1 pinned core "
                      f"(host has {os.cpu_count()} cpus)"}


def copy_calibration():
    import torch
    x = torch.ones((256 << 20) // 4, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(5):
        y.copy_(x)
    torch.cuda.synchronize()
    print("copied 256 MiB x 5", file=sys.stderr)


def profile_traffic(workload, dtype):
    """HBM bytes per leapfrog launch from the committed profile summary of
    this workload and dtype (lines 'workload: ...', 'dtype: ...' and
    'traffic_bytes_per_launch: N'), or None."""
    import glob
    for path in sorted(glob.glob(os.path.join(REPO, "profiles", "*.md"))):
        kv = {}
        for line in open(path):
            if ":" in line and not line.startswith((" ", "\t", "|", "`")):
                k, v = line.split(":", 1)
                kv[k.strip()] = v.strip()
        if kv.get("workload") == workload and kv.get("dtype") == dtype and "traffic_bytes_per_launch" in kv:
            try:
                return int(kv["traffic_bytes_per_launch"])
            except ValueError:
                return None
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--workload", default="c3", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true",
                    help="skip running to the full step count for the FP32-vs-FP64 comparison")
    ap.add_argument("--c5-train", type=int, default=20000, help="c5: training scenes")
    ap.add_argument("--c5-test", type=int, default=4096, help="c5: held-out scenes")
    ap.add_argument("--c5-epochs", type=int, default=150, help="c5: training epochs")
    ap.add_argument("--c5-lr", type=float, default=3e-4, help="c5: Adam learning rate (cosine-annealed)")
    ap.add_argument("--c5-batch", type=int, default=256, help="c5: training batch")
    ap.add_argument("--c5-label-batch", type=int, default=256, help="c5: scenes per packed FDTD batch")
    ap.add_argument("--c5-fdtd-steps", type=int, default=1024, help="c5: FDTD steps per label")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default tiling of the fused step instead of timing the candidates first")
    ap.add_argument("--c4-packed-only", action="store_true",
                    help="c4: time the packed batch only, without the one-stream-per-handle and sequential "
                         "comparison runs")
    ap.add_argument("--fp32-rows", type=int, default=0,
                    help="c1-c3: run the FP32 fused step at this many rows per work item (1: row mode, every z "
                         "column per item; 4 or 8: 62-column tiles) instead of autotuning its tiling")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 256 MiB device to device 5 times (the known-size kernel that "
                         "calibrates the FETCH_SIZE / WRITE_SIZE counters) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for {args.gpus}); "
              "multi-GPU sharding is designed but not built", file=sys.stderr)
        sys.exit(2)

    lib = rfsyn_amd.load_library()
    if args.workload == "c4":
        bench_c4(args, lib)
        return
    if args.workload == "c5":
        bench_c5(args, lib)
        return
    if args.workload == "c5train":
        bench_c5train(args, lib)
        return
    wl = WORKLOADS[args.workload]()
    timed_steps = args.warmup + args.steps
    total = max(timed_steps, 0 if args.no_accuracy else wl.steps_total)
    res = {}
    for prec in (wl.headline, capi.FDTD_FP32 if wl.headline == capi.FDTD_FP64 else capi.FDTD_FP64):
        tag = "fp32" if prec == capi.FDTD_FP32 else "fp64"
        fixed = args.fp32_rows and prec == capi.FDTD_FP32
        if fixed:   # read by the library at the handle's first run
            os.environ["RFSYN_LEAP_ROWS"] = str(args.fp32_rows)
        run, pids, mons = wl.build(lib, prec, total)
        chosen = (C.c_int * 4)(args.fp32_rows if fixed else 0, 0, 0, 0)
        two = C.c_int(0)
        if not args.no_autotune and not fixed:
            assert lib.fdtd_sim_autotune(run.sim, 3, chosen) == 0, lib.fdtd_last_error()
            assert lib.fdtd_sim_select_path(run.sim, 3, C.byref(two)) == 0, lib.fdtd_last_error()
        run.run(args.warmup)
        if fixed:
            del os.environ["RFSYN_LEAP_ROWS"]
        run.set_profiling(True)
        t0 = time.perf_counter()
        run.run(args.steps)
        wall = time.perf_counter() - t0
        (h_ms, e_ms, sh_ms), nk = run.kernel_times3()
        fused = lib.fdtd_sim_fused(run.sim) == 1
        box = None
        if fused:
            lo, hi = (C.c_int * 3)(), (C.c_int * 3)()
            assert lib.fdtd_sim_fused_box(run.sim, lo, hi) == 0
            box = [(lo[q], hi[q]) for q in range(3)]
        run.set_profiling(False)
        print(f"[bench] {args.workload} {tag}: {wall / args.steps * 1e3:.4f} ms/step, "
              f"H/fused {h_ms / nk * 1e3:.1f} us, E/ADE pass {e_ms / nk * 1e3:.1f} us, "
              f"shell {sh_ms / nk * 1e3:.1f} us", file=sys.stderr, flush=True)
        if total > timed_steps:
            run.run(total - timed_steps)
        dft = np.stack([run.dft(p) for p in pids])
        # the monitors' DFT of the source's component (Ex, the first tangential one)
        planes = [run.plane_dft(m, 0, (wl.mon_freqs, wl.n[0], wl.n[1])) for m in mons]
        block = lib.fdtd_sim_dft_block(run.sim)
        run.close()
        res[prec] = dict(wall=wall, h_ms=h_ms / nk, e_ms=e_ms / nk, sh_ms=sh_ms / nk, dft=dft, fused=fused, box=box,
                         planes=planes, block=block,
                         tiling=None if (args.no_autotune and not fixed) or not fused else
                         {"rows": chosen[0], "planes_per_item": chosen[1], "z_split": chosen[2], "coop": chosen[3],
                          "row_mode": prec == capi.FDTD_FP32 and chosen[0] == 1})

    n = wl.n
    cells = n[0] * n[1] * n[2]
    r = res[wl.headline]
    value = cells * args.steps / r["wall"] / 1e6
    s = 8 if wl.headline == capi.FDTD_FP64 else 4
    drude = getattr(wl, "drude_cells", 0)
    lorentz = getattr(wl, "lorentz_pole_cells", 0)
    if r["fused"]:
        # the fused launch on its box (first event slot): the two-sweep
        # accounting of one step over the cells of the box without the ADE
        # state, which the split ADE pass (second slot) moves; the shell
        # kernels around the box are the third slot
        box = r["box"]
        ade = r["e_ms"] > 0 and (drude or lorentz)
        byt = step_bytes(box_cells(box, n), box_slab_cells(box, n, wl.npml), s, wl.tag_bytes,
                         0 if ade else drude, 0 if ade else lorentz)
        dom, dur = "leapfrog", r["h_ms"]
        detail = {"box": box, "box_cells": box_cells(box, n), "shell_avg_launch_us": round(r["sh_ms"] * 1e3, 3)}
        if ade and hasattr(wl, "ade_pass_bytes"):
            ab = wl.ade_pass_bytes(s)
            detail["ade_pass"] = {"bytes_per_launch": ab, "avg_launch_us": round(r["e_ms"] * 1e3, 3),
                                  "achieved": round(ab / (r["e_ms"] * 1e-3) / 1e9, 2), "unit": "GB/s"}
    else:
        hb, eb = kernel_bytes(n, s, wl.npml, wl.tag_bytes, wl.ade_values)
        dom, dur, byt = ("update_e", r["e_ms"], eb) if r["e_ms"] >= r["h_ms"] else ("update_h", r["h_ms"], hb)
        detail = {"h_bytes_per_launch": hb, "h_avg_launch_us": round(r["h_ms"] * 1e3, 3),
                  "e_bytes_per_launch": eb, "e_avg_launch_us": round(r["e_ms"] * 1e3, 3)}
    achieved = byt / (dur * 1e-3) / 1e9
    full = step_bytes(cells, slab_cells(n, wl.npml), s, wl.tag_bytes, drude, lorentz)
    mon = None
    if r["planes"]:
        # per plane per step: 2 tangential components read, buffered and read
        # back (6 s per cell); per flush of K steps their FP64 complex
        # accumulators read and written (64 F per cell) over K steps
        n_plane = n[0] * n[1]
        mon_b = len(r["planes"]) * (6 * s * n_plane + 64 * wl.mon_freqs * n_plane // r["block"])
        full += mon_b
        mon = {"planes": len(r["planes"]), "freqs": wl.mon_freqs, "dft_block": r["block"], "bytes_per_step": mon_b}
    step_gbs = full / (r["wall"] / args.steps) / 1e9
    d64, d32 = res[capi.FDTD_FP64]["dft"], res[capi.FDTD_FP32]["dft"]
    peak = np.max(np.abs(d64), axis=1, keepdims=True)
    # a short run whose probes have not seen the pulse yet has no error to report
    err = float(np.max(np.abs(d32 - d64) / peak)) if np.all(peak > 0) else None
    err_detail = None
    if err is not None:
        e = np.abs(d32 - d64) / peak   # [probe, frequency], normalised by the probe's in-band peak
        sig = lambda a: [float(f"{x:.3g}") for x in a]
        err_detail = {"normalisation": "|X32 - X64| / max_f |X64(probe, f)|",
                      "per_probe_max": sig(e.max(axis=1)), "per_freq_max_over_probes": sig(e.max(axis=0))}
        if r["planes"]:
            mon_err = []
            for p64, p32 in zip(res[capi.FDTD_FP64]["planes"], res[capi.FDTD_FP32]["planes"]):
                num = np.sqrt(np.sum(np.abs(p32 - p64) ** 2, axis=(1, 2)))
                den = np.sqrt(np.sum(np.abs(p64) ** 2, axis=(1, 2)))
                rel = num / np.where(den > 0, den, 1.0)
                mon_err.append({"peak_normalised_max": float(f"{(num / den.max()).max():.3g}"),
                                "max": float(f"{rel.max():.3g}"), "per_freq": sig(rel)})
            err_detail["plane_monitors_rel_l2_per_freq"] = {
                "normalisation": "||X32 - X64||_2 / ||X64||_2 over the plane at each frequency, Ex "
                                 "(peak_normalised_max: divided by the largest ||X64||_2 over frequency instead, the "
                                 "probes' convention; the per-frequency ratio has no bound at a spectral null)",
                "R": mon_err[0], "T": mon_err[1] if len(mon_err) > 1 else None}
    other = res[capi.FDTD_FP32 if wl.headline == capi.FDTD_FP64 else capi.FDTD_FP64]
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
        "dtype": "f64" if wl.headline == capi.FDTD_FP64 else "f32",
        "data": "synthetic",
        "config": wl.config,
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": profile_traffic(args.workload, "f64" if wl.headline == capi.FDTD_FP64 else "f32")
                     if r["fused"] else None,
                     "bytes_per_launch": byt, "avg_launch_us": round(dur * 1e3, 3), **detail},
        "step": {"bytes_per_step": full, "achieved_gbs": round(step_gbs, 2),
                 "frac": round(step_gbs / HBM_PEAK_GBS, 4), "monitors": mon},
        "other_precision": {"dtype": "f32" if wl.headline == capi.FDTD_FP64 else "f64",
                            "value": round(cells * args.steps / other["wall"] / 1e6, 3)},
        "fp32_vs_fp64_probe_dft_err": err,
        "fp32_vs_fp64_err_detail": err_detail,
        "tiling": r["tiling"],
        "step_path": "fused" if r["fused"] else "two-kernel",
        "accuracy_steps": total,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(wl)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
