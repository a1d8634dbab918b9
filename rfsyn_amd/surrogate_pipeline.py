"""Dataset, training and spot-checked inference for the learned surrogate.

Scenes come from the periodic dispersive stack distribution at a fixed
size (the 512x512x256 stack's layer kinds on an n^3 grid): periodic x/y,
CPML on +-z, an Ex soft plane source, a Drude stripe grating, a 2-pole
Lorentz layer and an eps=4 substrate with PCG32-seeded thicknesses and
duty cycle (stream = scene index).  The GPU FDTD labels each scene: 8
seeded Ex probes, each with the transfer function (probe DFT divided by
the source's) at 64 frequencies, real and imaginary parts.

The network sees the scene as an (x, z) map of 3 channels (static
permittivity / 10, Drude mask, Lorentz mask; the scenes are invariant
along y) and per probe the query (probe x, y, z and source x, y, z over
n).  Inference runs the surrogate, spot-checks a seeded share of the
scenes with the FDTD, and falls back to the FDTD for the batch when a
spot check's relative L2 error exceeds the threshold.
"""
import ctypes as C
import json
import math
import os
import time

import numpy as np

from . import capi
from .driver import NativeRun, Pcg32, Surrogate, modgauss_params, run_many

N_PROBES, N_FREQS = 8, 64
F_LO, F_HI = 25e9, 55e9


def scene(index, n=64, npml=8, seed=21):
    """Seeded stack of scene `index` (cells along z) and its probes."""
    rng = Pcg32(seed, index)
    u = n / 64.0   # depths scale with the grid (n = 64: 2-5, 1-3, 3-8, 1-3, 4-10 cells from npml + 10)
    z0 = npml + max(3, int(round(10 * u)))
    t_g = max(1, int(round(u * (2 + rng.next_u32() % 4))))
    g1 = max(1, int(round(u * (1 + rng.next_u32() % 3))))
    t_l = max(1, int(round(u * (3 + rng.next_u32() % 6))))
    g2 = max(1, int(round(u * (1 + rng.next_u32() % 3))))
    t_s = max(1, int(round(u * (4 + rng.next_u32() % 7))))
    grating = (z0, z0 + t_g)
    lorentz = (grating[1] + g1, grating[1] + g1 + t_l)
    sub = (lorentz[1] + g2, lorentz[1] + g2 + t_s)
    duty = 0.3 + 0.4 * rng.uniform()
    period = max(4, n // 4)
    src_k = npml + max(2, int(round(4 * u)))
    assert sub[1] <= n - npml - 2, "stack does not fit the grid"
    probes = []
    for q in range(N_PROBES):
        x, y = rng.next_u32() % n, rng.next_u32() % n
        if q % 2 == 0:   # between the source and the stack (reflected + incident)
            z = src_k + 2 + rng.next_u32() % max(1, grating[0] - src_k - 3)
        else:            # behind the stack (transmitted)
            z = sub[1] + 1 + rng.next_u32() % max(1, n - npml - 2 - sub[1])
        probes.append((x, y, z))
    return {"n": n, "npml": npml, "grating": grating, "lorentz": lorentz, "substrate": sub, "duty": duty,
            "period": period, "src_k": src_k, "probes": probes}


def freqs():
    return np.linspace(F_LO, F_HI, N_FREQS)


def scene_map(sc):
    """[n(x), n(z), 3] input map of a scene."""
    n = sc["n"]
    m = np.zeros((n, n, 3), dtype=np.float32)
    m[:, :, 0] = 0.1
    width = max(1, int(round(sc["duty"] * sc["period"])))
    for i0 in range(0, n, sc["period"]):
        m[i0:i0 + width, sc["grating"][0]:sc["grating"][1], 1] = 1.0
    m[:, sc["lorentz"][0]:sc["lorentz"][1], 0] = 0.225
    m[:, sc["lorentz"][0]:sc["lorentz"][1], 2] = 1.0
    m[:, sc["substrate"][0]:sc["substrate"][1], 0] = 0.4
    return m


OCTAVES = 6
N_COORDS = 6 * (1 + 2 * OCTAVES)


def scene_coords(sc):
    """Query vector per probe: the probe and source positions over n,
    each with sin and cos of pi 2^k times it for k < OCTAVES (the transfer
    function's phase turns by about 0.4 rad per cell along z, which a
    plain coordinate input makes hard to fit)."""
    n = float(sc["n"])
    raw = np.array([[p[0] / n, p[1] / n, p[2] / n, 0.5, 0.5, sc["src_k"] / n] for p in sc["probes"]])
    w = np.pi * 2.0 ** np.arange(OCTAVES)
    ang = raw[:, :, None] * w[None, None, :]
    return np.concatenate([raw, np.sin(ang).reshape(len(raw), -1), np.cos(ang).reshape(len(raw), -1)],
                          axis=1).astype(np.float32)


def _materials(sc):
    def setup(lib, g, mat):
        d = capi.MaterialDispersive()
        d.eps_inf = 1.0
        d.n_drude = 1
        d.drude_wp[0] = 1.37e16
        d.drude_gamma[0] = 1e14
        did = C.c_uint8()
        assert lib.materials_register_dispersive(mat, C.byref(d), C.byref(did)) == 0
        lz = capi.MaterialDispersive()
        lz.eps_inf = 2.25
        lz.n_lorentz = 2
        for q, (de, f0) in enumerate(((1.0, 60e9), (0.5, 90e9))):
            lz.lorentz_deps[q] = de
            lz.lorentz_w0[q] = 2 * math.pi * f0
            lz.lorentz_delta[q] = 0.05 * 2 * math.pi * f0
        lid = C.c_uint8()
        assert lib.materials_register_dispersive(mat, C.byref(lz), C.byref(lid)) == 0
        n = sc["n"]
        width = max(1, int(round(sc["duty"] * sc["period"])))
        for i0 in range(0, n, sc["period"]):
            assert lib.materials_assign_box(mat, g, i0, min(i0 + width, n), 0, n, sc["grating"][0],
                                            sc["grating"][1], did.value) == 0
        assert lib.materials_assign_box(mat, g, 0, n, 0, n, sc["lorentz"][0], sc["lorentz"][1], lid.value) == 0
        assert lib.materials_set_box(mat, g, 0, n, 0, n, sc["substrate"][0], sc["substrate"][1], 4.0, 0.0) == 0
    return setup


def build_scene(lib, sc, precision=capi.FDTD_FP32):
    n = sc["n"]
    run = NativeRun(lib, (n, n, n), 0.5e-3, 0.99, periodic=3,
                    cpml=capi.CpmlParams(sc["npml"], 3.0, 1.0, 1.0, 0.05, 1e-8), precision=precision,
                    compensated=1 if precision == capi.FDTD_FP32 else 0, material_setup=_materials(sc))
    run.add_plane_source(capi.FIELD_EX, 2, sc["src_k"], "source_waveform_modulated_gaussian",
                         modgauss_params(40e9, 30e9))
    pids = [run.add_probe(capi.FIELD_EX, p, 0, freqs()) for p in sc["probes"]]
    return run, pids


def source_spectrum(lib, dt, steps):
    """DFT of the source waveform with the probes' convention (samples at
    t = n dt, weights at t = (n+1) dt as recorded after the E update)."""
    p = modgauss_params(40e9, 30e9)
    wf = capi.waveform_ptr(lib, "source_waveform_modulated_gaussian")
    fn = capi.WAVEFORM(wf)
    t = np.arange(steps) * dt
    w = np.array([fn(float(tt), C.cast(C.pointer(p), C.c_void_p)) for tt in t])
    om = 2 * np.pi * freqs()
    return (w[None, :] * np.exp(-1j * om[:, None] * (t[None, :] + dt))).sum(axis=1) * dt


def label_scenes(lib, indices, steps, batch=128, n=64, npml=8, seed=21, out_dir=None, log=None):
    """(maps [N, n, n, 3], coords [N, 8, 6], labels [N, 8, 128], seconds of
    FDTD) for the scenes, stepped together in batches (fdtd_sim_run_many,
    packed launches).  With out_dir, each scene's transfer functions and
    its JSON description are written there too (write_scene); log, if
    given, receives a line per batch."""
    maps, coords, labels = [], [], []
    spec = None
    elapsed = 0.0
    for b0 in range(0, len(indices), batch):
        scs = [scene(i, n, npml, seed) for i in indices[b0:b0 + batch]]
        runs = [build_scene(lib, sc) for sc in scs]
        t0 = time.perf_counter()
        run_many(lib, [r for r, _ in runs], steps)
        elapsed += time.perf_counter() - t0
        dt = runs[0][0].g.contents.dt
        if spec is None:
            spec = source_spectrum(lib, dt, steps)
        for index, sc, (run, pids) in zip(indices[b0:b0 + batch], scs, runs):
            tf = np.stack([run.dft(p) for p in pids]) / spec[None, :]
            if out_dir is not None:
                write_scene(lib, out_dir, index, sc, tf, dt, steps, seed)
            labels.append(np.concatenate([tf.real, tf.imag], axis=1).astype(np.float32))
            maps.append(scene_map(sc))
            coords.append(scene_coords(sc))
            run.close()
        if log:
            log(f"labelled {len(labels)}/{len(indices)} scenes, {elapsed:.1f} s of FDTD")
    return np.stack(maps), np.stack(coords), np.stack(labels), elapsed


SCENE_FORMAT = "rfsyn-fdtd-scene/1"


def scene_record(index, sc, dt, steps, seed, tf_file):
    """The JSON description of a labelled scene: the grid, the boundary,
    the source, every layer with its material, the probes and the layout of
    its transfer-function file (probe x frequency complex values, the
    probe's DFT over the source waveform's, as fdtd_write_transfer writes
    them)."""
    n = sc["n"]
    return {
        "format": SCENE_FORMAT,
        "index": int(index), "seed": int(seed),
        "grid": [n, n, n], "dx_m": 0.5e-3, "courant": 0.99, "dt_s": float(dt), "steps": int(steps),
        "periodic_axes": ["x", "y"],
        "cpml": {"axes": ["z"], "cells": sc["npml"], "m": 3.0, "ma": 1.0, "kappa_max": 1.0,
                 "alpha_max": 0.05, "r0": 1e-8},
        "source": {"kind": "plane", "soft": True, "component": "Ex", "axis": "z", "index": sc["src_k"],
                   "waveform": "modulated_gaussian", "f0_hz": 40e9, "bandwidth_20db_hz": 30e9},
        "layers": [
            {"name": "grating", "z": list(sc["grating"]), "period_cells": sc["period"], "duty": sc["duty"],
             "width_cells": max(1, int(round(sc["duty"] * sc["period"]))),
             "material": {"eps_inf": 1.0, "drude": [{"wp_rad_s": 1.37e16, "gamma_rad_s": 1e14}]}},
            {"name": "lorentz", "z": list(sc["lorentz"]),
             "material": {"eps_inf": 2.25, "lorentz": [
                 {"delta_eps": 1.0, "w0_rad_s": 2 * math.pi * 60e9, "delta_rad_s": 0.05 * 2 * math.pi * 60e9},
                 {"delta_eps": 0.5, "w0_rad_s": 2 * math.pi * 90e9, "delta_rad_s": 0.05 * 2 * math.pi * 90e9}]}},
            {"name": "substrate", "z": list(sc["substrate"]), "material": {"eps_r": 4.0, "sigma_s_m": 0.0}},
        ],
        "probes": [{"component": "Ex", "node": [int(v) for v in p]} for p in sc["probes"]],
        "freqs_hz": [float(f) for f in freqs()],
        "transfer": {"file": tf_file, "dtype": "complex128 as float64 (re, im) pairs, little-endian",
                     "shape": [len(sc["probes"]), N_FREQS], "order": "probe-major"},
    }


def write_scene(lib, out_dir, index, sc, tf, dt, steps, seed):
    """scene_<index>.tf (fdtd_write_transfer) and scene_<index>.json."""
    os.makedirs(out_dir, exist_ok=True)
    tf = np.ascontiguousarray(tf, dtype=np.complex128)
    re, im = np.ascontiguousarray(tf.real), np.ascontiguousarray(tf.imag)
    name = f"scene_{index:06d}"
    dp = C.POINTER(C.c_double)
    if lib.fdtd_write_transfer(os.path.join(out_dir, name + ".tf").encode(), re.ctypes.data_as(dp),
                               im.ctypes.data_as(dp), tf.shape[0], tf.shape[1]) != 0:
        raise OSError(f"writing {name}.tf failed")
    with open(os.path.join(out_dir, name + ".json"), "w") as f:
        json.dump(scene_record(index, sc, dt, steps, seed, name + ".tf"), f, indent=1)


def read_scene(lib, out_dir, index):
    """(record, transfer functions [probes, freqs] complex128) of a scene
    written by write_scene."""
    with open(os.path.join(out_dir, f"scene_{index:06d}.json")) as f:
        rec = json.load(f)
    if rec.get("format") != SCENE_FORMAT:
        raise ValueError(f"not a {SCENE_FORMAT} record")
    p, nf = rec["transfer"]["shape"]
    re, im = np.zeros((p, nf)), np.zeros((p, nf))
    dp = C.POINTER(C.c_double)
    if lib.fdtd_read_transfer(os.path.join(out_dir, rec["transfer"]["file"]).encode(), re.ctypes.data_as(dp),
                              im.ctypes.data_as(dp), p, nf) != 0:
        raise OSError(f"reading {rec['transfer']['file']} failed")
    return rec, re + 1j * im


def make_net(lib, n=64, batch=256, seed=1, width=512, channels=(16, 32, 64, 128)):
    return Surrogate(lib, n, 3, channels, N_COORDS, width, 2 * N_FREQS, N_PROBES, batch, seed=seed)


def train(net, maps, coords, labels, epochs, lr=3e-4, batch=256, seed=5, log=None, schedule="const"):
    """Adam over shuffled batches; schedule "cos" anneals the learning
    rate from lr to 0 over the run (half a cosine period)."""
    rng = np.random.default_rng(seed)
    N = maps.shape[0]
    losses = []
    total = epochs * max(1, N // batch)
    step = 0
    for ep in range(epochs):
        order = rng.permutation(N)
        tot, cnt = 0.0, 0
        for b0 in range(0, N - batch + 1, batch):
            idx = order[b0:b0 + batch]
            rate = lr * 0.5 * (1.0 + math.cos(math.pi * step / total)) if schedule == "cos" else lr
            tot += net.train_step(maps[idx], coords[idx], labels[idx], rate)
            cnt += 1
            step += 1
        losses.append(tot / max(cnt, 1))
        if log:
            log(f"epoch {ep}: train loss {losses[-1]:.4e}")
    return losses


def predict(net, maps, coords, batch):
    out = []
    for b0 in range(0, maps.shape[0], batch):
        out.append(net.predict(maps[b0:b0 + batch], coords[b0:b0 + batch]))
    return np.concatenate(out)


def rel_l2(pred, ref):
    """Relative L2 error per (scene, probe) over the 128 outputs."""
    return np.linalg.norm(pred - ref, axis=-1) / np.maximum(np.linalg.norm(ref, axis=-1), 1e-30)


def spot_checked_inference(lib, net, indices, maps, coords, steps, batch, share=0.05, threshold=2e-2,
                           n=64, npml=8, seed=21, check_seed=9, scale=1.0):
    """Surrogate outputs for the scenes, a seeded share spot-checked with
    the FDTD; a batch whose spot checks exceed the threshold (median
    relative L2 over their probes) is replaced by FDTD labels.  Returns
    (outputs, seconds {wall, network, fdtd}, fallback batches, spot-check
    errors).  The network predicts labels divided by `scale`."""
    rng = np.random.default_rng(check_seed)
    out = np.zeros((len(indices), N_PROBES, 2 * N_FREQS), dtype=np.float32)
    t0 = time.perf_counter()
    fallbacks, errs, t_net, t_fdtd = 0, [], 0.0, 0.0
    for b0 in range(0, len(indices), batch):
        sl = slice(b0, min(b0 + batch, len(indices)))
        t1 = time.perf_counter()
        pred = net.predict(maps[sl], coords[sl]) * scale
        t_net += time.perf_counter() - t1
        k = max(1, int(round(share * pred.shape[0])))
        pick = rng.choice(pred.shape[0], size=k, replace=False)
        _, _, lab, tf = label_scenes(lib, [indices[b0 + p] for p in pick], steps, n=n, npml=npml, seed=seed)
        t_fdtd += tf
        e = np.median(rel_l2(pred[pick], lab), axis=1)
        errs.extend(e.tolist())
        if np.max(e) > threshold:
            fallbacks += 1
            _, _, pred, tf = label_scenes(lib, list(indices[sl]), steps, n=n, npml=npml, seed=seed)
            t_fdtd += tf
        out[sl] = pred
    return out, {"wall": time.perf_counter() - t0, "network": t_net, "fdtd": t_fdtd}, fallbacks, errs
