"""Python driver for the C API: seeded scene builders shared by bench.py,
the tests and smoke().

Every builder takes a bound library (product, oracle or reference build)
and drives it only through the C API of include/*.h, so the same scene can
be run on each and compared.
"""
import ctypes as C
import time
import math

import numpy as np

from . import capi

EPS0 = 8.8541878128e-12
MU0 = 1.25663706212e-6
C0 = 299792458.0


# --------------------------------------------------------------- PCG32
class Pcg32:
    """engine/core/rng.c:3-23 (PCG-XSH-RR 64/32)."""

    MASK = (1 << 64) - 1

    def __init__(self, seed, stream=0):
        self.state = 0
        self.inc = ((stream << 1) | 1) & self.MASK
        self.next_u32()
        self.state = (self.state + seed) & self.MASK
        self.next_u32()

    def next_u32(self):
        old = self.state
        self.state = (old * 6364136223846793005 + self.inc) & self.MASK
        xs = (((old >> 18) ^ old) >> 27) & 0xFFFFFFFF
        rot = old >> 59
        return ((xs >> rot) | (xs << ((-rot) & 31))) & 0xFFFFFFFF

    def uniform(self):
        """rng_uniform: (r >> 8) / 2^24 as float32."""
        return float(np.float32((self.next_u32() >> 8) / float(1 << 24)))


# --------------------------------------------------------------- helpers
def field_arrays(g):
    """numpy views of the six host arrays of an fdtd_grid_t pointer."""
    gg = g.contents
    n = (gg.nx + 1) * (gg.ny + 1) * (gg.nz + 1)
    return {name: np.ctypeslib.as_array(getattr(gg, name), shape=(n,))
            for name in ("ex", "ey", "ez", "hx", "hy", "hz")}


def pattern_fill(g, scale=1.0e-4):
    """tests/periodic_wrapping_unit.c:23-43 fill pattern on every padded node."""
    gg = g.contents
    nx, ny, nz = gg.nx, gg.ny, gg.nz
    i, j, k = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    f = field_arrays(g)
    for c, name in enumerate(("ex", "ey", "ez", "hx", "hy", "hz")):
        v = (31 * i + 17 * j + 7 * k + 5 * c) % 29
        f[name][:] = ((v.astype(np.float64) - 14.0) * scale).ravel()


def random_fill(g, seed, amp=1.0e-3):
    rng = np.random.default_rng(seed)
    for name, a in field_arrays(g).items():
        a[:] = rng.uniform(-amp, amp, a.shape)


def same_values(a, b):
    """Bitwise equality treating +0 and -0 as equal."""
    a = np.asarray(a)
    b = np.asarray(b)
    return a.shape == b.shape and bool(np.all((a == b) | (np.isnan(a) & np.isnan(b))))


def c1_cpml_params(npml=10):
    """CPML of the 64^3 FP64 scene: m=3, ma=1, kappa_max=1, alpha_max=0.05, R0=1e-8."""
    return capi.CpmlParams(npml, 3.0, 1.0, 1.0, 0.05, 1e-8)


def modgauss_params(f0, bw):
    """Modulated Gaussian with -20 dB full bandwidth bw: tau = 2 sqrt(ln 10)/(pi bw),
    sigma = tau/sqrt 2, delay t0 = 4 sigma."""
    tau = 2.0 * math.sqrt(math.log(10.0)) / (math.pi * bw)
    sigma = tau / math.sqrt(2.0)
    return capi.ModGaussParams(f0, 4.0 * sigma, tau)


def c1_probes(n=64, src=(32, 32, 32), count=8, seed=1, lo=10, hi=54, min_dist=15.0):
    """Seeded Ez probe nodes in [lo,hi)^3 at least min_dist cells from the source.
    Raises ValueError when the box holds no such node (rejection sampling
    would never end)."""
    far = max(math.dist((x, y, z), src) for x in (lo, hi - 1) for y in (lo, hi - 1) for z in (lo, hi - 1))
    if hi <= lo or far < min_dist:
        raise ValueError(f"no probe node in [{lo},{hi})^3 is {min_dist} cells from {src}")
    rng = Pcg32(seed, 0)
    out = []
    while len(out) < count:
        p = tuple(lo + rng.next_u32() % (hi - lo) for _ in range(3))
        if math.dist(p, src) >= min_dist and p not in out:
            out.append(p)
    return out


def c1_freqs(nf=64, f_lo=20e9, f_hi=40e9):
    return np.array([f_lo + (f_hi - f_lo) * q / (nf - 1) for q in range(nf)], dtype=np.float64)


class NativeRun:
    """Holds a native fdtd_sim built on one library and the objects it needs."""

    def __init__(self, lib, n, dx=1e-3, courant=0.99, periodic=0, cpml=None,
                 precision=capi.FDTD_FP64, compensated=0, block=0, material_setup=None,
                 init_fields=None):
        self.lib = lib
        nx, ny, nz = n
        self.g = lib.fdtd_grid_create(nx, ny, nz, dx, courant)
        assert self.g
        if periodic:
            assert lib.fdtd_grid_set_periodic(self.g, periodic) == 0
        self.pml = lib.cpml_create(self.g, cpml) if cpml is not None else None
        if cpml is not None:
            assert self.pml
        self.mat = lib.materials_create(self.g)
        assert self.mat
        if material_setup:
            material_setup(lib, self.g, self.mat)
        if init_fields:
            init_fields(self.g)
        opt = capi.SimOptions(precision, compensated, block)
        self.sim = lib.fdtd_sim_create(self.g, self.pml, self.mat, C.byref(opt))
        assert self.sim, lib.fdtd_last_error()
        self._keep = []
        self.probe_nf = {}

    def add_point_source(self, comp, pos, wf_name, params, amp=1.0, hard=False):
        self._keep.append(params)
        wf = capi.waveform_ptr(self.lib, wf_name)
        src = self.lib.source_make_point(pos[0], pos[1], pos[2], comp,
                                         capi.SRC_HARD if hard else capi.SRC_SOFT, amp, wf,
                                         C.cast(C.pointer(params), C.c_void_p))
        assert self.lib.fdtd_sim_add_point_source(self.sim, C.byref(src)) == 0

    def add_plane_source(self, comp, axis, index, wf_name, params, amp=1.0, after_e=0, hard=False):
        self._keep.append(params)
        wf = capi.waveform_ptr(self.lib, wf_name)
        mode = capi.SRC_HARD if hard else capi.SRC_SOFT
        assert self.lib.fdtd_sim_add_plane_source(self.sim, comp, axis, index, mode, after_e, amp,
                                                  wf, C.cast(C.pointer(params), C.c_void_p)) == 0

    def add_probe(self, comp, pos, td_cap, freqs=None):
        nf = 0 if freqs is None else len(freqs)
        fp = None
        if nf:
            arr = np.ascontiguousarray(freqs, dtype=np.float64)
            self._keep.append(arr)
            fp = arr.ctypes.data_as(C.POINTER(C.c_double))
        pid = self.lib.fdtd_sim_add_probe(self.sim, comp, pos[0], pos[1], pos[2], td_cap, fp, nf)
        assert pid >= 0
        self.probe_nf[pid] = nf
        return pid

    def add_plane_monitor(self, axis, index, freqs):
        arr = np.ascontiguousarray(freqs, dtype=np.float64)
        self._keep.append(arr)
        mid = self.lib.fdtd_sim_add_plane_monitor(self.sim, axis, index,
                                                  arr.ctypes.data_as(C.POINTER(C.c_double)), len(arr))
        assert mid >= 0
        return mid

    def run(self, steps):
        assert self.lib.fdtd_sim_run(self.sim, steps) == 0, self.lib.fdtd_last_error()

    def td(self, pid, n):
        out = np.zeros(n)
        got = self.lib.fdtd_sim_probe_td(self.sim, pid, out.ctypes.data_as(C.POINTER(C.c_double)), n)
        assert got == n
        return out

    def dft(self, pid):
        nf = self.probe_nf[pid]
        re = np.zeros(nf)
        im = np.zeros(nf)
        assert self.lib.fdtd_sim_probe_dft(self.sim, pid, re.ctypes.data_as(C.POINTER(C.c_double)),
                                           im.ctypes.data_as(C.POINTER(C.c_double))) == 0
        return re + 1j * im

    def plane_dft(self, mid, which, shape):
        re = np.zeros(shape)
        im = np.zeros(shape)
        assert self.lib.fdtd_sim_plane_dft(self.sim, mid, which, re.ctypes.data_as(C.POINTER(C.c_double)),
                                           im.ctypes.data_as(C.POINTER(C.c_double))) == 0
        return re + 1j * im

    def fields(self):
        assert self.lib.fdtd_sim_download(self.sim, self.g) == 0
        return {k: v.copy() for k, v in field_arrays(self.g).items()}

    def set_profiling(self, on):
        assert self.lib.fdtd_sim_set_profiling(self.sim, 1 if on else 0) == 0

    def kernel_times(self):
        """(total H ms, total E ms, timed steps) since the previous call."""
        h = C.c_double()
        e = C.c_double()
        n = C.c_long()
        assert self.lib.fdtd_sim_kernel_times(self.sim, C.byref(h), C.byref(e), C.byref(n)) == 0
        return h.value, e.value, n.value

    def kernel_times3(self):
        """(ms per event slot: H / fused, E / ADE pass, shell; timed steps)
        since the previous call."""
        ms = (C.c_double * 3)()
        n = C.c_long()
        assert self.lib.fdtd_sim_kernel_times3(self.sim, ms, C.byref(n)) == 0
        return list(ms), n.value

    def energy(self):
        e = C.c_double()
        h = C.c_double()
        assert self.lib.fdtd_sim_energy(self.sim, C.byref(e), C.byref(h)) == 0
        return e.value, h.value

    def energy_audit(self):
        """(the five audit sums of fdtd_sim_energy_audit, audited step)."""
        out = (C.c_double * 5)()
        step = C.c_size_t()
        assert self.lib.fdtd_sim_energy_audit(self.sim, out, C.byref(step)) == 0, self.lib.fdtd_last_error()
        return list(out), step.value

    def add_flux_plane(self, axis, index):
        fid = self.lib.fdtd_sim_add_flux_plane(self.sim, axis, index)
        assert fid >= 0
        return fid

    def flux(self, fid):
        w = C.c_double()
        assert self.lib.fdtd_sim_flux(self.sim, fid, C.byref(w)) == 0
        return w.value

    def close(self):
        self.lib.fdtd_sim_destroy(self.sim)
        self.lib.materials_destroy(self.mat)
        if self.pml:
            self.lib.cpml_destroy(self.pml)
        self.lib.fdtd_grid_destroy(self.g)


def run_c1(lib, n=64, steps=2000, precision=capi.FDTD_FP64, compensated=0, n_probes=8, block=0):
    """64^3 vacuum, 10-cell CPML, Ez modulated-Gaussian dipole at the centre,
    8 seeded Ez probes with 64-frequency DFT (20-40 GHz)."""
    c = n // 2
    run = NativeRun(lib, (n, n, n), 1e-3, 0.99, cpml=c1_cpml_params(10), precision=precision,
                    compensated=compensated, block=block)
    run.add_point_source(capi.FIELD_EZ, (c, c, c), "source_waveform_modulated_gaussian",
                         modgauss_params(30e9, 20e9))
    freqs = c1_freqs()
    probes = c1_probes(n, (c, c, c), n_probes, lo=10, hi=n - 10)
    pids = [run.add_probe(capi.FIELD_EZ, p, steps, freqs) for p in probes]
    run.run(steps)
    td = np.stack([run.td(p, steps) for p in pids])
    dft = np.stack([run.dft(p) for p in pids])
    run.close()
    return td, dft


# ------------------------------------------------------------ C3 scene
def c3_layers(seed=3, nz=256, npml=16):
    """Seeded layer stack of the periodic dispersive scene along z (cell
    indices): Drude grating, gap, 2-pole Lorentz layer, gap, eps=4 substrate."""
    rng = Pcg32(seed, 0)
    z = 64
    t_g = 8 + rng.next_u32() % 17
    gap1 = 4 + rng.next_u32() % 9
    t_l = 16 + rng.next_u32() % 33
    gap2 = 4 + rng.next_u32() % 9
    t_s = 24 + rng.next_u32() % 41
    duty = 0.3 + 0.4 * rng.uniform()
    grating = (z, z + t_g)
    lorentz = (grating[1] + gap1, grating[1] + gap1 + t_l)
    sub = (lorentz[1] + gap2, lorentz[1] + gap2 + t_s)
    assert sub[1] <= nz - npml - 8
    return {"grating": grating, "lorentz": lorentz, "substrate": sub, "duty": duty, "period": 32}


def c3_materials(lib, g, mat, layers):
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
    gg = g.contents
    nx, ny = gg.nx, gg.ny
    per = layers["period"]
    width = max(1, int(round(layers["duty"] * per)))
    k0, k1 = layers["grating"]
    for i0 in range(0, nx, per):
        assert lib.materials_assign_box(mat, g, i0, min(i0 + width, nx), 0, ny, k0, k1, did.value) == 0
    k0, k1 = layers["lorentz"]
    assert lib.materials_assign_box(mat, g, 0, nx, 0, ny, k0, k1, lid.value) == 0
    k0, k1 = layers["substrate"]
    assert lib.materials_set_box(mat, g, 0, nx, 0, ny, k0, k1, 4.0, 0.0) == 0


def c3_probes(n, layers, count=8, seed=5):
    """Seeded Ex probes in the vacuum gaps and above/below the stack."""
    rng = Pcg32(seed, 0)
    zs = [layers["grating"][0] - 12, layers["grating"][1] + 1, layers["substrate"][1] + 6,
          layers["lorentz"][0] - 2]
    out = []
    for q in range(count):
        out.append((rng.next_u32() % n[0], rng.next_u32() % n[1], zs[q % len(zs)]))
    return out


def build_c3(lib, precision, td_cap, n=(512, 512, 256), npml=16, seed=3, monitors=True, block=0):
    """512x512x256, dx = 0.5 mm, periodic x/y, 16-cell CPML on +-z, Ex soft
    plane source at k = npml + 8 (40 GHz, 30 GHz -20 dB bandwidth), Drude
    grating / 2-pole Lorentz / eps=4 stack, two full-plane monitors at 128
    frequencies over 25-55 GHz, plus 8 Ex point probes with DFTs."""
    layers = c3_layers(seed, n[2], npml)
    run = NativeRun(lib, n, 0.5e-3, 0.99, periodic=3, cpml=capi.CpmlParams(npml, 3.0, 1.0, 1.0, 0.05, 1e-8),
                    precision=precision, compensated=1, block=block,
                    material_setup=lambda lb, g, m: c3_materials(lb, g, m, layers))
    run.add_plane_source(capi.FIELD_EX, 2, npml + 8, "source_waveform_modulated_gaussian",
                         modgauss_params(40e9, 30e9))
    freqs = np.linspace(25e9, 55e9, 128)
    mons = []
    if monitors:
        mons = [run.add_plane_monitor(2, npml + 4, freqs), run.add_plane_monitor(2, n[2] - npml - 4, freqs)]
    pids = [run.add_probe(capi.FIELD_EX, p, td_cap, freqs[::2]) for p in c3_probes(n, layers)]
    return run, pids, mons, layers


# ------------------------------------------------------------ C2 scene
def c2_shapes(n=256, npml=12, n_spheres=40, r_lo=4.0, r_hi=24.0, seed=2, slab_half=10.0):
    """Seeded heterogeneous dielectric: a slab of eps_r = 6 through the grid
    centre tilted 23 degrees about x (normal (0, -sin 23, cos 23)), then
    n_spheres spheres (radius U[r_lo, r_hi] cells, eps_r U[1, 12], centre
    uniform in the interior at least r + 2 cells from the CPML)."""
    rng = Pcg32(seed, 0)
    th = math.radians(23.0)
    shapes = []
    s = capi.Shape()
    s.kind = capi.SHAPE_SLAB
    s.eps_r = 6.0
    s.p[:] = (n / 2.0, n / 2.0, n / 2.0)
    s.normal[:] = (0.0, -math.sin(th), math.cos(th))
    s.half_thickness = slab_half
    s.lo[:] = (0.0, 0.0, 0.0)
    s.hi[:] = (float(n), float(n), float(n))
    shapes.append(s)
    for _ in range(n_spheres):
        r = r_lo + (r_hi - r_lo) * rng.uniform()
        eps = 1.0 + 11.0 * rng.uniform()
        lo, hi = npml + 2 + r, n - npml - 2 - r
        c = [lo + (hi - lo) * rng.uniform() for _ in range(3)]
        s = capi.Shape()
        s.kind = capi.SHAPE_SPHERE
        s.eps_r = eps
        s.p[:] = c
        s.radius = r
        shapes.append(s)
    return shapes


def c2_paint(lib, g, mat, shapes, npml, stats_out=None):
    arr = (capi.Shape * len(shapes))(*shapes)
    opt = capi.SceneOptions(4, npml + 2)
    st = capi.SceneStats()
    t0 = time.perf_counter()
    assert lib.fdtd_scene_paint(mat, g, arr, len(shapes), C.byref(opt), C.byref(st)) == 0
    if stats_out is not None:
        stats_out.update(nodes=st.nodes_painted, overrides=st.override_edges, tensors=st.tensor_edges,
                         scalar_mixed=st.scalar_mixed_edges, paint_s=round(time.perf_counter() - t0, 4))


def build_c2(lib, precision, td_cap, n=256, npml=12, n_spheres=40, r_lo=4.0, r_hi=24.0, n_probes=16,
             seed=2, stats_out=None):
    """256^3, dx = 1 mm, 12-cell CPML on all faces (C1 grading), the seeded
    sphere + tilted-slab scene with 3x3 subpixel tensors on interface
    edges, Ez modulated-Gaussian dipole at the centre, 16 seeded Ez probes
    with 64-frequency DFTs over 20-40 GHz."""
    shapes = c2_shapes(n, npml, n_spheres, r_lo, r_hi, seed)
    run = NativeRun(lib, (n, n, n), 1e-3, 0.99, cpml=c1_cpml_params(npml), precision=precision,
                    compensated=1,
                    material_setup=lambda lb, g, m: c2_paint(lb, g, m, shapes, npml, stats_out))
    c = n // 2
    run.add_point_source(capi.FIELD_EZ, (c, c, c), "source_waveform_modulated_gaussian",
                         modgauss_params(30e9, 20e9))
    freqs = c1_freqs()
    probes = c1_probes(n, (c, c, c), n_probes, seed=seed, lo=npml + 3, hi=n - npml - 3)
    pids = [run.add_probe(capi.FIELD_EZ, p, td_cap, freqs) for p in probes]
    return run, pids


def c4_scene(index, seed=4, npml=10, n_lo=48, n_hi=192):
    """Scene `index` of the seeded batch: sizes U{n_lo..n_hi} per axis;
    1-12 lossy boxes (eps U[1,10], sigma U[0,0.05] S/m) inside the CPML;
    one dipole on a random axis at a random interior node; 4 probes on the
    dipole's component at least 8 cells from it.  PCG32 stream = index."""
    rng = Pcg32(seed, index)
    n = tuple(n_lo + rng.next_u32() % (n_hi - n_lo + 1) for _ in range(3))
    lo_i = npml + 2
    boxes = []
    for _ in range(1 + rng.next_u32() % 12):
        b = []
        for a in range(3):
            span = n[a] - 2 * lo_i
            size = 4 + rng.next_u32() % max(1, span // 3)
            start = lo_i + rng.next_u32() % max(1, span - size)
            b.append((start, start + size))
        eps = 1.0 + 9.0 * rng.uniform()
        sigma = 0.05 * rng.uniform()
        boxes.append((b, eps, sigma))
    comp = rng.next_u32() % 3
    src = tuple(lo_i + rng.next_u32() % (n[a] - 2 * lo_i) for a in range(3))
    probes = []
    while len(probes) < 4:
        p = tuple(lo_i + rng.next_u32() % (n[a] - 2 * lo_i) for a in range(3))
        if sum((p[a] - src[a]) ** 2 for a in range(3)) >= 64:
            probes.append(p)
    return {"n": n, "boxes": boxes, "comp": comp, "src": src, "probes": probes}


def build_c4_scene(lib, spec, precision, td_cap, npml=10):
    """A handle for one c4 scene: dx = 1 mm, S = 0.99, CPML npml cells on
    every face (C1 grading), modulated-Gaussian dipole (30 GHz, 20 GHz -20 dB
    bandwidth), 4 probes with 64-frequency DFTs over 20-40 GHz."""
    def mats(lb, g, mat):
        for b, eps, sigma in spec["boxes"]:
            assert lb.materials_set_box(mat, g, b[0][0], b[0][1], b[1][0], b[1][1], b[2][0], b[2][1],
                                        eps, sigma) == 0
    run = NativeRun(lib, spec["n"], 1e-3, 0.99, cpml=c1_cpml_params(npml), precision=precision,
                    compensated=1 if precision == capi.FDTD_FP32 else 0, material_setup=mats)
    run.add_point_source(spec["comp"], spec["src"], "source_waveform_modulated_gaussian",
                         modgauss_params(30e9, 20e9))
    freqs = c1_freqs()
    pids = [run.add_probe(spec["comp"], p, td_cap, freqs) for p in spec["probes"]]
    return run, pids


def run_many(lib, runs, steps):
    """fdtd_sim_run_many over NativeRun handles."""
    arr = (C.c_void_p * len(runs))(*[r.sim for r in runs])
    assert lib.fdtd_sim_run_many(arr, len(runs), steps) == 0, lib.fdtd_last_error()


class Batch:
    """fdtd_batch_t over NativeRun handles (packed launches); close it
    before the handles."""

    def __init__(self, lib, runs):
        self.lib = lib
        self.n = len(runs)
        arr = (C.c_void_p * len(runs))(*[r.sim for r in runs])
        self.b = lib.fdtd_batch_create(arr, len(runs))
        assert self.b, lib.fdtd_last_error()

    def run(self, steps):
        assert self.lib.fdtd_batch_run(self.b, steps) == 0, self.lib.fdtd_last_error()

    def launches(self):
        """(packed leapfrog, shell, split ADE pass) launches per step."""
        lf, sh, ad = C.c_int(), C.c_int(), C.c_int()
        assert self.lib.fdtd_batch_launches(self.b, C.byref(lf), C.byref(sh), C.byref(ad)) == 0
        return lf.value, sh.value, ad.value

    def finite(self):
        """Per-scene failure flags: 1 where every field value is finite."""
        out = (C.c_int * self.n)()
        assert self.lib.fdtd_batch_fields_finite(self.b, out) == 0, self.lib.fdtd_last_error()
        return list(out)

    def close(self):
        if self.b:
            self.lib.fdtd_batch_destroy(self.b)
            self.b = None


class Surrogate:
    """ctypes wrapper of an fdtd_sur_t (include/fdtd_surrogate.h)."""

    def __init__(self, lib, grid, in_channels, channels, coords, width, outputs, queries, max_batch, seed=1):
        self.lib = lib
        self.cfg = capi.SurConfig(grid, in_channels, (C.c_int * 4)(*channels), coords, width, outputs, queries,
                                  max_batch)
        self.h = lib.fdtd_sur_create(C.byref(self.cfg), seed)
        assert self.h, lib.fdtd_last_error()
        self.n_params = lib.fdtd_sur_n_params(self.h)

    @staticmethod
    def _f(a):
        a = np.ascontiguousarray(a, dtype=np.float32)
        return a, a.ctypes.data_as(C.POINTER(C.c_float))

    def params(self):
        out = np.zeros(self.n_params, dtype=np.float32)
        assert self.lib.fdtd_sur_get_params(self.h, out.ctypes.data_as(C.POINTER(C.c_float))) == 0
        return out

    def set_params(self, p):
        a, pa = self._f(p)
        assert a.size == self.n_params
        assert self.lib.fdtd_sur_set_params(self.h, pa) == 0

    def predict(self, maps, coords):
        n = maps.shape[0]
        m, pm = self._f(maps)
        c, pc = self._f(coords)
        out = np.zeros((n, self.cfg.queries, self.cfg.outputs), dtype=np.float32)
        assert self.lib.fdtd_sur_predict(self.h, pm, pc, n, out.ctypes.data_as(C.POINTER(C.c_float))) == 0
        return out

    def gradient(self, maps, coords, labels):
        n = maps.shape[0]
        m, pm = self._f(maps)
        c, pc = self._f(coords)
        lb, pl = self._f(labels)
        g = np.zeros(self.n_params, dtype=np.float32)
        loss = C.c_double()
        assert self.lib.fdtd_sur_gradient(self.h, pm, pc, pl, n, g.ctypes.data_as(C.POINTER(C.c_float)),
                                          C.byref(loss)) == 0
        return loss.value, g

    def train_step(self, maps, coords, labels, lr):
        n = maps.shape[0]
        m, pm = self._f(maps)
        c, pc = self._f(coords)
        lb, pl = self._f(labels)
        loss = C.c_double()
        assert self.lib.fdtd_sur_train_step(self.h, pm, pc, pl, n, lr, C.byref(loss)) == 0
        return loss.value

    def close(self):
        self.lib.fdtd_sur_destroy(self.h)
