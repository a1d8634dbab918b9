"""2D subpixel helpers (include/subpixel_te.h): known answers of the
reference's tests/subpixel_te_unit.c, closed forms, and bitwise agreement
with the reference compiled from its sources (oracle/_ref) on seeded
random inputs."""
import ctypes as C
import math
import os

import numpy as np
import pytest

import rfsyn_amd

REF_SO = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref",
                      "libref_fdtd.so")
OK, INVALID, NUMERIC = 0, -1, -2
DBL_MAX = np.finfo(np.float64).max


class Rect(C.Structure):
    _fields_ = [("center_x", C.c_double), ("center_y", C.c_double), ("size_u", C.c_double),
                ("size_v", C.c_double), ("angle_rad", C.c_double)]


def _bind(path):
    lib = C.CDLL(path)
    d, pd, pr = C.c_double, C.POINTER(C.c_double), C.POINTER(Rect)
    sig = {
        "subpixel_te_rectangle_fraction": [pr, d, d, d, d, pd],
        "subpixel_te_rectangle_edge_fraction": [pr, d, d, C.c_int, d, pd],
        "subpixel_te_mix_epsilon_harmonic": [d, d, d, pd],
        "subpixel_te_mix_epsilon": [d, d, d, pd],
        "subpixel_te_rectangle_epsilon": [pr, d, d, d, d, d, d, pd, pd],
        "subpixel_te_mix_epsilon_tensor": [d, d, d, d, pd, pd, pd],
    }
    for name, args in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = C.c_int
    return lib


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(rfsyn_amd.LIB_PATH):
        pytest.skip("library not built")
    return _bind(rfsyn_amd.LIB_PATH)


@pytest.fixture(scope="module")
def ref():
    if not os.path.exists(REF_SO):
        pytest.skip("reference library not built here")
    return _bind(REF_SO)


def frac(lib, r, x, y, dx, dy):
    out = C.c_double(math.nan)
    assert lib.subpixel_te_rectangle_fraction(C.byref(Rect(*r)), x, y, dx, dy, C.byref(out)) == OK
    return out.value


def edge(lib, r, x, y, axis, length):
    out = C.c_double(math.nan)
    st = lib.subpixel_te_rectangle_edge_fraction(C.byref(Rect(*r)), x, y, axis, length, C.byref(out))
    return st, out.value


def mix(fn, *args):
    out = C.c_double(-1.0)
    st = fn(*args, C.byref(out))
    return st, out.value


def test_basic_geometry(lib):
    """subpixel_te_unit.c test_basic_geometry."""
    assert abs(frac(lib, (0, 0, 4, 4, 0), 0, 0, 1, 1) - 1) <= 1e-14
    assert abs(frac(lib, (0, 0, 4, 4, 0), 10, 0, 1, 1)) <= 1e-14
    assert abs(frac(lib, (1, 0, 2, 4, 0), 0, 0, 1, 1) - 0.5) <= 1e-14
    assert abs(frac(lib, (1, 1, 2, 2, 0), 0, 0, 1, 1) - 0.25) <= 1e-14
    assert abs(frac(lib, (0, 0, 0.5, 0.25, 0), 0, 0, 1, 1) - 0.125) <= 1e-14
    assert abs(frac(lib, (0, 0, 0.5, 0.25, math.pi / 6), 0, 0, 1, 1) - 0.125) <= 1e-14
    assert abs(frac(lib, (0, 0, 1, 1, 0), 0, 0, 2, 1) - 0.5) <= 1e-14
    assert abs(frac(lib, (1, 0, 1, 2, 0), 0, 0, 1, 1)) <= 1e-14
    assert abs(frac(lib, (0, 0, math.sqrt(0.5), math.sqrt(0.5), math.pi / 4), 0, 0, 1, 1) - 0.5) <= 2e-14
    assert abs(frac(lib, (1000001, -1000000, 2, 4, 0), 1000000, -1000000, 1, 1) - 0.5) <= 1e-14


def test_rotated_cuts(lib):
    for deg in (0, 15, 30, 45, 60, 90, 120):
        a = math.radians(deg)
        assert abs(frac(lib, (math.cos(a), math.sin(a), 2, 10, a), 0, 0, 1, 1) - 0.5) <= 1e-13
    for cut, want in zip((-1, -0.5, 0, 0.5, 1), (0, 0.125, 0.5, 0.875, 1)):
        cu = cut / math.sqrt(2) - 2
        r = (cu / math.sqrt(2), cu / math.sqrt(2), 4, 10, math.pi / 4)
        assert abs(frac(lib, r, 0, 0, 1, 1) - want) <= 1e-13


def test_translation_scale_and_period(lib):
    prev = 1.0
    for q in range(-4, 5):
        v = frac(lib, (1 + 0.1 * q, 0, 2, 4, 0), 0, 0, 1, 1)
        assert abs(v - (0.5 - 0.1 * q)) <= 1e-14 and v <= prev + 1e-14
        prev = v
    r = (0.2, -0.1, 0.7, 0.6, 0.43)
    base = frac(lib, r, 0, 0, 1, 1)
    for sc in (1e-3, 1e3):
        assert abs(frac(lib, (r[0] * sc, r[1] * sc, r[2] * sc, r[3] * sc, r[4]), 0, 0, sc, sc) - base) <= 1e-13
    assert abs(frac(lib, r[:4] + (r[4] + 2 * math.pi,), 0, 0, 1, 1) - base) <= 1e-13


def test_area_partition_and_subdivision(lib):
    """The cell fractions of a tiling sum to the rectangle's area, and a
    cell's fraction is the mean of its four quarter cells'."""
    for q in range(4):
        r = (0.13, -0.17, 0.8, 1.2, q * math.pi / 12)
        area = sum(frac(lib, r, -2 + (i + 0.5) * 0.25, -2 + (j + 0.5) * 0.25, 0.25, 0.25) * 0.0625
                   for i in range(16) for j in range(16))
        assert abs(area - 0.96) <= 2e-12
    rng = np.random.default_rng(7)
    for _ in range(300):
        r = (rng.uniform(-0.5, 0.5), rng.uniform(-0.5, 0.5), 0.01 + 1.5 * rng.uniform(), 0.01 + 1.5 * rng.uniform(),
             2 * math.pi * rng.uniform())
        coarse = frac(lib, r, 0, 0, 1, 1)
        fine = sum(0.25 * frac(lib, r, sx, sy, 0.5, 0.5) for sx in (-0.25, 0.25) for sy in (-0.25, 0.25))
        assert 0.0 <= coarse <= 1.0 and abs(fine - coarse) <= 1e-12


def test_axis_aligned_closed_form(lib):
    """Axis-aligned rectangles: the fraction is the product of the two
    interval overlaps divided by the cell area."""
    rng = np.random.default_rng(11)
    for _ in range(300):
        cx, cy = rng.uniform(-2, 2, 2)
        su, sv = rng.uniform(0.05, 3, 2)
        dx, dy = rng.uniform(0.2, 2, 2)
        ox = max(0.0, min(cx + su / 2, dx / 2) - max(cx - su / 2, -dx / 2))
        oy = max(0.0, min(cy + sv / 2, dy / 2) - max(cy - sv / 2, -dy / 2))
        assert abs(frac(lib, (cx, cy, su, sv, 0.0), 0, 0, dx, dy) - ox * oy / (dx * dy)) <= 1e-14


def test_mixing_rules(lib):
    assert mix(lib.subpixel_te_mix_epsilon, 0.0, 1.0, 4.0) == (OK, 1.0)
    assert mix(lib.subpixel_te_mix_epsilon, 1.0, 1.0, 4.0) == (OK, 4.0)
    assert abs(mix(lib.subpixel_te_mix_epsilon, 0.5, 1.0, 4.0)[1] - 2.5) <= 1e-14
    assert abs(mix(lib.subpixel_te_mix_epsilon, 0.25, 4.0, 1.0)[1] - 3.25) <= 1e-14
    assert abs(mix(lib.subpixel_te_mix_epsilon, 0.37, 7.0, 7.0)[1] - 7.0) <= 1e-14
    st, e = mix(lib.subpixel_te_mix_epsilon, 0.5, 1.0, DBL_MAX)
    assert st == OK and math.isfinite(e) and abs(e / DBL_MAX - 0.5) <= 1e-14
    assert mix(lib.subpixel_te_mix_epsilon_harmonic, 0.0, 1.0, 4.0) == (OK, 1.0)
    assert mix(lib.subpixel_te_mix_epsilon_harmonic, 1.0, 1.0, 4.0) == (OK, 4.0)
    assert abs(mix(lib.subpixel_te_mix_epsilon_harmonic, 0.5, 1.0, 4.0)[1] - 1.6) <= 1e-14
    assert mix(lib.subpixel_te_mix_epsilon_harmonic, 1.1, 1.0, 4.0)[0] == INVALID
    f, e = C.c_double(-1), C.c_double(-1)
    assert lib.subpixel_te_rectangle_epsilon(C.byref(Rect(1, 0, 2, 4, 0)), 0, 0, 1, 1, 1, 4, C.byref(f),
                                             C.byref(e)) == OK
    assert abs(f.value - 0.5) <= 1e-14 and abs(e.value - 2.5) <= 1e-14


def tensor(lib, f, eb, ei, a):
    xx, xy, yy = C.c_double(), C.c_double(), C.c_double()
    st = lib.subpixel_te_mix_epsilon_tensor(f, eb, ei, a, C.byref(xx), C.byref(xy), C.byref(yy))
    return st, xx.value, xy.value, yy.value


def test_tensor(lib):
    for a in (0.0, 0.77):
        st, xx, xy, yy = tensor(lib, 0.0, 1, 4, a)
        assert st == OK and abs(xx - 1) <= 1e-14 and abs(yy - 1) <= 1e-14 and xy == 0
        st, xx, xy, yy = tensor(lib, 1.0, 1, 4, a)
        assert st == OK and abs(xx - 0.25) <= 1e-14 and abs(yy - 0.25) <= 1e-14 and xy == 0
    _, xx, xy, yy = tensor(lib, 0.5, 1, 4, 0.0)
    assert abs(xx - 0.625) <= 1e-14 and abs(yy - 0.4) <= 1e-14 and xy == 0
    _, xx, xy, yy = tensor(lib, 0.5, 1, 4, math.pi / 2)
    assert abs(xx - 0.4) <= 1e-13 and abs(yy - 0.625) <= 1e-13 and abs(xy) <= 1e-13
    for q in range(1, 6):
        _, xx, xy, yy = tensor(lib, 0.5, 1, 4, 0.31 * q)
        # trace and determinant are rotation invariants: lam_n + lam_t, lam_n lam_t
        assert abs(xx + yy - 1.025) <= 1e-13 and abs(xx * yy - xy * xy - 0.25) <= 1e-13


def test_edge_fraction(lib):
    assert edge(lib, (0, 0, 2, 2, 0), -0.5, 0, 1, 1.0) == (OK, 1.0)
    assert edge(lib, (0, 0, 2, 2, 0), 0, -0.5, 2, 1.0) == (OK, 1.0)
    assert edge(lib, (0, 0, 2, 2, 0), 1.5, 0, 1, 1.0) == (OK, 0.0)
    # the rectangle spans x in [-1, 2]: the unit x edge from the origin lies
    # inside it (subpixel_te_unit.c:587-589 states 0.5; the reference
    # function returns 1, as here)
    assert edge(lib, (0.5, 0, 3, 1, 0), 0, 0, 1, 1.0) == (OK, 1.0)
    # axis-aligned closed form along x: overlap of [x, x+L] with [cx-su/2, cx+su/2]
    # when |y - cy| < sv/2
    rng = np.random.default_rng(3)
    for _ in range(200):
        cx, x = rng.uniform(-2, 2, 2)
        su, length = rng.uniform(0.1, 3, 2)
        want = max(0.0, min(x + length, cx + su / 2) - max(x, cx - su / 2)) / length
        st, v = edge(lib, (cx, 0.0, su, 1.0, 0.0), x, 0.2, 1, length)
        assert st == OK and abs(v - want) <= 1e-14


def test_errors_leave_outputs(lib):
    v = C.c_double(0.123)
    r = Rect(0, 0, 1, 1, 0)
    assert lib.subpixel_te_rectangle_fraction(None, 0, 0, 1, 1, C.byref(v)) == INVALID
    assert lib.subpixel_te_rectangle_fraction(C.byref(r), 0, 0, 1, 1, None) == INVALID
    assert lib.subpixel_te_rectangle_fraction(C.byref(r), math.nan, 0, 1, 1, C.byref(v)) == INVALID
    assert lib.subpixel_te_rectangle_fraction(C.byref(r), 0, 0, 0, 1, C.byref(v)) == INVALID
    assert lib.subpixel_te_rectangle_fraction(C.byref(r), 0, 0, 1, math.inf, C.byref(v)) == INVALID
    assert lib.subpixel_te_rectangle_fraction(C.byref(Rect(0, 0, -1, 1, 0)), 0, 0, 1, 1, C.byref(v)) == INVALID
    assert lib.subpixel_te_rectangle_fraction(C.byref(Rect(0, 0, 1, 1, math.nan)), 0, 0, 1, 1, C.byref(v)) == INVALID
    assert v.value == 0.123
    e = C.c_double(9.0)
    for args in ((-0.01, 1, 4), (1.01, 1, 4), (math.nan, 1, 4), (0.5, 0.5, 4), (0.5, 1, math.inf)):
        assert lib.subpixel_te_mix_epsilon(*args, C.byref(e)) == INVALID
    assert e.value == 9.0
    assert lib.subpixel_te_rectangle_epsilon(C.byref(r), 0, 0, 1, 1, 1, 4, C.byref(v), C.byref(v)) == INVALID
    assert lib.subpixel_te_rectangle_epsilon(C.byref(r), 0, 0, 1, 1, 1, 0, C.byref(v), C.byref(e)) == INVALID
    assert v.value == 0.123 and e.value == 9.0
    assert lib.subpixel_te_rectangle_fraction(C.byref(Rect(-DBL_MAX, 0, 1, 1, 0)), DBL_MAX, 0, 1, 1,
                                              C.byref(v)) == NUMERIC
    assert v.value == 0.123
    assert edge(lib, (0, 0, 1, 1, 0), 0, 0, 3, 1.0)[0] == INVALID
    assert edge(lib, (0, 0, 0.0, 1, 0), 0, 0, 1, 1.0)[0] == NUMERIC


def test_bitwise_against_reference(lib, ref):
    """Seeded random rectangles, cells, edges and mixes: the same status and
    the same bits as the reference compiled from its sources.  This also
    records the reference's edge parametrisation: for the rotated case of
    subpixel_te_unit.c:591-594 (rectangle (0, 0.5, 1, 3, pi/2), y edge from
    the origin) the reference returns 0, not the 0.5 its unit test states,
    and this library returns what the reference returns."""
    rng = np.random.default_rng(2024)
    assert edge(lib, (0, 0.5, 1, 3, math.pi / 2), 0, 0, 2, 1.0) == edge(ref, (0, 0.5, 1, 3, math.pi / 2), 0, 0, 2, 1.0)
    for _ in range(3000):
        r = (rng.uniform(-3, 3), rng.uniform(-3, 3), rng.uniform(0.01, 4), rng.uniform(0.01, 4),
             rng.uniform(-7, 7))
        x, y = rng.uniform(-3, 3, 2)
        dx, dy = rng.uniform(0.05, 2, 2)
        assert frac(lib, r, x, y, dx, dy) == frac(ref, r, x, y, dx, dy)
        ax = int(rng.integers(1, 3))
        length = rng.uniform(0.05, 2)
        assert edge(lib, r, x, y, ax, length) == edge(ref, r, x, y, ax, length)
        f = rng.uniform()
        eb, ei = rng.uniform(1, 12, 2)
        for fn in ("subpixel_te_mix_epsilon", "subpixel_te_mix_epsilon_harmonic"):
            assert mix(getattr(lib, fn), f, eb, ei) == mix(getattr(ref, fn), f, eb, ei)
        a = rng.uniform(-4, 4)
        assert tensor(lib, f, eb, ei, a) == tensor(ref, f, eb, ei, a)
