"""Times the CPU oracle (the port, liboracle_fdtd.so) and the rfsyn
reference compiled from its own sources (_ref/libref_fdtd.so) on the same
C1 scene, one pinned core each, interleaved, so the port's CPU baseline can
be read as the reference's speed: 64^3 vacuum, 10-cell CPML, Ez
modulated-Gaussian dipole, 8 Ez probes x 64-frequency DFTs, stepped through
the low-level calls the reference's validators use (H CPML, source at
t = n dt, E CPML + materials, probe DFTs at (n+1) dt).

usage: python oracle/time_port_vs_reference.py [steps] [repeats]
Prints a markdown table.  Test infrastructure: needs the reference built
(oracle/Makefile target ref)."""
import ctypes as C
import os
import platform
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

from rfsyn_amd import capi  # noqa: E402
import bench  # noqa: E402


def main():
    steps = int(sys.argv[1]) if len(sys.argv) > 1 else 300
    repeats = int(sys.argv[2]) if len(sys.argv) > 2 else 3
    libs = {"port (oracle/liboracle_fdtd.so)": os.path.join(HERE, "liboracle_fdtd.so"),
            "reference (oracle/_ref/libref_fdtd.so)": os.path.join(HERE, "_ref", "libref_fdtd.so")}
    bound = {k: capi.bind(C.CDLL(v)) for k, v in libs.items()}
    cpu = min(os.sched_getaffinity(0))
    os.sched_setaffinity(0, {cpu})
    wl = bench.C1()
    times = {k: [] for k in libs}
    for _ in range(repeats):
        for k, lib in bound.items():
            t, cells, _ = wl.cpu_sample(lib, steps)
            times[k].append(t)
    print(f"C1 scene (64^3, 10-cell CPML, 8 probes x 64 frequencies), {steps} steps, "
          f"1 pinned core of {os.cpu_count()} ({bench.cpu_model() or platform.processor()}), "
          f"{repeats} interleaved repeats\n")
    print("| library | best s | median s | Mcell-updates/s (best) |")
    print("|---|---|---|---|")
    for k, v in times.items():
        v = sorted(v)
        print(f"| {k} | {v[0]:.3f} | {v[len(v) // 2]:.3f} | {64 ** 3 * steps / v[0] / 1e6:.1f} |")
    a, b = (min(v) for v in times.values())
    print(f"\nport / reference time: {a / b:.3f}")


if __name__ == "__main__":
    main()
