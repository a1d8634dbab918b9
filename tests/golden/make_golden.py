"""Copies the reference's own output fixtures that pin the echo path into
this folder (run in the build container, where the reference is present).

example_000000.iq: the float32 interleaved I/Q output of the default preset
(50^3, 200 steps, 8-cell CPML, Ez Gaussian-derivative dipole at the
centre, probe 6 cells off it) through the reference's fdtd stage
(cli/fdtd_stage.c:10-29 writes I = (float) td[n], Q = 0; cli/main.c
writes the samples raw).  It is data: the expected output of
fdtd_simulate_echo at float32 rounding.
"""
import os
import shutil

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = "/root/reference/output_phase1_check/example_000000.iq"

if __name__ == "__main__":
    shutil.copyfile(SRC, os.path.join(HERE, "example_000000.iq"))
    print("copied", SRC)
