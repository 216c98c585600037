"""The device hypot (fteik_amd/csrc/hypot_borges.h) equals the C
library's hypot, which the reference's HYPOT calls (localSolver.F90:46).

The header is compiled for the host with the same no-contraction flag as
the device build; IEEE double arithmetic with correctly rounded sqrt and
division gives the same bits on both."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_hypot_borges_equals_libm(tmp_path):
    exe = str(tmp_path / "hypot_check")
    subprocess.run(["g++", "-O2", "-ffp-contract=off", "-o", exe,
                    os.path.join(ROOT, "tests", "hypot_check.cpp"), "-lm"],
                   check=True)
    out = subprocess.run([exe], check=True, capture_output=True,
                         text=True).stdout.split()
    bad, n = int(out[0]), int(out[1])
    assert n == 3000000 and bad == 0
