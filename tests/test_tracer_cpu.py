"""The GPU streamline tracer, built for the host, against the oracle's FP64 advector.

msceer_amd/csrc/tracer.hpp holds the device code that pass 3 runs in FP64
(one iteration of the adaptive Euler loop per call, voxel entries folded into
the step that leaves a voxel).  tests/tracer_harness.cpp compiles it with g++
next to oracle/integ_oracle.cpp (the reference's voxel-by-voxel loop restated
in FP64, gi_adaptive_euler_advector.h:51-305) and traces 331 776 streamlines
on a Gaussian-mixture field with random certain/assigned vertices, both stop
rules, step budgets 7, 60 and 4096 and both flow directions (sign -1 down,
+1 up): every path, end vertex and event must be identical.  Neither side fuses products and sums (-ffp-contract=off here,
`#pragma clang fp contract(off)` in the HIP build).
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_tracer_matches_oracle_fp64_streamlines(tmp_path):
    exe = str(tmp_path / "tracer_harness")
    cmd = ["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-w", "-fopenmp",
           os.path.join(ROOT, "tests", "tracer_harness.cpp"), os.path.join(ROOT, "oracle", "lstar_oracle.cpp"),
           "-o", exe]
    subprocess.run(cmd, check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    words = r.stdout.split()
    assert words[words.index("streamlines") + 1] == "331776"
    assert words[words.index("differing") + 1] == "0"
