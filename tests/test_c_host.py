"""A plain C99 host of the ABI (tests/c_host/swarm.c): it compiles against
include/particulator_hip.h alone, and with no GPU it builds the air electron table through
the host entry points (pt_sb_build, pt_table_build, pt_process_desc) and stops at
pt_create with PT_E_NOT_BUILT and its message in pt_last_error(NULL)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "particulator.jl_amd", "particulator_hip")
SB_DIR = os.path.join(ROOT, "particulator.jl_amd", "data", "brem_SB")


def build_c_host(outdir):
    if shutil.which("gcc") is None:
        pytest.skip("no gcc")
    if not os.path.exists(os.path.join(LIBDIR, "libparticulator_hip.so")):
        pytest.skip("library not built")
    exe = os.path.join(str(outdir), "swarm")
    cmd = ["gcc", "-std=c99", "-O2", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"),
           "-o", exe, os.path.join(ROOT, "tests", "c_host", "swarm.c"), "-L", LIBDIR, "-lparticulator_hip",
           "-Wl,-rpath," + LIBDIR, "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def test_c_host_builds_tables_without_a_gpu(tmp_path):
    exe = build_c_host(tmp_path)
    env = dict(os.environ, HIP_VISIBLE_DEVICES="")  # hide any device: only the host code runs
    r = subprocess.run([exe, SB_DIR, "100", "1", "1"], capture_output=True, text=True, env=env, timeout=120)
    assert r.returncode == 1, (r.stdout, r.stderr)
    assert "pt_create" in r.stderr and "-9 no HIP device available" in r.stderr, r.stderr
