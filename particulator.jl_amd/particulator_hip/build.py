"""Build libparticulator_hip.so in-tree with hipcc for gfx950.

Flags: -ffp-contract=off keeps the reference's unfused multiply/add order on the selection
and push paths; packed-FP32 instructions are disabled.  Machine LICM is disabled: in the
step kernel it hoists the 64-bit polynomial constants of log and sincos out of the
particle loop into vector registers, which then spill to scratch (256 VGPRs and 84 B of
spills per lane, reloaded on every collision); without it the kernel needs 187 VGPRs and no
scratch, and the step runs about 9% faster.  The max-ILP machine scheduler adds about 1%.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(ROOT, "csrc")
OUT = os.path.join(HERE, "libparticulator_hip.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
         "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-munsafe-fp-atomics", "-mllvm", "--disable-machine-licm",
         "-mllvm", "--amdgpu-sched-strategy=max-ilp", "-Wno-unused-command-line-argument"]


def sources():
    return [os.path.join(CSRC, "particulator.hip"), os.path.join(CSRC, "pt_tables.cpp")]


def build(verbose=False):
    deps = sources() + [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".h")]
    cmd = [HIPCC] + FLAGS + ["-o", OUT + ".tmp"] + sources()
    if verbose:
        print(" ".join(cmd))
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed (%d)" % r.returncode)
    os.replace(OUT + ".tmp", OUT)
    del deps
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
