"""Build the in-tree native libraries (gfx950 HIP engine + C host driver).

    python -m traffic_assignment_amd.build

Outputs (git-ignored, shipped with the tree):
  traffic_assignment_amd/libtapgpu.so      HIP engine, C ABI include/tapgpu.h
  traffic_assignment_amd/libtap_hip.so     reference-API host driver
                                           (tap_frankwolfe / fw_assign /
                                           pape_cuda / TNTP parsers)
  traffic_assignment_amd/tap_frankwolfe_hip  CLI (tap_main.c's flags)
  traffic_assignment_amd/tap_variants_hip    CLI (tap_mpi_main.c's flags: network
                                           variants from a modification file,
                                           solved as one batch)
  traffic_assignment_amd/libtapsynth.so    host-threaded free-flow zone times
                                           for the workload generator
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INC = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("TAP_OFFLOAD_ARCH", "gfx950")

# No FMA contraction anywhere (bit parity of every + and *); packed FP32
# instructions are disabled as a matter of policy on this hardware pool.
# Loop headers are aligned to 16 bytes: the per-origin kernels' round loops
# then run at the same speed whatever code sits before them (measured: an
# unaligned build of the same loops was up to 1.6% slower).
HIP_FLAGS = ["-O3", "-std=c++17", f"--offload-arch={ARCH}", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-fPIC",
             "-falign-loops=16"]
C_FLAGS = ["-O3", "-ffp-contract=off", "-fPIC", "-Wall", "-Wextra"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build failed: " + " ".join(cmd))
    return r


def build_engine():
    out = os.path.join(HERE, "libtapgpu.so")
    _run([HIPCC, *HIP_FLAGS, "-shared", "-I", INC, "-o", out, os.path.join(CSRC, "tapgpu.hip")])
    return out


def build_host():
    """C host driver keeping the reference API, linked against libtapgpu."""
    srcs = [os.path.join(CSRC, f) for f in ("tap_host.c", "tntp_io.c")]
    if not all(os.path.exists(s) for s in srcs):
        return None
    lib = os.path.join(HERE, "libtap_hip.so")
    _run(["gcc", *C_FLAGS, "-shared", "-I", INC, "-o", lib, *srcs,
          "-L", HERE, "-ltapgpu", "-Wl,-rpath,$ORIGIN", "-lm", "-lpthread"])
    for src, name in (("tap_main_hip.c", "tap_frankwolfe_hip"),
                      ("tap_greedy_main.c", "tap_greedy_hip"),
                      ("tap_variants_main.c", "tap_variants_hip")):
        main = os.path.join(CSRC, src)
        if os.path.exists(main):
            exe = os.path.join(HERE, name)
            _run(["gcc", *C_FLAGS, "-I", INC, "-o", exe, main, "-L", HERE, "-ltap_hip",
                  "-ltapgpu", "-Wl,-rpath,$ORIGIN", "-lm"])
    return lib


def build_synth():
    """Host-threaded zone-to-zone free-flow times for the workload generator
    (no HIP dependency)."""
    lib = os.path.join(HERE, "libtapsynth.so")
    _run(["gcc", *C_FLAGS, "-shared", "-o", lib, os.path.join(CSRC, "zone_times.c"), "-lpthread",
          "-lm"])
    return lib


def build_all():
    build_engine()
    build_host()
    build_synth()


if __name__ == "__main__":
    build_all()
    print("built", os.path.join(HERE, "libtapgpu.so"))
