"""Build libsadmc_amd.so in-tree with hipcc for gfx950.

Each system's kernels are a separate translation unit, compiled in parallel
and linked into one shared library.  Flags: -ffp-contract=off so that every
floating-point expression rounds as written (the device results are
reproducible bit-for-bit on a CPU), and packed FP32 instructions disabled.
"""
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
UNITS = ["sadmc_host.hip", "k_ising.hip", "k_lj64.hip", "k_lj32.hip", "k_wca.hip", "k_batched.hip",
         "k_learned64.hip", "k_learned32.hip", "k_train.hip", "k_entropy.hip", "k_pt.hip", "k_binning.hip", "k_zmc.hip",
         "sadmc_noalloc.hip"]
OUT = os.path.join(HERE, "libsadmc_amd.so")
ARCH = "--offload-arch=gfx950"
CFLAGS = [ARCH, "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang",
          "-packed-fp32-ops", "-Wno-unused-command-line-argument"]


def build(verbose=False, jobs=None):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    tmpdir = tempfile.mkdtemp(prefix="sadmc_build_")
    objs, procs = [], []
    jobs = jobs or min(len(UNITS), max(1, (os.cpu_count() or 4)))
    pending = list(UNITS)
    failed = []
    while pending or procs:
        while pending and len(procs) < jobs:
            u = pending.pop(0)
            obj = os.path.join(tmpdir, u.replace(".hip", ".o"))
            cmd = [hipcc] + CFLAGS + ["-c", "-o", obj, os.path.join(CSRC, u)]
            procs.append((u, obj, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
        u, obj, p = procs.pop(0)
        out, _ = p.communicate()
        if p.returncode != 0:
            failed.append((u, out))
        elif verbose and out:
            sys.stderr.write(out)
        objs.append(obj)
    if failed:
        for u, out in failed:
            sys.stderr.write(f"--- {u}\n{out}\n")
        raise RuntimeError("hipcc failed: " + ", ".join(u for u, _ in failed))
    tmp = OUT + ".tmp"
    r = subprocess.run([hipcc, ARCH, "-shared", "-fPIC", "-o", tmp] + objs, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link failed")
    os.replace(tmp, OUT)
    for o in objs:
        os.remove(o)
    os.rmdir(tmpdir)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
