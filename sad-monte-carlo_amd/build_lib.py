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


# host_asan: the host side of every unit under AddressSanitizer, device code
# untouched, linked without the runtime; load it with clang's ASan runtime
# preloaded and SADMC_LIB pointing at it:
#   python sad-monte-carlo_amd/build_lib.py --host-asan
#   SADMC_LIB=sad-monte-carlo_amd/libsadmc_amd_hostasan.so ASAN_OPTIONS=detect_leaks=0 \
#   LD_PRELOAD=$(/opt/rocm/llvm/bin/clang -print-file-name=libclang_rt.asan-x86_64.so) \
#   python -m pytest tests -m "not gpu" -p no:xdist
HOST_ASAN = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fno-omit-frame-pointer", "-Xarch_host", "-g"]
OUT_HOST_ASAN = os.path.join(HERE, "libsadmc_amd_hostasan.so")


def build(verbose=False, jobs=None, host_asan=False):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    out_path = OUT_HOST_ASAN if host_asan else OUT
    tmpdir = tempfile.mkdtemp(prefix="sadmc_build_")
    objs, procs = [], []
    jobs = jobs or min(len(UNITS), max(1, (os.cpu_count() or 4)))
    pending = list(UNITS)
    failed = []
    while pending or procs:
        while pending and len(procs) < jobs:
            u = pending.pop(0)
            obj = os.path.join(tmpdir, u.replace(".hip", ".o"))
            cmd = [hipcc] + CFLAGS + (HOST_ASAN if host_asan else []) + ["-c", "-o", obj, os.path.join(CSRC, u)]
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
    tmp = out_path + ".tmp"
    r = subprocess.run([hipcc, ARCH, "-shared", "-fPIC", "-o", tmp] + objs, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link failed")
    os.replace(tmp, out_path)
    for o in objs:
        os.remove(o)
    os.rmdir(tmpdir)
    return out_path


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv, host_asan="--host-asan" in sys.argv))
