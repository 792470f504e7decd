"""Build the MI355X tracking engine (libsimtrack_amd.so) for gfx950.

Compiles the HIP kernels and the C-ABI host layer with hipcc and links one
shared library into simtrack_amd/lib/.  Floating-point contraction is off
for device and host code: the element maps must round exactly like the
reference's CPU maps (explicit fma() appears only in st_libm.h, identically
on both sides).
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIBNAME = "libsimtrack_amd.so"
ARCH = os.environ.get("ST_OFFLOAD_ARCH", "gfx950")

COMMON = [
    "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + CSRC, "-I" + os.path.join(ROOT, "include"),
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build step failed: " + " ".join(cmd[:3]) + " ...")
    return r


def build(verbose=False):
    """Compile every HIP/C++ source and link the shared library; returns its path."""
    os.makedirs(LIBDIR, exist_ok=True)
    objdir = os.path.join(LIBDIR, "obj")
    os.makedirs(objdir, exist_ok=True)
    hc = hipcc()
    jobs = []
    # The tracking kernels' loops start on 64-byte boundaries: with one
    # wavefront per SIMD the instruction fetch of a misaligned loop head is
    # not hidden (the spin kernel on the 998-element beam-beam ring: 265 ->
    # 255 ms per launch, same instructions, bit-identical results).
    for src, extra in (("st_kernels.hip", ["--offload-arch=" + ARCH, "-falign-loops=64"]),
                       ("st_kernels_perf.hip", ["--offload-arch=" + ARCH, "-ffp-contract=fast", "-falign-loops=64"]),
                       ("st_tune.hip", ["--offload-arch=" + ARCH]),
                       ("st_mlp.hip", ["--offload-arch=" + ARCH]),
                       ("st_ssbb.hip", ["--offload-arch=" + ARCH]),
                       ("st_pic.hip", ["--offload-arch=" + ARCH]),
                       ("st_capi.cpp", ["--offload-arch=" + ARCH, "-x", "hip"]),
                       ("st_lower.cpp", ["--offload-arch=" + ARCH, "-x", "hip"]),
                       ("st_track.cpp", ["--offload-arch=" + ARCH, "-x", "hip"]),
                       ("st_scan.cpp", ["--offload-arch=" + ARCH, "-x", "hip"]),
                       ("st_batch.cpp", ["--offload-arch=" + ARCH, "-x", "hip"]),
                       ("st_ss_host.cpp", ["--offload-arch=" + ARCH, "-x", "hip"]),
                       ("st_da.cpp", [])):
        obj = os.path.join(objdir, src.rsplit(".", 1)[0] + ".o")
        jobs.append(([hc, *COMMON, *extra, "-c", os.path.join(CSRC, src), "-o", obj], obj))
    # independent translation units: compile them side by side
    from concurrent.futures import ThreadPoolExecutor
    workers = max(1, min(len(jobs), int(os.environ.get("MAX_JOBS", "0")) or (os.cpu_count() or 1)))
    with ThreadPoolExecutor(workers) as ex:
        for f in [ex.submit(_run, cmd, verbose) for cmd, _ in jobs]:
            f.result()
    objs = [obj for _, obj in jobs]
    out = os.path.join(LIBDIR, LIBNAME)
    tmp = out + ".tmp"
    _run([hc, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp, *objs], verbose)
    os.replace(tmp, out)
    return out


def library_path():
    return os.path.join(LIBDIR, LIBNAME)


if __name__ == "__main__":
    print(build(verbose=True))
