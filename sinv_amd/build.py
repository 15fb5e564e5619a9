"""Build libsinv_hip.so in-tree for gfx950 (hipcc cross-compiles without a GPU).

    python -m sinv_amd.build [-v]

Every .hip file under sinv_amd/csrc is compiled to an object with hipcc and
linked into sinv_amd/libsinv_hip.so.  The library's run path points at the HIP
runtime bundled with torch, so a process that has imported torch uses one HIP
runtime for both.
"""

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libsinv_hip.so")
ARCH = "gfx950"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

CFLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # no packed FP32 instructions anywhere in the device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
]


def _torch_libdir():
    try:
        import importlib.util
        spec = importlib.util.find_spec("torch")
        if spec and spec.origin:
            d = os.path.join(os.path.dirname(spec.origin), "lib")
            if os.path.exists(os.path.join(d, "libamdhip64.so")):
                return d
    except Exception:
        pass
    return None


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{out}")
    if verbose and out.strip():
        print(out)
    return out


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    srcs = sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))
    if not srcs:
        raise RuntimeError("no HIP sources found")
    jobs = []
    for f in srcs:
        obj = os.path.join(OBJ, f[:-4] + ".o")
        jobs.append((obj, [HIPCC, *CFLAGS, "-c", os.path.join(CSRC, f), "-o", obj]))
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        list(ex.map(lambda j: _run(j[1], verbose), jobs))
    link = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *[j[0] for j in jobs],
            "-o", LIB + ".tmp"]
    tl = _torch_libdir()
    if tl:
        link += [f"-Wl,-rpath,{tl}", "-Wl,--disable-new-dtags"]
    _run(link, verbose)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
