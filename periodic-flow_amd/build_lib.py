"""Build libpfhip.so (gfx950) in-tree with hipcc.

Each translation unit under csrc/ is compiled in parallel, then linked into
periodicflow/libpfhip.so next to the Python package that loads it.  Run
``python periodic-flow_amd/build_lib.py`` (or ``__graft_entry__.build()``).
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "periodicflow", "libpfhip.so")
OBJ = os.path.join(HERE, "_build")

# one translation unit per kernel family, precision and launch variant
SOURCES = [
    "pf_col_f32_v0.hip",
    "pf_col_f32_v1.hip",
    "pf_col_f32_v2.hip",
    "pf_col_f32_v3.hip",
    "pf_col_f32_v4.hip",
    "pf_col_f32_v5.hip",
    "pf_col_f32_v6.hip",
    "pf_col_f32_v7.hip",
    "pf_col_f64.hip",
    "pf_row_f32_v0.hip",
    "pf_row_f32_v1.hip",
    "pf_row_f32_v2.hip",
    "pf_row_f32_v3.hip",
    "pf_row_f32_v4.hip",
    "pf_row_f64.hip",
    "pf_api.hip",
    "pf_point.hip",
    "pf_pair.hip",
]

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    # packed FP32 VALU forms are kept out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # no fused multiply-adds the source does not write: every product is
    # rounded on its own, so a butterfly gives the same bits in every launch
    # variant and kernel it is inlined into (launch variants, paired launches
    # and the register-resident Stokes steps are bitwise equal to the
    # default path); about 1% of step time
    "-ffp-contract=off",
]


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _compile(src):
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc()] + FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    msgs = "\n".join(
        line for line in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in line
    )
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{msgs}")
    return obj, msgs


def build(verbose=True, jobs=None):
    os.makedirs(OBJ, exist_ok=True)
    jobs = jobs or min(len(SOURCES), max(1, (os.cpu_count() or 4)))
    with cf.ThreadPoolExecutor(jobs) as ex:
        results = list(ex.map(_compile, SOURCES))
    for obj, msgs in results:
        if verbose and msgs.strip():
            print(msgs)
    tmp = OUT + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + [o for o, _ in results]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, OUT)
    if verbose:
        print("built", OUT)
    return OUT


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
