"""Build the gfx950 HIP library ``libinelastic_hip.so`` in-tree with hipcc.

Also builds ``libinelastic_probe.so`` from ``probe/`` (a vector-ALU FMA
throughput probe the benchmark reports beside the kernel rooflines; not part
of the collision library).
"""
import concurrent.futures as cf
import os
import pathlib
import shutil
import subprocess

HERE = pathlib.Path(__file__).resolve().parent
CSRC = HERE / "csrc"
OUT = HERE / "libinelastic_hip.so"
PROBE_OUT = HERE / "libinelastic_probe.so"
BUILD = HERE / "_build"

COMMON = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
    # packed FP32 (v_pk_*_f32) stays off: the compiler must not form it
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", str(HERE.parent / "include"),
]
# per-file extra flags: transport kernels round exactly like numpy (no FMA contraction)
EXTRA = {"transport.hip": ["-ffp-contract=off"]}


def _hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose=False, jobs=8):
    hipcc = _hipcc()
    BUILD.mkdir(exist_ok=True)
    srcs = sorted(CSRC.glob("*.hip"))

    def compile_one(src):
        obj = BUILD / (src.stem + ".o")
        cmd = [hipcc, "-c", str(src), "-o", str(obj)] + COMMON + EXTRA.get(src.name, [])
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("hipcc failed for %s:\n%s" % (src.name, r.stderr[-4000:]))
        if verbose and r.stderr.strip():
            print(r.stderr)
        return obj

    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(compile_one, srcs))
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [hipcc, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", str(tmp)] + [str(o) for o in objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stderr[-4000:])
    os.replace(tmp, OUT)
    tmp = PROBE_OUT.with_suffix(".so.tmp")
    cmd = [hipcc, "-shared", "-o", str(tmp)] + [str(s) for s in sorted((HERE / "probe").glob("*.hip"))] \
        + COMMON
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("probe build failed:\n" + r.stderr[-4000:])
    os.replace(tmp, PROBE_OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
