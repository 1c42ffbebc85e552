"""Build libld.so (HIP, gfx950) in-tree with hipcc.

One object per (precision, N) instantiation of the step engine plus the C-ABI
object, compiled in parallel and linked into ``lucky_droplets_amd/libld.so``.
Usage: ``python -m lucky_droplets_amd.build [-j JOBS]``.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
LIB = HERE / "libld.so"
BUILD = HERE / "_build"

ARCH = "gfx950"
SIZES = (16, 32, 64, 128, 256, 512)
PRECISIONS = (("float", "f32"), ("double", "f64"))

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# Packed-FP32 instructions are disabled on purpose: some of their forms have
# produced wrong results on gfx950 without any error.
BASE_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    # hipcc's host pass does not know the packed-fp32 feature; that warning is harmless
    err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{err}")
    return err


def _sources_newer(obj: Path, deps) -> bool:
    if not obj.exists():
        return True
    t = obj.stat().st_mtime
    return any(d.stat().st_mtime > t for d in deps)


def build(jobs: int | None = None, force: bool = True, verbose: bool = False) -> Path:
    """Compile every HIP translation unit for gfx950 and link libld.so."""
    BUILD.mkdir(exist_ok=True)
    headers = sorted(CSRC.glob("*.hpp")) + [HERE.parent / "include" / "ld.h"]
    tasks = []
    for ctype, tn in PRECISIONS:
        for n in SIZES:
            obj = BUILD / f"ld_inst_{tn}_{n}.o"
            cmd = [HIPCC, *BASE_FLAGS, f"-DLD_T={ctype}", f"-DLD_TN={tn}", f"-DLD_N={n}",
                   "-c", str(CSRC / "ld_inst.hip"), "-o", str(obj)]
            tasks.append((obj, cmd, headers + [CSRC / "ld_inst.hip"]))
    api_obj = BUILD / "ld_api.o"
    tasks.append((api_obj, [HIPCC, *BASE_FLAGS, "-c", str(CSRC / "ld_api.hip"), "-o", str(api_obj)],
                  headers + [CSRC / "ld_api.hip"]))
    cnn_obj = BUILD / "ld_cnn.o"
    tasks.append((cnn_obj, [HIPCC, *BASE_FLAGS, "-c", str(CSRC / "ld_cnn.hip"), "-o", str(cnn_obj)],
                  headers + [CSRC / "ld_cnn.hip"]))
    todo = [(o, c) for (o, c, deps) in tasks if force or _sources_newer(o, deps)]
    jobs = jobs or min(8, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = {ex.submit(_run, c): o for (o, c) in todo}
        for f in cf.as_completed(futs):
            msg = f.result()
            if verbose and msg.strip():
                print(futs[f].name, msg, file=sys.stderr)
    objs = [str(o) for (o, _, _) in tasks]
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *objs])
    os.replace(tmp, LIB)
    return LIB


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", "--jobs", type=int, default=None)
    ap.add_argument("--incremental", action="store_true", help="only rebuild objects older than their sources")
    ap.add_argument("-v", "--verbose", action="store_true")
    a = ap.parse_args()
    lib = build(a.jobs, force=not a.incremental, verbose=a.verbose)
    print(lib)


if __name__ == "__main__":
    main()
