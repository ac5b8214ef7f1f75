"""Build of libkitamr_hip.so for gfx950 with hipcc (in-tree, no JIT cache).

The compiler's per-kernel resource report (VGPRs, scratch bytes per lane, occupancy) is kept
beside the library as ``lib/kernel_resources.json``: a hot kernel that spills to scratch memory
runs several times slower, and the CPU tests check the report for that.
"""
from __future__ import annotations

import json
import os
import re
import subprocess
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent.parent
CSRC = PKG / "csrc"
OUT = PKG / "lib" / "libkitamr_hip.so"
SOURCES = ["kah_api.cpp", "kah_kernels.hip"] + [
    f"kah_{k}_{d}d{p}.hip" for k in ("slope", "flux") for d in (2, 3) for p in ("", "_f32")]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-x", "hip",
    # every floating-point operation rounds as written (no contraction into FMA), so the
    # FP64 path keeps the reference's association order
    "-ffp-contract=off",
    # no packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]
RESOURCES = OUT.parent / "kernel_resources.json"


def parse_resources(text):
    """Per-kernel figures of hipcc's -Rpass-analysis=kernel-resource-usage remarks."""
    out, cur = {}, None
    for line in text.splitlines():
        m = re.search(r"remark: Function Name: (\S+)", line)
        if m:
            cur = out.setdefault(m.group(1), {})
            continue
        m = re.search(r"remark:\s+(VGPRs|AGPRs|SGPRs|ScratchSize \[bytes/lane\]|"
                      r"Occupancy \[waves/SIMD\]|LDS Size \[bytes/block\]): (\d+)", line)
        if m and cur is not None:
            cur[m.group(1).split()[0]] = int(m.group(2))
    return out


def build(verbose=False):
    OUT.parent.mkdir(parents=True, exist_ok=True)
    objdir = PKG / "lib" / "obj"
    objdir.mkdir(parents=True, exist_ok=True)

    def compile_one(src):
        obj = objdir / (Path(src).stem + ".o")
        cmd = [HIPCC, *FLAGS, "-c", str(CSRC / src), "-o", str(obj)]
        if src.endswith(".hip"):
            cmd.append("-Rpass-analysis=kernel-resource-usage")
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc failed on {src}:\n{r.stderr}")
        res = parse_resources(r.stderr)
        if verbose:
            rest = [ln for ln in r.stderr.splitlines() if "kernel-resource-usage" not in ln]
            if rest:
                print("\n".join(rest))
        return obj, res

    with ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        done = list(ex.map(compile_one, SOURCES))
    objs = [o for o, _ in done]
    resources = {}
    for _, res in done:
        resources.update(res)
    RESOURCES.write_text(json.dumps(resources, indent=0, sort_keys=True))
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
