"""The C ABI surface without a GPU: the library is built for gfx950, loads, and exports every
entry point include/kitamr_hip.h declares; the product path refuses to run without it."""
import re
from pathlib import Path

import pytest

from kitamr_hip import abi

ROOT = Path(__file__).resolve().parent.parent


def declared_entry_points():
    text = (ROOT / "include" / "kitamr_hip.h").read_text()
    return sorted(set(re.findall(r"\b(kah_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_entry_point():
    names = declared_entry_points()
    assert "kah_step" in names and "kah_set_topology" in names
    exported = abi.library_exports()
    assert set(names) <= set(abi.EXPORTS)
    missing = [n for n in names if n not in exported]
    assert not missing, missing


def test_library_is_gfx950_code_object():
    data = abi.LIB_PATH.read_bytes()
    assert b"gfx950" in data


def test_missing_library_fails_loudly(tmp_path):
    with pytest.raises(RuntimeError, match="not built"):
        abi.load_library(tmp_path / "libkitamr_hip.so")


def test_struct_layout_matches_header(tmp_path):
    """ctypes layouts of kah_params / kah_topology equal the C compiler's (offsetof each field)."""
    import subprocess
    fields = {"kah_params": [f for f, _ in abi.KahParams._fields_],
              "kah_topology": [f for f, _ in abi.KahTopology._fields_]}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "kitamr_hip.h"', "int main(){"]
    for st, fl in fields.items():
        lines.append(f'printf("{st} sizeof %zu\\n", sizeof({st}));')
        for f in fl:
            lines.append(f'printf("{st} {f} %zu\\n", offsetof({st}, {f}));')
    lines.append("return 0;}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    got = {}
    for line in out:
        if line:
            st, f, v = line.split()
            got[(st, f)] = int(v)
    for st, cls in (("kah_params", abi.KahParams), ("kah_topology", abi.KahTopology)):
        assert got[(st, "sizeof")] == __import__("ctypes").sizeof(cls)
        for f in fields[st]:
            assert getattr(cls, f).offset == got[(st, f)], (st, f)


def test_per_step_kernels_use_no_scratch_memory():
    """The compiler's resource report of the build (kernel_resources.json beside the library):
    no instantiation of the per-step kernels (slopes, flux/relaxation, immersed boundary)
    keeps arrays in scratch memory, which costs several times their run time. The walled
    slope kernel of cells with more than 8 half fluxes is the known exception (a handful of
    cells at wall corners of refined meshes)."""
    import json
    from kitamr_hip import build as B
    if not B.RESOURCES.exists():
        pytest.skip("library built without the resource report")
    res = json.loads(B.RESOURCES.read_text())
    hot = [k for k in res if any(s in k for s in ("k_slope_chunk", "k_flux_relax", "k_ib_solid",
                                                   "k_ib_face", "k_overlap_map", "k_cell_pre"))]
    assert len(hot) > 100
    spilled = sorted(k for k in hot if res[k].get("ScratchSize", 0) > 0)
    assert not spilled, spilled[:5]
    walled = [k for k in res if "7k_slopeI" in k and res[k].get("ScratchSize", 0) > 0]
    for k in walled:
        assert re.search(r"Li(16|24)ELi4EEEv", k), k
