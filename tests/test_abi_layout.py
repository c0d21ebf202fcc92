"""The ctypes structures of kerrray_amd._lib against the C header's: the same size and the same offset for
every field, from a small C program compiled against include/kerrray_hip.h (gcc, no GPU)."""

from __future__ import annotations

import ctypes
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PAIRS = {"KrParams": "kr_params", "KrOutputs": "kr_outputs", "KrImage": "kr_image", "KrDisk": "kr_disk",
         "KrJob": "kr_job", "KrPlannedOpts": "kr_planned_opts", "KrPlannedStats": "kr_planned_stats",
         "KrTune": "kr_tune", "KrImpactOpts": "kr_impact_opts", "KrImpactResult": "kr_impact_result",
         "KrBatchPrecision": "kr_batch_precision"}


@pytest.fixture(scope="module")
def c_layout(tmp_path_factory) -> dict[str, int]:
    """``{"kr_x": sizeof, "kr_x.field": offsetof}`` as the C compiler lays the header's structures out."""
    from kerrray_amd import _lib

    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "kerrray_hip.h"', "int main(void) {"]
    for py, c in PAIRS.items():
        lines.append(f'    printf("{c} %zu\\n", sizeof({c}));')
        for field, *_ in getattr(_lib, py)._fields_:
            lines.append(f'    printf("{c}.{field} %zu\\n", offsetof({c}, {field}));')
    lines += ["    return 0;", "}"]
    d = tmp_path_factory.mktemp("abi")
    src, exe = d / "layout.c", d / "layout"
    src.write_text("\n".join(lines) + "\n")
    subprocess.run([cc, "-std=c11", "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    return {k: int(v) for k, v in (row.split() for row in out.splitlines())}


@pytest.mark.parametrize("py", list(PAIRS))
def test_ctypes_structure_matches_the_header(c_layout, py) -> None:
    from kerrray_amd import _lib

    cls, c = getattr(_lib, py), PAIRS[py]
    assert ctypes.sizeof(cls) == c_layout[c]
    for field, *_ in cls._fields_:
        assert getattr(cls, field).offset == c_layout[f"{c}.{field}"], field


def test_every_header_structure_has_a_ctypes_twin() -> None:
    import re

    with open(os.path.join(ROOT, "include", "kerrray_hip.h")) as f:
        header = f.read()
    declared = set(re.findall(r"^} (kr_\w+);", header, re.M)) - {"kr_status"}
    assert declared == set(PAIRS.values())
