"""The ctypes mirrors of the C-ABI structures (solfec/_lib.py) against the
header they mirror (include/sfgs.h): a small C program compiled with gcc
prints sizeof and offsetof of every field the Python side names, and they must
equal ctypes' own.  A field added to one side only, or a type changed on one
side, fails here instead of corrupting a GPU call.  No GPU."""
import ctypes as C
import os
import shutil
import subprocess

import pytest

from solfec import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
STRUCTS = ("SFGS_PARAM", "SFGS_REPORT", "SFGS_NEWTON_PARAM", "SFGS_NEWTON_REPORT", "SFGS_BATCH_TUNE",
           "SFGS_SCENE_REPORT")


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_ctypes_structures_match_header(tmp_path):
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "sfgs.h"', "int main(void) {"]
    for name in STRUCTS:
        lines.append(f'  printf("{name} sizeof %zu\\n", sizeof({name}));')
        for field, _ in getattr(_lib, name)._fields_:
            lines.append(f'  printf("{name} {field} %zu\\n", offsetof({name}, {field}));')
    lines += ["  return 0;", "}"]
    src, exe = tmp_path / "layout.c", tmp_path / "layout"
    src.write_text("\n".join(lines) + "\n")
    subprocess.run(["gcc", "-std=gnu99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True, capture_output=True, timeout=120)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True, timeout=60).stdout.split("\n")
    got = {(a, b): int(c) for a, b, c in (ln.split() for ln in out if ln.strip())}
    for name in STRUCTS:
        cls = getattr(_lib, name)
        assert C.sizeof(cls) == got[(name, "sizeof")], name
        for field, _ in cls._fields_:
            assert getattr(cls, field).offset == got[(name, field)], (name, field)
