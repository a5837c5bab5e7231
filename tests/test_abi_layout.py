"""The ctypes binding mirrors include/pdfgrade.h exactly: every PDF_* constant
it re-exports has the header's value, and every struct it mirrors has the C
struct's size and field offsets (a probe compiled with gcc against the
header prints them).  No GPU needed."""
from __future__ import annotations

import ctypes as C
import os
import re
import shutil
import subprocess

import pytest

import pdfgrade as pg

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pdfgrade.h")

# ctypes mirror -> C type name in the header
MIRRORS = {
    pg._GenParams: "pdf_gen_params",
    pg._Desc: "pdf_netlist_desc",
    pg._Info: "pdf_netlist_info",
    pg._GraderTuning: "pdf_grader_tuning",
    pg.KernelTimes: "pdf_kernel_times",
    pg._BatchTuning: "pdf_batch_tuning",
}


def header_defines() -> dict:
    text = open(HEADER).read()
    pat = r"^#define PDF_([A-Z0-9_]+)\s+\(?(-?(?:0x[0-9A-Fa-f]+|\d+))[uU]?\)?"
    return {name: int(val, 0) for name, val in re.findall(pat, text, re.M)}


def test_binding_constants_equal_header():
    defs = header_defines()
    checked = []
    for name, val in defs.items():
        if hasattr(pg, name):
            assert getattr(pg, name) == val, name
            checked.append(name)
    # error codes, gate types, tuning keys, debug flags, sizes
    for must in ("OK", "ERR_ARG", "ERR_OVERFLOW", "INPT", "DFF", "NONE", "TUNE_LANES", "TUNE_REV_COMPACT", "TUNE_REV_TILES",
                 "TUNE_WORDS_PER_LANE", "DEBUG_PI_COVER", "FEATURES", "MLP_PARAMS"):
        assert must in checked, must


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not on PATH")
def test_struct_mirrors_match_c_layout(tmp_path):
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "pdfgrade.h"', "int main(void) {"]
    for cls, cname in MIRRORS.items():
        lines.append(f'  printf("{cname} sizeof %zu\\n", sizeof({cname}));')
        for field, _ in cls._fields_:
            lines.append(f'  printf("{cname} {field} %zu\\n", offsetof({cname}, {field}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "probe.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "probe"
    r = subprocess.run(["gcc", "-std=c11", "-I", os.path.dirname(HEADER), str(src), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    c_layout = {}
    for line in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.splitlines():
        cname, field, value = line.split()
        c_layout[(cname, field)] = int(value)
    for cls, cname in MIRRORS.items():
        assert C.sizeof(cls) == c_layout[(cname, "sizeof")], cname
        for field, _ in cls._fields_:
            assert getattr(cls, field).offset == c_layout[(cname, field)], (cname, field)
