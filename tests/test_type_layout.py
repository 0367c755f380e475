"""include/gamera_cs_types.h restates the reference's structs so that
libgcs.so reads a reference caller's grid, storage, options and contexts in
place. This test pins that claim: for every struct the header defines, the
size and every member's offset, compiled once against the header and once
against the reference's own headers (nonorthogonal/, cubedsphere/), must be
equal. A member name the reference lacks fails the reference compile.

Needs gcc and the reference tree; it runs on the CPU only."""
import os
import re
import shutil
import subprocess

import pytest

from conftest import ROOT

REF = "/root/reference"
HEADER = os.path.join(ROOT, "include", "gamera_cs_types.h")
REF_HEADERS = ["nonorthogonal_storage.h", "nonorthogonal_grid.h", "nonorthogonal_geometry.h",
               "nonorthogonal_state.h", "nonorthogonal_sweep.h", "nonorthogonal_step.h",
               "nonorthogonal_background.h", "nonorthogonal_radial_map.h",
               "nonorthogonal_mi.h", "solar_wind.h", "cubedsphere_step.h",
               "cubedsphere_earth_startup.h", "cubedsphere_earth_boundary.h",
               "cubedsphere_mi_dipole.h", "cubedsphere_decomposition.h",
               "cubedsphere_halo.h"]
REF_INC = [os.path.join(REF, "nonorthogonal", "src", "nonorthogonal"),
           os.path.join(REF, "nonorthogonal", "src"), os.path.join(REF, "cubedsphere")]


def structs(text):
    """(name, [members]) of every `typedef struct { ... } name;` whose body
    holds only plain member declarations (nested struct bodies are skipped)."""
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    out = []
    for m in re.finditer(r"typedef\s+struct\s*\{(.*?)\}\s*(\w+)\s*;", text, flags=re.S):
        body, name = m.group(1), m.group(2)
        if "{" in body:
            continue
        members = []
        for decl in body.split(";"):
            decl = decl.strip()
            if not decl:
                continue
            fn = re.search(r"\(\s*\*\s*(\w+)\s*\)", decl)  # function pointer member
            if fn:
                members.append(fn.group(1))
                continue
            decl = re.sub(r"\[[^\]]*\]", "", decl)          # array extents
            parts = [p.strip() for p in decl.split(",")]
            first = re.match(r".*?(\w+)\s*$", parts[0])
            members.append(first.group(1))
            members += [re.sub(r"[*\s]", "", p) for p in parts[1:]]
        out.append((name, members))
    return out


def probe_source(items, includes):
    lines = [*(f'#include "{h}"' for h in includes), "#include <stddef.h>",
             "#include <stdio.h>", "int main(void) {"]
    for name, members in items:
        lines.append(f'  printf("{name} %zu\\n", sizeof({name}));')
        for mem in members:
            lines.append(f'  printf("{name}.{mem} %zu\\n", offsetof({name}, {mem}));')
    lines += ["  return 0;", "}"]
    return "\n".join(lines) + "\n"


def run_probe(tmp_path, tag, source, include_dirs):
    src = tmp_path / f"layout_{tag}.c"
    exe = tmp_path / f"layout_{tag}"
    src.write_text(source)
    cc = subprocess.run(["gcc", "-std=gnu99", *(f"-I{d}" for d in include_dirs), str(src),
                         "-o", str(exe)], capture_output=True, text=True)
    assert cc.returncode == 0, cc.stderr[-3000:]
    return subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout


def test_types_header_matches_reference_layouts(tmp_path):
    if shutil.which("gcc") is None or not os.path.isdir(os.path.join(REF, "cubedsphere")):
        pytest.skip("needs gcc and the reference tree")
    with open(HEADER) as f:
        items = structs(f.read())
    assert len(items) >= 25, [n for n, _ in items]
    ours = run_probe(tmp_path, "ours", probe_source(items, ["gamera_cs_types.h"]),
                     [os.path.dirname(HEADER)])
    ref = run_probe(tmp_path, "ref", probe_source(items, REF_HEADERS), REF_INC)
    diff = [(a, b) for a, b in zip(ours.splitlines(), ref.splitlines()) if a != b]
    print(f"{len(items)} structs, {len(ours.splitlines())} sizes and offsets compared",
          flush=True)
    assert not diff, diff[:20]
