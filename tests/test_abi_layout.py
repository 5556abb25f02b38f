"""The Python binding's ctypes mirrors of the C-ABI structs have the C header's layout.

A small C program built with gcc from `include/drape_hip.h` prints each struct's size and every field's offset;
they must equal what ctypes computes for the mirror in `drape_amd._ffi`. A field added to one side only (or
reordered, or given another type) shifts an offset or the size and fails here, before any call could read or
write past a struct.
"""
import ctypes
import os
import subprocess

import pytest

from drape_amd import _ffi

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

MIRRORS = {
    "drape_fabric": _ffi.Fabric,
    "drape_settings": _ffi.Settings,
    "drape_mesh_desc": _ffi.MeshDesc,
    "drape_garment_desc": _ffi.GarmentDesc,
    "drape_field_desc": _ffi.FieldDesc,
    "drape_scene_desc": _ffi.SceneDesc,
    "drape_stats": _ffi.Stats,
    "drape_counters": _ffi.Counters,
    "drape_pcg_report": _ffi.PcgReport,
    "drape_structure_sizes": _ffi.StructureSizes,
    "drape_bake_options": _ffi.BakeOptions,
    "drape_body_mesh": _ffi.BodyMesh,
}


@pytest.fixture(scope="module")
def c_layout(tmp_path_factory):
    d = tmp_path_factory.mktemp("abi")
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "drape_hip.h"', "int main(void) {"]
    for cname, mirror in MIRRORS.items():
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for field, _ in mirror._fields_:
            lines.append(f'  printf("{cname} {field} %zu\\n", offsetof({cname}, {field}));')
    lines += ["  return 0;", "}"]
    src = d / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = d / "layout"
    subprocess.run(["gcc", "-std=c11", "-Wall", "-Werror", "-I", os.path.join(REPO, "include"), str(src), "-o",
                    str(exe)], check=True, capture_output=True, text=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    layout = {}
    for line in out.splitlines():
        cname, key, value = line.split()
        layout[(cname, key)] = int(value)
    return layout


@pytest.mark.parametrize("cname", sorted(MIRRORS))
def test_ctypes_mirror_matches_the_header(c_layout, cname):
    mirror = MIRRORS[cname]
    assert c_layout[(cname, "size")] == ctypes.sizeof(mirror), cname
    for field, _ in mirror._fields_:
        assert c_layout[(cname, field)] == getattr(mirror, field).offset, (cname, field)
