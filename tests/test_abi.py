"""CPU checks of the C ABI: the library loads, exports every declared entry
point, and its struct layouts match the numpy records the tests and bench use."""
import ctypes as C
import os
import re
import subprocess
import textwrap

import numpy as np
import pytest

from dbox_amd import native, scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "dbox_hip.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:dbx_status|void|const char \*)\s*(dbx_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    decl = declared_functions()
    assert len(decl) >= 18
    out = subprocess.run(["nm", "-D", "--defined-only", native.LIB_PATH], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\bT (dbx_\w+)", out))
    missing = [d for d in decl if d not in exported]
    assert not missing, missing
    assert set(decl) == set(native.EXPORTS)
    L = native.lib()
    for d in decl:
        assert getattr(L, d) is not None


def test_struct_layouts_match_numpy_records(tmp_path):
    src = tmp_path / "sizes.c"
    src.write_text(textwrap.dedent(f"""
        #include <stdio.h>
        #include <stddef.h>
        #include "{HEADER}"
        int main(void) {{
            printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu\\n", sizeof(dbx_body_def), sizeof(dbx_fixture_def),
                   sizeof(dbx_body_state), sizeof(dbx_contact), sizeof(dbx_profile), sizeof(dbx_config),
                   offsetof(dbx_fixture_def, p), sizeof(dbx_contact_event), sizeof(dbx_contact_setting));
            return 0;
        }}"""))
    exe = tmp_path / "sizes"
    subprocess.run(["gcc", "-std=c99", "-Wall", "-o", str(exe), str(src)], check=True)
    sizes = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    assert sizes[0] == scenes.BODY_DTYPE.itemsize
    assert sizes[1] == scenes.FIXTURE_DTYPE.itemsize
    assert sizes[2] == native.BODY_STATE_DTYPE.itemsize
    assert sizes[3] == native.CONTACT_DTYPE.itemsize
    assert sizes[4] == C.sizeof(native.DbxProfile)
    assert sizes[5] == C.sizeof(native.DbxConfig)
    assert sizes[6] == scenes.FIXTURE_DTYPE.fields["p"][1]
    assert sizes[7] == native.EVENT_DTYPE.itemsize
    assert sizes[8] == native.SETTING_DTYPE.itemsize


def test_create_without_device_fails_cleanly_or_succeeds():
    L = native.lib()
    h = C.c_void_p()
    cfg = native.DbxConfig(0, 0, 0, 0)
    st = L.dbx_create(C.byref(cfg), C.byref(h))
    assert st in (native.DBX_OK, native.DBX_E_HIP)
    if st == native.DBX_OK:
        L.dbx_destroy(h)
    bad = native.DbxConfig(0, 0, 100, 0)  # block size must be a multiple of 64
    assert L.dbx_create(C.byref(bad), C.byref(h)) == native.DBX_E_INVALID_ARG
    assert L.dbx_create(C.byref(cfg), None) == native.DBX_E_INVALID_ARG


def test_null_handle_calls_return_invalid_arg():
    L = native.lib()
    assert L.dbx_step(None, 0.016, 8, 3, 3) == native.DBX_E_INVALID_ARG
    assert L.dbx_sync(None) == native.DBX_E_INVALID_ARG
    assert L.dbx_finalize(None) == native.DBX_E_INVALID_ARG
    L.dbx_destroy(None)


def test_scene_generators_are_deterministic_and_sized():
    a = scenes.pyramid()
    b = scenes.pyramid()
    assert a.bodies.tobytes() == b.bodies.tobytes() and len(a.bodies) == 211
    sizes = scenes.world_sizes(4096, scenes.config_seed(3))
    assert sizes.min() >= 8 and sizes.max() <= 1024
    assert 55 <= np.median(sizes) <= 75
    w = scenes.batched_worlds(3, sizes=np.array([8, 64, 200]))
    assert [len(s.bodies) for s in w] == [9, 65, 201]
    st = scenes.stacks_world(1)
    assert 20 <= st.dynamic_count <= 240
    big = scenes.big_pyramid(50000)
    assert big.dynamic_count == 50000


def test_scene_definitions_build_in_the_oracle():
    from oracle import pyoracle as po
    for s in [scenes.pyramid(), scenes.mixed_container(200, 40, 80, 3), scenes.stacks_world(2), scenes.hello_world()]:
        w = po.OracleWorld(s)
        w.step()
        assert len(w.bodies()) == len(s.bodies)
