"""The C-ABI library loads and exports every function include/sem_capi.h
declares; argument validation runs before any device work (no GPU needed)."""
import ctypes as C
import os
import re

import pytest

from mun_amd import _lib as L

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "sem_capi.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sem_[a-z_0-9]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    lib = L.lib()
    names = declared_functions()
    assert len(names) >= 14
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(L.EXPORTS)


def _desc(**kw):
    d = dict(nelem_x=4, nelem_z=4, order=4, xmin=0.0, xmax=100.0, zmin=0.0, zmax=100.0, dt=1e-3, nt=10,
             sponge_width=20.0, sponge_cmax=3000.0, sponge_R=1e-8, sponge_sides=15, precision=0,
             adjoint_mode=0, record_mode=0, checkpoint_every=5, lagrange_eps=1e-12)
    d.update(kw)
    return L.SemDesc(**d)


@pytest.mark.parametrize("kw,code", [
    (dict(order=5), -2), (dict(nelem_x=0), -1), (dict(xmax=0.0), -1), (dict(dt=0.0), -1),
    (dict(nt=0), -1), (dict(precision=7), -1), (dict(adjoint_mode=3), -1), (dict(record_mode=2), -1),
    (dict(sponge_width=0.0), -1),
    # 32-bit byte offsets within one solve's padded array: 8192^2 elements pad to
    # 32769 x 32800 nodes, 4.3 GB in FP32; 5800^2 to 23201 x 23232 nodes, 4.3 GB in FP64
    (dict(nelem_x=11585, nelem_z=11585), -1), (dict(nelem_x=1 << 21), -1),
    (dict(nelem_x=8192, nelem_z=8192), -1), (dict(nelem_x=5800, nelem_z=5800, precision=1), -1)])
def test_plan_create_rejects_bad_descriptors(kw, code):
    lib = L.lib()
    h = C.c_void_p()
    rc = lib.sem_plan_create(C.byref(_desc(**kw)), C.byref(h))
    assert rc == code
    assert h.value is None
    assert lib.sem_last_error().decode()
    with pytest.raises(L.SemError):
        L.check(rc)


def test_null_arguments_fail_cleanly():
    lib = L.lib()
    assert lib.sem_plan_create(None, None) == -1
    assert lib.sem_plan_dims(None, None, None, None) == -1
    assert lib.sem_forward(None, None, None, 0, None) == -1
    assert b"null" in lib.sem_last_error()


@pytest.mark.parametrize("kw", [dict(win_begin=5, win_end=3), dict(win_end=11), dict(win_begin=-1),
                                dict(win_begin=10)])
def test_plan_create_rejects_bad_windows(kw):
    lib = L.lib()
    h = C.c_void_p()
    assert lib.sem_plan_create(C.byref(_desc(**kw)), C.byref(h)) == -1
    assert "window" in lib.sem_last_error().decode()


def test_struct_layouts_match_the_header(tmp_path):
    """sizeof / offsetof of every ABI struct as a C compiler lays them out
    equals the ctypes mirror."""
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    structs = {"sem_desc": L.SemDesc, "sem_batch": L.SemBatch, "sem_slot": L.SemSlot,
               "sem_curve_desc": L.SemCurveDesc, "sem_request": L.SemRequest}
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void) {"]
    for name, cls in structs.items():
        lines.append(f'printf("{name} size %zu\\n", sizeof({name}));')
        for fname, _ in cls._fields_:
            lines.append(f'printf("{name} {fname} %zu\\n", offsetof({name}, {fname}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run([cc, str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {}
    for ln in out:
        if ln.strip():
            parts = ln.split()
            got[tuple(parts[:-1])] = int(parts[-1])
    for name, cls in structs.items():
        assert got[(name, "size")] == C.sizeof(cls), name
        for fname, _ in cls._fields_:
            assert got[(name, fname)] == getattr(cls, fname).offset, (name, fname)


def test_schedule_and_streaming_entry_points_validate_arguments():
    """The schedule and streaming entry points reject a null plan or bad
    values before any device work."""
    lib = L.lib()
    assert lib.sem_plan_set_schedule(None, 0, 0, 1) == -1
    assert lib.sem_plan_get_schedule(None, None, None, None) == -1
    assert lib.sem_plan_schedule_auto(None, 1, 1, 1 << 30) == -1
    assert lib.sem_plan_set_streaming(None, 0) == -1
    assert lib.sem_plan_set_step_fusion(None, 2) == -1
    assert lib.sem_plan_get_step_fusion(None, None) == -1
    assert lib.sem_slots_step2(None, None, 1, 0, 0, 0, None) == -1
    assert lib.sem_workspace_bytes(None, 1, 1, 1, None) == -1
    assert "bad argument" in lib.sem_last_error().decode()


def test_noalloc_projection_validates_before_device_work():
    lib = L.lib()
    assert lib.sem_bspline_project_noalloc(None, 1, 8, 8, None, None, None) == -1
    assert "bad argument" in lib.sem_last_error().decode()


def test_loop_entry_points_validate_before_device_work():
    """The device-resident loop's entry points reject null handles and bad
    sizes before touching a device."""
    lib = L.lib()
    h = C.c_void_p()
    assert lib.sem_loop_create(None, 4, 100, 0, 64, 16, C.byref(h)) == -1
    assert lib.sem_loop_push(None, None, None) == -1
    assert lib.sem_loop_round(None, None) == -1
    tot = C.c_longlong()
    assert lib.sem_loop_poll(None, 0, None, None, 0, C.byref(tot)) == -1
    assert lib.sem_loop_counts(None, None, None, None, None) == -1
    assert lib.sem_loop_chunk(None, None) == -1
    assert lib.sem_loop_destroy(None) == 0
    assert "bad argument" in lib.sem_last_error().decode()


def test_stream_entry_points_validate_before_device_work():
    lib = L.lib()
    assert lib.sem_stream_create(None) == -1
    assert "bad argument" in lib.sem_last_error().decode()
    assert lib.sem_stream_destroy(None) == 0
