"""The C-ABI boundary without a GPU: libmarx_gpu.so loads, exports every
function and variable that include/*.h declares, the Python binding covers
all of them, the structure layouts seen from Python are the C compiler's,
the parameter-file reader behaves like the reference's pfile, and every
launching entry point fails loudly (-1 and a message) when no GPU is there
(no CPU fallback)."""
import ctypes
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

import marx_amd
from marx_amd import _lib
from cfgview import STRUCTS

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", h) for h in ("marx_gpu.h", "marx_gpu_cfg.h")]


def declared_symbols():
    src = "".join(open(h).read() for h in HEADERS)
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    src = re.sub(r"^\s*#.*$", "", src, flags=re.M)
    src = src.replace('extern "C" {', "")
    src = re.sub(r"typedef\s+struct\s*\w*\s*\{.*?\}\s*\w+\s*;", "", src, flags=re.S)
    src = re.sub(r"typedef[^;]*;", "", src, flags=re.S)
    names = []
    for d in (x.strip().lstrip("}").strip() for x in src.split(";")):
        if not d:
            continue
        m = re.search(r"([A-Za-z_]\w*)\s*\((?:[^()]|\([^()]*\))*\)\s*$", d, flags=re.S)
        if m:
            names.append(m.group(1))
        elif d.startswith("extern"):
            names.append(re.findall(r"(\w+)\s*$", d)[0])
    return names


def test_library_exports_every_declared_symbol():
    lib = marx_amd.load()
    names = declared_symbols()
    for must in ("marx_mirror_init", "marx_mirror_reflect", "marx_grating_init", "marx_grating_diffract",
                 "marx_detector_init", "marx_detect", "marx_create_photons", "mxg_trace_device",
                 "mxg_get_config", "Marx_Mirror_Geometric_Area"):
        assert must in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_every_function():
    names = set(declared_symbols()) - {"Marx_Mirror_Geometric_Area", "_Marx_TStart_MJDsecs", "_Marx_TStart_Yrs"}
    bound = set(_lib.exported_symbols())
    assert names - bound == set()


def _compile_and_run(tmp_path, body):
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    src = tmp_path / "layout.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "marx_gpu.h"\n'
                   '#include "marx_gpu_cfg.h"\nint main(void){\n' + body + "return 0;}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    return subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()


def test_config_layouts_match_the_compiler(tmp_path):
    body, expect = [], []
    for cname, cls in STRUCTS.items():
        body.append(f'printf("%zu\\n", sizeof({cname}));')
        expect.append(ctypes.sizeof(cls))
        for fname, _ in cls._fields_:
            body.append(f'printf("%zu\\n", offsetof({cname}, {fname}));')
            expect.append(getattr(cls, fname).offset)
    got = [int(v) for v in _compile_and_run(tmp_path, "\n".join(body))]
    assert got == expect


def test_photon_record_layout_matches_the_compiler(tmp_path):
    """ATTR_DTYPE (the numpy view of Marx_Photon_Attr_Type, marx.h:51-103)."""
    fields = {"energy": "energy", "x": "x", "p": "p", "arrival_time": "arrival_time", "flags": "flags",
              "y_pixel": "y_pixel", "z_pixel": "z_pixel", "u_pixel": "u_pixel", "v_pixel": "v_pixel",
              "dither": "dither_state", "pi": "pi", "pulse_height": "pulse_height",
              "mirror_shell": "mirror_shell", "ccd_num": "ccd_num", "detector_region": "detector_region",
              "order": "order", "support_orders": "support_orders", "tag": "tag"}
    body = ['printf("%zu\\n", sizeof(Marx_Photon_Attr_Type));']
    body += [f'printf("%zu\\n", offsetof(Marx_Photon_Attr_Type, {c}));' for c in fields.values()]
    body += ['printf("%zu\\n", sizeof(Marx_Photon_Type));']
    body += [f'printf("%zu\\n", offsetof(Marx_Photon_Type, {f}));' for f, _ in _lib.PhotonList._fields_]
    got = [int(v) for v in _compile_and_run(tmp_path, "\n".join(body))]
    want = [136] + [marx_amd.ATTR_DTYPE.fields[f][1] for f in fields]
    want += [ctypes.sizeof(_lib.PhotonList)] + [getattr(_lib.PhotonList, f).offset
                                               for f, _ in _lib.PhotonList._fields_]
    assert got == want


def _no_gpu():
    lib = marx_amd.load()
    return lib.mxg_init() != 0


def test_entry_points_fail_loudly_without_gpu():
    if not _no_gpu():
        pytest.skip("a GPU is present")
    from marx_amd import Photons, Tracer, on_axis_photons
    t = Tracer(GratingType="NONE", DetectorType="NONE")
    assert "device" in _lib.last_error() or "HIP" in _lib.last_error()
    ph = Photons(16)
    ph.load(on_axis_photons(16, 1.49))
    with pytest.raises(RuntimeError, match="marx_mirror_reflect failed"):
        t.reflect(ph)
    with pytest.raises(RuntimeError, match="mxg_trace_device failed"):
        t.trace_device(16)


# ---- parameter files (pfile subset: pf_open_parameter_file, pf_get_*, pf_set_value) ----

def test_parameter_file_parsing(tmp_path):
    p = tmp_path / "t.par"
    p.write_text('# comment line\n'
                 'Alpha,r,a,1.5,,,"a real"\n'
                 'Count,i,h,42,0,100,"an integer"\n'
                 'Flag,b,h,yes,,,"a boolean"\n'
                 'Name,s,a,"hello world",,,"a string"\n'
                 'Where,f,h,"$MXG_TEST_DIR",,,"a file from the environment"\n')
    os.environ["MXG_TEST_DIR"] = "/some/where"
    pf = marx_amd.ParamFile(str(p))
    lib = pf.lib
    assert pf.get_double("alpha") == 1.5            # names are case-insensitive
    v = ctypes.c_int()
    assert lib.pf_get_integer(pf.handle, b"Count", ctypes.byref(v)) == 0 and v.value == 42
    assert lib.pf_get_boolean(pf.handle, b"Flag", ctypes.byref(v)) == 0 and v.value == 1
    assert pf.get_string("Name") == "hello world"
    buf = ctypes.create_string_buffer(256)
    assert lib.pf_get_file(pf.handle, b"Where", buf, 256) == 0
    assert buf.value.decode() == "/some/where"
    pf.set("Alpha", 2.25)
    pf.set("Flag", False)
    assert pf.get_double("Alpha") == 2.25
    assert lib.pf_get_boolean(pf.handle, b"Flag", ctypes.byref(v)) == 0 and v.value == 0
    with pytest.raises(RuntimeError):
        pf.get_double("NoSuchParameter")
    with pytest.raises(RuntimeError):
        pf.set("NoSuchParameter", 1)


def test_standard_parameter_file_opens():
    pf = marx_amd.ParamFile()
    assert pf.get_string("GratingType") in ("HETG", "LETG", "NONE")
    assert pf.get_double("FocalLength") == pytest.approx(10070.0)


def test_launch_schedule_validation():
    """mxg_set_schedule: blocks per CU 1..32 for every stage kernel, and 0
    for the grating only (its default: one round of the blocks a CU holds
    at once); a rejected schedule leaves the installed one unchanged."""
    t = marx_amd.Tracer(seed=1, GratingType="NONE", DetectorType="NONE")
    default = t.get_schedule()
    assert default["grating"] == 0 and default["source"] >= 1
    try:
        t.set_schedule(grating=8)
        assert t.get_schedule()["grating"] == 8
        t.set_schedule(grating=0)
        assert t.get_schedule()["grating"] == 0
        for k in ("source", "enter", "reflect", "detect"):
            with pytest.raises(RuntimeError):
                t.set_schedule(**{k: 0})
        with pytest.raises(RuntimeError):
            t.set_schedule(grating=33)
        assert t.get_schedule() == default
    finally:
        t.set_schedule(**{k: default[k] for k in ("source", "enter", "reflect", "grating", "detect")})
