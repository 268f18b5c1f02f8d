"""The drop-in boundary from C: tests/driver/marx_driver.c follows the
reference driver's sequence (marx/src/marx.c:285-373, :641-656, :545-560,
:240-273) and is linked against libmarx_gpu.so with a parameter-file reader
of its own whose handle layout differs from the library's.  On the CPU:

- it links with the declared entry points alone;
- the init functions return the reference's module ids (HRMA 2, HETG 1 /
  LETG 2 / NONE 0, ACIS-S 3, ACIS-I 4, HRC-S 1, HRC-I 2) while reading every
  parameter through the executable's pf_* functions (ELF interposition: the
  library's own reader would not understand the driver's handle);
- without a GPU the run functions return -1 with a message, and nothing
  falls back to the CPU."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "marx_amd", "lib")
PAR = os.path.join(ROOT, "marx_amd", "data", "marx.par")


def build_driver(tmp_path):
    exe = str(tmp_path / "marx_driver")
    subprocess.run(["gcc", "-O1", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "driver", "marx_driver.c"), "-L", LIBDIR, "-lmarx_gpu",
                    f"-Wl,-rpath,{LIBDIR}", "-o", exe], check=True)
    return exe


def write_par(path, **over):
    """A copy of the standard parameter file with values replaced."""
    out = []
    for line in open(PAR):
        f = line.split(",")
        if f and f[0] in over:
            v = over[f[0]]
            f[3] = f'"{v}"' if isinstance(v, str) else str(v)
            line = ",".join(f)
        out.append(line)
    open(path, "w").write("".join(out))
    return path


def run_driver(exe, par, num, out):
    r = subprocess.run([exe, par, str(num), out], capture_output=True, text=True, timeout=600)
    lines = dict(l.split(" ", 1) for l in r.stdout.splitlines() if " " in l)
    return r.returncode, lines, r.stderr


@pytest.mark.parametrize("grating,det,ids", [("HETG", "ACIS-S", "2 1 3"), ("LETG", "HRC-S", "2 2 1"),
                                             ("NONE", "ACIS-I", "2 0 4"), ("HETG", "HRC-I", "2 1 2")])
def test_driver_links_and_initialises_through_its_own_pfile(tmp_path, grating, det, ids):
    exe = build_driver(tmp_path)
    par = write_par(str(tmp_path / "marx.par"), GratingType=grating, DetectorType=det, RandomSeed=5)
    rc, out, err = run_driver(exe, par, 1000, str(tmp_path / "out.bin"))
    assert out["modules"] == ids, (out, err)
    assert int(out["pf_calls"]) > 50


def test_driver_run_fails_loudly_without_a_gpu(tmp_path):
    import marx_amd
    if marx_amd.load().mxg_init() == 0:
        pytest.skip("a GPU is present: the run succeeds (tests/test_gpu_driver.py)")
    exe = build_driver(tmp_path)
    par = write_par(str(tmp_path / "marx.par"), RandomSeed=5)
    rc, out, err = run_driver(exe, par, 1000, str(tmp_path / "out.bin"))
    assert rc == 4
    assert out["process"] == "-1"
    assert "no HIP device" in out["error"]
    assert not os.path.exists(str(tmp_path / "out.bin"))
