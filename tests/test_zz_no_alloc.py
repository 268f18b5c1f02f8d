"""mxg_trace_device_noalloc (the allocation-free variant of mxg_trace_device:
lists and counters in the library's fixed device scratch, the configuration
in its device buffers, photons in the caller's columns) against
mxg_trace_device: the same records, bit for bit, on two calls in a row, the
second with fewer photons (the scratch keeps the first call's contents),
HETG + ACIS-S with dither, also after a new set-up; and a chunk larger than
the scratch is refused with a message."""
import ctypes

import numpy as np
import pytest
import torch

from marx_amd import Photons
from marx_amd._lib import ATTR_DTYPE, SoaView
from marx_amd.workloads import WORKLOADS
from tests.test_gpu_batch import assert_same_records

# (record field, torch dtype, components)
COLUMNS = [("energy", torch.float64, 1), ("x", torch.float64, 3), ("p", torch.float64, 3),
           ("arrival_time", torch.float64, 1), ("flags", torch.int32, 1), ("y_pixel", torch.float32, 1),
           ("z_pixel", torch.float32, 1), ("u_pixel", torch.float32, 1), ("v_pixel", torch.float32, 1),
           ("dither", torch.float32, 6), ("pi", torch.float32, 1), ("pulse_height", torch.int16, 1),
           ("mirror_shell", torch.uint8, 1), ("ccd_num", torch.int8, 1), ("detector_region", torch.int8, 1),
           ("order", torch.int8, 1), ("support_orders", torch.int8, 4), ("tag", torch.int32, 1)]


def _columns(capacity):
    """Caller-owned device columns, filled with a pattern the trace must
    overwrite, and their mxg_soa_view."""
    cols, view = {}, SoaView()
    view.capacity = capacity
    for name, dt, k in COLUMNS:
        t = torch.full((k, capacity), 7, dtype=dt, device="cuda")
        cols[name] = t
        if k == 1:
            setattr(view, name, t.data_ptr())
        else:
            arr = getattr(view, name)
            for j in range(k):
                arr[j] = t[j].data_ptr()
    torch.cuda.synchronize()
    return cols, view


def _records(cols, n):
    rec = np.zeros(n, dtype=ATTR_DTYPE)
    for name, dt, k in COLUMNS:
        a = cols[name][:, :n].cpu().numpy()
        if name in ("flags", "tag"):
            a = a.view(np.uint32)
        rec[name] = a[0] if k == 1 else a.T
    return rec


def _noalloc(lib, view, n, tag_start, start_time):
    ncol, ndet = ctypes.c_uint(), ctypes.c_uint()
    st = lib.mxg_trace_device_noalloc(ctypes.byref(view), n, tag_start, start_time, -1.0, ctypes.byref(ncol),
                                      ctypes.byref(ndet))
    return st, ncol.value, ndet.value


@pytest.mark.gpu
def test_noalloc_trace_equals_the_device_trace(gpu, tmp_path):
    t = WORKLOADS["hetg-acis-s"](str(tmp_path), photons=200_000, exposure=1.0e4)
    lib = t.lib
    big = 200_000
    cols, view = _columns(big)
    for n, tag_start, start_time in ((big, 11, 50.0), (big // 3, 777, 9000.0)):
        print(f"noalloc case: n={n} tag_start={tag_start}", flush=True)
        ncol_a, ndet_a, _ = t.trace_device(n, tag_start=tag_start, start_time=start_time)
        p = Photons(n)
        t.download(p)
        want = p.active.copy()
        st, ncol_b, ndet_b = _noalloc(lib, view, n, tag_start, start_time)
        assert st == 0, lib.mxg_last_error()
        assert (ncol_b, ndet_b) == (ncol_a, ndet_a)
        got = _records(cols, ncol_b)
        assert_same_records(got, want, f"noalloc n={n}")
        assert ndet_b == int(((got["flags"] & 0xFF) == 0).sum())

    # a changed configuration that fits the device buffers is copied in place
    # (uniform arrival times: no new tables), before any allocating call
    lib.mxg_set_uniform_times(3000.0)
    n = big // 2
    st, ncol_b, ndet_b = _noalloc(lib, view, n, 5, 0.0)
    assert st == 0, lib.mxg_last_error()
    got = _records(cols, ncol_b)
    ncol_a, ndet_a, _ = t.trace_device(n, tag_start=5, start_time=0.0)
    p = Photons(n)
    t.download(p)
    assert (ncol_b, ndet_b) == (ncol_a, ndet_a)
    assert_same_records(got, p.active.copy(), "noalloc after a configuration change")
    assert got["arrival_time"].max() < 3000.0

    # a new set-up whose tables outgrow the device buffers is refused until
    # an allocating call has made them
    t2 = WORKLOADS["letg-hrc-s"](str(tmp_path), photons=1000, exposure=1.0e4)
    st, _, _ = _noalloc(lib, view, 1000, 0, 0.0)
    assert st == -1 and b"allocating call" in lib.mxg_last_error()
    t2.trace_device(1000)
    st, ncol_b, _ = _noalloc(lib, view, 1000, 0, 0.0)
    assert st == 0, lib.mxg_last_error()
    t2.trace_device(1000)
    p = Photons(1000)
    t2.download(p)
    assert_same_records(_records(cols, ncol_b), p.active.copy(), "noalloc after a new set-up")


@pytest.mark.gpu
def test_noalloc_refuses_a_chunk_beyond_its_scratch(gpu, tmp_path):
    t = WORKLOADS["hetg-acis-s"](str(tmp_path), photons=1000, exposure=1.0e4)
    lib = t.lib
    t.trace_device(1000)
    cap = 1 << 20                                   # MXG_NOALLOC_MAX_PHOTONS
    view = SoaView()
    view.capacity = cap + 1
    st, _, _ = _noalloc(lib, view, cap + 1, 0, 0.0)
    assert st == -1
    assert "exceeds" in lib.mxg_last_error().decode()
