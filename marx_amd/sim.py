"""Python host API over the C-ABI of libmarx_gpu.

Mirrors the reference's operator interface: a parameter file drives
``marx_mirror_init`` / ``marx_grating_init`` / ``marx_detector_init`` /
``marx_create_source``, and a ``Marx_Photon_Type`` list flows through
``marx_mirror_reflect`` -> ``marx_grating_diffract`` -> ``marx_detect``
(marx/src/marx.c:240-273, :641-656).  Return codes follow the reference:
-1 raises ``RuntimeError`` carrying the library's error message.
"""
import ctypes
import os

import numpy as np

from . import _lib
from ._lib import ATTR_DTYPE, PAR_FILE, PhotonList, Schedule, check


class ParamFile:
    """An IRAF-style parameter file (pfile API subset) with overrides."""

    def __init__(self, path=PAR_FILE, **overrides):
        self.lib = _lib.load()
        self.handle = self.lib.pf_open_parameter_file(path.encode(), b"rH")
        if not self.handle:
            raise RuntimeError(f"cannot open {path}: {_lib.last_error()}")
        for k, v in overrides.items():
            self.set(k, v)

    def set(self, name, value):
        if isinstance(value, bool):
            value = "yes" if value else "no"
        name = name.replace("__", "-")
        check(self.lib.pf_set_value(self.handle, name.encode(), str(value).encode()),
              f"pf_set_value({name})")

    def get_string(self, name):
        buf = ctypes.create_string_buffer(1024)
        check(self.lib.pf_get_string(self.handle, name.encode(), buf, 1024), f"pf_get_string({name})")
        return buf.value.decode()

    def get_double(self, name):
        v = ctypes.c_double()
        check(self.lib.pf_get_double(self.handle, name.encode(), ctypes.byref(v)), f"pf_get_double({name})")
        return v.value

    def close(self):
        if self.handle:
            self.lib.pf_close_parameter_file(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Photons:
    """A Marx_Photon_Type owned by the library, with a numpy view of its
    136-byte attribute records."""

    def __init__(self, capacity):
        self.lib = _lib.load()
        self.ptr = self.lib.marx_alloc_photon_type(int(capacity))
        if not self.ptr:
            raise MemoryError("marx_alloc_photon_type failed")
        self.capacity = int(capacity)

    @property
    def pt(self):
        return self.ptr.contents

    @property
    def attrs(self):
        """All capacity records (structured array view, no copy)."""
        buf = (ctypes.c_char * (self.capacity * ATTR_DTYPE.itemsize)).from_address(self.pt.attributes)
        return np.frombuffer(buf, dtype=ATTR_DTYPE)

    @property
    def active(self):
        """The first n_photons records."""
        return self.attrs[: self.pt.n_photons]

    def sorted_index(self):
        n = self.pt.num_sorted
        return np.ctypeslib.as_array(self.pt.sorted_index, shape=(max(n, 1),))[:n].copy()

    def load(self, records, sort=True):
        """Fills the list with records (ATTR_DTYPE) and builds sorted_index in
        energy order with index tie-break (marx/libsrc/source.c:361-371)."""
        n = len(records)
        if n > self.capacity:
            raise ValueError("too many photons for this list")
        a = self.attrs
        a[:n] = records
        p = self.pt
        p.n_photons = n
        p.num_sorted = n
        p.history = 0x1F
        order = np.lexsort((np.arange(n), records["energy"])) if sort else np.arange(n)
        idx = np.ctypeslib.as_array(p.sorted_index, shape=(self.capacity,))
        idx[:n] = order.astype(np.uint32)
        en = np.ctypeslib.as_array(p.sorted_energies, shape=(self.capacity,))
        en[:n] = records["energy"][order]

    def write(self, directory, write_mask=_lib.WRITE_ALL, open_mode=1, total_time=0.0):
        """marx_write_photons: this list's column files into directory."""
        check(self.lib.marx_write_photons(self.ptr, int(write_mask), os.fsencode(directory), int(open_mode),
                                          float(total_time)), "marx_write_photons")

    def dump_to_rayfile(self, path, new_file=1, total_time=0.0):
        """marx_dump_to_rayfile (adds total_time to the written photons'
        arrival times in this list, as the reference does)."""
        check(self.lib.marx_dump_to_rayfile(os.fsencode(path), int(new_file), self.ptr, float(total_time)),
              "marx_dump_to_rayfile")

    def __del__(self):
        try:
            if self.ptr:
                self.lib.marx_dealloc_photon_type(self.ptr)
                self.ptr = None
        except Exception:
            pass


def set_time(lib, tstart):
    """marx_set_time from a TStart value as marx.c:285-373 reads it: years
    below 2100 (seconds from MJDref = 1998.0 of 365.25-day years), else
    seconds from MJDref; refused below 1999."""
    secs_per_year = 365.25 * 24.0 * 3600.0
    if tstart < 1999:
        raise ValueError("A tstart value less than 1999 is not permitted")
    if tstart < 2100:
        yrs, secs = tstart, (tstart - 1998.0) * secs_per_year
    else:
        yrs, secs = 1998.0 + tstart / secs_per_year, tstart
    check(lib.marx_set_time(yrs, secs), "marx_set_time")
    return yrs, secs


class Tracer:
    """Initialises the path's modules from a parameter file and runs them."""

    def __init__(self, par=None, seed=1, source=False, uniform_times=0.0, **overrides):
        self.lib = _lib.load()
        self.par = par if par is not None else ParamFile(**overrides)
        self.lib.mxg_set_seed(int(seed))
        set_time(self.lib, self.par.get_double("TStart"))
        self.mirror = check(self.lib.marx_mirror_init(self.par.handle), "marx_mirror_init")
        self.grating = check(self.lib.marx_grating_init(self.par.handle), "marx_grating_init")
        self.detector = check(self.lib.marx_detector_init(self.par.handle), "marx_detector_init")
        self.source = None
        if source:
            self.init_source(uniform_times=uniform_times)

    @property
    def geometric_area(self):
        """Open HRMA aperture in cm^2 (Marx_Mirror_Geometric_Area)."""
        return ctypes.c_double.in_dll(self.lib, "Marx_Mirror_Geometric_Area").value

    def init_source(self, uniform_times=0.0, **overrides):
        """marx_create_source with the parameter file plus overrides."""
        for k, v in overrides.items():
            self.par.set(k, v)
        if self.source:
            self.lib.marx_close_source(self.source)
        self.source = self.lib.marx_create_source(self.par.handle)
        if not self.source:
            raise RuntimeError("marx_create_source failed: " + _lib.last_error())
        # marx.c opens the source after creating it (a USER plug-in's
        # user_open_source runs here)
        check(self.lib.marx_open_source(self.source), "marx_open_source")
        self.lib.mxg_set_uniform_times(float(uniform_times))

    def stage_counts(self):
        c = (ctypes.c_uint * 4)()
        self.lib.mxg_last_stage_counts(c)
        return list(c)

    def stage_ms(self):
        """Device ms of [source, mirror, grating, detector] in the last timed,
        uncaptured pass."""
        m = (ctypes.c_float * 4)()
        self.lib.mxg_last_stage_ms(m)
        return list(m)

    def kernel_stats(self):
        """(ms, photons_in) of the mirror's two kernels [aperture + paraboloid,
        reflections] in the last timed trace_device call."""
        m = (ctypes.c_float * 2)()
        c = (ctypes.c_uint * 2)()
        self.lib.mxg_last_kernel_stats(m, c)
        return list(m), list(c)

    def create_photons(self, photons, num, exposure_time=None):
        """marx_create_photons: num photons of the source into photons."""
        n = ctypes.c_uint()
        exp = ctypes.byref(ctypes.c_double(exposure_time)) if exposure_time is not None else None
        check(self.lib.marx_create_photons(self.source, photons.ptr, int(num), ctypes.byref(n), exp),
              "marx_create_photons")
        return n.value

    def trace_device(self, num, tag_start=0, start_time=0.0, exposure_time=-1.0, timing=False):
        """Source -> mirror -> grating -> detector without leaving HBM.
        Returns (collected, detected, stage_ms or None)."""
        ncol, ndet = ctypes.c_uint(), ctypes.c_uint()
        ms = (ctypes.c_float * 6)() if timing else None
        check(self.lib.mxg_trace_device(int(num), int(tag_start), float(start_time), float(exposure_time),
                                        ctypes.byref(ncol), ctypes.byref(ndet), ms), "mxg_trace_device")
        return ncol.value, ndet.value, (list(ms) if timing else None)

    @staticmethod
    def _batch_args(dirs, counts):
        d = np.ascontiguousarray(dirs, dtype=np.float64).reshape(-1)
        c = np.ascontiguousarray(counts, dtype=np.uint32)
        if len(d) != 3 * len(c):
            raise ValueError("dirs must hold 3 components per source")
        return d, c

    def trace_batch(self, dirs, counts, tag_start=0, start_time=0.0, exposure=1.0e4, use_graph=True):
        """mxg_trace_batch: point sources with directions dirs (k, 3) and
        photon counts (k,), traced in one device pass.  Returns (detected,
        device ms)."""
        d, c = self._batch_args(dirs, counts)
        ndet, ms = ctypes.c_uint(), ctypes.c_float()
        check(self.lib.mxg_trace_batch(len(c), d.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                       c.ctypes.data_as(ctypes.POINTER(ctypes.c_uint)), int(tag_start),
                                       float(start_time), float(exposure), int(bool(use_graph)),
                                       ctypes.byref(ndet), ctypes.byref(ms)), "mxg_trace_batch")
        return ndet.value, ms.value

    def autotune_batch(self, dirs, counts, tag_start=0, start_time=0.0, exposure=1.0e4, repeats=2):
        """mxg_autotune_batch: installs and returns (schedule dict, graph pass ms)."""
        d, c = self._batch_args(dirs, counts)
        best, ms = Schedule(), ctypes.c_float()
        check(self.lib.mxg_autotune_batch(len(c), d.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                          c.ctypes.data_as(ctypes.POINTER(ctypes.c_uint)), int(tag_start),
                                          float(start_time), float(exposure), int(repeats),
                                          ctypes.byref(best), ctypes.byref(ms)), "mxg_autotune_batch")
        return best.as_dict(), ms.value

    def get_schedule(self):
        s = Schedule()
        check(self.lib.mxg_get_schedule(ctypes.byref(s)), "mxg_get_schedule")
        return s.as_dict()

    def set_schedule(self, blocks_per_cu=None, reflect_waves=None, persistent_tile=None,
                     persistent_blocks_per_cu=None, **per_kernel):
        """Sets the launch schedule: blocks_per_cu for every kernel, or per
        kernel (source=, enter=, reflect=, grating=, detect=; grating=0,
        the default, is one round of the blocks a CU holds at once), the
        reflection kernel's waves per SIMD, and the persistent pass's tile
        (0: the staged pass) and blocks per CU."""
        s = Schedule()
        check(self.lib.mxg_get_schedule(ctypes.byref(s)), "mxg_get_schedule")
        for i, k in enumerate(Schedule.KERNELS):
            if blocks_per_cu is not None:
                s.blocks_per_cu[i] = int(blocks_per_cu)
            if k in per_kernel:
                s.blocks_per_cu[i] = int(per_kernel[k])
        if reflect_waves is not None:
            s.reflect_waves = int(reflect_waves)
        if persistent_tile is not None:
            s.persistent_tile = int(persistent_tile)
        if persistent_blocks_per_cu is not None:
            s.persistent_blocks_per_cu = int(persistent_blocks_per_cu)
        check(self.lib.mxg_set_schedule(ctypes.byref(s)), "mxg_set_schedule")

    def download(self, photons):
        check(self.lib.mxg_download_photons(photons.ptr), "mxg_download_photons")

    def write_device(self, directory, write_mask=_lib.WRITE_ALL, open_mode=1, total_time=0.0):
        """mxg_write_device: the column files of the last device-resident
        trace into directory (which must exist).  Returns the rows written."""
        n = ctypes.c_uint()
        check(self.lib.mxg_write_device(int(write_mask), os.fsencode(directory), int(open_mode),
                                        float(total_time), ctypes.byref(n)), "mxg_write_device")
        return n.value

    # (name, numpy dtype) of the columns of event_coords, in mxg_event_columns order
    EVENT_COLUMNS = (("tag", np.uint32), ("ccd_id", np.int32), ("tdetx", np.int32), ("tdety", np.int32),
                     ("detx", np.float64), ("dety", np.float64), ("x", np.float64), ("y", np.float64),
                     ("fltgrade", np.int16))

    def event_cfg(self, pix_adj="EDSER", seed=1, total_time=0.0):
        """The mxg_event_cfg that marx2fits would use for this simulation:
        FocalLength, DetOffsetX/Y/Z, Roll_Nom, the ACIS frame times and
        whether DitherModel is other than NONE, from the parameter file."""
        c = _lib.EventCfg()
        c.focal_length = self.par.get_double("FocalLength")
        for k, a in enumerate("XYZ"):
            c.det_offset[k] = self.par.get_double("DetOffset" + a)
        c.roll_nom = self.par.get_double("Roll_Nom")
        c.exposure_time = self.par.get_double("ACIS_Exposure_Time")
        c.frame_transfer_time = self.par.get_double("ACIS_Frame_Transfer_Time")
        c.total_time = float(total_time)
        c.seed = int(seed)
        c.pix_adj = _lib.PIX_ADJ[pix_adj.upper()]
        c.dither_used = int(self.par.get_string("DitherModel") != "NONE")
        return c

    def event_coords(self, capacity, pix_adj="EDSER", seed=1, total_time=0.0, cfg=None):
        """mxg_event_coords: the marx2fits event columns (tag, ccd_id,
        tdetx/y, detx/y, sky x/y, fltgrade) of the last device-resident ACIS
        trace, as a dict of numpy arrays, in the rows of write_device."""
        cfg = cfg if cfg is not None else self.event_cfg(pix_adj, seed, total_time)
        arrays = {k: np.empty(int(capacity), dt) for k, dt in self.EVENT_COLUMNS}
        cols = _lib.EventColumns(*[a.ctypes.data for a in arrays.values()])
        n = ctypes.c_uint()
        check(self.lib.mxg_event_coords(ctypes.byref(cfg), ctypes.byref(cols), int(capacity), ctypes.byref(n)),
              "mxg_event_coords")
        return {k: a[:n.value] for k, a in arrays.items()}

    def event_kernel_ms(self):
        ms = ctypes.c_float()
        self.lib.mxg_event_kernel_ms(ctypes.byref(ms))
        return ms.value

    def set_seed(self, seed):
        self.lib.mxg_set_seed(int(seed))

    def reflect(self, photons, verbose=0):
        check(self.lib.marx_mirror_reflect(photons.ptr, verbose), "marx_mirror_reflect")

    def diffract(self, photons, verbose=0):
        check(self.lib.marx_grating_diffract(photons.ptr, verbose), "marx_grating_diffract")

    def detect(self, photons, verbose=0):
        check(self.lib.marx_detect(photons.ptr, verbose), "marx_detect")

    def config_pointers(self):
        """(cfg, farena, nf, darena, nd) raw pointers of the initialised tables."""
        cfg, fa, da = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        nf, nd = ctypes.c_uint(), ctypes.c_uint()
        self.lib.mxg_get_config(ctypes.byref(cfg), ctypes.byref(fa), ctypes.byref(nf),
                                ctypes.byref(da), ctypes.byref(nd))
        return cfg.value, fa.value, nf.value, da.value, nd.value


def offaxis_directions(theta, phi):
    """Unit directions of travel (k, 3) of photons from point sources at
    off-axis angles theta (radians) and azimuths phi: an on-axis source
    sends its photons along (-1, 0, 0)."""
    theta = np.asarray(theta, dtype=np.float64)
    phi = np.asarray(phi, dtype=np.float64)
    return -np.stack([np.cos(theta), np.sin(theta) * np.cos(phi), np.sin(theta) * np.sin(phi)], axis=-1)


def on_axis_photons(n, energy, tag_start=0, rng=None, time_span=1.0e4):
    """n records of an on-axis point source: p = (-1, 0, 0), given energy
    (scalar or array), arrival times uniform over time_span."""
    rec = np.zeros(n, dtype=ATTR_DTYPE)
    rec["energy"] = energy
    rec["p"] = (-1.0, 0.0, 0.0)
    rec["tag"] = np.arange(tag_start, tag_start + n, dtype=np.uint32)
    rng = rng if rng is not None else np.random.default_rng(0)
    rec["arrival_time"] = np.sort(rng.uniform(0.0, time_span, n))
    return rec
