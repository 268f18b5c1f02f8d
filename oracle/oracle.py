"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and
by bench.py's CPU-baseline leg, always as the checker.  The oracle reads the
tables of an initialised libmarx_gpu configuration (mxg_get_config) and runs
the reference's stage-major passes on the CPU over a Marx_Photon_Type list.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# MARX_ORACLE_LIB names another build of the oracle (a planted-mistake
# build of oracle/mutants.py); the default is the in-tree build.
LIB_PATH = os.environ.get("MARX_ORACLE_LIB") or os.path.join(HERE, "liboracle.so")


class OrcCtx(ctypes.Structure):
    _fields_ = [("cfg", ctypes.c_void_p), ("fa", ctypes.c_void_p), ("da", ctypes.c_void_p),
                ("seed", ctypes.c_ulonglong), ("source_id", ctypes.c_uint)]


_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} not built: run `make -C oracle`")
    lib = ctypes.CDLL(LIB_PATH)
    u32p = ctypes.POINTER(ctypes.c_uint)
    f32p = ctypes.POINTER(ctypes.c_float)
    f64p = ctypes.POINTER(ctypes.c_double)
    lib.orc_philox4x32_10.argtypes = [u32p, u32p, u32p]
    lib.orc_philox4x32_10.restype = None
    for name in ("orc_stream_uniforms", "orc_stream_gaussians"):
        fn = getattr(lib, name)
        fn.argtypes = [ctypes.c_ulonglong, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint, f64p, ctypes.c_uint]
        fn.restype = None
    lib.orc_binary_search_f.argtypes = [ctypes.c_float, f32p, ctypes.c_uint]
    lib.orc_binary_search_f.restype = ctypes.c_uint
    lib.orc_interpolate_f.argtypes = [ctypes.c_float, f32p, f32p, ctypes.c_uint]
    lib.orc_interpolate_f.restype = ctypes.c_float
    lib.orc_binary_search_d.argtypes = [ctypes.c_double, f64p, ctypes.c_uint]
    lib.orc_binary_search_d.restype = ctypes.c_uint
    lib.orc_interpolate_d.argtypes = [ctypes.c_double, f64p, f64p, ctypes.c_uint]
    lib.orc_interpolate_d.restype = ctypes.c_double
    lib.orc_quadratic_root.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double, f64p, f64p]
    lib.orc_quadratic_root.restype = ctypes.c_int
    lib.orc_reflectivity.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double]
    lib.orc_reflectivity.restype = ctypes.c_double
    lib.orc_wfold_interp.argtypes = [ctypes.POINTER(OrcCtx), ctypes.c_void_p, ctypes.c_double,
                                     ctypes.c_double, ctypes.c_double]
    lib.orc_wfold_interp.restype = ctypes.c_double
    lib.orc_interpolate_n_fvector.argtypes = [f32p, ctypes.POINTER(f32p), ctypes.c_uint, f32p,
                                              ctypes.POINTER(f32p), ctypes.c_uint, ctypes.c_uint]
    lib.orc_interpolate_n_fvector.restype = ctypes.c_int
    for name in ("orc_mirror_reflect", "orc_grating_diffract", "orc_detect"):
        fn = getattr(lib, name)
        fn.argtypes = [ctypes.POINTER(OrcCtx), ctypes.c_void_p]
        fn.restype = ctypes.c_int
    lib.orc_create_photons.argtypes = [ctypes.POINTER(OrcCtx), ctypes.c_void_p, ctypes.c_uint,
                                       u32p, f64p]
    lib.orc_create_photons.restype = ctypes.c_int
    lib.orc_create_batch.argtypes = [ctypes.POINTER(OrcCtx), ctypes.c_void_p, ctypes.c_uint, f64p, u32p,
                                     ctypes.c_uint, ctypes.c_double, ctypes.c_double]
    lib.orc_create_batch.restype = ctypes.c_int
    lib.orc_create_user.argtypes = [ctypes.POINTER(OrcCtx), ctypes.c_void_p, ctypes.c_uint, u32p, f64p,
                                    ctypes.c_double, ctypes.c_void_p]
    lib.orc_create_user.restype = ctypes.c_int
    lib.orc_trace_batch_parallel.argtypes = [ctypes.POINTER(OrcCtx), ctypes.c_uint, f64p, u32p, ctypes.c_uint,
                                             ctypes.c_double, ctypes.c_int]
    lib.orc_trace_batch_parallel.restype = ctypes.c_long
    lib.orc_diff_forward.argtypes = [ctypes.POINTER(OrcCtx), ctypes.c_void_p, ctypes.c_uint, ctypes.c_uint,
                                     ctypes.c_void_p, f64p, f64p, f64p, f64p, f64p, f64p, f64p,
                                     ctypes.POINTER(ctypes.c_ubyte)]
    lib.orc_diff_forward.restype = ctypes.c_int
    lib.orc_diff_adjoint.argtypes = [ctypes.c_uint, f64p, f64p, f64p, f64p, f64p, f64p,
                                     ctypes.POINTER(ctypes.c_ubyte), ctypes.c_void_p, f64p, f64p]
    lib.orc_diff_adjoint.restype = ctypes.c_int
    lib.orc_diff_forward_parallel.argtypes = lib.orc_diff_forward.argtypes + [ctypes.c_int]
    lib.orc_diff_forward_parallel.restype = ctypes.c_int
    lib.orc_diff_adjoint_parallel.argtypes = lib.orc_diff_adjoint.argtypes + [ctypes.c_int]
    lib.orc_diff_adjoint_parallel.restype = ctypes.c_int
    lib.orc_dither_photons.argtypes = [ctypes.POINTER(OrcCtx), ctypes.c_void_p]
    lib.orc_dither_photons.restype = ctypes.c_int
    lib.orc_sort_photons.argtypes = [ctypes.c_void_p]
    lib.orc_sort_photons.restype = ctypes.c_int
    lib.orc_trace_parallel.argtypes = [ctypes.POINTER(OrcCtx), ctypes.c_uint, ctypes.c_uint, ctypes.c_int]
    lib.orc_trace_parallel.restype = ctypes.c_long
    i32p = ctypes.POINTER(ctypes.c_int)
    lib.orc_event_coords.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint, f32p, u32p,
                                     ctypes.POINTER(ctypes.c_byte), f32p, f32p, f32p, ctypes.POINTER(f32p),
                                     f32p, u32p, u32p, i32p, i32p, f64p, f64p, f64p, f64p,
                                     ctypes.POINTER(ctypes.c_short)]
    lib.orc_event_coords.restype = ctypes.c_int
    lib.orc_set_strut_taps.argtypes = [f64p, ctypes.c_uint]
    lib.orc_set_strut_taps.restype = None
    lib.orc_subpix.argtypes = [f32p, ctypes.c_uint, ctypes.c_float, f32p, f32p]
    lib.orc_subpix.restype = ctypes.c_int
    _lib = lib
    return lib


def _arr(a, ctype):
    return a.ctypes.data_as(ctypes.POINTER(ctype))


def philox(ctr, key):
    lib = load()
    c = np.asarray(ctr, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    lib.orc_philox4x32_10(_arr(c, ctypes.c_uint), _arr(k, ctypes.c_uint), _arr(out, ctypes.c_uint))
    return out


def stream_uniforms(seed, tag, stage, n, source_id=0):
    out = np.zeros(n, dtype=np.float64)
    load().orc_stream_uniforms(seed, tag, stage, source_id, _arr(out, ctypes.c_double), n)
    return out


def stream_gaussians(seed, tag, stage, n, source_id=0):
    out = np.zeros(n, dtype=np.float64)
    load().orc_stream_gaussians(seed, tag, stage, source_id, _arr(out, ctypes.c_double), n)
    return out


def interpolate_f(x, xp, yp):
    xp = np.ascontiguousarray(xp, dtype=np.float32)
    yp = np.ascontiguousarray(yp, dtype=np.float32)
    return load().orc_interpolate_f(float(x), _arr(xp, ctypes.c_float), _arr(yp, ctypes.c_float), len(xp))


def binary_search_f(x, xp):
    xp = np.ascontiguousarray(xp, dtype=np.float32)
    return load().orc_binary_search_f(float(x), _arr(xp, ctypes.c_float), len(xp))


def interpolate_d(x, xp, yp):
    xp = np.ascontiguousarray(xp, dtype=np.float64)
    yp = np.ascontiguousarray(yp, dtype=np.float64)
    return load().orc_interpolate_d(float(x), _arr(xp, ctypes.c_double), _arr(yp, ctypes.c_double), len(xp))


def quadratic_root(a, b, c):
    rp, rm = ctypes.c_double(), ctypes.c_double()
    st = load().orc_quadratic_root(a, b, c, ctypes.byref(rp), ctypes.byref(rm))
    return st, rp.value, rm.value


def reflectivity(cos_theta, beta, delta):
    return load().orc_reflectivity(cos_theta, beta, delta)


def interpolate_n_fvector(x, oldx, oldy):
    """x (n,), oldx (m,), oldy (k, m) -> (k, n) float32."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    oldx = np.ascontiguousarray(oldx, dtype=np.float32)
    oldy = [np.ascontiguousarray(r, dtype=np.float32) for r in oldy]
    out = [np.zeros(len(x), dtype=np.float32) for _ in oldy]
    f32p = ctypes.POINTER(ctypes.c_float)
    yp = (f32p * len(out))(*[_arr(o, ctypes.c_float) for o in out])
    oyp = (f32p * len(oldy))(*[_arr(o, ctypes.c_float) for o in oldy])
    st = load().orc_interpolate_n_fvector(_arr(x, ctypes.c_float), yp, len(x), _arr(oldx, ctypes.c_float),
                                          oyp, len(oldx), len(oldy))
    if st != 0:
        raise ValueError("interpolate_n_fvector needs at least 2 table rows")
    return np.array(out)


def subpix(table, n, energy):
    """acis_subpix.c:283-322 over one table (n energies, then n chip x and n
    chip y offsets): (dx, dy)."""
    t = np.ascontiguousarray(table, dtype=np.float32)
    dx, dy = ctypes.c_float(), ctypes.c_float()
    load().orc_subpix(_arr(t, ctypes.c_float), int(n), float(energy), ctypes.byref(dx), ctypes.byref(dy))
    return dx.value, dy.value


class Oracle:
    """Runs the oracle passes with the tables of an initialised Tracer."""

    def __init__(self, tracer, seed=None, source_id=0):
        self.lib = load()
        cfg, fa, _, da, _ = tracer.config_pointers()
        s = tracer.lib.mxg_get_seed() if seed is None else seed
        self.ctx = OrcCtx(cfg, fa, da, int(s), int(source_id))

    def _run(self, fn, photons):
        st = fn(ctypes.byref(self.ctx), ctypes.cast(photons.ptr, ctypes.c_void_p))
        if st != 0:
            raise RuntimeError(f"oracle pass {fn.__name__} failed ({st})")

    def reflect(self, photons):
        self._run(self.lib.orc_mirror_reflect, photons)

    def reflect_with_strut_taps(self, photons):
        """orc_mirror_reflect, returning the ray (x, p) of each photon
        (attribute order) as it reached the pre-, cap and post-collimator
        strut sets: array (capacity, 3, 6), NaN where it did not reach one."""
        taps = np.full((photons.capacity, 3, 6), np.nan)
        self.lib.orc_set_strut_taps(_arr(taps, ctypes.c_double), photons.capacity)
        try:
            self._run(self.lib.orc_mirror_reflect, photons)
        finally:
            self.lib.orc_set_strut_taps(None, 0)
        return taps

    def dither(self, photons):
        """The dither of a replayed ray file applied to photons in place."""
        self._run(self.lib.orc_dither_photons, photons)

    def diffract(self, photons):
        self._run(self.lib.orc_grating_diffract, photons)

    def detect(self, photons):
        self._run(self.lib.orc_detect, photons)

    def create_photons(self, photons, num, exposure_time=None):
        n = ctypes.c_uint()
        exp = ctypes.c_double(exposure_time) if exposure_time is not None else None
        st = self.lib.orc_create_photons(ctypes.byref(self.ctx), ctypes.cast(photons.ptr, ctypes.c_void_p),
                                         num, ctypes.byref(n), ctypes.byref(exp) if exp is not None else None)
        if st != 0:
            raise RuntimeError("oracle create_photons failed")
        return n.value

    def create_user(self, photons, num, plugin, exposure_time=None, tstart=0.0):
        """The photons of a USER source whose plug-in is the loaded CDLL
        plugin (its user_create_ray)."""
        n = ctypes.c_uint()
        exp = ctypes.byref(ctypes.c_double(exposure_time)) if exposure_time is not None else None
        fn = ctypes.cast(plugin.user_create_ray, ctypes.c_void_p)
        st = self.lib.orc_create_user(ctypes.byref(self.ctx), ctypes.cast(photons.ptr, ctypes.c_void_p),
                                      num, ctypes.byref(n), exp, tstart, fn)
        if st != 0:
            raise RuntimeError("oracle create_user failed")
        return n.value

    def create_batch(self, photons, dirs, counts, tag_start=0, start_time=0.0, exposure=1.0e4):
        """The photons of a batch of point sources (mxg_trace_batch)."""
        d = np.ascontiguousarray(dirs, dtype=np.float64).reshape(-1)
        c = np.ascontiguousarray(counts, dtype=np.uint32)
        st = self.lib.orc_create_batch(ctypes.byref(self.ctx), ctypes.cast(photons.ptr, ctypes.c_void_p),
                                       len(c), _arr(d, ctypes.c_double), _arr(c, ctypes.c_uint),
                                       tag_start, start_time, exposure)
        if st != 0:
            raise RuntimeError("oracle create_batch failed")

    def diff_forward(self, diff_cfg, n, tag_start, spec, nthreads=1):
        """The weighted trace of n photons: (FP64 image, per-photon state);
        nthreads != 1 runs it on that many OpenMP threads (0: all)."""
        img = np.zeros((spec.npix, spec.npix), dtype=np.float64)
        st = {k: np.zeros(n) for k in ("w", "y", "z")}
        st.update({k: np.zeros(5 * n) for k in ("dw", "dy", "dz")})
        st["shell"] = np.zeros(n, dtype=np.uint8)
        args = [ctypes.byref(self.ctx), ctypes.byref(diff_cfg), n, tag_start, ctypes.byref(spec),
                _arr(img, ctypes.c_double), *[_arr(st[k], ctypes.c_double) for k in ("w", "y", "z", "dw", "dy", "dz")],
                _arr(st["shell"], ctypes.c_ubyte)]
        r = (self.lib.orc_diff_forward(*args) if nthreads == 1
             else self.lib.orc_diff_forward_parallel(*args, nthreads))
        if r != 0:
            raise RuntimeError("oracle diff_forward failed")
        st["n"] = n
        return img, st

    def diff_adjoint(self, st, spec, g, nthreads=1):
        """The FP64 gradient (20,) for d(loss)/d(image) = g."""
        g = np.ascontiguousarray(g, dtype=np.float64)
        grad = np.zeros(20)
        args = [st["n"], *[_arr(st[k], ctypes.c_double) for k in ("w", "y", "z", "dw", "dy", "dz")],
                _arr(st["shell"], ctypes.c_ubyte), ctypes.byref(spec), _arr(g, ctypes.c_double),
                _arr(grad, ctypes.c_double)]
        r = (self.lib.orc_diff_adjoint(*args) if nthreads == 1
             else self.lib.orc_diff_adjoint_parallel(*args, nthreads))
        if r != 0:
            raise RuntimeError("oracle diff_adjoint failed")
        return grad

    def event_coords(self, ecfg, cols, subtab=None, sub_off=None, sub_n=None):
        """marx2fits over the rows of a dump: ecfg an mxg_event_cfg (its
        total_time is unused: the time column holds it), cols the columns as
        read from the dump files (TIME, TAG, CCDID, CHIPX, CHIPY, B_ENERGY
        and, when the simulation dithered, RA, DEC, ROLL, DET_DY, DET_DZ,
        DET_THETA).  Returns the dict of event columns."""
        n = len(cols["TAG"])
        c = {k: np.ascontiguousarray(cols[k], dtype=dt) for k, dt in
             (("TIME", np.float32), ("TAG", np.uint32), ("CCDID", np.int8), ("CHIPX", np.float32),
              ("CHIPY", np.float32), ("B_ENERGY", np.float32))}
        names = ("RA", "DEC", "ROLL", "DET_DY", "DET_DZ", "DET_THETA")
        dcols = [np.ascontiguousarray(cols[k], dtype=np.float32) if k in cols else np.zeros(n, np.float32)
                 for k in names]
        f32p = ctypes.POINTER(ctypes.c_float)
        dptr = (f32p * 6)(*[_arr(a, ctypes.c_float) for a in dcols])
        if subtab is None:
            subtab, sub_off, sub_n = np.zeros(1, np.float32), np.zeros(512, np.uint32), np.zeros(512, np.uint32)
        subtab = np.ascontiguousarray(subtab, dtype=np.float32)
        sub_off = np.ascontiguousarray(sub_off, dtype=np.uint32)
        sub_n = np.ascontiguousarray(sub_n, dtype=np.uint32)
        out = {"tag": c["TAG"].copy(), "ccd_id": c["CCDID"].astype(np.int32),
               "tdetx": np.zeros(n, np.int32), "tdety": np.zeros(n, np.int32),
               "detx": np.zeros(n), "dety": np.zeros(n), "x": np.zeros(n), "y": np.zeros(n),
               "fltgrade": np.zeros(n, np.int16)}
        st = self.lib.orc_event_coords(
            ctypes.c_void_p(self.ctx.cfg), ctypes.byref(ecfg), n, _arr(c["TIME"], ctypes.c_float),
            _arr(c["TAG"], ctypes.c_uint), _arr(c["CCDID"], ctypes.c_byte), _arr(c["CHIPX"], ctypes.c_float),
            _arr(c["CHIPY"], ctypes.c_float), _arr(c["B_ENERGY"], ctypes.c_float), dptr,
            _arr(subtab, ctypes.c_float), _arr(sub_off, ctypes.c_uint), _arr(sub_n, ctypes.c_uint),
            _arr(out["tdetx"], ctypes.c_int), _arr(out["tdety"], ctypes.c_int), _arr(out["detx"], ctypes.c_double),
            _arr(out["dety"], ctypes.c_double), _arr(out["x"], ctypes.c_double), _arr(out["y"], ctypes.c_double),
            _arr(out["fltgrade"], ctypes.c_short))
        if st != 0:
            raise RuntimeError("oracle event_coords failed")
        return out

    def trace_batch_parallel(self, dirs, counts, tag_start=0, exposure=1.0e4, nthreads=0):
        d = np.ascontiguousarray(dirs, dtype=np.float64).reshape(-1)
        c = np.ascontiguousarray(counts, dtype=np.uint32)
        return self.lib.orc_trace_batch_parallel(ctypes.byref(self.ctx), len(c), _arr(d, ctypes.c_double),
                                                 _arr(c, ctypes.c_uint), tag_start, exposure, nthreads)

    def trace_parallel(self, num, chunk=100000, nthreads=0):
        return self.lib.orc_trace_parallel(ctypes.byref(self.ctx), num, chunk, nthreads)
