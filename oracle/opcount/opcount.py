"""Per-photon FP64 operation counts of the oracle's stages (test
infrastructure).

Runs the op-counting oracle build (liboracle_opcount.so) over a sample of a
workload and reports, per stage, the FP64 adds/subs, muls, divs, sqrts and
transcendental calls per photon entering the stage.  These are the
algorithmic work figures behind the benchmark's roofline; they count the
restated reference algorithm and nothing of the GPU code.

    python -m oracle.opcount.opcount [--photons 100000] [--workload hetg-acis-s|letg-hrc-s|batched-acis-i|diff]

The diff workload counts the weighted forward pass (orc_diff_forward, per
photon of the launch) in parts: the values, the tilt tangents, the
translation tangents, and the oracle's finite-difference reflectivity slope
(not the path's: the GPU differentiates the Fresnel equations); and the
adjoint (orc_diff_adjoint) per photon.
"""
import argparse
import ctypes
import json
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

NAMES = ["add", "mul", "div", "sqrt", "sin", "cos", "tan", "atan", "atan2", "asin", "acos",
         "exp", "expm1", "log", "log1p", "pow", "fmod", "erf"]


def load_opcount():
    from oracle import oracle as O
    lib = ctypes.CDLL(os.path.join(os.path.dirname(HERE), "liboracle_opcount.so"))
    for name in ("orc_mirror_reflect", "orc_grating_diffract", "orc_detect"):
        getattr(lib, name).argtypes = [ctypes.POINTER(O.OrcCtx), ctypes.c_void_p]
        getattr(lib, name).restype = ctypes.c_int
    lib.orc_create_photons.argtypes = [ctypes.POINTER(O.OrcCtx), ctypes.c_void_p, ctypes.c_uint,
                                       ctypes.POINTER(ctypes.c_uint), ctypes.c_void_p]
    lib.orc_create_photons.restype = ctypes.c_int
    lib.orc_create_batch.argtypes = [ctypes.POINTER(O.OrcCtx), ctypes.c_void_p, ctypes.c_uint,
                                     ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_uint),
                                     ctypes.c_uint, ctypes.c_double, ctypes.c_double]
    lib.orc_create_batch.restype = ctypes.c_int
    lib.orc_opcount_reset.restype = None
    lib.orc_opcount_get.argtypes = [ctypes.POINTER(ctypes.c_ulonglong)]
    lib.orc_opcount_get_phase.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_ulonglong)]
    lib.orc_opcount_marks.argtypes = [ctypes.c_int]
    lib.orc_opcount_marks.restype = ctypes.c_ulonglong
    return lib


def _per_photon(c, entering):
    per = {k: c[i] / max(entering, 1) for i, k in enumerate(NAMES)}
    per["flops"] = per["add"] + per["mul"] + per["div"] + per["sqrt"]
    per["transcendental"] = sum(per[k] for k in NAMES[4:])
    per["photons_in"] = entering
    return per


def count_stages(tracer, num, seed=1, batch=None):
    """Per-stage counts per entering photon for one chunk of num photons, or
    for a batch of point sources given as batch = (dirs, counts)."""
    from oracle import oracle as O
    from marx_amd import Photons
    lib = load_opcount()
    cfg, fa, _, da, _ = tracer.config_pointers()
    ctx = O.OrcCtx(cfg, fa, da, seed, 0)
    ph = Photons(num)
    out = {}

    def run(stage, call, entering):
        lib.orc_opcount_reset()
        if call() != 0:
            raise RuntimeError(f"op-count oracle stage {stage} failed")
        c = (ctypes.c_ulonglong * 17)()
        lib.orc_opcount_get(c)
        out[stage] = _per_photon(c, entering)
        if stage == "mirror":
            # the mirror's two kernels: up to the paraboloid hit (every photon
            # entering the stage), and after it (photons that hit) including
            # the energy-only optical-constant lookups
            ph = [(ctypes.c_ulonglong * 17)() for _ in range(3)]
            for k in range(3):
                lib.orc_opcount_get_phase(k, ph[k])
            hits = lib.orc_opcount_marks(1)
            looked_up = lib.orc_opcount_marks(2)
            out["mirror_enter"] = _per_photon(ph[0], entering)
            a = _per_photon(ph[1], hits)
            b = _per_photon(ph[2], looked_up)
            out["mirror_reflect"] = {k: a[k] + b[k] for k in a}
            out["mirror_reflect"]["photons_in"] = hits

    n = ctypes.c_uint()
    if batch is None:
        run("source", lambda: lib.orc_create_photons(ctypes.byref(ctx), ctypes.cast(ph.ptr, ctypes.c_void_p),
                                                     num, ctypes.byref(n), None), num)
    else:
        d = np.ascontiguousarray(batch[0], dtype=np.float64).reshape(-1)
        c = np.ascontiguousarray(batch[1], dtype=np.uint32)
        run("source", lambda: lib.orc_create_batch(
            ctypes.byref(ctx), ctypes.cast(ph.ptr, ctypes.c_void_p), len(c),
            d.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), c.ctypes.data_as(ctypes.POINTER(ctypes.c_uint)),
            0, 0.0, 1.0e4), int(c.sum()))
    alive = lambda: int(((ph.active["flags"] & 0xFF) == 0).sum())
    e = alive()
    run("mirror", lambda: lib.orc_mirror_reflect(ctypes.byref(ctx), ctypes.cast(ph.ptr, ctypes.c_void_p)), e)
    e = alive()
    run("grating", lambda: lib.orc_grating_diffract(ctypes.byref(ctx), ctypes.cast(ph.ptr, ctypes.c_void_p)), e)
    e = alive()
    run("detector", lambda: lib.orc_detect(ctypes.byref(ctx), ctypes.cast(ph.ptr, ctypes.c_void_p)), e)
    out["detected"] = alive()
    return out


def count_diff(tracer, num, params, npix=256, pixel_size=0.004, seed=1):
    """Per-photon counts of the weighted forward pass and of the adjoint."""
    from oracle import oracle as O
    from marx_amd._lib import ImageSpec
    from marx_amd.diff import setup
    lib = load_opcount()
    f64p = ctypes.POINTER(ctypes.c_double)
    lib.orc_diff_forward.argtypes = ([ctypes.POINTER(O.OrcCtx), ctypes.c_void_p, ctypes.c_uint, ctypes.c_uint,
                                      ctypes.c_void_p] + [f64p] * 7 + [ctypes.POINTER(ctypes.c_ubyte)])
    lib.orc_diff_adjoint.argtypes = ([ctypes.c_uint] + [f64p] * 6
                                     + [ctypes.POINTER(ctypes.c_ubyte), ctypes.c_void_p, f64p, f64p])
    cfg, fa, _, da, _ = tracer.config_pointers()
    ctx = O.OrcCtx(cfg, fa, da, seed, 0)
    spec = ImageSpec(npix, pixel_size, 0.0, 0.0)
    d = setup(params)
    img = np.zeros(npix * npix)
    st = {k: np.zeros(n) for k, n in (("w", num), ("y", num), ("z", num), ("dw", 5 * num), ("dy", 5 * num),
                                       ("dz", 5 * num))}
    shell = np.zeros(num, dtype=np.uint8)
    p = lambda a: a.ctypes.data_as(f64p)
    cols = [p(st[k]) for k in ("w", "y", "z", "dw", "dy", "dz")]
    sh = shell.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte))
    lib.orc_opcount_reset()
    if lib.orc_diff_forward(ctypes.byref(ctx), ctypes.byref(d), num, 0, ctypes.byref(spec), p(img), *cols, sh):
        raise RuntimeError("op-count oracle diff forward failed")
    ph = [(ctypes.c_ulonglong * 17)() for _ in range(8)]
    for k in range(8):
        lib.orc_opcount_get_phase(k, ph[k])
    add = lambda *xs: {k: sum(x[k] for x in xs) for k in xs[0]}
    # bucket part + kind: part 0 up to the paraboloid hit, 4 after it; kind 0
    # values, 1 tilt tangents, 2 translation tangents, 3 the oracle's slope
    c = [_per_photon(ph[k], num) for k in range(8)]
    out = {"values": add(c[0], c[4]), "tilt_tangents": add(c[1], c[5]),
           "translation_tangents": add(c[2], c[6]), "reflectivity_slope_fd": add(c[3], c[7]),
           # the whole forward pass of each tangent count (value only, tilts,
           # all five), per photon of the pass
           "forward<0>": add(c[0], c[4]), "forward<2>": add(c[0], c[1], c[4], c[5]),
           "forward<5>": add(c[0], c[1], c[2], c[4], c[5], c[6]),
           # its part up to the paraboloid hit (k_diff_enter, values only)
           "k_diff_enter": c[0]}
    g = np.random.default_rng(0).normal(size=npix * npix)
    grad = np.zeros(20)
    lib.orc_opcount_reset()
    lib.orc_diff_adjoint(num, *cols, sh, ctypes.byref(spec), p(g), p(grad))
    c = (ctypes.c_ulonglong * 17)()
    lib.orc_opcount_get(c)
    out["k_diff_adjoint"] = _per_photon(c, num)
    out["alive"] = int((st["w"] > 0).sum())
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--photons", type=int, default=100000)
    ap.add_argument("--json", default=None)
    ap.add_argument("--workload", default="hetg-acis-s")
    a = ap.parse_args()
    from marx_amd.workloads import WORKLOADS, batched_acis_i
    if a.workload == "diff":
        from marx_amd.workloads import diff_mirror
        tracer, params, _ = diff_mirror()
        res = count_diff(tracer, a.photons, params)
        for st in ("values", "tilt_tangents", "translation_tangents", "reflectivity_slope_fd",
                   "forward<0>", "forward<2>", "forward<5>", "k_diff_enter", "k_diff_adjoint"):
            r = res[st]
            print(f"{st:22s} flops/photon={r['flops']:8.1f} (add {r['add']:.1f} mul {r['mul']:.1f} "
                  f"div {r['div']:.1f} sqrt {r['sqrt']:.1f}) transcendental/photon={r['transcendental']:.2f}")
        print(f"alive {res['alive']} of {a.photons}")
        if a.json:
            with open(a.json, "w") as fp:
                json.dump(res, fp, indent=1)
        return
    with tempfile.TemporaryDirectory() as tmp:
        if a.workload == "batched-acis-i":
            # the batch's sources with their counts scaled to about a.photons
            tracer, dirs, counts = batched_acis_i(tmp)
            c = np.maximum(1, np.round(counts * (a.photons / counts.sum()))).astype(np.uint32)
            res = count_stages(tracer, int(c.sum()), batch=(dirs, c))
        else:
            tracer = WORKLOADS[a.workload](tmp, photons=a.photons)
            res = count_stages(tracer, a.photons)
    for st in ("source", "mirror", "mirror_enter", "mirror_reflect", "grating", "detector"):
        r = res[st]
        print(f"{st:14s} photons_in={r['photons_in']:8d} flops/photon={r['flops']:8.1f} "
              f"(add {r['add']:.1f} mul {r['mul']:.1f} div {r['div']:.1f} sqrt {r['sqrt']:.1f}) "
              f"transcendental/photon={r['transcendental']:.2f}")
    if a.json:
        with open(a.json, "w") as fp:
            json.dump(res, fp, indent=1)


if __name__ == "__main__":
    main()
