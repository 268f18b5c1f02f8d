"""The differentiable mirror from Python: focal-plane images of
survival-weighted photons as a function of the 20 shell misalignment
parameters, with their gradient by the library's adjoint kernel behind a
torch.autograd.Function.

Parameters (MXG_DIFF_NPAR = 5 per shell pair, pairs 1, 3, 4, 6 in order):
tilt az and tilt el in arcsec, decenter y, decenter z and despace x in mm.

    dm = DiffMirror(tracer, photons=10**7)
    img = dm.image(params)                     # (256, 256) float32, cuda
    loss = ((img - target) ** 2).sum()
    loss.backward()                            # params.grad (float64)
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import DIFF_NPAR, DiffCfg, ImageSpec, check

NUM_PARAMS = 4 * DIFF_NPAR
PARAM_NAMES = [f"{q}{s}" for s in (1, 3, 4, 6)
               for q in ("tilt_az_P", "tilt_el_P", "decenter_y_P", "decenter_z_P", "despace_P")]


def setup(params):
    """mxg_diff_setup: the shell geometry (DiffCfg) at the parameters."""
    p = np.ascontiguousarray(params, dtype=np.float64)
    assert p.shape == (NUM_PARAMS,)
    d = DiffCfg()
    check(_lib.load().mxg_diff_setup(p.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                     ctypes.byref(d)), "mxg_diff_setup")
    return d


class DiffMirror:
    """Weighted-photon images of a fixed photon set (photons tags
    tag_start ... of the tracer's source) through the misaligned mirror."""

    def __init__(self, tracer, photons, tag_start=0, npix=256, pixel_size=0.004, center=(0.0, 0.0)):
        self.tracer = tracer
        self.lib = tracer.lib
        self.photons = int(photons)
        self.tag_start = int(tag_start)
        self.spec = ImageSpec(int(npix), float(pixel_size), float(center[0]), float(center[1]))
        self.npix = int(npix)
        self.total_weight = 0.0

    def forward_into(self, params, image, tangents=True):
        """mxg_diff_forward (or, with tangents=False, mxg_diff_image: no
        derivatives, no adjoint) into a float32 CUDA tensor of (npix, npix)."""
        import torch
        p = np.ascontiguousarray(np.asarray(params, dtype=np.float64))
        assert p.shape == (NUM_PARAMS,)
        assert image.is_cuda and image.dtype == torch.float32 and image.is_contiguous()
        assert image.numel() == self.npix * self.npix
        torch.cuda.current_stream().synchronize()
        tw = ctypes.c_double()
        fn, name = ((self.lib.mxg_diff_forward, "mxg_diff_forward") if tangents
                    else (self.lib.mxg_diff_image, "mxg_diff_image"))
        check(fn(p.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), self.photons, self.tag_start,
                 ctypes.byref(self.spec), ctypes.c_void_p(image.data_ptr()), ctypes.byref(tw)), name)
        self.total_weight = tw.value
        # the library's id of this pass, over every DiffMirror of the process
        return int(self.lib.mxg_diff_last_pass())

    def adjoint(self, grad_image, token):
        """mxg_diff_adjoint of forward pass `token` for d(loss)/d(image).
        The device holds the state of the last forward pass of any
        DiffMirror; the library refuses a token of an older pass and a
        gradient image of another size."""
        import torch
        if tuple(grad_image.shape) != (self.npix, self.npix):
            raise ValueError(f"gradient image of shape {tuple(grad_image.shape)} for an image of "
                             f"{self.npix} x {self.npix} pixels")
        g = grad_image.detach().to(torch.float32).contiguous()
        torch.cuda.current_stream().synchronize()
        out = (ctypes.c_double * NUM_PARAMS)()
        check(self.lib.mxg_diff_adjoint(ctypes.c_ulonglong(token), self.npix, ctypes.c_void_p(g.data_ptr()),
                                        out), "mxg_diff_adjoint")
        return np.array(out)

    def kernel_ms(self):
        """(forward, adjoint, enter, listed): device ms of the last
        k_diff_forward, k_diff_adjoint, and k_diff_enter with its list
        compaction; photons that reached the paraboloid (k_diff_forward's)."""
        ms = (ctypes.c_float * 3)()
        listed = ctypes.c_uint()
        check(self.lib.mxg_diff_kernel_ms(ms, ctypes.byref(listed)), "mxg_diff_kernel_ms")
        return float(ms[0]), float(ms[1]), float(ms[2]), int(listed.value)

    def state(self, tangents=True):
        """Per-photon state of the last forward pass: dict of w, y, z (n,),
        shell (n,) and, with tangents, dw, dy, dz (5 n, parameter-major)."""
        n = self.photons
        st = {k: np.zeros(n) for k in ("w", "y", "z")}
        keys = ("dw", "dy", "dz") if tangents else ()
        st.update({k: np.zeros(DIFF_NPAR * n) for k in keys})
        st["shell"] = np.zeros(n, dtype=np.uint8)
        f64 = lambda k: st[k].ctypes.data_as(ctypes.POINTER(ctypes.c_double)) if k in st else None
        check(self.lib.mxg_diff_state(n, *[f64(k) for k in ("w", "y", "z", "dw", "dy", "dz")],
                                      st["shell"].ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte))),
              "mxg_diff_state")
        st["n"] = n
        return st

    def image(self, params):
        """Differentiable image (torch) of params, a (20,) float64 tensor."""
        return MirrorImage.apply(params, self)


def _image_function():
    import torch

    class _MirrorImage(torch.autograd.Function):
        """image = weighted focal-plane image(params); backward runs the
        library's adjoint kernel (mxg_diff_adjoint)."""

        @staticmethod
        def forward(ctx, params, dm):
            img = torch.empty((dm.npix, dm.npix), dtype=torch.float32, device="cuda")
            ctx.token = dm.forward_into(params.detach().cpu().numpy(), img)
            ctx.dm = dm
            ctx.device, ctx.dtype = params.device, params.dtype
            return img

        @staticmethod
        def backward(ctx, grad_img):
            g = ctx.dm.adjoint(grad_img, ctx.token)
            return torch.as_tensor(g, dtype=ctx.dtype, device=ctx.device), None

    return _MirrorImage


class _Lazy:
    """MirrorImage.apply without importing torch at module import."""
    _fn = None

    def apply(self, *args):
        if _Lazy._fn is None:
            _Lazy._fn = _image_function()
        return _Lazy._fn.apply(*args)


MirrorImage = _Lazy()
