"""The oracle of the differentiable mirror, pinned on the CPU:
- the tilt-matrix derivatives of mxg_diff_setup (complex step) equal central
  differences of the matrices themselves;
- decenter and despace move the geometry as defined (to_osac and the front
  position shift by the parameter);
- the oracle's forward-mode gradient of a binned-image loss equals central
  finite differences of the FP64 loss, parameter by parameter;
- moving a shell pair by a decenter moves its photons' focal-plane positions
  by exactly that decenter (a rigid shift of the optic)."""
import numpy as np
import pytest

from marx_amd import Tracer
from marx_amd._lib import ImageSpec
from marx_amd.diff import NUM_PARAMS, setup
from oracle.oracle import Oracle

ON_AXIS = dict(SourceRA=250.2134679741175, SourceDEC=-53.75743813458669)


@pytest.fixture(scope="module")
def tracer():
    t = Tracer(seed=7, source=True, GratingType="NONE", DetectorType="NONE", SourceType="POINT",
               SpectrumType="FLAT", MinEnergy=1.49, MaxEnergy=1.49, DitherModel="NONE", **ON_AXIS)
    return t


def _params(rng, scale=1.0):
    p = np.zeros(NUM_PARAMS)
    for s in range(4):
        p[5 * s:5 * s + 2] = rng.normal(0, 1.0 * scale, 2)        # arcsec
        p[5 * s + 2:5 * s + 5] = rng.normal(0, 0.01 * scale, 3)   # mm
    return p


def test_tilt_matrix_derivatives_are_exact(tracer):
    p = _params(np.random.default_rng(0))
    d = setup(p)
    for s in range(4):
        for j, h in ((0, 1e-2), (1, 1e-2)):
            dp = np.zeros(NUM_PARAMS)
            dp[5 * s + j] = h
            a, b = setup(p + dp), setup(p - dp)
            for k in range(2):
                for name, dname in (("fwd", "dfwd"), ("bwd", "dbwd")):
                    fd = (np.array(getattr(a.shell[s], name)[k]) - np.array(getattr(b.shell[s], name)[k])) / (2 * h)
                    cs = np.array(getattr(d.shell[s], dname)[k][j])
                    np.testing.assert_allclose(cs, fd, rtol=1e-6, atol=1e-14)
        # the forward and backward matrices are inverse rotations
        for k in range(2):
            f = np.array(d.shell[s].fwd[k]).reshape(3, 3)
            b = np.array(d.shell[s].bwd[k]).reshape(3, 3)
            np.testing.assert_allclose(f @ b, np.eye(3), atol=1e-14)


def test_shifts_move_the_geometry(tracer):
    base = setup(np.zeros(NUM_PARAMS))
    p = np.zeros(NUM_PARAMS)
    p[5 * 2 + 2:5 * 2 + 5] = (0.01, -0.02, 0.03)        # shell pair 4: dy, dz, dx
    d = setup(p)
    for k in range(2):
        np.testing.assert_allclose(np.array(d.shell[2].to_osac[k]) - np.array(base.shell[2].to_osac[k]),
                                   [-0.03, -0.01, 0.02], atol=1e-12)
    assert d.shell[2].front_position - base.shell[2].front_position == pytest.approx(0.03, abs=1e-12)
    assert np.array_equal(np.array(d.shell[0].to_osac), np.array(base.shell[0].to_osac))


def test_decenter_shifts_the_image(tracer):
    """A decenter of a shell pair translates the pair, so the focal-plane
    positions of its photons move by the decenter (1e-9 mm) and their
    weights do not change -- except for the few photons whose strut shadow
    changes, the struts staying where they are."""
    o = Oracle(tracer, seed=7)
    spec = ImageSpec(64, 0.01, 0.0, 0.0)
    n = 2000
    p0 = np.zeros(NUM_PARAMS)
    _, a = o.diff_forward(setup(p0), n, 0, spec)
    p1 = p0.copy()
    p1[5 * 1 + 2] = 0.004
    p1[5 * 1 + 3] = -0.007
    _, b = o.diff_forward(setup(p1), n, 0, spec)
    sel = (a["shell"] == 1) & (a["w"] > 0) & (b["w"] > 0)
    assert sel.sum() > 50
    changed = (a["shell"] == 1) & ((a["w"] > 0) != (b["w"] > 0))
    assert changed.sum() < 0.02 * (a["shell"] == 1).sum()
    np.testing.assert_allclose(b["y"][sel] - a["y"][sel], 0.004, atol=1e-9)
    np.testing.assert_allclose(b["z"][sel] - a["z"][sel], -0.007, atol=1e-9)
    np.testing.assert_allclose(b["w"][sel], a["w"][sel], rtol=1e-12)
    other = (a["shell"] != 1)
    np.testing.assert_array_equal(b["y"][other], a["y"][other])


def test_oracle_gradient_equals_finite_differences(tracer):
    """d loss / d params by the oracle's forward-mode derivatives vs central
    differences of the FP64 loss (loss = sum (image - target)^2)."""
    rng = np.random.default_rng(3)
    o = Oracle(tracer, seed=7)
    spec = ImageSpec(128, 0.004, 0.0, 0.0)
    n = 4000
    p0 = _params(rng)
    target, _ = o.diff_forward(setup(_params(rng)), n, 0, spec)

    def loss(p):
        img, _ = o.diff_forward(setup(p), n, 0, spec)
        return float(((img - target) ** 2).sum())

    img, st = o.diff_forward(setup(p0), n, 0, spec)
    grad = o.diff_adjoint(st, spec, 2.0 * (img - target))
    for j in range(NUM_PARAMS):
        h = 1e-4 if (j % 5) < 2 else 1e-6       # arcsec, mm
        dp = np.zeros(NUM_PARAMS)
        dp[j] = h
        fd = (loss(p0 + dp) - loss(p0 - dp)) / (2 * h)
        assert grad[j] == pytest.approx(fd, rel=2e-5, abs=1e-6 * np.abs(grad).max()), f"parameter {j}"


def test_threaded_oracle_equals_serial(tracer):
    """The OpenMP oracle (the CPU baseline's) gives the serial oracle's
    per-photon state bit for bit, its image and gradient to the FP64
    summation order."""
    o = Oracle(tracer, seed=7)
    spec = ImageSpec(64, 0.01, 0.0, 0.0)
    p = _params(np.random.default_rng(4))
    a_img, a = o.diff_forward(setup(p), 3000, 5, spec)
    b_img, b = o.diff_forward(setup(p), 3000, 5, spec, nthreads=4)
    for k in ("w", "y", "z", "dw", "dy", "dz", "shell"):
        assert np.array_equal(a[k], b[k]), k
    np.testing.assert_allclose(b_img, a_img, rtol=1e-12, atol=1e-12 * a_img.max())
    g = np.random.default_rng(6).normal(size=(64, 64))
    np.testing.assert_allclose(o.diff_adjoint(b, spec, g, nthreads=4), o.diff_adjoint(a, spec, g),
                               rtol=1e-12, atol=1e-12)


def test_adjoint_guard_without_a_forward_pass(tracer):
    """The adjoint's guard on the host: before any forward pass the pass id
    is 0 and the library refuses an adjoint (no forward state, no GPU
    call), and DiffMirror refuses a gradient image of the wrong shape
    before it reaches the library."""
    import ctypes
    import torch
    from marx_amd.diff import DiffMirror
    lib = tracer.lib
    assert lib.mxg_diff_last_pass() == 0
    g = np.zeros(16, dtype=np.float32)
    out = (ctypes.c_double * NUM_PARAMS)()
    assert lib.mxg_diff_adjoint(1, 4, g.ctypes.data_as(ctypes.c_void_p), out) == -1
    dm = DiffMirror(tracer, photons=100, npix=8)
    with pytest.raises(ValueError, match="shape"):
        dm.adjoint(torch.zeros((4, 4)), 1)
