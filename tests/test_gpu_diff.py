"""The differentiable mirror on the GPU (mxg_diff_forward / mxg_diff_adjoint
through the C-ABI and the torch.autograd.Function), against the FP64 CPU
oracle and finite differences.

Tolerances: the image is summed in FP32 on the GPU and in FP64 by the oracle
(relative 1e-5 of the peak); per-photon weights, positions and derivatives
are FP64 on both sides, so the gradient for the same d(loss)/d(image)
agrees to 1e-6 relative (of the largest component); central differences
of the FP64 loss of the GPU's per-photon state agree to 1e-4."""
import ctypes

import numpy as np
import pytest
import torch

from marx_amd import Tracer
from marx_amd._lib import ImageSpec, check
from marx_amd.diff import NUM_PARAMS, DiffMirror, setup
from oracle.oracle import Oracle

ON_AXIS = dict(SourceRA=250.2134679741175, SourceDEC=-53.75743813458669)
N = 20000


def _params(rng, scale=1.0):
    p = np.zeros(NUM_PARAMS)
    for s in range(4):
        p[5 * s:5 * s + 2] = rng.normal(0, 1.0 * scale, 2)
        p[5 * s + 2:5 * s + 5] = rng.normal(0, 0.01 * scale, 3)
    return p


@pytest.fixture(scope="module")
def setup_mirror(gpu):
    t = Tracer(seed=11, source=True, GratingType="NONE", DetectorType="NONE", SourceType="POINT",
               SpectrumType="FLAT", MinEnergy=1.49, MaxEnergy=1.49, DitherModel="NONE", **ON_AXIS)
    dm = DiffMirror(t, photons=N, tag_start=100, npix=128, pixel_size=0.004)
    return t, dm


def _gpu_image(dm, p):
    img = torch.empty((dm.npix, dm.npix), dtype=torch.float32, device="cuda")
    token = dm.forward_into(p, img)
    return img, token


@pytest.mark.gpu
def test_image_and_gradient_match_the_oracle(setup_mirror):
    t, dm = setup_mirror
    rng = np.random.default_rng(5)
    p = _params(rng)
    o = Oracle(t, seed=11)
    spec = ImageSpec(dm.npix, 0.004, 0.0, 0.0)
    target, _ = o.diff_forward(setup(_params(rng)), N, 100, spec)
    oimg, st = o.diff_forward(setup(p), N, 100, spec)

    img, token = _gpu_image(dm, p)
    gs = dm.state()
    # photon by photon: weights, positions and derivatives
    assert np.array_equal(gs["shell"], st["shell"])
    assert np.array_equal(gs["w"] > 0, st["w"] > 0), f"alive differs for photons {np.nonzero((gs['w'] > 0) != (st['w'] > 0))[0][:10]}"
    np.testing.assert_allclose(gs["w"], st["w"], rtol=1e-9, atol=0)
    alive = np.tile(st["w"] > 0, 5)
    for k in ("y", "z"):
        np.testing.assert_allclose(gs[k][st["w"] > 0], st[k][st["w"] > 0], rtol=0, atol=1e-9)
    for k, tol in (("dw", 1e-6), ("dy", 1e-6), ("dz", 1e-6)):
        a, b = gs[k][alive], st[k][alive]
        np.testing.assert_allclose(a, b, rtol=tol, atol=tol * np.abs(b).max(), err_msg=k)
    # photons of weight 0 have derivatives 0 (lost before the paraboloid,
    # their columns are filled in by mxg_diff_state)
    dead = np.tile(st["w"] == 0, 5)
    for k in ("dw", "dy", "dz"):
        assert not gs[k][dead].any(), k
    g_img = img.cpu().numpy().astype(np.float64)
    peak = oimg.max()
    assert np.abs(g_img - oimg).max() <= 1e-5 * peak
    assert dm.total_weight == pytest.approx(st["w"].sum(), rel=1e-12)

    G = (2.0 * (img - torch.as_tensor(target, dtype=torch.float32, device="cuda"))).contiguous()
    grad = dm.adjoint(G, token)
    ograd = o.diff_adjoint(st, spec, G.cpu().numpy().astype(np.float64))
    np.testing.assert_allclose(grad, ograd, rtol=1e-6, atol=1e-6 * np.abs(ograd).max())
    # the reduction is deterministic
    assert np.array_equal(grad, dm.adjoint(G, token))


def _splat64(st, npix, pix):
    """FP64 bilinear image of per-photon (w, y, z), the splat of the
    forward kernel."""
    img = np.zeros(npix * npix)
    w, y, z = st["w"], st["y"], st["z"]
    a = w > 0
    u = y[a] / pix + 0.5 * npix - 0.5
    v = z[a] / pix + 0.5 * npix - 0.5
    i0, j0 = np.floor(u), np.floor(v)
    fu, fv = u - i0, v - j0
    for dy_, dz_, b in ((0, 0, (1 - fu) * (1 - fv)), (1, 0, fu * (1 - fv)), (0, 1, (1 - fu) * fv), (1, 1, fu * fv)):
        iy, iz = i0.astype(int) + dy_, j0.astype(int) + dz_
        ok = (iy >= 0) & (iz >= 0) & (iy < npix) & (iz < npix)
        np.add.at(img, iz[ok] * npix + iy[ok], w[a][ok] * b[ok])
    return img.reshape(npix, npix)


@pytest.mark.gpu
def test_autograd_matches_finite_differences(setup_mirror):
    """loss.backward() through the autograd Function vs central differences
    of the FP64 loss of the GPU's per-photon positions and weights.  The
    bilinear splat makes the loss piecewise smooth, with a kink wherever a
    photon crosses a pixel boundary; the steps (2e-6 arcsec, 1e-7 mm, about
    1e-7 mm at the focal plane) are small enough that few photons cross one,
    and central differences converge to the derivative there (at 1e-5 mm
    they are still 3 percent off for a decenter on this case)."""
    t, dm = setup_mirror
    rng = np.random.default_rng(9)
    p0 = _params(rng)
    with torch.no_grad():
        target = dm.image(torch.as_tensor(_params(rng), dtype=torch.float64)).clone()
    t64 = target.cpu().numpy().astype(np.float64)

    params = torch.tensor(p0, dtype=torch.float64, requires_grad=True)
    loss = ((dm.image(params) - target) ** 2).sum()
    loss.backward()
    grad = params.grad.numpy()

    def run(p):
        img = torch.empty((dm.npix, dm.npix), dtype=torch.float32, device="cuda")
        dm.forward_into(p, img)
        return dm.state()

    def loss(st):
        return float(((_splat64(st, dm.npix, 0.004) - t64) ** 2).sum())

    s0 = run(p0)
    scale = np.abs(grad).max()
    for j in range(NUM_PARAMS):
        h = 2e-6 if (j % 5) < 2 else 1e-7
        dp = np.zeros(NUM_PARAMS)
        dp[j] = h
        sp, sm = run(p0 + dp), run(p0 - dp)
        # a photon that is vignetted on one side of p0 only (strut edge, end
        # of a mirror, scatter beyond pi/4) is a step in the loss, which a
        # central difference reads as a huge slope: hold those few at p0
        step = ((sp["w"] > 0) != (s0["w"] > 0)) | ((sm["w"] > 0) != (s0["w"] > 0))
        assert step.sum() <= 5, f"parameter {j}: {step.sum()} photons change state"
        for st in (sp, sm):
            for k in ("w", "y", "z"):
                st[k][step] = s0[k][step]
        fd = (loss(sp) - loss(sm)) / (2 * h)
        assert grad[j] == pytest.approx(fd, rel=1e-4, abs=1e-4 * scale), f"parameter {j}"


@pytest.mark.gpu
def test_adjoint_needs_the_latest_forward(setup_mirror):
    t, dm = setup_mirror
    params = torch.zeros(NUM_PARAMS, dtype=torch.float64, requires_grad=True)
    img = dm.image(params)
    with torch.no_grad():
        dm.image(torch.zeros(NUM_PARAMS, dtype=torch.float64))
    with pytest.raises(RuntimeError, match="another forward pass ran since"):
        img.sum().backward()


@pytest.mark.gpu
def test_adjoint_refuses_another_mirrors_pass(setup_mirror):
    """The device keeps one forward state for the process: a pass of a second
    DiffMirror (here with a larger image) between a forward pass and its
    backward makes the library refuse the backward instead of returning the
    other pass's gradient or reading past the smaller gradient image."""
    t, dm = setup_mirror
    other = DiffMirror(t, photons=dm.photons, npix=2 * dm.npix, pixel_size=dm.spec.pixel_size)
    params = torch.zeros(NUM_PARAMS, dtype=torch.float64, requires_grad=True)
    img = dm.image(params)
    with torch.no_grad():
        other.image(torch.zeros(NUM_PARAMS, dtype=torch.float64))
    with pytest.raises(RuntimeError, match="another forward pass ran since"):
        img.sum().backward()
    # the latest pass is the other mirror's, of another size
    g = torch.zeros((dm.npix, dm.npix), dtype=torch.float32, device="cuda")
    with pytest.raises(RuntimeError, match="pixels a side"):
        check(dm.lib.mxg_diff_adjoint(dm.lib.mxg_diff_last_pass(), dm.npix, ctypes.c_void_p(g.data_ptr()),
                                      (ctypes.c_double * NUM_PARAMS)()), "mxg_diff_adjoint")


@pytest.mark.gpu
def test_value_only_pass_equals_the_tangent_pass(setup_mirror):
    """mxg_diff_image computes the values of mxg_diff_forward with the
    derivative arithmetic left out: weights, positions and shells are equal
    bit for bit, the images to the FP32 summation order (1e-6 of the peak),
    and the adjoint refuses to follow it."""
    t, dm = setup_mirror
    p = _params(np.random.default_rng(21))
    a, _ = _gpu_image(dm, p)
    sa = dm.state()
    b = torch.empty_like(a)
    tok = dm.forward_into(p, b, tangents=False)
    sb = dm.state(tangents=False)
    for k in ("w", "y", "z", "shell"):
        assert np.array_equal(sa[k], sb[k]), k
    assert dm.total_weight == pytest.approx(sa["w"].sum(), rel=1e-12)
    peak = float(a.max())
    assert float((a - b).abs().max()) <= 1e-6 * peak
    with pytest.raises(RuntimeError, match="no derivatives"):
        dm.adjoint(torch.zeros_like(b), tok)
    with pytest.raises(RuntimeError, match="no derivatives"):
        dm.state()
