"""The benchmark configurations at their full sizes on the GPU, checked
against the oracle on a sample.  Every draw is keyed by (seed, tag, stage),
so a photon's trace does not depend on the chunk it is in: the GPU traces
the whole chunk, and the oracle traces a random sample of its photons from
the GPU's own post-source records (the drop-in marx_create_photons of the
same tags), photon by photon to the parity bar of tests/parity.py.

- HETG + ACIS-S: 1e7 photons, 2e5 sampled;
- LETG + HRC-S: 5e7 photons of a 2' disk, 2e5 sampled;
- batched: the 4096 point sources with 1e2-1e5 photons each (5.9e7), in one
  graph-launched pass; 40 whole sources traced by the oracle from its own
  batch source (the batch draws are keyed by tag like the single source);
- differentiable: 1e7 weighted photons on 256 x 256 pixels; 20 runs of
  1e4 consecutive tags against the oracle's forward pass, and the whole
  image against the FP64 splat of the GPU's own per-photon state."""
import numpy as np
import pytest
import torch

from marx_amd import Photons
from marx_amd._lib import ImageSpec
from marx_amd.diff import DiffMirror, setup
from marx_amd.workloads import WORKLOADS, batched_acis_i, diff_mirror
from oracle.oracle import Oracle
from tests.parity import compare
from tests.test_gpu_diff import _splat64

SAMPLE = 200_000


def _sampled_parity(t, n, label, exposure):
    rng = np.random.default_rng(12)
    print(f"full size {label}: {n} photons", flush=True)
    ncol, ndet, _ = t.trace_device(n, tag_start=0, start_time=0.0)
    assert ncol == n
    full = Photons(n)
    t.download(full)
    final = full.active
    assert np.array_equal(final["tag"], np.arange(n, dtype=np.uint32))
    src = Photons(n)
    assert t.create_photons(src, n) == n
    pick = np.sort(rng.choice(n, SAMPLE, replace=False))
    rec = src.active[pick].copy()
    del src
    po = Photons(SAMPLE)
    po.load(rec)
    o = Oracle(t)
    o.reflect(po)
    o.diffract(po)
    o.detect(po)
    got = po.active
    got = got[np.argsort(got["tag"])]
    want = final[pick].copy()
    del full
    compare(want, got, max_flips=3, check_pixels=True, label=label)
    assert ((want["flags"] & 0xFF) == 0).sum() > 1000
    assert ndet > 0


@pytest.mark.gpu
def test_hetg_acis_s_at_1e7(gpu, tmp_path):
    n = 10_000_000
    t = WORKLOADS["hetg-acis-s"](str(tmp_path), photons=n, exposure=5.0e4)
    _sampled_parity(t, n, "hetg-acis-s 1e7", 5.0e4)


@pytest.mark.gpu
def test_letg_hrc_s_at_5e7(gpu, tmp_path):
    n = 50_000_000
    t = WORKLOADS["letg-hrc-s"](str(tmp_path), photons=n, exposure=5.0e4)
    _sampled_parity(t, n, "letg-hrc-s 5e7", 5.0e4)


@pytest.mark.gpu
def test_batch_of_4096_sources(gpu, tmp_path):
    t, dirs, counts = batched_acis_i(str(tmp_path))
    n = int(counts.sum())
    print(f"full size batch: {len(counts)} sources, {n} photons", flush=True)
    assert len(counts) == 4096 and counts.min() >= 100 and counts.max() <= 100_000
    exposure = 5.0e4
    t.trace_batch(dirs, counts, tag_start=0, start_time=0.0, exposure=exposure, use_graph=True)
    full = Photons(n)
    t.download(full)
    final = full.active
    off = np.concatenate([[0], np.cumsum(counts.astype(np.int64))])
    rng = np.random.default_rng(3)
    o = Oracle(t)
    for k in sorted(rng.choice(len(counts), 40, replace=False)):
        m = int(counts[k])
        po = Photons(m)
        o.create_batch(po, dirs[k:k + 1], counts[k:k + 1], int(off[k]), 0.0, exposure)
        o.reflect(po)
        o.detect(po)
        got = po.active
        got = got[np.argsort(got["tag"])]
        compare(final[off[k]:off[k + 1]].copy(), got, max_flips=2, check_pixels=True, label=f"source {k}")


@pytest.mark.gpu
def test_diff_at_1e7_on_256(gpu):
    t, params, target_params = diff_mirror()
    n, npix, pix = 10_000_000, 256, 0.004
    print(f"full size diff: {n} photons, {npix}^2", flush=True)
    dm = DiffMirror(t, photons=n, npix=npix, pixel_size=pix)
    img = torch.empty((npix, npix), dtype=torch.float32, device="cuda")
    dm.forward_into(params, img)
    gs = dm.state()
    # the whole image: the FP32 device sum against the FP64 splat of the
    # same per-photon state
    ref = _splat64(gs, npix, pix)
    assert np.abs(img.cpu().numpy().astype(np.float64) - ref).max() <= 1e-5 * ref.max()
    o = Oracle(t)
    spec = ImageSpec(npix, pix, 0.0, 0.0)
    cfg = setup(params)
    rng = np.random.default_rng(9)
    m = 10_000
    for start in sorted(rng.choice(n // m, 20, replace=False) * m):
        _, st = o.diff_forward(cfg, m, int(start), spec)
        sl = slice(int(start), int(start) + m)
        w = gs["w"][sl]
        assert np.array_equal(gs["shell"][sl], st["shell"])
        assert np.array_equal(w > 0, st["w"] > 0)
        np.testing.assert_allclose(w, st["w"], rtol=1e-9, atol=0)
        a = st["w"] > 0
        for k in ("y", "z"):
            np.testing.assert_allclose(gs[k][sl][a], st[k][a], rtol=0, atol=1e-9)
        alive = np.tile(a, 5)
        for k in ("dw", "dy", "dz"):
            g = gs[k].reshape(5, n)[:, sl].reshape(-1)
            np.testing.assert_allclose(g[alive], st[k][alive], rtol=1e-6, atol=1e-6 * np.abs(st[k][alive]).max(),
                                       err_msg=k)
