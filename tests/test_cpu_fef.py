"""The ACIS FEF (acis_fef.c): PHA and PI of detected ACIS photons from the
Gaussian mixture of their chip region at their energy.  The reference's
CALDB FEF is missing from its data, so the tests write synthetic FEF
files (tests/fefgen.py) and pin the oracle by what the model fixes:

- the PHA of photons of one region at a grid energy is distributed as
  max(0, sum_k A_k exp(-(x - c_k)^2 / 2 sigma_k^2)) over x >= 0, truncated
  to integers (chi-square): this holds for positive mixtures (choice by
  area above 0, retry below 0), for a negative-amplitude notch (the
  rejection step) and for a component centred below 0 (its tail sampler);
- PI = the channel->energy interpolation of PHA - U, so it lies in
  [(PHA - 1) e, PHA e] for the file's linear gain e;
- the region is chosen by the 32-pixel cell of the chip pixel; photons off
  the chip are undetected unless DetExtendFlag clamps them;
- between grid energies the Gaussians are interpolated linearly."""
import numpy as np
import pytest
from scipy import special, stats

from marx_amd import Photons, Tracer
from marx_amd._lib import ATTR_DTYPE
from oracle.oracle import Oracle
from tests.cfgview import ConfigView
from tests.fefgen import ENERGIES, EV_PER_CHANNEL, gaussians, write_fef

SIG = 0.4246609001440095


@pytest.fixture
def acis_s(tmp_path):
    path = write_fef(str(tmp_path / "fef.fits"), range(4, 10))
    t = Tracer(seed=29, GratingType="NONE", DetectorType="ACIS-S", DitherModel="NONE", DetIdeal=True,
               ACISFEFFile=path)
    return t


def _hits(t, chip_id, px, py, energy, tag0=0):
    """Records whose rays hit chip chip_id at pixels (px, py), energy given."""
    d = ConfigView(t).cfg.det
    g = [d.chip[k] for k in range(d.num_chips) if d.chip[k].id == chip_id][0]
    n = len(px)
    P = (np.array(g.x_ll)[None, :] + (np.asarray(px) * g.x_pixel_size)[:, None] * np.array(g.xhat)[None, :]
         + (np.asarray(py) * g.y_pixel_size)[:, None] * np.array(g.yhat)[None, :])
    rec = np.zeros(n, ATTR_DTYPE)
    rec["x"] = P + np.array([5.0, 0.0, 0.0])
    rec["p"] = (-1.0, 0.0, 0.0)
    rec["energy"] = energy
    rec["tag"] = np.arange(tag0, tag0 + n)
    rec["mirror_shell"] = 0
    return rec


def _detect(t, rec):
    p = Photons(len(rec))
    p.load(rec)
    p.pt.history = 0x1 | 0x2 | 0x4 | 0x8 | 0x10 | 0x20000    # energy, time, x, p, tag, mirror
    Oracle(t).detect(p)
    a = p.active
    return a[np.argsort(a["tag"])]


def _density_bins(e, region, edges):
    """Integrals of max(0, sum_k A_k exp(-z^2/2)) over the bins, normalised
    over x >= 0."""
    gs = [(SIG * fwhm, pos, ampl) for fwhm, pos, ampl in gaussians(region, e)]
    x = np.linspace(0.0, edges[-1] + 400.0, 400001)
    f = np.zeros_like(x)
    for s, c, a in gs:
        f += a * np.exp(-0.5 * ((x - c) / s) ** 2)
    f = np.maximum(f, 0.0)
    cdf = np.concatenate([[0.0], np.cumsum(0.5 * (f[1:] + f[:-1]) * np.diff(x))])
    return np.diff(np.interp(edges, x, cdf)) / cdf[-1]


@pytest.mark.parametrize("region,ei", [(0, 12), (1, 14), (2, 6), (3, 18)])
def test_pha_follows_the_region_mixture(acis_s, region, ei):
    t = acis_s
    e = float(np.float32(ENERGIES[ei]))
    rng = np.random.default_rng(region)
    n = 40000
    qx, qy = divmod(region, 2)
    px = rng.uniform(qx * 512 + 1, qx * 512 + 511, n)
    py = rng.uniform(qy * 512 + 1, qy * 512 + 511, n)
    a = _detect(t, _hits(t, 7, px, py, e))
    ok = (a["flags"] & 0xFF) == 0
    # PHA 0 gives PI from PHA - U < 0, below the gain table: some of those
    # are undetected (acis_fef.c:1068-1070), so the comparison starts at 1
    assert ok.mean() > 0.9
    pha = a["pulse_height"][ok].astype(np.int64)
    lo, hi = np.percentile(pha, [0.1, 99.9])
    edges = np.unique(np.linspace(max(lo - 1, 1), hi + 2, 41).astype(int)).astype(float)
    inside = (pha >= edges[0]) & (pha < edges[-1])
    counts = np.histogram(pha[inside], edges)[0]
    # PHA = floor(x): the PHA bin [k, k + 1) of the density
    expect = _density_bins(e, region, edges)
    expect = expect / expect.sum() * counts.sum()
    keep = expect > 5
    chi2 = ((counts[keep] - expect[keep]) ** 2 / expect[keep]).sum()
    assert stats.chi2.sf(chi2, keep.sum() - 1) > 1e-4, (region, chi2, keep.sum())
    # PI: the channel->energy map of PHA - U (linear gain in the file)
    pi = a["pi"][ok].astype(np.float64)
    gain = EV_PER_CHANNEL / 1000.0
    assert np.all(pi >= (pha - 1) * gain * (1 - 1e-6) - 1e-6) and np.all(pi <= pha * gain * (1 + 1e-6) + 1e-6)


def test_regions_and_the_chip_edge(acis_s, tmp_path):
    t = acis_s
    e = float(np.float32(ENERGIES[15]))
    n = 20000
    rng = np.random.default_rng(3)
    means = []
    for region in (0, 3):
        qx, qy = divmod(region, 2)
        a = _detect(t, _hits(t, 6, rng.uniform(qx * 512 + 40, qx * 512 + 470, n),
                             rng.uniform(qy * 512 + 40, qy * 512 + 470, n), e))
        means.append(a["pulse_height"][(a["flags"] & 0xFF) == 0].mean())
    # region 3's main peak sits 3 percent higher (the synthetic file)
    c = 1000.0 * e / EV_PER_CHANNEL
    assert means[1] - means[0] == pytest.approx(0.03 * c * (1.0 / 1.25), rel=0.15)
    # between grid energies: the main peak moves linearly
    e_mid = 0.5 * (float(ENERGIES[15]) + float(ENERGIES[16]))
    a = _detect(t, _hits(t, 6, rng.uniform(40, 470, n), rng.uniform(40, 470, n), e_mid))
    m_mid = a["pulse_height"][(a["flags"] & 0xFF) == 0].mean()
    a16 = _detect(t, _hits(t, 6, rng.uniform(40, 470, n), rng.uniform(40, 470, n), float(np.float32(ENERGIES[16]))))
    m16 = a16["pulse_height"][(a16["flags"] & 0xFF) == 0].mean()
    assert m_mid == pytest.approx(0.5 * (means[0] + m16), rel=2e-3)


def test_fef_is_off_without_a_file(tmp_path):
    t = Tracer(seed=29, GratingType="NONE", DetectorType="ACIS-S", DitherModel="NONE", DetIdeal=True)
    assert ConfigView(t).cfg.det.fef_on == 0
    a = _detect(t, _hits(t, 7, np.full(100, 300.5), np.full(100, 300.5), 1.5))
    assert ((a["flags"] & 0xFF) == 0).all()
    assert (a["pulse_height"] == 0).all() and (a["pi"] == 0).all()
