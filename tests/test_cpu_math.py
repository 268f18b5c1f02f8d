"""The oracle's numerical building blocks, each checked against an independent
computation: numpy's searchsorted / textbook linear interpolation, the
closed forms of the quadratic, Fresnel's equations evaluated with numpy
complex arithmetic, the documented WFold table semantics, the WFold tables
the mirror loads against the test's own decoding of the files, and the
WFold draw as a rotation in the plane of incidence distributed as the table."""
import ctypes
import os

import numpy as np
import pytest

from oracle import oracle as O


def _sorted_f32(rng, n):
    x = np.unique(rng.uniform(0.1, 10.0, n).astype(np.float32))
    return x


# ---- binary search and linear interpolation (jdmath finterpo.c / dinterpo.c) ----

def test_binary_search_f_matches_searchsorted():
    """The result is the number of knots <= x, except that a knot equal to x
    met during the bisection is returned as its own index (finterpo.c:37-85)."""
    rng = np.random.default_rng(1)
    for n in (1, 2, 3, 17, 1426):
        xp = _sorted_f32(rng, n)
        for x in np.concatenate([rng.uniform(0.0, 11.0, 200).astype(np.float32), xp]):
            i = O.binary_search_f(x, xp)
            right = np.searchsorted(xp, x, side="right")
            if x in xp:
                assert i == right or (i < n and xp[i] == x)
            else:
                assert i == right


def _linear(x, xp, yp):
    """Textbook linear interpolation with linear extrapolation from the end
    segments, in float64."""
    xp = xp.astype(np.float64)
    yp = yp.astype(np.float64)
    if len(xp) == 1:
        return yp[0]
    i = int(np.clip(np.searchsorted(xp, x, side="right"), 1, len(xp) - 1))
    return yp[i - 1] + (yp[i] - yp[i - 1]) * (x - xp[i - 1]) / (xp[i] - xp[i - 1])


@pytest.mark.parametrize("n", [1, 2, 5, 339, 1426])
def test_interpolate_f_brute_force(n):
    rng = np.random.default_rng(n)
    xp = _sorted_f32(rng, n)
    yp = rng.uniform(-3.0, 3.0, len(xp)).astype(np.float32)
    xs = np.concatenate([rng.uniform(0.0, 11.0, 300), xp.astype(np.float64)]).astype(np.float32)
    for x in xs:
        got = O.interpolate_f(x, xp, yp)
        hit = np.nonzero(xp == x)[0]
        if len(hit):
            assert got == yp[hit[0]]          # exact-hit shortcut
        else:
            want = _linear(float(x), xp, yp)
            assert got == pytest.approx(want, rel=2e-6, abs=2e-6 * max(1.0, np.abs(yp).max()))


def test_interpolate_d_brute_force():
    rng = np.random.default_rng(7)
    xp = np.unique(rng.uniform(0.0, 5.0, 40))
    yp = rng.normal(size=len(xp))
    for x in rng.uniform(-1.0, 6.0, 500):
        assert O.interpolate_d(x, xp, yp) == pytest.approx(_linear(x, xp, yp), rel=1e-12, abs=1e-12)


def test_interpolate_n_fvector_equals_pointwise():
    """The merge-walk over energy-sorted abscissae (finterpo.c:134-193) gives
    the per-element interpolation of every column."""
    rng = np.random.default_rng(3)
    oldx = _sorted_f32(rng, 300)
    oldy = rng.uniform(0.0, 1.0, (5, len(oldx))).astype(np.float32)
    x = rng.uniform(0.0, 11.0, 2000).astype(np.float32)
    x[100:110] = oldx[20]                        # repeated exact hits
    x[200:210] = oldx[[0, 1, 2, 50, 51, -3, -2, -1, -1, 7]]
    x = np.sort(x)
    out = O.interpolate_n_fvector(x, oldx, oldy)
    for k in range(5):
        ref = np.array([O.interpolate_f(v, oldx, oldy[k]) for v in x], dtype=np.float32)
        np.testing.assert_allclose(out[k], ref, rtol=2e-6, atol=2e-7)


# ---- quadratic roots (jdmath qroot.c) ----

def test_quadratic_root_real_roots():
    rng = np.random.default_rng(11)
    for _ in range(2000):
        r1, r2 = rng.uniform(-100, 100, 2)
        a = rng.uniform(0.5, 2.0) * rng.choice([-1, 1])
        b, c = -a * (r1 + r2), a * r1 * r2
        st, rp, rm = O.quadratic_root(a, b, c)
        if b * b <= 4 * a * c:
            continue
        assert st == 1
        got = sorted([rp, rm])
        want = sorted([r1, r2])
        assert got == pytest.approx(want, rel=1e-9, abs=1e-9)


def test_quadratic_root_cancellation():
    """Roots of very different size: the small root keeps full relative
    accuracy (no b - sqrt(b^2 - 4ac) cancellation)."""
    a, b, c = 1.0, -1e8, 1.0          # roots 1e8 and 1e-8
    st, rp, rm = O.quadratic_root(a, b, c)
    assert st == 1
    small, big = sorted([rp, rm])
    assert small == pytest.approx(1e-8, rel=1e-14)
    assert big == pytest.approx(1e8, rel=1e-14)


def test_quadratic_root_degenerate():
    assert O.quadratic_root(0.0, 1.0, 1.0)[0] == -1
    st, rp, rm = O.quadratic_root(1.0, -2.0, 1.0)
    assert st == 1 and rp == rm == 1.0
    assert O.quadratic_root(1.0, 0.0, 1.0)[0] == 0      # complex pair


# ---- Fresnel reflectivity (marx reflect.c:39-77; Jackson eq. 7.39, 7.41) ----

def _fresnel(cos_i, beta, delta):
    n = complex(1.0 - delta, beta)
    sin2 = 1.0 - cos_i * cos_i
    root = np.sqrt(n * n - sin2)
    r_s = (cos_i - root) / (cos_i + root)
    r_p = (n * n * cos_i - root) / (n * n * cos_i + root)
    return 0.5 * (abs(r_s) ** 2 + abs(r_p) ** 2)


def test_reflectivity_normal_incidence():
    for beta, delta in ((1e-6, 1e-5), (3e-3, 2e-2), (0.0, 0.1)):
        n = complex(1.0 - delta, beta)
        assert O.reflectivity(1.0, beta, delta) == pytest.approx(abs((1 - n) / (1 + n)) ** 2, rel=1e-12)


def test_reflectivity_vs_fresnel_equations():
    rng = np.random.default_rng(5)
    for _ in range(2000):
        cos_i = rng.uniform(1e-4, 1.0)
        beta = 10 ** rng.uniform(-8, -2)
        delta = 10 ** rng.uniform(-6, -1)
        assert O.reflectivity(cos_i, beta, delta) == pytest.approx(_fresnel(cos_i, beta, delta),
                                                                   rel=1e-10, abs=1e-14)


def test_reflectivity_total_external_reflection():
    """No absorption and a grazing angle below the critical angle: R = 1."""
    delta = 1e-4
    crit = np.sqrt(2 * delta - delta * delta)       # sin of the critical grazing angle
    for s in np.linspace(0.05, 0.95, 10) * crit:
        assert O.reflectivity(s, 0.0, delta) == pytest.approx(1.0, abs=1e-12)
    assert O.reflectivity(0.5, 0.0, 0.0) == pytest.approx(0.0, abs=1e-30)   # vacuum


def test_reflectivity_falls_with_grazing_angle():
    vals = [O.reflectivity(np.sin(np.radians(a)), 2e-6, 3e-5) for a in np.linspace(0.2, 3.0, 30)]
    assert all(x > y for x, y in zip(vals, vals[1:]))


# ---- WFold scattering tables (wfold.c:304-369) on the real tables ----

@pytest.fixture(scope="module")
def mirror_tables():
    from marx_amd import Tracer
    from cfgview import ConfigView
    t = Tracer(GratingType="NONE", DetectorType="NONE")
    return t, ConfigView(t)


def _fold_theta(view, arr, p):
    """theta(p) of one fold array: 0 below p_min, (norm (1 - p))^expon above
    p_max, linear on the uniform p grid between."""
    e_alpha, p_min, delta_p, p_max, norm, expon, num, toff = arr
    if p < p_min:
        return 0.0
    if p > p_max:
        return (norm * (1.0 - p)) ** expon
    grid = p_min + delta_p * np.arange(int(num))
    return float(np.interp(p, grid, view.f(int(toff), int(num)).astype(np.float64)))


def test_wfold_table_semantics(mirror_tables):
    t, view = mirror_tables
    from oracle.oracle import Oracle, load
    o = Oracle(t)
    lib = load()
    rng = np.random.default_rng(2)
    for shell in range(4):
        for k in range(2):
            w = view.cfg.hrma.shell[shell].wfold[k]
            assert w.n_arrays > 1
            arrays = view.d(w.arrays_off, w.n_arrays * 8).reshape(w.n_arrays, 8)
            ea = view.d(w.e_alphas_off, w.n_arrays)
            assert np.all(np.diff(ea) > 0) and np.array_equal(ea, arrays[:, 0])
            wp = ctypes.addressof(w)
            for _ in range(60):
                j = rng.integers(1, w.n_arrays)
                e_alpha = rng.uniform(ea[j - 1], ea[j])
                energy = rng.uniform(0.3, 9.0)
                r = rng.uniform(0.0, 1.0)
                got = lib.orc_wfold_interp(ctypes.byref(o.ctx), wp, energy, e_alpha / energy, r)
                t0 = _fold_theta(view, arrays[j - 1], r)
                t1 = _fold_theta(view, arrays[j], r)
                want = t0 + (t1 - t0) * (e_alpha - ea[j - 1]) / (ea[j] - ea[j - 1])
                assert got == pytest.approx(want, rel=1e-6, abs=1e-12)


def _read_wfold_file(path):
    """The WFold binary format (wfold.c:110-170), decoded here: records of a
    big-endian int32 count n (0 or end of file ends the table), six float32
    (e_alpha, p_min, delta_p, p_max, norm, expon) and n float32 theta values;
    delta_p is re-derived as (p_max - p_min) / (n - 1)."""
    b = open(path, "rb").read()
    off, out = 0, []
    while off + 4 <= len(b):
        n = int(np.frombuffer(b, ">i4", 1, off)[0])
        if n == 0:
            break
        hdr = np.frombuffer(b, ">f4", 6, off + 4).astype(np.float64)
        theta = np.frombuffer(b, ">f4", n, off + 28).astype(np.float32)
        off += 28 + 4 * n
        e_alpha, p_min, _, p_max, norm, expon = hdr
        if n == 1:
            p_max, dp = p_min, 0.1
        else:
            dp = (p_max - p_min) / (n - 1)
        out.append(((e_alpha, p_min, dp, p_max, norm, expon, n), theta))
    assert off == len(b) or n == 0
    return out


def test_wfold_tables_are_the_files(mirror_tables):
    """What the mirror set-up loaded for each shell and surface is exactly
    the content of its WFold file (WFold_<P|H><shell>_File), decoded here."""
    from marx_amd._lib import DATA_DIR
    t, view = mirror_tables
    for shell, number in enumerate((1, 3, 4, 6)):
        for k, surf in enumerate("PH"):
            path = os.path.join(DATA_DIR, t.par.get_string(f"WFold_{surf}{number}_File"))
            recs = _read_wfold_file(path)
            w = view.cfg.hrma.shell[shell].wfold[k]
            assert w.n_arrays == len(recs) > 10, (number, surf)
            arrays = view.d(w.arrays_off, w.n_arrays * 8).reshape(w.n_arrays, 8)
            np.testing.assert_array_equal(view.d(w.e_alphas_off, w.n_arrays), [r[0][0] for r in recs])
            for j, (hdr, theta) in enumerate(recs):
                np.testing.assert_array_equal(arrays[j, :7], hdr, err_msg=f"{number}{surf} array {j}")
                np.testing.assert_array_equal(view.f(int(arrays[j, 7]), hdr[6]), theta)


def _fold_theta_grid(view, arr, r):
    """_fold_theta over an array of p values."""
    e_alpha, p_min, delta_p, p_max, norm, expon, num, toff = arr
    vals = view.f(int(toff), int(num)).astype(np.float64)
    mid = np.interp(r, p_min + delta_p * np.arange(int(num)), vals)
    with np.errstate(divide="ignore", invalid="ignore"):
        tail = (norm * (1.0 - r)) ** expon
    return np.where(r < p_min, 0.0, np.where(r > p_max, tail, mid))


def test_wfold_draw_rotates_in_the_plane_of_incidence():
    """The H-surface WFold scatter (wfold.c:334-369, applied as in hrma.c):
    with the P-surface factors 0, an unblurred mirror without struts, and the same
    photons traced with and without WFold, the only difference is a rotation
    of the reflected direction at the same H hit point, in the plane of the
    surface normal (n from the conic y^2 + z^2 = (a-1) x^2 + b x + c in the
    shell frame), by f_H theta(E sin graze, r) for r uniform.  The
    probability-integral transform of each photon's angle under its own
    table (randomized over theta's atom at 0) must be uniform; the same
    angles scaled by 0.8 or 1.25 (a scatter angle off by that factor) must
    not be."""
    from scipy import stats
    from marx_amd import Photons, Tracer, on_axis_photons
    from cfgview import ConfigView
    from oracle.oracle import Oracle
    zero_p = {f"P{k}ScatFactor": 0.0 for k in (1, 3, 4, 6)}
    common = dict(GratingType="NONE", DetectorType="NONE", HRMA_Ideal="no", HRMA_Use_Blur="no",
                  HRMA_Use_Struts="no", seed=21, **zero_p)
    n, energy = 150000, 1.49
    rec = on_axis_photons(n, energy)
    out = {}
    for wf in ("no", "yes"):
        t = Tracer(HRMA_Use_WFold=wf, **common)
        ph = Photons(n)
        ph.load(rec)
        Oracle(t).reflect(ph)
        a = ph.active
        out[wf] = a[np.argsort(a["tag"])]
    view = ConfigView(t)
    A, B = out["no"], out["yes"]
    ok = ((A["flags"] & 0xFF) == 0) & ((B["flags"] & 0xFF) == 0)
    A, B = A[ok], B[ok]
    np.testing.assert_array_equal(A["mirror_shell"], B["mirror_shell"])
    np.testing.assert_allclose(B["x"], A["x"], rtol=0, atol=1e-9)
    rng = np.random.default_rng(9)
    grid = (np.arange(4000) + 0.5) / 4000.0
    scales = (1.0, 0.8, 1.25)
    pits = {c: [] for c in scales}
    n_moved, expect_moved, var_moved = 0, 0.0, 0.0
    for shell in range(4):
        s = view.cfg.hrma.shell[shell]
        sel = A["mirror_shell"] == shell
        assert sel.sum() > 1000
        fwd = np.array(s.fwd[1]).reshape(3, 3)
        x = (np.asarray(A["x"][sel]) + np.array(s.to_osac[1])) @ fwd.T
        pa, pb = np.asarray(A["p"][sel]) @ fwd.T, np.asarray(B["p"][sel]) @ fwd.T
        nrm = np.stack([(s.conic_a[1] - 1) * x[:, 0] + 0.5 * s.conic_b[1], -x[:, 1], -x[:, 2]], axis=1)
        nrm /= np.linalg.norm(nrm, axis=1)[:, None]
        cr = np.cross(pa, pb)
        ang = np.arctan2(np.linalg.norm(cr, axis=1), np.sum(pa * pb, axis=1))
        moved = ang > 1e-12
        # rotation axis perpendicular to the normal: pb stays in the plane of (pa, n)
        assert np.abs(np.sum(cr[moved] * nrm[moved], axis=1)).max() < 1e-9 * ang[moved].max()
        sin_graze = np.abs(np.sum(pa * nrm, axis=1))
        f_h = s.scat_factor[1]
        assert f_h > 1.0 and s.scat_factor[0] == 0.0
        w = view.cfg.hrma.shell[shell].wfold[1]
        if True:
            arrays = view.d(w.arrays_off, w.n_arrays * 8).reshape(w.n_arrays, 8)
            ea = arrays[:, 0]
            theta = np.array([_fold_theta_grid(view, arr, grid) for arr in arrays])
            e_alpha = energy * sin_graze
            inside = (e_alpha > ea[0]) & (e_alpha < ea[-1])
            assert inside.mean() > 0.99
            j = np.searchsorted(ea, e_alpha[inside])
            wgt = (e_alpha[inside] - ea[j - 1]) / (ea[j] - ea[j - 1])
            obs = ang[inside] / f_h
            for i in range(len(obs)):
                th = np.maximum.accumulate((1 - wgt[i]) * theta[j[i] - 1] + wgt[i] * theta[j[i]])
                # F(0): the chance of no scatter (r below p_min)
                f0 = np.searchsorted(th, 1e-15, "right") / len(grid)
                expect_moved += 1.0 - f0
                var_moved += f0 * (1.0 - f0)
                if obs[i] <= 1e-12:        # unscattered: rounding of the zero P rotation only
                    continue
                n_moved += 1
                for c in scales:
                    o = c * obs[i]
                    lo = np.searchsorted(th, o * (1 - 1e-9), "left")
                    hi = np.searchsorted(th, o * (1 + 1e-9), "right")
                    u = (lo + rng.uniform() * max(hi - lo, 1)) / len(grid)
                    pits[c].append((u - f0) / (1.0 - f0))
    # the scattered fraction is 1 - F(0), and the angles of the scattered
    # photons follow theta given theta > 0
    assert abs(n_moved - expect_moved) < 5 * np.sqrt(var_moved) + 1, (n_moved, expect_moved)
    pv = {c: stats.kstest(pits[c], "uniform").pvalue for c in scales}
    assert n_moved > 1500
    assert pv[1.0] > 1e-3 and pv[0.8] < 1e-8 and pv[1.25] < 1e-8, pv
