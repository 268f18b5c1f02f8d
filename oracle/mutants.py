"""Planted mistakes: each pin of the oracle must fail when the mistake it
guards against is made.  TEST INFRASTRUCTURE ONLY (CPU).

For every mutant below, a scratch copy of the oracle (or of the library's
host set-up code, for mistakes in the set-up that the oracle reads) gets one
textual change, is built there, and the named pin tests run against that
build (MARX_ORACLE_LIB / MARX_GPU_LIB); a mutant "is caught" when the tests
fail.  The unchanged builds must pass the same tests first.

    python -m oracle.mutants [--only NAME ...]      (prints a markdown table)
"""
import argparse
import os
import shutil
import subprocess
import sys
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PM = "tests/test_cpu_pins_mirror.py"
PS = "tests/test_cpu_pins_source.py"
PD = "tests/test_cpu_pins_detector.py"
PC = "tests/test_cpu_contam.py"

# name: (where, file, old text, new text, tests, what the mistake is)
MUTANTS = {
    "blur_x2": ("oracle", "orc_hrma.c", "phi = blur * (1.0 / 3600.0 * PI / 180.0);",
                "phi = 2.0 * blur * (1.0 / 3600.0 * PI / 180.0);",
                [PM + "::test_mirror_blur_is_a_gaussian_tilt_of_the_normal"], "mirror normal blur x2"),
    "blur_about_normal": ("oracle", "orc_hrma.c", "*n = orc_v_rotate_unit_vector (*n, perp, phi);",
                          "*n = orc_v_rotate_unit_vector (*n, *n, phi);",
                          [PM + "::test_mirror_blur_is_a_gaussian_tilt_of_the_normal"],
                          "normal turned about itself (no blur)"),
    "struts_off": ("oracle", "orc_hrma.c", "if (0 == orc_strut_shadow (h, at->x, at->p, s, num)) return 0;",
                   "if (1) return 0;", [PM + "::test_strut_shadows_follow_the_strut_bands"],
                   "strut shadows removed"),
    # both cap struts: they lie 50 mm apart on the same bands, so one strut
    # alone shadows what the other would
    "strut_width": ("library", "host/hrma_init.c",
                    "h->cap_struts[0].half_width = 0.5 * 0.75 * 25.4;\n"
                    "   h->cap_struts[1].xpos = -0.5 * 1.965 * 25.4; h->cap_struts[1].half_width = 0.5 * 0.75 * 25.4;",
                    "h->cap_struts[0].half_width = 0.5 * 0.5 * 25.4;\n"
                    "   h->cap_struts[1].xpos = -0.5 * 1.965 * 25.4; h->cap_struts[1].half_width = 0.5 * 0.5 * 25.4;",
                    [PM + "::test_strut_shadows_follow_the_strut_bands"], "cap struts 1/2 inch wide instead of 3/4"),
    "tilt_sign": ("library", "host/hrma_init.c", "double el = -el_deg * (PI / 180.0);",
                  "double el = el_deg * (PI / 180.0);",
                  [PM + "::test_tilt_matrices_are_the_documented_rotations",
                   PM + "::test_one_surface_tilt_moves_the_image"], "elevation tilt with the wrong sign"),
    "fwd_bwd_swap": ("oracle", "orc_hrma.c", "at->p = orc_m3_vmul (s->fwd[0], at->p);",
                     "at->p = orc_m3_vmul (s->bwd[0], at->p);",
                     [PM + "::test_one_surface_tilt_moves_the_image"],
                     "paraboloid direction turned by the inverse tilt"),
    "dither_amp_x2": ("oracle", "orc_source.c", "ra = d->ra = (float) (s->ra_amp * sin",
                      "ra = d->ra = (float) (2.0 * s->ra_amp * sin",
                      [PS + "::test_internal_dither_is_the_lissajous_pattern"], "INTERNAL dither amplitude x2"),
    "aspect_blur_x2": ("oracle", "orc_source.c", "delta_ra = s->aspect_blur * orc_gaussian (&st);",
                       "delta_ra = 2.0 * s->aspect_blur * orc_gaussian (&st);",
                       [PS + "::test_aspect_blur_scatters_the_pointing"], "aspect blur x2"),
    "spectrum_sqrt": ("oracle", "orc_source.c", "return orc_interpolate_d (orc_uniform (st), c->da + s->file_cum_off",
                      "return orc_interpolate_d (sqrt (orc_uniform (st)), c->da + s->file_cum_off",
                      [PS + "::test_file_spectrum_follows_its_rectangle_rule"], "FILE inverse CDF warped to sqrt(U)"),
    "flat_shift": ("oracle", "orc_source.c", "return emin + de * orc_uniform (st);",
                   "return emin + de * orc_uniform (st) * orc_uniform (st);",
                   [PS + "::test_flat_spectrum_is_uniform"], "FLAT energies not uniform"),
    "poisson_x2": ("oracle", "orc_source.c", "t = s->mean_time * ((double) S * (1.0 / 4294967296.0));",
                   "t = 2.0 * s->mean_time * ((double) S * (1.0 / 4294967296.0));",
                   [PS + "::test_poisson_arrival_gaps"], "Poisson gap mean x2"),
    "disk_linear": ("oracle", "orc_source.c", "rnd = s->disk_theta * sqrt (x0 + x1 * orc_uniform (st));",
                    "rnd = s->disk_theta * (x0 + x1 * orc_uniform (st));",
                    [PS + "::test_disk_profile"], "DISK uniform in theta, not in area"),
    "streak_row": ("oracle", "orc_detect.c", "at->z_pixel = (float) (1.0 + 1022.0 * orc_uniform (st));",
                   "at->z_pixel = (float) (1022.0 * orc_uniform (st));",
                   [PD + "::test_streaked_photons_keep_their_column"], "streak row over [0, 1022)"),
    "uvis_off": ("oracle", "orc_detect.c", "   else\n     q = d->filter_qe + mcp_id;\n   if (q->n != 0)",
                 "   else\n     q = d->filter_qe + mcp_id;\n   if (0)",
                 [PD + "::test_hrc_i_detects_mcp_qe_times_uvis",
                  PD + "::test_hrc_s_detects_mcp_qe_times_the_region_transmission"], "HRC UVIS draw removed"),
    "region_swap": ("oracle", "orc_detect.c", "else if (y < d->shield_sl) *r = (z >= d->shield_t) ? 0 : 1;",
                    "else if (y < d->shield_sl) *r = (z >= d->shield_t) ? 1 : 0;",
                    [PD + "::test_hrc_s_detects_mcp_qe_times_the_region_transmission"],
                    "HRC-S shield regions 0 and 1 swapped"),
    "pha_width": ("oracle", "orc_detect.c", "energy = energy * (1.0 + 0.424661 * orc_gaussian (st));",
                  "energy = energy * (1.0 + 0.5 * orc_gaussian (st));",
                  [PD + "::test_hrc_pha_follows_its_gain_law"], "HRC PHA width 0.5 instead of 0.424661"),
    "contam_off": ("oracle", "orc_detect.c", "               qe_contam = compute_contamination (c, d->contam",
                   "               qe_contam = 1.0 + 0.0 * compute_contamination (c, d->contam",
                   [PC + "::test_detected_fraction_is_qe_times_the_transmission"],
                   "ACIS contamination not applied"),
    "contam_rows": ("oracle", "orc_detect.c", "return pow (fabs ((y - y_0) / (964.0 - y_0)), 4.5);",
                    "return pow (fabs ((y - y_0) / (964.0 - y_0)), 5.5);",
                    [PC + "::test_detected_fraction_is_qe_times_the_transmission"],
                    "ACIS-S row profile with the lower exponent above the middle row"),
    "contam_time": ("library", "host/detector_init.c", "   double t = _Marx_TStart_MJDsecs, w;",
                    "   double t = 0.0, w;", [PC + "::test_contam_tables_are_read"],
                    "contamination taken at MJDref, not at the observation start"),
    "chip_shift": ("library", "host/detector_init.c", "   ofs = mxg_v_sum (aim, origin);\n",
                   "   ofs = mxg_v_sum (aim, origin);\n   ofs.y += 0.5;\n",
                   [PD + "::test_acis_chip_corners_are_the_caldb_geometry",
                    PD + "::test_on_axis_image_lands_on_the_documented_aimpoint"], "every chip moved by 0.5 mm"),
}


def _replace_once(path, old, new):
    s = open(path).read()
    if s.count(old) != 1:
        raise SystemExit(f"{path}: the text of the mutant occurs {s.count(old)} times")
    open(path, "w").write(s.replace(old, new))


def _build(where, rel, old, new, tmp):
    """A built scratch copy with the change (none if rel is None); returns
    the environment that points the tests at it."""
    shutil.copytree(os.path.join(ROOT, "include"), os.path.join(tmp, "include"))
    if where == "oracle":
        src = os.path.join(tmp, "oracle")
        shutil.copytree(os.path.join(ROOT, "oracle"), src,
                        ignore=shutil.ignore_patterns("_build*", "*.so", "__pycache__", "_ref"))
        if rel:
            _replace_once(os.path.join(src, rel), old, new)
        subprocess.run(["make", "-C", src, "-j8", "liboracle.so"], check=True, capture_output=True)
        return {"MARX_ORACLE_LIB": os.path.join(src, "liboracle.so")}
    src = os.path.join(tmp, "marx_amd", "csrc")
    shutil.copytree(os.path.join(ROOT, "marx_amd", "csrc"), src,
                    ignore=shutil.ignore_patterns("_build", "__pycache__"))
    # the library finds its calibration data at ../data from its own file
    os.symlink(os.path.join(ROOT, "marx_amd", "data"), os.path.join(tmp, "marx_amd", "data"))
    if rel:
        _replace_once(os.path.join(src, rel), old, new)
    subprocess.run(["make", "-C", src, "-j8"], check=True, capture_output=True)
    return {"MARX_GPU_LIB": os.path.join(tmp, "marx_amd", "lib", "libmarx_gpu.so")}


def _run_tests(env, tests):
    e = dict(os.environ)
    e.update(env)
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider"] + tests,
                       cwd=ROOT, env=e, capture_output=True, text=True)
    summary = [l for l in r.stdout.splitlines() if " passed" in l or " failed" in l or " error" in l]
    return r.returncode == 0, (summary or [""])[-1].strip()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--only", nargs="*", default=None)
    args = ap.parse_args()
    names = args.only or list(MUTANTS)
    rows = []
    baseline = {}
    for name in names:
        where, rel, old, new, tests, what = MUTANTS[name]
        if where not in baseline:
            with tempfile.TemporaryDirectory() as tmp:
                env = _build(where, None, None, None, tmp)
                all_tests = sorted({t for m in MUTANTS.values() if m[0] == where for t in m[4]})
                ok, tail = _run_tests(env, all_tests)
            baseline[where] = ok
            print(f"unchanged {where} build: {'pass' if ok else 'FAIL'} ({tail})", flush=True)
            if not ok:
                raise SystemExit(f"the pins fail on the unchanged {where} build")
        with tempfile.TemporaryDirectory() as tmp:
            env = _build(where, rel, old, new, tmp)
            ok, tail = _run_tests(env, tests)
        rows.append((name, where, f"{rel}", what, "caught" if not ok else "NOT CAUGHT", tail))
        print(f"{name}: {'caught' if not ok else 'NOT CAUGHT'} ({tail})", flush=True)
    print()
    print("| mutant | planted in | mistake | result |")
    print("|---|---|---|---|")
    for name, where, rel, what, res, _ in rows:
        print(f"| {name} | {where} `{rel}` | {what} | {res} |")
    return 0 if all(r[4] == "caught" for r in rows) else 1


if __name__ == "__main__":
    sys.exit(main())
