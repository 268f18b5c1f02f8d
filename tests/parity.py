"""Attribute-by-attribute comparison of two photon lists (GPU vs oracle).

Positions must agree within 1e-9 mm and directions within 1e-12 for every
photon whose categorical path agrees; categorical fields (flags, shell,
order, chip, region, support orders) must be equal.  A categorical mismatch
is tolerated only as a counted threshold flip: the two codes differ in the
last bits of sin/cos/log/pow (glibc vs the device math library), which can
move a comparison that sits within ~1e-12 of its threshold.
"""
import numpy as np

POS_TOL_MM = 1e-9
DIR_TOL = 1e-12


def compare(gpu, orc, max_flips=0, check_pixels=True, label=""):
    assert len(gpu) == len(orc)
    cat_fields = ["flags", "mirror_shell", "order", "ccd_num", "detector_region"]
    same = np.ones(len(gpu), dtype=bool)
    for f in cat_fields:
        same &= gpu[f] == orc[f]
    same &= np.all(gpu["support_orders"] == orc["support_orders"], axis=1)
    flips = int((~same).sum())
    dx = np.abs(gpu["x"][same] - orc["x"][same]).max() if same.any() else 0.0
    dp = np.abs(gpu["p"][same] - orc["p"][same]).max() if same.any() else 0.0
    stats = {"n": len(gpu), "flips": flips, "max_dx_mm": float(dx), "max_dp": float(dp)}
    if check_pixels:
        for f in ["y_pixel", "z_pixel", "u_pixel", "v_pixel", "pi"]:
            a = gpu[f][same].astype(np.float32)
            b = orc[f][same].astype(np.float32)
            ulp = np.abs(a.view(np.int32).astype(np.int64) - b.view(np.int32).astype(np.int64))
            stats[f + "_ulp_max"] = int(ulp.max()) if len(ulp) else 0
            stats[f + "_ulp_count"] = int((ulp > 0).sum())
    print(label, stats)
    assert flips <= max_flips, f"{label}: {flips} categorical mismatches: {stats}"
    assert dx <= POS_TOL_MM, f"{label}: position mismatch {dx} mm"
    assert dp <= DIR_TOL, f"{label}: direction mismatch {dp}"
    assert np.array_equal(gpu["energy"], orc["energy"])
    assert np.array_equal(gpu["tag"], orc["tag"])
    return stats
