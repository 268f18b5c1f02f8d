"""ctypes mirror of include/marx_gpu_cfg.h, for tests that read the
initialised configuration (mxg_get_config) and its float/double arenas.
test_cpu_abi.py checks every size and field offset against the C compiler."""
import ctypes

import numpy as np

D, U, I = ctypes.c_double, ctypes.c_uint, ctypes.c_int
NUM_SHELLS, MAX_CHIPS, MAX_DRAKE = 4, 10, 8


class Bucket(ctypes.Structure):
    _fields_ = [("base", ctypes.c_ulonglong), ("off", U), ("nb", U), ("shift", U), ("ok", U)]


class LBucket(ctypes.Structure):
    _fields_ = [("x0", D), ("scale", D), ("off", U), ("nb", U), ("ok", U), ("pad", U)]


class Wfold(ctypes.Structure):
    _fields_ = [("n_arrays", U), ("arrays_off", U), ("e_alphas_off", U), ("pad", U)]


class ShellCfg(ctypes.Structure):
    _fields_ = [("conic_a", D * 2), ("conic_b", D * 2), ("conic_c", D * 2),
                ("conic_xmin", D * 2), ("conic_xmax", D * 2), ("to_osac", (D * 3) * 2),
                ("fwd", (D * 9) * 2), ("bwd", (D * 9) * 2), ("blur", D * 2), ("scat_factor", D * 2), ("tilt_deg", (D * 2) * 2),
                ("front_position", D), ("area_fraction", D), ("min_radius", D), ("max_radius", D),
                ("shutter_bitmap", U), ("num_open_shutters", U), ("corr_n", U), ("corr_e_off", U),
                ("corr_f_off", U), ("mirror_number", U), ("wfold", Wfold * 2), ("corr_bucket", Bucket),
                ("ea_bucket", Bucket * 2)]


class Strut(ctypes.Structure):
    _fields_ = [("xpos", D), ("half_width", D)]


class HrmaCfg(ctypes.Structure):
    _fields_ = [("enabled", I), ("ideal", I), ("use_blur", I), ("use_wfold", I), ("use_scale", I),
                ("use_struts", I), ("vignetting", D), ("cap_position", D), ("geometric_area", D),
                ("opt_n", U), ("opt_e_off", U), ("opt_beta_off", U), ("opt_delta_off", U),
                ("shell", ShellCfg * NUM_SHELLS), ("pre_struts", Strut * 2), ("cap_struts", Strut * 2),
                ("post_struts", Strut * 2), ("strut_cos", D), ("strut_sin", D), ("opt_bucket", Bucket)]


class Grating(ctypes.Structure):
    _fields_ = [("num_orders", U), ("num_energies", U), ("energies_off", U), ("cum_eff_off", U),
                ("orders_off", U), ("cum_sorted", U), ("dispersion_angle", D), ("period", D),
                ("dp_over_p", D), ("theta_blur", D), ("vig", D),
                ("disp_cos", D), ("disp_sin", D), ("e_bucket", Bucket)]


class Sectors(ctypes.Structure):
    _fields_ = [("n", U), ("off", U)]


class GratingCfg(ctypes.Structure):
    _fields_ = [("type", I), ("rowland", D * NUM_SHELLS), ("g", Grating * NUM_SHELLS),
                ("sectors", Sectors * NUM_SHELLS), ("has_fine", I), ("has_coarse", I),
                ("fine", Grating), ("coarse", Grating)]


class Chip(ctypes.Structure):
    _fields_ = [("id", I), ("pad", I), ("x_ll", D * 3), ("xhat", D * 3), ("yhat", D * 3),
                ("normal", D * 3), ("xlen", D), ("ylen", D), ("x_pixel_size", D), ("y_pixel_size", D),
                ("xpixel_offset", D), ("ypixel_offset", D), ("qe_n", U), ("qe_e_off", U), ("qe_off", U),
                ("filter_n", U), ("filter_e_off", U), ("filter_off", U), ("qe_bucket", Bucket)]


class QeTable(ctypes.Structure):
    _fields_ = [("n", U), ("e_off", U), ("eff_off", U), ("pad", U), ("bucket", Bucket)]


class Rect(ctypes.Structure):
    _fields_ = [("a", D * 3), ("e1", D * 3), ("e2", D * 3), ("normal", D * 3), ("len1", D), ("len2", D)]


class Contam(ctypes.Structure):
    _fields_ = [("num_layers", U), ("fxy_model", U), ("blocking", U), ("pad", U), ("x0", D), ("y0", D),
                ("tau0", D * 5), ("tau1", D * 5), ("mu_n", U * 5), ("mu_e_off", U * 5), ("mu_off", U * 5),
                ("fxy_off", U * 5)]


class DetectorCfg(ctypes.Structure):
    _fields_ = [("type", I), ("ideal", I), ("extend", I), ("num_chips", I), ("matrix", D * 9),
                ("offset", D * 3), ("focal_length", D), ("chip", Chip * MAX_CHIPS),
                ("frame_time", D), ("exposure_time", D), ("frame_transfer_time", D),
                ("mcp_qe", QeTable * 3), ("filter_qe", QeTable * 4), ("mcp_map", I * 4),
                ("shield_t", D), ("shield_l", D), ("shield_r", D), ("shield_x", D),
                ("shield_sl", D), ("shield_sr", D), ("shield_sl_gap", D), ("shield_sr_gap", D),
                ("shield_y_center", D), ("shield_z_center", D),
                ("g1_sigma", D), ("g1_xctr", D), ("g1_yctr", D), ("g1_wgt", D),
                ("g2_sigma", D), ("g2_xctr", D), ("g2_yctr", D), ("g2_wgt", D),
                ("l1_hwhm", D), ("l1_xctr", D), ("l1_yctr", D), ("l1_rmax", D), ("l1_wgt", D),
                ("u_pixel_size", D), ("v_pixel_size", D), ("pix_u", D * 4), ("pix_v", D * 4),
                ("pix_u0", D * 4), ("pix_v0", D * 4), ("pix_x0", D * 4), ("pix_y0", D * 4),
                ("use_drake", I), ("drake_n", I), ("drake_cr_width", D), ("drake", Rect * MAX_DRAKE),
                ("drake_c_n", U), ("drake_c_e_off", U), ("drake_c_beta_off", U), ("drake_c_delta_off", U),
                ("drake_cr_n", U), ("drake_cr_e_off", U), ("drake_cr_beta_off", U),
                ("drake_cr_delta_off", U), ("fef_on", I), ("fef_max_gaussians", I), ("fef_map_off", U),
                ("fef_desc_off", U), ("contam_on", I), ("contam_pad", I), ("contam", Contam * MAX_CHIPS)]


class SourceCfg(ctypes.Structure):
    _fields_ = [("type", I), ("spectrum", I), ("p", D * 3), ("p_normal", D * 3), ("distance", D),
                ("disk_theta", D), ("disk_theta_min", D), ("emin", D), ("emax", D), ("total_flux", D),
                ("file_n", U), ("file_e_off", U), ("file_cum_off", U), ("pad", U), ("mean_time", D),
                ("time_mode", I), ("dither_mode", I), ("uniform_tmax", D),
                ("ra_amp", D), ("dec_amp", D), ("roll_amp", D),
                ("ra_period", D), ("dec_period", D), ("roll_period", D),
                ("ra_phase", D), ("dec_phase", D), ("roll_phase", D),
                ("nominal_roll", D), ("aspect_blur", D), ("gauss_sigma", D), ("beta_core", D), ("beta_xpon", D),
                ("line_theta", D), ("line_cos_phi", D), ("line_sin_phi", D), ("line_alpha", D),
                ("line_normal", D * 3), ("image_n", U), ("image_nx", U), ("image_ny", U),
                ("image_off", U), ("image_xscale", D), ("image_yscale", D), ("file_bucket", LBucket)]


class Config(ctypes.Structure):
    _fields_ = [("hrma", HrmaCfg), ("grating", GratingCfg), ("det", DetectorCfg), ("src", SourceCfg)]


from marx_amd._lib import DiffCfg, DiffShell, EventCfg, EventColumns, ImageSpec  # noqa: E402

STRUCTS = {"mxg_bucket": Bucket, "mxg_lbucket": LBucket, "mxg_event_cfg": EventCfg, "mxg_event_columns": EventColumns, "mxg_diff_shell": DiffShell, "mxg_diff_cfg": DiffCfg, "mxg_image_spec": ImageSpec,"mxg_wfold": Wfold, "mxg_shell_cfg": ShellCfg, "mxg_strut": Strut, "mxg_hrma_cfg": HrmaCfg,
           "mxg_grating": Grating, "mxg_sectors": Sectors, "mxg_grating_cfg": GratingCfg,
           "mxg_chip": Chip, "mxg_qe_table": QeTable, "mxg_contam": Contam, "mxg_rect": Rect,
           "mxg_detector_cfg": DetectorCfg, "mxg_source_cfg": SourceCfg, "mxg_config": Config}


class ConfigView:
    """The configuration of an initialised Tracer, with its arenas as numpy
    arrays (views of the library's memory, valid until the next init)."""

    def __init__(self, tracer):
        cfg, fa, nf, da, nd = tracer.config_pointers()
        self.cfg = Config.from_address(cfg)
        self.fa = (np.ctypeslib.as_array((ctypes.c_float * nf).from_address(fa))
                   if fa and nf else np.zeros(0, np.float32))
        self.da = (np.ctypeslib.as_array((ctypes.c_double * nd).from_address(da))
                   if da and nd else np.zeros(0, np.float64))

    def f(self, off, n):
        return self.fa[off:off + n]

    def d(self, off, n):
        return self.da[off:off + n]
