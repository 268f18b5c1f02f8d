"""ctypes binding of libmarx_gpu (include/marx_gpu.h).

The library is loaded from ``marx_amd/lib/libmarx_gpu.so`` (built in-tree by
``__graft_entry__.build()`` / ``make -C marx_amd/csrc``).  There is no
fallback: a missing library raises ``ImportError`` and a missing GPU makes
every launching entry point return -1 with ``mxg_last_error()`` set.
"""
import ctypes
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
# MARX_GPU_LIB names another build of the library (used by the oracle's
# planted-mistake runs of the host set-up code); the default is the in-tree
# build.
LIB_PATH = os.environ.get("MARX_GPU_LIB") or os.path.join(PKG_DIR, "lib", "libmarx_gpu.so")
DATA_DIR = os.path.join(PKG_DIR, "data")
PAR_FILE = os.path.join(DATA_DIR, "marx.par")

# Marx_Photon_Attr_Type (marx/libsrc/marx.h:51-103), 136 bytes.
ATTR_DTYPE = np.dtype({
    "names": ["energy", "x", "p", "arrival_time", "flags", "y_pixel", "z_pixel",
              "u_pixel", "v_pixel", "dither", "pi", "pulse_height", "mirror_shell",
              "ccd_num", "detector_region", "order", "support_orders", "tag"],
    "formats": ["<f8", ("<f8", 3), ("<f8", 3), "<f8", "<u4", "<f4", "<f4", "<f4", "<f4",
                ("<f4", 6), "<f4", "<i2", "<u4", "i1", "i1", "i1", ("i1", 4), "<u4"],
    "offsets": [0, 8, 32, 56, 64, 68, 72, 76, 80, 84, 108, 112, 116, 120, 121, 122, 123, 128],
    "itemsize": 136,
})

PHOTON_UNDETECTED = 0x01
PHOTON_UNREFLECTED = 0x02
PHOTON_UNDIFFRACTED = 0x04
PHOTON_MISSED_DETECTOR = 0x08
PHOTON_MIRROR_VBLOCKED = 0x10
PHOTON_DRAKE_BLOCKED = 0x20
PHOTON_GRATING_VBLOCKED = 0x40
BAD_PHOTON_MASK = 0xFF
PHOTON_DRAKE_REFLECTED = 0x100
PHOTON_ACIS_STREAKED = 0x200

MARX_MIRROR_SHELL_OK = 0x00000800
MARX_ORDER_OK = 0x00100000
MARX_DET_NUM_OK = 0x00000100


class PhotonList(ctypes.Structure):
    """Marx_Photon_Type (marx/libsrc/marx.h:108-151)."""
    _fields_ = [
        ("source_distance", ctypes.c_double),
        ("attributes", ctypes.c_void_p),
        ("n_photons", ctypes.c_uint),
        ("max_n_photons", ctypes.c_uint),
        ("total_time", ctypes.c_double),
        ("start_time", ctypes.c_double),
        ("sorted_index", ctypes.POINTER(ctypes.c_uint)),
        ("sorted_energies", ctypes.POINTER(ctypes.c_double)),
        ("num_sorted", ctypes.c_uint),
        ("history", ctypes.c_ulong),
        ("tag_start", ctypes.c_uint),
    ]


class SoaView(ctypes.Structure):
    """mxg_soa_view: device pointers of the resident columns."""
    _fields_ = [
        ("n", ctypes.c_uint), ("capacity", ctypes.c_uint),
        ("energy", ctypes.c_void_p), ("x", ctypes.c_void_p * 3), ("p", ctypes.c_void_p * 3),
        ("arrival_time", ctypes.c_void_p), ("flags", ctypes.c_void_p),
        ("y_pixel", ctypes.c_void_p), ("z_pixel", ctypes.c_void_p),
        ("u_pixel", ctypes.c_void_p), ("v_pixel", ctypes.c_void_p),
        ("dither", ctypes.c_void_p * 6), ("pi", ctypes.c_void_p),
        ("pulse_height", ctypes.c_void_p), ("mirror_shell", ctypes.c_void_p),
        ("ccd_num", ctypes.c_void_p), ("detector_region", ctypes.c_void_p),
        ("order", ctypes.c_void_p), ("support_orders", ctypes.c_void_p * 4),
        ("tag", ctypes.c_void_p),
    ]


class Schedule(ctypes.Structure):
    """mxg_schedule: blocks per CU of the source, paraboloid, reflection,
    grating and detector kernels, the reflection kernel's waves per SIMD, and
    the persistent pass's tile (0: the staged pass) and blocks per CU."""
    _fields_ = [("blocks_per_cu", ctypes.c_uint * 5), ("reflect_waves", ctypes.c_uint),
                ("persistent_tile", ctypes.c_uint), ("persistent_blocks_per_cu", ctypes.c_uint)]

    KERNELS = ("source", "enter", "reflect", "grating", "detect")

    def as_dict(self):
        d = {k: int(v) for k, v in zip(self.KERNELS, self.blocks_per_cu)}
        d["reflect_waves"] = int(self.reflect_waves)
        d["persistent_tile"] = int(self.persistent_tile)
        d["persistent_blocks_per_cu"] = int(self.persistent_blocks_per_cu)
        return d


class ImageSpec(ctypes.Structure):
    """mxg_image_spec: npix x npix pixels of pixel_size mm centred on
    (center_y, center_z) mm in the focal plane."""
    _fields_ = [("npix", ctypes.c_uint), ("pixel_size", ctypes.c_double),
                ("center_y", ctypes.c_double), ("center_z", ctypes.c_double)]


class DiffShell(ctypes.Structure):
    """mxg_diff_shell"""
    _fields_ = [("fwd", (ctypes.c_double * 9) * 2), ("bwd", (ctypes.c_double * 9) * 2),
                ("dfwd", ((ctypes.c_double * 9) * 2) * 2), ("dbwd", ((ctypes.c_double * 9) * 2) * 2),
                ("to_osac", (ctypes.c_double * 3) * 2), ("front_position", ctypes.c_double)]


DIFF_NPAR = 5


class DiffCfg(ctypes.Structure):
    """mxg_diff_cfg: the shell geometry at given misalignment parameters."""
    _fields_ = [("params", ctypes.c_double * (4 * DIFF_NPAR)), ("shell", DiffShell * 4)]


PPhotonList = ctypes.POINTER(PhotonList)

# pixel adjustments of mxg_event_cfg.pix_adj (marx2fits --pixadj)
PIX_ADJ = {"NONE": 0, "RANDOMIZE": 1, "EDSER": 2, "EXACT": 3}


class EventCfg(ctypes.Structure):
    """mxg_event_cfg: the marx2fits settings of mxg_event_coords."""
    _fields_ = [("focal_length", ctypes.c_double), ("det_offset", ctypes.c_double * 3),
                ("roll_nom", ctypes.c_double), ("exposure_time", ctypes.c_double),
                ("frame_transfer_time", ctypes.c_double), ("total_time", ctypes.c_double),
                ("seed", ctypes.c_ulonglong), ("pix_adj", ctypes.c_int), ("dither_used", ctypes.c_int)]


class EventColumns(ctypes.Structure):
    """mxg_event_columns: host arrays (NULL ones skipped)."""
    _fields_ = [("tag", ctypes.c_void_p), ("ccd_id", ctypes.c_void_p), ("tdetx", ctypes.c_void_p),
                ("tdety", ctypes.c_void_p), ("detx", ctypes.c_void_p), ("dety", ctypes.c_void_p),
                ("x", ctypes.c_void_p), ("y", ctypes.c_void_p), ("fltgrade", ctypes.c_void_p)]


class DumpFile(ctypes.Structure):
    """Marx_Dump_File_Type (marx/libsrc/marx.h:161-173)."""


DumpFile._fields_ = [("fp", ctypes.c_void_p), ("type", ctypes.c_char), ("colname", ctypes.c_char * 16),
                     ("num_rows", ctypes.c_int), ("num_cols", ctypes.c_int),
                     ("next", ctypes.POINTER(DumpFile))]

# history bits of Marx_Photon_Type (marx/libsrc/marx.h:128-147): every column
WRITE_ALL = 0x01FFFFFF

# (name, restype, argtypes) of every exported C-ABI symbol of include/marx_gpu.h
_SIGNATURES = [
    ("marx_alloc_photon_type", PPhotonList, [ctypes.c_uint]),
    ("marx_dealloc_photon_type", ctypes.c_int, [PPhotonList]),
    ("marx_prune_photons", None, [PPhotonList]),
    ("pf_open_parameter_file", ctypes.c_void_p, [ctypes.c_char_p, ctypes.c_char_p]),
    ("pf_close_parameter_file", ctypes.c_int, [ctypes.c_void_p]),
    ("pf_get_double", ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_double)]),
    ("pf_get_integer", ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int)]),
    ("pf_get_boolean", ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int)]),
    ("pf_get_string", ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_uint]),
    ("pf_get_file", ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_uint]),
    ("pf_set_value", ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_char_p]),
    ("marx_set_data_directory", ctypes.c_int, [ctypes.c_char_p]),
    ("marx_set_time", ctypes.c_int, [ctypes.c_double, ctypes.c_double]),
    ("marx_mirror_init", ctypes.c_int, [ctypes.c_void_p]),
    ("marx_mirror_reflect", ctypes.c_int, [PPhotonList, ctypes.c_int]),
    ("marx_grating_init", ctypes.c_int, [ctypes.c_void_p]),
    ("marx_grating_diffract", ctypes.c_int, [PPhotonList, ctypes.c_int]),
    ("marx_detector_init", ctypes.c_int, [ctypes.c_void_p]),
    ("marx_detect", ctypes.c_int, [PPhotonList, ctypes.c_int]),
    ("mxg_last_error", ctypes.c_char_p, []),
    ("mxg_init", ctypes.c_int, []),
    ("mxg_set_seed", None, [ctypes.c_ulonglong]),
    ("mxg_get_seed", ctypes.c_ulonglong, []),
    ("mxg_set_lazy_sync", None, [ctypes.c_int]),
    ("mxg_get_soa", ctypes.c_int, [ctypes.POINTER(SoaView)]),
    ("mxg_trace_device_noalloc", ctypes.c_int, [ctypes.POINTER(SoaView), ctypes.c_uint, ctypes.c_uint,
                                                ctypes.c_double, ctypes.c_double, ctypes.POINTER(ctypes.c_uint),
                                                ctypes.POINTER(ctypes.c_uint)]),
    ("marx_create_source", ctypes.c_void_p, [ctypes.c_void_p]),
    ("marx_open_source", ctypes.c_int, [ctypes.c_void_p]),
    ("marx_close_source", ctypes.c_int, [ctypes.c_void_p]),
    ("marx_create_photons", ctypes.c_int, [ctypes.c_void_p, PPhotonList, ctypes.c_uint,
                                           ctypes.POINTER(ctypes.c_uint), ctypes.POINTER(ctypes.c_double)]),
    ("mxg_set_uniform_times", None, [ctypes.c_double]),
    ("mxg_trace_device", ctypes.c_int, [ctypes.c_uint, ctypes.c_uint, ctypes.c_double, ctypes.c_double,
                                        ctypes.POINTER(ctypes.c_uint), ctypes.POINTER(ctypes.c_uint),
                                        ctypes.POINTER(ctypes.c_float)]),
    ("mxg_download_photons", ctypes.c_int, [PPhotonList]),
    ("mxg_trace_batch", ctypes.c_int, [ctypes.c_uint, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_uint),
                                       ctypes.c_uint, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                       ctypes.POINTER(ctypes.c_uint), ctypes.POINTER(ctypes.c_float)]),
    ("mxg_set_schedule", ctypes.c_int, [ctypes.POINTER(Schedule)]),
    ("marx_write_photons", ctypes.c_int, [PPhotonList, ctypes.c_ulong, ctypes.c_char_p, ctypes.c_int,
                                          ctypes.c_double]),
    ("mxg_write_device", ctypes.c_int, [ctypes.c_ulong, ctypes.c_char_p, ctypes.c_int, ctypes.c_double,
                                        ctypes.POINTER(ctypes.c_uint)]),
    ("mxg_event_coords", ctypes.c_int, [ctypes.POINTER(EventCfg), ctypes.POINTER(EventColumns), ctypes.c_uint,
                                        ctypes.POINTER(ctypes.c_uint)]),
    ("mxg_event_kernel_ms", ctypes.c_int, [ctypes.POINTER(ctypes.c_float)]),
    ("marx_open_read_dump_file", ctypes.POINTER(DumpFile), [ctypes.c_char_p]),
    ("marx_close_read_dump_file", None, [ctypes.POINTER(DumpFile)]),
    ("marx_create_write_dump_file", ctypes.c_void_p, [ctypes.c_char_p, ctypes.POINTER(DumpFile)]),
    ("marx_close_write_dump_file", ctypes.c_int, [ctypes.c_void_p, ctypes.c_ulong]),
    ("marx_dump", ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ctypes.c_char_p)]),
    ("marx_dump_to_rayfile", ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, PPhotonList, ctypes.c_double]),
    ("marx_dump_rayfile", ctypes.c_int, [ctypes.c_char_p]),
    ("mxg_diff_setup", ctypes.c_int, [ctypes.POINTER(ctypes.c_double), ctypes.c_void_p]),
    ("mxg_diff_forward", ctypes.c_int, [ctypes.POINTER(ctypes.c_double), ctypes.c_uint, ctypes.c_uint,
                                        ctypes.POINTER(ImageSpec), ctypes.c_void_p,
                                        ctypes.POINTER(ctypes.c_double)]),
    ("mxg_diff_image", ctypes.c_int, [ctypes.POINTER(ctypes.c_double), ctypes.c_uint, ctypes.c_uint,
                                      ctypes.POINTER(ImageSpec), ctypes.c_void_p,
                                      ctypes.POINTER(ctypes.c_double)]),
    ("mxg_diff_kernel_ms", ctypes.c_int, [ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_uint)]),
    ("mxg_diff_last_pass", ctypes.c_ulonglong, []),
    ("mxg_diff_adjoint", ctypes.c_int, [ctypes.c_ulonglong, ctypes.c_uint, ctypes.c_void_p,
                                        ctypes.POINTER(ctypes.c_double)]),
    ("mxg_diff_state", ctypes.c_int, [ctypes.c_uint] + [ctypes.POINTER(ctypes.c_double)] * 6
                                      + [ctypes.POINTER(ctypes.c_ubyte)]),
    ("mxg_get_schedule", ctypes.c_int, [ctypes.POINTER(Schedule)]),
    ("mxg_autotune_batch", ctypes.c_int, [ctypes.c_uint, ctypes.POINTER(ctypes.c_double),
                                          ctypes.POINTER(ctypes.c_uint), ctypes.c_uint, ctypes.c_double,
                                          ctypes.c_double, ctypes.c_uint, ctypes.POINTER(Schedule),
                                          ctypes.POINTER(ctypes.c_float)]),
    ("mxg_last_stage_counts", ctypes.c_int, [ctypes.POINTER(ctypes.c_uint)]),
    ("mxg_last_stage_ms", ctypes.c_int, [ctypes.POINTER(ctypes.c_float)]),
    ("mxg_last_kernel_stats", ctypes.c_int, [ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_uint)]),
    ("mxg_device_info", ctypes.c_int, [ctypes.c_char_p, ctypes.c_uint, ctypes.POINTER(ctypes.c_int)]),
    ("mxg_get_config", ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_void_p),
                                      ctypes.POINTER(ctypes.c_uint), ctypes.POINTER(ctypes.c_void_p),
                                      ctypes.POINTER(ctypes.c_uint)]),
]

_lib = None


def load():
    """Loads libmarx_gpu.so once; raises ImportError when it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} not found: build it with `make -C marx_amd/csrc` "
            "(or __graft_entry__.build())")
    lib = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_GLOBAL)
    for name, res, args in _SIGNATURES:
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def exported_symbols():
    return [s[0] for s in _SIGNATURES]


def last_error():
    msg = load().mxg_last_error()
    return msg.decode() if msg else ""


def check(status, what):
    if status == -1:
        raise RuntimeError(f"{what} failed: {last_error()}")
    return status
