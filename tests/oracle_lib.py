"""ctypes binding of the CPU oracle (oracle/liboracle.so) for the tests and the CPU baseline.

The oracle is test infrastructure: it is loaded only by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker or the reported baseline, never as the product path.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "liboracle.so")

SUBSENSE, LOBSTER = 1, 2


class OrcParams(C.Structure):
    _fields_ = [
        ("desc_dist_threshold", C.c_size_t),
        ("color_dist_threshold", C.c_size_t),
        ("bg_samples", C.c_size_t),
        ("required_bg_samples", C.c_size_t),
        ("samples_for_mv_avgs", C.c_size_t),
        ("lbsp_threshold_offset", C.c_size_t),
        ("rel_lbsp_threshold", C.c_float),
        ("median_blur_kernel", C.c_int),
        ("rng_seed", C.c_uint64),
        ("stream_id", C.c_uint32),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built: run `make -C oracle` (or __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
        u16p = np.ctypeslib.ndpointer(np.uint16, flags="C_CONTIGUOUS")
        L.orc_default_params.argtypes = [C.c_int, C.POINTER(OrcParams)]
        L.orc_create.argtypes = [C.c_int, C.POINTER(OrcParams)]
        L.orc_create.restype = C.c_void_p
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_initialize.argtypes = [C.c_void_p, u8p, C.c_int, C.c_int, C.c_int, C.c_void_p]
        L.orc_apply.argtypes = [C.c_void_p, u8p, u8p, C.c_double]
        L.orc_get_background_image.argtypes = [C.c_void_p, u8p]
        L.orc_get_background_descriptors_image.argtypes = [C.c_void_p, u16p]
        L.orc_refresh_model.argtypes = [C.c_void_p, C.c_float, C.c_int]
        L.orc_set_automatic_model_reset.argtypes = [C.c_void_p, C.c_int]
        L.orc_last_error.argtypes = [C.c_void_p]
        L.orc_last_error.restype = C.c_char_p
        L.orc_state.argtypes = [C.c_void_p, C.c_char_p, C.c_void_p, C.c_size_t]
        L.orc_state.restype = C.c_longlong
        L.orc_scalars.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.c_int]
        L.orc_counters.argtypes = [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
        L.orc_philox4x32_10.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        L.orc_draw.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32]
        L.orc_draw.restype = C.c_uint32
        L.orc_lbsp_threshold.argtypes = [u8p, C.c_uint8, C.c_uint8]
        L.orc_lbsp_threshold.restype = C.c_uint16
        L.orc_lbsp_compute_abs.argtypes = [u8p, C.c_int, C.c_int, C.c_int, C.c_int, u16p]
        L.orc_build_lut.argtypes = [C.c_int, C.c_float, C.c_size_t, u8p]
        L.orc_sat_u8_f.argtypes = [C.c_float]
        L.orc_sat_u8_f.restype = C.c_uint8
        for fn in ("orc_sample_pos_7x7", "orc_neighbor_pos_3x3", "orc_neighbor_pos_5x5"):
            getattr(L, fn).argtypes = [C.c_int] * 6 + [C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.orc_l1dist3_u8.argtypes = [u8p, u8p]
        L.orc_l1dist3_u8.restype = C.c_uint8
        L.orc_morph.argtypes = [u8p, u8p, C.c_int, C.c_int, C.c_int, C.c_int]
        L.orc_median_binary.argtypes = [u8p, u8p, C.c_int, C.c_int, C.c_int]
        L.orc_flood_holes.argtypes = [u8p, u8p, C.c_int, C.c_int]
        L.orc_resize_area.argtypes = [u8p, C.c_int, C.c_int, C.c_int, u8p, C.c_int, C.c_int]
        L.orc_binclassif.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_size_t,
                                     C.POINTER(C.c_uint64)]
        _lib = L
    return _lib


def philox(ctr, key):
    L = lib()
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    L.orc_philox4x32_10(c, k, o)
    return list(o)


def default_params(algo):
    p = OrcParams()
    lib().orc_default_params(algo, C.byref(p))
    return p


SCALAR_NAMES = ["frame_idx", "frames_since_reset", "cooldown", "auto_reset", "lr_scaling", "use3x3", "median_k",
                "n_roi", "lower_cap", "upper_cap", "last_nz_ratio", "refresh_seq", "apply_count"]


class Oracle:
    """Sequential CPU restatement of SuBSENSE / LOBSTER (see oracle/orc.h)."""

    def __init__(self, algo=SUBSENSE, params=None, **kw):
        self.algo = algo
        p = params if params is not None else default_params(algo)
        for k, v in kw.items():
            setattr(p, k, v)
        self.params = p
        self.h = lib().orc_create(algo, C.byref(p))
        if not self.h:
            raise ValueError("invalid oracle parameters")
        self.shape = None

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def _err(self):
        return lib().orc_last_error(self.h).decode()

    def initialize(self, img, roi=None):
        img = np.ascontiguousarray(img, dtype=np.uint8)
        h, w = img.shape[:2]
        c = 1 if img.ndim == 2 else img.shape[2]
        r = None
        if roi is not None:
            roi = np.ascontiguousarray(roi, dtype=np.uint8)
            r = roi.ctypes.data
        rc = lib().orc_initialize(self.h, img, w, h, c, r)
        if rc != 0:
            raise RuntimeError(self._err())
        self.shape = (h, w, c)

    def apply(self, img, lr=None):
        if lr is None:
            lr = 0.0 if self.algo == SUBSENSE else 16.0
        img = np.ascontiguousarray(img, dtype=np.uint8)
        if self.shape is None:
            mask = np.zeros(1, np.uint8)  # orc_apply rejects the call before touching the mask
        else:
            h, w, c = self.shape
            if img.shape[:2] != (h, w) or (1 if img.ndim == 2 else img.shape[2]) != c:
                raise RuntimeError("input image type/size mismatch with initialization type/size")
            mask = np.zeros((h, w), np.uint8)
        rc = lib().orc_apply(self.h, img, mask, float(lr))
        if rc != 0:
            raise RuntimeError(self._err())
        return mask

    def background_image(self):
        h, w, c = self.shape
        out = np.zeros((h, w, c) if c > 1 else (h, w), np.uint8)
        lib().orc_get_background_image(self.h, out)
        return out

    def background_descriptors(self):
        h, w, c = self.shape
        out = np.zeros((h, w, c) if c > 1 else (h, w), np.uint16)
        lib().orc_get_background_descriptors_image(self.h, out)
        return out

    def refresh_model(self, frac, force_fg=False):
        if lib().orc_refresh_model(self.h, float(frac), int(force_fg)) != 0:
            raise RuntimeError(self._err())

    def set_automatic_model_reset(self, on):
        lib().orc_set_automatic_model_reset(self.h, int(bool(on)))

    def state(self, name):
        h, w, c = self.shape
        n = self.params.bg_samples
        spec = {
            "samples_color": (np.uint8, (n, h, w, c)), "samples_desc": (np.uint16, (n, h, w, c)),
            "last_color": (np.uint8, (h, w, c)), "last_desc": (np.uint16, (h, w, c)),
            "last_fg": (np.uint8, (h, w)), "roi": (np.uint8, (h, w)), "lut": (np.uint8, (256,)),
            "unstable": (np.uint8, (h, w)), "blinks": (np.uint8, (h, w)), "last_raw": (np.uint8, (h, w)),
            "last_raw_blink": (np.uint8, (h, w)), "last_dil_inv": (np.uint8, (h, w)),
            "dsLT": (np.float32, (h // 8, w // 8, c)), "dsST": (np.float32, (h // 8, w // 8, c)),
        }
        dt, shp = spec.get(name, (np.float32, (h, w)))
        out = np.zeros(shp, dt)
        rc = lib().orc_state(self.h, name.encode(), out.ctypes.data, out.nbytes)
        if rc < 0:
            raise KeyError(name)
        return out

    def scalars(self):
        buf = (C.c_double * 32)()
        n = lib().orc_scalars(self.h, buf, 32)
        return {SCALAR_NAMES[i]: buf[i] for i in range(n)}

    def counters(self):
        a, b = C.c_uint64(), C.c_uint64()
        lib().orc_counters(self.h, C.byref(a), C.byref(b))
        return a.value, b.value


def binclassif(classif, gt=None, roi=None):
    """The oracle's BinClassif counts (TP, TN, FP, FN, SE, DC) of 2-D uint8 masks."""
    classif = np.ascontiguousarray(classif, dtype=np.uint8)
    h, w = classif.shape
    out = (C.c_uint64 * 6)()
    g = None if gt is None else np.ascontiguousarray(gt, dtype=np.uint8)
    r = None if roi is None else np.ascontiguousarray(roi, dtype=np.uint8)
    lib().orc_binclassif(classif.ctypes.data, None if g is None else g.ctypes.data,
                         None if r is None else r.ctypes.data, w, h, w, out)
    return tuple(int(x) for x in out)
