"""Python mirror of litiv's IBackgroundSubtractor for the gfx950 SuBSENSE / LOBSTER library.

Method names follow the reference classes (modules/video/include/litiv/video/
BackgroundSubtractionUtils.hpp:24-97, BackgroundSubtractorSuBSENSE.hpp:48-135,
BackgroundSubtractorLOBSTER.hpp:132-155) so code written against the reference reads the same:

    bgs = BackgroundSubtractorSuBSENSE()
    bgs.initialize(frame0)            # numpy HxWx3 / HxW uint8
    mask = bgs.apply(frame)           # HxW uint8, 0/255
    bg = bgs.getBackgroundImage()

Errors raise LvbgsError carrying the library's message (the reference throws lv::Exception).
Every call goes through liblitiv_bgs.so (include/litiv_bgs.h); there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

from . import _native

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liblitiv_bgs.so")

SUBSENSE, LOBSTER = 1, 2
OK, ERR_ARG, ERR_STATE, ERR_HIP, ERR_ALLOC, ERR_DEVICE = 0, -1, -2, -3, -4, -5

SCALAR_NAMES = ["frame_idx", "frames_since_reset", "cooldown", "auto_reset", "lr_scaling", "use3x3", "median_k",
                "n_roi", "lower_cap", "upper_cap", "last_nz_ratio", "refresh_seq", "apply_count"]


class LvbgsError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (code %d)" % (msg, code))
        self.code = code


class Params(C.Structure):
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
        ("device", C.c_int),
    ]


class Stats(C.Structure):
    _fields_ = [("examined", C.c_uint64), ("writes", C.c_uint64), ("fixpoint_iters", C.c_uint32),
                ("max_fixpoint_iter", C.c_uint32), ("width", C.c_uint32), ("height", C.c_uint32),
                ("channels", C.c_uint32), ("n_roi", C.c_uint32), ("examined_speculative", C.c_uint64),
                ("fixpoint_work", C.c_uint32 * 8), ("examined_cooperative", C.c_uint64)]


class BatchOptions(C.Structure):
    _fields_ = [("use_graph", C.c_int), ("seq_limit", C.c_int), ("cont_blocks", C.c_int),
                ("cont_px_per_block", C.c_int), ("fixpoint_launches", C.c_int), ("fixpoint_blocks", C.c_int),
                ("fix_px_per_block", C.c_int), ("fetch_depth", C.c_int), ("groups", C.c_int)]


class ServerOptions(C.Structure):
    _fields_ = [("hbm_budget", C.c_size_t), ("max_inflight", C.c_int), ("max_batch", C.c_int),
                ("launch", BatchOptions), ("host_ring", C.c_int)]


class BinClassifCounts(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("tp", "tn", "fp", "fn", "se", "dc")]


class ServerStats(C.Structure):
    _fields_ = [("steps", C.c_uint64), ("frames_submitted", C.c_uint64), ("frames_launched", C.c_uint64),
                ("frames_done", C.c_uint64), ("bytes_used", C.c_uint64), ("pending_frames", C.c_uint64),
                ("open_streams", C.c_uint32), ("pooled_contexts", C.c_uint32), ("queued_steps", C.c_uint32),
                ("reserved", C.c_uint32)]


DONE_DTYPE = np.dtype([("tag", np.uint64), ("stream", np.int32), ("status", np.int32)])
ERR_BUDGET = -6
ERR_BUSY = -7
ERR_TOO_LARGE = -8
# include/litiv_bgs.h LVBGS_BINCLASSIF_NOALLOC_MAX_PIXELS: the no-alloc evaluator's frame limit
BINCLASSIF_NOALLOC_MAX_PIXELS = 67100672

_lib = None


def library_path():
    return _LIB_PATH


def lib():
    """Loads liblitiv_bgs.so; raises if it is missing (build it with __graft_entry__.build())."""
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError("liblitiv_bgs.so is not built (run __graft_entry__.build()); "
                               "litiv_amd has no CPU fallback")
        L = _native.load(_LIB_PATH)
        vp, sz, i = C.c_void_p, C.c_size_t, C.c_int
        L.lvbgs_default_params.argtypes = [i, C.POINTER(Params)]
        L.lvbgs_create.argtypes = [i, C.POINTER(Params), C.POINTER(vp)]
        L.lvbgs_destroy.argtypes = [vp]
        L.lvbgs_initialize.argtypes = [vp, vp, i, i, i, sz, vp, sz]
        L.lvbgs_apply.argtypes = [vp, vp, sz, vp, sz, C.c_double]
        L.lvbgs_apply_device.argtypes = [vp, vp, vp, C.c_double]
        L.lvbgs_synchronize.argtypes = [vp]
        L.lvbgs_get_background_image.argtypes = [vp, vp, sz]
        L.lvbgs_get_background_descriptors_image.argtypes = [vp, vp, sz]
        L.lvbgs_set_automatic_model_reset.argtypes = [vp, i]
        L.lvbgs_refresh_model.argtypes = [vp, C.c_float, i]
        L.lvbgs_default_learning_rate.argtypes = [vp]
        L.lvbgs_default_learning_rate.restype = C.c_double
        L.lvbgs_get_roi.argtypes = [vp, vp, sz]
        L.lvbgs_validate_roi.argtypes = [i, i, vp, sz]
        L.lvbgs_set_roi.argtypes = [vp, vp, sz]
        L.lvbgs_last_error.argtypes = [vp]
        L.lvbgs_last_error.restype = C.c_char_p
        L.lvbgs_lbsp_compute_abs.argtypes = [vp, i, i, i, i, vp, vp]
        L.lvbgs_debug_lbsp_packed.argtypes = [vp, i, i, i, i, vp, vp]
        L.lvbgs_get_stats.argtypes = [vp, C.POINTER(Stats)]
        L.lvbgs_get_state.argtypes = [vp, C.c_char_p, vp, sz]
        L.lvbgs_get_state.restype = C.c_longlong
        L.lvbgs_get_scalars.argtypes = [vp, C.POINTER(C.c_double), i]
        L.lvbgs_stream.argtypes = [vp]
        L.lvbgs_stream.restype = vp
        L.lvbgs_set_timing.argtypes = [vp, C.c_char_p]
        L.lvbgs_take_timing.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_int)]
        L.lvbgs_set_fixpoint_schedule.argtypes = [vp, i, i]
        L.lvbgs_set_match_schedule.argtypes = [vp, i, i]
        L.lvbgs_debug_rng.argtypes = [vp, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, vp]
        L.lvbgs_debug_postprocess.argtypes = [vp, vp, vp]
        L.lvbgs_initialize_device.argtypes = [vp, vp, i, i, i, vp, sz]
        L.lvbgs_model_bytes.argtypes = [i, C.POINTER(Params), i, i, i]
        L.lvbgs_model_bytes.restype = sz
        L.lvbgs_server_default_options.argtypes = [C.POINTER(ServerOptions)]
        L.lvbgs_server_create.argtypes = [i, C.POINTER(ServerOptions), C.POINTER(vp)]
        L.lvbgs_server_destroy.argtypes = [vp]
        L.lvbgs_server_last_error.argtypes = [vp]
        L.lvbgs_server_last_error.restype = C.c_char_p
        L.lvbgs_server_stream.argtypes = [vp]
        L.lvbgs_server_stream.restype = vp
        L.lvbgs_server_open.argtypes = [vp, C.POINTER(Params), vp, i, i, i, C.POINTER(i)]
        L.lvbgs_server_close.argtypes = [vp, i]
        L.lvbgs_server_submit.argtypes = [vp, i, vp, vp, C.c_double, C.c_uint64]
        L.lvbgs_server_submit_host.argtypes = [vp, i, vp, sz, vp, C.c_double, C.c_uint64]
        L.lvbgs_server_open_host.argtypes = [vp, C.POINTER(Params), vp, sz, i, i, i, C.POINTER(i)]
        L.lvbgs_server_stream_bytes.argtypes = [vp, C.POINTER(Params), i, i, i]
        L.lvbgs_server_stream_bytes.restype = sz
        L.lvbgs_host_alloc.argtypes = [sz]
        L.lvbgs_host_alloc.restype = vp
        L.lvbgs_host_free.argtypes = [vp]
        L.lvbgs_server_pump.argtypes = [vp]
        L.lvbgs_server_poll.argtypes = [vp, vp, i]
        L.lvbgs_server_drain.argtypes = [vp]
        L.lvbgs_server_get_stats.argtypes = [vp, C.POINTER(ServerStats)]
        L.lvbgs_server_context.argtypes = [vp, i]
        L.lvbgs_server_context.restype = vp
        L.lvbgs_server_debug_fixpoint_cap.argtypes = [vp, i]
        L.lvbgs_binclassif_accumulate.argtypes = [vp, vp, vp, i, i, C.c_size_t, C.POINTER(BinClassifCounts), vp]
        L.lvbgs_binclassif_accumulate_noalloc.argtypes = [vp, vp, vp, i, i, C.c_size_t, C.POINTER(BinClassifCounts),
                                                          vp]
        L.lvbgs_archive_write.argtypes = [C.c_char_p, i, i, C.POINTER(C.c_int32), vp]
        L.lvbgs_archive_read.argtypes = [C.c_char_p, C.POINTER(i), C.POINTER(i), C.POINTER(C.c_int32), i, vp,
                                         C.c_size_t]
        L.lvbgs_archive_read.restype = C.c_longlong
        L.lvbgs_checkpoint_save.argtypes = [vp, C.c_char_p]
        L.lvbgs_checkpoint_load.argtypes = [vp, C.c_char_p]
        L.lvbgs_batch_create.argtypes = [C.POINTER(vp), i, C.POINTER(vp)]
        L.lvbgs_batch_destroy.argtypes = [vp]
        L.lvbgs_batch_last_error.argtypes = [vp]
        L.lvbgs_batch_last_error.restype = C.c_char_p
        L.lvbgs_batch_stream.argtypes = [vp]
        L.lvbgs_batch_stream.restype = vp
        L.lvbgs_batch_default_options.argtypes = [C.POINTER(BatchOptions)]
        L.lvbgs_batch_get_options.argtypes = [vp, C.POINTER(BatchOptions)]
        L.lvbgs_batch_set_options.argtypes = [vp, C.POINTER(BatchOptions)]
        L.lvbgs_batch_set_stream_options.argtypes = [vp, i, i, i, i]
        L.lvbgs_batch_get_stream_options.argtypes = [vp, i, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.lvbgs_batch_apply_device.argtypes = [vp, C.POINTER(vp), C.POINTER(vp), C.POINTER(C.c_double)]
        L.lvbgs_batch_synchronize.argtypes = [vp]
        L.lvbgs_batch_set_timing.argtypes = [vp, C.c_char_p]
        L.lvbgs_batch_take_timing.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_int)]
        _lib = L
    return _lib


def default_params(algo):
    p = Params()
    if lib().lvbgs_default_params(algo, C.byref(p)) != OK:
        raise ValueError("unknown algorithm %r" % algo)
    return p


def _u8(img):
    a = np.ascontiguousarray(img)
    if a.dtype != np.uint8:
        raise TypeError("images must be uint8 (CV_8UC1 / CV_8UC3)")
    return a


def lbsp_compute_abs(d_img, width, height, channels, threshold, d_out, stream=None, packed=False):
    """LBSP::compute2 with an absolute threshold on device buffers (LBSP.cpp:102-123). packed=True
    runs it through the sample match's form of the threshold test (lvbgs_debug_lbsp_packed)."""
    fn = lib().lvbgs_debug_lbsp_packed if packed else lib().lvbgs_lbsp_compute_abs
    rc = fn(C.c_void_p(d_img), width, height, channels, int(threshold), C.c_void_p(d_out), C.c_void_p(stream or 0))
    if rc != OK:
        raise LvbgsError(rc, "lvbgs_lbsp_compute_abs failed")


# OpenCV depth codes of the archive format and the numpy types they hold
_CV_DEPTHS = {0: np.uint8, 1: np.int8, 2: np.uint16, 3: np.int16, 4: np.int32, 5: np.float32, 6: np.float64}


def write_archive(path, arr, channels=1):
    """lv::write(path, mat, lv::MatArchive_BINARY): arr's trailing axis holds the channels when
    channels > 1 (an H x W x 3 array with channels=3 is a 2-D CV_8UC3 matrix)."""
    arr = np.ascontiguousarray(arr)
    depth = {np.dtype(v): k for k, v in _CV_DEPTHS.items()}.get(arr.dtype)
    if depth is None:
        raise TypeError("unsupported element type %s" % arr.dtype)
    shape = arr.shape[:-1] if channels > 1 else arr.shape
    if channels > 1 and arr.shape[-1] != channels:
        raise ValueError("trailing axis must hold the channels")
    sizes = (C.c_int32 * max(1, len(shape)))(*shape)
    rc = lib().lvbgs_archive_write(os.fsencode(path), depth + ((channels - 1) << 3), len(shape), sizes,
                                   arr.ctypes.data)
    if rc:
        raise LvbgsError(rc, "cannot write archive %s" % path)


def read_archive(path):
    """lv::read(path, lv::MatArchive_BINARY): the matrix as a numpy array (channels on a trailing
    axis when there are more than one) and its OpenCV type code."""
    t, nd, sizes = C.c_int(), C.c_int(), (C.c_int32 * 8)()
    n = lib().lvbgs_archive_read(os.fsencode(path), C.byref(t), C.byref(nd), sizes, 8, None, 0)
    if n < 0:
        raise LvbgsError(int(n), "cannot read archive %s" % path)
    depth, ch = t.value & 7, (t.value >> 3) + 1
    if depth not in _CV_DEPTHS or nd.value > 8:
        raise LvbgsError(-1, "unsupported archive type %d in %s" % (t.value, path))
    shape = tuple(sizes[:nd.value]) + ((ch,) if ch > 1 else ())
    out = np.empty(shape, _CV_DEPTHS[depth])
    if lib().lvbgs_archive_read(os.fsencode(path), None, None, None, 0, out.ctypes.data, out.nbytes) != n:
        raise LvbgsError(-1, "cannot read archive %s" % path)
    return out, t.value


class BackgroundSubtractor:
    """IIBackgroundSubtractor (BackgroundSubtractionUtils.hpp:24-97) backed by the HIP library."""

    ALGO = None

    def __init__(self, params=None, **kw):
        p = params if params is not None else default_params(self.ALGO)
        for k, v in kw.items():
            if not hasattr(p, k):
                raise TypeError("unknown parameter %r" % k)
            setattr(p, k, v)
        self.params = p
        h = C.c_void_p()
        rc = lib().lvbgs_create(self.ALGO, C.byref(p), C.byref(h))
        if rc != OK:
            raise LvbgsError(rc, "lvbgs_create failed: invalid parameters" if rc == ERR_ARG else "lvbgs_create failed")
        self._h = h
        self.shape = None

    @classmethod
    def _borrow(cls, handle, params, shape):
        """A non-owning view of a context owned elsewhere (a server's stream)."""
        self = cls.__new__(cls)
        self.params = params
        self._h = C.c_void_p(handle)
        self._owned = False
        self.shape = shape
        return self

    def close(self):
        if getattr(self, "_h", None):
            if getattr(self, "_owned", True):
                lib().lvbgs_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # interpreter shutdown: the library may already be gone
            pass

    def _check(self, rc):
        if rc < 0:
            raise LvbgsError(rc, lib().lvbgs_last_error(self._h).decode())
        return rc

    def initialize_device(self, d_img, width, height, channels):
        """initialize() from a device-resident, tightly packed frame."""
        self._check(lib().lvbgs_initialize_device(self._h, C.c_void_p(d_img), width, height, channels, None, 0))
        self.shape = (height, width, channels)

    # ---- reference interface ----
    def initialize(self, img, roi=None):
        img = _u8(img)
        h, w = img.shape[:2]
        c = 1 if img.ndim == 2 else img.shape[2]
        r, rs = None, 0
        if roi is not None:
            roi = _u8(roi)
            r, rs = roi.ctypes.data, roi.strides[0]
        self._check(lib().lvbgs_initialize(self._h, img.ctypes.data, w, h, c, img.strides[0], r, rs))
        self.shape = (h, w, c)

    def apply(self, img, fgmask=None, learningRate=None):
        if learningRate is None:
            learningRate = self.getDefaultLearningRate()
        img = _u8(img)
        if self.shape is None:
            raise LvbgsError(ERR_STATE, "algo & model must be initialized first")
        h, w, c = self.shape
        if img.shape[:2] != (h, w) or (1 if img.ndim == 2 else img.shape[2]) != c:
            raise LvbgsError(ERR_ARG, "input image type/size mismatch with initialization type/size")
        if fgmask is None:
            fgmask = np.empty((h, w), np.uint8)
        self._check(lib().lvbgs_apply(self._h, img.ctypes.data, img.strides[0], fgmask.ctypes.data, fgmask.strides[0],
                                      float(learningRate)))
        return fgmask

    def apply_device(self, d_img, d_mask=None, learningRate=None):
        """apply() on device-resident buffers (tightly packed); asynchronous on this context's stream."""
        if learningRate is None:
            learningRate = self.getDefaultLearningRate()
        self._check(lib().lvbgs_apply_device(self._h, C.c_void_p(d_img), C.c_void_p(d_mask or 0), float(learningRate)))

    def synchronize(self):
        self._check(lib().lvbgs_synchronize(self._h))

    def getBackgroundImage(self):
        h, w, c = self.shape
        out = np.empty((h, w, c) if c > 1 else (h, w), np.uint8)
        self._check(lib().lvbgs_get_background_image(self._h, out.ctypes.data, out.strides[0]))
        return out

    def getBackgroundDescriptorsImage(self):
        h, w, c = self.shape
        out = np.empty((h, w, c) if c > 1 else (h, w), np.uint16)
        self._check(lib().lvbgs_get_background_descriptors_image(self._h, out.ctypes.data, out.strides[0]))
        return out

    def getDefaultLearningRate(self):
        return lib().lvbgs_default_learning_rate(self._h)

    def setAutomaticModelReset(self, enabled):
        self._check(lib().lvbgs_set_automatic_model_reset(self._h, int(bool(enabled))))

    def refreshModel(self, fSamplesRefreshFrac, bForceFGUpdate=False):
        self._check(lib().lvbgs_refresh_model(self._h, float(fSamplesRefreshFrac), int(bool(bForceFGUpdate))))

    @staticmethod
    def validateROI(roi):
        roi = np.ascontiguousarray(roi, dtype=np.uint8).copy()
        rc = lib().lvbgs_validate_roi(roi.shape[1], roi.shape[0], roi.ctypes.data, roi.strides[0])
        if rc != OK:
            raise LvbgsError(rc, "provided ROI must be non-empty and of type 8UC1")
        return roi

    def setROI(self, roi):
        roi = np.ascontiguousarray(roi, dtype=np.uint8).copy()
        self._check(lib().lvbgs_set_roi(self._h, roi.ctypes.data, roi.strides[0]))
        if self.shape is not None:
            self.shape = (roi.shape[0], roi.shape[1], self.shape[2])
        return roi

    def getROICopy(self):
        h, w, _ = self.shape
        out = np.empty((h, w), np.uint8)
        self._check(lib().lvbgs_get_roi(self._h, out.ctypes.data, out.strides[0]))
        return out

    # ---- introspection (tests / benchmarks) ----
    def stream(self):
        return lib().lvbgs_stream(self._h)

    def stats(self):
        s = Stats()
        self._check(lib().lvbgs_get_stats(self._h, C.byref(s)))
        d = {k: getattr(s, k) for k, _ in Stats._fields_}
        d["fixpoint_work"] = list(d["fixpoint_work"])
        return d

    def scalars(self):
        buf = (C.c_double * 32)()
        n = self._check(lib().lvbgs_get_scalars(self._h, buf, 32))
        return {SCALAR_NAMES[i]: buf[i] for i in range(n)}

    def state(self, name):
        h, w, c = self.shape
        n = self.params.bg_samples
        spec = {
            "samples_color": (np.uint8, (n, h, w, c)), "samples_desc": (np.uint16, (n, h, w, c)),
            "last_color": (np.uint8, (h, w, c)), "last_desc": (np.uint16, (h, w, c)), "desc_cur": (np.uint16, (h, w, c)),
            "lut": (np.uint8, (256,)), "dsLT": (np.float32, (h // 8, w // 8, c)), "dsST": (np.float32, (h // 8, w // 8, c)),
            "mres": (np.uint32, (h, w)), "rec": (np.uint32, (h, w)),
        }
        u8 = {"last_fg", "roi", "unstable", "blinks", "last_raw", "last_raw_blink", "last_dil_inv", "raw", "dec"}
        dt, shp = spec.get(name, ((np.uint8, (h, w)) if name in u8 else (np.float32, (h, w))))
        out = np.empty(shp, dt)
        rc = lib().lvbgs_get_state(self._h, name.encode(), out.ctypes.data, out.nbytes)
        if rc < 0:
            raise KeyError(name)
        return out

    # ---- checkpoint / resume ----
    def save_checkpoint(self, directory):
        """Writes the whole model into directory (created if missing), one binary archive per state
        buffer plus meta.bin; load_checkpoint on a subtractor created with the same parameters
        continues the sequence exactly."""
        os.makedirs(directory, exist_ok=True)
        self._check(lib().lvbgs_checkpoint_save(self._h, os.fsencode(directory)))

    def load_checkpoint(self, directory):
        self._check(lib().lvbgs_checkpoint_load(self._h, os.fsencode(directory)))
        meta, _ = read_archive(os.path.join(directory, "meta.bin"))
        self.shape = (int(meta[3]), int(meta[2]), int(meta[4]))

    def set_timing(self, stage):
        self._check(lib().lvbgs_set_timing(self._h, (stage or "").encode()))

    def take_timing(self):
        ms, n = C.c_double(), C.c_int()
        self._check(lib().lvbgs_take_timing(self._h, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def set_fixpoint_schedule(self, launches, blocks):
        self._check(lib().lvbgs_set_fixpoint_schedule(self._h, int(launches), int(blocks)))

    def set_match_schedule(self, seq_limit, cont_blocks=1024):
        """Samples scanned per pixel by the sequential match phase (1..255; 0 = automatic: 8 for
        SuBSENSE RGB frames of 1280x720 or more, else 255) and the continuation grid; results do
        not depend on it."""
        self._check(lib().lvbgs_set_match_schedule(self._h, int(seq_limit), int(cont_blocks)))

    def debug_rng(self, frame, phase, event, n):
        out = np.empty((n, 4), np.uint32)
        self._check(lib().lvbgs_debug_rng(self._h, frame, phase, event, n, out.ctypes.data))
        return out

    def debug_postprocess(self, raw):
        """Hole filling + median of a raw 0/255 mask (SuBSENSE .cpp:540-547) on the device; returns
        the filtered mask and leaves it as the context's last foreground mask."""
        raw = np.ascontiguousarray(raw, dtype=np.uint8)
        out = np.empty_like(raw)
        self._check(lib().lvbgs_debug_postprocess(self._h, raw.ctypes.data, out.ctypes.data))
        return out


class BackgroundSubtractorSuBSENSE(BackgroundSubtractor):
    """BackgroundSubtractorSuBSENSE (BackgroundSubtractorSuBSENSE.hpp:48-135); default learning rate 0."""
    ALGO = SUBSENSE


class BackgroundSubtractorLOBSTER(BackgroundSubtractor):
    """BackgroundSubtractorLOBSTER (BackgroundSubtractorLOBSTER.hpp:132-155); default learning rate 16."""
    ALGO = LOBSTER


class Batch:
    """Several initialized SuBSENSE subtractors on one device stepped together: one launch per kernel
    covers every stream (streams may differ in size and channel count). Their results equal those
    of stepping each subtractor alone. While the batch exists its members are stepped only through it;
    getters and state reads of a member are valid after synchronize()."""

    def __init__(self, subtractors, **options):
        self.members = list(subtractors)
        hs = (C.c_void_p * len(self.members))(*[m._h.value for m in self.members])
        h = C.c_void_p()
        rc = lib().lvbgs_batch_create(hs, len(self.members), C.byref(h))
        if rc != OK:
            raise LvbgsError(rc, "lvbgs_batch_create failed: members must be distinct, initialized SuBSENSE "
                                 "subtractors on one device")
        self._h = h
        self._imgs = (C.c_void_p * len(self.members))()
        self._masks = (C.c_void_p * len(self.members))()
        self._lrs = (C.c_double * len(self.members))()
        if options:
            self.set_options(**options)

    def close(self):
        if getattr(self, "_h", None):
            lib().lvbgs_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # interpreter shutdown: the library may already be gone
            pass

    def _check(self, rc):
        if rc < 0:
            raise LvbgsError(rc, lib().lvbgs_batch_last_error(self._h).decode())
        return rc

    def options(self):
        o = BatchOptions()
        self._check(lib().lvbgs_batch_get_options(self._h, C.byref(o)))
        return {k: getattr(o, k) for k, _ in BatchOptions._fields_}

    def set_options(self, **kw):
        """Launch options (use_graph, seq_limit, cont_blocks, cont_px_per_block, fixpoint_launches,
        fixpoint_blocks, fix_px_per_block, fetch_depth, groups); unspecified ones keep their current
        value. Options never change results."""
        o = BatchOptions()
        self._check(lib().lvbgs_batch_get_options(self._h, C.byref(o)))
        for k, v in kw.items():
            if not hasattr(o, k):
                raise TypeError("unknown batch option %r" % k)
            setattr(o, k, int(v))
        self._check(lib().lvbgs_batch_set_options(self._h, C.byref(o)))

    STREAM_OPTIONS = ("seq_limit", "cont_px_per_block", "fix_px_per_block")

    def stream_options(self, index):
        """Member `index`'s overrides of the batch options (0 = the batch's value)."""
        vals = [C.c_int() for _ in self.STREAM_OPTIONS]
        self._check(lib().lvbgs_batch_get_stream_options(self._h, int(index), *[C.byref(x) for x in vals]))
        return {k: x.value for k, x in zip(self.STREAM_OPTIONS, vals)}

    def set_stream_options(self, index, **kw):
        """Overrides for member `index` (seq_limit, cont_px_per_block, fix_px_per_block; 0 = the batch's
        value); unspecified ones keep their current override."""
        cur = self.stream_options(index)
        for k, v in kw.items():
            if k not in cur:
                raise TypeError("unknown stream option %r" % k)
            cur[k] = int(v)
        self._check(lib().lvbgs_batch_set_stream_options(self._h, int(index), *[cur[k] for k in self.STREAM_OPTIONS]))

    def shape_classes(self):
        """Member indices grouped by (width, height, channels), largest total pixel count first."""
        groups = {}
        for i, m in enumerate(self.members):
            groups.setdefault(m.shape, []).append(i)
        return sorted(groups.items(), key=lambda kv: -kv[0][0] * kv[0][1] * len(kv[1]))

    def apply_device(self, d_imgs, d_masks=None, learningRates=None):
        """One apply() of every member: device frame pointers (and optional mask pointers) in member order."""
        n = len(self.members)
        if len(d_imgs) != n or (d_masks is not None and len(d_masks) != n):
            raise LvbgsError(ERR_ARG, "one frame (and mask) per member is required")
        for i in range(n):
            self._imgs[i] = d_imgs[i]
            self._masks[i] = d_masks[i] if d_masks is not None else None
            self._lrs[i] = 0.0 if learningRates is None else float(learningRates[i])
        self._check(lib().lvbgs_batch_apply_device(self._h, self._imgs, self._masks, self._lrs))

    def synchronize(self):
        self._check(lib().lvbgs_batch_synchronize(self._h))

    def stream(self):
        return lib().lvbgs_batch_stream(self._h)

    def set_timing(self, stage):
        """"match" | "update" | "step" | "" (off): HIP-event timing of each following step (kernel stages
        need groups = 1)."""
        self._check(lib().lvbgs_batch_set_timing(self._h, (stage or "").encode()))

    def take_timing(self):
        ms, n = C.c_double(), C.c_int()
        self._check(lib().lvbgs_batch_take_timing(self._h, C.byref(ms), C.byref(n)))
        return ms.value, n.value


class Server:
    """Serving loop over resident SuBSENSE models (continuous batching, device-memory admission).
    Frames and masks are device pointers; completions come back through poll()."""

    def __init__(self, device=0, hbm_budget=250 << 30, max_inflight=2, max_batch=256, host_ring=0, **launch):
        o = ServerOptions()
        lib().lvbgs_server_default_options(C.byref(o))
        o.hbm_budget = int(hbm_budget)
        o.max_inflight = int(max_inflight)
        o.max_batch = int(max_batch)
        o.host_ring = int(host_ring)
        self.host_ring = int(host_ring)
        for k, v in launch.items():
            if not hasattr(o.launch, k):
                raise TypeError("unknown launch option %r" % k)
            setattr(o.launch, k, int(v))
        h = C.c_void_p()
        rc = lib().lvbgs_server_create(int(device), C.byref(o), C.byref(h))
        if rc != OK:
            raise LvbgsError(rc, "lvbgs_server_create failed")
        self._h = h
        self._params = {}
        self._shape = {}
        self._done = np.zeros(4096, DONE_DTYPE)

    def close_server(self):
        if getattr(self, "_h", None):
            lib().lvbgs_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close_server()
        except Exception:  # interpreter shutdown: the library may already be gone
            pass

    def _check(self, rc):
        if rc < 0:
            raise LvbgsError(rc, lib().lvbgs_server_last_error(self._h).decode())
        return rc

    @staticmethod
    def model_bytes(width, height, channels, **params):
        p = default_params(SUBSENSE)
        for k, v in params.items():
            setattr(p, k, v)
        return lib().lvbgs_model_bytes(SUBSENSE, C.byref(p), width, height, channels)

    def stream_bytes(self, width, height, channels, **params):
        """Device bytes a stream of this geometry holds in this server (model, per-step buffers and its
        host-ingest slots): what admission counts against the budget."""
        p = default_params(SUBSENSE)
        for k, v in params.items():
            setattr(p, k, v)
        return lib().lvbgs_server_stream_bytes(self._h, C.byref(p), width, height, channels)

    def open(self, d_img0, width, height, channels, **params):
        """Admits a stream; returns its id, or None when its model does not fit the budget."""
        p = default_params(SUBSENSE)
        for k, v in params.items():
            if not hasattr(p, k):
                raise TypeError("unknown parameter %r" % k)
            setattr(p, k, v)
        sid = C.c_int()
        rc = lib().lvbgs_server_open(self._h, C.byref(p), C.c_void_p(d_img0), width, height, channels, C.byref(sid))
        if rc == ERR_BUDGET:
            return None
        self._check(rc)
        self._params[sid.value] = p
        self._shape[sid.value] = (height, width, channels)
        return sid.value

    def open_host(self, img0, **params):
        """Admits a stream initialized from a host frame (numpy HxWxC / HxW uint8); id or None."""
        img0 = _u8(img0)
        h, w = img0.shape[:2]
        c = 1 if img0.ndim == 2 else img0.shape[2]
        p = default_params(SUBSENSE)
        for k, v in params.items():
            if not hasattr(p, k):
                raise TypeError("unknown parameter %r" % k)
            setattr(p, k, v)
        sid = C.c_int()
        rc = lib().lvbgs_server_open_host(self._h, C.byref(p), img0.ctypes.data, img0.strides[0], w, h, c,
                                          C.byref(sid))
        if rc == ERR_BUDGET:
            return None
        self._check(rc)
        self._params[sid.value] = p
        self._shape[sid.value] = (h, w, c)
        return sid.value

    def submit_host(self, sid, host_ptr, d_mask=None, lr=0.0, tag=0, stride=0):
        """Submits a frame in host memory (address host_ptr, kept unchanged until it completes);
        returns False when the stream's ingest slots are all busy (poll, then retry)."""
        rc = lib().lvbgs_server_submit_host(self._h, int(sid), C.c_void_p(host_ptr), int(stride),
                                            C.c_void_p(d_mask or 0), float(lr), int(tag))
        if rc == ERR_BUSY:
            return False
        self._check(rc)
        return True

    def close(self, sid):
        self._check(lib().lvbgs_server_close(self._h, int(sid)))

    def submit(self, sid, d_img, d_mask=None, lr=0.0, tag=0):
        self._check(lib().lvbgs_server_submit(self._h, int(sid), C.c_void_p(d_img), C.c_void_p(d_mask or 0),
                                              float(lr), int(tag)))

    def pump(self):
        return self._check(lib().lvbgs_server_pump(self._h))

    def poll(self):
        """Completions since the last call: structured array (tag, stream, status)."""
        out = []
        while True:
            n = self._check(lib().lvbgs_server_poll(self._h, self._done.ctypes.data, len(self._done)))
            out.append(self._done[:n].copy())
            if n < len(self._done):
                break
        return np.concatenate(out)

    def drain(self):
        self._check(lib().lvbgs_server_drain(self._h))

    def stats(self):
        st = ServerStats()
        self._check(lib().lvbgs_server_get_stats(self._h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in ServerStats._fields_}

    def debug_fixpoint_cap(self, tail_cap):
        """Test hook: iteration cap of the fix-point tail of later steps (default 65536); a stream
        needing more fails with ERR_DEVICE and is closed by the server."""
        self._check(lib().lvbgs_server_debug_fixpoint_cap(self._h, int(tail_cap)))

    def retired(self, sid):
        """True once a closed stream's last frame has completed (its id may then be reused)."""
        return not lib().lvbgs_server_context(self._h, int(sid))

    def member(self, sid):
        """The open stream's subtractor (state reads and getters after drain()); not owned."""
        h = lib().lvbgs_server_context(self._h, int(sid))
        if not h:
            raise LvbgsError(ERR_ARG, "unknown stream")
        return BackgroundSubtractorSuBSENSE._borrow(h, self._params[sid], self._shape[sid])


class BinClassif:
    """lv::BinClassif (modules/datasets/include/litiv/datasets/metrics.hpp:32-66): classification
    counters of binary masks against CDnet-style ground truth, accumulated on the device."""

    FIELDS = ("tp", "tn", "fp", "fn", "se", "dc")

    def __init__(self):
        self._c = BinClassifCounts()

    def __getattr__(self, name):
        if name in BinClassif.FIELDS:
            return getattr(self._c, name)
        raise AttributeError(name)

    def counts(self):
        return {k: getattr(self._c, k) for k in self.FIELDS}

    def total(self, with_dont_care=False):
        c = self._c
        return c.tp + c.tn + c.fp + c.fn + (c.dc if with_dont_care else 0)

    def accumulate_device(self, d_classif, d_gt, d_roi, width, height, stride=0, stream=None):
        """accumulate(oClassif, oGT, oROI) (metrics.cpp:21-50) on device masks; d_gt / d_roi may be None."""
        rc = lib().lvbgs_binclassif_accumulate(C.c_void_p(d_classif), C.c_void_p(d_gt or 0), C.c_void_p(d_roi or 0),
                                               int(width), int(height), int(stride), C.byref(self._c),
                                               C.c_void_p(stream or 0))
        if rc != OK:
            raise LvbgsError(rc, "lvbgs_binclassif_accumulate failed")
        return self

    def metrics(self):
        """BinClassifMetrics (metrics.hpp:203-255) of the accumulated counters."""
        c = self._c
        tp, tn, fp, fn = c.tp, c.tn, c.fp, c.fn
        recall = tp / (tp + fn) if tp + fn > 0 else 0.0
        precision = tp / (tp + fp) if tp + fp > 0 else 0.0
        total = tp + tn + fp + fn
        mcc = 0.0
        if tp + fp > 0 and tp + fn > 0 and tn + fp > 0 and tn + fn > 0:
            mcc = (float(tp) * tn - float((fp * fn) % (1 << 64))) / np.sqrt(
                (float(tp) + fp) * (tp + fn) * (tn + fp) * (tn + fn))
        return {
            "recall": recall,
            "specificity": tn / (tn + fp) if tn + fp > 0 else 0.0,
            "fpr": fp / (fp + tn) if fp + tn > 0 else 0.0,
            "fnr": fn / (tp + fn) if tp + fn > 0 else 0.0,
            "pbc": 100.0 * (fn + fp) / total if total > 0 else 0.0,
            "precision": precision,
            "fmeasure": 2.0 * (recall * precision) / (recall + precision) if recall + precision > 0 else 0.0,
            "mcc": mcc,
        }


class PinnedBuffer:
    """Page-locked host memory (lvbgs_host_alloc) viewed as a numpy uint8 array, for frames submitted
    with Server.submit_host (asynchronous copies)."""

    def __init__(self, nbytes):
        self._p = lib().lvbgs_host_alloc(int(nbytes))
        if not self._p:
            raise LvbgsError(ERR_ALLOC, "lvbgs_host_alloc failed")
        self.nbytes = int(nbytes)
        self.array = np.ctypeslib.as_array((C.c_uint8 * self.nbytes).from_address(self._p))

    @property
    def ptr(self):
        return self._p

    def free(self):
        if getattr(self, "_p", None):
            self.array = None
            lib().lvbgs_host_free(self._p)
            self._p = None

    def __del__(self):
        try:
            self.free()
        except Exception:  # interpreter shutdown: the library may already be gone
            pass
