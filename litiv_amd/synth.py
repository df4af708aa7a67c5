"""ctypes wrapper of the synthetic workload generator (liblitiv_synth.so, include/litiv_synth.h)."""
import ctypes as C
import os

import numpy as np

from . import _native

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblitiv_synth.so")

QVGA, HD1080, UHD4K, THERMAL, BATCH64, SERVE = 0, 1, 2, 3, 4, 5
MAX_OBJECTS = 32


class Config(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("width", "height", "channels", "frames")] + [("seed", C.c_uint32)] + [
        (n, C.c_int) for n in ("n_objects", "obj_min", "obj_max", "speed_min", "speed_max", "noise_sigma", "bg_lo",
                               "bg_hi", "bg_cell", "octaves", "drift", "flicker_size", "flicker_amp", "flicker_pmin",
                               "flicker_pmax", "shape", "obj_lo", "obj_hi")]


class Object(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("x0", "y0", "w", "h", "vx", "vy")] + [("color", C.c_int * 3)]


class Scene(C.Structure):
    _fields_ = [("cfg", Config), ("n_obj", C.c_int), ("obj", Object * MAX_OBJECTS), ("fx0", C.c_int), ("fy0", C.c_int)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("liblitiv_synth.so is not built: run __graft_entry__.build()")
        L = _native.load(LIB_PATH)
        L.lvsynth_preset.argtypes = [C.c_int, C.c_int, C.POINTER(Config)]
        L.lvsynth_build_scene.argtypes = [C.POINTER(Config), C.POINTER(Scene)]
        L.lvsynth_frame_host.argtypes = [C.POINTER(Scene), C.c_int, C.c_void_p]
        L.lvsynth_frames_device.argtypes = [C.POINTER(Scene), C.c_int, C.c_int, C.c_void_p, C.c_size_t, C.c_void_p]
        _lib = L
    return _lib


def preset(kind, stream=0, **overrides):
    cfg = Config()
    if lib().lvsynth_preset(kind, stream, C.byref(cfg)) != 0:
        raise ValueError("unknown preset %r/%r" % (kind, stream))
    for k, v in overrides.items():
        setattr(cfg, k, v)
    return cfg


class Sequence:
    """A deterministic synthetic sequence; frames are HWC uint8 (HW for one channel)."""

    def __init__(self, cfg):
        self.cfg = cfg
        self.scene = Scene()
        if lib().lvsynth_build_scene(C.byref(cfg), C.byref(self.scene)) != 0:
            raise ValueError("invalid synthetic config")
        self.shape = (cfg.height, cfg.width) if cfg.channels == 1 else (cfg.height, cfg.width, cfg.channels)
        self.frame_bytes = cfg.height * cfg.width * cfg.channels

    def frame(self, t):
        out = np.empty(self.shape, np.uint8)
        lib().lvsynth_frame_host(C.byref(self.scene), int(t), out.ctypes.data)
        return out

    def frames_device(self, t0, count, dev_ptr, frame_stride=None, stream=None):
        """Writes frames t0..t0+count-1 into device memory at dev_ptr (stride in bytes)."""
        stride = self.frame_bytes if frame_stride is None else frame_stride
        done = 0
        while done < count:
            n = min(count - done, 65535)
            rc = lib().lvsynth_frames_device(C.byref(self.scene), int(t0 + done), n, C.c_void_p(dev_ptr + done * stride),
                                             stride, C.c_void_p(stream or 0))
            if rc != 0:
                raise RuntimeError("lvsynth_frames_device failed (%d)" % rc)
            done += n
