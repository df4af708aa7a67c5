"""Loading of the package's HIP libraries."""
import ctypes as C


def load(path):
    """ctypes handle of one of the package's HIP libraries.

    When PyTorch is installed it is imported first. Its wheel ships its own libamdhip64 under the
    soname of the system runtime (libamdhip64.so.7) but links it by the unversioned name: loaded
    after torch, this library binds to the runtime torch already holds, and device pointers and
    streams pass freely between the two; loaded before torch, the process ends up with two HIP
    runtimes and torch's finds no device."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    return C.CDLL(path)
