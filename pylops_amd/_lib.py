"""ctypes binding of libkirchx.so (include/kirchx.h).

The library is loaded after ``import torch`` so that it binds to the HIP
runtime torch already loaded (both carry the soname libamdhip64.so.7).
Nothing here falls back to a CPU path: a missing library or device raises.
"""
import ctypes
import os

import torch  # noqa: F401  (must precede the HIP library)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "libkirchx.so")

KX_ABI_VERSION = 1
KX_F32, KX_F64, KX_F16 = 0, 1, 2
KX_EINVAL, KX_ESHAPE, KX_ETYPE, KX_EABI, KX_ELDS = -1, -2, -3, -4, -5

EXPORTS = (
    "kx_version", "kx_last_error", "kx_workspace_bytes", "kx_encode_table",
    "kx_tile_table", "kx_spread", "kx_gather", "kx_conv", "kx_forward",
    "kx_adjoint", "kx_reduce_workspace_bytes", "kx_dot",
)


class KxPlan(ctypes.Structure):
    """Mirror of ``kx_plan`` in include/kirchx.h."""
    _fields_ = [
        ("abi_version", ctypes.c_int32),
        ("dtype", ctypes.c_int32),
        ("tab_fmt", ctypes.c_int32),
        ("dynamic", ctypes.c_int32),
        ("ns", ctypes.c_int32), ("nr", ctypes.c_int32), ("nt", ctypes.c_int32),
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
        ("tz", ctypes.c_int32), ("tx", ctypes.c_int32), ("ty", ctypes.c_int32),
        ("ntz", ctypes.c_int32), ("ntx", ctypes.c_int32), ("nty", ctypes.c_int32),
        ("ntiles", ctypes.c_int64),
        ("neg_trav", ctypes.c_int32),
        ("rblock", ctypes.c_int32),
        ("nh", ctypes.c_int32),
        ("pad0", ctypes.c_int32),
        ("off_s", ctypes.c_void_p),
        ("base_s", ctypes.c_void_p),
        ("off_r", ctypes.c_void_p),
        ("base_r", ctypes.c_void_p),
        ("wp_s", ctypes.c_void_p), ("wq_s", ctypes.c_void_p),
        ("wp_r", ctypes.c_void_p), ("wq_r", ctypes.c_void_p),
        ("izmin", ctypes.c_void_p),
        ("h", ctypes.c_void_p),
        ("hstar", ctypes.c_void_p),
    ]


class KirchxError(RuntimeError):
    pass


_lib = None


def _declare(L):
    P, I32, I64, D = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_double
    PP = ctypes.POINTER(KxPlan)
    sig = {
        "kx_version": ([], ctypes.c_int),
        "kx_last_error": ([], ctypes.c_char_p),
        "kx_workspace_bytes": ([PP], I64),
        "kx_encode_table": ([PP, P, I32, D, I32, P, P, P], ctypes.c_int),
        "kx_tile_table": ([PP, P, I32, I32, P, P], ctypes.c_int),
        "kx_spread": ([PP, P, P, P], ctypes.c_int),
        "kx_gather": ([PP, P, P, P], ctypes.c_int),
        "kx_conv": ([P, P, I64, I32, P, I32, I32, P], ctypes.c_int),
        "kx_forward": ([PP, P, P, P, P], ctypes.c_int),
        "kx_adjoint": ([PP, P, P, P, P], ctypes.c_int),
        "kx_reduce_workspace_bytes": ([], I64),
        "kx_dot": ([P, P, I64, I32, P, P, P], ctypes.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res


def load():
    """Load libkirchx.so (no GPU needed to load it)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise ImportError(
                f"{LIBPATH} is missing: build it with `python -m pylops_amd.build` "
                "(there is no CPU fallback)")
        L = ctypes.CDLL(LIBPATH)
        _declare(L)
        if L.kx_version() != KX_ABI_VERSION:
            raise ImportError(f"libkirchx ABI {L.kx_version()} != {KX_ABI_VERSION}")
        _lib = L
    return _lib


def check(rc, what=""):
    """Map a library status to the reference's exception types: argument
    errors raise ValueError / NotImplementedError, device errors raise
    KirchxError (a RuntimeError)."""
    if rc == 0:
        return
    msg = load().kx_last_error().decode(errors="replace")
    if rc in (KX_EINVAL, KX_ESHAPE, KX_EABI):
        raise ValueError(f"{what}: {msg}")
    if rc in (KX_ETYPE, KX_ELDS):
        raise NotImplementedError(f"{what}: {msg}")
    raise KirchxError(f"{what}: HIP error {rc}: {msg}")


def require_device():
    """The HIP engine needs a GPU; fail loudly otherwise."""
    load()
    if not torch.cuda.is_available():
        raise RuntimeError("pylops_amd needs a HIP device (MI355X); none is visible")


def stream_ptr(device=None):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None
