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

KX_ABI_VERSION = 5
KX_AUX_BYTES = 4096
KX_F32, KX_F64, KX_F16 = 0, 1, 2
KX_ORDER_AUTO, KX_ORDER_NATURAL, KX_ORDER_ROTATED = 0, 1, 2
KX_SIDE_AUTO, KX_SIDE_RECEIVERS, KX_SIDE_SOURCES = 0, 1, 2
KX_DYNAMIC_NOGATE = 2
KX_EINVAL, KX_ESHAPE, KX_ETYPE, KX_EABI, KX_ELDS = -1, -2, -3, -4, -5

EXPORTS = (
    "kx_version", "kx_last_error", "kx_workspace_bytes", "kx_encode_table",
    "kx_tile_table", "kx_spread", "kx_gather", "kx_conv", "kx_forward",
    "kx_adjoint", "kx_reduce_workspace_bytes", "kx_dot",
    "kx_solver_workspace_bytes", "kx_vec", "kx_dot_dev", "kx_lsqr_xw",
    "kx_cgls_scalar", "kx_lsqr_scalar", "kx_batch_dev_bytes", "kx_batch_build",
    "kx_batch_workspace_bytes", "kx_batch_forward", "kx_batch_adjoint",
    "kx_seg_workspace_bytes", "kx_seg_vec", "kx_seg_dot", "kx_cgls_scalar_batch",
    "kx_eikonal", "kx_conv_rows", "kx_xcorr_rows", "kx_spread_models",
    "kx_gather_models", "kx_forward_noalloc",
)

# solver scalar layouts and vector ops (include/kirchx.h)
KX_CG_KOLD = 0
KX_CG_QQ = 1
KX_CG_CC = 2
KX_CG_A = 3
KX_CG_K = 4
KX_CG_B = 5
KX_CG_SS = 6
KX_CG_XX = 7
KX_CG_DAMP2 = 8
KX_CG_TOL = 9
KX_CG_ACTIVE = 10
KX_CG_ITER = 11
KX_CG_NITER = 12
KX_CG_NSC = 16
KX_LS_ALFA = 0
KX_LS_BETA = 1
KX_LS_ANORM = 2
KX_LS_ACOND = 3
KX_LS_DAMP = 4
KX_LS_DAMPSQ = 5
KX_LS_DDNORM = 6
KX_LS_RES2 = 7
KX_LS_XNORM = 8
KX_LS_XXNORM = 9
KX_LS_Z = 10
KX_LS_CS2 = 11
KX_LS_SN2 = 12
KX_LS_RHOBAR = 13
KX_LS_PHIBAR = 14
KX_LS_BNORM = 15
KX_LS_RNORM = 16
KX_LS_R1NORM = 17
KX_LS_R2NORM = 18
KX_LS_ARNORM = 19
KX_LS_ARNORM0 = 20
KX_LS_TEST1 = 21
KX_LS_TEST2 = 22
KX_LS_TEST3 = 23
KX_LS_ATOL = 24
KX_LS_BTOL = 25
KX_LS_CTOL = 26
KX_LS_ISTOP = 27
KX_LS_ITER = 28
KX_LS_NITER = 29
KX_LS_ACTIVE = 30
KX_LS_T1 = 31
KX_LS_T2 = 32
KX_LS_RHO = 33
KX_LS_THETA = 34
KX_LS_PHI = 35
KX_LS_TAU = 36
KX_LS_PSI = 37
KX_LS_UU = 38
KX_LS_VV = 39
KX_LS_WW = 40
KX_LS_VAR = 41
KX_LS_RTOL = 42
KX_LS_BETAPOS = 43
KX_LS_ALFAPOS = 44
KX_LS_NSC = 48
KX_VEC_XPBY = 0
KX_VEC_AXPY = 1
KX_VEC_SCALE = 2



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
        ("sorder", ctypes.c_int32),
        ("off_s", ctypes.c_void_p),
        ("base_s", ctypes.c_void_p),
        ("off_r", ctypes.c_void_p),
        ("base_r", ctypes.c_void_p),
        ("wp_s", ctypes.c_void_p), ("wq_s", ctypes.c_void_p),
        ("wp_r", ctypes.c_void_p), ("wq_r", ctypes.c_void_p),
        ("izmin", ctypes.c_void_p),
        ("h", ctypes.c_void_p),
        ("hstar", ctypes.c_void_p),
        ("wamp", ctypes.c_double),
        ("aux", ctypes.c_void_p),
        ("gsplit", ctypes.c_int32),
        ("sside", ctypes.c_int32),
        ("gtiles", ctypes.c_int32),
        ("pad0", ctypes.c_int32),
        ("wtyp", ctypes.c_double),
    ]


class KxBatch(ctypes.Structure):
    """Mirror of ``kx_batch`` in include/kirchx.h."""
    _fields_ = [
        ("abi_version", ctypes.c_int32),
        ("nprob", ctypes.c_int32),
        ("dtype", ctypes.c_int32),
        ("tab_fmt", ctypes.c_int32),
        ("dynamic", ctypes.c_int32),
        ("neg_trav", ctypes.c_int32),
        ("rblock", ctypes.c_int32),
        ("sorder", ctypes.c_int32),
        ("gsplit", ctypes.c_int32),
        ("nt_max", ctypes.c_int32),
        ("nh_max", ctypes.c_int32),
        ("gtiles", ctypes.c_int32),
        ("model_len", ctypes.c_int64),
        ("data_len", ctypes.c_int64),
        ("n_spread", ctypes.c_int64),
        ("n_gather", ctypes.c_int64),
        ("n_conv", ctypes.c_int64),
        ("grid_spread", ctypes.c_int32),
        ("grid_gather", ctypes.c_int32),
        ("dev", ctypes.c_void_p),
    ]


class KxSegs(ctypes.Structure):
    """Mirror of ``kx_segs`` in include/kirchx.h."""
    _fields_ = [
        ("nseg", ctypes.c_int32),
        ("nchunk", ctypes.c_int32),
        ("off", ctypes.c_void_p),
        ("chunk_seg", ctypes.c_void_p),
        ("chunk_lo", ctypes.c_void_p),
        ("seg_chunk", ctypes.c_void_p),
    ]


class KirchxError(RuntimeError):
    pass


_lib = None


def _declare(L):
    P, I32, I64, D = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_double
    PP = ctypes.POINTER(KxPlan)
    PB = ctypes.POINTER(KxBatch)
    PS = ctypes.POINTER(KxSegs)
    sig = {
        "kx_version": ([], ctypes.c_int),
        "kx_last_error": ([], ctypes.c_char_p),
        "kx_workspace_bytes": ([PP], I64),
        "kx_encode_table": ([PP, P, I32, D, I32, P, P, P], ctypes.c_int),
        "kx_tile_table": ([PP, P, I32, I32, P, P], ctypes.c_int),
        "kx_spread": ([PP, P, P, P], ctypes.c_int),
        "kx_gather": ([PP, P, P, P], ctypes.c_int),
        "kx_spread_models": ([PP, P, P, I32, P], ctypes.c_int),
        "kx_gather_models": ([PP, P, P, I32, P], ctypes.c_int),
        "kx_conv": ([P, P, I64, I32, P, I32, I32, P], ctypes.c_int),
        "kx_conv_rows": ([P, I64, P, I64, I32, P, I64, I32, I32, P], ctypes.c_int),
        "kx_xcorr_rows": ([P, I64, I32, P, I64, I32, P, I64, I32, I32, I32, P], ctypes.c_int),
        "kx_forward": ([PP, P, P, P, P], ctypes.c_int),
        "kx_adjoint": ([PP, P, P, P, P], ctypes.c_int),
        "kx_forward_noalloc": ([PP, P, P, P], ctypes.c_int),
        "kx_reduce_workspace_bytes": ([], I64),
        "kx_dot": ([P, P, I64, I32, P, P, P], ctypes.c_int),
        "kx_solver_workspace_bytes": ([], I64),
        "kx_vec": ([I32, I64, I32, P, P, P, D, P, P, P, P, P], ctypes.c_int),
        "kx_dot_dev": ([I64, I32, P, P, P, P, P, P], ctypes.c_int),
        "kx_lsqr_xw": ([I64, I32, P, P, P, P, P, P], ctypes.c_int),
        "kx_cgls_scalar": ([P, P, P, I32, P], ctypes.c_int),
        "kx_lsqr_scalar": ([P, P, I32, P], ctypes.c_int),
        "kx_batch_dev_bytes": ([PP, I32], I64),
        "kx_batch_build": ([PB, PP, I32, P, P], ctypes.c_int),
        "kx_batch_workspace_bytes": ([PB], I64),
        "kx_batch_forward": ([PB, P, P, P, P], ctypes.c_int),
        "kx_batch_adjoint": ([PB, P, P, P, P], ctypes.c_int),
        "kx_seg_workspace_bytes": ([PS], I64),
        "kx_seg_vec": ([I32, PS, I32, P, P, P, I32, I32, D, I32, I32, P, P], ctypes.c_int),
        "kx_seg_dot": ([PS, I32, P, P, P, I32, I32, I32, P, P], ctypes.c_int),
        "kx_cgls_scalar_batch": ([P, I32, P, P, I32, I32, P], ctypes.c_int),
        "kx_eikonal": ([I32, I32, I32, I32, D, D, D, P, I32, P, P, I32, I32, P,
                        ctypes.POINTER(I32), P], ctypes.c_int),  # iters: int32[2]
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
