"""Tapers used by the Kirchhoff aperture (pylops/utils/tapers.py:18-56,
115-151)."""
import numpy as np

__all__ = ["hanningtaper", "taper"]


def hanningtaper(nmask, ntap):
    """Unit mask of length nmask with Hanning edges of ntap samples."""
    if ntap > 0 and (nmask // ntap) < 2:
        ntap_min = nmask // 2 if nmask % 2 == 0 else (nmask - 1) // 2
        raise ValueError(f"ntap={ntap} must be smaller or equal than {ntap_min}")
    han = np.hanning(ntap * 2 - 1)
    edge = han[:ntap]
    return np.concatenate([edge, np.ones(nmask - 2 * ntap), np.flipud(edge)])


def taper(nmask, ntap, tapertype):
    if tapertype == "hanning":
        return hanningtaper(nmask, ntap)
    if tapertype is None or tapertype == "none":
        return np.ones(nmask)
    raise NotImplementedError(f"taper type {tapertype} is not provided by pylops_amd")
