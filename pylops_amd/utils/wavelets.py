"""Wavelets (pylops/utils/wavelets.py:17-22, 167-206)."""
import warnings

import numpy as np

__all__ = ["ricker"]


def _tcrop(t):
    """Drop the last sample of an even-length time axis."""
    if len(t) % 2 == 0:
        t = t[:-1]
        warnings.warn("One sample removed from time axis...")
    return t


def ricker(t, f0=10, taper=None):
    """Ricker wavelet on the symmetric axis [-t[:0:-1], t]; returns
    (w, t_symmetric, index of the peak)."""
    t = _tcrop(np.asarray(t))
    t = np.concatenate((np.flipud(-t[1:]), t), axis=0)
    w = (1 - 2 * (np.pi * f0 * t) ** 2) * np.exp(-((np.pi * f0 * t) ** 2))
    wcenter = np.argmax(np.abs(w))
    if taper is not None:
        w *= taper(len(t))
    return w, t, wcenter
