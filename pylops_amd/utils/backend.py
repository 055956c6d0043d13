"""Array-module helpers: numpy on the host, torch tensors on the device.

Mirrors the reference rule "the output lives in the module of the input"
(pylops/utils/backend.py:126-150): numpy in -> numpy out, torch in -> torch
out (on the tensor's device).
"""
import numpy as np
import torch

__all__ = ["is_torch", "to_device", "like_input", "vstack", "device_of"]


def is_torch(x):
    return isinstance(x, torch.Tensor)


def torch_dtype(dtype):
    dt = np.dtype(dtype)
    if dt == np.float32:
        return torch.float32
    if dt == np.float64:
        return torch.float64
    raise NotImplementedError(f"dtype {dt} is not supported (float32 or float64)")


def device_of(x):
    return x.device if is_torch(x) else None


def to_device(x, dtype, device):
    """Contiguous torch tensor of ``dtype`` on ``device`` (a copy only when
    needed)."""
    tdt = torch_dtype(dtype)
    if is_torch(x):
        return x.to(device=device, dtype=tdt).contiguous()
    return torch.as_tensor(np.ascontiguousarray(x), dtype=tdt, device=device)


def like_input(y, x):
    """Return device tensor ``y`` in the array module of ``x``."""
    if is_torch(x):
        return y if y.device == x.device else y.to(x.device)
    return y.detach().cpu().numpy()


def vstack(arrs):
    if arrs and is_torch(arrs[0]):
        return torch.vstack(arrs)
    return np.vstack(arrs)
