"""Shared small geometries: the reference test geometry PAR
(pytests/test_kirchhoff.py:22-65) and the 3D ns == nr variant used by the
golden vectors (tests/golden/make_golden.py)."""
import numpy as np

from pylops_amd.utils.wavelets import ricker

PAR = dict(ny=3, nx=12, nz=20, nt=50, dy=3, dx=1, dz=2, dt=0.004,
           nsy=4, nry=3, nsx=6, nrx=2)
v0 = 500.0
y = np.arange(PAR["ny"]) * PAR["dy"]
x = np.arange(PAR["nx"]) * PAR["dx"]
z = np.arange(PAR["nz"]) * PAR["dz"]
t = np.arange(PAR["nt"]) * PAR["dt"]

s2d = np.vstack((np.linspace(x.min(), x.max(), PAR["nsx"]), 2 * np.ones(PAR["nsx"])))
r2d = np.vstack((np.linspace(x.min(), x.max(), PAR["nrx"]), 2 * np.ones(PAR["nrx"])))

_sy, _sx = np.meshgrid(np.linspace(y.min(), y.max(), PAR["nsy"]),
                       np.linspace(x.min(), x.max(), PAR["nsx"]), indexing="ij")
s3d_par = np.vstack((_sy.ravel(), _sx.ravel(), 2 * np.ones(PAR["nsx"] * PAR["nsy"])))
_ry, _rx = np.meshgrid(np.linspace(y.min(), y.max(), PAR["nry"]),
                       np.linspace(x.min(), x.max(), PAR["nrx"]), indexing="ij")
r3d_par = np.vstack((_ry.ravel(), _rx.ravel(), 2 * np.ones(PAR["nrx"] * PAR["nry"])))

_sy, _sx = np.meshgrid(np.linspace(y.min(), y.max(), 2),
                       np.linspace(x.min(), x.max(), 2), indexing="ij")
s3d = np.vstack((_sy.ravel(), _sx.ravel(), 2 * np.ones(4)))
_ry, _rx = np.meshgrid(np.array([y.min() + 1.0, y.max() - 1.0]),
                       np.array([x.min() + 3.0, x.max() - 2.0]), indexing="ij")
r3d = np.vstack((_ry.ravel(), _rx.ravel(), 2 * np.ones(4)))

wav, _, wavc = ricker(t[:21], f0=40)

GOLDEN = __import__("os").path.join(__import__("os").path.dirname(__file__), "golden")


def relerr(a, b):
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


# keyword arguments of each golden case (tests/golden/make_golden.py)
def golden_cases():
    f32 = wav.astype(np.float32)
    return {
        "kin2d_f32": dict(dtype="float32", wav=f32),
        "kin2d_f32w64": dict(dtype="float32"),
        "dyn2d_f64": dict(dynamic=True),
        "dyn2d_ap_f64": dict(dynamic=True, aperture=2, angleaperture=50),
        "dyn2d_ap_f32": dict(dynamic=True, aperture=(1.0, 3.0), angleaperture=(20, 60),
                             dtype="float32", wav=f32),
        "dyn2d_byot_f64": dict(mode="byot", dynamic=True, angleaperture=70),
        "dyn3d_f64": dict(dim3=True, dynamic=True, angleaperture=60),
        "wavfilter2d_f64": dict(wavfilter=True),
        "wavfilter3d_f64": dict(dim3=True, wavfilter=True),
    }


def golden_kwargs(name, g):
    """Constructor keywords (z, x, t, srcs, recs, vel, wav, wavcenter + rest)
    for golden case ``name`` given its loaded npz ``g``."""
    kw = dict(golden_cases()[name])
    dim3 = kw.pop("dim3", False)
    w = kw.pop("wav", wav)
    args = dict(z=z, x=x, t=t, srcs=s3d if dim3 else s2d, recs=r3d if dim3 else r2d,
                vel=v0, wav=w, wavcenter=wavc)
    if dim3:
        args["y"] = y
    if kw.get("mode") == "byot":
        kw["trav"] = (g["in_trav_srcs"], g["in_trav_recs"])
        kw["amp"] = (g["in_amp_srcs"], g["in_amp_recs"])
    else:
        kw.setdefault("mode", "analytic")
    args.update(kw)
    return args
