"""Write the golden input/output vectors under tests/golden/.

Run once, where the reference sources are present (they are not part of this
repository):  PYTHONDONTWRITEBYTECODE=1 python tests/golden/make_golden.py

Only behaviour that no closed form, invariant or library routine pins is
recorded here (the reference-specific quirks):
  * float32 operator arithmetic: float64 tables/weights stored into float32
    (pylops/waveeqprocessing/kirchhoff.py:743-751, 1051, 1101),
  * dynamic weights, angle/offset aperture taper indexing, amplitude epsilon,
    angle tables (kirchhoff.py:316-393, 783-1001),
  * byot unit-spacing gradients (kirchhoff.py:351-360),
  * 3D receiver angle with the source-side sign (kirchhoff.py:387-393),
  * wavelet filtering (kirchhoff.py:659-681),
  * CGLS r1norm/damp conventions and LSQR `var` (cls_basic.py:418-630,
    835-1186).
The reference runs with engine="numpy", which executes the same kernel
functions as its numba engine (kirchhoff.py:1035-1046).
"""
import os
import sys
import warnings

import numpy as np

REF = "/root/reference"
OUT = os.path.dirname(os.path.abspath(__file__))

sys.dont_write_bytecode = True
sys.path.insert(0, REF)
warnings.simplefilter("ignore")

from pylops.optimization.basic import cgls, lsqr  # noqa: E402
from pylops.utils.wavelets import ricker  # noqa: E402
from pylops.waveeqprocessing.kirchhoff import Kirchhoff  # noqa: E402

# small geometry of the reference's Kirchhoff tests (pytests/test_kirchhoff.py:22-65)
PAR = dict(ny=3, nx=12, nz=20, nt=50, dy=3, dx=1, dz=2, dt=0.004,
           nsy=4, nry=3, nsx=6, nrx=2)
v0 = 500.0
y = np.arange(PAR["ny"]) * PAR["dy"]
x = np.arange(PAR["nx"]) * PAR["dx"]
z = np.arange(PAR["nz"]) * PAR["dz"]
t = np.arange(PAR["nt"]) * PAR["dt"]
s2d = np.vstack((np.linspace(x.min(), x.max(), PAR["nsx"]), 2 * np.ones(PAR["nsx"])))
r2d = np.vstack((np.linspace(x.min(), x.max(), PAR["nrx"]), 2 * np.ones(PAR["nrx"])))
wav, _, wavc = ricker(t[:21], f0=40)

# 3D geometry with as many sources as receivers (2x2 each): the reference's
# receiver-angle broadcast (kirchhoff.py:387-393) needs ns == nr
sy3 = np.linspace(y.min(), y.max(), 2)
sx3 = np.linspace(x.min(), x.max(), 2)
syy, sxx = np.meshgrid(sy3, sx3, indexing="ij")
s3d = np.vstack((syy.ravel(), sxx.ravel(), 2 * np.ones(4)))
ry3 = np.array([y.min() + 1.0, y.max() - 1.0])
rx3 = np.array([x.min() + 3.0, x.max() - 2.0])
ryy, rxx = np.meshgrid(ry3, rx3, indexing="ij")
r3d = np.vstack((ryy.ravel(), rxx.ravel(), 2 * np.ones(4)))


def fwd_adj(Op, seed, dtype):
    rng = np.random.default_rng(seed)
    m = rng.standard_normal(Op.shape[1]).astype(dtype)
    d = rng.standard_normal(Op.shape[0]).astype(dtype)
    return m, d, np.asarray(Op.matvec(m)), np.asarray(Op.rmatvec(d))


def kernels(Op, m, dk):
    """The engine kernels alone, with the argument tuples of
    kirchhoff.py:1052-1090 / 1102-1140 (y pre-zeroed in the operator dtype)."""
    y = np.zeros((Op.nsnr, Op.nt), dtype=Op.dtype)
    yi = np.zeros(Op.ni, dtype=Op.dtype)
    dk = dk.reshape(Op.nsnr, Op.nt)
    if Op.dynamic:
        rest = (Op.ns, Op.nr, Op.nt, Op.ni, Op.dt, Op.vel, Op.trav_srcs,
                Op.trav_recs, Op.amp_srcs, Op.amp_recs, Op.aperture[0],
                Op.aperture[1], Op.aperturetap, Op.nz, Op.six, Op.rix,
                Op.angleaperture[0], Op.angleaperture[1], Op.angle_srcs,
                Op.angle_recs)
    else:
        rest = (Op.ns, Op.nr, Op.nt, Op.ni, Op.dt, Op.trav_srcs, Op.trav_recs)
    spread = Op._kirch_matvec(m.ravel(), y, *rest)
    gather = Op._kirch_rmatvec(dk, yi, *rest)
    return dict(spread=np.asarray(spread).ravel(), gather=np.asarray(gather))


def tables(Op):
    out = dict(trav_srcs=Op.trav_srcs, trav_recs=Op.trav_recs)
    if Op.dynamic:
        out.update(amp_srcs=Op.amp_srcs, amp_recs=Op.amp_recs,
                   angle_srcs=Op.angle_srcs, angle_recs=Op.angle_recs,
                   six=np.asarray(Op.six, dtype=np.float64),
                   rix=np.asarray(Op.rix, dtype=np.float64),
                   aperture=np.asarray(Op.aperture, dtype=np.float64),
                   angleaperture=np.asarray(Op.angleaperture, dtype=np.float64))
    return out


cases = {}

# 1. kinematic 2D, float32 operator and float32 wavelet
Op = Kirchhoff(z, x, t, s2d, r2d, v0, wav.astype(np.float32), wavc,
               mode="analytic", dynamic=False, engine="numpy", dtype="float32")
m, d, fwd, adj = fwd_adj(Op, 11, np.float32)
dk = np.random.default_rng(111).standard_normal(Op.shape[0])
cases["kin2d_f32"] = dict(m=m, d=d, fwd=fwd, adj=adj, dk=dk, **kernels(Op, m, dk))

# 2. kinematic 2D, float32 operator, float64 wavelet (upcast path)
Op = Kirchhoff(z, x, t, s2d, r2d, v0, wav, wavc, mode="analytic",
               dynamic=False, engine="numpy", dtype="float32")
m, d, fwd, adj = fwd_adj(Op, 12, np.float32)
dk = np.random.default_rng(112).standard_normal(Op.shape[0])
cases["kin2d_f32w64"] = dict(m=m, d=d, fwd=fwd, adj=adj, dk=dk, **kernels(Op, m, dk))

# 3. dynamic 2D, default apertures, float64
Op = Kirchhoff(z, x, t, s2d, r2d, v0, wav, wavc, mode="analytic",
               dynamic=True, engine="numpy", dtype="float64")
m, d, fwd, adj = fwd_adj(Op, 13, np.float64)
dk = np.random.default_rng(113).standard_normal(Op.shape[0])
cases["dyn2d_f64"] = dict(m=m, d=d, fwd=fwd, adj=adj, dk=dk, **kernels(Op, m, dk), **tables(Op))

# 4. dynamic 2D, offset aperture 2 and angle aperture 50 deg, float64
Op = Kirchhoff(z, x, t, s2d, r2d, v0, wav, wavc, mode="analytic",
               dynamic=True, aperture=2, angleaperture=50, engine="numpy",
               dtype="float64")
m, d, fwd, adj = fwd_adj(Op, 14, np.float64)
dk = np.random.default_rng(114).standard_normal(Op.shape[0])
cases["dyn2d_ap_f64"] = dict(m=m, d=d, fwd=fwd, adj=adj, dk=dk, **kernels(Op, m, dk), **tables(Op))

# 5. dynamic 2D, apertures as tuples, float32 operator and float32 wavelet
Op = Kirchhoff(z, x, t, s2d, r2d, v0, wav.astype(np.float32), wavc,
               mode="analytic", dynamic=True, aperture=(1.0, 3.0),
               angleaperture=(20, 60), engine="numpy", dtype="float32")
m, d, fwd, adj = fwd_adj(Op, 15, np.float32)
dk = np.random.default_rng(115).standard_normal(Op.shape[0])
cases["dyn2d_ap_f32"] = dict(m=m, d=d, fwd=fwd, adj=adj, dk=dk, **kernels(Op, m, dk), **tables(Op))

# 6. dynamic 2D byot with separate trav and amp tables (unit-spacing gradients)
Opa = Kirchhoff(z, x, t, s2d, r2d, v0, wav, wavc, mode="analytic",
                dynamic=True, engine="numpy")
rng = np.random.default_rng(16)
amp_s = Opa.amp_srcs * (1 + 0.1 * rng.random(Opa.amp_srcs.shape))
amp_r = Opa.amp_recs * (1 + 0.1 * rng.random(Opa.amp_recs.shape))
ts_b = Opa.trav_srcs * 1.1
tr_b = Opa.trav_recs * 0.95
Op = Kirchhoff(z, x, t, s2d, r2d, v0, wav, wavc, mode="byot", dynamic=True,
               trav=(ts_b, tr_b), amp=(amp_s, amp_r), angleaperture=70,
               engine="numpy", dtype="float64")
m, d, fwd, adj = fwd_adj(Op, 17, np.float64)
dk = np.random.default_rng(117).standard_normal(Op.shape[0])
cases["dyn2d_byot_f64"] = dict(m=m, d=d, fwd=fwd, adj=adj, dk=dk, **kernels(Op, m, dk), in_trav_srcs=ts_b,
                               in_trav_recs=tr_b, in_amp_srcs=amp_s,
                               in_amp_recs=amp_r, **tables(Op))

# 7. dynamic 3D analytic, ns == nr, float64 (source-sign receiver angle)
Op = Kirchhoff(z, x, t, s3d, r3d, v0, wav, wavc, y=y, mode="analytic",
               dynamic=True, angleaperture=60, engine="numpy", dtype="float64")
m, d, fwd, adj = fwd_adj(Op, 18, np.float64)
dk = np.random.default_rng(118).standard_normal(Op.shape[0])
cases["dyn3d_f64"] = dict(m=m, d=d, fwd=fwd, adj=adj, dk=dk, **kernels(Op, m, dk), **tables(Op))

# 8. wavelet filter, 2D kinematic, float64
Op = Kirchhoff(z, x, t, s2d, r2d, v0, wav, wavc, mode="analytic",
               dynamic=False, wavfilter=True, engine="numpy", dtype="float64")
m, d, fwd, adj = fwd_adj(Op, 19, np.float64)
cases["wavfilter2d_f64"] = dict(m=m, d=d, fwd=fwd, adj=adj, wav=np.asarray(Op.wav))

# 9. wavelet filter, 3D kinematic, float64
Op = Kirchhoff(z, x, t, s3d, r3d, v0, wav, wavc, y=y, mode="analytic",
               dynamic=False, wavfilter=True, engine="numpy", dtype="float64")
m, d, fwd, adj = fwd_adj(Op, 20, np.float64)
cases["wavfilter3d_f64"] = dict(m=m, d=d, fwd=fwd, adj=adj, wav=np.asarray(Op.wav))

# 10. CGLS and LSQR on the 2D dynamic operator, float64
Op = Kirchhoff(z, x, t, s2d, r2d, v0, wav, wavc, mode="analytic",
               dynamic=True, engine="numpy", dtype="float64")
rng = np.random.default_rng(21)
refl = np.zeros((PAR["nx"], PAR["nz"]))
refl[:, PAR["nz"] // 2] = 1.0
refl[:, 3 * PAR["nz"] // 4] = -0.5
dobs = Op @ refl.ravel()
x0 = 0.01 * rng.standard_normal(Op.shape[1])
out = {}
xc, istop, iiter, r1, r2, cost = cgls(Op, dobs, niter=10, tol=0.0)
out.update(cgls_x=xc, cgls_istop=istop, cgls_iiter=iiter, cgls_r1=r1,
           cgls_r2=r2, cgls_cost=cost)
xc, istop, iiter, r1, r2, cost = cgls(Op, dobs, x0=x0, niter=8, damp=0.3, tol=0.0)
out.update(cglsd_x=xc, cglsd_istop=istop, cglsd_iiter=iiter, cglsd_r1=r1,
           cglsd_r2=r2, cglsd_cost=cost)
res = lsqr(Op, dobs, niter=10, calc_var=True)
for k, v in zip(["x", "istop", "iiter", "r1", "r2", "anorm", "acond", "arnorm",
                 "xnorm", "var", "cost"], res):
    out["lsqr_" + k] = v
res = lsqr(Op, dobs, x0=x0, damp=0.2, niter=7, calc_var=True)
for k, v in zip(["x", "istop", "iiter", "r1", "r2", "anorm", "acond", "arnorm",
                 "xnorm", "var", "cost"], res):
    out["lsqrd_" + k] = v
cases["solvers_dyn2d_f64"] = dict(refl=refl, dobs=dobs, x0=x0, **out)

for name, arrs in cases.items():
    np.savez_compressed(os.path.join(OUT, name + ".npz"),
                        **{k: np.asarray(v) for k, v in arrs.items()})
    print(name, {k: np.asarray(v).shape for k, v in arrs.items()
                 if np.asarray(v).size > 1})
