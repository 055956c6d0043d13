"""GPU parity of the HIP Kirchhoff operator against the CPU oracle and the
reference's golden vectors.

Tolerances (relative L2):
  float64 operator vs float64 oracle / golden     1e-10
  float32 operator vs oracle                      1e-5
  float16-offset tables vs float64 oracle         1e-3
  dottest                                         1e-6 (float64), 1e-5 (float32)
"""
import os

import numpy as np
import pytest

from oracle import oracle as O
from tests import geometry as G

pytestmark = pytest.mark.gpu


def _K(**kw):
    import pylops_amd
    return pylops_amd.Kirchhoff(**kw)


def _load(name):
    return np.load(os.path.join(G.GOLDEN, name + ".npz"))


def _dottest(Op, rtol, seed=0):
    rng = np.random.default_rng(seed)
    u = rng.standard_normal(Op.shape[1]).astype(Op.dtype)
    v = rng.standard_normal(Op.shape[0]).astype(Op.dtype)
    yy = np.dot(np.asarray(Op.matvec(u), dtype=np.float64), v.astype(np.float64))
    xx = np.dot(u.astype(np.float64), np.asarray(Op.rmatvec(v), dtype=np.float64))
    assert abs(xx - yy) <= rtol * max(abs(xx), abs(yy)), (xx, yy)


@pytest.mark.parametrize("name", sorted(G.golden_cases()))
def test_golden(name):
    g = _load(name)
    kw = G.golden_kwargs(name, g)
    Op = _K(**kw)
    oracle = O.KirchhoffOracle(**kw)
    f32 = Op.dtype == np.float32
    tol = 1e-5 if f32 else 1e-10
    fwd = Op @ g["m"]
    adj = Op.H @ g["d"]
    # output types as the reference's: a float32 operator with a float64
    # wavelet returns float64 data (scipy's upcast, C1D:99-106) and a
    # float32 image
    assert fwd.dtype == g["fwd"].dtype and adj.dtype == g["adj"].dtype, (fwd.dtype, adj.dtype)
    assert G.relerr(fwd, g["fwd"]) < tol
    assert G.relerr(adj, g["adj"]) < tol
    if "spread" in g.files:
        assert G.relerr(Op.spread(g["m"]), g["spread"]) < tol
        assert G.relerr(Op.gather(g["dk"]), oracle.gather(g["dk"])) < tol
    _dottest(Op, 1e-5 if f32 else 1e-9)


@pytest.mark.parametrize("dtype", ["float64", "float32"])
@pytest.mark.parametrize("dynamic", [False, True])
@pytest.mark.parametrize("dim3", [False, True])
def test_vs_oracle_par(dtype, dynamic, dim3):
    kw = dict(z=G.z, x=G.x, t=G.t, srcs=G.s3d if dim3 else G.s2d,
              recs=G.r3d if dim3 else G.r2d, vel=G.v0, wav=G.wav, wavcenter=G.wavc,
              mode="analytic", dynamic=dynamic, dtype=dtype, angleaperture=60)
    if dim3:
        kw["y"] = G.y
    Op = _K(**kw)
    orc = O.KirchhoffOracle(**kw)
    rng = np.random.default_rng(5)
    m = rng.standard_normal(Op.shape[1])
    d = rng.standard_normal(Op.shape[0])
    tol = 1e-5 if dtype == "float32" else 1e-10
    assert G.relerr(Op.matvec(m.astype(dtype)), orc.matvec(m)) < tol
    assert G.relerr(Op.rmatvec(d.astype(dtype)), orc.rmatvec(d)) < tol
    _dottest(Op, 1e-5 if dtype == "float32" else 1e-9)


def test_reference_dottests():
    """pytests/test_kirchhoff.py:215-317 (analytic and byot, 2D and 3D)."""
    from pylops_amd.utils import dottest
    for dim3 in (False, True):
        s, r = (G.s3d_par, G.r3d_par) if dim3 else (G.s2d, G.r2d)
        yy = G.y if dim3 else None
        Op = _K(z=G.z, x=G.x, t=G.t, srcs=s, recs=r, vel=G.v0, wav=G.wav,
                wavcenter=G.wavc, y=yy, mode="analytic")
        assert dottest(Op, Op.shape[0], Op.shape[1])
        trav = (Op.trav_srcs[:, :, None] + Op.trav_recs[:, None, :]).reshape(Op.ni, -1)
        Op1 = _K(z=G.z, x=G.x, t=G.t, srcs=s, recs=r, vel=G.v0, wav=G.wav,
                 wavcenter=G.wavc, y=yy, mode="byot", trav=trav)
        assert dottest(Op1, Op1.shape[0], Op1.shape[1])
        rng = np.random.default_rng(1)
        m = rng.standard_normal(Op.shape[1])
        d = rng.standard_normal(Op.shape[0])
        np.testing.assert_allclose(Op @ m, Op1 @ m, atol=1e-10)
        np.testing.assert_allclose(Op.H @ d, Op1.H @ d, atol=1e-10)


@pytest.mark.parametrize("dtype", ["float32", "float64"])
@pytest.mark.parametrize("dim3", [False, True])
def test_single_table_vs_oracle(dtype, dim3):
    """One combined (ni, ns*nr) byot table (K:683-725, the reference's
    _trav_kirch_matvec / _trav_kirch_rmatvec) against the oracle's
    single-table kernels on the same table, forward and adjoint."""
    s, r = (G.s3d_par, G.r3d_par) if dim3 else (G.s2d, G.r2d)
    yy = G.y if dim3 else None
    ts, tr = O.analytic_tables(G.z, G.x, s, r, G.v0, yy)[:2]
    rng = np.random.default_rng(31)
    # not separable: a per-(source, receiver) shift on top of ts + tr
    trav = (ts[:, :, None] + tr[:, None, :]).reshape(ts.shape[0], -1)
    trav = trav + 0.002 * rng.random(trav.shape[1])[None, :]
    wav = G.wav.astype(dtype)
    kw = dict(z=G.z, x=G.x, t=G.t, srcs=s, recs=r, vel=G.v0, wav=wav, wavcenter=G.wavc,
              mode="byot", trav=trav, dtype=dtype)
    if dim3:
        kw["y"] = G.y
    Op = _K(**kw)
    orc = O.KirchhoffOracle(**kw)
    assert not orc.travsrcrec
    m = rng.standard_normal(Op.shape[1])
    d = rng.standard_normal(Op.shape[0])
    tol = 1e-5 if dtype == "float32" else 1e-10
    assert G.relerr(Op @ m.astype(dtype), orc.matvec(m.astype(dtype))) < tol
    assert G.relerr(Op.H @ d.astype(dtype), orc.rmatvec(d.astype(dtype))) < tol


def test_engines_and_their_errors():
    """engine: the reference names ("numpy", "numba", "cuda") run the HIP
    kernels (bitwise the default engine's products), an unknown name raises
    KeyError and "cuda" with one combined table NotImplementedError, as in
    the reference (K:1003-1046, pytests/test_kirchhoff.py:219)."""
    kw = dict(z=G.z, x=G.x, t=G.t, srcs=G.s2d, recs=G.r2d, vel=G.v0, wav=G.wav,
              wavcenter=G.wavc, mode="analytic", dtype="float32")
    Op = _K(**kw)
    rng = np.random.default_rng(41)
    m = rng.standard_normal(Op.shape[1]).astype(np.float32)
    d = rng.standard_normal(Op.shape[0]).astype(np.float32)
    for engine in ("numpy", "numba", "cuda"):
        Oe = _K(engine=engine, **kw)
        assert Oe.engine == engine
        np.testing.assert_array_equal(Oe @ m, Op @ m)
        np.testing.assert_array_equal(Oe.H @ d, Op.H @ d)
    with pytest.raises(KeyError):
        _K(engine="fortran", **kw)
    trav = (Op.trav_srcs[:, :, None] + Op.trav_recs[:, None, :]).reshape(Op.ni, -1)
    kws = dict(kw, mode="byot", trav=trav)
    with pytest.raises(NotImplementedError):
        _K(engine="cuda", **kws)
    _K(engine="numba", **kws)  # the other engines take a single table


@pytest.mark.parametrize("dynamic,dim3", [(False, False), (True, False), (False, True)])
def test_trav_vs_travsrcrec(dynamic, dim3):
    """pytests/test_kirchhoff.py:320-460 restated (2D kinematic and dynamic,
    3D kinematic, as there): the operator built with the separable tables
    and the one given a combined table (plus the amplitude tables when
    dynamic) with angleaperture=None agree, forward and adjoint."""
    s, r = (G.s3d_par, G.r3d_par) if dim3 else (G.s2d, G.r2d)
    kw = dict(z=G.z, x=G.x, t=G.t, srcs=s, recs=r, vel=G.v0, wav=G.wav, wavcenter=G.wavc,
              mode="analytic", dynamic=dynamic, angleaperture=None)
    if dim3:
        kw["y"] = G.y
    Dop = _K(**kw)
    ni = Dop.ni
    trav = (np.asarray(Dop.trav_srcs)[:, :, None]
            + np.asarray(Dop.trav_recs)[:, None, :]).reshape(ni, -1)
    amp = (np.asarray(Dop.amp_srcs), np.asarray(Dop.amp_recs)) if dynamic else None
    D1op = _K(**dict(kw, trav=trav, amp=amp))
    rng = np.random.default_rng(43)
    xx = rng.normal(0, 1, Dop.shape[1])
    yy = rng.normal(0, 1, Dop.shape[0])
    np.testing.assert_array_almost_equal(Dop @ xx, D1op @ xx, decimal=2)
    np.testing.assert_array_almost_equal(Dop.H @ yy, D1op.H @ yy, decimal=2)


@pytest.mark.parametrize("dim3", [False, True])
@pytest.mark.parametrize("dynamic", [False, True])
@pytest.mark.parametrize("nt", [50, 41])
def test_one_source_one_receiver(dim3, dynamic, nt):
    """The smallest geometry: one source and one receiver (one trace, one
    receiver per workgroup, one source per adjoint pass), with traces of 50
    samples and of the wavelet's own length (41 taps), against the oracle.
    (A trace shorter than the wavelet makes the reference's Convolve1D take
    its long-filter branch, whose data take the filter's shape: no
    Kirchhoff operator is built that way.)"""
    t = np.arange(nt) * G.PAR["dt"]
    if dim3:
        s = np.array([[G.y[1]], [G.x[3]], [2.0]])
        r = np.array([[G.y[2]], [G.x[9]], [2.0]])
    else:
        s = np.array([[G.x[3]], [2.0]])
        r = np.array([[G.x[9]], [2.0]])
    kw = dict(z=G.z, x=G.x, t=t, srcs=s, recs=r, vel=G.v0, wav=G.wav.astype(np.float32),
              wavcenter=G.wavc, mode="analytic", dynamic=dynamic, dtype="float32")
    if dim3:
        kw["y"] = G.y
    if dim3 and dynamic:
        kw["angleaperture"] = 60
    Op = _K(**kw)
    orc = O.KirchhoffOracle(**kw)
    rng = np.random.default_rng(47)
    m = rng.standard_normal(Op.shape[1]).astype(np.float32)
    d = rng.standard_normal(Op.shape[0]).astype(np.float32)
    assert Op.shape == (nt, Op.ni)
    assert G.relerr(Op @ m, orc.matvec(m)) < 1e-5
    assert G.relerr(Op.H @ d, orc.rmatvec(d)) < 1e-5


def test_attribute_reassignment_rebuilds_tables():
    Op = _K(z=G.z, x=G.x, t=G.t, srcs=G.s2d, recs=G.r2d, vel=G.v0, wav=G.wav,
            wavcenter=G.wavc, mode="analytic")
    m = np.random.default_rng(2).standard_normal(Op.shape[1])
    d0 = Op @ m
    Op.trav_srcs = np.asarray(Op.trav_srcs) * 1.05
    orc = O.KirchhoffOracle(G.z, G.x, G.t, G.s2d, G.r2d, G.v0, G.wav, G.wavc, mode="byot",
                            trav=(np.asarray(Op.trav_srcs), np.asarray(Op.trav_recs)))
    d1 = Op @ m
    assert G.relerr(d1, orc.matvec(m)) < 1e-10
    assert G.relerr(d0, d1) > 1e-3
    # an edit in place could not reach the device tables: it raises
    with pytest.raises(ValueError, match="read-only"):
        Op.trav_srcs[:, 0] += 0.1
    # float64 LDS atomics add in a launch-dependent order: equal to rounding
    assert G.relerr(Op @ m, d1) < 1e-13


@pytest.mark.parametrize("scale", [1e-15, 1e15])
def test_weights_in_extreme_units(scale):
    """Amplitude tables in tiny or huge units (weights of 1e-30 or 1e30):
    the float32 forward's fixed-point scale would leave the float range
    (m 2^e overflows, or the small values underflow), so the launch
    accumulates in float64; the result still matches the oracle."""
    import pylops_amd
    ts, tr, ds, dr, _, _ = pylops_amd.Kirchhoff._traveltime_table(G.z, G.x, G.s2d, G.r2d,
                                                                  G.v0, mode="analytic")
    amp = (scale / np.maximum(ds, 1.0), scale / np.maximum(dr, 1.0))
    kw = dict(z=G.z, x=G.x, t=G.t, srcs=G.s2d, recs=G.r2d, vel=G.v0, wav=G.wav.astype(np.float32),
              wavcenter=G.wavc, mode="byot", trav=(ts, tr), amp=amp, dynamic=True,
              angleaperture=70, dtype="float32")
    Op = _K(**kw)
    orc = O.KirchhoffOracle(**kw)
    m = np.random.default_rng(8).standard_normal(Op.shape[1]).astype(np.float32)
    got = np.asarray(Op @ m)
    assert np.all(np.isfinite(got))
    assert G.relerr(got, orc.matvec(m)) < 1e-5


def _fixed_point_used(Op):
    """Whether the operator's last float32 forward accumulated in fixed point
    (the scale the forward wrote to its scratch: [e, fixed])."""
    return int(Op._dev_tables["aux"][1].item()) == 1


def _per_trace_relerr(got, want, keep):
    """Largest relative L2 error over the traces (rows) on the samples
    `keep` marks, among traces whose kept samples are not all zero."""
    got = np.where(keep, np.asarray(got, np.float64), 0.0)
    want = np.where(keep, np.asarray(want, np.float64), 0.0)
    nw = np.linalg.norm(want, axis=-1)
    ok = nw > 0
    return float((np.linalg.norm(got - want, axis=-1)[ok] / nw[ok]).max())


def test_forward_wide_dynamic_range_model():
    """A model of 1e-5-sized values with one outlier of 1 (max/median ~1.5e5):
    the fixed-point rounding, set by the outlier, would be ~1e-4 of every
    sample the outlier does not reach; such a launch accumulates in float64,
    and each trace away from the outlier's isochron matches the float64
    oracle as the reference's float32 sums do.  The layered model of the
    workload below keeps the fixed-point path."""
    import pylops_amd
    kw, _ = _config1()
    Op = pylops_amd.Kirchhoff(**kw)
    orc = O.KirchhoffOracle(**{**kw, "dtype": "float64"})
    orc32 = O.KirchhoffOracle(**kw)
    nx, nz = 101, 100
    rng = np.random.default_rng(21)
    m = 1e-5 * rng.standard_normal((nx, nz))
    i0 = (50, 60)
    m[i0] = 1.0
    assert np.abs(m).max() / np.median(np.abs(m)) > 1e5
    m = m.astype(np.float32).ravel()
    got = np.asarray(Op @ m).reshape(11 * 51, -1)
    assert not _fixed_point_used(Op)
    want = orc.matvec(m.astype(np.float64)).reshape(11 * 51, -1)
    ref32 = orc32.matvec(m).reshape(11 * 51, -1)
    # samples farther from the outlier's arrival than the wavelet reaches
    ts, tr, _, _, _, _ = pylops_amd.Kirchhoff._traveltime_table(
        kw["z"], kw["x"], kw["srcs"], kw["recs"], kw["vel"], mode="analytic")
    ii = i0[0] * nz + i0[1]
    t_out = (ts[ii][:, None] + tr[ii][None, :]).reshape(-1, 1) / 0.004
    k = np.arange(got.shape[1])[None, :]
    keep = np.abs(k - t_out) > len(Op.cop.h) // 2 + 3
    err = _per_trace_relerr(got, want, keep)
    err32 = _per_trace_relerr(ref32, want, keep)
    assert err < 1e-5, (err, err32)
    # the layered workload model stays in fixed point
    refl = 0.01 * rng.standard_normal((nx, nz))
    refl[:, 40] += 0.7
    Op @ refl.astype(np.float32).ravel()
    assert _fixed_point_used(Op)


def test_forward_wide_amplitude_spread():
    """Dynamic byot tables whose amplitudes fall by 1e4 with depth on both
    sides: per trace (every trace has its own scale), the float32 forward
    matches the float64 oracle to 1e-5 whichever accumulation the launch
    picks."""
    import pylops_amd
    kw, _ = _config1()
    ts, tr, ds, dr, _, _ = pylops_amd.Kirchhoff._traveltime_table(
        kw["z"], kw["x"], kw["srcs"], kw["recs"], kw["vel"], mode="analytic")
    depth = np.tile(kw["z"], len(kw["x"]))[:, None] / kw["z"][-1]
    fade = 10.0 ** (-4.0 * depth)
    amp = (fade / np.sqrt(ds + 10.0), fade / np.sqrt(dr + 10.0))
    kwd = dict(kw, mode="byot", trav=(ts, tr), amp=amp, dynamic=True, angleaperture=80)
    Op = pylops_amd.Kirchhoff(**kwd)
    orc = O.KirchhoffOracle(**{**kwd, "dtype": "float64"})
    m = np.random.default_rng(22).standard_normal(Op.shape[1]).astype(np.float32)
    got = np.asarray(Op @ m).reshape(11 * 51, -1)
    want = orc.matvec(m.astype(np.float64)).reshape(11 * 51, -1)
    err = _per_trace_relerr(got, want, np.ones_like(got, dtype=bool))
    assert err < 1e-5, (err, _fixed_point_used(Op))


def _config1(dtype="float32", tabledtype=None):
    """2D constant-velocity LSM workload: grid 101 x 100 at 10 m, nt=651 at
    4 ms, 11 sources and 51 receivers on the surface, 20 Hz Ricker."""
    x = np.arange(101) * 10.0
    z = np.arange(100) * 10.0
    t = np.arange(651) * 0.004
    srcs = np.vstack((np.linspace(x[0], x[-1], 11), np.zeros(11)))
    recs = np.vstack((np.linspace(x[0], x[-1], 51), np.zeros(51)))
    wav, _, wc = O.ricker(t[:41], f0=20)
    return dict(z=z, x=x, t=t, srcs=srcs, recs=recs, vel=2000.0,
                wav=wav.astype(np.float32), wavcenter=wc, mode="analytic",
                dtype=dtype), tabledtype


def test_config1_parity_fp32_fp16():
    kw, _ = _config1()
    rng = np.random.default_rng(1001)
    refl = 0.01 * rng.standard_normal((101, 100))
    for iz in rng.integers(10, 90, 3):
        refl[:, iz] += rng.uniform(-1, 1)
    orc64 = O.KirchhoffOracle(**{**kw, "dtype": "float64"})
    d64 = orc64.matvec(refl.ravel())
    a64 = orc64.rmatvec(d64)
    Op = _K(**kw)
    d32 = Op @ refl.astype(np.float32)
    a32 = Op.H @ d32
    assert G.relerr(d32, d64) < 1e-5
    assert G.relerr(a32, a64) < 1e-5
    orc32 = O.KirchhoffOracle(**kw)
    assert G.relerr(d32, orc32.matvec(refl.astype(np.float32))) < 1e-5
    Op16 = _K(**kw, tabledtype="float16")
    d16 = Op16 @ refl.astype(np.float32)
    a16 = Op16.H @ d64.astype(np.float32)
    assert G.relerr(d16, d64) < 1e-3
    assert G.relerr(a16, a64) < 1e-3
    _dottest(Op, 1e-5)
    _dottest(Op16, 1e-5)


@pytest.mark.parametrize("case", ["2d", "3d", "tiny_units", "huge_units"])
def test_dynamic_fp16_tables_vs_oracle(case):
    """Float16 tables of a dynamic operator: float16 receiver offsets and
    receiver weights (scaled to at most 1, the source weights by the same
    factor, so that amplitude tables in any units fit float16) against the
    float64 oracle at the float16-table bar, forward and adjoint; the dot
    test holds to float32 rounding."""
    import torch
    import pylops_amd
    if case == "3d":
        # a 20 Hz wavelet: at 40 Hz and 4 ms a sample is ~1 rad of phase, and
        # the float16 offsets alone leave ~1e-3 on this 16-trace geometry
        wav20, _, wc20 = O.ricker(G.t[:21], f0=20)
        kw = dict(z=G.z, x=G.x, y=G.y, t=G.t, srcs=G.s3d, recs=G.r3d, vel=G.v0,
                  wav=wav20.astype(np.float32), wavcenter=wc20, mode="analytic",
                  dynamic=True, angleaperture=60, dtype="float32")
    else:
        kw, _ = _config1()
        kw = dict(kw, dynamic=True, angleaperture=70)
        if case != "2d":
            scale = 1e-15 if case == "tiny_units" else 1e15
            ts, tr, ds, dr, _, _ = pylops_amd.Kirchhoff._traveltime_table(
                kw["z"], kw["x"], kw["srcs"], kw["recs"], kw["vel"], mode="analytic")
            kw.update(mode="byot", trav=(ts, tr),
                      amp=(scale / np.sqrt(ds + 10.0), scale / np.sqrt(dr + 10.0)))
    Op = pylops_amd.Kirchhoff(tabledtype="float16", **kw)
    assert Op.plan.tab_fmt == 2 and Op._dev_tables["wp_r"].dtype == torch.float16
    orc = O.KirchhoffOracle(**{**kw, "dtype": "float64"})
    rng = np.random.default_rng(51)
    m = rng.standard_normal(Op.shape[1]).astype(np.float32)
    d = rng.standard_normal(Op.shape[0]).astype(np.float32)
    assert G.relerr(Op @ m, orc.matvec(m.astype(np.float64))) < 1e-3
    assert G.relerr(Op.H @ d, orc.rmatvec(d.astype(np.float64))) < 1e-3
    _dottest(Op, 1e-5)


def test_negative_traveltimes_truncate_like_reference():
    """byot tables with negative times take the truncation path (K:743-751):
    a time in (-1, 0) samples lands on sample 0 with a negative weight.  The
    shifts are chosen off the sample grid: a time within one float64 ulp of
    -1 sample is a discontinuity of the truncation rule and rounds either
    way."""
    Op0 = _K(z=G.z, x=G.x, t=G.t, srcs=G.s2d, recs=G.r2d, vel=G.v0, wav=G.wav,
             wavcenter=G.wavc, mode="analytic")
    ts = np.asarray(Op0.trav_srcs) - 0.01113
    tr = np.asarray(Op0.trav_recs) - 0.00871
    kw = dict(z=G.z, x=G.x, t=G.t, srcs=G.s2d, recs=G.r2d, vel=G.v0, wav=G.wav,
              wavcenter=G.wavc, mode="byot", trav=(ts, tr))
    Op = _K(**kw)
    orc = O.KirchhoffOracle(**kw)
    rng = np.random.default_rng(4)
    m = rng.standard_normal(Op.shape[1])
    d = rng.standard_normal(Op.shape[0])
    assert G.relerr(Op.spread(m), orc.spread(m)) < 1e-10
    assert G.relerr(Op.gather(d), orc.gather(d)) < 1e-10


def test_conv_matches_oracle():
    import torch
    from pylops_amd.signalprocessing import Convolve1D
    rng = np.random.default_rng(6)
    x = rng.standard_normal((7, 300))
    for nh, off in ((81, 40), (21, 3), (20, 12)):
        h = rng.standard_normal(nh)
        C = Convolve1D((7, 300), h=h, offset=off, axis=1, dtype="float64")
        hp = O.padded_filter(h, off)
        np.testing.assert_allclose(C @ x, O.conv_same(x, hp), atol=1e-12)
        np.testing.assert_allclose(C.H @ x, O.conv_same(x, hp[::-1]), atol=1e-12)
        xt = torch.as_tensor(x, dtype=torch.float32, device="cuda")
        C32 = Convolve1D((7, 300), h=h, offset=off, axis=1, dtype="float32")
        y = C32 @ xt
        assert isinstance(y, torch.Tensor) and y.is_cuda
        assert G.relerr(y.cpu().numpy(), O.conv_same(x, hp)) < 1e-6


@pytest.mark.parametrize("nt", [3, 300, 512, 513, 1501])
@pytest.mark.parametrize("nh", [1, 3, 81, 601, 1023])
def test_conv_lengths(nt, nh):
    """The convolution kernel against numpy's direct convolution for traces
    shorter and longer than one workgroup's 512 outputs and filters shorter
    and longer than them (float32 and float64)."""
    import torch
    from pylops_amd import _lib
    L = _lib.load()
    rng = np.random.default_rng(nt * 7 + nh)
    x = rng.standard_normal((5, nt))
    h = rng.standard_normal(nh)
    ref = O.conv_same(x, h)
    for dt, kx, tol in ((torch.float32, _lib.KX_F32, 1e-6), (torch.float64, _lib.KX_F64, 1e-13)):
        xt = torch.as_tensor(x, dtype=dt, device="cuda")
        ht = torch.as_tensor(h, dtype=dt, device="cuda")
        yt = torch.full_like(xt, float("nan"))
        rc = L.kx_conv(_lib.ptr(xt), _lib.ptr(yt), 5, nt, _lib.ptr(ht), nh, kx,
                       _lib.stream_ptr())
        _lib.check(rc, "kx_conv")
        y = yt.cpu().numpy()
        assert np.all(np.isfinite(y))
        assert G.relerr(y, ref) < tol, (nt, nh, dt)


def test_dot_fp64_accumulation():
    import torch
    from pylops_amd import _lib
    L = _lib.load()
    rng = np.random.default_rng(7)
    for n in (1, 1000, 1_000_003):
        a = rng.standard_normal(n).astype(np.float32)
        b = rng.standard_normal(n).astype(np.float32)
        ta = torch.as_tensor(a, device="cuda")
        tb = torch.as_tensor(b, device="cuda")
        out = torch.empty(1, dtype=torch.float64, device="cuda")
        work = torch.empty(L.kx_reduce_workspace_bytes() // 8, dtype=torch.float64, device="cuda")
        rc = L.kx_dot(_lib.ptr(ta), _lib.ptr(tb), n, _lib.KX_F32, _lib.ptr(out), _lib.ptr(work),
                      _lib.stream_ptr())
        _lib.check(rc)
        ref = np.dot(a.astype(np.float64), b.astype(np.float64))
        assert abs(float(out) - ref) <= 1e-12 * max(1.0, np.abs(a).dot(np.abs(b)))


@pytest.mark.parametrize("dynamic", [False, True])
def test_spread_lane_orders_and_receiver_blocks(dynamic):
    """Every receiver block size and lane order of the forward, with a last
    block that is only partly filled (11 receivers), matches the oracle; the
    fixed-point sums do not depend on the order of the adds, so all
    variants agree bit for bit."""
    from pylops_amd import _lib
    recs = np.vstack((np.linspace(G.x.min(), G.x.max(), 11), 2 * np.ones(11)))
    srcs = np.vstack((np.linspace(G.x.min(), G.x.max(), 3), 2 * np.ones(3)))
    kw = dict(z=G.z, x=G.x, t=G.t, srcs=srcs, recs=recs, vel=G.v0, wav=G.wav,
              wavcenter=G.wavc, mode="analytic", dynamic=dynamic, dtype="float32",
              angleaperture=60)
    Op = _K(**kw)
    orc = O.KirchhoffOracle(**kw)
    m = np.random.default_rng(9).standard_normal(Op.shape[1]).astype(np.float32)
    want = orc.spread(m)
    outs = []
    for order in (_lib.KX_ORDER_NATURAL, _lib.KX_ORDER_ROTATED):
        for rb in (1, 2, 4, 8, 16):
            Op.plan.sorder = order
            Op.plan.rblock = rb
            got = np.asarray(Op.spread(m))
            assert G.relerr(got, want) < 1e-5, (order, rb)
            outs.append(got)
    for got in outs[1:]:
        np.testing.assert_array_equal(got, outs[0])


@pytest.mark.parametrize("case", ["kin2d", "kin2d_f16", "kin3d", "neg_byot", "nonfinite",
                                  "dynamic", "dynamic_f16_neg", "dynamic3d"])
def test_spread_source_blocks_bitwise(case):
    """Workgroups blocking sources (KX_SIDE_SOURCES: one receiver and a block
    of sources, the side tables exchanged) give traces bitwise equal to
    blocks of receivers, for every receiver block and lane order, with more
    and fewer sources than receivers, float16 tables, 3D and the
    negative-time path and dynamic weights (their products formed per step,
    so that they round as with blocks of receivers); a non-finite model
    keeps its NaN pattern."""
    import pylops_amd
    from pylops_amd import _lib
    srcs = np.vstack((np.linspace(G.x.min(), G.x.max(), 5), 2 * np.ones(5)))
    recs = np.vstack((np.linspace(G.x.min(), G.x.max(), 7), 2 * np.ones(7)))
    kw = dict(z=G.z, x=G.x, t=G.t, srcs=srcs, recs=recs, vel=G.v0, wav=G.wav,
              wavcenter=G.wavc, mode="analytic", dtype="float32")
    if case == "kin2d_f16":
        kw.update(srcs=recs, recs=srcs, tabledtype="float16")
    elif case == "kin3d":
        kw.update(y=G.y, srcs=G.s3d_par, recs=G.r3d_par)
    elif case == "neg_byot":
        ts, tr, _, _, _, _ = pylops_amd.Kirchhoff._traveltime_table(G.z, G.x, srcs, recs,
                                                                    G.v0, mode="analytic")
        kw.update(mode="byot", trav=(ts - 0.01113, tr - 0.00871), vel=None)
    elif case == "dynamic":
        kw.update(dynamic=True, angleaperture=60)
    elif case == "dynamic_f16_neg":
        ts, tr, ds, dr, _, _ = pylops_amd.Kirchhoff._traveltime_table(G.z, G.x, srcs, recs,
                                                                      G.v0, mode="analytic")
        kw.update(mode="byot", trav=(ts - 0.01113, tr - 0.00871),
                  amp=(1.0 / np.maximum(ds, 1.0), 1.0 / np.maximum(dr, 1.0)),
                  dynamic=True, tabledtype="float16")
    elif case == "dynamic3d":
        kw.update(y=G.y, srcs=G.s3d, recs=G.r3d, dynamic=True, angleaperture=70)
    Op = _K(**kw)
    orc = O.KirchhoffOracle(**{k: v for k, v in kw.items() if k != "tabledtype"})
    m = np.random.default_rng(21).standard_normal(Op.shape[1]).astype(np.float32)
    if case == "nonfinite":
        m[40] = np.nan
    Op.plan.sside = _lib.KX_SIDE_RECEIVERS
    Op.plan.sorder = _lib.KX_ORDER_NATURAL
    Op.plan.rblock = 1
    want = np.asarray(Op.spread(m))
    if case not in ("nonfinite", "kin2d_f16", "dynamic_f16_neg"):  # float16: operator tests
        assert G.relerr(want, orc.spread(m)) < 1e-5
    for order in (_lib.KX_ORDER_NATURAL, _lib.KX_ORDER_ROTATED):
        for rb in (1, 2, 4, 8, 16):
            Op.plan.sside = _lib.KX_SIDE_SOURCES
            Op.plan.sorder = order
            Op.plan.rblock = rb
            got = np.asarray(Op.spread(m))
            if case == "nonfinite":
                np.testing.assert_array_equal(np.isnan(got), np.isnan(want))
                ok = ~np.isnan(want)
                assert G.relerr(got[ok], want[ok]) < 1e-12, (order, rb)
            else:
                np.testing.assert_array_equal(got, want, err_msg=f"order={order} rblock={rb}")


def test_spread_nonfinite_model_propagates_like_reference():
    """inf / NaN in the model switch the launch to float64 accumulation:
    the samples they reach become non-finite exactly where the oracle's
    do, the others keep oracle parity."""
    kw = dict(z=G.z, x=G.x, t=G.t, srcs=G.s2d, recs=G.r2d, vel=G.v0, wav=G.wav,
              wavcenter=G.wavc, mode="analytic", dynamic=False, dtype="float32")
    Op = _K(**kw)
    orc = O.KirchhoffOracle(**kw)
    m = np.random.default_rng(3).standard_normal(Op.shape[1]).astype(np.float32)
    m[17] = np.nan
    m[Op.shape[1] // 2] = np.inf
    got = np.asarray(Op.spread(m))
    want = orc.spread(m)
    bad = ~np.isfinite(want)
    assert bad.any()
    np.testing.assert_array_equal(~np.isfinite(got), bad)
    np.testing.assert_array_equal(np.isnan(got), np.isnan(want))
    assert G.relerr(got[~bad], want[~bad]) < 1e-5
    # a finite model right after still takes the fixed-point path
    m2 = np.nan_to_num(m, nan=0.0, posinf=0.0)
    assert G.relerr(np.asarray(Op.spread(m2)), orc.spread(m2)) < 1e-5


@pytest.mark.parametrize("dynamic", [False, True])
@pytest.mark.parametrize("dtype", ["float32", "float64"])
def test_gather_source_splits(dynamic, dtype):
    """Every split of a tile's sources over waves (kx_plan.gsplit) matches
    the oracle; 19 sources leave a partly filled source block."""
    srcs = np.vstack((np.linspace(G.x.min(), G.x.max(), 19), 2 * np.ones(19)))
    recs = np.vstack((np.linspace(G.x.min(), G.x.max(), 13), 2 * np.ones(13)))
    kw = dict(z=G.z, x=G.x, t=G.t, srcs=srcs, recs=recs, vel=G.v0, wav=G.wav,
              wavcenter=G.wavc, mode="analytic", dynamic=dynamic, dtype=dtype,
              angleaperture=60)
    Op = _K(**kw)
    orc = O.KirchhoffOracle(**kw)
    d = np.random.default_rng(4).standard_normal(Op.shape[0]).astype(dtype)
    want = orc.gather(d)
    tol = 1e-5 if dtype == "float32" else 1e-10
    for split in (1, 2, 4):
        Op.plan.gsplit = split
        assert G.relerr(np.asarray(Op.gather(d)), want) < tol, split
    Op.plan.gsplit = 0


@pytest.mark.parametrize("split", [1, 4])
def test_gather_samples_outside_the_staged_window(split):
    """A workgroup's tiles whose traveltimes span far more than the staged
    window (steep, finely sampled times over a deep image) read the samples
    outside it from global memory."""
    z = np.arange(256) * 4.0
    x = np.arange(6) * 4.0
    t = np.arange(2400) * 0.0005
    srcs = np.vstack((np.array([0.0, 20.0]), np.zeros(2)))
    recs = np.vstack((np.array([4.0, 12.0, 20.0]), np.zeros(3)))
    kw = dict(z=z, x=x, t=t, srcs=srcs, recs=recs, vel=2000.0, wav=G.wav,
              wavcenter=G.wavc, mode="analytic", dynamic=False, dtype="float32")
    Op = _K(**kw, tile=(16, 4, 1))
    orc = O.KirchhoffOracle(**kw)
    Op.plan.gsplit = split
    d = np.random.default_rng(6).standard_normal(Op.shape[0]).astype(np.float32)
    assert G.relerr(np.asarray(Op.gather(d)), orc.gather(d)) < 1e-5


@pytest.mark.parametrize("case", ["kin2d", "dyn2d", "kin3d", "f16", "nonfinite", "neg_byot"])
def test_forward_convolves_in_the_spread_launch(case):
    """kx_forward of a float32 plan convolves the workgroups' LDS traces in
    the spread launch; its data are bitwise kx_spread followed by kx_conv,
    for every receiver block, lane order and the float64 fallback of a
    non-finite model (nt = 120: the padded float rows and the 41-tap filter
    fit the LDS of the 64-bit sums; at the reference geometry's nt = 50
    they do not, and the golden tests run the two-launch path)."""
    import torch
    from pylops_amd import _lib
    t = np.arange(120) * G.PAR["dt"]
    kw = dict(z=G.z, x=G.x, t=t, srcs=G.s2d, recs=G.r2d, vel=G.v0,
              wav=G.wav.astype(np.float32), wavcenter=G.wavc, mode="analytic",
              dtype="float32")
    if case == "dyn2d":
        kw.update(dynamic=True, angleaperture=60)
    elif case == "kin3d":
        kw.update(y=G.y, srcs=G.s3d_par, recs=G.r3d_par)
    elif case == "f16":
        kw.update(tabledtype="float16")
    elif case == "neg_byot":
        import pylops_amd
        ts, tr, _, _, _, _ = pylops_amd.Kirchhoff._traveltime_table(G.z, G.x, G.s2d, G.r2d,
                                                                    G.v0, mode="analytic")
        kw.update(mode="byot", trav=(ts - 0.01113, tr - 0.00871), vel=None)
    if case == "kin3d":
        kw["t"] = G.t  # nt = 50: the two-launch path
    Op = _K(**kw)
    dev = torch.device("cuda", 0)
    m = np.random.default_rng(4).standard_normal(Op.shape[1]).astype(np.float32)
    if case == "nonfinite":
        m[17] = np.inf
    L = _lib.load()
    st = _lib.stream_ptr(dev)
    for order in (_lib.KX_ORDER_NATURAL, _lib.KX_ORDER_ROTATED):
        for rb in (1, 2, 4, 8):
            Op.plan.sorder, Op.plan.rblock = order, rb
            fused = np.asarray(Op @ m)
            traces = torch.as_tensor(np.asarray(Op.spread(m)), device=dev)
            two = torch.empty_like(traces)
            _lib.check(L.kx_conv(_lib.ptr(traces), _lib.ptr(two), Op.nsnr, Op.nt, Op.plan.h,
                                 Op.plan.nh, _lib.KX_F32, st), "kx_conv")
            two = two.cpu().numpy()
            if case == "nonfinite":
                np.testing.assert_array_equal(np.isnan(fused), np.isnan(two))
                ok = np.isfinite(two)
                np.testing.assert_array_equal(fused[ok], two[ok])
            else:
                np.testing.assert_array_equal(fused, two, err_msg=f"order={order} rb={rb}")


@pytest.mark.parametrize("dynamic", [False, True])
def test_long_traces_in_time_segments(dynamic):
    """Traces longer than a workgroup's LDS (24000 samples; on gfx950 one
    workgroup holds 20416) are spread in time segments, one per workgroup,
    the pair of samples of a contribution split where it straddles a
    segment boundary; the forward (segments, then the wavelet in HBM), the
    adjoint, a float64 operator and a batch of models (matmat, which then
    spreads model by model) match the oracle.  A 1 m image 2 km deep puts
    its samples on both sides of sample 20416 (2.04 s at 0.1 ms)."""
    from oracle import oracle as Ox
    nt, dt = 24000, 0.0001
    t = np.arange(nt) * dt
    x = np.arange(8) * 1.0
    z = 1990.0 + np.arange(8) * 1.0
    srcs = np.vstack((np.array([-600.0, 600.0]), np.zeros(2)))
    recs = np.vstack((np.array([-300.0, 0.0, 300.0]), np.zeros(3)))
    wav, _, wc = Ox.ricker(t[:41], f0=500)
    kw = dict(z=z, x=x, t=t, srcs=srcs, recs=recs, vel=2000.0, wav=wav.astype(np.float32),
              wavcenter=wc, mode="analytic", dynamic=dynamic, dtype="float32")
    if dynamic:
        # the offset aperture is |six - rix| / (iz + 1) in grid indices (K:836-880):
        # sources 600 grid points away over 8 depths need a wide one
        kw.update(angleaperture=89, aperture=1e6)
    Op = _K(**kw)
    orc = O.KirchhoffOracle(**kw)
    rng = np.random.default_rng(41)
    m = rng.standard_normal(Op.shape[1]).astype(np.float32)
    d = np.asarray(Op @ m)
    want = orc.matvec(m)
    k = np.nonzero(np.abs(want.reshape(-1, nt)).max(axis=0) > 0)[0]
    assert k.min() < 20416 < k.max(), (k.min(), k.max())  # both segments hold samples
    assert G.relerr(d, want) < 1e-5
    assert G.relerr(np.asarray(Op.H @ want.astype(np.float32)), orc.rmatvec(want)) < 1e-5
    M = rng.standard_normal((Op.shape[1], 2)).astype(np.float32)
    D = np.asarray(Op.matmat(M))
    for j in range(2):
        assert G.relerr(D[:, j], orc.matvec(M[:, j])) < 1e-5
    kw64 = dict(kw, dtype="float64", wav=wav)
    Op64 = _K(**kw64)
    orc64 = O.KirchhoffOracle(**kw64)
    assert G.relerr(np.asarray(Op64 @ m.astype(np.float64)),
                    orc64.matvec(m.astype(np.float64))) < 1e-10


@pytest.mark.parametrize("dtype", ["float32", "float64"])
def test_unreachable_traveltimes_contribute_nothing(dtype):
    """Traveltimes far past every trace (1e6 s: unreachable points, whose
    tile bases the table encoding clamps to 2^29 samples) fall outside the
    trace and add nothing, as the reference's bounds test skips them
    (K:743-751, 754-781).  One source and one receiver are unreachable
    everywhere, so tile groups whose source and receiver bases are both
    clamped meet in the adjoint, and a band of depths is unreachable for
    every source; every adjoint tile-group width is checked."""
    Op0 = _K(z=G.z, x=G.x, t=G.t, srcs=G.s2d, recs=G.r2d, vel=G.v0, wav=G.wav,
             wavcenter=G.wavc, mode="analytic")
    ts = np.array(Op0.trav_srcs, dtype=np.float64)
    tr = np.array(Op0.trav_recs, dtype=np.float64)
    ts[:, 0] = 1e6
    tr[:, -1] = 1e6
    band = (np.arange(ts.shape[0]) % len(G.z)) >= 12
    ts[band, :] = 1e6
    kw = dict(z=G.z, x=G.x, t=G.t, srcs=G.s2d, recs=G.r2d, vel=G.v0, wav=G.wav,
              wavcenter=G.wavc, mode="byot", trav=(ts, tr), dtype=dtype)
    Op = _K(**kw)
    orc = O.KirchhoffOracle(**kw)
    rng = np.random.default_rng(12)
    m = rng.standard_normal(Op.shape[1]).astype(dtype)
    d = rng.standard_normal(Op.shape[0]).astype(dtype)
    tol = 1e-5 if dtype == "float32" else 1e-10
    want_s = orc.spread(m)
    want_g = orc.gather(d)
    assert np.linalg.norm(want_s) > 0 and np.linalg.norm(want_g) > 0
    assert G.relerr(np.asarray(Op.spread(m)), want_s) < tol
    for gt in (1, 2, 4):
        for gs in (1, 4):
            Op.plan.gtiles, Op.plan.gsplit = gt, gs
            assert G.relerr(np.asarray(Op.gather(d)), want_g) < tol, (gt, gs)
    Op.plan.gtiles = Op.plan.gsplit = 0


def test_unreachable_traveltimes_float16_tables():
    """The same unreachable times with float16 receiver offsets (65504
    samples for an unreachable entry): the unreachable traces are exactly
    zero, and the other entries lose no precision to the marking (the
    error against the oracle stays within 1.5x that of float16 tables of
    the unmodified times, ~1e-3 on this grid: offsets of tens of samples
    at 2^-6 sample resolution); traces too long for that offset to clear
    them are refused."""
    Op0 = _K(z=G.z, x=G.x, t=G.t, srcs=G.s2d, recs=G.r2d, vel=G.v0, wav=G.wav,
             wavcenter=G.wavc, mode="analytic")
    ts0 = np.array(Op0.trav_srcs, dtype=np.float64)
    tr0 = np.array(Op0.trav_recs, dtype=np.float64)
    ts, tr = ts0.copy(), tr0.copy()
    ts[:, 0] = 1e6
    tr[:, -1] = 1e6
    tr[(np.arange(tr.shape[0]) % len(G.z)) >= 12, 1] = 1e6
    base = dict(z=G.z, x=G.x, t=G.t, srcs=G.s2d, recs=G.r2d, vel=G.v0, wav=G.wav,
                wavcenter=G.wavc, mode="byot", dtype="float32")
    rng = np.random.default_rng(13)
    ns, nr, nt = G.s2d.shape[1], G.r2d.shape[1], len(G.t)
    m = rng.standard_normal(len(G.x) * len(G.z)).astype(np.float32)
    d = rng.standard_normal(ns * nr * nt).astype(np.float32)
    errs = {}
    for name, tabs in (("plain", (ts0, tr0)), ("unreachable", (ts, tr))):
        kw = dict(base, trav=tabs)
        Op = _K(**kw, tabledtype="float16")
        orc = O.KirchhoffOracle(**kw)
        fwd = np.asarray(Op.spread(m))
        errs[name] = (G.relerr(fwd, orc.spread(m)), G.relerr(np.asarray(Op.gather(d)), orc.gather(d)))
    f = fwd.reshape(ns, nr, nt)
    assert not f[0].any() and not f[:, -1].any()
    for k in range(2):
        assert errs["unreachable"][k] < 1.5 * errs["plain"][k] + 1e-4, errs
    kw = dict(base, trav=(ts, tr))
    tlong = np.arange(70000) * float(G.t[1] - G.t[0])
    with pytest.raises(NotImplementedError, match="float16"):
        _K(**dict(kw, t=tlong), tabledtype="float16")


@pytest.mark.parametrize("tabledtype", [None, "float16"])
def test_dynamic_without_aperture_gates_skips_the_gate(tabledtype):
    """A dynamic plan whose offset aperture gates no trace (the default
    aperture) is flagged KX_DYNAMIC_NOGATE, and its adjoint runs without the
    per-sample gate test: bitwise the image of the gated kernel, and the
    oracle's; a plan with gates keeps them."""
    from pylops_amd import _lib
    kw = dict(z=G.z, x=G.x, t=G.t, srcs=G.s2d, recs=G.r2d, vel=G.v0, wav=G.wav,
              wavcenter=G.wavc, mode="analytic", dynamic=True, dtype="float32",
              angleaperture=60)
    Op = _K(**kw, tabledtype=tabledtype)
    assert Op.plan.dynamic == _lib.KX_DYNAMIC_NOGATE
    d = np.random.default_rng(14).standard_normal(Op.shape[0]).astype(np.float32)
    got = np.asarray(Op.gather(d))
    Op.plan.dynamic = 1
    gated = np.asarray(Op.gather(d))
    Op.plan.dynamic = _lib.KX_DYNAMIC_NOGATE
    np.testing.assert_array_equal(got, gated)
    if tabledtype is None:
        assert G.relerr(got, O.KirchhoffOracle(**kw).gather(d)) < 1e-5
    Opg = _K(**dict(kw, aperture=2.0), tabledtype=tabledtype)
    assert Opg.plan.dynamic == 1
