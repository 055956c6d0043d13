"""CPU tests of the host side: the C-ABI library loads and exports every
declared symbol, the ctypes plan matches the C struct layout, and the
table / weight construction reproduces the reference's per-triplet weights
(checked against the oracle, itself pinned by golden vectors)."""
import os
import re
import subprocess

import numpy as np
import pytest
import torch

from oracle import oracle as O
from tests import geometry as G

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "kirchx.h")


def _declared():
    src = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|int64_t|const char \*)\s*(kx_\w+)\(", src, re.M)))


def test_library_exports_every_declared_symbol():
    from pylops_amd import _lib
    L = _lib.load()
    names = _declared()
    assert len(names) >= 12
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(_lib.EXPORTS)
    assert L.kx_version() == _lib.KX_ABI_VERSION
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIBPATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\sT\s(kx_\w+)", out))
    assert set(names) <= exported


@pytest.mark.parametrize("cname,pyname", [("kx_plan", "KxPlan"), ("kx_batch", "KxBatch"),
                                           ("kx_segs", "KxSegs")])
def test_struct_layout_matches_header(tmp_path, cname, pyname):
    """Compile the header with gcc and compare sizeof/offsetof with ctypes."""
    import ctypes
    from pylops_amd import _lib
    cls = getattr(_lib, pyname)
    fields = [f[0] for f in cls._fields_]
    prog = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"',
            'int main(void){', f'printf("%zu\\n", sizeof({cname}));']
    prog += [f'printf("%zu\\n", offsetof({cname}, {f}));' for f in fields]
    prog += ['return 0;}']
    c = tmp_path / "layout.c"
    c.write_text("\n".join(prog))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", str(c), "-o", str(exe)], check=True)
    vals = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True,
                                           check=True).stdout.split()]
    assert vals[0] == ctypes.sizeof(cls)
    assert len(vals) == len(fields) + 1
    for f, off in zip(fields, vals[1:]):
        assert getattr(cls, f).offset == off, f


def test_header_constants_match_binding():
    from pylops_amd import _lib
    text = open(HEADER).read()
    for name in ("KX_ABI_VERSION", "KX_AUX_BYTES", "KX_F32", "KX_F64", "KX_F16", "KX_ORDER_AUTO",
                 "KX_ORDER_NATURAL", "KX_ORDER_ROTATED", "KX_SIDE_AUTO", "KX_SIDE_RECEIVERS",
                 "KX_SIDE_SOURCES",
                 "KX_EINVAL", "KX_ESHAPE", "KX_ETYPE", "KX_EABI", "KX_ELDS"):
        m = re.search(rf"#define {name} (-?\d+)", text)
        assert m is not None, name
        assert int(m.group(1)) == getattr(_lib, name), name


def test_spread_requires_aux():
    """float data needs the caller's aux scratch; the check precedes any
    device call."""
    import ctypes
    from pylops_amd import _lib
    from pylops_amd.waveeqprocessing import _tables as T
    L = _lib.load()
    p = T.make_geom_plan(8, 1, 16, (16, 4, 1))
    p.dtype = p.tab_fmt = _lib.KX_F32
    p.ns, p.nr, p.nt = 1, 1, 16
    p.off_s = p.base_s = p.off_r = p.base_r = 16  # never dereferenced
    rc = L.kx_spread(ctypes.byref(p), ctypes.c_void_p(16), ctypes.c_void_p(16), None)
    assert rc == _lib.KX_EINVAL
    assert "aux" in L.kx_last_error().decode()


def test_error_mapping():
    from pylops_amd import _lib
    import ctypes
    L = _lib.load()
    p = _lib.KxPlan()
    p.abi_version = 99
    rc = L.kx_spread(ctypes.byref(p), None, None, None)
    assert rc == _lib.KX_EABI
    with pytest.raises(ValueError, match="abi_version"):
        _lib.check(rc, "kx_spread")
    p.abi_version = _lib.KX_ABI_VERSION
    p.nx, p.ny, p.nz, p.tz, p.tx, p.ty = 4, 1, 4, 8, 4, 1
    rc = L.kx_spread(ctypes.byref(p), None, None, None)
    assert rc == _lib.KX_ESHAPE
    rc = L.kx_conv(None, None, 1, 10, None, 3, 0, None)
    assert rc == _lib.KX_EINVAL


def test_analytic_tables_match_oracle():
    from pylops_amd.waveeqprocessing import _tables as T
    for dim3 in (False, True):
        s, r = (G.s3d, G.r3d) if dim3 else (G.s2d, G.r2d)
        y = G.y if dim3 else None
        geo = O.identify_geometry(G.z, G.x, s, r, y)
        ts, tr, ds, dr, gs, gr = T.analytic(G.z, G.x, s, r, G.v0, y, geo["dims"],
                                            geo["dsamp"], torch.device("cpu"))
        ots, otr, ods, odr, ogs, ogr = O.analytic_tables(G.z, G.x, s, r, G.v0, y)
        # torch's CPU sqrt is not always correctly rounded (1 ulp); the device
        # path uses IEEE sqrt and division
        np.testing.assert_allclose(ts.numpy(), ots, rtol=5e-16, atol=0)
        np.testing.assert_allclose(tr.numpy(), otr, rtol=5e-16, atol=0)
        for a, b in zip(gs + gr, ogs + ogr):
            np.testing.assert_allclose(a.numpy(), b, rtol=1e-13, atol=1e-15)


@pytest.mark.parametrize("name", ["dyn2d_f64", "dyn2d_ap_f64", "dyn2d_ap_f32",
                                  "dyn2d_byot_f64", "dyn3d_f64"])
def test_dynamic_weight_factorisation(name):
    """Ps*Pr + Qs*Qr gated by the depth aperture equals the reference's
    per-triplet weight 2 cos((ts-tr)/2) As Ar / v * aptap (K:820-889) for
    every (image point, source, receiver) of the golden geometries."""
    from pylops_amd.waveeqprocessing import _tables as T
    g = np.load(os.path.join(G.GOLDEN, name + ".npz"))
    kw = G.golden_kwargs(name, g)
    orc = O.KirchhoffOracle(**kw)
    cpu = torch.device("cpu")
    amin, amax = orc.angleaperture
    vel = torch.as_tensor(orc.vel)
    ps, qs = T.side_weights(torch.as_tensor(g["amp_srcs"]), torch.as_tensor(g["angle_srcs"]),
                            amin, amax, orc.aptap, scale=2.0 / vel)
    pr, qr = T.side_weights(torch.as_tensor(g["amp_recs"]), torch.as_tensor(g["angle_recs"]),
                            amin, amax, orc.aptap)
    w = (ps[:, :, None] * pr[:, None, :] + qs[:, :, None] * qr[:, None, :]).numpy()
    ni, ns, nr = w.shape
    izmin = T.aperture_izmin(g["six"], g["rix"], orc.nz, float(orc.aperture[1]))
    iz = np.arange(ni) % orc.nz
    gate = iz[:, None, None] >= izmin.reshape(ns, nr)[None]
    w = np.where(gate, w, 0.0)
    # reference weight, evaluated with the oracle's restatement of K:820-880
    a_s = np.abs(g["angle_srcs"])[:, :, None]
    a_r = np.abs(g["angle_recs"])[:, None, :]
    dang = amax - amin
    tap = orc.aptap

    def taper_of(a):
        idx = np.clip(np.floor_divide(20 * (a - amin), dang), 0, 20).astype(int)
        return np.where(a >= amin, tap[idx], 1.0)

    ok_ang = (a_s < amax) & (a_r < amax)
    ap = np.abs(g["six"] - g["rix"]).reshape(ns, nr)[None] / (iz[:, None, None] + 1)
    ok = ok_ang & (ap < orc.aperture[1])
    cosang = np.cos((g["angle_srcs"][:, :, None] - g["angle_recs"][:, None, :]) / 2.0)
    ref = (2.0 * cosang * g["amp_srcs"][:, :, None] * g["amp_recs"][:, None, :]
           / orc.vel[:, None, None]) * taper_of(a_s) * taper_of(a_r)
    ref = np.where(ok, ref, 0.0)
    np.testing.assert_allclose(w, ref, rtol=1e-12, atol=1e-15 * np.abs(ref).max())
    del cpu


def test_aperture_izmin_is_exact_threshold():
    from pylops_amd.waveeqprocessing import _tables as T
    six = np.array([0, 3, 7, 11, 5])
    rix = np.array([0, 0, 2, 0, 11])
    for apmax in (0.5, 1.0, 2.0, 13.0):
        izmin = T.aperture_izmin(six, rix, 20, apmax)
        for k in range(5):
            for iz in range(20):
                assert (iz >= izmin[k]) == (abs(six[k] - rix[k]) / (iz + 1) < apmax)


def test_kirchhoff_needs_a_device():
    """No CPU fallback: construction fails loudly without a HIP device."""
    if torch.cuda.is_available():
        pytest.skip("a HIP device is visible")
    import pylops_amd
    with pytest.raises(RuntimeError, match="HIP device"):
        pylops_amd.Kirchhoff(G.z, G.x, G.t, G.s2d, G.r2d, G.v0, G.wav, G.wavc,
                             mode="analytic")


def test_linearoperator_semantics_numpy():
    """pytests/test_linearoperator.py:331-388 restated on a dense operator."""
    from pylops_amd import LinearOperator

    class Dense(LinearOperator):
        def __init__(self, A, dims, dimsd):
            super().__init__(dtype=A.dtype, dims=dims, dimsd=dimsd)
            self.A = A

        def _matvec(self, x):
            return self.A @ x

        def _rmatvec(self, y):
            return self.A.T @ y

    A = np.arange(24.0).reshape(6, 4)
    Op = Dense(A, (2, 2), (3, 2))
    x = np.ones((2, 2))
    y = Op @ x
    assert y.shape == (3, 2)
    np.testing.assert_allclose(y.ravel(), A @ np.ones(4))
    assert (Op.H @ y).shape == (2, 2)
    assert Op.matvec_count == 1 and Op.rmatvec_count == 0
    X = np.ones((4, 3))
    Y = Op.matmat(X)
    assert Y.shape == (6, 3)
    assert Op.matmat_count == 1 and Op.matvec_count == 1
    with pytest.raises(ValueError):
        Op.matvec(np.ones(5))
    np.testing.assert_allclose((2 * Op) @ np.ones(4), 2 * A @ np.ones(4))
    np.testing.assert_allclose((Op.H * Op) @ np.ones(4), A.T @ A @ np.ones(4))


def test_segments_cover_each_segment_in_chunks():
    """The chunk tables of the segment-wise solver kernels: every chunk lies
    inside one segment, chunks tile each segment in order, empty segments
    get one empty chunk."""
    from pylops_amd.optimization.cls_batch import Segments
    off = np.array([0, 5, 5, 40, 41, 100])
    S = Segments(off, torch.device("cpu"), chunk=16)
    lo = S._chunk_lo.numpy()
    seg = S._chunk_seg.numpy()
    first = S._seg_chunk.numpy()
    assert S.nseg == 5 and S.nchunk == len(seg) == len(lo) - 1
    assert first[0] == 0 and first[-1] == S.nchunk
    assert lo[-1] == off[-1] and np.all(np.diff(lo) >= 0) and np.all(np.diff(lo) <= 16)
    for p in range(S.nseg):
        cs = range(first[p], first[p + 1])
        assert len(cs) >= 1 and all(seg[c] == p for c in cs)
        assert lo[first[p]] == off[p] and lo[first[p + 1]] == off[p + 1]
    assert S.work.numel() == S.nchunk


def test_convolve1d_construction_rules():
    """Convolve1D's host-side rules (C1D:21-47, 57-199): the method rule by
    signal dimension, the offset bounds, the data shape of a long filter,
    and filter shapes that do not broadcast."""
    from pylops_amd.signalprocessing import Convolve1D
    with pytest.raises(NotImplementedError):
        Convolve1D(50, np.ones(5), method="overlapadd")
    with pytest.raises(NotImplementedError):
        Convolve1D((4, 50), np.ones(5), method="direct")
    Convolve1D(50, np.ones(5), method="fft")
    Convolve1D((4, 50), np.ones(5), method="overlapadd")
    with pytest.raises(ValueError):
        Convolve1D(50, np.ones(5), offset=5)
    with pytest.raises(ValueError):
        Convolve1D(10, np.ones(30), offset=10)
    C = Convolve1D(10, np.ones(30), offset=9)
    assert C.shape == (30, 10) and C.long
    C = Convolve1D((3, 10), np.ones((3, 30)), axis=1)
    assert C.dimsd == (3, 30) and C.dims == (3, 10)
    with pytest.raises(ValueError):
        Convolve1D((3, 10), np.ones(30), axis=1)  # 1-D long filter, 2-D signal
    with pytest.raises(ValueError):
        Convolve1D((3, 50), np.ones((2, 5)), axis=1)  # 2 filters for 3 traces
    C = Convolve1D((3, 50), np.ones((1, 5)), axis=1)
    assert C.h.shape == (3, 9)  # 5 taps padded for offset 0


def test_solvers_are_freed_without_cyclic_gc():
    """A solver holds no reference cycle (its callback wrappers reach it
    weakly), so it and its device buffers are released as soon as it goes
    out of scope, before a next solve allocates its own."""
    import gc
    import types
    import weakref
    from pylops_amd.optimization.cls_basic import CGLS, LSQR
    from pylops_amd.optimization.cls_sparsity import FISTA, ISTA
    Op = types.SimpleNamespace(shape=(3, 3))
    enabled = gc.isenabled()
    gc.disable()
    try:
        for cls in (CGLS, LSQR, ISTA, FISTA):
            s = cls(Op)
            r = weakref.ref(s)
            del s
            assert r() is None, cls.__name__
    finally:
        if enabled:
            gc.enable()


def test_solver_callback_hooks_wrap_methods():
    """The on_setup/step/run_begin/end hooks (basesolver.py:18-330) see the
    solver itself and the method's vector, begin hooks in list order and end
    hooks in reverse order."""
    import types
    from pylops_amd.optimization.basesolver import Solver
    from pylops_amd.optimization.callback import Callbacks

    class Counting(Solver):
        def setup(self, y, x0=None, show=False):
            return 0 if x0 is None else x0

        def step(self, x, show=False):
            self.iiter += 1
            return x + 1

        def run(self, x, niter=2, show=False):
            for _ in range(niter):
                x = self.step(x)
            return x

        def finalize(self, show=False):
            pass

        def memory_usage(self, show=False, unit="B"):
            return 0

        def solve(self, y, x0=None, niter=2, show=False):
            x = self.setup(y, x0=x0)
            x = self.run(x, niter)
            self.finalize()
            return x

    seen = []

    class Rec(Callbacks):
        def __init__(self, tag):
            self.tag = tag

        def on_setup_begin(self, solver, x0):
            seen.append((self.tag, "setup_begin", solver, x0))

        def on_step_end(self, solver, x):
            seen.append((self.tag, "step_end", solver, x))

        def on_run_end(self, solver, x):
            seen.append((self.tag, "run_end", solver, x))

    s = Counting(types.SimpleNamespace(shape=(1, 1)), callbacks=[Rec("a"), Rec("b")])
    assert s.solve(None, x0=5, niter=2) == 7
    assert [(t, e, x) for t, e, _, x in seen] == [
        ("a", "setup_begin", 5), ("b", "setup_begin", 5),
        ("b", "step_end", 5), ("a", "step_end", 5),
        ("b", "step_end", 6), ("a", "step_end", 6),
        ("b", "run_end", 5), ("a", "run_end", 5)]
    assert all(sv is s for _, _, sv, _ in seen)


@pytest.mark.parametrize("nt", [21, 20])
def test_ricker(nt):
    """pytests/test_wavelets.py:58-72 restated: size of the symmetric axis
    and wavelet (one sample dropped from an even axis) and a unit peak at
    the centre; plus the closed form (1 - 2a) exp(-a), a = (pi f0 t)^2, and
    the mirror symmetry."""
    import warnings
    from pylops_amd.utils.wavelets import ricker
    t = np.arange(nt) * 0.004
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        wav, twav, wcenter = ricker(t, f0=20)
    n = (nt - 1 if nt % 2 == 0 else nt) * 2 - 1
    assert twav.size == n
    assert wav.shape[0] == n
    assert wav[wcenter] == 1
    assert wcenter == n // 2 and twav[wcenter] == 0
    a = (np.pi * 20 * twav) ** 2
    np.testing.assert_allclose(wav, (1 - 2 * a) * np.exp(-a), rtol=0, atol=1e-15)
    np.testing.assert_array_equal(wav, wav[::-1])


def test_identify_geometry():
    """pytests/test_kirchhoff.py:78-137 restated on the product's
    Kirchhoff._identify_geometry (K:457-500)."""
    from pylops_amd.waveeqprocessing import Kirchhoff
    P = G.PAR
    (ndims, shiftdim, dims, ny, nx, nz, ns, nr, dy, dx, dz, dsamp,
     origin) = Kirchhoff._identify_geometry(G.z, G.x, G.s2d, G.r2d)
    assert ndims == 2 and shiftdim == 0
    assert list(dims) == [P["nx"], P["nz"]]
    assert (ny, nx, nz) == (1, P["nx"], P["nz"])
    assert (ns, nr) == (P["nsx"], P["nrx"])
    assert list(dsamp) == [P["dx"], P["dz"]] and list(origin) == [0, 0]
    (ndims, shiftdim, dims, ny, nx, nz, ns, nr, dy, dx, dz, dsamp,
     origin) = Kirchhoff._identify_geometry(G.z, G.x, G.s3d_par, G.r3d_par, y=G.y)
    assert ndims == 3 and shiftdim == 1
    assert list(dims) == [P["ny"], P["nx"], P["nz"]]
    assert (ny, nx, nz) == (P["ny"], P["nx"], P["nz"])
    assert (ns, nr) == (P["nsy"] * P["nsx"], P["nry"] * P["nrx"])
    assert list(dsamp) == [P["dy"], P["dx"], P["dz"]] and list(origin) == [0, 0, 0]


def test_traveltime_ana():
    """pytests/test_kirchhoff.py:139-157 restated: horizontal and vertical
    straight paths in a homogeneous medium, exactly."""
    from pylops_amd.waveeqprocessing import Kirchhoff
    src = np.array([100, 0])[:, np.newaxis]
    ts, tr, ds, dr, _, _ = Kirchhoff._traveltime_table(
        np.arange(0, 200, 1), np.arange(0, 200, 1), src, src, G.v0, mode="analytic")
    assert ds[0, 0] + dr[0, 0] == 200
    assert ts[0, 0] == 100 / G.v0
    assert tr[0, 0] == 100 / G.v0
