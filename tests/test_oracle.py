"""Pins the CPU oracle before it is trusted as the checker.

* golden vectors written by the reference itself (tests/golden/) -- kernels
  bit-exact, tables bit-exact, full operator to FFT-vs-direct rounding;
* closed forms: analytic traveltime (pytests/test_kirchhoff.py:139-157),
  single-point impulse response, convolution against numpy.convolve;
* brute force: the dense matrix of the forward equals the transpose of the
  dense matrix of the adjoint;
* invariants: single table == separable tables (test_kirchhoff.py:320-462);
* library routines: LSQR against scipy.sparse.linalg.lsqr, CGLS against
  numpy.linalg.lstsq (pytests/test_solver.py:170-232).
"""
import os

import numpy as np
import pytest

from oracle import oracle as O
from tests import geometry as G


def _load(name):
    return np.load(os.path.join(G.GOLDEN, name + ".npz"))


@pytest.mark.parametrize("name", sorted(G.golden_cases()))
def test_oracle_vs_golden(name):
    g = _load(name)
    Op = O.KirchhoffOracle(**G.golden_kwargs(name, g))
    for k in ["trav_srcs", "trav_recs", "amp_srcs", "amp_recs", "angle_srcs",
              "angle_recs"]:
        if k in g.files:
            np.testing.assert_array_equal(getattr(Op, k), g[k], err_msg=k)
    if "spread" in g.files:
        # the engine kernels reproduce the reference's float64 arithmetic and
        # per-`+=` float32 rounding exactly
        np.testing.assert_array_equal(Op.spread(g["m"]), g["spread"])
        np.testing.assert_array_equal(Op.gather(g["dk"], in_f64=True), g["gather"])
    tol = 1e-6 if Op.dtype == np.float32 else 1e-13
    assert G.relerr(Op.matvec(g["m"]), g["fwd"]) < tol
    assert G.relerr(Op.rmatvec(g["d"]), g["adj"]) < tol


def test_traveltime_analytic_closed_form():
    """pytests/test_kirchhoff.py:139-157."""
    src = np.array([100, 0])[:, np.newaxis]
    ts, tr, ds, dr, _, _ = O.analytic_tables(np.arange(0, 200, 1), np.arange(0, 200, 1),
                                             src, src, G.v0)
    assert ds[0, 0] + dr[0, 0] == 200
    assert ts[0, 0] == 100 / G.v0
    assert tr[0, 0] == 100 / G.v0


@pytest.mark.parametrize("dtype", ["float64", "float32"])
def test_impulse_response_closed_form(dtype):
    """A unit reflector at one image point spreads into two taps per trace at
    (t_s + t_r)/dt with linear-interpolation weights."""
    Op = O.KirchhoffOracle(G.z, G.x, G.t, G.s2d, G.r2d, G.v0, G.wav, G.wavc,
                           mode="analytic", dtype=dtype)
    ix, iz = 5, 13
    ii = ix * G.PAR["nz"] + iz
    m = np.zeros(Op.ni)
    m[ii] = 1.0
    y = Op.spread(m).reshape(Op.ns * Op.nr, Op.nt)
    px, pz = G.x[ix], G.z[iz]
    for isrc in range(Op.ns):
        for irec in range(Op.nr):
            T = (np.hypot(px - G.s2d[0, isrc], pz - G.s2d[1, isrc])
                 + np.hypot(px - G.r2d[0, irec], pz - G.r2d[1, irec])) / G.v0
            q = T / G.PAR["dt"]
            it = int(q)
            f = q - it
            expect = np.zeros(Op.nt)
            if 0 <= it < Op.nt - 1:
                expect[it], expect[it + 1] = 1 - f, f
            np.testing.assert_allclose(y[isrc * Op.nr + irec], expect, atol=1e-7)


def test_conv_same_vs_numpy_convolve():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((3, 57))
    for nh in (1, 5, 21, 41):
        h = rng.standard_normal(nh)
        full = np.stack([np.convolve(r, h) for r in x])
        c = (nh - 1) // 2
        np.testing.assert_allclose(O.conv_same(x, h), full[:, c:c + 57], atol=1e-12)


@pytest.mark.parametrize("nh,offset", [(21, 10), (21, 3), (21, 17), (20, 4), (20, 12)])
def test_filter_offset_centres_wavelet(nh, offset):
    """After Convolve1D's padding (convolve1d.py:72-86), an impulse at n0
    returns h with h[offset] at n0."""
    rng = np.random.default_rng(1)
    h = rng.standard_normal(nh)
    hp = O.padded_filter(h, offset)
    assert hp.size % 2 == 1
    n0 = 40
    x = np.zeros((1, 90))
    x[0, n0] = 1.0
    y = O.conv_same(x, hp)[0]
    np.testing.assert_allclose(y[n0 - offset:n0 - offset + nh], h, atol=1e-14)
    assert abs(y).sum() == pytest.approx(abs(h).sum())


@pytest.mark.parametrize("dynamic", [False, True])
def test_dense_forward_is_transpose_of_adjoint(dynamic):
    Op = O.KirchhoffOracle(G.z, G.x, G.t, G.s2d, G.r2d, G.v0, G.wav, G.wavc,
                           mode="analytic", dynamic=dynamic, angleaperture=50)
    n, m = Op.shape[1], Op.shape[0]
    A = np.stack([Op.matvec(e) for e in np.eye(n)], axis=1)
    sel = np.random.default_rng(2).choice(m, 60, replace=False)
    AT = np.stack([Op.rmatvec(np.eye(m)[i]) for i in sel], axis=0)
    np.testing.assert_allclose(AT, A[sel], atol=1e-12 * np.abs(A).max())


def test_single_table_equals_separable():
    Op = O.KirchhoffOracle(G.z, G.x, G.t, G.s2d, G.r2d, G.v0, G.wav, G.wavc,
                           mode="analytic")
    trav = (Op.trav_srcs[:, :, None] + Op.trav_recs[:, None, :]).reshape(Op.ni, -1)
    Op1 = O.KirchhoffOracle(G.z, G.x, G.t, G.s2d, G.r2d, G.v0, G.wav, G.wavc,
                            mode="byot", trav=trav)
    rng = np.random.default_rng(3)
    m = rng.standard_normal(Op.ni)
    d = rng.standard_normal(Op.shape[0])
    np.testing.assert_allclose(Op.matvec(m), Op1.matvec(m), atol=1e-12)
    np.testing.assert_allclose(Op.rmatvec(d), Op1.rmatvec(d), atol=1e-12)


def test_solvers_vs_golden():
    g = _load("solvers_dyn2d_f64")
    Op = O.KirchhoffOracle(G.z, G.x, G.t, G.s2d, G.r2d, G.v0, G.wav, G.wavc,
                           mode="analytic", dynamic=True)
    y = g["dobs"]
    x, istop, iiter, r1, r2, cost = O.cgls(Op, y, niter=10, tol=0.0)
    np.testing.assert_allclose(x, g["cgls_x"].ravel(), rtol=1e-9, atol=1e-12)
    assert (istop, iiter) == (int(g["cgls_istop"]), int(g["cgls_iiter"]))
    np.testing.assert_allclose([r1, r2], [g["cgls_r1"], g["cgls_r2"]], rtol=1e-9)
    np.testing.assert_allclose(cost, g["cgls_cost"], rtol=1e-9)
    x, istop, iiter, r1, r2, cost = O.cgls(Op, y, x0=g["x0"], niter=8, damp=0.3, tol=0.0)
    np.testing.assert_allclose(x, g["cglsd_x"], rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose([r1, r2], [g["cglsd_r1"], g["cglsd_r2"]], rtol=1e-9)
    np.testing.assert_allclose(cost, g["cglsd_cost"], rtol=1e-9)
    for pre, kw in (("lsqr_", dict(niter=10)),
                    ("lsqrd_", dict(x0=g["x0"], damp=0.2, niter=7))):
        res = O.lsqr(Op, y, calc_var=True, **kw)
        names = ["x", "istop", "iiter", "r1", "r2", "anorm", "acond", "arnorm",
                 "xnorm", "var", "cost"]
        for k, v in zip(names, res):
            np.testing.assert_allclose(np.ravel(v), np.ravel(g[pre + k]), rtol=1e-8,
                                       atol=1e-12, err_msg=pre + k)


@pytest.mark.parametrize("shape", [(11, 11), (31, 11)])
def test_solvers_vs_library(shape):
    """pytests/test_solver.py:170-232 restated: x = ones is recovered; LSQR
    matches scipy.sparse.linalg.lsqr."""
    from scipy.sparse.linalg import lsqr as sp_lsqr
    rng = np.random.default_rng(10)
    A = rng.normal(0, 10, shape) + 10 * np.eye(*shape)
    Op = O.DenseOp(A)
    xt = np.ones(shape[1])
    y = A @ xt
    x = O.cgls(Op, y, niter=2 * shape[1], tol=0.0)[0]
    np.testing.assert_allclose(x, np.linalg.lstsq(A, y, rcond=None)[0], rtol=1e-6)
    res = O.lsqr(Op, y, niter=shape[1], atol=0, btol=0, conlim=0)
    ref = sp_lsqr(A, y, iter_lim=shape[1], atol=0, btol=0, conlim=0)
    np.testing.assert_allclose(res[0], ref[0], rtol=1e-6)
    np.testing.assert_allclose(res[0], xt, rtol=1e-4)


# ---------------------------------------------------------------------------
# ISTA / FISTA restatement (oracle.shrinkage and thresholds)
# ---------------------------------------------------------------------------
def _brute_prox(x, penalty):
    """argmin_u 0.5 (u - x)^2 + penalty(u) on a fine grid."""
    u = np.linspace(-3 * abs(x) - 1, 3 * abs(x) + 1, 400001)
    return u[np.argmin(0.5 * (u - x) ** 2 + penalty(u))]


@pytest.mark.parametrize("thresh", [0.1, 0.7, 2.0])
def test_thresholds_are_proximal_maps(thresh):
    """soft = prox of t|u|, hard = prox of t [u != 0] (cut at sqrt(2t)),
    half = prox of (t/2) |u|^(1/2), i.e. argmin (u - x)^2 + t |u|^(1/2)
    (Xu et al. 2012; cut at 54^(1/3)/4 t^(2/3)): each checked by brute-force
    minimisation away from the jump."""
    cases = {"soft": (O.soft_threshold, lambda u: thresh * np.abs(u), 0.0),
             "hard": (O.hard_threshold, lambda u: thresh * (u != 0), np.sqrt(2 * thresh)),
             "half": (O.half_threshold, lambda u: 0.5 * thresh * np.sqrt(np.abs(u)),
                      54 ** (1 / 3) / 4 * thresh ** (2 / 3))}
    for name, (f, pen, cut) in cases.items():
        for x in (-3.1, -1.3, -0.4, 0.05, 0.6, 1.7, 4.2):
            if cut and abs(abs(x) - cut) < 0.05 * cut:
                continue  # the global minimiser jumps at the cut
            got = f(np.array([x]), thresh)[0]
            want = _brute_prox(x, pen)
            assert abs(got - want) < 2e-4 * max(1, abs(x)), (name, x, got, want)


def test_fista_ista_fixed_point_is_the_lasso_optimum():
    """With alpha = 1/L, many iterations reach a point satisfying the
    optimality conditions of 0.5|y - Ax|^2 + (eps/2)|x|_1 (the threshold is
    eps*alpha/2): A^T r = (eps/2) sign(x) on the support, |A^T r| <= eps/2
    off it; eps -> 0 gives the least-squares solution."""
    rng = np.random.default_rng(4)
    A = rng.standard_normal((31, 11))
    y = rng.standard_normal(31)
    Op = O.DenseOp(A)
    L = np.linalg.eigvalsh(A.T @ A).max()
    eps = 2.0
    for fast in (True, False):
        x, it, _ = O.shrinkage(Op, y, 1.0 / L, niter=20000, eps=eps, tol=1e-14, fast=fast)
        g = A.T @ (y - A @ x)
        on = x != 0
        np.testing.assert_allclose(g[on], eps / 2 * np.sign(x[on]), atol=1e-7)
        assert np.all(np.abs(g[~on]) <= eps / 2 + 1e-7)
        xls, _, _ = O.shrinkage(Op, y, 1.0 / L, niter=20000, eps=0.0, tol=1e-15, fast=fast)
        np.testing.assert_allclose(xls, np.linalg.lstsq(A, y, rcond=None)[0], atol=1e-8)


@pytest.mark.parametrize("shape", [(11, 11), (31, 11), (21, 41)])
def test_fista_ista_recovery_as_reference_tests(shape):
    """pytests/test_sparsity.py:259-343: a 3-spike model recovered to one
    decimal with eps=0.5, 2000 iterations, every threshold kind."""
    np.random.seed(42)
    A = np.random.randn(*shape)
    x = np.zeros(shape[1])
    x[shape[1] // 2] = 1
    x[3] = 1
    x[shape[1] - 4] = -1
    y = A @ x
    alpha = 1.0 / np.linalg.eigvalsh(A.T @ A).max()
    for kind in ("hard", "soft", "half"):
        for fast in (False, True):
            xinv, _, _ = O.shrinkage(O.DenseOp(A), y, alpha, niter=2000, eps=0.5, tol=0,
                                     threshkind=kind, fast=fast)
            np.testing.assert_array_almost_equal(x, xinv, decimal=1)


# ---------------------------------------------------------------------------
# eikonal traveltimes (oracle.eikonal_fsm); the reference's scikit-fmm is
# absent, so the restatement is pinned by closed forms
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("order", [1, 2])
def test_eikonal_one_dimensional_is_exact(order):
    """On a one-cell-deep grid only the x update is active: T = s (|k| h -
    h/2), the contour half a cell from the source centre; both difference
    orders reproduce a linear function exactly."""
    T, _ = O.eikonal_fsm(np.full((1, 30, 1), 1 / 500.0), [(10, 0, 0)], (2.0, 1.0, 3.0),
                         order=order)
    k = np.arange(30) - 10
    np.testing.assert_allclose(T[0, 0, :, 0], np.maximum(np.abs(k) * 2.0 - 1.0, 0) / 500.0,
                               rtol=0, atol=1e-15)


def test_eikonal_constant_velocity_matches_analytic():
    """pytests/test_kirchhoff.py:163-212 restated: eikonal and analytic
    traveltimes agree to two decimals (seconds) on the reference's grids,
    2D and 3D, with unequal spacings."""
    from tests import geometry as G
    for dim3 in (False, True):
        s = G.s3d if dim3 else G.s2d
        if dim3:
            ny, nx, nz = len(G.y), len(G.x), len(G.z)
            h = (G.x[1] - G.x[0], G.y[1] - G.y[0], G.z[1] - G.z[0])
            src = [(int(round(p[1] / h[0])), int(round(p[0] / h[1])), int(round(p[2] / h[2])))
                   for p in s.T]
        else:
            ny, nx, nz = 1, len(G.x), len(G.z)
            h = (G.x[1] - G.x[0], 1.0, G.z[1] - G.z[0])
            src = [(int(round(p[0] / h[0])), 0, int(round(p[1] / h[2]))) for p in s.T]
        T, _ = O.eikonal_fsm(np.full((ny, nx, nz), 1 / G.v0), src, h)
        ana = O.analytic_tables(G.z, G.x, s, s, G.v0, G.y if dim3 else None)[0]
        np.testing.assert_array_almost_equal(T.reshape(len(src), -1).T, ana, decimal=2)


def test_eikonal_source_cell_convention():
    """The cells around a point start at the straight-ray time to the zero
    contour of phi (-1 at the point's cell, +1 elsewhere), the diamond
    through the half-way points to the neighbouring centres, and stay
    fixed: 0 at the point's own cell (inside the contour: this build keeps
    it at 0, where scikit-fmm's signed result would hold a negative time),
    s hx / 2 and s hz / 2 one cell away along x and z, and s rho (1 - 1/S)
    on a diagonal, rho the distance and S = 2 (|dx|/hx + |dz|/hz); 3D adds
    the y term."""
    s = 1 / 1800.0
    hx, hz = 4.0, 3.0
    T, _ = O.eikonal_fsm(np.full((1, 9, 9), s), [(4, 0, 4)], (hx, 1.0, hz))
    t = T[0, 0]
    assert t[4, 4] == 0.0
    np.testing.assert_allclose([t[3, 4], t[5, 4]], s * hx / 2, rtol=1e-15)
    np.testing.assert_allclose([t[4, 3], t[4, 5]], s * hz / 2, rtol=1e-15)
    rho = np.hypot(hx, hz)
    np.testing.assert_allclose([t[3, 3], t[5, 5], t[3, 5], t[5, 3]], s * rho * 0.75, rtol=1e-15)
    rho2 = np.hypot(2 * hx, hz)
    np.testing.assert_allclose(t[6, 5], s * rho2 * (1 - 1 / 6.0), rtol=1e-15)
    hy = 5.0
    T3, _ = O.eikonal_fsm(np.full((5, 5, 5), s), [(2, 2, 2)], (hx, hy, hz))
    t3 = T3[0]
    assert t3[2, 2, 2] == 0.0
    np.testing.assert_allclose(t3[3, 2, 2], s * hy / 2, rtol=1e-15)
    rho3 = np.sqrt(hx * hx + hy * hy + hz * hz)
    np.testing.assert_allclose(t3[3, 3, 3], s * rho3 * (1 - 1 / 6.0), rtol=1e-15)


@pytest.mark.parametrize("order", [1, 2])
def test_eikonal_symmetry(order):
    """Constant velocity, source at the centre of a symmetric grid: every
    sweep order is used, so the converged field is mirror symmetric."""
    T, _ = O.eikonal_fsm(np.full((1, 21, 17), 1 / 2000.0), [(10, 0, 8)], (5.0, 1.0, 7.0),
                         order=order)
    np.testing.assert_allclose(T[0, 0], T[0, 0, ::-1, :], rtol=1e-14)
    np.testing.assert_allclose(T[0, 0], T[0, 0, :, ::-1], rtol=1e-14)


@pytest.mark.parametrize("order", [1, 2])
def test_eikonal_linear_gradient_converges_first_order(order):
    """v(z) = v0 + g z has the closed-form first arrival
    t = acosh(1 + g^2 r^2 / (2 v(z_s) v(z))) / g from the source point;
    away from the source the error halves with the spacing (the contour
    half a cell from the source and the source singularity keep both
    difference orders at first order against the point-source time)."""
    def err(h):
        n = int(400.0 / h) + 1
        x = np.arange(n) * h
        X, Z = np.meshgrid(x, x, indexing="ij")
        v = 1500.0 + Z
        ix, iz = n // 2, n // 4
        T, _ = O.eikonal_fsm((1 / v)[None], [(ix, 0, iz)], (h, 1.0, h), order=order)
        r2 = (X - x[ix]) ** 2 + (Z - x[iz]) ** 2
        ex = np.arccosh(1 + r2 / (2 * (1500.0 + x[iz]) * v))
        far = np.sqrt(r2) > 100.0
        return np.max(np.abs(T[0, 0] - ex)[far] / ex[far])
    e1, e2, e3 = err(10.0), err(5.0), err(2.5)
    assert e1 < 0.06 and 1.7 < e1 / e2 < 2.3 and 1.7 < e2 / e3 < 2.3, (e1, e2, e3)


def _diamond_distance(px, pz, hx, hz):
    """Distance from points (px, pz) outside it to the diamond
    |x|/hx + |z|/hz = 1/2 (the zero contour of phi around one cell)."""
    V = np.array([[hx / 2, 0.0], [0.0, hz / 2], [-hx / 2, 0.0], [0.0, -hz / 2]])
    P = np.stack((px.ravel(), pz.ravel()), 1)
    best = np.full(len(P), np.inf)
    for k in range(4):
        A, B = V[k], V[(k + 1) % 4]
        AB = B - A
        t = np.clip(((P - A) @ AB) / (AB @ AB), 0.0, 1.0)
        best = np.minimum(best, np.linalg.norm(P - (A + t[:, None] * AB), axis=1))
    return best.reshape(px.shape)


def test_eikonal_second_order_tracks_the_contour_distance():
    """Constant velocity: the exact answer is the distance to the zero
    contour divided by v.  Second-order sweeps cut the median error of the
    first-order ones by more than five and stay below 0.3% (median) and
    2.5% (max) beyond five cells; a wrong second-order term (sign, factor,
    side) fails this."""
    n, h, v = 121, 5.0, 2000.0
    x = np.arange(n) * h
    X, Z = np.meshgrid(x, x, indexing="ij")
    ix, iz = 60, 30
    ex = _diamond_distance(X - x[ix], Z - x[iz], h, h) / v
    far = np.hypot(X - x[ix], Z - x[iz]) > 5 * h
    med = {}
    for order in (1, 2):
        T, _ = O.eikonal_fsm(np.full((1, n, n), 1 / v), [(ix, 0, iz)], (h, 1.0, h), order=order)
        e = np.abs(T[0, 0] - ex)[far] / ex[far]
        med[order] = np.median(e)
        if order == 2:
            assert med[2] < 3e-3 and e.max() < 2.5e-2, (med[2], e.max())
    assert med[1] > 5 * med[2], med


# ---------------------------------------------------------------------------
# Convolve1D with a filter longer than the signal (oracle.conv_long_*)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("n,offset", [(10, 0), (10, 4), (10, 9), (7, 0), (7, 3), (7, 6)])
def test_conv_long_forward_is_scipy_same(n, offset):
    """The forward is the reference's own call: fftconvolve(h, pad(x),
    mode="same") (C1D:165-174), with the model padded by the offset rule."""
    from scipy.signal import fftconvolve
    rng = np.random.default_rng(n + 10 * offset)
    x, h = rng.standard_normal(n), rng.standard_normal(31)
    sh = 2 * (n // 2 - offset) - (1 if n % 2 == 0 else 0)
    xp = np.pad(x, (max(sh, 0), -min(sh, 0)))
    np.testing.assert_allclose(O.conv_long_forward(x, h, offset),
                               fftconvolve(h, xp, mode="same"), rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("n,offset", [(10, 0), (10, 4), (10, 9), (7, 0), (7, 3), (7, 6)])
def test_conv_long_pair_is_adjoint(n, offset):
    """The centre samples the adjoint keeps (C1D:188-199) make it the
    adjoint of the forward: <A x, d> = <x, A^T d>."""
    rng = np.random.default_rng(100 + n + offset)
    h = rng.standard_normal(23)
    x, d = rng.standard_normal(n), rng.standard_normal(23)
    lhs = O.conv_long_forward(x, h, offset) @ d
    rhs = x @ O.conv_long_adjoint(d, h, offset, n)
    assert abs(lhs - rhs) < 1e-12 * max(1.0, abs(lhs)), (lhs, rhs)


def test_conv_long_delta_filter_shifts_the_model():
    """A unit impulse at p as the long filter returns the padded model
    shifted: y[i] = xp[i + (m - 1)//2 - p]."""
    n, offset, nh, p = 8, 2, 40, 17
    x = np.arange(1.0, n + 1)
    h = np.zeros(nh)
    h[p] = 1.0
    sh = 2 * (n // 2 - offset) - 1
    xp = np.pad(x, (max(sh, 0), -min(sh, 0)))
    m = xp.size
    want = np.array([xp[i + (m - 1) // 2 - p] if 0 <= i + (m - 1) // 2 - p < m else 0.0
                     for i in range(nh)])
    np.testing.assert_array_equal(O.conv_long_forward(x, h, offset), want)


def test_conv_same_rows_matches_numpy_per_row():
    """One filter per trace: each row is numpy's convolve, centred."""
    rng = np.random.default_rng(5)
    x = rng.standard_normal((4, 50))
    h = rng.standard_normal((4, 9))
    got = O.conv_same_rows(x, h)
    for r in range(4):
        np.testing.assert_allclose(got[r], np.convolve(x[r], h[r])[4:54], rtol=1e-12,
                                   atol=1e-12)



@pytest.mark.parametrize("fast", [True, False])
def test_shrinkage_analysis_form_with_orthonormal_s(fast):
    """With an orthonormal S the analysis form on A is the synthesis form
    on A S mapped back by S: x = S c (a transposed S fails this)."""
    rng = np.random.default_rng(4)
    A = rng.standard_normal((30, 20))
    S, _ = np.linalg.qr(rng.standard_normal((20, 20)))
    y = rng.standard_normal(30)
    alpha = 1.0 / np.linalg.eigvalsh(A.T @ A).max()
    xa, _, _ = O.shrinkage(O.DenseOp(A), y, alpha, niter=15, eps=0.3, tol=0, fast=fast, S=S)
    c, _, _ = O.shrinkage(O.DenseOp(A @ S), y, alpha, niter=15, eps=0.3, tol=0, fast=fast)
    np.testing.assert_allclose(xa, S @ c, rtol=1e-10, atol=1e-12)


def serpentine(nlegs=48, width=9):
    """A 2D medium whose first arrivals follow a serpentine corridor: rows
    of slowness 1e-3 s/m at every second depth, joined at alternate ends,
    walls of slowness 1 s/m (h = 1 m), the source at the corridor's start.
    Fast sweeping advances about two corridor legs per round of its four
    sweep directions, so the corridor needs more than 20 rounds."""
    nz = 2 * nlegs - 1
    slow = np.ones((1, width, nz))
    for leg in range(nlegs):
        slow[0, :, 2 * leg] = 1e-3
        if leg + 1 < nlegs:
            slow[0, width - 1 if leg % 2 == 0 else 0, 2 * leg + 1] = 1e-3
    return slow, [(0, 0, 0)], (1.0, 1.0, 1.0)


def test_eikonal_needs_more_rounds_in_a_serpentine():
    """The oracle's fast sweeping on the serpentine corridor: 20 first-order
    rounds leave the far end of the corridor unconverged (far above its
    converged time), and the converged times follow the corridor (the
    arrival at its end is the corridor length times its slowness, to the
    accuracy of the first-order scheme along a straight corridor)."""
    slow, src, h = serpentine()
    T20, r20 = O.eikonal_fsm(slow, src, h, max_iter=20, order=1)
    T, r = O.eikonal_fsm(slow, src, h, max_iter=200, order=1)
    assert r20 == 20 and 20 < r < 200
    nz = slow.shape[2]
    end = T[0, 0, 0 if (nz // 2) % 2 else slow.shape[1] - 1, nz - 1]
    path = (slow.shape[1] - 1) * (nz + 1) // 2 + (nz - 1)   # cells along the corridor
    assert abs(end - path * 1e-3) < 0.05 * path * 1e-3
    assert T20[0, 0, :, nz - 1].min() > 2 * T[0, 0, :, nz - 1].max()
