"""GPU tests of the batched path: many uneven Kirchhoff problems through one
persistent launch per product (KirchhoffBatch) and CGLS per problem
(cgls_batch).

Bars:
  batched forward vs each operator's own forward   identical (float32: the
      fixed-point sum does not depend on the work split), 1e-12 (float64)
  batched adjoint vs each operator's own adjoint   1e-6 (float32; the split
      of a tile's sources over waves may differ), 1e-12 (float64)
  batched products vs the oracle                   1e-5 (float32)
  cgls_batch vs cgls per problem                   1e-4 relative L2 on x,
      1e-4 on the cost history (float32 CGLS amplifies the reordering above)
  graph replays vs eager                           identical
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests import geometry as G

pytestmark = pytest.mark.gpu


def _problem(rng, dynamic, dtype, tabledtype=None):
    nz = int(rng.integers(12, 40))
    nx = int(rng.integers(12, 56))
    ns = int(rng.integers(1, 5))
    nr = int(rng.integers(3, 19))
    nt = int(rng.integers(60, 180))
    dx = dz = 10.0
    x = np.arange(nx) * dx
    z = np.arange(nz) * dz
    t = np.arange(nt) * 0.004
    srcs = np.vstack((np.linspace(x[0], x[-1], ns), np.zeros(ns)))
    recs = np.vstack((np.linspace(x[0], x[-1], nr), np.zeros(nr)))
    f0 = float(rng.uniform(15, 30))
    wav, _, wc = O.ricker(t[:21], f0=f0)
    vel = float(rng.uniform(1500, 3500))
    kw = dict(z=z, x=x, t=t, srcs=srcs, recs=recs, vel=vel, wav=wav.astype(dtype),
              wavcenter=wc, mode="analytic", dynamic=dynamic, dtype=dtype)
    if tabledtype is not None:
        kw["tabledtype"] = tabledtype
    return kw


def _ops(n, seed, dynamic, dtype, tabledtype=None):
    import pylops_amd
    rng = np.random.default_rng(seed)
    kws = [_problem(rng, dynamic, dtype, tabledtype) for _ in range(n)]
    return kws, [pylops_amd.Kirchhoff(**kw) for kw in kws]


CASES = [(False, "float32", None), (True, "float32", None), (False, "float32", "float16"),
         (True, "float64", None)]


@pytest.mark.parametrize("dynamic,dtype,tab", CASES)
def test_batch_products_match_single_operators(dynamic, dtype, tab):
    from pylops_amd.waveeqprocessing import KirchhoffBatch
    kws, ops = _ops(7, 11, dynamic, dtype, tab)
    B = KirchhoffBatch(ops)
    rng = np.random.default_rng(3)
    m = rng.standard_normal(B.shape[1]).astype(dtype)
    d = rng.standard_normal(B.shape[0]).astype(dtype)
    fb = B @ m
    ab = B.H @ d
    f32 = dtype == "float32"
    for p, op in enumerate(ops):
        mo, do = B.model_offsets, B.data_offsets
        f1 = op @ m[mo[p]:mo[p + 1]]
        a1 = op.H @ d[do[p]:do[p + 1]]
        if f32:
            np.testing.assert_array_equal(fb[do[p]:do[p + 1]], f1)
            assert G.relerr(ab[mo[p]:mo[p + 1]], a1) < 1e-6
        else:
            assert G.relerr(fb[do[p]:do[p + 1]], f1) < 1e-12
            assert G.relerr(ab[mo[p]:mo[p + 1]], a1) < 1e-12


@pytest.mark.parametrize("dynamic", [False, True])
@pytest.mark.parametrize("gsplit,gtiles", [(1, 2), (4, 2), (2, 1)])
def test_batch_adjoint_bitwise_with_matched_shape(dynamic, gsplit, gtiles):
    """The batch adjoint runs the single-problem adjoint body on each
    problem: with the same tile groups and split, each problem's image is
    bitwise its own operator's."""
    from pylops_amd.waveeqprocessing import KirchhoffBatch
    _, ops = _ops(5, 17, dynamic, "float32")
    B = KirchhoffBatch(ops, gsplit=gsplit, gtiles=gtiles)
    d = np.random.default_rng(9).standard_normal(B.shape[0]).astype(np.float32)
    ab = B.H @ d
    mo, do = B.model_offsets, B.data_offsets
    for p, op in enumerate(ops):
        op.gsplit, op.gtiles = B.gsplit, B.gtiles
        np.testing.assert_array_equal(ab[mo[p]:mo[p + 1]], op.H @ d[do[p]:do[p + 1]])


def test_batch_products_vs_oracle():
    from pylops_amd.waveeqprocessing import KirchhoffBatch
    kws, ops = _ops(5, 7, True, "float32")
    B = KirchhoffBatch(ops)
    rng = np.random.default_rng(4)
    m = rng.standard_normal(B.shape[1]).astype(np.float32)
    d = rng.standard_normal(B.shape[0]).astype(np.float32)
    fb, ab = B @ m, B.H @ d
    for p, kw in enumerate(kws):
        orc = O.KirchhoffOracle(**kw)
        mo, do = B.model_offsets, B.data_offsets
        assert G.relerr(fb[do[p]:do[p + 1]], orc.matvec(m[mo[p]:mo[p + 1]])) < 1e-5
        assert G.relerr(ab[mo[p]:mo[p + 1]], orc.rmatvec(d[do[p]:do[p + 1]])) < 1e-5


@pytest.mark.parametrize("rblock,sorder,gsplit,gtiles",
                         [(0, 0, 0, 0), (1, 1, 1, 1), (4, 2, 2, 2), (8, 1, 4, 1), (2, 2, 4, 2)])
def test_batch_launch_shapes_agree(rblock, sorder, gsplit, gtiles):
    """Every receiver block, lane order and adjoint shape gives the same
    forward (exactly, fixed point) and the same adjoint (to rounding)."""
    from pylops_amd.waveeqprocessing import KirchhoffBatch
    _, ops = _ops(6, 21, False, "float32")
    B0 = KirchhoffBatch(ops)
    B = KirchhoffBatch(ops, rblock=rblock, sorder=sorder, gsplit=gsplit, gtiles=gtiles)
    if rblock:
        assert B.rblock == rblock
    if gsplit:
        assert B.gsplit == gsplit
    if gtiles:
        assert B.gtiles == gtiles
    rng = np.random.default_rng(8)
    m = rng.standard_normal(B.shape[1]).astype(np.float32)
    d = rng.standard_normal(B.shape[0]).astype(np.float32)
    np.testing.assert_array_equal(B @ m, B0 @ m)
    assert G.relerr(B.H @ d, B0.H @ d) < 1e-6


@pytest.mark.parametrize("dtype", ["float64", "float32"])
def test_batch_dottest_and_repeated_launches(dtype):
    """Adjointness, and the work counters reset themselves: repeated
    products agree (exactly in float32, where the forward sums in fixed
    point; float64 LDS atomics add in arrival order)."""
    from pylops_amd.waveeqprocessing import KirchhoffBatch
    _, ops = _ops(9, 5, True, dtype)
    B = KirchhoffBatch(ops)
    rng = np.random.default_rng(1)
    u = rng.standard_normal(B.shape[1]).astype(dtype)
    v = rng.standard_normal(B.shape[0]).astype(dtype)
    Bu = B @ u
    BHv = B.H @ v
    yy, xx = np.dot(Bu.astype(np.float64), v), np.dot(u, BHv.astype(np.float64))
    assert abs(xx - yy) <= (1e-10 if dtype == "float64" else 1e-5) * max(abs(xx), abs(yy))
    for _ in range(3):
        if dtype == "float32":
            np.testing.assert_array_equal(B @ u, Bu)
        else:
            assert G.relerr(B @ u, Bu) < 1e-13
        np.testing.assert_array_equal(B.H @ v, BHv)


def test_batch_graph_replays_match_eager():
    """A captured forward + adjoint, replayed with eager work and
    synchronisations between the replays, matches the eager products on
    every replay."""
    import torch
    from pylops_amd.waveeqprocessing import KirchhoffBatch
    _, ops = _ops(6, 9, False, "float32", "float16")
    B = KirchhoffBatch(ops)
    dev = B.device
    g = torch.Generator(device="cpu").manual_seed(0)
    m = torch.randn(B.shape[1], generator=g).to(dev)
    d = torch.empty(B.shape[0], device=dev)
    a = torch.empty(B.shape[1], device=dev)
    B._into(m, d, False)
    B._into(d, a, True)
    d_ref, a_ref = d.clone(), a.clone()
    graph = torch.cuda.CUDAGraph()
    torch.cuda.synchronize()
    with torch.cuda.graph(graph):
        B._into(m, d, False)
        B._into(d, a, True)
    for k in range(4):
        d.zero_()
        a.zero_()
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(d, d_ref), k
        assert torch.equal(a, a_ref), k
        # eager work between replays
        B._into(a, d, False)
        torch.cuda.synchronize()


def test_cgls_batch_matches_cgls_per_problem():
    import torch
    from pylops_amd.optimization.basic import cgls, cgls_batch
    kws, ops = _ops(6, 13, False, "float32")
    rng = np.random.default_rng(2)
    ms = [rng.standard_normal(op.shape[1]).astype(np.float32) for op in ops]
    ys = [np.asarray(op @ mp) for op, mp in zip(ops, ms)]
    ys[2] = np.zeros_like(ys[2])  # converged before the first iteration
    damp = np.array([0.0, 0.1, 0.0, 0.5, 0.0, 0.0])
    res = cgls_batch(ops, ys, niter=8, damp=damp, tol=0.0)
    for p, op in enumerate(ops):
        x1, istop1, it1, r1n1, r2n1, cost1 = cgls(op, ys[p], niter=8, damp=float(damp[p]),
                                                  tol=0.0)
        xb, istopb, itb, r1nb, r2nb, costb = res[p]
        assert itb == it1 and istopb == istop1, (p, itb, it1)
        if p == 2:
            assert it1 == 0 and np.all(xb == 0)
            continue
        assert G.relerr(xb, x1) < 1e-4, p
        np.testing.assert_allclose(costb, cost1, rtol=1e-4)
        assert abs(r2nb - r2n1) <= 1e-4 * abs(r2n1)
    torch.cuda.synchronize()


def test_cgls_batch_graph_equals_eager():
    import torch
    from pylops_amd.optimization.cls_batch import CGLSBatch
    from pylops_amd.waveeqprocessing import KirchhoffBatch
    _, ops = _ops(5, 17, True, "float32")
    B = KirchhoffBatch(ops)
    rng = np.random.default_rng(6)
    y = torch.as_tensor(rng.standard_normal(B.shape[0]).astype(np.float32), device=B.device)
    s1 = CGLSBatch(B)
    x1 = s1.solve(y, niter=6, damp=0.05, tol=0.0, graph=False)
    s2 = CGLSBatch(B)
    x2 = s2.solve(y, niter=6, damp=0.05, tol=0.0, graph=True)
    assert torch.equal(x1, x2)
    for p in range(B.nprob):
        np.testing.assert_array_equal(s1.cost[p], s2.cost[p])
    assert np.all(s2.iiter == 6)


def test_batch_long_traces_beyond_64k_lds():
    """Traces of 12000 samples (96 KB of LDS each, more than the 64 KB every
    CDNA device offers a workgroup, within gfx950's 160 KB) batch like short
    ones: the batched forward is bitwise each operator's and the adjoint
    agrees; traces that fit no workgroup are refused with the reason."""
    import pylops_amd
    from pylops_amd.waveeqprocessing import KirchhoffBatch
    kws = []
    for nt, nx in ((12000, 10), (11000, 14)):
        t = np.arange(nt) * 0.0005
        x = np.arange(nx) * 10.0
        z = 500.0 + np.arange(12) * 10.0
        wav, _, wc = O.ricker(t[:41], f0=40)
        kws.append(dict(z=z, x=x, t=t, srcs=np.vstack((np.array([0.0, x[-1]]), np.zeros(2))),
                        recs=np.vstack((np.linspace(0.0, x[-1], 3), np.zeros(3))), vel=300.0,
                        wav=wav.astype(np.float32), wavcenter=wc, mode="analytic",
                        dtype="float32"))
    ops = [pylops_amd.Kirchhoff(**kw) for kw in kws]
    B = KirchhoffBatch(ops)
    rng = np.random.default_rng(5)
    m = rng.standard_normal(B.shape[1]).astype(np.float32)
    d = rng.standard_normal(B.shape[0]).astype(np.float32)
    fb, ab = B @ m, B.H @ d
    mo, do = B.model_offsets, B.data_offsets
    assert np.abs(fb).max() > 0
    for p, op in enumerate(ops):
        np.testing.assert_array_equal(fb[do[p]:do[p + 1]], op @ m[mo[p]:mo[p + 1]])
        assert G.relerr(ab[mo[p]:mo[p + 1]], op.H @ d[do[p]:do[p + 1]]) < 1e-6
    long_kw = dict(kws[0], t=np.arange(24000) * 0.0005)
    with pytest.raises(NotImplementedError, match="does not fit one trace in LDS"):
        KirchhoffBatch([pylops_amd.Kirchhoff(**long_kw)])


def test_batch_rejects_mixed_table_formats():
    import pylops_amd
    from pylops_amd.waveeqprocessing import KirchhoffBatch
    rng = np.random.default_rng(0)
    a = pylops_amd.Kirchhoff(**_problem(rng, False, "float32"))
    b = pylops_amd.Kirchhoff(**_problem(rng, False, "float32", "float16"))
    with pytest.raises(ValueError, match="differs from problem 0"):
        KirchhoffBatch([a, b])


def test_autotune_single_operator(tmp_path):
    """The tuner returns a valid shape set, caches it, and the tuned
    operator gives the default operator's results (to rounding)."""
    import pylops_amd
    from pylops_amd.waveeqprocessing import autotune
    rng = np.random.default_rng(31)
    kw = _problem(rng, True, "float32")
    cache = str(tmp_path / "tune.json")
    best = autotune(cache=cache, reps=1, **kw)
    assert best["tile"][0] * best["tile"][1] * best["tile"][2] == 64
    assert best["rblock"] in (1, 2, 4, 8) and best["gsplit"] in (1, 2, 4)
    assert best["sside"] in (1, 2)
    assert best["ms"] > 0
    assert autotune(cache=cache, reps=1, **kw) == best  # from the cache
    op0 = pylops_amd.Kirchhoff(**kw)
    op = pylops_amd.Kirchhoff(**kw, tile=best["tile"], rblock=best["rblock"],
                              sorder=best["sorder"], sside=best["sside"], gsplit=best["gsplit"])
    m = rng.standard_normal(op.shape[1]).astype(np.float32)
    d = rng.standard_normal(op.shape[0]).astype(np.float32)
    # another tile splits traveltimes into other integer bases and offsets:
    # equal to float32 rounding of the interpolation weights
    assert G.relerr(op @ m, op0 @ m) < 1e-6
    assert G.relerr(op.H @ d, op0.H @ d) < 1e-6


def test_autotune_batch(tmp_path):
    import pylops_amd
    from pylops_amd.waveeqprocessing import KirchhoffBatch, autotune_batch
    rng = np.random.default_rng(32)
    kws = [_problem(rng, False, "float32", "float16") for _ in range(4)]
    best = autotune_batch(kws, reps=1, cache=False)
    ops = [pylops_amd.Kirchhoff(**kw, tile=best["tile"]) for kw in kws]
    B = KirchhoffBatch(ops, rblock=best["rblock"], sorder=best["sorder"], gsplit=best["gsplit"],
                       gtiles=best["gtiles"])
    assert B.rblock == best["rblock"] and B.gsplit == best["gsplit"]
    assert B.gtiles == best["gtiles"]
    B0 = KirchhoffBatch([pylops_amd.Kirchhoff(**kw) for kw in kws])
    m = rng.standard_normal(B.shape[1]).astype(np.float32)
    d = rng.standard_normal(B.shape[0]).astype(np.float32)
    # float16 offsets from another tile's bases round differently: the
    # float16-table bar
    assert G.relerr(B @ m, B0 @ m) < 2e-3
    assert G.relerr(B.H @ d, B0.H @ d) < 2e-3


def test_launch_attributes_update_the_plan():
    import pylops_amd
    rng = np.random.default_rng(33)
    op = pylops_amd.Kirchhoff(**_problem(rng, False, "float32"))
    m = rng.standard_normal(op.shape[1]).astype(np.float32)
    d0 = op @ m
    op.rblock, op.sorder, op.gsplit = 2, 1, 4
    assert (op.plan.rblock, op.plan.sorder, op.plan.gsplit) == (2, 1, 4)
    np.testing.assert_array_equal(op @ m, d0)
