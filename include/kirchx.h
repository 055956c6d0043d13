/*
 * kirchx.h -- C ABI of the MI355X Kirchhoff demigration/migration library
 * (libkirchx.so).  Plain C types only: device pointers are `void *`,
 * streams are `hipStream_t` passed as `void *`, sizes are int32/int64.
 *
 * Every entry point is stream-ordered and asynchronous: it enqueues kernels
 * on `stream` and returns.  Return value 0 means success; a negative value is
 * an argument error (KX_EINVAL ...) and a positive value is a hipError_t.
 * kx_last_error() gives a message for the calling thread's last failure.
 * The library allocates no device memory; scratch space is a caller-owned
 * workspace whose size kx_workspace_bytes() reports (kx_forward_noalloc
 * alone uses a KX_AUX_BYTES device array of the library instead).
 *
 * Reference interfaces replaced (PyLops, pylops/waveeqprocessing/kirchhoff.py
 * = K, pylops/signalprocessing/convolve1d.py = C1D,
 * pylops/optimization/cls_basic.py = CB):
 *   kx_spread   <- Kirchhoff._travsrcrec_kirch_matvec   K:727-752
 *                  Kirchhoff._ampsrcrec_kirch_matvec    K:783-889
 *                  Kirchhoff._trav_kirch_matvec         K:683-702
 *   kx_gather   <- Kirchhoff._travsrcrec_kirch_rmatvec  K:754-781
 *                  Kirchhoff._ampsrcrec_kirch_rmatvec   K:891-1001
 *                  Kirchhoff._trav_kirch_rmatvec        K:704-725
 *   kx_conv     <- _Convolve1Dshort._matvec/_rmatvec    C1D:99-115
 *                  (scipy.signal.fftconvolve(mode="same"))
 *   kx_conv_rows, kx_xcorr_rows
 *               <- N-D filters of _Convolve1Dshort      C1D:57-115
 *                  _Convolve1Dlong._matvec/_rmatvec     C1D:118-199
 *   kx_forward  <- Kirchhoff._matvec                    K:1048-1094
 *   kx_adjoint  <- Kirchhoff._rmatvec                   K:1096-1143
 *   kx_encode_table / kx_tile_table
 *               <- the table setup of Kirchhoff.__init__ K:304-398
 *                  (device layout only; values come from the caller)
 *   kx_dot, kx_dot_dev, kx_vec, kx_lsqr_xw, kx_cgls_scalar, kx_lsqr_scalar
 *               <- the vector algebra and host scalar updates of CGLS.step
 *                  CB:510-564 and LSQR.step CB:987-1136 (float64 reductions)
 */
#ifndef KIRCHX_H
#define KIRCHX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KX_ABI_VERSION 5

/* Bytes of the per-plan device scratch `aux` (forward accumulation scale). */
#define KX_AUX_BYTES 4096

/* lane order of the forward (kx_plan.sorder): at step r of a tile the lane
 * of point p treats receiver r (natural) or (p + r) mod rblock (rotated) of
 * its workgroup's block; both orders give bitwise the same traces */
#define KX_ORDER_AUTO 0
#define KX_ORDER_NATURAL 1
#define KX_ORDER_ROTATED 2

/* what a forward workgroup blocks (kx_plan.sside): one source and a block of
 * rblock receivers, or one receiver and a block of rblock sources (float32
 * operators; float64 ones block receivers).  Both give bitwise the same
 * traces. */
#define KX_SIDE_AUTO 0
#define KX_SIDE_RECEIVERS 1
#define KX_SIDE_SOURCES 2

/* plan.dynamic value for dynamic plans whose aperture gates are all 0 */
#define KX_DYNAMIC_NOGATE 2

/* element types */
#define KX_F32 0
#define KX_F64 1
#define KX_F16 2 /* table offsets only */

/* error codes (negative: argument errors) */
#define KX_OK 0
#define KX_EINVAL -1
#define KX_ESHAPE -2
#define KX_ETYPE -3
#define KX_EABI -4
#define KX_ELDS -5

/* Operator plan.  All table pointers are device pointers in the tiled
 * layout produced by kx_encode_table / kx_tile_table:
 *   offsets  off[ntiles][n][64]   element type tab_fmt (f32, f16 or f64;
 *            with KX_F16 the source side off_s is f32), traveltime in
 *            samples minus base,
 *   bases    base[ntiles][n]      int32, samples,
 *   weights  wp/wq[ntiles][n][64] element type dtype (dynamic only; with
 *            KX_F16 the receiver side wp_r, wq_r is f16, the source side
 *            f32: scale the receiver weights to at most 1 and the source
 *            weights by the same factor, so that their products keep
 *            their values and the f16 range fits any units),
 * n = ns for the source side and nr for the receiver side.
 * A tile is a tz*tx*ty = 64 block of image points; tiles are enumerated
 * z fastest, then x, then y; lanes inside a tile likewise. */
typedef struct kx_plan {
    int32_t abi_version; /* KX_ABI_VERSION */
    int32_t dtype;       /* KX_F32 or KX_F64: model, data and accumulation */
    int32_t tab_fmt;     /* KX_F32, KX_F16 (dtype F32 only) or KX_F64 */
    int32_t dynamic;     /* 1: amplitude/obliquity/aperture weights;
                            KX_DYNAMIC_NOGATE (2): the same with every
                            izmin entry 0 (no offset-aperture gate), which
                            lets the adjoint skip the gate test */
    int32_t ns, nr, nt;  /* sources, receivers, time samples */
    int32_t nx, ny, nz;  /* image grid, ny = 1 in 2D */
    int32_t tz, tx, ty;  /* tile shape, tz*tx*ty == 64 */
    int32_t ntz, ntx, nty;
    int64_t ntiles;      /* ntz*ntx*nty */
    int32_t neg_trav;    /* 1 if a traveltime may be negative (truncation path) */
    int32_t rblock;      /* receivers (sources, see sside) per spread workgroup, 0 = automatic */
    int32_t nh;          /* wavelet length after offset padding (odd) */
    int32_t sorder;      /* spread lane order KX_ORDER_*, 0 = automatic */
    const void *off_s;
    const int32_t *base_s;
    const void *off_r;
    const int32_t *base_r;
    const void *wp_s, *wq_s, *wp_r, *wq_r; /* dynamic weights or NULL */
    const int32_t *izmin;                  /* [ns*nr] aperture gate or NULL */
    const void *h;                         /* wavelet, nh elements of dtype */
    const void *hstar;                     /* flipped wavelet */
    double wamp;  /* bound on |weight| of one (point, trace) pair:
                     1 kinematic, max_s |(wp_s, wq_s)| * max_r |(wp_r, wq_r)| dynamic */
    void *aux;    /* KX_AUX_BYTES of device memory owned by the caller; the
                     forward writes its accumulation scale there, so calls that
                     share one aux must be ordered on one stream */
    int32_t gsplit; /* waves per tile in the adjoint, 1, 2 or 4 (each sums a
                       share of the sources), 0 = automatic */
    int32_t sside;  /* forward workgroup blocks KX_SIDE_*, 0 = automatic */
    int32_t gtiles; /* neighbouring image tiles per wave in the adjoint (1, 2
                       or 4; they share the staged trace windows),
                       0 = automatic */
    int32_t pad0;
    double wtyp;  /* typical |weight| of one (point, trace) pair: 1 kinematic,
                     mean_s |(wp_s, wq_s)| * mean_r |(wp_r, wq_r)| over the
                     nonzero weights dynamic (0: taken as wamp).  A float32
                     forward accumulates in float64 instead of fixed point
                     when max|m| wamp exceeds 1024 mean|m| wtyp (mean over the
                     nonzero m), where fixed-point rounding would no longer be
                     small against a typical contribution */
} kx_plan;

int kx_version(void);
const char *kx_last_error(void);

/* Bytes of workspace kx_forward / kx_adjoint need (one [ns*nr][nt] trace
 * buffer of dtype). */
int64_t kx_workspace_bytes(const kx_plan *p);

/* Table layout: reads a [n][ni] float64 table tabT (reference table
 * transposed), multiplies by `scale` (1/dt), and writes off/base in the
 * tiled layout; fmt = KX_F32 / KX_F16 / KX_F64.  Padding lanes replicate
 * the nearest grid point.  A NaN entry marks a time that no trace sample
 * can take whatever the other side's time (past sample nt - 1 or before
 * -1 even with the other side's extreme): it gets an offset of 2^29
 * samples (65504 with KX_F16) and no part in its tile's base, so the
 * kernels skip it and its tile's other entries keep small offsets.  Bases
 * are clamped to +-2^27 samples (-2^27 for a tile of NaN entries only). */
int kx_encode_table(const kx_plan *geom, const double *tabT, int32_t n,
                    double scale, int32_t fmt, void *off, int32_t *base,
                    void *stream);

/* Reorders a [n][ni] float64 table into [ntiles][n][64] of dtype `fmt`
 * (KX_F32, KX_F64 or KX_F16), zero padded (dynamic weights). */
int kx_tile_table(const kx_plan *geom, const double *tabT, int32_t n,
                  int32_t fmt, void *out, void *stream);

/* The engine kernels alone.
 * kx_spread: traces[ns*nr][nt] = sum over image points of m spread at the
 *            two-sample linear interpolation of t_s + t_r (overwrites).
 * kx_gather: m[ni] = sum over traces of the interpolated trace samples
 *            (overwrites). */
int kx_spread(const kx_plan *p, const void *m, void *traces, void *stream);
int kx_gather(const kx_plan *p, const void *traces, void *m, void *stream);

/* kx_spread for nmodel models at once (a float32 operator; 1 <= nmodel <=
 * 64): model j at m + j*ni, its traces at traces + j*ns*nr*nt.  The
 * column-by-column forward of LinearOperator.matmat (LO:436-447) with the
 * table loads and index arithmetic of a tile shared by 2, 4 or 8 models per
 * workgroup; each model's traces are bitwise its kx_spread.  Uses plan.aux
 * for the per-model scales (bytes [2048, 2048 + 8*nmodel)). */
int kx_spread_models(const kx_plan *p, const void *m, void *traces, int32_t nmodel, void *stream);

/* kx_gather for nmodel models at once (a float32 operator): traces of
 * model j at traces + j*ns*nr*nt, its image at m + j*ni.  The adjoint
 * kernels of LinearOperator.rmatmat's column-by-column rmatvecs
 * (LO:449-460) with the tables, sample indices, weights and window
 * positions of a wave shared by 4 or 2 models; bitwise each column's
 * kx_gather. */
int kx_gather_models(const kx_plan *p, const void *traces, void *m, int32_t nmodel,
                     void *stream);

/* "same"-mode convolution of each of ntr rows of length nt with an odd
 * filter h[nh]:  y[n] = sum_k h[k] x[n + (nh-1)/2 - k].  x != y. */
int kx_conv(const void *x, void *y, int64_t ntr, int32_t nt, const void *h,
            int32_t nh, int32_t dtype, void *stream);

/* kx_conv with row strides: row r convolves x + r*xstride (nt samples) with
 * the odd filter h + r*hstride (nh taps) into y + r*nt; a stride of 0
 * shares one row.  Per-trace filters of Convolve1D (an N-D h,
 * C1D:57-115) and the forward of a filter longer than the signal
 * (_Convolve1Dlong._matvec, C1D:165-174: the long filter is the signal and
 * the padded model the filter). */
int kx_conv_rows(const void *x, int64_t xstride, void *y, int64_t ntr, int32_t nt,
                 const void *h, int64_t hstride, int32_t nh, int32_t dtype, void *stream);

/* Correlation at nlag lags: y[r*nlag + j] = sum_{u<la} a[r*astride + u] *
 * b[r*bstride + lag0 + j + u], b zero outside [0, lb); summed in float64.
 * The adjoint of a filter longer than the signal
 * (_Convolve1Dlong._rmatvec, C1D:176-199: full convolution with the flipped
 * filter, then the centre samples). */
int kx_xcorr_rows(const void *a, int64_t astride, int32_t la, const void *b, int64_t bstride,
                  int32_t lb, void *y, int64_t ntr, int32_t nlag, int32_t lag0, int32_t dtype,
                  void *stream);

/* Full operator: forward = spread then convolve with h; adjoint =
 * correlate (convolve with hstar) then gather.  work: kx_workspace_bytes.
 * A float32 forward whose workgroups hold their traces as 64-bit sums
 * convolves them in LDS in the same launch (the data are bitwise those of
 * kx_spread followed by kx_conv) and leaves work untouched. */
int kx_forward(const kx_plan *p, const void *m, void *d, void *work, void *stream);
int kx_adjoint(const kx_plan *p, const void *d, void *m, void *work, void *stream);

/* kx_forward with no caller memory beyond m and d, for callers that cannot
 * allocate (code under HIP graph capture, serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  The forward's only
 * scratch -- the per-launch fixed-point scale and partial maxima,
 * KX_AUX_BYTES -- lives in a device array of the library, one per GPU, held
 * until the process ends; plan.aux is not used, and calls that overlap on
 * one GPU race for the array.  Float32 plans whose forward convolves in the
 * spread launch (see kx_forward) only; others fail with KX_ETYPE (float64)
 * or KX_ELDS.  The data are bitwise those of kx_forward. */
int kx_forward_noalloc(const kx_plan *p, const void *m, void *d, void *stream);

/* ---- vector algebra for the solvers (float64 accumulation) ---- */
/* Bytes of workspace for kx_dot. */
int64_t kx_reduce_workspace_bytes(void);
/* out[0] = sum a[i]*b[i] accumulated in float64; out is a device double. */
int kx_dot(const void *a, const void *b, int64_t n, int32_t dtype, double *out,
           void *work, void *stream);

/* Solver scalars live in a device float64 array `sc`.
 * CGLS layout (CGLS.step, CB:510-564): */
#define KX_CG_KOLD 0
#define KX_CG_QQ 1
#define KX_CG_CC 2
#define KX_CG_A 3
#define KX_CG_K 4
#define KX_CG_B 5
#define KX_CG_SS 6
#define KX_CG_XX 7
#define KX_CG_DAMP2 8
#define KX_CG_TOL 9
#define KX_CG_ACTIVE 10
#define KX_CG_ITER 11
#define KX_CG_NITER 12
#define KX_CG_NSC 16
/* LSQR layout (LSQR.step, CB:987-1136): */
#define KX_LS_ALFA 0
#define KX_LS_BETA 1
#define KX_LS_ANORM 2
#define KX_LS_ACOND 3
#define KX_LS_DAMP 4
#define KX_LS_DAMPSQ 5
#define KX_LS_DDNORM 6
#define KX_LS_RES2 7
#define KX_LS_XNORM 8
#define KX_LS_XXNORM 9
#define KX_LS_Z 10
#define KX_LS_CS2 11
#define KX_LS_SN2 12
#define KX_LS_RHOBAR 13
#define KX_LS_PHIBAR 14
#define KX_LS_BNORM 15
#define KX_LS_RNORM 16
#define KX_LS_R1NORM 17
#define KX_LS_R2NORM 18
#define KX_LS_ARNORM 19
#define KX_LS_ARNORM0 20
#define KX_LS_TEST1 21
#define KX_LS_TEST2 22
#define KX_LS_TEST3 23
#define KX_LS_ATOL 24
#define KX_LS_BTOL 25
#define KX_LS_CTOL 26
#define KX_LS_ISTOP 27
#define KX_LS_ITER 28
#define KX_LS_NITER 29
#define KX_LS_ACTIVE 30
#define KX_LS_T1 31
#define KX_LS_T2 32
#define KX_LS_RHO 33
#define KX_LS_THETA 34
#define KX_LS_PHI 35
#define KX_LS_TAU 36
#define KX_LS_PSI 37
#define KX_LS_UU 38
#define KX_LS_VV 39
#define KX_LS_WW 40
#define KX_LS_VAR 41
#define KX_LS_RTOL 42
#define KX_LS_BETAPOS 43
#define KX_LS_ALFAPOS 44
#define KX_LS_NSC 48

/* kx_vec operations (c = sign * coef[0], or 1 / coef[0] for SCALE):
 *   KX_VEC_XPBY   y = x + c y
 *   KX_VEC_AXPY   y = y + c x
 *   KX_VEC_SCALE  y = y / coef[0]
 * and, when out != NULL, out[0] = y.y in float64.  The kernel does nothing
 * unless every non-NULL gate points at a non-zero value. */
#define KX_VEC_XPBY 0
#define KX_VEC_AXPY 1
#define KX_VEC_SCALE 2

/* Bytes of the reduction workspace of the calls below; zero it once after
 * allocation, the kernels leave it zeroed. */
int64_t kx_solver_workspace_bytes(void);
int kx_vec(int32_t op, int64_t n, int32_t dtype, const void *x, void *y,
           const double *coef, double sign, double *out, void *work,
           const double *gate0, const double *gate1, void *stream);
/* out[0] = a.b in float64, skipped when *gate0 == 0. */
int kx_dot_dev(int64_t n, int32_t dtype, const void *a, const void *b, double *out,
               void *work, const double *gate0, void *stream);
/* LSQR x / w update: sc[KX_LS_WW] = w.w, x += t1 w, w = v + t2 w. */
int kx_lsqr_xw(int64_t n, int32_t dtype, void *x, void *w, const void *v, double *sc,
               void *work, void *stream);
/* Scalar recurrences, one thread: CGLS phases 0 (step size a), 1 (beta),
 * 2 (end of iteration: kold, counters, cost[iter], cost1[iter], loop flag);
 * LSQR phases 0 (beta, anorm), 1 (alfa), 2 (rotations), 3 (norms, tests,
 * cost[iter], loop flag). */
int kx_cgls_scalar(double *sc, double *cost, double *cost1, int32_t phase, void *stream);
int kx_lsqr_scalar(double *sc, double *cost, int32_t phase, void *stream);

/* ---- batches of independent problems ----
 * Replace the reference's one-problem-at-a-time use of Kirchhoff._matvec /
 * _rmatvec (K:1048-1143) and CGLS.step (CB:510-564) for many problems: every
 * operator application of a batch is one persistent launch over all
 * problems' work, and the CGLS vector algebra runs segment-wise, one segment
 * per problem, with per-problem float64 scalars.  Every problem of a batch
 * shares dtype, tab_fmt, dynamic and neg_trav; models and data are
 * concatenated in problem order. */
typedef struct kx_batch {
    int32_t abi_version;
    int32_t nprob;
    int32_t dtype, tab_fmt, dynamic, neg_trav;
    int32_t rblock;            /* forward receivers per workgroup (all problems) */
    int32_t sorder;            /* forward lane order (all problems) */
    int32_t gsplit;            /* adjoint waves per tile (all problems) */
    int32_t nt_max, nh_max;
    int32_t gtiles;            /* adjoint tiles per wave, 1 or 2 (all problems) */
    int64_t model_len;         /* elements of the concatenated models */
    int64_t data_len;          /* elements of the concatenated data */
    int64_t n_spread, n_gather, n_conv; /* work items */
    int32_t grid_spread, grid_gather;   /* workgroups of the persistent launches */
    void *dev;                 /* kx_batch_dev_bytes of device memory, caller-owned */
} kx_batch;

/* Device bytes a batch of these plans needs (descriptors, work lists,
 * counters), or a negative KX_E* code when the plans do not form a batch
 * (kx_last_error says why).  rblock, sorder, gsplit and gtiles of plans[0]
 * apply to the whole batch (0 = automatic; gtiles above 2 counts as 2), the
 * other plans' values are ignored. */
int64_t kx_batch_dev_bytes(const kx_plan *plans, int32_t nprob);
/* Fills *b and writes the descriptors and work lists to dev; returns after
 * the copies are done (not capturable).  The plans' tables and wavelets must
 * outlive the batch; their aux buffers are not used.  Launches on one batch
 * share its work counters: order them on one stream. */
int kx_batch_build(kx_batch *b, const kx_plan *plans, int32_t nprob, void *dev,
                   void *stream);
/* Bytes of workspace of kx_batch_forward / kx_batch_adjoint (data_len
 * elements of dtype). */
int64_t kx_batch_workspace_bytes(const kx_batch *b);
int kx_batch_forward(const kx_batch *b, const void *m, void *d, void *work, void *stream);
int kx_batch_adjoint(const kx_batch *b, const void *d, void *m, void *work, void *stream);

/* Segments of a concatenated vector, cut into chunks inside segments (all
 * arrays on the device): segment p = [off[p], off[p+1]), chunk c =
 * [chunk_lo[c], chunk_lo[c+1]) of segment chunk_seg[c], and the chunks of
 * segment p are [seg_chunk[p], seg_chunk[p+1]). */
typedef struct kx_segs {
    int32_t nseg, nchunk;
    const int64_t *off;
    const int32_t *chunk_seg;
    const int64_t *chunk_lo;
    const int32_t *seg_chunk;
} kx_segs;

/* Bytes of reduction workspace for a segment set. */
int64_t kx_seg_workspace_bytes(const kx_segs *s);
/* Segment-wise kx_vec / kx_dot_dev: the scalars of segment p are
 * sc[p * nsc + column]; coef, out and gate are columns (out, gate: -1 for
 * none).  A segment whose gate scalar is 0 is left untouched. */
int kx_seg_vec(int32_t op, const kx_segs *s, int32_t dtype, const void *x, void *y, double *sc,
               int32_t nsc, int32_t coef, double sign, int32_t out, int32_t gate, void *work,
               void *stream);
int kx_seg_dot(const kx_segs *s, int32_t dtype, const void *a, const void *b, double *sc,
               int32_t nsc, int32_t out, int32_t gate, void *work, void *stream);
/* kx_cgls_scalar for nseg problems at once: scalars sc[p * KX_CG_NSC + ...],
 * costs cost[p * ncost + iter], cost1 likewise. */
int kx_cgls_scalar_batch(double *sc, int32_t nseg, double *cost, double *cost1, int32_t ncost,
                         int32_t phase, void *stream);

/* ---- traveltime tables by the eikonal equation ----
 * Replaces the scikit-fmm calls of Kirchhoff._traveltime_table, mode
 * "eikonal" (K:609-638): T[ii * n + s] (the reference's (ni, n) table,
 * point order ii) is the first-arrival time from the zero contour of
 * phi = -1 at source cell src[s] = (ix, iy, iz) and +1 elsewhere
 * (1 <= n <= 2^22), for the slowness field slow[ni]
 * (unit slowness gives the distance, skfmm.distance).  Fast sweeping with
 * Godunov updates: first order until no value changes, then, for order 2
 * (scikit-fmm's default), second-order one-sided differences until no value
 * changes; each phase runs at most max_iter rounds of all sweep directions.
 * iters (host, two int32): iters[0] receives the rounds run, iters[1] 1 when
 * every phase ended with no value changing and 0 when a phase ran out of
 * rounds (the tables are then upper bounds of the converged times).  2D: ndims = 2 and ny is
 * ignored (iy = 0).  src (device) is copied back and every cell checked
 * against the grid before any launch (KX_ESHAPE).  flag: one device int32
 * of scratch.  Synchronises the stream (not capturable). */
int kx_eikonal(int32_t ndims, int32_t nx, int32_t ny, int32_t nz, double hx, double hy,
               double hz, const double *slow, int32_t n, const int32_t *src, double *T,
               int32_t max_iter, int32_t order, int32_t *flag, int32_t *iters, void *stream);

#ifdef __cplusplus
}
#endif
#endif
