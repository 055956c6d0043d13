// kirchx.hip -- MI355X (gfx950) Kirchhoff demigration / migration library
// (include/kirchx.h): error reporting, plan checks, the tiled table layout,
// the wavelet convolution, the operator composition (kx_forward /
// kx_adjoint) and the float64 reductions.  The forward launches live in
// kx_spread.hip, the adjoint launches in kx_gather.hip, the kernel bodies in
// kx_device.h.
//
// Data layout in HBM (all caller-owned):
//   model  m[ni]              reference order ii = (iy*nx + ix)*nz + iz
//   data   d[ns*nr][nt]       trace isrc*nr + irec
//   tables off[ntiles][n][64] traveltime in samples minus a per-(tile, n)
//                             integer base[ntiles][n]; a tile is a
//                             tz*tx*ty = 64 block of image points, one lane
//                             of a wavefront per point
// The split into an integer base and a small offset keeps the fractional
// sample position exact to ~1e-6 samples in fp32 (and to ~4e-3 in fp16),
// independent of how long the traveltime is.
//
// Forward (spread, replaces K:727-752 / K:783-889): one workgroup owns one
// source and a block of receivers; their traces live in LDS, image tiles are
// streamed through the waves, every (point, trace) pair adds two linearly
// interpolated samples with LDS atomics (64-bit fixed point for float data),
// and the traces are written to HBM once.  No global atomics.
// Adjoint (gather, replaces K:754-781 / K:891-1001): one wavefront owns a
// group of neighbouring image tiles and a block of sources; receiver offsets
// are streamed coalesced, each source's trace window is staged in LDS as
// (sample, difference) pairs, and each point reads one pair per trace.
// Every output element is written once.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>

#include "kx_internal.h"

// ---------------------------------------------------------------------------
// error reporting
// ---------------------------------------------------------------------------
static thread_local char g_err[512] = "";

static int kx_fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

#define KX_HIP(call)                                                           \
    do {                                                                       \
        hipError_t e_ = (call);                                                \
        if (e_ != hipSuccess)                                                  \
            return kx_fail((int)e_, "%s: %s", #call, hipGetErrorString(e_));   \
    } while (0)

// error hooks for the other translation units of the library (not exported)
extern "C" __attribute__((visibility("hidden"))) int kx_solver_fail(int code, const char *msg)
{
    return kx_fail(code, "%s", msg);
}

extern "C" __attribute__((visibility("hidden"))) int kx_failf(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

extern "C" int kx_version(void) { return KX_ABI_VERSION; }

namespace {
int check_plan(const kx_plan *p);
}

// plan validation for the other translation units of the library (not exported)
extern "C" __attribute__((visibility("hidden"))) int kx_plan_check(const kx_plan *p)
{
    return check_plan(p);
}
extern "C" const char *kx_last_error(void) { return g_err; }

#include "kx_device.h"

// ---------------------------------------------------------------------------
// single-problem kernels
// ---------------------------------------------------------------------------
namespace {

template <typename TD>
__global__ __launch_bounds__(kConvThreads) void conv_kernel(const TD *__restrict__ x, int64_t xstride,
                                                      TD *__restrict__ y, int nt,
                                                      const TD *__restrict__ h, int64_t hstride,
                                                      int nh, int nseg)
{
    conv_segment<TD>(x, xstride, y, nt, h, hstride, nh, blockIdx.x / nseg,
                     (blockIdx.x % nseg) * kConvOut);
}

// y[r * nlag + j] = sum_u a[r * astride + u] b[r * bstride + lag0 + j + u]
template <typename TD>
__global__ __launch_bounds__(kXcorrThreads) void xcorr_kernel(const TD *__restrict__ a,
                                                        int64_t astride, int la,
                                                        const TD *__restrict__ b,
                                                        int64_t bstride, int lb,
                                                        TD *__restrict__ y, int nlag, int lag0)
{
    const int64_t r = blockIdx.x / nlag;
    const int j = (int)(blockIdx.x % nlag);
    const TD v = xcorr_lag<TD>(a + r * astride, la, b + r * bstride, lb, lag0 + j);
    if (threadIdx.x == 0)
        y[r * nlag + j] = v;
}

// ---------------------------------------------------------------------------
// table layout
// ---------------------------------------------------------------------------
template <typename TT> __device__ __forceinline__ TT to_tab(double v);
template <> __device__ __forceinline__ float to_tab<float>(double v) { return (float)v; }
template <> __device__ __forceinline__ double to_tab<double>(double v) { return v; }
template <> __device__ __forceinline__ __half to_tab<__half>(double v) { return __float2half((float)v); }

// Offset of an entry that no trace sample can take (marked NaN by the
// caller: its time lies past the last sample, or before -1, whatever the
// other side's time): finite, exact, and so large that every sum with the
// other side's offset and both bases (each within +-2^27) falls past the
// trace while two such offsets still add up to less than 2^31, so the
// kernels' bounds tests skip it without any conversion saturating.  Such
// entries take no part in their tile's base, which keeps the others'
// offsets small.
template <typename TT> __device__ __forceinline__ TT unreached_offset() { return TT(0x1p29); }
template <> __device__ __forceinline__ __half unreached_offset<__half>()
{
    return __float2half(65504.0f);  // largest finite float16
}

template <typename TT>
__global__ __launch_bounds__(kBlock) void encode_kernel(KArgs a, const double *__restrict__ tabT,
                                                        int n, int64_t ni, double scale,
                                                        TT *__restrict__ off, int *__restrict__ base)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t w = (int64_t)blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave;
    if (w >= (int64_t)n * a.ntiles)
        return;
    const int j = (int)(w / a.ntiles);
    const int t = (int)(w % a.ntiles);
    TilePoint p = tile_point(a, t, lane);
    // padding lanes replicate the nearest grid point: keeps offsets small
    p.iz = min(p.iz, a.nz - 1);
    p.ix = min(p.ix, a.nx - 1);
    p.iy = min(p.iy, a.ny - 1);
    const double q = tabT[(size_t)j * ni + point_index(a, p)] * scale;  // NaN: unreached
    double mn = q, mx = q;  // fmin / fmax pass over NaN
#pragma unroll
    for (int o = kWave / 2; o >= 1; o >>= 1) {
        mn = fmin(mn, __shfl_xor(mn, o, kWave));
        mx = fmax(mx, __shfl_xor(mx, o, kWave));
    }
    // NaN when no entry is reached, and then -2^27.  Bases lie in +-2^27
    // (134M samples, far past any trace), so that the adjoint's sums of the
    // source and receiver bases of four tiles stay inside 32 bits
    double b = floor(0.5 * (mn + mx) + 0.5);
    b = fmin(fmax(b, -134217728.0), 134217728.0);
    const size_t o = (size_t)t * n + j;  // tile-major output
    off[o * kWave + lane] = q == q ? to_tab<TT>(q - b) : unreached_offset<TT>();
    if (lane == 0)
        base[o] = (int)b;
}

template <typename TD>
__global__ __launch_bounds__(kBlock) void tile_kernel(KArgs a, const double *__restrict__ tabT,
                                                      int n, int64_t ni, TD *__restrict__ out)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t w = (int64_t)blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave;
    if (w >= (int64_t)n * a.ntiles)
        return;
    const int j = (int)(w / a.ntiles);
    const int t = (int)(w % a.ntiles);
    const TilePoint p = tile_point(a, t, lane);
    const double v = p.valid ? tabT[(size_t)j * ni + point_index(a, p)] : 0.0;
    if constexpr (sizeof(TD) == 2)
        out[((size_t)t * n + j) * kWave + lane] = __float2half_rn((float)v);
    else
        out[((size_t)t * n + j) * kWave + lane] = (TD)v;
}

// ---------------------------------------------------------------------------
// reductions (float64 accumulation)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double block_sum(double v)
{
    __shared__ double part[kBlock / kWave];
#pragma unroll
    for (int o = kWave / 2; o >= 1; o >>= 1)
        v += __shfl_xor(v, o, kWave);
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    if (lane == 0)
        part[wave] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0) {
        for (int i = 0; i < (int)(blockDim.x / kWave); ++i)
            s += part[i];
    }
    return s;  // valid in thread 0
}

template <typename TD>
__global__ __launch_bounds__(kBlock) void dot_partial_kernel(const TD *__restrict__ a,
                                                             const TD *__restrict__ b, int64_t n,
                                                             double *__restrict__ partial)
{
    double acc = 0.0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        acc += (double)a[i] * (double)b[i];
    const double s = block_sum(acc);
    if (threadIdx.x == 0)
        partial[blockIdx.x] = s;
}

__global__ __launch_bounds__(kBlock) void sum_final_kernel(const double *__restrict__ partial,
                                                           int np, double *__restrict__ out)
{
    double acc = 0.0;
    for (int i = threadIdx.x; i < np; i += blockDim.x)
        acc += partial[i];
    const double s = block_sum(acc);
    if (threadIdx.x == 0)
        out[0] = s;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
int dsize(int dtype) { return dtype == KX_F64 ? 8 : 4; }

int check_geom(const kx_plan *p)
{
    if (!p)
        return kx_fail(KX_EINVAL, "plan is NULL");
    if (p->abi_version != KX_ABI_VERSION)
        return kx_fail(KX_EABI, "plan abi_version %d, library %d", p->abi_version, KX_ABI_VERSION);
    if (p->nx <= 0 || p->ny <= 0 || p->nz <= 0)
        return kx_fail(KX_ESHAPE, "image grid must be positive (nx=%d ny=%d nz=%d)", p->nx, p->ny, p->nz);
    if (p->tz <= 0 || p->tx <= 0 || p->ty <= 0 || p->tz * p->tx * p->ty != kWave)
        return kx_fail(KX_ESHAPE, "tile %dx%dx%d must hold 64 points", p->tz, p->tx, p->ty);
    if (p->ntz != (p->nz + p->tz - 1) / p->tz || p->ntx != (p->nx + p->tx - 1) / p->tx ||
        p->nty != (p->ny + p->ty - 1) / p->ty ||
        p->ntiles != (int64_t)p->ntz * p->ntx * p->nty)
        return kx_fail(KX_ESHAPE, "tile counts do not match the grid");
    if (p->ntiles * kWave > INT32_MAX)
        return kx_fail(KX_ESHAPE, "too many image points for 32-bit tile indexing");
    return KX_OK;
}

int check_plan(const kx_plan *p)
{
    int rc = check_geom(p);
    if (rc)
        return rc;
    if (p->dtype != KX_F32 && p->dtype != KX_F64)
        return kx_fail(KX_ETYPE, "dtype must be KX_F32 or KX_F64");
    const bool ok_fmt = (p->dtype == KX_F32 && (p->tab_fmt == KX_F32 || p->tab_fmt == KX_F16)) ||
                        (p->dtype == KX_F64 && p->tab_fmt == KX_F64);
    if (!ok_fmt)
        return kx_fail(KX_ETYPE, "table format %d not supported with dtype %d", p->tab_fmt, p->dtype);
    if (p->ns <= 0 || p->nr <= 0 || p->nt < 2)
        return kx_fail(KX_ESHAPE, "need ns, nr >= 1 and nt >= 2 (ns=%d nr=%d nt=%d)", p->ns, p->nr, p->nt);
    if ((int64_t)p->ns * p->nr > INT32_MAX)
        return kx_fail(KX_ESHAPE, "too many traces");
    if (!p->off_s || !p->base_s || !p->off_r || !p->base_r)
        return kx_fail(KX_EINVAL, "traveltime tables missing");
    if (p->dynamic && (!p->wp_s || !p->wq_s || !p->wp_r || !p->wq_r || !p->izmin))
        return kx_fail(KX_EINVAL, "dynamic weights missing");
    return KX_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" int64_t kx_workspace_bytes(const kx_plan *p)
{
    if (check_plan(p))
        return -1;
    return (int64_t)p->ns * p->nr * p->nt * dsize(p->dtype);
}

extern "C" int kx_encode_table(const kx_plan *geom, const double *tabT, int32_t n, double scale,
                               int32_t fmt, void *off, int32_t *base, void *stream)
{
    int rc = check_geom(geom);
    if (rc)
        return rc;
    if (!tabT || !off || !base || n <= 0)
        return kx_fail(KX_EINVAL, "encode_table: bad arguments");
    const KArgs a = make_args(geom);
    const int64_t ni = (int64_t)geom->nx * geom->ny * geom->nz;
    const int64_t nwaves = (int64_t)n * geom->ntiles;
    const int64_t nblk = (nwaves + 3) / 4;
    if (nblk > INT32_MAX)
        return kx_fail(KX_ESHAPE, "encode_table: table too large");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    switch (fmt) {
    case KX_F32:
        hipLaunchKernelGGL(encode_kernel<float>, dim3((int)nblk), dim3(kBlock), 0, st, a, tabT, n,
                           ni, scale, reinterpret_cast<float *>(off), base);
        break;
    case KX_F16:
        hipLaunchKernelGGL(encode_kernel<__half>, dim3((int)nblk), dim3(kBlock), 0, st, a, tabT, n,
                           ni, scale, reinterpret_cast<__half *>(off), base);
        break;
    case KX_F64:
        hipLaunchKernelGGL(encode_kernel<double>, dim3((int)nblk), dim3(kBlock), 0, st, a, tabT, n,
                           ni, scale, reinterpret_cast<double *>(off), base);
        break;
    default:
        return kx_fail(KX_ETYPE, "encode_table: unknown format %d", fmt);
    }
    KX_LAUNCHED("encode_table");
    return KX_OK;
}

extern "C" int kx_tile_table(const kx_plan *geom, const double *tabT, int32_t n, int32_t fmt,
                             void *out, void *stream)
{
    int rc = check_geom(geom);
    if (rc)
        return rc;
    if (!tabT || !out || n <= 0)
        return kx_fail(KX_EINVAL, "tile_table: bad arguments");
    const KArgs a = make_args(geom);
    const int64_t ni = (int64_t)geom->nx * geom->ny * geom->nz;
    const int64_t nblk = ((int64_t)n * geom->ntiles + 3) / 4;
    if (nblk > INT32_MAX)
        return kx_fail(KX_ESHAPE, "tile_table: table too large");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    if (fmt == KX_F32)
        hipLaunchKernelGGL(tile_kernel<float>, dim3((int)nblk), dim3(kBlock), 0, st, a, tabT, n,
                           ni, reinterpret_cast<float *>(out));
    else if (fmt == KX_F64)
        hipLaunchKernelGGL(tile_kernel<double>, dim3((int)nblk), dim3(kBlock), 0, st, a, tabT, n,
                           ni, reinterpret_cast<double *>(out));
    else if (fmt == KX_F16)
        hipLaunchKernelGGL(tile_kernel<__half>, dim3((int)nblk), dim3(kBlock), 0, st, a, tabT, n,
                           ni, reinterpret_cast<__half *>(out));
    else
        return kx_fail(KX_ETYPE, "tile_table: unknown format %d", fmt);
    KX_LAUNCHED("tile_table");
    return KX_OK;
}

extern "C" int kx_conv_rows(const void *x, int64_t xstride, void *y, int64_t ntr, int32_t nt,
                            const void *h, int64_t hstride, int32_t nh, int32_t dtype,
                            void *stream)
{
    if (!x || !y || !h || x == y)
        return kx_fail(KX_EINVAL, "conv: bad buffers");
    if (ntr <= 0 || nt <= 0 || nh <= 0 || (nh % 2) == 0)
        return kx_fail(KX_ESHAPE, "conv: need ntr, nt > 0 and odd nh (nh=%d)", nh);
    if (xstride < 0 || hstride < 0)
        return kx_fail(KX_ESHAPE, "conv: negative row stride");
    const int nseg = (nt + kConvOut - 1) / kConvOut;
    if (ntr * nseg > INT32_MAX)
        return kx_fail(KX_ESHAPE, "conv: too many traces");
    const size_t lds = (size_t)conv_lds_elems(nh) * (dtype == KX_F64 ? 8 : 4);
    if (lds > (size_t)kLdsBudget)
        return kx_fail(KX_ELDS, "conv: filter of %d taps does not fit LDS", nh);
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const dim3 grid((unsigned)(ntr * nseg));
    if (dtype == KX_F32)
        hipLaunchKernelGGL(conv_kernel<float>, grid, dim3(kConvThreads), lds, st,
                           reinterpret_cast<const float *>(x), xstride, reinterpret_cast<float *>(y),
                           nt, reinterpret_cast<const float *>(h), hstride, nh, nseg);
    else if (dtype == KX_F64)
        hipLaunchKernelGGL(conv_kernel<double>, grid, dim3(kConvThreads), lds, st,
                           reinterpret_cast<const double *>(x), xstride,
                           reinterpret_cast<double *>(y), nt, reinterpret_cast<const double *>(h),
                           hstride, nh, nseg);
    else
        return kx_fail(KX_ETYPE, "conv: dtype %d", dtype);
    KX_LAUNCHED("conv");
    return KX_OK;
}

extern "C" int kx_conv(const void *x, void *y, int64_t ntr, int32_t nt, const void *h, int32_t nh,
                       int32_t dtype, void *stream)
{
    return kx_conv_rows(x, nt, y, ntr, nt, h, 0, nh, dtype, stream);
}

extern "C" int kx_xcorr_rows(const void *a, int64_t astride, int32_t la, const void *b,
                             int64_t bstride, int32_t lb, void *y, int64_t ntr, int32_t nlag,
                             int32_t lag0, int32_t dtype, void *stream)
{
    if (!a || !b || !y || y == a || y == b)
        return kx_fail(KX_EINVAL, "xcorr: bad buffers");
    if (ntr <= 0 || nlag <= 0 || la <= 0 || lb <= 0 || astride < 0 || bstride < 0)
        return kx_fail(KX_ESHAPE, "xcorr: need ntr, nlag, la, lb > 0 and strides >= 0");
    if (ntr * nlag > INT32_MAX)
        return kx_fail(KX_ESHAPE, "xcorr: too many outputs");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const dim3 grid((unsigned)(ntr * nlag));
    if (dtype == KX_F32)
        hipLaunchKernelGGL(xcorr_kernel<float>, grid, dim3(kXcorrThreads), 0, st,
                           reinterpret_cast<const float *>(a), astride, la,
                           reinterpret_cast<const float *>(b), bstride, lb,
                           reinterpret_cast<float *>(y), nlag, lag0);
    else if (dtype == KX_F64)
        hipLaunchKernelGGL(xcorr_kernel<double>, grid, dim3(kXcorrThreads), 0, st,
                           reinterpret_cast<const double *>(a), astride, la,
                           reinterpret_cast<const double *>(b), bstride, lb,
                           reinterpret_cast<double *>(y), nlag, lag0);
    else
        return kx_fail(KX_ETYPE, "xcorr: dtype %d", dtype);
    KX_LAUNCHED("xcorr");
    return KX_OK;
}

extern "C" int kx_forward(const kx_plan *p, const void *m, void *d, void *work, void *stream)
{
    int rc = check_plan(p);
    if (rc)
        return rc;
    if (!p->h || p->nh <= 0)
        return kx_fail(KX_EINVAL, "forward: wavelet missing");
    if (!work)
        return kx_fail(KX_EINVAL, "forward: workspace missing");
    // one launch where the workgroups can convolve their traces in LDS
    rc = kx_spread_wavelet(p, m, d, stream);
    if (rc != KX_ENOFUSE)
        return rc;
    rc = kx_spread(p, m, work, stream);
    if (rc)
        return rc;
    return kx_conv(work, d, (int64_t)p->ns * p->nr, p->nt, p->h, p->nh, p->dtype, stream);
}

extern "C" int kx_adjoint(const kx_plan *p, const void *d, void *m, void *work, void *stream)
{
    int rc = check_plan(p);
    if (rc)
        return rc;
    if (!p->hstar || p->nh <= 0)
        return kx_fail(KX_EINVAL, "adjoint: wavelet missing");
    if (!work)
        return kx_fail(KX_EINVAL, "adjoint: workspace missing");
    rc = kx_conv(d, work, (int64_t)p->ns * p->nr, p->nt, p->hstar, p->nh, p->dtype, stream);
    if (rc)
        return rc;
    return kx_gather(p, work, m, stream);
}

extern "C" int64_t kx_reduce_workspace_bytes(void) { return (int64_t)kReduceBlocks * 8; }

extern "C" int kx_dot(const void *a, const void *b, int64_t n, int32_t dtype, double *out,
                      void *work, void *stream)
{
    if (!a || !b || !out || !work || n < 0)
        return kx_fail(KX_EINVAL, "dot: bad arguments");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    double *partial = reinterpret_cast<double *>(work);
    const int64_t want = (n + kBlock - 1) / kBlock;
    const int nblk = (int)std::max<int64_t>(1, std::min<int64_t>(kReduceBlocks, want));
    if (dtype == KX_F32)
        hipLaunchKernelGGL(dot_partial_kernel<float>, dim3(nblk), dim3(kBlock), 0, st,
                           reinterpret_cast<const float *>(a), reinterpret_cast<const float *>(b),
                           n, partial);
    else if (dtype == KX_F64)
        hipLaunchKernelGGL(dot_partial_kernel<double>, dim3(nblk), dim3(kBlock), 0, st,
                           reinterpret_cast<const double *>(a), reinterpret_cast<const double *>(b),
                           n, partial);
    else
        return kx_fail(KX_ETYPE, "dot: dtype %d", dtype);
    KX_LAUNCHED("dot");
    hipLaunchKernelGGL(sum_final_kernel, dim3(1), dim3(kBlock), 0, st, partial, nblk, out);
    KX_LAUNCHED("dot final");
    return KX_OK;
}
