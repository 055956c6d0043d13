// kirchx.hip -- MI355X (gfx950) Kirchhoff demigration / migration kernels
// and their C ABI (include/kirchx.h).
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
// Adjoint (gather, replaces K:754-781 / K:891-1001): one wavefront owns one
// image tile; source offsets for a block of sources are held in registers,
// receiver offsets are streamed coalesced, and each lane gathers its two
// samples per trace.  Every output element is written once.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>

#include "kirchx.h"

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

#define KX_LAUNCHED(name)                                                      \
    do {                                                                       \
        hipError_t e_ = hipGetLastError();                                     \
        if (e_ != hipSuccess)                                                  \
            return kx_fail((int)e_, "%s launch: %s", name, hipGetErrorString(e_)); \
    } while (0)

// error hook for the other translation units of the library (not exported)
extern "C" __attribute__((visibility("hidden"))) int kx_solver_fail(int code, const char *msg)
{
    return kx_fail(code, "%s", msg);
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

template <typename TD, typename TT, bool DYN, bool NEG, int ORD, int RB, bool XS>
__global__ __launch_bounds__(kSpreadBlock) void spread_kernel(KArgs a, const TD *__restrict__ m,
                                                              TD *__restrict__ y,
                                                              const SpreadScale *__restrict__ sc)
{
    const int L = xcd_remap(blockIdx.x, gridDim.x);
    const SpreadScale scale = *sc;
    spread_item<TD, TT, DYN, NEG, ORD, RB, XS>(a, m, y, scale, L % a.ns, (L / a.ns) * RB);
}

// K models of a float32 operator; workgroup L (after the XCD renumbering) =
// source L mod ns, receiver L / ns.  The launch accumulates in fixed point
// when every model's scale allows it, in float64 otherwise.
template <typename TT, bool DYN, bool NEG, int K>
__global__ __launch_bounds__(kSpreadBlock) void spread_models_kernel(
    KArgs a, const float *__restrict__ m, int64_t ni, float *__restrict__ y, int64_t ystride,
    const SpreadScale *__restrict__ sc)
{
    const int L = xcd_remap(blockIdx.x, gridDim.x);
    bool fixed = true;
#pragma unroll
    for (int j = 0; j < K; ++j)
        fixed = fixed && sc[j].fixed;
    if (fixed)
        spread_models_body<TT, DYN, NEG, true, K>(a, m, ni, y, ystride, sc, L % a.ns, L / a.ns);
    else
        spread_models_body<TT, DYN, NEG, false, K>(a, m, ni, y, ystride, sc, L % a.ns, L / a.ns);
}

// K models (K > 1: model k's traces at d + k*dstride, its image at
// m + k*mstride); groups of NP tiles per wave
template <typename TD, typename TT, bool DYN, bool NEG, int JW, int P, int NP, int K>
__global__ __launch_bounds__(kGatherBlock) void gather_kernel(KArgs a, const TD *__restrict__ d,
                                                              TD *__restrict__ m,
                                                              int64_t dstride, int64_t mstride)
{
    gather_tiles<TD, TT, DYN, NEG, JW, P, NP, K>(a, d, m, xcd_remap(blockIdx.x, gridDim.x),
                                                 dstride, mstride);
}

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
    const double q = tabT[(size_t)j * ni + point_index(a, p)] * scale;
    double mn = q, mx = q;
#pragma unroll
    for (int o = kWave / 2; o >= 1; o >>= 1) {
        mn = fmin(mn, __shfl_xor(mn, o, kWave));
        mx = fmax(mx, __shfl_xor(mx, o, kWave));
    }
    double b = floor(0.5 * (mn + mx) + 0.5);
    b = fmin(fmax(b, -536870912.0), 536870912.0);
    const size_t o = (size_t)t * n + j;  // tile-major output
    off[o * kWave + lane] = to_tab<TT>(q - b);
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
    out[((size_t)t * n + j) * kWave + lane] =
        p.valid ? (TD)tabT[(size_t)j * ni + point_index(a, p)] : TD(0);
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

constexpr int kRblocks[] = {16, 8, 4, 2, 1};
constexpr int kLdsPerCu = 160 * 1024;
constexpr int kMaxWavesPerCu = 32;

size_t spread_lds_bytes(int rb, int nt) { return (size_t)rb * nt * 8; }

template <typename TD, typename TT, bool DYN, bool NEG, bool XS, int ORD, int RB>
int launch_spread_rb(const KArgs &a, const void *m, void *y, const SpreadScale *sc,
                     hipStream_t st)
{
    const size_t lds = spread_lds_bytes(RB, a.nt);
    const int nrb = (a.nr + RB - 1) / RB;
    const int nblk = a.ns * nrb;
    allow_lds(reinterpret_cast<const void *>(&spread_kernel<TD, TT, DYN, NEG, ORD, RB, XS>), lds);
    hipLaunchKernelGGL((spread_kernel<TD, TT, DYN, NEG, ORD, RB, XS>), dim3(nblk),
                       dim3(kSpreadBlock), lds, st, a, reinterpret_cast<const TD *>(m),
                       reinterpret_cast<TD *>(y), sc);
    KX_LAUNCHED("spread");
    return KX_OK;
}

template <typename TD, typename TT, bool DYN, bool NEG, bool XS, int ORD>
const void *spread_fn(int rb)
{
    switch (rb) {
    case 16:
        return reinterpret_cast<const void *>(&spread_kernel<TD, TT, DYN, NEG, ORD, 16, XS>);
    case 8:
        return reinterpret_cast<const void *>(&spread_kernel<TD, TT, DYN, NEG, ORD, 8, XS>);
    case 4:
        return reinterpret_cast<const void *>(&spread_kernel<TD, TT, DYN, NEG, ORD, 4, XS>);
    case 2:
        return reinterpret_cast<const void *>(&spread_kernel<TD, TT, DYN, NEG, ORD, 2, XS>);
    default:
        return reinterpret_cast<const void *>(&spread_kernel<TD, TT, DYN, NEG, kOrdNatural, 1, XS>);
    }
}

// Workgroups of one spread kernel resident per CU (registers, LDS, waves);
// without a device, the LDS and wave limits alone.
int spread_per_cu(const void *fn, int rb, int nt)
{
    const size_t lds = spread_lds_bytes(rb, nt);
    allow_lds(fn, lds);
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fn, kSpreadBlock, lds) != hipSuccess ||
        n <= 0) {
        (void)hipGetLastError();
        n = std::min((int)(kLdsPerCu / lds), kMaxWavesPerCu / kSpreadWaves);
    }
    return std::max(n, 1);
}

// Receivers per workgroup.  A requested value is rounded down to a supported
// one whose traces fit in LDS.  Automatic: the block that keeps the most
// traces resident per CU, weighted by how far the resident waves (up to 24)
// can hide the latency of the table loads; ties go to the larger block,
// which re-reads the image and the source table less often.
template <typename TD, typename TT, bool DYN, bool NEG, bool XS, int ORD>
int pick_rblock(int requested, int nr, int nt, int max_rb)
{
    auto fits = [&](int rb) {
        return rb <= max_rb && spread_lds_bytes(rb, nt) <= lds_budget();
    };
    if (requested > 0) {
        for (int rb : kRblocks)
            if (rb <= requested && fits(rb))
                return rb;
        return 1;
    }
    int best = 1;
    double best_score = -1.0;
    for (int rb : kRblocks) {
        if (!fits(rb) || (rb > 1 && rb / 2 >= nr))
            continue;
        const int per_cu = spread_per_cu(spread_fn<TD, TT, DYN, NEG, XS, ORD>(rb), rb, nt);
        const double waves = std::min(24.0, (double)per_cu * kSpreadWaves);
        const double score = waves / 24.0 * per_cu * std::min(rb, nr);
        if (score > best_score * 1.0001) {
            best = rb;
            best_score = score;
        }
    }
    return best;
}

template <typename TD, typename TT, bool DYN, bool NEG, bool XS, int ORD>
int launch_spread_o(KArgs a, int requested, int max_rb, const void *m, void *y,
                    const SpreadScale *sc, hipStream_t st)
{
    a.rblock = pick_rblock<TD, TT, DYN, NEG, XS, ORD>(requested, a.nr, a.nt, max_rb);
    switch (a.rblock) {
    case 16:
        return launch_spread_rb<TD, TT, DYN, NEG, XS, ORD, 16>(a, m, y, sc, st);
    case 8:
        return launch_spread_rb<TD, TT, DYN, NEG, XS, ORD, 8>(a, m, y, sc, st);
    case 4:
        return launch_spread_rb<TD, TT, DYN, NEG, XS, ORD, 4>(a, m, y, sc, st);
    case 2:
        return launch_spread_rb<TD, TT, DYN, NEG, XS, ORD, 2>(a, m, y, sc, st);
    default:
        return launch_spread_rb<TD, TT, DYN, NEG, XS, kOrdNatural, 1>(a, m, y, sc, st);
    }
}

template <typename TD, typename TT, bool DYN, bool NEG, bool XS>
int launch_spread_t(const KArgs &a, int order, int requested, int max_rb, const void *m, void *y,
                    const SpreadScale *sc, hipStream_t st)
{
    switch (order) {
    case KX_ORDER_ROTATED:
        return launch_spread_o<TD, TT, DYN, NEG, XS, kOrdRotated>(a, requested, max_rb, m, y, sc,
                                                                  st);
    case KX_ORDER_GROUPED:
        return launch_spread_o<TD, TT, DYN, NEG, XS, kOrdGrouped>(a, requested, max_rb, m, y, sc,
                                                                  st);
    default:
        return launch_spread_o<TD, TT, DYN, NEG, XS, kOrdNatural>(a, requested, max_rb, m, y, sc,
                                                                  st);
    }
}


template <typename TD, typename TT>
int launch_spread_d(const kx_plan *p, KArgs a, const void *m, void *y, hipStream_t st)
{
    if (spread_lds_bytes(1, p->nt) > lds_budget())
        return kx_fail(KX_ELDS, "nt=%d does not fit one trace in LDS", p->nt);
    if (sizeof(TD) == 4 && !p->aux)
        return kx_fail(KX_EINVAL, "spread: plan.aux (KX_AUX_BYTES of device scratch) is NULL");
    // registers per lane grow with the block: keep two workgroups per CU
    const int max_rb = sizeof(TD) == 8 ? 4 : (p->dynamic ? 8 : 16);
    SpreadScale *sc = reinterpret_cast<SpreadScale *>(p->aux);
    if constexpr (sizeof(TD) == 4) {
        // fixed-point scale of this launch from max|m|
        const int64_t ni = (int64_t)p->nx * p->ny * p->nz;
        float *part = reinterpret_cast<float *>(p->aux) + 16;
        const int nb = (int)std::min<int64_t>(kScaleBlocks, (ni + kBlock - 1) / kBlock);
        hipLaunchKernelGGL(absmax_partial_kernel, dim3(nb), dim3(kBlock), 0, st,
                           reinterpret_cast<const float *>(m), ni, part);
        KX_LAUNCHED("spread scale");
        hipLaunchKernelGGL(scale_final_kernel, dim3(1), dim3(kWave), 0, st, part, nb, p->wamp, sc);
        KX_LAUNCHED("spread scale");
    }
    // A workgroup blocking sources (float32 operators) runs the spread with
    // the two sides' tables exchanged and writes its traces to the
    // transposed positions: T = os + or and the integer bases add in either
    // order, and the dynamic kernel (XS) forms the source-side products per
    // step, so the traces are bitwise those of blocks of receivers.
    // Automatic: a 2D kinematic plan blocks the side with fewer members,
    // whose rows then serve many workgroups through L2 (13% faster on a
    // 61 x 301 geometry, either way round); 3D and dynamic plans block
    // receivers (measured no faster otherwise).
    int side = p->sside;
    if (side == KX_SIDE_AUTO)
        side = (!p->dynamic && p->ny == 1 && p->ns < p->nr) ? KX_SIDE_SOURCES : KX_SIDE_RECEIVERS;
    const bool xs = sizeof(TD) == 4 && side == KX_SIDE_SOURCES;
    if (xs)
        exchange_sides(a);
    // automatic lane order: rotated when a step loads one receiver table,
    // natural when it loads three (dynamic weights), where the scattered
    // loads cost more than the collisions they avoid
    int order = p->sorder;
    if (order == KX_ORDER_AUTO)
        order = p->dynamic ? KX_ORDER_NATURAL : KX_ORDER_ROTATED;
    // exchanged-side dynamic kernels exist for float32 operators only (xs
    // is false otherwise)
    constexpr bool kXS = sizeof(TD) == 4;
    if (p->dynamic && xs)
        return p->neg_trav ? launch_spread_t<TD, TT, true, true, kXS>(a, order, p->rblock, max_rb,
                                                                      m, y, sc, st)
                           : launch_spread_t<TD, TT, true, false, kXS>(a, order, p->rblock,
                                                                       max_rb, m, y, sc, st);
    if (p->dynamic)
        return p->neg_trav ? launch_spread_t<TD, TT, true, true, false>(a, order, p->rblock,
                                                                        max_rb, m, y, sc, st)
                           : launch_spread_t<TD, TT, true, false, false>(a, order, p->rblock,
                                                                         max_rb, m, y, sc, st);
    return p->neg_trav ? launch_spread_t<TD, TT, false, true, false>(a, order, p->rblock, max_rb,
                                                                     m, y, sc, st)
                       : launch_spread_t<TD, TT, false, false, false>(a, order, p->rblock,
                                                                      max_rb, m, y, sc, st);
}

// sources per wave and pass (measured): every pass re-reads the receiver
// table -- three tables when dynamic -- and more sources hold more
// registers, lowering occupancy
template <bool DYN, int NP, int K> struct GatherJW {
    static constexpr int v = (DYN && NP == 1 && K == 1) ? 8 : 4;
};

template <typename TD, typename TT, bool DYN, bool NEG, int P, int NP, int K>
int launch_gather_p(KArgs a, const void *d, int64_t dstride, void *m, int64_t mstride,
                    hipStream_t st)
{
    constexpr int JW = GatherJW<DYN, NP, K>::v;
    constexpr int NG = kGatherWaves / P;
    if ((int64_t)JW * P * a.nr * a.nt >= ((int64_t)1 << 31))
        return kx_fail(KX_ESHAPE, "gather: %d sources x nr x nt exceeds 2^31 samples", JW * P);
    set_tile_groups(a, NP);
    const int nblk = (a.ngroups + NG - 1) / NG;
    hipLaunchKernelGGL((gather_kernel<TD, TT, DYN, NEG, JW, P, NP, K>), dim3(nblk),
                       dim3(kGatherBlock), 0, st, a, reinterpret_cast<const TD *>(d),
                       reinterpret_cast<TD *>(m), dstride, mstride);
    KX_LAUNCHED("gather");
    return KX_OK;
}

// Waves per tile group: the fewest that give every CU kGatherFill waves.
constexpr int kGatherFill = 24;

int device_cus()
{
    int cus = 256, dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
        (void)hipGetLastError();
        cus = 256;
    }
    return cus;
}

// Tiles per wave (NP): requested (rounded down to 1, 2 or 4), or the most
// whose groups still give every CU kGatherFill waves with at most
// kGatherWaves waves per group.
int gather_tiles_per_wave(const KArgs &a, int gtiles)
{
    if (gtiles > 0)
        return gtiles >= 4 ? 4 : (gtiles >= 2 ? 2 : 1);
    const int64_t want = (int64_t)kGatherFill * device_cus();
    for (int np : {4, 2}) {
        KArgs b = a;
        set_tile_groups(b, np);
        if ((int64_t)b.ngroups * kGatherWaves >= want)
            return np;
    }
    return 1;
}

int gather_split(const KArgs &a, int np, int gsplit)
{
    int P = 1;
    if (gsplit > 0) {
        while (P * 2 <= std::min(gsplit, kGatherWaves))
            P *= 2;
    } else {
        KArgs b = a;
        set_tile_groups(b, np);
        const int64_t want = (int64_t)kGatherFill * device_cus();
        while (P < kGatherWaves && (int64_t)b.ngroups * P < want)
            P *= 2;
    }
    return P;
}

template <typename TD, typename TT, bool DYN, bool NEG, int NP, int K>
int launch_gather_np(const KArgs &a, int P, const void *d, void *m, hipStream_t st,
                     int64_t dstride, int64_t mstride)
{
    switch (P) {
    case 1:
        return launch_gather_p<TD, TT, DYN, NEG, 1, NP, K>(a, d, dstride, m, mstride, st);
    case 2:
        return launch_gather_p<TD, TT, DYN, NEG, 2, NP, K>(a, d, dstride, m, mstride, st);
    default:
        return launch_gather_p<TD, TT, DYN, NEG, 4, NP, K>(a, d, dstride, m, mstride, st);
    }
}

template <typename TD, typename TT, bool DYN, bool NEG, int K = 1>
int launch_gather_t(const KArgs &a, int gsplit, int gtiles, const void *d, void *m,
                    hipStream_t st, int64_t dstride = 0, int64_t mstride = 0)
{
    const int np = gather_tiles_per_wave(a, gtiles);
    const int P = gather_split(a, np, gsplit);
    switch (np) {
    case 4:
        return launch_gather_np<TD, TT, DYN, NEG, 4, K>(a, P, d, m, st, dstride, mstride);
    case 2:
        return launch_gather_np<TD, TT, DYN, NEG, 2, K>(a, P, d, m, st, dstride, mstride);
    default:
        return launch_gather_np<TD, TT, DYN, NEG, 1, K>(a, P, d, m, st, dstride, mstride);
    }
}

template <typename TD, typename TT>
int launch_gather_d(const kx_plan *p, const KArgs &a, const void *d, void *m, hipStream_t st)
{
    const int g = p->gsplit, n = p->gtiles;
    if (p->dynamic)
        return p->neg_trav ? launch_gather_t<TD, TT, true, true>(a, g, n, d, m, st)
                           : launch_gather_t<TD, TT, true, false>(a, g, n, d, m, st);
    return p->neg_trav ? launch_gather_t<TD, TT, false, true>(a, g, n, d, m, st)
                       : launch_gather_t<TD, TT, false, false>(a, g, n, d, m, st);
}

// K float32 models per launch, K in {4, 2}, then one by one.
template <typename TT, bool DYN, bool NEG>
int launch_gather_models(const kx_plan *p, const KArgs &a, const float *d, float *m, int nmodel,
                         hipStream_t st)
{
    const int64_t ni = (int64_t)p->nx * p->ny * p->nz;
    const int64_t dstride = (int64_t)p->ns * p->nr * p->nt;
    const int g = p->gsplit, n = p->gtiles;
    int j = 0, rc = KX_OK;
    while (j < nmodel && rc == KX_OK) {
        const int left = nmodel - j;
        const float *dj = d + j * dstride;
        float *mj = m + j * ni;
        if (left >= 4) {
            rc = launch_gather_t<float, TT, DYN, NEG, 4>(a, g, n, dj, mj, st, dstride, ni);
            j += 4;
        } else if (left >= 2) {
            rc = launch_gather_t<float, TT, DYN, NEG, 2>(a, g, n, dj, mj, st, dstride, ni);
            j += 2;
        } else {
            rc = launch_gather_t<float, TT, DYN, NEG, 1>(a, g, n, dj, mj, st);
            j += 1;
        }
    }
    return rc;
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
    else
        return kx_fail(KX_ETYPE, "tile_table: unknown format %d", fmt);
    KX_LAUNCHED("tile_table");
    return KX_OK;
}

extern "C" int kx_spread(const kx_plan *p, const void *m, void *traces, void *stream)
{
    int rc = check_plan(p);
    if (rc)
        return rc;
    if (!m || !traces)
        return kx_fail(KX_EINVAL, "spread: NULL buffer");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const KArgs a = make_args(p);
    if (p->dtype == KX_F64)
        return launch_spread_d<double, double>(p, a, m, traces, st);
    if (p->tab_fmt == KX_F16)
        return launch_spread_d<float, __half>(p, a, m, traces, st);
    return launch_spread_d<float, float>(p, a, m, traces, st);
}

namespace {

constexpr int kMaxModels = 64;
constexpr size_t kModelScaleOffset = 2048;  // bytes into plan.aux

template <typename TT, bool DYN, bool NEG, int K>
void launch_models_k(const KArgs &a, const float *m, int64_t ni, float *y, int64_t ystride,
                     const SpreadScale *sc, hipStream_t st)
{
    const size_t lds = spread_lds_bytes(K, a.nt);
    hipLaunchKernelGGL((spread_models_kernel<TT, DYN, NEG, K>), dim3(a.ns * a.nr),
                       dim3(kSpreadBlock), lds, st, a, m, ni, y, ystride, sc);
}

template <typename TT, bool DYN, bool NEG>
int launch_models(const kx_plan *p, const KArgs &a, const float *m, float *y, int nmodel,
                  hipStream_t st)
{
    const int64_t ni = (int64_t)p->nx * p->ny * p->nz;
    const int64_t ystride = (int64_t)p->ns * p->nr * p->nt;
    SpreadScale *scs = reinterpret_cast<SpreadScale *>(static_cast<char *>(p->aux) +
                                                       kModelScaleOffset);
    float *part = reinterpret_cast<float *>(p->aux) + 16;
    const int nb = (int)std::min<int64_t>(kScaleBlocks, (ni + kBlock - 1) / kBlock);
    for (int j = 0; j < nmodel; ++j) {
        hipLaunchKernelGGL(absmax_partial_kernel, dim3(nb), dim3(kBlock), 0, st, m + j * ni, ni,
                           part);
        hipLaunchKernelGGL(scale_final_kernel, dim3(1), dim3(kWave), 0, st, part, nb, p->wamp,
                           scs + j);
    }
    KX_LAUNCHED("spread models scale");
    int j = 0;
    while (j < nmodel) {
        const int left = nmodel - j;
        auto fits = [&](int k) {
            return k <= left && spread_lds_bytes(k, a.nt) <= (size_t)kLdsBudget;
        };
        const float *mj = m + j * ni;
        float *yj = y + j * ystride;
        if (fits(8)) {
            launch_models_k<TT, DYN, NEG, 8>(a, mj, ni, yj, ystride, scs + j, st);
            j += 8;
        } else if (fits(4)) {
            launch_models_k<TT, DYN, NEG, 4>(a, mj, ni, yj, ystride, scs + j, st);
            j += 4;
        } else if (fits(2)) {
            launch_models_k<TT, DYN, NEG, 2>(a, mj, ni, yj, ystride, scs + j, st);
            j += 2;
        } else {
            const int rc = launch_spread_d<float, TT>(p, a, mj, yj, st);
            if (rc)
                return rc;
            j += 1;
        }
        KX_LAUNCHED("spread models");
    }
    return KX_OK;
}

}  // namespace

extern "C" int kx_spread_models(const kx_plan *p, const void *m, void *traces, int32_t nmodel,
                                void *stream)
{
    int rc = check_plan(p);
    if (rc)
        return rc;
    if (!m || !traces)
        return kx_fail(KX_EINVAL, "spread_models: NULL buffer");
    if (p->dtype != KX_F32)
        return kx_fail(KX_ETYPE, "spread_models: float32 operators only");
    if (nmodel < 1 || nmodel > kMaxModels)
        return kx_fail(KX_ESHAPE, "spread_models: nmodel=%d outside [1, %d]", nmodel, kMaxModels);
    if (!p->aux)
        return kx_fail(KX_EINVAL, "spread_models: plan.aux (KX_AUX_BYTES of device scratch) is NULL");
    if (spread_lds_bytes(1, p->nt) > (size_t)kLdsBudget)
        return kx_fail(KX_ELDS, "nt=%d does not fit one trace in LDS", p->nt);
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const KArgs a = make_args(p);
    const float *mf = reinterpret_cast<const float *>(m);
    float *yf = reinterpret_cast<float *>(traces);
    if (p->tab_fmt == KX_F16)
        return p->dynamic ? (p->neg_trav ? launch_models<__half, true, true>(p, a, mf, yf, nmodel, st)
                                         : launch_models<__half, true, false>(p, a, mf, yf, nmodel, st))
                          : (p->neg_trav ? launch_models<__half, false, true>(p, a, mf, yf, nmodel, st)
                                         : launch_models<__half, false, false>(p, a, mf, yf, nmodel, st));
    return p->dynamic ? (p->neg_trav ? launch_models<float, true, true>(p, a, mf, yf, nmodel, st)
                                     : launch_models<float, true, false>(p, a, mf, yf, nmodel, st))
                      : (p->neg_trav ? launch_models<float, false, true>(p, a, mf, yf, nmodel, st)
                                     : launch_models<float, false, false>(p, a, mf, yf, nmodel, st));
}

extern "C" int kx_gather(const kx_plan *p, const void *traces, void *m, void *stream)
{
    int rc = check_plan(p);
    if (rc)
        return rc;
    if (!m || !traces)
        return kx_fail(KX_EINVAL, "gather: NULL buffer");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const KArgs a = make_args(p);
    if (p->dtype == KX_F64)
        return launch_gather_d<double, double>(p, a, traces, m, st);
    if (p->tab_fmt == KX_F16)
        return launch_gather_d<float, __half>(p, a, traces, m, st);
    return launch_gather_d<float, float>(p, a, traces, m, st);
}

extern "C" int kx_gather_models(const kx_plan *p, const void *traces, void *m, int32_t nmodel,
                                void *stream)
{
    int rc = check_plan(p);
    if (rc)
        return rc;
    if (!m || !traces)
        return kx_fail(KX_EINVAL, "gather_models: NULL buffer");
    if (p->dtype != KX_F32)
        return kx_fail(KX_ETYPE, "gather_models: float32 operators only");
    if (nmodel < 1)
        return kx_fail(KX_ESHAPE, "gather_models: nmodel=%d", nmodel);
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const KArgs a = make_args(p);
    const float *df = reinterpret_cast<const float *>(traces);
    float *mf = reinterpret_cast<float *>(m);
    if (p->tab_fmt == KX_F16)
        return p->dynamic ? (p->neg_trav ? launch_gather_models<__half, true, true>(p, a, df, mf, nmodel, st)
                                         : launch_gather_models<__half, true, false>(p, a, df, mf, nmodel, st))
                          : (p->neg_trav ? launch_gather_models<__half, false, true>(p, a, df, mf, nmodel, st)
                                         : launch_gather_models<__half, false, false>(p, a, df, mf, nmodel, st));
    return p->dynamic ? (p->neg_trav ? launch_gather_models<float, true, true>(p, a, df, mf, nmodel, st)
                                     : launch_gather_models<float, true, false>(p, a, df, mf, nmodel, st))
                      : (p->neg_trav ? launch_gather_models<float, false, true>(p, a, df, mf, nmodel, st)
                                     : launch_gather_models<float, false, false>(p, a, df, mf, nmodel, st));
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
