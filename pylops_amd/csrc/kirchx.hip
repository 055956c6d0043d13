// kirchx.hip -- MI355X (gfx950) Kirchhoff demigration / migration kernels
// and their C ABI (include/kirchx.h).
//
// Data layout in HBM (all caller-owned):
//   model  m[ni]              reference order ii = (iy*nx + ix)*nz + iz
//   data   d[ns*nr][nt]       trace isrc*nr + irec
//   tables off[n][ntiles][64] traveltime in samples minus a per-(n, tile)
//                             integer base[n][ntiles]; a tile is a
//                             tz*tx*ty = 64 block of image points, one lane
//                             of a wavefront per point
// The split into an integer base and a small offset keeps the fractional
// sample position exact to ~1e-6 samples in fp32 (and to ~4e-3 in fp16),
// independent of how long the traveltime is.
//
// Forward (spread, replaces K:727-752 / K:783-889): one workgroup owns one
// source and a block of receivers; their traces live in LDS, image tiles are
// streamed through the waves, every (point, trace) pair adds two linearly
// interpolated samples with LDS float atomics, and the traces are written to
// HBM once.  No global atomics.
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
extern "C" const char *kx_last_error(void) { return g_err; }

// ---------------------------------------------------------------------------
// device helpers
// ---------------------------------------------------------------------------
namespace {

constexpr int kWave = 64;
constexpr int kBlock = 256;
constexpr int kLdsBudget = 64 * 1024;
constexpr int kReduceBlocks = 1024;

template <typename TD> struct TimeType { using type = float; };
template <> struct TimeType<double> { using type = double; };

// LDS trace accumulator of the forward.  float64: on gfx950 an LDS float64
// atomic add sustains ~3 lanes/clk/CU against ~0.3 for float32, and the
// traces come out more accurate.
using AccT = double;

__device__ __forceinline__ float ld_off(const float *p) { return *p; }
__device__ __forceinline__ float ld_off(const __half *p) { return __half2float(*p); }
__device__ __forceinline__ double ld_off(const double *p) { return *p; }

// Hardware dispatches workgroup b to XCD (b mod 8).  Renumber so that each
// XCD receives a contiguous range of logical blocks: neighbouring image
// tiles (gather) or the same receiver block (spread) then share one L2.
__device__ __forceinline__ int xcd_remap(int b, int n)
{
    const int k = b & 7;
    const int q = n >> 3;
    const int rem = n & 7;
    return k * q + (k < rem ? k : rem) + (b >> 3);
}

struct KArgs {
    int ns, nr, nt;
    int nx, ny, nz;
    int tz, tx, ty;
    int ntz, ntx, nty;
    int ntiles;
    int rblock;
    const void *off_s;
    const int *base_s;
    const void *off_r;
    const int *base_r;
    const void *wp_s, *wq_s, *wp_r, *wq_r;
    const int *izmin;
};

struct TilePoint {
    int ix, iy, iz;
    bool valid;
};

__device__ __forceinline__ TilePoint tile_point(const KArgs &a, int t, int lane)
{
    const int lz = lane % a.tz;
    const int lx = (lane / a.tz) % a.tx;
    const int ly = lane / (a.tz * a.tx);
    const int tzi = t % a.ntz;
    const int rest = t / a.ntz;
    const int txi = rest % a.ntx;
    const int tyi = rest / a.ntx;
    TilePoint p;
    p.iz = tzi * a.tz + lz;
    p.ix = txi * a.tx + lx;
    p.iy = tyi * a.ty + ly;
    p.valid = (p.iz < a.nz) & (p.ix < a.nx) & (p.iy < a.ny);
    return p;
}

__device__ __forceinline__ int64_t point_index(const KArgs &a, const TilePoint &p)
{
    return ((int64_t)p.iy * a.nx + p.ix) * a.nz + p.iz;
}

// ---------------------------------------------------------------------------
// forward: spread into LDS-resident traces
// ---------------------------------------------------------------------------
// Everything one tile of one wave needs: the image value and the source /
// receiver offsets, bases and weights of its 64 points.
template <typename TD, typename TT, int RB> struct SpreadTile {
    TD mv;
    TT os;  // raw table elements: converted at use, so a prefetch never waits
    int bs;
    TT orr[RB];
    int br[RB];
    TD ps, qs;
    TD pr[RB], qr[RB];
    int iz;
};

// Opaque zero held in a VGPR: indexing with it keeps a wave-uniform load on
// the vector memory path.  Scalar loads share the lgkm counter with LDS
// operations, so waiting on a scalar load would also wait for every LDS
// atomic still in flight.
__device__ __forceinline__ int vgpr_zero()
{
    int z = 0;
    asm volatile("" : "+v"(z));
    return z;
}

template <typename TD, typename TT, bool DYN, int RB>
__device__ __forceinline__ void spread_load(const KArgs &a, int t, int lane, int nrl, int vz,
                                            const TD *__restrict__ m, const TT *offs,
                                            const int *bss, const TT *offr, const int *bsr,
                                            const TD *wps, const TD *wqs, const TD *wpr,
                                            const TD *wqr, size_t rstride,
                                            SpreadTile<TD, TT, RB> &L)
{
    const TilePoint p = tile_point(a, t, lane);
    L.iz = p.iz;
    L.mv = p.valid ? m[point_index(a, p)] : TD(0);
    const size_t tl = (size_t)t * kWave + lane;
    L.os = offs[tl];
    L.bs = bss[t + vz];
    if constexpr (DYN) {
        L.ps = wps[tl];
        L.qs = wqs[tl];
    }
#pragma unroll
    for (int j = 0; j < RB; ++j) {
        if (j < nrl) {
            L.orr[j] = offr[j * rstride + tl];
            L.br[j] = bsr[(size_t)j * a.ntiles + t + vz];
            if constexpr (DYN) {
                L.pr[j] = wpr[j * rstride + tl];
                L.qr[j] = wqr[j * rstride + tl];
            }
        }
    }
}

template <typename TD, typename TT, bool DYN, bool NEG, int RB>
__global__ __launch_bounds__(kBlock) void spread_kernel(KArgs a, const TD *__restrict__ m,
                                                        TD *__restrict__ y)
{
    using TC = typename TimeType<TD>::type;
    extern __shared__ __align__(16) unsigned char kx_smem[];
    AccT *acc = reinterpret_cast<AccT *>(kx_smem);

    // one workgroup = one source s and receivers [r0, r0 + RB)
    const int L = xcd_remap(blockIdx.x, gridDim.x);
    const int s = L % a.ns;
    const int r0 = (L / a.ns) * RB;
    const int nrl = min(RB, a.nr - r0);
    const int ntr = nrl * a.nt;

    for (int i = threadIdx.x; i < ntr; i += blockDim.x)
        acc[i] = AccT(0);
    __syncthreads();

    const int lane = threadIdx.x & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int nw = blockDim.x / kWave;
    const int vz = vgpr_zero();
    const size_t rstride = (size_t)a.ntiles * kWave;  // next receiver row
    const TT *offs = reinterpret_cast<const TT *>(a.off_s) + (size_t)s * rstride;
    const TT *offr = reinterpret_cast<const TT *>(a.off_r) + (size_t)r0 * rstride;
    const int *bss = a.base_s + (size_t)s * a.ntiles;
    const int *bsr = a.base_r + (size_t)r0 * a.ntiles;
    const TD *wps = reinterpret_cast<const TD *>(a.wp_s) + (size_t)s * rstride;
    const TD *wqs = reinterpret_cast<const TD *>(a.wq_s) + (size_t)s * rstride;
    const TD *wpr = reinterpret_cast<const TD *>(a.wp_r) + (size_t)r0 * rstride;
    const TD *wqr = reinterpret_cast<const TD *>(a.wq_r) + (size_t)r0 * rstride;

    int izmin[RB];
    if constexpr (DYN) {
#pragma unroll
        for (int j = 0; j < RB; ++j)
            izmin[j] = j < nrl ? a.izmin[s * a.nr + r0 + j] : 0;
    }

    SpreadTile<TD, TT, RB> cur, nxt;
    int t = wave;
    if (t < a.ntiles)
        spread_load<TD, TT, DYN, RB>(a, t, lane, nrl, vz, m, offs, bss, offr, bsr, wps, wqs,
                                     wpr, wqr, rstride, cur);
    for (; t < a.ntiles; t += nw) {
        // prefetch the next tile while this one is accumulated
        if (t + nw < a.ntiles)
            spread_load<TD, TT, DYN, RB>(a, t + nw, lane, nrl, vz, m, offs, bss, offr, bsr,
                                         wps, wqs, wpr, wqr, rstride, nxt);
        if (__ballot(cur.mv != TD(0)) != 0) {
            int it[RB];
            AccT w0[RB], w1[RB];
            const TC os = (TC)ld_off(&cur.os);
#pragma unroll
            for (int j = 0; j < RB; ++j) {
                const TC T = os + (TC)ld_off(&cur.orr[j]);
                const TC fl = floor(T);
                TC f = T - fl;
                int k = (int)fl + cur.bs + cur.br[j];
                if constexpr (NEG) {
                    // the reference truncates toward zero: a time in (-1, 0)
                    // lands on sample 0 with a negative weight
                    if (k == -1 && f > TC(0)) {
                        k = 0;
                        f -= TC(1);
                    }
                }
                TD w = cur.mv;
                if constexpr (DYN) {
                    w = cur.mv * (cur.ps * cur.pr[j] + cur.qs * cur.qr[j]);
                    if (cur.iz < izmin[j])
                        w = TD(0);
                }
                it[j] = (j < nrl && (unsigned)k < (unsigned)(a.nt - 1)) ? j * a.nt + k : -1;
                w0[j] = (AccT)(w * (TD)(TC(1) - f));
                w1[j] = (AccT)(w * (TD)f);
            }
#pragma unroll
            for (int j = 0; j < RB; ++j) {
                if (it[j] >= 0) {
                    atomicAdd(&acc[it[j]], w0[j]);
                    atomicAdd(&acc[it[j] + 1], w1[j]);
                }
            }
        }
        cur = nxt;
    }
    __syncthreads();
    TD *yo = y + ((size_t)s * a.nr + r0) * a.nt;
    for (int i = threadIdx.x; i < ntr; i += blockDim.x)
        yo[i] = (TD)acc[i];
}

// ---------------------------------------------------------------------------
// adjoint: gather per image point
// ---------------------------------------------------------------------------
template <typename TD, typename TT, bool DYN, bool NEG, int SB>
__global__ __launch_bounds__(kBlock) void gather_kernel(KArgs a, const TD *__restrict__ d,
                                                        TD *__restrict__ m)
{
    using TC = typename TimeType<TD>::type;
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int L = xcd_remap(blockIdx.x, gridDim.x);
    const int t = L * (kBlock / kWave) + wave;
    if (t >= a.ntiles)
        return;  // no barrier in this kernel
    const TilePoint p = tile_point(a, t, lane);
    const TT *offs = reinterpret_cast<const TT *>(a.off_s);
    const TT *offr = reinterpret_cast<const TT *>(a.off_r);
    const TD *wps = reinterpret_cast<const TD *>(a.wp_s);
    const TD *wqs = reinterpret_cast<const TD *>(a.wq_s);
    const TD *wpr = reinterpret_cast<const TD *>(a.wp_r);
    const TD *wqr = reinterpret_cast<const TD *>(a.wq_r);
    const unsigned ntm1 = (unsigned)(a.nt - 1);

    TD acc = TD(0);
    for (int s0 = 0; s0 < a.ns; s0 += SB) {
        const int nsc = min(SB, a.ns - s0);
        TC os[SB];
        int bs[SB];
        TD ps[SB], qs[SB];
#pragma unroll
        for (int j = 0; j < SB; ++j) {
            if (j < nsc) {
                const size_t so = (size_t)(s0 + j) * a.ntiles + t;
                os[j] = (TC)ld_off(offs + so * kWave + lane);
                bs[j] = a.base_s[so];
                if constexpr (DYN) {
                    ps[j] = wps[so * kWave + lane];
                    qs[j] = wqs[so * kWave + lane];
                }
            } else {
                os[j] = TC(0);
                bs[j] = 0;
                ps[j] = qs[j] = TD(0);
            }
        }
        for (int r = 0; r < a.nr; ++r) {
            const size_t ro = (size_t)r * a.ntiles + t;
            const TC orr = (TC)ld_off(offr + ro * kWave + lane);
            const int br = a.base_r[ro];
            TD pr = TD(0), qr = TD(0);
            if constexpr (DYN) {
                pr = wpr[ro * kWave + lane];
                qr = wqr[ro * kWave + lane];
            }
#pragma unroll
            for (int j = 0; j < SB; ++j) {
                if (j < nsc) {
                    const TC T = os[j] + orr;
                    const TC fl = floor(T);
                    TC f = T - fl;
                    int it = (int)fl + bs[j] + br;
                    if constexpr (NEG) {
                        if (it == -1 && f > TC(0)) {
                            it = 0;
                            f -= TC(1);
                        }
                    }
                    const int trace = (s0 + j) * a.nr + r;
                    if ((unsigned)it < ntm1) {
                        const TD *q = d + (size_t)trace * a.nt + it;
                        TD v = q[0] * (TD)(TC(1) - f) + q[1] * (TD)f;
                        if constexpr (DYN) {
                            TD w = ps[j] * pr + qs[j] * qr;
                            if (p.iz < a.izmin[trace])
                                w = TD(0);
                            v *= w;
                        }
                        acc += v;
                    }
                }
            }
        }
    }
    if (p.valid)
        m[point_index(a, p)] = acc;
}

// ---------------------------------------------------------------------------
// "same" convolution along time, odd filter
// ---------------------------------------------------------------------------
constexpr int kConvOut = 256;

template <typename TD>
__global__ __launch_bounds__(kBlock) void conv_kernel(const TD *__restrict__ x, TD *__restrict__ y,
                                                      int nt, const TD *__restrict__ h, int nh,
                                                      int nseg)
{
    extern __shared__ __align__(16) unsigned char kx_smem[];
    TD *hs = reinterpret_cast<TD *>(kx_smem);
    TD *win = hs + nh;
    const int64_t tr = blockIdx.x / nseg;
    const int n0 = (blockIdx.x % nseg) * kConvOut;
    const int c = (nh - 1) / 2;
    const int start = n0 + c - (nh - 1);
    const int nwin = kConvOut + nh - 1;
    const TD *xr = x + tr * nt;
    for (int k = threadIdx.x; k < nh; k += blockDim.x)
        hs[k] = h[nh - 1 - k];  // flipped: out[i] = sum_k hs[k] win[i + k]
    for (int j = threadIdx.x; j < nwin; j += blockDim.x) {
        const int n = start + j;
        win[j] = (n >= 0 && n < nt) ? xr[n] : TD(0);
    }
    __syncthreads();
    const int i = threadIdx.x;
    if (i < kConvOut && n0 + i < nt) {
        TD acc = TD(0);
        for (int k = 0; k < nh; ++k)
            acc += hs[k] * win[i + k];
        y[tr * nt + n0 + i] = acc;
    }
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
    off[(size_t)w * kWave + lane] = to_tab<TT>(q - b);
    if (lane == 0)
        base[w] = (int)b;
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
    out[(size_t)w * kWave + lane] =
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

KArgs make_args(const kx_plan *p)
{
    KArgs a;
    a.ns = p->ns;
    a.nr = p->nr;
    a.nt = p->nt;
    a.nx = p->nx;
    a.ny = p->ny;
    a.nz = p->nz;
    a.tz = p->tz;
    a.tx = p->tx;
    a.ty = p->ty;
    a.ntz = p->ntz;
    a.ntx = p->ntx;
    a.nty = p->nty;
    a.ntiles = (int)p->ntiles;
    a.rblock = p->rblock;
    a.off_s = p->off_s;
    a.base_s = p->base_s;
    a.off_r = p->off_r;
    a.base_r = p->base_r;
    a.wp_s = p->wp_s;
    a.wq_s = p->wq_s;
    a.wp_r = p->wp_r;
    a.wq_r = p->wq_r;
    a.izmin = p->izmin;
    return a;
}

template <typename TD, typename TT, bool DYN, bool NEG, int RB>
int launch_spread_rb(const KArgs &a, const void *m, void *y, hipStream_t st)
{
    const size_t lds = (size_t)RB * a.nt * sizeof(AccT);
    const int nrb = (a.nr + RB - 1) / RB;
    const int nblk = a.ns * nrb;
    hipLaunchKernelGGL((spread_kernel<TD, TT, DYN, NEG, RB>), dim3(nblk), dim3(kBlock), lds, st,
                       a, reinterpret_cast<const TD *>(m), reinterpret_cast<TD *>(y));
    KX_LAUNCHED("spread");
    return KX_OK;
}

template <typename TD, typename TT, bool DYN, bool NEG>
int launch_spread_t(const KArgs &a, const void *m, void *y, hipStream_t st)
{
    switch (a.rblock) {
    case 16:
        return launch_spread_rb<TD, TT, DYN, NEG, 16>(a, m, y, st);
    case 8:
        return launch_spread_rb<TD, TT, DYN, NEG, 8>(a, m, y, st);
    case 4:
        return launch_spread_rb<TD, TT, DYN, NEG, 4>(a, m, y, st);
    case 2:
        return launch_spread_rb<TD, TT, DYN, NEG, 2>(a, m, y, st);
    default:
        return launch_spread_rb<TD, TT, DYN, NEG, 1>(a, m, y, st);
    }
}

// receivers per workgroup: a power of two in [1, 16] whose traces fit in LDS
int pick_rblock(int requested, int nr, int per_trace_bytes)
{
    int rb = requested > 0 ? requested : 8;
    rb = std::min(rb, kLdsBudget / per_trace_bytes);
    int p2 = 1;
    while (p2 * 2 <= rb && p2 * 2 <= 16)
        p2 *= 2;
    while (p2 > 1 && p2 / 2 >= nr)
        p2 /= 2;  // do not leave a block mostly empty
    return p2;
}

template <typename TD, typename TT>
int launch_spread_d(const kx_plan *p, KArgs a, const void *m, void *y, hipStream_t st)
{
    const int per_trace = p->nt * (int)sizeof(AccT);
    if (per_trace > kLdsBudget)
        return kx_fail(KX_ELDS, "nt=%d does not fit one trace in LDS", p->nt);
    a.rblock = pick_rblock(p->rblock, p->nr, per_trace);
    if (p->dynamic)
        return p->neg_trav ? launch_spread_t<TD, TT, true, true>(a, m, y, st)
                           : launch_spread_t<TD, TT, true, false>(a, m, y, st);
    return p->neg_trav ? launch_spread_t<TD, TT, false, true>(a, m, y, st)
                       : launch_spread_t<TD, TT, false, false>(a, m, y, st);
}

template <typename TD, typename TT, bool DYN, bool NEG>
int launch_gather_t(const KArgs &a, const void *d, void *m, hipStream_t st)
{
    constexpr int SB = sizeof(TD) == 8 ? 8 : 16;
    const int nblk = (a.ntiles + (kBlock / kWave) - 1) / (kBlock / kWave);
    hipLaunchKernelGGL((gather_kernel<TD, TT, DYN, NEG, SB>), dim3(nblk), dim3(kBlock), 0, st, a,
                       reinterpret_cast<const TD *>(d), reinterpret_cast<TD *>(m));
    KX_LAUNCHED("gather");
    return KX_OK;
}

template <typename TD, typename TT>
int launch_gather_d(const kx_plan *p, const KArgs &a, const void *d, void *m, hipStream_t st)
{
    if (p->dynamic)
        return p->neg_trav ? launch_gather_t<TD, TT, true, true>(a, d, m, st)
                           : launch_gather_t<TD, TT, true, false>(a, d, m, st);
    return p->neg_trav ? launch_gather_t<TD, TT, false, true>(a, d, m, st)
                       : launch_gather_t<TD, TT, false, false>(a, d, m, st);
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

extern "C" int kx_conv(const void *x, void *y, int64_t ntr, int32_t nt, const void *h, int32_t nh,
                       int32_t dtype, void *stream)
{
    if (!x || !y || !h || x == y)
        return kx_fail(KX_EINVAL, "conv: bad buffers");
    if (ntr <= 0 || nt <= 0 || nh <= 0 || (nh % 2) == 0)
        return kx_fail(KX_ESHAPE, "conv: need ntr, nt > 0 and odd nh (nh=%d)", nh);
    const int nseg = (nt + kConvOut - 1) / kConvOut;
    if (ntr * nseg > INT32_MAX)
        return kx_fail(KX_ESHAPE, "conv: too many traces");
    const size_t lds = (size_t)(nh + kConvOut + nh - 1) * (dtype == KX_F64 ? 8 : 4);
    if (lds > (size_t)kLdsBudget)
        return kx_fail(KX_ELDS, "conv: filter of %d taps does not fit LDS", nh);
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const dim3 grid((unsigned)(ntr * nseg));
    if (dtype == KX_F32)
        hipLaunchKernelGGL(conv_kernel<float>, grid, dim3(kBlock), lds, st,
                           reinterpret_cast<const float *>(x), reinterpret_cast<float *>(y), nt,
                           reinterpret_cast<const float *>(h), nh, nseg);
    else if (dtype == KX_F64)
        hipLaunchKernelGGL(conv_kernel<double>, grid, dim3(kBlock), lds, st,
                           reinterpret_cast<const double *>(x), reinterpret_cast<double *>(y), nt,
                           reinterpret_cast<const double *>(h), nh, nseg);
    else
        return kx_fail(KX_ETYPE, "conv: dtype %d", dtype);
    KX_LAUNCHED("conv");
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
