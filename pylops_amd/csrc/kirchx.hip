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

template <typename T, int N> struct alignas(sizeof(T) * N) VecN {
    T v[N];
};

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
// LDS accumulation.  For float data every contribution c (rounded to float
// as the reference rounds it) is added as the integer rint(c * 2^e) to a
// 64-bit fixed-point sample, with one power-of-two scale 2^e per launch
// chosen from max|m| so that |c| 2^e <= 2^30: the rounding of a contribution
// is at most 2^-31 of the largest possible one, the sum itself is exact and
// does not depend on the order of the adds, and it cannot overflow (at most
// 2^32 contributions per sample).  On gfx950 LDS integer atomics sustain
// ~1.5x the rate of float64 ones and float32 LDS atomics ~1/6 of float64.
// A launch whose m holds inf or NaN accumulates in float64 instead, so that
// they propagate as in the reference; double data always does.
constexpr int kSpreadWaves = 8;
constexpr int kSpreadBlock = kSpreadWaves * kWave;
constexpr int kScaleBlocks = 256;

// aux layout: [0] int exponent e, [1] int 1 = fixed point / 0 = float64,
// [16, 16 + kScaleBlocks) float partial maxima
struct SpreadScale {
    int e;
    int fixed;
};

// Running tile coordinates of a wave: tiles are visited t, t + step, ...;
// advancing carries through z, then x, then y without a division.
struct TileCursor {
    int t, tzi, txi, tyi;
    __device__ __forceinline__ void start(const KArgs &a, int t0)
    {
        t = t0;
        tzi = t0 % a.ntz;
        const int rest = t0 / a.ntz;
        txi = rest % a.ntx;
        tyi = rest / a.ntx;
    }
    __device__ __forceinline__ void advance(const KArgs &a, int step)
    {
        t += step;
        tzi += step;
        while (tzi >= a.ntz) {
            tzi -= a.ntz;
            if (++txi == a.ntx) {
                txi = 0;
                ++tyi;
            }
        }
    }
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

// Lane roles.  Lane p holds point p of the tile and reads the source side
// of its point once per tile.  At step r (r < RB) it treats receiver j of
// the workgroup's block: j = r in natural order, so that the table loads of
// a step are coalesced rows, or j = (p + r) mod RB in rotated order, so that
// one atomic instruction covers all RB trace rows at 64 / RB points RB apart
// each.  Samples of neighbouring points of one trace often coincide, and the
// lanes of an LDS atomic that hit one address are serialised: the rotated
// order trades scattered table loads for fewer such collisions.
template <int RB, bool ROT> __device__ __forceinline__ int step_receiver(int lane, int r)
{
    return ROT ? (lane + r) & (RB - 1) : r;
}

// One lane's share of one tile: the image value and the table entries of
// its point, receivers in the lane's step order.
template <typename TD, typename TT, int RB> struct SpreadTile {
    TD mv;
    TT os;       // raw table elements: converted at use, so that a prefetch
    TT orr[RB];  // never waits for them
    TD ps, qs;
    TD pr[RB], qr[RB];
    int bs;
    int br[RB];
};

template <typename TD, typename TT, bool DYN, bool ROT, int RB>
__device__ __forceinline__ void spread_load(const KArgs &a, const TileCursor &c, int lane, int rel,
                                            int lz, int lx, int ly, int s, int r0, int nrl,
                                            const TD *__restrict__ m, int vz,
                                            SpreadTile<TD, TT, RB> &L)
{
    const int t = c.t;
    const int iz = c.tzi * a.tz + lz, ix = c.txi * a.tx + lx, iy = c.tyi * a.ty + ly;
    const bool valid = (iz < a.nz) & (ix < a.nx) & (iy < a.ny);
    const int orig = (c.tyi * a.ty * a.nx + c.txi * a.tx) * a.nz + c.tzi * a.tz;
    L.mv = valid ? m[orig + rel] : TD(0);
    const size_t ts = (size_t)t * a.ns + s;        // (tile, source) row
    const size_t tr = (size_t)t * a.nr + r0;       // (tile, first receiver) row
    L.os = reinterpret_cast<const TT *>(a.off_s)[ts * kWave + lane];
    L.bs = a.base_s[ts + vz];
    if constexpr (DYN) {
        L.ps = reinterpret_cast<const TD *>(a.wp_s)[ts * kWave + lane];
        L.qs = reinterpret_cast<const TD *>(a.wq_s)[ts * kWave + lane];
    }
#pragma unroll
    for (int r = 0; r < RB; ++r) {
        int j = step_receiver<RB, ROT>(lane, r);
        j = j < nrl ? j : 0;  // rows past the last receiver read row 0
        const size_t e = (tr + j) * kWave + lane;
        L.orr[r] = reinterpret_cast<const TT *>(a.off_r)[e];
        L.br[r] = a.base_r[tr + j];
        if constexpr (DYN) {
            L.pr[r] = reinterpret_cast<const TD *>(a.wp_r)[e];
            L.qr[r] = reinterpret_cast<const TD *>(a.wq_r)[e];
        }
    }
}

template <typename TD, typename TT, bool DYN, bool NEG, bool ROT, bool FIX, int RB>
__device__ __forceinline__ void spread_body(const KArgs &a, const TD *__restrict__ m,
                                            TD *__restrict__ y, int e)
{
    using TC = typename TimeType<TD>::type;
    extern __shared__ __align__(16) unsigned char kx_smem[];
    double *accd = reinterpret_cast<double *>(kx_smem);
    unsigned long long *acci = reinterpret_cast<unsigned long long *>(kx_smem);

    // one workgroup = one source s and receivers [r0, r0 + RB)
    const int L = xcd_remap(blockIdx.x, gridDim.x);
    const int s = L % a.ns;
    const int r0 = (L / a.ns) * RB;
    const int nrl = min(RB, a.nr - r0);
    const int ntr = nrl * a.nt;

    for (int i = threadIdx.x; i < ntr; i += blockDim.x)
        acci[i] = 0ull;  // 0.0 and integer 0 share the bit pattern
    __syncthreads();

    const int lane = threadIdx.x & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int nw = blockDim.x / kWave;
    const int vz = vgpr_zero();
    const int lz = lane % a.tz, lx = (lane / a.tz) % a.tx, ly = lane / (a.tz * a.tx);
    const int rel = (ly * a.nx + lx) * a.nz + lz;  // image offset from the tile origin
    int lrow[RB];                                  // LDS row of the receiver of step r
    int izm[RB];
#pragma unroll
    for (int r = 0; r < RB; ++r) {
        const int j = step_receiver<RB, ROT>(lane, r);
        lrow[r] = j < nrl ? j * a.nt : -1;
        if constexpr (DYN)
            izm[r] = a.izmin[s * a.nr + r0 + (j < nrl ? j : 0)];
    }
    const unsigned ntm1 = (unsigned)(a.nt - 1);

    SpreadTile<TD, TT, RB> cur, nxt;
    TileCursor c;
    c.start(a, wave);
    if (c.t < a.ntiles)
        spread_load<TD, TT, DYN, ROT, RB>(a, c, lane, rel, lz, lx, ly, s, r0, nrl, m, vz, cur);
    while (c.t < a.ntiles) {
        // prefetch the next tile while this one is accumulated
        TileCursor cn = c;
        cn.advance(a, nw);
        if (cn.t < a.ntiles)
            spread_load<TD, TT, DYN, ROT, RB>(a, cn, lane, rel, lz, lx, ly, s, r0, nrl, m, vz, nxt);
        if (__ballot(cur.mv != TD(0)) != 0) {
            const int iz = c.tzi * a.tz + lz;
            const TC os = (TC)ld_off(&cur.os);
            // the image value, scaled by 2^e for the fixed-point sum
            TD mv = cur.mv;
            if constexpr (FIX)
                mv = ldexpf(mv, e);
#pragma unroll
            for (int r = 0; r < RB; ++r) {
                const TC T = os + (TC)ld_off(&cur.orr[r]);
                const TC fl = floor(T);
                TC f = T - fl;
                int k = (int)fl + cur.bs + cur.br[r];
                if constexpr (NEG) {
                    // the reference truncates toward zero: a time in (-1, 0)
                    // lands on sample 0 with a negative weight
                    if (k == -1 && f > TC(0)) {
                        k = 0;
                        f -= TC(1);
                    }
                }
                TD w = mv;
                if constexpr (DYN) {
                    w = w * (cur.ps * cur.pr[r] + cur.qs * cur.qr[r]);
                    if (iz < izm[r])
                        w = TD(0);
                }
                if (lrow[r] >= 0 && (unsigned)k < ntm1 && w != TD(0)) {
                    const int it = lrow[r] + k;
                    const TD w0 = w * (TD)(TC(1) - f);
                    const TD w1 = w * (TD)f;
                    if constexpr (FIX) {
                        atomicAdd(&acci[it], (unsigned long long)(long long)__float2int_rn(w0));
                        atomicAdd(&acci[it + 1], (unsigned long long)(long long)__float2int_rn(w1));
                    } else {
                        atomicAdd(&accd[it], (double)w0);
                        atomicAdd(&accd[it + 1], (double)w1);
                    }
                }
            }
        }
        cur = nxt;
        c = cn;
    }
    __syncthreads();
    TD *yo = y + ((size_t)s * a.nr + r0) * a.nt;
    for (int i = threadIdx.x; i < ntr; i += blockDim.x) {
        if constexpr (FIX)
            yo[i] = (TD)ldexp((double)(long long)acci[i], -e);
        else
            yo[i] = (TD)accd[i];
    }
}

template <typename TD, typename TT, bool DYN, bool NEG, bool ROT, int RB>
__global__ __launch_bounds__(kSpreadBlock) void spread_kernel(KArgs a, const TD *__restrict__ m,
                                                              TD *__restrict__ y,
                                                              const SpreadScale *__restrict__ sc)
{
    if constexpr (sizeof(TD) == 4) {
        const int e = sc->e;
        if (sc->fixed)
            spread_body<TD, TT, DYN, NEG, ROT, true, RB>(a, m, y, e);
        else
            spread_body<TD, TT, DYN, NEG, ROT, false, RB>(a, m, y, 0);
    } else {
        spread_body<TD, TT, DYN, NEG, ROT, false, RB>(a, m, y, 0);
    }
}

// Per-launch scale of the fixed-point accumulation: partial max|m| per block
// (inf for a non-finite value), then one block turns the maximum into e.
__global__ __launch_bounds__(kBlock) void absmax_partial_kernel(const float *__restrict__ m,
                                                                int64_t n,
                                                                float *__restrict__ part)
{
    float mx = 0.f;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        float v = fabsf(m[i]);
        if (!(v <= 3.402823466e38f))
            v = INFINITY;  // NaN or inf
        mx = fmaxf(mx, v);
    }
    __shared__ float wmax[kBlock / kWave];
#pragma unroll
    for (int o = kWave / 2; o >= 1; o >>= 1)
        mx = fmaxf(mx, __shfl_xor(mx, o, kWave));
    if ((threadIdx.x & (kWave - 1)) == 0)
        wmax[threadIdx.x / kWave] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        float b = wmax[0];
        for (int i = 1; i < (int)(blockDim.x / kWave); ++i)
            b = fmaxf(b, wmax[i]);
        part[blockIdx.x] = b;
    }
}

__global__ __launch_bounds__(kWave) void scale_final_kernel(const float *__restrict__ part, int np,
                                                            double wamp,
                                                            SpreadScale *__restrict__ sc)
{
    float mx = 0.f;
    for (int i = threadIdx.x; i < np; i += kWave)
        mx = fmaxf(mx, part[i]);
#pragma unroll
    for (int o = kWave / 2; o >= 1; o >>= 1)
        mx = fmaxf(mx, __shfl_xor(mx, o, kWave));
    if (threadIdx.x == 0) {
        // a contribution is at most |m| * wamp * 2 (the truncation path of a
        // negative time weighs a sample by up to 2)
        const double bound = 2.0 * (double)mx * wamp;
        if (!(bound <= 1.7976931348623157e308)) {
            sc->e = 0;
            sc->fixed = 0;
        } else if (bound == 0.0) {
            sc->e = 0;
            sc->fixed = 1;
        } else {
            int ex;
            frexp(bound, &ex);  // bound < 2^ex
            sc->e = 30 - ex;
            sc->fixed = 1;
        }
    }
}

// ---------------------------------------------------------------------------
// adjoint: gather per image point from wave-staged trace windows
// ---------------------------------------------------------------------------
// A wave owns one tile and a share of the sources (P waves of a workgroup
// split a tile's sources; their partial sums are added in a fixed order at
// the end).  It sweeps the receivers r; for each of its JW sources the trace
// window of kWin samples centred on the tile's base sum (source plus
// receiver base: the tile's own offsets span a few tens of samples) is read
// with one coalesced load -- issued a round ahead -- into the wave's own LDS
// region, and each lane reads its two samples there.  A 64-lane gather from
// global memory costs the texture path ~10x the coalesced load and LDS reads
// that replace it.  Samples outside the window are read from global memory,
// so the window size affects speed only.  No barrier until the final sum.
constexpr int kGatherWaves = 4;
constexpr int kGatherBlock = kGatherWaves * kWave;
constexpr int kWin = kWave;  // samples per staged window, one per lane

template <typename TD, typename TT, bool DYN, bool NEG, int JW, int P>
__global__ __launch_bounds__(kGatherBlock) void gather_kernel(KArgs a, const TD *__restrict__ d,
                                                              TD *__restrict__ m)
{
    using TC = typename TimeType<TD>::type;
    constexpr int NT = kGatherWaves / P;  // tiles per workgroup
    constexpr int SB = JW * P;            // sources per block (all parts)
    __shared__ TD win[kGatherWaves][JW][kWin];

    const int lane = threadIdx.x & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int part = wave / NT;  // this wave sums sources s0 + part + P i
    const int L = xcd_remap(blockIdx.x, gridDim.x);
    const int tt = L * NT + wave % NT;
    const bool active = tt < a.ntiles;
    const int t = active ? tt : a.ntiles - 1;
    const TilePoint p = tile_point(a, t, lane);
    const TT *offs = reinterpret_cast<const TT *>(a.off_s);
    const TT *offr = reinterpret_cast<const TT *>(a.off_r);
    const TD *wps = reinterpret_cast<const TD *>(a.wp_s);
    const TD *wqs = reinterpret_cast<const TD *>(a.wq_s);
    const TD *wpr = reinterpret_cast<const TD *>(a.wp_r);
    const TD *wqr = reinterpret_cast<const TD *>(a.wq_r);
    const unsigned ntm1 = (unsigned)(a.nt - 1);
    const int vz = vgpr_zero();
    TD(&w)[JW][kWin] = win[wave];

    TD acc = TD(0);
    if (active) {
        for (int s0 = 0; s0 < a.ns; s0 += SB) {
            // this wave's sources j = s0 + part + P i
            TC os[JW];
            int bs[JW];
            TD ps[JW], qs[JW];
            // trace (j, receiver 0) as an element offset from the block's
            // first trace (fits 31 bits: checked at launch), held in vector
            // registers: JW 64-bit pointers would exhaust the scalar ones
            const TD *dblk = d + (size_t)s0 * a.nr * a.nt;
            int soff[JW], izoff[JW];
#pragma unroll
            for (int i = 0; i < JW; ++i) {
                const int j = s0 + part + P * i;
                if (j < a.ns) {
                    const size_t so = (size_t)t * a.ns + j;
                    os[i] = (TC)ld_off(offs + so * kWave + lane);
                    bs[i] = a.base_s[so + vz];
                    if constexpr (DYN) {
                        ps[i] = wps[so * kWave + lane];
                        qs[i] = wqs[so * kWave + lane];
                    }
                } else {  // past the last source: every sample before the trace
                    os[i] = TC(0);
                    bs[i] = -(1 << 30);
                    ps[i] = qs[i] = TD(0);
                }
                soff[i] = (min(j, a.ns - 1) - s0) * a.nr * a.nt + vz;
                izoff[i] = min(j, a.ns - 1) * a.nr + vz;
            }

            // the window of source i in round r starts at bs + br(r) - kWin / 2;
            // lane l stages its sample l
            TD sv[JW];
#define KX_STAGE(R, BR)                                                                        \
    _Pragma("unroll") for (int i = 0; i < JW; ++i)                                              \
    {                                                                                          \
        const int e = bs[i] + (BR) - kWin / 2 + lane;                                          \
        sv[i] = (unsigned)e < (unsigned)a.nt ? dblk[soff[i] + (R) * a.nt + e] : TD(0);          \
    }
            // receiver side of round r, and receiver base of round r + 1
            TT orr_c = offr[((size_t)t * a.nr) * kWave + lane];
            int br_c = a.base_r[(size_t)t * a.nr + vz];
            TD pr_c = TD(0), qr_c = TD(0);
            if constexpr (DYN) {
                pr_c = wpr[((size_t)t * a.nr) * kWave + lane];
                qr_c = wqr[((size_t)t * a.nr) * kWave + lane];
            }
            KX_STAGE(0, br_c)
            for (int r = 0; r < a.nr; ++r) {
                // this round's samples into the wave's LDS region
#pragma unroll
                for (int i = 0; i < JW; ++i)
                    w[i][lane] = sv[i];
                const TC orr = (TC)ld_off(&orr_c);
                const int br = br_c;
                const TD pr = pr_c, qr = qr_c;
                // next round's receiver side and samples, in flight meanwhile
                if (r + 1 < a.nr) {
                    const size_t ro = (size_t)t * a.nr + r + 1;
                    orr_c = offr[ro * kWave + lane];
                    br_c = a.base_r[ro + vz];
                    if constexpr (DYN) {
                        pr_c = wpr[ro * kWave + lane];
                        qr_c = wqr[ro * kWave + lane];
                    }
                    KX_STAGE(r + 1, br_c)
                }
#pragma unroll
                for (int i = 0; i < JW; ++i) {
                    const TC T = os[i] + orr;
                    const TC fl = floor(T);
                    TC f = T - fl;
                    int it = (int)fl + bs[i] + br;
                    if constexpr (NEG) {
                        // the reference truncates toward zero: a time in
                        // (-1, 0) reads sample 0 with a negative weight
                        if (it == -1 && f > TC(0)) {
                            it = 0;
                            f -= TC(1);
                        }
                    }
                    // The LDS read is unconditional (clamped index); a sample
                    // outside the window overwrites it from global memory (a
                    // branch between the two loads would merge them into one
                    // generic-address load).
                    const int kk = it - (bs[i] + br - kWin / 2);
                    const int kc = min(max(kk, 0), kWin - 2);
                    TD q0 = w[i][kc];
                    TD q1 = w[i][kc + 1];
                    const bool in_trace = (unsigned)it < ntm1;
                    if (in_trace && kk != kc) {
                        const TD *q = dblk + (soff[i] + r * a.nt + it);
                        q0 = q[0];
                        q1 = q[1];
                    }
                    TD v = q0 * (TD)(TC(1) - f) + q1 * (TD)f;
                    if constexpr (DYN) {
                        TD wgt = ps[i] * pr + qs[i] * qr;
                        if (p.iz < a.izmin[izoff[i] + r])
                            wgt = TD(0);
                        v *= wgt;
                    }
                    acc += in_trace ? v : TD(0);
                }
            }
#undef KX_STAGE
        }
    }
    if constexpr (P > 1) {
        // the partial sums of a tile's points, added in part order
        __syncthreads();
        TD *red = &win[0][0][0];
        red[threadIdx.x] = acc;
        __syncthreads();
        if (part == 0) {
            acc = TD(0);
#pragma unroll
            for (int q = 0; q < P; ++q)
                acc += red[(q * NT + wave) * kWave + lane];
        }
    }
    if (part == 0 && active && p.valid)
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

constexpr int kRblocks[] = {16, 8, 4, 2, 1};
constexpr int kLdsPerCu = 160 * 1024;
constexpr int kMaxWavesPerCu = 32;

size_t spread_lds_bytes(int rb, int nt) { return (size_t)rb * nt * 8; }

template <typename TD, typename TT, bool DYN, bool NEG, bool ROT, int RB>
int launch_spread_rb(const KArgs &a, const void *m, void *y, const SpreadScale *sc,
                     hipStream_t st)
{
    const size_t lds = spread_lds_bytes(RB, a.nt);
    const int nrb = (a.nr + RB - 1) / RB;
    const int nblk = a.ns * nrb;
    hipLaunchKernelGGL((spread_kernel<TD, TT, DYN, NEG, ROT, RB>), dim3(nblk), dim3(kSpreadBlock),
                       lds, st, a, reinterpret_cast<const TD *>(m), reinterpret_cast<TD *>(y), sc);
    KX_LAUNCHED("spread");
    return KX_OK;
}

template <typename TD, typename TT, bool DYN, bool NEG, bool ROT>
const void *spread_fn(int rb)
{
    switch (rb) {
    case 16:
        return reinterpret_cast<const void *>(&spread_kernel<TD, TT, DYN, NEG, ROT, 16>);
    case 8:
        return reinterpret_cast<const void *>(&spread_kernel<TD, TT, DYN, NEG, ROT, 8>);
    case 4:
        return reinterpret_cast<const void *>(&spread_kernel<TD, TT, DYN, NEG, ROT, 4>);
    case 2:
        return reinterpret_cast<const void *>(&spread_kernel<TD, TT, DYN, NEG, ROT, 2>);
    default:
        return reinterpret_cast<const void *>(&spread_kernel<TD, TT, DYN, NEG, false, 1>);
    }
}

// Workgroups of one spread kernel resident per CU (registers, LDS, waves);
// without a device, the LDS and wave limits alone.
int spread_per_cu(const void *fn, int rb, int nt)
{
    const size_t lds = spread_lds_bytes(rb, nt);
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
template <typename TD, typename TT, bool DYN, bool NEG, bool ROT>
int pick_rblock(int requested, int nr, int nt, int max_rb)
{
    auto fits = [&](int rb) {
        return rb <= max_rb && spread_lds_bytes(rb, nt) <= (size_t)kLdsBudget;
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
        const int per_cu = spread_per_cu(spread_fn<TD, TT, DYN, NEG, ROT>(rb), rb, nt);
        const double waves = std::min(24.0, (double)per_cu * kSpreadWaves);
        const double score = waves / 24.0 * per_cu * std::min(rb, nr);
        if (score > best_score * 1.0001) {
            best = rb;
            best_score = score;
        }
    }
    return best;
}

template <typename TD, typename TT, bool DYN, bool NEG, bool ROT>
int launch_spread_o(KArgs a, int requested, int max_rb, const void *m, void *y,
                    const SpreadScale *sc, hipStream_t st)
{
    a.rblock = pick_rblock<TD, TT, DYN, NEG, ROT>(requested, a.nr, a.nt, max_rb);
    switch (a.rblock) {
    case 16:
        return launch_spread_rb<TD, TT, DYN, NEG, ROT, 16>(a, m, y, sc, st);
    case 8:
        return launch_spread_rb<TD, TT, DYN, NEG, ROT, 8>(a, m, y, sc, st);
    case 4:
        return launch_spread_rb<TD, TT, DYN, NEG, ROT, 4>(a, m, y, sc, st);
    case 2:
        return launch_spread_rb<TD, TT, DYN, NEG, ROT, 2>(a, m, y, sc, st);
    default:
        return launch_spread_rb<TD, TT, DYN, NEG, false, 1>(a, m, y, sc, st);
    }
}

template <typename TD, typename TT, bool DYN, bool NEG>
int launch_spread_t(const KArgs &a, int order, int requested, int max_rb, const void *m, void *y,
                    const SpreadScale *sc, hipStream_t st)
{
    return order == KX_ORDER_ROTATED
               ? launch_spread_o<TD, TT, DYN, NEG, true>(a, requested, max_rb, m, y, sc, st)
               : launch_spread_o<TD, TT, DYN, NEG, false>(a, requested, max_rb, m, y, sc, st);
}


template <typename TD, typename TT>
int launch_spread_d(const kx_plan *p, KArgs a, const void *m, void *y, hipStream_t st)
{
    if (spread_lds_bytes(1, p->nt) > (size_t)kLdsBudget)
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
    // automatic lane order: rotated when a step loads one receiver table,
    // natural when it loads three (dynamic weights), where the scattered
    // loads cost more than the collisions they avoid
    int order = p->sorder;
    if (order == KX_ORDER_AUTO)
        order = p->dynamic ? KX_ORDER_NATURAL : KX_ORDER_ROTATED;
    if (p->dynamic)
        return p->neg_trav
                   ? launch_spread_t<TD, TT, true, true>(a, order, p->rblock, max_rb, m, y, sc, st)
                   : launch_spread_t<TD, TT, true, false>(a, order, p->rblock, max_rb, m, y, sc, st);
    return p->neg_trav
               ? launch_spread_t<TD, TT, false, true>(a, order, p->rblock, max_rb, m, y, sc, st)
               : launch_spread_t<TD, TT, false, false>(a, order, p->rblock, max_rb, m, y, sc, st);
}

template <typename TD, typename TT, bool DYN, bool NEG, int P>
int launch_gather_p(const KArgs &a, const void *d, void *m, hipStream_t st)
{
    // sources per wave and pass (measured): every pass re-reads the receiver
    // table -- three tables when dynamic -- and more sources hold more
    // registers, lowering occupancy
    constexpr int JW = DYN ? 8 : 4;
    constexpr int NT = kGatherWaves / P;
    if ((int64_t)JW * P * a.nr * a.nt >= ((int64_t)1 << 31))
        return kx_fail(KX_ESHAPE, "gather: %d sources x nr x nt exceeds 2^31 samples", JW * P);
    const int nblk = (a.ntiles + NT - 1) / NT;
    hipLaunchKernelGGL((gather_kernel<TD, TT, DYN, NEG, JW, P>), dim3(nblk), dim3(kGatherBlock),
                       0, st, a, reinterpret_cast<const TD *>(d), reinterpret_cast<TD *>(m));
    KX_LAUNCHED("gather");
    return KX_OK;
}

// Waves per tile: the fewest that give every CU kGatherFill waves.
constexpr int kGatherFill = 24;

template <typename TD, typename TT, bool DYN, bool NEG>
int launch_gather_t(const KArgs &a, int gsplit, const void *d, void *m, hipStream_t st)
{
    int P = 1;
    if (gsplit > 0) {
        while (P * 2 <= std::min(gsplit, kGatherWaves))
            P *= 2;
    } else {
        int cus = 256, dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            cus = 256;
        while (P < kGatherWaves && (int64_t)a.ntiles * P < (int64_t)kGatherFill * cus)
            P *= 2;
    }
    switch (P) {
    case 1:
        return launch_gather_p<TD, TT, DYN, NEG, 1>(a, d, m, st);
    case 2:
        return launch_gather_p<TD, TT, DYN, NEG, 2>(a, d, m, st);
    default:
        return launch_gather_p<TD, TT, DYN, NEG, 4>(a, d, m, st);
    }
}

template <typename TD, typename TT>
int launch_gather_d(const kx_plan *p, const KArgs &a, const void *d, void *m, hipStream_t st)
{
    const int g = p->gsplit;
    if (p->dynamic)
        return p->neg_trav ? launch_gather_t<TD, TT, true, true>(a, g, d, m, st)
                           : launch_gather_t<TD, TT, true, false>(a, g, d, m, st);
    return p->neg_trav ? launch_gather_t<TD, TT, false, true>(a, g, d, m, st)
                       : launch_gather_t<TD, TT, false, false>(a, g, d, m, st);
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
