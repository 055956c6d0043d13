// kx_device.h -- device code shared by the single-problem (kirchx.hip) and
// batched (kx_batch.hip) launches of the Kirchhoff kernels: data types, tile
// geometry and the bodies of the forward, adjoint and convolution kernels,
// each taking its work item explicitly.
#pragma once
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "kirchx.h"

// ---------------------------------------------------------------------------
// device helpers
// ---------------------------------------------------------------------------
namespace {

constexpr int kWave = 64;
constexpr int kBlock = 256;
constexpr int kLdsBudget = 64 * 1024;  // LDS of one workgroup on every CDNA device

// LDS one workgroup of the current device may use: the device's limit
// (160 KB on gfx950), queried once per device, or kLdsBudget without one.
inline size_t lds_budget()
{
    static int cached[64] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) {
        (void)hipGetLastError();
        return kLdsBudget;
    }
    if (cached[dev] == 0) {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) !=
                hipSuccess ||
            v < kLdsBudget) {
            (void)hipGetLastError();
            v = kLdsBudget;
        }
        cached[dev] = v;
    }
    return (size_t)cached[dev];
}

// Lets kernel fn take `lds` bytes of dynamic LDS (beyond 64 KB the
// attribute must be raised first).
inline void allow_lds(const void *fn, size_t lds)
{
    if (lds > (size_t)kLdsBudget &&
        hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) !=
            hipSuccess)
        (void)hipGetLastError();
}
constexpr int kReduceBlocks = 1024;

template <typename TD> struct TimeType { using type = float; };
template <> struct TimeType<double> { using type = double; };

template <typename T, int N> struct alignas(sizeof(T) * N) VecN {
    T v[N];
};

// Element type of the source-side offsets: float16 plans keep the source
// tables (a few rows per tile) in float32, so that a traveltime sum carries
// one float16 rounding, the receiver's.
template <typename TT> struct SrcTab { using type = TT; };
template <> struct SrcTab<__half> { using type = float; };
// Receiver-side dynamic weights: float16 beside float16 offsets (scaled to
// at most 1 by the caller), the data type otherwise; wv reads one as a
// float of the arithmetic.
template <typename TD, typename TT> struct RecW { using type = TD; };
template <> struct RecW<float, __half> { using type = __half; };
__device__ __forceinline__ float wv(float x) { return x; }
__device__ __forceinline__ float wv(__half x) { return __half2float(x); }
__device__ __forceinline__ double wv(double x) { return x; }

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
    // adjoint tile groups: gx x gy tiles, ngx x ngy groups per z row,
    // ngroups in all (set_tile_groups; 1 x 1 by default)
    int gx, gy, ngx, ngy, ngroups;
    const void *off_s;
    const int *base_s;
    const void *off_r;
    const int *base_r;
    const void *wp_s, *wq_s, *wp_r, *wq_r;
    const int *izmin;
    // forward output: the trace of (source s, block member j of a block
    // starting at r0) is trace s * trs + (r0 + j) * trb; (nr, 1) unless the
    // forward blocks sources (the side tables exchanged, then (1, ns))
    int trs, trb;
    // forward traces held in LDS: samples [t0, t0 + tlen) of every trace,
    // the whole trace (0, nt) unless a trace does not fit in LDS
    int t0, tlen;
};

inline KArgs make_args(const kx_plan *p)
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
    a.gx = a.gy = 1;
    a.ngx = p->ntx;
    a.ngy = p->nty;
    a.ngroups = (int)p->ntiles;
    a.off_s = p->off_s;
    a.base_s = p->base_s;
    a.off_r = p->off_r;
    a.base_r = p->base_r;
    a.wp_s = p->wp_s;
    a.wq_s = p->wq_s;
    a.wp_r = p->wp_r;
    a.wq_r = p->wq_r;
    a.izmin = p->izmin;
    a.trs = p->nr;
    a.trb = 1;
    a.t0 = 0;
    a.tlen = p->nt;
    return a;
}

// The forward's arguments with the two sides exchanged: its workgroups then
// hold one receiver and a block of sources, and write each trace to its
// transposed position (kx_plan.sside).
inline void exchange_sides(KArgs &a)
{
    const int ns = a.ns, nr = a.nr;
    a.ns = nr;
    a.nr = ns;
    const void *o = a.off_s;
    a.off_s = a.off_r;
    a.off_r = o;
    const int *b = a.base_s;
    a.base_s = a.base_r;
    a.base_r = b;
    const void *w = a.wp_s;
    a.wp_s = a.wp_r;
    a.wp_r = w;
    w = a.wq_s;
    a.wq_s = a.wq_r;
    a.wq_r = w;
    a.trs = 1;
    a.trb = nr;
}

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
// as the reference rounds it) is added as the integer floor(c * 2^e + 1/2)
// to a 64-bit fixed-point sample, with one power-of-two scale 2^e per launch
// chosen from max|m| so that |c| 2^e <= 2^30: the rounding of a contribution
// is at most 2^-31 of the largest possible one, the sum itself is exact and
// does not depend on the order of the adds, and it cannot overflow (at most
// 2^32 contributions per sample).  On gfx950 LDS integer atomics sustain
// ~1.5x the rate of float64 ones and float32 LDS atomics ~1/6 of float64.
// A launch whose m holds inf or NaN accumulates in float64 instead, so that
// they propagate as in the reference; double data always does.  So does a
// launch whose typical contribution lies far below the largest one (see
// scale_from_stats), where the fixed-point rounding would no longer be
// small against it.
constexpr int kSpreadWaves = 8;
constexpr int kSpreadBlock = kSpreadWaves * kWave;
constexpr int kScaleBlocks = 128;

// aux layout: [0] int exponent e, [1] int 1 = fixed point / 0 = float64,
// floats [16, 16 + 3 kScaleBlocks): partial max|m|, sum |m| and count of
// nonzero m per block
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
// the workgroup's block:
//   natural   j = r: the table loads of a step are coalesced rows;
//   rotated   j = (p + r) mod RB: one atomic instruction covers all RB trace
//             rows at 64 / RB points RB apart each; a step's table load
//             reads a scattered element of every row.
// Samples of neighbouring points of one trace often coincide, and an LDS
// atomic instruction whose lanes crowd a few banks of one row is
// serialised: the rotated order spreads each instruction over RB rows.
constexpr int kOrdNatural = 0;
constexpr int kOrdRotated = 1;

// The receiver lane p treats at step r.
template <int RB, int ORD> __device__ __forceinline__ int step_receiver(int lane, int r)
{
    if constexpr (ORD == kOrdNatural)
        return r;
    else
        return (lane + r) & (RB - 1);
}

// Receiver bases held per lane: the base of receiver (lane mod RB) of the
// block, one load for all RB receivers, handed to the lanes that need it by
// a lane read (natural order) or a quad permutation (rotated order,
// RB <= 4); rotated blocks of 8 and 16 receivers load one base per
// step instead.
template <int ORD, int RB> struct BaseLoads {
    static constexpr int n = (ORD == kOrdRotated && RB >= 8) ? RB : 1;
};

// One lane's share of one tile: the image value and the table entries of
// its point, receivers in the lane's step order.
template <typename TD, typename TT, int ORD, int RB> struct SpreadTile {
    TD mv;
    typename SrcTab<TT>::type os;  // raw table elements: converted at use, so that a prefetch
    TT orr[RB];  // never waits for them
    TD ps, qs;
    typename RecW<TD, TT>::type pr[RB], qr[RB];
    int bs;
    int br[BaseLoads<ORD, RB>::n];
};

// The receiver base of step r for this lane.
template <int ORD, int RB>
__device__ __forceinline__ int step_base(const int (&br)[BaseLoads<ORD, RB>::n], int r)
{
    if constexpr (BaseLoads<ORD, RB>::n > 1) {
        return br[r];
    } else if constexpr (ORD == kOrdNatural || RB == 1) {
        // every lane treats receiver r: the base held by lane r
        return __builtin_amdgcn_readlane(br[0], r);
    } else if constexpr (RB == 2) {
        // lane l treats receiver (l + r) mod 2: swap lane pairs on odd steps
        // (a quad permutation reads no lane outside the wave: mov_dpp needs
        // no old value)
        return (r & 1) ? __builtin_amdgcn_mov_dpp(br[0], 0xB1, 0xf, 0xf, true) : br[0];
    } else {
        // RB == 4: lane l treats receiver (l + r) mod 4, held by lane
        // (l & ~3) | ((l + r) & 3): quad_perm [r, r+1, r+2, r+3] mod 4
        switch (r & 3) {
        case 1:
            return __builtin_amdgcn_mov_dpp(br[0], 0x39, 0xf, 0xf, true);
        case 2:
            return __builtin_amdgcn_mov_dpp(br[0], 0x4E, 0xf, 0xf, true);
        case 3:
            return __builtin_amdgcn_mov_dpp(br[0], 0x93, 0xf, 0xf, true);
        default:
            return br[0];
        }
    }
}

template <typename TD, typename TT, bool DYN, int ORD, int RB>
__device__ __forceinline__ void spread_load(const KArgs &a, const TileCursor &c, int lane, int rel,
                                            int lz, int lx, int ly, int s, int r0, int nrl,
                                            const TD *__restrict__ m, int vz,
                                            SpreadTile<TD, TT, ORD, RB> &L)
{
    const int t = c.t;
    const int iz = c.tzi * a.tz + lz, ix = c.txi * a.tx + lx, iy = c.tyi * a.ty + ly;
    const bool valid = (iz < a.nz) & (ix < a.nx) & (iy < a.ny);
    const int orig = (c.tyi * a.ty * a.nx + c.txi * a.tx) * a.nz + c.tzi * a.tz;
    // Row pointers are wave-uniform and the lane offsets 32-bit unsigned, so
    // every load takes the scalar-base form (one VALU op of addressing).
    const TD mval = (m + orig)[valid ? (unsigned)rel : 0u];
    L.mv = valid ? mval : TD(0);
    const size_t ts = (size_t)t * a.ns + s;        // (tile, source) row
    const size_t tr = (size_t)t * a.nr + r0;       // (tile, first receiver) row
    L.os = (reinterpret_cast<const typename SrcTab<TT>::type *>(a.off_s) + ts * kWave)[(unsigned)lane];
    L.bs = a.base_s[ts + vz];
    if constexpr (DYN) {
        L.ps = (reinterpret_cast<const TD *>(a.wp_s) + ts * kWave)[(unsigned)lane];
        L.qs = (reinterpret_cast<const TD *>(a.wq_s) + ts * kWave)[(unsigned)lane];
    }
    const TT *orow = reinterpret_cast<const TT *>(a.off_r) + tr * kWave;
    const int *brow = a.base_r + tr;
    if constexpr (BaseLoads<ORD, RB>::n == 1) {
        const int jl = lane & (RB - 1);
        L.br[0] = brow[(unsigned)((jl < nrl ? jl : 0) + vz)];
    }
#pragma unroll
    for (int r = 0; r < RB; ++r) {
        int j = step_receiver<RB, ORD>(lane, r);
        j = j < nrl ? j : 0;  // rows past the last receiver read row 0
        const unsigned e = (unsigned)(j * kWave + lane);
        L.orr[r] = orow[e];
        if constexpr (BaseLoads<ORD, RB>::n > 1)
            L.br[r] = brow[(unsigned)(j + vz)];
        if constexpr (DYN) {
            using RW = typename RecW<TD, TT>::type;
            L.pr[r] = (reinterpret_cast<const RW *>(a.wp_r) + tr * kWave)[e];
            L.qr[r] = (reinterpret_cast<const RW *>(a.wq_r) + tr * kWave)[e];
        }
    }
}

// floor(x + 1/2) as an integer, one conversion instruction (|x| < 2^31)
__device__ __forceinline__ int round_half_up(float x)
{
    int i;
    asm("v_cvt_rpi_i32_f32 %0, %1" : "=v"(i) : "v"(x));
    return i;
}
__device__ __forceinline__ int round_half_up(double x) { return (int)floor(x + 0.5); }

// min(max(x, 0), hi) in one instruction (hi wave-uniform)
__device__ __forceinline__ int clamp0(int x, int hi)
{
    int r;
    asm("v_med3_i32 %0, %1, 0, %2" : "=v"(r) : "v"(x), "s"(hi));
    return r;
}

// (int)floor(x) in one conversion instruction (|x| < 2^31)
__device__ __forceinline__ int floor_int(float x)
{
    int i;
    asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(i) : "v"(x));
    return i;
}

constexpr int kConvThreads = 128;
constexpr int kConvPer = 4;
constexpr int kConvOut = kConvThreads * kConvPer;
// filter length padded to a whole number of 4-element groups
__host__ __device__ __forceinline__ int conv_taps_padded(int nh) { return (nh + 3) & ~3; }

// LDS accumulation of the forward's traces (ACC): kAccFix, 64-bit fixed
// point with one 64-bit integer atomic per sample (float32 data), or
// kAccF64, float64 atomics (float64 data; float32 launches whose model
// holds inf or NaN).
constexpr int kAccF64 = 0;
constexpr int kAccFix = 1;

// One contribution of weight w at time T samples past the integer base:
// w (1 - f) to sample k and w f to sample k + 1 of the LDS trace at lrow,
// k = floor(T) + base, f = T - floor(T).  Lanes with lrow < 0 (no receiver),
// with `on` false (aperture), or whose k + 1 falls past the trace add
// nothing.  The index and fraction of a float time come from one conversion
// and one fract instruction: fract(T) is T - floor(T) rounded as the
// subtraction rounds it, except that a result that would round to 1 (T in
// (-2^-24, 0)) stays just below 1.
// SEG: the LDS traces hold samples [t0, t0 + tlen) only; each of the two
// samples is added where it lies in them.
template <typename TD, typename TC, bool NEG, int ACC, bool SEG = false>
__device__ __forceinline__ void spread_sample(TC T, int base, TD w, bool on, int lrow,
                                              unsigned ntm1, void *lds, int t0 = 0,
                                              unsigned tlen = 0)
{
    TC f;
    int k;
    if constexpr (!NEG && sizeof(TC) == 4) {
        f = __builtin_amdgcn_fractf(T);
        k = floor_int(T) + base;
    } else {
        const TC fl = floor(T);
        f = T - fl;
        k = (int)fl + base;
        if constexpr (NEG) {
            // the reference truncates toward zero: a time in (-1, 0)
            // lands on sample 0 with a negative weight
            if (k == -1 && f > TC(0)) {
                k = 0;
                f -= TC(1);
            }
        }
    }
    if (on && lrow >= 0 && (unsigned)k < ntm1 && w != TD(0)) {
        int it = lrow + k;
        bool in0 = true, in1 = true;
        if constexpr (SEG) {
            const int kl = k - t0;
            it = lrow + kl;
            in0 = (unsigned)kl < tlen;
            in1 = (unsigned)(kl + 1) < tlen;
        }
        const TD w1 = w * (TD)f;
        if constexpr (ACC == kAccFix) {
            // w (1 - f) in one rounding; fixed point by floor(x + 1/2) in
            // one conversion
            unsigned long long *acci = static_cast<unsigned long long *>(lds);
            const TD w0 = fma(-w, (TD)f, w);
            if (in0)
                atomicAdd(&acci[it], (unsigned long long)(long long)round_half_up(w0));
            if (in1)
                atomicAdd(&acci[it + 1], (unsigned long long)(long long)round_half_up(w1));
        } else {
            double *accd = static_cast<double *>(lds);
            const TD w0 = w * (TD)(TC(1) - f);
            if (in0)
                atomicAdd(&accd[it], (double)w0);
            if (in1)
                atomicAdd(&accd[it + 1], (double)w1);
        }
    }
}

// Accumulates one tile (64 points of a wave, RB receivers) into the LDS
// traces.  Dynamic weights are (m ps) pr + (m qs) qr with the source-side
// products formed once per tile; with the sides exchanged (XS: the block
// members are sources) the same products are formed per step, so that the
// weights round as they do with blocks of receivers.
template <typename TD, typename TT, bool DYN, bool NEG, int ORD, int ACC, int RB,
          bool XS = false, bool SEG = false>
__device__ __forceinline__ void spread_accumulate(const SpreadTile<TD, TT, ORD, RB> &cur, int iz,
                                                  int e,
                                                  const int (&lrow)[RB], const int (&izm)[RB],
                                                  unsigned ntm1, void *lds, int t0 = 0,
                                                  unsigned tlen = 0)
{
    using TC = typename TimeType<TD>::type;
    if (__ballot(cur.mv != TD(0)) == 0)
        return;
    const TC os = (TC)ld_off(&cur.os);
    // the image value, scaled by 2^e for the fixed-point sum
    TD mv = cur.mv;
    if constexpr (ACC == kAccFix)
        mv = ldexpf(mv, e);
    TD mps = mv, mqs = TD(0);
    if constexpr (DYN && !XS) {
        mps = mv * cur.ps;
        mqs = mv * cur.qs;
    }
#pragma unroll
    for (int r = 0; r < RB; ++r) {
        const TC T = os + (TC)ld_off(&cur.orr[r]);
        TD w = mv;
        bool on = true;
        if constexpr (DYN) {
            if constexpr (XS)  // (ps, qs) here are the receiver's, pr, qr the source's
                w = fma(mv * (TD)wv(cur.qr[r]), cur.qs, (mv * (TD)wv(cur.pr[r])) * cur.ps);
            else
                w = fma(mqs, (TD)wv(cur.qr[r]), mps * (TD)wv(cur.pr[r]));
            on = iz >= izm[r];
        }
        spread_sample<TD, TC, NEG, ACC, SEG>(T, cur.bs + step_base<ORD, RB>(cur.br, r), w, on,
                                             lrow[r], ntm1, lds, t0, tlen);
    }
}

// LDS row stride of the forward's traces, in samples: nt rounded up to 2
// mod 16.  A 64-bit sample spans two of the 32 banks that serve one 16-lane
// quarter of an atomic instruction, so the rows of a workgroup's receivers
// start 2 bank pairs apart; in rotated order the lanes of a quarter on
// neighbouring receivers, whose samples lie close together, then fall
// between each other's banks (the 3D spread's own addresses replayed: 5.07
// atomic lanes per clock and CU at a stride of 1010, 4.18 at 1001, 4.0-4.8
// at the other residues).  Monotone in nt.
__host__ __device__ __forceinline__ int trace_stride(int nt) { return nt + ((2 - nt) & 15); }

// Zeroes the ntr samples of a workgroup's LDS traces (0.0 and integer 0
// share the bit pattern).
__device__ __forceinline__ void traces_clear(void *lds, int ntr)
{
    unsigned long long *p = static_cast<unsigned long long *>(lds);
    for (int i = threadIdx.x; i < ntr; i += blockDim.x)
        p[i] = 0ull;
    __syncthreads();
}

// Writes the workgroup's nrl LDS traces (rows lts samples apart, samples
// [t0, t0 + tlen) of traces of nt samples) to y: row j to trace first +
// j * trb; fixed point scaled back by 2^-e, or float64 sums rounded to TD.
template <typename TD, int ACC>
__device__ __forceinline__ void traces_store(TD *__restrict__ y, size_t first, int trb, int nt,
                                             int lts, int nrl, int e, const void *lds,
                                             int t0 = 0, int tlen = -1)
{
    __syncthreads();
    if (tlen < 0)
        tlen = nt;
    TD *yo = y + first * nt + t0;
    const int ntr = nrl * tlen;
    for (int i = threadIdx.x; i < ntr; i += blockDim.x) {
        const int j = i / tlen, n = i - j * tlen;
        TD v;
        if constexpr (ACC == kAccFix)
            v = (TD)ldexp(
                (double)(long long)static_cast<const unsigned long long *>(lds)[j * lts + n], -e);
        else
            v = (TD) static_cast<const double *>(lds)[j * lts + n];
        yo[(size_t)j * trb * nt + n] = v;
    }
    __syncthreads();  // the traces are read out before a next item clears them
}

// The forward's wavelet convolution on the workgroup's LDS traces (the
// reference convolves after the spread, K:1048-1094 / C1D:99-115): the
// 64-bit sums become float32 rows padded with zeros in the same LDS --
// converted in passes of kFusePass samples per thread, each pass read into
// registers before any of it is written, so that a pass writes only over
// samples already read (fused_conv_ok checks this for the launch's shape)
// -- and each row is convolved as kx_conv convolves a trace, with the
// operations in the same order, so the data are bitwise those of the
// spread followed by kx_conv.  Row j holds x[p - L] at position p (L =
// (nh - 1) / 2, zeros outside the trace) in S = fused_row(nt, nh) floats;
// the flipped filter, zero padded to whole 4-tap groups, follows the rows.
constexpr int kFusePass = 8;

__host__ __device__ __forceinline__ int fused_row(int nt, int nh)
{
    return (nt + conv_taps_padded(nh) + 3 + 3) & ~3;
}

// Whether a launch of rb-trace workgroups with lds bytes of LDS can convolve
// in place: the float rows and the filter fit, and no conversion pass
// writes past the first sample of the next pass (the samples are 8-byte
// sums, the rows 4-byte floats).
inline bool fused_conv_ok(int rb, int nt, int nh, size_t lds, int block)
{
    const int64_t S = fused_row(nt, nh), L = (nh - 1) / 2, lts = trace_stride(nt);
    if (((int64_t)rb * S + conv_taps_padded(nh)) * 4 > (int64_t)lds)
        return false;
    const int64_t ntr = (int64_t)rb * nt, B = (int64_t)kFusePass * block;
    for (int64_t q = 0; (q + 1) * B < ntr; ++q) {
        const int64_t i = (q + 1) * B - 1;  // last sample of pass q
        const int64_t j = i / nt, n = i - j * nt;
        const int64_t i1 = i + 1;  // first sample of pass q + 1, at 8 (j1 lts + n1) bytes
        const int64_t j1 = i1 / nt, n1 = i1 - j1 * nt;
        if ((j * S + L + n + 1) * 4 > (j1 * lts + n1) * 8)
            return false;
    }
    return true;
}

template <typename TD, int ACC>
__device__ __forceinline__ void traces_conv_store(TD *__restrict__ y, size_t first, int trb,
                                                  int nt, int nrl, int e, void *lds,
                                                  const TD *__restrict__ h, int nh)
{
    if constexpr (sizeof(TD) == 4) {
        const int ntr = nrl * nt, lts = trace_stride(nt);
        const int S = fused_row(nt, nh), L = (nh - 1) / 2, nhp = conv_taps_padded(nh);
        float *win = static_cast<float *>(lds);
        const unsigned long long *acci = static_cast<const unsigned long long *>(lds);
        const double *accd = static_cast<const double *>(lds);
        __syncthreads();  // every contribution added
        for (int p0 = 0; p0 < ntr; p0 += kFusePass * (int)blockDim.x) {
            float v[kFusePass];
#pragma unroll
            for (int q = 0; q < kFusePass; ++q) {
                const int i = p0 + q * (int)blockDim.x + (int)threadIdx.x;
                if (i < ntr) {
                    const int j = i / nt, li = j * lts + (i - j * nt);
                    if constexpr (ACC == kAccF64)
                        v[q] = (float)accd[li];
                    else
                        v[q] = (float)ldexp((double)(long long)acci[li], -e);
                }
            }
            __syncthreads();  // the pass is read before it is overwritten
#pragma unroll
            for (int q = 0; q < kFusePass; ++q) {
                const int i = p0 + q * (int)blockDim.x + (int)threadIdx.x;
                if (i < ntr) {
                    const int j = i / nt;
                    win[j * S + L + (i - j * nt)] = v[q];
                }
            }
            __syncthreads();
        }
        // zero margins of the rows, then the flipped filter
        const int F = nrl * S;
        for (int p = threadIdx.x; p < F; p += blockDim.x) {
            const int n = p % S - L;
            if (n < 0 || n >= nt)
                win[p] = 0.f;
        }
        for (int k = threadIdx.x; k < nhp; k += blockDim.x)
            win[F + k] = k < nh ? h[nh - 1 - k] : 0.f;
        __syncthreads();
        const int G = (nt + 3) / 4;  // groups of 4 outputs per row
        const float4 *h4 = reinterpret_cast<const float4 *>(win + F);
        for (int g = threadIdx.x; g < nrl * G; g += blockDim.x) {
            const int j = g / G, n0 = (g - j * G) * 4;
            const float4 *w4 = reinterpret_cast<const float4 *>(win + j * S + n0);
            float acc[kConvPer] = {};
            float4 cur = w4[0];
            for (int gq = 0; gq < nhp / 4; ++gq) {
                const float4 nxt = w4[gq + 1];
                const float4 hq = h4[gq];
                const float w[8] = {cur.x, cur.y, cur.z, cur.w, nxt.x, nxt.y, nxt.z, nxt.w};
                const float hh[4] = {hq.x, hq.y, hq.z, hq.w};
#pragma unroll
                for (int q = 0; q < 4; ++q)
#pragma unroll
                    for (int r = 0; r < kConvPer; ++r)
                        acc[r] += hh[q] * w[q + r];
                cur = nxt;
            }
            TD *yo = y + (first + (size_t)j * trb) * nt;
#pragma unroll
            for (int r = 0; r < kConvPer; ++r)
                if (n0 + r < nt)
                    yo[n0 + r] = acc[r];
        }
        __syncthreads();  // the rows are read before a next item clears them
    } else {
        traces_store<TD, ACC>(y, first, trb, nt, trace_stride(nt), nrl, e, lds);
    }
}

// SEG: the LDS traces hold samples [a.t0, a.t0 + a.tlen) of each trace (a
// trace longer than LDS, spread one segment per workgroup; no wavelet).
template <typename TD, typename TT, bool DYN, bool NEG, int ORD, int ACC, int RB,
          bool XS = false, bool SEG = false>
__device__ __forceinline__ void spread_body(const KArgs &a, const TD *__restrict__ m,
                                            TD *__restrict__ y, int e, int s, int r0,
                                            const TD *__restrict__ h = nullptr, int nh = 0)
{
    extern __shared__ __align__(16) unsigned char kx_smem[];
    void *lds = kx_smem;

    // one workgroup = one source s and receivers [r0, r0 + RB)
    const int nrl = min(RB, a.nr - r0);
    const int lts = trace_stride(SEG ? a.tlen : a.nt);

    traces_clear(lds, nrl * lts);

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
        const int j = step_receiver<RB, ORD>(lane, r);
        lrow[r] = j < nrl ? j * lts : -1;
        if constexpr (DYN)  // the gate of a pair is indexed like its trace
            izm[r] = a.izmin[(size_t)s * a.trs + (size_t)(r0 + (j < nrl ? j : 0)) * a.trb];
    }
    const unsigned ntm1 = (unsigned)(a.nt - 1);

    // The next tile is loaded while this one is accumulated.  Float32
    // operators take two tile buffers in turn, the loop unrolled twice, no
    // copies; float64 operators (more registers per tile) load the next
    // tile into nxt and copy it over cur at the end of the loop.  Measured
    // against those copies, and against loads that only ever fill nxt with
    // a copy to cur at the top of the loop: 3D 128x128x120 kinematic 2%
    // (float32 tables) and 1.5% (float16 tables) faster, 2D 601x300 dynamic
    // 1% and 2% faster.  The loads-ahead form also lets the waves of a
    // workgroup drift apart, so that the workgroups of one receiver block
    // (one per source) stop sharing its rows in L2 (102 against 57 GB from
    // HBM per launch in 3D).
    // Every iteration issues the same loads: past the last tile the
    // current one is loaded again, unused, and a point past the grid's
    // edge reads the tile's first image value.  With a load under a branch
    // the compiler cannot count the loads in flight on every path and
    // waits for the prefetch just issued before each receiver's atomics
    // (4% slower in 3D, 3% in 2D).
    constexpr bool kPair = sizeof(TD) == 4;
    TileCursor c;
    c.start(a, wave);
    if constexpr (kPair) {
        SpreadTile<TD, TT, ORD, RB> ta, tb;
        if (c.t < a.ntiles)
            spread_load<TD, TT, DYN, ORD, RB>(a, c, lane, rel, lz, lx, ly, s, r0, nrl, m, vz, ta);
        while (c.t < a.ntiles) {
            TileCursor cn = c;
            cn.advance(a, nw);
            spread_load<TD, TT, DYN, ORD, RB>(a, cn.t < a.ntiles ? cn : c, lane, rel, lz, lx, ly,
                                              s, r0, nrl, m, vz, tb);
            spread_accumulate<TD, TT, DYN, NEG, ORD, ACC, RB, XS, SEG>(
                ta, c.tzi * a.tz + lz, e, lrow, izm, ntm1, lds, a.t0, (unsigned)a.tlen);
            c = cn;
            if (c.t >= a.ntiles)
                break;
            cn.advance(a, nw);
            spread_load<TD, TT, DYN, ORD, RB>(a, cn.t < a.ntiles ? cn : c, lane, rel, lz, lx, ly,
                                              s, r0, nrl, m, vz, ta);
            spread_accumulate<TD, TT, DYN, NEG, ORD, ACC, RB, XS, SEG>(
                tb, c.tzi * a.tz + lz, e, lrow, izm, ntm1, lds, a.t0, (unsigned)a.tlen);
            c = cn;
        }
    } else {
        SpreadTile<TD, TT, ORD, RB> cur, nxt;
        if (c.t < a.ntiles)
            spread_load<TD, TT, DYN, ORD, RB>(a, c, lane, rel, lz, lx, ly, s, r0, nrl, m, vz, cur);
        while (c.t < a.ntiles) {
            TileCursor cn = c;
            cn.advance(a, nw);
            spread_load<TD, TT, DYN, ORD, RB>(a, cn.t < a.ntiles ? cn : c, lane, rel, lz, lx, ly,
                                              s, r0, nrl, m, vz, nxt);
            spread_accumulate<TD, TT, DYN, NEG, ORD, ACC, RB, XS, SEG>(
                cur, c.tzi * a.tz + lz, e, lrow, izm, ntm1, lds, a.t0, (unsigned)a.tlen);
            cur = nxt;
            c = cn;
        }
    }
    // with a wavelet: convolved in LDS, y receives the data (see
    // traces_conv_store); without: y receives the traces
    if (!SEG && h != nullptr)
        traces_conv_store<TD, ACC>(y, (size_t)s * a.trs + (size_t)r0 * a.trb, a.trb, a.nt, nrl,
                                   e, lds, h, nh);
    else
        traces_store<TD, ACC>(y, (size_t)s * a.trs + (size_t)r0 * a.trb, a.trb, a.nt, lts, nrl,
                              e, lds, SEG ? a.t0 : 0, SEG ? a.tlen : a.nt);
}

// One (source, receiver block) item with the launch's accumulation scale:
// float32 data in fixed point, or in float64 when the launch's scale says
// so (a non-finite model); float64 data always in float64.
template <typename TD, typename TT, bool DYN, bool NEG, int ORD, int RB, bool XS = false,
          bool SEG = false>
__device__ __forceinline__ void spread_item(const KArgs &a, const TD *__restrict__ m,
                                            TD *__restrict__ y, const SpreadScale &sc, int s,
                                            int r0, const TD *__restrict__ h = nullptr,
                                            int nh = 0)
{
    if constexpr (sizeof(TD) == 4) {
        if (sc.fixed)
            spread_body<TD, TT, DYN, NEG, ORD, kAccFix, RB, XS, SEG>(a, m, y, sc.e, s, r0, h,
                                                                     nh);
        else
            spread_body<TD, TT, DYN, NEG, ORD, kAccF64, RB, XS, SEG>(a, m, y, 0, s, r0, h, nh);
    } else {
        spread_body<TD, TT, DYN, NEG, ORD, kAccF64, RB, XS, SEG>(a, m, y, 0, s, r0);
    }
}

// ---------------------------------------------------------------------------
// forward for K models at once (float32 operators)
// ---------------------------------------------------------------------------
// One workgroup = one (source, receiver) trace of K models: the tables of a
// tile are loaded once and the index, fraction and aperture gate of every
// point computed once for all K models; model j accumulates into LDS row j
// with its own fixed-point scale.  Each model's contributions are the
// integers its single-model launch adds, so every model's traces are bitwise
// those of kx_spread (the sums are exact).
template <typename TT, int K> struct ModelsTile {
    float mv[K];
    typename SrcTab<TT>::type os;
    TT orr;
    float ps, qs;
    typename RecW<float, TT>::type pr, qr;
    int bs, br;
};

template <typename TT, bool DYN, int K>
__device__ __forceinline__ void models_load(const KArgs &a, const TileCursor &c, int lane, int rel,
                                            int lz, int lx, int ly, int s, int r,
                                            const float *__restrict__ m, int64_t ni, int vz,
                                            ModelsTile<TT, K> &L)
{
    const int t = c.t;
    const int iz = c.tzi * a.tz + lz, ix = c.txi * a.tx + lx, iy = c.tyi * a.ty + ly;
    const bool valid = (iz < a.nz) & (ix < a.nx) & (iy < a.ny);
    const int orig = (c.tyi * a.ty * a.nx + c.txi * a.tx) * a.nz + c.tzi * a.tz;
#pragma unroll
    for (int j = 0; j < K; ++j) {
        // unconditional loads, as in spread_load
        const float mval = (m + j * ni + orig)[valid ? (unsigned)rel : 0u];
        L.mv[j] = valid ? mval : 0.f;
    }
    const size_t ts = (size_t)t * a.ns + s;
    const size_t tr = (size_t)t * a.nr + r;
    L.os = (reinterpret_cast<const typename SrcTab<TT>::type *>(a.off_s) + ts * kWave)[(unsigned)lane];
    L.orr = (reinterpret_cast<const TT *>(a.off_r) + tr * kWave)[(unsigned)lane];
    L.bs = a.base_s[ts + vz];
    L.br = a.base_r[tr + vz];
    if constexpr (DYN) {
        L.ps = (reinterpret_cast<const float *>(a.wp_s) + ts * kWave)[(unsigned)lane];
        L.qs = (reinterpret_cast<const float *>(a.wq_s) + ts * kWave)[(unsigned)lane];
        using RW = typename RecW<float, TT>::type;
        L.pr = (reinterpret_cast<const RW *>(a.wp_r) + tr * kWave)[(unsigned)lane];
        L.qr = (reinterpret_cast<const RW *>(a.wq_r) + tr * kWave)[(unsigned)lane];
    }
}

template <typename TT, bool DYN, bool NEG, bool FIX, int K>
__device__ __forceinline__ void models_accumulate(const ModelsTile<TT, K> &cur, int iz,
                                                  const int (&e)[K], int izm, unsigned ntm1,
                                                  int nt, double *accd, unsigned long long *acci)
{
    bool any = false;
#pragma unroll
    for (int j = 0; j < K; ++j)
        any = any || cur.mv[j] != 0.f;
    if (__ballot(any) == 0)
        return;
    const float T = ld_off(&cur.os) + ld_off(&cur.orr);
    float f;
    int k;
    if constexpr (!NEG) {
        f = __builtin_amdgcn_fractf(T);
        k = floor_int(T) + cur.bs + cur.br;
    } else {
        const float fl = floorf(T);
        f = T - fl;
        k = (int)fl + cur.bs + cur.br;
        if (k == -1 && f > 0.f) {
            k = 0;
            f -= 1.f;
        }
    }
    bool on = (unsigned)k < ntm1;
    if constexpr (DYN)
        on = on && iz >= izm;
    if (!on)
        return;
#pragma unroll
    for (int j = 0; j < K; ++j) {
        float mv = cur.mv[j];
        if constexpr (FIX)
            mv = ldexpf(mv, e[j]);
        float w = mv;
        if constexpr (DYN)
            w = fmaf(mv * cur.qs, wv(cur.qr), (mv * cur.ps) * wv(cur.pr));
        if (w != 0.f) {
            const int it = j * nt + k;
            const float w1 = w * f;
            if constexpr (FIX) {
                const float w0 = fmaf(-w, f, w);
                atomicAdd(&acci[it], (unsigned long long)(long long)round_half_up(w0));
                atomicAdd(&acci[it + 1], (unsigned long long)(long long)round_half_up(w1));
            } else {
                const float w0 = w * (1.f - f);
                atomicAdd(&accd[it], (double)w0);
                atomicAdd(&accd[it + 1], (double)w1);
            }
        }
    }
}

// Traces (s, r) of models [0, K): m + j*ni is model j, y + j*ystride its
// traces; sc[j] its fixed-point scale.
template <typename TT, bool DYN, bool NEG, bool FIX, int K>
__device__ __forceinline__ void spread_models_body(const KArgs &a, const float *__restrict__ m,
                                                   int64_t ni, float *__restrict__ y,
                                                   int64_t ystride,
                                                   const SpreadScale *__restrict__ sc, int s,
                                                   int r)
{
    extern __shared__ __align__(16) unsigned char kx_smem[];
    double *accd = reinterpret_cast<double *>(kx_smem);
    unsigned long long *acci = reinterpret_cast<unsigned long long *>(kx_smem);
    const int ntr = K * a.nt;
    for (int i = threadIdx.x; i < ntr; i += blockDim.x)
        acci[i] = 0ull;
    __syncthreads();

    const int lane = threadIdx.x & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int nw = blockDim.x / kWave;
    const int vz = vgpr_zero();
    const int lz = lane % a.tz, lx = (lane / a.tz) % a.tx, ly = lane / (a.tz * a.tx);
    const int rel = (ly * a.nx + lx) * a.nz + lz;
    int e[K];
#pragma unroll
    for (int j = 0; j < K; ++j)
        e[j] = sc[j].e;
    const int izm = DYN ? a.izmin[s * a.nr + r] : 0;
    const unsigned ntm1 = (unsigned)(a.nt - 1);

    ModelsTile<TT, K> cur, nxt;
    TileCursor c;
    c.start(a, wave);
    if (c.t < a.ntiles)
        models_load<TT, DYN, K>(a, c, lane, rel, lz, lx, ly, s, r, m, ni, vz, cur);
    while (c.t < a.ntiles) {
        TileCursor cn = c;
        cn.advance(a, nw);
        models_load<TT, DYN, K>(a, cn.t < a.ntiles ? cn : c, lane, rel, lz, lx, ly, s, r, m, ni,
                                vz, nxt);
        models_accumulate<TT, DYN, NEG, FIX, K>(cur, c.tzi * a.tz + lz, e, izm, ntm1, a.nt,
                                                accd, acci);
        cur = nxt;
        c = cn;
    }
    __syncthreads();
    const size_t off = ((size_t)s * a.nr + r) * a.nt;
    for (int i = threadIdx.x; i < ntr; i += blockDim.x) {
        const int j = i / a.nt, q = i - j * a.nt;
        float *yo = y + j * ystride + off;
        if constexpr (FIX)
            yo[q] = (float)ldexp((double)(long long)acci[i], -e[j]);
        else
            yo[q] = (float)accd[i];
    }
}

// Per-launch scale of the fixed-point accumulation: per block the partial
// max|m| (inf for a non-finite value), sum of |m| over the finite values and
// count of nonzero values, then one wave turns them into e.  Block b of nb
// strides over m[0, n) and writes part[b], part[nb + b], part[2 nb + b].
struct ModelStats {
    float mx, sum, cnt;
};

__device__ __forceinline__ void stats_block(const float *__restrict__ m, int64_t n, int b, int nb,
                                            float *__restrict__ part)
{
    ModelStats s{0.f, 0.f, 0.f};
    const int64_t stride = (int64_t)nb * blockDim.x;
    for (int64_t i = (int64_t)b * blockDim.x + threadIdx.x; i < n; i += stride) {
        float v = fabsf(m[i]);
        if (!(v <= 3.402823466e38f)) {
            v = INFINITY;  // NaN or inf
        } else {
            s.sum += v;
            s.cnt += v != 0.f ? 1.f : 0.f;
        }
        s.mx = fmaxf(s.mx, v);
    }
#pragma unroll
    for (int o = kWave / 2; o >= 1; o >>= 1) {
        s.mx = fmaxf(s.mx, __shfl_xor(s.mx, o, kWave));
        s.sum += __shfl_xor(s.sum, o, kWave);
        s.cnt += __shfl_xor(s.cnt, o, kWave);
    }
    __shared__ ModelStats ws[kBlock / kWave];
    if ((threadIdx.x & (kWave - 1)) == 0)
        ws[threadIdx.x / kWave] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        ModelStats r = ws[0];
        for (int i = 1; i < (int)(blockDim.x / kWave); ++i) {
            r.mx = fmaxf(r.mx, ws[i].mx);
            r.sum += ws[i].sum;
            r.cnt += ws[i].cnt;
        }
        part[b] = r.mx;
        part[nb + b] = r.sum;
        part[2 * nb + b] = r.cnt;
    }
}

// The np partial statistics combined by one wave (sums in float64): the
// maximum and the mean of the nonzero |m| (0 without any); valid in every
// lane.
__device__ __forceinline__ void stats_wave(const float *__restrict__ part, int np, float &mx,
                                           double &mean)
{
    float x = 0.f;
    double sum = 0.0, cnt = 0.0;
    for (int i = threadIdx.x; i < np; i += kWave) {
        x = fmaxf(x, part[i]);
        sum += part[np + i];
        cnt += part[2 * np + i];
    }
#pragma unroll
    for (int o = kWave / 2; o >= 1; o >>= 1) {
        x = fmaxf(x, __shfl_xor(x, o, kWave));
        sum += __shfl_xor(sum, o, kWave);
        cnt += __shfl_xor(cnt, o, kWave);
    }
    mx = x;
    mean = cnt > 0.0 ? sum / cnt : 0.0;
}

// A contribution is at most |m| * wamp * 2 (the truncation path of a
// negative time weighs a sample by up to 2): e puts that bound below 2^30.
// The scaled image values m 2^e are float32: when max|m| 2^e would not be a
// finite float, or values 2^-40 below the maximum would no longer be normal
// floats (weights in extreme units), the launch accumulates in float64.
// A contribution then rounds by at most 2^-30 max|m| wamp, i.e. by
// 2^-30 R of a typical one, R = max|m| wamp / (mean|m| wtyp) (mean over the
// nonzero m; wtyp a typical weight, wamp when not given): when R exceeds
// kRangeMax = 2^10 (a typical contribution rounding by more than ~2^-20 of
// itself, and a few 1e-6 of the sums it makes), the launch accumulates in
// float64 too.
constexpr double kRangeMax = 1024.0;

__device__ __forceinline__ SpreadScale scale_from_stats(float mx, double mean, double wamp,
                                                        double wtyp)
{
    const double bound = 2.0 * (double)mx * wamp;
    SpreadScale sc;
    if (!(bound <= 1.7976931348623157e308)) {
        sc.e = 0;
        sc.fixed = 0;
    } else if (bound == 0.0) {
        sc.e = 0;
        sc.fixed = 1;
    } else {
        int ex;
        frexp(bound, &ex);  // bound < 2^ex
        sc.e = 30 - ex;
        const double top = ldexp((double)mx, sc.e);
        sc.fixed = (top <= 0x1p126 && ldexp(top, -40) >= 0x1p-126) ? 1 : 0;
        const double wt = wtyp > 0.0 ? wtyp : wamp;
        if (mean > 0.0 && (double)mx * wamp > kRangeMax * mean * wt)
            sc.fixed = 0;
    }
    return sc;
}

__global__ __launch_bounds__(kBlock) void absmax_partial_kernel(const float *__restrict__ m,
                                                                int64_t n,
                                                                float *__restrict__ part)
{
    stats_block(m, n, blockIdx.x, gridDim.x, part);
}

__global__ __launch_bounds__(kWave) void scale_final_kernel(const float *__restrict__ part, int np,
                                                            double wamp, double wtyp,
                                                            SpreadScale *__restrict__ sc)
{
    float mx;
    double mean;
    stats_wave(part, np, mx, mean);
    if (threadIdx.x == 0)
        *sc = scale_from_stats(mx, mean, wamp, wtyp);
}

// ---------------------------------------------------------------------------
// adjoint: gather per image point from wave-staged trace windows
// ---------------------------------------------------------------------------
// A wave owns a group of NP neighbouring tiles (NP = 1, 2 or 4: adjacent
// along x, or 2 x 2 in x and y; KArgs.gx x KArgs.gy tiles) and a share of
// the sources (P waves of a workgroup split a group's sources; their partial
// sums are added in a fixed order at the end).  It sweeps the receivers r.
// For each of its JW sources the trace window of kWin samples starting at
// the group's mean base sum (source plus receiver base) minus kWin / 2 is
// staged a round ahead: lane l loads window sample l with one coalesced
// load, and at the start of the round writes the pair (d[k], d[k+1] - d[k])
// of window position l to the wave's own LDS region, d[k+1] coming from lane
// l + 1 by a DPP lane shift.  Pairs outside [0, nt - 2] (the reference's
// 0 <= it < nt - 1) and the last position (no neighbour) are zero.  A
// point's contribution is then one 8-byte LDS read and one fma,
// fma(f, d[k+1] - d[k], d[k]), the reference's d[k] (1 - f) + d[k+1] f.  A
// sample outside its window reads the zero pair and is marked; the wave
// adds the marked samples afterwards from global memory (rare: a tile's
// offsets span a few tens of samples).  A 64-lane gather from global memory
// costs the texture path ~10x the coalesced load and LDS read that replace
// it, and a window serves the NP tiles of the group, so the staging is paid
// once per NP points.  No barrier until the final sum.
constexpr int kGatherWaves = 4;
constexpr int kGatherBlock = kGatherWaves * kWave;
constexpr int kWin = kWave;  // samples per staged window, one per lane

// lane l + 1's value (DPP wave_shl:1; lane 63 receives 0)
__device__ __forceinline__ float lane_next(float v)
{
    return __builtin_bit_cast(
        float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x130, 0xf, 0xf, false));
}
__device__ __forceinline__ double lane_next(double v)
{
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)u, 0x130, 0xf,
                                                              0xf, false);
    const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)(u >> 32), 0x130,
                                                              0xf, 0xf, false);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

template <int N> struct IntC {
    static constexpr int value = N;
};

template <int NP> struct Log2 {
    static constexpr int v = NP >= 4 ? 2 : (NP >= 2 ? 1 : 0);
};

// Tile-group layout of the adjoint: groups of gx x gy tiles (gx * gy = np),
// z fastest, then x, then y.
inline void set_tile_groups(KArgs &a, int np)
{
    if (np <= 1) {
        a.gx = a.gy = 1;
    } else if (np == 2 || a.nty == 1) {
        a.gx = np;
        a.gy = 1;
    } else {
        a.gx = 2;
        a.gy = np / 2;
    }
    a.ngx = (a.ntx + a.gx - 1) / a.gx;
    a.ngy = (a.nty + a.gy - 1) / a.gy;
    a.ngroups = a.ntz * a.ngx * a.ngy;
}

// K > 1: the data of K models (model k at d + k*dstride, its image at
// m + k*mstride) share the tables, sample indices, weights and window
// positions; each model has its own staged windows and sums, with the
// single-model operations in the single-model order, so each image is
// bitwise the one K = 1 gives.
template <typename TD, typename TT, bool DYN, bool NEG, int JW, int P, int NP, int K = 1,
          bool GATE = true>
__device__ __forceinline__ void gather_tiles(const KArgs &a, const TD *__restrict__ d,
                                             TD *__restrict__ m, int L, int64_t dstride = 0,
                                             int64_t mstride = 0)
{
    using TC = typename TimeType<TD>::type;
    using TP = VecN<TD, 2>;               // (d[k], d[k+1] - d[k])
    constexpr int NG = kGatherWaves / P;  // tile groups per workgroup
    constexpr int SB = JW * P;            // sources per block (all parts)
    constexpr int kFar = -(1 << 26);      // base of a missing source: before every trace
    __shared__ TP win[kGatherWaves][K][JW][kWin];
    static_assert(K * NP * kGatherBlock * sizeof(TD) <= sizeof(win), "reduction space");

    const int lane = threadIdx.x & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int part = wave / NG;  // this wave sums sources s0 + part + P i
    const int gq = L * NG + wave % NG;
    const bool active = gq < a.ngroups;
    const int g = active ? gq : a.ngroups - 1;
    const int tzi = g % a.ntz;
    const int grest = g / a.ntz;
    const int gx0 = (grest % a.ngx) * a.gx, gy0 = (grest / a.ngx) * a.gy;
    int t[NP];
    bool tok[NP];
    TilePoint pt[NP];
#pragma unroll
    for (int j = 0; j < NP; ++j) {
        const int txi = gx0 + j % a.gx, tyi = gy0 + j / a.gx;
        tok[j] = txi < a.ntx && tyi < a.nty;
        // a tile past the grid's edge repeats the group's first tile and
        // stores nothing
        t[j] = tzi + a.ntz * (tok[j] ? txi + a.ntx * tyi : gx0 + a.ntx * gy0);
        pt[j] = tile_point(a, t[j], lane);
    }
    const typename SrcTab<TT>::type *offs =
        reinterpret_cast<const typename SrcTab<TT>::type *>(a.off_s);
    const TT *offr = reinterpret_cast<const TT *>(a.off_r);
    const TD *wps = reinterpret_cast<const TD *>(a.wp_s);
    const TD *wqs = reinterpret_cast<const TD *>(a.wq_s);
    using RW = typename RecW<TD, TT>::type;
    const RW *wpr = reinterpret_cast<const RW *>(a.wp_r);
    const RW *wqr = reinterpret_cast<const RW *>(a.wq_r);
    const unsigned ntm1 = (unsigned)(a.nt - 1);
    const int vz = vgpr_zero();
    TP(&w)[K][JW][kWin] = win[wave];

    TD acc[K][NP];
#pragma unroll
    for (int k = 0; k < K; ++k)
#pragma unroll
        for (int j = 0; j < NP; ++j)
            acc[k][j] = TD(0);
    // One block of SB sources, of which this wave's slots i < JB are
    // summed: JB = JW, or JW / 2 in a last block where at most half of
    // the wave's slots hold a source (an empty slot adds zero pairs at the
    // full arithmetic cost of a source).
    auto source_block = [&](auto jw, const int s0) __attribute__((always_inline)) {
        constexpr int JB = decltype(jw)::value;
        // this wave's sources s0 + part + P i: offsets per point, bases
        // per tile (wave-uniform, held in scalar registers)
        TC os[JB][NP];
        int bs[JB][NP];
        TD ps[JB][NP], qs[JB][NP];
        // trace (source, receiver 0) as an element offset from the
        // block's first trace (fits 31 bits: checked at launch), held in
        // vector registers: JB 64-bit pointers would exhaust the scalar
        // ones
        const TD *dblk = d + (size_t)s0 * a.nr * a.nt;
        int soff[JB], izoff[JB];
#pragma unroll
        for (int i = 0; i < JB; ++i) {
            const int js = s0 + part + P * i;
#pragma unroll
            for (int j = 0; j < NP; ++j) {
                if (js < a.ns) {
                    const size_t so = (size_t)t[j] * a.ns + js;
                    os[i][j] = (TC)ld_off(offs + so * kWave + lane);
                    bs[i][j] = __builtin_amdgcn_readfirstlane(a.base_s[so + vz]);
                    if constexpr (DYN) {
                        ps[i][j] = wps[so * kWave + lane];
                        qs[i][j] = wqs[so * kWave + lane];
                    }
                } else {  // past the last source: every sample before the trace
                    os[i][j] = TC(0);
                    bs[i][j] = kFar;
                    ps[i][j] = qs[i][j] = TD(0);
                }
            }
            soff[i] = (min(js, a.ns - 1) - s0) * a.nr * a.nt + vz;
            izoff[i] = min(js, a.ns - 1) * a.nr + vz;
        }

        // the window of source i in round r starts at ws = mean over the
        // group's tiles of bs + br(r), minus kWin / 2; lane l stages
        // sample ws + l (a clamped load; svok: the pair of position l lies
        // in [0, nt - 2] and has a neighbour), the aperture gate of the
        // round is loaded with it
        TD sv[K][JB];
        bool svok[JB];
        int wsn[JB], izn[JB] = {};
#define KX_STAGE(R, BR)                                                                        \
    _Pragma("unroll") for (int i = 0; i < JB; ++i)                                              \
    {                                                                                          \
        int sum = 0;                                                                           \
        _Pragma("unroll") for (int j = 0; j < NP; ++j) sum += bs[i][j] + (BR)[j];              \
        wsn[i] = (sum >> Log2<NP>::v) - kWin / 2;                                              \
        const int e = wsn[i] + lane;                                                           \
        const int ec = clamp0(e, a.nt - 1);                                                    \
        _Pragma("unroll") for (int k = 0; k < K; ++k)                                         \
            sv[k][i] = dblk[k * dstride + (unsigned)(soff[i] + (R) * a.nt + ec)];              \
        svok[i] = (unsigned)e < ntm1 && lane < kWin - 1;                                       \
        if constexpr (DYN && GATE)                                                             \
            izn[i] = a.izmin[(unsigned)(izoff[i] + (R))];                                      \
    }
        // receiver side of the round ahead; receiver bases two rounds
        // ahead (the window positions of the round ahead depend on them)
        TT orr_c[NP];
        int br_c[NP], br_n[NP];
        RW pr_c[NP], qr_c[NP];
#pragma unroll
        for (int j = 0; j < NP; ++j) {
            const size_t ro = (size_t)t[j] * a.nr;
            orr_c[j] = offr[ro * kWave + lane];
            br_c[j] = __builtin_amdgcn_readfirstlane(a.base_r[ro + vz]);
            br_n[j] = a.nr > 1 ? a.base_r[ro + 1 + vz] : 0;
            pr_c[j] = qr_c[j] = RW(0.0f);
            if constexpr (DYN) {
                pr_c[j] = wpr[ro * kWave + lane];
                qr_c[j] = wqr[ro * kWave + lane];
            }
        }
        KX_STAGE(0, br_c)
        for (int r = 0; r < a.nr; ++r) {
            // this round's pairs into the wave's LDS region
            int ws[JB], izr[JB];
#pragma unroll
            for (int i = 0; i < JB; ++i) {
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    const TD nx = lane_next(sv[k][i]);
                    TP q;
                    q.v[0] = svok[i] ? sv[k][i] : TD(0);
                    q.v[1] = svok[i] ? nx - sv[k][i] : TD(0);
                    w[k][i][lane] = q;
                }
                ws[i] = wsn[i];
                izr[i] = izn[i];
            }
            TC orr[NP];
            int br[NP];
            TD pr[NP], qr[NP];
#pragma unroll
            for (int j = 0; j < NP; ++j) {
                orr[j] = (TC)ld_off(&orr_c[j]);
                br[j] = br_c[j];
                pr[j] = (TD)wv(pr_c[j]);
                qr[j] = (TD)wv(qr_c[j]);
            }
            // next round's receiver side and samples, in flight meanwhile.
            // Every round issues the same loads (the last ones reload the
            // last receiver, unused): with loads under a condition the
            // compiler cannot count them and waits for this round's
            // first staged samples before reading the windows.
            {
                const int r1 = min(r + 1, a.nr - 1), r2 = min(r + 2, a.nr - 1);
#pragma unroll
                for (int j = 0; j < NP; ++j) {
                    const size_t ro = (size_t)t[j] * a.nr + r1;
                    orr_c[j] = offr[ro * kWave + lane];
                    if constexpr (DYN) {
                        pr_c[j] = wpr[ro * kWave + lane];
                        qr_c[j] = wqr[ro * kWave + lane];
                    }
                    br_c[j] = __builtin_amdgcn_readfirstlane(br_n[j]);
                }
                KX_STAGE(r1, br_c)
#pragma unroll
                for (int j = 0; j < NP; ++j)
                    br_n[j] = a.base_r[(size_t)t[j] * a.nr + r2 + vz];
            }
            // Each lane reads one pair per (source, tile) from the staged
            // windows; a sample outside its window reads the zero pair,
            // and the wave adds it afterwards from global memory, so that
            // the common path has no per-sample branch.  The common path
            // keeps only the largest window index of the lane: the rare
            // wave with a lane outside its windows finds the samples
            // again with the same index computation.
            auto window_index = [&](int i, int j, TC &f) -> int {
                const int db = bs[i][j] + br[j] - ws[i];  // wave-uniform
                const TC T = os[i][j] + orr[j];
                int kk;
                if constexpr (!NEG && sizeof(TC) == 4) {
                    f = __builtin_amdgcn_fractf(T);
                    kk = floor_int(T) + db;
                } else {
                    const TC fl = floor(T);
                    f = T - fl;
                    kk = (int)fl + db;
                    if constexpr (NEG) {
                        // the reference truncates toward zero: a time in
                        // (-1, 0) reads sample 0 with a negative weight
                        if (kk + ws[i] == -1 && f > TC(0)) {
                            kk += 1;
                            f -= TC(1);
                        }
                    }
                }
                return kk;
            };
            unsigned kmax = 0;
#pragma unroll
            for (int i = 0; i < JB; ++i) {
#pragma unroll
                for (int j = 0; j < NP; ++j) {
                    TC f;
                    const unsigned kk = (unsigned)window_index(i, j, f);
                    kmax = max(kmax, kk);
                    // outside the window: position kWin - 1, the zero pair
                    const unsigned kc = min(kk, (unsigned)(kWin - 1));
                    TD wgt = TD(1);
                    if constexpr (DYN) {
                        wgt = ps[i][j] * pr[j] + qs[i][j] * qr[j];
                        if (GATE && pt[j].iz < izr[i])
                            wgt = TD(0);
                    }
#pragma unroll
                    for (int k = 0; k < K; ++k) {
                        const TP q = w[k][i][kc];
                        TD v = fma((TD)f, q.v[1], q.v[0]);
                        if constexpr (DYN)
                            v *= wgt;
                        acc[k][j] += v;
                    }
                }
            }
            if (__ballot(kmax > (unsigned)(kWin - 2)) != 0) {
                // the samples outside the windows, from global memory
#pragma unroll
                for (int i = 0; i < JB; ++i) {
#pragma unroll
                    for (int j = 0; j < NP; ++j) {
                        TC fw;
                        if ((unsigned)window_index(i, j, fw) > (unsigned)(kWin - 2)) {
                            const TC T = os[i][j] + orr[j];
                            const TC fl = floor(T);
                            TC f = T - fl;
                            int it = (int)fl + bs[i][j] + br[j];
                            if constexpr (NEG) {
                                if (it == -1 && f > TC(0)) {
                                    it = 0;
                                    f -= TC(1);
                                }
                            }
                            if ((unsigned)it < ntm1) {
                                TD wgt = TD(1);
                                if constexpr (DYN) {
                                    wgt = ps[i][j] * pr[j] + qs[i][j] * qr[j];
                                    if (GATE && pt[j].iz < izr[i])
                                        wgt = TD(0);
                                }
#pragma unroll
                                for (int k = 0; k < K; ++k) {
                                    const TD *q = dblk + k * dstride + (soff[i] + r * a.nt + it);
                                    const TD g0 = q[0], g1 = q[1];
                                    TD v = fma((TD)f, g1 - g0, g0);
                                    if constexpr (DYN)
                                        v *= wgt;
                                    acc[k][j] += v;
                                }
                            }
                        }
                    }
                }
            }
        }
#undef KX_STAGE
    };
    if (active) {
        for (int s0 = 0; s0 < a.ns; s0 += SB) {
            // slots i with a source s0 + part + P i < ns
            const int nj = (a.ns - s0 - part + P - 1) / P;
            if (nj > JW / 2)
                source_block(IntC<JW>{}, s0);
            else if (nj > 0)
                source_block(IntC<(JW > 1 ? JW / 2 : 1)>{}, s0);
        }
    }
    if constexpr (P > 1) {
        // the partial sums of a group's points, added in part order
        __syncthreads();
        TD *red = reinterpret_cast<TD *>(&win[0][0][0][0]);
#pragma unroll
        for (int k = 0; k < K; ++k)
#pragma unroll
            for (int j = 0; j < NP; ++j)
                red[(k * NP + j) * kGatherBlock + threadIdx.x] = acc[k][j];
        __syncthreads();
        if (part == 0) {
#pragma unroll
            for (int k = 0; k < K; ++k)
#pragma unroll
                for (int j = 0; j < NP; ++j) {
                    acc[k][j] = TD(0);
#pragma unroll
                    for (int q = 0; q < P; ++q)
                        acc[k][j] += red[(k * NP + j) * kGatherBlock + (q * NG + wave) * kWave + lane];
                }
        }
    }
    if (part == 0 && active) {
#pragma unroll
        for (int j = 0; j < NP; ++j) {
            if (tok[j] && pt[j].valid) {
                const int64_t pi = point_index(a, pt[j]);
#pragma unroll
                for (int k = 0; k < K; ++k)
                    m[k * mstride + pi] = acc[k][j];
            }
        }
    }
    if constexpr (P > 1)
        __syncthreads();  // the partial sums are read before a next item stages
}

// ---------------------------------------------------------------------------
// "same" convolution along time, odd filter
// ---------------------------------------------------------------------------
// One workgroup of kConvThreads computes kConvOut consecutive outputs of one
// trace, kConvPer adjacent outputs per thread.  The flipped filter and the
// trace window (outputs plus the filter's halo) are staged in LDS, zero
// padded to whole 4-element groups; a float thread reads both as float4
// (consecutive lanes read consecutive 16 bytes: no bank conflicts) and keeps
// the 8 window samples of two groups in registers, so 16 multiply-adds cost
// two LDS reads.


// LDS elements of one convolution workgroup: filter + window + one group of
// read-ahead
__host__ __device__ __forceinline__ int conv_lds_elems(int nh)
{
    return conv_taps_padded(nh) + kConvOut + conv_taps_padded(nh) + 4;
}

// Output samples [n0, n0 + kConvOut) of trace tr:
// y[n] = sum_k h[k] x[n + (nh - 1) / 2 - k], with the trace's signal row
// at x + tr * xstride and its filter at h + tr * hstride (a stride of 0
// shares one row between all traces); y rows are nt apart.
template <typename TD>
__device__ __forceinline__ void conv_segment(const TD *__restrict__ x, int64_t xstride,
                                             TD *__restrict__ y, int nt, const TD *__restrict__ h,
                                             int64_t hstride, int nh, int64_t tr, int n0)
{
    extern __shared__ __align__(16) unsigned char kx_smem[];
    const int nhp = conv_taps_padded(nh);
    TD *hs = reinterpret_cast<TD *>(kx_smem);
    TD *win = hs + nhp;
    const int c = (nh - 1) / 2;
    const int start = n0 + c - (nh - 1);
    const int nwin = kConvOut + nhp + 4;
    const TD *xr = x + tr * xstride;
    const TD *hr = h + tr * hstride;
    for (int k = threadIdx.x; k < nhp; k += blockDim.x)
        hs[k] = k < nh ? hr[nh - 1 - k] : TD(0);  // flipped: out[i] = sum_k hs[k] win[i + k]
    for (int j = threadIdx.x; j < nwin; j += blockDim.x) {
        const int n = start + j;
        win[j] = (n >= 0 && n < nt) ? xr[n] : TD(0);
    }
    __syncthreads();
    const int i0 = threadIdx.x * kConvPer;  // first output of this thread
    TD acc[kConvPer] = {};
    if constexpr (sizeof(TD) == 4) {
        const float4 *w4 = reinterpret_cast<const float4 *>(win) + threadIdx.x;
        const float4 *h4 = reinterpret_cast<const float4 *>(hs);
        float4 cur = w4[0];
        for (int g = 0; g < nhp / 4; ++g) {
            const float4 nxt = w4[g + 1];
            const float4 hq = h4[g];
            const float w[8] = {cur.x, cur.y, cur.z, cur.w, nxt.x, nxt.y, nxt.z, nxt.w};
            const float hh[4] = {hq.x, hq.y, hq.z, hq.w};
#pragma unroll
            for (int q = 0; q < 4; ++q)
#pragma unroll
                for (int r = 0; r < kConvPer; ++r)
                    acc[r] += hh[q] * w[q + r];
            cur = nxt;
        }
    } else {
        for (int k = 0; k < nhp; ++k) {
            const TD hk = hs[k];
#pragma unroll
            for (int r = 0; r < kConvPer; ++r)
                acc[r] += hk * win[i0 + r + k];
        }
    }
#pragma unroll
    for (int r = 0; r < kConvPer; ++r)
        if (n0 + i0 + r < nt)
            y[tr * nt + n0 + i0 + r] = acc[r];
    __syncthreads();  // the window is read before a next item loads its own
}

// One lag of one trace: y = sum_{u < la} a[u] b[lag + u], b zero outside
// [0, lb); summed in float64 by the workgroup (kXcorrThreads threads).
constexpr int kXcorrThreads = 256;
template <typename TD>
__device__ __forceinline__ TD xcorr_lag(const TD *__restrict__ a, int la, const TD *__restrict__ b,
                                        int lb, int lag)
{
    __shared__ double part[kXcorrThreads / kWave];
    double acc = 0.0;
    const int u0 = max(0, -lag), u1 = min(la, lb - lag);
    for (int u = u0 + (int)threadIdx.x; u < u1; u += blockDim.x)
        acc += (double)a[u] * (double)b[lag + u];
#pragma unroll
    for (int o = kWave / 2; o >= 1; o >>= 1)
        acc += __shfl_xor(acc, o, kWave);
    if ((threadIdx.x & (kWave - 1)) == 0)
        part[threadIdx.x / kWave] = acc;
    __syncthreads();
    double t = 0.0;
    for (int w = 0; w < (int)(blockDim.x / kWave); ++w)
        t += part[w];
    __syncthreads();  // part is read before a next call writes it
    return (TD)t;
}

}  // namespace
