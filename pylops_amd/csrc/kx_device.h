// kx_device.h -- device code shared by the single-problem (kirchx.hip) and
// batched (kx_batch.hip) launches of the Kirchhoff kernels: data types, tile
// geometry and the bodies of the forward, adjoint and convolution kernels,
// each taking its work item explicitly.
#pragma once
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cstdint>

#include "kirchx.h"

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
    // forward output: the trace of (source s, block member j of a block
    // starting at r0) is trace s * trs + (r0 + j) * trb; (nr, 1) unless the
    // forward blocks sources (the side tables exchanged, then (1, ns))
    int trs, trb;
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

// Receiver bases held per lane: the base of receiver (lane mod RB) of the
// block, one load for all RB receivers, handed to the lanes that need it by
// a lane read (natural order) or a quad permutation (rotated order, RB <= 4);
// rotated blocks of 8 and 16 receivers load one base per step instead.
template <bool ROT, int RB> struct BaseLoads {
    static constexpr int n = (ROT && RB >= 8) ? RB : 1;
};

// One lane's share of one tile: the image value and the table entries of
// its point, receivers in the lane's step order.
template <typename TD, typename TT, bool ROT, int RB> struct SpreadTile {
    TD mv;
    TT os;       // raw table elements: converted at use, so that a prefetch
    TT orr[RB];  // never waits for them
    TD ps, qs;
    TD pr[RB], qr[RB];
    int bs;
    int br[BaseLoads<ROT, RB>::n];
};

// The receiver base of step r for this lane.
template <bool ROT, int RB>
__device__ __forceinline__ int step_base(const int (&br)[BaseLoads<ROT, RB>::n], int r)
{
    if constexpr (BaseLoads<ROT, RB>::n > 1) {
        return br[r];
    } else if constexpr (!ROT || RB == 1) {
        // every lane treats receiver r: the base held by lane r
        return __builtin_amdgcn_readlane(br[0], r);
    } else if constexpr (RB == 2) {
        // lane l treats receiver (l + r) mod 2: swap lane pairs on odd steps
        return (r & 1) ? __builtin_amdgcn_update_dpp(0, br[0], 0xB1, 0xf, 0xf, false) : br[0];
    } else {
        // RB == 4: lane l treats receiver (l + r) mod 4, held by lane
        // (l & ~3) | ((l + r) & 3): quad_perm [r, r+1, r+2, r+3] mod 4
        switch (r & 3) {
        case 1:
            return __builtin_amdgcn_update_dpp(0, br[0], 0x39, 0xf, 0xf, false);
        case 2:
            return __builtin_amdgcn_update_dpp(0, br[0], 0x4E, 0xf, 0xf, false);
        case 3:
            return __builtin_amdgcn_update_dpp(0, br[0], 0x93, 0xf, 0xf, false);
        default:
            return br[0];
        }
    }
}

template <typename TD, typename TT, bool DYN, bool ROT, int RB>
__device__ __forceinline__ void spread_load(const KArgs &a, const TileCursor &c, int lane, int rel,
                                            int lz, int lx, int ly, int s, int r0, int nrl,
                                            const TD *__restrict__ m, int vz,
                                            SpreadTile<TD, TT, ROT, RB> &L)
{
    const int t = c.t;
    const int iz = c.tzi * a.tz + lz, ix = c.txi * a.tx + lx, iy = c.tyi * a.ty + ly;
    const bool valid = (iz < a.nz) & (ix < a.nx) & (iy < a.ny);
    const int orig = (c.tyi * a.ty * a.nx + c.txi * a.tx) * a.nz + c.tzi * a.tz;
    // Row pointers are wave-uniform and the lane offsets 32-bit unsigned, so
    // every load takes the scalar-base form (one VALU op of addressing).
    L.mv = valid ? (m + orig)[(unsigned)rel] : TD(0);
    const size_t ts = (size_t)t * a.ns + s;        // (tile, source) row
    const size_t tr = (size_t)t * a.nr + r0;       // (tile, first receiver) row
    L.os = (reinterpret_cast<const TT *>(a.off_s) + ts * kWave)[(unsigned)lane];
    L.bs = a.base_s[ts + vz];
    if constexpr (DYN) {
        L.ps = (reinterpret_cast<const TD *>(a.wp_s) + ts * kWave)[(unsigned)lane];
        L.qs = (reinterpret_cast<const TD *>(a.wq_s) + ts * kWave)[(unsigned)lane];
    }
    const TT *orow = reinterpret_cast<const TT *>(a.off_r) + tr * kWave;
    const int *brow = a.base_r + tr;
    if constexpr (BaseLoads<ROT, RB>::n == 1) {
        const int jl = lane & (RB - 1);
        L.br[0] = brow[(unsigned)((jl < nrl ? jl : 0) + vz)];
    }
#pragma unroll
    for (int r = 0; r < RB; ++r) {
        int j = step_receiver<RB, ROT>(lane, r);
        j = j < nrl ? j : 0;  // rows past the last receiver read row 0
        const unsigned e = (unsigned)(j * kWave + lane);
        L.orr[r] = orow[e];
        if constexpr (BaseLoads<ROT, RB>::n > 1)
            L.br[r] = brow[(unsigned)(j + vz)];
        if constexpr (DYN) {
            L.pr[r] = (reinterpret_cast<const TD *>(a.wp_r) + tr * kWave)[e];
            L.qr[r] = (reinterpret_cast<const TD *>(a.wq_r) + tr * kWave)[e];
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

// (int)floor(x) in one conversion instruction (|x| < 2^31)
__device__ __forceinline__ int floor_int(float x)
{
    int i;
    asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(i) : "v"(x));
    return i;
}

// One contribution of weight w at time T samples past the integer base:
// w (1 - f) to sample k and w f to sample k + 1 of the LDS trace at lrow,
// k = floor(T) + base, f = T - floor(T).  Lanes with lrow < 0 (no receiver),
// with `on` false (aperture), or whose k + 1 falls past the trace add
// nothing.  The index and fraction of a float time come from one conversion
// and one fract instruction: fract(T) is T - floor(T) rounded as the
// subtraction rounds it, except that a result that would round to 1 (T in
// (-2^-24, 0)) stays just below 1.
template <typename TD, typename TC, bool NEG, bool FIX>
__device__ __forceinline__ void spread_sample(TC T, int base, TD w, bool on, int lrow,
                                              unsigned ntm1, double *accd,
                                              unsigned long long *acci)
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
        const int it = lrow + k;
        const TD w1 = w * (TD)f;
        if constexpr (FIX) {
            // w (1 - f) in one rounding; fixed point by floor(x + 1/2) in
            // one conversion
            const TD w0 = fma(-w, (TD)f, w);
            atomicAdd(&acci[it], (unsigned long long)(long long)round_half_up(w0));
            atomicAdd(&acci[it + 1], (unsigned long long)(long long)round_half_up(w1));
        } else {
            const TD w0 = w * (TD)(TC(1) - f);
            atomicAdd(&accd[it], (double)w0);
            atomicAdd(&accd[it + 1], (double)w1);
        }
    }
}

// Accumulates one tile (64 points of a wave, RB receivers) into the LDS
// traces.  Dynamic weights are (m ps) pr + (m qs) qr with the source-side
// products formed once per tile; with the sides exchanged (XS: the block
// members are sources) the same products are formed per step, so that the
// weights round as they do with blocks of receivers.
template <typename TD, typename TT, bool DYN, bool NEG, bool ROT, bool FIX, int RB,
          bool XS = false>
__device__ __forceinline__ void spread_accumulate(const SpreadTile<TD, TT, ROT, RB> &cur, int iz,
                                                  int e,
                                                  const int (&lrow)[RB], const int (&izm)[RB],
                                                  unsigned ntm1, double *accd,
                                                  unsigned long long *acci)
{
    using TC = typename TimeType<TD>::type;
    if (__ballot(cur.mv != TD(0)) == 0)
        return;
    const TC os = (TC)ld_off(&cur.os);
    // the image value, scaled by 2^e for the fixed-point sum
    TD mv = cur.mv;
    if constexpr (FIX)
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
            if constexpr (XS)  // (ps, qs) here are the receiver's, pr[r], qr[r] the source's
                w = fma(mv * cur.qr[r], cur.qs, (mv * cur.pr[r]) * cur.ps);
            else
                w = fma(mqs, cur.qr[r], mps * cur.pr[r]);
            on = iz >= izm[r];
        }
        spread_sample<TD, TC, NEG, FIX>(T, cur.bs + step_base<ROT, RB>(cur.br, r), w, on,
                                        lrow[r], ntm1, accd, acci);
    }
}

// Zeroes the ntr samples of a workgroup's LDS traces (0.0 and integer 0
// share the bit pattern).
__device__ __forceinline__ void traces_clear(unsigned long long *acci, int ntr)
{
    for (int i = threadIdx.x; i < ntr; i += blockDim.x)
        acci[i] = 0ull;
    __syncthreads();
}

// Writes the workgroup's LDS traces (ntr = rows * nt samples) to y: row j
// to trace first + j * trb; fixed point scaled back by 2^-e, or float64
// sums rounded to TD.
template <typename TD, bool FIX>
__device__ __forceinline__ void traces_store(TD *__restrict__ y, size_t first, int trb, int nt,
                                             int ntr, int e, const double *accd,
                                             const unsigned long long *acci)
{
    __syncthreads();
    TD *yo = y + first * nt;
    for (int i = threadIdx.x; i < ntr; i += blockDim.x) {
        TD v;
        if constexpr (FIX)
            v = (TD)ldexp((double)(long long)acci[i], -e);
        else
            v = (TD)accd[i];
        if (trb == 1) {
            yo[i] = v;  // the block's traces are adjacent
        } else {
            const int j = i / nt;
            yo[(size_t)j * trb * nt + (i - j * nt)] = v;
        }
    }
    __syncthreads();  // the traces are read out before a next item clears them
}

template <typename TD, typename TT, bool DYN, bool NEG, bool ROT, bool FIX, int RB,
          bool XS = false>
__device__ __forceinline__ void spread_body(const KArgs &a, const TD *__restrict__ m,
                                            TD *__restrict__ y, int e, int s, int r0)
{
    extern __shared__ __align__(16) unsigned char kx_smem[];
    double *accd = reinterpret_cast<double *>(kx_smem);
    unsigned long long *acci = reinterpret_cast<unsigned long long *>(kx_smem);

    // one workgroup = one source s and receivers [r0, r0 + RB)
    const int nrl = min(RB, a.nr - r0);
    const int ntr = nrl * a.nt;

    traces_clear(acci, ntr);

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
        if constexpr (DYN)  // the gate of a pair is indexed like its trace
            izm[r] = a.izmin[(size_t)s * a.trs + (size_t)(r0 + (j < nrl ? j : 0)) * a.trb];
    }
    const unsigned ntm1 = (unsigned)(a.nt - 1);

    // the next tile is loaded while this one is accumulated (a twice
    // unrolled loop with two buffers and no copies measured slower: the
    // second buffer costs registers and occupancy)
    SpreadTile<TD, TT, ROT, RB> cur, nxt;
    TileCursor c;
    c.start(a, wave);
    if (c.t < a.ntiles)
        spread_load<TD, TT, DYN, ROT, RB>(a, c, lane, rel, lz, lx, ly, s, r0, nrl, m, vz, cur);
    while (c.t < a.ntiles) {
        TileCursor cn = c;
        cn.advance(a, nw);
        if (cn.t < a.ntiles)
            spread_load<TD, TT, DYN, ROT, RB>(a, cn, lane, rel, lz, lx, ly, s, r0, nrl, m, vz, nxt);
        spread_accumulate<TD, TT, DYN, NEG, ROT, FIX, RB, XS>(cur, c.tzi * a.tz + lz, e, lrow,
                                                              izm, ntm1, accd, acci);
        cur = nxt;
        c = cn;
    }
    traces_store<TD, FIX>(y, (size_t)s * a.trs + (size_t)r0 * a.trb, a.trb, a.nt, ntr, e, accd,
                          acci);
}

// One (source, receiver block) item with the launch's accumulation scale.
template <typename TD, typename TT, bool DYN, bool NEG, bool ROT, int RB, bool XS = false>
__device__ __forceinline__ void spread_item(const KArgs &a, const TD *__restrict__ m,
                                            TD *__restrict__ y, const SpreadScale &sc, int s,
                                            int r0)
{
    if constexpr (sizeof(TD) == 4) {
        if (sc.fixed)
            spread_body<TD, TT, DYN, NEG, ROT, true, RB, XS>(a, m, y, sc.e, s, r0);
        else
            spread_body<TD, TT, DYN, NEG, ROT, false, RB, XS>(a, m, y, 0, s, r0);
    } else {
        spread_body<TD, TT, DYN, NEG, ROT, false, RB, XS>(a, m, y, 0, s, r0);
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
    TT os, orr;
    float ps, qs, pr, qr;
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
    for (int j = 0; j < K; ++j)
        L.mv[j] = valid ? (m + j * ni + orig)[(unsigned)rel] : 0.f;
    const size_t ts = (size_t)t * a.ns + s;
    const size_t tr = (size_t)t * a.nr + r;
    L.os = (reinterpret_cast<const TT *>(a.off_s) + ts * kWave)[(unsigned)lane];
    L.orr = (reinterpret_cast<const TT *>(a.off_r) + tr * kWave)[(unsigned)lane];
    L.bs = a.base_s[ts + vz];
    L.br = a.base_r[tr + vz];
    if constexpr (DYN) {
        L.ps = (reinterpret_cast<const float *>(a.wp_s) + ts * kWave)[(unsigned)lane];
        L.qs = (reinterpret_cast<const float *>(a.wq_s) + ts * kWave)[(unsigned)lane];
        L.pr = (reinterpret_cast<const float *>(a.wp_r) + tr * kWave)[(unsigned)lane];
        L.qr = (reinterpret_cast<const float *>(a.wq_r) + tr * kWave)[(unsigned)lane];
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
            w = fmaf(mv * cur.qs, cur.qr, (mv * cur.ps) * cur.pr);
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
        if (cn.t < a.ntiles)
            models_load<TT, DYN, K>(a, cn, lane, rel, lz, lx, ly, s, r, m, ni, vz, nxt);
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

// Per-launch scale of the fixed-point accumulation: partial max|m| per block
// (inf for a non-finite value), then one wave turns the maximum into e.
// Block b of nb strides over m[0, n); the result is valid in thread 0.
__device__ __forceinline__ float absmax_block(const float *__restrict__ m, int64_t n, int b, int nb)
{
    float mx = 0.f;
    const int64_t stride = (int64_t)nb * blockDim.x;
    for (int64_t i = (int64_t)b * blockDim.x + threadIdx.x; i < n; i += stride) {
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
    float r = wmax[0];
    for (int i = 1; i < (int)(blockDim.x / kWave); ++i)
        r = fmaxf(r, wmax[i]);
    return r;
}

// Maximum of np partial maxima, by one wave; valid in every lane.
__device__ __forceinline__ float absmax_wave(const float *__restrict__ part, int np)
{
    float mx = 0.f;
    for (int i = threadIdx.x; i < np; i += kWave)
        mx = fmaxf(mx, part[i]);
#pragma unroll
    for (int o = kWave / 2; o >= 1; o >>= 1)
        mx = fmaxf(mx, __shfl_xor(mx, o, kWave));
    return mx;
}

// A contribution is at most |m| * wamp * 2 (the truncation path of a
// negative time weighs a sample by up to 2): e puts that bound below 2^30.
__device__ __forceinline__ SpreadScale scale_from_max(float mx, double wamp)
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
        sc.fixed = 1;
    }
    return sc;
}

__global__ __launch_bounds__(kBlock) void absmax_partial_kernel(const float *__restrict__ m,
                                                                int64_t n,
                                                                float *__restrict__ part)
{
    const float b = absmax_block(m, n, blockIdx.x, gridDim.x);
    if (threadIdx.x == 0)
        part[blockIdx.x] = b;
}

__global__ __launch_bounds__(kWave) void scale_final_kernel(const float *__restrict__ part, int np,
                                                            double wamp,
                                                            SpreadScale *__restrict__ sc)
{
    const float mx = absmax_wave(part, np);
    if (threadIdx.x == 0)
        *sc = scale_from_max(mx, wamp);
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

// K > 1: the data of K models (model k at d + k*dstride, its image at
// m + k*mstride) share the tables, sample indices, weights and window
// positions; each model has its own staged windows and sum, with the
// single-model operations in the single-model order, so each image is
// bitwise the one K = 1 gives.
template <typename TD, typename TT, bool DYN, bool NEG, int JW, int P, int K = 1>
__device__ __forceinline__ void gather_tiles(const KArgs &a, const TD *__restrict__ d,
                                             TD *__restrict__ m, int L, int64_t dstride = 0,
                                             int64_t mstride = 0)
{
    using TC = typename TimeType<TD>::type;
    constexpr int NT = kGatherWaves / P;  // tiles per workgroup
    constexpr int SB = JW * P;            // sources per block (all parts)
    __shared__ TD win[kGatherWaves][K][JW][kWin];

    const int lane = threadIdx.x & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int part = wave / NT;  // this wave sums sources s0 + part + P i
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
    TD(&w)[K][JW][kWin] = win[wave];

    TD acc[K];
#pragma unroll
    for (int k = 0; k < K; ++k)
        acc[k] = TD(0);
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
            // lane l stages its sample l (a clamped load, flagged in svok when
            // the sample lies inside the trace; outside ones become zero when
            // the round begins, so that the loads are waited for only then);
            // the aperture gate of the round is loaded with it
            TD sv[K][JW];
            bool svok[JW];
            int izn[JW] = {};
#define KX_STAGE(R, BR)                                                                        \
    _Pragma("unroll") for (int i = 0; i < JW; ++i)                                              \
    {                                                                                          \
        const int e = bs[i] + (BR) - kWin / 2 + lane;                                          \
        const int ec = min(max(e, 0), a.nt - 1);                                               \
        _Pragma("unroll") for (int k = 0; k < K; ++k)                                         \
            sv[k][i] = dblk[k * dstride + (unsigned)(soff[i] + (R) * a.nt + ec)];              \
        svok[i] = (unsigned)e < (unsigned)a.nt;                                                \
        if constexpr (DYN)                                                                     \
            izn[i] = a.izmin[(unsigned)(izoff[i] + (R))];                                      \
    }
            // receiver side of the round ahead; receiver bases two rounds
            // ahead (the window addresses of the round ahead depend on them)
            const size_t t0r = (size_t)t * a.nr;
            TT orr_c = offr[t0r * kWave + lane];
            int br_c = a.base_r[t0r + vz];
            int br_n = a.nr > 1 ? a.base_r[t0r + 1 + vz] : 0;
            TD pr_c = TD(0), qr_c = TD(0);
            if constexpr (DYN) {
                pr_c = wpr[t0r * kWave + lane];
                qr_c = wqr[t0r * kWave + lane];
            }
            KX_STAGE(0, br_c)
            for (int r = 0; r < a.nr; ++r) {
                // this round's samples into the wave's LDS region
                int izr[JW];
#pragma unroll
                for (int i = 0; i < JW; ++i) {
#pragma unroll
                    for (int k = 0; k < K; ++k)
                        w[k][i][lane] = svok[i] ? sv[k][i] : TD(0);
                    izr[i] = izn[i];
                }
                const TC orr = (TC)ld_off(&orr_c);
                const int br = br_c;
                const TD pr = pr_c, qr = qr_c;
                // next round's receiver side and samples, in flight meanwhile
                if (r + 1 < a.nr) {
                    const size_t ro = t0r + r + 1;
                    orr_c = offr[ro * kWave + lane];
                    if constexpr (DYN) {
                        pr_c = wpr[ro * kWave + lane];
                        qr_c = wqr[ro * kWave + lane];
                    }
                    br_c = br_n;
                    KX_STAGE(r + 1, br_c)
                    if (r + 2 < a.nr)
                        br_n = a.base_r[ro + 1 + vz];
                }
                // Each lane reads its two samples from the staged windows.  A
                // sample outside its window (rare: the tile's own offsets span
                // a few tens of samples) is marked, and the wave fixes the
                // round's sum afterwards from global memory, so that the
                // common path has no per-source branch.
                unsigned miss = 0;
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
                    const int kk = it - (bs[i] + br - kWin / 2);
                    const int kc = min(max(kk, 0), kWin - 2);
                    const bool in_trace = (unsigned)it < ntm1;
                    miss |= (in_trace && kk != kc) ? (1u << i) : 0u;
                    TD wgt = TD(1);
                    if constexpr (DYN) {
                        wgt = ps[i] * pr + qs[i] * qr;
                        if (p.iz < izr[i])
                            wgt = TD(0);
                    }
#pragma unroll
                    for (int k = 0; k < K; ++k) {
                        const TD q0 = w[k][i][kc];
                        const TD q1 = w[k][i][kc + 1];
                        TD v = fma((TD)f, q1 - q0, q0);
                        if constexpr (DYN)
                            v *= wgt;
                        acc[k] += in_trace ? v : TD(0);
                    }
                }
                if (__ballot(miss != 0) != 0) {
                    // replace the window values of the marked sources by the
                    // samples themselves
#pragma unroll
                    for (int i = 0; i < JW; ++i) {
                        if ((miss >> i) & 1u) {
                            const TC T = os[i] + orr;
                            const TC fl = floor(T);
                            TC f = T - fl;
                            int it = (int)fl + bs[i] + br;
                            if constexpr (NEG) {
                                if (it == -1 && f > TC(0)) {
                                    it = 0;
                                    f -= TC(1);
                                }
                            }
                            const int kk = it - (bs[i] + br - kWin / 2);
                            const int kc = min(max(kk, 0), kWin - 2);
                            TD wgt = TD(1);
                            if constexpr (DYN) {
                                wgt = ps[i] * pr + qs[i] * qr;
                                if (p.iz < izr[i])
                                    wgt = TD(0);
                            }
#pragma unroll
                            for (int k = 0; k < K; ++k) {
                                const TD *q = dblk + k * dstride + (soff[i] + r * a.nt + it);
                                const TD g0 = q[0], g1 = q[1];
                                const TD l0 = w[k][i][kc], l1 = w[k][i][kc + 1];
                                TD dv = fma((TD)f, g1 - g0, g0) - fma((TD)f, l1 - l0, l0);
                                if constexpr (DYN)
                                    dv *= wgt;
                                acc[k] += dv;
                            }
                        }
                    }
                }
            }
#undef KX_STAGE
        }
    }
    if constexpr (P > 1) {
        // the partial sums of a tile's points, added in part order
        __syncthreads();
        TD *red = &win[0][0][0][0];
#pragma unroll
        for (int k = 0; k < K; ++k)
            red[k * kGatherBlock + threadIdx.x] = acc[k];
        __syncthreads();
        if (part == 0) {
#pragma unroll
            for (int k = 0; k < K; ++k) {
                acc[k] = TD(0);
#pragma unroll
                for (int q = 0; q < P; ++q)
                    acc[k] += red[k * kGatherBlock + (q * NT + wave) * kWave + lane];
            }
        }
    }
    if (part == 0 && active && p.valid) {
        const int64_t pi = point_index(a, p);
#pragma unroll
        for (int k = 0; k < K; ++k)
            m[k * mstride + pi] = acc[k];
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
constexpr int kConvThreads = 128;
constexpr int kConvPer = 4;
constexpr int kConvOut = kConvThreads * kConvPer;

// filter length padded to a whole number of 4-element groups
__host__ __device__ __forceinline__ int conv_taps_padded(int nh) { return (nh + 3) & ~3; }

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
