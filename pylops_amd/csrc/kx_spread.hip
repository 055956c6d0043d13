// kx_spread.hip -- the forward (spread) launches of the Kirchhoff operator
// and their C ABI entry points kx_spread / kx_spread_models
// (include/kirchx.h).  Kernel bodies: kx_device.h.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <mutex>

#include "kx_internal.h"
#include "kx_device.h"

namespace {

// h != NULL: y receives the traces convolved with the nh-tap wavelet h
// (the forward operator), otherwise the traces.
template <typename TD, typename TT, bool DYN, bool NEG, int ORD, int RB, bool XS>
__global__ __launch_bounds__(kSpreadBlock) void spread_kernel(KArgs a, const TD *__restrict__ m,
                                                              TD *__restrict__ y,
                                                              const SpreadScale *__restrict__ sc,
                                                              const TD *__restrict__ h, int nh)
{
    const int L = xcd_remap(blockIdx.x, gridDim.x);
    const SpreadScale scale = *sc;
    spread_item<TD, TT, DYN, NEG, ORD, RB, XS>(a, m, y, scale, L % a.ns, (L / a.ns) * RB, h, nh);
}

// Traces longer than a workgroup's LDS: workgroup L (after the XCD
// renumbering) = segment L mod nseg of the trace of source (L / nseg) mod
// ns and receiver L / (nseg ns), samples [seg tseg, seg tseg + tseg); the
// segments of one trace are neighbours and share its table rows in L2.
template <typename TD, typename TT, bool DYN, bool NEG>
__global__ __launch_bounds__(kSpreadBlock) void spread_seg_kernel(KArgs a,
                                                                  const TD *__restrict__ m,
                                                                  TD *__restrict__ y,
                                                                  const SpreadScale *__restrict__ sc,
                                                                  int tseg, int nseg)
{
    const int L = xcd_remap(blockIdx.x, gridDim.x);
    const int seg = L % nseg, sr = L / nseg;
    a.t0 = seg * tseg;
    a.tlen = min(tseg, a.nt - a.t0);
    const SpreadScale scale = *sc;
    spread_item<TD, TT, DYN, NEG, kOrdNatural, 1, false, true>(a, m, y, scale, sr % a.ns,
                                                               sr / a.ns);
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

constexpr int kRblocks[] = {16, 8, 4, 2, 1};
constexpr int kLdsPerCu = 160 * 1024;
constexpr int kMaxWavesPerCu = 32;

// LDS of a workgroup's traces: 8 bytes per sample, rows trace_stride(nt)
// samples apart (the models kernel keeps rows of nt samples)
size_t spread_lds_bytes(int rb, int nt) { return (size_t)rb * trace_stride(nt) * 8; }
size_t models_lds_bytes(int k, int nt) { return (size_t)k * nt * 8; }

template <typename TD, typename TT, bool DYN, bool NEG, bool XS, int ORD, int RB>
int launch_spread_rb(const KArgs &a, const void *m, void *y, const SpreadScale *sc,
                     hipStream_t st, const void *h, int nh)
{
    const size_t lds = spread_lds_bytes(RB, a.nt);
    if (h != nullptr && (sizeof(TD) != 4 || !fused_conv_ok(RB, a.nt, nh, lds, kSpreadBlock)))
        return KX_ENOFUSE;
    const int nrb = (a.nr + RB - 1) / RB;
    const int nblk = a.ns * nrb;
    allow_lds(reinterpret_cast<const void *>(&spread_kernel<TD, TT, DYN, NEG, ORD, RB, XS>),
              lds);
    hipLaunchKernelGGL((spread_kernel<TD, TT, DYN, NEG, ORD, RB, XS>), dim3(nblk),
                       dim3(kSpreadBlock), lds, st, a, reinterpret_cast<const TD *>(m),
                       reinterpret_cast<TD *>(y), sc, reinterpret_cast<const TD *>(h), nh);
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
        return reinterpret_cast<const void *>(
            &spread_kernel<TD, TT, DYN, NEG, kOrdNatural, 1, XS>);
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
        const int per_cu =
            spread_per_cu(spread_fn<TD, TT, DYN, NEG, XS, ORD>(rb), rb, nt);
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
                    const SpreadScale *sc, hipStream_t st, const void *h, int nh)
{
    a.rblock = pick_rblock<TD, TT, DYN, NEG, XS, ORD>(requested, a.nr, a.nt, max_rb);
    switch (a.rblock) {
    case 16:
        return launch_spread_rb<TD, TT, DYN, NEG, XS, ORD, 16>(a, m, y, sc, st, h, nh);
    case 8:
        return launch_spread_rb<TD, TT, DYN, NEG, XS, ORD, 8>(a, m, y, sc, st, h, nh);
    case 4:
        return launch_spread_rb<TD, TT, DYN, NEG, XS, ORD, 4>(a, m, y, sc, st, h, nh);
    case 2:
        return launch_spread_rb<TD, TT, DYN, NEG, XS, ORD, 2>(a, m, y, sc, st, h, nh);
    default:
        return launch_spread_rb<TD, TT, DYN, NEG, XS, kOrdNatural, 1>(a, m, y, sc, st, h, nh);
    }
}

template <typename TD, typename TT, bool DYN, bool NEG, bool XS>
int launch_spread_t(const KArgs &a, int order, int requested, int max_rb, const void *m, void *y,
                    const SpreadScale *sc, hipStream_t st, const void *h, int nh)
{
    switch (order) {
    case KX_ORDER_ROTATED:
        return launch_spread_o<TD, TT, DYN, NEG, XS, kOrdRotated>(a, requested, max_rb, m, y,
                                                                  sc, st, h, nh);
    default:
        return launch_spread_o<TD, TT, DYN, NEG, XS, kOrdNatural>(a, requested, max_rb, m, y,
                                                                  sc, st, h, nh);
    }
}

// h != NULL: the spread and the wavelet convolution in one launch (y
// receives the data), or KX_ENOFUSE when this plan's launch cannot
// convolve in LDS (float64 data, too little LDS).
// Segments of traces that do not fit a workgroup's LDS: the longest
// multiple of 64 samples that does, or 0 if not even 64 fit.
int trace_segment(int nt)
{
    int t = (int)std::min<size_t>((size_t)nt, lds_budget() / 8) & ~63;
    while (t > 0 && spread_lds_bytes(1, t) > lds_budget())
        t -= 64;
    return t;
}

template <typename TD, typename TT, bool DYN, bool NEG>
int launch_spread_seg(const KArgs &a, const void *m, void *y, const SpreadScale *sc,
                      hipStream_t st)
{
    const int tseg = trace_segment(a.nt);
    if (tseg <= 0)
        return kx_failf(KX_ELDS, "nt=%d: not even 64 samples of a trace fit in LDS", a.nt);
    const int nseg = (a.nt + tseg - 1) / tseg;
    const int64_t nblk = (int64_t)a.ns * a.nr * nseg;
    if (nblk > INT32_MAX)
        return kx_failf(KX_ESHAPE, "spread: %lld trace segments exceed the grid",
                        (long long)nblk);
    const size_t lds = spread_lds_bytes(1, tseg);
    allow_lds(reinterpret_cast<const void *>(&spread_seg_kernel<TD, TT, DYN, NEG>), lds);
    hipLaunchKernelGGL((spread_seg_kernel<TD, TT, DYN, NEG>), dim3((unsigned)nblk),
                       dim3(kSpreadBlock), lds, st, a, reinterpret_cast<const TD *>(m),
                       reinterpret_cast<TD *>(y), sc, tseg, nseg);
    KX_LAUNCHED("spread segments");
    return KX_OK;
}

template <typename TD, typename TT>
int launch_spread_d(const kx_plan *p, KArgs a, const void *m, void *y, hipStream_t st,
                    const void *h = nullptr)
{
    const int nh = h != nullptr ? p->nh : 0;
    if (h != nullptr && sizeof(TD) != 4)
        return KX_ENOFUSE;
    // a trace longer than LDS: spread in time segments (the wavelet then
    // convolves the traces in HBM)
    const bool seg = spread_lds_bytes(1, p->nt) > lds_budget();
    if (seg && h != nullptr)
        return KX_ENOFUSE;
    if (sizeof(TD) == 4 && !p->aux)
        return kx_failf(KX_EINVAL, "spread: plan.aux (KX_AUX_BYTES of device scratch) is NULL");
    // registers per lane grow with the block: keep two workgroups per CU
    const int max_rb = sizeof(TD) == 8 ? 4 : (p->dynamic ? 8 : 16);
    SpreadScale *sc = reinterpret_cast<SpreadScale *>(p->aux);
    int side = p->sside;
    if (side == KX_SIDE_AUTO)
        side = (p->ny == 1 && p->ns < p->nr) ? KX_SIDE_SOURCES : KX_SIDE_RECEIVERS;
    // (float16 plans keep the two sides' offsets in different types and
    // block receivers)
    const bool xs = sizeof(TD) == 4 && side == KX_SIDE_SOURCES && p->tab_fmt != KX_F16;
    if (sizeof(TD) == 4) {
        // fixed-point scale of this launch from max|m|
        const int64_t ni = (int64_t)p->nx * p->ny * p->nz;
        float *part = reinterpret_cast<float *>(p->aux) + 16;
        const int nb = (int)std::min<int64_t>(kScaleBlocks, (ni + kBlock - 1) / kBlock);
        hipLaunchKernelGGL(absmax_partial_kernel, dim3(nb), dim3(kBlock), 0, st,
                           reinterpret_cast<const float *>(m), ni, part);
        KX_LAUNCHED("spread scale");
        hipLaunchKernelGGL(scale_final_kernel, dim3(1), dim3(kWave), 0, st, part, nb, p->wamp,
                           p->wtyp, sc);
        KX_LAUNCHED("spread scale");
    }
    // A workgroup blocking sources (float32 operators) runs the spread with
    // the two sides' tables exchanged and writes its traces to the
    // transposed positions: T = os + or and the integer bases add in either
    // order, and the dynamic kernel (XS) forms the source-side products per
    // step, so the traces are bitwise those of blocks of receivers.
    // Automatic: a 2D plan blocks the side with fewer members, whose rows
    // then serve many workgroups through L2 (on a 61 x 301 geometry 13%
    // faster kinematic, either way round, and 6% with dynamic weights: 3.68
    // against 3.91 ms for the forward with its wavelet); 3D plans block
    // receivers (measured slower otherwise: 128x128x120 with 25 x 400, 24.2
    // against 22.8 ms for the natural order, 17.4 against 15.8 ms rotated).
    if (seg)
        return p->dynamic ? (p->neg_trav ? launch_spread_seg<TD, TT, true, true>(a, m, y, sc, st)
                                         : launch_spread_seg<TD, TT, true, false>(a, m, y, sc, st))
                          : (p->neg_trav ? launch_spread_seg<TD, TT, false, true>(a, m, y, sc, st)
                                         : launch_spread_seg<TD, TT, false, false>(a, m, y, sc,
                                                                                   st));
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
                                                                      m, y, sc, st, h, nh)
                           : launch_spread_t<TD, TT, true, false, kXS>(a, order, p->rblock,
                                                                       max_rb, m, y, sc, st, h, nh);
    if (p->dynamic)
        return p->neg_trav ? launch_spread_t<TD, TT, true, true, false>(a, order, p->rblock,
                                                                        max_rb, m, y, sc, st, h, nh)
                           : launch_spread_t<TD, TT, true, false, false>(a, order, p->rblock,
                                                                         max_rb, m, y, sc, st, h, nh);
    return p->neg_trav ? launch_spread_t<TD, TT, false, true, false>(a, order, p->rblock, max_rb,
                                                                     m, y, sc, st, h, nh)
                       : launch_spread_t<TD, TT, false, false, false>(a, order, p->rblock,
                                                                      max_rb, m, y, sc, st, h, nh);
}

}  // namespace

extern "C" int kx_spread(const kx_plan *p, const void *m, void *traces, void *stream)
{
    int rc = kx_plan_check(p);
    if (rc)
        return rc;
    if (!m || !traces)
        return kx_failf(KX_EINVAL, "spread: NULL buffer");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const KArgs a = make_args(p);
    if (p->dtype == KX_F64)
        return launch_spread_d<double, double>(p, a, m, traces, st);
    if (p->tab_fmt == KX_F16)
        return launch_spread_d<float, __half>(p, a, m, traces, st);
    return launch_spread_d<float, float>(p, a, m, traces, st);
}

extern "C" __attribute__((visibility("hidden"))) int kx_spread_wavelet(const kx_plan *p,
                                                                      const void *m, void *d,
                                                                      void *stream)
{
    int rc = kx_plan_check(p);
    if (rc)
        return rc;
    if (!m || !d || !p->h || p->nh <= 0)
        return kx_failf(KX_EINVAL, "spread_wavelet: NULL buffer or wavelet");
    if (p->dtype != KX_F32)
        return KX_ENOFUSE;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const KArgs a = make_args(p);
    if (p->tab_fmt == KX_F16)
        return launch_spread_d<float, __half>(p, a, m, d, st, p->h);
    return launch_spread_d<float, float>(p, a, m, d, st, p->h);
}

namespace {

constexpr int kMaxModels = 64;
constexpr size_t kModelScaleOffset = 2048;  // bytes into plan.aux

template <typename TT, bool DYN, bool NEG, int K>
void launch_models_k(const KArgs &a, const float *m, int64_t ni, float *y, int64_t ystride,
                     const SpreadScale *sc, hipStream_t st)
{
    const size_t lds = models_lds_bytes(K, a.nt);
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
                           p->wtyp, scs + j);
    }
    KX_LAUNCHED("spread models scale");
    int j = 0;
    while (j < nmodel) {
        const int left = nmodel - j;
        auto fits = [&](int k) {
            return k <= left && models_lds_bytes(k, a.nt) <= (size_t)kLdsBudget;
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
    int rc = kx_plan_check(p);
    if (rc)
        return rc;
    if (!m || !traces)
        return kx_failf(KX_EINVAL, "spread_models: NULL buffer");
    if (p->dtype != KX_F32)
        return kx_failf(KX_ETYPE, "spread_models: float32 operators only");
    if (nmodel < 1 || nmodel > kMaxModels)
        return kx_failf(KX_ESHAPE, "spread_models: nmodel=%d outside [1, %d]", nmodel, kMaxModels);
    if (!p->aux)
        return kx_failf(KX_EINVAL, "spread_models: plan.aux (KX_AUX_BYTES of device scratch) is NULL");
    // (traces of a few models per workgroup that do not fit LDS fall back to
    // single spreads, in time segments if need be)
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

// ---------------------------------------------------------------------------
// the forward operator without caller memory
// ---------------------------------------------------------------------------
// The one-launch forward keeps only its accumulation scale and the partial
// maxima of |m| in scratch memory (plan.aux for kx_forward).  For callers
// that cannot allocate, that scratch is this device array: every GPU the
// library runs on holds its own copy (KX_AUX_BYTES) until the process ends.
static __device__ __align__(256) unsigned char scratch[KX_AUX_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// The address of scratch on the current device, looked up once per device
// (hipGetSymbolAddress with that device current).
int scratch_address(void **out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess)
        return kx_failf((int)e, "forward_noalloc: hipGetDevice: %s", hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxDevices)
        return kx_failf(KX_EINVAL, "forward_noalloc: device %d beyond the %d supported", dev,
                        kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[dev] == nullptr) {
        void *ptr = nullptr;
        e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess)
            return kx_failf((int)e, "forward_noalloc: hipGetSymbolAddress(scratch): %s",
                            hipGetErrorString(e));
        if (ptr == nullptr)
            return kx_failf(KX_EINVAL, "forward_noalloc: scratch has no device address");
        g_scratch_ptr[dev] = ptr;
    }
    *out = g_scratch_ptr[dev];
    return KX_OK;
}

}  // namespace

extern "C" int kx_forward_noalloc(const kx_plan *p, const void *m, void *d, void *stream)
{
    int rc = kx_plan_check(p);
    if (rc)
        return rc;
    if (!m || !d || !p->h || p->nh <= 0)
        return kx_failf(KX_EINVAL, "forward_noalloc: NULL buffer or wavelet");
    if (p->dtype != KX_F32)
        return kx_failf(KX_ETYPE, "forward_noalloc: float32 plans only (a float64 forward "
                                  "needs kx_forward and its workspace)");
    void *s = nullptr;
    rc = scratch_address(&s);
    if (rc)
        return rc;
    // the plan as given, its scale kept in this device's scratch (every
    // launch writes the scale before the spread reads it: nothing needs
    // clearing)
    kx_plan q = *p;
    q.aux = s;
    rc = kx_spread_wavelet(&q, m, d, stream);
    if (rc == KX_ENOFUSE)
        return kx_failf(KX_ELDS,
                        "forward_noalloc: this plan's forward cannot convolve in the spread "
                        "launch (nt=%d traces with %d wavelet taps do not fit a workgroup's "
                        "LDS, or float32 atomics were requested); use kx_forward and its "
                        "workspace",
                        p->nt, p->nh);
    return rc;
}
