// kx_gather.hip -- the adjoint (gather) launches of the Kirchhoff operator
// and their C ABI entry points kx_gather / kx_gather_models
// (include/kirchx.h).  Kernel bodies: kx_device.h.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "kx_internal.h"
#include "kx_device.h"

namespace {

// K models (K > 1: model k's traces at d + k*dstride, its image at
// m + k*mstride); groups of NP tiles per wave; GATE: apply the per-trace
// aperture gate (dynamic plans whose gates are all 0 skip it)
template <typename TD, typename TT, bool DYN, bool NEG, int JW, int P, int NP, int K,
          bool GATE = true>
__global__ __launch_bounds__(kGatherBlock) void gather_kernel(KArgs a, const TD *__restrict__ d,
                                                              TD *__restrict__ m,
                                                              int64_t dstride, int64_t mstride)
{
    gather_tiles<TD, TT, DYN, NEG, JW, P, NP, K, GATE>(a, d, m, xcd_remap(blockIdx.x, gridDim.x),
                                                       dstride, mstride);
}

// sources per wave and pass (measured): every pass re-reads the receiver
// table -- three tables when dynamic -- and more sources hold more
// registers, lowering occupancy
template <bool DYN, int NP, int K> struct GatherJW {
    static constexpr int v = (DYN && NP == 1 && K == 1) ? 8 : 4;
};

template <typename TD, typename TT, bool DYN, bool NEG, int P, int NP, int K, bool GATE = true>
int launch_gather_p(KArgs a, const void *d, int64_t dstride, void *m, int64_t mstride,
                    hipStream_t st)
{
    constexpr int JW = GatherJW<DYN, NP, K>::v;
    constexpr int NG = kGatherWaves / P;
    if ((int64_t)JW * P * a.nr * a.nt >= ((int64_t)1 << 31))
        return kx_failf(KX_ESHAPE, "gather: %d sources x nr x nt exceeds 2^31 samples", JW * P);
    set_tile_groups(a, NP);
    const int nblk = (a.ngroups + NG - 1) / NG;
    hipLaunchKernelGGL((gather_kernel<TD, TT, DYN, NEG, JW, P, NP, K, GATE>), dim3(nblk),
                       dim3(kGatherBlock), 0, st, a, reinterpret_cast<const TD *>(d),
                       reinterpret_cast<TD *>(m), dstride, mstride);
    KX_LAUNCHED("gather");
    return KX_OK;
}

// Launch shape of the adjoint: P waves per tile group (each sums a share of
// the sources) and NP tiles per wave (sharing the staged windows).  Four
// waves per group read one group's receiver rows together; fewer when the
// sources do not fill their passes.  NP: the most tiles per wave that still
// give every CU kGatherFill waves (measured: 2D 601x300 dynamic 2 x 4,
// 2.5 ms; 3D 128x128x120 4 x 4, 8.8 ms).
constexpr int kGatherFill = 20;

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

int gather_jw(bool dyn, int np, int k) { return (dyn && np == 1 && k == 1) ? 8 : 4; }

void gather_shape(const KArgs &a, bool dyn, int k, int gsplit, int gtiles, int *np_out,
                  int *p_out)
{
    auto split_for = [&](int np) {
        int P = 1;
        if (gsplit > 0) {
            while (P * 2 <= std::min(gsplit, kGatherWaves))
                P *= 2;
            return P;
        }
        const int jw = gather_jw(dyn, np, k);
        P = kGatherWaves;
        while (P > 1 && (int64_t)jw * (P / 2) >= a.ns)
            P /= 2;
        return P;
    };
    int np = 1;
    if (gtiles > 0) {
        np = gtiles >= 4 ? 4 : (gtiles >= 2 ? 2 : 1);
    } else {
        const int64_t want = (int64_t)kGatherFill * device_cus();
        for (int c : {4, 2}) {
            KArgs b = a;
            set_tile_groups(b, c);
            if ((int64_t)b.ngroups * split_for(c) >= want) {
                np = c;
                break;
            }
        }
    }
    *np_out = np;
    *p_out = split_for(np);
}

template <typename TD, typename TT, bool DYN, bool NEG, int NP, int K, bool GATE = true>
int launch_gather_np(const KArgs &a, int P, const void *d, void *m, hipStream_t st,
                     int64_t dstride, int64_t mstride)
{
    switch (P) {
    case 1:
        return launch_gather_p<TD, TT, DYN, NEG, 1, NP, K, GATE>(a, d, dstride, m, mstride, st);
    case 2:
        return launch_gather_p<TD, TT, DYN, NEG, 2, NP, K, GATE>(a, d, dstride, m, mstride, st);
    default:
        return launch_gather_p<TD, TT, DYN, NEG, 4, NP, K, GATE>(a, d, dstride, m, mstride, st);
    }
}

template <typename TD, typename TT, bool DYN, bool NEG, int K = 1, bool GATE = true>
int launch_gather_t(const KArgs &a, int gsplit, int gtiles, const void *d, void *m,
                    hipStream_t st, int64_t dstride = 0, int64_t mstride = 0)
{
    int np = 1, P = 1;
    gather_shape(a, DYN, K, gsplit, gtiles, &np, &P);
    switch (np) {
    case 4:
        return launch_gather_np<TD, TT, DYN, NEG, 4, K, GATE>(a, P, d, m, st, dstride, mstride);
    case 2:
        return launch_gather_np<TD, TT, DYN, NEG, 2, K, GATE>(a, P, d, m, st, dstride, mstride);
    default:
        return launch_gather_np<TD, TT, DYN, NEG, 1, K, GATE>(a, P, d, m, st, dstride, mstride);
    }
}

template <typename TD, typename TT>
int launch_gather_d(const kx_plan *p, const KArgs &a, const void *d, void *m, hipStream_t st)
{
    const int g = p->gsplit, n = p->gtiles;
    // dynamic == KX_DYNAMIC_NOGATE: every aperture gate is 0, and float32
    // operators run a gather without the per-sample gate test (2 VALU
    // instructions of about 20 per sample, and the gate loads); the image
    // is bitwise that of the gated kernel
    if constexpr (sizeof(TD) == 4)
        if (p->dynamic == KX_DYNAMIC_NOGATE)
            return p->neg_trav
                       ? launch_gather_t<TD, TT, true, true, 1, false>(a, g, n, d, m, st)
                       : launch_gather_t<TD, TT, true, false, 1, false>(a, g, n, d, m, st);
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

extern "C" int kx_gather(const kx_plan *p, const void *traces, void *m, void *stream)
{
    int rc = kx_plan_check(p);
    if (rc)
        return rc;
    if (!m || !traces)
        return kx_failf(KX_EINVAL, "gather: NULL buffer");
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
    int rc = kx_plan_check(p);
    if (rc)
        return rc;
    if (!m || !traces)
        return kx_failf(KX_EINVAL, "gather_models: NULL buffer");
    if (p->dtype != KX_F32)
        return kx_failf(KX_ETYPE, "gather_models: float32 operators only");
    if (nmodel < 1)
        return kx_failf(KX_ESHAPE, "gather_models: nmodel=%d", nmodel);
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
