// kx_batch.hip -- batches of independent Kirchhoff problems (include/kirchx.h).
//
// Replaces the reference's one-problem-at-a-time use of Kirchhoff._matvec /
// _rmatvec (pylops/waveeqprocessing/kirchhoff.py:1048-1143) when many
// problems of uneven sizes are inverted together.  Every operator application
// of a batch is one persistent launch per engine kernel: a grid sized to fill
// the chip pulls work items -- (problem, source, receiver block) for the
// spread, (problem, tile group) for the gather -- from atomic counters, so
// that small and large problems share the CUs without a launch per problem.
// The item bodies are the single-problem ones (kx_device.h), so a problem's
// results do not depend on the batch it is in.
//
// Locality: each XCD has its own L2.  Items that read the same table rows --
// one receiver block of a problem (spread), a run of neighbouring tiles
// (gather), and the same items of problems sharing their tables (the columns
// of a multi-column product) -- form a group; groups are dealt to eight
// per-XCD queues (largest first, to the least loaded queue), and workgroup b,
// which the hardware places on XCD b mod 8, drains queue b mod 8 before
// taking items from the others.
//
// Device memory (caller-owned, kx_batch_dev_bytes), 256-byte aligned parts:
//   BatchProb   probs[nprob]       geometry, tables, offsets into m and d
//   SpreadScale scales[nprob]      fixed-point scale of each problem's forward
//   float       part[nprob][3][kBatchScaleBlocks]  partial max|m|, sum |m|,
//                                  count of nonzero m
//   unsigned    ctr[kCounters]     per-XCD work counters and queue offsets
//   SpreadItem  sitems[n_spread]   the eight spread queues, one after the other
//   GatherItem  gitems[n_gather]   the eight gather queues
//   int64_t     cprefix[nprob + 1] convolution blocks before each problem
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <map>
#include <numeric>
#include <utility>
#include <vector>

#include "kirchx.h"

extern "C" int kx_solver_fail(int code, const char *msg);
extern "C" int kx_plan_check(const kx_plan *p);

#include "kx_device.h"

namespace {

int bfail(int code, const char *fmt, ...)
{
    char buf[480];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    return kx_solver_fail(code, buf);
}

#define KB_HIP(call)                                                                       \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess)                                                              \
            return bfail((int)e_, "%s: %s", #call, hipGetErrorString(e_));                 \
    } while (0)

#define KB_LAUNCHED(name)                                                                  \
    do {                                                                                   \
        hipError_t e_ = hipGetLastError();                                                 \
        if (e_ != hipSuccess)                                                              \
            return bfail((int)e_, "%s launch: %s", name, hipGetErrorString(e_));           \
    } while (0)

constexpr int kBatchScaleBlocks = 32;
constexpr int kXcds = 8;
// counters: [0, 8) spread next per queue, [8] spread done, [16, 24) gather
// next, [24] gather done; queue offsets: [32, 41) spread, [48, 57) gather
constexpr int kCounters = 64;
constexpr int kSpreadQ = 0, kGatherQ = 16, kQoff = 32;
constexpr int kGatherFill = 24;
constexpr int kLdsPerCu = 160 * 1024;
constexpr int kMaxWavesPerCu = 32;
constexpr int kRblocks[] = {16, 8, 4, 2, 1};

struct BatchProb {
    KArgs a;
    int64_t moff, doff;  // first element of the problem's model and data
    int64_t ni;          // image points
    const void *h, *hstar;
    int nh;
    int nseg;            // convolution blocks per trace
    double wamp, wtyp;
};

struct SpreadItem {
    int p, s, r0, pad;
};

struct GatherItem {
    int p, L;
};

// ---------------------------------------------------------------------------
// persistent work loop
// ---------------------------------------------------------------------------
// Next item of the launch, or -1 when every queue is empty.  Thread 0 takes
// it from the current queue's counter, moving on to the next queue when that
// one is exhausted (q and left live in thread 0 only); the block reads the
// item from LDS.  ctr: the launch's counters (next[8], done), qoff its
// queue offsets.
__device__ __forceinline__ int next_item(unsigned *ctr, const int *qoff, int *cur, int &q,
                                         int &left)
{
    if (threadIdx.x == 0) {
        int it = -1;
        while (left > 0) {
            const unsigned k = atomicAdd(&ctr[q], 1u);
            if (k < (unsigned)(qoff[q + 1] - qoff[q])) {
                it = qoff[q] + (int)k;
                break;
            }
            q = (q + 1) & (kXcds - 1);
            --left;
        }
        *cur = it;
    }
    __syncthreads();
    const int it = __builtin_amdgcn_readfirstlane(*cur);
    __syncthreads();  // read by every wave before thread 0 takes the next
    return it;
}

// The last workgroup to finish resets the counters for the next launch: by
// then every workgroup has found every queue empty.
__device__ __forceinline__ void launch_done(unsigned *ctr)
{
    if (threadIdx.x == 0) {
        __threadfence();
        if (atomicAdd(&ctr[kXcds], 1u) == gridDim.x - 1) {
            for (int k = 0; k <= kXcds; ++k)
                atomicExch(&ctr[k], 0u);
        }
    }
}

template <typename TD, typename TT, bool DYN, bool NEG, bool ROT, int RB>
__global__ __launch_bounds__(kSpreadBlock) void batch_spread_kernel(
    const BatchProb *__restrict__ probs, const SpreadScale *__restrict__ scales,
    const SpreadItem *__restrict__ items, unsigned *ctr, const int *__restrict__ qoff,
    const TD *__restrict__ m, TD *__restrict__ y)
{
    __shared__ int cur;
    int q = blockIdx.x & (kXcds - 1), left = kXcds;
    for (;;) {
        const int it = next_item(ctr, qoff, &cur, q, left);
        if (it < 0)
            break;
        const SpreadItem w = items[it];
        const KArgs a = probs[w.p].a;
        const SpreadScale sc = scales[w.p];
        spread_item<TD, TT, DYN, NEG, ROT ? kOrdRotated : kOrdNatural, RB>(a, m + probs[w.p].moff, y + probs[w.p].doff, sc,
                                               w.s, w.r0);
    }
    launch_done(ctr);
}

template <typename TD, typename TT, bool DYN, bool NEG, int JW, int P, int NP>
__global__ __launch_bounds__(kGatherBlock) void batch_gather_kernel(
    const BatchProb *__restrict__ probs, const GatherItem *__restrict__ items, unsigned *ctr,
    const int *__restrict__ qoff, const TD *__restrict__ d, TD *__restrict__ m)
{
    __shared__ int cur;
    int q = blockIdx.x & (kXcds - 1), left = kXcds;
    for (;;) {
        const int it = next_item(ctr, qoff, &cur, q, left);
        if (it < 0)
            break;
        const GatherItem w = items[it];
        const KArgs a = probs[w.p].a;
        gather_tiles<TD, TT, DYN, NEG, JW, P, NP>(a, d + probs[w.p].doff, m + probs[w.p].moff,
                                                  w.L);
    }
    launch_done(ctr);
}

// One convolution block (kConvOut samples of one trace) per workgroup; the
// problem is found by bisection of the block prefix.
template <typename TD>
__global__ __launch_bounds__(kConvThreads) void batch_conv_kernel(const BatchProb *__restrict__ probs,
                                                            const int64_t *__restrict__ cprefix,
                                                            int nprob, const TD *__restrict__ x,
                                                            TD *__restrict__ y, int adjoint)
{
    const int64_t b = blockIdx.x;
    int lo = 0, hi = nprob;  // cprefix[lo] <= b < cprefix[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (cprefix[mid] <= b)
            lo = mid;
        else
            hi = mid;
    }
    const BatchProb &P = probs[lo];
    const int64_t k = b - cprefix[lo];
    const int nt = P.a.nt;
    const TD *h = reinterpret_cast<const TD *>(adjoint ? P.hstar : P.h);
    conv_segment<TD>(x + P.doff, nt, y + P.doff, nt, h, 0, P.nh, k / P.nseg,
                     (int)(k % P.nseg) * kConvOut);
}

__global__ __launch_bounds__(kBlock) void batch_absmax_kernel(const BatchProb *__restrict__ probs,
                                                              const float *__restrict__ m,
                                                              float *__restrict__ part)
{
    const int p = blockIdx.y;
    stats_block(m + probs[p].moff, probs[p].ni, blockIdx.x, kBatchScaleBlocks,
                part + (size_t)p * 3 * kBatchScaleBlocks);
}

__global__ __launch_bounds__(kWave) void batch_scale_kernel(const BatchProb *__restrict__ probs,
                                                            const float *__restrict__ part,
                                                            SpreadScale *__restrict__ scales)
{
    const int p = blockIdx.x;
    float mx;
    double mean;
    stats_wave(part + (size_t)p * 3 * kBatchScaleBlocks, kBatchScaleBlocks, mx, mean);
    if (threadIdx.x == 0)
        scales[p] = scale_from_stats(mx, mean, probs[p].wamp, probs[p].wtyp);
}

// ---------------------------------------------------------------------------
// host side: choices, layout, work lists
// ---------------------------------------------------------------------------
size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }
int dsize(int dtype) { return dtype == KX_F64 ? 8 : 4; }

struct Choice {
    int rb, rot, P, np;
};

// Adjoint tile groups of NP tiles over all problems.
int64_t batch_groups(const kx_plan *plans, int nprob, int np)
{
    int64_t g = 0;
    for (int p = 0; p < nprob; ++p) {
        KArgs a = make_args(&plans[p]);
        set_tile_groups(a, np);
        g += a.ngroups;
    }
    return g;
}

struct Layout {
    size_t probs, scales, part, ctr, sitems, gitems, cprefix, total;
    int64_t n_spread, n_gather, n_conv;
};

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

// Largest receiver block per element type: registers per lane grow with the
// block, and two workgroups per CU are kept.
template <typename TD, bool DYN> constexpr int max_rblock()
{
    return sizeof(TD) == 8 ? 4 : (DYN ? 8 : 16);
}

template <typename TD, typename TT, bool DYN, bool NEG, bool ROT, int RB>
const void *bspread_ptr()
{
    if constexpr (RB <= max_rblock<TD, DYN>())
        return reinterpret_cast<const void *>(
            &batch_spread_kernel<TD, TT, DYN, NEG, ROT && (RB > 1), RB>);
    else
        return nullptr;
}

template <typename TD, typename TT, bool DYN, bool NEG, bool ROT>
const void *bspread_fn(int rb)
{
    switch (rb) {
    case 16:
        return bspread_ptr<TD, TT, DYN, NEG, ROT, 16>();
    case 8:
        return bspread_ptr<TD, TT, DYN, NEG, ROT, 8>();
    case 4:
        return bspread_ptr<TD, TT, DYN, NEG, ROT, 4>();
    case 2:
        return bspread_ptr<TD, TT, DYN, NEG, ROT, 2>();
    default:
        return bspread_ptr<TD, TT, DYN, NEG, ROT, 1>();
    }
}

// sources per wave and pass of the batch adjoint, as for one problem
template <bool DYN, int NP> struct BatchJW {
    static constexpr int v = (DYN && NP == 1) ? 8 : 4;
};
int batch_jw(bool dyn, int np) { return (dyn && np == 1) ? 8 : 4; }

template <typename TD, typename TT, bool DYN, bool NEG, int NP>
const void *bgather_fn_np(int P)
{
    constexpr int JW = BatchJW<DYN, NP>::v;
    switch (P) {
    case 1:
        return reinterpret_cast<const void *>(&batch_gather_kernel<TD, TT, DYN, NEG, JW, 1, NP>);
    case 2:
        return reinterpret_cast<const void *>(&batch_gather_kernel<TD, TT, DYN, NEG, JW, 2, NP>);
    default:
        return reinterpret_cast<const void *>(&batch_gather_kernel<TD, TT, DYN, NEG, JW, 4, NP>);
    }
}

template <typename TD, typename TT, bool DYN, bool NEG>
const void *bgather_fn(int P, int np)
{
    return np == 2 ? bgather_fn_np<TD, TT, DYN, NEG, 2>(P) : bgather_fn_np<TD, TT, DYN, NEG, 1>(P);
}

// Resident workgroups of a kernel per CU; without a device, the LDS and
// wave limits alone.
int per_cu(const void *fn, int block, size_t lds)
{
    allow_lds(fn, lds);
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fn, block, lds) != hipSuccess || n <= 0) {
        (void)hipGetLastError();
        n = kMaxWavesPerCu / (block / kWave);
        if (lds > 0)
            n = std::min(n, (int)(kLdsPerCu / (lds + 64)));
    }
    return std::max(n, 1);
}

// rows trace_stride(nt) samples apart; trace_stride is monotone, so the
// largest nt of a batch sizes every problem's rows
size_t spread_lds(int rb, int nt) { return (size_t)rb * trace_stride(nt) * 8; }

// Kernel entry points of a batch, by its element and table types.
struct Fns {
    const void *(*spread)(int rb, bool rot, bool dyn, bool neg);
    const void *(*gather)(int P, int np, bool dyn, bool neg);
};

template <typename TD, typename TT>
const void *spread_of(int rb, bool rot, bool dyn, bool neg)
{
    if (dyn)
        return neg ? (rot ? bspread_fn<TD, TT, true, true, true>(rb)
                          : bspread_fn<TD, TT, true, true, false>(rb))
                   : (rot ? bspread_fn<TD, TT, true, false, true>(rb)
                          : bspread_fn<TD, TT, true, false, false>(rb));
    return neg ? (rot ? bspread_fn<TD, TT, false, true, true>(rb)
                      : bspread_fn<TD, TT, false, true, false>(rb))
               : (rot ? bspread_fn<TD, TT, false, false, true>(rb)
                      : bspread_fn<TD, TT, false, false, false>(rb));
}

template <typename TD, typename TT>
const void *gather_of(int P, int np, bool dyn, bool neg)
{
    if (dyn)
        return neg ? bgather_fn<TD, TT, true, true>(P, np) : bgather_fn<TD, TT, true, false>(P, np);
    return neg ? bgather_fn<TD, TT, false, true>(P, np) : bgather_fn<TD, TT, false, false>(P, np);
}

Fns fns_of(int dtype, int tab_fmt)
{
    if (dtype == KX_F64)
        return {spread_of<double, double>, gather_of<double, double>};
    if (tab_fmt == KX_F16)
        return {spread_of<float, __half>, gather_of<float, __half>};
    return {spread_of<float, float>, gather_of<float, float>};
}

// Deals groups of work items to the eight XCD queues: largest group first,
// each to the least loaded queue; a queue lists its groups in that order and
// each group's items in `within` order.  Returns the item permutation and the
// queue offsets qoff[0..8].
template <typename Within>
std::vector<int> deal(const std::vector<int> &group_of, int ngroups,
                      const std::vector<int64_t> &cost, int qoff[kXcds + 1], Within within)
{
    std::vector<int64_t> gsum(ngroups, 0);
    std::vector<std::vector<int>> members(ngroups);
    for (size_t i = 0; i < group_of.size(); ++i) {
        gsum[group_of[i]] += cost[i];
        members[group_of[i]].push_back((int)i);
    }
    std::vector<int> gorder(ngroups);
    std::iota(gorder.begin(), gorder.end(), 0);
    std::stable_sort(gorder.begin(), gorder.end(),
                     [&](int a, int b) { return gsum[a] > gsum[b]; });
    int64_t load[kXcds] = {0};
    std::vector<int> queue_groups[kXcds];
    for (int g : gorder) {
        int k = 0;
        for (int j = 1; j < kXcds; ++j)
            if (load[j] < load[k])
                k = j;
        load[k] += gsum[g];
        queue_groups[k].push_back(g);
    }
    std::vector<int> perm;
    perm.reserve(group_of.size());
    for (int k = 0; k < kXcds; ++k) {
        qoff[k] = (int)perm.size();
        for (int g : queue_groups[k]) {
            std::vector<int> &m = members[g];
            std::sort(m.begin(), m.end(), within);
            perm.insert(perm.end(), m.begin(), m.end());
        }
    }
    qoff[kXcds] = (int)perm.size();
    return perm;
}

int check_batch(const kx_plan *plans, int nprob)
{
    if (!plans || nprob <= 0)
        return bfail(KX_EINVAL, "batch: no plans");
    const kx_plan &q = plans[0];
    for (int p = 0; p < nprob; ++p) {
        const kx_plan &a = plans[p];
        int rc = kx_plan_check(&a);
        if (rc)
            return rc;
        if (a.dtype != q.dtype || a.tab_fmt != q.tab_fmt || (a.dynamic != 0) != (q.dynamic != 0) ||
            a.neg_trav != q.neg_trav)
            return bfail(KX_EINVAL,
                         "batch: problem %d differs from problem 0 in dtype, table format, "
                         "dynamic or neg_trav", p);
        if (!a.h || !a.hstar || a.nh <= 0 || (a.nh % 2) == 0)
            return bfail(KX_EINVAL, "batch: problem %d has no odd-length wavelet", p);
        if (spread_lds(1, a.nt) > lds_budget())
            return bfail(KX_ELDS, "batch: problem %d: nt=%d does not fit one trace in LDS "
                                  "(the single-problem operator spreads such traces in time "
                                  "segments)", p, a.nt);
        const size_t clds = (size_t)conv_lds_elems(a.nh) * dsize(a.dtype);
        if (clds > (size_t)kLdsBudget)
            return bfail(KX_ELDS, "batch: problem %d: filter of %d taps does not fit LDS", p,
                         a.nh);
    }
    return KX_OK;
}

Choice choose(const kx_plan *plans, int nprob)
{
    const kx_plan &q = plans[0];
    const Fns f = fns_of(q.dtype, q.tab_fmt);
    int nt_max = 0, nr_max = 0;
    int64_t tiles = 0;
    for (int p = 0; p < nprob; ++p) {
        nt_max = std::max(nt_max, plans[p].nt);
        nr_max = std::max(nr_max, plans[p].nr);
        tiles += plans[p].ntiles;
    }
    Choice c;
    int order = q.sorder;
    if (order == KX_ORDER_AUTO)
        order = q.dynamic ? KX_ORDER_NATURAL : KX_ORDER_ROTATED;
    c.rot = order == KX_ORDER_ROTATED;
    // receivers per workgroup: as for one problem, on the batch's longest trace
    const int max_rb = q.dtype == KX_F64 ? max_rblock<double, false>()
                                         : (q.dynamic ? max_rblock<float, true>()
                                                      : max_rblock<float, false>());
    auto fits = [&](int rb) {
        return rb <= max_rb && spread_lds(rb, nt_max) <= lds_budget();
    };
    c.rb = 1;
    if (q.rblock > 0) {
        for (int rb : kRblocks)
            if (rb <= q.rblock && fits(rb)) {
                c.rb = rb;
                break;
            }
    } else {
        double best = -1.0;
        for (int rb : kRblocks) {
            if (!fits(rb) || (rb > 1 && rb / 2 >= nr_max))
                continue;
            const int n = per_cu(f.spread(rb, c.rot, q.dynamic, q.neg_trav), kSpreadBlock,
                                 spread_lds(rb, nt_max));
            const double waves = std::min(24.0, (double)n * kSpreadWaves);
            const double score = waves / 24.0 * n * std::min(rb, nr_max);
            if (score > best * 1.0001) {
                best = score;
                c.rb = rb;
            }
        }
    }
    if (c.rb == 1)
        c.rot = 0;
    // adjoint tiles per wave (1 or 2) and waves per tile group: as for one
    // problem, on the batch's group count
    const int64_t fill = (int64_t)kGatherFill * device_cus();
    if (q.gtiles > 0)
        c.np = q.gtiles >= 2 ? 2 : 1;
    else
        c.np = batch_groups(plans, nprob, 2) * kGatherWaves >= fill ? 2 : 1;
    const int64_t groups = c.np == 1 ? tiles : batch_groups(plans, nprob, c.np);
    c.P = 1;
    if (q.gsplit > 0) {
        while (c.P * 2 <= std::min(q.gsplit, kGatherWaves))
            c.P *= 2;
    } else {
        while (c.P < kGatherWaves && groups * c.P < fill)
            c.P *= 2;
    }
    return c;
}

// Offsets of the parts of the device buffer.
Layout place(int nprob, int64_t n_spread, int64_t n_gather, int64_t n_conv)
{
    Layout L;
    L.n_spread = n_spread;
    L.n_gather = n_gather;
    L.n_conv = n_conv;
    size_t o = 0;
    L.probs = o;
    o = align256(o + sizeof(BatchProb) * nprob);
    L.scales = o;
    o = align256(o + sizeof(SpreadScale) * nprob);
    L.part = o;
    o = align256(o + sizeof(float) * 3 * kBatchScaleBlocks * nprob);
    L.ctr = o;
    o = align256(o + sizeof(unsigned) * kCounters);
    L.sitems = o;
    o = align256(o + sizeof(SpreadItem) * n_spread);
    L.gitems = o;
    o = align256(o + sizeof(GatherItem) * n_gather);
    L.cprefix = o;
    o = align256(o + sizeof(int64_t) * (nprob + 1));
    L.total = o;
    return L;
}

Layout layout(const kx_plan *plans, int nprob, const Choice &c)
{
    int64_t ns = 0, ng = 0, nc = 0;
    const int NT = kGatherWaves / c.P;
    for (int p = 0; p < nprob; ++p) {
        const kx_plan &a = plans[p];
        KArgs ka = make_args(&a);
        set_tile_groups(ka, c.np);
        ns += (int64_t)a.ns * ((a.nr + c.rb - 1) / c.rb);
        ng += (ka.ngroups + NT - 1) / NT;
        nc += (int64_t)a.ns * a.nr * ((a.nt + kConvOut - 1) / kConvOut);
    }
    return place(nprob, ns, ng, nc);
}

struct View {
    const BatchProb *probs;
    SpreadScale *scales;
    float *part;
    unsigned *ctr;
    const int *ctr_const;  // the same words, read as queue offsets
    const SpreadItem *sitems;
    const GatherItem *gitems;
    const int64_t *cprefix;
};

View view(const kx_batch *b)
{
    const Layout L = place(b->nprob, b->n_spread, b->n_gather, b->n_conv);
    char *base = reinterpret_cast<char *>(b->dev);
    View v;
    v.probs = reinterpret_cast<const BatchProb *>(base + L.probs);
    v.scales = reinterpret_cast<SpreadScale *>(base + L.scales);
    v.part = reinterpret_cast<float *>(base + L.part);
    v.ctr = reinterpret_cast<unsigned *>(base + L.ctr);
    v.ctr_const = reinterpret_cast<const int *>(base + L.ctr);
    v.sitems = reinterpret_cast<const SpreadItem *>(base + L.sitems);
    v.gitems = reinterpret_cast<const GatherItem *>(base + L.gitems);
    v.cprefix = reinterpret_cast<const int64_t *>(base + L.cprefix);
    return v;
}

int check_built(const kx_batch *b)
{
    if (!b || b->abi_version != KX_ABI_VERSION || b->nprob <= 0 || !b->dev)
        return bfail(KX_EINVAL, "batch: not built");
    return KX_OK;
}

// ---------------------------------------------------------------------------
// launches
// ---------------------------------------------------------------------------
template <typename TD, typename TT, bool DYN, bool NEG, bool ROT, int RB>
int launch_bspread(const kx_batch *b, const View &v, const void *m, void *y, hipStream_t st)
{
    if constexpr (RB <= max_rblock<TD, DYN>()) {
        allow_lds(reinterpret_cast<const void *>(
                      &batch_spread_kernel<TD, TT, DYN, NEG, ROT && (RB > 1), RB>),
                  spread_lds(RB, b->nt_max));
        hipLaunchKernelGGL((batch_spread_kernel<TD, TT, DYN, NEG, ROT && (RB > 1), RB>),
                           dim3(b->grid_spread), dim3(kSpreadBlock), spread_lds(RB, b->nt_max),
                           st, v.probs, v.scales, v.sitems, v.ctr + kSpreadQ,
                           v.ctr_const + kQoff, reinterpret_cast<const TD *>(m),
                           reinterpret_cast<TD *>(y));
        KB_LAUNCHED("batch spread");
        return KX_OK;
    } else {
        return bfail(KX_EINVAL, "batch: receiver block %d not supported for this type", RB);
    }
}

template <typename TD, typename TT, bool DYN, bool NEG, bool ROT>
int launch_bspread_rb(const kx_batch *b, const View &v, const void *m, void *y, hipStream_t st)
{
    switch (b->rblock) {
    case 16:
        return launch_bspread<TD, TT, DYN, NEG, ROT, 16>(b, v, m, y, st);
    case 8:
        return launch_bspread<TD, TT, DYN, NEG, ROT, 8>(b, v, m, y, st);
    case 4:
        return launch_bspread<TD, TT, DYN, NEG, ROT, 4>(b, v, m, y, st);
    case 2:
        return launch_bspread<TD, TT, DYN, NEG, ROT, 2>(b, v, m, y, st);
    default:
        return launch_bspread<TD, TT, DYN, NEG, ROT, 1>(b, v, m, y, st);
    }
}

template <typename TD, typename TT>
int launch_bspread_d(const kx_batch *b, const View &v, const void *m, void *y, hipStream_t st)
{
    const bool rot = b->sorder == KX_ORDER_ROTATED;
    if (b->dynamic) {
        if (b->neg_trav)
            return rot ? launch_bspread_rb<TD, TT, true, true, true>(b, v, m, y, st)
                       : launch_bspread_rb<TD, TT, true, true, false>(b, v, m, y, st);
        return rot ? launch_bspread_rb<TD, TT, true, false, true>(b, v, m, y, st)
                   : launch_bspread_rb<TD, TT, true, false, false>(b, v, m, y, st);
    }
    if (b->neg_trav)
        return rot ? launch_bspread_rb<TD, TT, false, true, true>(b, v, m, y, st)
                   : launch_bspread_rb<TD, TT, false, true, false>(b, v, m, y, st);
    return rot ? launch_bspread_rb<TD, TT, false, false, true>(b, v, m, y, st)
               : launch_bspread_rb<TD, TT, false, false, false>(b, v, m, y, st);
}

template <typename TD, typename TT, bool DYN, bool NEG, int P, int NP>
int launch_bgather(const kx_batch *b, const View &v, const void *d, void *m, hipStream_t st)
{
    constexpr int JW = BatchJW<DYN, NP>::v;
    hipLaunchKernelGGL((batch_gather_kernel<TD, TT, DYN, NEG, JW, P, NP>), dim3(b->grid_gather),
                       dim3(kGatherBlock), 0, st, v.probs, v.gitems, v.ctr + kGatherQ,
                       v.ctr_const + kQoff + 16, reinterpret_cast<const TD *>(d),
                       reinterpret_cast<TD *>(m));
    KB_LAUNCHED("batch gather");
    return KX_OK;
}

template <typename TD, typename TT, bool DYN, bool NEG, int NP>
int launch_bgather_np(const kx_batch *b, const View &v, const void *d, void *m, hipStream_t st)
{
    switch (b->gsplit) {
    case 1:
        return launch_bgather<TD, TT, DYN, NEG, 1, NP>(b, v, d, m, st);
    case 2:
        return launch_bgather<TD, TT, DYN, NEG, 2, NP>(b, v, d, m, st);
    default:
        return launch_bgather<TD, TT, DYN, NEG, 4, NP>(b, v, d, m, st);
    }
}

template <typename TD, typename TT, bool DYN, bool NEG>
int launch_bgather_p(const kx_batch *b, const View &v, const void *d, void *m, hipStream_t st)
{
    return b->gtiles == 2 ? launch_bgather_np<TD, TT, DYN, NEG, 2>(b, v, d, m, st)
                          : launch_bgather_np<TD, TT, DYN, NEG, 1>(b, v, d, m, st);
}

template <typename TD, typename TT>
int launch_bgather_d(const kx_batch *b, const View &v, const void *d, void *m, hipStream_t st)
{
    if (b->dynamic)
        return b->neg_trav ? launch_bgather_p<TD, TT, true, true>(b, v, d, m, st)
                           : launch_bgather_p<TD, TT, true, false>(b, v, d, m, st);
    return b->neg_trav ? launch_bgather_p<TD, TT, false, true>(b, v, d, m, st)
                       : launch_bgather_p<TD, TT, false, false>(b, v, d, m, st);
}

int launch_bconv(const kx_batch *b, const View &v, const void *x, void *y, int adjoint,
                 hipStream_t st)
{
    const size_t lds = (size_t)conv_lds_elems(b->nh_max) * dsize(b->dtype);
    if (b->dtype == KX_F64)
        hipLaunchKernelGGL(batch_conv_kernel<double>, dim3((unsigned)b->n_conv), dim3(kConvThreads), lds,
                           st, v.probs, v.cprefix, b->nprob, reinterpret_cast<const double *>(x),
                           reinterpret_cast<double *>(y), adjoint);
    else
        hipLaunchKernelGGL(batch_conv_kernel<float>, dim3((unsigned)b->n_conv), dim3(kConvThreads), lds,
                           st, v.probs, v.cprefix, b->nprob, reinterpret_cast<const float *>(x),
                           reinterpret_cast<float *>(y), adjoint);
    KB_LAUNCHED("batch conv");
    return KX_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" int64_t kx_batch_dev_bytes(const kx_plan *plans, int32_t nprob)
{
    const int rc = check_batch(plans, nprob);
    if (rc)
        return rc;
    const Layout L = layout(plans, nprob, choose(plans, nprob));
    return (int64_t)L.total;
}

extern "C" int kx_batch_build(kx_batch *b, const kx_plan *plans, int32_t nprob, void *dev,
                              void *stream)
{
    if (!b || !dev)
        return bfail(KX_EINVAL, "batch: NULL batch or device buffer");
    int rc = check_batch(plans, nprob);
    if (rc)
        return rc;
    const Choice c = choose(plans, nprob);
    const Layout L = layout(plans, nprob, c);
    if (L.n_spread > INT32_MAX || L.n_gather > INT32_MAX || L.n_conv > INT32_MAX)
        return bfail(KX_ESHAPE, "batch: more than 2^31 work items");
    const kx_plan &q = plans[0];
    const int JW = batch_jw(q.dynamic, c.np);
    const int NT = kGatherWaves / c.P;

    // descriptors
    std::vector<BatchProb> probs(nprob);
    std::vector<int64_t> cprefix(nprob + 1, 0);
    int64_t moff = 0, doff = 0;
    int nt_max = 0, nh_max = 0;
    for (int p = 0; p < nprob; ++p) {
        const kx_plan &a = plans[p];
        if ((int64_t)JW * c.P * a.nr * a.nt >= ((int64_t)1 << 31))
            return bfail(KX_ESHAPE, "batch: problem %d: %d sources x nr x nt exceeds 2^31 samples",
                         p, JW * c.P);
        BatchProb &B = probs[p];
        B.a = make_args(&a);
        B.a.rblock = c.rb;
        set_tile_groups(B.a, c.np);
        B.moff = moff;
        B.doff = doff;
        B.ni = (int64_t)a.nx * a.ny * a.nz;
        B.h = a.h;
        B.hstar = a.hstar;
        B.nh = a.nh;
        B.nseg = (a.nt + kConvOut - 1) / kConvOut;
        B.wamp = a.wamp;
        B.wtyp = a.wtyp;
        moff += B.ni;
        doff += (int64_t)a.ns * a.nr * a.nt;
        cprefix[p + 1] = cprefix[p] + (int64_t)a.ns * a.nr * B.nseg;
        nt_max = std::max(nt_max, a.nt);
        nh_max = std::max(nh_max, a.nh);
    }

    // work lists: spread groups = (receiver table, receiver block), gather
    // groups = (source table, run of neighbouring tile groups), dealt to the
    // eight XCD queues
    std::vector<SpreadItem> sitems;
    std::vector<int64_t> scost;
    std::vector<int> sgroup;
    sitems.reserve(L.n_spread);
    scost.reserve(L.n_spread);
    sgroup.reserve(L.n_spread);
    std::map<std::pair<uintptr_t, int>, int> skeys;
    for (int p = 0; p < nprob; ++p) {
        const kx_plan &a = plans[p];
        for (int s = 0; s < a.ns; ++s)
            for (int r0 = 0; r0 < a.nr; r0 += c.rb) {
                const int nrl = std::min(c.rb, a.nr - r0);
                sitems.push_back({p, s, r0, 0});
                // tile sweeps plus clearing and writing out the traces
                scost.push_back(a.ntiles * kWave * nrl + (int64_t)nrl * a.nt * 4);
                const auto key = std::make_pair(reinterpret_cast<uintptr_t>(a.off_r), r0);
                auto it = skeys.emplace(key, (int)skeys.size()).first;
                sgroup.push_back(it->second);
            }
    }
    int sqoff[kXcds + 1];
    const std::vector<int> sperm = deal(sgroup, (int)skeys.size(), scost, sqoff,
                                        [&](int i, int j) {
                                            const SpreadItem &u = sitems[i], &v = sitems[j];
                                            return u.s != v.s ? u.s < v.s : u.p < v.p;
                                        });
    std::vector<SpreadItem> ssorted(sitems.size());
    for (size_t i = 0; i < sperm.size(); ++i)
        ssorted[i] = sitems[sperm[i]];

    std::vector<GatherItem> gitems;
    std::vector<int64_t> gcost;
    std::vector<int> ggroup;
    gitems.reserve(L.n_gather);
    gcost.reserve(L.n_gather);
    ggroup.reserve(L.n_gather);
    std::map<std::pair<uintptr_t, int>, int> gkeys;
    for (int p = 0; p < nprob; ++p) {
        const kx_plan &a = plans[p];
        const int64_t groups = probs[p].a.ngroups;  // of c.np tiles
        const int ng = (int)((groups + NT - 1) / NT);
        // neighbouring tiles read neighbouring trace windows: runs of tile
        // groups, about one eighth of the image each
        const int span = std::max(16, (ng + kXcds - 1) / kXcds);
        for (int g = 0; g < ng; ++g) {
            const int nti = std::min<int64_t>(NT, groups - (int64_t)g * NT);
            gitems.push_back({p, g});
            gcost.push_back((int64_t)nti * c.np * a.ns * a.nr);
            const auto key = std::make_pair(reinterpret_cast<uintptr_t>(a.off_s), g / span);
            auto it = gkeys.emplace(key, (int)gkeys.size()).first;
            ggroup.push_back(it->second);
        }
    }
    int gqoff[kXcds + 1];
    const std::vector<int> gperm = deal(ggroup, (int)gkeys.size(), gcost, gqoff,
                                        [&](int i, int j) {
                                            const GatherItem &u = gitems[i], &v = gitems[j];
                                            return u.L != v.L ? u.L < v.L : u.p < v.p;
                                        });
    std::vector<GatherItem> gsorted(gitems.size());
    for (size_t i = 0; i < gperm.size(); ++i)
        gsorted[i] = gitems[gperm[i]];
    std::vector<int> words(kCounters, 0);
    for (int k = 0; k <= kXcds; ++k) {
        words[kQoff + k] = sqoff[k];
        words[kQoff + 16 + k] = gqoff[k];
    }

    // persistent grids: the resident workgroups of the whole chip
    const Fns f = fns_of(q.dtype, q.tab_fmt);
    const int cus = device_cus();
    const int sp_cu = per_cu(f.spread(c.rb, c.rot, q.dynamic, q.neg_trav), kSpreadBlock,
                             spread_lds(c.rb, nt_max));
    const int ga_cu = per_cu(f.gather(c.P, c.np, q.dynamic, q.neg_trav), kGatherBlock, 0);

    b->abi_version = KX_ABI_VERSION;
    b->nprob = nprob;
    b->dtype = q.dtype;
    b->tab_fmt = q.tab_fmt;
    b->dynamic = q.dynamic;
    b->neg_trav = q.neg_trav;
    b->rblock = c.rb;
    b->sorder = c.rot ? KX_ORDER_ROTATED : KX_ORDER_NATURAL;
    b->gsplit = c.P;
    b->nt_max = nt_max;
    b->nh_max = nh_max;
    b->gtiles = c.np;
    b->model_len = moff;
    b->data_len = doff;
    b->n_spread = L.n_spread;
    b->n_gather = L.n_gather;
    b->n_conv = L.n_conv;
    b->grid_spread = (int)std::min<int64_t>(L.n_spread, (int64_t)cus * sp_cu);
    b->grid_gather = (int)std::min<int64_t>(L.n_gather, (int64_t)cus * ga_cu);
    b->dev = dev;

    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    char *base = reinterpret_cast<char *>(dev);
    KB_HIP(hipMemsetAsync(base, 0, L.total, st));
    KB_HIP(hipMemcpyAsync(base + L.probs, probs.data(), sizeof(BatchProb) * nprob,
                          hipMemcpyHostToDevice, st));
    KB_HIP(hipMemcpyAsync(base + L.sitems, ssorted.data(), sizeof(SpreadItem) * ssorted.size(),
                          hipMemcpyHostToDevice, st));
    KB_HIP(hipMemcpyAsync(base + L.gitems, gsorted.data(), sizeof(GatherItem) * gsorted.size(),
                          hipMemcpyHostToDevice, st));
    KB_HIP(hipMemcpyAsync(base + L.cprefix, cprefix.data(), sizeof(int64_t) * (nprob + 1),
                          hipMemcpyHostToDevice, st));
    KB_HIP(hipMemcpyAsync(base + L.ctr, words.data(), sizeof(int) * kCounters,
                          hipMemcpyHostToDevice, st));
    KB_HIP(hipStreamSynchronize(st));  // the host vectors end here
    return KX_OK;
}

extern "C" int64_t kx_batch_workspace_bytes(const kx_batch *b)
{
    if (check_built(b))
        return -1;
    return b->data_len * dsize(b->dtype);
}

extern "C" int kx_batch_forward(const kx_batch *b, const void *m, void *d, void *work, void *stream)
{
    int rc = check_built(b);
    if (rc)
        return rc;
    if (!m || !d || !work)
        return bfail(KX_EINVAL, "batch forward: NULL buffer");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const View v = view(b);
    if (b->dtype == KX_F32) {
        hipLaunchKernelGGL(batch_absmax_kernel, dim3(kBatchScaleBlocks, b->nprob), dim3(kBlock), 0,
                           st, v.probs, reinterpret_cast<const float *>(m), v.part);
        KB_LAUNCHED("batch scale");
        hipLaunchKernelGGL(batch_scale_kernel, dim3(b->nprob), dim3(kWave), 0, st, v.probs, v.part,
                           v.scales);
        KB_LAUNCHED("batch scale");
    }
    if (b->dtype == KX_F64)
        rc = launch_bspread_d<double, double>(b, v, m, work, st);
    else if (b->tab_fmt == KX_F16)
        rc = launch_bspread_d<float, __half>(b, v, m, work, st);
    else
        rc = launch_bspread_d<float, float>(b, v, m, work, st);
    if (rc)
        return rc;
    return launch_bconv(b, v, work, d, 0, st);
}

extern "C" int kx_batch_adjoint(const kx_batch *b, const void *d, void *m, void *work, void *stream)
{
    int rc = check_built(b);
    if (rc)
        return rc;
    if (!m || !d || !work)
        return bfail(KX_EINVAL, "batch adjoint: NULL buffer");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const View v = view(b);
    rc = launch_bconv(b, v, d, work, 1, st);
    if (rc)
        return rc;
    if (b->dtype == KX_F64)
        return launch_bgather_d<double, double>(b, v, work, m, st);
    if (b->tab_fmt == KX_F16)
        return launch_bgather_d<float, __half>(b, v, work, m, st);
    return launch_bgather_d<float, float>(b, v, work, m, st);
}
