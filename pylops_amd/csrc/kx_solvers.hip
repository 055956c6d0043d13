// kx_solvers.hip -- device-side vector algebra and scalar recurrences of
// CGLS and LSQR (pylops/optimization/cls_basic.py: CGLS.step CB:510-564,
// LSQR.step CB:987-1136).
//
// Vectors are float32 or float64; every inner product and norm is
// accumulated in float64.  The scalars of a solver live in one device
// float64 array `sc` (layouts KX_CG_* / KX_LS_* in kirchx.h) and are updated
// by single-thread "scalar step" kernels, so an iteration never needs the
// host: it can be enqueued (or graph-captured) as a fixed sequence.  The
// reference's loop exit (CGLS `iiter < niter and kold > tol`, LSQR `iiter <
// niter and istop == 0`) becomes the flag sc[ACTIVE]: once it drops, every
// later kernel of the solve returns without touching its vectors.
//
// Reductions: each block writes its partial sum, and the last block to
// finish (counted with a device-scope atomic in the caller's workspace)
// adds the partials and resets the counter to zero, so the workspace stays
// ready for the next reduction without a memset.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>

#include "kirchx.h"

extern "C" __attribute__((visibility("hidden"))) int kx_solver_fail(int code, const char *msg);

namespace {

constexpr int kBlock = 256;
constexpr int kWave = 64;
constexpr int kMaxBlocks = 1024;

struct Gate {
    const double *g0;  // run only if *g0 != 0 (NULL: no condition)
    const double *g1;
};

__device__ __forceinline__ bool gate_open(const Gate &g)
{
    if (g.g0 && *g.g0 == 0.0)
        return false;
    if (g.g1 && *g.g1 == 0.0)
        return false;
    return true;
}

__device__ __forceinline__ double block_reduce(double v, double *part)
{
#pragma unroll
    for (int o = kWave / 2; o >= 1; o >>= 1)
        v += __shfl_xor(v, o, kWave);
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    if (lane == 0)
        part[wave] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < (int)(blockDim.x / kWave); ++i)
            s += part[i];
    __syncthreads();
    return s;  // valid in thread 0
}

// Sum `v` over the grid into *out (float64).  work: kMaxBlocks partials,
// then one unsigned counter.
__device__ void grid_sum(double v, double *work, double *out)
{
    __shared__ double part[kBlock / kWave];
    __shared__ bool last;
    double *partial = work;
    unsigned *counter = reinterpret_cast<unsigned *>(work + kMaxBlocks);
    const double bs = block_reduce(v, part);
    if (threadIdx.x == 0) {
        partial[blockIdx.x] = bs;
        __threadfence();
        const unsigned prev = atomicAdd(counter, 1u);
        last = (prev == gridDim.x - 1);
    }
    __syncthreads();
    if (last) {
        __threadfence();
        double s = 0.0;
        for (int i = threadIdx.x; i < (int)gridDim.x; i += blockDim.x)
            s += __hip_atomic_load(&partial[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        s = block_reduce(s, part);
        if (threadIdx.x == 0) {
            *out = s;
            __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// y = x + c*y   (mode 0)      c = sign * sc[ic]
// y = y + c*x   (mode 1)
// y = c*y       (mode 2)      c = 1 / sc[ic]
// then, if out != NULL, *out = y.y
template <typename T, int MODE>
__global__ __launch_bounds__(kBlock) void vec_kernel(int64_t n, const T *__restrict__ x,
                                                     T *__restrict__ y, const double *coef,
                                                     double sign, double *out, double *work,
                                                     Gate g)
{
    if (!gate_open(g))
        return;  // uniform over the grid
    double c = coef ? *coef : 0.0;
    if (MODE == 2)
        c = 1.0 / c;
    else
        c *= sign;
    const T ct = (T)c;
    double acc = 0.0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        T v;
        if (MODE == 0)
            v = x[i] + ct * y[i];
        else if (MODE == 1)
            v = y[i] + ct * x[i];
        else
            v = ct * y[i];
        y[i] = v;
        acc += (double)v * (double)v;
    }
    if (out)
        grid_sum(acc, work, out);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void dot_kernel(int64_t n, const T *__restrict__ a,
                                                     const T *__restrict__ b, double *out,
                                                     double *work, Gate g)
{
    if (!gate_open(g))
        return;
    double acc = 0.0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        acc += (double)a[i] * (double)b[i];
    grid_sum(acc, work, out);
}

// LSQR x / w update (CB:1048-1069):  out = w.w (old w),  x += t1 w,
// w = v + t2 w
template <typename T>
__global__ __launch_bounds__(kBlock) void lsqr_xw_kernel(int64_t n, T *__restrict__ x,
                                                         T *__restrict__ w,
                                                         const T *__restrict__ v,
                                                         const double *sc, double *out,
                                                         double *work, Gate g)
{
    if (!gate_open(g))
        return;
    const T t1 = (T)sc[KX_LS_T1];
    const T t2 = (T)sc[KX_LS_T2];
    double acc = 0.0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const T wi = w[i];
        acc += (double)wi * (double)wi;
        x[i] = x[i] + t1 * wi;
        w[i] = v[i] + t2 * wi;
    }
    grid_sum(acc, work, out);
}

__device__ __forceinline__ double norm2(double a, double b) { return sqrt(a * a + b * b); }

// CGLS scalar recurrences (CB:518-563)
__device__ __forceinline__ void cgls_scalars(double *sc, double *cost, double *cost1, int phase)
{
    if (sc[KX_CG_ACTIVE] == 0.0)
        return;
    if (phase == 0) {
        const double a = sc[KX_CG_KOLD] / (sc[KX_CG_QQ] + sc[KX_CG_DAMP2] * sc[KX_CG_CC]);
        sc[KX_CG_A] = a;
    } else if (phase == 1) {
        sc[KX_CG_B] = fabs(sc[KX_CG_K]) / sc[KX_CG_KOLD];
    } else {
        sc[KX_CG_KOLD] = fabs(sc[KX_CG_K]);
        const double it = sc[KX_CG_ITER] + 1.0;
        sc[KX_CG_ITER] = it;
        const int k = (int)it;
        const double cs = sqrt(sc[KX_CG_SS]);
        cost[k] = cs;
        cost1[k] = sqrt(cs * cs + sc[KX_CG_DAMP2] * fabs(sc[KX_CG_XX]));
        sc[KX_CG_ACTIVE] = (it < sc[KX_CG_NITER] && sc[KX_CG_KOLD] > sc[KX_CG_TOL]) ? 1.0 : 0.0;
    }
}

__global__ void cgls_scalar_kernel(double *sc, double *cost, double *cost1, int phase)
{
    cgls_scalars(sc, cost, cost1, phase);
}

// one thread per problem of a batch
__global__ void cgls_scalar_batch_kernel(double *sc, int nseg, double *cost, double *cost1,
                                         int ncost, int phase)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < nseg)
        cgls_scalars(sc + (size_t)p * KX_CG_NSC, cost + (size_t)p * ncost,
                     cost1 + (size_t)p * ncost, phase);
}

// ---------------------------------------------------------------------------
// segment-wise vector algebra: one block per chunk, a chunk inside one
// segment; partial sums per chunk, then one block per segment adds them
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool seg_on(const double *sc, int nsc, int p, int gate)
{
    return gate < 0 || sc[(size_t)p * nsc + gate] != 0.0;
}

template <typename T, int MODE>
__global__ __launch_bounds__(kBlock) void seg_vec_kernel(kx_segs s, const T *__restrict__ x,
                                                         T *__restrict__ y, const double *sc,
                                                         int nsc, int coef, double sign,
                                                         int gate, double *partial)
{
    __shared__ double part[kBlock / kWave];
    const int c = blockIdx.x;
    const int p = s.chunk_seg[c];
    double acc = 0.0;
    if (seg_on(sc, nsc, p, gate)) {  // uniform over the block
        double cf = sc[(size_t)p * nsc + coef];
        cf = MODE == 2 ? 1.0 / cf : cf * sign;
        const T ct = (T)cf;
        for (int64_t i = s.chunk_lo[c] + threadIdx.x; i < s.chunk_lo[c + 1]; i += blockDim.x) {
            T v;
            if (MODE == 0)
                v = x[i] + ct * y[i];
            else if (MODE == 1)
                v = y[i] + ct * x[i];
            else
                v = ct * y[i];
            y[i] = v;
            acc += (double)v * (double)v;
        }
    }
    const double b = block_reduce(acc, part);
    if (threadIdx.x == 0)
        partial[c] = b;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void seg_dot_kernel(kx_segs s, const T *__restrict__ a,
                                                         const T *__restrict__ b,
                                                         const double *sc, int nsc, int gate,
                                                         double *partial)
{
    __shared__ double part[kBlock / kWave];
    const int c = blockIdx.x;
    const int p = s.chunk_seg[c];
    double acc = 0.0;
    if (seg_on(sc, nsc, p, gate)) {
        for (int64_t i = s.chunk_lo[c] + threadIdx.x; i < s.chunk_lo[c + 1]; i += blockDim.x)
            acc += (double)a[i] * (double)b[i];
    }
    const double r = block_reduce(acc, part);
    if (threadIdx.x == 0)
        partial[c] = r;
}

__global__ __launch_bounds__(kBlock) void seg_final_kernel(kx_segs s, const double *partial,
                                                           double *sc, int nsc, int out,
                                                           int gate)
{
    __shared__ double part[kBlock / kWave];
    const int p = blockIdx.x;
    if (!seg_on(sc, nsc, p, gate))
        return;  // uniform over the block
    double acc = 0.0;
    for (int c = s.seg_chunk[p] + threadIdx.x; c < s.seg_chunk[p + 1]; c += blockDim.x)
        acc += partial[c];
    const double r = block_reduce(acc, part);
    if (threadIdx.x == 0)
        sc[(size_t)p * nsc + out] = r;
}

// LSQR scalar recurrences (CB:987-1136)
__global__ void lsqr_scalar_kernel(double *sc, double *cost, int phase)
{
    if (sc[KX_LS_ACTIVE] == 0.0)
        return;
    if (phase == 0) {
        // beta = ||u|| after u = Op v - alfa u
        const double beta = sqrt(sc[KX_LS_UU]);
        sc[KX_LS_BETA] = beta;
        sc[KX_LS_BETAPOS] = beta > 0.0 ? 1.0 : 0.0;
        if (beta > 0.0) {
            const double an = sc[KX_LS_ANORM], al = sc[KX_LS_ALFA], dp = sc[KX_LS_DAMP];
            sc[KX_LS_ANORM] = sqrt(an * an + al * al + beta * beta + dp * dp);
        }
        sc[KX_LS_ALFAPOS] = 0.0;
    } else if (phase == 1) {
        // alfa = ||v|| after v = Op^H u - beta v (only when beta > 0)
        if (sc[KX_LS_BETAPOS] != 0.0) {
            const double alfa = sqrt(sc[KX_LS_VV]);
            sc[KX_LS_ALFA] = alfa;
            sc[KX_LS_ALFAPOS] = alfa > 0.0 ? 1.0 : 0.0;
        }
    } else if (phase == 2) {
        // plane rotations
        const double damp = sc[KX_LS_DAMP];
        const double alfa = sc[KX_LS_ALFA];
        const double beta = sc[KX_LS_BETA];
        double rhobar = sc[KX_LS_RHOBAR];
        double phibar = sc[KX_LS_PHIBAR];
        const double rhobar1 = norm2(rhobar, damp);
        const double cs1 = rhobar / rhobar1;
        const double sn1 = damp / rhobar1;
        const double psi = sn1 * phibar;
        phibar = cs1 * phibar;
        const double rho = norm2(rhobar1, beta);
        const double cs = rhobar1 / rho;
        const double sn = beta / rho;
        const double theta = sn * alfa;
        rhobar = -cs * alfa;
        const double phi = cs * phibar;
        phibar = sn * phibar;
        const double tau = sn * phi;
        sc[KX_LS_RHOBAR] = rhobar;
        sc[KX_LS_PHIBAR] = phibar;
        sc[KX_LS_PSI] = psi;
        sc[KX_LS_RHO] = rho;
        sc[KX_LS_THETA] = theta;
        sc[KX_LS_PHI] = phi;
        sc[KX_LS_TAU] = tau;
        sc[KX_LS_T1] = phi / rho;
        sc[KX_LS_T2] = -theta / rho;
    } else {
        // norms, condition estimate and stopping tests
        const double rho = sc[KX_LS_RHO];
        const double dk2 = sc[KX_LS_WW] / (rho * rho);
        sc[KX_LS_DDNORM] += dk2;
        sc[KX_LS_VAR] += dk2;
        const double delta = sc[KX_LS_SN2] * rho;
        const double gambar = -sc[KX_LS_CS2] * rho;
        const double rhs = sc[KX_LS_PHI] - delta * sc[KX_LS_Z];
        const double zbar = rhs / gambar;
        sc[KX_LS_XNORM] = sqrt(sc[KX_LS_XXNORM] + zbar * zbar);
        const double gamma = norm2(gambar, sc[KX_LS_THETA]);
        sc[KX_LS_CS2] = gambar / gamma;
        sc[KX_LS_SN2] = sc[KX_LS_THETA] / gamma;
        sc[KX_LS_Z] = rhs / gamma;
        sc[KX_LS_XXNORM] += sc[KX_LS_Z] * sc[KX_LS_Z];
        const double anorm = sc[KX_LS_ANORM];
        sc[KX_LS_ACOND] = anorm * sqrt(sc[KX_LS_DDNORM]);
        const double res1 = sc[KX_LS_PHIBAR] * sc[KX_LS_PHIBAR];
        sc[KX_LS_RES2] += sc[KX_LS_PSI] * sc[KX_LS_PSI];
        const double rnorm = sqrt(res1 + sc[KX_LS_RES2]);
        sc[KX_LS_RNORM] = rnorm;
        sc[KX_LS_ARNORM] = sc[KX_LS_ALFA] * fabs(sc[KX_LS_TAU]);
        const double r1sq = rnorm * rnorm - sc[KX_LS_DAMPSQ] * sc[KX_LS_XXNORM];
        double r1norm = sqrt(fabs(r1sq));
        const int it = (int)sc[KX_LS_ITER];
        cost[it + 1] = r1norm;
        if (r1sq < 0)
            r1norm = -r1norm;
        sc[KX_LS_R1NORM] = r1norm;
        sc[KX_LS_R2NORM] = rnorm;
        const double bnorm = sc[KX_LS_BNORM];
        const double xnorm = sc[KX_LS_XNORM];
        const double test1 = rnorm / bnorm;
        const double test2 = sc[KX_LS_ARNORM] / sc[KX_LS_ARNORM0];
        const double test3 = 1.0 / sc[KX_LS_ACOND];
        const double t1 = test1 / (1.0 + anorm * xnorm / bnorm);
        const double rtol = sc[KX_LS_BTOL] + sc[KX_LS_ATOL] * anorm * xnorm / bnorm;
        sc[KX_LS_TEST1] = test1;
        sc[KX_LS_TEST2] = test2;
        sc[KX_LS_TEST3] = test3;
        sc[KX_LS_RTOL] = rtol;
        double istop = sc[KX_LS_ISTOP];
        if (it >= (int)sc[KX_LS_NITER])
            istop = 7;
        if (1 + test3 <= 1)
            istop = 6;
        if (1 + test2 <= 1)
            istop = 5;
        if (1 + t1 <= 1)
            istop = 4;
        if (test3 <= sc[KX_LS_CTOL])
            istop = 3;
        if (test2 <= sc[KX_LS_ATOL])
            istop = 2;
        if (test1 <= rtol)
            istop = 1;
        sc[KX_LS_ISTOP] = istop;
        sc[KX_LS_ITER] = it + 1;
        sc[KX_LS_ACTIVE] = (it + 1 < (int)sc[KX_LS_NITER] && istop == 0) ? 1.0 : 0.0;
    }
}

int grid_for(int64_t n)
{
    const int64_t want = (n + kBlock * 4 - 1) / (kBlock * 4);
    return (int)std::max<int64_t>(1, std::min<int64_t>(kMaxBlocks, want));
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" int64_t kx_solver_workspace_bytes(void)
{
    return (int64_t)kMaxBlocks * 8 + 64;
}

extern "C" int kx_vec(int32_t op, int64_t n, int32_t dtype, const void *x, void *y,
                      const double *coef, double sign, double *out, void *work,
                      const double *gate0, const double *gate1, void *stream)
{
    if (n < 0 || !y || (op != KX_VEC_SCALE && !x) || (op == KX_VEC_SCALE && !coef))
        return kx_solver_fail(KX_EINVAL, "kx_vec: bad arguments");
    if (out && !work)
        return kx_solver_fail(KX_EINVAL, "kx_vec: reduction needs the workspace");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const Gate g{gate0, gate1};
    const int nb = grid_for(n);
    double *w = reinterpret_cast<double *>(work);
#define KX_VEC_LAUNCH(T, MODE)                                                            \
    hipLaunchKernelGGL((vec_kernel<T, MODE>), dim3(nb), dim3(kBlock), 0, st, n,            \
                       reinterpret_cast<const T *>(x), reinterpret_cast<T *>(y), coef, sign, \
                       out, w, g)
    if (dtype == KX_F32) {
        if (op == KX_VEC_XPBY)
            KX_VEC_LAUNCH(float, 0);
        else if (op == KX_VEC_AXPY)
            KX_VEC_LAUNCH(float, 1);
        else if (op == KX_VEC_SCALE)
            KX_VEC_LAUNCH(float, 2);
        else
            return kx_solver_fail(KX_EINVAL, "kx_vec: unknown op");
    } else if (dtype == KX_F64) {
        if (op == KX_VEC_XPBY)
            KX_VEC_LAUNCH(double, 0);
        else if (op == KX_VEC_AXPY)
            KX_VEC_LAUNCH(double, 1);
        else if (op == KX_VEC_SCALE)
            KX_VEC_LAUNCH(double, 2);
        else
            return kx_solver_fail(KX_EINVAL, "kx_vec: unknown op");
    } else {
        return kx_solver_fail(KX_ETYPE, "kx_vec: dtype");
    }
#undef KX_VEC_LAUNCH
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? KX_OK : kx_solver_fail((int)e, hipGetErrorString(e));
}

extern "C" int kx_dot_dev(int64_t n, int32_t dtype, const void *a, const void *b, double *out,
                          void *work, const double *gate0, void *stream)
{
    if (n < 0 || !a || !b || !out || !work)
        return kx_solver_fail(KX_EINVAL, "kx_dot_dev: bad arguments");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const Gate g{gate0, nullptr};
    const int nb = grid_for(n);
    double *w = reinterpret_cast<double *>(work);
    if (dtype == KX_F32)
        hipLaunchKernelGGL(dot_kernel<float>, dim3(nb), dim3(kBlock), 0, st, n,
                           reinterpret_cast<const float *>(a), reinterpret_cast<const float *>(b),
                           out, w, g);
    else if (dtype == KX_F64)
        hipLaunchKernelGGL(dot_kernel<double>, dim3(nb), dim3(kBlock), 0, st, n,
                           reinterpret_cast<const double *>(a),
                           reinterpret_cast<const double *>(b), out, w, g);
    else
        return kx_solver_fail(KX_ETYPE, "kx_dot_dev: dtype");
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? KX_OK : kx_solver_fail((int)e, hipGetErrorString(e));
}

extern "C" int kx_lsqr_xw(int64_t n, int32_t dtype, void *x, void *w, const void *v, double *sc,
                          void *work, void *stream)
{
    if (n < 0 || !x || !w || !v || !sc || !work)
        return kx_solver_fail(KX_EINVAL, "kx_lsqr_xw: bad arguments");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const Gate g{sc + KX_LS_ACTIVE, nullptr};
    const int nb = grid_for(n);
    double *wk = reinterpret_cast<double *>(work);
    if (dtype == KX_F32)
        hipLaunchKernelGGL(lsqr_xw_kernel<float>, dim3(nb), dim3(kBlock), 0, st, n,
                           reinterpret_cast<float *>(x), reinterpret_cast<float *>(w),
                           reinterpret_cast<const float *>(v), sc, sc + KX_LS_WW, wk, g);
    else if (dtype == KX_F64)
        hipLaunchKernelGGL(lsqr_xw_kernel<double>, dim3(nb), dim3(kBlock), 0, st, n,
                           reinterpret_cast<double *>(x), reinterpret_cast<double *>(w),
                           reinterpret_cast<const double *>(v), sc, sc + KX_LS_WW, wk, g);
    else
        return kx_solver_fail(KX_ETYPE, "kx_lsqr_xw: dtype");
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? KX_OK : kx_solver_fail((int)e, hipGetErrorString(e));
}

extern "C" int kx_cgls_scalar(double *sc, double *cost, double *cost1, int32_t phase,
                              void *stream)
{
    if (!sc || !cost || !cost1 || phase < 0 || phase > 2)
        return kx_solver_fail(KX_EINVAL, "kx_cgls_scalar: bad arguments");
    hipLaunchKernelGGL(cgls_scalar_kernel, dim3(1), dim3(1), 0,
                       reinterpret_cast<hipStream_t>(stream), sc, cost, cost1, phase);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? KX_OK : kx_solver_fail((int)e, hipGetErrorString(e));
}

extern "C" int kx_lsqr_scalar(double *sc, double *cost, int32_t phase, void *stream)
{
    if (!sc || !cost || phase < 0 || phase > 3)
        return kx_solver_fail(KX_EINVAL, "kx_lsqr_scalar: bad arguments");
    hipLaunchKernelGGL(lsqr_scalar_kernel, dim3(1), dim3(1), 0,
                       reinterpret_cast<hipStream_t>(stream), sc, cost, phase);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? KX_OK : kx_solver_fail((int)e, hipGetErrorString(e));
}

// ---------------------------------------------------------------------------
// segment-wise entry points
// ---------------------------------------------------------------------------
namespace {

int check_segs(const kx_segs *s)
{
    if (!s || s->nseg <= 0 || s->nchunk <= 0 || !s->off || !s->chunk_seg || !s->chunk_lo ||
        !s->seg_chunk)
        return kx_solver_fail(KX_EINVAL, "segments: bad descriptor");
    return KX_OK;
}

int seg_final(const kx_segs *s, double *partial, double *sc, int nsc, int out, int gate,
              hipStream_t st)
{
    hipLaunchKernelGGL(seg_final_kernel, dim3(s->nseg), dim3(kBlock), 0, st, *s, partial, sc, nsc,
                       out, gate);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? KX_OK : kx_solver_fail((int)e, hipGetErrorString(e));
}

}  // namespace

extern "C" int64_t kx_seg_workspace_bytes(const kx_segs *s)
{
    return s && s->nchunk > 0 ? (int64_t)s->nchunk * 8 : -1;
}

extern "C" int kx_seg_vec(int32_t op, const kx_segs *s, int32_t dtype, const void *x, void *y,
                          double *sc, int32_t nsc, int32_t coef, double sign, int32_t out,
                          int32_t gate, void *work, void *stream)
{
    int rc = check_segs(s);
    if (rc)
        return rc;
    if (!y || !sc || !work || coef < 0 || coef >= nsc || out >= nsc || gate >= nsc ||
        (op != KX_VEC_SCALE && !x))
        return kx_solver_fail(KX_EINVAL, "kx_seg_vec: bad arguments");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    double *partial = reinterpret_cast<double *>(work);
    const dim3 grid(s->nchunk), block(kBlock);
#define KX_SEG_VEC(T, M)                                                                       \
    hipLaunchKernelGGL((seg_vec_kernel<T, M>), grid, block, 0, st, *s,                         \
                       reinterpret_cast<const T *>(x), reinterpret_cast<T *>(y), sc, nsc, coef,  \
                       sign, gate, partial)
    if (dtype == KX_F32) {
        if (op == KX_VEC_XPBY)
            KX_SEG_VEC(float, 0);
        else if (op == KX_VEC_AXPY)
            KX_SEG_VEC(float, 1);
        else
            KX_SEG_VEC(float, 2);
    } else if (dtype == KX_F64) {
        if (op == KX_VEC_XPBY)
            KX_SEG_VEC(double, 0);
        else if (op == KX_VEC_AXPY)
            KX_SEG_VEC(double, 1);
        else
            KX_SEG_VEC(double, 2);
    } else {
        return kx_solver_fail(KX_ETYPE, "kx_seg_vec: dtype");
    }
#undef KX_SEG_VEC
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return kx_solver_fail((int)e, hipGetErrorString(e));
    return out >= 0 ? seg_final(s, partial, sc, nsc, out, gate, st) : KX_OK;
}

extern "C" int kx_seg_dot(const kx_segs *s, int32_t dtype, const void *a, const void *b,
                          double *sc, int32_t nsc, int32_t out, int32_t gate, void *work,
                          void *stream)
{
    int rc = check_segs(s);
    if (rc)
        return rc;
    if (!a || !b || !sc || !work || out < 0 || out >= nsc || gate >= nsc)
        return kx_solver_fail(KX_EINVAL, "kx_seg_dot: bad arguments");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    double *partial = reinterpret_cast<double *>(work);
    if (dtype == KX_F32)
        hipLaunchKernelGGL(seg_dot_kernel<float>, dim3(s->nchunk), dim3(kBlock), 0, st, *s,
                           reinterpret_cast<const float *>(a), reinterpret_cast<const float *>(b),
                           sc, nsc, gate, partial);
    else if (dtype == KX_F64)
        hipLaunchKernelGGL(seg_dot_kernel<double>, dim3(s->nchunk), dim3(kBlock), 0, st, *s,
                           reinterpret_cast<const double *>(a),
                           reinterpret_cast<const double *>(b), sc, nsc, gate, partial);
    else
        return kx_solver_fail(KX_ETYPE, "kx_seg_dot: dtype");
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return kx_solver_fail((int)e, hipGetErrorString(e));
    return seg_final(s, partial, sc, nsc, out, gate, st);
}

extern "C" int kx_cgls_scalar_batch(double *sc, int32_t nseg, double *cost, double *cost1,
                                    int32_t ncost, int32_t phase, void *stream)
{
    if (!sc || !cost || !cost1 || nseg <= 0 || ncost <= 0 || phase < 0 || phase > 2)
        return kx_solver_fail(KX_EINVAL, "kx_cgls_scalar_batch: bad arguments");
    hipLaunchKernelGGL(cgls_scalar_batch_kernel, dim3((nseg + kWave - 1) / kWave), dim3(kWave), 0,
                       reinterpret_cast<hipStream_t>(stream), sc, nseg, cost, cost1, ncost, phase);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? KX_OK : kx_solver_fail((int)e, hipGetErrorString(e));
}
