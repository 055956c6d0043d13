// kx_eikonal.hip -- first-arrival traveltimes by fast sweeping
// (include/kirchx.h: kx_eikonal).
//
// Replaces the scikit-fmm calls of Kirchhoff._traveltime_table's eikonal
// mode (pylops/waveeqprocessing/kirchhoff.py:609-638): for every source
// cell, the traveltime (or, with unit slowness, the distance) from the zero
// contour of phi = -1 at that cell and +1 elsewhere.
//
// Method (Zhao 2005, fast sweeping): first-order Godunov upwind updates of
// |grad T| = s, swept over the grid in the 4 (2D) or 8 (3D) diagonal
// orderings, repeated until no value changes.  Within one sweep every cell
// of an anti-diagonal ix' + iy' + iz' = d (primed: index counted along the
// sweep's direction) depends only on the diagonals before it, so one launch
// updates a whole diagonal for every source at once.  Cells within two
// cells of the source (per axis) start at the straight-ray time from the
// contour -- a diamond through the half-way points to the neighbouring cell
// centres, where phi interpolates to zero -- and stay fixed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <vector>

#include "kirchx.h"

extern "C" int kx_solver_fail(int code, const char *msg);

namespace {

constexpr double kBig = 1e30;
constexpr int kEikBlock = 256;
constexpr int kInitRadius = 2;

int efail(int code, const char *fmt, ...)
{
    char buf[480];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    return kx_solver_fail(code, buf);
}

struct EikArgs {
    int nx, ny, nz;
    double hx, hy, hz;
    int three;     // 1 in 3D
    int64_t ni;
};

// Godunov update from the smaller neighbour value a of each axis (kBig when
// there is none): the largest t solving sum_k ((t - a_k) / h_k)^2 = s^2 over
// the axes whose a_k lies below t.
__device__ __forceinline__ double godunov(double ax, double ay, double az, double hx, double hy,
                                          double hz, double s, int three)
{
    double a0 = ax, a1 = az, a2 = ay;
    double h0 = hx, h1 = hz, h2 = hy;
    // order by a (three elements, the third only in 3D)
    if (a1 < a0) {
        double t = a0; a0 = a1; a1 = t;
        t = h0; h0 = h1; h1 = t;
    }
    if (three) {
        if (a2 < a1) {
            double t = a1; a1 = a2; a2 = t;
            t = h1; h1 = h2; h2 = t;
        }
        if (a1 < a0) {
            double t = a0; a0 = a1; a1 = t;
            t = h0; h0 = h1; h1 = t;
        }
    }
    double t = a0 + s * h0;
    if (t > a1) {
        const double w0 = 1.0 / (h0 * h0), w1 = 1.0 / (h1 * h1);
        double A = w0 + w1;
        double B = a0 * w0 + a1 * w1;
        double C = a0 * a0 * w0 + a1 * a1 * w1 - s * s;
        t = (B + sqrt(fmax(B * B - A * C, 0.0))) / A;
        if (three && t > a2) {
            const double w2 = 1.0 / (h2 * h2);
            A += w2;
            B += a2 * w2;
            C += a2 * a2 * w2;
            t = (B + sqrt(fmax(B * B - A * C, 0.0))) / A;
        }
    }
    return t;
}

// T = straight-ray time from the source contour inside the fixed box, kBig
// elsewhere.  One thread per (cell, source).
__global__ __launch_bounds__(kEikBlock) void eik_init_kernel(EikArgs g,
                                                             const double *__restrict__ slow,
                                                             const int *__restrict__ src,
                                                             double *__restrict__ T)
{
    const int s = blockIdx.y;
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.ni)
        return;
    const int iz = (int)(c % g.nz);
    const int ix = (int)((c / g.nz) % g.nx);
    const int iy = (int)(c / ((int64_t)g.nz * g.nx));
    const int sx = src[3 * s], sy = src[3 * s + 1], sz = src[3 * s + 2];
    double v = kBig;
    if (abs(ix - sx) <= kInitRadius && abs(iy - sy) <= kInitRadius &&
        abs(iz - sz) <= kInitRadius) {
        const double dx = (ix - sx) * g.hx, dy = (iy - sy) * g.hy, dz = (iz - sz) * g.hz;
        const double rho = sqrt(dx * dx + dy * dy + dz * dz);
        // the contour in this direction lies at rho / S
        const double S = 2.0 * (fabs(dx) / g.hx + fabs(dz) / g.hz + (g.three ? fabs(dy) / g.hy : 0.0));
        const int64_t cs = ((int64_t)sy * g.nx + sx) * g.nz + sz;
        const double sl = 0.5 * (slow[c] + slow[cs]);
        v = S > 0.0 ? sl * rho * fmax(0.0, 1.0 - 1.0 / S) : 0.0;
    }
    T[(int64_t)s * g.ni + c] = v;
}

// One anti-diagonal d of one sweep direction (dx, dy, dz = +1 / -1), every
// source: thread (pair, source) with pair = (ix' - ixlo) * ny + iy'.
__global__ __launch_bounds__(kEikBlock) void eik_sweep_kernel(EikArgs g,
                                                              const double *__restrict__ slow,
                                                              const int *__restrict__ src,
                                                              double *__restrict__ T, int d,
                                                              int dx, int dy, int dz, int ixlo,
                                                              int nix, int *changed)
{
    const int s = blockIdx.y;
    const int64_t pair = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (pair >= (int64_t)nix * g.ny)
        return;
    const int px = ixlo + (int)(pair / g.ny);
    const int py = (int)(pair % g.ny);
    const int pz = d - px - py;
    if (pz < 0 || pz >= g.nz)
        return;
    const int ix = dx > 0 ? px : g.nx - 1 - px;
    const int iy = dy > 0 ? py : g.ny - 1 - py;
    const int iz = dz > 0 ? pz : g.nz - 1 - pz;
    const int sx = src[3 * s], sy = src[3 * s + 1], sz = src[3 * s + 2];
    if (abs(ix - sx) <= kInitRadius && abs(iy - sy) <= kInitRadius &&
        abs(iz - sz) <= kInitRadius)
        return;  // fixed
    double *Ts = T + (int64_t)s * g.ni;
    const int64_t c = ((int64_t)iy * g.nx + ix) * g.nz + iz;
    const int64_t sxs = g.nz, sys = (int64_t)g.nx * g.nz;
    const double ax = fmin(ix > 0 ? Ts[c - sxs] : kBig, ix < g.nx - 1 ? Ts[c + sxs] : kBig);
    const double az = fmin(iz > 0 ? Ts[c - 1] : kBig, iz < g.nz - 1 ? Ts[c + 1] : kBig);
    double ay = kBig;
    if (g.three)
        ay = fmin(iy > 0 ? Ts[c - sys] : kBig, iy < g.ny - 1 ? Ts[c + sys] : kBig);
    const double t = godunov(ax, ay, az, g.hx, g.hy, g.hz, slow[c], g.three);
    const double old = Ts[c];
    if (t < old) {
        Ts[c] = t;
        if (old - t > 1e-13 * t)
            *changed = 1;
    }
}

}  // namespace

extern "C" int kx_eikonal(int32_t ndims, int32_t nx, int32_t ny, int32_t nz, double hx, double hy,
                          double hz, const double *slow, int32_t n, const int32_t *src, double *T,
                          int32_t max_iter, int32_t *flag, int32_t *iters, void *stream)
{
    if (ndims != 2 && ndims != 3)
        return efail(KX_EINVAL, "eikonal: ndims must be 2 or 3");
    if (ndims == 2)
        ny = 1;
    if (nx <= 0 || ny <= 0 || nz <= 0 || n <= 0 || n > 65535 || max_iter <= 0)
        return efail(KX_ESHAPE, "eikonal: bad sizes (nx=%d ny=%d nz=%d n=%d)", nx, ny, nz, n);
    if (!(hx > 0) || !(hz > 0) || (ndims == 3 && !(hy > 0)))
        return efail(KX_EINVAL, "eikonal: spacings must be positive");
    if (!slow || !src || !T || !flag)
        return efail(KX_EINVAL, "eikonal: NULL buffer");
    EikArgs g;
    g.nx = nx;
    g.ny = ny;
    g.nz = nz;
    g.hx = hx;
    g.hy = ndims == 3 ? hy : 1.0;
    g.hz = hz;
    g.three = ndims == 3;
    g.ni = (int64_t)nx * ny * nz;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    // the kernels index the slowness at each source cell: check the cells
    // on the host before any launch
    std::vector<int32_t> hs((size_t)3 * n);
    hipError_t e = hipMemcpyAsync(hs.data(), src, hs.size() * sizeof(int32_t),
                                  hipMemcpyDeviceToHost, st);
    if (e == hipSuccess)
        e = hipStreamSynchronize(st);
    if (e != hipSuccess)
        return efail((int)e, "eikonal: %s", hipGetErrorString(e));
    for (int32_t k = 0; k < n; ++k) {
        const int32_t sx = hs[3 * k], sy = hs[3 * k + 1], sz = hs[3 * k + 2];
        if (sx < 0 || sx >= nx || sy < 0 || sy >= ny || sz < 0 || sz >= nz)
            return efail(KX_ESHAPE, "eikonal: source %d at cell (%d, %d, %d) outside the "
                         "%d x %d x %d grid", k, sx, sy, sz, nx, ny, nz);
    }
    const int64_t nb = (g.ni + kEikBlock - 1) / kEikBlock;
    if (nb > INT32_MAX)
        return efail(KX_ESHAPE, "eikonal: grid too large");
    hipLaunchKernelGGL(eik_init_kernel, dim3((unsigned)nb, n), dim3(kEikBlock), 0, st, g, slow, src, T);
    e = hipGetLastError();
    if (e != hipSuccess)
        return efail((int)e, "eikonal init: %s", hipGetErrorString(e));
    const int ndiag = (nx - 1) + (ny - 1) + (nz - 1) + 1;
    const int ndir = ndims == 3 ? 8 : 4;
    int it = 0;
    for (; it < max_iter; ++it) {
        e = hipMemsetAsync(flag, 0, sizeof(int32_t), st);
        if (e != hipSuccess)
            return efail((int)e, "eikonal: %s", hipGetErrorString(e));
        for (int q = 0; q < ndir; ++q) {
            const int dx = (q & 1) ? -1 : 1, dz = (q & 2) ? -1 : 1, dy = (q & 4) ? -1 : 1;
            for (int d = 0; d < ndiag; ++d) {
                const int ixlo = std::max(0, d - (ny - 1) - (nz - 1));
                const int ixhi = std::min(nx - 1, d);
                const int nix = ixhi - ixlo + 1;
                const int64_t npair = (int64_t)nix * ny;
                const int64_t nbd = (npair + kEikBlock - 1) / kEikBlock;
                hipLaunchKernelGGL(eik_sweep_kernel, dim3((unsigned)nbd, n), dim3(kEikBlock), 0, st,
                                   g, slow, src, T, d, dx, dy, dz, ixlo, nix, flag);
            }
        }
        e = hipGetLastError();
        if (e != hipSuccess)
            return efail((int)e, "eikonal sweep: %s", hipGetErrorString(e));
        int32_t h = 0;
        e = hipMemcpyAsync(&h, flag, sizeof(int32_t), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess)
            e = hipStreamSynchronize(st);
        if (e != hipSuccess)
            return efail((int)e, "eikonal: %s", hipGetErrorString(e));
        if (!h) {
            ++it;
            break;
        }
    }
    if (iters)
        *iters = it;
    return KX_OK;
}
