// kx_eikonal.hip -- first-arrival traveltimes by fast sweeping
// (include/kirchx.h: kx_eikonal).
//
// Replaces the scikit-fmm calls of Kirchhoff._traveltime_table's eikonal
// mode (pylops/waveeqprocessing/kirchhoff.py:609-638): for every source
// cell, the traveltime (or, with unit slowness, the distance) from the zero
// contour of phi = -1 at that cell and +1 elsewhere.
//
// Method (Zhao 2005, fast sweeping): Godunov upwind updates of |grad T| = s,
// swept over the grid in the 4 (2D) or 8 (3D) diagonal orderings, repeated
// until no value changes: first-order updates to convergence, then (order
// 2, scikit-fmm's default order) second-order one-sided differences where
// the two upwind cells are monotone, lowering values until they converge.  Within one sweep every cell
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
// a workgroup is kEikSrc sources (one wave, adjacent in memory) x
// kEikCells cells
constexpr int kEikSrc = 64;
constexpr int kEikCells = 4;
constexpr int kEikBlock = kEikSrc * kEikCells;
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
    int n;         // sources
    int64_t ni;
};

// One axis of the upwind update: upwind value a (the smaller neighbour,
// kBig when there is none), difference base b and step g, so that the
// one-sided derivative is (t - b) / g.  First order: b = a, g = h.  Second
// order, where the second cell on the upwind side is known and not above
// the first: (3t - 4a + a2) / (2h), i.e. b = (4a - a2) / 3, g = 2h / 3.
struct Axis {
    double a, b, g;
};

__device__ __forceinline__ Axis axis_terms(const double *Ts, int64_t c, int i, int n,
                                           int64_t stride, double h, int order)
{
    const double l1 = i > 0 ? Ts[c - stride] : kBig;
    const double r1 = i < n - 1 ? Ts[c + stride] : kBig;
    Axis x;
    x.g = h;
    if (l1 <= r1) {
        x.a = l1;
        x.b = l1;
        if (order == 2 && i > 1 && l1 < kBig) {
            const double a2 = Ts[c - 2 * stride];
            if (a2 < kBig && a2 <= l1) {
                x.b = (4.0 * l1 - a2) / 3.0;
                x.g = (2.0 / 3.0) * h;
            }
        }
    } else {
        x.a = r1;
        x.b = r1;
        if (order == 2 && i < n - 2 && r1 < kBig) {
            const double a2 = Ts[c + 2 * stride];
            if (a2 < kBig && a2 <= r1) {
                x.b = (4.0 * r1 - a2) / 3.0;
                x.g = (2.0 / 3.0) * h;
            }
        }
    }
    return x;
}

__device__ __forceinline__ void order_axes(Axis &u, Axis &v)
{
    if (v.a < u.a) {
        const Axis t = u;
        u = v;
        v = t;
    }
}

// Godunov update: the largest t with sum_k ((t - b_k) / g_k)^2 = s^2 over
// the axes whose upwind value a_k lies below t, taken in increasing a_k.
// *ok is cleared when a quadratic has no real root (possible only with
// second-order terms).
__device__ __forceinline__ double solve_upwind(Axis x0, Axis x1, Axis x2, double s, int three,
                                               bool *ok)
{
    order_axes(x0, x1);
    if (three) {
        order_axes(x1, x2);
        order_axes(x0, x1);
    }
    *ok = true;
    double t = x0.b + s * x0.g;
    if (t > x1.a) {
        const double w0 = 1.0 / (x0.g * x0.g), w1 = 1.0 / (x1.g * x1.g);
        double A = w0 + w1;
        double B = x0.b * w0 + x1.b * w1;
        double C = x0.b * x0.b * w0 + x1.b * x1.b * w1 - s * s;
        double D = B * B - A * C;
        *ok = D >= 0.0;
        t = (B + sqrt(fmax(D, 0.0))) / A;
        if (three && t > x2.a) {
            const double w2 = 1.0 / (x2.g * x2.g);
            A += w2;
            B += x2.b * w2;
            C += x2.b * x2.b * w2;
            D = B * B - A * C;
            *ok = *ok && D >= 0.0;
            t = (B + sqrt(fmax(D, 0.0))) / A;
        }
    }
    return t;
}

// T = straight-ray time from the source contour inside the fixed box, kBig
// elsewhere.  One thread per (cell, source); T[c * n + s].
__global__ __launch_bounds__(kEikBlock) void eik_init_kernel(EikArgs g,
                                                             const double *__restrict__ slow,
                                                             const int *__restrict__ src,
                                                             double *__restrict__ T)
{
    const int s = blockIdx.y * kEikSrc + threadIdx.x;
    const int64_t c = (int64_t)blockIdx.x * kEikCells + threadIdx.y;
    if (c >= g.ni || s >= g.n)
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
    T[c * g.n + s] = v;
}

// One anti-diagonal d of one sweep direction (dx, dy, dz = +1 / -1), every
// source: thread (pair, source) with pair = (ix' - ixlo) * ny + iy'.  The
// cells of a diagonal are never adjacent in memory, so the tables are
// point-major (T[c * n + s]) and the lanes of a wave take adjacent sources
// of one cell: every load and store is one contiguous 512-byte row.
// order 2 takes second-order terms where they apply and falls back to the
// first-order update of the cell when a quadratic has no root.
__global__ __launch_bounds__(kEikBlock) void eik_sweep_kernel(EikArgs g,
                                                              const double *__restrict__ slow,
                                                              const int *__restrict__ src,
                                                              double *__restrict__ T, int d,
                                                              int dx, int dy, int dz, int ixlo,
                                                              int nix, int order, int *changed)
{
    const int s = blockIdx.y * kEikSrc + threadIdx.x;
    const int64_t pair = (int64_t)blockIdx.x * kEikCells + threadIdx.y;
    if (pair >= (int64_t)nix * g.ny || s >= g.n)
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
    const int64_t c = ((int64_t)iy * g.nx + ix) * g.nz + iz;
    const int64_t e = c * g.n + s;
    const int64_t szs = g.n, sxs = (int64_t)g.nz * g.n, sys = (int64_t)g.nx * g.nz * g.n;
    const double sl = slow[c];
    const Axis ax = axis_terms(T, e, ix, g.nx, sxs, g.hx, order);
    const Axis az = axis_terms(T, e, iz, g.nz, szs, g.hz, order);
    Axis ay = {kBig, kBig, g.hy};
    if (g.three)
        ay = axis_terms(T, e, iy, g.ny, sys, g.hy, order);
    bool ok;
    // axes in the order x, z, y (the third only in 3D)
    double t = solve_upwind(ax, az, ay, sl, g.three, &ok);
    if (!ok) {
        const Axis fx = {ax.a, ax.a, g.hx}, fy = {ay.a, ay.a, g.hy}, fz = {az.a, az.a, g.hz};
        t = solve_upwind(fx, fz, fy, sl, g.three, &ok);
    }
    const double old = T[e];
    if (t < old) {
        T[e] = t;
        if (old - t > 1e-13 * t)
            *changed = 1;
    }
}

}  // namespace

extern "C" int kx_eikonal(int32_t ndims, int32_t nx, int32_t ny, int32_t nz, double hx, double hy,
                          double hz, const double *slow, int32_t n, const int32_t *src, double *T,
                          int32_t max_iter, int32_t order, int32_t *flag, int32_t *iters,
                          void *stream)
{
    if (ndims != 2 && ndims != 3)
        return efail(KX_EINVAL, "eikonal: ndims must be 2 or 3");
    if (ndims == 2)
        ny = 1;
    if (order != 1 && order != 2)
        return efail(KX_EINVAL, "eikonal: order must be 1 or 2");
    if (nx <= 0 || ny <= 0 || nz <= 0 || n <= 0 || n > (1 << 22) || max_iter <= 0)
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
    g.n = n;
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
    const int64_t nb = (g.ni + kEikCells - 1) / kEikCells;
    const unsigned nsb = (unsigned)((n + kEikSrc - 1) / kEikSrc);
    if (nb > INT32_MAX)
        return efail(KX_ESHAPE, "eikonal: grid too large");
    hipLaunchKernelGGL(eik_init_kernel, dim3((unsigned)nb, nsb), dim3(kEikSrc, kEikCells), 0, st, g,
                       slow, src, T);
    e = hipGetLastError();
    if (e != hipSuccess)
        return efail((int)e, "eikonal init: %s", hipGetErrorString(e));
    const int ndiag = (nx - 1) + (ny - 1) + (nz - 1) + 1;
    const int ndir = ndims == 3 ? 8 : 4;
    int rounds = 0, converged = 1;
    // first-order sweeps to convergence, then (order 2) second-order ones
    for (int phase = 1; phase <= order; ++phase) {
        for (int it = 0; it < max_iter; ++it) {
            ++rounds;
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
                    const int64_t nbd = (npair + kEikCells - 1) / kEikCells;
                    hipLaunchKernelGGL(eik_sweep_kernel, dim3((unsigned)nbd, nsb),
                                       dim3(kEikSrc, kEikCells), 0,
                                       st, g, slow, src, T, d, dx, dy, dz, ixlo, nix, phase, flag);
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
            if (!h)
                break;
            if (it == max_iter - 1)
                converged = 0;  // values still changed in the last round
        }
    }
    if (iters) {
        iters[0] = rounds;
        iters[1] = converged;
    }
    return KX_OK;
}
