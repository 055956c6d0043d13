/*
 * kirch_oracle.c -- CPU oracle for the Kirchhoff hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is a plain-C restatement of the
 * reference PyLops loop kernels, used as the checker for the HIP library and
 * as the timed CPU baseline.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product never does.
 *
 * Every function keeps the reference's loop nest, parallel axis and float64
 * index/weight arithmetic:
 *   - forward kernels parallelise over sources (numba prange(ns)),
 *   - adjoint kernels parallelise over image points (numba prange(ni)),
 *   - trav/dt, the int32 truncation, the interpolation weight and the dynamic
 *     amplitude are evaluated in float64 in the reference's operand order,
 *   - the output element type is the operator dtype: the *_f32 entry points
 *     round every `+=` to float32 exactly like a float32 numpy/numba array;
 *     *_x64y32 take a float64 input with a float32 output, which is what the
 *     reference's float32 adjoint sees after correlating with a float64
 *     wavelet (scipy upcasts, convolve1d.py:108-115).
 * Tables are the reference's row-major float64 tables trav[ii*n + i].
 *
 * Reference: pylops/waveeqprocessing/kirchhoff.py (K:) lines cited per
 * function; Convolve1D "same" convolution: pylops/signalprocessing/
 * convolve1d.py:57-115 (scipy.signal.fftconvolve(mode="same") centring).
 */
#include <math.h>
#include <stdint.h>

/* CPython / numpy float floor division (float_divmod / npy_divmod), used by
 * the reference's `//` on float64 scalars in the aperture taper index
 * (K:845-879, K:954-988). */
static double py_floordiv(double vx, double wx)
{
    double mod = fmod(vx, wx);
    double div = (vx - mod) / wx;
    if (mod != 0.0) {
        if ((wx < 0) != (mod < 0)) {
            mod += wx;
            div -= 1.0;
        }
    }
    double floordiv;
    if (div != 0.0) {
        floordiv = floor(div);
        if (div - floordiv > 0.5)
            floordiv += 1.0;
    } else {
        floordiv = copysign(0.0, vx / wx);
    }
    return floordiv;
}

/* numpy astype("int32") on x86: truncation, out-of-range / NaN -> INT32_MIN. */
static int32_t np_astype_i32(double q)
{
    if (!(q > -2147483649.0 && q < 2147483648.0))
        return INT32_MIN;
    return (int32_t)q;
}

/* Python int(q) followed by the reference's `0 <= it < nt-1` test: returns 1
 * and sets *it when the test passes.  int() truncates toward zero. */
static int py_int_in_range(double q, int nt, int64_t *it)
{
    if (!(q > -1.0 && q < (double)(nt - 1)))
        return 0;
    *it = (int64_t)q;
    return (*it >= 0 && *it < nt - 1);
}

/* ------------------------------------------------------------------------
 * Kinematic, separable tables.
 * Forward  K:727-752 (_travsrcrec_kirch_matvec)
 * Adjoint  K:754-781 (_travsrcrec_kirch_rmatvec)
 * ---------------------------------------------------------------------- */
#define KIN_FWD(NAME, TX, TY)                                                       \
    void NAME(const TX *x, TY *y, int ns, int nr, int nt, int64_t ni, double dt, \
              const double *ts, const double *tr)                              \
    {                                                                          \
        _Pragma("omp parallel for schedule(dynamic,1)")                        \
        for (int isrc = 0; isrc < ns; ++isrc) {                                \
            for (int irec = 0; irec < nr; ++irec) {                            \
                TY *yrow = y + (int64_t)(isrc * nr + irec) * nt;                \
                for (int64_t ii = 0; ii < ni; ++ii) {                          \
                    double trav = ts[ii * ns + isrc] + tr[ii * nr + irec];     \
                    int32_t itr = np_astype_i32(trav / dt);                    \
                    double travd = trav / dt - itr;                            \
                    if (0 <= itr && itr < nt - 1) {                            \
                        yrow[itr] += x[ii] * (1 - travd);                      \
                        yrow[itr + 1] += x[ii] * travd;                        \
                    }                                                          \
                }                                                              \
            }                                                                  \
        }                                                                      \
    }

#define KIN_ADJ(NAME, TX, TY)                                                       \
    void NAME(const TX *x, TY *y, int ns, int nr, int nt, int64_t ni, double dt, \
              const double *ts, const double *tr)                              \
    {                                                                          \
        _Pragma("omp parallel for schedule(static)")                           \
        for (int64_t ii = 0; ii < ni; ++ii) {                                  \
            const double *tsi = ts + ii * ns;                                  \
            const double *tri = tr + ii * nr;                                  \
            for (int isrc = 0; isrc < ns; ++isrc) {                            \
                double tsv = tsi[isrc];                                        \
                for (int irec = 0; irec < nr; ++irec) {                        \
                    double trav = tsv + tri[irec];                             \
                    int64_t itr;                                               \
                    if (py_int_in_range(trav / dt, nt, &itr)) {                \
                        double travd = trav / dt - (double)itr;                \
                        const TX *xr = x + (int64_t)(isrc * nr + irec) * nt;    \
                        y[ii] += xr[itr] * (1 - travd) + xr[itr + 1] * travd;  \
                    }                                                          \
                }                                                              \
            }                                                                  \
        }                                                                      \
    }

KIN_FWD(or_kin_fwd_f64, double, double)
KIN_FWD(or_kin_fwd_f32, float, float)
KIN_FWD(or_kin_fwd_x64y32, double, float)
KIN_ADJ(or_kin_adj_f64, double, double)
KIN_ADJ(or_kin_adj_f32, float, float)
KIN_ADJ(or_kin_adj_x64y32, double, float)

/* ------------------------------------------------------------------------
 * Kinematic, single combined table itrav/travd[ii*nsnr + isrcrec]
 * (precomputed as in K:395-398).
 * Forward  K:683-702 (_trav_kirch_matvec)
 * Adjoint  K:704-725 (_trav_kirch_rmatvec)
 * ---------------------------------------------------------------------- */
#define TRAV_FWD(NAME, TX, TY)                                                      \
    void NAME(const TX *x, TY *y, int nsnr, int nt, int64_t ni,                  \
              const int32_t *itrav, const double *travd)                       \
    {                                                                          \
        _Pragma("omp parallel for schedule(dynamic,8)")                        \
        for (int isr = 0; isr < nsnr; ++isr) {                                 \
            TY *yrow = y + (int64_t)isr * nt;                                   \
            for (int64_t ii = 0; ii < ni; ++ii) {                              \
                int32_t itr = itrav[ii * nsnr + isr];                          \
                double d = travd[ii * nsnr + isr];                             \
                if (0 <= itr && itr < nt - 1) {                                \
                    yrow[itr] += x[ii] * (1 - d);                              \
                    yrow[itr + 1] += x[ii] * d;                                \
                }                                                              \
            }                                                                  \
        }                                                                      \
    }

#define TRAV_ADJ(NAME, TX, TY)                                                      \
    void NAME(const TX *x, TY *y, int nsnr, int nt, int64_t ni,                  \
              const int32_t *itrav, const double *travd)                       \
    {                                                                          \
        _Pragma("omp parallel for schedule(static)")                           \
        for (int64_t ii = 0; ii < ni; ++ii) {                                  \
            for (int isr = 0; isr < nsnr; ++isr) {                             \
                int32_t itr = itrav[ii * nsnr + isr];                          \
                double d = travd[ii * nsnr + isr];                             \
                if (0 <= itr && itr < nt - 1) {                                \
                    const TX *xr = x + (int64_t)isr * nt;                       \
                    y[ii] += xr[itr] * (1 - d) + xr[itr + 1] * d;              \
                }                                                              \
            }                                                                  \
        }                                                                      \
    }

TRAV_FWD(or_trav_fwd_f64, double, double)
TRAV_FWD(or_trav_fwd_f32, float, float)
TRAV_FWD(or_trav_fwd_x64y32, double, float)
TRAV_ADJ(or_trav_adj_f64, double, double)
TRAV_ADJ(or_trav_adj_f32, float, float)
TRAV_ADJ(or_trav_adj_x64y32, double, float)

/* ------------------------------------------------------------------------
 * Dynamic (amplitude, obliquity and aperture weighted), separable tables.
 * Forward  K:783-889 (_ampsrcrec_kirch_matvec)
 * Adjoint  K:891-1001 (_ampsrcrec_kirch_rmatvec)
 * six/rix are passed as float64 (the reference holds int64 in 2D and float64
 * zeros in 3D; |six-rix| is exact either way for grid indices < 2^53).
 * ---------------------------------------------------------------------- */
static double dyn_aptap(double as, double ar, double apmin_ang, double apmax_ang,
                        double dang, double aperture, double apmin, double apmax,
                        double dap, const double *tap, int *pass)
{
    /* Returns the taper product and sets *pass like the nested ifs of
     * K:836-880: both |angles| below the max angle, then offset aperture
     * below the max aperture. */
    double aptap = 1.0;
    *pass = 0;
    if (as < apmax_ang && ar < apmax_ang) {
        if (as >= apmin_ang)
            aptap = aptap * tap[(int)py_floordiv(20 * (as - apmin_ang), dang)];
        if (ar >= apmin_ang)
            aptap = aptap * tap[(int)py_floordiv(20 * (ar - apmin_ang), dang)];
        if (aperture < apmax) {
            if (aperture >= apmin)
                aptap = aptap * tap[(int)(20 * py_floordiv(aperture - apmin, dap))];
            *pass = 1;
        }
    }
    return aptap;
}

#define DYN_FWD(NAME, TX, TY)                                                        \
    void NAME(const TX *x, TY *y, int ns, int nr, int nt, int64_t ni, double dt,  \
              const double *vel, const double *ts, const double *tr,            \
              const double *amps, const double *ampr, double apmin,             \
              double apmax, const double *tap, int nz, const double *six,       \
              const double *rix, double angmin, double angmax,                  \
              const double *angs, const double *angr)                           \
    {                                                                           \
        double dap = apmax - apmin;                                             \
        double dang = angmax - angmin;                                          \
        _Pragma("omp parallel for schedule(dynamic,1)")                         \
        for (int isrc = 0; isrc < ns; ++isrc) {                                 \
            for (int irec = 0; irec < nr; ++irec) {                             \
                int64_t isr = (int64_t)isrc * nr + irec;                        \
                TY *yrow = y + isr * nt;                                         \
                double sixv = six[isr], rixv = rix[isr];                        \
                for (int64_t ii = 0; ii < ni; ++ii) {                           \
                    double trav = ts[ii * ns + isrc] + tr[ii * nr + irec];      \
                    int32_t itr = np_astype_i32(trav / dt);                     \
                    double travd = trav / dt - itr;                             \
                    double a_s = angs[ii * ns + isrc];                          \
                    double a_r = angr[ii * nr + irec];                          \
                    double cosangle = cos((a_s - a_r) / 2.0);                   \
                    double damp = 2.0 * cosangle * amps[ii * ns + isrc] *       \
                                  ampr[ii * nr + irec] / vel[ii];               \
                    int iz = (int)(ii % nz);                                    \
                    double aperture = fabs(sixv - rixv) / (double)(iz + 1);     \
                    int pass;                                                   \
                    double aptap = dyn_aptap(fabs(a_s), fabs(a_r), angmin,      \
                                             angmax, dang, aperture, apmin,     \
                                             apmax, dap, tap, &pass);           \
                    if (pass && 0 <= itr && itr < nt - 1) {                     \
                        yrow[itr] += x[ii] * (1 - travd) * damp * aptap;        \
                        yrow[itr + 1] += x[ii] * travd * damp * aptap;          \
                    }                                                           \
                }                                                               \
            }                                                                   \
        }                                                                       \
    }

#define DYN_ADJ(NAME, TX, TY)                                                        \
    void NAME(const TX *x, TY *y, int ns, int nr, int nt, int64_t ni, double dt,  \
              const double *vel, const double *ts, const double *tr,            \
              const double *amps, const double *ampr, double apmin,             \
              double apmax, const double *tap, int nz, const double *six,       \
              const double *rix, double angmin, double angmax,                  \
              const double *angs, const double *angr)                           \
    {                                                                           \
        double dap = apmax - apmin;                                             \
        double dang = angmax - angmin;                                          \
        _Pragma("omp parallel for schedule(static)")                            \
        for (int64_t ii = 0; ii < ni; ++ii) {                                   \
            double velii = vel[ii];                                             \
            int iz = (int)(ii % nz);                                            \
            for (int isrc = 0; isrc < ns; ++isrc) {                             \
                double tsv = ts[ii * ns + isrc];                                \
                double asv = amps[ii * ns + isrc];                              \
                double a_s = angs[ii * ns + isrc];                              \
                for (int irec = 0; irec < nr; ++irec) {                         \
                    int64_t isr = (int64_t)isrc * nr + irec;                    \
                    double trav = tsv + tr[ii * nr + irec];                     \
                    double a_r = angr[ii * nr + irec];                          \
                    double cosangle = cos((a_s - a_r) / 2.0);                   \
                    double damp = 2.0 * cosangle * asv * ampr[ii * nr + irec] / \
                                  velii;                                        \
                    double aperture = fabs(six[isr] - rix[isr]) / (double)(iz + 1); \
                    int pass;                                                   \
                    double aptap = dyn_aptap(fabs(a_s), fabs(a_r), angmin,      \
                                             angmax, dang, aperture, apmin,     \
                                             apmax, dap, tap, &pass);           \
                    int64_t itr;                                                \
                    if (pass && py_int_in_range(trav / dt, nt, &itr)) {         \
                        double travd = trav / dt - (double)itr;                 \
                        const TX *xr = x + isr * nt;                             \
                        y[ii] += (xr[itr] * (1 - travd) + xr[itr + 1] * travd) * \
                                 damp * aptap;                                  \
                    }                                                           \
                }                                                               \
            }                                                                   \
        }                                                                       \
    }

DYN_FWD(or_dyn_fwd_f64, double, double)
DYN_FWD(or_dyn_fwd_f32, float, float)
DYN_FWD(or_dyn_fwd_x64y32, double, float)
DYN_ADJ(or_dyn_adj_f64, double, double)
DYN_ADJ(or_dyn_adj_f32, float, float)
DYN_ADJ(or_dyn_adj_x64y32, double, float)

/* ------------------------------------------------------------------------
 * "same"-mode 1-D convolution along the last axis of a [ntr, nt] array with
 * an odd-length filter h[nh] (Convolve1D always presents an odd filter after
 * its offset padding, convolve1d.py:78-86):
 *     y[n] = sum_k h[k] * x[n + (nh-1)/2 - k]
 * computed directly in float64.  The adjoint is the same call with h flipped
 * (convolve1d.py:94, 108-115).
 * ---------------------------------------------------------------------- */
void or_conv_same_f64(const double *x, double *y, int64_t ntr, int nt,
                      const double *h, int nh)
{
    int c = (nh - 1) / 2;
    _Pragma("omp parallel for schedule(static)")
    for (int64_t tr = 0; tr < ntr; ++tr) {
        const double *xr = x + tr * nt;
        double *yr = y + tr * nt;
        for (int n = 0; n < nt; ++n) {
            double acc = 0.0;
            for (int k = 0; k < nh; ++k) {
                int j = n + c - k;
                if (j >= 0 && j < nt)
                    acc += h[k] * xr[j];
            }
            yr[n] = acc;
        }
    }
}

int or_num_threads(void)
{
#ifdef _OPENMP
    extern int omp_get_max_threads(void);
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* Threads of the following parallel regions (the timed CPU baseline runs
 * at all host cores and at one thread). */
void or_set_num_threads(int n)
{
#ifdef _OPENMP
    extern void omp_set_num_threads(int);
    omp_set_num_threads(n);
#else
    (void)n;
#endif
}
