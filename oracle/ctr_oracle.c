/*
 * ctr_oracle.c -- CPU restatement of CTAnalyzer's reconstruction hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  See ctr_oracle.h for the shared
 * conventions and the reference citations of every entry point.  Compiled
 * with -ffp-contract=off so that the FFT restatement rounds like the
 * reference's own C++ (no fused multiply-adds).
 */
#include "ctr_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_MAX(a, b) ((a) > (b) ? (a) : (b))
#define ORC_MIN(a, b) ((a) < (b) ? (a) : (b))

static void set_threads(int threads)
{
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
}

double orc_reference_pi(void) { return (double)3.14154965f; }

/* ------------------------------------------------------------------------ */
/* FFT                                                                       */
/* ------------------------------------------------------------------------ */

static int floor_log2_i(int n)
{
    /* FastFourierTransformGPU.cpp:269-275 / FastFourierTransform.cpp:52-54 */
    int levels = 0;
    for (unsigned long t = (unsigned long)n; t > 1UL; t >>= 1) levels++;
    return levels;
}

static unsigned long reverse_bits_ul(unsigned long x, int levels)
{
    /* FastFourierTransform.cpp:42-47 */
    unsigned long r = 0;
    for (int i = 0; i < levels; i++, x >>= 1) r = (r << 1) | (x & 1UL);
    return r;
}

int orc_fft_radix2(double *re, double *im, int n, double pi)
{
    /* FastFourierTransform.cpp:49-88 */
    if (n < 1) return -1;
    int levels = floor_log2_i(n);
    if ((1L << levels) != (long)n) return -1;
    if (n == 1) return 0;

    int half = n / 2;
    double *tr = (double *)malloc(sizeof(double) * (size_t)half);
    double *ti = (double *)malloc(sizeof(double) * (size_t)half);
    double *br = (double *)malloc(sizeof(double) * (size_t)n);
    double *bi = (double *)malloc(sizeof(double) * (size_t)n);
    /* twiddle table polar(1, -2 pi k / n) (FastFourierTransform.cpp:58-61) */
    for (int k = 0; k < half; k++) {
        double phase = -2.0 * pi * (double)k / (double)n;
        tr[k] = cos(phase);
        ti[k] = sin(phase);
    }
    /* bit-reversal permutation by swaps (FastFourierTransform.cpp:63-67) */
    for (unsigned long i = 0; i < (unsigned long)n; i++) {
        unsigned long j = reverse_bits_ul(i, levels);
        if (j > i) {
            double t = re[i]; re[i] = re[j]; re[j] = t;
            t = im[i]; im[i] = im[j]; im[j] = t;
        }
    }
    /* stages, each reading a full copy of the previous stage
     * (FastFourierTransform.cpp:69-87) */
    memcpy(br, re, sizeof(double) * (size_t)n);
    memcpy(bi, im, sizeof(double) * (size_t)n);
    for (int size = 2; size <= n; size *= 2) {
        int hs = size / 2, step = n / size;
        for (int i = 0; i < n; i += size) {
            for (int j = i, k = 0; j < i + hs; j++, k += step) {
                double xr = br[j + hs], xi = bi[j + hs];
                double pr = xr * tr[k] - xi * ti[k];
                double pi_ = xr * ti[k] + xi * tr[k];
                re[j + hs] = br[j] + (-pr);
                im[j + hs] = bi[j] + (-pi_);
                re[j] = br[j] + pr;
                im[j] = bi[j] + pi_;
            }
        }
        if (size == n) break;
        memcpy(br, re, sizeof(double) * (size_t)n);
        memcpy(bi, im, sizeof(double) * (size_t)n);
    }
    free(tr); free(ti); free(br); free(bi);
    return 0;
}

int orc_fft_bit_reverse(double *re, double *im, int n)
{
    /* The swap loop alone: FastFourierTransform.cpp:63-67, which
     * FFFT::reverse_bit_swap runs on every texture row
     * (FastFourierTransformGPU.cpp:238-244). */
    if (n < 1) return -1;
    int levels = floor_log2_i(n);
    if ((1L << levels) != (long)n) return -1;
    for (unsigned long i = 0; i < (unsigned long)n; i++) {
        unsigned long j = reverse_bits_ul(i, levels);
        if (j > i) {
            double t = re[i]; re[i] = re[j]; re[j] = t;
            t = im[i]; im[i] = im[j]; im[j] = t;
        }
    }
    return 0;
}

int orc_fft_single_stage(const double *rre, const double *rim, double *wre,
                         double *wim, int n, int step_index, double pi)
{
    /* One butterfly stage, one output per element:
     * ParallelFastFourierTransform.cpp:262-277 (the CPU guide of
     * FFFT::fft_single_step, FastFourierTransformGPU.cpp:246-256), with the
     * n/2-entry table polar(1, -2 pi k / n) of :113-115. */
    if (n < 2) return -1;
    int levels = floor_log2_i(n);
    if ((1L << levels) != (long)n || step_index < 0 || (2L << step_index) > (long)n)
        return -1;
    const int size = 2 << step_index, half = size / 2, tablestep = n / size;
    for (int i = 0; i < n; i++) {
        const int local = i - (i / size) * size;
        const int k = (local % half) * tablestep;
        const double phase = -2.0 * pi * (double)k / (double)n;
        const double tr = cos(phase), ti = sin(phase);
        if (local >= half) {
            /* read[i - half] + (-1) * (read[i] * w) */
            const double pr = rre[i] * tr - rim[i] * ti;
            const double pi_ = rre[i] * ti + rim[i] * tr;
            wre[i] = rre[i - half] + (-pr);
            wim[i] = rim[i - half] + (-pi_);
        } else {
            /* read[i] + read[i + half] * w */
            const double pr = rre[i + half] * tr - rim[i + half] * ti;
            const double pi_ = rre[i + half] * ti + rim[i + half] * tr;
            wre[i] = rre[i] + pr;
            wim[i] = rim[i] + pi_;
        }
    }
    return 0;
}

int orc_inverse_fft_radix2(double *re, double *im, int n, double pi)
{
    /* FastFourierTransform.cpp:90-106: conj, FFT, conj, divide by n */
    for (int i = 0; i < n; i++) im[i] *= -1.0;
    if (orc_fft_radix2(re, im, n, pi) != 0) return -1;
    for (int i = 0; i < n; i++) im[i] *= -1.0;
    for (int i = 0; i < n; i++) {
        re[i] = re[i] / (double)n;
        im[i] = im[i] / (double)n;
    }
    return 0;
}

void orc_fft_shift(double *re, double *im, int n, int shift)
{
    /* FastFourierTransform.cpp:113-120.  The reference computes
     * (i + shift) % n in size_t, which equals the true modulo for
     * power-of-two n; the true modulo is used here for every n. */
    double *cr = (double *)malloc(sizeof(double) * (size_t)n);
    double *ci = (double *)malloc(sizeof(double) * (size_t)n);
    memcpy(cr, re, sizeof(double) * (size_t)n);
    memcpy(ci, im, sizeof(double) * (size_t)n);
    for (int i = 0; i < n; i++) {
        long t = ((long)i + (long)shift) % (long)n;
        if (t < 0) t += n;
        re[t] = cr[i];
        im[t] = ci[i];
    }
    free(cr); free(ci);
}

void orc_dft(const double *re, const double *im, double *ore, double *oim,
             int n, int sign)
{
    const long double two_pi = 6.283185307179586476925286766559L;
    for (int k = 0; k < n; k++) {
        long double sr = 0.0L, si = 0.0L;
        for (int t = 0; t < n; t++) {
            long m = ((long)k * (long)t) % (long)n;
            long double ph = (long double)sign * two_pi * (long double)m /
                             (long double)n;
            long double c = cosl(ph), s = sinl(ph);
            sr += (long double)re[t] * c - (long double)im[t] * s;
            si += (long double)re[t] * s + (long double)im[t] * c;
        }
        ore[k] = (double)sr;
        oim[k] = (double)si;
    }
}

/* ------------------------------------------------------------------------ */
/* Filter                                                                    */
/* ------------------------------------------------------------------------ */

double orc_filter_gain(int i, int n, int type, double pi)
{
    /* FastFourierTransformImage.cpp:35-37,45: the argument is formed in
     * float, the gain in double and multiplied by n. */
    float xf = (float)i / (float)n - 0.5f;
    double x = (double)xf;
    double g;
    switch (type) {
    case 0: g = fabs(x); break;                    /* RAM_LAK     */
    case 1: g = sin(fabs(x)); break;               /* SHEPP_LOGAN */
    case 2: g = fabs(sin(pi * x)) / pi; break;     /* COSINE      */
    default: return NAN;
    }
    return g * (double)n;
}

int orc_filter_rows(const float *in, int width, int rows, int npad, int type,
                    double pi, double *out, int threads)
{
    if (npad < width || width < 1 || type < 0 || type > 2) return -1;
    int lv = floor_log2_i(npad);
    if ((1L << lv) != (long)npad) return -1;
    double *gain = (double *)malloc(sizeof(double) * (size_t)npad);
    for (int i = 0; i < npad; i++) gain[i] = orc_filter_gain(i, npad, type, pi);
    set_threads(threads);
#pragma omp parallel
    {
        double *re = (double *)malloc(sizeof(double) * (size_t)npad);
        double *im = (double *)malloc(sizeof(double) * (size_t)npad);
#pragma omp for schedule(static)
        for (int r = 0; r < rows; r++) {
            /* blit real->complex, zero-pad (FastFourierTransformGPU.cpp:48-49) */
            for (int i = 0; i < npad; i++) {
                re[i] = i < width ? (double)in[(size_t)r * width + i] : 0.0;
                im[i] = 0.0;
            }
            orc_fft_radix2(re, im, npad, pi);
            orc_fft_shift(re, im, npad, npad / 2);
            for (int i = 0; i < npad; i++) { re[i] *= gain[i]; im[i] *= gain[i]; }
            orc_fft_shift(re, im, npad, -(npad / 2));
            orc_inverse_fft_radix2(re, im, npad, pi);
            /* crop the first width samples, real part
             * (FastFourierTransformGPU.cpp:79, FBP3D.cpp:411) */
            for (int i = 0; i < width; i++) out[(size_t)r * width + i] = re[i];
        }
        free(re); free(im);
    }
    free(gain);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Weights                                                                   */
/* ------------------------------------------------------------------------ */

void orc_fdk_weight(double *proj, int nu, int nv, int n_views, double sod,
                    double odd, double detector_width)
{
    double sdd = sod + odd, delta = detector_width / nu;
    for (int n = 0; n < n_views; n++)
        for (int iv = 0; iv < nv; iv++) {
            double v = (iv + 0.5 - 0.5 * nv) * delta;
            for (int iu = 0; iu < nu; iu++) {
                double u = (iu + 0.5 - 0.5 * nu) * delta;
                proj[((size_t)n * nv + iv) * nu + iu] *=
                    sdd / sqrt(sdd * sdd + u * u + v * v);
            }
        }
}

void orc_log_normalize(double *proj, long long count, double base,
                       double floor_value)
{
    for (long long i = 0; i < count; i++)
        proj[i] = -log(ORC_MAX(proj[i], floor_value) / base);
}

/* ------------------------------------------------------------------------ */
/* Backprojection                                                            */
/* ------------------------------------------------------------------------ */

static double bilinear0(const float *img, int nu, int nv, double fu, double fv)
{
    if (!(fu > -1.0 && fu < (double)nu && fv > -1.0 && fv < (double)nv))
        return 0.0;
    double fu0 = floor(fu), fv0 = floor(fv);
    int iu = (int)fu0, iv = (int)fv0;
    double wu = fu - fu0, wv = fv - fv0;
    double acc = 0.0;
    for (int b = 0; b < 2; b++) {
        int r = iv + b;
        if (r < 0 || r >= nv) continue;
        double wr = b ? wv : 1.0 - wv;
        for (int a = 0; a < 2; a++) {
            int c = iu + a;
            if (c < 0 || c >= nu) continue;
            double wc = a ? wu : 1.0 - wu;
            acc += wr * wc * (double)img[(size_t)r * nu + c];
        }
    }
    return acc;
}

/* Bilinear sample of a view given per cell: cell c = iv + 1 (iv = -1 ..
 * nv - 1) spans detector rows iv and iv + 1, with top[c] the values of row
 * iv and bot[c] the values of row iv + 1 as that cell sees them ((nv + 1) x
 * nu arrays each).  Zero outside the detector as bilinear0.  With
 * top[c] = row c - 1 and bot[c] = row c (zero rows outside) it is bilinear0
 * up to the order of the FP64 sum. */
static double bilinear_cells(const float *top, const float *bot, int nu, int nv,
                             double fu, double fv)
{
    if (!(fu > -1.0 && fu < (double)nu && fv > -1.0 && fv < (double)nv))
        return 0.0;
    double fu0 = floor(fu), fv0 = floor(fv);
    int iu = (int)fu0, c = (int)fv0 + 1;
    double wu = fu - fu0, wv = fv - fv0;
    const float *rt = top + (size_t)c * nu, *rb = bot + (size_t)c * nu;
    double acc = 0.0;
    for (int a = 0; a < 2; a++) {
        int col = iu + a;
        if (col < 0 || col >= nu) continue;
        double wc = a ? wu : 1.0 - wu;
        acc += wc * ((1.0 - wv) * (double)rt[col] + wv * (double)rb[col]);
    }
    return acc;
}

/* One view's sample: bilinear0 of proj_n, or bilinear_cells of the n-th
 * (top, bot) pair when bot is given. */
static double sample_view(const float *proj, const float *bot, int nu, int nv, int n,
                          double fu, double fv)
{
    if (!bot)
        return bilinear0(proj + (size_t)n * nu * nv, nu, nv, fu, fv);
    const size_t cells = (size_t)(nv + 1) * nu;
    return bilinear_cells(proj + (size_t)n * cells, bot + (size_t)n * cells, nu, nv, fu, fv);
}

static double view_angle(int n, int n_views, double rot, double offset)
{
    /* FBP3D.cpp:536-540 */
    return 2.0 * M_PI * rot * (double)n / (double)n_views + offset;
}

static void backproject_cone_box(const float *proj, const float *bot, int nu, int nv,
                                 int n_views, double sod, double odd,
                                 double detector_width, double volume_width,
                                 double rotation_count, double rotation_offset,
                                 double scale, int X, int Y, int Z, int i0,
                                 int i1, int j0, int j1, int k0, int k1,
                                 double *out, int threads)
{
    double sdd = sod + odd, delta = detector_width / nu, d = volume_width / X;
    int ni = i1 - i0, nj = j1 - j0;
    set_threads(threads);
#pragma omp parallel for schedule(dynamic, 1)
    for (int k = k0; k < k1; k++) {
        double z = (k + 0.5 - 0.5 * Z) * d;
        for (int n = 0; n < n_views; n++) {
            double th = view_angle(n, n_views, rotation_count, rotation_offset);
            double c = cos(th), s = sin(th);
            for (int j = j0; j < j1; j++) {
                double y = (j + 0.5 - 0.5 * Y) * d;
                for (int i = i0; i < i1; i++) {
                    double x = (i + 0.5 - 0.5 * X) * d;
                    double t = x * c + z * s;          /* depth toward source */
                    double sl = -x * s + z * c;        /* lateral coordinate  */
                    double L = sod - t;
                    double u = sdd * sl / L, v = sdd * y / L;
                    double fu = u / delta + 0.5 * nu - 0.5;
                    double fv = v / delta + 0.5 * nv - 0.5;
                    double w = sod / L;
                    out[((size_t)(k - k0) * nj + (j - j0)) * ni + (i - i0)] +=
                        scale * w * w * sample_view(proj, bot, nu, nv, n, fu, fv);
                }
            }
        }
    }
}

void orc_backproject_cone_region(const float *proj, int nu, int nv,
                                 int n_views, double sod, double odd,
                                 double detector_width, double volume_width,
                                 double rotation_count, double rotation_offset,
                                 double scale, int X, int Y, int Z, int i0,
                                 int i1, int j0, int j1, int k0, int k1,
                                 double *out, int threads)
{
    backproject_cone_box(proj, NULL, nu, nv, n_views, sod, odd, detector_width,
                         volume_width, rotation_count, rotation_offset, scale, X,
                         Y, Z, i0, i1, j0, j1, k0, k1, out, threads);
}

void orc_backproject_cone_cells(const float *top, const float *bot, int nu,
                                int nv, int n_views, double sod, double odd,
                                double detector_width, double volume_width,
                                double rotation_count, double rotation_offset,
                                double scale, int X, int Y, int Z, double *vol,
                                int threads)
{
    backproject_cone_box(top, bot, nu, nv, n_views, sod, odd, detector_width,
                         volume_width, rotation_count, rotation_offset, scale, X,
                         Y, Z, 0, X, 0, Y, 0, Z, vol, threads);
}

void orc_backproject_cone(const float *proj, int nu, int nv, int n_views,
                          double sod, double odd, double detector_width,
                          double volume_width, double rotation_count,
                          double rotation_offset, double scale, int X, int Y,
                          int Z, double *vol, int threads)
{
    orc_backproject_cone_region(proj, nu, nv, n_views, sod, odd,
                                detector_width, volume_width, rotation_count,
                                rotation_offset, scale, X, Y, Z, 0, X, 0, Y, 0,
                                Z, vol, threads);
}

static void backproject_parallel_all(const float *proj, const float *bot, int nu, int nv,
                                     int n_views, double detector_width,
                                     double volume_width, double rotation_count,
                                     double rotation_offset, double scale, int X,
                                     int Y, int Z, double *vol, int threads)
{
    double delta = detector_width / nu, d = volume_width / X;
    set_threads(threads);
#pragma omp parallel for schedule(dynamic, 1)
    for (int k = 0; k < Z; k++) {
        double z = (k + 0.5 - 0.5 * Z) * d;
        for (int n = 0; n < n_views; n++) {
            double th = view_angle(n, n_views, rotation_count, rotation_offset);
            double c = cos(th), s = sin(th);
            for (int j = 0; j < Y; j++) {
                double y = (j + 0.5 - 0.5 * Y) * d;
                for (int i = 0; i < X; i++) {
                    double x = (i + 0.5 - 0.5 * X) * d;
                    double sl = -x * s + z * c;
                    double fu = sl / delta + 0.5 * nu - 0.5;
                    double fv = y / delta + 0.5 * nv - 0.5;
                    vol[((size_t)k * Y + j) * X + i] +=
                        scale * sample_view(proj, bot, nu, nv, n, fu, fv);
                }
            }
        }
    }
}

void orc_backproject_parallel(const float *proj, int nu, int nv, int n_views,
                              double detector_width, double volume_width,
                              double rotation_count, double rotation_offset,
                              double scale, int X, int Y, int Z, double *vol,
                              int threads)
{
    backproject_parallel_all(proj, NULL, nu, nv, n_views, detector_width,
                             volume_width, rotation_count, rotation_offset, scale,
                             X, Y, Z, vol, threads);
}

void orc_backproject_parallel_cells(const float *top, const float *bot, int nu,
                                    int nv, int n_views, double detector_width,
                                    double volume_width, double rotation_count,
                                    double rotation_offset, double scale, int X,
                                    int Y, int Z, double *vol, int threads)
{
    backproject_parallel_all(top, bot, nu, nv, n_views, detector_width,
                             volume_width, rotation_count, rotation_offset, scale,
                             X, Y, Z, vol, threads);
}

void orc_backproject_cone_f32(const float *proj, int nu, int nv, int n_views,
                              double sod, double odd, double detector_width,
                              double volume_width, double rotation_count,
                              double rotation_offset, double scale, int X,
                              int Y, int Z, int z_begin, int z_end, float *vol,
                              int threads)
{
    /* Per-view projective rows in voxel-index space, built in FP64 and used
     * in FP32: den = L/SOD, fu*den, fv*den are affine in (i, j, k). */
    double sdd = sod + odd, delta = detector_width / nu, d = volume_width / X;
    double ox = 0.5 - 0.5 * X, oy = 0.5 - 0.5 * Y, oz = 0.5 - 0.5 * Z;
    double cu = 0.5 * nu - 0.5, cv = 0.5 * nv - 0.5, mag = sdd / (sod * delta);
    float *g = (float *)malloc(sizeof(float) * 9 * (size_t)n_views);
    for (int n = 0; n < n_views; n++) {
        double th = view_angle(n, n_views, rotation_count, rotation_offset);
        double c = cos(th), s = sin(th);
        double di = -c * d / sod, dk = -s * d / sod;
        double d0 = 1.0 - (c * d * ox + s * d * oz) / sod;
        float *q = g + 9 * (size_t)n;
        q[0] = (float)di; q[1] = (float)dk; q[2] = (float)d0;
        q[3] = (float)(-s * d * mag + cu * di);
        q[4] = (float)(c * d * mag + cu * dk);
        q[5] = (float)((-s * d * ox + c * d * oz) * mag + cu * d0);
        q[6] = (float)(cv * di);
        q[7] = (float)(cv * dk);
        q[8] = (float)(d * oy * mag + cv * d0);
    }
    float fvj = (float)(d * mag), fscale = (float)scale;
    set_threads(threads);
#pragma omp parallel for schedule(dynamic, 1)
    for (int k = z_begin; k < z_end; k++) {
        float *slice = vol + (size_t)(k - z_begin) * X * Y;
        for (int n = 0; n < n_views; n++) {
            const float *q = g + 9 * (size_t)n;
            const float *pv = proj + (size_t)n * nu * nv;
            for (int i = 0; i < X; i++) {
                float den = q[0] * i + q[1] * k + q[2];
                float inv = 1.0f / den;
                float fu = (q[3] * i + q[4] * k + q[5]) * inv;
                float w2 = inv * inv * fscale;
                float fv0 = (q[6] * i + q[7] * k + q[8]) * inv;
                float dv = fvj * inv;
                if (!(fu > -1.0f && fu < (float)nu)) continue;
                float fu0 = floorf(fu);
                int iu = (int)fu0;
                float wu = fu - fu0;
                for (int j = 0; j < Y; j++) {
                    float fv = fv0 + dv * (float)j;
                    if (!(fv > -1.0f && fv < (float)nv)) continue;
                    float fv0f = floorf(fv);
                    int iv = (int)fv0f;
                    float wv = fv - fv0f;
                    float t00 = 0, t01 = 0, t10 = 0, t11 = 0;
                    if (iv >= 0) {
                        const float *r = pv + (size_t)iv * nu;
                        if (iu >= 0) t00 = r[iu];
                        if (iu + 1 < nu) t01 = r[iu + 1];
                    }
                    if (iv + 1 < nv) {
                        const float *r = pv + (size_t)(iv + 1) * nu;
                        if (iu >= 0) t10 = r[iu];
                        if (iu + 1 < nu) t11 = r[iu + 1];
                    }
                    float top = t00 + wu * (t01 - t00);
                    float bot = t10 + wu * (t11 - t10);
                    slice[(size_t)j * X + i] += w2 * (top + wv * (bot - top));
                }
            }
        }
    }
    free(g);
}

/* ------------------------------------------------------------------------ */
/* Forward projection                                                        */
/* ------------------------------------------------------------------------ */

static double trilinear0(const float *vol, int X, int Y, int Z, double fi,
                         double fj, double fk)
{
    if (!(fi > -1.0 && fi < X && fj > -1.0 && fj < Y && fk > -1.0 && fk < Z))
        return 0.0;
    double i0 = floor(fi), j0 = floor(fj), k0 = floor(fk);
    int ii = (int)i0, jj = (int)j0, kk = (int)k0;
    double wi = fi - i0, wj = fj - j0, wk = fk - k0;
    double acc = 0.0;
    for (int c = 0; c < 2; c++) {
        int k = kk + c;
        if (k < 0 || k >= Z) continue;
        double wc = c ? wk : 1.0 - wk;
        for (int b = 0; b < 2; b++) {
            int j = jj + b;
            if (j < 0 || j >= Y) continue;
            double wb = b ? wj : 1.0 - wj;
            for (int a = 0; a < 2; a++) {
                int i = ii + a;
                if (i < 0 || i >= X) continue;
                double wa = a ? wi : 1.0 - wi;
                acc += wa * wb * wc * (double)vol[((size_t)k * Y + j) * X + i];
            }
        }
    }
    return acc;
}

/* Clip origin + t*dir (unit dir) against the box |x|<=hx, |y|<=hy, |z|<=hz. */
static int clip_box(const double *o, const double *dir, const double *h,
                    double *t_in, double *t_out)
{
    double tmin = -INFINITY, tmax = INFINITY;
    for (int a = 0; a < 3; a++) {
        if (dir[a] == 0.0) {
            if (o[a] < -h[a] || o[a] > h[a]) return 0;
            continue;
        }
        double t0 = (-h[a] - o[a]) / dir[a], t1 = (h[a] - o[a]) / dir[a];
        if (t0 > t1) { double t = t0; t0 = t1; t1 = t; }
        tmin = ORC_MAX(tmin, t0);
        tmax = ORC_MIN(tmax, t1);
    }
    if (!(tmax > tmin)) return 0;
    *t_in = tmin;
    *t_out = tmax;
    return 1;
}

static double march(const float *vol, int X, int Y, int Z, double d,
                    const double *o, const double *dir, double h)
{
    double half[3] = {(0.5 * X + 1.0) * d, (0.5 * Y + 1.0) * d,
                      (0.5 * Z + 1.0) * d};
    double t_in, t_out;
    if (!clip_box(o, dir, half, &t_in, &t_out)) return 0.0;
    int K = (int)ceil((t_out - t_in) / h);
    double acc = 0.0;
    for (int k = 0; k < K; k++) {
        double t = t_in + (k + 0.5) * h;
        double px = o[0] + t * dir[0], py = o[1] + t * dir[1],
               pz = o[2] + t * dir[2];
        acc += trilinear0(vol, X, Y, Z, px / d + 0.5 * X - 0.5,
                          py / d + 0.5 * Y - 0.5, pz / d + 0.5 * Z - 0.5);
    }
    return acc * h;
}

void orc_forward_cone(const float *vol, int X, int Y, int Z,
                      double volume_width, double sod, double odd,
                      double detector_width, int nu, int nv, int n_views,
                      double rotation_count, double rotation_offset,
                      double step, double *proj, int threads)
{
    double d = volume_width / X, delta = detector_width / nu;
    set_threads(threads);
#pragma omp parallel for collapse(2) schedule(dynamic, 4)
    for (int n = 0; n < n_views; n++)
        for (int iv = 0; iv < nv; iv++) {
            double th = view_angle(n, n_views, rotation_count, rotation_offset);
            double c = cos(th), s = sin(th);
            double src[3] = {sod * c, 0.0, sod * s};
            double v = (iv + 0.5 - 0.5 * nv) * delta;
            for (int iu = 0; iu < nu; iu++) {
                double u = (iu + 0.5 - 0.5 * nu) * delta;
                double pd[3] = {-odd * c - u * s, v, -odd * s + u * c};
                double dir[3] = {pd[0] - src[0], pd[1] - src[1], pd[2] - src[2]};
                double len = sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
                dir[0] /= len; dir[1] /= len; dir[2] /= len;
                proj[((size_t)n * nv + iv) * nu + iu] =
                    march(vol, X, Y, Z, d, src, dir, step * d);
            }
        }
}

void orc_forward_parallel(const float *vol, int X, int Y, int Z,
                          double volume_width, double detector_width, int nu,
                          int nv, int n_views, double rotation_count,
                          double rotation_offset, double step, double *proj,
                          int threads)
{
    double d = volume_width / X, delta = detector_width / nu;
    set_threads(threads);
#pragma omp parallel for collapse(2) schedule(dynamic, 4)
    for (int n = 0; n < n_views; n++)
        for (int iv = 0; iv < nv; iv++) {
            double th = view_angle(n, n_views, rotation_count, rotation_offset);
            double c = cos(th), s = sin(th);
            double v = (iv + 0.5 - 0.5 * nv) * delta;
            double dir[3] = {-c, 0.0, -s};
            for (int iu = 0; iu < nu; iu++) {
                double u = (iu + 0.5 - 0.5 * nu) * delta;
                double o[3] = {-u * s, v, u * c};
                proj[((size_t)n * nv + iv) * nu + iu] =
                    march(vol, X, Y, Z, d, o, dir, step * d);
            }
        }
}

void orc_rebin_fan_to_parallel(const float *fan, int nu_fan, int n_views,
                               double sod, double odd, double detector_width,
                               double rotation_count, double *par, int nu_par,
                               double parallel_width)
{
    /* Parallel ray (theta, s) is the fan ray of source angle
     * beta = theta + gamma at detector u = SDD tan(gamma), sin(gamma) =
     * s / SOD (flat detector at SDD, same angle convention as the
     * projectors).  Bilinear in (view, u); periodic in views for one full
     * turn, zero outside otherwise. */
    double sdd = sod + odd, delta = detector_width / nu_fan;
    double dth = 2.0 * M_PI * rotation_count / n_views;
    double pitch = parallel_width / nu_par;
    int periodic = fabs(rotation_count - 1.0) < 1e-12;
    for (int n = 0; n < n_views; n++)
        for (int is = 0; is < nu_par; is++) {
            double s = (is + 0.5 - 0.5 * nu_par) * pitch, out = 0.0;
            if (fabs(s) < sod) {
                double gam = asin(s / sod);
                double fb = (dth * n + gam) / dth;
                double fu = sdd * tan(gam) / delta + 0.5 * nu_fan - 0.5;
                if (fu > -1.0 && fu < nu_fan) {
                    double fb0 = floor(fb), fu0 = floor(fu);
                    for (int b = 0; b < 2; b++) {
                        long vb = (long)fb0 + b;
                        if (periodic) {
                            vb %= n_views;
                            if (vb < 0) vb += n_views;
                        } else if (vb < 0 || vb >= n_views) {
                            continue;
                        }
                        double wb = b ? fb - fb0 : 1.0 - (fb - fb0);
                        for (int a = 0; a < 2; a++) {
                            int c = (int)fu0 + a;
                            if (c < 0 || c >= nu_fan) continue;
                            double wa = a ? fu - fu0 : 1.0 - (fu - fu0);
                            out += wb * wa * (double)fan[(size_t)vb * nu_fan + c];
                        }
                    }
                }
            }
            par[(size_t)n * nu_par + is] = out;
        }
}

/* ------------------------------------------------------------------------ */
/* Analytic phantoms                                                         */
/* ------------------------------------------------------------------------ */

static void rotation(const double *e, double R[3][3])
{
    double ca = cos(e[6]), sa = sin(e[6]);
    double cb = cos(e[7]), sb = sin(e[7]);
    double cg = cos(e[8]), sg = sin(e[8]);
    double Rx[3][3] = {{1, 0, 0}, {0, ca, -sa}, {0, sa, ca}};
    double Ry[3][3] = {{cb, 0, sb}, {0, 1, 0}, {-sb, 0, cb}};
    double Rz[3][3] = {{cg, -sg, 0}, {sg, cg, 0}, {0, 0, 1}};
    double T[3][3];
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) {
            T[r][c] = 0;
            for (int m = 0; m < 3; m++) T[r][c] += Ry[r][m] * Rx[m][c];
        }
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) {
            R[r][c] = 0;
            for (int m = 0; m < 3; m++) R[r][c] += Rz[r][m] * T[m][c];
        }
}

double orc_ellipsoid_line_integral(const double *ellipsoids, int count,
                                   const double *origin, const double *dir)
{
    double total = 0.0;
    for (int e = 0; e < count; e++) {
        const double *E = ellipsoids + 10 * (size_t)e;
        double R[3][3];
        rotation(E, R);
        double p[3] = {origin[0] - E[0], origin[1] - E[1], origin[2] - E[2]};
        double P[3], D[3];
        for (int a = 0; a < 3; a++) { /* body = R^T world, then / semi-axis */
            P[a] = (R[0][a] * p[0] + R[1][a] * p[1] + R[2][a] * p[2]) / E[3 + a];
            D[a] = (R[0][a] * dir[0] + R[1][a] * dir[1] + R[2][a] * dir[2]) / E[3 + a];
        }
        double A = D[0] * D[0] + D[1] * D[1] + D[2] * D[2];
        double B = 2.0 * (P[0] * D[0] + P[1] * D[1] + P[2] * D[2]);
        double C = P[0] * P[0] + P[1] * P[1] + P[2] * P[2] - 1.0;
        double disc = B * B - 4.0 * A * C;
        if (disc <= 0.0 || A <= 0.0) continue;
        double dt = sqrt(disc) / A; /* t2 - t1 */
        double dl = sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
        total += E[9] * dt * dl;
    }
    return total;
}

void orc_analytic_cone(const double *ellipsoids, int count, double sod,
                       double odd, double detector_width, int nu, int nv,
                       int n_views, double rotation_count,
                       double rotation_offset, double *proj, int threads)
{
    double delta = detector_width / nu;
    set_threads(threads);
#pragma omp parallel for collapse(2) schedule(dynamic, 4)
    for (int n = 0; n < n_views; n++)
        for (int iv = 0; iv < nv; iv++) {
            double th = view_angle(n, n_views, rotation_count, rotation_offset);
            double c = cos(th), s = sin(th);
            double src[3] = {sod * c, 0.0, sod * s};
            double v = (iv + 0.5 - 0.5 * nv) * delta;
            for (int iu = 0; iu < nu; iu++) {
                double u = (iu + 0.5 - 0.5 * nu) * delta;
                double dir[3] = {-odd * c - u * s - src[0], v - src[1],
                                 -odd * s + u * c - src[2]};
                proj[((size_t)n * nv + iv) * nu + iu] =
                    orc_ellipsoid_line_integral(ellipsoids, count, src, dir);
            }
        }
}

void orc_analytic_parallel(const double *ellipsoids, int count,
                           double detector_width, int nu, int nv, int n_views,
                           double rotation_count, double rotation_offset,
                           double *proj, int threads)
{
    double delta = detector_width / nu;
    set_threads(threads);
#pragma omp parallel for collapse(2) schedule(dynamic, 4)
    for (int n = 0; n < n_views; n++)
        for (int iv = 0; iv < nv; iv++) {
            double th = view_angle(n, n_views, rotation_count, rotation_offset);
            double c = cos(th), s = sin(th);
            double v = (iv + 0.5 - 0.5 * nv) * delta;
            double dir[3] = {-c, 0.0, -s};
            for (int iu = 0; iu < nu; iu++) {
                double u = (iu + 0.5 - 0.5 * nu) * delta;
                double o[3] = {-u * s, v, u * c};
                proj[((size_t)n * nv + iv) * nu + iu] =
                    orc_ellipsoid_line_integral(ellipsoids, count, o, dir);
            }
        }
}

void orc_render_ellipsoids(const double *ellipsoids, int count, int X, int Y,
                           int Z, double volume_width, double *vol,
                           int threads)
{
    double d = volume_width / X;
    set_threads(threads);
#pragma omp parallel for schedule(dynamic, 1)
    for (int k = 0; k < Z; k++) {
        double z = (k + 0.5 - 0.5 * Z) * d;
        for (int e = 0; e < count; e++) {
            const double *E = ellipsoids + 10 * (size_t)e;
            double R[3][3];
            rotation(E, R);
            for (int j = 0; j < Y; j++) {
                double y = (j + 0.5 - 0.5 * Y) * d;
                for (int i = 0; i < X; i++) {
                    double x = (i + 0.5 - 0.5 * X) * d;
                    double p[3] = {x - E[0], y - E[1], z - E[2]};
                    double q = 0.0;
                    for (int a = 0; a < 3; a++) {
                        double b = (R[0][a] * p[0] + R[1][a] * p[1] +
                                    R[2][a] * p[2]) / E[3 + a];
                        q += b * b;
                    }
                    if (q <= 1.0) vol[((size_t)k * Y + j) * X + i] += E[9];
                }
            }
        }
    }
}
