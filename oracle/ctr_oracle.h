/*
 * ctr_oracle.h -- CPU restatement of CTAnalyzer's reconstruction hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: tests/, the smoke
 * check in __graft_entry__.py and the cpu_baseline leg of bench.py may load
 * it.  The product library never links, loads or calls it.
 *
 * Every routine cites the reference file:line it restates.  Paths are
 * relative to the reference repository root (furkanunsal42/
 * 3DFilteredBackProjection-Public).  The GLSL kernels of the reference are
 * empty files (README.md:2), so the arithmetic of the projectors,
 * backprojectors and weights follows the host code's uniform names plus
 * textbook FDK (Feldkamp, Davis, Kress 1984; Kak & Slaney ch. 3); those
 * conventions are stated next to each function.
 *
 * Conventions shared with the HIP library (all lengths in the caller's unit,
 * usually mm):
 *   view angle      theta_n = 2*pi*rotation_count*n/N + rotation_offset
 *                   (Source/FilteredBackProjection/FBP3D.cpp:536-540)
 *   volume          (X, Y, Z), x fastest, Y = rotation axis, isotropic pitch
 *                   d = volume_width / X (FBP3D.cpp:26,531); voxel centre
 *                   x = (i + 0.5 - X/2) d, likewise y, z
 *   projections     (nu, nv, N), u fastest (FBP3D.cpp:40); detector pitch
 *                   delta = detector_width / nu, square pixels; pixel centre
 *                   u = (iu + 0.5 - nu/2) delta, v = (iv + 0.5 - nv/2) delta
 *   sinograms       (nu, N, nv) (FBP3D.cpp:54)
 *   cone geometry   source at SOD (cos t, 0, sin t); flat detector at
 *                   distance detector_panel_distance on the far side of the
 *                   rotation axis, SDD = SOD + detector_panel_distance
 *                   (FBP3D.cpp:290-293), u axis (-sin t, 0, cos t), v axis +y
 *   2D              an image (W, W) is the y = 0 slice of a volume (W, 1, W);
 *                   a sinogram (nu, N) is projections with nv = 1
 *   interpolation   bilinear (backprojection) / trilinear (forward
 *                   projection) with zero outside the sampled grid
 */
#ifndef CTR_ORACLE_H
#define CTR_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* pi of the reference CPU FFT: (double)3.14154965f
 * (Source/FastFourierTransformCPU/FastFourierTransform.h:9). */
double orc_reference_pi(void);

/* ---- FFT (Source/FastFourierTransformCPU/FastFourierTransform.cpp:49-125) */
/* In-place radix-2 DIT FFT with the reference's n/2 twiddle table,
 * bit-reversal swap and per-stage buffer copy.  `pi` selects the twiddle
 * constant (M_PI or orc_reference_pi()).  Returns -1 if n is not a power of
 * two (FastFourierTransform.cpp:55-56), else 0. */
int orc_fft_radix2(double *re, double *im, int n, double pi);
/* The bit-reversal swap pass of the FFT on its own
 * (FastFourierTransform.cpp:63-67; FFFT::reverse_bit_swap,
 * FastFourierTransformGPU.cpp:238-244).  -1 unless n is a power of two. */
int orc_fft_bit_reverse(double *re, double *im, int n);
/* One radix-2 stage of span 2 << step_index from read into write, one
 * output per element (ParallelFastFourierTransform.cpp:262-277;
 * FFFT::fft_single_step, FastFourierTransformGPU.cpp:246-256).  Bit
 * reversal followed by stages 0 .. log2(n) - 1 is the FFT. */
int orc_fft_single_stage(const double *rre, const double *rim, double *wre,
                         double *wim, int n, int step_index, double pi);
/* conj -> FFT -> conj -> /n (FastFourierTransform.cpp:90-106). */
int orc_inverse_fft_radix2(double *re, double *im, int n, double pi);
/* dst[(i + shift) mod n] = src[i] (FastFourierTransform.cpp:108-125). */
void orc_fft_shift(double *re, double *im, int n, int shift);
/* Direct O(n^2) DFT in long double; sign = -1 forward, +1 backward
 * (unnormalised).  Not a restatement: the closed form that pins the FFT. */
void orc_dft(const double *re, const double *im, double *ore, double *oim,
             int n, int sign);

/* ---- Fourier-space filter ------------------------------------------------ */
/* Filter type ids follow FBP2D::FilterType (Source/FilteredBackProjection/
 * FBP2D.h:10-14): 0 RAM_LAK, 1 SHEPP_LOGAN, 2 COSINE. */
/* Gain at index i of the fft-shifted spectrum of length n:
 * n * g((float)i/n - 0.5f) with g = |x|, sin|x|, |sin(pi x)|/pi
 * (Source/FastFourierTransformCPU/FastFourierTransformImage.cpp:35-37,45). */
double orc_filter_gain(int i, int n, int type, double pi);
/* Full row filter of Source/FilteredBackProjection/FBP3D.cpp:399-413:
 * real -> complex, zero-pad to npad (FastFourierTransformGPU.cpp:38-52),
 * FFT, fft_shift, x gain, inverse shift, inverse FFT, crop the first width
 * samples (FastFourierTransformGPU.cpp:66-81), real part.
 * in: rows x width float; out: rows x width double.  Returns 0, or -1 when
 * npad is not a power of two or npad < width, or type is unknown. */
int orc_filter_rows(const float *in, int width, int rows, int npad, int type,
                    double pi, double *out, int threads);

/* ---- Detector-domain weights ------------------------------------------- */
/* FDK cosine pre-weight SDD / sqrt(SDD^2 + u^2 + v^2) applied in place to
 * projections (nu, nv, N) (FBP3D.cpp:369-387, ct3d_fdk_weight). */
void orc_fdk_weight(double *proj, int nu, int nv, int n_views, double sod,
                    double odd, double detector_width);
/* p = -ln(max(I, floor) / base) in place (FBP3D.cpp:333-349,
 * ct3d_log_normalize). */
void orc_log_normalize(double *proj, long long count, double base,
                       double floor_value);

/* ---- Backprojection (voxel-driven) ---------------------------------------
 * vol[(k*Y + j)*X + i] += scale * w^2 * bilinear(proj_n, fu, fv) summed over
 * views, FP64 throughout.  Cone: w = SOD / (SOD - t) with t the voxel's
 * depth toward the source (FBP3D.cpp:524-559,
 * ..._cone_equidistance_fast).  Parallel: w = 1 (FBP3D.cpp:460-490). */
void orc_backproject_cone(const float *proj, int nu, int nv, int n_views,
                          double sod, double odd, double detector_width,
                          double volume_width, double rotation_count,
                          double rotation_offset, double scale, int X, int Y,
                          int Z, double *vol, int threads);
/* The same sum over the voxel box [i0,i1) x [j0,j1) x [k0,k1) of an
 * (X, Y, Z) grid only; out is (k1-k0, j1-j0, i1-i0), accumulated into. */
void orc_backproject_cone_region(const float *proj, int nu, int nv,
                                 int n_views, double sod, double odd,
                                 double detector_width, double volume_width,
                                 double rotation_count, double rotation_offset,
                                 double scale, int X, int Y, int Z, int i0,
                                 int i1, int j0, int j1, int k0, int k1,
                                 double *out, int threads);
void orc_backproject_parallel(const float *proj, int nu, int nv, int n_views,
                              double detector_width, double volume_width,
                              double rotation_count, double rotation_offset,
                              double scale, int X, int Y, int Z, double *vol,
                              int threads);
/* The same sums with each view given per interpolation cell: top and bot
 * are (n_views, nv + 1, nu); cell c = iv + 1 (iv = -1 .. nv - 1) spans
 * detector rows iv and iv + 1, top[n][c] holding row iv and bot[n][c] row
 * iv + 1 as that cell sees them.  This is the dequantized model of the
 * backprojector's 16-bit value-and-step texels (row iv + 1 seen from cell
 * iv is value + step of row iv, not row iv + 1's own value).  With
 * top[n][c] = row c - 1 and bot[n][c] = row c (zero outside the detector)
 * it is orc_backproject_cone / _parallel. */
void orc_backproject_cone_cells(const float *top, const float *bot, int nu,
                                int nv, int n_views, double sod, double odd,
                                double detector_width, double volume_width,
                                double rotation_count, double rotation_offset,
                                double scale, int X, int Y, int Z, double *vol,
                                int threads);
void orc_backproject_parallel_cells(const float *top, const float *bot, int nu,
                                    int nv, int n_views, double detector_width,
                                    double volume_width, double rotation_count,
                                    double rotation_offset, double scale, int X,
                                    int Y, int Z, double *vol, int threads);
/* FP32 voxel loop with FP64 per-view geometry: the CPU baseline timing path
 * (same conventions, float accumulation). */
void orc_backproject_cone_f32(const float *proj, int nu, int nv, int n_views,
                              double sod, double odd, double detector_width,
                              double volume_width, double rotation_count,
                              double rotation_offset, double scale, int X,
                              int Y, int Z, int z_begin, int z_end, float *vol,
                              int threads);

/* ---- Forward projection (ray-driven) ------------------------------------
 * One ray per detector pixel centre.  The ray is clipped against the volume
 * box grown by one voxel on every side; samples sit at
 * t_k = t_in + (k + 1/2) h, k < ceil((t_out - t_in)/h), h = step * d, and
 * the line integral is h * sum(trilinear(volume, P(t_k))).
 * Cone: FBP3D.cpp:269-299; parallel: FBP3D.cpp:212-240. */
void orc_forward_cone(const float *vol, int X, int Y, int Z,
                      double volume_width, double sod, double odd,
                      double detector_width, int nu, int nv, int n_views,
                      double rotation_count, double rotation_offset,
                      double step, double *proj, int threads);
void orc_forward_parallel(const float *vol, int X, int Y, int Z,
                          double volume_width, double detector_width, int nu,
                          int nv, int n_views, double rotation_count,
                          double rotation_offset, double step, double *proj,
                          int threads);

/* ---- Fan -> parallel rebinning (FBP2D::parallelize_fan_beam_sinogram,
 * Source/FilteredBackProjection/FBP2D.cpp:31-52; the reference's GLSL is
 * empty, the resampling rule is stated in ctr_oracle.c).
 * fan (N, nu_fan) over 2*pi*rotation_count -> par (N, nu_par) with pitch
 * parallel_width / nu_par, same view angles. */
void orc_rebin_fan_to_parallel(const float *fan, int nu_fan, int n_views,
                               double sod, double odd, double detector_width,
                               double rotation_count, double *par, int nu_par,
                               double parallel_width);

/* ---- Analytic phantoms ----------------------------------------------------
 * Ellipsoid record (10 doubles): cx cy cz  a b c  alpha beta gamma  density,
 * world units.  Body-to-world rotation R = Rz(gamma) Ry(beta) Rx(alpha).
 * (FBP3D.cpp:609-628 passes a centre, three semi-axes, three angles and a
 * value; the reference's Euler order is not recoverable.) */
/* Exact line integral sum_e density_e * chord_e along the ray
 * origin + t*dir (dir need not be unit; the integral is over length). */
double orc_ellipsoid_line_integral(const double *ellipsoids, int count,
                                   const double *origin, const double *dir);
/* Analytic cone / parallel projections of ellipsoids, same ray layout as
 * the forward projectors above. */
void orc_analytic_cone(const double *ellipsoids, int count, double sod,
                       double odd, double detector_width, int nu, int nv,
                       int n_views, double rotation_count,
                       double rotation_offset, double *proj, int threads);
void orc_analytic_parallel(const double *ellipsoids, int count,
                           double detector_width, int nu, int nv, int n_views,
                           double rotation_count, double rotation_offset,
                           double *proj, int threads);
/* Voxelise: add density to every voxel whose centre lies inside
 * (FBP3D.cpp:609-628, ct3d_render_ellipsoid). */
void orc_render_ellipsoids(const double *ellipsoids, int count, int X, int Y,
                           int Z, double volume_width, double *vol,
                           int threads);

#ifdef __cplusplus
}
#endif
#endif
