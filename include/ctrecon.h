/*
 * ctrecon.h -- C ABI of the MI355X (gfx950) CT reconstruction library.
 *
 * This is the drop-in boundary under the C++ classes FBP2D, FBP3D and FFFT
 * (the reference's Source/FilteredBackProjection/FBP2D.h, FBP3D.h and
 * Source/FastFourierTransformGPU/FastFourierTransformGPU.h).  Every entry
 * point below names the reference interface it replaces.  Signatures hold
 * only plain C types: device pointers are `float*` / `void*` into HIP device
 * memory owned by the caller, streams are `void*` (a hipStream_t, or NULL for
 * the default stream), geometry is plain doubles with the reference's
 * parameter meaning.  All calls are asynchronous on the given stream unless
 * stated otherwise.
 *
 * Return value: CTR_OK (0) or a negative CTR_ERR_* code; ctr_last_error()
 * returns the calling thread's last message.
 *
 * Conventions (shared with the CPU oracle):
 *   view angle    theta_n = 2*pi*rotation_count*n/N + rotation_offset
 *   volume        (X, Y, Z), x fastest, Y = rotation axis, pitch
 *                 d = volume_width / X; voxel centre (i + 1/2 - X/2) d, ...
 *   projections   layout CTR_LAYOUT_PROJECTIONS = (nu, nv, N), u fastest;
 *                 layout CTR_LAYOUT_SINOGRAMS   = (nu, N, nv)
 *   detector      pitch delta = detector_width / nu (square pixels);
 *                 pixel centre u = (iu + 1/2 - nu/2) delta
 *   cone          source at SOD (cos t, 0, sin t); detector plane at
 *                 detector_distance beyond the axis; SDD = SOD + that
 *   2D            an image (W, W) is a volume (W, 1, W); a sinogram (nu, N)
 *                 is projections with nv = 1
 */
#ifndef CTRECON_H
#define CTRECON_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CTR_OK 0
#define CTR_ERR_INVALID_ARGUMENT (-1)
#define CTR_ERR_NOT_POWER_OF_TWO (-2)
#define CTR_ERR_OUT_OF_MEMORY (-3)
#define CTR_ERR_HIP (-4)
#define CTR_ERR_NULL_POINTER (-5)
#define CTR_ERR_OVER_BUDGET (-6)

/* element types of staged (filtered) projections: FP32, or 16-bit
 * value-and-step texels with a per-view scale (layout below) */
#define CTR_F32 0
#define CTR_F16 1
#define CTR_BF16 2

/* workspace options (ctr_workspace_set_option), all diagnostic kernel
 * choices */
#define CTR_OPTION_BACKPROJECT_KERNEL 0 /* CTR_KERNEL_AUTO (default), _LDS or _GATHER */
#define CTR_OPTION_FORWARD_KERNEL 1     /* CTR_KERNEL_AUTO (default: slab) or _RAY */
#define CTR_OPTION_FORWARD_ADDRESSING 2 /* CTR_KERNEL_AUTO (default) or _POINTER */
#define CTR_OPTION_COUNT 3

/* kernel choices of the diagnostic options */
#define CTR_KERNEL_AUTO 0
#define CTR_KERNEL_LDS 1     /* backprojection: LDS-staged kernel at any size */
#define CTR_KERNEL_GATHER 2  /* backprojection: direct-tap kernel */
#define CTR_KERNEL_RAY 3     /* forward projection: ray kernel instead of slab */
#define CTR_KERNEL_POINTER 4 /* forward projection: 64-bit addressing */

/* FBP2D::FilterType (Source/FilteredBackProjection/FBP2D.h:10-14) */
#define CTR_RAM_LAK 0
#define CTR_SHEPP_LOGAN 1
#define CTR_COSINE 2

#define CTR_GEOMETRY_CONE 0
#define CTR_GEOMETRY_PARALLEL 1

#define CTR_LAYOUT_PROJECTIONS 0
#define CTR_LAYOUT_SINOGRAMS 1

/* blit modes of FFFT::blit_texture_* (FastFourierTransformGPU.cpp:144-186) */
#define CTR_BLIT_REAL_TO_COMPLEX 0
#define CTR_BLIT_COMPLEX_TO_COMPLEX 1
#define CTR_BLIT_COMPLEX_TO_REAL_R 2
#define CTR_BLIT_COMPLEX_TO_REAL_I 3

typedef struct ctr_workspace ctr_workspace;

/* ---- library ------------------------------------------------------------ */
/* Thread-local text of the last error ("" if none).  Stands in for the
 * reference's "[CTAnalyzer Error] ..." console messages (e.g.
 * Source/FilteredBackProjection/FBP3D.cpp:214-217). */
const char *ctr_last_error(void);
int ctr_version(void);
/* Waits for all work on the device (no reference counterpart; the GL path
 * synchronised implicitly). */
int ctr_synchronize(void);

/* ---- workspace (scratch owned by the library) ------------------------------
 * Per-view geometry tables, tile footprint tables, per-row maxima and the
 * FP32 staging buffer of 16-bit filtering live here so that callers never
 * see allocations inside hot loops.  Passing NULL to any entry point uses a
 * per-device default workspace.  A workspace grows on demand (hipMalloc,
 * synchronous) and never shrinks: one warm-up call of a problem leaves later
 * calls of that size allocation-free (ctr_workspace_reserve pre-sizes the
 * per-view geometry tables for n_views).  No reference
 * counterpart: the reference allocated textures per call
 * (FastFourierTransformGPU.cpp:15,74,115,138). */
int ctr_workspace_create(ctr_workspace **out);
int ctr_workspace_destroy(ctr_workspace *ws);
int ctr_workspace_reserve(ctr_workspace *ws, int n_views, int max_fft_length);

/* Per-workspace options: diagnostic kernel choices (which backprojection or
 * forward-projection kernel runs, and the forward projector's addressing);
 * results are the same under every choice, to the tolerances the kernels'
 * tests state.  Every entry point reads the options of the workspace it is
 * given when it launches, so two threads driving two streams with two
 * workspaces can use different settings at the same time; a HIP graph keeps
 * the kernels it captured.  A new workspace -- and the per-device default
 * workspace that NULL selects -- starts from the environment as read once
 * when the library is loaded: CTR_BACKPROJECT_KERNEL=lds|gather,
 * CTR_FORWARD_KERNEL=ray, CTR_FORWARD_ADDRESSING=pointer.  The default
 * workspace is shared by every NULL caller on a device, so its options
 * cannot be changed (CTR_ERR_INVALID_ARGUMENT for ws == NULL).
 * ctr_workspace_set_option returns the previous value.  No reference
 * counterpart (one GL context, one set of shaders). */
int ctr_workspace_set_option(ctr_workspace *ws, int option, int value);
/* The option's value for ws (NULL: the default workspaces' value, which is
 * the one read from the environment at library load). */
int ctr_workspace_get_option(ctr_workspace *ws, int option);

/* ---- FFT family: FFFT (Source/FastFourierTransformGPU/) --------------------
 * Complex data are interleaved float pairs (RG32F texture), rows of n
 * complex values, `rows` rows. */

/* FFFT::fft_radix2 (FastFourierTransformGPU.cpp:3-28): in-place forward FFT
 * of every row, exp(-2 pi i k/n) kernel, unnormalised.  n must be a power of
 * two in [1, 8192] (CTR_ERR_NOT_POWER_OF_TWO otherwise, as at :10-13). */
int ctr_fft_radix2(void *stream, float *data, int n, int rows);
/* FFFT::inverse_fft_radix2 (FastFourierTransformGPU.cpp:30-36):
 * conj -> FFT -> conj -> divide by n, in place. */
int ctr_inverse_fft_radix2(void *stream, float *data, int n, int rows);
/* FFFT::fft_shift(src, dst, shift) (FastFourierTransformGPU.cpp:103-111):
 * dst[(i + shift) mod n] = src[i] per row; src == dst is allowed (the
 * reference's in-place overload, :113-119). */
int ctr_fft_shift(void *stream, const float *src, float *dst, int n, int rows,
                  int shift);
/* FFFT::blit_texture_* (FastFourierTransformGPU.cpp:144-186): copies the
 * min(src, dst) extent; real textures hold 1 float per texel, complex 2.
 * Texels of dst outside the copied extent are left untouched. */
int ctr_blit(void *stream, int mode, const float *src, int src_width,
             int src_height, float *dst, int dst_width, int dst_height);
/* FFFT::conjugate_complex_texture / multiply_complex_texture /
 * divide_complex_texture (FastFourierTransformGPU.cpp:212-236); count is
 * the number of complex values. */
int ctr_conjugate(void *stream, float *data, size_t count);
int ctr_scale_complex(void *stream, float *data, size_t count,
                      float multiplier);

/* FFFT::reverse_bit_swap (FastFourierTransformGPU.cpp:238-244): in-place
 * bit-reversal permutation of every row (n a power of two). */
int ctr_fft_bit_reverse(void *stream, float *data, int n, int rows);
/* FFFT::fft_single_step (FastFourierTransformGPU.cpp:246-256): one radix-2
 * DIT stage of span 2 << step_index from `read` into `write` (distinct
 * buffers), one output per element, as in the reference's CPU guide
 * (Source/FastFourierTransformCPU/ParallelFastFourierTransform.cpp:274-290). */
int ctr_fft_single_stage(void *stream, const float *read, float *write, int n, int rows,
                         int step_index);
/* Real-texture scaling, as FBP2D::_int_texture_div divides by `denominator`
 * (FBP2D.cpp:189-195): data[i] *= multiplier. */
int ctr_scale_real(void *stream, float *data, size_t count, float multiplier);

/* FBP2D::apply_filter (Source/FilteredBackProjection/FBP2D.cpp:124-144):
 * multiplies the fft-SHIFTED spectrum rows in place by
 * H[i] = n * g(i/n - 1/2), g = |x| (RAM_LAK), sin|x| (SHEPP_LOGAN),
 * |sin(pi x)|/pi (COSINE) (FastFourierTransformCPU/
 * FastFourierTransformImage.cpp:35-37,45).  CTR_ERR_INVALID_ARGUMENT for an
 * unknown filter (FBP2D.cpp:135-137). */
int ctr_apply_filter(void *stream, float *spectrum, int n, int rows,
                     int filter);

/* ---- fused row filter ------------------------------------------------------
 * The whole per-row chain of FBP3D::apply_filter_to_projections /
 * apply_filter_to_sinograms (FBP3D.cpp:389-442) in one LDS-resident pass:
 * zero-pad to fft_length, FFT, x H (unshifted index form of the shifted
 * gain), inverse FFT, crop the first `width` samples, real part.  Two real
 * rows share one complex FFT (H is real and even).  fft_length 0 selects the
 * reference default 2^ceil(log2 width) (FastFourierTransformGPU.cpp:258-280);
 * any power of two >= width up to 8192 is accepted.  in may equal out. */
int ctr_filter_rows(void *stream, const float *in, float *out, int width,
                    int rows, int fft_length, int filter);

/* FDK pre-weight and ramp filter of all projections, written in the staged
 * layout the backprojector reads (below).  fdk_weight != 0 multiplies each
 * sample by SDD / sqrt(SDD^2 + u^2 + v^2) (FP64) before filtering
 * (FBP3D.cpp:369-387).  For CTR_F16 / CTR_BF16 every view is scaled by its
 * max |value| (stored in view_scale[N], device memory; 1 for a view that is
 * all zero) before rounding; for CTR_F32 view_scale may be NULL.  `layout`
 * gives the input layout; both layouts give the same staged buffer bit for
 * bit (rows are paired for the shared complex FFTs in projection order,
 * view by view, in either layout). */
int ctr_filter_projections(void *stream, ctr_workspace *ws, const float *proj,
                           int layout, int nu, int nv, int n_views,
                           int fft_length, int filter, int fdk_weight,
                           double source_distance, double detector_distance,
                           double detector_width, void *staged,
                           int staged_dtype, float *view_scale);

/* ---- staged projections ------------------------------------------------------
 * View-major, zero-bordered copy of the filtered projections that the
 * backprojector samples.  For CTR_F32 texel (iu, iv, n) is
 *   staged[(n * pitch_v + iv + 1) * pitch_u + iu + 1],
 * with one zero column/row before the data, two zero rows and at least two
 * zero columns after it; pitch_u is a multiple of 32 texels (128-byte rows).
 * For CTR_F16 and CTR_BF16 each 32-bit texel at staged (c, r) is a
 * value-and-step texel: its lower 16 bits hold v = q[r][c] / view_scale[n]
 * rounded to the 16-bit type, its upper 16 bits the step d to the next
 * staged row, q[r + 1][c] / view_scale[n] - v, so one load gives a column's
 * two vertical taps v and v + d (q taken as 0 past the view's last staged
 * row; the border row r = 0 therefore holds (0, first data row)).  FP16:
 * d rounded to FP16.  BF16: the upper half is chosen so that the whole word
 * read as an FP32 value is nearest to d.  No reference counterpart (the
 * reference sampled its Texture2DArray directly, FBP3D.cpp:543). */
int ctr_staged_layout(int nu, int nv, int n_views, int dtype, int *pitch_u,
                      int *pitch_v, size_t *bytes);
/* Copies plain projections (either layout) into the staged layout,
 * computing the per-view scale for CTR_F16 / CTR_BF16. */
int ctr_stage_projections(void *stream, ctr_workspace *ws, const float *proj,
                          int layout, int nu, int nv, int n_views,
                          void *staged, int staged_dtype, float *view_scale);

/* ---- backprojection ----------------------------------------------------------
 * Voxel-driven backprojection over all views in ONE launch (the reference
 * issues one full-volume read-modify-write dispatch per view,
 * FBP3D.cpp:538-556).  Replaces
 *   FBP3D::project_backward_cone_fdk_from_projections   (FBP3D.cpp:524-559)
 *   FBP3D::project_backward_cone_fdk_from_sinograms     (FBP3D.cpp:561-593)
 *   FBP3D::project_backward_parallel_from_projections   (FBP3D.cpp:460-490)
 *   FBP3D::project_backward_parallel_from_sinograms     (FBP3D.cpp:492-522)
 *   FBP2D::project_backward                             (FBP2D.cpp:96-122)
 * For every voxel:
 *   vol = (accumulate ? vol : 0)
 *       + scale * sum_n view_scale[n] * w_n^2 * bilinear(staged_n, fu, fv)
 * with w = SOD / (SOD - depth) for cone geometry and 1 for parallel;
 * bilinear is zero outside the detector.  clip_negative != 0 then applies
 * max(v, 0) (FBP3D.cpp:444-458, run after every reference 3D
 * backprojection).  view_scale may be NULL for CTR_F32.  Coordinates are set
 * up in FP64 per column and view and stepped in FP32 along y; interpolation
 * and accumulation are FP32 in every format (16-bit texels widened
 * exactly): for CTR_F16 / CTR_BF16 the vertical taps of a column are v and
 * v + d of its value-and-step texel.  Row coordinates are FP32 numbers;
 * for CTR_F16 the LDS-staged kernel holds them biased by 128, rounded to
 * 2^-16 of a row (the bias gives the LDS address in one instruction).  The
 * sum over views runs in view order, so results are reproducible bit for
 * bit. */
int ctr_backproject(void *stream, ctr_workspace *ws, int geometry,
                    const void *staged, int staged_dtype,
                    const float *view_scale, int nu, int nv, int n_views,
                    double source_distance, double detector_distance,
                    double detector_width, double volume_width,
                    double rotation_count, double rotation_offset,
                    float *volume, int X, int Y, int Z, float scale,
                    int accumulate, int clip_negative);

/* ctr_backproject without a workspace, allocating nothing: no hipMalloc,
 * hipFree or allocating library call, so it can run where allocation is
 * not allowed (inside HIP graph capture, tight serving loops, harnesses that
 * forbid it).  No reference counterpart.  Same arguments and results as
 * ctr_backproject (the same kernels, bit for bit), with the options of the
 * default workspace (the environment at library load).  Its per-view
 * geometry and tile-footprint tables live in a fixed 128 MiB device array of
 * the library, one per GPU, kept until the process ends: jobs needing more
 * (144 B per view plus 16 B per view and 16 x 16 column tile; 512^3 x 720
 * views takes 12 MB, 1024^3 x 1200 views 79 MB) fail with
 * CTR_ERR_OVER_BUDGET before any launch.  Calls overlapping on one GPU share
 * that array and race: serialise them (one stream, or events). */
int ctr_backproject_noalloc(void *stream, int geometry, const void *staged,
                            int staged_dtype, const float *view_scale, int nu, int nv,
                            int n_views, double source_distance,
                            double detector_distance, double detector_width,
                            double volume_width, double rotation_count,
                            double rotation_offset, float *volume, int X, int Y, int Z,
                            float scale, int accumulate, int clip_negative);

/* One reconstruction of a packed backprojection (fields as in
 * ctr_backproject). */
typedef struct ctr_backproject_job {
    const void *staged;
    const float *view_scale;
    float *volume;
    int geometry, nu, nv, n_views, X, Y, Z;
    double source_distance, detector_distance, detector_width, volume_width;
    double rotation_count, rotation_offset;
    float scale;
    int accumulate, clip_negative;
} ctr_backproject_job;

/* Up to 16 independent backprojections (any sizes and geometries, one
 * staged dtype) in ONE persistent launch: the serving loop packs small
 * reconstructions this way so that they fill the GPU together.  Each job's
 * result equals ctr_backproject on that job alone.  No reference
 * counterpart (the reference reconstructs one volume at a time). */
int ctr_backproject_packed(void *stream, ctr_workspace *ws, int staged_dtype,
                           const ctr_backproject_job *jobs, int count);

/* ---- forward projection ------------------------------------------------------
 * Ray-driven projector, one ray per detector pixel centre, replacing
 *   FBP3D::project_forward_cone_to_projections   (FBP3D.cpp:269-299)
 *   FBP3D::project_forward_cone_to_sinogram      (FBP3D.cpp:301-330)
 *   FBP3D::project_forward_parallel_to_projections (FBP3D.cpp:212-240)
 *   FBP3D::project_forward_parallel_to_sinogram  (FBP3D.cpp:242-267)
 *   FBP2D::project_forward / project_forward_equiangular (FBP2D.cpp:54-94)
 * The ray is clipped to the volume box grown by one voxel per side; samples
 * at t_in + (k + 1/2) h, h = step * d, trilinear (zero outside), times h.
 * Output overwrites `out` in the given layout.  ws supplies the kernel
 * options (CTR_OPTION_FORWARD_*; NULL: the default workspace). */
int ctr_forward_project(void *stream, ctr_workspace *ws, int geometry, const float *volume,
                        int X, int Y, int Z, double source_distance,
                        double detector_distance, double detector_width,
                        double volume_width, double rotation_count,
                        double rotation_offset, float *out, int layout,
                        int nu, int nv, int n_views, double step);

/* ---- elementwise detector / volume passes ------------------------------------ */
/* FBP3D::apply_fdk_weights_to_projections (FBP3D.cpp:369-387), in place,
 * either layout; weight SDD / sqrt(SDD^2 + u^2 + v^2) evaluated in FP64. */
int ctr_fdk_weight(void *stream, float *proj, int layout, int nu, int nv,
                   int n_views, double source_distance,
                   double detector_distance, double detector_width);
/* FBP3D::log_normalize_projections (FBP3D.cpp:333-349), in place:
 * p = -ln(max(I, floor_value) / base). */
int ctr_log_normalize(void *stream, float *data, size_t count, float base,
                      float floor_value);
/* FBP3D::clip_negatives_of_volume (FBP3D.cpp:444-458): v = max(v, 0). */
int ctr_clip_negative(void *stream, float *data, size_t count);
/* Texture*::clear(value) as used by generate_blank_* (FBP3D.cpp:16-56). */
int ctr_fill(void *stream, float *data, size_t count, float value);

/* FBP2D::parallelize_fan_beam_sinogram (FBP2D.cpp:31-52): resamples a flat-
 * detector fan sinogram (nu_fan, N) over 2*pi*rotation_count into a parallel
 * sinogram (nu_par, N) over the same angles with detector pitch
 * parallel_width / nu_par, by bilinear interpolation in (fan angle,
 * detector) space; parallel rays outside the fan are 0. */
int ctr_rebin_fan_to_parallel(void *stream, const float *fan, int nu_fan,
                              int n_views, double source_distance,
                              double detector_distance, double detector_width,
                              double rotation_count, float *parallel,
                              int nu_par, double parallel_width);

/* ---- test-input generators -------------------------------------------------- */
/* FBP3D::render_ellipsoid (FBP3D.cpp:609-628): adds density to every voxel
 * whose centre lies inside.  ellipsoids: host array of `count` records of 10
 * doubles (cx cy cz a b c alpha beta gamma density) in world units, body-to-
 * world rotation Rz(gamma) Ry(beta) Rx(alpha).  count <= 4096. */
int ctr_render_ellipsoids(void *stream, ctr_workspace *ws,
                          const double *ellipsoids, int count, float *volume,
                          int X, int Y, int Z, double volume_width);
/* Exact line integrals of ellipsoids (records as in ctr_render_ellipsoids)
 * along the rays of ctr_forward_project: p = sum_e density_e * chord_e,
 * written to (accumulate = 0) or added to (accumulate != 0) `out` in the
 * given layout.  The chords are FP64; each 16 x 16 pixel tile of a view
 * first culls the ellipsoids whose bounding sphere no line of the tile can
 * meet (FP64, conservative).  Makes the
 * seeded input of each serving request without a voxelised phantom.  No
 * reference counterpart (the reference projects voxel volumes,
 * FBP3D.cpp:269-330). */
int ctr_project_ellipsoids(void *stream, int geometry, const double *ellipsoids, int count,
                           double source_distance, double detector_distance,
                           double detector_width, double rotation_count,
                           double rotation_offset, float *out, int layout, int nu, int nv,
                           int n_views, int accumulate);
/* Replaces every line integral p by -ln(max(K, 1) / i0) with
 * K ~ Poisson(i0 * exp(-p)) from a counter-based generator keyed by
 * (seed, element index): the noise model of the benchmark's cone config.
 * No reference counterpart. */
int ctr_poisson_log(void *stream, float *data, size_t count, double i0,
                    uint64_t seed);

/* ---- formats either side of the path -------------------------------------- */
/* Raw scanner frames -> float projections.  Replaces the per-view decode,
 * resize and R16 upload of FBP3D::read_projections /
 * read_projections_as_sinograms (FBP3D.cpp:74-176; the host reads the files).
 * raw: device buffer of n_frames frames of file_w x file_h pixels, each of
 * `channels` interleaved little-endian unsigned channels of
 * bytes_per_channel (1 or 2) bytes.  Channel 0 is converted as an R8 / R16
 * texture upload converts it (value / 255 or / 65535) and resampled to
 * target_w x target_h per axis: box (area) filter when shrinking, linear
 * between pixel centres when enlarging.  Frame f becomes view first_view + f
 * of `out`, a (n_views, target_h, target_w) projection stack or a
 * (target_h, n_views, target_w) sinogram stack per `layout`. */
int ctr_decode_frames(void *stream, const void *raw, int file_w, int file_h, int channels,
                      int bytes_per_channel, int n_frames, float *out, int layout,
                      int target_w, int target_h, int n_views, int first_view);
/* Minimum and maximum of `count` floats (NaN ignored; (0, 0) when none) into
 * the device array min_max[2].  Replaces
 * FBP3D::compute_{min,max}_value_of_volume (FBP3D.cpp:780-800). */
int ctr_min_max(void *stream, ctr_workspace *ws, const float *data, size_t count,
                float *min_max);
/* data = (data - min) / (max - min) from the device array min_max[2] (0 when
 * max == min).  Replaces FBP3D::normalize_histogram (FBP3D.cpp:802-826). */
int ctr_normalize_range(void *stream, float *data, size_t count, const float *min_max);
/* y-slices y0 .. y0 + count - 1 of a (Z, Y, X) float volume as unorm16
 * (round(clamp(v, 0, 1) * 65535)), each slice Z rows of X values: the pixels
 * FBP3D::write_volume (FBP3D.cpp:178-195) writes per slice file. */
int ctr_slices_unorm16(void *stream, const float *volume, int X, int Y, int Z, int y0,
                       int count, unsigned short *out);

#ifdef __cplusplus
}
#endif
#endif
