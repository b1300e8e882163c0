/*
 * teamtomo_amd.h -- C ABI of the MI355X (gfx950) Fourier-slice gather/scatter library.
 *
 * Every entry point takes plain device pointers, sizes and a hipStream_t
 * passed as `void *` (NULL = the null stream).  Launches are asynchronous and
 * never synchronise the stream.  Outputs are written (or accumulated, where
 * stated) into buffers the caller allocated.  The return value is TT_OK or an
 * error code; tt_last_error() returns a message for the calling thread.
 *
 * Layouts (C-contiguous, zyx):
 *   complex64 volume  (D, H, Wr)  interleaved (re, im) float pairs, DC at (D/2, H/2, 0)
 *   complex64 slices  (n, Hi, Wri) interleaved, DC at (Hi/2, 0)
 *   rotations         (n, 3, 3) float32, already cast, near-zero clamped and
 *                     flipped to zyx by the caller (rows/cols ordered z, y, x)
 *   rotation grads    (n, 3, 3) float64, accumulated (+=)
 *
 * Reference interfaces replaced (paths relative to
 * packages/primitives/ of teamtomo/teamtomo):
 *   torch-fourier-slice/src/torch_fourier_slice/slice_extraction/_extract_central_slices_rfft_3d.py:84
 *   torch-fourier-slice/src/torch_fourier_slice/slice_insertion/_insert_central_slices_rfft_3d.py:9,289
 *   torch-image-interpolation/src/torch_image_interpolation/image_interpolation_3d.py:11,104
 *   torch-fourier-slice/src/torch_fourier_slice/slice_extraction/_extract_central_slices_rfft_2d.py:9
 *   torch-image-interpolation/src/torch_image_interpolation/image_interpolation_2d.py:11,104
 *   torch-transform-image/src/torch_transform_image/transforms_3d.py:10, transforms_2d.py:11
 *   torch-subpixel-crop/src/torch_subpixel_crop/subpixel_crop_3d.py:12, subpixel_crop_2d.py:12
 */
#ifndef TEAMTOMO_AMD_H
#define TEAMTOMO_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TT_OK 0
#define TT_ERR_INVALID 1 /* bad argument (shape, size, null pointer)        */
#define TT_ERR_HIP 2     /* a HIP runtime call or kernel launch failed       */

#define TT_INTERP_NEAREST 0
#define TT_INTERP_TRILINEAR 1
#define TT_INTERP_BILINEAR 1 /* the 2-D name of the linear mode  */
#define TT_INTERP_BICUBIC 2  /* 2-D samplers only                */

/* `flags` bits of the scatter entry points */
#define TT_FLAG_DETERMINISTIC 1 /* fixed-point accumulation: bitwise reproducible      */
#define TT_FLAG_OVERWRITE 2     /* write the outputs instead of adding to them          */
#define TT_FLAG_PIXEL_ATOMIC 4  /* force the pixel-driven global-atomic scatter         */

/* Extraction gather kernels (process-wide choice, see tt_set_gather_path) */
#define TT_GATHER_AUTO 0  /* brick gather, pixel gather beyond the brick limits */
#define TT_GATHER_PIXEL 1 /* one lane per pixel, corners read through L2/L1      */
#define TT_GATHER_BRICK 2 /* volume bricks staged in LDS, serving all slices     */

/* Scatter kernels of the adjoint and the insertion (see tt_set_scatter_path) */
#define TT_SCATTER_AUTO 0   /* binned for few pixels per brick, else the brick engine */
#define TT_SCATTER_BRICK 1  /* brick engine: bricks find and walk their slice rows    */
#define TT_SCATTER_BINNED 2 /* pixels sorted into per-brick bins, one pass per brick  */

/* Message for the last error on the calling thread ("" if none). */
const char *tt_last_error(void);

/* ABI version; bumped whenever a signature below changes. */
int tt_abi_version(void);

/*
 * Selects the kernel tt_extract_c64 uses (TT_GATHER_AUTO by default).  The
 * two agree to within ~1e-6 of max|V| (FP32 brick frames against FP64 pixel
 * coordinates), not bit for bit; the choice moves the cost between L2 traffic
 * (pixel-major) and LDS staging (brick).  Has no counterpart in the
 * reference: it is a tuning and testing switch of this library.
 */
int tt_set_gather_path(int path);

/*
 * Selects the scatter the complex64 adjoint and insertion entry points use
 * (TT_SCATTER_AUTO by default: the binned scatter when the call has at most
 * 800 slice pixels per 4x8x64 brick of the volume, such as a tilt series
 * into a large volume, and its record buffer, 8 x 24 bytes per slice pixel,
 * stays within 8 GiB and can be allocated; the brick engine otherwise).  TT_SCATTER_BINNED uses the binned scatter wherever it fits.
 * Both accumulate in fixed point and are bitwise reproducible; they agree to
 * ~1e-6 relative (different FP32 roundings of the coordinates and scales).
 * Process-wide; a tuning and testing switch with no counterpart in the
 * reference.
 */
int tt_set_scatter_path(int path);

/*
 * Host-side check of the brick engine's work plan for one call (kind 0: the
 * gather of tt_extract_c64, 1: the adjoint of tt_extract_bwd_c64, 2: the
 * insertion of tt_insert_c64, 3 and 4: the 8x8x8-cell gather and insertion
 * tiles a batch may choose; rotations are not read, only their number):
 * every unit's brick, rotation segment and partial slot, every split
 * record's slots, the extents of the segment and halo buffers the kernels
 * index, that each brick's units cover its rotation segments exactly once,
 * and (kinds 0, 2, 3, 4) that the batch builder's buffers for such an item
 * are as large.  Needs no GPU.  stats[8] = bricks, units, split bricks, partial
 * slots, segment-buffer bytes, halo-buffer bytes, empty units, largest
 * segment count.  TT_ERR_INVALID with the failed condition in tt_last_error.
 * No counterpart in the reference: a diagnostics entry of this library.
 */
int tt_plan_check(int kind, int vol_d, int vol_h, int vol_wr, int img_d, int img_h, int img_w, int64_t n_rot,
                  double fftfreq_max, int64_t *stats);

/*
 * Rotation preparation: out (n, 3, 3) float32 = the (n, 3, 3) float32 or
 * float64 matrices cast to float32, entries with |R| < rot_tolerance set to
 * zero (rot_tolerance < 0: no clamp; the comparison in float32), and both
 * matrix axes reversed when flip != 0 (xyz -> zyx).  This is the form every
 * rot_zyx argument below takes.  Replaces the preparation steps of
 * torch-fourier-slice/.../slice_extraction/_extract_central_slices_rfft_3d.py:147-154
 * and .../slice_insertion/_insert_central_slices_rfft_3d.py:16,66-67, without
 * the reference's in-place clamp of the caller's tensor.
 */
int tt_prepare_rotations_f32(const float *rot, int64_t n, int flip, double rot_tolerance, float *out,
                             void *stream);
int tt_prepare_rotations_f64(const double *rot, int64_t n, int flip, double rot_tolerance, float *out,
                             void *stream);

/*
 * Central-slice extraction (the gather).
 * Replaces extract_central_slices_rfft_3d
 * (torch-fourier-slice/.../slice_extraction/_extract_central_slices_rfft_3d.py:84-222)
 * and its sample_image_3d / grid_sample call (.../image_interpolation_3d.py:11-101).
 *   volume      (vol_d, vol_h, vol_wr) complex64
 *   img_*       image_shape (D, H, W) that drives the frequency grid
 *   slices      (n_rot, img_h, img_w/2+1) complex64, written (every pixel)
 *   fftfreq_max < 0 means None; x_tolerance as in the reference.
 */
int tt_extract_c64(const void *volume, int vol_d, int vol_h, int vol_wr,
                   int img_d, int img_h, int img_w,
                   const float *rot_zyx, int64_t n_rot,
                   double fftfreq_max, double x_tolerance,
                   void *slices, void *stream);

/*
 * tt_extract_c64 without allocation: the same brick gather (the same slices,
 * bit for bit) for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops).  It has no counterpart in the
 * reference.  Its plane records and unit table live in a fixed file-scope
 * device array per GPU (about 3.0 MB, held until the process ends) instead of
 * the allocations tt_extract_c64 makes; the first call for a geometry plans
 * its units on the host (host memory only) and copies them in, stream-ordered.
 *   n_rot <= 16384, and the volume's brick units <= 65536 (256^3 needs
 *   about 9000): larger calls are refused with TT_ERR_INVALID.
 *   Every call on one GPU shares the array: calls that overlap on one GPU
 *   (different streams, threads) race; calls on one stream are ordered.
 * Geometries beyond the brick engine's limits are refused (no pixel path).
 */
int tt_extract_c64_noalloc(const void *volume, int vol_d, int vol_h, int vol_wr,
                           int img_d, int img_h, int img_w,
                           const float *rot_zyx, int64_t n_rot,
                           double fftfreq_max, double x_tolerance,
                           void *slices, void *stream);

/*
 * Multichannel extraction: replaces extract_central_slices_rfft_3d_multichannel
 * (torch-fourier-slice/.../slice_extraction/_extract_central_slices_rfft_3d.py:225-367).
 *   volume  (channels, vol_d, vol_h, vol_wr) complex64
 *   slices  (n_rot, channels, img_h, img_w/2+1) complex64, written
 * The brick gather serves two channels per launch (one enumeration of the
 * slice pixels for both); otherwise as tt_extract_c64.
 */
int tt_extract_mc_c64(const void *volume, int channels, int vol_d, int vol_h, int vol_wr,
                      int img_d, int img_h, int img_w, const float *rot_zyx, int64_t n_rot,
                      double fftfreq_max, double x_tolerance, void *slices, void *stream);

/*
 * Backward of the extraction: the adjoint scatter into grad_volume
 * (complex64 (vol_d, vol_h, vol_wr); accumulated, or written with
 * TT_FLAG_OVERWRITE) and/or the analytic rotation gradient (accumulated into
 * grad_rot, float64 (n_rot, 3, 3) zyx).  Either output may be NULL.
 * `volume` is read only when grad_rot != NULL.
 * Replaces torch autograd through grid_sample (the grid_sampler_3d_backward
 * reached from .../image_interpolation_3d.py:69-75).
 * flags: TT_FLAG_* bits.
 */
int tt_extract_bwd_c64(const void *grad_slices, const void *volume,
                       int vol_d, int vol_h, int vol_wr,
                       int img_d, int img_h, int img_w,
                       const float *rot_zyx, int64_t n_rot,
                       double fftfreq_max, double x_tolerance,
                       void *grad_volume, double *grad_rot,
                       int flags, void *stream);

/*
 * Central-slice insertion (the scatter) into a volume and a weights volume
 * of shape (vol_d, vol_h, vol_w/2+1), accumulated (or written with
 * TT_FLAG_OVERWRITE).
 * Replaces insert_central_slices_rfft_3d and _insert_into_3d_dft_linear
 * (torch-fourier-slice/.../slice_insertion/_insert_central_slices_rfft_3d.py:9-106,289-406),
 * including the x=0-plane Friedel mirror.
 *   slices      (n_rot, vol_h, vol_w/2+1)
 *   _c64 : slices/volume complex64, weights float32
 *   _c128: slices/volume complex128, weights float64
 */
int tt_insert_c64(const void *slices, int vol_d, int vol_h, int vol_w,
                  const float *rot_zyx, int64_t n_rot, double fftfreq_max,
                  void *volume, float *weights, int flags, void *stream);
int tt_insert_c128(const void *slices, int vol_d, int vol_h, int vol_w,
                   const float *rot_zyx, int64_t n_rot, double fftfreq_max,
                   void *volume, double *weights, int flags, void *stream);

/*
 * Multichannel insertion: replaces insert_central_slices_rfft_3d_multichannel
 * (torch-fourier-slice/.../slice_insertion/_insert_central_slices_rfft_3d.py:109-210).
 *   slices   (n_rot, channels, vol_h, vol_w/2+1) complex64, read in place
 *   volume   (channels, vol_d, vol_h, vol_w/2+1) complex64
 *   weights  (vol_d, vol_h, vol_w/2+1) float32, shared by the channels (as in
 *            the reference, one weight per inserted point, not per channel)
 * Accumulates, or writes with TT_FLAG_OVERWRITE.  Brick engine only
 * (TT_FLAG_PIXEL_ATOMIC and geometries beyond its limits are refused).
 */
int tt_insert_mc_c64(const void *slices, int channels, int vol_d, int vol_h, int vol_w,
                     const float *rot_zyx, int64_t n_rot, double fftfreq_max,
                     void *volume, float *weights, int flags, void *stream);

/*
 * Backward of the insertion: the gradient w.r.t. the slices (the adjoint
 * gather, written into grad_slices) and/or w.r.t. the rotations (accumulated
 * into grad_rot).  grad_weights may be NULL (treated as 0); `slices` is read
 * only when grad_rot != NULL.
 */
int tt_insert_bwd_c64(const void *grad_volume, const float *grad_weights,
                      const void *slices, int vol_d, int vol_h, int vol_w,
                      const float *rot_zyx, int64_t n_rot, double fftfreq_max,
                      void *grad_slices, double *grad_rot, void *stream);
int tt_insert_bwd_c128(const void *grad_volume, const double *grad_weights,
                       const void *slices, int vol_d, int vol_h, int vol_w,
                       const float *rot_zyx, int64_t n_rot, double fftfreq_max,
                       void *grad_slices, double *grad_rot, void *stream);

/*
 * Generic sampler: replaces sample_image_3d
 * (torch-image-interpolation/.../image_interpolation_3d.py:11-101).
 *   image   (channels, d, h, w) real, or complex interleaved if is_complex
 *   coords  (n, 3) zyx array coordinates
 *   out     (n, channels) values (complex interleaved if is_complex)
 *   mode    TT_INTERP_TRILINEAR or TT_INTERP_NEAREST
 * _f32: float32 image/coords/out; _f64: float64.
 */
int tt_sample3d_f32(const float *image, int channels, int d, int h, int w, int is_complex,
                    const float *coords, int64_t n, int mode, float *out, void *stream);
int tt_sample3d_f64(const double *image, int channels, int d, int h, int w, int is_complex,
                    const double *coords, int64_t n, int mode, double *out, void *stream);

/*
 * Backward of the generic sampler: replaces the autograd of sample_image_3d,
 * i.e. grid_sample's backward with border padding and align_corners=True and
 * the zeroing of outside samples (.../image_interpolation_3d.py:64-89).
 *   grad_out    (n, channels) gradient of the samples (complex interleaved)
 *   grad_image  (channels, d, h, w), ACCUMULATED (the caller zeroes it), or NULL
 *   grad_coords (n, 3) written, or NULL; Re(conj(g) ds/dc) summed over the
 *               channels, 0 along an axis where c <= 0 or c >= N-1, 0 for
 *               nearest sampling and for samples outside the image
 */
int tt_sample3d_bwd_f32(const float *image, int channels, int d, int h, int w, int is_complex,
                        const float *coords, int64_t n, int mode, const float *grad_out, float *grad_image,
                        float *grad_coords, void *stream);
int tt_sample3d_bwd_f64(const double *image, int channels, int d, int h, int w, int is_complex,
                        const double *coords, int64_t n, int mode, const double *grad_out, double *grad_image,
                        double *grad_coords, void *stream);

/*
 * Generic splat, accumulating in place: replaces insert_into_image_3d
 * (torch-image-interpolation/.../image_interpolation_3d.py:104-275).
 *   values  (n, channels) (complex interleaved if is_complex)
 *   coords  (n, 3) float32 zyx
 *   image   (channels, d, h, w) accumulated
 *   weights (d, h, w) accumulated; float32 if weights_f64 == 0 else float64
 * _f32: float32 values/image; _f64: float64 values/image.
 */
int tt_splat3d_f32(const float *values, const float *coords, int64_t n, int channels,
                   int is_complex, int d, int h, int w, int mode,
                   float *image, void *weights, int weights_f64, void *stream);
int tt_splat3d_f64(const double *values, const float *coords, int64_t n, int channels,
                   int is_complex, int d, int h, int w, int mode,
                   double *image, void *weights, int weights_f64, void *stream);

/*
 * Central lines of a 2-D rfft (the 2-D analogue of tt_extract_c64).
 * Replaces extract_central_slices_rfft_2d
 * (torch-fourier-slice/.../slice_extraction/_extract_central_slices_rfft_2d.py:9-92)
 * and its sample_image_2d / grid_sample call (.../image_interpolation_2d.py:11-101).
 *   image       (h, wr) complex64 fftshifted rfft, DC at (h/2, 0)
 *   img_h/img_w image_shape (h, w) that drives the frequency line
 *   rot_yx      (n_rot, 2, 2) float32, already cast and flipped to yx
 *   lines       (n_rot, img_w/2+1) complex64, written (every pixel)
 *   fftfreq_max < 0 means None; otherwise pixels with f_x > fftfreq_max are 0
 *               (the reference raises for every fftfreq_max: its elementwise
 *               (w, 2) mask flattens the grid, :34-36).
 */
int tt_extract2d_c64(const void *image, int h, int wr, int img_h, int img_w,
                     const float *rot_yx, int64_t n_rot, double fftfreq_max,
                     void *lines, void *stream);

/*
 * Backward of tt_extract2d_c64: the image gradient (adjoint bilinear splat,
 * accumulated into grad_image, complex64 (h, wr)) and/or the rotation
 * gradient (accumulated into grad_rot, float64 (n_rot, 2, 2) yx).  Either
 * output may be NULL; `image` is read only when grad_rot != NULL.
 * Replaces torch autograd through grid_sample (grid_sampler_2d_backward
 * reached from .../image_interpolation_2d.py:69-75).
 */
int tt_extract2d_bwd_c64(const void *grad_lines, const void *image, int h, int wr,
                         int img_h, int img_w, const float *rot_yx, int64_t n_rot,
                         double fftfreq_max, void *grad_image, double *grad_rot,
                         void *stream);

/*
 * Generic 2-D sampler: replaces sample_image_2d
 * (torch-image-interpolation/.../image_interpolation_2d.py:11-101).
 *   image   (channels, h, w) real, or complex interleaved if is_complex
 *   coords  (n, 2) yx array coordinates
 *   out     (n, channels) values (complex interleaved if is_complex)
 *   mode    TT_INTERP_BILINEAR, TT_INTERP_NEAREST or TT_INTERP_BICUBIC
 * _f32: float32 image/coords/out; _f64: float64.
 */
int tt_sample2d_f32(const float *image, int channels, int h, int w, int is_complex,
                    const float *coords, int64_t n, int mode, float *out, void *stream);
int tt_sample2d_f64(const double *image, int channels, int h, int w, int is_complex,
                    const double *coords, int64_t n, int mode, double *out, void *stream);

/*
 * Generic 2-D splat, accumulating in place: replaces insert_into_image_2d
 * (torch-image-interpolation/.../image_interpolation_2d.py:104-275).
 *   values  (n, channels) (complex interleaved if is_complex)
 *   coords  (n, 2) float32 yx
 *   image   (channels, h, w) accumulated
 *   weights (h, w) accumulated; float32 if weights_f64 == 0 else float64
 *   mode    TT_INTERP_BILINEAR or TT_INTERP_NEAREST
 */
int tt_splat2d_f32(const float *values, const float *coords, int64_t n, int channels,
                   int is_complex, int h, int w, int mode,
                   float *image, void *weights, int weights_f64, void *stream);
int tt_splat2d_f64(const double *values, const float *coords, int64_t n, int channels,
                   int is_complex, int h, int w, int mode,
                   double *image, void *weights, int weights_f64, void *stream);

/*
 * Affine warps: replace affine_transform_image_3d / _2d
 * (torch-transform-image/src/torch_transform_image/transforms_3d.py:10-40,
 * transforms_2d.py:11-41) with the coordinate grid computed in the kernel.
 * Output voxel (z, y, x) of leading batch index t samples the image at
 * M (z, y, x, 1)^T, M the block of matrix ((t md + z') mh + y') mw + x'
 * (z' = z if md > 1 else 0, ...): the matrices broadcast against the output
 * grid as the reference's matmul does (md, mh, mw each 1 or d, h, w).
 *   image   (channels, D, H, W) / (channels, H, W) float32, or complex64
 *           interleaved if is_complex
 *   mats    (n_lead * md * mh * mw, 3, 4) / (n_lead * mh * mw, 2, 3) float32
 *           zyx / yx blocks (row-major)
 *   out     (n_lead, d, h, w, channels) / (n_lead, h, w, channels)
 *   mode    TT_INTERP_TRILINEAR (= BILINEAR), TT_INTERP_NEAREST, or (2-D)
 *           TT_INTERP_BICUBIC
 */
int tt_affine3d_f32(const float *image, int channels, int D, int H, int W, int is_complex,
                    const float *mats, int64_t n_lead, int md, int mh, int mw,
                    int d, int h, int w, int mode, float *out, void *stream);
int tt_affine2d_f32(const float *image, int channels, int H, int W, int is_complex,
                    const float *mats, int64_t n_lead, int mh, int mw, int h, int w, int mode,
                    float *out, void *stream);

/*
 * Subpixel crops: replace subpixel_crop_3d / _2d
 * (torch-subpixel-crop/src/torch_subpixel_crop/subpixel_crop_3d.py:12-108,
 * subpixel_crop_2d.py:12-163) together with torch-fourier-shift's phase ramp
 * (torch-fourier-shift/src/torch_fourier_shift/phase_shift_grids.py:60-127);
 * the FFTs in between stay with the caller.
 * tt_crop3d_f32: patches (n, p, p, p) = image (D, H, W) around the int32
 *   positions (n, 3) zyx (patch centre at index p/2, zeros outside), times
 *   mask (p, p, p) unless NULL.
 * tt_crop2d_f32: patches (n, B, p, p) = images (B, H, W), patch (m, b) of image
 *   b around positions[m, b] (n, B, 2) yx, times mask (p, p) unless NULL.
 * The phase ramp is tt_fourier_shift_c64 (in place, rfft, not fftshifted).
 */
int tt_crop3d_f32(const float *image, int D, int H, int W, const int *positions, int64_t n,
                  int sidelength, const float *mask, float *patches, void *stream);
int tt_crop2d_f32(const float *images, int B, int H, int W, const int *positions, int64_t n,
                  int sidelength, const float *mask, float *patches, void *stream);

/*
 * Fourier shifts: replace fourier_shift_dft_1d/_2d/_3d
 * (torch-fourier-shift/src/torch_fourier_shift/fourier_shift_dft.py:10-115,
 * with the phase grids of phase_shift_grids.py:9-127) and are the middle step
 * of fourier_shift_image_1d/_2d/_3d (fourier_shift_image.py:10-91).
 * tt_fourier_shift_c64 / _c128: out[t] = dft[t] * exp(-2 pi i f . s_t) for
 *   t < n, s_t = shifts[t] (n, ndim) in (z,) (y,) x order.  One DFT holds
 *   ([d,] [h,] ws) elements, ws = w/2 + 1 if rfft else w, of an image of shape
 *   ([d,] [h,] w) (d, h ignored below ndim 3, 2).  f is fftfreq on every axis
 *   (rfftfreq on the last if rfft); fftshifted != 0: the DFTs are fftshifted
 *   over every axis but an rfft's last.  dft_stride is 0 (one DFT for every
 *   t) or the element count of one DFT.  out may equal dft.
 * tt_fourier_shift_bwd_c64 / _c128: the backward of that product for the
 *   incoming gradient grad_out (n DFTs): grad_dft[t] = grad_out[t] *
 *   conj(ramp) (n DFTs, skipped when NULL) and grad_shifts[t, a] =
 *   2 pi sum_v f_a(v) Im(conj(grad_out) out) (n, ndim), skipped when NULL,
 *   which reads dft.  Stream-ordered scratch (hipMallocAsync) holds the
 *   per-block partial sums.
 */
int tt_fourier_shift_c64(const void *dft, int64_t dft_stride, const float *shifts, int64_t n,
                         int ndim, int d, int h, int w, int rfft, int fftshifted, void *out,
                         void *stream);
int tt_fourier_shift_c128(const void *dft, int64_t dft_stride, const double *shifts, int64_t n,
                          int ndim, int d, int h, int w, int rfft, int fftshifted, void *out,
                          void *stream);
int tt_fourier_shift_bwd_c64(const void *grad_out, const void *dft, int64_t dft_stride,
                             const float *shifts, int64_t n, int ndim, int d, int h, int w,
                             int rfft, int fftshifted, void *grad_dft, float *grad_shifts,
                             void *stream);
int tt_fourier_shift_bwd_c128(const void *grad_out, const void *dft, int64_t dft_stride,
                              const double *shifts, int64_t n, int ndim, int d, int h, int w,
                              int rfft, int fftshifted, void *grad_dft, double *grad_shifts,
                              void *stream);

/*
 * The elementwise steps around the FFTs of project_3d_to_2d and
 * backproject_2d_to_3d (torch-fourier-slice/.../project.py:44-70,
 * backproject.py:75-91), one pass each.  sinc^2 is sinc(|f|)^2 on the centred
 * grid f_i = (i - N/2)/N of the padded N^3 volume.
 * tt_pad_sinc2_fftshift_*: out (N^3, N = d + 2 pad) = fftshift(pad(vol) * sinc^2).
 * tt_ifftshift_sinc2_crop_*: out (d^3) = crop(ifftshift(x) / sinc^2) (divide
 *   != 0: the backprojection's gridding correction) or crop(ifftshift(x) *
 *   sinc^2) (divide == 0: the adjoint of tt_pad_sinc2_fftshift_*).
 * tt_wdiv_ifftshift_*: out = ifftshift over the two full axes of
 *   vol / clamp(w, 1), (D, H, Wr) rfft volume and weights.
 */
int tt_pad_sinc2_fftshift_f32(const float *vol, int d, int pad, float *out, void *stream);
int tt_pad_sinc2_fftshift_f64(const double *vol, int d, int pad, double *out, void *stream);
int tt_ifftshift_sinc2_crop_f32(const float *x, int N, int pad, int divide, float *out, void *stream);
int tt_ifftshift_sinc2_crop_f64(const double *x, int N, int pad, int divide, double *out, void *stream);
int tt_wdiv_ifftshift_c64(const void *vol, const float *w, int D, int H, int Wr, void *out, void *stream);
int tt_wdiv_ifftshift_c128(const void *vol, const double *w, int D, int H, int Wr, void *out, void *stream);

/*
 * Batched extraction or insertion over many volumes of their own box size and
 * rotation count.  No single reference function corresponds: it runs the
 * work of a loop of extract_central_slices_rfft_3d
 * (.../slice_extraction/_extract_central_slices_rfft_3d.py:84-222) or
 * insert_central_slices_rfft_3d (.../slice_insertion/_insert_central_slices_rfft_3d.py:9-106)
 * calls, one per volume, as one persistent work-queue launch of the brick
 * kernel per tile configuration (plus batched helper kernels).
 *
 * tt_batch_create plans the batch and allocates everything a run needs
 * (hipMalloc), optionally timing the candidate tile configurations per box
 * size (autotune != 0: runs the batch, synchronises the stream).
 * tt_batch_run only launches kernels on `stream`: no allocation, no host
 * planning and no memset, so it can be captured in a graph and replayed.  The
 * item buffers are bound at creation; a run reads the current contents.
 * Outputs are written, not accumulated.  Insertion: deterministic fixed-point
 * sums, as the one-volume entry point.
 */
#define TT_BATCH_EXTRACT 0
#define TT_BATCH_INSERT 1

typedef struct tt_batch_item {
  const void *input;          /* extract: volume (vol_d, vol_h, vol_wr) c64;
                                 insert: slices (n_rot, img_h, img_w/2+1) c64   */
  int vol_d, vol_h, vol_wr;   /* rfft volume shape                              */
  int img_d, img_h, img_w;    /* image shape (insertion: the full volume shape) */
  const float *rot_zyx;       /* (n_rot, 3, 3) prepared as for the one-call API */
  int64_t n_rot;
  double fftfreq_max;         /* < 0 means None                                 */
  double x_tolerance;         /* extraction only                                */
  void *output;               /* extract: slices (n_rot, img_h, img_w/2+1) c64;
                                 insert: volume (vol_d, vol_h, vol_wr) c64       */
  float *weights;             /* insert: (vol_d, vol_h, vol_wr) float32          */
} tt_batch_item;

int tt_batch_create(int kind, const tt_batch_item *items, int n_items, int autotune, void *stream,
                    void **handle);
int tt_batch_run(void *handle, void *stream);
/* Tile (brick of cells) the batch uses for one item, as picked at creation. */
int tt_batch_tile(void *handle, int item, int *bz, int *by, int *bx);
int tt_batch_destroy(void *handle);

#ifdef __cplusplus
}
#endif

#endif /* TEAMTOMO_AMD_H */
