// Rotation preparation in one pass: float32 cast, the near-zero clamp and the
// xyz -> zyx flip that every slice entry point expects of its rotations.
//
// Reference: packages/primitives/torch-fourier-slice/src/torch_fourier_slice/
//   slice_extraction/_extract_central_slices_rfft_3d.py:147-154 (cast, zero
//   |R| < rot_tolerance, flip of both matrix axes unless zyx_matrices);
//   slice_insertion/_insert_central_slices_rfft_3d.py:16,66-67 (cast, flip).
// The reference clamps the caller's float32 tensor in place; here the result
// is written to a separate buffer.  The comparison is made in float32, as
// torch compares a float32 tensor with a Python float.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "../../include/teamtomo_amd.h"
#include "tt_common.h"

namespace tt {

namespace {

template <typename T>
__global__ __launch_bounds__(256) void prepare_rot_kernel(const T *__restrict__ in, int64_t n, int flip, int clamp,
                                                          float tol, float *__restrict__ out) {
  const int64_t o = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (o >= n * 9) return;
  const int64_t m = o / 9;
  const int e = (int)(o - m * 9);
  const int src = flip ? 8 - e : e;  // (2 - i) * 3 + (2 - j) = 8 - (3 i + j)
  float v = (float)in[m * 9 + src];
  if (clamp && fabsf(v) < tol) v = 0.0f;
  out[o] = v;
}

template <typename T>
int prepare_impl(const T *in, int64_t n, int flip, double tol, float *out, void *stream) {
  if (n < 0) return fail_invalid("n must be >= 0");
  if (n == 0) return TT_OK;
  if (!in || !out) return fail_invalid("NULL pointer");
  const int64_t total = n * 9;
  if ((total + 255) / 256 >= ((int64_t)1 << 31)) return fail_invalid("too many rotations for one call");
  hipLaunchKernelGGL(prepare_rot_kernel<T>, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, in, n, flip ? 1 : 0, tol >= 0.0 ? 1 : 0, (float)tol, out);
  return check_launch("prepare_rot_kernel");
}

}  // namespace

}  // namespace tt

using namespace tt;

extern "C" {

int tt_prepare_rotations_f32(const float *rot, int64_t n, int flip, double rot_tolerance, float *out,
                             void *stream) {
  return prepare_impl<float>(rot, n, flip, rot_tolerance, out, stream);
}

int tt_prepare_rotations_f64(const double *rot, int64_t n, int flip, double rot_tolerance, float *out,
                             void *stream) {
  return prepare_impl<double>(rot, n, flip, rot_tolerance, out, stream);
}

}  // extern "C"
