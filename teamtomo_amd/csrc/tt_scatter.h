// Scatter drivers shared by the extraction adjoint and the insertion.
#pragma once

#include "tt_common.h"

namespace tt {

// gV += A^T gS, where A is the extraction gather of geometry g.
int scatter_extract_adjoint(const SliceGeom &g, const f2v *gS, const float *rot, int64_t n_rot,
                            f2v *gV, int flags, hipStream_t s);

// (vol, w) += insertion of `slices` (with the x=0 Friedel mirror).
template <typename T>
int scatter_insert(const SliceGeom &g, const typename Cx<T>::type *slices, const float *rot,
                   int64_t n_rot, typename Cx<T>::type *vol, T *w, int flags, hipStream_t s);

// Fill n elements of type T with zeros (a kernel, so it can be graph-captured).
template <typename T> int zero_fill(T *p, int64_t n, hipStream_t s);

}  // namespace tt
