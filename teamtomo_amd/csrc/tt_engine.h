// Brick engine: brick-stationary central-slice gather and scatter (complex64).
#pragma once

#include "tt_common.h"

namespace tt {

// Geometries the engine handles: fewer than 2^30 rotations, slice height
// below 32768 and rfft width below 65536 (the item and row records hold the
// row and column in 16 bits).
bool engine_supports(const SliceGeom &g, int64_t n_rot);

// slices = extraction of `vol` at the rotations; writes every pixel of the
// n_rot slices in one launch (the brick gather, and zero-row blocks for the
// pixels outside the volume or the disc).
int engine_extract(const SliceGeom &g, const f2v *vol, const float *rot, int64_t n_rot, f2v *out,
                   hipStream_t s);

// As engine_extract, with no allocation: the plane records and unit table
// live in a fixed file-scope device array (at most 16384 rotations and 65536
// brick units per call; calls that overlap on one GPU race).
int engine_extract_noalloc(const SliceGeom &g, const f2v *vol, const float *rot, int64_t n_rot, f2v *out,
                           hipStream_t s);

// Two channels of a multichannel volume (channel stride g.vstride) into a
// multichannel slice stack (rotation stride g.ostride, channel stride g.P),
// sharing one plan, plane-record pass and brick launch.
int engine_extract2(const SliceGeom &g, const f2v *vol, const float *rot, int64_t n_rot, f2v *out,
                    hipStream_t s);

// gV (+)= A^T gS, A the extraction of geometry g.  64-bit fixed-point LDS
// accumulation per brick: bitwise reproducible.
int engine_adjoint(const SliceGeom &g, const f2v *gS, const float *rot, int64_t n_rot, f2v *gV,
                   int accumulate, hipStream_t s);

// (vol, w) (+)= insertion of `slices`, including the x=0-plane Friedel mirror.
int engine_insert(const SliceGeom &g, const f2v *slices, const float *rot, int64_t n_rot, f2v *vol,
                  float *w, int accumulate, hipStream_t s);

// Host-only check of the brick engine's plan for a call (kind 0 gather, 1
// adjoint, 2 insertion): see tt_plan_check.
int engine_plan_check(const SliceGeom &g, int kind, int64_t n_rot, int64_t stats[8]);

// The insertion's x=0-plane Friedel mirror pass: out[z, y, 0] (+)= plane[z, y]
// + conj(plane[-z, -y]) and likewise for the weights (DC not mirrored).
int launch_mirror(const f2v *plane, const float *wplane, const SliceGeom &g, f2v *out, float *wout, int accumulate,
                  hipStream_t s);

// Binned scatter (tt_binned.hip) for calls with few slice pixels per volume
// brick: pixels sorted into per-brick bins, one LDS accumulation per brick,
// 64-bit fixed point (bitwise reproducible).  set_scatter_path takes a
// TT_SCATTER_* value; binned_selected says whether engine_adjoint and
// engine_insert take the binned path for this call.
int set_scatter_path(int path);
bool binned_selected(const SliceGeom &g, int64_t n_rot);
// binned_scatter's result when its record buffer cannot be allocated: nothing
// was launched, and the caller takes the brick engine
constexpr int kBinnedNoMemory = -1000;
int binned_scatter(bool insert, const SliceGeom &g, const f2v *src, const float *rot, int64_t n_rot, f2v *out,
                   float *wout, int accumulate, hipStream_t s);

}  // namespace tt
