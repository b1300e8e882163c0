// Brick-stationary central-slice gather (extraction forward).
#pragma once

#include "tt_common.h"

namespace tt {

// True when the brick gather is expected to beat the pixel-major gather:
// at least two samples per volume voxel, and a volume of at least 2 voxels
// per axis.
bool brick_gather_eligible(const SliceGeom &g, int64_t n_rot);

// slices = extraction of `vol` (complex64) at the rotations; writes every
// pixel of the n_rot slices.
int brick_gather(const SliceGeom &g, const f2v *vol, const float *rot, int64_t n_rot, f2v *out,
                 hipStream_t s);

}  // namespace tt
