// Drop-in for split/include/split/device/morph/erode.cuh:9-22.
#if !defined(SPLIT_DEVICE_INCLUDED_MORPH_ERODE)
#define SPLIT_DEVICE_INCLUDED_MORPH_ERODE

#include "split/detail/internal.h"
#include "split/device/views.h"

SPLIT_DEVICE_NAMESPACE_BEGIN

namespace morph
{
/// In each row of the labels, pixels closer than niterations to either end
/// of their run of equal labels become -1 (in place).  Throws
/// split::device_error on failure; returns with the work finished.
SPLIT_API void erode_horizontal_sweep(view2d<int> dio_labels,
                                      const int niterations = 1);

/// The sweep over the columns, then over the rows: a pixel keeps its label
/// iff the (2 niterations + 1)^2 square around it lies in the image and
/// carries one label; others become -1 (in place).
SPLIT_API void erode(view2d<int> dio_labels, const int niterations = 1);

}  // namespace morph

SPLIT_DEVICE_NAMESPACE_END

#endif  // SPLIT_DEVICE_INCLUDED_MORPH_ERODE
