// Drop-in for split/include/split/device/color/beta_feature.cuh:9-18.
#if !defined(SPLIT_DEVICE_INCLUDED_COLOR_BETA_FEATURE)
#define SPLIT_DEVICE_INCLUDED_COLOR_BETA_FEATURE

#include "split/detail/internal.h"
#include "split/device/views.h"

SPLIT_DEVICE_NAMESPACE_BEGIN

namespace color
{
/// beta = -mu where |a|, |b| < 0.2 max|a| and L < 0.2 max L, +mu where
/// |a|, |b| < 0.2 max|a| and L > 0.95 max L, 0 elsewhere (3 x N Lab points
/// in; do_beta may be row 0 of them).  Throws split::device_error on
/// failure; returns with the work finished.
SPLIT_API void beta_feature(const_view2d<real> di_lab_points,
                            view1d<real> do_beta,
                            const real mu = 1e+5);

}  // namespace color

SPLIT_DEVICE_NAMESPACE_END

#endif  // SPLIT_DEVICE_INCLUDED_COLOR_BETA_FEATURE
