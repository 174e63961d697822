// Drop-in for split/include/split/device/kmeans/propagation.cuh:24-27.
#if !defined(SPLIT_DEVICE_INCLUDED_KMEANS_PROPAGATION)
#define SPLIT_DEVICE_INCLUDED_KMEANS_PROPAGATION

#include "split/detail/internal.h"
#include "split/device/views.h"

SPLIT_DEVICE_NAMESPACE_BEGIN

namespace kmeans
{
/// Writes each point's centroid (column labels[i] of di_centroids) into
/// do_points (3 x N).  Points whose label is outside [0, K) are left as
/// they are.  Throws split::device_error on failure.
SPLIT_API void propagate_centroids(const_view1d<int> di_cluster_labels,
                                   const_view2d<real> di_centroids,
                                   view2d<real> do_points);
}  // namespace kmeans

SPLIT_DEVICE_NAMESPACE_END

#endif  // SPLIT_DEVICE_INCLUDED_KMEANS_PROPAGATION
