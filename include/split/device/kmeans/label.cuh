// Drop-in for split/include/split/device/kmeans/label.cuh:27-40.
#if !defined(SPLIT_DEVICE_INCLUDED_KMEANS_LABEL)
#define SPLIT_DEVICE_INCLUDED_KMEANS_LABEL

#include <cstddef>

#include "split/detail/internal.h"
#include "split/device/views.h"

SPLIT_DEVICE_NAMESPACE_BEGIN

namespace kmeans
{
/// Labels each point with the index of its nearest centroid (squared
/// Euclidean distance, ties to the lowest index).  di_centroids: 3 x K,
/// di_points: 3 x N (SoA planes), do_cluster_labels: N ints; all device
/// memory.  The fused kernel needs no temp; do_temp is accepted for
/// signature parity.  Throws split::device_error on failure.
SPLIT_API void label_points(const_view2d<real> di_centroids,
                            const_view2d<real> di_points,
                            view1d<int> do_cluster_labels,
                            device_ptr<void> do_temp);

/// Temp bytes callers allocate for label_points: the reference's
/// N*K*sizeof(real) (label.cu:68-72), kept so that a block sized by it still
/// serves every other entry point.
SPLIT_API std::size_t label_points_workspace(const int i_npoints,
                                             const int i_nclusters);
}  // namespace kmeans

SPLIT_DEVICE_NAMESPACE_END

#endif  // SPLIT_DEVICE_INCLUDED_KMEANS_LABEL
