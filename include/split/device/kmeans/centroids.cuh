// Drop-in for split/include/split/device/kmeans/centroids.cuh:26-40.
#if !defined(SPLIT_DEVICE_INCLUDED_KMEANS_CENTROIDS)
#define SPLIT_DEVICE_INCLUDED_KMEANS_CENTROIDS

#include <cstddef>

#include "split/detail/internal.h"
#include "split/device/views.h"

SPLIT_DEVICE_NAMESPACE_BEGIN

namespace kmeans
{
/// Mean of the points carrying each label; labels outside [0, K) are
/// ignored and a label with no points gets the zero centroid.
/// do_centroids: 3 x K.  do_temp must hold calculate_centroids_workspace(N,
/// K) bytes (a cluster_workspace(N, ...) block does for any K <= N).
/// Throws split::device_error on failure.
SPLIT_API void calculate_centroids(const_view1d<int> di_labels,
                                   const_view2d<real> di_points,
                                   view2d<real> do_centroids,
                                   device_ptr<void> do_temp);

/// Temp bytes for calculate_centroids: at least the reference's
/// N*2*sizeof(int) + K*sizeof(real) (centroids.cu:16-20); about 28 bytes
/// per label for K > 1024.
SPLIT_API std::size_t calculate_centroids_workspace(const int i_npoints,
                                                    const int i_ncentroids);
}  // namespace kmeans

SPLIT_DEVICE_NAMESPACE_END

#endif  // SPLIT_DEVICE_INCLUDED_KMEANS_CENTROIDS
