// Drop-in for split/include/split/device/kmeans/cluster.cuh:31-47.
#if !defined(SPLIT_DEVICE_INCLUDED_KMEANS_CLUSTER)
#define SPLIT_DEVICE_INCLUDED_KMEANS_CLUSTER

#include <cstddef>

#include "split/detail/internal.h"
#include "split/device/views.h"

SPLIT_DEVICE_NAMESPACE_BEGIN

namespace kmeans
{
/// What a Lloyd run did.
struct cluster_result
{
  int iterations;  ///< Lloyd iterations performed
  real shift;      ///< Frobenius norm of the last centroid update
};

/// Lloyd k-means.  dio_centroids holds the initial centroids (3 x K) and
/// receives the final ones; do_cluster_labels receives the assignment of the
/// last iteration (made against the centroids before the last update, as in
/// the reference).  Stops when the centroid update's Frobenius norm is
/// <= i_threshold or after i_max_iter iterations; i_threshold < 0 runs
/// exactly i_max_iter.  do_temp must hold cluster_workspace(N, K, 3) bytes.
/// Returns once the result is on the host, like the reference.
/// Throws split::device_error on failure.
SPLIT_API cluster_result
cluster(const_view2d<real> di_points,
        view2d<real> dio_centroids,
        view1d<int> do_cluster_labels,
        device_ptr<void> do_temp,
        int i_max_iter,
        real i_threshold = 1.f);

/// Temp bytes for cluster: at least the reference's formula
/// (cluster.cu:12-25), and enough that the same block also serves
/// connected_components, compress_labels and calculate_centroids with any
/// K <= N afterwards (the sample's reuse, main.cu:240-307).
SPLIT_API std::size_t
cluster_workspace(const int i_npoints, const int i_nclusters, const int i_dim);

}  // namespace kmeans

SPLIT_DEVICE_NAMESPACE_END

#endif  // SPLIT_DEVICE_INCLUDED_KMEANS_CLUSTER
