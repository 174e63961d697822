// Drop-in for
// split/include/split/device/kmeans/uniform_random_initialize.cuh:20-22.
#if !defined(SPLIT_DEVICE_INCLUDED_KMEANS_UNIFORM_RANDOM_INITIALIZE)
#define SPLIT_DEVICE_INCLUDED_KMEANS_UNIFORM_RANDOM_INITIALIZE

#include <cstdint>

#include "split/detail/internal.h"
#include "split/device/views.h"

SPLIT_DEVICE_NAMESPACE_BEGIN

namespace kmeans
{
/// Copies the colours of K distinct pixels, drawn with SplitMix64 from
/// `seed`, into do_centroids (3 x K).  Reproducible for a given seed.
SPLIT_API void uniform_random_initialize(const_view2d<real> di_points,
                                         view2d<real> do_centroids,
                                         std::uint64_t seed);

/// The reference's two-argument form (seeded from the clock there); here it
/// uses a fixed default seed so that runs repeat.
SPLIT_API void uniform_random_initialize(const_view2d<real> di_points,
                                         view2d<real> do_centroids);

}  // namespace kmeans

SPLIT_DEVICE_NAMESPACE_END

#endif  // SPLIT_DEVICE_INCLUDED_KMEANS_UNIFORM_RANDOM_INITIALIZE
