// Drop-in for split/include/split/device/kmeans/kmeans.cuh:4-8.
#if !defined(SPLIT_DEVICE_INCLUDED_KMEANS)
#define SPLIT_DEVICE_INCLUDED_KMEANS

#include "split/device/kmeans/uniform_random_initialize.cuh"
#include "split/device/kmeans/centroids.cuh"
#include "split/device/kmeans/label.cuh"
#include "split/device/kmeans/cluster.cuh"
#include "split/device/kmeans/propagation.cuh"

#endif  // SPLIT_DEVICE_INCLUDED_KMEANS
