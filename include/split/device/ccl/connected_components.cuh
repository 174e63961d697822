// Drop-in for split/include/split/device/ccl/connected_components.cuh:25-29.
#if !defined(SPLIT_DEVICE_INCLUDED_CCL_CONNECTED_COMPONENTS)
#define SPLIT_DEVICE_INCLUDED_CCL_CONNECTED_COMPONENTS

#include "split/detail/internal.h"
#include "split/device/views.h"

SPLIT_DEVICE_NAMESPACE_BEGIN

namespace ccl
{
/// Connected components of equal cluster labels.  di_labels: H x W
/// row-major; do_labels[p] = the largest linear pixel index of p's component
/// (the fixed point the reference's max-propagation converges to).  The
/// union-find always completes, so i_max_iterations is accepted and unused.
/// i_connectivity: 4 (the reference's) or 8.  Uses the first
/// split_hip_connected_components_workspace(W, H) bytes of dio_temp, about
/// 0.375 bytes per pixel (a cluster_workspace block, which the reference's
/// callers pass, is larger).
/// Throws split::device_error on failure.
SPLIT_API void connected_components(const_view2d<int> di_labels,
                                    device_ptr<void> dio_temp,
                                    view1d<int> do_labels,
                                    int i_max_iterations = 10,
                                    int i_connectivity = 4);

}  // namespace ccl

SPLIT_DEVICE_NAMESPACE_END

#endif  // SPLIT_DEVICE_INCLUDED_CCL_CONNECTED_COMPONENTS
