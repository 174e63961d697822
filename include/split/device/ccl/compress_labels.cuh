// Drop-in for split/include/split/device/ccl/compress_labels.cuh:21-23.
#if !defined(SPLIT_DEVICE_INCLUDED_CCL_COMPRESS_LABELS)
#define SPLIT_DEVICE_INCLUDED_CCL_COMPRESS_LABELS

#include <cstddef>

#include "split/detail/internal.h"
#include "split/device/views.h"

SPLIT_DEVICE_NAMESPACE_BEGIN

namespace ccl
{
/// In place: each label becomes the rank of its value among the sorted
/// distinct values; returns the number of distinct values.  dio_temp must
/// hold compress_labels_workspace(N, max - min + 1) bytes (N ints, as the
/// reference uses, always suffice for labels from connected_components).
/// Throws split::device_error on failure.
SPLIT_API int compress_labels(view1d<int> dio_labels,
                              device_ptr<void> dio_temp);

/// Temp bytes compress_labels needs for N labels spanning `value_span`
/// distinct integer values (max - min + 1).
SPLIT_API std::size_t compress_labels_workspace(const int i_nlabels,
                                                const long long value_span);

}  // namespace ccl

SPLIT_DEVICE_NAMESPACE_END

#endif  // SPLIT_DEVICE_INCLUDED_CCL_COMPRESS_LABELS
