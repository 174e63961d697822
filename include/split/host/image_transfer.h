// Host <-> device image transfer of the reference's sample
// (samples/src/main.cu:84-132, make_device_image / make_host_image, which
// the sample defines for itself; a caller replaces those definitions with
// these).  The host image is pixel-interleaved (C channels per pixel, as
// split::host::stbi::loadf returns it); the device image is C x N planes.
#if !defined(SPLIT_HOST_INCLUDED_IMAGE_TRANSFER)
#define SPLIT_HOST_INCLUDED_IMAGE_TRANSFER

#include "split/detail/internal.h"
#include "split/device/views.h"

SPLIT_HOST_NAMESPACE_BEGIN

/// Interleaved host image -> device planes (d_image.num_rows channels of
/// d_image.num_cols pixels).  One contiguous copy and a device deinterleave;
/// returns when h_image may be reused.  Throws split::device_error.
SPLIT_API void make_device_image(const real* h_image,
                                 split::device::view2d<real> d_image);

/// Device planes -> interleaved host image; returns when h_image is
/// complete.  Throws split::device_error.
SPLIT_API void make_host_image(split::device::view2d<real> d_image,
                               real* h_image);

SPLIT_HOST_NAMESPACE_END

#endif  // SPLIT_HOST_INCLUDED_IMAGE_TRANSFER
