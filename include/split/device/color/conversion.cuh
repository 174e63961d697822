// Drop-in for split/include/split/device/color/conversion.cuh:10-82.
//
// The reference's converters are thrust functors (operator() on a
// thrust::tuple<real, real, real>) applied with thrust::transform.  Here
// each converter is a tag type with the same name and template parameter
// that names the gfx950 conversion convert_color_space runs; the colour
// arithmetic lives in the device kernels only.
#if !defined(SPLIT_DEVICE_INCLUDED_COLOR_CONVERSION)
#define SPLIT_DEVICE_INCLUDED_COLOR_CONVERSION

#include <type_traits>

#include "split/detail/internal.h"
#include "split/device/views.h"
#include "split_hip.h"

SPLIT_DEVICE_NAMESPACE_BEGIN

namespace color
{
///@brief RGB -> CIE XYZ (sRGB companding, D65, scaled to 100)
template <typename ColorT = void>
struct rgb_to_xyz { static constexpr int conversion = SPLIT_HIP_RGB_TO_XYZ; };
///@brief CIE XYZ -> CIE Lab
template <typename ColorT = void>
struct xyz_to_lab { static constexpr int conversion = SPLIT_HIP_XYZ_TO_LAB; };
///@brief CIE XYZ -> RGB
template <typename ColorT = void>
struct xyz_to_rgb { static constexpr int conversion = SPLIT_HIP_XYZ_TO_RGB; };
///@brief CIE Lab -> CIE XYZ
template <typename ColorT = void>
struct lab_to_xyz { static constexpr int conversion = SPLIT_HIP_LAB_TO_XYZ; };
///@brief RGB -> CIE Lab
template <typename ColorT = void>
struct rgb_to_lab { static constexpr int conversion = SPLIT_HIP_RGB_TO_LAB; };
///@brief CIE Lab -> RGB
template <typename ColorT = void>
struct lab_to_rgb { static constexpr int conversion = SPLIT_HIP_LAB_TO_RGB; };
///@brief RGB -> (intensity, r / intensity, g / intensity)
template <typename ColorT = void>
struct rgb_to_ic { static constexpr int conversion = SPLIT_HIP_RGB_TO_IC; };

/// Runs split_hip_color_conversion `conversion` over the 3 x N points
/// (out may be in).  Throws split::device_error on failure.
SPLIT_API void convert_color_space(const_view2d<real> di_in,
                                   view2d<real> do_out,
                                   int conversion);

/***
   @brief Transforms a set of values from one color space to another via a
   converter (one of the tag types above)
   ***/
template <typename Converter,
          typename = decltype(std::decay_t<Converter>::conversion)>
SPLIT_API void convert_color_space(const_view2d<real> di_in,
                                   view2d<real> do_out,
                                   Converter&& di_converter)
{
  (void)di_converter;
  convert_color_space(di_in, do_out, std::decay_t<Converter>::conversion);
}
}  // namespace color

SPLIT_DEVICE_NAMESPACE_END

#endif  // SPLIT_DEVICE_INCLUDED_COLOR_CONVERSION
