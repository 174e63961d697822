// Error type thrown by the drop-in entry points when the device path fails.
// The reference surfaces device failures as thrust::system_error (a
// std::runtime_error) thrown out of its Thrust calls; this keeps "throws on
// device failure" with a status code from the C ABI attached.
#pragma once

#include <stdexcept>
#include <string>

#include "split/detail/internal.h"

SPLIT_NAMESPACE_BEGIN

class device_error : public std::runtime_error {
 public:
  device_error(int status, const std::string& what)
      : std::runtime_error(what), status_(status) {}
  int status() const noexcept { return status_; }

 private:
  int status_;
};

SPLIT_NAMESPACE_END
