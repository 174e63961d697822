// Core definitions of the split drop-in headers: namespaces, API markers and
// the `real` type, as in split/include/split/detail/internal.h:4-39.
#if !defined(SPLIT_INCLUDED)

#define SPLIT_VERSION_MAJOR 0
#define SPLIT_VERSION_MINOR 1
#define SPLIT_VERSION_PATCH 0
#define SPLIT_VERSION 10
#define SPLIT_VERSION_MESSAGE "SPLIT: version 0.1.0 (MI355X)"
#define SPLIT_INCLUDED SPLIT_VERSION

#define SPLIT_NAMESPACE split
#define SPLIT_NAMESPACE_BEGIN namespace SPLIT_NAMESPACE {
#define SPLIT_NAMESPACE_END }

#define SPLIT_HOST_NAMESPACE_BEGIN SPLIT_NAMESPACE_BEGIN namespace host {
#define SPLIT_HOST_NAMESPACE_END }}

#define SPLIT_DEVICE_NAMESPACE_BEGIN SPLIT_NAMESPACE_BEGIN namespace device {
#define SPLIT_DEVICE_NAMESPACE_END }}

// Host entry points; the kernels behind them are reached through the C ABI
// in split_hip.h, so these headers compile with any C++17 host compiler.
#define SPLIT_API

SPLIT_NAMESPACE_BEGIN

// The gfx950 path computes in float32 (split's default `real`).
typedef float real;

SPLIT_NAMESPACE_END

#endif  // SPLIT_INCLUDED
