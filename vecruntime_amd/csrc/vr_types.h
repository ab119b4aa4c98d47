// Declarations shared by host code and device code, including the kernels compiled at run time
// with hipRTC: nothing here needs the C++ standard library.
#pragma once

#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>
#else
typedef __hip_internal::int8_t int8_t;
typedef __hip_internal::uint8_t uint8_t;
typedef __hip_internal::int16_t int16_t;
typedef __hip_internal::uint16_t uint16_t;
typedef __hip_internal::int32_t int32_t;
typedef __hip_internal::uint32_t uint32_t;
typedef __hip_internal::int64_t int64_t;
typedef __hip_internal::uint64_t uint64_t;
#define INT64_MAX 0x7fffffffffffffffLL
#define INT64_MIN (-INT64_MAX - 1)
#define INT32_MAX 0x7fffffff
#define INT32_MIN (-INT32_MAX - 1)
#define UINT32_MAX 0xffffffffu
#endif

#include "../../include/vr.h"

#define VR_MAX_KEYS 8
#define VR_MAX_VALUES 16
#define VR_MAX_OPS 24
#define VR_MAX_AGGS 24

namespace vr {

// Device-side view of one input column.
struct DCol {
    const uint8_t* valid;  // nullptr: no nulls
    const void* data;
    const int32_t* remap;  // dictionary-encoded UTF8: input code -> table dictionary id
    int32_t type;
    int32_t remap_len;     // entries of `remap` (codes outside it read as null)
};

}  // namespace vr
