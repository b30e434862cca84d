// SPDX-License-Identifier: GPL-3.0-or-later
// Host-side helpers shared by the C-ABI translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>

#include "../../include/dgmr.h"
#include "device_math.h"

namespace dgmr {

void set_error(const std::string& msg);

#define DGMR_HIP_CHECK(expr)                                                                                   \
    do {                                                                                                       \
        hipError_t e_ = (expr);                                                                                \
        if (e_ != hipSuccess) {                                                                                \
            ::dgmr::set_error(std::string(#expr) + ": " + hipGetErrorString(e_));                             \
            return DGMR_ERR_DEVICE;                                                                            \
        }                                                                                                      \
    } while (0)

#define DGMR_ARG_CHECK(cond, msg)                                                                              \
    do {                                                                                                       \
        if (!(cond)) {                                                                                         \
            ::dgmr::set_error(msg);                                                                            \
            return DGMR_ERR_ARGUMENT;                                                                          \
        }                                                                                                      \
    } while (0)

// bump allocator over a caller-provided workspace (256-byte aligned slices)
struct Carver {
    char* base;
    size_t cap;
    size_t used = 0;
    Carver(void* b, size_t c) : base(static_cast<char*>(b)), cap(c) {}
    template <class T>
    T* take(size_t count)
    {
        const size_t bytes = ((count * sizeof(T) + 255) / 256) * 256;
        T* p = reinterpret_cast<T*>(base ? base + used : nullptr);
        used += bytes;
        return p;
    }
    bool fits() const { return used <= cap; }
};

inline int64_t ceil_div64(int64_t a, int64_t b) { return (a + b - 1) / b; }

bool valid_tiling(const dgmr_tiling* t);

// fills the device camera: focal lengths as splat_forward.cu:35-36, inverse
// projection for the marcher's rays (vol_marcher_forward.cu:302)
template <class S>
DevCamera<S> make_dev_camera(const dgmr_camera& c, const dgmr_tiling& t);

uint32_t higher_msb(uint32_t n);

} // namespace dgmr
