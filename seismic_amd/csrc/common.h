// common.h — error handling and small helpers shared by the engine sources.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdio>
#include <string>

#include "../../include/seisfd.h"

struct EngineError {
    int code;
    std::string msg;
    EngineError(int c, std::string m) : code(c), msg(std::move(m)) {}
};

#define HIPCHK(x)                                                                      \
    do {                                                                               \
        hipError_t e_ = (x);                                                           \
        if (e_ != hipSuccess) {                                                        \
            char b_[256];                                                              \
            snprintf(b_, sizeof b_, "%s failed: %s (%s:%d)", #x, hipGetErrorString(e_), \
                     __FILE__, __LINE__);                                              \
            throw EngineError(SFD_EHIP, b_);                                           \
        }                                                                              \
    } while (0)

// Bump allocator over one device buffer: plan offsets, then allocate once.
struct Layout {
    size_t off = 0;
    size_t take(size_t bytes) {
        size_t o = (off + 255) & ~(size_t)255;
        off = o + bytes;
        return o;
    }
};

