// common.h — error handling and small helpers shared by the engine sources.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdio>
#include <new>
#include <stdexcept>
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


// Runs f for a C entry point so that no exception crosses the ABI (a Fortran
// or Python host would terminate): engine errors keep their code, failed host
// allocations are SFD_ENOMEM, anything else SFD_EINVAL.  err (optional)
// receives the message, cleared on success.
template <class F>
static int abi_call(std::string *err, F &&f) {
    try {
        f();
        if (err) err->clear();
        return SFD_OK;
    } catch (const EngineError &e) {
        if (err) *err = e.msg;
        return e.code;
    } catch (const std::bad_alloc &) {
        if (err) *err = "host allocation failed";
        return SFD_ENOMEM;
    } catch (const std::exception &e) {
        if (err) *err = e.what();
        return SFD_EINVAL;
    }
}
