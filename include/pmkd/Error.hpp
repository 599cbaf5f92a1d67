// Error.hpp -- exceptions of the C++ API (the reference throws
// Core::Exception / NotImplemented, TriangleMapUpperCreator.cu:65-76,551-552,606-607).
#pragma once
#include <hip/hip_runtime_api.h>

#include <stdexcept>
#include <string>

#include "../pmkd.h"

namespace pmkd {

class Exception : public std::runtime_error {
public:
    Exception(int code, const std::string &msg) : std::runtime_error(msg), code_(code) {}
    int code() const { return code_; }

private:
    int code_;
};

class NotImplemented : public Exception {
public:
    explicit NotImplemented(const std::string &msg) : Exception(PM_ERR_NOT_IMPLEMENTED, msg) {}
};

inline void check(pm_ctx *ctx, int rc) {
    if (rc == PM_OK) return;
    const std::string msg = pm_last_error(ctx);
    if (rc == PM_ERR_NOT_IMPLEMENTED) throw NotImplemented(msg);
    throw Exception(rc, msg);
}

inline void hip_ok(hipError_t e, const char *what) {
    if (e != hipSuccess) throw Exception(PM_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

}  // namespace pmkd
