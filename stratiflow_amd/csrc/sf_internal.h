// sf_internal.h -- declarations shared by the library's host translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <memory>
#include <stdexcept>
#include <string>

#include "../../include/sfhip.h"

namespace sf {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define SF_HIP(expr)                                                                                    \
    do {                                                                                                \
        hipError_t e_ = (expr);                                                                         \
        if (e_ != hipSuccess) throw ::sf::Error(SF_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define SF_REQUIRE(cond, code, msg)                  \
    do {                                             \
        if (!(cond)) throw ::sf::Error(code, msg);   \
    } while (0)

// geometry of a context, for the units outside sf_api.hip
struct CtxInfo {
    int precision;
    int N1, N2, N3, A1, A2;
    bool three;
    hipStream_t stream;
    const void* kx;  // A1 active x wavenumbers (type of the context's precision)
    const void* ky;  // A2 active y wavenumbers
};

}  // namespace sf

struct sf_field {
    sf_ctx* ctx;
    int kind;
    int bc;
    size_t n;       // scalars of type T (complex counts twice)
    void* d;        // device pointer
    bool owned;     // false for solver-owned fields
};

struct sf_ctx {
    int precision;
    std::string err;
    std::shared_ptr<void> impl;  // sf::Ctx<T>
    void* raw;
};

sf::CtxInfo sf_ctx_info(const sf_ctx* c);
