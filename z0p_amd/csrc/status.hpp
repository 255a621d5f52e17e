// Internal error carrier: thrown inside the library, translated to the C
// ABI status code at every extern "C" boundary.
#pragma once
#include "../../include/z0p.h"

namespace z0p {

struct Status {
    int code;
    explicit Status(int c) : code(c) {}
};

inline void require(bool ok, int code) {
    if (!ok) throw Status(code);
}

}  // namespace z0p
