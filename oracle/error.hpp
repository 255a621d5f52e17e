// ORACLE — test infrastructure only (see bb.hpp).
//
// Error codes, in the declaration order of src/error.rs:7-40 (Ok = 0).
#pragma once

namespace oracle {

enum ErrorCode : int {
    OK = 0,
    ERR_EMPTY_INPUT = 1,
    ERR_INDEX_OUT_OF_BOUNDS = 2,
    ERR_INVALID_DOMAIN_SIZE = 3,
    ERR_INVALID_BLOWUP = 4,
    ERR_NO_QUERIES = 5,
    ERR_BAD_MERKLE_PATH = 6,
    ERR_FRI_VERIFICATION_FAILED = 7,
    ERR_MALFORMED_INPUT = 8,
    ERR_RNG_UNAVAILABLE = 9,
};

// The reference returns Result<_, Error>; the oracle throws and the C entry
// points translate back to the code.
struct Z0pError {
    int code;
    const char* reason;
    explicit Z0pError(int c, const char* r = "") : code(c), reason(r) {}
};

}  // namespace oracle
