// Shared host/device helpers of the nect_amd library (gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

#include "../../include/nect_amd.h"

namespace nect {

// Thread-local error message behind nect_last_error().
void set_error(const std::string &msg);
void clear_error();

// Returns 0 if the last launch succeeded, else records the error and returns 1.
int32_t check_launch(const char *what);

#define NECT_REQUIRE(cond, msg)                 \
    do {                                        \
        if (!(cond)) {                          \
            ::nect::set_error(msg);             \
            return 1;                           \
        }                                       \
    } while (0)

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

constexpr int kWave = 64;       // CDNA wavefront width
constexpr int kTile = 32;       // samples per tile (the N dimension of a 32x32 MFMA)
constexpr int kXcds = 8;        // accelerator complex dies (each with its own L2)

}  // namespace nect
