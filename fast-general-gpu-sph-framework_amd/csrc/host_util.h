// host_util.h - error reporting shared by the host translation units.
#pragma once

#include <string>

#include "gsph.h"

namespace gsph {

// Records msg as the last error of ctx (or of the calling thread when ctx is
// null) and returns code.
void set_error(gsph_ctx *ctx, const std::string &msg);

inline int fail(gsph_ctx *ctx, int code, const std::string &msg) {
    set_error(ctx, msg);
    return code;
}

}  // namespace gsph
