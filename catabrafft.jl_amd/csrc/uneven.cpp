// uneven.cpp -- many transforms of different lengths in one plan (placeholder until the
// multi-plan kernels land).
#include "catabrafft.h"
#include "library.h"

namespace cfft {

int uneven_create(Ctx*, int64_t, const int64_t*, const int64_t*, int, int,
                  std::shared_ptr<Plan>*) {
    return fail(CFFT_ERR_UNSUPPORTED, "uneven plans are not available yet");
}

int uneven_exec(const Plan&, const void*, void*, int, int, hipStream_t) {
    return fail(CFFT_ERR_UNSUPPORTED, "uneven plans are not available yet");
}

}  // namespace cfft
