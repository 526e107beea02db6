// k_resolve_f32_polish.hip — the FP32 resolver kernel, kind KIND_POLISH (cbz_resolve.h).
#define CBZ_RESOLVE_INSTANTIATE
#include "cbz_resolve.h"

namespace cbz {
CBZ_RESOLVE_KIND(, float, KIND_POLISH)
}  // namespace cbz
