// k_resolve_f32_ppolish.hip — the F32 resolver kernel, kind KIND_PPOLISH (cbz_resolve.h).
#define CBZ_RESOLVE_INSTANTIATE
#include "cbz_resolve.h"

namespace cbz {
CBZ_RESOLVE_KIND(, float, KIND_PPOLISH)
}  // namespace cbz
