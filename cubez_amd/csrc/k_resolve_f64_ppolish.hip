// k_resolve_f64_ppolish.hip — the F64 resolver kernel, kind KIND_PPOLISH (cbz_resolve.h).
#define CBZ_RESOLVE_INSTANTIATE
#include "cbz_resolve.h"

namespace cbz {
CBZ_RESOLVE_KIND(, double, KIND_PPOLISH)
}  // namespace cbz
