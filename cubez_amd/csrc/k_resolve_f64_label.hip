// k_resolve_f64_label.hip — the FP64 resolver kernel, kind KIND_LABEL (cbz_resolve.h).
#define CBZ_RESOLVE_INSTANTIATE
#include "cbz_resolve.h"

namespace cbz {
CBZ_RESOLVE_KIND(, double, KIND_LABEL)
}  // namespace cbz
