// k_learned.hip -- learned pressure-correction network (filled in below).
#include "wpf_internal.h"

namespace wpf {

void free_learned(wpf_ctx* c) { (void)c; }

}  // namespace wpf
