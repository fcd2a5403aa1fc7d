// Population dynamics (Moran process) host driver.
#include "runtime_internal.h"

namespace scrna {

int population_launch(scrna_plan* P, hipStream_t st) {
    (void)P;
    (void)st;
    return set_error(SCRNA_ERR_UNSUPPORTED, "population dynamics are not available in this build");
}

}  // namespace scrna
