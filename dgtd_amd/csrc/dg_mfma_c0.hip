// Matrix-core stage, tuning configuration 0 ("wide", mfma_stage.hpp).
#include "mfma_stage.hpp"

namespace dgtd {
DGTD_INSTANTIATE_CFG(0, float)
DGTD_INSTANTIATE_CFG(0, double)
}  // namespace dgtd
