// Matrix-core stage, tuning configuration 1 ("deep", mfma_stage.hpp).
#include "mfma_stage.hpp"

namespace dgtd {
DGTD_INSTANTIATE_CFG(1, float)
DGTD_INSTANTIATE_CFG(1, double)
}  // namespace dgtd
