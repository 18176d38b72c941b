// Rung "tb<k>": temporal blocking (placeholder until the kernels land; reports unsupported).
#include "sweep_impl.h"

namespace ph3 {

bool tb_supported(int, int) { return false; }

int launch_tb(const SweepArgs&, int k) {
    ph3host::set_error("temporal blocking k=%d not built", k);
    return PH3_EUNSUP;
}

}  // namespace ph3
