// hsmc_mfma.hip -- split-BF16 MFMA form of the HSMC step (chosen form).
#include "pump_internal.hpp"

namespace pump {
int propagate_bits_mfma(int, const int64_t *, int, const int *, const int *, const float *,
                        const float *, const float *, const float *, float, int, const float *,
                        int, float, int, const uint64_t *, uint64_t *, int64_t, int, float *,
                        int64_t, unsigned long long *, unsigned long long *, void *, size_t,
                        hipStream_t)
{
    return PUMP_E_RANGE;
}
} // namespace pump
