// Stable key-value sort used by the set-up passes (hipCUB LSD radix sort: equal
// keys keep their input order, which the deterministic CSR builds rely on).
#include <hipcub/hipcub.hpp>

#include "internal.hpp"

namespace aphip {

void stable_sort_pairs(Ctx &c, const int32_t *keys_in, int32_t *keys_out, const int32_t *vals_in, int32_t *vals_out,
                       int64_t n, int key_bits) {
    if (n == 0) return;
    APHIP_REQUIRE(n < (int64_t)INT32_MAX, "sort size exceeds int32");
    size_t temp = 0;
    APHIP_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, temp, keys_in, keys_out, vals_in, vals_out, (int)n, 0,
                                                 key_bits, c.stream));
    DevBuf<char> tmp;
    tmp.alloc(temp);
    APHIP_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, temp, keys_in, keys_out, vals_in, vals_out, (int)n, 0,
                                                 key_bits, c.stream));
    APHIP_HIP(hipStreamSynchronize(c.stream));
}

}  // namespace aphip
