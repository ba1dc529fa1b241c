// RuntimeFeatures / version (src/stub.cpp:17-43) and the host entry of the
// GPU fwht_inplace (include/bmmpy/math/fwht.hpp:8-9).
#include "bmmpy.hpp"

#include "bmmpy_hip.h"

#include <thread>

namespace bmmpy {

RuntimeFeatures get_runtime_features() {
    RuntimeFeatures f;
    f.max_threads = (int)std::max(1u, std::thread::hardware_concurrency());
    // What the host CPU offers (src/stub.cpp:24); no host path here uses it.
    __builtin_cpu_init();
    f.avx2_supported = __builtin_cpu_supports("avx2") != 0;
    bmm_device_info info{};
    if (bmm_hip_device_info(&info) == BMM_OK && info.device_count > 0) {
        f.hip_available = true;
        f.device_count = info.device_count;
        f.compute_units = info.compute_units;
        f.device_name = info.name;
        f.device_arch = info.arch;
    }
    return f;
}

const char* get_version() noexcept { return bmm_hip_version(); }

void fwht_inplace(std::int32_t* data, std::size_t size) {
    throw_status(bmm_hip_fwht_i32(data, size));
}

void fwht_inplace(std::int16_t* data, std::size_t size) {
    throw_status(bmm_hip_fwht_i16(data, size));
}

void calc_scores_and_order(const std::int16_t* h, std::size_t m, std::int32_t n,
                           std::int16_t* s_by_mask, std::int32_t* order, std::int32_t* cnt,
                           std::int32_t* off) {
    throw_status(bmm_hip_scores_order_i16(h, m, n, s_by_mask, order, cnt, off));
}

void calc_scores_and_order(const std::int32_t* h, std::size_t m, std::int32_t n,
                           std::int32_t* s_by_mask, std::int32_t* order, std::int32_t* cnt,
                           std::int32_t* off) {
    throw_status(bmm_hip_scores_order_i32(h, m, n, s_by_mask, order, cnt, off));
}

void fixed_weight_masks_u32(std::uint32_t n, std::uint32_t k, std::vector<std::uint32_t>& out) {
    out.clear();
    std::size_t count = 0;
    throw_status(bmm_hip_fixed_weight_masks_u32(n, k, nullptr, 0, &count));
    out.resize(count);
    if (count)
        throw_status(bmm_hip_fixed_weight_masks_u32(n, k, out.data(), count, &count));
}

void fixed_weight_masks_u64(std::uint32_t n, std::uint32_t k, std::vector<std::uint64_t>& out) {
    out.clear();
    std::size_t count = 0;
    throw_status(bmm_hip_fixed_weight_masks_u64(n, k, nullptr, 0, &count));
    out.resize(count);
    if (count)
        throw_status(bmm_hip_fixed_weight_masks_u64(n, k, out.data(), count, &count));
}

} // namespace bmmpy
