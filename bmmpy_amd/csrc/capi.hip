// extern "C" boundary of libbmmpy_hip.so (declared in include/bmmpy_hip.h).
// Validation happens before any device query; device and host errors become
// BMM_* status codes with a thread-local message.
#include "common.hpp"

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

namespace bmmhip {

namespace {
thread_local std::string g_last_error;
}

void set_last_error(const char* msg) { g_last_error = msg ? msg : ""; }

KernelTimer& kernel_timer() {
    static KernelTimer t;
    return t;
}

__global__ void fill_u64_kernel(uint64_t* p, size_t n, uint64_t v) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

void fill_u64(uint64_t* p, size_t n, uint64_t v, hipStream_t s) {
    if (n == 0)
        return;
    const unsigned grid = (unsigned)std::min<size_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(fill_u64_kernel, dim3(grid), dim3(256), 0, s, p, n, v);
    BMM_HIP(hipGetLastError());
}

__global__ void fill_u32_kernel(uint32_t* p, size_t n, uint32_t v) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

void fill_u32(uint32_t* p, size_t n, uint32_t v, hipStream_t s) {
    if (n == 0)
        return;
    const unsigned grid = (unsigned)std::min<size_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(fill_u32_kernel, dim3(grid), dim3(256), 0, s, p, n, v);
    BMM_HIP(hipGetLastError());
}

// Fails with BMM_ERR_DEVICE when no device is visible (checked after argument
// validation, like the reference searchers).
void require_device() {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
        fail(BMM_ERR_DEVICE, "no HIP device is available");
}

// Implemented in the kernel translation units.
std::vector<bmm_topk_entry> ranked_topk(const uint32_t* pat, const int32_t* mult, size_t U, int t,
                                        int32_t total, size_t k);
template <typename T> void fwht_array_dev(T* d, size_t n, hipStream_t stream);
void gray_coset_min_dev(const uint64_t* d_rows, size_t n_rows, size_t words_per_row,
                        size_t block_rows, const int32_t* targets, size_t n_targets,
                        size_t chunk_bits, uint64_t* d_first_key, uint64_t* d_mask_key,
                        hipStream_t stream);
void gray_coset_min_noalloc(const uint64_t* d_rows, size_t n_rows, size_t words_per_row,
                        size_t block_rows, const int32_t* targets, size_t n_targets,
                        size_t chunk_bits, uint64_t* d_first_key, uint64_t* d_mask_key,
                        hipStream_t stream);
void coset_min_batch_dev(const uint64_t* d_rows, const int64_t* word_offset, const int32_t* rows,
                         const int32_t* cols, size_t n_mat, const int64_t* out_offset,
                         uint32_t* d_out_w, uint64_t* d_out_idx, int persistent_wgs,
                         hipStream_t stream);
int default_persistent_wgs();
void window_coset_min(const uint64_t* rows, size_t n_rows, size_t words, size_t cols,
                      const int32_t* window_rows, size_t width, const int32_t* targets,
                      size_t n_targets, uint32_t* out_w, uint64_t* out_mask);
void replay_host(uint64_t* rows, size_t n_rows, size_t words, const uint16_t* ops, size_t n_ops,
                 uint64_t* weight);
struct Ga;
Ga* ga_create(const uint64_t* matrix, size_t rows, size_t cols, const bmm_ga_params* params);
void ga_destroy(Ga* G);
void ga_run(Ga* G, size_t max_generations, uint64_t target, double time_budget_s,
            bmm_ga_status* st);
void ga_best(Ga* G, uint64_t* out, uint64_t* weight);
void ga_island_best(Ga* G, int island, uint64_t* out, uint64_t* weight);
void ga_best_ops(Ga* G, int island, uint64_t* base, uint16_t* ops, size_t* n_ops);
void ga_population(Ga* G, int island, uint16_t* ops_next, int32_t* len_next, uint16_t* ops_prev,
                   int32_t* len_prev, uint32_t* fit_prev, uint64_t* rng, bmm_ga_island_state* st);
void ga_set_island_base(Ga* G, int island, const uint64_t* matrix);
void ga_proposals(Ga* G, int island, uint16_t* out, size_t cap, size_t* n);
void ga_configure_islands(Ga* G, const uint64_t* seeds, const int32_t* migrate, size_t n);
struct Mlp;
Mlp* mlp_create(uint64_t seed);
void mlp_destroy(Mlp* M);
void mlp_set_params(Mlp* M, const float* host);
void mlp_get_params(Mlp* M, float* host);
void mlp_forward_host(Mlp* M, const float* x, size_t B, float* y);
size_t weight_order(const uint64_t* rows, size_t t, size_t words, size_t cols, size_t k,
                    uint32_t* out_mask, uint32_t* out_weight);
void pair_overlaps(const uint64_t* rows, size_t n, size_t words, uint32_t* out);
void row_combine(const uint64_t* rows, size_t n_rows, size_t words, const uint64_t* masks,
                 size_t mask_words, size_t n_out, uint64_t* out, uint64_t* out_weight);
template <typename S>
void scores_order(const S* h, size_t m, int32_t n, S* s_by_mask, int32_t* order, int32_t* cnt,
                  int32_t* off);
uint64_t binom_u64(uint32_t n, uint32_t k);
template <typename T>
void fixed_weight_masks(uint32_t n, uint32_t k, T* out);
void mlp_score_pairs_host(Mlp* M, const uint64_t* sketch, size_t rows, const uint16_t* pairs,
                          size_t n, float* y);
float mlp_grads_host(Mlp* M, const float* x, const float* t, size_t B, float* grads);
float mlp_train_host(Mlp* M, const float* x, const float* t, size_t B, float lr);
size_t mlp_num_params();
struct BatchPlan;
BatchPlan* batch_plan_create(const uint64_t* d_words, const int64_t* word_offset,
                             const int32_t* rows, const int32_t* cols, size_t n_mat,
                             const int64_t* out_offset, uint32_t* d_out_w, uint64_t* d_out_idx,
                             int persistent_wgs, int autotune);
void batch_plan_run(BatchPlan* P, hipStream_t stream, bool use_graph);
void batch_plan_destroy(BatchPlan* P);
void batch_plan_set_shape(BatchPlan* P, int max_block_bits, int persistent_wgs, int chunk);
void batch_plan_info(const BatchPlan* P, bmm_batch_info* out);
std::vector<bmm_topk_entry> fwht_topk_rows(const uint64_t* rows, size_t t, size_t words,
                                           size_t cols, size_t k);
std::vector<bmm_topk_entry> mitm_topk_split(const uint64_t* q, const uint64_t* r,
                                            const int32_t* mult, size_t unique, size_t low_bits,
                                            size_t high_bits, int32_t total, size_t k);
void gray_check_args(size_t n_rows, size_t words_per_row, size_t block_rows,
                     const int32_t* targets, size_t n_targets, size_t chunk_bits);

namespace {
struct CapiWorkspace {
    DevBuf a, b, c, d;
};
CapiWorkspace& capi_ws() {
    thread_local CapiWorkspace ws;
    return ws;
}

template <typename T> void fwht_host(T* data, size_t n) {
    if (n <= 1)
        return;
    if ((n & (n - 1)) != 0)
        fail(BMM_ERR_INVALID, "fwht_inplace: size must be a power of two");
    require_device();
    T* d = capi_ws().a.get<T>(n);
    BMM_HIP(hipMemcpy(d, data, n * sizeof(T), hipMemcpyHostToDevice));
    fwht_array_dev<T>(d, n, nullptr);
    BMM_HIP(hipMemcpy(data, d, n * sizeof(T), hipMemcpyDeviceToHost));
}

// fixed_weight_masks_u32 / _u64 (src/math/comb.cpp:13-82): the reference's
// checks in its order (k = 0 first, then the mask width, then k > n).
template <typename T>
int fixed_weight_entry(uint32_t n, uint32_t k, T* out, size_t capacity, size_t* count,
                       const char* name) {
    return guarded([&] {
        constexpr uint32_t width = 8 * sizeof(T);
        if (!count)
            fail(BMM_ERR_INVALID, std::string(name) + ": null count");
        if (k == 0) {
            *count = 1;
            if (out && capacity >= 1)
                out[0] = 0;
            return;
        }
        if (n > width)
            fail(BMM_ERR_INVALID, std::string(name) + ": n exceeds " + std::to_string(width) +
                                      "-bit mask width");
        const uint64_t c = binom_u64(n, k);
        *count = (size_t)c;
        if (c == 0 || !out)
            return;
        if (capacity < c)
            fail(BMM_ERR_INVALID, std::string(name) + ": output capacity below C(n, k)");
        require_device();
        fixed_weight_masks<T>(n, k, out);
    });
}

template <typename S>
int scores_order_entry(const S* h, size_t m, int32_t n, S* s_by_mask, int32_t* order,
                       int32_t* cnt, int32_t* off) {
    return guarded([&] {
        if (n < 0)
            fail(BMM_ERR_INVALID, "calc_scores_and_order: n must be non-negative");
        if (!cnt || !off || (m > 0 && (!h || !s_by_mask)) || (m > 1 && !order))
            fail(BMM_ERR_INVALID, "calc_scores_and_order: null pointer");
        if (n > (1 << 28))
            fail(BMM_ERR_INVALID, "calc_scores_and_order: n must be at most 2^28");
        if (m > ((size_t)1 << 31))
            fail(BMM_ERR_INVALID, "calc_scores_and_order: m must be at most 2^31");
        if (m == 0) {   // empty spectrum: zero counts and offsets (fwht.cpp:71-76)
            std::fill(cnt, cnt + (size_t)n + 1, 0);
            std::fill(off, off + (size_t)n + 1, 0);
            return;
        }
        require_device();
        scores_order<S>(h, m, n, s_by_mask, order, cnt, off);
    });
}
} // namespace

} // namespace bmmhip

using namespace bmmhip;

extern "C" {

const char* bmm_hip_last_error(void) { return g_last_error.c_str(); }

const char* bmm_hip_version(void) { return "0.1.0"; }

int bmm_hip_timing_begin(void) {
    return guarded([&] {
        KernelTimer& t = kernel_timer();
        for (auto& e : t.events) {
            (void)hipEventDestroy(e.first);
            (void)hipEventDestroy(e.second);
        }
        t.events.clear();
        t.enabled = true;
    });
}

int bmm_hip_timing_end(double* total_ms, size_t* launches) {
    return guarded([&] {
        if (!total_ms || !launches)
            fail(BMM_ERR_INVALID, "timing_end: null output");
        KernelTimer& t = kernel_timer();
        t.enabled = false;
        double sum = 0.0;
        for (auto& e : t.events) {
            BMM_HIP(hipEventSynchronize(e.second));
            float ms = 0.f;
            BMM_HIP(hipEventElapsedTime(&ms, e.first, e.second));
            sum += ms;
            (void)hipEventDestroy(e.first);
            (void)hipEventDestroy(e.second);
        }
        *total_ms = sum;
        *launches = t.events.size();
        t.events.clear();
    });
}

int bmm_hip_device_info(bmm_device_info* out) {
    return guarded([&] {
        if (!out)
            fail(BMM_ERR_INVALID, "device_info: null output");
        std::memset(out, 0, sizeof(*out));
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
            (void)hipGetLastError();
            out->device_count = 0;
            out->device = -1;
            return;
        }
        out->device_count = n;
        int dev = 0;
        BMM_HIP(hipGetDevice(&dev));
        hipDeviceProp_t p{};
        BMM_HIP(hipGetDeviceProperties(&p, dev));
        out->device = dev;
        out->compute_units = p.multiProcessorCount;
        out->wave_size = p.warpSize;
        out->lds_per_block = (int)p.sharedMemPerBlock;
        out->clock_khz = p.clockRate;
        std::strncpy(out->arch, p.gcnArchName, sizeof(out->arch) - 1);
        std::strncpy(out->name, p.name, sizeof(out->name) - 1);
    });
}

int bmm_hip_fwht_i32(int32_t* data, size_t n) {
    return guarded([&] {
        if (n > 1 && !data)
            fail(BMM_ERR_INVALID, "fwht_inplace: null data");
        fwht_host<int32_t>(data, n);
    });
}

int bmm_hip_fwht_i16(int16_t* data, size_t n) {
    return guarded([&] {
        if (n > 1 && !data)
            fail(BMM_ERR_INVALID, "fwht_inplace: null data");
        fwht_host<int16_t>(data, n);
    });
}

int bmm_hip_fwht_i32_dev(int32_t* d_data, size_t n, void* stream) {
    return guarded([&] { fwht_array_dev<int32_t>(d_data, n, (hipStream_t)stream); });
}

int bmm_hip_fwht_i16_dev(int16_t* d_data, size_t n, void* stream) {
    return guarded([&] { fwht_array_dev<int16_t>(d_data, n, (hipStream_t)stream); });
}

int bmm_hip_gray_coset_min_dev(const uint64_t* d_rows, size_t n_rows, size_t words_per_row,
                               size_t block_rows, const int32_t* targets, size_t n_targets,
                               size_t chunk_bits, uint64_t* d_first_key, uint64_t* d_mask_key,
                               void* stream) {
    return guarded([&] {
        gray_coset_min_dev(d_rows, n_rows, words_per_row, block_rows, targets, n_targets,
                           chunk_bits, d_first_key, d_mask_key, (hipStream_t)stream);
    });
}

int bmm_hip_gray_coset_min_noalloc(const uint64_t* d_rows, size_t n_rows, size_t words_per_row,
                                   size_t block_rows, const int32_t* targets, size_t n_targets,
                                   size_t chunk_bits, uint64_t* d_first_key,
                                   uint64_t* d_mask_key, void* stream) {
    return guarded([&] {
        gray_coset_min_noalloc(d_rows, n_rows, words_per_row, block_rows, targets, n_targets,
                               chunk_bits, d_first_key, d_mask_key, (hipStream_t)stream);
    });
}

int bmm_hip_gray_coset_min(const uint64_t* row_words, size_t n_rows, size_t words_per_row,
                           size_t block_rows, const int32_t* targets, size_t n_targets,
                           uint32_t* out_weight, uint64_t* out_first, uint64_t* out_mask) {
    return guarded([&] {
        if (!row_words || !targets || !out_weight || !out_first || !out_mask)
            fail(BMM_ERR_INVALID, "gray_coset_min: null pointer");
        gray_check_args(n_rows, words_per_row, block_rows, targets, n_targets, 0);
        require_device();
        CapiWorkspace& ws = capi_ws();
        uint64_t* d_rows = ws.a.get<uint64_t>(n_rows * words_per_row);
        uint64_t* d_k = ws.b.get<uint64_t>(2 * n_targets);
        BMM_HIP(hipMemcpy(d_rows, row_words, n_rows * words_per_row * 8, hipMemcpyHostToDevice));
        gray_coset_min_dev(d_rows, n_rows, words_per_row, block_rows, targets, n_targets, 0,
                           d_k, d_k + n_targets, nullptr);
        std::vector<uint64_t> k(2 * n_targets);
        BMM_HIP(hipMemcpy(k.data(), d_k, k.size() * 8, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n_targets; ++i) {
            out_weight[i] = (uint32_t)(k[i] >> 32);
            out_first[i] = k[i] & 0xffffffffull;
            out_mask[i] = k[n_targets + i] & 0xffffffffull;
        }
    });
}

int bmm_hip_coset_min_batch(const uint64_t* row_words, const int64_t* word_offset,
                            const int32_t* rows, const int32_t* cols, size_t n_mat,
                            const int64_t* out_offset, uint32_t* out_weight,
                            uint64_t* out_index) {
    return guarded([&] {
        if (n_mat == 0)
            return;
        if (!row_words || !word_offset || !rows || !cols || !out_offset || !out_weight ||
            !out_index)
            fail(BMM_ERR_INVALID, "coset_min_batch: null pointer");
        size_t words = 0, slots = 0;
        for (size_t m = 0; m < n_mat; ++m) {
            if (rows[m] < 2 || rows[m] > 31)
                fail(BMM_ERR_INVALID, "coset_min: rows must be in [2, 31]");
            if (cols[m] < 0 || word_offset[m] < 0 || out_offset[m] < 0)
                fail(BMM_ERR_INVALID, "coset_min: negative size or offset");
            words = std::max(words, (size_t)word_offset[m] +
                                        (size_t)rows[m] * (size_t)((cols[m] + 63) / 64));
            slots = std::max(slots, (size_t)out_offset[m] + (size_t)rows[m]);
        }
        require_device();
        CapiWorkspace& ws = capi_ws();
        uint64_t* d_rows = ws.a.get<uint64_t>(words);
        uint32_t* d_w = ws.b.get<uint32_t>(slots);
        uint64_t* d_i = ws.c.get<uint64_t>(slots);
        BMM_HIP(hipMemcpy(d_rows, row_words, words * 8, hipMemcpyHostToDevice));
        coset_min_batch_dev(d_rows, word_offset, rows, cols, n_mat, out_offset, d_w, d_i,
                            default_persistent_wgs(), nullptr);
        BMM_HIP(hipMemcpy(out_weight, d_w, slots * 4, hipMemcpyDeviceToHost));
        BMM_HIP(hipMemcpy(out_index, d_i, slots * 8, hipMemcpyDeviceToHost));
    });
}

int bmm_hip_coset_min_batch_dev(const uint64_t* d_row_words, const int64_t* word_offset,
                                const int32_t* rows, const int32_t* cols, size_t n_mat,
                                const int64_t* out_offset, uint32_t* d_out_weight,
                                uint64_t* d_out_index, size_t persistent_wgs, void* stream) {
    return guarded([&] {
        if (n_mat == 0)
            return;
        if (!d_row_words || !word_offset || !rows || !cols || !out_offset || !d_out_weight ||
            !d_out_index)
            fail(BMM_ERR_INVALID, "coset_min_batch: null pointer");
        coset_min_batch_dev(d_row_words, word_offset, rows, cols, n_mat, out_offset, d_out_weight,
                            d_out_index,
                            persistent_wgs ? (int)persistent_wgs : default_persistent_wgs(),
                            (hipStream_t)stream);
    });
}

int bmm_hip_batch_plan_create(const uint64_t* d_row_words, const int64_t* word_offset,
                              const int32_t* rows, const int32_t* cols, size_t n_mat,
                              const int64_t* out_offset, uint32_t* d_out_weight,
                              uint64_t* d_out_index, size_t persistent_wgs, int autotune,
                              void** plan) {
    return guarded([&] {
        if (!plan)
            fail(BMM_ERR_INVALID, "batch_plan_create: null handle");
        *plan = nullptr;
        if (!d_row_words || !word_offset || !rows || !cols || !out_offset || !d_out_weight ||
            !d_out_index)
            fail(BMM_ERR_INVALID, "batch_plan_create: null pointer");
        *plan = batch_plan_create(d_row_words, word_offset, rows, cols, n_mat, out_offset,
                                  d_out_weight, d_out_index, (int)persistent_wgs, autotune);
    });
}

int bmm_hip_batch_plan_run(void* plan, int use_graph, void* stream) {
    return guarded([&] {
        if (!plan)
            fail(BMM_ERR_INVALID, "batch_plan_run: null handle");
        batch_plan_run(static_cast<BatchPlan*>(plan), (hipStream_t)stream, use_graph != 0);
    });
}

int bmm_hip_batch_plan_info(const void* plan, bmm_batch_info* out) {
    return guarded([&] {
        if (!plan || !out)
            fail(BMM_ERR_INVALID, "batch_plan_info: null pointer");
        batch_plan_info(static_cast<const BatchPlan*>(plan), out);
    });
}

int bmm_hip_batch_plan_set_shape(void* plan, int max_block_bits, size_t persistent_wgs,
                                 size_t chunk) {
    return guarded([&] {
        if (!plan)
            fail(BMM_ERR_INVALID, "batch_plan_set_shape: null handle");
        if (persistent_wgs > (1u << 20) || chunk > (1u << 20))
            fail(BMM_ERR_INVALID, "batch_plan_set_shape: launch size too large");
        batch_plan_set_shape(static_cast<BatchPlan*>(plan), max_block_bits, (int)persistent_wgs,
                             (int)chunk);
    });
}

int bmm_hip_batch_plan_destroy(void* plan) {
    return guarded([&] { batch_plan_destroy(static_cast<BatchPlan*>(plan)); });
}

int bmm_hip_window_coset_min(const uint64_t* row_words, size_t n_rows, size_t words_per_row,
                             size_t cols, const int32_t* window_rows, size_t width,
                             const int32_t* targets, size_t n_targets, uint32_t* out_weight,
                             uint64_t* out_mask) {
    return guarded([&] {
        if (n_targets == 0)
            return;
        if (!row_words || !window_rows || !targets || !out_weight || !out_mask)
            fail(BMM_ERR_INVALID, "window_coset_min: null pointer");
        if (width == 0 || width > 47)
            fail(BMM_ERR_INVALID, "window_coset_min: window width must be in [1, 47]");
        for (size_t i = 0; i < n_targets; ++i) {
            if (targets[i] < 0 || (size_t)targets[i] >= n_rows)
                fail(BMM_ERR_RANGE, "window_coset_min: target row out of range");
            for (size_t j = 0; j < width; ++j)
                if (window_rows[i * width + j] < 0 ||
                    (size_t)window_rows[i * width + j] >= n_rows)
                    fail(BMM_ERR_RANGE, "window_coset_min: window row out of range");
        }
        require_device();
        window_coset_min(row_words, n_rows, words_per_row, cols, window_rows, width, targets,
                         n_targets, out_weight, out_mask);
    });
}

int bmm_hip_coset_min(const uint64_t* row_words, size_t n_rows, size_t words_per_row,
                      size_t cols, uint32_t* out_weight, uint64_t* out_index) {
    if (words_per_row != (cols + 63) / 64) {
        set_last_error("coset_min: words_per_row must equal ceil(cols / 64)");
        return BMM_ERR_INVALID;
    }
    const int64_t woff = 0, ooff = 0;
    const int32_t r = (int32_t)n_rows, c = (int32_t)cols;
    if (n_rows > 31 || cols > (size_t)INT32_MAX) {
        set_last_error("coset_min: rows must be in [2, 31]");
        return BMM_ERR_INVALID;
    }
    return bmm_hip_coset_min_batch(row_words, &woff, &r, &c, 1, &ooff, out_weight, out_index);
}

// Column patterns of a window (bit r = row r) for the full-ranking path,
// one entry per active column among the first `cols` bit positions.
void window_patterns(const uint64_t* row_words, size_t rows, size_t words_per_row, size_t cols,
                     std::vector<uint32_t>& pat, std::vector<int32_t>& mult) {
    for (size_t wi = 0; wi < words_per_row && wi * 64 < cols; ++wi) {
        uint64_t active = 0;
        for (size_t r = 0; r < rows; ++r)
            active |= row_words[r * words_per_row + wi];
        const size_t left = cols - wi * 64;
        if (left < 64)
            active &= (1ull << left) - 1;
        while (active) {
            const int bit = __builtin_ctzll(active);
            uint32_t p = 0;
            for (size_t r = 0; r < rows; ++r)
                p |= (uint32_t)((row_words[r * words_per_row + wi] >> bit) & 1u) << r;
            pat.push_back(p);
            active &= active - 1;
        }
    }
    if (pat.size() > (size_t)INT32_MAX)
        fail(BMM_ERR_OVERFLOW, "topk: total weight exceeds int32 range");
    mult.assign(pat.size(), 1);
}

std::vector<bmm_topk_entry> ranked_rows(const uint64_t* row_words, size_t rows,
                                        size_t words_per_row, size_t cols, size_t k) {
    if (rows > 30)
        fail(BMM_ERR_INVALID, "max_candidates above 4096 needs a window of at most 30 rows");
    std::vector<uint32_t> pat;
    std::vector<int32_t> mult;
    window_patterns(row_words, rows, words_per_row, cols, pat, mult);
    return ranked_topk(pat.data(), mult.data(), pat.size(), (int)rows, (int32_t)pat.size(), k);
}

int bmm_hip_fwht_topk(const uint64_t* row_words, size_t rows, size_t words_per_row,
                      size_t cols, size_t k, bmm_topk_entry* out, size_t* n_out) {
    return guarded([&] {
        if (!n_out || (k && !out))
            fail(BMM_ERR_INVALID, "fwht_topk: null output");
        *n_out = 0;
        if (rows == 0 || k == 0 || words_per_row == 0)
            return;
        if (!row_words)
            fail(BMM_ERR_INVALID, "fwht_topk: null rows");
        if (rows > 48)
            fail(BMM_ERR_INVALID, "fwht_topk: window rows must be <= 48");
        if (cols > words_per_row * 64)
            fail(BMM_ERR_INVALID, "fwht_topk: cols exceeds words_per_row * 64");
        if (k > BMM_TOPK_MAX && rows > 30)
            fail(BMM_ERR_INVALID, "fwht_topk: max_candidates above 4096 needs a window of at "
                                  "most 30 rows");
        require_device();
        auto res = k > BMM_TOPK_MAX ? ranked_rows(row_words, rows, words_per_row, cols, k)
                                    : fwht_topk_rows(row_words, rows, words_per_row, cols, k);
        std::copy(res.begin(), res.end(), out);
        *n_out = res.size();
    });
}

int bmm_hip_mitm_topk(const uint64_t* q, const uint64_t* r, const int32_t* mult, size_t unique,
                      size_t low_bits, size_t high_bits, int32_t total, size_t k,
                      bmm_topk_entry* out, size_t* n_out) {
    return guarded([&] {
        if (!n_out || (k && !out))
            fail(BMM_ERR_INVALID, "mitm_topk: null output");
        *n_out = 0;
        if (k == 0 || unique == 0)
            return;
        if (!q || !r || !mult)
            fail(BMM_ERR_INVALID, "mitm_topk: null input");
        const size_t t = low_bits + high_bits;
        if (t == 0 || t > 48 || low_bits > 48 || high_bits > 48)
            fail(BMM_ERR_INVALID, "mitm_topk: low_bits + high_bits must be in [1, 48]");
        if (k > BMM_TOPK_MAX && t > 30)
            fail(BMM_ERR_INVALID, "mitm_topk: max_candidates above 4096 needs a window of at "
                                  "most 30 rows");
        int64_t sum = 0;
        for (size_t i = 0; i < unique; ++i) {
            if ((low_bits < 64 && r[i] >> low_bits) || (high_bits < 64 && q[i] >> high_bits))
                fail(BMM_ERR_INVALID,
                     "mitm_topk: pattern index is outside the selected split");
            if (mult[i] <= 0)
                fail(BMM_ERR_INVALID, "mitm_topk: multiplicity must be positive");
            sum += mult[i];
        }
        if (sum != total)
            fail(BMM_ERR_INVALID, "mitm_topk: total_weight differs from the multiplicity sum");
        require_device();
        std::vector<bmm_topk_entry> res;
        if (k > BMM_TOPK_MAX) {
            std::vector<uint32_t> pat(unique);
            for (size_t i = 0; i < unique; ++i)
                pat[i] = (uint32_t)((q[i] << low_bits) | r[i]);
            res = ranked_topk(pat.data(), mult, unique, (int)t, total, k);
        } else {
            res = mitm_topk_split(q, r, mult, unique, low_bits, high_bits, total, k);
        }
        std::copy(res.begin(), res.end(), out);
        *n_out = res.size();
    });
}

int bmm_hip_bruteforce_topk(const uint64_t* row_words, size_t rows, size_t words_per_row,
                            size_t chunk_bits, size_t k, bmm_topk_entry* out, size_t* n_out) {
    return guarded([&] {
        if (!n_out || (k && !out))
            fail(BMM_ERR_INVALID, "bruteforce_topk: null output");
        *n_out = 0;
        if (rows == 0 || k == 0 || words_per_row == 0)
            return;
        if (!row_words)
            fail(BMM_ERR_INVALID, "bruteforce_topk: null rows");
        if (rows > 48)
            fail(BMM_ERR_INVALID, "bruteforce_topk: window rows must be <= 48");
        if (chunk_bits != 0 && (chunk_bits > rows || chunk_bits >= 64))
            fail(BMM_ERR_INVALID, "bruteforce_topk: chunk_bits must be in [1, rows] or 0");
        if (k > BMM_TOPK_MAX && rows > 30)
            fail(BMM_ERR_INVALID, "bruteforce_topk: max_candidates above 4096 needs a window of "
                                  "at most 30 rows");
        require_device();
        // Popcount over all words_per_row words, tail bits included.
        auto res = k > BMM_TOPK_MAX
                       ? ranked_rows(row_words, rows, words_per_row, words_per_row * 64, k)
                       : fwht_topk_rows(row_words, rows, words_per_row, words_per_row * 64, k);
        std::copy(res.begin(), res.end(), out);
        *n_out = res.size();
    });
}

int bmm_hip_replay(uint64_t* rows, size_t n_rows, size_t words_per_row, const uint16_t* ops,
                   size_t n_ops, uint64_t* out_weight) {
    return guarded([&] {
        if (!out_weight || (n_ops && !ops) || (n_rows && words_per_row && !rows))
            fail(BMM_ERR_INVALID, "replay: null pointer");
        if (n_rows > 65536 || n_rows * words_per_row > ((size_t)1 << 34))
            fail(BMM_ERR_INVALID, "replay: matrix too large");
        if (n_ops > (size_t)INT32_MAX)
            fail(BMM_ERR_INVALID, "replay: too many ops");
        for (size_t i = 0; i < 2 * n_ops; ++i)
            if (ops[i] >= n_rows)
                fail(BMM_ERR_RANGE, "replay: row index out of range");
        *out_weight = 0;
        if (n_rows == 0 || words_per_row == 0)
            return;
        require_device();
        replay_host(rows, n_rows, words_per_row, ops, n_ops, out_weight);
    });
}

int bmm_hip_ga_default_params(bmm_ga_params* p) {
    return guarded([&] {
        if (!p)
            fail(BMM_ERR_INVALID, "ga_default_params: null output");
        std::memset(p, 0, sizeof(*p));
        p->islands = 64;
        p->population = 256;
        p->max_ops = 512;
        p->tournament = 3;
        p->elite = 2;
        p->migrate_every = 20;
        p->mutations = 2;
        p->learned = 0;
        p->n_candidates = 256;
        p->train_steps = 8;
        p->check_every = 10;
        p->use_graph = 1;
        p->li_sweeps = 1;
        p->p_learned = 0.5f;
        p->p_append = 0.6f;
        p->p_replace = 0.25f;
        p->p_delete = 0.15f;
        p->learning_rate = 1e-3f;
        p->seed = 0x5EED0005ull;
        // Six information-set trials (one round, side by side) and no window
        // rounds per epoch: on the planted 256 x 8192 instance the trials set
        // the number of epochs (one with five or more, two with three or
        // four, two or three with two; median time to 1.01 x planted
        // 0.0076 s with six or five, 0.0116 with four), and window rounds
        // (8.6 ms each) never saved one.
        p->win_rows = 20;
        p->win_rounds = 0;
        p->win_candidates = 32;
        p->isd_trials = 6;
        p->num_parents = 2;
        p->num_offspring = 1;
        p->mutation_rate = 0.f;
        p->catastrophe_threshold = 0;
        p->catastrophe_survival = 0.2f;
        p->restart_after = 0;
        p->restart_p = 0.2f;
        // The improvement steps once before the first generation: on the
        // planted 256 x 8192 instance 1.01 x planted in a median 0.018 s
        // instead of 0.053 s (the first epoch no longer evolves long op
        // lists on the heavy input).
        p->init_rounds = 1;
    });
}

int bmm_hip_ga_create(const uint64_t* matrix, size_t rows, size_t cols,
                      const bmm_ga_params* params, void** ga) {
    return guarded([&] {
        if (!matrix || !params || !ga)
            fail(BMM_ERR_INVALID, "ga_create: null pointer");
        *ga = nullptr;
        if (rows < 2 || rows > 256 || cols == 0 || cols % 2048 != 0)
            fail(BMM_ERR_INVALID, "ga_create: rows in [2, 256] and cols a multiple of 2048");
        require_device();
        *ga = ga_create(matrix, rows, cols, params);
    });
}

int bmm_hip_ga_run(void* ga, size_t max_generations, uint64_t target_weight,
                   double time_budget_s, bmm_ga_status* out) {
    return guarded([&] {
        if (!ga || !out)
            fail(BMM_ERR_INVALID, "ga_run: null pointer");
        ga_run(static_cast<Ga*>(ga), max_generations, target_weight, time_budget_s, out);
    });
}

int bmm_hip_ga_best(void* ga, uint64_t* out_matrix, uint64_t* out_weight) {
    return guarded([&] {
        if (!ga || !out_matrix || !out_weight)
            fail(BMM_ERR_INVALID, "ga_best: null pointer");
        ga_best(static_cast<Ga*>(ga), out_matrix, out_weight);
    });
}

int bmm_hip_ga_island_best(void* ga, int island, uint64_t* out_matrix, uint64_t* out_weight) {
    return guarded([&] {
        if (!ga || !out_matrix || !out_weight)
            fail(BMM_ERR_INVALID, "ga_island_best: null pointer");
        ga_island_best(static_cast<Ga*>(ga), island, out_matrix, out_weight);
    });
}

int bmm_hip_ga_configure_islands(void* ga, const uint64_t* seeds, const int32_t* migrate,
                                 size_t n) {
    return guarded([&] {
        if (!ga)
            fail(BMM_ERR_INVALID, "ga_configure_islands: null pointer");
        ga_configure_islands(static_cast<Ga*>(ga), seeds, migrate, n);
    });
}

int bmm_hip_ga_best_ops(void* ga, int island, uint64_t* out_base, uint16_t* out_ops,
                        size_t* n_ops) {
    return guarded([&] {
        if (!ga || !out_base || !out_ops || !n_ops)
            fail(BMM_ERR_INVALID, "ga_best_ops: null pointer");
        ga_best_ops(static_cast<Ga*>(ga), island, out_base, out_ops, n_ops);
    });
}

int bmm_hip_ga_proposals(void* ga, int island, uint16_t* out_pairs, size_t capacity,
                         size_t* n_pairs) {
    return guarded([&] {
        if (!ga || !n_pairs || (capacity && !out_pairs))
            fail(BMM_ERR_INVALID, "ga_proposals: null pointer");
        ga_proposals(static_cast<Ga*>(ga), island, out_pairs, capacity, n_pairs);
    });
}

int bmm_hip_ga_set_island_base(void* ga, int island, const uint64_t* matrix) {
    return guarded([&] {
        if (!ga || !matrix)
            fail(BMM_ERR_INVALID, "ga_set_island_base: null pointer");
        ga_set_island_base(static_cast<Ga*>(ga), island, matrix);
    });
}

int bmm_hip_ga_population(void* ga, int island, uint16_t* ops_next, int32_t* len_next,
                          uint16_t* ops_prev, int32_t* len_prev, uint32_t* fit_prev,
                          uint64_t* rng, bmm_ga_island_state* state) {
    return guarded([&] {
        if (!ga)
            fail(BMM_ERR_INVALID, "ga_population: null pointer");
        ga_population(static_cast<Ga*>(ga), island, ops_next, len_next, ops_prev, len_prev,
                      fit_prev, rng, state);
    });
}

int bmm_hip_ga_destroy(void* ga) {
    return guarded([&] { ga_destroy(static_cast<Ga*>(ga)); });
}

size_t bmm_hip_mlp_num_params(void) { return mlp_num_params(); }

int bmm_hip_mlp_create(uint64_t seed, void** mlp) {
    return guarded([&] {
        if (!mlp)
            fail(BMM_ERR_INVALID, "mlp_create: null handle");
        require_device();
        *mlp = mlp_create(seed);
    });
}

int bmm_hip_mlp_destroy(void* mlp) {
    return guarded([&] { mlp_destroy(static_cast<Mlp*>(mlp)); });
}

int bmm_hip_mlp_set_params(void* mlp, const float* params) {
    return guarded([&] {
        if (!mlp || !params)
            fail(BMM_ERR_INVALID, "mlp_set_params: null pointer");
        mlp_set_params(static_cast<Mlp*>(mlp), params);
    });
}

int bmm_hip_mlp_get_params(void* mlp, float* params) {
    return guarded([&] {
        if (!mlp || !params)
            fail(BMM_ERR_INVALID, "mlp_get_params: null pointer");
        mlp_get_params(static_cast<Mlp*>(mlp), params);
    });
}

int bmm_hip_mlp_forward(void* mlp, const float* x, size_t batch, float* y) {
    return guarded([&] {
        if (!mlp || !x || !y || batch == 0)
            fail(BMM_ERR_INVALID, "mlp_forward: null pointer or empty batch");
        mlp_forward_host(static_cast<Mlp*>(mlp), x, batch, y);
    });
}

int bmm_hip_pair_overlaps(const uint64_t* row_words, size_t n, size_t words, uint32_t* out) {
    return guarded([&] {
        if ((n > 0 && !out) || (n > 0 && words > 0 && !row_words))
            fail(BMM_ERR_INVALID, "pair_overlaps: null pointer");
        if (n > 65536)
            fail(BMM_ERR_INVALID, "pair_overlaps: at most 65536 rows");
        if (words >= ((size_t)1 << 25))
            fail(BMM_ERR_OVERFLOW, "pair_overlaps: overlaps would exceed 31 bits");
        if (n == 0)
            return;
        require_device();
        pair_overlaps(row_words, n, words, out);
    });
}

int bmm_hip_row_combine(const uint64_t* row_words, size_t n_rows, size_t words,
                        const uint64_t* masks, size_t mask_words, size_t n_out,
                        uint64_t* out_words, uint64_t* out_weight) {
    return guarded([&] {
        if (!out_weight || (n_rows * words > 0 && !row_words) ||
            (n_out * mask_words > 0 && !masks) || (n_out * words > 0 && !out_words))
            fail(BMM_ERR_INVALID, "row_combine: null pointer");
        *out_weight = 0;
        if (words >= ((size_t)1 << 31) || n_out >= ((size_t)1 << 31) ||
            n_rows >= ((size_t)1 << 31) || mask_words >= ((size_t)1 << 25))
            fail(BMM_ERR_INVALID, "row_combine: dimensions too large");
        if (n_out * words >= ((size_t)1 << 34))
            fail(BMM_ERR_INVALID, "row_combine: output larger than 2^34 words");
        for (size_t i = 0; i < n_out * mask_words; ++i) {
            const size_t w = i % mask_words;
            const uint64_t allowed = w * 64 >= n_rows
                                         ? 0ull
                                         : (n_rows - w * 64 >= 64 ? ~0ull
                                                                  : (1ull << (n_rows - w * 64)) - 1);
            if (masks[i] & ~allowed)
                fail(BMM_ERR_RANGE, "row_combine: mask selects a row beyond n_rows");
        }
        if (n_out == 0 || words == 0)
            return;
        require_device();
        row_combine(row_words, n_rows, words, masks, mask_words, n_out, out_words, out_weight);
    });
}

int bmm_hip_fixed_weight_masks_u32(uint32_t n, uint32_t k, uint32_t* out, size_t capacity,
                                   size_t* count) {
    return fixed_weight_entry<uint32_t>(n, k, out, capacity, count, "fixed_weight_masks_u32");
}

int bmm_hip_fixed_weight_masks_u64(uint32_t n, uint32_t k, uint64_t* out, size_t capacity,
                                   size_t* count) {
    return fixed_weight_entry<uint64_t>(n, k, out, capacity, count, "fixed_weight_masks_u64");
}

int bmm_hip_scores_order_i32(const int32_t* h, size_t m, int32_t n, int32_t* s_by_mask,
                             int32_t* order, int32_t* cnt, int32_t* off) {
    return scores_order_entry<int32_t>(h, m, n, s_by_mask, order, cnt, off);
}

int bmm_hip_scores_order_i16(const int16_t* h, size_t m, int32_t n, int16_t* s_by_mask,
                             int32_t* order, int32_t* cnt, int32_t* off) {
    return scores_order_entry<int16_t>(h, m, n, s_by_mask, order, cnt, off);
}

int bmm_hip_weight_order(const uint64_t* row_words, size_t t, size_t words, size_t cols,
                         size_t k, uint32_t* out_mask, uint32_t* out_weight, size_t* n_out) {
    return guarded([&] {
        if (!n_out || (t > 0 && words > 0 && !row_words) || (k > 0 && (!out_mask || !out_weight)))
            fail(BMM_ERR_INVALID, "weight_order: null pointer");
        *n_out = 0;
        if (t > 32)
            fail(BMM_ERR_INVALID, "weight_order: at most 32 rows");
        if (cols > words * 64)
            fail(BMM_ERR_INVALID, "weight_order: cols exceeds words_per_row * 64");
        if (cols >= (size_t)1 << 31)
            fail(BMM_ERR_OVERFLOW, "weight_order: too many columns");
        if (t == 0 || k == 0)
            return;
        require_device();
        *n_out = weight_order(row_words, t, words, cols, k, out_mask, out_weight);
    });
}

int bmm_hip_mlp_score_pairs(void* mlp, const uint64_t* sketch, size_t rows,
                            const uint16_t* pairs, size_t n, float* y) {
    return guarded([&] {
        if (!mlp || !sketch || !pairs || !y || n == 0)
            fail(BMM_ERR_INVALID, "mlp_score_pairs: null pointer or empty batch");
        if (rows < 1 || rows > 256)
            fail(BMM_ERR_INVALID, "mlp_score_pairs: rows must be in [1, 256]");
        for (size_t i = 0; i < n; ++i)
            if ((size_t)(pairs[i] & 0xff) >= rows || (size_t)(pairs[i] >> 8) >= rows)
                fail(BMM_ERR_RANGE, "mlp_score_pairs: pair row out of range");
        mlp_score_pairs_host(static_cast<Mlp*>(mlp), sketch, rows, pairs, n, y);
    });
}

int bmm_hip_mlp_gradients(void* mlp, const float* x, const float* target, size_t batch,
                          float* grads, float* loss) {
    return guarded([&] {
        if (!mlp || !x || !target || !loss || batch == 0)
            fail(BMM_ERR_INVALID, "mlp_gradients: null pointer or empty batch");
        *loss = mlp_grads_host(static_cast<Mlp*>(mlp), x, target, batch, grads);
    });
}

int bmm_hip_mlp_train_step(void* mlp, const float* x, const float* target, size_t batch,
                           float learning_rate, float* loss) {
    return guarded([&] {
        if (!mlp || !x || !target || !loss || batch == 0)
            fail(BMM_ERR_INVALID, "mlp_train_step: null pointer or empty batch");
        *loss = mlp_train_host(static_cast<Mlp*>(mlp), x, target, batch, learning_rate);
    });
}

} // extern "C"
