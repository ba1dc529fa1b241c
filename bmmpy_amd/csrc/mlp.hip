// Move-proposal MLP: 256 -> 256 -> 256 -> 1, ReLU, bf16 weights and
// activations for the matrix products, fp32 accumulation (MFMA), fp32 master
// weights and Adam moments.  No reference counterpart: it only biases which
// row pairs the GA proposes; every proposed move is rescored exactly.
//
// Products run on v_mfma_f32_32x32x16_bf16: one wavefront computes one 32x32
// output tile, K in steps of 16.  Operand layout of that instruction: lane l
// supplies A[l % 32][8 (l / 32) + 0..7] and B[8 (l / 32) + 0..7][l % 32]; the
// 16 accumulators of lane l are D[8 (j / 4) + 4 (l / 32) + j % 4][l % 32].
// Operands are read through explicit (row, col) strides, so transposed
// products (weight gradients, backward activations) need no copies.
#include "common.hpp"

#include <algorithm>
#include <cmath>
#include <memory>
#include <vector>

namespace bmmhip {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kHid = 256;
constexpr int kIn = 256;
constexpr size_t kNumParams = (size_t)kIn * kHid + kHid + (size_t)kHid * kHid + kHid + kHid + 1;

__device__ inline __bf16 to_bf16(float x) { return (__bf16)x; }

// C[m][n] = sum_k A(m, k) B(k, n) (+ bias[n]); optional ReLU.  Writes fp32
// (pre-activation) to Cf and/or bf16 (post-activation) to Cb.  M, N multiples
// of 32, K a multiple of 16.  AK / BK: A / B contiguous along k (stride 1,
// 16-byte aligned rows), read as one 8-element vector per lane and step.
// With gridDim.y > 1 the K range is cut into gridDim.y slices of k_chunk:
// slice y writes its raw partial tile to part[y][M][N] and
// gemm_reduce_kernel adds the slices in order (no bias, no ReLU, fp32 only).
template <bool AK, bool BK>
__global__ __launch_bounds__(64) void gemm_bf16_kernel(
    const __bf16* __restrict__ A, long sam, long sak, const __bf16* __restrict__ B, long sbk,
    long sbn, int M, int N, int K, int k_chunk, const float* __restrict__ bias, int relu,
    float* __restrict__ Cf, long ldcf, __bf16* __restrict__ Cb, long ldcb,
    float* __restrict__ part) {
    const int tiles_n = N / 32;
    const int tm = blockIdx.x / tiles_n, tn = blockIdx.x % tiles_n;
    const int lane = threadIdx.x;
    const int r = lane % 32, half = lane / 32;
    const long am = (long)(tm * 32 + r);
    const long bn = (long)(tn * 32 + r);
    const int kb = blockIdx.y * k_chunk;
    const int ke = min(K, kb + k_chunk);
    f32x16 acc = {};
    for (int k0 = kb; k0 < ke; k0 += 16) {
        const long kh = k0 + 8 * half;
        bf16x8 va, vb;
        if constexpr (AK) {
            va = *reinterpret_cast<const bf16x8*>(A + am * sam + kh);
        } else {
#pragma unroll
            for (int i = 0; i < 8; ++i)
                va[i] = A[am * sam + (kh + i) * sak];
        }
        if constexpr (BK) {
            vb = *reinterpret_cast<const bf16x8*>(B + kh + bn * sbn);
        } else {
#pragma unroll
            for (int i = 0; i < 8; ++i)
                vb[i] = B[(kh + i) * sbk + bn * sbn];
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(va, vb, acc, 0, 0, 0);
    }
    const int col = tn * 32 + r;
    if (gridDim.y > 1) {
        float* P = part + (size_t)blockIdx.y * M * N;
#pragma unroll
        for (int j = 0; j < 16; ++j)
            P[(long)(tm * 32 + 8 * (j / 4) + 4 * half + (j % 4)) * N + col] = acc[j];
        return;
    }
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        const int row = tm * 32 + 8 * (j / 4) + 4 * half + (j % 4);
        float v = acc[j] + (bias ? bias[col] : 0.f);
        if (Cf)
            Cf[(long)row * ldcf + col] = v;
        if (relu)
            v = v > 0.f ? v : 0.f;
        if (Cb)
            Cb[(long)row * ldcb + col] = to_bf16(v);
    }
}

// Cf[m][n] = sum over s of part[s][m][n], s in increasing order.
__global__ void gemm_reduce_kernel(const float* __restrict__ part, int slices, int M, int N,
                                   float* __restrict__ Cf, long ldcf) {
    const size_t MN = (size_t)M * N;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < MN;
         e += (size_t)gridDim.x * blockDim.x) {
        float v = 0.f;
        for (int s = 0; s < slices; ++s)
            v += part[s * MN + e];
        Cf[(e / N) * ldcf + e % N] = v;
    }
}

// Launches the product.  A product with fewer than 256 output tiles, a long
// K (>= 1024) and only an fp32 output (the weight gradients, K = the batch)
// is cut along K into up to 16 slices of at least 256 so that the launch has
// enough wavefronts; `part` must then hold slices * M * N floats (pass null
// to keep one slice).  The slice sums are added in a fixed order, so results
// are deterministic for a given K.
void gemm_bf16(const __bf16* A, long sam, long sak, const __bf16* B, long sbk, long sbn, int M,
               int N, int K, const float* bias, bool relu, float* Cf, long ldcf, __bf16* Cb,
               long ldcb, hipStream_t s, DevBuf* part = nullptr) {
    if (M % 32 || N % 32 || K % 16)
        fail(BMM_ERR_INVALID, "gemm_bf16: shapes must be multiples of 32 x 32 x 16");
    const int tiles = (M / 32) * (N / 32);
    int slices = 1;
    if (part && tiles < 256 && K >= 1024 && !bias && !relu && !Cb && Cf)
        slices = std::min(16, K / 256);
    const int chunk = (K / 16 + slices - 1) / slices * 16;
    float* P = slices > 1 ? part->get<float>((size_t)slices * M * N) : nullptr;
    const auto al16 = [](const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; };
    const bool ak = sak == 1 && sam % 8 == 0 && al16(A);
    const bool bk = sbk == 1 && sbn % 8 == 0 && al16(B);
    const dim3 grid((unsigned)tiles, (unsigned)slices);
#define BMM_GEMM_LAUNCH(AKV, BKV)                                                              \
    hipLaunchKernelGGL((gemm_bf16_kernel<AKV, BKV>), grid, dim3(64), 0, s, A, sam, sak, B, sbk,  \
                       sbn, M, N, K, chunk, bias, relu ? 1 : 0, Cf, ldcf, Cb, ldcb, P)
    if (ak && bk)
        BMM_GEMM_LAUNCH(true, true);
    else if (ak)
        BMM_GEMM_LAUNCH(true, false);
    else if (bk)
        BMM_GEMM_LAUNCH(false, true);
    else
        BMM_GEMM_LAUNCH(false, false);
#undef BMM_GEMM_LAUNCH
    BMM_HIP(hipGetLastError());
    if (slices > 1) {
        hipLaunchKernelGGL(gemm_reduce_kernel, dim3(256), dim3(256), 0, s, P, slices, M, N, Cf,
                           ldcf);
        BMM_HIP(hipGetLastError());
    }
}

__global__ void f32_to_bf16_kernel(const float* __restrict__ x, __bf16* __restrict__ y, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        y[i] = to_bf16(x[i]);
}

// y[n][k] = bf16(x[k][n]) for a 256 x 256 fp32 matrix.
__global__ void transpose_bf16_kernel(const float* __restrict__ x, __bf16* __restrict__ y) {
    const int n = blockIdx.x, k = threadIdx.x;
    y[(size_t)n * kHid + k] = to_bf16(x[(size_t)k * kHid + n]);
}

// Fused scoring of (src -> dst) row pairs of the GA: the pair's features
// (sketch(src) | sketch(dst), one bit per feature, as +-1 bf16) are built in
// LDS; both hidden layers run on v_mfma_f32_32x32x16_bf16 with the 64 x 256
// activations kept in LDS; weights come from the transposed bf16 copies
// (lane l's fragment of W^T row n is 16 contiguous bytes); then the head.
// One workgroup (4 wavefronts) per 64 pairs; wavefront v computes output
// columns [64 v, 64 v + 64) of every layer as 2 x 2 tiles of 32 x 32.
constexpr int kSB = 64;            // pairs per workgroup
constexpr int kLdA = kHid + 8;     // LDS row stride (bf16): rows 16 bytes apart in bank space

typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));

__device__ inline void score_layer(const __bf16* __restrict__ in, const __bf16* __restrict__ Wt,
                                   const float* __restrict__ bias, __bf16* __restrict__ out) {
    const int lane = threadIdx.x % kWave, wv = threadIdx.x / kWave;
    const int r = lane % 32, half = lane / 32;
    f32x16 acc[2][2] = {};
    for (int k0 = 0; k0 < kHid; k0 += 16) {
        bf16x8 a[2], b[2];
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            a[t] = *reinterpret_cast<const bf16x8*>(&in[(32 * t + r) * kLdA + k0 + 8 * half]);
            b[t] = *reinterpret_cast<const bf16x8*>(
                &Wt[(size_t)(64 * wv + 32 * t + r) * kHid + k0 + 8 * half]);
        }
#pragma unroll
        for (int tm = 0; tm < 2; ++tm)
#pragma unroll
            for (int tn = 0; tn < 2; ++tn)
                acc[tm][tn] =
                    __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[tm], b[tn], acc[tm][tn], 0, 0, 0);
    }
#pragma unroll
    for (int tn = 0; tn < 2; ++tn) {
        const int col = 64 * wv + 32 * tn + r;
        const float bc = bias[col];
#pragma unroll
        for (int tm = 0; tm < 2; ++tm)
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                const int row = 32 * tm + 8 * (j / 4) + 4 * half + (j % 4);
                const float v = acc[tm][tn][j] + bc;
                out[row * kLdA + col] = to_bf16(v > 0.f ? v : 0.f);
            }
    }
}

__global__ __launch_bounds__(256) void mlp_score_kernel(
    const uint64_t* __restrict__ sketch, int rows, const uint16_t* __restrict__ pairs,
    const int32_t* __restrict__ pair_isl, int n, const __bf16* __restrict__ W1t,
    const float* __restrict__ b1, const __bf16* __restrict__ W2t, const float* __restrict__ b2,
    const __bf16* __restrict__ w3, const float* __restrict__ b3, float* __restrict__ y) {
    __shared__ __align__(16) __bf16 X[kSB * kLdA];
    __shared__ __align__(16) __bf16 H[kSB * kLdA];
    const int p0 = blockIdx.x * kSB;
    {
        // Thread t: pair t / 4, features [64 q, 64 q + 64) for q = t % 4
        // (src sketch words 0, 1, then dst sketch words 0, 1).
        const int bi = threadIdx.x / 4, q = threadIdx.x % 4;
        const int i = p0 + bi;
        uint64_t w = 0;
        if (i < n) {
            const int isl = pair_isl[i];
            const int row = q < 2 ? (pairs[i] & 0xff) : (pairs[i] >> 8);
            w = sketch[((size_t)isl * rows + row) * 2 + (q & 1)];
        }
        u16x8* dst = reinterpret_cast<u16x8*>(&X[bi * kLdA + 64 * q]);
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            u16x8 v;
#pragma unroll
            for (int f = 0; f < 8; ++f)   // bf16 +1 = 0x3f80, -1 = 0xbf80
                v[f] = (unsigned short)(((w >> (8 * c + f)) & 1u) ? 0x3f80u : 0xbf80u);
            dst[c] = v;
        }
    }
    __syncthreads();
    score_layer(X, W1t, b1, H);
    __syncthreads();
    score_layer(H, W2t, b2, X);
    __syncthreads();
    {
        const int bi = threadIdx.x / 4, q = threadIdx.x % 4;
        float acc = 0.f;
        for (int h = 64 * q; h < 64 * q + 64; ++h)
            acc += (float)X[bi * kLdA + h] * (float)w3[h];
        acc += __shfl_xor(acc, 1, kWave);
        acc += __shfl_xor(acc, 2, kWave);
        if (q == 0 && p0 + bi < n)
            y[p0 + bi] = acc + b3[0];
    }
}

// Output head (N = 1): y[b] = sum_h A2[b][h] w3[h] + b3, and the loss terms.
__global__ void head_kernel(const __bf16* __restrict__ A2, const __bf16* __restrict__ w3,
                            const float* __restrict__ b3, int B, int valid,
                            const float* __restrict__ target, float* __restrict__ y,
                            float* __restrict__ dy, float* __restrict__ loss) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B)
        return;
    float acc = 0.f;
    for (int h = 0; h < kHid; ++h)
        acc += (float)A2[(long)b * kHid + h] * (float)w3[h];
    acc += b3[0];
    y[b] = acc;
    if (target && dy) {
        const float e = b < valid ? acc - target[b] : 0.f;
        dy[b] = 2.f * e / (float)valid;
        if (loss && b < valid)
            atomicAdd(loss, e * e / (float)valid);
    }
}

// dZ2[b][h] = dy[b] w3[h] * (Z2[b][h] > 0); also dW3, db3 partials.
__global__ void head_backward_kernel(const float* __restrict__ dy, const __bf16* __restrict__ w3,
                                     const float* __restrict__ Z2, const __bf16* __restrict__ A2,
                                     int B, __bf16* __restrict__ dZ2b, float* __restrict__ dZ2f,
                                     float* __restrict__ g_w3, float* __restrict__ g_b3) {
    const int h = blockIdx.x;   // one block per hidden unit, threads over the batch
    float gw = 0.f, gb = 0.f;
    const float w = (float)w3[h];
    for (int b = threadIdx.x; b < B; b += blockDim.x) {
        const float d = dy[b];
        const float z = Z2[(long)b * kHid + h];
        const float g = z > 0.f ? d * w : 0.f;
        dZ2b[(long)b * kHid + h] = to_bf16(g);
        dZ2f[(long)b * kHid + h] = g;
        gw += (float)A2[(long)b * kHid + h] * d;
        gb += d;
    }
    __shared__ float r1[256], r2[256];
    r1[threadIdx.x] = gw;
    r2[threadIdx.x] = gb;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            r1[threadIdx.x] += r1[threadIdx.x + s];
            r2[threadIdx.x] += r2[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        g_w3[h] = r1[0];
        if (h == 0)
            g_b3[0] = r2[0];
    }
}

// Column sums of an fp32 [B x 256] matrix (bias gradients).
__global__ void colsum_kernel(const float* __restrict__ X, int B, float* __restrict__ out) {
    const int h = blockIdx.x;
    float s = 0.f;
    for (int b = threadIdx.x; b < B; b += blockDim.x)
        s += X[(long)b * kHid + h];
    __shared__ float r[256];
    r[threadIdx.x] = s;
    __syncthreads();
    for (int k = blockDim.x / 2; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k)
            r[threadIdx.x] += r[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        out[h] = r[0];
}

// dZ1 = dA1 * (Z1 > 0), in place on the fp32 buffer, plus its bf16 copy.
__global__ void relu_mask_kernel(float* __restrict__ d, const float* __restrict__ Z,
                                 __bf16* __restrict__ db, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const float g = Z[i] > 0.f ? d[i] : 0.f;
        d[i] = g;
        db[i] = to_bf16(g);
    }
}

__global__ void adam_kernel(float* __restrict__ w, const float* __restrict__ g,
                            float* __restrict__ m, float* __restrict__ v, size_t n, float lr,
                            float b1, float b2, float eps, float c1, float c2,
                            __bf16* __restrict__ wb) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
        w[i] = wi;
        if (wb)
            wb[i] = to_bf16(wi);
    }
}

// ------------------------------------------------------------- model -----
struct Mlp {
    // fp32 master parameters (flat: W1, b1, W2, b2, w3, b3) + Adam moments.
    DevBuf p, g, m, v;
    DevBuf wb;    // bf16 copies of W1, W2, w3 (W1 | W2 | w3)
    DevBuf wbt;   // bf16 transposes W1^T | W2^T ([out][in]) for the scoring kernel
    DevBuf score; // outputs of mlp_score_pairs_dev
    DevBuf act;   // activations / scratch for the current batch
    DevBuf part;  // K-slice partial tiles of the weight-gradient products
    int cap_batch = 0;
    long step = 0;
    float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
};

namespace {
struct Offs {
    size_t W1 = 0, b1 = (size_t)kIn * kHid, W2 = b1 + kHid, b2 = W2 + (size_t)kHid * kHid,
           w3 = b2 + kHid, b3 = w3 + kHid;
};
constexpr Offs kOff{};

struct Acts {
    __bf16 *X, *A1, *A2, *dZ2b, *dZ1b;
    float *Z1, *Z2, *y, *dy, *dZ2f, *dZ1f, *loss;
};

Acts acts_of(Mlp& M, int B) {
    const size_t bh = (size_t)B * kHid;
    if (B > M.cap_batch) {
        M.act.get<unsigned char>(bh * (2 * 5) + bh * 4 * 6 + (size_t)B * 8 + 64);
        M.cap_batch = B;
    }
    unsigned char* base = static_cast<unsigned char*>(M.act.ptr);
    Acts a;
    a.X = reinterpret_cast<__bf16*>(base);
    a.A1 = a.X + bh;
    a.A2 = a.A1 + bh;
    a.dZ2b = a.A2 + bh;
    a.dZ1b = a.dZ2b + bh;
    float* f = reinterpret_cast<float*>(a.dZ1b + bh);
    a.Z1 = f;
    a.Z2 = a.Z1 + bh;
    a.dZ2f = a.Z2 + bh;
    a.dZ1f = a.dZ2f + bh;
    a.y = a.dZ1f + bh;
    a.dy = a.y + B;
    a.loss = a.dy + B;
    return a;
}

void sync_transposes(Mlp& M, hipStream_t s) {
    const float* p = static_cast<const float*>(M.p.ptr);
    __bf16* t = static_cast<__bf16*>(M.wbt.ptr);
    hipLaunchKernelGGL(transpose_bf16_kernel, dim3(kHid), dim3(kHid), 0, s, p + kOff.W1, t);
    hipLaunchKernelGGL(transpose_bf16_kernel, dim3(kHid), dim3(kHid), 0, s, p + kOff.W2,
                       t + (size_t)kIn * kHid);
    BMM_HIP(hipGetLastError());
}

void sync_bf16(Mlp& M, hipStream_t s) {
    float* p = static_cast<float*>(M.p.ptr);
    __bf16* wb = static_cast<__bf16*>(M.wb.ptr);
    hipLaunchKernelGGL(f32_to_bf16_kernel, dim3(256), dim3(256), 0, s, p + kOff.W1, wb,
                       (size_t)kIn * kHid);
    hipLaunchKernelGGL(f32_to_bf16_kernel, dim3(256), dim3(256), 0, s, p + kOff.W2,
                       wb + (size_t)kIn * kHid, (size_t)kHid * kHid);
    hipLaunchKernelGGL(f32_to_bf16_kernel, dim3(1), dim3(256), 0, s, p + kOff.w3,
                       wb + (size_t)kIn * kHid + (size_t)kHid * kHid, (size_t)kHid);
    sync_transposes(M, s);
}
} // namespace

Mlp* mlp_create(uint64_t seed) {
    auto M = std::make_unique<Mlp>();
    M->p.get<float>(kNumParams);
    M->g.get<float>(kNumParams);
    M->m.get<float>(kNumParams);
    M->v.get<float>(kNumParams);
    M->wb.get<__bf16>((size_t)kIn * kHid + (size_t)kHid * kHid + kHid);
    M->wbt.get<__bf16>((size_t)kIn * kHid + (size_t)kHid * kHid);
    // He-uniform init from XorShift64.
    std::vector<float> h(kNumParams, 0.f);
    uint64_t s = seed == 0 ? 0x9E3779B97F4A7C15ull : seed;
    auto unit = [&]() {
        s ^= s << 13;
        s ^= s >> 7;
        s ^= s << 17;
        return (double)(s >> 11) * (1.0 / 9007199254740992.0);
    };
    const float a1 = std::sqrt(6.f / kIn), a2 = std::sqrt(6.f / kHid);
    for (size_t i = 0; i < (size_t)kIn * kHid; ++i)
        h[kOff.W1 + i] = (float)((2 * unit() - 1) * a1);
    for (size_t i = 0; i < (size_t)kHid * kHid; ++i)
        h[kOff.W2 + i] = (float)((2 * unit() - 1) * a2);
    for (size_t i = 0; i < (size_t)kHid; ++i)
        h[kOff.w3 + i] = (float)((2 * unit() - 1) * a2);
    BMM_HIP(hipMemcpy(M->p.ptr, h.data(), kNumParams * 4, hipMemcpyHostToDevice));
    std::vector<float> z(kNumParams, 0.f);
    BMM_HIP(hipMemcpy(M->m.ptr, z.data(), kNumParams * 4, hipMemcpyHostToDevice));
    BMM_HIP(hipMemcpy(M->v.ptr, z.data(), kNumParams * 4, hipMemcpyHostToDevice));
    sync_bf16(*M, nullptr);
    BMM_HIP(hipDeviceSynchronize());
    return M.release();
}

void mlp_destroy(Mlp* M) { delete M; }

void mlp_set_params(Mlp* M, const float* host) {
    BMM_HIP(hipMemcpy(M->p.ptr, host, kNumParams * 4, hipMemcpyHostToDevice));
    sync_bf16(*M, nullptr);
    BMM_HIP(hipDeviceSynchronize());
}

void mlp_get_params(Mlp* M, float* host) {
    BMM_HIP(hipMemcpy(host, M->p.ptr, kNumParams * 4, hipMemcpyDeviceToHost));
}

// Forward on bf16 features d_x [B x 256] already on the device (B multiple of
// 32).  Writes y (fp32, B) into the activation buffer; returns it.
float* mlp_forward_dev(Mlp* M, const __bf16* d_x, int B, hipStream_t s) {
    Acts a = acts_of(*M, B);
    const float* p = static_cast<const float*>(M->p.ptr);
    const __bf16* wb = static_cast<const __bf16*>(M->wb.ptr);
    const __bf16* W1 = wb;
    const __bf16* W2 = wb + (size_t)kIn * kHid;
    const __bf16* w3 = W2 + (size_t)kHid * kHid;
    gemm_bf16(d_x, kIn, 1, W1, kHid, 1, B, kHid, kIn, p + kOff.b1, true, a.Z1, kHid, a.A1, kHid,
              s);
    gemm_bf16(a.A1, kHid, 1, W2, kHid, 1, B, kHid, kHid, p + kOff.b2, true, a.Z2, kHid, a.A2,
              kHid, s);
    hipLaunchKernelGGL(head_kernel, dim3((B + 255) / 256), dim3(256), 0, s, a.A2, w3,
                       p + kOff.b3, B, B, nullptr, a.y, nullptr, nullptr);
    BMM_HIP(hipGetLastError());
    return a.y;
}

// Scores n (src -> dst) pairs from per-row sketches (2 u64 words per row,
// row-major per island); returns y on the device (valid until the next call).
float* mlp_score_pairs_dev(Mlp* M, const uint64_t* d_sketch, int rows, const uint16_t* d_pairs,
                           const int32_t* d_pair_isl, int n, hipStream_t s) {
    float* y = M->score.get<float>((size_t)std::max(n, 1));
    if (n == 0)
        return y;
    const float* p = static_cast<const float*>(M->p.ptr);
    const __bf16* wb = static_cast<const __bf16*>(M->wb.ptr);
    const __bf16* t = static_cast<const __bf16*>(M->wbt.ptr);
    hipLaunchKernelGGL(mlp_score_kernel, dim3((unsigned)((n + kSB - 1) / kSB)), dim3(256), 0, s,
                       d_sketch, rows, d_pairs, d_pair_isl, n, t, p + kOff.b1,
                       t + (size_t)kIn * kHid, p + kOff.b2,
                       wb + (size_t)kIn * kHid + (size_t)kHid * kHid, p + kOff.b3, y);
    BMM_HIP(hipGetLastError());
    return y;
}

// Forward + backward for MSE against d_t (first `valid` rows count); leaves
// gradients in M->g and the loss in the activation buffer.
float* mlp_backward_dev(Mlp* M, const __bf16* d_x, const float* d_t, int B, int valid,
                        hipStream_t s) {
    Acts a = acts_of(*M, B);
    const float* p = static_cast<const float*>(M->p.ptr);
    float* g = static_cast<float*>(M->g.ptr);
    const __bf16* wb = static_cast<const __bf16*>(M->wb.ptr);
    const __bf16* W1 = wb;
    const __bf16* W2 = wb + (size_t)kIn * kHid;
    const __bf16* w3 = W2 + (size_t)kHid * kHid;
    gemm_bf16(d_x, kIn, 1, W1, kHid, 1, B, kHid, kIn, p + kOff.b1, true, a.Z1, kHid, a.A1, kHid,
              s);
    gemm_bf16(a.A1, kHid, 1, W2, kHid, 1, B, kHid, kHid, p + kOff.b2, true, a.Z2, kHid, a.A2,
              kHid, s);
    fill_u32(reinterpret_cast<uint32_t*>(a.loss), 1, 0u, s);
    hipLaunchKernelGGL(head_kernel, dim3((B + 255) / 256), dim3(256), 0, s, a.A2, w3,
                       p + kOff.b3, B, valid, d_t, a.y, a.dy, a.loss);
    hipLaunchKernelGGL(head_backward_kernel, dim3(kHid), dim3(256), 0, s, a.dy, w3, a.Z2, a.A2,
                       B, a.dZ2b, a.dZ2f, g + kOff.w3, g + kOff.b3);
    BMM_HIP(hipGetLastError());
    // dW2 = A1^T dZ2 ; db2 = colsum(dZ2)
    gemm_bf16(a.A1, 1, kHid, a.dZ2b, kHid, 1, kHid, kHid, B, nullptr, false, g + kOff.W2, kHid,
              nullptr, 0, s, &M->part);
    hipLaunchKernelGGL(colsum_kernel, dim3(kHid), dim3(256), 0, s, a.dZ2f, B, g + kOff.b2);
    // dA1 = dZ2 W2^T, masked by Z1 > 0
    gemm_bf16(a.dZ2b, kHid, 1, W2, 1, kHid, B, kHid, kHid, nullptr, false, a.dZ1f, kHid,
              nullptr, 0, s);
    hipLaunchKernelGGL(relu_mask_kernel, dim3(1024), dim3(256), 0, s, a.dZ1f, a.Z1, a.dZ1b,
                       (size_t)B * kHid);
    // dW1 = X^T dZ1 ; db1 = colsum(dZ1)
    gemm_bf16(d_x, 1, kIn, a.dZ1b, kHid, 1, kIn, kHid, B, nullptr, false, g + kOff.W1, kHid,
              nullptr, 0, s, &M->part);
    hipLaunchKernelGGL(colsum_kernel, dim3(kHid), dim3(256), 0, s, a.dZ1f, B, g + kOff.b1);
    BMM_HIP(hipGetLastError());
    return a.loss;
}

void mlp_adam_dev(Mlp* M, float lr, hipStream_t s) {
    ++M->step;
    const float c1 = 1.f - std::pow(M->beta1, (float)M->step);
    const float c2 = 1.f - std::pow(M->beta2, (float)M->step);
    hipLaunchKernelGGL(adam_kernel, dim3(512), dim3(256), 0, s, static_cast<float*>(M->p.ptr),
                       static_cast<const float*>(M->g.ptr), static_cast<float*>(M->m.ptr),
                       static_cast<float*>(M->v.ptr), kNumParams, lr, M->beta1, M->beta2, M->eps,
                       c1, c2, static_cast<__bf16*>(nullptr));
    BMM_HIP(hipGetLastError());
    sync_bf16(*M, s);
}

namespace {
// Host features (float, B x 256) -> padded bf16 device buffer.
struct MlpIo {
    DevBuf xf, xb, t, sk, pr, is;
};
MlpIo& mlp_io() {
    thread_local MlpIo io;
    return io;
}
int padded(int B) { return (B + 31) / 32 * 32; }

__bf16* upload_features(const float* x, int B, int Bp) {
    MlpIo& io = mlp_io();
    float* xf = io.xf.get<float>((size_t)Bp * kIn);
    __bf16* xb = io.xb.get<__bf16>((size_t)Bp * kIn);
    BMM_HIP(hipMemset(xf, 0, (size_t)Bp * kIn * 4));
    BMM_HIP(hipMemcpy(xf, x, (size_t)B * kIn * 4, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(f32_to_bf16_kernel, dim3(256), dim3(256), 0, nullptr, xf, xb,
                       (size_t)Bp * kIn);
    BMM_HIP(hipGetLastError());
    return xb;
}
} // namespace

void mlp_forward_host(Mlp* M, const float* x, size_t B, float* y) {
    const int Bp = padded((int)B);
    __bf16* xb = upload_features(x, (int)B, Bp);
    float* d_y = mlp_forward_dev(M, xb, Bp, nullptr);
    BMM_HIP(hipMemcpy(y, d_y, B * 4, hipMemcpyDeviceToHost));
}

void mlp_score_pairs_host(Mlp* M, const uint64_t* sketch, size_t rows, const uint16_t* pairs,
                          size_t n, float* y) {
    MlpIo& io = mlp_io();
    uint64_t* d_sk = io.sk.get<uint64_t>(rows * 2);
    uint16_t* d_pr = io.pr.get<uint16_t>(n);
    int32_t* d_is = io.is.get<int32_t>(n);
    BMM_HIP(hipMemcpy(d_sk, sketch, rows * 16, hipMemcpyHostToDevice));
    BMM_HIP(hipMemcpy(d_pr, pairs, n * 2, hipMemcpyHostToDevice));
    BMM_HIP(hipMemset(d_is, 0, n * 4));
    float* d_y = mlp_score_pairs_dev(M, d_sk, (int)rows, d_pr, d_is, (int)n, nullptr);
    BMM_HIP(hipMemcpy(y, d_y, n * 4, hipMemcpyDeviceToHost));
}

float mlp_grads_host(Mlp* M, const float* x, const float* t, size_t B, float* grads) {
    const int Bp = padded((int)B);
    __bf16* xb = upload_features(x, (int)B, Bp);
    float* d_t = mlp_io().t.get<float>((size_t)Bp);
    BMM_HIP(hipMemset(d_t, 0, (size_t)Bp * 4));
    BMM_HIP(hipMemcpy(d_t, t, B * 4, hipMemcpyHostToDevice));
    float* d_loss = mlp_backward_dev(M, xb, d_t, Bp, (int)B, nullptr);
    float loss = 0.f;
    BMM_HIP(hipMemcpy(&loss, d_loss, 4, hipMemcpyDeviceToHost));
    if (grads)
        BMM_HIP(hipMemcpy(grads, M->g.ptr, kNumParams * 4, hipMemcpyDeviceToHost));
    return loss;
}

float mlp_train_host(Mlp* M, const float* x, const float* t, size_t B, float lr) {
    const float loss = mlp_grads_host(M, x, t, B, nullptr);
    mlp_adam_dev(M, lr, nullptr);
    BMM_HIP(hipDeviceSynchronize());
    return loss;
}

size_t mlp_num_params() { return kNumParams; }

} // namespace bmmhip
