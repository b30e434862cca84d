// SPDX-License-Identifier: GPL-3.0-or-later
// Fused L1 photometric loss for the training step: the reference leaves the
// loss to its caller (python_bindings/vol_marcher.py returns the image and
// takes dL/dimage); a 3DGS step computes L = mean |image - target| and its
// gradient sign(image - target) / n.  The forward reads both images once and
// leaves one partial sum per workgroup (summed in fixed order by the caller,
// so the loss is deterministic); the backward reads them once more and
// writes the gradient scaled by dL/dloss, which it reads from device memory.
#include "device_math.h"
#include "host_common.h"

namespace dgmr {
namespace {

constexpr int kLossThreads = 256;
constexpr int kLossBlocks = 1024;

__global__ __launch_bounds__(kLossThreads) void k_l1_forward(const float* __restrict__ a, const float* __restrict__ b, int64_t n, float scale,
    float* __restrict__ partials)
{
    __shared__ float s_wave[kLossThreads / 64];
    float acc = 0.f;
    const int64_t stride = int64_t(gridDim.x) * kLossThreads;
    const int64_t n4 = n / 4;
    const float4* a4 = reinterpret_cast<const float4*>(a);
    const float4* b4 = reinterpret_cast<const float4*>(b);
    for (int64_t i = int64_t(blockIdx.x) * kLossThreads + threadIdx.x; i < n4; i += stride) {
        const float4 x = a4[i], y = b4[i];
        acc += (fabsf(x.x - y.x) + fabsf(x.y - y.y)) + (fabsf(x.z - y.z) + fabsf(x.w - y.w));
    }
    for (int64_t i = 4 * n4 + int64_t(blockIdx.x) * kLossThreads + threadIdx.x; i < n; i += stride)
        acc += fabsf(a[i] - b[i]);
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1)
        acc += __shfl_xor(acc, d, 64);
    if ((threadIdx.x & 63) == 0)
        s_wave[threadIdx.x / 64] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        float t = 0.f;
#pragma unroll
        for (int w = 0; w < kLossThreads / 64; ++w)
            t += s_wave[w];
        partials[blockIdx.x] = t * scale;
    }
}


__global__ __launch_bounds__(kLossThreads) void k_l1_backward(const float* __restrict__ a, const float* __restrict__ b, int64_t n, float scale,
    const float* __restrict__ grad_loss, float* __restrict__ grad)
{
    const float s = scale * grad_loss[0];
    const int64_t stride = int64_t(gridDim.x) * kLossThreads;
    const int64_t n4 = n / 4;
    const float4* a4 = reinterpret_cast<const float4*>(a);
    const float4* b4 = reinterpret_cast<const float4*>(b);
    float4* g4 = reinterpret_cast<float4*>(grad);
    for (int64_t i = int64_t(blockIdx.x) * kLossThreads + threadIdx.x; i < n4; i += stride) {
        const float4 x = a4[i], y = b4[i];
        g4[i] = make_float4(l1_sign(x.x - y.x) * s, l1_sign(x.y - y.y) * s, l1_sign(x.z - y.z) * s, l1_sign(x.w - y.w) * s);
    }
    for (int64_t i = 4 * n4 + int64_t(blockIdx.x) * kLossThreads + threadIdx.x; i < n; i += stride)
        grad[i] = l1_sign(a[i] - b[i]) * s;
}

// Many images in one launch (the views of a batch): each workgroup takes
// kManyChunk consecutive floats of one image, found by a search over the
// items' first workgroups.  The descriptors travel as a kernel argument.
constexpr int kManyItems = 32;
constexpr int kManyVec = 4; // float4 per thread
constexpr int kManyChunk = kLossThreads * kManyVec * 4;

struct L1Many {
    const float* image[kManyItems];
    const float* target[kManyItems];
    float* grad[kManyItems];
    const float* grad_loss[kManyItems];
    int64_t n[kManyItems];
    float scale[kManyItems];
    uint32_t first_block[kManyItems + 1];
    int count;
};

__global__ __launch_bounds__(kLossThreads) void k_l1_backward_many(const L1Many m)
{
    const uint32_t blk = blockIdx.x;
    int lo = 0, hi = m.count - 1; // the last item whose first workgroup is <= blk (items without workgroups are passed over)
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (m.first_block[mid] <= blk)
            lo = mid;
        else
            hi = mid - 1;
    }
    const int64_t n = m.n[lo];
    const float s = m.scale[lo] * m.grad_loss[lo][0];
    const float* __restrict__ a = m.image[lo];
    const float* __restrict__ b = m.target[lo];
    float* __restrict__ g = m.grad[lo];
    const int64_t base = int64_t(blk - m.first_block[lo]) * kManyChunk;
#pragma unroll
    for (int r = 0; r < kManyVec; ++r) {
        const int64_t i = base + (int64_t(r) * kLossThreads + threadIdx.x) * 4;
        if (i + 4 <= n) {
            const float4 x = *reinterpret_cast<const float4*>(a + i), y = *reinterpret_cast<const float4*>(b + i);
            *reinterpret_cast<float4*>(g + i) =
                make_float4(l1_sign(x.x - y.x) * s, l1_sign(x.y - y.y) * s, l1_sign(x.z - y.z) * s, l1_sign(x.w - y.w) * s);
        } else {
            for (int64_t j = i; j < n; ++j)
                g[j] = l1_sign(a[j] - b[j]) * s;
        }
    }
}

bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

unsigned loss_blocks(int64_t n) { return unsigned(std::min<int64_t>(kLossBlocks, std::max<int64_t>(1, ceil_div64(n, 4 * kLossThreads)))); }

} // namespace
} // namespace dgmr

using namespace dgmr;

extern "C" {

int dgmr_l1_partials_size(int64_t n, int32_t* count)
{
    DGMR_ARG_CHECK(count && n >= 0, "dgmr_l1_partials_size: bad argument");
    *count = int32_t(loss_blocks(n));
    return DGMR_OK;
}

int dgmr_l1_loss_forward(const float* image, const float* target, int64_t n, float scale, float* partials, dgmr_stream stream)
{
    DGMR_ARG_CHECK(image && target && partials && n >= 0, "dgmr_l1_loss_forward: bad argument");
    DGMR_ARG_CHECK(aligned16(image) && aligned16(target), "dgmr_l1_loss_forward: images must be 16-byte aligned");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    StageScope sc("loss.l1_forward", st, 8.0 * double(n));
    k_l1_forward<<<loss_blocks(n), kLossThreads, 0, st>>>(image, target, n, scale, partials);
    DGMR_HIP_CHECK(hipGetLastError());
    return DGMR_OK;
}

int dgmr_l1_loss_backward(const float* image, const float* target, int64_t n, float scale, const float* grad_loss, float* grad,
    dgmr_stream stream)
{
    DGMR_ARG_CHECK(image && target && grad_loss && grad && n >= 0, "dgmr_l1_loss_backward: bad argument");
    DGMR_ARG_CHECK(aligned16(image) && aligned16(target) && aligned16(grad), "dgmr_l1_loss_backward: images must be 16-byte aligned");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    StageScope sc("loss.l1_backward", st, 12.0 * double(n));
    k_l1_backward<<<loss_blocks(n), kLossThreads, 0, st>>>(image, target, n, scale, grad_loss, grad);
    DGMR_HIP_CHECK(hipGetLastError());
    return DGMR_OK;
}

int dgmr_l1_loss_backward_many(const dgmr_l1_item* items, int32_t n_items, dgmr_stream stream)
{
    DGMR_ARG_CHECK(n_items >= 0 && (items || n_items == 0), "dgmr_l1_loss_backward_many: bad argument");
    double elems = 0;
    for (int32_t k = 0; k < n_items; ++k) {
        const dgmr_l1_item& it = items[k];
        DGMR_ARG_CHECK(it.image && it.target && it.grad && it.grad_loss && it.n >= 0, "dgmr_l1_loss_backward_many: bad item");
        DGMR_ARG_CHECK(aligned16(it.image) && aligned16(it.target) && aligned16(it.grad),
            "dgmr_l1_loss_backward_many: images must be 16-byte aligned");
        DGMR_ARG_CHECK(it.n <= int64_t(kManyChunk) * 0x7fffffff, "dgmr_l1_loss_backward_many: image too large");
        elems += double(it.n);
    }
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    StageScope sc("loss.l1_backward", st, 12.0 * elems);
    for (int32_t k0 = 0; k0 < n_items; k0 += kManyItems) {
        L1Many m {};
        m.count = std::min<int32_t>(kManyItems, n_items - k0);
        int64_t blocks = 0;
        for (int j = 0; j < m.count; ++j) {
            const dgmr_l1_item& it = items[k0 + j];
            m.image[j] = it.image;
            m.target[j] = it.target;
            m.grad[j] = it.grad;
            m.grad_loss[j] = it.grad_loss;
            m.n[j] = it.n;
            m.scale[j] = it.scale;
            m.first_block[j] = uint32_t(blocks);
            blocks += ceil_div64(it.n, kManyChunk);
        }
        DGMR_ARG_CHECK(blocks < 0x7fffffff, "dgmr_l1_loss_backward_many: too many elements in one call");
        m.first_block[m.count] = uint32_t(blocks);
        if (blocks == 0)
            continue;
        k_l1_backward_many<<<unsigned(blocks), kLossThreads, 0, st>>>(m);
        DGMR_HIP_CHECK(hipGetLastError());
    }
    return DGMR_OK;
}

} // extern "C"
