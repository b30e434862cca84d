// SPDX-License-Identifier: GPL-3.0-or-later
// Fused Adam step over the Gaussian parameter tensors.  The reference leaves
// the optimiser to its caller (python_bindings/vol_marcher.py returns
// gradients only); a 3DGS training step follows the backward with Adam over
// means, log-scales, quaternions, opacity logits and SH (59 floats per
// Gaussian).  One launch per tensor, each element read once (parameter,
// gradient, two moments: 16 B) and written once (parameter, moments: 12 B),
// four elements per thread when the tensor is 16-byte aligned.
#include <cmath>

#include "host_common.h"

namespace dgmr {
namespace {

struct AdamScalars {
    float lr, beta1, beta2, eps, weight_decay;
    float omb1, omb2;        // 1 - beta1, 1 - beta2, formed in double from the host values
    float bias1, bias2_sqrt; // 1 - beta1^t, sqrt(1 - beta2^t)
};

__device__ __forceinline__ void adam_one(float& p, float g, float& m, float& v, const AdamScalars& a)
{
    if (a.weight_decay != 0.f)
        g = fmaf(a.weight_decay, p, g);
    m = fmaf(a.beta1, m, a.omb1 * g);
    v = fmaf(a.beta2, v, a.omb2 * g * g);
    const float denom = sqrtf(v) / a.bias2_sqrt + a.eps;
    p -= (a.lr / a.bias1) * (m / denom);
}

__global__ __launch_bounds__(256) void k_adam4(float4* __restrict__ p, const float4* __restrict__ g, float4* __restrict__ m,
    float4* __restrict__ v, int64_t n4, AdamScalars a)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n4)
        return;
    float4 P = p[i], M = m[i], V = v[i];
    const float4 G = g[i];
    adam_one(P.x, G.x, M.x, V.x, a);
    adam_one(P.y, G.y, M.y, V.y, a);
    adam_one(P.z, G.z, M.z, V.z, a);
    adam_one(P.w, G.w, M.w, V.w, a);
    p[i] = P;
    m[i] = M;
    v[i] = V;
}

__global__ __launch_bounds__(256) void k_adam1(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
    int64_t n, AdamScalars a)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    float P = p[i], M = m[i], V = v[i];
    adam_one(P, g[i], M, V, a);
    p[i] = P;
    m[i] = M;
    v[i] = V;
}

bool aligned16(const void* q) { return (reinterpret_cast<uintptr_t>(q) & 15u) == 0; }

} // namespace
} // namespace dgmr

using namespace dgmr;

extern "C" {

int dgmr_adam_step(const dgmr_adam_tensor* tensors, int32_t n_tensors, double beta1, double beta2, double eps, double weight_decay,
    int64_t step, dgmr_stream stream)
{
    DGMR_ARG_CHECK(tensors && n_tensors >= 0, "dgmr_adam_step: null tensor table");
    DGMR_ARG_CHECK(step >= 1, "dgmr_adam_step: step counts from 1");
    DGMR_ARG_CHECK(beta1 >= 0 && beta1 < 1 && beta2 >= 0 && beta2 < 1 && eps > 0, "dgmr_adam_step: bad hyper-parameters");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    // scalars formed in double, then rounded once (as torch forms 1 - beta from Python floats)
    AdamScalars a { 0.f, float(beta1), float(beta2), float(eps), float(weight_decay), float(1.0 - beta1), float(1.0 - beta2),
        float(1.0 - std::pow(beta1, double(step))), float(std::sqrt(1.0 - std::pow(beta2, double(step)))) };
    for (int t = 0; t < n_tensors; ++t) {
        const dgmr_adam_tensor& x = tensors[t];
        DGMR_ARG_CHECK(x.numel >= 0, "dgmr_adam_step: negative size");
        if (x.numel == 0)
            continue;
        DGMR_ARG_CHECK(x.param && x.grad && x.exp_avg && x.exp_avg_sq, "dgmr_adam_step: null tensor");
        a.lr = x.lr;
        StageScope sc("optim.adam", st, 28.0 * double(x.numel));
        if (x.numel % 4 == 0 && aligned16(x.param) && aligned16(x.grad) && aligned16(x.exp_avg) && aligned16(x.exp_avg_sq)) {
            const int64_t n4 = x.numel / 4;
            k_adam4<<<unsigned(ceil_div64(n4, 256)), 256, 0, st>>>(reinterpret_cast<float4*>(x.param), reinterpret_cast<const float4*>(x.grad),
                reinterpret_cast<float4*>(x.exp_avg), reinterpret_cast<float4*>(x.exp_avg_sq), n4, a);
        } else {
            k_adam1<<<unsigned(ceil_div64(x.numel, 256)), 256, 0, st>>>(x.param, x.grad, x.exp_avg, x.exp_avg_sq, x.numel, a);
        }
        DGMR_HIP_CHECK(hipGetLastError());
    }
    return DGMR_OK;
}

} // extern "C"
