// corr_ext: spatial correlation sampler (FlowNet-style local correlation),
// csrc/corr_ext/correlation_sampler.cpp:38-85 and correlation_cuda_kernel.cu.
//
//   out[n][ph][pw][h][w] = sum_c sum_{i<kH, j<kW} in1[n][c][y1][x1] * in2[n][c][y1 + sh][x1 + sw]
//   y1 = h*dH - padH + i*dilH,  x1 = w*dW - padW + j*dilW,
//   sh = ph*dpH - dpH*(patchH-1)/2,  sw = pw*dpW - dpW*(patchW-1)/2,
// a term counting only when both positions lie inside the H x W maps.  Inputs
// are [N,C,H,W]; the output is [N,patchH,patchW,oH,oW] with
// oH = (H + 2 padH - ((kH-1) dilH + 1)) / dH + 1 (same for W).
//
// gfx950 layout: one lane per output element, consecutive lanes on
// consecutive output columns, so every channel step reads two coalesced rows;
// the channel sum accumulates in FP32 (FP64 for f64 inputs).  The gradients are
// gathers: one lane per input element sums the output gradients that read it,
// so there are no atomics and the result is deterministic.  For odd patch
// sizes (the usual case) the backward uses the forward's patch centre; the
// reference's backward centres even patches at ((patch-1)/2)*dilation_patch,
// and so does this one (correlation_cuda_kernel.cu:94-95, 165-166).

#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "vh_common.hpp"

namespace vh {
namespace {

template <typename T> struct AccOf { using type = float; };
template <> struct AccOf<double> { using type = double; };

template <typename T> __device__ __forceinline__ float to_f(T v) { return (float)v; }
template <> __device__ __forceinline__ float to_f<__half>(__half v) { return __half2float(v); }
template <> __device__ __forceinline__ float to_f<__hip_bfloat16>(__hip_bfloat16 v) { return __bfloat162float(v); }
template <typename A, typename T> __device__ __forceinline__ A load_acc(const T* p) { return (A)to_f<T>(*p); }
template <> __device__ __forceinline__ double load_acc<double, double>(const double* p) { return *p; }

template <typename T> __device__ __forceinline__ T from_acc(float v) { return (T)v; }
template <> __device__ __forceinline__ __half from_acc<__half>(float v) { return __float2half(v); }
template <> __device__ __forceinline__ __hip_bfloat16 from_acc<__hip_bfloat16>(float v) { return __float2bfloat16(v); }
template <typename T> __device__ __forceinline__ T store_acc(typename AccOf<T>::type v) { return from_acc<T>(v); }
template <> __device__ __forceinline__ double store_acc<double>(double v) { return v; }

struct SCArgs {
    int N, C, H, W, oH, oW;
    int kH, kW, patchH, patchW, padH, padW, dilH, dilW, dpH, dpW, dH, dW;
};

// grid (ceil(oW/64) * oH, patchH * patchW, N), 64 lanes along w
template <typename T>
__global__ __launch_bounds__(64) void spatial_corr_fwd_kernel(const T* __restrict__ in1, const T* __restrict__ in2,
                                                              T* __restrict__ out, SCArgs a) {
    using A = typename AccOf<T>::type;
    const int ntw = (a.oW + 63) / 64;
    const int h = blockIdx.x / ntw;
    const int w = (blockIdx.x % ntw) * 64 + threadIdx.x;
    const int p = blockIdx.y, ph = p / a.patchW, pw = p % a.patchW;
    const int n = blockIdx.z;
    if (w >= a.oW) return;
    const int sh = ph * a.dpH - a.dpH * (a.patchH - 1) / 2;
    const int sw = pw * a.dpW - a.dpW * (a.patchW - 1) / 2;
    const int64_t HW = (int64_t)a.H * a.W;
    const T* b1 = in1 + (int64_t)n * a.C * HW;
    const T* b2 = in2 + (int64_t)n * a.C * HW;
    A acc = 0;
    for (int i = 0; i < a.kH; i++) {
        const int y1 = h * a.dH - a.padH + i * a.dilH, y2 = y1 + sh;
        if (y1 < 0 || y1 >= a.H || y2 < 0 || y2 >= a.H) continue;
        for (int j = 0; j < a.kW; j++) {
            const int x1 = w * a.dW - a.padW + j * a.dilW, x2 = x1 + sw;
            if (x1 < 0 || x1 >= a.W || x2 < 0 || x2 >= a.W) continue;
            const T* p1 = b1 + (int64_t)y1 * a.W + x1;
            const T* p2 = b2 + (int64_t)y2 * a.W + x2;
            A s[4] = {0, 0, 0, 0};
            int c = 0;
            for (; c + 4 <= a.C; c += 4) {
#pragma unroll
                for (int u = 0; u < 4; u++) s[u] += load_acc<A>(p1 + (c + u) * HW) * load_acc<A>(p2 + (c + u) * HW);
            }
            for (; c < a.C; c++) s[0] += load_acc<A>(p1 + c * HW) * load_acc<A>(p2 + c * HW);
            acc += (s[0] + s[1]) + (s[2] + s[3]);
        }
    }
    out[(((int64_t)n * a.patchH + ph) * a.patchW + pw) * a.oH * a.oW + (int64_t)h * a.oW + w] = store_acc<T>(acc);
}

// Output rows / columns whose kernel window reads input row y at tap i:
// h with h*dH = y + padH - i*dilH, 0 <= h < oH.
__device__ __forceinline__ bool out_of(int y, int pad, int i, int dil, int stride, int on, int& h) {
    const int t = y + pad - i * dil;
    if (t < 0 || t % stride) return false;
    h = t / stride;
    return h < on;
}

// grad1[n][c][y][x] = sum over (ph, pw, i, j) of gout[n][ph][pw][h][w] * in2[n][c][y+sh][x+sw],
// (h, w) the output read at tap (i, j); grid (ceil(W/64) * H, C, N)
// grad2[n][c][y][x] = sum over (ph, pw, i, j) of gout[...][h][w] * in1[n][c][y-sh][x-sw]
template <typename T, bool SECOND>
__global__ __launch_bounds__(64) void spatial_corr_bwd_kernel(const T* __restrict__ gout, const T* __restrict__ other,
                                                              T* __restrict__ grad, SCArgs a) {
    using A = typename AccOf<T>::type;
    const int ntw = (a.W + 63) / 64;
    const int y = blockIdx.x / ntw;
    const int x = (blockIdx.x % ntw) * 64 + threadIdx.x;
    const int c = blockIdx.y, n = blockIdx.z;
    if (x >= a.W) return;
    const int64_t HW = (int64_t)a.H * a.W;
    const T* ob = other + ((int64_t)n * a.C + c) * HW;
    const int64_t oHW = (int64_t)a.oH * a.oW;
    const int radH = (a.patchH - 1) / 2, radW = (a.patchW - 1) / 2;
    A acc = 0;
    for (int ph = 0; ph < a.patchH; ph++) {
        const int sh = a.dpH * (ph - radH);
        const int yo = SECOND ? y - sh : y + sh;  // the other input's row
        const int y1 = SECOND ? yo : y;            // the row of input 1 that the output taps
        if (yo < 0 || yo >= a.H) continue;
        for (int pw = 0; pw < a.patchW; pw++) {
            const int sw = a.dpW * (pw - radW);
            const int xo = SECOND ? x - sw : x + sw;
            const int x1 = SECOND ? xo : x;
            if (xo < 0 || xo >= a.W) continue;
            const A v = load_acc<A>(ob + (int64_t)yo * a.W + xo);
            const T* g = gout + (((int64_t)n * a.patchH + ph) * a.patchW + pw) * oHW;
            A s = 0;
            for (int i = 0; i < a.kH; i++) {
                int h;
                if (!out_of(y1, a.padH, i, a.dilH, a.dH, a.oH, h)) continue;
                for (int j = 0; j < a.kW; j++) {
                    int w;
                    if (!out_of(x1, a.padW, j, a.dilW, a.dW, a.oW, w)) continue;
                    s += load_acc<A>(g + (int64_t)h * a.oW + w);
                }
            }
            acc += s * v;
        }
    }
    grad[((int64_t)n * a.C + c) * HW + (int64_t)y * a.W + x] = store_acc<T>(acc);
}

int check_args(const SCArgs& a, const char* what) {
    if (a.N < 0 || a.C <= 0 || a.H <= 0 || a.W <= 0 || a.kH <= 0 || a.kW <= 0 || a.patchH <= 0 || a.patchW <= 0 ||
        a.padH < 0 || a.padW < 0 || a.dilH <= 0 || a.dilW <= 0 || a.dpH <= 0 || a.dpW <= 0 || a.dH <= 0 ||
        a.dW <= 0)
        return fail(VH_ERR_ARG, "%s: bad sizes or parameters", what);
    if (a.oH <= 0 || a.oW <= 0) return fail(VH_ERR_ARG, "%s: empty output (%d x %d)", what, a.oH, a.oW);
    if ((int64_t)a.patchH * a.patchW > 65535 || a.C > 65535 || a.N > 65535)
        return fail(VH_ERR_UNSUPPORTED, "%s: grid too large", what);
    return VH_OK;
}

template <typename T>
int fwd_launch(const void* in1, const void* in2, void* out, const SCArgs& a, hipStream_t s) {
    const dim3 grid((unsigned)((a.oW + 63) / 64 * a.oH), (unsigned)(a.patchH * a.patchW), (unsigned)a.N);
    hipLaunchKernelGGL(spatial_corr_fwd_kernel<T>, grid, dim3(64), 0, s, (const T*)in1, (const T*)in2, (T*)out, a);
    VH_LAUNCH_CHECK("spatial_corr_forward");
    return VH_OK;
}

template <typename T>
int bwd_launch(const void* in1, const void* in2, const void* gout, void* g1, void* g2, const SCArgs& a,
               hipStream_t s) {
    const dim3 grid((unsigned)((a.W + 63) / 64 * a.H), (unsigned)a.C, (unsigned)a.N);
    hipLaunchKernelGGL((spatial_corr_bwd_kernel<T, false>), grid, dim3(64), 0, s, (const T*)gout, (const T*)in2,
                       (T*)g1, a);
    VH_LAUNCH_CHECK("spatial_corr_backward input1");
    hipLaunchKernelGGL((spatial_corr_bwd_kernel<T, true>), grid, dim3(64), 0, s, (const T*)gout, (const T*)in1,
                       (T*)g2, a);
    VH_LAUNCH_CHECK("spatial_corr_backward input2");
    return VH_OK;
}

SCArgs make_args(int N, int C, int H, int W, int kH, int kW, int patchH, int patchW, int padH, int padW, int dilH,
                 int dilW, int dpH, int dpW, int dH, int dW) {
    SCArgs a{N, C, H, W, 0, 0, kH, kW, patchH, patchW, padH, padW, dilH, dilW, dpH, dpW, dH, dW};
    if (dH > 0 && dW > 0) {
        a.oH = (H + 2 * padH - ((kH - 1) * dilH + 1)) / dH + 1;
        a.oW = (W + 2 * padW - ((kW - 1) * dilW + 1)) / dW + 1;
    }
    return a;
}

}  // namespace
}  // namespace vh

using namespace vh;

extern "C" {

int vh_spatial_corr_out_size(int H, int W, int kH, int kW, int padH, int padW, int dilH, int dilW, int dH, int dW,
                             int* oH, int* oW) {
    if (!oH || !oW || dH <= 0 || dW <= 0) return fail(VH_ERR_ARG, "spatial_corr_out_size: bad arguments");
    const SCArgs a = make_args(1, 1, H, W, kH, kW, 1, 1, padH, padW, dilH, dilW, 1, 1, dH, dW);
    *oH = a.oH;
    *oW = a.oW;
    return VH_OK;
}

int vh_spatial_corr_forward(int dtype, const void* in1, const void* in2, int N, int C, int H, int W, int kH, int kW,
                            int patchH, int patchW, int padH, int padW, int dilH, int dilW, int dpH, int dpW, int dH,
                            int dW, void* out, void* stream) {
    const SCArgs a = make_args(N, C, H, W, kH, kW, patchH, patchW, padH, padW, dilH, dilW, dpH, dpW, dH, dW);
    if (int rc = check_args(a, "spatial_corr_forward")) return rc;
    if (N == 0) return VH_OK;
    if (!in1 || !in2 || !out) return fail(VH_ERR_ARG, "spatial_corr_forward: null pointer");
    hipStream_t s = (hipStream_t)stream;
    switch (dtype) {
        case VH_F32: return fwd_launch<float>(in1, in2, out, a, s);
        case VH_F64: return fwd_launch<double>(in1, in2, out, a, s);
        case VH_F16: return fwd_launch<__half>(in1, in2, out, a, s);
        case VH_BF16: return fwd_launch<__hip_bfloat16>(in1, in2, out, a, s);
        default: return fail(VH_ERR_UNSUPPORTED, "spatial_corr_forward: dtype %d", dtype);
    }
}

int vh_spatial_corr_backward(int dtype, const void* in1, const void* in2, const void* grad_out, int N, int C, int H,
                             int W, int kH, int kW, int patchH, int patchW, int padH, int padW, int dilH, int dilW,
                             int dpH, int dpW, int dH, int dW, void* grad1, void* grad2, void* stream) {
    const SCArgs a = make_args(N, C, H, W, kH, kW, patchH, patchW, padH, padW, dilH, dilW, dpH, dpW, dH, dW);
    if (int rc = check_args(a, "spatial_corr_backward")) return rc;
    if (N == 0) return VH_OK;
    if (!in1 || !in2 || !grad_out || !grad1 || !grad2) return fail(VH_ERR_ARG, "spatial_corr_backward: null pointer");
    hipStream_t s = (hipStream_t)stream;
    switch (dtype) {
        case VH_F32: return bwd_launch<float>(in1, in2, grad_out, grad1, grad2, a, s);
        case VH_F64: return bwd_launch<double>(in1, in2, grad_out, grad1, grad2, a, s);
        case VH_F16: return bwd_launch<__half>(in1, in2, grad_out, grad1, grad2, a, s);
        case VH_BF16: return bwd_launch<__hip_bfloat16>(in1, in2, grad_out, grad1, grad2, a, s);
        default: return fail(VH_ERR_UNSUPPORTED, "spatial_corr_backward: dtype %d", dtype);
    }
}

}  // extern "C"
