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
// Kernel 1 / stride 1 (the local-attention form) runs on the matrix cores (see
// spatial_corr_mfma_kernel) when the caller passes the workspace for the
// channels-last copies.  Otherwise, one lane per output element, consecutive
// lanes on consecutive output columns, so every channel step reads two coalesced rows;
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


// ---- matrix-core path: kernel 1, stride 1 (the local-attention form) ----------
// For one output row h and shift row ph the outputs over (w, pw) are the band
// x' = x1(w) + sw(pw) of the product S = A^T B, A = in1 row y1 (C x W),
// B = in2 row y1 + sh (C x W).  A wave computes S for 16 positions w and
// 16 * NT columns x' (NT = ceil((16 + (patchW-1) dpW) / 16)) on the matrix cores,
// and writes the band's diagonals to out[.., ph, pw, h, w].  The inputs are first
// copied channels-last with the channels padded to 32, so one 16-byte load per
// lane feeds four v_mfma_f32_16x16x4f32 or one v_mfma_f32_16x16x32_{f16,bf16}.
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

// channel padding of the channels-last copies: a multiple of 128 bytes, so that
// the mfma kernel's count of 64-byte steps is even
constexpr int kClPadBytes = 128;
constexpr int kMaxNT = 4;
constexpr uint32_t kScRange = 0x80000000u;    // num_records of the operand resources (bytes)
constexpr uint32_t kScInvalid = 0x80000000u;  // offset of an off-map pixel: past kScRange

__device__ __forceinline__ __amdgpu_buffer_rsrc_t sc_rsrc(const void* base) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)kScRange, 0x00020000);
}

struct alignas(16) Q16 {
    uint32_t w[4];
};

__device__ __forceinline__ Q16 sc_load(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    return __builtin_bit_cast(Q16, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
}

template <typename T> struct Mfma16;
template <> struct Mfma16<float> {
    __device__ static f32x4 run(const Q16& a, const Q16& b, f32x4 acc) {
#pragma unroll
        for (int q = 0; q < 4; q++)
            acc = __builtin_amdgcn_mfma_f32_16x16x4f32(__builtin_bit_cast(float, a.w[q]),
                                                       __builtin_bit_cast(float, b.w[q]), acc, 0, 0, 0);
        return acc;
    }
};
template <> struct Mfma16<__half> {
    __device__ static f32x4 run(const Q16& a, const Q16& b, f32x4 acc) {
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(f16x8, a), __builtin_bit_cast(f16x8, b), acc,
                                                      0, 0, 0);
    }
};
template <> struct Mfma16<__hip_bfloat16> {
    __device__ static f32x4 run(const Q16& a, const Q16& b, f32x4 acc) {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b),
                                                       acc, 0, 0, 0);
    }
};

// [N,C,H,W] -> [N,H*W,Cp] with zeros for c >= C, converted to TO; 32 channels
// x 64 pixels per block (Cp a multiple of 32).
template <typename T, typename TO = T>
__global__ __launch_bounds__(256) void spatial_corr_to_cl_kernel(const T* __restrict__ in, TO* __restrict__ out, int C,
                                                                 int HW, int Cp) {
    __shared__ TO tile[32][65];
    const int p0 = blockIdx.x * 64, c0 = blockIdx.y * 32, n = blockIdx.z;
    const int tid = threadIdx.x;
    const T* src = in + (int64_t)n * C * HW;
    for (int i = tid; i < 32 * 64; i += 256) {
        const int c = i / 64, p = i % 64;
        const int cc = c0 + c, pp = p0 + p;
        tile[c][p] = from_acc<TO>((cc < C && pp < HW) ? to_f<T>(src[(int64_t)cc * HW + pp]) : 0.f);
    }
    __syncthreads();
    TO* dst = out + (int64_t)n * HW * Cp;
    for (int i = tid; i < 32 * 64; i += 256) {
        const int p = i / 32, c = i % 32;
        const int pp = p0 + p;
        if (pp < HW) dst[(int64_t)pp * Cp + c0 + c] = tile[c][p];
    }
}

// grid: ceil(oW/64) * oH * patchH * N workgroups in one dimension, 4 waves;
// logical block (n, h, ph, wx), wx fastest; wave v covers w0 = 64 wx + 16 v.
template <typename T, int NT>
__global__ __launch_bounds__(256) void spatial_corr_mfma_kernel(const T* __restrict__ a_cl, const T* __restrict__ b_cl,
                                                                T* __restrict__ out, SCArgs a, int Cp) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int nwx = (a.oW + 63) / 64;
    const int64_t nb = (int64_t)nwx * a.oH * a.patchH * a.N;
    int64_t lb = xcd_logical(blockIdx.x, nb);
    const int wx = (int)(lb % nwx);
    lb /= nwx;
    const int ph = (int)(lb % a.patchH);
    lb /= a.patchH;
    const int h = (int)(lb % a.oH);
    const int n = (int)(lb / a.oH);
    const int w0 = wx * 64 + wave * 16;
    if (w0 >= a.oW) return;   // no barriers below
    const int radH = a.dpH * (a.patchH - 1) / 2, radW = a.dpW * (a.patchW - 1) / 2;
    const int y1 = h - a.padH, y2 = y1 + ph * a.dpH - radH;
    const int g = lane >> 4, c16 = lane & 15;
    f32x4 acc[NT];
#pragma unroll
    for (int t = 0; t < NT; t++) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
    if (y1 >= 0 && y1 < a.H && y2 >= 0 && y2 < a.H) {
        // Operands come through buffer resources based at the channels-last
        // copies: a lane whose pixel lies off the map gets an offset past the
        // resource's range and reads zeros without a memory access.  At each
        // step the four lane groups take the four 16-byte chunks of the next 64
        // bytes of a pixel's channels (the same channel-to-k map for both
        // operands), so the four groups of a pixel read 64 contiguous bytes.
        const int x1 = w0 + c16 - a.padW;                 // A row i = c16
        const int xb = w0 - a.padW - radW;                // x' of tile 0, column 0
        const __amdgpu_buffer_rsrc_t ra = sc_rsrc(a_cl), rb = sc_rsrc(b_cl);
        const uint32_t lane_off = (uint32_t)g * 16u;
        const uint32_t aoff = (x1 >= 0 && x1 < a.W)
                                  ? (uint32_t)((((int64_t)n * a.H + y1) * a.W + x1) * Cp * (int64_t)sizeof(T)) + lane_off
                                  : kScInvalid;
        uint32_t boff[NT];
#pragma unroll
        for (int t = 0; t < NT; t++) {
            const int x2 = xb + 16 * t + c16;
            boff[t] = (x2 >= 0 && x2 < a.W)
                          ? (uint32_t)((((int64_t)n * a.H + y2) * a.W + x2) * Cp * (int64_t)sizeof(T)) + lane_off
                          : kScInvalid;
        }
        const int steps = Cp * (int)sizeof(T) / 64;   // 64 bytes of channels per step and pixel; even
        // two register sets in turn: the loads of one step are in flight while
        // the other step's products run
        Q16 a0 = sc_load(ra, aoff), b0[NT], a1, b1[NT];
#pragma unroll
        for (int t = 0; t < NT; t++) b0[t] = sc_load(rb, boff[t]);
        for (int s = 0; s < steps; s += 2) {
            const uint32_t o1 = 64u * (uint32_t)(s + 1);
            a1 = sc_load(ra, aoff + o1);
#pragma unroll
            for (int t = 0; t < NT; t++) b1[t] = sc_load(rb, boff[t] + o1);
#pragma unroll
            for (int t = 0; t < NT; t++) acc[t] = Mfma16<T>::run(a0, b0[t], acc[t]);
            // the step after next, or this step again at the end (in range, unused)
            const uint32_t o2 = 64u * (uint32_t)(s + 2 < steps ? s + 2 : s);
            a0 = sc_load(ra, aoff + o2);
#pragma unroll
            for (int t = 0; t < NT; t++) b0[t] = sc_load(rb, boff[t] + o2);
#pragma unroll
            for (int t = 0; t < NT; t++) acc[t] = Mfma16<T>::run(a1, b1[t], acc[t]);
        }
    }
    // D[row][col] of tile t sits in lane col + 16 (row / 4), register row % 4.
    const int64_t plane = (int64_t)a.oH * a.oW;
    T* ob = out + (((int64_t)n * a.patchH + ph) * a.patchW) * plane + (int64_t)h * a.oW;
#pragma unroll
    for (int t = 0; t < NT; t++) {
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int row = 4 * g + r;
            const int diff = 16 * t + c16 - row;   // = pw * dpW
            const int w = w0 + row;
            if (diff < 0 || diff % a.dpW || w >= a.oW) continue;
            const int pw = diff / a.dpW;
            if (pw < a.patchW) ob[(int64_t)pw * plane + w] = from_acc<T>(acc[t][r]);
        }
    }
}

int cl_channels(int C, size_t elem) {
    const int pad = kClPadBytes / (int)elem;
    return (C + pad - 1) / pad * pad;
}

bool mfma_path(int dtype, const SCArgs& a) {
    if (dtype != VH_F32 && dtype != VH_F16 && dtype != VH_BF16) return false;
    if (a.kH != 1 || a.kW != 1 || a.dH != 1 || a.dW != 1) return false;
    const int nt = (16 + (a.patchW - 1) * a.dpW + 15) / 16;
    const size_t elem = dtype == VH_F32 ? 4 : 2;
    const int64_t copy_bytes = (int64_t)a.N * a.H * a.W * cl_channels(a.C, elem) * (int64_t)elem;
    return nt <= kMaxNT && (int64_t)((a.oW + 63) / 64) * a.oH * a.patchH * a.N < ((int64_t)1 << 31) &&
           copy_bytes + 64 < (int64_t)kScRange;
}

size_t mfma_workspace(int dtype, const SCArgs& a) {
    if (!mfma_path(dtype, a)) return 0;
    const size_t elem = dtype == VH_F32 ? 4 : 2;
    const size_t one = (size_t)a.N * a.H * a.W * cl_channels(a.C, elem) * elem;
    return 2 * ((one + 255) / 256 * 256);
}

template <typename T>
int fwd_mfma_launch(const void* in1, const void* in2, void* out, const SCArgs& a, void* ws, hipStream_t s) {
    const int Cp = cl_channels(a.C, sizeof(T));
    const int HW = a.H * a.W;
    const size_t one = ((size_t)a.N * HW * Cp * sizeof(T) + 255) / 256 * 256;
    T* a_cl = (T*)ws;
    T* b_cl = (T*)((char*)ws + one);
    const dim3 tg((unsigned)((HW + 63) / 64), (unsigned)(Cp / 32), (unsigned)a.N);
    hipLaunchKernelGGL(spatial_corr_to_cl_kernel<T>, tg, dim3(256), 0, s, (const T*)in1, a_cl, a.C, HW, Cp);
    VH_LAUNCH_CHECK("spatial_corr channels-last copy");
    hipLaunchKernelGGL(spatial_corr_to_cl_kernel<T>, tg, dim3(256), 0, s, (const T*)in2, b_cl, a.C, HW, Cp);
    VH_LAUNCH_CHECK("spatial_corr channels-last copy");
    const dim3 grid((unsigned)((int64_t)((a.oW + 63) / 64) * a.oH * a.patchH * a.N));
    const int nt = (16 + (a.patchW - 1) * a.dpW + 15) / 16;
    switch (nt) {
        case 1: hipLaunchKernelGGL((spatial_corr_mfma_kernel<T, 1>), grid, dim3(256), 0, s, a_cl, b_cl, (T*)out, a, Cp); break;
        case 2: hipLaunchKernelGGL((spatial_corr_mfma_kernel<T, 2>), grid, dim3(256), 0, s, a_cl, b_cl, (T*)out, a, Cp); break;
        case 3: hipLaunchKernelGGL((spatial_corr_mfma_kernel<T, 3>), grid, dim3(256), 0, s, a_cl, b_cl, (T*)out, a, Cp); break;
        default: hipLaunchKernelGGL((spatial_corr_mfma_kernel<T, 4>), grid, dim3(256), 0, s, a_cl, b_cl, (T*)out, a, Cp); break;
    }
    VH_LAUNCH_CHECK("spatial_corr_forward (matrix cores)");
    return VH_OK;
}


// ---- matrix-core backward: kernel 1, stride 1 ----------------------------------
// With the backward's patch centre rb = ((patch-1)/2) dp, for one row y and
// shift row ph:
//   grad1[c][y][x]   = sum_x2 G1[x][x2] in2[c][y + sh][x2],
//                      G1[x][x2] = gout[ph][pw][y + padH][x + padW], pw dpW = x2 - x + rbW
//   grad2[c][y2][x2] = sum_x1 G2[x2][x1] in1[c][y2 - sh][x1],
//                      G2[x2][x1] = gout[ph][pw][y2 - sh + padH][x1 + padW], pw dpW = x2 - x1 + rbW
// a banded product over 16 + (patchW-1) dpW columns.  A wave owns 16 output
// positions x and up to kBwdCT tiles of 16 channels; per shift row it gathers its
// band of G once (the A operand, zero off the band) and multiplies it with
// channels-last FP32 copies of the other input (B) on v_mfma_f32_16x16x4f32,
// accumulating over all shift rows in registers.  Each lane then writes four
// consecutive positions of one channel: no atomics, one writer per element.
constexpr int kBwdCT = 4;     // channel tiles of 16 per wave
constexpr int kBwdKS = 16;    // at most 64 band columns (16 steps of 4)

__device__ __forceinline__ float buf_f32(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}
__device__ __forceinline__ uint32_t buf_u16(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    return (uint32_t)__builtin_amdgcn_raw_buffer_load_b16(r, off, 0, 0);
}
template <typename T> __device__ __forceinline__ float load_g(__amdgpu_buffer_rsrc_t r, uint32_t off);
template <> __device__ __forceinline__ float load_g<float>(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    return buf_f32(r, off);
}
template <> __device__ __forceinline__ float load_g<__half>(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    return __half2float(__builtin_bit_cast(__half, (unsigned short)buf_u16(r, off)));
}
template <> __device__ __forceinline__ float load_g<__hip_bfloat16>(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    return __uint_as_float(buf_u16(r, off) << 16);
}

// grid: ceil(ceil(W/16) / 4) * H * N * ncg workgroups in one dimension
// (ncg = ceil(Cp / (16 kBwdCT))), 4 waves; logical block (n, cg, y, bx), bx
// fastest, spread over the XCDs in contiguous ranges (xcd_logical) so that the
// rows neighbouring y share stay in one L2.
template <typename T, bool SECOND>
__global__ __launch_bounds__(256) void spatial_corr_bwd_mfma_kernel(const T* __restrict__ gout,
                                                                    const float* __restrict__ other_cl,
                                                                    T* __restrict__ grad, SCArgs a, int Cp) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int ncg = (Cp + 16 * kBwdCT - 1) / (16 * kBwdCT);
    const int nbx = ((a.W + 15) / 16 + 3) / 4;
    const int64_t nb = (int64_t)nbx * a.H * a.N * ncg;
    int64_t lb = xcd_logical(blockIdx.x, nb);
    const int bx = (int)(lb % nbx);
    lb /= nbx;
    const int y = (int)(lb % a.H);
    lb /= a.H;
    const int cg = (int)(lb % ncg);
    const int n = (int)(lb / ncg);
    const int x0 = (bx * 4 + wave) * 16;
    if (x0 >= a.W) return;   // no barriers below
    const int nct = min(kBwdCT, Cp / 16 - cg * kBwdCT);
    const int i16 = lane & 15, g = lane >> 4;
    const int rbH = (a.patchH - 1) / 2 * a.dpH, rbW = (a.patchW - 1) / 2 * a.dpW;
    const int span = (a.patchW - 1) * a.dpW;
    const int ksteps = (16 + span + 3) / 4;
    const int kb = SECOND ? x0 + rbW - span : x0 - rbW;   // band column of k = 0
    const int64_t plane = (int64_t)a.oH * a.oW;
    const __amdgpu_buffer_rsrc_t rg = sc_rsrc(gout + (int64_t)n * a.patchH * a.patchW * plane);
    const __amdgpu_buffer_rsrc_t ro = sc_rsrc(other_cl + (int64_t)n * a.H * a.W * Cp);
    f32x4 acc[kBwdCT];
#pragma unroll
    for (int t = 0; t < kBwdCT; t++) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int ph = 0; ph < a.patchH; ph++) {
        const int sh = ph * a.dpH - rbH;
        const int yo = SECOND ? y - sh : y + sh;          // row of the other input
        if (yo < 0 || yo >= a.H) continue;
        const int hrow = (SECOND ? yo : y) + a.padH;      // output row of gout
        // A: G[i][k] for k = 4 s + g
        float af[kBwdKS];
#pragma unroll
        for (int st = 0; st < kBwdKS; st++) {
            af[st] = 0.f;
            if (st < ksteps) {
                const int kk = 4 * st + g;
                const int xk = kb + kk;                    // x2 (first) / x1 (second)
                const int xo = x0 + i16;                   // x (first) / x2 (second)
                const int pwd = SECOND ? span + i16 - kk : kk - i16;
                const int gcol = (SECOND ? xk : xo) + a.padW;
                const bool ok = pwd >= 0 && pwd % a.dpW == 0 && pwd / a.dpW < a.patchW && xo < a.W &&
                                (!SECOND || (xk >= 0 && xk < a.W));
                const int64_t off = ((int64_t)(ph * a.patchW + (ok ? pwd / a.dpW : 0)) * plane +
                                     (int64_t)hrow * a.oW + (ok ? gcol : 0)) * (int64_t)sizeof(T);
                af[st] = load_g<T>(rg, ok ? (uint32_t)off : kScInvalid);
            }
        }
        // band step outer, channel tiles inner: kBwdCT independent accumulations
#pragma unroll
        for (int st = 0; st < kBwdKS; st++) {
            if (st >= ksteps) break;
            const int xk = kb + 4 * st + g;
            const bool ok = xk >= 0 && xk < a.W;
            const uint32_t base = (uint32_t)((((int64_t)yo * a.W + (ok ? xk : 0)) * Cp + cg * kBwdCT * 16 + i16) * 4);
            float bv[kBwdCT];
#pragma unroll
            for (int t = 0; t < kBwdCT; t++) bv[t] = buf_f32(ro, ok && t < nct ? base + 64u * t : kScInvalid);
#pragma unroll
            for (int t = 0; t < kBwdCT; t++) acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[st], bv[t], acc[t], 0, 0, 0);
        }
    }
    // D[row = x][col = c]: lane c + 16 (row / 4), register row % 4
    const int64_t HW = (int64_t)a.H * a.W;
#pragma unroll
    for (int t = 0; t < kBwdCT; t++) {
        if (t >= nct) break;
        const int c = (cg * kBwdCT + t) * 16 + i16;
        if (c >= a.C) continue;
        T* gp = grad + ((int64_t)n * a.C + c) * HW + (int64_t)y * a.W;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int x = x0 + 4 * g + r;
            if (x < a.W) gp[x] = from_acc<T>(acc[t][r]);
        }
    }
}

bool bwd_mfma_path(int dtype, const SCArgs& a) {
    if (dtype != VH_F32 && dtype != VH_F16 && dtype != VH_BF16) return false;
    if (a.kH != 1 || a.kW != 1 || a.dH != 1 || a.dW != 1) return false;
    if ((a.patchW - 1) * a.dpW > 4 * kBwdKS - 16) return false;
    const int Cp = cl_channels(a.C, 4);
    const int64_t copy_bytes = (int64_t)a.N * a.H * a.W * Cp * 4;
    const int64_t gout_bytes = (int64_t)a.patchH * a.patchW * a.oH * a.oW * (dtype == VH_F32 ? 4 : 2);
    const int64_t ncg = (Cp + 16 * kBwdCT - 1) / (16 * kBwdCT);
    return copy_bytes / a.N + 64 < (int64_t)kScRange && gout_bytes + 64 < (int64_t)kScRange &&
           (int64_t)(((a.W + 15) / 16 + 3) / 4) * a.H * a.N * ncg < ((int64_t)1 << 31);
}

size_t bwd_mfma_workspace(int dtype, const SCArgs& a) {
    if (!bwd_mfma_path(dtype, a)) return 0;
    const size_t one = (size_t)a.N * a.H * a.W * cl_channels(a.C, 4) * 4;
    return 2 * ((one + 255) / 256 * 256);
}

template <typename T>
int bwd_mfma_launch(const void* in1, const void* in2, const void* gout, void* g1, void* g2, const SCArgs& a, void* ws,
                    hipStream_t s) {
    const int Cp = cl_channels(a.C, 4);
    const int HW = a.H * a.W;
    const size_t one = ((size_t)a.N * HW * Cp * 4 + 255) / 256 * 256;
    float* in1_cl = (float*)ws;
    float* in2_cl = (float*)((char*)ws + one);
    const dim3 tg((unsigned)((HW + 63) / 64), (unsigned)(Cp / 32), (unsigned)a.N);
    hipLaunchKernelGGL((spatial_corr_to_cl_kernel<T, float>), tg, dim3(256), 0, s, (const T*)in1, in1_cl, a.C, HW, Cp);
    VH_LAUNCH_CHECK("spatial_corr channels-last copy");
    hipLaunchKernelGGL((spatial_corr_to_cl_kernel<T, float>), tg, dim3(256), 0, s, (const T*)in2, in2_cl, a.C, HW, Cp);
    VH_LAUNCH_CHECK("spatial_corr channels-last copy");
    const int ncg = (Cp + 16 * kBwdCT - 1) / (16 * kBwdCT);
    const dim3 grid((unsigned)((int64_t)(((a.W + 15) / 16 + 3) / 4) * a.H * a.N * ncg));
    hipLaunchKernelGGL((spatial_corr_bwd_mfma_kernel<T, false>), grid, dim3(256), 0, s, (const T*)gout,
                       (const float*)in2_cl, (T*)g1, a, Cp);
    VH_LAUNCH_CHECK("spatial_corr_backward input1 (matrix cores)");
    hipLaunchKernelGGL((spatial_corr_bwd_mfma_kernel<T, true>), grid, dim3(256), 0, s, (const T*)gout,
                       (const float*)in1_cl, (T*)g2, a, Cp);
    VH_LAUNCH_CHECK("spatial_corr_backward input2 (matrix cores)");
    return VH_OK;
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

size_t vh_spatial_corr_workspace_size(int dtype, int N, int C, int H, int W, int kH, int kW, int patchH, int patchW,
                                      int padH, int padW, int dilH, int dilW, int dpH, int dpW, int dH, int dW) {
    const SCArgs a = make_args(N, C, H, W, kH, kW, patchH, patchW, padH, padW, dilH, dilW, dpH, dpW, dH, dW);
    if (check_args(a, "spatial_corr_workspace_size") != VH_OK) return 0;
    return mfma_workspace(dtype, a);
}

int vh_spatial_corr_forward(int dtype, const void* in1, const void* in2, int N, int C, int H, int W, int kH, int kW,
                            int patchH, int patchW, int padH, int padW, int dilH, int dilW, int dpH, int dpW, int dH,
                            int dW, void* out, void* workspace, size_t workspace_bytes, void* stream) {
    const SCArgs a = make_args(N, C, H, W, kH, kW, patchH, patchW, padH, padW, dilH, dilW, dpH, dpW, dH, dW);
    if (int rc = check_args(a, "spatial_corr_forward")) return rc;
    if (N == 0) return VH_OK;
    if (!in1 || !in2 || !out) return fail(VH_ERR_ARG, "spatial_corr_forward: null pointer");
    hipStream_t s = (hipStream_t)stream;
    const size_t need = mfma_workspace(dtype, a);
    if (need > 0 && workspace && workspace_bytes >= need) {
        switch (dtype) {
            case VH_F32: return fwd_mfma_launch<float>(in1, in2, out, a, workspace, s);
            case VH_F16: return fwd_mfma_launch<__half>(in1, in2, out, a, workspace, s);
            case VH_BF16: return fwd_mfma_launch<__hip_bfloat16>(in1, in2, out, a, workspace, s);
            default: break;
        }
    }
    switch (dtype) {
        case VH_F32: return fwd_launch<float>(in1, in2, out, a, s);
        case VH_F64: return fwd_launch<double>(in1, in2, out, a, s);
        case VH_F16: return fwd_launch<__half>(in1, in2, out, a, s);
        case VH_BF16: return fwd_launch<__hip_bfloat16>(in1, in2, out, a, s);
        default: return fail(VH_ERR_UNSUPPORTED, "spatial_corr_forward: dtype %d", dtype);
    }
}

size_t vh_spatial_corr_backward_workspace_size(int dtype, int N, int C, int H, int W, int kH, int kW, int patchH,
                                               int patchW, int padH, int padW, int dilH, int dilW, int dpH, int dpW,
                                               int dH, int dW) {
    const SCArgs a = make_args(N, C, H, W, kH, kW, patchH, patchW, padH, padW, dilH, dilW, dpH, dpW, dH, dW);
    if (check_args(a, "spatial_corr_backward_workspace_size") != VH_OK) return 0;
    return bwd_mfma_workspace(dtype, a);
}

int vh_spatial_corr_backward(int dtype, const void* in1, const void* in2, const void* grad_out, int N, int C, int H,
                             int W, int kH, int kW, int patchH, int patchW, int padH, int padW, int dilH, int dilW,
                             int dpH, int dpW, int dH, int dW, void* grad1, void* grad2, void* workspace,
                             size_t workspace_bytes, void* stream) {
    const SCArgs a = make_args(N, C, H, W, kH, kW, patchH, patchW, padH, padW, dilH, dilW, dpH, dpW, dH, dW);
    if (int rc = check_args(a, "spatial_corr_backward")) return rc;
    if (N == 0) return VH_OK;
    if (!in1 || !in2 || !grad_out || !grad1 || !grad2) return fail(VH_ERR_ARG, "spatial_corr_backward: null pointer");
    hipStream_t s = (hipStream_t)stream;
    const size_t need = bwd_mfma_workspace(dtype, a);
    if (need > 0 && workspace && workspace_bytes >= need) {
        switch (dtype) {
            case VH_F32: return bwd_mfma_launch<float>(in1, in2, grad_out, grad1, grad2, a, workspace, s);
            case VH_F16: return bwd_mfma_launch<__half>(in1, in2, grad_out, grad1, grad2, a, workspace, s);
            case VH_BF16: return bwd_mfma_launch<__hip_bfloat16>(in1, in2, grad_out, grad1, grad2, a, workspace, s);
            default: break;
        }
    }
    switch (dtype) {
        case VH_F32: return bwd_launch<float>(in1, in2, grad_out, grad1, grad2, a, s);
        case VH_F64: return bwd_launch<double>(in1, in2, grad_out, grad1, grad2, a, s);
        case VH_F16: return bwd_launch<__half>(in1, in2, grad_out, grad1, grad2, a, s);
        case VH_BF16: return bwd_launch<__hip_bfloat16>(in1, in2, grad_out, grad1, grad2, a, s);
        default: return fail(VH_ERR_UNSUPPORTED, "spatial_corr_backward: dtype %d", dtype);
    }
}

}  // extern "C"
