// Correlation volume, pyramid and lookups on gfx950.
//
//   corr_index fwd/bwd   csrc/droid_net_ext/correlation_kernels.cu:22-158
//   volume + pyramid     vipe/slam/networks/droid_net.py:56-69, 94-102
//   altcorr fwd/bwd      csrc/droid_net_ext/altcorr_kernel.cu:26-138, 242-366
//   altcorr indexed      csrc/droid_net_ext/altcorr_kernel.cu:140-240, 394-424
//
// Window convention shared by every lookup: taps (i, j) in [0, 2r+1]^2 sit at
// (floor(x)-r+i, floor(y)-r+j); output channel c = i*(2r+1) + j (x offset
// major) receives the bilinear mix of its four neighbouring taps; taps
// outside the plane read zero.
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include <cstdint>
#include <type_traits>

#include "vh_common.hpp"

namespace vh {
namespace {

template <typename T>
__device__ __forceinline__ float to_f(T v) { return (float)v; }
template <>
__device__ __forceinline__ float to_f<__half>(__half v) { return __half2float(v); }
template <>
__device__ __forceinline__ float to_f<__hip_bfloat16>(__hip_bfloat16 v) { return __bfloat162float(v); }

template <typename T>
__device__ __forceinline__ T from_f(float v) { return (T)v; }
template <>
__device__ __forceinline__ __half from_f<__half>(float v) { return __float2half(v); }
template <>
__device__ __forceinline__ __hip_bfloat16 from_f<__hip_bfloat16>(float v) { return __float2bfloat16(v); }

// Accumulation type: float for 16/32-bit volumes, double for double volumes.
template <typename T> struct Acc { using type = float; };
template <> struct Acc<double> { using type = double; };

template <typename A, typename T>
__device__ __forceinline__ A load_as(const T* p) { return (A)to_f(*p); }
template <>
__device__ __forceinline__ double load_as<double, double>(const double* p) { return *p; }

template <typename T, typename A>
__device__ __forceinline__ T store_from(A v) { return from_f<T>((float)v); }
template <>
__device__ __forceinline__ double store_from<double, double>(double v) { return v; }

// ---------------------------------------------------------------------------
// corr_index: one lane per source pixel (b, y, x); the lane gathers the
// (2r+2)^2 taps of its own plane and writes its (2r+1)^2 outputs, channel
// stride H1*W1 (coalesced across x).
// ---------------------------------------------------------------------------
// Output column j (y offset) needs only tap rows j and j+1, so each lane
// streams its 2r+2 tap rows keeping two of them live and stores the 2r+1
// outputs of a column as soon as both rows are in: ~40 live registers instead
// of (2r+1)^2 accumulators, which keeps occupancy high for the scattered loads.
template <typename T, typename A, int R>
__device__ __forceinline__ void load_tap_row(const T* __restrict__ plane, int y1, int xb, int H2, int W2,
                                             A (&row)[2 * R + 2]) {
    const bool yok = (y1 >= 0) && (y1 < H2);
    const T* rp = plane + (int64_t)(yok ? y1 : 0) * W2;
#pragma unroll
    for (int i = 0; i < 2 * R + 2; i++) {
        const int x1 = xb + i;
        row[i] = (yok && x1 >= 0 && x1 < W2) ? load_as<A>(rp + x1) : A(0);
    }
}

template <typename T, int R>
__device__ __forceinline__ void lookup_level(const T* __restrict__ plane, int H2, int W2, float x0, float y0,
                                             T* __restrict__ o, int64_t ostride) {
    using A = typename Acc<T>::type;
    constexpr int RD = 2 * R + 1;
    const float fx0 = floorf(x0), fy0 = floorf(y0);
    const float dx = x0 - fx0, dy = y0 - fy0;
    const int xb = (int)fx0 - R, yb = (int)fy0 - R;
    // bilinear weights are float products, as the reference forms them (scalar_t(dx * dy))
    const A w00 = (A)((1.f - dx) * (1.f - dy)), w10 = (A)(dx * (1.f - dy)), w01 = (A)((1.f - dx) * dy),
            w11 = (A)(dx * dy);
    A r0[RD + 1], r1[RD + 1];
    load_tap_row<T, A, R>(plane, yb, xb, H2, W2, r0);
#pragma unroll
    for (int j = 0; j < RD; j++) {
        A (&top)[RD + 1] = (j & 1) ? r1 : r0;
        A (&bot)[RD + 1] = (j & 1) ? r0 : r1;
        load_tap_row<T, A, R>(plane, yb + j + 1, xb, H2, W2, bot);
#pragma unroll
        for (int i = 0; i < RD; i++) {
            const A v = top[i] * w00 + top[i + 1] * w10 + bot[i] * w01 + bot[i + 1] * w11;
            o[(int64_t)(i * RD + j) * ostride] = store_from<T>(v);
        }
    }
}

// Vector form of one lookup: each tap row comes from NL aligned loads of VEC
// elements starting at xa = floor(xb / VEC) * VEC; the row is then shifted by
// xb - xa in registers.  Needs W2 % VEC == 0 and a plane base aligned to
// VEC * sizeof(T), so that a vector lies wholly inside or wholly outside the
// row and the zero padding stays exact.  The 2r+2 taps of a row then cost
// NL = ceil((2r+2 + VEC-1) / VEC) loads instead of 2r+2 scalar ones: 3 instead
// of 8 at r = 3 and VEC = 4, 5 at VEC = 2.
template <typename T, int VEC> struct VecOf;
template <> struct VecOf<float, 4> { using type = float4; };
template <> struct VecOf<float, 2> { using type = float2; };
template <> struct VecOf<__half, 4> { using type = uint2; };
template <> struct VecOf<__half, 2> { using type = unsigned int; };

template <typename T, int VEC>
__device__ __forceinline__ void unpack_vec(const typename VecOf<T, VEC>::type& v, float (&e)[VEC]);
template <>
__device__ __forceinline__ void unpack_vec<float, 4>(const float4& v, float (&e)[4]) {
    e[0] = v.x; e[1] = v.y; e[2] = v.z; e[3] = v.w;
}
template <>
__device__ __forceinline__ void unpack_vec<float, 2>(const float2& v, float (&e)[2]) {
    e[0] = v.x; e[1] = v.y;
}
template <>
__device__ __forceinline__ void unpack_vec<__half, 4>(const uint2& v, float (&e)[4]) {
    const __half2 a = *reinterpret_cast<const __half2*>(&v.x), b = *reinterpret_cast<const __half2*>(&v.y);
    const float2 fa = __half22float2(a), fb = __half22float2(b);
    e[0] = fa.x; e[1] = fa.y; e[2] = fb.x; e[3] = fb.y;
}
template <>
__device__ __forceinline__ void unpack_vec<__half, 2>(const unsigned int& v, float (&e)[2]) {
    const float2 f = __half22float2(*reinterpret_cast<const __half2*>(&v));
    e[0] = f.x; e[1] = f.y;
}

template <typename T, int R, int VEC>
__device__ __forceinline__ void load_tap_row_vec(const T* __restrict__ plane, int y1, int xb, int H2, int W2,
                                                 float (&row)[2 * R + 2]) {
    using V = typename VecOf<T, VEC>::type;
    constexpr int NT = 2 * R + 2;
    constexpr int NL = (NT + VEC - 1 + VEC - 1) / VEC;  // vectors covering NT taps at any shift < VEC
    const bool yok = (y1 >= 0) && (y1 < H2);
    const T* rp = plane + (int64_t)(yok ? y1 : 0) * W2;
    const int xa = xb & ~(VEC - 1);
    float g[NL][VEC];
#pragma unroll
    for (int q = 0; q < NL; q++) {
        // out-of-plane vectors read the (aligned, valid) plane start and are zeroed
        const int x = xa + q * VEC;
        const bool ok = yok && x >= 0 && x < W2;
        const V v = *reinterpret_cast<const V*>(ok ? rp + x : plane);
        unpack_vec<T, VEC>(v, g[q]);
#pragma unroll
        for (int e = 0; e < VEC; e++) g[q][e] = ok ? g[q][e] : 0.f;
    }
    // Shift by sh = xb - xa in two select stages.  The candidates pass through an
    // empty asm so that the compiler keeps per-lane selects instead of turning
    // them into a dynamically indexed (scratch) array access.
    const int sh = xb - xa;  // 0 .. VEC-1
    float h[NL * VEC - 1];
#pragma unroll
    for (int k = 0; k < NL * VEC - 1; k++) {
        float a = g[k / VEC][k % VEC], b = g[(k + 1) / VEC][(k + 1) % VEC];
        asm("" : "+v"(a), "+v"(b));
        h[k] = (sh & 1) ? b : a;
    }
    if (VEC == 4) {
#pragma unroll
        for (int i = 0; i < NT; i++) {
            float a = h[i], b = h[i + 2];
            asm("" : "+v"(a), "+v"(b));
            row[i] = (sh & 2) ? b : a;
        }
    } else {
#pragma unroll
        for (int i = 0; i < NT; i++) row[i] = h[i];
    }
}

template <typename T, int R, int VEC>
__device__ __forceinline__ void lookup_level_vec(const T* __restrict__ plane, int H2, int W2, float x0, float y0,
                                                 T* __restrict__ o, int64_t ostride) {
    constexpr int RD = 2 * R + 1;
    const float fx0 = floorf(x0), fy0 = floorf(y0);
    const float dx = x0 - fx0, dy = y0 - fy0;
    const int xb = (int)fx0 - R, yb = (int)fy0 - R;
    const float w00 = (1.f - dx) * (1.f - dy), w10 = dx * (1.f - dy), w01 = (1.f - dx) * dy, w11 = dx * dy;
    float r0[RD + 1], r1[RD + 1];
    load_tap_row_vec<T, R, VEC>(plane, yb, xb, H2, W2, r0);
#pragma unroll
    for (int j = 0; j < RD; j++) {
        float (&top)[RD + 1] = (j & 1) ? r1 : r0;
        float (&bot)[RD + 1] = (j & 1) ? r0 : r1;
        load_tap_row_vec<T, R, VEC>(plane, yb + j + 1, xb, H2, W2, bot);
#pragma unroll
        for (int i = 0; i < RD; i++) {
            const float v = top[i] * w00 + top[i + 1] * w10 + bot[i] * w01 + bot[i + 1] * w11;
            o[(int64_t)(i * RD + j) * ostride] = store_from<T>(v);
        }
    }
}

// Lookup of one level choosing the widest aligned vector the plane allows
// (vec: 4, 2, or 1 for the scalar path; decided on the host per tensor).
template <typename T, int R>
__device__ __forceinline__ void lookup_level_any(const T* __restrict__ plane, int H2, int W2, float x0, float y0,
                                                 T* __restrict__ o, int64_t ostride, int vec) {
    if constexpr (std::is_same<T, float>::value || std::is_same<T, __half>::value) {
        if (vec == 4) return lookup_level_vec<T, R, 4>(plane, H2, W2, x0, y0, o, ostride);
        if (vec == 2) return lookup_level_vec<T, R, 2>(plane, H2, W2, x0, y0, o, ostride);
    }
    lookup_level<T, R>(plane, H2, W2, x0, y0, o, ostride);
}

// Widest vector width (4, 2 or 1 elements) usable for planes of width W2
// starting at base with plane size H2*W2.
inline int lookup_vec_width(const void* base, int H2, int W2, int elem) {
    for (int v = 4; v >= 2; v /= 2) {
        if (W2 % v == 0 && ((uintptr_t)base % (uintptr_t)(v * elem)) == 0) return v;
    }
    (void)H2;
    return 1;
}

template <typename T, int R>
__global__ __launch_bounds__(256) void corr_index_fwd_kernel(const T* __restrict__ vol,
                                                             const float* __restrict__ coords, T* __restrict__ out,
                                                             int64_t B, int H1, int W1, int H2, int W2, int vec) {
    constexpr int RD = 2 * R + 1;
    const int64_t HW1 = (int64_t)H1 * W1;
    const int64_t pix = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (pix >= B * HW1) return;
    const int64_t b = pix / HW1;
    const int64_t yx = pix - b * HW1;
    const float x0 = coords[(b * 2 + 0) * HW1 + yx];
    const float y0 = coords[(b * 2 + 1) * HW1 + yx];
    lookup_level_any<T, R>(vol + pix * (int64_t)H2 * W2, H2, W2, x0, y0, out + b * RD * RD * HW1 + yx, HW1, vec);
}

// Generic-radius variant (any r): same maths, runtime loops.
template <typename T>
__global__ __launch_bounds__(256) void corr_index_fwd_kernel_rt(const T* __restrict__ vol,
                                                                const float* __restrict__ coords,
                                                                T* __restrict__ out, int64_t B, int H1, int W1,
                                                                int H2, int W2, int r) {
    using A = typename Acc<T>::type;
    const int rd = 2 * r + 1;
    const int64_t HW1 = (int64_t)H1 * W1;
    const int64_t pix = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (pix >= B * HW1) return;
    const int64_t b = pix / HW1;
    const int64_t yx = pix - b * HW1;
    const float x0 = coords[(b * 2 + 0) * HW1 + yx];
    const float y0 = coords[(b * 2 + 1) * HW1 + yx];
    const float fx0 = floorf(x0), fy0 = floorf(y0);
    const float dx = x0 - fx0, dy = y0 - fy0;
    const int xb = (int)fx0 - r, yb = (int)fy0 - r;
    const T* plane = vol + pix * (int64_t)H2 * W2;
    T* o = out + b * rd * rd * HW1 + yx;
    for (int i = 0; i < rd; i++) {
        for (int j = 0; j < rd; j++) {
            A v = 0;
            for (int dj = 0; dj < 2; dj++)
                for (int di = 0; di < 2; di++) {
                    const int x1 = xb + i + di, y1 = yb + j + dj;
                    if (x1 < 0 || x1 >= W2 || y1 < 0 || y1 >= H2) continue;
                    const A w = (A)((di ? dx : 1.f - dx) * (dj ? dy : 1.f - dy));
                    v += w * load_as<A>(plane + (int64_t)y1 * W2 + x1);
                }
            o[(i * rd + j) * HW1] = store_from<T>(v);
        }
    }
}

// Backward: each lane owns its plane, so the tap writes do not collide; the
// caller provides a zero-filled volume_grad.
template <typename T>
__global__ __launch_bounds__(256) void corr_index_bwd_kernel(const float* __restrict__ coords,
                                                             const T* __restrict__ gout, T* __restrict__ vgrad,
                                                             int64_t B, int H1, int W1, int H2, int W2, int r) {
    using A = typename Acc<T>::type;
    const int rd = 2 * r + 1;
    const int64_t HW1 = (int64_t)H1 * W1;
    const int64_t pix = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (pix >= B * HW1) return;
    const int64_t b = pix / HW1;
    const int64_t yx = pix - b * HW1;
    const float x0 = coords[(b * 2 + 0) * HW1 + yx];
    const float y0 = coords[(b * 2 + 1) * HW1 + yx];
    const float fx0 = floorf(x0), fy0 = floorf(y0);
    const float dx = x0 - fx0, dy = y0 - fy0;
    const A w11 = (A)(dx * dy), w10 = (A)(dx * (1.f - dy)), w01 = (A)((1.f - dx) * dy),
            w00 = (A)((1.f - dx) * (1.f - dy));
    const int xb = (int)fx0 - r, yb = (int)fy0 - r;
    T* plane = vgrad + pix * (int64_t)H2 * W2;
    const T* g = gout + b * rd * rd * HW1 + yx;
    for (int j = 0; j < rd + 1; j++) {
        const int y1 = yb + j;
        if (y1 < 0 || y1 >= H2) continue;
        for (int i = 0; i < rd + 1; i++) {
            const int x1 = xb + i;
            if (x1 < 0 || x1 >= W2) continue;
            A s = 0;
            if (i > 0 && j > 0) s += load_as<A>(g + ((i - 1) * rd + (j - 1)) * HW1) * w11;
            if (i > 0 && j < rd) s += load_as<A>(g + ((i - 1) * rd + j) * HW1) * w10;
            if (i < rd && j > 0) s += load_as<A>(g + (i * rd + (j - 1)) * HW1) * w01;
            if (i < rd && j < rd) s += load_as<A>(g + (i * rd + j) * HW1) * w00;
            plane[(int64_t)y1 * W2 + x1] = store_from<T>(s);
        }
    }
}

template <typename T>
int corr_index_fwd_launch(int64_t B, int H1, int W1, int H2, int W2, const void* vol, const float* coords, int r,
                          void* out, hipStream_t s) {
    const int64_t n = B * H1 * (int64_t)W1;
    if (n == 0) return VH_OK;
    const unsigned grid = ceil_div(n, 256);
    if (r == 3)
        hipLaunchKernelGGL((corr_index_fwd_kernel<T, 3>), dim3(grid), dim3(256), 0, s, (const T*)vol, coords, (T*)out,
                           B, H1, W1, H2, W2, lookup_vec_width(vol, H2, W2, (int)sizeof(T)));
    else
        hipLaunchKernelGGL(corr_index_fwd_kernel_rt<T>, dim3(grid), dim3(256), 0, s, (const T*)vol, coords, (T*)out,
                           B, H1, W1, H2, W2, r);
    VH_LAUNCH_CHECK("corr_index_forward");
    return VH_OK;
}

template <typename T>
int corr_index_bwd_launch(int64_t B, int H1, int W1, int H2, int W2, const float* coords, const void* g, int r,
                          void* vg, hipStream_t s) {
    const int64_t n = B * H1 * (int64_t)W1;
    if (n == 0) return VH_OK;
    hipLaunchKernelGGL(corr_index_bwd_kernel<T>, dim3(ceil_div(n, 256)), dim3(256), 0, s, coords, (const T*)g, (T*)vg,
                       B, H1, W1, H2, W2, r);
    VH_LAUNCH_CHECK("corr_index_backward");
    return VH_OK;
}


// Fused pyramid lookup (CorrBlock.__call__, droid_net.py:71-84): one lane per
// source pixel handles every level, reading coords in [B,H,W,2] layout,
// scaling them by 2^-l in-register and writing channel l*(2r+1)^2 + i*(2r+1)+j
// of one [B, L*(2r+1)^2, H, W] output - no coordinate copies, no concat.
struct LevelPtrs {
    const void* p[8];
    int h[8];
    int w[8];
    int vec[8];
};

template <typename T, int R>
__global__ __launch_bounds__(256) void corr_lookup_pyramid_kernel(LevelPtrs lv, int L, const float* __restrict__ coords,
                                                                  T* __restrict__ out, int64_t B, int H1, int W1) {
    constexpr int RD = 2 * R + 1;
    const int64_t HW1 = (int64_t)H1 * W1;
    const int64_t pix = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (pix >= B * HW1) return;
    const int64_t b = pix / HW1;
    const int64_t yx = pix - b * HW1;
    const float cxf = coords[pix * 2 + 0];
    const float cyf = coords[pix * 2 + 1];
    T* o = out + b * (int64_t)L * RD * RD * HW1 + yx;
#pragma unroll 1
    for (int l = 0; l < L; l++) {
        const float sc = 1.0f / (float)(1 << l);
        const int H2 = lv.h[l], W2 = lv.w[l];
        const T* plane = (const T*)lv.p[l] + pix * (int64_t)H2 * W2;
        lookup_level_any<T, R>(plane, H2, W2, cxf * sc, cyf * sc, o + (int64_t)l * RD * RD * HW1, HW1, lv.vec[l]);
    }
}

template <typename T>
int corr_lookup_pyramid_launch(int64_t B, int H1, int W1, int L, const void* const* levels, const int* Hl,
                               const int* Wl, const float* coords, void* out, hipStream_t s) {
    LevelPtrs lv;
    for (int l = 0; l < L; l++) {
        lv.p[l] = levels[l];
        lv.h[l] = Hl[l];
        lv.w[l] = Wl[l];
        lv.vec[l] = lookup_vec_width(levels[l], Hl[l], Wl[l], (int)sizeof(T));
    }
    const int64_t n = B * H1 * (int64_t)W1;
    if (n == 0) return VH_OK;
    hipLaunchKernelGGL((corr_lookup_pyramid_kernel<T, 3>), dim3(ceil_div(n, 256)), dim3(256), 0, s, lv, L, coords,
                       (T*)out, B, H1, W1);
    VH_LAUNCH_CHECK("corr_lookup_pyramid");
    return VH_OK;
}

// ---------------------------------------------------------------------------
// altcorr: on-the-fly correlation.  One lane per output pixel; channels are
// processed in chunks of 32 held in registers; the 64 tap dot products
// accumulate in registers and are splatted at the end.
// ---------------------------------------------------------------------------
constexpr int kChunk = 32;

template <typename T, int R>
__device__ __forceinline__ void altcorr_pixel(const T* __restrict__ f1, const T* __restrict__ f2, int H2, int W2,
                                              int C, float x, float y, float* __restrict__ o, int64_t ostride) {
    constexpr int RD = 2 * R + 1;
    constexpr int NT = RD + 1;
    const float fx0 = floorf(x), fy0 = floorf(y);
    const float dx = x - fx0, dy = y - fy0;
    const int xb = (int)fx0 - R, yb = (int)fy0 - R;
    float s[NT][NT];  // [iy][ix]
#pragma unroll
    for (int a = 0; a < NT; a++)
#pragma unroll
        for (int b = 0; b < NT; b++) s[a][b] = 0.f;
    for (int c0 = 0; c0 < C; c0 += kChunk) {
        const int cn = (C - c0) < kChunk ? (C - c0) : kChunk;
        float a1[kChunk];
#pragma unroll
        for (int k = 0; k < kChunk; k++) a1[k] = (k < cn) ? to_f(f1[c0 + k]) : 0.f;
#pragma unroll
        for (int iy = 0; iy < NT; iy++) {
            const int h2 = yb + iy;
            if (h2 < 0 || h2 >= H2) continue;
#pragma unroll
            for (int ix = 0; ix < NT; ix++) {
                const int w2 = xb + ix;
                if (w2 < 0 || w2 >= W2) continue;
                const T* p = f2 + ((int64_t)h2 * W2 + w2) * C + c0;
                float acc = 0.f;
#pragma unroll
                for (int k = 0; k < kChunk; k++)
                    if (k < cn) acc += a1[k] * to_f(p[k]);
                s[iy][ix] += acc;
            }
        }
    }
#pragma unroll
    for (int ix = 0; ix < RD; ix++)
#pragma unroll
        for (int iy = 0; iy < RD; iy++) {
            const float v = s[iy][ix] * (1 - dy) * (1 - dx) + s[iy][ix + 1] * (1 - dy) * dx +
                            s[iy + 1][ix] * dy * (1 - dx) + s[iy + 1][ix + 1] * dy * dx;
            o[(ix * RD + iy) * ostride] = v;
        }
}

template <typename T, int R>
__global__ __launch_bounds__(128) void altcorr_fwd_kernel(const T* __restrict__ fmap1, const T* __restrict__ fmap2,
                                                          const float* __restrict__ coords, T* __restrict__ out,
                                                          int64_t B, int N, int H1, int W1, int H2, int W2, int C) {
    constexpr int RD = 2 * R + 1;
    const int64_t HW1 = (int64_t)H1 * W1;
    const int64_t gid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;  // over B*N*HW1
    if (gid >= B * N * HW1) return;
    const int64_t bn = gid / HW1;
    const int64_t yx = gid - bn * HW1;
    const int64_t b = bn / N;
    const float x = coords[gid * 2 + 0];
    const float y = coords[gid * 2 + 1];
    float tmp[RD * RD];
    altcorr_pixel<T, R>(fmap1 + (b * HW1 + yx) * C, fmap2 + b * (int64_t)H2 * W2 * C, H2, W2, C, x, y, tmp, 1);
    T* o = out + bn * RD * RD * HW1 + yx;
#pragma unroll
    for (int c = 0; c < RD * RD; c++) o[c * HW1] = from_f<T>(tmp[c]);
}

// Indexed multi-level forward: one lane per (b, m, pixel, level).
template <typename T, int R>
__global__ __launch_bounds__(128) void altcorr_index_kernel(const T* __restrict__ fmap1, const T* __restrict__ f2l,
                                                            const float* __restrict__ coords,
                                                            const int64_t* __restrict__ ii,
                                                            const int64_t* __restrict__ jj, float* __restrict__ out,
                                                            int64_t B, int Nf, int M, int H, int W, int C, int H2,
                                                            int W2, int level, int L) {
    constexpr int RD = 2 * R + 1;
    const int64_t HW = (int64_t)H * W;
    const int64_t gid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;  // over B*M*HW
    if (gid >= B * M * HW) return;
    const int64_t bm = gid / HW;
    const int64_t yx = gid - bm * HW;
    const int64_t b = bm / M;
    const int m = (int)(bm - b * M);
    const int64_t src = ii[m], dst = jj[m];
    const float scale = (float)(1 << level);
    const float x = coords[gid * 2 + 0] / scale;
    const float y = coords[gid * 2 + 1] / scale;
    altcorr_pixel<T, R>(fmap1 + ((b * Nf + src) * HW + yx) * C, f2l + (b * Nf + dst) * (int64_t)H2 * W2 * C, H2, W2,
                        C, x, y, out + (bm * L * RD * RD + level * RD * RD) * HW + yx, HW);
}

// Backward (float only, as the reference): one lane per (b, n, pixel) and tap;
// fmap1_grad accumulates per lane then atomically, fmap2_grad atomically.
template <int R>
__global__ __launch_bounds__(128) void altcorr_bwd_kernel(const float* __restrict__ fmap1,
                                                          const float* __restrict__ fmap2,
                                                          const float* __restrict__ coords,
                                                          const float* __restrict__ gout, float* __restrict__ g1,
                                                          float* __restrict__ g2, int64_t B, int N, int H1, int W1,
                                                          int H2, int W2, int C) {
    constexpr int RD = 2 * R + 1;
    constexpr int NT = RD + 1;
    const int64_t HW1 = (int64_t)H1 * W1;
    const int64_t gid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (gid >= B * N * HW1) return;
    const int64_t bn = gid / HW1;
    const int64_t yx = gid - bn * HW1;
    const int64_t b = bn / N;
    const float x = coords[gid * 2 + 0];
    const float y = coords[gid * 2 + 1];
    const float fx0 = floorf(x), fy0 = floorf(y);
    const float dx = x - fx0, dy = y - fy0;
    const int xb = (int)fx0 - R, yb = (int)fy0 - R;
    const float* go = gout + bn * RD * RD * HW1 + yx;
    // tap weights g[iy][ix] = sum of output gradients routed to this tap
    float gt[NT][NT];
#pragma unroll
    for (int a = 0; a < NT; a++)
#pragma unroll
        for (int c = 0; c < NT; c++) gt[a][c] = 0.f;
#pragma unroll
    for (int ix = 0; ix < RD; ix++)
#pragma unroll
        for (int iy = 0; iy < RD; iy++) {
            const float g = go[(ix * RD + iy) * HW1];
            gt[iy][ix] += g * (1 - dy) * (1 - dx);
            gt[iy][ix + 1] += g * (1 - dy) * dx;
            gt[iy + 1][ix] += g * dy * (1 - dx);
            gt[iy + 1][ix + 1] += g * dy * dx;
        }
    const float* f1 = fmap1 + (b * HW1 + yx) * C;
    float* gf1 = g1 + (b * HW1 + yx) * C;
    const float* f2b = fmap2 + b * (int64_t)H2 * W2 * C;
    float* g2b = g2 + b * (int64_t)H2 * W2 * C;
    for (int c = 0; c < C; c++) {
        const float a = f1[c];
        float acc = 0.f;
#pragma unroll
        for (int iy = 0; iy < NT; iy++) {
            const int h2 = yb + iy;
            if (h2 < 0 || h2 >= H2) continue;
#pragma unroll
            for (int ix = 0; ix < NT; ix++) {
                const int w2 = xb + ix;
                if (w2 < 0 || w2 >= W2) continue;
                const int64_t off = ((int64_t)h2 * W2 + w2) * C + c;
                acc += gt[iy][ix] * f2b[off];
                atomicAdd(g2b + off, gt[iy][ix] * a);
            }
        }
        atomicAdd(gf1 + c, acc);
    }
}

// ---------------------------------------------------------------------------
// All-pairs volume (level 0) on MFMA + 2x2 mean pyramid.
//
// Level 0 for one edge is a [HW x C] x [C x HW] product.  A workgroup of
// 256 lanes (4 waves) computes a 64 x 64 output tile: each wave a 32 x 32
// sub-tile with v_mfma_f32_32x32x8 (f16 / bf16 inputs, K = 8 per issue) or
// v_mfma_f32_32x32x2 (f32 inputs).  Operands are staged through LDS in
// K-slices of 32 channels.  The 1/16 scale is applied in the epilogue.
// ---------------------------------------------------------------------------
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef _Float16 f16x4 __attribute__((ext_vector_type(4)));
typedef short bf16x4 __attribute__((ext_vector_type(4)));

constexpr int kTM = 64, kTN = 64, kTK = 32;

template <typename Tin>
struct Mfma;

template <>
struct Mfma<__half> {
    static constexpr int KSTEP = 8;  // per issue
    __device__ static f32x16 run(const __half* a, const __half* b, f32x16 c) {
        f16x4 av, bv;
        for (int i = 0; i < 4; i++) {
            av[i] = (_Float16)__half2float(a[i]);
            bv[i] = (_Float16)__half2float(b[i]);
        }
        return __builtin_amdgcn_mfma_f32_32x32x8f16(av, bv, c, 0, 0, 0);
    }
};

template <>
struct Mfma<__hip_bfloat16> {
    static constexpr int KSTEP = 8;
    __device__ static f32x16 run(const __hip_bfloat16* a, const __hip_bfloat16* b, f32x16 c) {
        bf16x4 av, bv;
        for (int i = 0; i < 4; i++) {
            av[i] = *reinterpret_cast<const short*>(&a[i]);
            bv[i] = *reinterpret_cast<const short*>(&b[i]);
        }
        return __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(av, bv, c, 0, 0, 0);
    }
};

template <>
struct Mfma<float> {
    static constexpr int KSTEP = 2;
    __device__ static f32x16 run(const float* a, const float* b, f32x16 c) {
        return __builtin_amdgcn_mfma_f32_32x32x2f32(a[0], b[0], c, 0, 0, 0);
    }
};

// LDS tiles: As[k][m], Bs[k][n] with K-slice kTK.  MFMA 32x32xK operand layout:
// lane l supplies row (l % 32) of A and column (l % 32) of B for the K-group
// (l / 32): f16/bf16 K=8 -> elements k = 4*(l/32) .. +3; f32 K=2 -> k = l/32.
// Accumulator layout: lane l holds column (l % 32), rows
// 8*(v/4) + 4*(l/32) + (v%4) for v = 0..15.
template <typename Tin, typename Tout>
__global__ __launch_bounds__(256) void corr_gemm_kernel(const Tin* __restrict__ f1, const Tin* __restrict__ f2,
                                                        Tout* __restrict__ out, int C, int HW) {
    constexpr int KS = Mfma<Tin>::KSTEP;
    constexpr int PER = KS / 2;  // operand elements per lane per issue
    __shared__ Tin As[kTK][kTM + 4];
    __shared__ Tin Bs[kTK][kTN + 4];
    const int e = blockIdx.z;
    const int m0 = blockIdx.y * kTM;
    const int n0 = blockIdx.x * kTN;
    const int tid = threadIdx.x;
    const int wave = tid / 64, lane = tid % 64;
    const int wm = (wave / 2) * 32, wn = (wave % 2) * 32;
    const Tin* A = f1 + (int64_t)e * C * HW;  // [C][HW] : A[k][m]
    const Tin* Bm = f2 + (int64_t)e * C * HW; // [C][HW] : B[k][n]
    f32x16 acc;
    for (int v = 0; v < 16; v++) acc[v] = 0.f;
    for (int k0 = 0; k0 < C; k0 += kTK) {
        // stage K-slice: kTK x 64 for A and B, coalesced along HW
        for (int idx = tid; idx < kTK * kTM; idx += 256) {
            const int k = idx / kTM, m = idx % kTM;
            const int kk = k0 + k;
            As[k][m] = (kk < C && m0 + m < HW) ? A[(int64_t)kk * HW + m0 + m] : from_f<Tin>(0.f);
            Bs[k][m] = (kk < C && n0 + m < HW) ? Bm[(int64_t)kk * HW + n0 + m] : from_f<Tin>(0.f);
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < kTK; ks += KS) {
            Tin a[PER > 0 ? PER * 2 : 1], b[PER > 0 ? PER * 2 : 1];
            const int kb = ks + (lane / 32) * (KS / 2);
#pragma unroll
            for (int i = 0; i < KS / 2; i++) {
                a[i] = As[kb + i][wm + lane % 32];
                b[i] = Bs[kb + i][wn + lane % 32];
            }
            acc = Mfma<Tin>::run(a, b, acc);
        }
        __syncthreads();
    }
    Tout* O = out + (int64_t)e * HW * HW;
#pragma unroll
    for (int v = 0; v < 16; v++) {
        const int row = m0 + wm + 8 * (v / 4) + 4 * (lane / 32) + (v % 4);
        const int col = n0 + wn + lane % 32;
        if (row < HW && col < HW) O[(int64_t)row * HW + col] = from_f<Tout>(acc[v] * 0.0625f);
    }
}

// 2x2 mean over the last two dims of [planes, H, W] -> [planes, H/2, W/2].
template <typename T>
__global__ __launch_bounds__(256) void pool2x2_kernel(const T* __restrict__ in, T* __restrict__ out, int64_t planes,
                                                      int H, int W) {
    const int Ho = H / 2, Wo = W / 2;
    const int64_t n = planes * Ho * Wo;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t p = i / ((int64_t)Ho * Wo);
        const int r = (int)(i - p * Ho * Wo);
        const int y = r / Wo, x = r % Wo;
        const T* s = in + p * H * (int64_t)W + (int64_t)(2 * y) * W + 2 * x;
        const float v = 0.25f * ((to_f(s[0]) + to_f(s[1])) + (to_f(s[W]) + to_f(s[W + 1])));
        out[i] = from_f<T>(v);
    }
}

template <typename Tin, typename Tout>
int corr_build_launch(int64_t E, int C, int H, int W, const void* f1, const void* f2, int L, void* const* levels,
                      hipStream_t s) {
    const int HW = H * W;
    dim3 grid(ceil_div(HW, kTN), ceil_div(HW, kTM), (unsigned)E);
    hipLaunchKernelGGL((corr_gemm_kernel<Tin, Tout>), grid, dim3(256), 0, s, (const Tin*)f1, (const Tin*)f2,
                       (Tout*)levels[0], C, HW);
    VH_LAUNCH_CHECK("corr_build gemm");
    int h = H, w = W;
    for (int l = 1; l < L; l++) {
        const int64_t planes = E * (int64_t)HW;
        const int64_t n = planes * (h / 2) * (int64_t)(w / 2);
        int64_t blocks = (n + 255) / 256;
        if (blocks > 1 << 20) blocks = 1 << 20;
        if (n > 0)
            hipLaunchKernelGGL(pool2x2_kernel<Tout>, dim3((unsigned)blocks), dim3(256), 0, s,
                               (const Tout*)levels[l - 1], (Tout*)levels[l], planes, h, w);
        VH_LAUNCH_CHECK("corr_build pool");
        h /= 2;
        w /= 2;
    }
    return VH_OK;
}

}  // namespace
}  // namespace vh

using namespace vh;

extern "C" {

int vh_corr_index_forward(int dtype, int64_t B, int H1, int W1, int H2, int W2, const void* volume,
                          const float* coords, int radius, void* corr, void* stream) {
    if (B < 0 || H1 < 0 || W1 < 0 || H2 < 0 || W2 < 0 || radius < 0)
        return fail(VH_ERR_ARG, "corr_index_forward: negative size");
    hipStream_t s = (hipStream_t)stream;
    switch (dtype) {
        case VH_F16: return corr_index_fwd_launch<__half>(B, H1, W1, H2, W2, volume, coords, radius, corr, s);
        case VH_F32: return corr_index_fwd_launch<float>(B, H1, W1, H2, W2, volume, coords, radius, corr, s);
        case VH_F64: return corr_index_fwd_launch<double>(B, H1, W1, H2, W2, volume, coords, radius, corr, s);
        default: return fail(VH_ERR_UNSUPPORTED, "corr_index_forward: dtype %d", dtype);
    }
}

int vh_corr_index_backward(int dtype, int64_t B, int H1, int W1, int H2, int W2, const float* coords,
                           const void* corr_grad, int radius, void* volume_grad, void* stream) {
    if (B < 0 || H1 < 0 || W1 < 0 || H2 < 0 || W2 < 0 || radius < 0)
        return fail(VH_ERR_ARG, "corr_index_backward: negative size");
    hipStream_t s = (hipStream_t)stream;
    switch (dtype) {
        case VH_F16: return corr_index_bwd_launch<__half>(B, H1, W1, H2, W2, coords, corr_grad, radius, volume_grad, s);
        case VH_F32: return corr_index_bwd_launch<float>(B, H1, W1, H2, W2, coords, corr_grad, radius, volume_grad, s);
        case VH_F64: return corr_index_bwd_launch<double>(B, H1, W1, H2, W2, coords, corr_grad, radius, volume_grad, s);
        default: return fail(VH_ERR_UNSUPPORTED, "corr_index_backward: dtype %d", dtype);
    }
}

int vh_corr_lookup_pyramid(int dtype, int64_t B, int H1, int W1, int num_levels, const void* const* levels,
                           const int* Hl, const int* Wl, const float* coords, int radius, void* out, void* stream) {
    if (B < 0 || H1 < 0 || W1 < 0) return fail(VH_ERR_ARG, "corr_lookup_pyramid: negative size");
    if (num_levels < 1 || num_levels > 8) return fail(VH_ERR_ARG, "corr_lookup_pyramid: 1..8 levels");
    if (radius != 3) return fail(VH_ERR_UNSUPPORTED, "corr_lookup_pyramid: radius %d (built for r=3)", radius);
    hipStream_t s = (hipStream_t)stream;
    switch (dtype) {
        case VH_F16: return corr_lookup_pyramid_launch<__half>(B, H1, W1, num_levels, levels, Hl, Wl, coords, out, s);
        case VH_F32: return corr_lookup_pyramid_launch<float>(B, H1, W1, num_levels, levels, Hl, Wl, coords, out, s);
        case VH_F64: return corr_lookup_pyramid_launch<double>(B, H1, W1, num_levels, levels, Hl, Wl, coords, out, s);
        default: return fail(VH_ERR_UNSUPPORTED, "corr_lookup_pyramid: dtype %d", dtype);
    }
}

int vh_corr_build(int in_dtype, int out_dtype, int64_t E, int C, int H, int W, const void* fmap1, const void* fmap2,
                  int num_levels, void* const* levels, void* stream) {
    if (E < 0 || C <= 0 || H <= 0 || W <= 0 || num_levels < 1)
        return fail(VH_ERR_ARG, "corr_build: bad sizes");
    if (E == 0) return VH_OK;
    hipStream_t s = (hipStream_t)stream;
    if (out_dtype == VH_F32) {
        if (in_dtype == VH_F32) return corr_build_launch<float, float>(E, C, H, W, fmap1, fmap2, num_levels, levels, s);
        if (in_dtype == VH_F16) return corr_build_launch<__half, float>(E, C, H, W, fmap1, fmap2, num_levels, levels, s);
        if (in_dtype == VH_BF16)
            return corr_build_launch<__hip_bfloat16, float>(E, C, H, W, fmap1, fmap2, num_levels, levels, s);
    } else if (out_dtype == VH_F16) {
        if (in_dtype == VH_F32) return corr_build_launch<float, __half>(E, C, H, W, fmap1, fmap2, num_levels, levels, s);
        if (in_dtype == VH_F16) return corr_build_launch<__half, __half>(E, C, H, W, fmap1, fmap2, num_levels, levels, s);
        if (in_dtype == VH_BF16)
            return corr_build_launch<__hip_bfloat16, __half>(E, C, H, W, fmap1, fmap2, num_levels, levels, s);
    }
    return fail(VH_ERR_UNSUPPORTED, "corr_build: dtypes in=%d out=%d", in_dtype, out_dtype);
}

size_t vh_corr_build_fp8_workspace(int64_t E, int C, int H, int W) { return 0; }

int vh_corr_build_fp8(int out_dtype, int64_t E, int C, int H, int W, const float* fmap1, const float* fmap2,
                      int num_levels, void* const* levels, void* workspace, void* stream) {
    return fail(VH_ERR_UNSUPPORTED, "corr_build_fp8: not built yet");
}

int vh_altcorr_forward(int dtype, int64_t B, int N, int H1, int W1, int H2, int W2, int C, const void* fmap1,
                       const void* fmap2, const float* coords, int radius, void* corr, void* stream) {
    if (radius != 3) return fail(VH_ERR_UNSUPPORTED, "altcorr_forward: radius %d (kernels built for r=3)", radius);
    const int64_t n = B * N * (int64_t)H1 * W1;
    if (n == 0) return VH_OK;
    hipStream_t s = (hipStream_t)stream;
    dim3 grid(ceil_div(n, 128));
    switch (dtype) {
        case VH_F16:
            hipLaunchKernelGGL((altcorr_fwd_kernel<__half, 3>), grid, dim3(128), 0, s, (const __half*)fmap1,
                               (const __half*)fmap2, coords, (__half*)corr, B, N, H1, W1, H2, W2, C);
            break;
        case VH_F32:
            hipLaunchKernelGGL((altcorr_fwd_kernel<float, 3>), grid, dim3(128), 0, s, (const float*)fmap1,
                               (const float*)fmap2, coords, (float*)corr, B, N, H1, W1, H2, W2, C);
            break;
        default: return fail(VH_ERR_UNSUPPORTED, "altcorr_forward: dtype %d", dtype);
    }
    VH_LAUNCH_CHECK("altcorr_forward");
    return VH_OK;
}

int vh_altcorr_backward(int dtype, int64_t B, int N, int H1, int W1, int H2, int W2, int C, const void* fmap1,
                        const void* fmap2, const float* coords, const void* corr_grad, int radius, void* fmap1_grad,
                        void* fmap2_grad, void* stream) {
    if (dtype != VH_F32) return fail(VH_ERR_UNSUPPORTED, "altcorr_backward: float32 only");
    if (radius != 3) return fail(VH_ERR_UNSUPPORTED, "altcorr_backward: radius %d (kernels built for r=3)", radius);
    const int64_t n = B * N * (int64_t)H1 * W1;
    if (n == 0) return VH_OK;
    hipLaunchKernelGGL(altcorr_bwd_kernel<3>, dim3(ceil_div(n, 128)), dim3(128), 0, (hipStream_t)stream,
                       (const float*)fmap1, (const float*)fmap2, coords, (const float*)corr_grad, (float*)fmap1_grad,
                       (float*)fmap2_grad, B, N, H1, W1, H2, W2, C);
    VH_LAUNCH_CHECK("altcorr_backward");
    return VH_OK;
}

int vh_altcorr_index_forward(int dtype, int64_t B, int Nf, int M, int H, int W, int C, const void* fmap1,
                             int num_levels, const void* const* levels, const int* Hl, const int* Wl,
                             const float* coords, const int64_t* ii, const int64_t* jj, int radius, void* corr,
                             void* stream) {
    if (radius != 3) return fail(VH_ERR_UNSUPPORTED, "altcorr_index_forward: radius %d (built for r=3)", radius);
    const int64_t n = B * M * (int64_t)H * W;
    if (n == 0) return VH_OK;
    hipStream_t s = (hipStream_t)stream;
    for (int l = 0; l < num_levels; l++) {
        dim3 grid(ceil_div(n, 128));
        switch (dtype) {
            case VH_F16:
                hipLaunchKernelGGL((altcorr_index_kernel<__half, 3>), grid, dim3(128), 0, s, (const __half*)fmap1,
                                   (const __half*)levels[l], coords, ii, jj, (float*)corr, B, Nf, M, H, W, C, Hl[l],
                                   Wl[l], l, num_levels);
                break;
            case VH_F32:
                hipLaunchKernelGGL((altcorr_index_kernel<float, 3>), grid, dim3(128), 0, s, (const float*)fmap1,
                                   (const float*)levels[l], coords, ii, jj, (float*)corr, B, Nf, M, H, W, C, Hl[l],
                                   Wl[l], l, num_levels);
                break;
            default: return fail(VH_ERR_UNSUPPORTED, "altcorr_index_forward: dtype %d", dtype);
        }
        VH_LAUNCH_CHECK("altcorr_index_forward");
    }
    return VH_OK;
}

}  // extern "C"
