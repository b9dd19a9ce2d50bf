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

#include <algorithm>
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

// Volume plane layouts.  Planar: [H2][W2] row-major (the reference's).
// Tiled: each plane is padded to Hp = ceil(H2/2)*2 rows and Wp = ceil(W2/tw)*tw
// columns (tw = 4 for f32, 8 for f16: 16 bytes), the padding holding zeros.
// Planes are grouped by four, pm / 4 (source pixels side by side in a row), and
// a group is stored as [Hp/2][Wp/tw] tiles of 128 bytes, one cache line each; a
// tile holds rows 2a, 2a + 1 and columns tw b .. tw b + tw - 1 of the four
// planes as eight 16-byte pieces, piece (y % 2) * 4 + pm % 4.  The neighbouring
// source pixels of a group look up nearly the same windows (the flow is
// smooth), so the lines a window touches serve all four; per source and level
// the 8 x 8 windows of the bench's graph touch 3.4 (f32) / 2.2 (f16) GB of
// lines over 384 edges against 3.8 / 2.8 GB with a 4-row tile per plane, the
// best of the 128-byte line contents compared.  The build writes whole pieces.
// The number of planes of an edge (H1 * W1) must be a multiple of 4.
constexpr int kPieceBytes = 16;                 // one tile row of one plane: tw columns
constexpr int kTilePlanes = 4;                  // source planes sharing a tile
constexpr int kTileRows = 2;                    // rows of a plane in a tile
constexpr int kTilePieces = kTileRows * kTilePlanes;
constexpr int kTileBytes = kTilePieces * kPieceBytes;

struct LevelGeom {
    int Hl, Wl, Hp, Wp, tw;
    int tiled;
    int64_t plane;  // elements per source pixel
};

// Element offset of the origin of plane pm of a tiled level (plane elements
// per source): its pair's start plus its 16-byte piece within a tile row.
__host__ __device__ __forceinline__ int64_t tiled_src(int64_t pm, int64_t plane, int tw) {
    return (pm / kTilePlanes) * kTilePlanes * plane + (pm % kTilePlanes) * tw;
}

// Element offset of (y, x) from a plane's origin in a tiled level whose padded
// row has ntx tiles of tw columns.
__host__ __device__ __forceinline__ int tiled_yx(int y, int x, int ntx, int tw) {
    return ((y / kTileRows) * ntx + x / tw) * (kTilePieces * tw) + (y % kTileRows) * (kTilePlanes * tw) + x % tw;
}

inline LevelGeom level_geom(int Hl, int Wl, int elem, bool tiled) {
    LevelGeom g;
    g.Hl = Hl;
    g.Wl = Wl;
    g.tiled = tiled ? 1 : 0;
    g.tw = kPieceBytes / elem;
    g.Hp = tiled ? (Hl + kTileRows - 1) / kTileRows * kTileRows : Hl;
    g.Wp = tiled ? (Wl + g.tw - 1) / g.tw * g.tw : Wl;
    g.plane = (int64_t)g.Hp * g.Wp;
    return g;
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

// One tap row of a window in two steps: the raw vector loads (tap_row_load),
// then the unpack and the shift to the window's first column (tap_row_finish).
// A caller issues every row's loads before finishing any row: the shift's
// empty asm statements keep the compiler from moving loads across them.
template <typename T, int R, int VEC>
struct TapRow {
    static constexpr int NT = 2 * R + 2;
    static constexpr int NL = (NT + VEC - 1 + VEC - 1) / VEC;  // vectors covering NT taps at any shift < VEC
    using V = typename VecOf<T, VEC>::type;
    V v[NL];
    bool ok[NL];
};

template <typename T, int R, int VEC, bool TILED = false>
__device__ __forceinline__ void tap_row_load(const T* __restrict__ plane, int y1, int xb, int H2, int W2,
                                             TapRow<T, R, VEC>& t) {
    using V = typename TapRow<T, R, VEC>::V;
    constexpr int NL = TapRow<T, R, VEC>::NL;
    constexpr int TW = kPieceBytes / (int)sizeof(T);  // tile width (tiled layout)
    // tiled: H2, W2 are the padded sizes, whose extra rows and columns hold
    // zeros, and plane is the source's origin (tiled_src)
    const bool yok = (y1 >= 0) && (y1 < H2);
    const int yy = yok ? y1 : 0;
    const T* rp = TILED ? plane + tiled_yx(yy, 0, W2 / TW, TW) : plane + (int64_t)yy * W2;
    const int xa = xb & ~(VEC - 1);
#pragma unroll
    for (int q = 0; q < NL; q++) {
        // out-of-plane vectors read the (aligned, valid) plane start and are zeroed
        const int x = xa + q * VEC;
        t.ok[q] = yok && x >= 0 && x < W2;
        const T* src = TILED ? rp + (x / TW) * (kTilePieces * TW) + (x % TW) : rp + x;
        t.v[q] = *reinterpret_cast<const V*>(t.ok[q] ? src : plane);
    }
}

template <typename T, int R, int VEC>
__device__ __forceinline__ void tap_row_finish(const TapRow<T, R, VEC>& t, int xb, float (&row)[2 * R + 2]) {
    constexpr int NT = TapRow<T, R, VEC>::NT;
    constexpr int NL = TapRow<T, R, VEC>::NL;
    const int xa = xb & ~(VEC - 1);
    float g[NL][VEC];
#pragma unroll
    for (int q = 0; q < NL; q++) {
        unpack_vec<T, VEC>(t.v[q], g[q]);
#pragma unroll
        for (int e = 0; e < VEC; e++) g[q][e] = t.ok[q] ? g[q][e] : 0.f;
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

template <typename T, int R, int VEC, bool TILED = false>
__device__ __forceinline__ void lookup_level_vec(const T* __restrict__ plane, int H2, int W2, float x0, float y0,
                                                 T* __restrict__ o, int64_t ostride) {
    constexpr int RD = 2 * R + 1;
    const float fx0 = floorf(x0), fy0 = floorf(y0);
    const float dx = x0 - fx0, dy = y0 - fy0;
    const int xb = (int)fx0 - R, yb = (int)fy0 - R;
    const float w00 = (1.f - dx) * (1.f - dy), w10 = dx * (1.f - dy), w01 = (1.f - dx) * dy, w11 = dx * dy;
    if constexpr (sizeof(T) == 2) {
        // 16-bit planes: one row at a time, each output column stored as soon
        // as its two rows are in (measured faster than loading all rows first
        // for the half-width rows: 140 against 146 us on a 96-edge level)
        float rr[2][RD + 1];
        TapRow<T, R, VEC> t;
        tap_row_load<T, R, VEC, TILED>(plane, yb, xb, H2, W2, t);
        tap_row_finish<T, R, VEC>(t, xb, rr[0]);
#pragma unroll
        for (int j = 0; j < RD; j++) {
            tap_row_load<T, R, VEC, TILED>(plane, yb + j + 1, xb, H2, W2, t);
            tap_row_finish<T, R, VEC>(t, xb, rr[(j + 1) & 1]);
            const float (&top)[RD + 1] = rr[j & 1];
            const float (&bot)[RD + 1] = rr[(j + 1) & 1];
#pragma unroll
            for (int i = 0; i < RD; i++) {
                const float v = top[i] * w00 + top[i + 1] * w10 + bot[i] * w01 + bot[i + 1] * w11;
                o[(int64_t)(i * RD + j) * ostride] = store_from<T>(v);
            }
        }
        return;
    }
    // 32-bit planes: every tap row is loaded before the first output is stored:
    // stores count against the same vector-memory counter as loads, and the
    // compiler may not move a load above a store it cannot prove disjoint, so
    // interleaving them would serialise the rows' round trips
    TapRow<T, R, VEC> raw[RD + 1];
#pragma unroll
    for (int j = 0; j <= RD; j++) tap_row_load<T, R, VEC, TILED>(plane, yb + j, xb, H2, W2, raw[j]);
    float rows[RD + 1][RD + 1];
#pragma unroll
    for (int j = 0; j <= RD; j++) tap_row_finish<T, R, VEC>(raw[j], xb, rows[j]);
#pragma unroll
    for (int j = 0; j < RD; j++) {
#pragma unroll
        for (int i = 0; i < RD; i++) {
            const float v = rows[j][i] * w00 + rows[j][i + 1] * w10 + rows[j + 1][i] * w01 + rows[j + 1][i + 1] * w11;
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
        // vec == 0: tiled plane (H2, W2 padded)
        if (vec == 0) return lookup_level_vec<T, R, 4, true>(plane, H2, W2, x0, y0, o, ostride);
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

// TILED: every level is tiled (only that path is compiled).  The output
// offsets of the 49 channels are computed once and stay live across levels;
// the resulting ~150 VGPRs (3 waves per SIMD) measured faster than forcing 7-8
// waves, where the extra scattered gathers in flight slowed the lookup down.
template <typename T, int R, bool TILED>
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
        const T* plane = (const T*)lv.p[l] + (TILED ? tiled_src(pix, (int64_t)H2 * W2, kPieceBytes / (int)sizeof(T))
                                                     : pix * (int64_t)H2 * W2);
        if constexpr (TILED)
            lookup_level_vec<T, R, 4, true>(plane, H2, W2, cxf * sc, cyf * sc, o + (int64_t)l * RD * RD * HW1, HW1);
        else
            lookup_level_any<T, R>(plane, H2, W2, cxf * sc, cyf * sc, o + (int64_t)l * RD * RD * HW1, HW1, lv.vec[l]);
    }
}

// Tiled pyramid lookup through buffer resources.  One lane per source pixel.
// Per level a lane issues every load of its (2r+2)-row window before it uses
// any of them, so the rows and the neighbouring planes that share a tile are
// fetched back to back; read row by row, a tile could leave L2 between its
// rows and be fetched again.  The load resource starts at the wave's first
// plane (a multiple of 4) and a lane's plane origin is tiled_src(lane) into it;
// a vector outside the plane gets an offset past the resource's range, which
// reads zeros without touching memory.
// The outputs go through a second resource: one 32-bit lane offset, and the
// channel offset c * H1 * W1 in a scalar register.
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
constexpr uint32_t kBufRange = 0x80000000u;    // num_records of both resources (bytes)
constexpr uint32_t kBufInvalid = 0x80000000u;  // offset of an out-of-plane vector: past kBufRange
constexpr int kBufWord3 = 0x00020000;          // word 3 of a gfx9 raw buffer resource

__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void* base) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)kBufRange, kBufWord3);
}

// the 2 (f32) or 4 (f16) elements of one dword as floats
template <typename T> struct TapDword;
template <> struct TapDword<float> {
    static constexpr int n = 1;
    static __device__ __forceinline__ void unpack(uint32_t u, float* e) { e[0] = __uint_as_float(u); }
    static __device__ __forceinline__ void store(float v, __amdgpu_buffer_rsrc_t r, uint32_t vo, uint32_t so) {
        __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, vo, so, 0);
    }
};
template <> struct TapDword<__half> {
    static constexpr int n = 2;
    static __device__ __forceinline__ void unpack(uint32_t u, float* e) {
        const float2 f = __half22float2(*reinterpret_cast<const __half2*>(&u));
        e[0] = f.x;
        e[1] = f.y;
    }
    static __device__ __forceinline__ void store(float v, __amdgpu_buffer_rsrc_t r, uint32_t vo, uint32_t so) {
        __builtin_amdgcn_raw_buffer_store_b16(__half_as_ushort(__float2half(v)), r, vo, so, 0);
    }
};

template <typename T, int R>
__global__ __launch_bounds__(256) void corr_lookup_tiled_kernel(LevelPtrs lv, int L, const float* __restrict__ coords,
                                                                T* __restrict__ out, int64_t B, int H1, int W1) {
    constexpr int RD = 2 * R + 1, NT = 2 * R + 2;
    constexpr int TW = kPieceBytes / (int)sizeof(T);  // elements of one 16-byte tile row
    constexpr int LOG_TW = TW == 4 ? 2 : 3;
    constexpr int NL = (NT + TW - 1 + TW - 1) / TW;      // 16-byte vectors per tap row at any shift
    constexpr int NE = NL * TW;                          // elements loaded per tap row
    const int64_t HW1 = (int64_t)H1 * W1;
    const int lane = threadIdx.x & 63;
    const int64_t pix0 = (int64_t)blockIdx.x * 256 + __builtin_amdgcn_readfirstlane(threadIdx.x & ~63);
    const int64_t pix = pix0 + lane;
    if (pix >= B * HW1) return;
    const int64_t b0 = pix0 / HW1;
    const int64_t b = pix / HW1;
    const int64_t yx = pix - b * HW1;
    const float cx = coords[pix * 2 + 0], cy = coords[pix * 2 + 1];
    const __amdgpu_buffer_rsrc_t ro = buf_rsrc(out + b0 * (int64_t)L * RD * RD * HW1);
    const uint32_t ovo = (uint32_t)(((b - b0) * (int64_t)L * RD * RD * HW1 + yx) * (int64_t)sizeof(T));
#pragma unroll 1
    for (int l = 0; l < L; l++) {
        const float sc = 1.0f / (float)(1 << l);
        const int Hp = lv.h[l], Wp = lv.w[l], ntx = Wp >> LOG_TW;
        const int64_t plane = (int64_t)Hp * Wp;
        const __amdgpu_buffer_rsrc_t rl = buf_rsrc((const T*)lv.p[l] + pix0 * plane);
        const uint32_t lvo = (uint32_t)(tiled_src(lane, plane, TW) * (int64_t)sizeof(T));
        const float x0 = cx * sc, y0 = cy * sc;
        const float fx0 = floorf(x0), fy0 = floorf(y0);
        const float dx = x0 - fx0, dy = y0 - fy0;
        const int xb = (int)fx0 - R, yb = (int)fy0 - R;
        const int xt = xb >> LOG_TW, sh = xb - (xt << LOG_TW);  // first tile column, shift 0 .. TW-1
        u32x4 g[NT][NL];
#pragma unroll
        for (int r = 0; r < NT; r++) {
            const int y = yb + r;
            const bool yok = (unsigned)y < (unsigned)Hp;
            const uint32_t rowoff = lvo + (uint32_t)(tiled_yx(y, 0, ntx, TW) * (int)sizeof(T));
#pragma unroll
            for (int q = 0; q < NL; q++) {
                // a window on a piece boundary (sh == 0) needs one piece fewer:
                // the last one reads nothing and stays zero
                const bool ok = yok && (unsigned)(xt + q) < (unsigned)ntx && (q + 1 < NL || sh != 0);
                g[r][q] = __builtin_amdgcn_raw_buffer_load_b128(rl, ok ? rowoff + (uint32_t)(xt + q) * kTileBytes
                                                                       : kBufInvalid, 0, 0);
            }
        }
        const float w00 = (1.f - dx) * (1.f - dy), w10 = dx * (1.f - dy), w01 = (1.f - dx) * dy, w11 = dx * dy;
        float top[NT], bot[NT];
#pragma unroll
        for (int r = 0; r < NT; r++) {
            float e[NE];
#pragma unroll
            for (int q = 0; q < NL; q++)
#pragma unroll
                for (int k = 0; k < 4; k++) TapDword<T>::unpack(g[r][q][k], e + (q * 4 + k) * TapDword<T>::n);
            // taps e[sh .. sh+NT-1] by select stages of 1, 2 (and 4); the empty asm
            // keeps per-lane selects rather than a dynamically indexed array
            float h[NE];
#pragma unroll
            for (int k = 0; k < NE; k++) h[k] = e[k];
#pragma unroll
            for (int st = 1; st < TW; st <<= 1) {
#pragma unroll
                for (int k = 0; k + st < NE; k++) {
                    float a = h[k], bb = h[k + st];
                    asm("" : "+v"(a), "+v"(bb));
                    h[k] = (sh & st) ? bb : a;
                }
            }
            float(&cur)[NT] = (r & 1) ? bot : top;
#pragma unroll
            for (int i = 0; i < NT; i++) cur[i] = h[i];
            if (r == 0) continue;
            // output row j = r - 1 from tap rows r - 1 and r
            const int j = r - 1;
            const float(&t0)[NT] = (j & 1) ? bot : top;
            const float(&t1)[NT] = (j & 1) ? top : bot;
#pragma unroll
            for (int i = 0; i < RD; i++) {
                const float v = t0[i] * w00 + t0[i + 1] * w10 + t1[i] * w01 + t1[i + 1] * w11;
                TapDword<T>::store(v, ro, ovo, (uint32_t)(((l * RD + i) * RD + j) * HW1 * (int64_t)sizeof(T)));
            }
        }
    }
}

// Whether corr_lookup_tiled_kernel's 32-bit buffer offsets cover this shape: a
// wave's 64 planes of the largest level, and the outputs of the batches a wave
// can span.
inline bool lookup_tiled_offsets_fit(int64_t B, int64_t HW1, int L, int rd2, int64_t plane0_bytes, int elem) {
    const int64_t nb = std::min<int64_t>(B, 64 / std::max<int64_t>(HW1, 1) + 2);
    return 64 * plane0_bytes < (int64_t)kBufRange - kTileBytes &&
           nb * L * rd2 * HW1 * elem < (int64_t)kBufRange;
}

template <typename T>
int corr_lookup_pyramid_launch(int64_t B, int H1, int W1, int L, const void* const* levels, const int* Hl,
                               const int* Wl, bool tiled, const float* coords, void* out, hipStream_t s) {
    LevelPtrs lv;
    for (int l = 0; l < L; l++) {
        lv.p[l] = levels[l];
        if (tiled) {
            const LevelGeom g = level_geom(Hl[l], Wl[l], (int)sizeof(T), true);
            lv.h[l] = g.Hp;
            lv.w[l] = g.Wp;
            lv.vec[l] = 0;
        } else {
            lv.h[l] = Hl[l];
            lv.w[l] = Wl[l];
            lv.vec[l] = lookup_vec_width(levels[l], Hl[l], Wl[l], (int)sizeof(T));
        }
    }
    const int64_t n = B * H1 * (int64_t)W1;
    if (n == 0) return VH_OK;
    if constexpr (std::is_same<T, float>::value || std::is_same<T, __half>::value) {
        if (tiled) {
            if (lookup_tiled_offsets_fit(B, (int64_t)H1 * W1, L, 49, (int64_t)lv.h[0] * lv.w[0] * (int)sizeof(T),
                                         (int)sizeof(T)))
                hipLaunchKernelGGL((corr_lookup_tiled_kernel<T, 3>), dim3(ceil_div(n, 256)), dim3(256), 0, s, lv, L,
                                   coords, (T*)out, B, H1, W1);
            else
                hipLaunchKernelGGL((corr_lookup_pyramid_kernel<T, 3, true>), dim3(ceil_div(n, 256)), dim3(256), 0, s,
                                   lv, L, coords, (T*)out, B, H1, W1);
            VH_LAUNCH_CHECK("corr_lookup_pyramid");
            return VH_OK;
        }
    }
    hipLaunchKernelGGL((corr_lookup_pyramid_kernel<T, 3, false>), dim3(ceil_div(n, 256)), dim3(256), 0, s, lv, L,
                       coords, (T*)out, B, H1, W1);
    VH_LAUNCH_CHECK("corr_lookup_pyramid");
    return VH_OK;
}

// ---------------------------------------------------------------------------
// altcorr: on-the-fly correlation.  One lane per output pixel; channels are
// processed in chunks of 32 held in registers; the 64 tap dot products
// accumulate in registers and are splatted at the end.
// ---------------------------------------------------------------------------
constexpr int kChunk = 32;

// Sums and bilinear mix in Acc<T> (double for double features); o receives
// them converted to TO.
template <typename T, int R, typename TO>
__device__ __forceinline__ void altcorr_pixel(const T* __restrict__ f1, const T* __restrict__ f2, int H2, int W2,
                                              int C, float x, float y, TO* __restrict__ o, int64_t ostride) {
    using A = typename Acc<T>::type;
    constexpr int RD = 2 * R + 1;
    constexpr int NT = RD + 1;
    const float fx0 = floorf(x), fy0 = floorf(y);
    const A dx = x - fx0, dy = y - fy0;
    const int xb = (int)fx0 - R, yb = (int)fy0 - R;
    A s[NT][NT];  // [iy][ix]
#pragma unroll
    for (int a = 0; a < NT; a++)
#pragma unroll
        for (int b = 0; b < NT; b++) s[a][b] = A(0);
    for (int c0 = 0; c0 < C; c0 += kChunk) {
        const int cn = (C - c0) < kChunk ? (C - c0) : kChunk;
        A a1[kChunk];
#pragma unroll
        for (int k = 0; k < kChunk; k++) a1[k] = (k < cn) ? load_as<A>(f1 + c0 + k) : A(0);
#pragma unroll
        for (int iy = 0; iy < NT; iy++) {
            const int h2 = yb + iy;
            if (h2 < 0 || h2 >= H2) continue;
#pragma unroll
            for (int ix = 0; ix < NT; ix++) {
                const int w2 = xb + ix;
                if (w2 < 0 || w2 >= W2) continue;
                const T* p = f2 + ((int64_t)h2 * W2 + w2) * C + c0;
                A acc = A(0);
#pragma unroll
                for (int k = 0; k < kChunk; k++)
                    if (k < cn) acc += a1[k] * load_as<A>(p + k);
                s[iy][ix] += acc;
            }
        }
    }
#pragma unroll
    for (int ix = 0; ix < RD; ix++)
#pragma unroll
        for (int iy = 0; iy < RD; iy++) {
            const A v = s[iy][ix] * (1 - dy) * (1 - dx) + s[iy][ix + 1] * (1 - dy) * dx +
                        s[iy + 1][ix] * dy * (1 - dx) + s[iy + 1][ix + 1] * dy * dx;
            o[(ix * RD + iy) * ostride] = (TO)v;
        }
}

// Vector form of altcorr_pixel for C a multiple of kChunk, 16-byte aligned
// rows and levels under 2 GiB.  f1 is read as 16-byte vectors; each tap's
// chunk of f2 as 16-byte loads through a resource based at the level's first
// byte (f2off: the byte offset of this lane's plane), and a tap outside the
// plane gets an offset past the resource's range, which reads zeros without
// a memory access.  FP32 features use FMA chains, FP16 v_dot2_f32_f16; both
// accumulate in FP32.
typedef unsigned int acu32x4 __attribute__((ext_vector_type(4)));
typedef _Float16 acf16x2 __attribute__((ext_vector_type(2)));

template <typename T> struct AltDot;
template <> struct AltDot<float> {
    static constexpr int NV = kChunk / 4;  // 16-byte vectors per chunk
    static __device__ __forceinline__ float dot(const acu32x4 (&a)[NV], const acu32x4 (&b)[NV]) {
        float p0 = 0.f, p1 = 0.f;
#pragma unroll
        for (int q = 0; q < NV; q++) {
            p0 = fmaf(__uint_as_float(a[q][0]), __uint_as_float(b[q][0]), p0);
            p1 = fmaf(__uint_as_float(a[q][1]), __uint_as_float(b[q][1]), p1);
            p0 = fmaf(__uint_as_float(a[q][2]), __uint_as_float(b[q][2]), p0);
            p1 = fmaf(__uint_as_float(a[q][3]), __uint_as_float(b[q][3]), p1);
        }
        return p0 + p1;
    }
};
template <> struct AltDot<__half> {
    static constexpr int NV = kChunk / 8;
    static __device__ __forceinline__ float dot(const acu32x4 (&a)[NV], const acu32x4 (&b)[NV]) {
        float p0 = 0.f, p1 = 0.f;
#pragma unroll
        for (int q = 0; q < NV; q++) {
            // each element goes through a named scalar: a bit_cast of a vector
            // subscript would read element 0 (clang takes the vector's address)
#pragma unroll
            for (int k = 0; k < 4; k += 2) {
                const uint32_t a0 = a[q][k], b0 = b[q][k], a1 = a[q][k + 1], b1 = b[q][k + 1];
                p0 = __builtin_amdgcn_fdot2(__builtin_bit_cast(acf16x2, a0), __builtin_bit_cast(acf16x2, b0), p0, false);
                p1 = __builtin_amdgcn_fdot2(__builtin_bit_cast(acf16x2, a1), __builtin_bit_cast(acf16x2, b1), p1, false);
            }
        }
        return p0 + p1;
    }
};

template <typename T, int R>
__device__ __forceinline__ void altcorr_pixel_vec(const T* __restrict__ f1, __amdgpu_buffer_rsrc_t f2r, uint32_t f2off,
                                                  int H2, int W2, int C, float x, float y, float* __restrict__ o,
                                                  int64_t ostride) {
    constexpr int RD = 2 * R + 1;
    constexpr int NT = RD + 1;
    constexpr int NV = AltDot<T>::NV;
    const float fx0 = floorf(x), fy0 = floorf(y);
    const float dx = x - fx0, dy = y - fy0;
    const int xb = (int)fx0 - R, yb = (int)fy0 - R;
    const uint32_t rowb = (uint32_t)C * (uint32_t)sizeof(T);
    float s[NT][NT];
#pragma unroll
    for (int a = 0; a < NT; a++)
#pragma unroll
        for (int b = 0; b < NT; b++) s[a][b] = 0.f;
    for (int c0 = 0; c0 < C; c0 += kChunk) {
        acu32x4 a1[NV];
        const acu32x4* f1v = reinterpret_cast<const acu32x4*>(f1 + c0);
#pragma unroll
        for (int q = 0; q < NV; q++) a1[q] = f1v[q];
        const uint32_t cb = f2off + (uint32_t)c0 * (uint32_t)sizeof(T);
#pragma unroll
        for (int iy = 0; iy < NT; iy++) {
            const int h2 = yb + iy;
            const bool yok = (unsigned)h2 < (unsigned)H2;
#pragma unroll
            for (int ix = 0; ix < NT; ix++) {
                const int w2 = xb + ix;
                const bool ok = yok && (unsigned)w2 < (unsigned)W2;
                const uint32_t off = ok ? cb + (uint32_t)(h2 * W2 + w2) * rowb : kBufInvalid;
                acu32x4 g[NV];
#pragma unroll
                for (int q = 0; q < NV; q++) g[q] = __builtin_amdgcn_raw_buffer_load_b128(f2r, off + 16u * q, 0, 0);
                s[iy][ix] += AltDot<T>::dot(a1, g);
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

template <typename T, int R, bool VEC>
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
    using A = typename Acc<T>::type;
    A tmp[RD * RD];
    if constexpr (VEC)
        altcorr_pixel_vec<T, R>(fmap1 + (b * HW1 + yx) * C, buf_rsrc(fmap2),
                                (uint32_t)(b * (int64_t)H2 * W2 * C * (int64_t)sizeof(T)), H2, W2, C, x, y, tmp, 1);
    else
        altcorr_pixel<T, R>(fmap1 + (b * HW1 + yx) * C, fmap2 + b * (int64_t)H2 * W2 * C, H2, W2, C, x, y, tmp, 1);
    T* o = out + bn * RD * RD * HW1 + yx;
#pragma unroll
    for (int c = 0; c < RD * RD; c++) o[c * HW1] = store_from<T, A>(tmp[c]);
}

// Indexed multi-level forward, one wave per (b, m, pixel) at one level.  Lane
// (ix, cg) = (lane / 8, lane % 8) takes tap column ix and, in channel step k,
// the 16 bytes at channel 32k + 4 cg (FP32) or 64k + 8 cg (FP16) of each tap
// row: one load instruction reads 8 whole 128-byte lines, one per tap column.
// The 8 channel groups are summed with three DPP steps; the 8 x 8 tap sums go
// through LDS to 49 lanes that form the bilinear outputs.  The window
// geometry is the same for the whole wave.
template <typename T> struct AltStep;
template <> struct AltStep<float> {
    static constexpr int CH = 32;  // channels per step (8 lanes x 4)
    static __device__ __forceinline__ float dot(acu32x4 a, acu32x4 b) {
        float p = __uint_as_float(a[0]) * __uint_as_float(b[0]);
        p = fmaf(__uint_as_float(a[1]), __uint_as_float(b[1]), p);
        p = fmaf(__uint_as_float(a[2]), __uint_as_float(b[2]), p);
        return fmaf(__uint_as_float(a[3]), __uint_as_float(b[3]), p);
    }
};
template <> struct AltStep<__half> {
    static constexpr int CH = 64;  // 8 lanes x 8
    static __device__ __forceinline__ float dot(acu32x4 a, acu32x4 b) {
        float p = 0.f;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const uint32_t x = a[k], y = b[k];  // through named scalars (see AltDot)
            p = __builtin_amdgcn_fdot2(__builtin_bit_cast(acf16x2, x), __builtin_bit_cast(acf16x2, y), p, false);
        }
        return p;
    }
};

template <int CTRL>
__device__ __forceinline__ float alt_dpp(float x) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), CTRL, 0xf, 0xf, true));
}

// One wave's (pixel, level): f1 is the pixel's feature row, f2 the target
// plane [H2, W2, C]; writes the 49 bilinear outputs o[c * ostride] from lanes
// c < 49.  S is this wave's [8][9] LDS slot.
template <typename T, int R, typename TO>
__device__ __forceinline__ void altcorr_wave_pixel(const T* __restrict__ f1, const T* __restrict__ f2, int H2, int W2,
                                                   int C, float x, float y, float (*S)[2 * R + 3], TO* __restrict__ o,
                                                   int64_t ostride) {
    constexpr int RD = 2 * R + 1;
    constexpr int NT = RD + 1;
    static_assert(NT == 8, "8 tap columns x 8 channel groups per wave");
    constexpr int CH = AltStep<T>::CH;
    const int lane = threadIdx.x & 63;
    const int ix = lane >> 3, cg = lane & 7;
    const float fx0 = floorf(x), fy0 = floorf(y);
    const float dx = x - fx0, dy = y - fy0;
    const int xb = (int)fx0 - R, yb = (int)fy0 - R;
    const int w2 = xb + ix;
    const bool xok = (unsigned)w2 < (unsigned)W2;
    const int lc = cg * (CH / 8);  // this lane's first channel of a step
    const T* f1l = f1 + lc;
    const T* f2l = f2 + (int64_t)(xok ? w2 : 0) * C + lc;
    float sv[NT];
#pragma unroll
    for (int iy = 0; iy < NT; iy++) sv[iy] = 0.f;
    for (int c0 = 0; c0 < C; c0 += CH) {
        const acu32x4 a = *reinterpret_cast<const acu32x4*>(f1l + c0);
        acu32x4 g[NT];
#pragma unroll
        for (int iy = 0; iy < NT; iy++) {
            // a tap outside the plane reads row 0 / column 0 of it and is zeroed
            const int h2 = yb + iy;
            const bool yok = (unsigned)h2 < (unsigned)H2;
            const acu32x4 v = *reinterpret_cast<const acu32x4*>(f2l + (int64_t)(yok ? h2 : 0) * W2 * C + c0);
            g[iy] = (xok && yok) ? v : acu32x4{0u, 0u, 0u, 0u};
        }
#pragma unroll
        for (int iy = 0; iy < NT; iy++) sv[iy] += AltStep<T>::dot(a, g[iy]);
    }
    // sum over the 8 channel groups (lanes differing in bits 0..2)
#pragma unroll
    for (int iy = 0; iy < NT; iy++) {
        float v = sv[iy];
        v += alt_dpp<0xB1>(v);   // quad_perm [1,0,3,2]
        v += alt_dpp<0x4E>(v);   // quad_perm [2,3,0,1]
        v += alt_dpp<0x141>(v);  // row_half_mirror
        sv[iy] = v;
    }
    if (cg == 0) {
#pragma unroll
        for (int iy = 0; iy < NT; iy++) S[iy][ix] = sv[iy];
    }
    // a wavefront's LDS accesses complete in order, so its own writes are
    // visible to its reads; the wave barrier keeps the compiler from moving them
    __builtin_amdgcn_wave_barrier();
    if (lane < RD * RD) {
        const int oix = lane / RD, oiy = lane - oix * RD;
        const float v = S[oiy][oix] * (1 - dy) * (1 - dx) + S[oiy][oix + 1] * (1 - dy) * dx +
                        S[oiy + 1][oix] * dy * (1 - dx) + S[oiy + 1][oix + 1] * dy * dx;
        o[lane * ostride] = from_f<TO>(v);
    }
}

template <typename T, int R>
__global__ __launch_bounds__(256) void altcorr_index_wave_kernel(const T* __restrict__ fmap1, const T* __restrict__ f2l,
                                                                 const float* __restrict__ coords,
                                                                 const int64_t* __restrict__ ii,
                                                                 const int64_t* __restrict__ jj, float* __restrict__ out,
                                                                 int64_t B, int Nf, int M, int H, int W, int C, int H2,
                                                                 int W2, int level, int L) {
    constexpr int RD = 2 * R + 1;
    __shared__ float S[4][RD + 1][RD + 2];
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t HW = (int64_t)H * W;
    // XCD-contiguous: neighbouring pixels read overlapping target rows
    const int64_t wid = xcd_logical(blockIdx.x, gridDim.x) * 4 + wv;  // over B*M*HW
    if (wid >= B * M * HW) return;
    const int64_t bm = wid / HW;
    const int64_t yx = wid - bm * HW;
    const int64_t b = bm / M;
    const int m = (int)(bm - b * M);
    const int64_t src = ii[m], dst = jj[m];
    const float scale = (float)(1 << level);
    altcorr_wave_pixel<T, R, float>(fmap1 + ((b * Nf + src) * HW + yx) * C, f2l + (b * Nf + dst) * (int64_t)H2 * W2 * C,
                                    H2, W2, C, coords[wid * 2 + 0] / scale, coords[wid * 2 + 1] / scale, S[wv],
                                    out + ((bm * L + level) * RD * RD) * HW + yx, HW);
}

// CorrLayer forward, one wave per (b, n, pixel): out [B, N, 49, H1, W1].
template <typename T, int R>
__global__ __launch_bounds__(256) void altcorr_fwd_wave_kernel(const T* __restrict__ fmap1, const T* __restrict__ fmap2,
                                                               const float* __restrict__ coords, T* __restrict__ out,
                                                               int64_t B, int N, int H1, int W1, int H2, int W2,
                                                               int C) {
    constexpr int RD = 2 * R + 1;
    __shared__ float S[4][RD + 1][RD + 2];
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t HW1 = (int64_t)H1 * W1;
    // XCD-contiguous: neighbouring pixels read overlapping target rows
    const int64_t wid = xcd_logical(blockIdx.x, gridDim.x) * 4 + wv;  // over B*N*HW1
    if (wid >= B * N * HW1) return;
    const int64_t bn = wid / HW1;
    const int64_t yx = wid - bn * HW1;
    const int64_t b = bn / N;
    altcorr_wave_pixel<T, R, T>(fmap1 + (b * HW1 + yx) * C, fmap2 + b * (int64_t)H2 * W2 * C, H2, W2, C,
                                coords[wid * 2 + 0], coords[wid * 2 + 1], S[wv], out + bn * RD * RD * HW1 + yx, HW1);
}

// Indexed multi-level forward: one lane per (b, m, pixel, level).
template <typename T, int R, bool VEC>
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
    float* o = out + (bm * L * RD * RD + level * RD * RD) * HW + yx;
    if constexpr (VEC)
        altcorr_pixel_vec<T, R>(fmap1 + ((b * Nf + src) * HW + yx) * C, buf_rsrc(f2l),
                                (uint32_t)((b * Nf + dst) * (int64_t)H2 * W2 * C * (int64_t)sizeof(T)), H2, W2, C, x, y,
                                o, HW);
    else
        altcorr_pixel<T, R>(fmap1 + ((b * Nf + src) * HW + yx) * C, f2l + (b * Nf + dst) * (int64_t)H2 * W2 * C, H2,
                            W2, C, x, y, o, HW);
}

// Backward with one wave per (batch b, source pixel): the window geometry and
// tap weights of each of the N coordinate sets are the same for the whole wave
// (scalar loads), and the lanes run along the channels, so a tap's f2 row and
// its gradient row are contiguous accesses.  fmap1_grad of the pixel is summed
// over the N sets and taps in registers and stored once; fmap2_grad takes FP32
// atomic adds, the same scatter as the reference's kernel.
// GATHER: no f2 atomics; instead tap gradient t of (b, n, source pixel) goes to
// taps[((b N + n) H1 W1 + pixel) 64 + t] (t = 8 row + column of the window) and
// the window origin (xb, yb) to org[(b N + n) H1 W1 + pixel], for
// altcorr_bwd_gather_kernel.
template <int R, bool GATHER = false>
__global__ __launch_bounds__(256) void altcorr_bwd_wave_kernel(const float* __restrict__ fmap1,
                                                               const float* __restrict__ fmap2,
                                                               const float* __restrict__ coords,
                                                               const float* __restrict__ gout, float* __restrict__ g1,
                                                               float* __restrict__ g2, int64_t B, int N, int H1,
                                                               int W1, int H2, int W2, int C, bool vec2,
                                                               float* __restrict__ taps = nullptr,
                                                               int2* __restrict__ org = nullptr) {
    constexpr int RD = 2 * R + 1;
    constexpr int NT = RD + 1;
    const int64_t HW1 = (int64_t)H1 * W1;
    // XCD-contiguous: neighbouring pixels read overlapping target rows
    const int64_t wid = xcd_logical(blockIdx.x, gridDim.x) * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (wid >= B * HW1) return;
    const int lane = threadIdx.x & 63;
    const int64_t b = wid / HW1;
    const int64_t yx = wid - b * HW1;
    const float* f1 = fmap1 + (b * HW1 + yx) * C;
    const float* f2b = fmap2 + b * (int64_t)H2 * W2 * C;
    float* g2b = g2 + b * (int64_t)H2 * W2 * C;
    // even C: lanes take channel pairs (one 8-byte load per tap row), 128 channels a pass
    const int cstep = vec2 ? 128 : 64;
    for (int c0 = 0; c0 < C; c0 += cstep) {
        if (vec2) {
            const int c = c0 + 2 * lane;
            const bool cok = c < C;
            const float2 a = cok ? *reinterpret_cast<const float2*>(f1 + c) : make_float2(0.f, 0.f);
            float2 acc = make_float2(0.f, 0.f);
            for (int n = 0; n < N; n++) {
                const int64_t bn = b * N + n;
                const int64_t gid = bn * HW1 + yx;
                const float x = coords[gid * 2 + 0], y = coords[gid * 2 + 1];
                const float fx0 = floorf(x), fy0 = floorf(y);
                const float dx = x - fx0, dy = y - fy0;
                const int xb = (int)fx0 - R, yb = (int)fy0 - R;
                const float* go = gout + bn * RD * RD * HW1 + yx;
                float gt[NT][NT];
#pragma unroll
                for (int i = 0; i < NT; i++)
#pragma unroll
                    for (int j = 0; j < NT; j++) gt[i][j] = 0.f;
#pragma unroll
                for (int ix = 0; ix < RD; ix++)
#pragma unroll
                    for (int iy = 0; iy < RD; iy++) {
                        const float gv = go[(ix * RD + iy) * HW1];
                        gt[iy][ix] += gv * (1 - dy) * (1 - dx);
                        gt[iy][ix + 1] += gv * (1 - dy) * dx;
                        gt[iy + 1][ix] += gv * dy * (1 - dx);
                        gt[iy + 1][ix + 1] += gv * dy * dx;
                    }
#pragma unroll
                for (int iy = 0; iy < NT; iy++) {
                    const int h2 = yb + iy;
                    if (h2 < 0 || h2 >= H2) continue;
#pragma unroll
                    for (int ix = 0; ix < NT; ix++) {
                        const int w2 = xb + ix;
                        if (w2 < 0 || w2 >= W2) continue;
                        if (cok) {
                            const int64_t off = ((int64_t)h2 * W2 + w2) * C + c;
                            const float2 v = *reinterpret_cast<const float2*>(f2b + off);
                            acc.x += gt[iy][ix] * v.x;
                            acc.y += gt[iy][ix] * v.y;
                            if (!GATHER) {
                                atomicAdd(g2b + off, gt[iy][ix] * a.x);
                                atomicAdd(g2b + off + 1, gt[iy][ix] * a.y);
                            }
                        }
                    }
                }
            }
            if (cok) *reinterpret_cast<float2*>(g1 + (b * HW1 + yx) * C + c) = acc;
            continue;
        }
        const int c = c0 + lane;
        const bool cok = c < C;
        const float a = cok ? f1[c] : 0.f;
        float acc = 0.f;
        for (int n = 0; n < N; n++) {
            const int64_t bn = b * N + n;
            const int64_t gid = bn * HW1 + yx;
            const float x = coords[gid * 2 + 0], y = coords[gid * 2 + 1];
            const float fx0 = floorf(x), fy0 = floorf(y);
            const float dx = x - fx0, dy = y - fy0;
            const int xb = (int)fx0 - R, yb = (int)fy0 - R;
            const float* go = gout + bn * RD * RD * HW1 + yx;
            float gt[NT][NT];
#pragma unroll
            for (int i = 0; i < NT; i++)
#pragma unroll
                for (int j = 0; j < NT; j++) gt[i][j] = 0.f;
#pragma unroll
            for (int ix = 0; ix < RD; ix++)
#pragma unroll
                for (int iy = 0; iy < RD; iy++) {
                    const float gv = go[(ix * RD + iy) * HW1];
                    gt[iy][ix] += gv * (1 - dy) * (1 - dx);
                    gt[iy][ix + 1] += gv * (1 - dy) * dx;
                    gt[iy + 1][ix] += gv * dy * (1 - dx);
                    gt[iy + 1][ix + 1] += gv * dy * dx;
                }
#pragma unroll
            for (int iy = 0; iy < NT; iy++) {
                const int h2 = yb + iy;
                if (h2 < 0 || h2 >= H2) continue;
#pragma unroll
                for (int ix = 0; ix < NT; ix++) {
                    const int w2 = xb + ix;
                    if (w2 < 0 || w2 >= W2) continue;
                    if (cok) {
                        const int64_t off = ((int64_t)h2 * W2 + w2) * C + c;
                        acc += gt[iy][ix] * f2b[off];
                        if (!GATHER) atomicAdd(g2b + off, gt[iy][ix] * a);
                    }
                }
            }
        }
        if (cok) g1[(b * HW1 + yx) * C + c] = acc;
    }
    if (GATHER) {
        // lane t: tap (row t / 8, column t % 8) from the (up to) four outputs it feeds
        const int ty = lane >> 3, tx = lane & 7;
        for (int n = 0; n < N; n++) {
            const int64_t bn = b * N + n;
            const int64_t gid = bn * HW1 + yx;
            const float x = coords[gid * 2 + 0], y = coords[gid * 2 + 1];
            const float fx0 = floorf(x), fy0 = floorf(y);
            const float dx = x - fx0, dy = y - fy0;
            const float* go = gout + bn * RD * RD * HW1 + yx;
            float gv = 0.f;
            if (tx < RD && ty < RD) gv += go[(tx * RD + ty) * HW1] * (1 - dy) * (1 - dx);
            if (tx >= 1 && ty < RD) gv += go[((tx - 1) * RD + ty) * HW1] * (1 - dy) * dx;
            if (tx < RD && ty >= 1) gv += go[(tx * RD + ty - 1) * HW1] * dy * (1 - dx);
            if (tx >= 1 && ty >= 1) gv += go[((tx - 1) * RD + ty - 1) * HW1] * dy * dx;
            taps[gid * NT * NT + lane] = gv;
            if (lane == 0) org[gid] = make_int2((int)fx0 - R, (int)fy0 - R);
        }
    }
}

// The fmap1 gradient and the per-window tap gradients of the atomic-free
// backward, four neighbouring source pixels per wave (H1 * W1 a multiple of 4,
// C even).  Lanes take channel pairs, 128 channels a pass.  Lane t first forms
// tap t's gradient of each of the four windows (kept in a register, stored to
// the workspace for the gather); the wave then walks the bounding box of the
// four windows and loads each target pixel's features once for all the windows
// that contain it, broadcasting their tap weights with readlane.  Groups whose
// windows lie far apart (bounding box above kAltGroupArea pixels) walk each
// window separately.
constexpr int kAltGroupArea = 144;

__device__ __forceinline__ float readlane_f32(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

template <int R>
__global__ __launch_bounds__(256) void altcorr_bwd_group_kernel(const float* __restrict__ fmap1,
                                                                const float* __restrict__ fmap2,
                                                                const float* __restrict__ coords,
                                                                const float* __restrict__ gout,
                                                                float* __restrict__ g1, int64_t B, int N, int H1,
                                                                int W1, int H2, int W2, int C,
                                                                float* __restrict__ taps, int2* __restrict__ org) {
    constexpr int RD = 2 * R + 1;
    constexpr int NT = RD + 1;
    static_assert(NT * NT == 64, "one tap per lane");
    constexpr int PG = 4;  // source pixels per wave
    const int64_t HW1 = (int64_t)H1 * W1;
    // each XCD takes a contiguous range of groups: neighbouring groups read
    // overlapping target rows, which then come from that XCD's L2
    const int64_t grp = xcd_logical(blockIdx.x, gridDim.x) * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (grp * PG >= B * HW1) return;
    const int lane = threadIdx.x & 63;
    const int64_t b = grp * PG / HW1;
    const int64_t yx0 = grp * PG - b * HW1;
    const float* f2b = fmap2 + b * (int64_t)H2 * W2 * C;
    const int ty = lane >> 3, tx = lane & 7;
    for (int cb = 0; cb < C; cb += 128) {
        const int c = cb + 2 * lane;
        const bool cok = c < C;
        float2 acc[PG];
#pragma unroll
        for (int p = 0; p < PG; p++) acc[p] = make_float2(0.f, 0.f);
        for (int n = 0; n < N; n++) {
            const int64_t bn = b * N + n;
            int xb[PG], yb[PG];
            float gtv[PG];  // tap (ty, tx) gradient of window p
#pragma unroll
            for (int p = 0; p < PG; p++) {
                const int64_t gid = bn * HW1 + yx0 + p;
                const float x = coords[gid * 2 + 0], y = coords[gid * 2 + 1];
                const float fx0 = floorf(x), fy0 = floorf(y);
                const float dx = x - fx0, dy = y - fy0;
                xb[p] = (int)fx0 - R;
                yb[p] = (int)fy0 - R;
                const float* go = gout + bn * RD * RD * HW1 + yx0 + p;
                float gv = 0.f;
                if (tx < RD && ty < RD) gv += go[(tx * RD + ty) * HW1] * (1 - dy) * (1 - dx);
                if (tx >= 1 && ty < RD) gv += go[((tx - 1) * RD + ty) * HW1] * (1 - dy) * dx;
                if (tx < RD && ty >= 1) gv += go[(tx * RD + ty - 1) * HW1] * dy * (1 - dx);
                if (tx >= 1 && ty >= 1) gv += go[((tx - 1) * RD + ty - 1) * HW1] * dy * dx;
                gtv[p] = gv;
                if (cb == 0) {
                    taps[gid * NT * NT + lane] = gv;
                    if (lane == 0) org[gid] = make_int2(xb[p], yb[p]);
                }
            }
            int ymin = yb[0], ymax = yb[0], xmin = xb[0], xmax = xb[0];
#pragma unroll
            for (int p = 1; p < PG; p++) {
                ymin = min(ymin, yb[p]);
                ymax = max(ymax, yb[p]);
                xmin = min(xmin, xb[p]);
                xmax = max(xmax, xb[p]);
            }
            ymax += NT - 1;
            xmax += NT - 1;
            if ((int64_t)(ymax - ymin + 1) * (xmax - xmin + 1) <= kAltGroupArea) {
                const int y0 = max(ymin, 0), y1 = min(ymax, H2 - 1), x0 = max(xmin, 0), x1 = min(xmax, W2 - 1);
                for (int y = y0; y <= y1; y++)
                    for (int x = x0; x <= x1; x++) {
                        const float2 v = cok ? *reinterpret_cast<const float2*>(f2b + ((int64_t)y * W2 + x) * C + c)
                                             : make_float2(0.f, 0.f);
#pragma unroll
                        for (int p = 0; p < PG; p++) {
                            const int ry = y - yb[p], rx = x - xb[p];
                            if ((unsigned)ry < (unsigned)NT && (unsigned)rx < (unsigned)NT) {
                                const float w = readlane_f32(gtv[p], ry * NT + rx);
                                acc[p].x += w * v.x;
                                acc[p].y += w * v.y;
                            }
                        }
                    }
            } else {
#pragma unroll
                for (int p = 0; p < PG; p++)
                    for (int ry = 0; ry < NT; ry++) {
                        const int y = yb[p] + ry;
                        if (y < 0 || y >= H2) continue;
                        for (int rx = 0; rx < NT; rx++) {
                            const int x = xb[p] + rx;
                            if (x < 0 || x >= W2) continue;
                            const float2 v = cok ? *reinterpret_cast<const float2*>(f2b + ((int64_t)y * W2 + x) * C + c)
                                                 : make_float2(0.f, 0.f);
                            const float w = readlane_f32(gtv[p], ry * NT + rx);
                            acc[p].x += w * v.x;
                            acc[p].y += w * v.y;
                        }
                    }
            }
        }
        if (cok) {
#pragma unroll
            for (int p = 0; p < PG; p++) *reinterpret_cast<float2*>(g1 + (b * HW1 + yx0 + p) * C + c) = acc[p];
        }
    }
}

// fmap2 gradient without atomics: one workgroup per target row (b, y2).  The
// windows (n, source pixel) of batch b are scanned in ascending order, those
// that cover row y2 are compacted (order kept), and kAltWin at a time their
// tap gradients of this row and their source features are staged in LDS.
// Element (x2, channel group cg) of the row is owned by the thread of slot
// (x2 mod 8) * C/4 + cg, which adds each window's term in window order: the
// sums need no atomics and come out the same on every run.  The row
// accumulates in LDS and is stored once (zeros where no window reaches).
constexpr int kAltWin = 64;  // windows staged per pass

template <int R>
__global__ __launch_bounds__(256) void altcorr_bwd_gather_kernel(const float* __restrict__ fmap1,
                                                                 const float* __restrict__ taps,
                                                                 const int2* __restrict__ org, float* __restrict__ g2,
                                                                 int64_t B, int N, int H1, int W1, int H2, int W2,
                                                                 int C, int nwin) {
    constexpr int NT = 2 * R + 2;
    static_assert((NT & (NT - 1)) == 0, "window width a power of two (column phases)");
    extern __shared__ __align__(16) float gdyn[];
    const int C4 = C / 4;
    float4* acc = reinterpret_cast<float4*>(gdyn);              // [W2][C4] the row's gradient
    float4* sF = acc + (int64_t)W2 * C4;                         // [nwin][C4] source features
    float* sG = reinterpret_cast<float*>(sF + (int64_t)nwin * C4);  // [nwin][8] tap gradients of the row
    int* sX = reinterpret_cast<int*>(sG + nwin * NT);            // [nwin] window column origin
    int* sP = sX + nwin;                                         // [nwin] source pixel
    __shared__ int wl[256];
    __shared__ int wsum[4];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    // each XCD takes a contiguous range of rows: neighbouring rows stage the
    // same windows, whose features then come from that XCD's L2
    const int64_t blk = xcd_logical(blockIdx.x, gridDim.x);
    const int64_t b = blk / H2;
    const int y2 = (int)(blk - b * H2);
    const int64_t HW1 = (int64_t)H1 * W1;
    const int64_t nw = (int64_t)N * HW1;  // windows of batch b
    for (int i = tid; i < W2 * C4; i += 256) acc[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    for (int64_t s0 = 0; s0 < nw; s0 += 256) {
        // compact the covering windows of this scan chunk, in order
        const int64_t w = s0 + tid;
        bool cov = false;
        if (w < nw) {
            const int yb = org[b * nw + w].y;
            cov = y2 >= yb && y2 < yb + NT;
        }
        const uint64_t bal = __ballot(cov);
        const int before = __popcll(bal & ((1ull << lane) - 1ull));
        if (lane == 0) wsum[wv] = __popcll(bal);
        __syncthreads();
        int base = 0, total = 0;
        for (int q = 0; q < 4; q++) {
            base += q < wv ? wsum[q] : 0;
            total += wsum[q];
        }
        if (cov) wl[base + before] = tid;
        __syncthreads();
        for (int sub = 0; sub < total; sub += nwin) {
            const int cnt = min(nwin, total - sub);
            for (int j = tid; j < cnt; j += 256) {
                const int64_t id = s0 + wl[sub + j];  // n * HW1 + pixel (below 2^31)
                const int2 o = org[b * nw + id];
                sX[j] = o.x;
                sP[j] = (int)((uint32_t)id % (uint32_t)HW1);
                const float* tp = taps + (b * nw + id) * NT * NT + (y2 - o.y) * NT;
#pragma unroll
                for (int t = 0; t < NT; t++) sG[j * NT + t] = tp[t];
            }
            for (int i = tid; i < cnt * C4; i += 256) {
                const int j = i / C4, cg = i - j * C4;
                const uint32_t pix = (uint32_t)(s0 + wl[sub + j]) % (uint32_t)HW1;
                sF[i] = *reinterpret_cast<const float4*>(fmap1 + (b * HW1 + pix) * C + 4 * cg);
            }
            __syncthreads();
            for (int sl = tid; sl < NT * C4; sl += 256) {
                const int ph = sl / C4, cg = sl - ph * C4;
                for (int j = 0; j < cnt; j++) {
                    const int xb = sX[j];
                    const int tx = (ph - xb) & (NT - 1);  // the window column on this slot's phase
                    const int x2 = xb + tx;
                    if (x2 < 0 || x2 >= W2) continue;
                    const float gv = sG[j * NT + tx];
                    const float4 f = sF[j * C4 + cg];
                    float4 a = acc[x2 * C4 + cg];
                    a.x += gv * f.x;
                    a.y += gv * f.y;
                    a.z += gv * f.z;
                    a.w += gv * f.w;
                    acc[x2 * C4 + cg] = a;
                }
            }
            __syncthreads();
        }
    }
    float4* out = reinterpret_cast<float4*>(g2 + ((b * H2 + y2) * (int64_t)W2) * C);
    for (int i = tid; i < W2 * C4; i += 256) out[i] = acc[i];
}

// LDS bytes of altcorr_bwd_gather_kernel with nwin windows per pass
inline size_t altcorr_gather_lds(int W2, int C, int nwin) {
    return (size_t)W2 * C * 4 + (size_t)nwin * C * 4 + (size_t)nwin * (8 + 2) * 4;
}

// ---------------------------------------------------------------------------
// All-pairs volume and pyramid (CorrBlock, droid_net.py:56-69, 94-102).
//
// Level 0 of edge e is O[m][n] = sum_c f1[c][m] f2[c][n] / 16 over the HW
// source pixels m and HW target pixels n; level l > 0 is the 2x2 mean of
// level l-1 over the target dimensions (floor of odd sizes), each level
// rounded to the output type before the next is pooled, as avg_pool2d over
// a stored pyramid does.
//
//   1. corr_to_cl_kernel       features -> channels-last [E][HW][Cp] in the
//                              contraction type (zero channel padding).  For
//                              FP8 each channel is first balanced between the
//                              two maps (f1 / a_c, f2 * a_c, a_c = sqrt of the
//                              ratio of their per-edge amax), then each pixel
//                              row gets its own scale amax/448 (e4m3).
//   2. corr_build_kernel       one workgroup per (edge, 128 source pixels, band
//                              of 8 target rows); it walks the band's 8 x 16
//                              target patches left to right, so that whole
//                              output row segments come from one CU.  Each
//                              patch is a 128 x 128 tile: 4 waves x 2 x 2 tiles
//                              of 32 x 32 on MFMA (v_mfma_f32_32x32x16_{f16,
//                              bf16,fp8_fp8} or 32x32x2f32) with operands staged
//                              in LDS 128 channels at a time (A once per band
//                              when C <= 128).  Target pixels are the MFMA rows,
//                              ordered so that a lane's four accumulator rows
//                              are four adjacent columns (one vector LDS write).
//                              The epilogue keeps the level-0 tile in LDS and
//                              writes level 0 and the pooled levels 1-3 of the
//                              patch directly: the pyramid costs no second pass
//                              over level 0.
// ---------------------------------------------------------------------------
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr int kBM = 128;            // source pixels per tile
constexpr int kPY = 8, kPX = 16;    // target patch per tile
constexpr int kBN = kPY * kPX;      // 128 target pixels
constexpr int kKC = 128;            // channels staged per pass
constexpr int kOS = kBN + 4;        // FP32 tile row stride in LDS
constexpr float kFp8Max = 448.f;    // e4m3 (OCP) largest finite value

// Contraction types: element type in LDS, K per MFMA issue (KI), elements per
// lane per issue (KI / 2 lanes groups).
struct MmaF32 {
    using T = float;
    static constexpr int KI = 2;
    __device__ static f32x16 mma(const T* a, const T* b, f32x16 c) {
        return __builtin_amdgcn_mfma_f32_32x32x2f32(a[0], b[0], c, 0, 0, 0);
    }
};
struct MmaF16 {
    using T = _Float16;
    static constexpr int KI = 16;
    __device__ static f32x16 mma(const T* a, const T* b, f32x16 c) {
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(*reinterpret_cast<const f16x8*>(a),
                                                       *reinterpret_cast<const f16x8*>(b), c, 0, 0, 0);
    }
};
struct MmaBF16 {
    using T = __bf16;
    static constexpr int KI = 16;
    __device__ static f32x16 mma(const T* a, const T* b, f32x16 c) {
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(*reinterpret_cast<const bf16x8*>(a),
                                                        *reinterpret_cast<const bf16x8*>(b), c, 0, 0, 0);
    }
};
struct MmaFP8 {
    using T = unsigned char;  // e4m3
    static constexpr int KI = 16;
    __device__ static f32x16 mma(const T* a, const T* b, f32x16 c) {
        return __builtin_amdgcn_mfma_f32_32x32x16_fp8_fp8(*reinterpret_cast<const long*>(a),
                                                           *reinterpret_cast<const long*>(b), c, 0, 0, 0);
    }
};

template <typename Mma>
__device__ __forceinline__ typename Mma::T to_mma(float v) { return (typename Mma::T)v; }

// Per-(edge, channel) amax of both maps over the pixels (FP8 balancing).
template <typename Ts>
__global__ __launch_bounds__(256) void corr_chan_amax_kernel(const Ts* __restrict__ f1, const Ts* __restrict__ f2,
                                                             float* __restrict__ am1, float* __restrict__ am2,
                                                             int HW) {
    __shared__ float red[2][4];
    const int64_t ec = blockIdx.x;
    const Ts* p1 = f1 + ec * HW;
    const Ts* p2 = f2 + ec * HW;
    float a = 0.f, b = 0.f;
    for (int i = threadIdx.x; i < HW; i += 256) {
        a = fmaxf(a, fabsf(to_f(p1[i])));
        b = fmaxf(b, fabsf(to_f(p2[i])));
    }
    for (int o = 32; o >= 1; o >>= 1) {
        a = fmaxf(a, __shfl_xor(a, o, 64));
        b = fmaxf(b, __shfl_xor(b, o, 64));
    }
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = a;
        red[1][threadIdx.x >> 6] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        am1[ec] = fmaxf(fmaxf(red[0][0], red[0][1]), fmaxf(red[0][2], red[0][3]));
        am2[ec] = fmaxf(fmaxf(red[1][0], red[1][1]), fmaxf(red[1][2], red[1][3]));
    }
}

// Channels-last copy of one map in the contraction type: 32 pixels per
// workgroup, [C][32] staged in LDS as FP32.  side = 0 for fmap1 (divide by the
// balance factor), 1 for fmap2 (multiply); FP8 only.
template <typename Ts, typename Mma, bool FP8>
__global__ __launch_bounds__(256) void corr_to_cl_kernel(const Ts* __restrict__ f, typename Mma::T* __restrict__ out,
                                                         const float* __restrict__ am1, const float* __restrict__ am2,
                                                         float* __restrict__ pscale, int C, int Cp, int HW, int side) {
    using T = typename Mma::T;
    extern __shared__ float tile[];  // [Cp][33]
    __shared__ float sc[32];
    const int nblk = (HW + 31) / 32;
    const int e = blockIdx.x / nblk;
    const int p0 = (blockIdx.x % nblk) * 32;
    const int tid = threadIdx.x;
    const Ts* fe = f + (int64_t)e * C * HW;
    for (int i = tid; i < Cp * 32; i += 256) {
        const int c = i >> 5, px = i & 31;
        float v = 0.f;
        if (c < C && p0 + px < HW) {
            v = to_f(fe[(int64_t)c * HW + p0 + px]);
            if (FP8) {
                const float a1 = am1[(int64_t)e * C + c], a2 = am2[(int64_t)e * C + c];
                const float bal = (a1 > 0.f && a2 > 0.f) ? sqrtf(a1 / a2) : 1.f;
                v = side == 0 ? v / bal : v * bal;
            }
        }
        tile[c * 33 + px] = v;
    }
    __syncthreads();
    if (FP8) {
        // per-pixel amax over channels: 8 lanes per pixel
        const int px = tid >> 3, part = tid & 7;
        float a = 0.f;
        for (int c = part; c < Cp; c += 8) a = fmaxf(a, fabsf(tile[c * 33 + px]));
        for (int o = 4; o >= 1; o >>= 1) a = fmaxf(a, __shfl_xor(a, o, 64));
        if (part == 0) {
            const float s = a > 0.f ? a / kFp8Max : 1.f;
            sc[px] = s;
            if (p0 + px < HW) pscale[(int64_t)e * HW + p0 + px] = s;
        }
        __syncthreads();
        // four channels per lane, packed into one dword
        for (int i = tid; i < 32 * (Cp / 4); i += 256) {
            const int px = i / (Cp / 4), c4 = (i % (Cp / 4)) * 4;
            if (p0 + px >= HW) continue;
            const float inv = 1.f / sc[px];
            int pk = __builtin_amdgcn_cvt_pk_fp8_f32(tile[c4 * 33 + px] * inv, tile[(c4 + 1) * 33 + px] * inv, 0, false);
            pk = __builtin_amdgcn_cvt_pk_fp8_f32(tile[(c4 + 2) * 33 + px] * inv, tile[(c4 + 3) * 33 + px] * inv, pk, true);
            *reinterpret_cast<int*>(out + ((int64_t)e * HW + p0 + px) * Cp + c4) = pk;
        }
    } else {
        for (int i = tid; i < 32 * Cp; i += 256) {
            const int px = i / Cp, c = i % Cp;
            if (p0 + px < HW) out[((int64_t)e * HW + p0 + px) * Cp + c] = to_mma<Mma>(tile[c * 33 + px]);
        }
    }
}

template <typename Tout>
__device__ __forceinline__ float round_to(float v) { return to_f(from_f<Tout>(v)); }

// Four consecutive LDS values as float (one 8- or 16-byte read).
__device__ __forceinline__ void load4(const float* p, float* v) {
    const float4 t = *reinterpret_cast<const float4*>(p);
    v[0] = t.x; v[1] = t.y; v[2] = t.z; v[3] = t.w;
}
__device__ __forceinline__ void load4(const __half* p, float* v) {
    const uint2 t = *reinterpret_cast<const uint2*>(p);
    const float2 a = __half22float2(*reinterpret_cast<const __half2*>(&t.x));
    const float2 b = __half22float2(*reinterpret_cast<const __half2*>(&t.y));
    v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y;
}

// Store n <= V consecutive values; one vector store when all V are valid and
// the row layout keeps the vector aligned (vec), element stores otherwise.
template <typename Tout, int V>
__device__ __forceinline__ void store_run(Tout* dst, const float (&v)[V], int n, bool vec);
template <>
__device__ __forceinline__ void store_run<float, 4>(float* dst, const float (&v)[4], int n, bool vec) {
    if (vec && n == 4) {
        *reinterpret_cast<float4*>(dst) = make_float4(v[0], v[1], v[2], v[3]);
    } else {
        for (int i = 0; i < n; i++) dst[i] = v[i];
    }
}
template <>
__device__ __forceinline__ void store_run<float, 2>(float* dst, const float (&v)[2], int n, bool vec) {
    if (vec && n == 2) {
        *reinterpret_cast<float2*>(dst) = make_float2(v[0], v[1]);
    } else {
        for (int i = 0; i < n; i++) dst[i] = v[i];
    }
}
template <>
__device__ __forceinline__ void store_run<__half, 4>(__half* dst, const float (&v)[4], int n, bool vec) {
    if (vec && n == 4) {
        const __half2 a = __floats2half2_rn(v[0], v[1]), b = __floats2half2_rn(v[2], v[3]);
        uint2 u;
        u.x = *reinterpret_cast<const unsigned int*>(&a);
        u.y = *reinterpret_cast<const unsigned int*>(&b);
        *reinterpret_cast<uint2*>(dst) = u;
    } else {
        for (int i = 0; i < n; i++) dst[i] = __float2half(v[i]);
    }
}
template <>
__device__ __forceinline__ void store_run<__half, 2>(__half* dst, const float (&v)[2], int n, bool vec) {
    if (vec && n == 2) {
        *reinterpret_cast<__half2*>(dst) = __floats2half2_rn(v[0], v[1]);
    } else {
        for (int i = 0; i < n; i++) dst[i] = __float2half(v[i]);
    }
}

// A whole FP16 tile row (8 elements, 16 bytes; the tile's padding columns
// hold zeros in LDS as well).  Other output types never take this path.
template <typename Tout>
__device__ __forceinline__ void store_run8(Tout* dst, const float (&v)[8]) {
    if constexpr (sizeof(Tout) == 2) {
        uint4 u;
        unsigned int w[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const Tout a = from_f<Tout>(v[2 * k]), b = from_f<Tout>(v[2 * k + 1]);
            w[k] = (unsigned int)__builtin_bit_cast(unsigned short, a) |
                   ((unsigned int)__builtin_bit_cast(unsigned short, b) << 16);
        }
        u.x = w[0];
        u.y = w[1];
        u.z = w[2];
        u.w = w[3];
        *reinterpret_cast<uint4*>(dst) = u;
    }
}

// Workgroup barrier that orders LDS only.  The build kernel never reads back
// what it writes to global memory, so its stores need not drain at each
// barrier (__syncthreads would wait for them).
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// Target row r (0..127) of a patch is pixel (patch_y(r), patch_x(r)) of its
// 8 x 16 block: r = 64 X + 32 Y + 8 a + c covers rows 4 Y + a and columns
// 8 X + c, so that a lane's four consecutive accumulator rows of an MFMA
// sub-tile are four adjacent columns of one row.
__device__ __forceinline__ int patch_y(int r) { return 4 * ((r >> 5) & 1) + ((r >> 3) & 3); }
__device__ __forceinline__ int patch_x(int r) { return 8 * (r >> 6) + (r & 7); }

// Four consecutive values into LDS in the output type (one 8- or 16-byte write).
__device__ __forceinline__ void lds_store4(float* p, const float (&v)[4]) {
    *reinterpret_cast<float4*>(p) = make_float4(v[0], v[1], v[2], v[3]);
}
__device__ __forceinline__ void lds_store4(__half* p, const float (&v)[4]) {
    const __half2 a = __floats2half2_rn(v[0], v[1]), b = __floats2half2_rn(v[2], v[3]);
    uint2 u;
    u.x = *reinterpret_cast<const unsigned int*>(&a);
    u.y = *reinterpret_cast<const unsigned int*>(&b);
    *reinterpret_cast<uint2*>(p) = u;
}

// Eight consecutive LDS values (16-byte aligned) as float, and back.
__device__ __forceinline__ void lds_load8(const float* p, float (&v)[8]) {
    const float4 a = reinterpret_cast<const float4*>(p)[0], b = reinterpret_cast<const float4*>(p)[1];
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
    v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
}
__device__ __forceinline__ void lds_load8(const __half* p, float (&v)[8]) {
    const uint4 u = *reinterpret_cast<const uint4*>(p);
    const unsigned int w[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const float2 f = __half22float2(*reinterpret_cast<const __half2*>(&w[k]));
        v[2 * k] = f.x;
        v[2 * k + 1] = f.y;
    }
}
__device__ __forceinline__ void lds_store8(float* p, const float (&v)[8]) {
    reinterpret_cast<float4*>(p)[0] = make_float4(v[0], v[1], v[2], v[3]);
    reinterpret_cast<float4*>(p)[1] = make_float4(v[4], v[5], v[6], v[7]);
}
__device__ __forceinline__ void lds_store8(__half* p, const float (&v)[8]) {
    unsigned int w[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const __half2 h = __floats2half2_rn(v[2 * k], v[2 * k + 1]);
        w[k] = *reinterpret_cast<const unsigned int*>(&h);
    }
    *reinterpret_cast<uint4*>(p) = make_uint4(w[0], w[1], w[2], w[3]);
}

struct LevelGeoms {
    LevelGeom g[4];
};

// Element offset of (y, x) of source plane pm in a level of geometry g, with
// the tile width TW (elements of a 16-byte tile row of the output type) known
// at compile time, so the tile column and offsets are shifts and masks.
template <int TW>
__device__ __forceinline__ int64_t level_off(const LevelGeom& g, int64_t pm, int y, int x) {
    constexpr int LT = TW == 2 ? 1 : TW == 4 ? 2 : 3;
    static_assert(TW == (1 << LT), "tile width");
    if (!g.tiled) return pm * g.plane + y * g.Wl + x;
    return tiled_src(pm, g.plane, TW) + tiled_yx(y, 0, g.Wp >> LT, TW) + ((x >> LT) * (kTilePieces * TW) + (x & (TW - 1)));
}

template <typename Mma, typename Tout, bool FP8>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void corr_build_kernel(const typename Mma::T* __restrict__ A,
                                                         const typename Mma::T* __restrict__ B,
                                                         const float* __restrict__ s1, const float* __restrict__ s2,
                                                         Tout* __restrict__ o0, Tout* __restrict__ o1,
                                                         Tout* __restrict__ o2, Tout* __restrict__ o3, int L, int Cp,
                                                         int H, int W, LevelGeoms lg) {
    using T = typename Mma::T;
    constexpr int kTW = kPieceBytes / (int)sizeof(Tout);  // tile width of the tiled layout
    constexpr int RS = kKC + 16 / (int)sizeof(T);  // operand row stride (elements): 16 B of padding
    constexpr int AB = kBM * RS * (int)sizeof(T);  // bytes of one staged operand
    constexpr int OB = kBM * kOS * (int)sizeof(Tout);
    __shared__ __align__(16) unsigned char smem[AB + (AB > OB ? AB : OB)];
    __shared__ __align__(16) Tout P1[kBM][kPY / 2 * kPX / 2];  // pooled level 1 of the patch
    __shared__ __align__(16) Tout P2[kBM][kPY / 4 * kPX / 4];  // pooled level 2
    __shared__ float ss1[kBM], ss2[kBN];
    T* As = reinterpret_cast<T*>(smem);
    T* Bs = reinterpret_cast<T*>(smem + AB);
    Tout* Os = reinterpret_cast<Tout*>(smem + AB);  // level-0 tile in the output type, over Bs
    const int HW = H * W;
    const int npx = (W + kPX - 1) / kPX, npy = (H + kPY - 1) / kPY;
    const int nm = (HW + kBM - 1) / kBM;
    int64_t bid = blockIdx.x;
    const int ty = (int)(bid % npy);
    bid /= npy;
    const int mt = (int)(bid % nm);
    const int e = (int)(bid / nm);
    const int y0 = ty * kPY;
    const int m0 = mt * kBM;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int wm = (wv >> 1) * 64, wn = (wv & 1) * 64;
    const T* Ae = A + (int64_t)e * HW * Cp;
    const T* Be = B + (int64_t)e * HW * Cp;
    const int64_t ob = (int64_t)e * HW;
    constexpr int CH = 16 / (int)sizeof(T);  // elements per 16-byte chunk
    constexpr int CPR = kKC / CH;             // chunks per full row
    constexpr int NQ = kBM * CPR / 256;       // chunks per lane and operand
    constexpr int PER = Mma::KI / 2;
    if (FP8 && tid < kBM) ss1[tid] = (m0 + tid < HW) ? s1[(int64_t)e * HW + m0 + tid] : 0.f;
    const bool single_pass = Cp <= kKC;
    // Stage rows [0, kBM) of one operand: rows are source pixels (A) or the
    // patch's target pixels (B).  Loads go to registers first (all of a group
    // in flight), then to LDS; B for the next patch is loaded before the
    // current patch's epilogue so that its latency hides behind the stores.
    auto load_rows = [&](uint4 (&v)[NQ], bool isA, int x0, int c0, int cpr) {
#pragma unroll
        for (int u = 0; u < NQ; u++) {
            const int i = tid + u * 256;
            const int r = i / CPR, q = i % CPR;
            v[u] = make_uint4(0, 0, 0, 0);
            if (q < cpr) {
                if (isA) {
                    if (m0 + r < HW) v[u] = *reinterpret_cast<const uint4*>(Ae + (int64_t)(m0 + r) * Cp + c0 + q * CH);
                } else {
                    const int y = y0 + patch_y(r), x = x0 + patch_x(r);
                    if (y < H && x < W)
                        v[u] = *reinterpret_cast<const uint4*>(Be + (int64_t)(y * W + x) * Cp + c0 + q * CH);
                }
            }
        }
    };
    auto store_rows = [&](T* dst, const uint4 (&v)[NQ]) {
#pragma unroll
        for (int u = 0; u < NQ; u++) {
            const int i = tid + u * 256;
            *reinterpret_cast<uint4*>(dst + (i / CPR) * RS + (i % CPR) * CH) = v[u];
        }
    };
    uint4 breg[NQ];
    if (single_pass) {
        load_rows(breg, true, 0, 0, Cp / CH);
        store_rows(As, breg);
        load_rows(breg, false, 0, 0, Cp / CH);
    }
    for (int tx = 0; tx < npx; tx++) {
        const int x0 = tx * kPX;
        if (FP8 && tid < kBN) {
            const int y = y0 + tid / kPX, x = x0 + tid % kPX;
            ss2[tid] = (y < H && x < W) ? s2[(int64_t)e * HW + y * W + x] : 0.f;
        }
        f32x16 acc[2][2];
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++)
#pragma unroll
                for (int v = 0; v < 16; v++) acc[i][j][v] = 0.f;
        for (int c0 = 0; c0 < Cp; c0 += kKC) {
            const int kc = min(kKC, Cp - c0);
            if (!single_pass) {
                load_rows(breg, true, 0, c0, kc / CH);
                store_rows(As, breg);
                load_rows(breg, false, x0, c0, kc / CH);
            }
            store_rows(Bs, breg);  // single pass: loaded during the previous patch
            lds_barrier();
            // acc[i][j]: target rows wn + 32 i + (0..31) (MFMA rows) x source rows wm + 32 j + (0..31)
            for (int kk = 0; kk < kc; kk += Mma::KI) {
                const int k = kk + PER * (lane >> 5);
                const T* b0 = Bs + (wn + (lane & 31)) * RS + k;
                const T* b1 = b0 + 32 * RS;
                const T* a0 = As + (wm + (lane & 31)) * RS + k;
                const T* a1 = a0 + 32 * RS;
                acc[0][0] = Mma::mma(b0, a0, acc[0][0]);
                acc[0][1] = Mma::mma(b0, a1, acc[0][1]);
                acc[1][0] = Mma::mma(b1, a0, acc[1][0]);
                acc[1][1] = Mma::mma(b1, a1, acc[1][1]);
            }
            lds_barrier();
        }
        if (single_pass && tx + 1 < npx) load_rows(breg, false, x0 + kPX, 0, Cp / CH);
        // level-0 tile, rounded to the output type, over Bs: for source sub-tile j
        // the lane holds source row wm + 32 j + lane % 32 and, in acc[i][j][4 a + c],
        // target row 4 i + a, column wn / 8 + 4 (lane / 32) + c of the patch, so
        // each four values are one vector write
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++) {
                const int m = wm + 32 * j + (lane & 31);
#pragma unroll
                for (int a4 = 0; a4 < 4; a4++) {
                    const int n = (4 * i + a4) * kPX + (wn >> 3) + 4 * (lane >> 5);
                    float v4[4];
#pragma unroll
                    for (int c = 0; c < 4; c++) {
                        v4[c] = acc[i][j][4 * a4 + c] * 0.0625f;
                        if (FP8) v4[c] *= ss1[m] * ss2[n + c];
                    }
                    lds_store4(Os + m * kOS + n, v4);
                }
            }
        lds_barrier();
        const bool vec0 = lg.g[0].tiled || (W % 4) == 0;
        if (sizeof(Tout) == 2 && lg.g[0].tiled) {
            // level 0, FP16 tiles: 8 target columns (one 16-byte tile row) per store
#pragma unroll 4
            for (int q = 0; q < kBM * kBN / 8 / 256; q++) {
                const int idx = tid + q * 256;
                const int px = (idx & 1) * 8, py = (idx >> 1) & (kPY - 1), ml = idx >> 4;
                const int m = m0 + ml, y = y0 + py, x = x0 + px;
                if (m >= HW || y >= H || x >= W) continue;
                // the rounded FP16 values are copied as they are (two 8-byte LDS reads)
                const uint2 a = *reinterpret_cast<const uint2*>(Os + ml * kOS + py * kPX + px);
                const uint2 b = *reinterpret_cast<const uint2*>(Os + ml * kOS + py * kPX + px + 4);
                *reinterpret_cast<uint4*>(o0 + level_off<kTW>(lg.g[0], ob + m, y, x)) =
                    make_uint4(a.x, a.y, b.x, b.y);
            }
        } else {
            // level 0: 4 target columns per store
#pragma unroll 4
            for (int q = 0; q < kBM * kBN / 4 / 256; q++) {
                const int idx = tid + q * 256;
                const int px = (idx & 3) * 4, py = (idx >> 2) & (kPY - 1), ml = idx >> 5;
                const int m = m0 + ml, y = y0 + py, x = x0 + px;
                if (m >= HW || y >= H || x >= W) continue;
                float vv[4];
                load4(Os + ml * kOS + py * kPX + px, vv);
                store_run<Tout, 4>(o0 + level_off<kTW>(lg.g[0], ob + m, y, x), vv, min(4, W - x), vec0);
            }
        }
        if (L >= 2 && sizeof(Tout) == 2 && lg.g[1].tiled) {
            // level 1 of the patch (4 x 8), FP16 tiles: one whole tile row (8
            // columns, 16 bytes) per store, from rounded level 0
            const int Hl = H / 2, Wl = W / 2;
#pragma unroll
            for (int q = 0; q < kBM * 4 / 256; q++) {
                const int idx = tid + q * 256;
                const int py = idx & 3, ml = idx >> 2;
                float r0[16], r1[16];
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    load4(Os + ml * kOS + (2 * py) * kPX + 4 * k, r0 + 4 * k);
                    load4(Os + ml * kOS + (2 * py + 1) * kPX + 4 * k, r1 + 4 * k);
                }
                float vv[8];
#pragma unroll
                for (int u = 0; u < 8; u++)
                    vv[u] = round_to<Tout>((((r0[2 * u] + r0[2 * u + 1]) + r1[2 * u]) + r1[2 * u + 1]) * 0.25f);
                lds_store8(&P1[ml][py * 8], vv);
                const int m = m0 + ml, y = y0 / 2 + py, x = x0 / 2;
                // columns past Wl are the tile's zero padding (with W odd, column
                // Wl would pool the last level-0 column)
#pragma unroll
                for (int u = 0; u < 8; u++)
                    if (x + u >= Wl) vv[u] = 0.f;
                if (m < HW && y < Hl && x < Wl) store_run8(o1 + level_off<kTW>(lg.g[1], ob + m, y, x), vv);
            }
        } else if (L >= 2) {
            // level 1 of the patch (4 x 8), once each, from rounded level 0
            const int Hl = H / 2, Wl = W / 2;
            const bool vec = lg.g[1].tiled || (Wl % 4) == 0;
#pragma unroll
            for (int q = 0; q < kBM * 8 / 256; q++) {
                const int idx = tid + q * 256;
                const int hx = idx & 1, py = (idx >> 1) & 3, ml = idx >> 3;  // four L1 columns 4 hx ..
                float r0[8], r1[8];
                load4(Os + ml * kOS + (2 * py) * kPX + 8 * hx, r0);
                load4(Os + ml * kOS + (2 * py) * kPX + 8 * hx + 4, r0 + 4);
                load4(Os + ml * kOS + (2 * py + 1) * kPX + 8 * hx, r1);
                load4(Os + ml * kOS + (2 * py + 1) * kPX + 8 * hx + 4, r1 + 4);
                float vv[4];
#pragma unroll
                for (int u = 0; u < 4; u++)
                    vv[u] = round_to<Tout>((((r0[2 * u] + r0[2 * u + 1]) + r1[2 * u]) + r1[2 * u + 1]) * 0.25f);
                lds_store4(&P1[ml][py * 8 + 4 * hx], vv);
                const int m = m0 + ml, y = y0 / 2 + py, x = x0 / 2 + 4 * hx;
                if (m < HW && y < Hl && x < Wl) store_run<Tout, 4>(o1 + level_off<kTW>(lg.g[1], ob + m, y, x), vv, min(4, Wl - x), vec);
            }
        }
        if (L >= 3) {
            lds_barrier();
            // level 2 (2 x 4) from level 1
            const int Hl = H / 4, Wl = W / 4;
            const bool vec = lg.g[2].tiled || (Wl % 4) == 0;
            const int py = tid & 1, ml = tid >> 1;
            float r0[8], r1[8];
            lds_load8(&P1[ml][(2 * py) * 8], r0);
            lds_load8(&P1[ml][(2 * py + 1) * 8], r1);
            float vv[4];
#pragma unroll
            for (int u = 0; u < 4; u++)
                vv[u] = round_to<Tout>((((r0[2 * u] + r0[2 * u + 1]) + r1[2 * u]) + r1[2 * u + 1]) * 0.25f);
            lds_store4(&P2[ml][py * 4], vv);
            const int m = m0 + ml, y = y0 / 4 + py, x = x0 / 4;
            if (m < HW && y < Hl && x < Wl) store_run<Tout, 4>(o2 + level_off<kTW>(lg.g[2], ob + m, y, x), vv, min(4, Wl - x), vec);
        }
        if (L >= 4) {
            lds_barrier();
            // level 3 (1 x 2) from level 2
            if (tid < kBM) {
                const int Hl = H / 8, Wl = W / 8;
                const bool vec = lg.g[3].tiled || (Wl % 2) == 0;
                float p2[8], vv[2];
                lds_load8(&P2[tid][0], p2);
#pragma unroll
                for (int u = 0; u < 2; u++)
                    vv[u] = round_to<Tout>((((p2[2 * u] + p2[2 * u + 1]) + p2[4 + 2 * u]) + p2[4 + 2 * u + 1]) * 0.25f);
                const int m = m0 + tid, y = y0 / 8, x = x0 / 8;
                if (m < HW && y < Hl && x < Wl) store_run<Tout, 2>(o3 + level_off<kTW>(lg.g[3], ob + m, y, x), vv, min(2, Wl - x), vec);
            }
        }
        lds_barrier();  // the next patch stages over Os
    }
}

// 2x2 mean over the last two dims of [planes, H, W] -> [planes, H/2, W/2]
// (levels beyond the fused four).
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
        const float v = (((to_f(s[0]) + to_f(s[1])) + to_f(s[W])) + to_f(s[W + 1])) * 0.25f;
        out[i] = from_f<T>(v);
    }
}

// ---------------------------------------------------------------------------
// Feature maps for the on-the-fly correlation and for the feature gradients of
// the volume lookup (AltCorrBlock's pyramid, droid_net.py:146-160; the adjoint
// of CorrBlock's pyramid, droid_net.py:56-102).
//   feat_to_nhwc_kernel     [E][C][HW] (f32 / f16 / bf16) * scale -> [E][HW][C]
//                           (f32 or the input type)
//   pool2x2_nhwc_kernel     [E][H][W][C] -> [E][H/2][W/2][C], 2x2 means in FP32,
//                           rounded to the map type
//   corr_feature_grad_kernel  the level gradients of fmap1 (NHWC, full size)
//                           summed, and of fmap2 (NHWC, level size) spread over
//                           their 2^l x 2^l blocks with weight 4^-l, both times
//                           scale and stored NCHW.
// The transposes go through 32-channel x 64-pixel LDS tiles, so the NCHW side is
// read or written along pixels and the NHWC side along channels.
// ---------------------------------------------------------------------------
template <typename T, typename TO>
__global__ __launch_bounds__(256) void feat_to_nhwc_kernel(const T* __restrict__ in, int C, int HW, float scale,
                                                           TO* __restrict__ out) {
    using A = typename Acc<TO>::type;
    __shared__ A tile[32][65];
    const int p0 = blockIdx.x * 64, c0 = blockIdx.y * 32;
    const int64_t e = blockIdx.z;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const int c = ty + 4 * k, p = p0 + tx;
        tile[c][tx] = (c0 + c < C && p < HW) ? load_as<A>(in + (e * C + c0 + c) * HW + p) * (A)scale : (A)0;
    }
    __syncthreads();
    const int tc = threadIdx.x & 31, tp = threadIdx.x >> 5;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const int p = tp + 8 * k;
        if (c0 + tc < C && p0 + p < HW) out[(e * HW + p0 + p) * C + c0 + tc] = store_from<TO>(tile[tc][p]);
    }
}

template <typename T>
__global__ __launch_bounds__(256) void pool2x2_nhwc_kernel(const T* __restrict__ in, int64_t E, int H, int W, int C,
                                                           T* __restrict__ out) {
    const int Ho = H / 2, Wo = W / 2;
    const int64_t n = E * Ho * Wo * C;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int c = (int)(i % C);
        const int64_t q = i / C;
        const int x = (int)(q % Wo);
        const int64_t r = q / Wo;
        const int y = (int)(r % Ho);
        const int64_t e = r / Ho;
        using A = typename Acc<T>::type;
        const T* s = in + ((e * H + 2 * y) * W + 2 * x) * C + c;
        out[i] = store_from<T>((((load_as<A>(s) + load_as<A>(s + C)) + load_as<A>(s + (int64_t)W * C)) +
                                load_as<A>(s + (int64_t)W * C + C)) * (A)0.25);
    }
}

struct FeatGradLevels {
    const float* g1[4];  // [E][H][W][C]
    const float* g2[4];  // [E][H >> l][W >> l][C]
};

template <typename T>
__global__ __launch_bounds__(256) void corr_feature_grad_kernel(FeatGradLevels gl, int L, int C, int H, int W,
                                                                float scale, T* __restrict__ df1, T* __restrict__ df2) {
    __shared__ float t1[32][65], t2[32][65];
    const int HW = H * W;
    const int p0 = blockIdx.x * 64, c0 = blockIdx.y * 32;
    const int64_t e = blockIdx.z;
    const int tc = threadIdx.x & 31, tp = threadIdx.x >> 5;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const int p = p0 + tp + 8 * k, c = c0 + tc;
        float s1 = 0.f, s2 = 0.f;
        if (p < HW && c < C) {
            const int y = p / W, x = p - y * W;
            float w = 1.f;
            for (int l = 0; l < L; l++, w *= 0.25f) {
                s1 += gl.g1[l][(e * HW + p) * C + c];
                const int Hl = H >> l, Wl = W >> l, yl = y >> l, xl = x >> l;
                if (yl < Hl && xl < Wl) s2 += w * gl.g2[l][((e * Hl + yl) * Wl + xl) * C + c];
            }
        }
        t1[tc][tp + 8 * k] = s1 * scale;
        t2[tc][tp + 8 * k] = s2 * scale;
    }
    __syncthreads();
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const int c = ty + 4 * k;
        if (c0 + c < C && p0 + tx < HW) {
            const int64_t o = (e * C + c0 + c) * HW + p0 + tx;
            df1[o] = from_f<T>(t1[c][tx]);
            df2[o] = from_f<T>(t2[c][tx]);
        }
    }
}

inline size_t align256(size_t x) { return (x + 255) / 256 * 256; }
inline int chan_pad(int C) { return (C + 15) / 16 * 16; }

size_t corr_build_ws(int mma, int64_t E, int C, int H, int W) {
    const int Cp = chan_pad(C);
    const size_t el = mma == VH_F32 ? 4 : mma == VH_FP8E4M3 ? 1 : 2;
    const size_t HW = (size_t)H * W;
    size_t b = 2 * align256(E * HW * Cp * el);
    if (mma == VH_FP8E4M3) b += 2 * align256(E * HW * 4) + 2 * align256((size_t)E * C * 4);
    return b;
}

template <typename T>
__global__ __launch_bounds__(256) void zero_fill_kernel(T* __restrict__ p, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = from_f<T>(0.f);
}

template <typename Ts, typename Mma, typename Tout>
int corr_build_run(int64_t E, int C, int H, int W, const Ts* f1, const Ts* f2, int L, void* const* levels,
                   bool tiled, void* ws, hipStream_t s) {
    constexpr bool FP8 = std::is_same<Mma, MmaFP8>::value;
    using T = typename Mma::T;
    const int Cp = chan_pad(C);
    const int64_t HW = (int64_t)H * W;
    unsigned char* w = (unsigned char*)ws;
    T* A = (T*)w;
    w += align256(E * HW * Cp * sizeof(T));
    T* B = (T*)w;
    w += align256(E * HW * Cp * sizeof(T));
    float *ps1 = nullptr, *ps2 = nullptr, *am1 = nullptr, *am2 = nullptr;
    if (FP8) {
        ps1 = (float*)w; w += align256(E * HW * 4);
        ps2 = (float*)w; w += align256(E * HW * 4);
        am1 = (float*)w; w += align256((size_t)E * C * 4);
        am2 = (float*)w; w += align256((size_t)E * C * 4);
        hipLaunchKernelGGL(corr_chan_amax_kernel<Ts>, dim3((unsigned)(E * C)), dim3(256), 0, s, f1, f2, am1, am2,
                           (int)HW);
        VH_LAUNCH_CHECK("corr_build amax");
    }
    const unsigned cl_blocks = (unsigned)(E * ((HW + 31) / 32));
    const size_t cl_lds = (size_t)Cp * 33 * 4;
    hipLaunchKernelGGL((corr_to_cl_kernel<Ts, Mma, FP8>), dim3(cl_blocks), dim3(256), cl_lds, s, f1, A, am1, am2,
                       ps1, C, Cp, (int)HW, 0);
    VH_LAUNCH_CHECK("corr_build channels-last");
    hipLaunchKernelGGL((corr_to_cl_kernel<Ts, Mma, FP8>), dim3(cl_blocks), dim3(256), cl_lds, s, f2, B, am1, am2,
                       ps2, C, Cp, (int)HW, 1);
    VH_LAUNCH_CHECK("corr_build channels-last");
    const int64_t nblk = E * ((HW + kBM - 1) / kBM) * (int64_t)((H + kPY - 1) / kPY);
    if (nblk > 0x7fffffffll) return fail(VH_ERR_UNSUPPORTED, "corr_build: %lld tiles", (long long)nblk);
    Tout* lv[4] = {nullptr, nullptr, nullptr, nullptr};
    LevelGeoms lg;
    for (int l = 0; l < 4; l++) {
        lg.g[l] = level_geom(H >> l, W >> l, (int)sizeof(Tout), tiled);
        if (l < L) {
            lv[l] = (Tout*)levels[l];
            // tiled padding must read as zero; the tiles' valid part is written below
            const int64_t n = E * HW * lg.g[l].plane;
            if (tiled && (lg.g[l].Hp != lg.g[l].Hl || lg.g[l].Wp != lg.g[l].Wl) && n > 0) {
                int64_t blocks = (n + 255) / 256;
                if (blocks > 1 << 20) blocks = 1 << 20;
                hipLaunchKernelGGL(zero_fill_kernel<Tout>, dim3((unsigned)blocks), dim3(256), 0, s, lv[l], n);
                VH_LAUNCH_CHECK("corr_build pad");
            }
        }
    }
    hipLaunchKernelGGL((corr_build_kernel<Mma, Tout, FP8>), dim3((unsigned)nblk), dim3(256), 0, s, A, B, ps1, ps2,
                       lv[0], lv[1], lv[2], lv[3], L, Cp, H, W, lg);
    VH_LAUNCH_CHECK("corr_build");
    int h = H >> 3, wdt = W >> 3;
    for (int l = 4; l < L; l++) {
        const int64_t planes = E * HW;
        const int64_t n = planes * (h / 2) * (int64_t)(wdt / 2);
        int64_t blocks = (n + 255) / 256;
        if (blocks > 1 << 20) blocks = 1 << 20;
        if (n > 0)
            hipLaunchKernelGGL(pool2x2_kernel<Tout>, dim3((unsigned)blocks), dim3(256), 0, s, (const Tout*)levels[l - 1],
                               (Tout*)levels[l], planes, h, wdt);
        VH_LAUNCH_CHECK("corr_build pool");
        h /= 2;
        wdt /= 2;
    }
    return VH_OK;
}

template <typename Ts, typename Tout>
int corr_build_dispatch(int mma, int64_t E, int C, int H, int W, const void* f1, const void* f2, int L,
                        void* const* levels, bool tiled, void* ws, hipStream_t s) {
    const Ts* a = (const Ts*)f1;
    const Ts* b = (const Ts*)f2;
    if (tiled && L > 4) return fail(VH_ERR_UNSUPPORTED, "corr_build: tiled pyramids have at most 4 levels");
    switch (mma) {
        case VH_F32: return corr_build_run<Ts, MmaF32, Tout>(E, C, H, W, a, b, L, levels, tiled, ws, s);
        case VH_F16: return corr_build_run<Ts, MmaF16, Tout>(E, C, H, W, a, b, L, levels, tiled, ws, s);
        case VH_BF16: return corr_build_run<Ts, MmaBF16, Tout>(E, C, H, W, a, b, L, levels, tiled, ws, s);
        case VH_FP8E4M3: return corr_build_run<Ts, MmaFP8, Tout>(E, C, H, W, a, b, L, levels, tiled, ws, s);
        default: return fail(VH_ERR_UNSUPPORTED, "corr_build: contraction dtype %d", mma);
    }
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
                           const int* Hl, const int* Wl, int layout, const float* coords, int radius, void* out,
                           void* stream) {
    if (B < 0 || H1 < 0 || W1 < 0) return fail(VH_ERR_ARG, "corr_lookup_pyramid: negative size");
    if (num_levels < 1 || num_levels > 8) return fail(VH_ERR_ARG, "corr_lookup_pyramid: 1..8 levels");
    if (radius != 3) return fail(VH_ERR_UNSUPPORTED, "corr_lookup_pyramid: radius %d (built for r=3)", radius);
    const bool tiled = layout == VH_LAYOUT_TILED;
    if (layout != VH_LAYOUT_PLANAR && !tiled) return fail(VH_ERR_ARG, "corr_lookup_pyramid: layout %d", layout);
    if (tiled && dtype == VH_F64) return fail(VH_ERR_UNSUPPORTED, "corr_lookup_pyramid: tiled f64 volume");
    if (tiled && ((int64_t)H1 * W1) % kTilePlanes != 0)
        return fail(VH_ERR_UNSUPPORTED,
                    "corr_lookup_pyramid: tiled levels group the planes of an edge by %d (H1*W1 = %lld)", kTilePlanes,
                    (long long)H1 * W1);
    hipStream_t s = (hipStream_t)stream;
    switch (dtype) {
        case VH_F16: return corr_lookup_pyramid_launch<__half>(B, H1, W1, num_levels, levels, Hl, Wl, tiled, coords, out, s);
        case VH_F32: return corr_lookup_pyramid_launch<float>(B, H1, W1, num_levels, levels, Hl, Wl, tiled, coords, out, s);
        case VH_F64: return corr_lookup_pyramid_launch<double>(B, H1, W1, num_levels, levels, Hl, Wl, tiled, coords, out, s);
        default: return fail(VH_ERR_UNSUPPORTED, "corr_lookup_pyramid: dtype %d", dtype);
    }
}

size_t vh_corr_build_workspace(int mma_dtype, int64_t E, int C, int H, int W) {
    if (E < 0 || C <= 0 || H <= 0 || W <= 0) return 0;
    return corr_build_ws(mma_dtype, E, C, H, W);
}

int vh_corr_build(int in_dtype, int mma_dtype, int out_dtype, int64_t E, int C, int H, int W, const void* fmap1,
                  const void* fmap2, int num_levels, void* const* levels, int layout, void* workspace,
                  void* stream) {
    if (E < 0 || C <= 0 || H <= 0 || W <= 0 || num_levels < 1)
        return fail(VH_ERR_ARG, "corr_build: bad sizes");
    if (C > 448) return fail(VH_ERR_UNSUPPORTED, "corr_build: %d channels (at most 448)", C);
    if (E == 0) return VH_OK;
    if (!workspace) return fail(VH_ERR_ARG, "corr_build: null workspace");
    if (layout != VH_LAYOUT_PLANAR && layout != VH_LAYOUT_TILED) return fail(VH_ERR_ARG, "corr_build: layout %d", layout);
    if (layout == VH_LAYOUT_TILED && ((int64_t)H * W) % kTilePlanes != 0)
        return fail(VH_ERR_UNSUPPORTED, "corr_build: tiled levels group the planes of an edge by %d (H*W = %lld)",
                    kTilePlanes, (long long)H * W);
    hipStream_t s = (hipStream_t)stream;
    if (out_dtype == VH_F32) {
        if (in_dtype == VH_F32) return corr_build_dispatch<float, float>(mma_dtype, E, C, H, W, fmap1, fmap2, num_levels, levels, layout == VH_LAYOUT_TILED, workspace, s);
        if (in_dtype == VH_F16) return corr_build_dispatch<__half, float>(mma_dtype, E, C, H, W, fmap1, fmap2, num_levels, levels, layout == VH_LAYOUT_TILED, workspace, s);
        if (in_dtype == VH_BF16) return corr_build_dispatch<__hip_bfloat16, float>(mma_dtype, E, C, H, W, fmap1, fmap2, num_levels, levels, layout == VH_LAYOUT_TILED, workspace, s);
    } else if (out_dtype == VH_F16) {
        if (in_dtype == VH_F32) return corr_build_dispatch<float, __half>(mma_dtype, E, C, H, W, fmap1, fmap2, num_levels, levels, layout == VH_LAYOUT_TILED, workspace, s);
        if (in_dtype == VH_F16) return corr_build_dispatch<__half, __half>(mma_dtype, E, C, H, W, fmap1, fmap2, num_levels, levels, layout == VH_LAYOUT_TILED, workspace, s);
        if (in_dtype == VH_BF16) return corr_build_dispatch<__hip_bfloat16, __half>(mma_dtype, E, C, H, W, fmap1, fmap2, num_levels, levels, layout == VH_LAYOUT_TILED, workspace, s);
    }
    return fail(VH_ERR_UNSUPPORTED, "corr_build: dtypes in=%d out=%d", in_dtype, out_dtype);
}

}  // extern "C"

namespace vh {
namespace {
// Whether the vector altcorr path applies: channels in whole 32-element chunks,
// 16-byte aligned feature rows, and a level whose bytes (all frames) stay under
// the buffer resource's 31-bit offsets.
bool altcorr_vec_ok(int C, int elem, const void* f1, const void* f2, int64_t level_bytes) {
    return C % kChunk == 0 && (C * elem) % 16 == 0 && reinterpret_cast<uintptr_t>(f1) % 16 == 0 &&
           reinterpret_cast<uintptr_t>(f2) % 16 == 0 && level_bytes < (int64_t)kBufRange - (1 << 20);
}

// ---------------------------------------------------------------------------
// altcorr indexed forward on the matrix cores (FP16 features).  One wave per
// group of 32 consecutive source pixels of one (b, m, source row) at one
// level; the four waves of a workgroup take four groups and never wait for
// each other.  The wave walks the 32 windows' union of target rows, one
// 32-column tile at a time: D = F1 F2^T (32 pixels x 32 target columns) on
// v_mfma_f32_32x32x16_f16 with FP32 accumulation, both operands loaded
// straight into the MFMA lane layout (the pixels' features once per group,
// the next tile's target columns while a tile is computed).  Accumulator
// elements that are taps of their pixel's window go to that pixel's 8 x 8
// taps in the wave's LDS slot, and at the end every pixel mixes its taps
// bilinearly into its 49 outputs.  Work per union row is 32 x 32 x C MACs per
// tile, of which 32 x 8 x C are taps.  A group whose union spans more than 64
// columns or kAltUY rows (coordinates that diverge across the 32 pixels) takes
// the per-pixel wave path instead.
// ---------------------------------------------------------------------------
constexpr int kAltPix = 32;   // source pixels per group
constexpr int kAltUX = 64;    // union columns handled on the matrix cores (two tiles)
constexpr int kAltUY = 40;    // union rows handled on the matrix cores

template <int R, int C>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void altcorr_index_mma_kernel(const __half* __restrict__ fmap1,
                                                                const __half* __restrict__ f2l,
                                                                const float* __restrict__ coords,
                                                                const int64_t* __restrict__ ii,
                                                                const int64_t* __restrict__ jj, float* __restrict__ out,
                                                                int64_t B, int Nf, int M, int H, int W, int H2, int W2,
                                                                int level, int L) {
    constexpr int RD = 2 * R + 1, NT = RD + 1;
    static_assert(NT == 8, "8 x 8 taps per window");
    constexpr int KS = C / 16;  // MFMA steps over the channels
    __shared__ float Sxs[4][kAltPix * NT * NT];  // taps, one slot per wave
    __shared__ int sxbs[4][kAltPix], sybs[4][kAltPix];
    __shared__ float sdxs[4][kAltPix], sdys[4][kAltPix];
    __shared__ float Sw[4][NT][NT + 1];  // per-wave slots of the fallback path

    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    float* Sx = Sxs[wv];
    int* sxb = sxbs[wv];
    int* syb = sybs[wv];
    // group decomposition in 32 bits (the launch checks that the groups fit)
    const uint32_t nxs = (uint32_t)(W + kAltPix - 1) / kAltPix;
    const uint32_t ngroups = (uint32_t)(B * M) * (uint32_t)H * nxs;
    const uint32_t gid = blockIdx.x * 4u + (uint32_t)wv;  // over (b * M + m, row, 32-pixel segment)
    if (gid >= ngroups) return;
    const uint32_t bm32 = gid / ((uint32_t)H * nxs);
    const uint32_t rem = gid - bm32 * (uint32_t)H * nxs;
    const int y = (int)(rem / nxs);
    const int x0 = (int)(rem - (uint32_t)y * nxs) * kAltPix;
    const int64_t bm = bm32;
    const int64_t b = bm32 / (uint32_t)M;
    const int m = (int)(bm32 - (uint32_t)b * (uint32_t)M);
    const int64_t HW = (int64_t)H * W;
    const int64_t src = ii[m], dst = jj[m];
    const float scale = (float)(1 << level);
    const __half* f1row = fmap1 + ((b * Nf + src) * HW + (int64_t)y * W) * C;
    const __half* f2 = f2l + (b * Nf + dst) * (int64_t)H2 * W2 * C;
    float* obase = out + ((bm * L + level) * RD * RD) * HW + (int64_t)y * W + x0;
    const int np = min(kAltPix, W - x0);
    const int pl = lane & 31;

    // this lane's pixel (both half-waves hold the same 32) and the union box
    int xb = 1 << 29, yb = 1 << 29;
    float dx = 0.f, dy = 0.f;
    if (pl < np) {
        const int64_t ci = ((bm * H + y) * W + x0 + pl) * 2;
        const float xs = coords[ci] / scale, ys = coords[ci + 1] / scale;
        const float fx0 = floorf(xs), fy0 = floorf(ys);
        dx = xs - fx0;
        dy = ys - fy0;
        xb = (int)fx0 - R;
        yb = (int)fy0 - R;
    }
    int xmn = pl < np ? xb : INT_MAX, xmx = pl < np ? xb : INT_MIN;
    int ymn = pl < np ? yb : INT_MAX, ymx = pl < np ? yb : INT_MIN;
#pragma unroll
    for (int o = 16; o >= 1; o >>= 1) {
        xmn = min(xmn, __shfl_xor(xmn, o, 32));
        xmx = max(xmx, __shfl_xor(xmx, o, 32));
        ymn = min(ymn, __shfl_xor(ymn, o, 32));
        ymx = max(ymx, __shfl_xor(ymx, o, 32));
    }
    xmn = __builtin_amdgcn_readfirstlane(xmn);
    xmx = __builtin_amdgcn_readfirstlane(xmx);
    ymn = __builtin_amdgcn_readfirstlane(ymn);
    ymx = __builtin_amdgcn_readfirstlane(ymx);
    if (xmx - xmn + NT > kAltUX || ymx - ymn + NT > kAltUY) {
        // diverging coordinates: pixel by pixel, as altcorr_index_wave_kernel
        for (int p = 0; p < np; p++) {
            const int64_t ci = ((bm * H + y) * W + x0 + p) * 2;
            altcorr_wave_pixel<__half, R, float>(f1row + (int64_t)(x0 + p) * C, f2, H2, W2, C, coords[ci] / scale,
                                                 coords[ci + 1] / scale, Sw[wv], obase + p, HW);
        }
        return;
    }
    if (lane < kAltPix) {
        sxb[lane] = xb;
        syb[lane] = yb;
        sdxs[wv][lane] = dx;
        sdys[wv][lane] = dy;
    }
    for (int i = lane; i < kAltPix * NT * NT; i += 64) Sx[i] = 0.f;
    const int X0 = max(xmn, 0), X1 = min(xmx + NT - 1, W2 - 1);
    const int Y0 = max(ymn, 0), Y1 = min(ymx + NT - 1, H2 - 1);
    const int ncol = X1 - X0 + 1;
    const int tsh = ncol > 32 ? 1 : 0;  // log2 of the 32-column tiles per row
    const int nsteps = (Y1 >= Y0 && ncol > 0) ? (Y1 - Y0 + 1) << tsh : 0;
    const int kh = 8 * (lane >> 5);  // this lane's 8 channels of each 16-channel MFMA step
    // A: pixel lane % 32, channels 16 k + kh
    f16x8 av[KS];
    {
        // lanes past the group read its last pixel: their rows of D are never used
        const __half* fp = f1row + (int64_t)(x0 + min(pl, np - 1)) * C + kh;
#pragma unroll
        for (int k = 0; k < KS; k++) av[k] = *reinterpret_cast<const f16x8*>(fp + 16 * k);
    }
    // B of step st: union row Y0 + st >> tsh, tile st & tsh; column 32 t + lane % 32.
    // Loads are unconditional (a column past the union reads column X0; its
    // column of D is never used), so that no branch waits for them.
    auto loadB = [&](f16x8 (&bv)[KS], int st) {
        const int yy = Y0 + (st >> tsh), col = 32 * (st & tsh) + pl;
        const __half* fp = f2 + ((int64_t)yy * W2 + X0 + (col < ncol ? col : 0)) * C + kh;
#pragma unroll
        for (int k = 0; k < KS; k++) bv[k] = *reinterpret_cast<const f16x8*>(fp + 16 * k);
    };
    // dot products of step st; elements that are taps of their pixel go to Sx
    auto stepTaps = [&](const f16x8 (&bv)[KS], int st) {
        f32x16 acc;
#pragma unroll
        for (int v = 0; v < 16; v++) acc[v] = 0.f;
#pragma unroll
        for (int k = 0; k < KS; k++) acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(av[k], bv[k], acc, 0, 0, 0);
        // lane holds union column 32 t + lane % 32 of pixels 8 (v / 4) + 4 (lane / 32) + v % 4
        const int yy = Y0 + (st >> tsh), col = 32 * (st & tsh) + pl;
        const int xc = X0 + col;
#pragma unroll
        for (int v = 0; v < 16; v++) {
            const int pm = 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
            const int c = xc - sxb[pm], r = yy - syb[pm];
            if (col < ncol && (unsigned)c < (unsigned)NT && (unsigned)r < (unsigned)NT)
                Sx[(pm * NT + r) * NT + c] = acc[v];
        }
    };
    // a wavefront's LDS accesses complete in order: the slot writes above are
    // seen by the reads below; the wave barrier keeps the compiler in order
    __builtin_amdgcn_wave_barrier();
    // the next step's loads are issued on every path (the last step reloads
    // itself), so the wait before a step's MFMAs covers its own loads only
    if (nsteps > 0) {
        f16x8 b0[KS], b1[KS];
        loadB(b0, 0);
        for (int st = 0; st < nsteps; st += 2) {
            loadB(b1, min(st + 1, nsteps - 1));
            stepTaps(b0, st);
            if (st + 1 >= nsteps) break;
            loadB(b0, min(st + 2, nsteps - 1));
            stepTaps(b1, st + 1);
        }
    }
    __builtin_amdgcn_wave_barrier();
    // 49 outputs per pixel: lane (p, h) forms channels h, h + 2, ...
    if (pl < np) {
        const float* S = Sx + pl * NT * NT;
        const float w00 = (1.f - dy) * (1.f - dx), w01 = (1.f - dy) * dx, w10 = dy * (1.f - dx), w11 = dy * dx;
        for (int ch = lane >> 5; ch < RD * RD; ch += 2) {
            const int ix = ch / RD, iy = ch - ix * RD;
            const float v = S[iy * NT + ix] * w00 + S[iy * NT + ix + 1] * w01 + S[(iy + 1) * NT + ix] * w10 +
                            S[(iy + 1) * NT + ix + 1] * w11;
            obase[ch * HW + pl] = v;
        }
    }
}

}  // namespace
}  // namespace vh

extern "C" {

int vh_altcorr_forward(int dtype, int64_t B, int N, int H1, int W1, int H2, int W2, int C, const void* fmap1,
                       const void* fmap2, const float* coords, int radius, void* corr, void* stream) {
    if (radius != 3) return fail(VH_ERR_UNSUPPORTED, "altcorr_forward: radius %d (kernels built for r=3)", radius);
    const int64_t n = B * N * (int64_t)H1 * W1;
    if (n == 0) return VH_OK;
    hipStream_t s = (hipStream_t)stream;
    dim3 grid(ceil_div(n, 128));
    if (dtype == VH_F64) {  // per-element path with FP64 sums
        hipLaunchKernelGGL((altcorr_fwd_kernel<double, 3, false>), grid, dim3(128), 0, s, (const double*)fmap1,
                           (const double*)fmap2, coords, (double*)corr, B, N, H1, W1, H2, W2, C);
        VH_LAUNCH_CHECK("altcorr_forward");
        return VH_OK;
    }
    const int elem = dtype == VH_F16 ? 2 : 4;
    const bool vec = altcorr_vec_ok(C, elem, fmap1, fmap2, B * (int64_t)H2 * W2 * C * elem);
    // one wave per pixel when a channel step (8 lanes x 16 bytes) divides C
    if (vec && C % (8 * 16 / elem) == 0) {
        const dim3 wgrid((unsigned)ceil_div(n, 4));
        if (dtype == VH_F16)
            hipLaunchKernelGGL((altcorr_fwd_wave_kernel<__half, 3>), wgrid, dim3(256), 0, s, (const __half*)fmap1,
                               (const __half*)fmap2, coords, (__half*)corr, B, N, H1, W1, H2, W2, C);
        else if (dtype == VH_F32)
            hipLaunchKernelGGL((altcorr_fwd_wave_kernel<float, 3>), wgrid, dim3(256), 0, s, (const float*)fmap1,
                               (const float*)fmap2, coords, (float*)corr, B, N, H1, W1, H2, W2, C);
        else
            return fail(VH_ERR_UNSUPPORTED, "altcorr_forward: dtype %d", dtype);
        VH_LAUNCH_CHECK("altcorr_forward");
        return VH_OK;
    }
    switch (dtype) {
        case VH_F16:
            if (vec)
                hipLaunchKernelGGL((altcorr_fwd_kernel<__half, 3, true>), grid, dim3(128), 0, s, (const __half*)fmap1,
                                   (const __half*)fmap2, coords, (__half*)corr, B, N, H1, W1, H2, W2, C);
            else
                hipLaunchKernelGGL((altcorr_fwd_kernel<__half, 3, false>), grid, dim3(128), 0, s, (const __half*)fmap1,
                                   (const __half*)fmap2, coords, (__half*)corr, B, N, H1, W1, H2, W2, C);
            break;
        case VH_F32:
            if (vec)
                hipLaunchKernelGGL((altcorr_fwd_kernel<float, 3, true>), grid, dim3(128), 0, s, (const float*)fmap1,
                                   (const float*)fmap2, coords, (float*)corr, B, N, H1, W1, H2, W2, C);
            else
                hipLaunchKernelGGL((altcorr_fwd_kernel<float, 3, false>), grid, dim3(128), 0, s, (const float*)fmap1,
                                   (const float*)fmap2, coords, (float*)corr, B, N, H1, W1, H2, W2, C);
            break;
        default: return fail(VH_ERR_UNSUPPORTED, "altcorr_forward: dtype %d", dtype);
    }
    VH_LAUNCH_CHECK("altcorr_forward");
    return VH_OK;
}

int vh_feat_to_nhwc(int dtype, const void* in, int64_t E, int C, int HW, float scale, int out_dtype, void* out,
                    void* stream) {
    if (E < 0 || C < 0 || HW < 0) return fail(VH_ERR_ARG, "feat_to_nhwc: negative size");
    if (out_dtype != VH_F32 && out_dtype != dtype)
        return fail(VH_ERR_UNSUPPORTED, "feat_to_nhwc: output dtype %d (F32 or the input's)", out_dtype);
    if (E * C * HW == 0) return VH_OK;
    if (E > 65535) return fail(VH_ERR_UNSUPPORTED, "feat_to_nhwc: %lld maps (at most 65535)", (long long)E);
    const dim3 grid(ceil_div(HW, 64), ceil_div(C, 32), (unsigned)E);
    hipStream_t s = (hipStream_t)stream;
    const bool f32 = out_dtype == VH_F32;
#define FEAT_LAUNCH(T, TO) \
    hipLaunchKernelGGL((feat_to_nhwc_kernel<T, TO>), grid, dim3(256), 0, s, (const T*)in, C, HW, scale, (TO*)out)
    switch (dtype) {
        case VH_F32: FEAT_LAUNCH(float, float); break;
        case VH_F64:
            if (f32) FEAT_LAUNCH(double, float);
            else FEAT_LAUNCH(double, double);
            break;
        case VH_F16:
            if (f32) FEAT_LAUNCH(__half, float);
            else FEAT_LAUNCH(__half, __half);
            break;
        case VH_BF16:
            if (f32) FEAT_LAUNCH(__hip_bfloat16, float);
            else FEAT_LAUNCH(__hip_bfloat16, __hip_bfloat16);
            break;
        default: return fail(VH_ERR_UNSUPPORTED, "feat_to_nhwc: dtype %d", dtype);
    }
#undef FEAT_LAUNCH
    VH_LAUNCH_CHECK("feat_to_nhwc");
    return VH_OK;
}

int vh_pool2x2_nhwc(int dtype, const void* in, int64_t E, int H, int W, int C, void* out, void* stream) {
    if (E < 0 || H < 0 || W < 0 || C < 0) return fail(VH_ERR_ARG, "pool2x2_nhwc: negative size");
    const int64_t n = E * (H / 2) * (int64_t)(W / 2) * C;
    if (n == 0) return VH_OK;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 1 << 20) blocks = 1 << 20;
    hipStream_t s = (hipStream_t)stream;
    switch (dtype) {
        case VH_F32:
            hipLaunchKernelGGL(pool2x2_nhwc_kernel<float>, dim3((unsigned)blocks), dim3(256), 0, s, (const float*)in, E,
                               H, W, C, (float*)out);
            break;
        case VH_F16:
            hipLaunchKernelGGL(pool2x2_nhwc_kernel<__half>, dim3((unsigned)blocks), dim3(256), 0, s, (const __half*)in,
                               E, H, W, C, (__half*)out);
            break;
        case VH_F64:
            hipLaunchKernelGGL(pool2x2_nhwc_kernel<double>, dim3((unsigned)blocks), dim3(256), 0, s, (const double*)in,
                               E, H, W, C, (double*)out);
            break;
        case VH_BF16:
            hipLaunchKernelGGL(pool2x2_nhwc_kernel<__hip_bfloat16>, dim3((unsigned)blocks), dim3(256), 0, s,
                               (const __hip_bfloat16*)in, E, H, W, C, (__hip_bfloat16*)out);
            break;
        default: return fail(VH_ERR_UNSUPPORTED, "pool2x2_nhwc: dtype %d", dtype);
    }
    VH_LAUNCH_CHECK("pool2x2_nhwc");
    return VH_OK;
}

int vh_corr_feature_grad(int dtype, int64_t E, int C, int H, int W, int num_levels, const float* const* g1,
                         const float* const* g2, float scale, void* df1, void* df2, void* stream) {
    if (E < 0 || C < 0 || H < 0 || W < 0) return fail(VH_ERR_ARG, "corr_feature_grad: negative size");
    if (num_levels < 1 || num_levels > 4) return fail(VH_ERR_ARG, "corr_feature_grad: 1..4 levels");
    if (E * C * H * W == 0) return VH_OK;
    if (E > 65535) return fail(VH_ERR_UNSUPPORTED, "corr_feature_grad: %lld maps (at most 65535)", (long long)E);
    FeatGradLevels gl{};
    for (int l = 0; l < num_levels; l++) {
        if (!g1[l] || (!g2[l] && (H >> l) * (W >> l) > 0)) return fail(VH_ERR_ARG, "corr_feature_grad: level %d", l);
        gl.g1[l] = g1[l];
        gl.g2[l] = g2[l];
    }
    const dim3 grid(ceil_div((int64_t)H * W, 64), ceil_div(C, 32), (unsigned)E);
    hipStream_t s = (hipStream_t)stream;
    switch (dtype) {
        case VH_F32:
            hipLaunchKernelGGL(corr_feature_grad_kernel<float>, grid, dim3(256), 0, s, gl, num_levels, C, H, W, scale,
                               (float*)df1, (float*)df2);
            break;
        case VH_F16:
            hipLaunchKernelGGL(corr_feature_grad_kernel<__half>, grid, dim3(256), 0, s, gl, num_levels, C, H, W, scale,
                               (__half*)df1, (__half*)df2);
            break;
        case VH_BF16:
            hipLaunchKernelGGL(corr_feature_grad_kernel<__hip_bfloat16>, grid, dim3(256), 0, s, gl, num_levels, C, H, W,
                               scale, (__hip_bfloat16*)df1, (__hip_bfloat16*)df2);
            break;
        default: return fail(VH_ERR_UNSUPPORTED, "corr_feature_grad: dtype %d", dtype);
    }
    VH_LAUNCH_CHECK("corr_feature_grad");
    return VH_OK;
}

size_t vh_altcorr_backward_workspace_size(int64_t B, int N, int H1, int W1) {
    const int64_t nwin = B * (int64_t)N * H1 * W1;
    return (size_t)(nwin > 0 ? nwin : 0) * (64 * sizeof(float) + sizeof(int2)) + 256;
}

int vh_altcorr_backward(int dtype, int64_t B, int N, int H1, int W1, int H2, int W2, int C, const void* fmap1,
                        const void* fmap2, const float* coords, const void* corr_grad, int radius, void* fmap1_grad,
                        void* fmap2_grad, void* workspace, void* stream) {
    if (dtype != VH_F32) return fail(VH_ERR_UNSUPPORTED, "altcorr_backward: float32 only");
    if (radius != 3) return fail(VH_ERR_UNSUPPORTED, "altcorr_backward: radius %d (kernels built for r=3)", radius);
    hipStream_t s = (hipStream_t)stream;
    const int64_t nwin = B * (int64_t)N * H1 * W1;
    // channel pairs per lane where C is even and the maps are 8-byte aligned
    const bool vec2 = C % 2 == 0 && (reinterpret_cast<uintptr_t>(fmap1) & 7) == 0 &&
                      (reinterpret_cast<uintptr_t>(fmap2) & 7) == 0 && (reinterpret_cast<uintptr_t>(fmap1_grad) & 7) == 0 &&
                      (reinterpret_cast<uintptr_t>(fmap2_grad) & 7) == 0;
    // the gather path: C in whole float4 groups, a row of fmap2_grad in LDS and
    // a batch's windows numbered in 31 bits
    // (windows per pass: 64 or 32 if that leaves room for three workgroups per
    // CU, else the most that fit one; fewer than 32 cost more in passes than
    // the occupancy gains)
    int gw = 0;
    if (workspace && C % 4 == 0 && (reinterpret_cast<uintptr_t>(fmap1) & 15) == 0 &&
        (reinterpret_cast<uintptr_t>(fmap2_grad) & 15) == 0 && (int64_t)N * H1 * W1 < (int64_t{1} << 31)) {
        for (int w = kAltWin; w >= 32 && !gw; w /= 2)
            if (altcorr_gather_lds(W2, C, w) <= 53 * 1024) gw = w;
        for (int w = kAltWin; w >= 8 && !gw; w /= 2)
            if (altcorr_gather_lds(W2, C, w) <= 150 * 1024) gw = w;
    }
    if (gw) {
        // fmap1_grad and the tap gradients per window, then fmap2_grad by target
        // rows: both gradients fully written, no atomics
        float* taps = (float*)workspace;
        int2* org = (int2*)((char*)workspace + ((size_t)std::max<int64_t>(nwin, 0) * 64 * sizeof(float) + 255) / 256 * 256);
        if (B * (int64_t)H1 * W1 * C > 0) {
            if (vec2 && ((int64_t)H1 * W1) % 4 == 0)
                hipLaunchKernelGGL(altcorr_bwd_group_kernel<3>, dim3(ceil_div(B * (int64_t)H1 * W1 / 4, 4)), dim3(256),
                                   0, s, (const float*)fmap1, (const float*)fmap2, coords, (const float*)corr_grad,
                                   (float*)fmap1_grad, B, N, H1, W1, H2, W2, C, taps, org);
            else
                hipLaunchKernelGGL((altcorr_bwd_wave_kernel<3, true>), dim3(ceil_div(B * (int64_t)H1 * W1, 4)),
                                   dim3(256), 0, s, (const float*)fmap1, (const float*)fmap2, coords,
                                   (const float*)corr_grad, (float*)fmap1_grad, (float*)fmap2_grad, B, N, H1, W1, H2,
                                   W2, C, vec2, taps, org);
            VH_LAUNCH_CHECK("altcorr_backward");
        }
        if (B * (int64_t)H2 * W2 * C > 0) {
            hipLaunchKernelGGL(altcorr_bwd_gather_kernel<3>, dim3((unsigned)(B * H2)), dim3(256),
                               altcorr_gather_lds(W2, C, gw), s, (const float*)fmap1, taps, org, (float*)fmap2_grad,
                               B, N, H1, W1, H2, W2, C, gw);
            VH_LAUNCH_CHECK("altcorr_backward gather");
        }
        return VH_OK;
    }
    // without a workspace (or a row too wide for LDS): fmap2_grad is cleared here
    // (a fill kernel, so the call can be captured), then accumulated with FP32
    // atomic adds; fmap1_grad is stored by the wave that owns its pixel
    const int64_t n2 = B * (int64_t)H2 * W2 * C;
    if (n2 > 0) {
        hipLaunchKernelGGL(zero_fill_kernel<float>, dim3((unsigned)std::min<int64_t>(ceil_div(n2, 256), 65535)),
                           dim3(256), 0, s, (float*)fmap2_grad, n2);
        VH_LAUNCH_CHECK("altcorr_backward clear");
    }
    if (B * (int64_t)H1 * W1 * C == 0) return VH_OK;
    hipLaunchKernelGGL(altcorr_bwd_wave_kernel<3>, dim3(ceil_div(B * (int64_t)H1 * W1, 4)), dim3(256), 0, s,
                       (const float*)fmap1, (const float*)fmap2, coords, (const float*)corr_grad, (float*)fmap1_grad,
                       (float*)fmap2_grad, B, N, H1, W1, H2, W2, C, vec2);
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
    const int elem = dtype == VH_F16 ? 2 : 4;
    for (int l = 0; l < num_levels; l++) {
        dim3 grid(ceil_div(n, 128));
        if (dtype == VH_F64) {  // per-element path with FP64 sums, FP32 output
            hipLaunchKernelGGL((altcorr_index_kernel<double, 3, false>), grid, dim3(128), 0, s,
                               (const double*)fmap1, (const double*)levels[l], coords, ii, jj, (float*)corr, B, Nf,
                               M, H, W, C, Hl[l], Wl[l], l, num_levels);
            VH_LAUNCH_CHECK("altcorr_index_forward");
            continue;
        }
        const bool vec = altcorr_vec_ok(C, elem, fmap1, levels[l], B * (int64_t)Nf * Hl[l] * Wl[l] * C * elem);
        // FP16 on the matrix cores: 32-pixel groups; C = 64 or 128 (operands
        // held in registers, whole channel steps of the per-pixel fallback)
        if (dtype == VH_F16 && vec && (C == 64 || C == 128) &&
            B * M * H * ceil_div(W, kAltPix) < ((int64_t)1 << 31)) {
            const dim3 mgrid((unsigned)ceil_div(B * M * H * ceil_div(W, kAltPix), 4));
#define VH_ALT_MMA(CC)                                                                                          \
    hipLaunchKernelGGL((altcorr_index_mma_kernel<3, CC>), mgrid, dim3(256), 0, s, (const __half*)fmap1,          \
                       (const __half*)levels[l], coords, ii, jj, (float*)corr, B, Nf, M, H, W, Hl[l], Wl[l], l, \
                       num_levels)
            if (C == 64) VH_ALT_MMA(64);
            else VH_ALT_MMA(128);
#undef VH_ALT_MMA
            VH_LAUNCH_CHECK("altcorr_index_forward");
            continue;
        }
        // one wave per pixel when a channel step (8 lanes x 16 bytes) divides C
        if (vec && C % (8 * 16 / elem) == 0) {
            const dim3 wgrid((unsigned)ceil_div(n, 4));
            if (dtype == VH_F16)
                hipLaunchKernelGGL((altcorr_index_wave_kernel<__half, 3>), wgrid, dim3(256), 0, s, (const __half*)fmap1,
                                   (const __half*)levels[l], coords, ii, jj, (float*)corr, B, Nf, M, H, W, C, Hl[l],
                                   Wl[l], l, num_levels);
            else if (dtype == VH_F32)
                hipLaunchKernelGGL((altcorr_index_wave_kernel<float, 3>), wgrid, dim3(256), 0, s, (const float*)fmap1,
                                   (const float*)levels[l], coords, ii, jj, (float*)corr, B, Nf, M, H, W, C, Hl[l],
                                   Wl[l], l, num_levels);
            else
                return fail(VH_ERR_UNSUPPORTED, "altcorr_index_forward: dtype %d", dtype);
            VH_LAUNCH_CHECK("altcorr_index_forward");
            continue;
        }
        switch (dtype) {
            case VH_F16:
                if (vec)
                    hipLaunchKernelGGL((altcorr_index_kernel<__half, 3, true>), grid, dim3(128), 0, s,
                                       (const __half*)fmap1, (const __half*)levels[l], coords, ii, jj, (float*)corr, B,
                                       Nf, M, H, W, C, Hl[l], Wl[l], l, num_levels);
                else
                    hipLaunchKernelGGL((altcorr_index_kernel<__half, 3, false>), grid, dim3(128), 0, s,
                                       (const __half*)fmap1, (const __half*)levels[l], coords, ii, jj, (float*)corr, B,
                                       Nf, M, H, W, C, Hl[l], Wl[l], l, num_levels);
                break;
            case VH_F32:
                if (vec)
                    hipLaunchKernelGGL((altcorr_index_kernel<float, 3, true>), grid, dim3(128), 0, s,
                                       (const float*)fmap1, (const float*)levels[l], coords, ii, jj, (float*)corr, B,
                                       Nf, M, H, W, C, Hl[l], Wl[l], l, num_levels);
                else
                    hipLaunchKernelGGL((altcorr_index_kernel<float, 3, false>), grid, dim3(128), 0, s,
                                       (const float*)fmap1, (const float*)levels[l], coords, ii, jj, (float*)corr, B,
                                       Nf, M, H, W, C, Hl[l], Wl[l], l, num_levels);
                break;
            default: return fail(VH_ERR_UNSUPPORTED, "altcorr_index_forward: dtype %d", dtype);
        }
        VH_LAUNCH_CHECK("altcorr_index_forward");
    }
    return VH_OK;
}

}  // extern "C"
