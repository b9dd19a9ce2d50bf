// The DroidNet update operator (vipe/slam/networks/droid_net.py:393-519:
// ConvGRU, GraphAgg, UpdateModule) on gfx950, FP16 activations and weights with
// FP32 accumulation, as the reference runs it under autocast.
//
// Activations live channels-last (NHWC: [image][pixel][channel], f16), so the
// channels of a pixel are one contiguous row and a convolution is an implicit
// GEMM: M = pixels, N = output channels, K = taps x input channels.
//
//   uo_conv_kernel        the implicit GEMM on v_mfma_f32_32x32x16_f16.  One
//                         workgroup of 8 waves per 256 pixels x 128 output
//                         channels (128 x 256 where the layer has 256); each
//                         wave owns 64 x 64 as 2 x 2 tiles of 32 x 32.  K
//                         advances one tap of one 32-channel chunk of one input
//                         segment per step: the step's pixel rows (shifted by the
//                         tap, zero outside the image) and weight rows are copied
//                         from memory straight into a three-stage LDS ring, two
//                         steps ahead of the matrix cores.  Up to four input
//                         segments stand in for torch.cat along channels.  The
//                         epilogue adds the bias (and a per-image addend) and
//                         applies the layer's activation, or the ConvGRU gates /
//                         state update.
//   uo_conv_small_vec_kernel / uo_conv_small_kernel  convolutions with 1 or 2
//                         output channels (the delta, weight and eta heads) on
//                         the VALU: lanes along the input channels with 16-byte
//                         loads (4-byte loads where the weights are not 16-byte
//                         aligned), cross-lane sums per batch of pixels.
//   uo_to_nhwc / uo_to_nchw_kernel  layout changes at the module boundary.
//   uo_mean_kernel / uo_fc_kernel   ConvGRU's global context: the per-image mean
//                         of sigmoid(w(net)) * net and the three 1x1 convolutions
//                         applied to it.

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>

#include <cmath>

#include "vh_common.hpp"

namespace vh {
namespace {

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));

constexpr int kUoKC = 32;         // channels per K step
constexpr int kUoCoutAlign = 128;  // the weights' output channels are padded to a multiple of this

struct ConvArgs {
    const _Float16* seg[4];  // NHWC inputs
    int seg_c[4];            // channels taken from each (a multiple of 32)
    int seg_stride[4];       // channels per pixel of each buffer
    int seg_off[4];          // first channel taken
    int nseg, cin;           // segments, total input channels (the weights' row length)
    const _Float16* w;       // [KH * KW][Cout_pad][cin]
    const float* bias;       // [Cout]
    int nimg, H, W, Cout, Cout_pad, KH, KW;
    int epi;
    void* out;
    int out_f32, out_stride, out_off;
    const float* glo;        // [nimg][glo_stride] addend at channel n, or null
    int glo_stride;
    const _Float16* aux0;    // gates: net; gru: z; sigmul: net   ([P][aux_stride])
    const _Float16* aux1;    // gru: net
    int aux_stride;
    _Float16* out2;          // gates: r * net
};

__device__ __forceinline__ float sigmoidf_(float x) { return 1.f / (1.f + __expf(-x)); }
__device__ __forceinline__ float softplusf_(float x) { return x > 20.f ? x : log1pf(__expf(x)); }

// LDS ring of the implicit GEMM.  A stage holds one K step: the tile's BM
// pixel rows and BN weight rows of 64 bytes (32 channels), unpadded; piece q
// (16 bytes) of row r sits in slot q ^ ((r >> 2) & 3), which keeps the MFMA
// fragment reads (32 consecutive rows, one piece) on distinct banks.  The
// stages are filled by global_load_lds_dwordx4 (lane l's 16 bytes land at the
// instruction's LDS base + 16 l), so no registers stage the operands and the
// loads of the next kUoStages - 1 steps are in flight while a step computes.
constexpr int kUoStages = 3;
__device__ __align__(16) uint4 uo_zero16[1];  // zero-initialised source of taps outside the image

__device__ __forceinline__ void lds_load16(const void* g, unsigned lds_byte_addr) {
    __builtin_amdgcn_global_load_lds(g, (__attribute__((address_space(3))) void*)(uintptr_t)lds_byte_addr, 16, 0, 0);
}

// s_waitcnt immediate that waits until at most n vector memory operations are
// outstanding (vmcnt in bits 3:0 and 15:14; expcnt and lgkmcnt left at their
// maxima, i.e. not waited for)
constexpr int uo_vmcnt(int n) { return (n & 15) | (7 << 4) | (15 << 8) | ((n >> 4) << 14); }

// Bias, per-image addend, activation or ConvGRU update, and the store of one
// wave's 64 x 64 block at pixel m0 and channel n0.  Lane l holds output channel
// n0 + 32 j + l % 32 of pixels m0 + 32 i + 8 (v / 4) + 4 (l / 32) + v % 4.
template <int EPI>
__device__ __forceinline__ void uo_epilogue(const ConvArgs& a, const f32x16 (&acc)[2][2], int m0, int n0, int lane,
                                            int P, int HW) {
    const bool nchw = (a.epi & VH_UO_NCHW) != 0;
#pragma unroll
    for (int j = 0; j < 2; j++) {
        const int n = n0 + 32 * j + (lane & 31);
        if (n >= a.Cout) continue;
        const float bn = a.bias[n];
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int v = 0; v < 16; v++) {
                const int p = m0 + 32 * i + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
                if (p >= P) continue;
                float x = acc[i][j][v] + bn;
                int img = 0;
                if (a.glo || nchw) img = (unsigned)p / (unsigned)HW;
                if (a.glo) x += a.glo[img * a.glo_stride + n];
                if constexpr (EPI == VH_UO_GATES) {
                    const float g = sigmoidf_(x);
                    if (n < 128) {
                        ((_Float16*)a.out)[(int64_t)p * a.out_stride + a.out_off + n] = (_Float16)g;
                    } else {
                        const float h = (float)a.aux0[(int64_t)p * a.aux_stride + (n - 128)];
                        a.out2[(int64_t)p * 128 + (n - 128)] = (_Float16)(g * h);
                    }
                    continue;
                }
                if constexpr (EPI == VH_UO_GRU) {
                    const float qv = tanhf(x);
                    const float z = (float)a.aux0[(int64_t)p * a.aux_stride + n];
                    const float h = (float)a.aux1[(int64_t)p * a.aux_stride + n];
                    x = (1.f - z) * h + z * qv;
                } else if constexpr (EPI == VH_UO_RELU) {
                    x = fmaxf(x, 0.f);
                } else if constexpr (EPI == VH_UO_SIGMOID) {
                    x = sigmoidf_(x);
                } else if constexpr (EPI == VH_UO_SIGMUL) {
                    x = sigmoidf_(x) * (float)a.aux0[(int64_t)p * a.aux_stride + n];
                }
                const int64_t o = nchw ? ((int64_t)img * a.out_stride + a.out_off + n) * HW + (p - img * HW)
                                       : (int64_t)p * a.out_stride + a.out_off + n;
                if (a.out_f32) ((float*)a.out)[o] = x;
                else ((_Float16*)a.out)[o] = (_Float16)x;
            }
    }
}

// One K step's loads into the LDS stage at base: the wave's LA pixel-row
// pieces (the tap's shifted pixels of channels sp[0, 32) at channel stride ss,
// the zero page outside the image) and its LB weight-row pieces of wt.
template <int LA, int LB, int BM>
__device__ __forceinline__ void uo_issue_step(unsigned base, int wv, const _Float16* sp, int ss, int dy, int dx, int H,
                                              int W, int HW, const int (&qa)[LA], const int (&img)[LA],
                                              const int (&py)[LA], const int (&px)[LA], const bool (&pok)[LA],
                                              const _Float16* zp, const _Float16* wt, int cin, const int (&rb)[LB],
                                              const int (&qb)[LB]) {
#pragma unroll
    for (int u = 0; u < LA; u++) {
        const int yy = py[u] + dy, xx = px[u] + dx;
        const bool ok = pok[u] && yy >= 0 && yy < H && xx >= 0 && xx < W;
        const _Float16* g = sp + (int64_t)(img[u] * HW + yy * W + xx) * ss + 8 * qa[u];
        lds_load16(ok ? g : zp, base + (wv * LA + u) * 1024);
    }
#pragma unroll
    for (int u = 0; u < LB; u++) lds_load16(wt + (int64_t)rb[u] * cin + 8 * qb[u], base + BM * 64 + (wv * LB + u) * 1024);
}

// The implicit GEMM.  A workgroup of WM x WN waves computes BM = 64 WM pixels
// x BN = 64 WN output channels; each wave owns 64 x 64 as 2 x 2 tiles of
// 32 x 32.  Per K step each wave issues LA pixel-row and LB weight-row loads
// of 16 rows each.
template <int EPI, int WM, int WN>
__global__ __launch_bounds__(64 * WM * WN) __attribute__((amdgpu_waves_per_eu(WM * WN == 8 ? 4 : 3))) void
uo_conv_kernel(ConvArgs a) {
    constexpr int NW = WM * WN, BM = 64 * WM, BN = 64 * WN;
    constexpr int LA = BM / (16 * NW), LB = BN / (16 * NW);
    constexpr int kStage = (BM + BN) * 64;
    static_assert(LA >= 1 && LB >= 1 && LA * 16 * NW == BM && LB * 16 * NW == BN, "tile");
    static_assert(kUoStages == 3, "the waits below count one younger step");
    extern __shared__ __align__(16) unsigned char lds_[];
    const unsigned lds0 = (unsigned)(uintptr_t)lds_;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int wm = (wv / WN) * 64, wn = (wv % WN) * 64;
    const int HW = a.H * a.W;
    const int P = a.nimg * HW;  // < 2^31, checked by the host
    // tiles in contiguous ranges per XCD, pixel tiles fastest
    const int lb = (int)xcd_logical((int64_t)blockIdx.y * gridDim.x + blockIdx.x, (int64_t)gridDim.x * gridDim.y);
    const int by = lb / (int)gridDim.x, bx = lb - by * (int)gridDim.x;
    const int m0 = bx * BM;
    const int n0 = by * BN;
    const int ph = a.KH / 2, pw = a.KW / 2;
    // this lane's LDS slots: load u of the wave fills slots s = (wv L + u) 64 + lane,
    // row s / 4, slot s % 4, with piece (s % 4) ^ ((row >> 2) & 3) of the row
    int qa[LA], img[LA], py[LA], px[LA], rb[LB], qb[LB];
    bool pok[LA];
#pragma unroll
    for (int u = 0; u < LA; u++) {
        const int s = (wv * LA + u) * 64 + lane, r = s >> 2;
        qa[u] = (s & 3) ^ ((r >> 2) & 3);
        pok[u] = m0 + r < P;
        const int pc = pok[u] ? m0 + r : 0;
        img[u] = pc / HW;
        py[u] = (pc - img[u] * HW) / a.W;
        px[u] = pc - img[u] * HW - py[u] * a.W;
    }
#pragma unroll
    for (int u = 0; u < LB; u++) {
        const int s = (wv * LB + u) * 64 + lane;
        rb[u] = s >> 2;
        qb[u] = (s & 3) ^ ((rb[u] >> 2) & 3);
    }
    // K walks (segment, 32-channel chunk, tap) in that nesting: the taps of one
    // chunk read the same input lines (shifted by at most a row and a column),
    // so they come from the caches instead of HBM.  The segment is picked by
    // integer sums over the segment table, not by selecting among its entries:
    // selects of kernel-argument loads become a pointer table in scratch memory.
    const uint64_t sb0 = (uint64_t)a.seg[0];
    const uint64_t sd1 = (uint64_t)a.seg[1] - sb0, sd2 = (uint64_t)a.seg[2] - (uint64_t)a.seg[1],
                   sd3 = (uint64_t)a.seg[3] - (uint64_t)a.seg[2];
    const int sc0 = a.seg_c[0], sc1 = a.seg_c[1], sc2 = a.seg_c[2], sc3 = a.seg_c[3];
    const int ss0 = a.seg_stride[0], ss1 = a.seg_stride[1] - ss0, ss2 = a.seg_stride[2] - a.seg_stride[1],
              ss3 = a.seg_stride[3] - a.seg_stride[2];
    const int so0 = a.seg_off[0], so1 = a.seg_off[1] - so0, so2 = a.seg_off[2] - a.seg_off[1],
              so3 = a.seg_off[3] - a.seg_off[2];
    const int nseg = a.nseg, H = a.H, W = a.W, Cout_pad = a.Cout_pad, cin = a.cin;
    const _Float16* wp = a.w;
    const int per_tap = (sc0 + (nseg > 1 ? sc1 : 0) + (nseg > 2 ? sc2 : 0) + (nseg > 3 ? sc3 : 0)) / kUoKC;
    const int ntap = a.KH * a.KW;
    const int nchunks = ntap * per_tap;
    const _Float16* zp = reinterpret_cast<const _Float16*>(uo_zero16);
    // the issue cursor (uniform): segment, chunk and tap of the next step to load,
    // the tap also as its offsets (dy, dx)
    int cur_tap = 0, cur_dy = -ph, cur_dx = -pw, cur_sg = 0, cur_c0 = 0, cur_wc0 = 0;
    uint64_t cur_sp = sb0;
    int cur_ss = ss0, cur_so = so0, cur_sc = sc0;
    // the cursor's step into stage ST, then the cursor's advance
#define UO_ISSUE(ST)                                                                                             \
    {                                                                                                            \
        const unsigned base = lds0 + (ST) * kStage;                                                              \
        uo_issue_step<LA, LB, BM>(base, wv, (const _Float16*)cur_sp + cur_so + cur_c0, cur_ss, cur_dy, cur_dx, H, \
                                  W, HW, qa, img, py, px, pok, zp,                                               \
                                  wp + ((int64_t)cur_tap * Cout_pad + n0) * cin + cur_wc0, cin, rb, qb);          \
        ++cur_tap;                                                                                               \
        if (++cur_dx > pw) {                                                                                     \
            cur_dx = -pw;                                                                                        \
            ++cur_dy;                                                                                            \
        }                                                                                                        \
        if (cur_tap == ntap) {                                                                                   \
            cur_tap = 0;                                                                                         \
            cur_dy = -ph;                                                                                        \
            cur_c0 += kUoKC;                                                                                     \
            cur_wc0 += kUoKC;                                                                                    \
            if (cur_c0 == cur_sc && cur_sg + 1 < nseg) {                                                         \
                cur_c0 = 0;                                                                                      \
                const int sg = ++cur_sg;                                                                         \
                cur_sp = sb0 + (sg >= 1 ? sd1 : 0) + (sg >= 2 ? sd2 : 0) + (sg >= 3 ? sd3 : 0);                  \
                cur_sc = sg == 1 ? sc1 : sg == 2 ? sc2 : sc3;                                                    \
                cur_ss = ss0 + (sg >= 1 ? ss1 : 0) + (sg >= 2 ? ss2 : 0) + (sg >= 3 ? ss3 : 0);                  \
                cur_so = so0 + (sg >= 1 ? so1 : 0) + (sg >= 2 ? so2 : 0) + (sg >= 3 ? so3 : 0);                  \
            }                                                                                                    \
        }                                                                                                        \
    }
    f32x16 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int v = 0; v < 16; v++) acc[i][j][v] = 0.f;
    // prologue: steps 0 .. kUoStages - 2 in flight
#pragma unroll
    for (int s = 0; s < kUoStages - 1; s++)
        if (s < nchunks) UO_ISSUE(s)
    const int ra0 = wm + (lane & 31), rb0 = wn + (lane & 31);  // fragment rows (+ 32 i)
    for (int kc = 0; kc < nchunks; kc++) {
        // step kc's loads are older than the LA + LB of step kc + 1, if issued
        if (kc + 1 < nchunks) __builtin_amdgcn_s_waitcnt(uo_vmcnt(LA + LB));
        else __builtin_amdgcn_s_waitcnt(uo_vmcnt(0));
        // every wave's part of step kc is in LDS, and stage (kc - 1) % kUoStages is
        // free (its fragments were consumed by the MFMAs).  A bare barrier, with
        // compiler-only fences: __syncthreads would also wait for the younger
        // step's loads.
        asm volatile("" ::: "memory");
        __builtin_amdgcn_s_barrier();
        asm volatile("" ::: "memory");
        if (kc + kUoStages - 1 < nchunks) UO_ISSUE((kc + kUoStages - 1) % kUoStages)
        const unsigned char* st = lds_ + (kc % kUoStages) * kStage;
#pragma unroll
        for (int kk = 0; kk < kUoKC; kk += 16) {
            const int q = (kk >> 3) + (lane >> 5);
            f16x8 af[2], bf[2];
#pragma unroll
            for (int i = 0; i < 2; i++) {
                const int r = ra0 + 32 * i;
                af[i] = *reinterpret_cast<const f16x8*>(st + r * 64 + ((q ^ ((r >> 2) & 3)) << 4));
            }
#pragma unroll
            for (int j = 0; j < 2; j++) {
                const int r = rb0 + 32 * j;
                bf[j] = *reinterpret_cast<const f16x8*>(st + BM * 64 + r * 64 + ((q ^ ((r >> 2) & 3)) << 4));
            }
#pragma unroll
            for (int i = 0; i < 2; i++)
#pragma unroll
                for (int j = 0; j < 2; j++)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(af[i], bf[j], acc[i][j], 0, 0, 0);
        }
    }
#undef UO_ISSUE
    uo_epilogue<EPI>(a, acc, m0 + wm, n0 + wn, lane, P, HW);
}

// Convolution with Cout <= 2 (the delta, weight and eta heads) on the VALU.
// Lanes run along the input channels: lane l holds channels 2l, 2l + 1 of each
// 128-channel group, so one load instruction reads a pixel's whole channel row
// (256 contiguous bytes) and the weights of the lane's channels stay in
// registers.  A wave takes 64 consecutive pixels, kSmallBatch / Cout at a time:
// every tap of every pixel of the batch is loaded before any is used, each
// lane accumulates its channels' share of the kSmallBatch (pixel, output)
// values, and one cross-lane sum of the kSmallBatch values leaves the total of
// value (lane & 31) in lane l.
constexpr int kUoSmallMaxCin = 256;  // two 128-channel groups
constexpr int kUoSmallMaxTaps = 9;   // 3 x 3 kernels
constexpr int kSmallBatch = 32;      // (pixel, output) values reduced together

template <int CTRL>
__device__ __forceinline__ float uo_dpp(float x) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), CTRL, 0xf, 0xf, true));
}

// Sum over the 64 lanes of each of 32 values: each halving step keeps the half
// selected by one lane bit and adds the partner lane's copy (the partner of a
// DPP mirror or swap step differs in that bit); after the last step lane l
// holds the total of value (l & 31).  Every lane of the wave must be active.
__device__ __forceinline__ float uo_wave_sum32(float (&v)[32]) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(v[i]), __float_as_uint(v[i + 16]), false,
                                                        false);
        v[i] = __uint_as_float(r[0]) + __uint_as_float(r[1]);
    }
#define UO_HALVE(S, CTRL, BIT)                                          \
    _Pragma("unroll") for (int i = 0; i < S; i++) {                     \
        const float a = v[i] + uo_dpp<CTRL>(v[i]);                      \
        const float b = v[i + S] + uo_dpp<CTRL>(v[i + S]);              \
        v[i] = (lane & BIT) ? b : a;                                    \
    }
    UO_HALVE(8, 0x140, 8)
    UO_HALVE(4, 0x141, 4)
    UO_HALVE(2, 0x4E, 2)
    UO_HALVE(1, 0xB1, 1)
#undef UO_HALVE
    const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(v[0]), __float_as_uint(v[0]), false, false);
    return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}

template <int COUT>
__global__ __launch_bounds__(256) void uo_conv_small_kernel(const _Float16* __restrict__ in, int in_stride, int in_off,
                                                            int cin, const float* __restrict__ w,
                                                            const float* __restrict__ bias, int nimg, int H, int W,
                                                            int KH, int KW, int epi, float* __restrict__ out,
                                                            int out_stride) {
    constexpr int NPB = kSmallBatch / COUT;  // pixels per batch
    const int lane = threadIdx.x & 63;
    const int HW = H * W;
    const int P = nimg * HW;  // < 2^31, checked by the host
    const int pw0 = (blockIdx.x * 4 + (threadIdx.x >> 6)) * 64;  // this wave's first pixel
    const int ntap = KH * KW;
    const int ngrp = (cin + 127) / 128;
    for (int b0 = 0; b0 < 64; b0 += NPB) {
        // the batch's pixels (wave-uniform): image base, row and column; a
        // pixel past P gets a row that fails every bounds test
        int pb[NPB], py[NPB], px[NPB];
#pragma unroll
        for (int i = 0; i < NPB; i++) {
            const int p = pw0 + b0 + i;
            const int img = p / HW;
            const int rem = p - img * HW;
            pb[i] = img * HW;
            py[i] = p < P ? rem / W : -(1 << 20);
            px[i] = rem - (rem / W) * W;
        }
        float acc[32];
#pragma unroll
        for (int i = 0; i < 32; i++) acc[i] = 0.f;
#pragma unroll
        for (int g = 0; g < ngrp; g++) {
            const int c = 128 * g + 2 * lane;
            const bool cok = c < cin;
            for (int t = 0; t < ntap; t++) {
                const int dy = t / KW - KH / 2, dx = t % KW - KW / 2;
                // this lane's two channels' weights for the tap (the same for
                // every wave: served by the caches)
                float w0[COUT], w1[COUT];
#pragma unroll
                for (int o = 0; o < COUT; o++) {
                    w0[o] = cok ? w[((int64_t)t * cin + c) * COUT + o] : 0.f;
                    w1[o] = cok ? w[((int64_t)t * cin + c + 1) * COUT + o] : 0.f;
                }
                __half2 x[NPB];
#pragma unroll
                for (int i = 0; i < NPB; i++) {
                    const int yy = py[i] + dy, xx = px[i] + dx;
                    const bool ok = yy >= 0 && yy < H && xx >= 0 && xx < W && cok;
                    const int64_t q = ok ? (int64_t)(pb[i] + yy * W + xx) : 0;
                    const __half2 v = *reinterpret_cast<const __half2*>(in + q * in_stride + in_off + (cok ? c : 0));
                    x[i] = ok ? v : __half2(__float2half(0.f), __float2half(0.f));
                }
#pragma unroll
                for (int i = 0; i < NPB; i++) {
                    const float x0 = __low2float(x[i]), x1 = __high2float(x[i]);
#pragma unroll
                    for (int o = 0; o < COUT; o++) acc[i * COUT + o] += x0 * w0[o] + x1 * w1[o];
                }
            }
        }
        const float tot = uo_wave_sum32(acc);
        // lane l < 32 holds value l: pixel b0 + l / COUT, output l % COUT
        const int p = pw0 + b0 + (lane & 31) / COUT;
        const int o = (lane & 31) % COUT;
        if (lane < 32 && p < P) {
            float x = tot + bias[o];
            if (epi == VH_UO_SIGMOID) x = sigmoidf_(x);
            else if (epi == VH_UO_ETA) x = 0.01f * softplusf_(x);
            else if (epi == VH_UO_RELU) x = fmaxf(x, 0.f);
            out[(int64_t)p * out_stride + o] = x;
        }
    }
}

// The same convolution with 16-byte loads, where the channel counts and offsets
// are multiples of 8: lane l takes channels 8 (l % 16) .. +7 of a 128-channel
// group and pixel slot l / 16, so one load instruction reads four pixels' whole
// channel rows.  A wave takes its 64 pixels 32 at a time (8 per lane), loads
// every pixel's row of a tap before using any, and sums the 8 x Cout values of
// each lane over its 16-lane row by halving: lane j of a row ends with value
// j % (8 Cout), pixel j / Cout of the row's slot, output j % Cout.
template <int NV>
__device__ __forceinline__ float uo_row_sum(float (&v)[16], int lane) {
    if constexpr (NV == 8) {
#pragma unroll
        for (int i = 0; i < 8; i++) v[i] += __shfl_xor(v[i], 8, 16);
    } else {
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const float a = v[i] + __shfl_xor(v[i], 8, 16), b = v[i + 8] + __shfl_xor(v[i + 8], 8, 16);
            v[i] = (lane & 8) ? b : a;
        }
    }
#pragma unroll
    for (int m = 4; m >= 1; m >>= 1)
#pragma unroll
        for (int i = 0; i < m; i++) {
            const float a = v[i] + __shfl_xor(v[i], m, 16), b = v[i + m] + __shfl_xor(v[i + m], m, 16);
            v[i] = (lane & m) ? b : a;
        }
    return v[0];
}

template <int COUT>
__global__ __launch_bounds__(256) void uo_conv_small_vec_kernel(const _Float16* __restrict__ in, int in_stride,
                                                                int in_off, int cin, const float* __restrict__ w,
                                                                const float* __restrict__ bias, int nimg, int H, int W,
                                                                int KH, int KW, int epi, float* __restrict__ out,
                                                                int out_stride) {
    constexpr int PPL = 8;        // pixels per lane and batch
    constexpr int NV = PPL * COUT;  // values per lane
    const int lane = threadIdx.x & 63, cg = lane & 15, ps = lane >> 4;
    const int HW = H * W;
    const int P = nimg * HW;  // < 2^31, checked by the host
    const int pw0 = (blockIdx.x * 4 + (threadIdx.x >> 6)) * 64;
    const int ntap = KH * KW;
    const int ngrp = (cin + 127) / 128;
    for (int b0 = 0; b0 < 64; b0 += 4 * PPL) {
        // pixel b0 + 4 k + ps of the wave for k < PPL; one past P gets a row
        // that fails every bounds test
        int pb[PPL], py[PPL], px[PPL];
#pragma unroll
        for (int k = 0; k < PPL; k++) {
            const int p = pw0 + b0 + 4 * k + ps;
            const int img = p / HW;
            const int rem = p - img * HW;
            pb[k] = img * HW;
            py[k] = p < P ? rem / W : -(1 << 20);
            px[k] = rem - (rem / W) * W;
        }
        float acc[16];
#pragma unroll
        for (int i = 0; i < 16; i++) acc[i] = 0.f;
        for (int g = 0; g < ngrp; g++) {
            const int c = 128 * g + 8 * cg;
            const bool cok = c < cin;  // cin is a multiple of 8
            for (int t = 0; t < ntap; t++) {
                const int dy = t / KW - KH / 2, dx = t % KW - KW / 2;
                float4 wq[2 * COUT];  // weights [8 channels][COUT] of the tap
#pragma unroll
                for (int u = 0; u < 2 * COUT; u++)
                    wq[u] = cok ? reinterpret_cast<const float4*>(w + ((int64_t)t * cin + c) * COUT)[u]
                                : make_float4(0.f, 0.f, 0.f, 0.f);
                const float* wv = reinterpret_cast<const float*>(wq);
                uint4 x[PPL];
#pragma unroll
                for (int k = 0; k < PPL; k++) {
                    const int yy = py[k] + dy, xx = px[k] + dx;
                    const bool ok = yy >= 0 && yy < H && xx >= 0 && xx < W && cok;
                    const int64_t o = ok ? (int64_t)(pb[k] + yy * W + xx) * in_stride + in_off + c : 0;
                    const uint4 v = *reinterpret_cast<const uint4*>(in + o);
                    x[k] = ok ? v : make_uint4(0u, 0u, 0u, 0u);
                }
#pragma unroll
                for (int k = 0; k < PPL; k++) {
                    const unsigned d[4] = {x[k].x, x[k].y, x[k].z, x[k].w};
#pragma unroll
                    for (int e2 = 0; e2 < 4; e2++) {
                        const __half2 h = *reinterpret_cast<const __half2*>(&d[e2]);
                        const float x0 = __low2float(h), x1 = __high2float(h);
#pragma unroll
                        for (int o = 0; o < COUT; o++)
                            acc[k * COUT + o] += x0 * wv[(2 * e2) * COUT + o] + x1 * wv[(2 * e2 + 1) * COUT + o];
                    }
                }
            }
        }
        const float tot = uo_row_sum<NV>(acc, lane);
        const int j = lane & 15;
        const int p = pw0 + b0 + 4 * ((j % NV) / COUT) + ps;
        const int o = j % COUT;
        if (j < NV && p < P) {
            float v = tot + bias[o];
            if (epi == VH_UO_SIGMOID) v = sigmoidf_(v);
            else if (epi == VH_UO_ETA) v = 0.01f * softplusf_(v);
            else if (epi == VH_UO_RELU) v = fmaxf(v, 0.f);
            out[(int64_t)p * out_stride + o] = v;
        }
    }
}

// Layout changes at the module boundary, as register transposes of 8 x 8
// blocks.  Lane l of a wave takes channel group l % 8 (8 channels) and pixel
// group l / 8 (8 pixels), so a wave covers 64 channels x 64 pixels: each of its
// 8 loads reads one channel row of 64 pixels (whole lines) and each of its 8
// stores one pixel row of 64 channels.  The vector form needs 8-pixel groups
// inside an image and 16-byte aligned rows (checked by the host); other shapes
// take one lane per output element.

// Two 16-bit values as one dword: element e (0 or 1) of a and of b.
__device__ __forceinline__ unsigned pick16(unsigned a, unsigned b, int e) {
    return e ? ((a >> 16) | (b & 0xffff0000u)) : ((a & 0xffffu) | (b << 16));
}

__device__ __forceinline__ unsigned dw(const uint4& v, int k) { return k == 0 ? v.x : k == 1 ? v.y : k == 2 ? v.z : v.w; }

// [nimg][C][HW] (f32 or f16) -> f16 [nimg][HW][Cpad], channels past C zero.
template <typename T>
__global__ __launch_bounds__(256) void uo_to_nhwc_kernel(const T* __restrict__ in, int nimg, int C, int HW, int Cpad,
                                                         _Float16* __restrict__ out) {
    const int64_t P = (int64_t)nimg * HW;
    const int lane = threadIdx.x & 63;
    const int64_t p = (int64_t)blockIdx.x * 256 + (threadIdx.x & ~63) + 8 * (lane >> 3);
    const int c0 = blockIdx.y * 64 + 8 * (lane & 7);
    if (p >= P || c0 >= Cpad) return;
    const int64_t img = p / HW, q = p - img * HW;
    // v[i]: channel c0 + i at pixels p .. p + 7, as 8 halves (4 dwords)
    uint4 v[8];
#pragma unroll
    for (int i = 0; i < 8; i++) {
        const int c = c0 + i;
        v[i] = make_uint4(0, 0, 0, 0);
        if (c < C) {
            const T* src = in + ((int64_t)img * C + c) * HW + q;
            if constexpr (sizeof(T) == 2) {
                v[i] = *reinterpret_cast<const uint4*>(src);
            } else {
                const float4 a = reinterpret_cast<const float4*>(src)[0], b = reinterpret_cast<const float4*>(src)[1];
                const __half2 h0 = __floats2half2_rn(a.x, a.y), h1 = __floats2half2_rn(a.z, a.w);
                const __half2 h2 = __floats2half2_rn(b.x, b.y), h3 = __floats2half2_rn(b.z, b.w);
                v[i] = make_uint4(*reinterpret_cast<const unsigned*>(&h0), *reinterpret_cast<const unsigned*>(&h1),
                                  *reinterpret_cast<const unsigned*>(&h2), *reinterpret_cast<const unsigned*>(&h3));
            }
        }
    }
#pragma unroll
    for (int j = 0; j < 8; j++) {
        // pixel p + j: channels c0 .. c0 + 7
        const uint4 o = make_uint4(pick16(dw(v[0], j >> 1), dw(v[1], j >> 1), j & 1),
                                   pick16(dw(v[2], j >> 1), dw(v[3], j >> 1), j & 1),
                                   pick16(dw(v[4], j >> 1), dw(v[5], j >> 1), j & 1),
                                   pick16(dw(v[6], j >> 1), dw(v[7], j >> 1), j & 1));
        *reinterpret_cast<uint4*>(out + (p + j) * Cpad + c0) = o;
    }
}

template <typename T>
__global__ __launch_bounds__(256) void uo_to_nhwc_elem_kernel(const T* __restrict__ in, int nimg, int C, int HW,
                                                              int Cpad, _Float16* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (int64_t)nimg * HW * Cpad) return;
    const int c = (int)(i % Cpad);
    const int64_t p = i / Cpad, img = p / HW, q = p - img * HW;
    out[i] = c < C ? (_Float16)(float)in[((int64_t)img * C + c) * HW + q] : (_Float16)0.f;
}

// f16 [nimg][HW][stride] (channels [off, off + C)) -> [nimg][C][HW] (f32 or f16).
template <typename T>
__global__ __launch_bounds__(256) void uo_to_nchw_kernel(const _Float16* __restrict__ in, int nimg, int C, int HW,
                                                         int stride, int off, T* __restrict__ out) {
    const int64_t P = (int64_t)nimg * HW;
    const int lane = threadIdx.x & 63;
    const int64_t p = (int64_t)blockIdx.x * 256 + (threadIdx.x & ~63) + 8 * (lane >> 3);
    const int c0 = blockIdx.y * 64 + 8 * (lane & 7);
    if (p >= P || c0 >= C) return;
    const int64_t img = p / HW, q = p - img * HW;
    // w[j]: pixel p + j at channels c0 .. c0 + 7
    uint4 w[8];
#pragma unroll
    for (int j = 0; j < 8; j++) w[j] = *reinterpret_cast<const uint4*>(in + (p + j) * stride + off + c0);
#pragma unroll
    for (int i = 0; i < 8; i++) {
        // channel c0 + i: pixels p .. p + 7
        const uint4 o = make_uint4(pick16(dw(w[0], i >> 1), dw(w[1], i >> 1), i & 1),
                                   pick16(dw(w[2], i >> 1), dw(w[3], i >> 1), i & 1),
                                   pick16(dw(w[4], i >> 1), dw(w[5], i >> 1), i & 1),
                                   pick16(dw(w[6], i >> 1), dw(w[7], i >> 1), i & 1));
        T* dst = out + ((int64_t)img * C + c0 + i) * HW + q;
        if constexpr (sizeof(T) == 2) {
            *reinterpret_cast<uint4*>(dst) = o;
        } else {
            const unsigned d[4] = {o.x, o.y, o.z, o.w};
            float f[8];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const float2 t = __half22float2(*reinterpret_cast<const __half2*>(&d[k]));
                f[2 * k] = t.x;
                f[2 * k + 1] = t.y;
            }
            reinterpret_cast<float4*>(dst)[0] = make_float4(f[0], f[1], f[2], f[3]);
            reinterpret_cast<float4*>(dst)[1] = make_float4(f[4], f[5], f[6], f[7]);
        }
    }
}

template <typename T>
__global__ __launch_bounds__(256) void uo_to_nchw_elem_kernel(const _Float16* __restrict__ in, int nimg, int C, int HW,
                                                              int stride, int off, T* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (int64_t)nimg * C * HW) return;
    const int64_t q = i % HW, ic = i / HW, img = ic / C;
    const int c = (int)(ic - img * C);
    out[i] = (T)(float)in[((int64_t)img * HW + q) * stride + off + c];
}

// Per-image mean over the pixels of [nimg][HW][C] f32: one workgroup per
// (image, 64 channels).  With V4 a lane reads 4 channels as one 16-byte load, 16
// lanes cover the 64 channels of a pixel and the workgroup's 16 lane groups take
// pixels g, g + 16, ...; each lane keeps 4 running sums over its pixels (taken
// in turn), and fixed-order trees combine them, so the result is the same on
// every run.  Without V4 (C not a multiple of 4) a lane takes one channel.
template <bool V4>
__global__ __launch_bounds__(256) void uo_mean_kernel(const float* __restrict__ in, int nimg, int HW, int C,
                                                      float* __restrict__ out) {
    constexpr int NG = V4 ? 16 : 4;  // pixel groups of the workgroup
    __shared__ float part[NG][64];
    const int ncb = (C + 63) / 64;
    const int img = blockIdx.x / ncb;
    const int cb = (blockIdx.x - img * ncb) * 64;
    const float* base = in + (int64_t)img * HW * C;
    if constexpr (V4) {
        const int g = threadIdx.x >> 4, c = cb + 4 * (threadIdx.x & 15);
        float4 acc[4];
#pragma unroll
        for (int u = 0; u < 4; u++) acc[u] = make_float4(0.f, 0.f, 0.f, 0.f);
        if (c < C) {
            int p = g;
            for (; p + 3 * NG < HW; p += 4 * NG) {
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const float4 v = *reinterpret_cast<const float4*>(base + (int64_t)(p + u * NG) * C + c);
                    acc[u].x += v.x; acc[u].y += v.y; acc[u].z += v.z; acc[u].w += v.w;
                }
            }
#pragma unroll
            for (int u = 0; u < 3; u++) {
                const int pu = p + u * NG;
                if (pu < HW) {
                    const float4 v = *reinterpret_cast<const float4*>(base + (int64_t)pu * C + c);
                    acc[u].x += v.x; acc[u].y += v.y; acc[u].z += v.z; acc[u].w += v.w;
                }
            }
        }
        const float4 t = make_float4((acc[0].x + acc[1].x) + (acc[2].x + acc[3].x), (acc[0].y + acc[1].y) + (acc[2].y + acc[3].y),
                                     (acc[0].z + acc[1].z) + (acc[2].z + acc[3].z), (acc[0].w + acc[1].w) + (acc[2].w + acc[3].w));
        const int cl = 4 * (threadIdx.x & 15);
        part[g][cl] = t.x; part[g][cl + 1] = t.y; part[g][cl + 2] = t.z; part[g][cl + 3] = t.w;
    } else {
        const int g = threadIdx.x >> 6, c = cb + (threadIdx.x & 63);
        float sum = 0.f;
        if (c < C)
            for (int p = g; p < HW; p += NG) sum += base[(int64_t)p * C + c];
        part[g][threadIdx.x & 63] = sum;
    }
    __syncthreads();
    if (threadIdx.x < 64 && cb + (int)threadIdx.x < C) {
        float r[NG];
#pragma unroll
        for (int k = 0; k < NG; k++) r[k] = part[k][threadIdx.x];
#pragma unroll
        for (int w = NG / 2; w >= 1; w /= 2)
#pragma unroll
            for (int k = 0; k < w; k++) r[k] = r[k] + r[k + w];
        out[(int64_t)img * C + cb + threadIdx.x] = r[0] / (float)HW;
    }
}

// out[img][k] = sum_c w[k][c] in[img][c] + b[k] (1x1 convolutions of a 1x1 map).
__global__ __launch_bounds__(256) void uo_fc_kernel(const float* __restrict__ in, int nimg, int Cin,
                                                    const float* __restrict__ w, const float* __restrict__ b, int Cout,
                                                    float* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (int64_t)nimg * Cout) return;
    const int img = (int)(i / Cout), k = (int)(i - (int64_t)img * Cout);
    float s = 0.f;
    for (int c = 0; c < Cin; c++) s += w[(int64_t)k * Cin + c] * in[(int64_t)img * Cin + c];
    out[i] = s + b[k];
}

// Launches the implicit GEMM with WM x WN waves per workgroup (a tile of
// 64 WM pixels x 64 WN output channels; Cout_pad is a multiple of 64 WN).
template <int WM, int WN>
int uo_launch(const ConvArgs& a, int epi, int64_t P, hipStream_t st) {
    const dim3 grid(ceil_div(P, 64 * WM), a.Cout_pad / (64 * WN)), block(64 * WM * WN);
    const size_t lds = (size_t)kUoStages * (64 * WM + 64 * WN) * 64;
    switch (epi) {
#define UO_CASE(E)                                                                                         \
    case E: {                                                                                              \
        const void* fn = reinterpret_cast<const void*>(&uo_conv_kernel<E, WM, WN>);                        \
        if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess)   \
            return fail(VH_ERR_HIP, "uo_conv: %zu bytes of LDS", lds);                                     \
        hipLaunchKernelGGL((uo_conv_kernel<E, WM, WN>), grid, block, lds, st, a);                          \
        return VH_OK;                                                                                      \
    }
        UO_CASE(VH_UO_NONE)
        UO_CASE(VH_UO_RELU)
        UO_CASE(VH_UO_SIGMOID)
        UO_CASE(VH_UO_SIGMUL)
        UO_CASE(VH_UO_GATES)
        UO_CASE(VH_UO_GRU)
#undef UO_CASE
        default: return fail(VH_ERR_ARG, "uo_conv: epilogue %d", a.epi);
    }
}

}  // namespace
}  // namespace vh

using namespace vh;

extern "C" {

size_t vh_uo_conv_args_size(void) { return sizeof(vh_uo_conv_args); }

int vh_uo_conv(const vh_uo_conv_args* args, void* stream) {
    if (!args) return fail(VH_ERR_ARG, "uo_conv: null arguments");
    const vh_uo_conv_args& g = *args;
    if (g.nseg < 1 || g.nseg > 4) return fail(VH_ERR_ARG, "uo_conv: %d input segments (1 to 4)", g.nseg);
    if (g.KH < 1 || g.KW < 1 || (g.KH & 1) == 0 || (g.KW & 1) == 0)
        return fail(VH_ERR_ARG, "uo_conv: kernel %dx%d (odd sizes, 'same' padding)", g.KH, g.KW);
    if (g.Cout < 1 || g.Cout_pad < g.Cout || g.Cout_pad % kUoCoutAlign != 0)
        return fail(VH_ERR_ARG, "uo_conv: Cout %d, Cout_pad %d (a multiple of %d)", g.Cout, g.Cout_pad,
                    kUoCoutAlign);
    ConvArgs a{};
    int cin = 0;
    for (int s = 0; s < g.nseg; s++) {
        if (!g.seg[s] || ((uintptr_t)g.seg[s] & 15) || g.seg_c[s] <= 0 || g.seg_c[s] % kUoKC != 0 ||
            g.seg_off[s] % 8 != 0 || g.seg_stride[s] % 8 != 0 || g.seg_off[s] < 0 ||
            g.seg_off[s] + g.seg_c[s] > g.seg_stride[s])
            return fail(VH_ERR_ARG, "uo_conv: segment %d (channels %d at %d of %d; multiples of 32 / 8)", s,
                        g.seg_c[s], g.seg_off[s], g.seg_stride[s]);
        a.seg[s] = (const _Float16*)g.seg[s];
        a.seg_c[s] = g.seg_c[s];
        a.seg_stride[s] = g.seg_stride[s];
        a.seg_off[s] = g.seg_off[s];
        cin += g.seg_c[s];
    }
    const int epi = g.epilogue & 15;
    if (epi > VH_UO_GRU || epi == VH_UO_ETA || (g.epilogue & ~(15 | VH_UO_NCHW)))
        return fail(VH_ERR_ARG, "uo_conv: epilogue %d", g.epilogue);
    if ((epi == VH_UO_GATES || epi == VH_UO_GRU || epi == VH_UO_SIGMUL) && !g.aux0)
        return fail(VH_ERR_ARG, "uo_conv: the gate epilogues need aux0");
    if (epi == VH_UO_GATES && (g.Cout != 256 || !g.out2 || g.out_f32))
        return fail(VH_ERR_ARG, "uo_conv: the gates epilogue takes 256 f16 outputs (z, r) and out2");
    if (epi == VH_UO_GRU && (!g.aux1 || g.Cout != 128))
        return fail(VH_ERR_ARG, "uo_conv: the GRU epilogue takes 128 outputs and aux1");
    if (!g.weight || ((uintptr_t)g.weight & 15) || !g.bias || !g.out)
        return fail(VH_ERR_ARG, "uo_conv: weight (16-byte aligned), bias and out are required");
    const int nout = epi == VH_UO_GATES ? 128 : g.Cout;  // channels written per pixel to out
    const bool nchw = (g.epilogue & VH_UO_NCHW) != 0;
    if (g.out_off < 0 || g.out_off + nout > g.out_stride || (nchw && epi == VH_UO_GATES))
        return fail(VH_ERR_ARG, "uo_conv: output channels [%d, %d) of %d per pixel", g.out_off, g.out_off + nout,
                    g.out_stride);
    if ((epi == VH_UO_GATES || epi == VH_UO_GRU || epi == VH_UO_SIGMUL) && g.aux_stride < nout)
        return fail(VH_ERR_ARG, "uo_conv: aux_stride %d < %d", g.aux_stride, nout);
    if (g.glo && g.glo_stride < g.Cout) return fail(VH_ERR_ARG, "uo_conv: glo_stride %d < Cout", g.glo_stride);
    if (g.nimg < 0 || g.H < 0 || g.W < 0) return fail(VH_ERR_ARG, "uo_conv: image size");
    a.nseg = g.nseg;
    a.cin = cin;
    a.w = (const _Float16*)g.weight;
    a.bias = g.bias;
    a.nimg = g.nimg; a.H = g.H; a.W = g.W; a.Cout = g.Cout; a.Cout_pad = g.Cout_pad; a.KH = g.KH; a.KW = g.KW;
    a.epi = g.epilogue;
    a.out = g.out; a.out_f32 = g.out_f32; a.out_stride = g.out_stride; a.out_off = g.out_off;
    a.glo = g.glo; a.glo_stride = g.glo_stride;
    a.aux0 = (const _Float16*)g.aux0; a.aux1 = (const _Float16*)g.aux1; a.aux_stride = g.aux_stride;
    a.out2 = (_Float16*)g.out2;
    const int64_t P = (int64_t)g.nimg * g.H * g.W;
    if (P == 0) return VH_OK;
    if (P + 256 >= (int64_t)1 << 31) return fail(VH_ERR_UNSUPPORTED, "uo_conv: %lld pixels", (long long)P);
    const hipStream_t st = (hipStream_t)stream;
    // 128 pixels x 256 channels where the output has a multiple of 256 channels
    // (each pixel row is staged once for all of them), else 256 pixels x 128
    const int rc = g.Cout_pad % 256 == 0 ? uo_launch<2, 4>(a, epi, P, st) : uo_launch<4, 2>(a, epi, P, st);
    if (rc != VH_OK) return rc;
    VH_LAUNCH_CHECK("uo_conv");
    return VH_OK;
}

int vh_uo_conv_small(const void* in, int in_stride, int in_off, int cin, const float* weight, const float* bias,
                     int nimg, int H, int W, int KH, int KW, int Cout, int epilogue, float* out, int out_stride,
                     void* stream) {
    if (Cout < 1 || Cout > 2) return fail(VH_ERR_ARG, "uo_conv_small: Cout %d (1 or 2)", Cout);
    if (cin <= 0 || cin % 8 != 0 || in_off < 0 || in_off % 8 != 0 || in_stride % 8 != 0 || in_off + cin > in_stride)
        return fail(VH_ERR_ARG, "uo_conv_small: channels %d at %d of %d (multiples of 8)", cin, in_off, in_stride);
    if (KH < 1 || KW < 1 || (KH & 1) == 0 || (KW & 1) == 0)
        return fail(VH_ERR_ARG, "uo_conv_small: kernel %dx%d (odd sizes)", KH, KW);
    if (epilogue != VH_UO_NONE && epilogue != VH_UO_RELU && epilogue != VH_UO_SIGMOID && epilogue != VH_UO_ETA)
        return fail(VH_ERR_ARG, "uo_conv_small: epilogue %d", epilogue);
    if (!in || ((uintptr_t)in & 15) || !weight || !bias || !out || out_stride < Cout || nimg < 0 || H < 0 || W < 0)
        return fail(VH_ERR_ARG, "uo_conv_small: null or misaligned buffer, or out_stride %d < %d", out_stride, Cout);
    if (KH * KW > kUoSmallMaxTaps || cin > kUoSmallMaxCin)
        return fail(VH_ERR_UNSUPPORTED, "uo_conv_small: %dx%d kernel with %d input channels (limits %d taps, %d)", KH,
                    KW, cin, kUoSmallMaxTaps, kUoSmallMaxCin);
    const int64_t P = (int64_t)nimg * H * W;
    if (P == 0) return VH_OK;
    if (P + 256 >= (int64_t)1 << 31) return fail(VH_ERR_UNSUPPORTED, "uo_conv_small: %lld pixels", (long long)P);
    if (cin % 8 == 0 && in_stride % 8 == 0 && in_off % 8 == 0 && ((uintptr_t)weight & 15) == 0) {
        if (Cout == 2)
            hipLaunchKernelGGL(uo_conv_small_vec_kernel<2>, dim3(ceil_div(P, 256)), dim3(256), 0,
                               (hipStream_t)stream, (const _Float16*)in, in_stride, in_off, cin, weight, bias, nimg, H,
                               W, KH, KW, epilogue, out, out_stride);
        else
            hipLaunchKernelGGL(uo_conv_small_vec_kernel<1>, dim3(ceil_div(P, 256)), dim3(256), 0,
                               (hipStream_t)stream, (const _Float16*)in, in_stride, in_off, cin, weight, bias, nimg, H,
                               W, KH, KW, epilogue, out, out_stride);
    } else if (Cout == 2)
        hipLaunchKernelGGL(uo_conv_small_kernel<2>, dim3(ceil_div(P, 256)), dim3(256), 0, (hipStream_t)stream,
                           (const _Float16*)in, in_stride, in_off, cin, weight, bias, nimg, H, W, KH, KW, epilogue,
                           out, out_stride);
    else
        hipLaunchKernelGGL(uo_conv_small_kernel<1>, dim3(ceil_div(P, 256)), dim3(256), 0, (hipStream_t)stream,
                           (const _Float16*)in, in_stride, in_off, cin, weight, bias, nimg, H, W, KH, KW, epilogue,
                           out, out_stride);
    VH_LAUNCH_CHECK("uo_conv_small");
    return VH_OK;
}

int vh_uo_to_nhwc(int dtype, const void* in, int nimg, int C, int HW, int Cpad, void* out, void* stream) {
    if (Cpad < C || C < 0 || nimg < 0 || HW < 0) return fail(VH_ERR_ARG, "uo_to_nhwc: Cpad %d < C %d", Cpad, C);
    if (dtype != VH_F32 && dtype != VH_F16) return fail(VH_ERR_UNSUPPORTED, "uo_to_nhwc: dtype %d", dtype);
    const int64_t P = (int64_t)nimg * HW;
    if (P * Cpad == 0) return VH_OK;
    const hipStream_t st = (hipStream_t)stream;
    // the 8 x 8 transpose: 8-pixel groups inside an image, 16-byte rows
    const bool vec = HW % 8 == 0 && Cpad % 8 == 0 && ((uintptr_t)in & 15) == 0 && ((uintptr_t)out & 15) == 0;
    if (vec) {
        const dim3 grid(ceil_div(P, 256), ceil_div(Cpad, 64));
        if (dtype == VH_F32)
            hipLaunchKernelGGL(uo_to_nhwc_kernel<float>, grid, dim3(256), 0, st, (const float*)in, nimg, C, HW, Cpad,
                               (_Float16*)out);
        else
            hipLaunchKernelGGL(uo_to_nhwc_kernel<_Float16>, grid, dim3(256), 0, st, (const _Float16*)in, nimg, C, HW,
                               Cpad, (_Float16*)out);
    } else {
        const dim3 grid(ceil_div(P * Cpad, 256));
        if (dtype == VH_F32)
            hipLaunchKernelGGL(uo_to_nhwc_elem_kernel<float>, grid, dim3(256), 0, st, (const float*)in, nimg, C, HW,
                               Cpad, (_Float16*)out);
        else
            hipLaunchKernelGGL(uo_to_nhwc_elem_kernel<_Float16>, grid, dim3(256), 0, st, (const _Float16*)in, nimg, C,
                               HW, Cpad, (_Float16*)out);
    }
    VH_LAUNCH_CHECK("uo_to_nhwc");
    return VH_OK;
}

int vh_uo_to_nchw(int dtype, const void* in, int nimg, int C, int HW, int stride, int off, void* out, void* stream) {
    if (off < 0 || C < 0 || off + C > stride || nimg < 0 || HW < 0)
        return fail(VH_ERR_ARG, "uo_to_nchw: channels [%d, %d) of %d", off, off + C, stride);
    if (dtype != VH_F32 && dtype != VH_F16) return fail(VH_ERR_UNSUPPORTED, "uo_to_nchw: dtype %d", dtype);
    const int64_t P = (int64_t)nimg * HW;
    if (P * C == 0) return VH_OK;
    const hipStream_t st = (hipStream_t)stream;
    const bool vec = HW % 8 == 0 && C % 8 == 0 && stride % 8 == 0 && off % 8 == 0 && ((uintptr_t)in & 15) == 0 &&
                     ((uintptr_t)out & 15) == 0;
    if (vec) {
        const dim3 grid(ceil_div(P, 256), ceil_div(C, 64));
        if (dtype == VH_F32)
            hipLaunchKernelGGL(uo_to_nchw_kernel<float>, grid, dim3(256), 0, st, (const _Float16*)in, nimg, C, HW,
                               stride, off, (float*)out);
        else
            hipLaunchKernelGGL(uo_to_nchw_kernel<_Float16>, grid, dim3(256), 0, st, (const _Float16*)in, nimg, C, HW,
                               stride, off, (_Float16*)out);
    } else {
        const dim3 grid(ceil_div(P * C, 256));
        if (dtype == VH_F32)
            hipLaunchKernelGGL(uo_to_nchw_elem_kernel<float>, grid, dim3(256), 0, st, (const _Float16*)in, nimg, C, HW,
                               stride, off, (float*)out);
        else
            hipLaunchKernelGGL(uo_to_nchw_elem_kernel<_Float16>, grid, dim3(256), 0, st, (const _Float16*)in, nimg, C,
                               HW, stride, off, (_Float16*)out);
    }
    VH_LAUNCH_CHECK("uo_to_nchw");
    return VH_OK;
}

int vh_uo_mean(const float* in, int nimg, int HW, int C, float* out, void* stream) {
    if ((int64_t)nimg * C == 0) return VH_OK;
    const dim3 grid((unsigned)((int64_t)nimg * ((C + 63) / 64)));
    if (C % 4 == 0 && ((uintptr_t)in & 15) == 0)
        hipLaunchKernelGGL(uo_mean_kernel<true>, grid, dim3(256), 0, (hipStream_t)stream, in, nimg, HW, C, out);
    else
        hipLaunchKernelGGL(uo_mean_kernel<false>, grid, dim3(256), 0, (hipStream_t)stream, in, nimg, HW, C, out);
    VH_LAUNCH_CHECK("uo_mean");
    return VH_OK;
}

int vh_uo_fc(const float* in, int nimg, int Cin, const float* weight, const float* bias, int Cout, float* out,
             void* stream) {
    const int64_t n = (int64_t)nimg * Cout;
    if (n == 0) return VH_OK;
    hipLaunchKernelGGL(uo_fc_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, (hipStream_t)stream, in, nimg, Cin, weight,
                       bias, Cout, out);
    VH_LAUNCH_CHECK("uo_fc");
    return VH_OK;
}

}  // extern "C"
