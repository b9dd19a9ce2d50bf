// slam_ext geometry siblings of the BA (csrc/slam_ext/slam.cpp:52-55):
// iproj (back-projection), projmap (reprojection coordinates and validity),
// frame_distance (mean flow magnitude between frame pairs, used for edge
// selection) and depth_filter (multi-view consistency counts).  Semantics follow
// geom_kernels.cu:526-953 and :1849-1950; the kernels are laid out for gfx950:
//   * one lane per pixel, pixels contiguous across a wave (coalesced
//     disparity reads and output writes);
//   * frame poses are per-workgroup uniform and read through the scalar path;
//   * depth_filter visits its six neighbour frames inside one lane and writes
//     each count once (no atomics, deterministic);
//   * frame_distance reduces per pair in FP64 (wave shuffles, then LDS).
// Index tensors are bounds-checked on the device: an out-of-range frame
// index gives NaN outputs (iproj / projmap / frame_distance) or a zero count
// (depth_filter) for that item instead of reading outside the tensors.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "vh_common.hpp"

namespace vh {
namespace {

constexpr float kMinDepth = 0.25f;  // MIN_DEPTH (geom_kernels.cu:33)

struct Pose {
    float t[3], q[4];  // translation, unit quaternion (x, y, z, w)
};

__device__ __forceinline__ Pose load_pose(const float* __restrict__ poses, int i) {
    const float* p = poses + 7 * (int64_t)i;
    Pose o;
    o.t[0] = p[0]; o.t[1] = p[1]; o.t[2] = p[2];
    o.q[0] = p[3]; o.q[1] = p[4]; o.q[2] = p[5]; o.q[3] = p[6];
    return o;
}

// Rotate X by unit quaternion q: with u = 2 (q_v x X), X + w u + q_v x u.
__device__ __forceinline__ void rotate(const float* q, const float* X, float* Y) {
    const float u0 = 2.0f * (q[1] * X[2] - q[2] * X[1]);
    const float u1 = 2.0f * (q[2] * X[0] - q[0] * X[2]);
    const float u2 = 2.0f * (q[0] * X[1] - q[1] * X[0]);
    Y[0] = X[0] + q[3] * u0 + (q[1] * u2 - q[2] * u1);
    Y[1] = X[1] + q[3] * u1 + (q[2] * u0 - q[0] * u2);
    Y[2] = X[2] + q[3] * u2 + (q[0] * u1 - q[1] * u0);
}

// T_ij = T_j T_i^-1 (relSE3, geom_kernels.cu:101-114).
__device__ __forceinline__ Pose relative(const Pose& a, const Pose& b) {
    Pose r;
    const float* qi = a.q;
    const float* qj = b.q;
    r.q[0] = -qj[3] * qi[0] + qj[0] * qi[3] - qj[1] * qi[2] + qj[2] * qi[1];
    r.q[1] = -qj[3] * qi[1] + qj[1] * qi[3] - qj[2] * qi[0] + qj[0] * qi[2];
    r.q[2] = -qj[3] * qi[2] + qj[2] * qi[3] - qj[0] * qi[1] + qj[1] * qi[0];
    r.q[3] = qj[3] * qi[3] + qj[0] * qi[0] + qj[1] * qi[1] + qj[2] * qi[2];
    float rt[3];
    rotate(r.q, a.t, rt);
    r.t[0] = b.t[0] - rt[0];
    r.t[1] = b.t[1] - rt[1];
    r.t[2] = b.t[2] - rt[2];
    return r;
}

// Homogeneous point (X, 1, d) mapped by T: (R X + t d, d)  (actSE3, :78-84).
__device__ __forceinline__ void transform(const Pose& T, const float* X, float d, float* Y) {
    rotate(T.q, X, Y);
    Y[0] += d * T.t[0];
    Y[1] += d * T.t[1];
    Y[2] += d * T.t[2];
}

// ------------------------------------------------------------------ iproj
__global__ __launch_bounds__(256) void iproj_kernel(const float* __restrict__ poses, const float* __restrict__ disps,
                                                    const float* __restrict__ intr, int HW, int W,
                                                    float* __restrict__ points) {
    const int f = blockIdx.y;
    const int pix = blockIdx.x * 256 + threadIdx.x;
    if (pix >= HW) return;
    const Pose T = load_pose(poses, f);
    const float fx = intr[0], fy = intr[1], cx = intr[2], cy = intr[3];
    const int v = pix / W, u = pix - v * W;
    const float X[3] = {((float)u - cx) / fx, ((float)v - cy) / fy, 1.0f};
    const float d = disps[(int64_t)f * HW + pix];
    float Y[3];
    transform(T, X, d, Y);
    float* o = points + ((int64_t)f * HW + pix) * 3;
    o[0] = Y[0] / d;
    o[1] = Y[1] / d;
    o[2] = Y[2] / d;
}

// ---------------------------------------------------------------- projmap
__global__ __launch_bounds__(256) void projmap_kernel(const float* __restrict__ poses, const float* __restrict__ disps,
                                                      const float* __restrict__ intr, const int64_t* __restrict__ ii,
                                                      const int64_t* __restrict__ jj, int N, int HW, int W,
                                                      float* __restrict__ coords, float* __restrict__ valid) {
    const int e = blockIdx.y;
    const int pix = blockIdx.x * 256 + threadIdx.x;
    if (pix >= HW) return;
    const int64_t a = ii[e], b = jj[e];
    float* c = coords + ((int64_t)e * HW + pix) * 3;
    float* vo = valid + (int64_t)e * HW + pix;
    if (a < 0 || a >= N || b < 0 || b >= N) {
        c[0] = c[1] = c[2] = NAN;
        *vo = 0.0f;
        return;
    }
    const Pose T = relative(load_pose(poses, (int)a), load_pose(poses, (int)b));
    const float fx = intr[0], fy = intr[1], cx = intr[2], cy = intr[3];
    const int v = pix / W, u = pix - v * W;
    const float X[3] = {((float)u - cx) / fx, ((float)v - cy) / fy, 1.0f};
    float Y[3];
    transform(T, X, disps[a * HW + pix], Y);
    // pixels whose point ends up behind z = 0.01 keep their own coordinates
    const bool front = Y[2] > 0.01f;
    c[0] = front ? fx * (Y[0] / Y[2]) + cx : (float)u;
    c[1] = front ? fy * (Y[1] / Y[2]) + cy : (float)v;
    c[2] = 0.0f;  // the third channel is allocated but never set (zeros)
    *vo = (Y[2] > kMinDepth) ? 1.0f : 0.0f;
}

// --------------------------------------------------------- frame_distance
// One workgroup per pair b: frames (pi, pj), disparity map di, intrinsics rows
// (qi, qj).  Each pixel adds, with weight beta, its full-motion flow magnitude
// and, with weight 1 - beta, its translation-only flow magnitude; a pixel term
// counts as valid when its point lies in front of MIN_DEPTH.  dist = 1000 when
// the valid weight is below 0.75 of the total, else the valid-weighted mean
// (geom_kernels.cu:613-768).
__global__ __launch_bounds__(256) void frame_distance_kernel(
    const float* __restrict__ poses, const float* __restrict__ disps, const float* __restrict__ intr, int R,
    const int64_t* __restrict__ pi, const int64_t* __restrict__ pj, const int64_t* __restrict__ qi,
    const int64_t* __restrict__ qj, const int64_t* __restrict__ di, int N, int HW, int W, float beta,
    float* __restrict__ dist) {
    __shared__ double red[3][4];
    const int b = blockIdx.x;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int64_t fi = pi[b], fj = pj[b], fd = di[b], ri = qi[b], rj = qj[b];
    if (fi < 0 || fi >= N || fj < 0 || fj >= N || fd < 0 || fd >= N || ri < 0 || ri >= R || rj < 0 || rj >= R) {
        if (tid == 0) dist[b] = NAN;
        return;
    }
    const Pose T = relative(load_pose(poses, (int)fi), load_pose(poses, (int)fj));
    const float* Ii = intr + 4 * ri;
    const float* Ij = intr + 4 * rj;
    const float fxi = Ii[0], fyi = Ii[1], cxi = Ii[2], cyi = Ii[3];
    const float fxj = Ij[0], fyj = Ij[1], cxj = Ij[2], cyj = Ij[3];
    const float* dmap = disps + fd * HW;
    double acc = 0.0, val = 0.0, tot = 0.0;
    for (int pix = tid; pix < HW; pix += 256) {
        const int v = pix / W, u = pix - v * W;
        const float X[3] = {((float)u - cxi) / fxi, ((float)v - cyi) / fyi, 1.0f};
        const float d = dmap[pix];
        float Y[3];
        transform(T, X, d, Y);
        float du = fxj * (Y[0] / Y[2]) + cxj - (float)u;
        float dv = fyj * (Y[1] / Y[2]) + cyj - (float)v;
        const float m_full = sqrtf(du * du + dv * dv);
        tot += beta;
        if (Y[2] > kMinDepth) {
            acc += (double)(beta * m_full);
            val += beta;
        }
        // translation only
        const float Z0 = X[0] + d * T.t[0], Z1 = X[1] + d * T.t[1], Z2 = X[2] + d * T.t[2];
        du = fxj * (Z0 / Z2) + cxj - (float)u;
        dv = fyj * (Z1 / Z2) + cyj - (float)v;
        const float m_tr = sqrtf(du * du + dv * dv);
        tot += (1.0f - beta);
        if (Z2 > kMinDepth) {
            acc += (double)((1.0f - beta) * m_tr);
            val += (1.0f - beta);
        }
    }
    for (int o = 32; o >= 1; o >>= 1) {
        acc += __shfl_xor(acc, o, 64);
        val += __shfl_xor(val, o, 64);
        tot += __shfl_xor(tot, o, 64);
    }
    if (lane == 0) {
        red[0][wv] = acc;
        red[1][wv] = val;
        red[2][wv] = tot;
    }
    __syncthreads();
    if (tid == 0) {
        const double A = (red[0][0] + red[0][1]) + (red[0][2] + red[0][3]);
        const double V = (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]);
        const double S = (red[2][0] + red[2][1]) + (red[2][2] + red[2][3]);
        dist[b] = (V / (S + 1e-8) < 0.75) ? 1000.0f : (float)(A / V);
    }
}

// ----------------------------------------------------------- depth_filter
// Lane = pixel of frame ix = inds[b].  For each neighbour frame
// jx in {ix-1, ix-2, ix-3, ix+1, ix+2, ix+3} inside [0, N): reproject the pixel
// into jx; when the projection's top-left integer corner (u0, v0) has
// 0 <= u0 < W-1 and 0 <= v0 < H-1, the count grows by one if the projected
// depth 1/dj agrees within thresh[b] with the depth of any of the four corner
// disparities (tested in double precision, as the reference does)
// (geom_kernels.cu:770-885).
__global__ __launch_bounds__(256) void depth_filter_kernel(const float* __restrict__ poses,
                                                           const float* __restrict__ disps,
                                                           const float* __restrict__ intr,
                                                           const int64_t* __restrict__ inds,
                                                           const float* __restrict__ thresh, int N, int H, int W,
                                                           float* __restrict__ counter) {
    const int b = blockIdx.y;
    const int HW = H * W;
    const int pix = blockIdx.x * 256 + threadIdx.x;
    if (pix >= HW) return;
    const int64_t ix = inds[b];
    float* out = counter + (int64_t)b * HW + pix;
    if (ix < 0 || ix >= N) {
        *out = 0.0f;
        return;
    }
    const double t = (double)thresh[b];
    const float fx = intr[0], fy = intr[1], cx = intr[2], cy = intr[3];
    const int v = pix / W, u = pix - v * W;
    const float X[3] = {((float)u - cx) / fx, ((float)v - cy) / fy, 1.0f};
    const float d = disps[ix * HW + pix];
    const Pose Pi = load_pose(poses, (int)ix);
    float count = 0.0f;
#pragma unroll 1
    for (int n = 0; n < 6; n++) {
        const int64_t jx = (n < 3) ? ix - n - 1 : ix + n - 2;
        if (jx < 0 || jx >= N) continue;
        const Pose T = relative(Pi, load_pose(poses, (int)jx));
        float Y[3];
        transform(T, X, d, Y);
        const float uj = fx * (Y[0] / Y[2]) + cx;
        const float vj = fy * (Y[1] / Y[2]) + cy;
        const float dj = d / Y[2];
        const int u0 = (int)floorf(uj), v0 = (int)floorf(vj);
        if (!(u0 >= 0 && v0 >= 0 && u0 < W - 1 && v0 < H - 1)) continue;
        const float* dm = disps + jx * HW + (int64_t)v0 * W + u0;
        const double inv = 1.0 / (double)dj;
        const bool hit = fabs(inv - 1.0 / (double)dm[0]) < t || fabs(inv - 1.0 / (double)dm[1]) < t ||
                         fabs(inv - 1.0 / (double)dm[W]) < t || fabs(inv - 1.0 / (double)dm[W + 1]) < t;
        count += hit ? 1.0f : 0.0f;
    }
    *out = count;
}

}  // namespace
}  // namespace vh

using namespace vh;

extern "C" {

int vh_geom_iproj(const float* poses, const float* disps, const float* intrinsics, int N, int H, int W, float* points,
                  void* stream) {
    if (N < 0 || H <= 0 || W <= 0) return fail(VH_ERR_ARG, "iproj: bad sizes N=%d H=%d W=%d", N, H, W);
    if (N == 0) return VH_OK;
    if (!poses || !disps || !intrinsics || !points) return fail(VH_ERR_ARG, "iproj: null pointer");
    const int HW = H * W;
    hipLaunchKernelGGL(iproj_kernel, dim3(ceil_div(HW, 256), N), dim3(256), 0, (hipStream_t)stream, poses, disps,
                       intrinsics, HW, W, points);
    VH_LAUNCH_CHECK("iproj");
    return VH_OK;
}

int vh_geom_projmap(const float* poses, const float* disps, const float* intrinsics, const int64_t* ii,
                    const int64_t* jj, int E, int N, int H, int W, float* coords, float* valid, void* stream) {
    if (E < 0 || N <= 0 || H <= 0 || W <= 0) return fail(VH_ERR_ARG, "projmap: bad sizes E=%d N=%d H=%d W=%d", E, N, H, W);
    if (E == 0) return VH_OK;
    if (!poses || !disps || !intrinsics || !ii || !jj || !coords || !valid)
        return fail(VH_ERR_ARG, "projmap: null pointer");
    const int HW = H * W;
    hipLaunchKernelGGL(projmap_kernel, dim3(ceil_div(HW, 256), E), dim3(256), 0, (hipStream_t)stream, poses, disps,
                       intrinsics, ii, jj, N, HW, W, coords, valid);
    VH_LAUNCH_CHECK("projmap");
    return VH_OK;
}

int vh_geom_frame_distance(const float* poses, const float* disps, const float* intrinsics, int n_intrinsics,
                           const int64_t* pi, const int64_t* pj, const int64_t* qi, const int64_t* qj,
                           const int64_t* di, int num, int N, int H, int W, float beta, float* dist, void* stream) {
    if (num < 0 || N <= 0 || H <= 0 || W <= 0 || n_intrinsics <= 0)
        return fail(VH_ERR_ARG, "frame_distance: bad sizes num=%d N=%d H=%d W=%d intrinsics rows=%d", num, N, H, W,
                    n_intrinsics);
    if (num == 0) return VH_OK;
    if (!poses || !disps || !intrinsics || !pi || !pj || !qi || !qj || !di || !dist)
        return fail(VH_ERR_ARG, "frame_distance: null pointer");
    hipLaunchKernelGGL(frame_distance_kernel, dim3(num), dim3(256), 0, (hipStream_t)stream, poses, disps, intrinsics,
                       n_intrinsics, pi, pj, qi, qj, di, N, H * W, W, beta, dist);
    VH_LAUNCH_CHECK("frame_distance");
    return VH_OK;
}

int vh_geom_depth_filter(const float* poses, const float* disps, const float* intrinsics, const int64_t* inds,
                         const float* thresh, int num, int N, int H, int W, float* counter, void* stream) {
    if (num < 0 || N <= 0 || H <= 0 || W <= 0)
        return fail(VH_ERR_ARG, "depth_filter: bad sizes num=%d N=%d H=%d W=%d", num, N, H, W);
    if (num == 0) return VH_OK;
    if (!poses || !disps || !intrinsics || !inds || !thresh || !counter)
        return fail(VH_ERR_ARG, "depth_filter: null pointer");
    hipLaunchKernelGGL(depth_filter_kernel, dim3(ceil_div(H * W, 256), num), dim3(256), 0, (hipStream_t)stream, poses,
                       disps, intrinsics, inds, thresh, N, H, W, counter);
    VH_LAUNCH_CHECK("depth_filter");
    return VH_OK;
}

}  // extern "C"
