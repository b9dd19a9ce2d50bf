// Factor-graph update glue around the update operator
// (vipe/slam/components/factor_graph.py:229-313, 315-392): the motion features
// before the network and the target / weight epilogue after it, one fused pass
// each, for a single-view pinhole graph.
//
// motion (factor_graph.py:247-260 with buffer.py:705-726, geom.py:187-263,
// cameras.py:131-181):
//   X0 = ((u - cx)/fx, (v - cy)/fy, 1, d_ii),  X1 = R_ij X0[:3] + t_ij d_ii,
//   T_ij = T_jj T_ii^-1, intrinsics already at the feature resolution;
//   z = X1.z < 0.1 ? 1 : X1.z;  coords1 = (fx X1.x / z + cx, fy X1.y / z + cy);
//   motn = clamp([coords1 - (u, v), target - coords1], -64, 64) as [E,4,H,W].
// targets (factor_graph.py:268-277):
//   target = coords1 + delta;  weight = mask[frame] ? 0 : weight_net;
// written as the graph state [E,H,W,2] and, optionally, in the BA layout
// [E,2,H,W] that vh_ba_run reads, so no rearrangement follows.
// One lane per (edge, pixel); the edge's relative pose is computed once per
// workgroup into LDS.

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cstdint>

#include "vh_common.hpp"

namespace vh {
namespace {

constexpr float kReprojMinDepth = 0.1f;   // BaseCameraModel.MIN_DEPTH (cameras.py:48)

__device__ __forceinline__ void quat_to_rot(float x, float y, float z, float w, float* R) {
    R[0] = 1 - 2 * (y * y + z * z); R[1] = 2 * (x * y - w * z);     R[2] = 2 * (x * z + w * y);
    R[3] = 2 * (x * y + w * z);     R[4] = 1 - 2 * (x * x + z * z); R[5] = 2 * (y * z - w * x);
    R[6] = 2 * (x * z - w * y);     R[7] = 2 * (y * z + w * x);     R[8] = 1 - 2 * (x * x + y * y);
}

// T_j T_i^-1 as (R, t): R = R_j R_i^T, t = t_j - R t_i (quaternions composed first).
__device__ void relative_pose(const float* __restrict__ poses, int64_t i, int64_t j, float* Rt) {
    const float* Pi = poses + 7 * i;
    const float* Pj = poses + 7 * j;
    // q = q_j * conj(q_i)
    const float ax = Pj[3], ay = Pj[4], az = Pj[5], aw = Pj[6];
    const float bx = -Pi[3], by = -Pi[4], bz = -Pi[5], bw = Pi[6];
    const float qx = aw * bx + ax * bw + ay * bz - az * by;
    const float qy = aw * by - ax * bz + ay * bw + az * bx;
    const float qz = aw * bz + ax * by - ay * bx + az * bw;
    const float qw = aw * bw - ax * bx - ay * by - az * bz;
    quat_to_rot(qx, qy, qz, qw, Rt);
    for (int r = 0; r < 3; r++) Rt[9 + r] = Pj[r] - (Rt[3 * r] * Pi[0] + Rt[3 * r + 1] * Pi[1] + Rt[3 * r + 2] * Pi[2]);
}

// grid (ceil(HW/256), E)
__global__ __launch_bounds__(256) void graph_motion_kernel(const float* __restrict__ poses,
                                                           const float* __restrict__ disps,
                                                           const float* __restrict__ intr,
                                                           const int64_t* __restrict__ ii,
                                                           const int64_t* __restrict__ jj,
                                                           const float* __restrict__ target, int N, int H, int W,
                                                           float* __restrict__ coords1, float* __restrict__ motn) {
    __shared__ float sRt[12];
    __shared__ int sbad;
    const int e = blockIdx.y;
    const int64_t HW = (int64_t)H * W;
    if (threadIdx.x == 0) {
        const int64_t a = ii[e], b = jj[e];
        sbad = (a < 0 || a >= N || b < 0 || b >= N);
        if (!sbad) relative_pose(poses, a, b, sRt);
    }
    __syncthreads();
    const int64_t pix = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= HW) return;
    const int v = (int)(pix / W), u = (int)(pix - (int64_t)v * W);
    const int64_t ep = (int64_t)e * HW + pix;
    float cu, cv;
    if (sbad) {
        cu = cv = __int_as_float(0x7fc00000);   // an out-of-range frame index: NaN, nothing read
    } else {
        const float fx = intr[0], fy = intr[1], cx = intr[2], cy = intr[3];
        const float d = disps[ii[e] * HW + pix];
        const float X = ((float)u - cx) / fx, Y = ((float)v - cy) / fy;
        const float x1 = sRt[0] * X + sRt[1] * Y + sRt[2] + sRt[9] * d;
        const float y1 = sRt[3] * X + sRt[4] * Y + sRt[5] + sRt[10] * d;
        float z1 = sRt[6] * X + sRt[7] * Y + sRt[8] + sRt[11] * d;
        if (z1 < kReprojMinDepth) z1 = 1.f;
        const float iz = 1.f / z1;
        cu = fx * (x1 * iz) + cx;
        cv = fy * (y1 * iz) + cy;
    }
    const float2 t = reinterpret_cast<const float2*>(target)[ep];
    reinterpret_cast<float2*>(coords1)[ep] = make_float2(cu, cv);
    float* m = motn + (int64_t)e * 4 * HW + pix;
    m[0] = fminf(fmaxf(cu - (float)u, -64.f), 64.f);
    m[HW] = fminf(fmaxf(cv - (float)v, -64.f), 64.f);
    m[2 * HW] = fminf(fmaxf(t.x - cu, -64.f), 64.f);
    m[3 * HW] = fminf(fmaxf(t.y - cv, -64.f), 64.f);
}

template <typename T> __device__ __forceinline__ float tof(T v) { return (float)v; }
template <> __device__ __forceinline__ float tof<__half>(__half v) { return __half2float(v); }

// grid (ceil(HW/256), E)
template <typename TD>
__global__ __launch_bounds__(256) void graph_targets_kernel(const float* __restrict__ coords1,
                                                            const TD* __restrict__ delta, const TD* __restrict__ wnet,
                                                            const uint8_t* __restrict__ mask,
                                                            const int64_t* __restrict__ frame, int NM, int H, int W,
                                                            float* __restrict__ target, float* __restrict__ weight,
                                                            float* __restrict__ target_ba,
                                                            float* __restrict__ weight_ba) {
    const int e = blockIdx.y;
    const int64_t HW = (int64_t)H * W;
    const int64_t pix = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= HW) return;
    const int64_t ep = (int64_t)e * HW + pix;
    const float2 c = reinterpret_cast<const float2*>(coords1)[ep];
    const float tx = c.x + tof(delta[2 * ep]), ty = c.y + tof(delta[2 * ep + 1]);
    float wx = tof(wnet[2 * ep]), wy = tof(wnet[2 * ep + 1]);
    if (mask) {
        const int64_t f = frame[e];
        if (f >= 0 && f < NM && mask[f * HW + pix]) wx = wy = 0.f;
    }
    reinterpret_cast<float2*>(target)[ep] = make_float2(tx, ty);
    reinterpret_cast<float2*>(weight)[ep] = make_float2(wx, wy);
    if (target_ba) {
        target_ba[(int64_t)e * 2 * HW + pix] = tx;
        target_ba[(int64_t)e * 2 * HW + HW + pix] = ty;
    }
    if (weight_ba) {
        weight_ba[(int64_t)e * 2 * HW + pix] = wx;
        weight_ba[(int64_t)e * 2 * HW + HW + pix] = wy;
    }
}

}  // namespace
}  // namespace vh

using namespace vh;

extern "C" {

int vh_graph_motion(const float* poses, const float* disps, const float* intrinsics, const int64_t* ii,
                    const int64_t* jj, const float* target, int E, int N, int H, int W, float* coords1, float* motn,
                    void* stream) {
    if (E < 0 || N < 0 || H <= 0 || W <= 0) return fail(VH_ERR_ARG, "graph_motion: bad sizes");
    if (E == 0) return VH_OK;
    if (!poses || !disps || !intrinsics || !ii || !jj || !target || !coords1 || !motn)
        return fail(VH_ERR_ARG, "graph_motion: null pointer");
    if (E > 65535) return fail(VH_ERR_UNSUPPORTED, "graph_motion: more than 65535 edges");
    const int64_t HW = (int64_t)H * W;
    hipLaunchKernelGGL(graph_motion_kernel, dim3((unsigned)((HW + 255) / 256), (unsigned)E), dim3(256), 0,
                       (hipStream_t)stream, poses, disps, intrinsics, ii, jj, target, N, H, W, coords1, motn);
    VH_LAUNCH_CHECK("graph_motion_kernel");
    return VH_OK;
}

int vh_graph_targets(int dtype, const float* coords1, const void* delta, const void* weight_net, const uint8_t* mask,
                     const int64_t* frame, int NM, int E, int H, int W, float* target, float* weight,
                     float* target_ba, float* weight_ba, void* stream) {
    if (E < 0 || H <= 0 || W <= 0 || NM < 0) return fail(VH_ERR_ARG, "graph_targets: bad sizes");
    if (E == 0) return VH_OK;
    if (!coords1 || !delta || !weight_net || !target || !weight || (mask && !frame))
        return fail(VH_ERR_ARG, "graph_targets: null pointer");
    if (E > 65535) return fail(VH_ERR_UNSUPPORTED, "graph_targets: more than 65535 edges");
    const int64_t HW = (int64_t)H * W;
    const dim3 grid((unsigned)((HW + 255) / 256), (unsigned)E);
    hipStream_t s = (hipStream_t)stream;
    if (dtype == VH_F32)
        hipLaunchKernelGGL(graph_targets_kernel<float>, grid, dim3(256), 0, s, coords1, (const float*)delta,
                           (const float*)weight_net, mask, frame, NM, H, W, target, weight, target_ba, weight_ba);
    else if (dtype == VH_F16)
        hipLaunchKernelGGL(graph_targets_kernel<__half>, grid, dim3(256), 0, s, coords1, (const __half*)delta,
                           (const __half*)weight_net, mask, frame, NM, H, W, target, weight, target_ba, weight_ba);
    else
        return fail(VH_ERR_UNSUPPORTED, "graph_targets: dtype %d", dtype);
    VH_LAUNCH_CHECK("graph_targets_kernel");
    return VH_OK;
}

}  // extern "C"
