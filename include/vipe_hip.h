/*
 * libvipe_hip — C ABI of the MI355X (gfx950) SLAM inner loop.
 *
 * Plain pointers and sizes only.  Every compute entry point takes the HIP
 * stream to launch on (a hipStream_t passed as void*; NULL = default stream),
 * never synchronises unless stated, and returns VH_OK or an error code whose
 * text is available from vh_last_error().  Tensors are dense row-major
 * (C-contiguous) buffers in device memory unless the argument says host.
 *
 * Each declaration names the reference interface it replaces
 * (nv-tlabs/vipe, paths relative to its source tree).
 */
#ifndef VIPE_HIP_H
#define VIPE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status / dtype / placement codes ---------------------------------- */
#define VH_OK 0
#define VH_ERR_ARG 1         /* bad argument: shape, range, null pointer      */
#define VH_ERR_HIP 2         /* HIP runtime error (text in vh_last_error)     */
#define VH_ERR_UNSUPPORTED 3 /* dtype / size outside what a kernel handles    */
#define VH_ERR_UNDEFINED 4   /* reference returns an undefined tensor here    */

#define VH_F16 1
#define VH_F32 2
#define VH_F64 3
#define VH_BF16 4
#define VH_FP8E4M3 5
#define VH_U8 6
#define VH_I8 7
#define VH_I16 8
#define VH_I32 9
#define VH_I64 10

/* volume plane layouts (vh_corr_build / vh_corr_lookup_pyramid) */
#define VH_LAYOUT_PLANAR 0 /* [Hl][Wl] row-major: the reference's */
#define VH_LAYOUT_TILED 1  /* each plane padded to [ceil(Hl/2)*2][ceil(Wl/tw)*tw] (tw = 4 for F32, 8 for
                            * F16; padding zero); planes grouped by 4 (p / 4), a group stored as row-major
                            * 128-byte tiles of rows 2a, 2a+1 x columns tw b .. tw b + tw-1 of its planes,
                            * 16-byte piece (row % 2) * 4 + p % 4; needs H*W per edge a multiple of 4 */

#define VH_DEVICE 0 /* pointers are device memory, kernels on `stream` */
#define VH_HOST 1   /* pointers are host memory, computed on the CPU   */

/* Text of the last error raised on the calling thread. */
const char* vh_last_error(void);
/* ABI version; bumped when a signature changes. */
int vh_abi_version(void);
/* Stamp kernel: writes `value` into out[0..n) on `stream` (plumbing probe). */
int vh_fill_i32(int32_t* out, int64_t n, int32_t value, void* stream);

/* ---- lietorch_ext (csrc/lietorch_ext/lietorch.cpp:23-335) ---------------
 * group: 1 SO3 [qx,qy,qz,qw], 2 RxSO3 [q,s], 3 SE3 [t,q], 4 Sim3 [t,q,s]
 *        (vipe/ext/lietorch/groups.py:244,260,276,298).
 * dtype: VH_F32 or VH_F64.  place: VH_DEVICE or VH_HOST (the reference's
 * lietorch_cpu.cpp path for CPU tensors).  n = number of elements.
 * Embedding N and tangent K per group: (4,3) (5,4) (7,6) (8,7).
 * Gradients w.r.t. group elements are K-vectors written at stride N into
 * buffers the caller zero-fills; upstream gradients of group outputs are read
 * at stride N (lietorch_gpu.cu:38-46, 95-98).  An unknown group id or dtype
 * returns VH_ERR_UNDEFINED (the reference returns an undefined tensor,
 * dispatch.h:40-50).
 */
int vh_lie_expm(int group, int dtype, int64_t n, const void* a, void* X, int place, void* stream);          /* lietorch.cpp:23  expm          */
int vh_lie_expm_backward(int group, int dtype, int64_t n, const void* grad, const void* a, void* da, int place, void* stream); /* :36 expm_backward */
int vh_lie_logm(int group, int dtype, int64_t n, const void* X, void* a, int place, void* stream);          /* :50  logm          */
int vh_lie_logm_backward(int group, int dtype, int64_t n, const void* grad, const void* X, void* dX, int place, void* stream); /* :63 logm_backward */
int vh_lie_inv(int group, int dtype, int64_t n, const void* X, void* Y, int place, void* stream);           /* :78  inv           */
int vh_lie_inv_backward(int group, int dtype, int64_t n, const void* grad, const void* X, void* dX, int place, void* stream); /* :91 inv_backward */
int vh_lie_mul(int group, int dtype, int64_t n, const void* X, const void* Y, void* Z, int place, void* stream); /* :108 mul       */
int vh_lie_mul_backward(int group, int dtype, int64_t n, const void* grad, const void* X, const void* Y, void* dX, void* dY, int place, void* stream); /* :123 */
int vh_lie_adj(int group, int dtype, int64_t n, const void* X, const void* a, void* b, int place, void* stream); /* :139 adj        */
int vh_lie_adj_backward(int group, int dtype, int64_t n, const void* grad, const void* X, const void* a, void* dX, void* da, int place, void* stream); /* :154 */
int vh_lie_adjT(int group, int dtype, int64_t n, const void* X, const void* a, void* b, int place, void* stream); /* :170 adjT      */
int vh_lie_adjT_backward(int group, int dtype, int64_t n, const void* grad, const void* X, const void* a, void* dX, void* da, int place, void* stream); /* :185 */
int vh_lie_act(int group, int dtype, int64_t n, const void* X, const void* p, void* q, int place, void* stream); /* :201 act (3-pt) */
int vh_lie_act_backward(int group, int dtype, int64_t n, const void* grad, const void* X, const void* p, void* dX, void* dp, int place, void* stream); /* :216 */
int vh_lie_act4(int group, int dtype, int64_t n, const void* X, const void* p, void* q, int place, void* stream); /* :232 act4 */
int vh_lie_act4_backward(int group, int dtype, int64_t n, const void* grad, const void* X, const void* p, void* dX, void* dp, int place, void* stream); /* :247 */
int vh_lie_as_matrix(int group, int dtype, int64_t n, const void* X, void* T, int place, void* stream);     /* :277 as_matrix (row-major 4x4) */
int vh_lie_projector(int group, int dtype, int64_t n, const void* X, void* Pm, int place, void* stream);    /* :263 projector (row-major NxN) */
int vh_lie_Jinv(int group, int dtype, int64_t n, const void* X, const void* a, void* b, int place, void* stream); /* :290 Jinv */

/* ---- droid_net_ext correlation (csrc/droid_net_ext/droid.cpp:24-82) ----- */

/* corr_index_forward (droid.cpp:25, correlation_kernels.cu:22-66, 115-136).
 * volume [B,H1,W1,H2,W2] (dtype F16/F32/F64), coords [B,2,H1,W1] f32,
 * corr out [B,2r+1,2r+1,H1,W1] (dtype of volume), fully written. */
int vh_corr_index_forward(int dtype, int64_t B, int H1, int W1, int H2, int W2, const void* volume,
                          const float* coords, int radius, void* corr, void* stream);
/* corr_index_backward (droid.cpp:32, correlation_kernels.cu:68-113, 138-158).
 * volume_grad [B,H1,W1,H2,W2] must be zero-filled by the caller (the reference
 * allocates it with zeros_like); the kernel writes the gradient of every tap. */
int vh_corr_index_backward(int dtype, int64_t B, int H1, int W1, int H2, int W2, const float* coords,
                           const void* corr_grad, int radius, void* volume_grad, void* stream);

/* Fused pyramid lookup: CorrBlock.__call__ (vipe/slam/networks/droid_net.py:71-84,
 * i.e. the per-level corr_index_forward calls plus their concatenation).
 * levels[l] [B,H1,W1,<plane of Hl[l] x Wl[l]>] in `layout` (dtype F16/F32/F64;
 * tiled: F16/F32), coords [B,H1,W1,2] f32 at level-0 resolution (level l
 * samples at coords / 2^l), out [B, L*(2r+1)^2, H1, W1] in the volume dtype,
 * channel l*(2r+1)^2 + i*(2r+1) + j, fully written. */
int vh_corr_lookup_pyramid(int dtype, int64_t B, int H1, int W1, int num_levels, const void* const* levels,
                           const int* Hl, const int* Wl, int layout, const float* coords, int radius, void* out,
                           void* stream);

/* All-pairs correlation volume and its avg-pool pyramid: the CorrBlock
 * constructor (vipe/slam/networks/droid_net.py:56-69, 94-102), which the
 * reference forms with torch.matmul + F.avg_pool2d (no native counterpart).
 * fmap1, fmap2 [E,C,H,W] in `in_dtype` (F32, F16, BF16), C <= 448; levels[l]
 * is the level-l output [E,H,W,H>>l,W>>l] in `out_dtype` (F32 or F16).
 * Values are (f1/4).(f2/4) summed over C with FP32 accumulation; level l > 0
 * is the 2x2 mean of level l-1 as stored (floor of odd sizes).
 * `mma_dtype` selects the contraction: F32, F16, BF16, or FP8E4M3 (each
 * channel balanced between the two maps by sqrt(amax1/amax2), then one e4m3
 * scale per pixel row; the scales are applied to the FP32 accumulator).
 * `layout` VH_LAYOUT_PLANAR or VH_LAYOUT_TILED (at most 4 levels) sets the
 * plane layout of every level.  The caller provides
 * vh_corr_build_workspace(mma_dtype, E, C, H, W) bytes of device memory
 * (channels-last operand copies and scales). */
size_t vh_corr_build_workspace(int mma_dtype, int64_t E, int C, int H, int W);
int vh_corr_build(int in_dtype, int mma_dtype, int out_dtype, int64_t E, int C, int H, int W, const void* fmap1,
                  const void* fmap2, int num_levels, void* const* levels, int layout, void* workspace, void* stream);

/* altcorr_forward (droid.cpp:41, altcorr_kernel.cu:26-138, 368-392).
 * fmap1 [B,H1,W1,C], fmap2 [B,H2,W2,C] (F16/F32/F64; sums in FP64 for F64),
 * coords [B,N,H1,W1,2] f32, corr out [B,N,(2r+1)^2,H1,W1] in the fmap dtype,
 * fully written. */
int vh_altcorr_forward(int dtype, int64_t B, int N, int H1, int W1, int H2, int W2, int C, const void* fmap1,
                       const void* fmap2, const float* coords, int radius, void* corr, void* stream);
/* altcorr_backward (droid.cpp:62, altcorr_kernel.cu:242-366, 426-455); F32 only
 * like the reference.  fmap1_grad [B,H1,W1,C], fmap2_grad [B,H2,W2,C] out, fully
 * written whatever N is, zeros when N == 0; coords_grad is not produced (the
 * reference returns zeros).  With a workspace of
 * vh_altcorr_backward_workspace_size bytes (device) the f2 gradient is gathered
 * per target row with no atomics, the same on every run (C a multiple of 4,
 * 16-byte aligned fmap1 and fmap2_grad, a row of W2 x C floats within LDS,
 * N x H1 x W1 below 2^31);
 * otherwise (workspace NULL) it is cleared and accumulated with FP32 atomic adds
 * like the reference, whose summation order varies from run to run. */
size_t vh_altcorr_backward_workspace_size(int64_t B, int N, int H1, int W1);
int vh_altcorr_backward(int dtype, int64_t B, int N, int H1, int W1, int H2, int W2, int C, const void* fmap1,
                        const void* fmap2, const float* coords, const void* corr_grad, int radius, void* fmap1_grad,
                        void* fmap2_grad, void* workspace, void* stream);
/* Feature maps for the on-the-fly correlation and the feature gradients of the
 * volume lookup.  vh_feat_to_nhwc: in [E][C][HW] (F32, F64, F16 or BF16) times
 * scale -> out [E][HW][C] in out_dtype (F32 or dtype; AltCorrBlock's fmaps / 4
 * in channels-last, droid_net.py:146-160).  vh_pool2x2_nhwc: [E][H][W][C] ->
 * [E][H/2][W/2][C] (F32, F64, F16 or BF16), each output the mean of its 2x2
 * block taken in FP32 (FP64 for F64) (avg_pool2d(2, stride 2) on channels-last
 * maps, droid_net.py:157).  vh_corr_feature_grad: the gradients
 * of the two feature maps of CorrBlock's lookup (the adjoint of the volume,
 * its pooled pyramid and the lookup, droid_net.py:56-102, which the reference
 * gets from torch autograd), from the per-level altcorr backward outputs:
 * df1[e][c][y][x] = scale * sum_l g1[l][e][y][x][c] and
 * df2[e][c][y][x] = scale * sum_l 4^-l g2[l][e][y >> l][x >> l][c] (levels whose
 * (y >> l, x >> l) lie inside H >> l, W >> l), stored NCHW in dtype (F32, F16 or
 * BF16); g1[l] f32 [E][H][W][C], g2[l] f32 [E][H >> l][W >> l][C], 1 to 4 levels.
 * E is at most 65535. */
int vh_feat_to_nhwc(int dtype, const void* in, int64_t E, int C, int HW, float scale, int out_dtype, void* out,
                    void* stream);
int vh_pool2x2_nhwc(int dtype, const void* in, int64_t E, int H, int W, int C, void* out, void* stream);
int vh_corr_feature_grad(int dtype, int64_t E, int C, int H, int W, int num_levels, const float* const* g1,
                         const float* const* g2, float scale, void* df1, void* df2, void* stream);
/* altcorr_index_forward (droid.cpp:50, altcorr_kernel.cu:140-240, 394-424).
 * fmap1 [B,Nf,H,W,C]; levels[l] [B,Nf,Hl[l],Wl[l],C]; coords [B,M,H,W,2] f32;
 * ii, jj [M] int64 (device); features F16/F32/F64 (sums in FP64 for F64);
 * corr out [B,M,L*(2r+1)^2,H,W] f32, fully written. */
int vh_altcorr_index_forward(int dtype, int64_t B, int Nf, int M, int H, int W, int C, const void* fmap1,
                             int num_levels, const void* const* levels, const int* Hl, const int* Wl,
                             const float* coords, const int64_t* ii, const int64_t* jj, int radius, void* corr,
                             void* stream);

/* ---- slam_ext dense bundle adjustment (csrc/slam_ext/slam.cpp:41-55,
 *      geom_kernels.cu:1626-1847) ---------------------------------------- */

typedef struct vh_ba_plan vh_ba_plan;

/* Builds the graph bookkeeping of ba_cuda_impl once per edge list:
 * kx = unique(cat(arange(t0,t1), ii)) and its inverse, the per-depth-frame
 * edge lists, the Schur row pairs and the block-assembly lists
 * (geom_kernels.cu:1648-1655, 1437-1468).  ii/jj are HOST arrays.  Allocates
 * and uploads device index tables (synchronously).  The window [t0, t1) holds
 * at most 1023 poses (a whole 1024-keyframe buffer; the reduced system is
 * dense, (6 or 7) x P square).  Returns NULL on error. */
vh_ba_plan* vh_ba_plan_create(const int64_t* ii_host, const int64_t* jj_host, int E, int N, int H, int W, int t0,
                              int t1);
/* vh_ba_plan_create with the frame kernel's pixel tile chosen by the caller
 * (no counterpart in the reference; for autotuners that time the candidates):
 * pixel_tile is 64, 128 or 256 pixels per pass, or 0 for the automatic choice.
 * Returns NULL when the tile's E rows do not fit the workgroup's LDS. */
vh_ba_plan* vh_ba_plan_create_tiled(const int64_t* ii_host, const int64_t* jj_host, int E, int N, int H, int W,
                                    int t0, int t1, int pixel_tile);
/* Pixels per pass of the plan's frame kernel. */
int vh_ba_plan_pixel_tile(const vh_ba_plan* plan);
void vh_ba_plan_destroy(vh_ba_plan* plan);
/* Number of depth frames |kx| and the frame ids (host copy, kx_out[|kx|]). */
int vh_ba_plan_num_depth_frames(const vh_ba_plan* plan);
int vh_ba_plan_depth_frames(const vh_ba_plan* plan, int64_t* kx_out);
/* Device workspace bytes vh_ba_run needs for this plan. */
size_t vh_ba_workspace_size(const vh_ba_plan* plan);

#define VH_BA_MOTION_ONLY 1
#define VH_BA_OPTIMIZE_INTRINSICS 2
#define VH_BA_COMPUTE_ENERGY 4
/* Sim3 poses: poses [N,8] (t, q xyzw, s), dx [P,7] (tau, phi, sigma) with
 * 7-dof pose blocks and the retraction T <- Exp(dx) T of lietorch's Sim3
 * (sim3.h; vipe/slam/maths/geom.py:147-182 gives the Jacobian of the point
 * map).  The reference's fused BA takes SE3 only; this flag extends it.  At
 * most 18 edges may leave one depth frame. */
#define VH_BA_SIM3 8

/* ba / ba_extended (slam.cpp:41-51).  All pointers device, f32 unless noted:
 * poses [N,7] (in place), disps [N,H,W] (in place), intrinsics [4]
 * (fx,fy,cx,cy; in place when optimising intrinsics), disps_sens [N,H,W],
 * targets/weights [E,2,H,W], eta [|kx|,H,W], depth_active [N].
 * Outputs: dx [P,6] (P = t1-t0), dz [|kx|,H*W] (untouched when motion only),
 * energy [iterations] when VH_BA_COMPUTE_ENERGY.  The reduced camera system
 * is factorised in FP64 on the device; a non-positive pivot gives dx = 0
 * (geom_kernels.cu:1283-1285).  `workspace` holds vh_ba_workspace_size bytes. */
int vh_ba_run(const vh_ba_plan* plan, float* poses, float* disps, float* intrinsics, const float* disps_sens,
              const float* targets, const float* weights, const float* eta, const float* depth_active,
              int iterations, float lm, float ep, float alpha, float intrinsics_lm, float intrinsics_ep,
              float intrinsics_scale, int flags, float* dx, float* dz, float* energy, void* workspace,
              void* stream);

/* Allocation-free vh_ba_run.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments and
 * results as vh_ba_run without `workspace`: the workspace is a file-scope
 * device array of VH_BA_STATIC_SCRATCH_BYTES that each GPU the library runs on
 * holds until the process ends (its address is looked up on the first call on
 * each GPU).  A plan whose vh_ba_workspace_size exceeds that limit is rejected
 * with VH_ERR_UNSUPPORTED.  The plan itself is created beforehand
 * (vh_ba_plan_create allocates).  All calls on one GPU share its array, so
 * calls that overlap on one GPU (other streams or threads) race: serialise
 * them.  vh_ba_adjoint cannot follow this entry point (its workspace is not
 * the caller's). */
#define VH_BA_STATIC_SCRATCH_BYTES ((size_t)64 << 20)
int vh_ba_run_static(const vh_ba_plan* plan, float* poses, float* disps, float* intrinsics, const float* disps_sens,
                     const float* targets, const float* weights, const float* eta, const float* depth_active,
                     int iterations, float lm, float ep, float alpha, float intrinsics_lm, float intrinsics_ep,
                     float intrinsics_scale, int flags, float* dx, float* dz, float* energy, void* stream);

/* Robust kernels of the generic solver (vipe/slam/ba/kernel.py:21-162,
 * build_robust_kernel :142-162), applied inside the fused BA: each flow
 * residual component's weight is multiplied by kernel(r) at the
 * linearisation point (terms.py:73-75).  The reference's fused BA rejects
 * robust kernels (buffer.py:540-545); this entry point has no counterpart
 * there and is what lets the fused path serve those configurations.
 * threshold: Huber k, Tukey c, GNC-TLS c_bar; gnc_mu / gnc_mu_max: the GNC
 * continuation parameter of this call (mu >= mu_max: the hard TLS indicator). */
#define VH_ROBUST_NONE 0
#define VH_ROBUST_HUBER 1
#define VH_ROBUST_TUKEY 2
#define VH_ROBUST_GNC_TLS 3
int vh_ba_run_robust(const vh_ba_plan* plan, float* poses, float* disps, float* intrinsics, const float* disps_sens,
                     const float* targets, const float* weights, const float* eta, const float* depth_active,
                     int iterations, float lm, float ep, float alpha, float intrinsics_lm, float intrinsics_ep,
                     float intrinsics_scale, int flags, int robust_type, float threshold, float gnc_mu,
                     float gnc_mu_max, float* dx, float* dz, float* energy, void* workspace, void* stream);
/* Fused BA on multi-view rig buffers and on the Mei camera: the layouts the
 * reference's fused path rejects and sends to its generic solver
 * (buffer.py:384-410; the generic model is DenseDepthFlowTerm, terms.py:98-262,
 * with geom.iproj_i_proj_j_disp, geom.py:187-298, and the pinhole / Mei
 * models of vipe/utils/cameras.py:123-354).  No counterpart in the reference's
 * fused kernel; it keeps that kernel's conventions (MIN_DEPTH 0.25, weight
 * scale 0.001, the back-substitution skip of local pose index <= 0, damping,
 * zeros on a failed factorisation).
 *
 * The plan takes HOST arrays: edge e runs from view qi[e] of pose pi[e] to
 * view qj[e] of pose pj[e] (pi == pj with qi != qj: an edge within one frame,
 * which constrains the depth only).  Depth frames are pi V + qi over the
 * flattened disparities [N V, H, W]; kx = every view of the window's frames
 * and the edges' source depth frames, sorted.
 *
 * The run: poses [N,7] SE3 (in place), disps [N V, H, W] (in place),
 * intrinsics [V][4] (pinhole fx fy cx cy) or [V][5] (Mei: fx fy cx cy k1) at
 * the maps' resolution, fixed; rig [V][7] camera-to-rig SE3 (t, q xyzw) of
 * each view; disps_sens [N V, H, W]; eta [|kx|, H, W]; depth_active [N V];
 * flags VH_BA_MOTION_ONLY / VH_BA_COMPUTE_ENERGY; robust_* as
 * vh_ba_run_robust (VH_ROBUST_NONE for least squares).  Outputs as vh_ba_run. */
#define VH_CAM_PINHOLE 0
#define VH_CAM_MEI 1
vh_ba_plan* vh_ba_plan_create_rig(const int64_t* pi_host, const int64_t* pj_host, const int64_t* qi_host,
                                  const int64_t* qj_host, int E, int N, int V, int H, int W, int t0, int t1,
                                  int pixel_tile);
int vh_ba_run_rig(const vh_ba_plan* plan, float* poses, float* disps, const float* intrinsics, int camera,
                  const float* rig, const float* disps_sens, const float* targets, const float* weights,
                  const float* eta, const float* depth_active, int iterations, float lm, float ep, float alpha,
                  int flags, int robust_type, float threshold, float gnc_mu, float gnc_mu_max, float* dx, float* dz,
                  float* energy, void* workspace, void* stream);
/* RobustKernel.apply (kernel.py:41-131): w[i] = kernel(x[i]) for n FP32 values. */
int vh_robust_weights(int robust_type, float threshold, float gnc_mu, float gnc_mu_max, const float* x, float* w,
                      int64_t n, void* stream);

/* Implicit adjoint of one Gauss-Newton iteration of vh_ba_run (no counterpart
 * in the reference, whose BA layer has no backward).  Call it after a
 * vh_ba_run with iterations = 1 on the same plan and workspace, with nothing
 * else run on that workspace in between: it reuses that iteration's Cholesky
 * factor, Q, w and dx.  poses, disps, intrinsics are the values BEFORE that
 * iteration (vh_ba_run updates them in place); targets, weights, depth_active,
 * lm and flags are the ones the iteration used.  grad_dx [P,6] and grad_dz
 * [|kx|,H*W] (unused when motion only) are dL/d(dx), dL/d(dz); grad_targets and
 * grad_weights [E,2,H,W] receive dL/d(targets), dL/d(weights) (fully written).
 * Intrinsics optimisation is not supported.  With lm != 0 every non-stereo
 * edge of a depth frame must reach a different pose. */
int vh_ba_adjoint(const vh_ba_plan* plan, const float* poses, const float* disps, const float* intrinsics,
                  const float* targets, const float* weights, const float* depth_active, float lm, int flags,
                  const float* grad_dx, const float* grad_dz, float* grad_targets, float* grad_weights,
                  void* workspace, void* stream);

/* ---- slam_ext geometry siblings (csrc/slam_ext/slam.cpp:52-55,
 *      geom_kernels.cu:526-953, 1849-1950) ---------------------------------
 * All pointers device; poses [N,7] (t, q xyzw), disps [N,H,W] f32, index
 * tensors int64.  An out-of-range frame index gives NaN (iproj, projmap,
 * frame_distance) or a zero count (depth_filter) for that item. */

/* iproj_cuda (geom_kernels.cu:1931-1950): points [N,H,W,3] = T_n applied to
 * the back-projected pixel, divided by its disparity. */
int vh_geom_iproj(const float* poses, const float* disps, const float* intrinsics, int N, int H, int W, float* points,
                  void* stream);
/* projmap_cuda (geom_kernels.cu:1879-1903): coords [E,H,W,3] (third channel
 * zero), valid [E,H,W] for edges ii -> jj; intrinsics [4]. */
int vh_geom_projmap(const float* poses, const float* disps, const float* intrinsics, const int64_t* ii,
                    const int64_t* jj, int E, int N, int H, int W, float* coords, float* valid, void* stream);
/* frame_distance_cuda (geom_kernels.cu:1849-1877): dist [num] for frame pairs
 * (pi, pj), disparity maps di and intrinsics rows (qi, qj) of intrinsics
 * [n_intrinsics,4]; beta weighs full motion against translation only. */
int vh_geom_frame_distance(const float* poses, const float* disps, const float* intrinsics, int n_intrinsics,
                           const int64_t* pi, const int64_t* pj, const int64_t* qi, const int64_t* qj,
                           const int64_t* di, int num, int N, int H, int W, float beta, float* dist, void* stream);
/* depth_filter_cuda (geom_kernels.cu:1905-1929): counter [num,H,W] of the six
 * neighbour frames of inds[b] that agree with each pixel's depth within
 * thresh[b]; intrinsics [4]. */
int vh_geom_depth_filter(const float* poses, const float* disps, const float* intrinsics, const int64_t* inds,
                         const float* thresh, int num, int N, int H, int W, float* counter, void* stream);

/* ---- corr_ext: spatial correlation sampler ---------------------------------
 * Replaces correlation_sample_forward / correlation_sample_backward
 * (csrc/corr_ext/correlation_sampler.cpp:41-85, kernels in
 * correlation_cuda_kernel.cu:27-286).  in1, in2: [N,C,H,W] contiguous, same
 * shape, dtype VH_F32 / VH_F64 / VH_F16 / VH_BF16 (the sum accumulates in FP32,
 * FP64 for VH_F64).  out: [N,patchH,patchW,oH,oW] with
 * oH = (H + 2 padH - ((kH-1) dilH + 1)) / dH + 1 (vh_spatial_corr_out_size).
 * Every output / gradient element is written (no zero-fill needed). */
int vh_spatial_corr_out_size(int H, int W, int kH, int kW, int padH, int padW, int dilH, int dilW, int dH, int dW,
                             int* oH, int* oW);
/* Bytes of device workspace that let vh_spatial_corr_forward use the matrix
 * cores (kernel 1, stride 1, F32 / F16 / BF16, (patchW-1)*dpW <= 48): two
 * channels-last copies of the inputs.  0 when that path does not apply. */
size_t vh_spatial_corr_workspace_size(int dtype, int N, int C, int H, int W, int kH, int kW, int patchH, int patchW,
                                      int padH, int padW, int dilH, int dilW, int dpH, int dpW, int dH, int dW);
int vh_spatial_corr_forward(int dtype, const void* in1, const void* in2, int N, int C, int H, int W, int kH, int kW,
                            int patchH, int patchW, int padH, int padW, int dilH, int dilW, int dpH, int dpW, int dH,
                            int dW, void* out, void* workspace, size_t workspace_bytes, void* stream);
/* Bytes of device workspace that let vh_spatial_corr_backward use the matrix
 * cores (kernel 1, stride 1, F32 / F16 / BF16, (patchW-1)*dpW <= 48): two
 * channels-last FP32 copies of the inputs.  0 when that path does not apply. */
size_t vh_spatial_corr_backward_workspace_size(int dtype, int N, int C, int H, int W, int kH, int kW, int patchH,
                                               int patchW, int padH, int padW, int dilH, int dilW, int dpH, int dpW,
                                               int dH, int dW);
/* grad_out: [N,patchH,patchW,oH,oW] contiguous; grad1, grad2: [N,C,H,W]. */
int vh_spatial_corr_backward(int dtype, const void* in1, const void* in2, const void* grad_out, int N, int C, int H,
                             int W, int kH, int kW, int patchH, int patchW, int padH, int padW, int dilH, int dilW,
                             int dpH, int dpW, int dH, int dW, void* grad1, void* grad2, void* workspace,
                             size_t workspace_bytes, void* stream);

/* ---- scatter_ext -----------------------------------------------------------
 * Replaces scatter_cuda (csrc/scatter_ext/cuda/scatter_cuda.cu:56-129), the
 * kernel behind scatter_sum / scatter_mul / scatter_mean / scatter_min /
 * scatter_max (scatter.cpp:54-237).  reduce: 0 sum, 1 mul, 2 min, 3 max (mean
 * is a sum divided by a count, composed by the caller as the reference does).
 * dtype: VH_F32 / VH_F64 / VH_F16 / VH_BF16 / VH_U8 / VH_I8 / VH_I16 / VH_I32 /
 * VH_I64.  src: contiguous [B,E,K] (the axes before, at and after dim);
 * index: int64, element (b,e,k) at b*sB + e*sE + k*sK; out: contiguous
 * [B,N,K].  out_given = 0: out is initialised here (identity; min / max
 * entries left at the identity become 0); 1: out's values take part.
 * arg_out [B,N,K] int64 (min / max only, NULL otherwise): the smallest e whose
 * src equals the result, E if none.  Indices outside [0,N) are skipped.
 * workspace: device memory of vh_scatter_workspace_size(...) bytes; when that
 * size is non-zero (index independent of b and k, K >= 32, N <= 12000,
 * E <= 65536) the reduction runs atomic-free over a stable sort of the index
 * and gives the same bits on every run; with no workspace it uses atomics. */
size_t vh_scatter_workspace_size(int64_t B, int64_t E, int64_t K, int64_t sB, int64_t sE, int64_t sK, int64_t N);
int vh_scatter(int reduce, int dtype, const void* src, const int64_t* index, int64_t B, int64_t E, int64_t K,
               int64_t sB, int64_t sE, int64_t sK, int64_t N, void* out, int64_t* arg_out, int out_given,
               void* workspace, size_t workspace_bytes, void* stream);

/* ---- factor-graph update glue (vipe/slam/components/factor_graph.py) -----
 * Motion features of FactorGraph.update / update_batch (:247-260, :339-345),
 * with the reprojection of GraphBuffer.reproject_dense_disp (buffer.py:705-726)
 * for a single-view pinhole graph: poses [N,7], disps [N,H,W], intrinsics [4]
 * at the feature resolution, ii / jj [E] int64, target [E,H,W,2] (the graph's
 * current targets).  Writes coords1 [E,H,W,2] and motn [E,4,H,W] =
 * clamp([coords1 - grid, target - coords1], -64, 64).  Points with z < 0.1
 * project with z = 1 (cameras.py:175-176).  An out-of-range frame index gives
 * NaN coordinates for that edge and reads nothing. */
int vh_graph_motion(const float* poses, const float* disps, const float* intrinsics, const int64_t* ii,
                    const int64_t* jj, const float* target, int E, int N, int H, int W, float* coords1, float* motn,
                    void* stream);
/* The update's epilogue (:268-277, :359-362): target = coords1 + delta,
 * weight = weight_net with the pixels of mask[frame[e]] (uint8 [NM,H,W], may be
 * NULL) zeroed; delta / weight_net [E,H,W,2] in dtype VH_F32 or VH_F16 (the
 * network's autocast output).  target / weight [E,H,W,2] f32 (graph state);
 * target_ba / weight_ba [E,2,H,W] f32 (the layout vh_ba_run reads) when not NULL. */
int vh_graph_targets(int dtype, const float* coords1, const void* delta, const void* weight_net, const uint8_t* mask,
                     const int64_t* frame, int NM, int E, int H, int W, float* target, float* weight,
                     float* target_ba, float* weight_ba, void* stream);

/* ---- DroidNet update operator (vipe/slam/networks/droid_net.py:393-519) ---
 * The layers of UpdateModule.forward (:487-519), ConvGRU.forward (:406-420)
 * and GraphAgg.forward (:434-449), FP16 activations and weights with FP32
 * accumulation (the reference runs them under autocast, factor_graph.py:228).
 * Activations are channels-last f16: [image][pixel][channel], pixel = y*W + x.
 *
 * vh_uo_conv: a 'same'-padded KHxKW convolution (odd sizes, stride 1) of the
 * channel concatenation of up to 4 input segments (torch.cat(dim=1), :407-408,
 * :417), each seg_c channels (a multiple of 32) at channel seg_off of a buffer
 * with seg_stride channels per pixel (multiples of 8; 16-byte aligned
 * pointers).  weight: f16 [KH*KW][Cout_pad][sum seg_c] (Cout_pad a multiple of
 * 128, rows past Cout zero); bias f32 [Cout].  glo (may be NULL): f32
 * [nimg][glo_stride], added at channel n before the activation (the
 * convz_glo / convr_glo / convq_glo terms, :415-417).  Output element
 * (pixel p, channel n) goes to out[p*out_stride + out_off + n] (f32 when
 * out_f32, else f16), or with VH_UO_NCHW to
 * out[((img*out_stride) + out_off + n)*H*W + pixel].  Epilogues:
 *   VH_UO_NONE, VH_UO_RELU, VH_UO_SIGMOID   the activation
 *   VH_UO_SIGMUL  sigmoid(x) * aux0[p][n]                  (:413)
 *   VH_UO_GATES   Cout 256 = [convz; convr]: out[p][n] = z = sigmoid(x) for
 *                 n < 128, out2[p*128 + n-128] = sigmoid(x) * aux0[p][n-128]
 *                 (r * net) for the rest                   (:415-417)
 *   VH_UO_GRU     Cout 128: (1 - aux0) * aux1 + aux0 * tanh(x), with aux0 = z
 *                 and aux1 = net                           (:417-419)
 * aux buffers are f16 with aux_stride channels per pixel. */
#define VH_UO_NONE 0
#define VH_UO_RELU 1
#define VH_UO_SIGMOID 2
#define VH_UO_ETA 3 /* 0.01 * softplus(x): GraphAgg's damping head, :430, :449 */
#define VH_UO_SIGMUL 4
#define VH_UO_GATES 5
#define VH_UO_GRU 6
#define VH_UO_NCHW 16 /* flag or-ed into the epilogue */
typedef struct {
    const void* seg[4];
    int seg_c[4];
    int seg_stride[4];
    int seg_off[4];
    int nseg;
    const void* weight;
    const float* bias;
    int nimg, H, W, Cout, Cout_pad, KH, KW;
    int epilogue;
    void* out;
    int out_f32, out_stride, out_off;
    const float* glo;
    int glo_stride;
    const void* aux0;
    const void* aux1;
    int aux_stride;
    void* out2;
} vh_uo_conv_args;
/* sizeof(vh_uo_conv_args), for bindings that lay the struct out themselves. */
size_t vh_uo_conv_args_size(void);
int vh_uo_conv(const vh_uo_conv_args* args, void* stream);
/* The 1- and 2-channel heads (delta.2, weight.2 :475-485, eta :430): in f16
 * channels [in_off, in_off + cin) of [P][in_stride]; weight f32
 * [KH*KW][cin][Cout] (KH*KW <= 9, cin <= 256, Cout 1 or 2), bias f32 [Cout];
 * out f32 [P][out_stride]; epilogue VH_UO_NONE / RELU / SIGMOID / ETA. */
int vh_uo_conv_small(const void* in, int in_stride, int in_off, int cin, const float* weight, const float* bias,
                     int nimg, int H, int W, int KH, int KW, int Cout, int epilogue, float* out, int out_stride,
                     void* stream);
/* Layout changes at the module boundary (the reference's [B*N, C, H, W]
 * views, :499-502): in [nimg][C][HW] (dtype VH_F32 / VH_F16) -> out f16
 * [nimg][HW][Cpad], channels past C zero; and f16 [nimg][HW][stride]
 * channels [off, off + C) -> out [nimg][C][HW] in dtype. */
int vh_uo_to_nhwc(int dtype, const void* in, int nimg, int C, int HW, int Cpad, void* out, void* stream);
int vh_uo_to_nchw(int dtype, const void* in, int nimg, int C, int HW, int stride, int off, void* out, void* stream);
/* ConvGRU's global context (:413-414): out[img][c] = mean over the HW pixels
 * of in f32 [nimg][HW][C], summed in a fixed order; and the 1x1 convolutions
 * of that 1x1 map (:415-417): out[img][k] = sum_c weight[k][c] in[img][c] +
 * bias[k], f32. */
int vh_uo_mean(const float* in, int nimg, int HW, int C, float* out, void* stream);
int vh_uo_fc(const float* in, int nimg, int Cin, const float* weight, const float* bias, int Cout, float* out,
             void* stream);

#ifdef __cplusplus
}
#endif

#endif /* VIPE_HIP_H */
