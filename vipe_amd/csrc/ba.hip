// Dense DROID bundle adjustment on gfx950 (slam_ext.ba / ba_extended).
//
// Semantics: csrc/slam_ext/geom_kernels.cu:178-524 (per-edge terms),
// :1197-1424 (block system, damping, zeros on failed factorisation),
// :1426-1624 (Schur over rows that share a depth frame), :1179-1195
// (back-substitution that skips rows whose local pose index is <= 0),
// :974-1041 (retractions), :1626-1820 (iteration order).
//
// Per Gauss-Newton iteration, all on the device, no host round trip:
//   1. ba_frame_kernel   one workgroup per (depth frame k, pixel tile).  For
//      every edge leaving k it evaluates residuals and Jacobians per pixel and
//      accumulates (a) the edge's j-side pose terms Hjj, vj (+ intrinsics
//      terms, energy) - the i-side follows from Ji = -Adj(Tij)^T Jj; (b) the
//      depth terms C, w; (c) the E rows (self row + one per edge) in LDS, and
//      from them the frame's Schur Gram sum_p Q E_a E_b^T on the f32 matrix
//      cores (16 x 16 tiles of E rows), plus [w, Ef] as a pseudo row so the Schur right-hand side and the
//      intrinsics coupling come out of the same Gram.  Writes per-tile
//      partials and Q, w per pixel.  E is never written to memory.
//   2. ba_finalize_kernel  sums partials over tiles in FP64, forms
//      Hii = A^T Hjj A, Hij = -A^T Hjj, vi = -A^T vj (A = Adj(Tij)).
//   3. ba_assemble_kernel  dense lower-triangular reduced system H - S in
//      FP64 from host-built contribution lists (fixed order: deterministic),
//      damping diag += ep + lm * diag.
//   4. chol_*               blocked right-looking FP64 Cholesky (64 x 64 blocks);
//      a non-positive pivot raises a flag and the solve returns zeros.
//   5. ba_backsub_kernel   recomputes E rows from weights and geometry,
//      dz = Q (w - E^T dx) - Q Ef df, disps += dz.
//   6. ba_retract_kernel   T <- Exp(dx) T in FP64, intrinsics += df * scale.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "lie_math.hpp"
#include "vh_common.hpp"

#define MIN_DEPTH 0.25f

namespace vh {
namespace ba {

constexpr int kThreads = 256;
// A depth frame with at most kMaxDeg outgoing edges keeps its E rows in LDS
// (the frame kernel); one with more takes the wide-frame kernel, which stages
// them in the workspace, so the edge count of a frame is not limited.
constexpr int kMaxDeg = 18;   // edges leaving one depth frame on the LDS path
constexpr int kMaxRext = 20;  // self row + edges + pseudo row
constexpr int kWideGroup = 16;  // edges of a wide frame set up and reduced together
constexpr int kNV = 40;       // per-edge partial stride (floats)
constexpr int kCB = 64;       // Cholesky block size
constexpr int kMaxPoses = 1023;  // poses in the window: a whole 1024-keyframe buffer optimised over [1, 1024)
constexpr int kMaxLd = (7 * kMaxPoses + 1 + kCB - 1) / kCB * kCB;  // leading dimension of the reduced system
constexpr int kGramSplit = 8; // finalize workgroups per depth frame
constexpr int kLdsBudget = 64 * 1024;  // dynamic E-row tile; with ~15 KB static, two frame workgroups fit a CU
constexpr int kLdsMax = 128 * 1024;    // largest dynamic E-row tile a caller may ask for (one workgroup per CU)

// Poses: SE3 [t, q] with 6-dof blocks (PD = 6), or Sim3 [t, q, s] with 7-dof
// blocks (PD = 7, tangent [tau, phi, sigma]).  The projection of a point does
// not change when the target frame's scale does (a Sim3 scale moves X, Y, Z
// together), so every j-side Jacobian has a zero sigma column: the per-pixel
// partials below are the same 6-dof sums for both groups, and sigma enters
// through the i side, Ji = -Adj(Tij)^T Jj.
//
// per-edge partial layout (floats)
//   [0,21)  Hjj upper triangle, index(n,m) n>=m: n(n+1)/2 + m
//   [21,27) vj   27 energy  28 Hff  29 vf  [32,38) Fj
// finalised edge record (doubles), PD x PD blocks:
//   Hii, Hij, Hjj, vi, vj, Fi, Fj, Hff, vf, energy
template <int PD>
struct Rec {
    static constexpr int HII = 0, HIJ = PD * PD, HJJ = 2 * PD * PD, VI = 3 * PD * PD, VJ = VI + PD, FI = VJ + PD,
                         FJ = FI + PD, HFF = FJ + PD, VF = HFF + 1, EN = HFF + 2;
    static constexpr int SIZE = (EN + 1 + 7) / 8 * 8;  // 136 (SE3), 184 (Sim3)
};
constexpr int kEDSMax = Rec<7>::SIZE;

// contribution item kinds for the assembly lists
enum { K_HII = 0, K_HJJ = 1, K_HIJ = 2, K_HIJT = 3, K_G = 4, K_GT = 5, K_VI = 6, K_VJ = 7, K_ROW = 8 };

struct DevPlan {
    int E, N, H, W, HW, t0, t1, P, K;
    int TP, SUB, T, Rext_max, GT;  // GT: Gram floats per tile (blocks x PD^2, set per run)
    int n, ld, nb, nlb;  // system size, leading dim, blocks, lower pose blocks (set per run)
    int intr;            // system includes the intrinsics scalar (set per run)
    int pd;              // pose dof: 6 (SE3) or 7 (Sim3), set per run
    const int* kx;       // [K] frame id of slot k
    const int* fr_ptr;   // [K+1] CSR into fr_edges
    const int* fr_edges; // [E]
    const int* fr_self;  // [K]
    const int* gram_off; // [K+1] first Gram block of each frame (blocks of PD x PD)
    const int* ii;       // [E]
    const int* jj;       // [E]
    const int* blk_ab;   // [nlb*2]
    const int* blk_ptr;  // [nlb+1]
    const int* blk_items;// packed (kind << 27 | index): edge index, or Gram block index
    const int* rhs_ptr;  // [P+1]
    const int* rhs_items;
    const int* pseudo_off; // [K] Gram block index of the (pseudo, pseudo) block
    const int* wide;     // [n_wide] slots of the depth frames with more than kMaxDeg outgoing edges
    int n_wide;          // number of such frames
    int wide_rows;       // stored E rows per pixel of the largest of them (zero row included)
    // rig plans (V > 0): depth frame ids are p V + q over the flattened
    // [N V, H, W] disparities; ii / jj hold the edges' poses
    int V;               // views (0: a single-view plan, depth frame id = pose id)
    const int* qi;       // [E] source view of each edge
    const int* qj;       // [E] target view
    const int* kpose;    // [K] pose of depth frame k
};

struct Work {
    float* edge_part;  // [T][E][kNV]
    float* gram_part;  // [T][GT]
    float* fen_part;   // [T][K]
    float* Qb;         // [K][HW]
    float* Wb;         // [K][HW]
    double* ed;        // [E][Rec<PD>::SIZE]
    double* gd;        // [GT]
    double* fen;       // [K]
    double* A;         // [ld][ld]
    double* b;         // [ld]
    double* x;         // [ld]
    double* Li;        // [nb][64][64] L_kk^-T of each diagonal block
    double* y;         // [ld] forward-substitution result
    int* flag;         // [4]
    float* wide_e;     // [n_wide][T][wide_rows][256] E rows of the wide frames
};

// Robust kernels of the generic solver (vipe/slam/ba/kernel.py:21-162) as an
// iteratively reweighted factor on each flow-residual component's weight
// (terms.py:73-75: w <- w * kernel(r)), evaluated at the linearisation point:
//   1 Huber   w = 1 if |r| <= c else c / |r|
//   2 Tukey   w = (max(1 - (r/c)^2, 0))^2
//   3 GNC-TLS at mu, c2 = c^2: 1 if r^2 < mu/(mu+1) c2, 0 if r^2 > (mu+1)/mu c2,
//             else sqrt(c2 mu (mu+1) / r^2) - mu (clamped at 0); mu >= mu_max
//             gives the indicator r^2 < c2.
struct RobustK {
    int type;
    float c, mu, mu_max;
};

__host__ __device__ inline float robust_weight(float r, const RobustK& k) {
    if (k.type == 1) {
        const float ar = fabsf(r);
        return ar <= k.c ? 1.f : k.c / fmaxf(ar, 1e-8f);
    }
    if (k.type == 2) {
        const float u = (r / k.c) * (r / k.c);
        const float v = fmaxf(1.f - u, 0.f);
        return v * v;
    }
    if (k.type == 3) {
        const float r2 = r * r, c2 = k.c * k.c;
        if (k.mu >= k.mu_max) return r2 < c2 ? 1.f : 0.f;
        const float lo = (k.mu / (k.mu + 1.f)) * c2, hi = ((k.mu + 1.f) / k.mu) * c2;
        if (r2 < lo) return 1.f;
        if (r2 > hi) return 0.f;
        return fmaxf(sqrtf(c2 * k.mu * (k.mu + 1.f) / fmaxf(r2, 1e-12f)) - k.mu, 0.f);
    }
    return 1.f;
}

struct RunArgs {
    float* poses;
    float* disps;
    float* intrinsics;
    const float* disps_sens;
    const float* targets;
    const float* weights;
    const float* eta;
    const float* depth_active;
    float lm, ep, alpha, ilm, iep, iscale;
    float* dx;
    float* dz;
    float* energy;
    int itr;
    int compute_energy;
    RobustK rk;   // robust kernel on the flow residuals (type 0: plain least squares)
    const float* rig;  // rig runs: [V][7] camera-to-rig SE3 (t, q xyzw) of each view
    int kstride;       // rig runs: intrinsics per view, 4 (pinhole) or 5 (Mei: fx fy cx cy k1)
};

__host__ __device__ inline int pk(int kind, int idx) { return (kind << 27) | idx; }

// Relative pose Tij = Tj * Ti^-1 as rotation matrix, translation and scale
// (relSE3, geom_kernels.cu:101-114; for Sim3 the group product of sim3.h:
// s = s_j / s_i, t = t_j - s R t_i); stereo edges use (t = (-0.1,0,0), R = I,
// s = 1) (:231-241).
template <int PD>
__device__ __forceinline__ void edge_transform(const float* __restrict__ poses, int i, int j, float* R, float* t,
                                               float& sc) {
    sc = 1.f;
    if (i == j) {
        R[0] = 1; R[1] = 0; R[2] = 0; R[3] = 0; R[4] = 1; R[5] = 0; R[6] = 0; R[7] = 0; R[8] = 1;
        t[0] = -0.1f; t[1] = 0.f; t[2] = 0.f;
        return;
    }
    const float* Pi = poses + (PD + 1) * i;
    const float* Pj = poses + (PD + 1) * j;
    if (PD == 7) sc = Pj[7] / Pi[7];
    const float qi[4] = {Pi[3], Pi[4], Pi[5], Pi[6]};
    const float qj[4] = {Pj[3], Pj[4], Pj[5], Pj[6]};
    float q[4];
    q[0] = -qj[3] * qi[0] + qj[0] * qi[3] - qj[1] * qi[2] + qj[2] * qi[1];
    q[1] = -qj[3] * qi[1] + qj[1] * qi[3] - qj[2] * qi[0] + qj[0] * qi[2];
    q[2] = -qj[3] * qi[2] + qj[2] * qi[3] - qj[0] * qi[1] + qj[1] * qi[0];
    q[3] = qj[3] * qi[3] + qj[0] * qi[0] + qj[1] * qi[1] + qj[2] * qi[2];
    const float x = q[0], y = q[1], z = q[2], w = q[3];
    R[0] = 1 - 2 * (y * y + z * z); R[1] = 2 * (x * y - w * z);     R[2] = 2 * (x * z + w * y);
    R[3] = 2 * (x * y + w * z);     R[4] = 1 - 2 * (x * x + z * z); R[5] = 2 * (y * z - w * x);
    R[6] = 2 * (x * z - w * y);     R[7] = 2 * (y * z + w * x);     R[8] = 1 - 2 * (x * x + y * y);
    const float ti[3] = {Pi[0], Pi[1], Pi[2]};
    t[0] = Pj[0] - sc * (R[0] * ti[0] + R[1] * ti[1] + R[2] * ti[2]);
    t[1] = Pj[1] - sc * (R[3] * ti[0] + R[4] * ti[1] + R[5] * ti[2]);
    t[2] = Pj[2] - sc * (R[6] * ti[0] + R[7] * ti[1] + R[8] * ti[2]);
}

// Per-pixel terms of one edge (projective_transform_kernel :298-435).
struct PixTerms {
    float ru, rv, wu, wv, Jzu, Jzv, Jfu, Jfv;
    float Ju[6], Jv[6];
    bool ok;  // depth in front of MIN_DEPTH: the residual carries weight
};

// The homogeneous point [X0, Y0, 1, d] maps to [s R X + t d, d]; sc = s is 1
// for SE3 (PD = 6), where the products by it are left out.
template <bool INTR, bool RES, int PD = 6>
__device__ __forceinline__ void pixel_terms(const float* R, const float* t, float sc, float X0, float Y0, float d,
                                            float fx, float fy, float cx, float cy, float wgt_u, float wgt_v,
                                            float tgt_u, float tgt_v, float iscale, PixTerms& o) {
    float Xj0 = R[0] * X0 + R[1] * Y0 + R[2];
    float Xj1 = R[3] * X0 + R[4] * Y0 + R[5];
    float Xj2 = R[6] * X0 + R[7] * Y0 + R[8];
    if (PD == 7) {
        Xj0 *= sc;
        Xj1 *= sc;
        Xj2 *= sc;
    }
    Xj0 += t[0] * d;
    Xj1 += t[1] * d;
    Xj2 += t[2] * d;
    const bool ok = !(Xj2 < MIN_DEPTH);
    o.ok = ok;
    const float dinv = ok ? 1.0f / Xj2 : 0.0f;
    const float d2 = dinv * dinv;
    const float x = Xj0, y = Xj1, h = d;
    o.wu = ok ? 0.001f * wgt_u : 0.0f;
    o.wv = ok ? 0.001f * wgt_v : 0.0f;
    if (RES) {
        o.ru = tgt_u - (fx * dinv * x + cx);
        o.rv = tgt_v - (fy * dinv * y + cy);
    }
    o.Ju[0] = fx * (h * dinv);
    o.Ju[1] = 0.f;
    o.Ju[2] = fx * (-x * h * d2);
    o.Ju[3] = fx * (-x * y * d2);
    o.Ju[4] = fx * (1 + x * x * d2);
    o.Ju[5] = fx * (-y * dinv);
    o.Jv[0] = 0.f;
    o.Jv[1] = fy * (h * dinv);
    o.Jv[2] = fy * (-y * h * d2);
    o.Jv[3] = fy * (-1 - y * y * d2);
    o.Jv[4] = fy * (x * y * d2);
    o.Jv[5] = fy * (x * dinv);
    o.Jzu = fx * (t[0] * dinv - t[2] * (x * d2));
    o.Jzv = fy * (t[1] * dinv - t[2] * (y * d2));
    if (INTR) {
        // dXi/df = (-X0/fx, -Y0/fy, 0, 0) mapped by s R only (:325-331)
        const float a0 = -X0 / fx, a1 = -Y0 / fy;
        float e0 = R[0] * a0 + R[1] * a1;
        float e1 = R[3] * a0 + R[4] * a1;
        float e2 = R[6] * a0 + R[7] * a1;
        if (PD == 7) {
            e0 *= sc;
            e1 *= sc;
            e2 *= sc;
        }
        o.Jfu = iscale * (fx * dinv * e0 - fx * x * d2 * e2 + x * dinv);
        o.Jfv = iscale * (fy * dinv * e1 - fy * y * d2 * e2 + y * dinv);
    } else {
        o.Jfu = o.Jfv = 0.f;
    }
}

// -Adj(T)^T e.  SE3: Adj = [[R, [t]x R], [0, R]], so top = -R^T e_top and
// bottom = -(R^T (e_top x t) + R^T e_bot) (adjSE3, geom_kernels.cu:81-99).
// Sim3: Adj = [[sR, [t]x R, -t], [0, R, 0], [0, 0, 1]] (sim3.h Adj), so the top
// is scaled by s and o[6] = t . e_top - e[6].  e has PD entries (e[6] read only
// for Sim3).
template <int PD = 6>
__device__ __forceinline__ void neg_adjT(const float* R, const float* t, float sc, const float* e, float* o) {
    const float c0 = e[1] * t[2] - e[2] * t[1];
    const float c1 = e[2] * t[0] - e[0] * t[2];
    const float c2 = e[0] * t[1] - e[1] * t[0];
    const float b0 = e[3] + c0, b1 = e[4] + c1, b2 = e[5] + c2;
    o[0] = -(R[0] * e[0] + R[3] * e[1] + R[6] * e[2]);
    o[1] = -(R[1] * e[0] + R[4] * e[1] + R[7] * e[2]);
    o[2] = -(R[2] * e[0] + R[5] * e[1] + R[8] * e[2]);
    o[3] = -(R[0] * b0 + R[3] * b1 + R[6] * b2);
    o[4] = -(R[1] * b0 + R[4] * b1 + R[7] * b2);
    o[5] = -(R[2] * b0 + R[5] * b1 + R[8] * b2);
    if (PD == 7) {
        o[0] *= sc;
        o[1] *= sc;
        o[2] *= sc;
        o[PD - 1] = (t[0] * e[0] + t[1] * e[1] + t[2] * e[2]) - e[PD - 1];
    }
}

// ---------------------------------------------------------------------------
// Multi-view rigs and the Mei camera (rig plans; CAM 1 pinhole, 2 Mei)
// ---------------------------------------------------------------------------
// Edge e of a rig plan runs from view qi of pose pi to view qj of pose pj.
// With G = Pj Pi^-1 (the identity when pi == pj) and rig[q] the camera-to-rig
// SE3 of view q (the generic path's model, vipe/slam/maths/geom.py:187-298):
//   camera transform    T = rig[qj]^-1 G rig[qi]   (maps the point)
//   self-row transform  M = rig[qj]^-1 G           (Ji = -Jc Adj(M))
//   j-side transform    A = Adj(rig[qj]^-1)        (Jj = Jc A)
// with Jc the camera-frame Jacobian of a left perturbation of T.  The frame
// kernels sum camera-frame partials and E rows; finalize applies A to the
// edge blocks and ba_gram_rig_kernel to the Gram blocks of the edge rows, in
// FP64; the back-substitution dots the camera-frame E rows with
// A (dx_j - Adj(G) dx_i).  An edge within one frame (pi == pj) has Ji = -Jj,
// so its pose terms cancel: it carries none and no Schur rows, and constrains
// the depth only.  Intrinsics are per view and fixed.
constexpr int kRigEP = 36;  // per-edge LDS record: T (R 0-8, t 9-11), flags 12-14, M (R 16-24, t 25-27), Kj 28-32

__device__ __forceinline__ void quat_rot(const float* q, float* R) {
    const float x = q[0], y = q[1], z = q[2], w = q[3];
    R[0] = 1 - 2 * (y * y + z * z); R[1] = 2 * (x * y - w * z);     R[2] = 2 * (x * z + w * y);
    R[3] = 2 * (x * y + w * z);     R[4] = 1 - 2 * (x * x + z * z); R[5] = 2 * (y * z - w * x);
    R[6] = 2 * (x * z - w * y);     R[7] = 2 * (y * z + w * x);     R[8] = 1 - 2 * (x * x + y * y);
}

// G of rig-plan edge e (R, t); the identity for an edge within one frame
__device__ __forceinline__ void rig_pose_transform(const DevPlan& p, const RunArgs& a, int e, float* RG, float* tG) {
    float sc;
    if (p.ii[e] == p.jj[e]) {
        for (int i = 0; i < 9; i++) RG[i] = (i % 4 == 0) ? 1.f : 0.f;
        tG[0] = tG[1] = tG[2] = 0.f;
    } else {
        edge_transform<6>(a.poses, p.ii[e], p.jj[e], RG, tG, sc);
    }
}

// M (Rm, tm) and T (Rc, tc) of rig-plan edge e
__device__ __forceinline__ void rig_edge_transforms(const DevPlan& p, const RunArgs& a, int e, float* Rc, float* tc,
                                                    float* Rm, float* tm) {
    float RG[9], tG[3], Rj[9], Ri[9];
    rig_pose_transform(p, a, e, RG, tG);
    const float* rj = a.rig + 7 * p.qj[e];
    const float* ri = a.rig + 7 * p.qi[e];
    quat_rot(rj + 3, Rj);
    quat_rot(ri + 3, Ri);
    const float dt[3] = {tG[0] - rj[0], tG[1] - rj[1], tG[2] - rj[2]};
    for (int r = 0; r < 3; r++) {
        for (int c = 0; c < 3; c++) Rm[3 * r + c] = Rj[r] * RG[c] + Rj[3 + r] * RG[3 + c] + Rj[6 + r] * RG[6 + c];
        tm[r] = Rj[r] * dt[0] + Rj[3 + r] * dt[1] + Rj[6 + r] * dt[2];
    }
    for (int r = 0; r < 3; r++) {
        for (int c = 0; c < 3; c++)
            Rc[3 * r + c] = Rm[3 * r] * Ri[c] + Rm[3 * r + 1] * Ri[3 + c] + Rm[3 * r + 2] * Ri[6 + c];
        tc[r] = Rm[3 * r] * ri[0] + Rm[3 * r + 1] * ri[1] + Rm[3 * r + 2] * ri[2] + tm[r];
    }
}

// Adj(rig[q]^-1) = [[R^T, hat(-R^T t) R^T], [0, R^T]] in FP64, row-major
__device__ __forceinline__ void rig_adj_inv(const float* rig, int q, double* A) {
    float Rf[9];
    const float* r = rig + 7 * q;
    quat_rot(r + 3, Rf);
    double Rt[9], u[3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) Rt[3 * i + j] = Rf[3 * j + i];
    for (int i = 0; i < 3; i++) u[i] = -(Rt[3 * i] * r[0] + Rt[3 * i + 1] * r[1] + Rt[3 * i + 2] * r[2]);
    const double hu[9] = {0.0, -u[2], u[1], u[2], 0.0, -u[0], -u[1], u[0], 0.0};
    for (int i = 0; i < 36; i++) A[i] = 0.0;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            A[6 * i + j] = Rt[3 * i + j];
            A[6 * (i + 3) + j + 3] = Rt[3 * i + j];
            A[6 * i + j + 3] = hu[3 * i] * Rt[j] + hu[3 * i + 1] * Rt[3 + j] + hu[3 * i + 2] * Rt[6 + j];
        }
}

// The unit-depth ray [X0, Y0, 1] of pixel (u, v) in a view with intrinsics K
// (vipe/utils/cameras.py:131-159 pinhole, :228-251 Mei)
template <int CAM>
__device__ __forceinline__ void rig_iproj(const float* K, float u, float v, float& X0, float& Y0) {
    const float ub = (u - K[2]) / K[0], vb = (v - K[3]) / K[1];
    if (CAM == 2) {
        const float k1 = K[4];
        const float r2 = ub * ub + vb * vb;
        const float q = sqrtf(1.f + (1.f - k1 * k1) * r2);
        const float f = (k1 + q) / (1.f + r2);
        const float s = f / (f - k1);
        X0 = ub * s;
        Y0 = vb * s;
    } else {
        X0 = ub;
        Y0 = vb;
    }
}

// Per-pixel terms of one edge under the Mei model (cameras.py:283-337): the
// point X = R [X0, Y0, 1] + t d projects to fx X / (Z + k1 |X|) + cx; g = d(u,
// v)/dX gives Jc = [d g, X x g] and Jz = g . t.  Below MIN_DEPTH the projection
// and every Jacobian are zero, as in the fused kernel.
template <bool RES>
__device__ __forceinline__ void pixel_terms_mei(const float* R, const float* t, float X0, float Y0, float d,
                                                const float* K, float wgt_u, float wgt_v, float tgt_u, float tgt_v,
                                                PixTerms& o) {
    const float X = R[0] * X0 + R[1] * Y0 + R[2] + t[0] * d;
    const float Y = R[3] * X0 + R[4] * Y0 + R[5] + t[1] * d;
    const float Z = R[6] * X0 + R[7] * Y0 + R[8] + t[2] * d;
    const bool ok = !(Z < MIN_DEPTH);
    o.ok = ok;
    const float fx = K[0], fy = K[1], cx = K[2], cy = K[3], k1 = K[4];
    const float r = sqrtf(X * X + Y * Y + Z * Z);
    const float rb = Z + k1 * r;
    const float inv = ok ? 1.0f / rb : 0.0f;
    const float den = ok ? inv * inv / r : 0.0f;
    o.wu = ok ? 0.001f * wgt_u : 0.0f;
    o.wv = ok ? 0.001f * wgt_v : 0.0f;
    if (RES) {
        o.ru = tgt_u - (fx * X * inv + cx);
        o.rv = tgt_v - (fy * Y * inv + cy);
    }
    const float fu = fx * den, fv = fy * den;
    const float gu[3] = {fu * (rb * r - k1 * X * X), -fu * k1 * X * Y, -fu * X * (k1 * Z + r)};
    const float gv[3] = {-fv * k1 * X * Y, fv * (rb * r - k1 * Y * Y), -fv * Y * (k1 * Z + r)};
    o.Ju[0] = d * gu[0]; o.Ju[1] = d * gu[1]; o.Ju[2] = d * gu[2];
    o.Ju[3] = Y * gu[2] - Z * gu[1]; o.Ju[4] = Z * gu[0] - X * gu[2]; o.Ju[5] = X * gu[1] - Y * gu[0];
    o.Jv[0] = d * gv[0]; o.Jv[1] = d * gv[1]; o.Jv[2] = d * gv[2];
    o.Jv[3] = Y * gv[2] - Z * gv[1]; o.Jv[4] = Z * gv[0] - X * gv[2]; o.Jv[5] = X * gv[1] - Y * gv[0];
    o.Jzu = gu[0] * t[0] + gu[1] * t[1] + gu[2] * t[2];
    o.Jzv = gv[0] * t[0] + gv[1] * t[1] + gv[2] * t[2];
    o.Jfu = o.Jfv = 0.f;
}

// The pixel terms of a rig-plan edge from its LDS record ep (kRigEP layout)
template <int CAM, bool RES>
__device__ __forceinline__ void rig_pixel_terms(const float* ep, float X0, float Y0, float d, float wgt_u, float wgt_v,
                                                float tgt_u, float tgt_v, PixTerms& o) {
    const float* K = ep + 28;
    if (CAM == 2) pixel_terms_mei<RES>(ep, ep + 9, X0, Y0, d, K, wgt_u, wgt_v, tgt_u, tgt_v, o);
    else pixel_terms<false, RES, 6>(ep, ep + 9, 1.f, X0, Y0, d, K[0], K[1], K[2], K[3], wgt_u, wgt_v, tgt_u, tgt_v,
                                    1.f, o);
}

// Cross-lane exchanges without the LDS crossbar.  Every stage pairs each lane
// with one whose stage bit differs; the partner need not be lane ^ S, only a
// lane whose partial so far covers a disjoint set of lanes, which any such
// pairing (xor, or a mirror that also flips the lower bits) gives.
//   S = 1, 2: DPP quad_perm [1,0,3,2], [2,3,0,1]
//   S = 4, 8: DPP row_half_mirror, row_mirror
//   S = 16:   v_permlane16_swap (odd rows of one register <-> even rows of the other)
//   S = 32:   v_permlane32_swap (upper half of one register <-> lower half of the other)
// Every lane of the wave must be active at these calls (an inactive lane's
// register would be read as a partner's value); the callers reach them with
// the whole wave, masking their contributions by value instead.
template <int CTRL>
__device__ __forceinline__ float dpp_partner(float x) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), CTRL, 0xf, 0xf, true));
}
constexpr int kDppXor1 = 0xB1, kDppXor2 = 0x4E, kDppHalfMirror = 0x141, kDppMirror = 0x140;

// x + y after the 16-lane-row swap: rows 0, 2 hold x of the row pair, rows
// 1, 3 hold y of the row pair
__device__ __forceinline__ float swap16_sum(float x, float y) {
    const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(x), __float_as_uint(y), false, false);
    return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}
__device__ __forceinline__ float swap32_sum(float x) {
    const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(x), __float_as_uint(x), false, false);
    return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}

// The same exchanges on doubles, one 32-bit half at a time.
template <int CTRL>
__device__ __forceinline__ double dpp_partner_f64(double x) {
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_mov_dpp((int)(b & 0xffffffffll), CTRL, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xf, 0xf, true);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ double swap16_sum_f64(double x, double y) {
    const long long bx = __double_as_longlong(x), by = __double_as_longlong(y);
    const auto lo = __builtin_amdgcn_permlane16_swap((unsigned)(bx & 0xffffffffll), (unsigned)(by & 0xffffffffll), false, false);
    const auto hi = __builtin_amdgcn_permlane16_swap((unsigned)(bx >> 32), (unsigned)(by >> 32), false, false);
    return __longlong_as_double(((long long)hi[0] << 32) | lo[0]) + __longlong_as_double(((long long)hi[1] << 32) | lo[1]);
}
__device__ __forceinline__ double swap32_sum_f64(double x, double y) {
    const long long bx = __double_as_longlong(x), by = __double_as_longlong(y);
    const auto lo = __builtin_amdgcn_permlane32_swap((unsigned)(bx & 0xffffffffll), (unsigned)(by & 0xffffffffll), false, false);
    const auto hi = __builtin_amdgcn_permlane32_swap((unsigned)(bx >> 32), (unsigned)(by >> 32), false, false);
    return __longlong_as_double(((long long)hi[0] << 32) | lo[0]) + __longlong_as_double(((long long)hi[1] << 32) | lo[1]);
}
// sum over the 16 lanes of a row, in every lane of it
__device__ __forceinline__ double row_sum16_f64(double v) {
    v += dpp_partner_f64<0x140>(v);
    v += dpp_partner_f64<0x141>(v);
    v += dpp_partner_f64<0x4E>(v);
    v += dpp_partner_f64<0xB1>(v);
    return v;
}
__device__ __forceinline__ double wave_sum_f64(double v) {
    v = row_sum16_f64(v);
    v = swap16_sum_f64(v, v);
    return swap32_sum_f64(v, v);
}

// Sum of 32 values over the 64 lanes of a wave: each halving step S keeps
// half of the live values (the half selected by lane bit S) and adds the
// partner lane's copy of the same half, so after S = 16..1 lane l holds the
// partial of value (l & 31); a last step adds the other half-wave.
template <int S, int CTRL>
__device__ __forceinline__ void halve_step_dpp(float (&v)[32], bool hi) {
#pragma unroll
    for (int i = 0; i < S; i++) {
        const float a = v[i] + dpp_partner<CTRL>(v[i]);
        const float b = v[i + S] + dpp_partner<CTRL>(v[i + S]);
        v[i] = hi ? b : a;
    }
}

__device__ __forceinline__ float wave_sum32(float (&v)[32]) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int i = 0; i < 16; i++) v[i] = swap16_sum(v[i], v[i + 16]);
    halve_step_dpp<8, kDppMirror>(v, (lane & 8) != 0);
    halve_step_dpp<4, kDppHalfMirror>(v, (lane & 4) != 0);
    halve_step_dpp<2, kDppXor2>(v, (lane & 2) != 0);
    halve_step_dpp<1, kDppXor1>(v, (lane & 1) != 0);
    return swap32_sum(v[0]);
}

__device__ __forceinline__ float wave_sum(float v) {
    v += dpp_partner<kDppXor1>(v);
    v += dpp_partner<kDppXor2>(v);
    v += dpp_partner<kDppHalfMirror>(v);
    v += dpp_partner<kDppMirror>(v);
    v = swap16_sum(v, v);
    return swap32_sum(v);
}

// Decode Gram block b of a frame with Rext rows into (ra, rb), ra <= rb.
__device__ __forceinline__ void gram_rows(int b, int Rext, int& ra, int& rb) {
    int a = 0;
    while (b >= Rext - a) {
        b -= Rext - a;
        a++;
    }
    ra = a;
    rb = a + b;
}

// Schur Gram of a frame on v_mfma_f32_16x16x4f32.  The Rext*6 E rows are cut
// into nT = ceil(Rext*6 / 16) row tiles; the upper tile pairs (ti <= tj),
// np = nT (nT+1) / 2 of them, are 16 x 16 MFMA accumulators over the pixels.
// A sub-tile's TP pixels are split into ks parts; the 4 / ks waves of one part
// share its pairs (at most kGramSlots each).  ks minimises the per-wave work
// (pairs x pixels), the smaller ks on a tie.
constexpr int kGramSlots = 12;
constexpr int kRowPad = 4;  // E-row stride TP + 4: each row starts 16-byte aligned

typedef float f32x4 __attribute__((ext_vector_type(4)));

__host__ __device__ inline int gram_tiles(int Rext, int pd) { return (Rext * pd + 15) / 16; }

__host__ __device__ inline int gram_ksplit(int Rext, int pd) {
    const int nT = gram_tiles(Rext, pd);
    const int np = nT * (nT + 1) / 2;
    int best = 1, bestc = 1 << 30;
    for (int ks = 1; ks <= 4; ks *= 2) {
        const int wpk = 4 / ks;
        const int slots = (np + wpk - 1) / wpk;
        if (slots > kGramSlots) continue;
        const int c = slots * wpk;  // pairs x (TP / ks) in units of TP / 4
        if (c < bestc) {
            bestc = c;
            best = ks;
        }
    }
    return best;
}

// floats of the cross-part reduction buffer [ks][np][256]
__host__ __device__ inline int gram_red_floats(int Rext, int pd) {
    const int nT = gram_tiles(Rext, pd);
    return gram_ksplit(Rext, pd) * (nT * (nT + 1) / 2) * 256;
}

// whether a frame of Rext E-row blocks fits the Gram's per-wave slots
__host__ __device__ inline bool gram_fits(int Rext, int pd) {
    const int nT = gram_tiles(Rext, pd);
    const int np = nT * (nT + 1) / 2;
    const int wpk = 4 / gram_ksplit(Rext, pd);
    return (np + wpk - 1) / wpk <= kGramSlots;
}

// E rows a frame stores in LDS (see RowMap): at most 6 Rext + pd - 9 (a frame
// with its self block) or 6 Rext - 3 (without)
__host__ __device__ inline int frame_phys_rows(int Rext, int pd) {
    return max(6 * Rext + pd - 9, 6 * Rext - 3);
}

// dynamic LDS of the frame kernel: the stored E rows of a sub-tile, or the Gram
// reduction buffer of the largest frame, whichever is larger
inline size_t frame_lds_bytes(int Rext_max, int TP, int pd) {
    size_t fl = (size_t)frame_phys_rows(Rext_max, pd) * (TP + kRowPad);
    for (int r = 1; r <= Rext_max; r++) fl = std::max(fl, (size_t)gram_red_floats(r, pd));
    return fl * 4;
}

// The Gram works on Rext blocks of PD logical E rows, but LDS stores only the
// rows that can be non-zero: the self block's PD rows, the first 6 rows of each
// edge block (an edge's Sim3 sigma entry is zero) and the pseudo block's two
// rows [w, Ef]; every other logical row reads one shared zero row.  Layout:
// self rows [0, hs PD), edge s at hs PD + 6 s, pseudo at pbase = hs PD + 6 deg,
// the zero row at pbase + 2.
struct RowMap {
    int hs, R, pbase, zrow;
};

template <int PD>
__device__ __forceinline__ int phys_row(const RowMap& m, int r) {
    const int blk = r / PD, n = r - blk * PD;
    if (m.hs && blk == 0) return n;
    if (blk < m.R) return n < 6 ? m.hs * PD + (blk - m.hs) * 6 + n : m.zrow;
    return n < 2 ? m.pbase + n : m.zrow;
}

// One sub-tile's Gram pass of a wave that owns S tile pairs (pair0 + pstride j):
// G[16ti + r][16tj + c] += sum_q Q_q E[16ti + r][q] E[16tj + c][q] over pixels
// [q0, q1), 16 pixels (four k-steps) per step.  Lane (gi, gk) feeds row gi of a
// tile at pixels q + 4gk .. q + 4gk + 3.  A lane whose row is at or past nrows
// reads the last row instead: row r of A reaches only row r of D and row c of B
// only column c, so those padding rows and columns of D, never read back, are
// the only entries it touches.  The loads of a group of up to four
// pairs are issued before its MFMAs, which interleave over the group's
// independent accumulators.
template <int S, int PD>
__device__ __forceinline__ void gram_pass(f32x4 (&acc)[kGramSlots], const float* dyn, const float* sQ, int q0, int q1,
                                          int pair0, int pstride, int nT, int nrows, const RowMap& rm, int ldr, int gi,
                                          int gk) {
    int oA[S], oB[S];
#pragma unroll
    for (int j = 0; j < S; j++) {
        int ti, tj;
        gram_rows(pair0 + pstride * j, nT, ti, tj);
        oA[j] = phys_row<PD>(rm, min(16 * ti + gi, nrows - 1)) * ldr + 4 * gk;
        oB[j] = phys_row<PD>(rm, min(16 * tj + gi, nrows - 1)) * ldr + 4 * gk;
    }
    for (int q = q0; q < q1; q += 16) {
        const float4 qv = *reinterpret_cast<const float4*>(&sQ[q + 4 * gk]);
        // groups of four pairs bound the registers the loads hold
#pragma unroll
        for (int j0 = 0; j0 < S; j0 += 4) {
            float4 av[4], bv[4];
#pragma unroll
            for (int j = 0; j < 4; j++) {
                if (j0 + j < S) {
                    const float4 x = *reinterpret_cast<const float4*>(dyn + oA[j0 + j] + q);
                    const float4 y = *reinterpret_cast<const float4*>(dyn + oB[j0 + j] + q);
                    av[j] = make_float4(x.x * qv.x, x.y * qv.y, x.z * qv.z, x.w * qv.w);
                    bv[j] = y;
                }
            }
#pragma unroll
            for (int j = 0; j < 4; j++)
                if (j0 + j < S) acc[j0 + j] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[j].x, bv[j].x, acc[j0 + j], 0, 0, 0);
#pragma unroll
            for (int j = 0; j < 4; j++)
                if (j0 + j < S) acc[j0 + j] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[j].y, bv[j].y, acc[j0 + j], 0, 0, 0);
#pragma unroll
            for (int j = 0; j < 4; j++)
                if (j0 + j < S) acc[j0 + j] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[j].z, bv[j].z, acc[j0 + j], 0, 0, 0);
#pragma unroll
            for (int j = 0; j < 4; j++)
                if (j0 + j < S) acc[j0 + j] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[j].w, bv[j].w, acc[j0 + j], 0, 0, 0);
        }
    }
}

// One edge at this lane's pixel: the j-side pose partials v[0, 28) (Hjj upper
// triangle, vj, energy; v[28], v[29] Hff, vf with INTR; the rest zero), the
// focal partials fj, the j-side E row er (its Sim3 sigma entry is zero), and the
// frame's depth sums Csum, wsum, Ef and self E row es, which it adds to.  The
// edge is (R, t, sc) with targets / weights `in` = (wu, wv, tu, tv); an invalid
// pixel gives zeros.
template <bool INTR, bool MOTION, bool ROBUST, int PD>
__device__ __forceinline__ void edge_pixel(const float* Rm, const float* t, float sc, bool stereo, const float4& in,
                                           bool valid, float X0, float Y0, float d, float act, int has_self,
                                           const RunArgs& a, float fx, float fy, float cx, float cy, float (&v)[32],
                                           float (&fj)[6], float (&er)[PD], float& Csum, float& wsum, float& Ef,
                                           float (&es)[PD]) {
#pragma unroll
    for (int i = 0; i < 32; i++) v[i] = 0.f;
#pragma unroll
    for (int n = 0; n < 6; n++) fj[n] = 0.f;
#pragma unroll
    for (int n = 0; n < PD; n++) er[n] = 0.f;
    if (valid) {
        PixTerms T;
        pixel_terms<INTR, true, PD>(Rm, t, sc, X0, Y0, d, fx, fy, cx, cy, in.x, in.y, in.z, in.w, a.iscale, T);
        if (ROBUST) {
            T.wu *= robust_weight(T.ru, a.rk);
            T.wv *= robust_weight(T.rv, a.rk);
        }
        const float wdu = T.wu * act, wdv = T.wv * act;
        const float pu = stereo ? 0.f : T.wu;
        const float pv = stereo ? 0.f : T.wv;
        // pose j-side terms; Ju[1] and Jv[0] are structurally zero, so
        // the products they enter are left out (no fast-math folds them)
        int l = 0;
#pragma unroll
        for (int n = 0; n < 6; n++)
#pragma unroll
            for (int m = 0; m <= n; m++) {
                const float hu = (n != 1 && m != 1) ? pu * T.Ju[n] * T.Ju[m] : 0.f;
                const float hv = (n != 0 && m != 0) ? pv * T.Jv[n] * T.Jv[m] : 0.f;
                v[l] = (n != 1 && m != 1) ? ((n != 0 && m != 0) ? hu + hv : hu) : hv;
                l++;
            }
#pragma unroll
        for (int n = 0; n < 6; n++)
            v[21 + n] = n == 1 ? pv * T.rv * T.Jv[n] : n == 0 ? pu * T.ru * T.Ju[n]
                                                              : pu * T.ru * T.Ju[n] + pv * T.rv * T.Jv[n];
        v[27] = T.wu * T.ru * T.ru + T.wv * T.rv * T.rv;
        if (INTR) {
            v[28] = pu * T.Jfu * T.Jfu + pv * T.Jfv * T.Jfv;
            v[29] = pu * T.ru * T.Jfu + pv * T.rv * T.Jfv;
#pragma unroll
            for (int n = 0; n < 6; n++) fj[n] = pu * T.Ju[n] * T.Jfu + pv * T.Jv[n] * T.Jfv;
        }
        if (!MOTION) {
            Csum += wdu * T.Jzu * T.Jzu + wdv * T.Jzv * T.Jzv;
            wsum += wdu * T.ru * T.Jzu + wdv * T.rv * T.Jzv;
            if (INTR) Ef += wdu * T.Jzu * T.Jfu + wdv * T.Jzv * T.Jfv;
            const float cu = act * wdu * T.Jzu, cv = act * wdv * T.Jzv;
#pragma unroll
            for (int n = 0; n < 6; n++)
                er[n] = n == 1 ? cv * T.Jv[n] : n == 0 ? cu * T.Ju[n] : cu * T.Ju[n] + cv * T.Jv[n];
            if (has_self) {
                float o[PD];
                neg_adjT<PD>(Rm, t, sc, er, o);
#pragma unroll
                for (int n = 0; n < PD; n++) es[n] += o[n];
            }
        }
    }
}

// One rig-plan edge at this lane's pixel: as edge_pixel, with camera-frame
// pose partials v and E row er, the self row taking -Adj(M)^T er, and an edge
// within one frame contributing to C and w only.  Under the pinhole model
// Ju[1] and Jv[0] are structurally zero and their products are left out, as in
// edge_pixel; the Mei model has no such zeros.
template <bool MOTION, bool ROBUST, int CAM>
__device__ __forceinline__ void edge_pixel_rig(const float* ep, const float4& in, bool valid, float X0, float Y0,
                                               float d, float act, int has_self, const RunArgs& a, float (&v)[32],
                                               float (&er)[6], float& Csum, float& wsum, float (&es)[6]) {
    constexpr bool SP = CAM == 1;
#pragma unroll
    for (int i = 0; i < 32; i++) v[i] = 0.f;
#pragma unroll
    for (int n = 0; n < 6; n++) er[n] = 0.f;
    if (valid) {
        PixTerms T;
        rig_pixel_terms<CAM, true>(ep, X0, Y0, d, in.x, in.y, in.z, in.w, T);
        if (ROBUST) {
            T.wu *= robust_weight(T.ru, a.rk);
            T.wv *= robust_weight(T.rv, a.rk);
        }
        const bool inframe = ep[12] != 0.f;
        const float wdu = T.wu * act, wdv = T.wv * act;
        const float pu = inframe ? 0.f : T.wu;
        const float pv = inframe ? 0.f : T.wv;
        int l = 0;
#pragma unroll
        for (int n = 0; n < 6; n++)
#pragma unroll
            for (int m = 0; m <= n; m++) {
                if (SP) {
                    const float hu = (n != 1 && m != 1) ? pu * T.Ju[n] * T.Ju[m] : 0.f;
                    const float hv = (n != 0 && m != 0) ? pv * T.Jv[n] * T.Jv[m] : 0.f;
                    v[l] = (n != 1 && m != 1) ? ((n != 0 && m != 0) ? hu + hv : hu) : hv;
                } else {
                    v[l] = pu * T.Ju[n] * T.Ju[m] + pv * T.Jv[n] * T.Jv[m];
                }
                l++;
            }
#pragma unroll
        for (int n = 0; n < 6; n++) {
            if (SP) v[21 + n] = n == 1 ? pv * T.rv * T.Jv[n] : n == 0 ? pu * T.ru * T.Ju[n]
                                                                      : pu * T.ru * T.Ju[n] + pv * T.rv * T.Jv[n];
            else v[21 + n] = pu * T.ru * T.Ju[n] + pv * T.rv * T.Jv[n];
        }
        v[27] = T.wu * T.ru * T.ru + T.wv * T.rv * T.rv;
        if (!MOTION) {
            Csum += wdu * T.Jzu * T.Jzu + wdv * T.Jzv * T.Jzv;
            wsum += wdu * T.ru * T.Jzu + wdv * T.rv * T.Jzv;
            if (!inframe) {
                const float cu = act * wdu * T.Jzu, cv = act * wdv * T.Jzv;
#pragma unroll
                for (int n = 0; n < 6; n++) {
                    if (SP) er[n] = n == 1 ? cv * T.Jv[n] : n == 0 ? cu * T.Ju[n] : cu * T.Ju[n] + cv * T.Jv[n];
                    else er[n] = cu * T.Ju[n] + cv * T.Jv[n];
                }
                if (has_self) {
                    float o[6];
                    neg_adjT<6>(ep + 16, ep + 25, 1.f, er, o);
#pragma unroll
                    for (int n = 0; n < 6; n++) es[n] += o[n];
                }
            }
        }
    }
}

// The LDS record of rig-plan edge e (kRigEP layout)
__device__ __forceinline__ void rig_edge_record(const DevPlan& p, const RunArgs& a, int e, float* ep) {
    rig_edge_transforms(p, a, e, ep, ep + 9, ep + 16, ep + 25);
    ep[12] = (p.ii[e] == p.jj[e]) ? 1.f : 0.f;
    ep[13] = __int_as_float(e);
    ep[14] = 1.f;
    const float* Kj = a.intrinsics + p.qj[e] * a.kstride;
    for (int i = 0; i < 4; i++) ep[28 + i] = Kj[i];
    ep[32] = a.kstride > 4 ? Kj[4] : 0.f;
}

// ---------------------------------------------------------------------------
// 1. frame kernel
// ---------------------------------------------------------------------------
// CAM 0: single-view plans (DROID's pinhole model); CAM 1 / 2: rig plans with
// the pinhole / Mei model (SE3, fixed intrinsics).
template <bool INTR, bool MOTION, bool ROBUST = false, int PD = 6, int CAM = 0>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(2))) void ba_frame_kernel(DevPlan p, RunArgs a, Work w) {
    static_assert(CAM == 0 || (PD == 6 && !INTR), "rig plans: SE3 poses, fixed intrinsics");
    extern __shared__ f32x4 dyn4[];  // stored E rows (RowMap) [.][TP+kRowPad]; reused for the Gram reduction
    float* dyn = reinterpret_cast<float*>(dyn4);
    __shared__ float sEP[kMaxDeg][CAM ? kRigEP : 16];  // R(9) t(3) stereo(12) edge(13) scale(14); rig: kRigEP
    __shared__ __align__(16) float sQ[256];
    __shared__ float sAcc[4][kMaxDeg][kNV];  // per-wave edge partials (each wave adds into its own slot)
    __shared__ float sFen[4];

    const int k = blockIdx.y;
    const int tile = blockIdx.x;
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int TP = p.TP;
    const int ldr = TP + kRowPad;
    const int kf = p.kx[k];
    const int e0 = p.fr_ptr[k];
    const int deg = p.fr_ptr[k + 1] - e0;
    if (deg > kMaxDeg) return;  // a wide frame: ba_frame_wide_kernel
    const int has_self = p.fr_self[k];
    const int R = has_self + deg;
    const int Rext = R + 1;
    const int NB = Rext * (Rext + 1) / 2;
    const int HW = p.HW;
    const float act = a.depth_active[kf];
    // intrinsics of the frame's view (rig plans) or of the single camera
    const float* Kf = CAM ? a.intrinsics + (kf % max(p.V, 1)) * a.kstride : a.intrinsics;
    const float fx = Kf[0], fy = Kf[1], cx = Kf[2], cy = Kf[3];

    if (tid < deg) {
        const int e = p.fr_edges[e0 + tid];
        if (CAM) {
            rig_edge_record(p, a, e, sEP[tid]);
        } else {
            float Rm[9], t[3], sc;
            edge_transform<PD>(a.poses, p.ii[e], p.jj[e], Rm, t, sc);
            for (int i = 0; i < 9; i++) sEP[tid][i] = Rm[i];
            for (int i = 0; i < 3; i++) sEP[tid][9 + i] = t[i];
            sEP[tid][12] = (p.ii[e] == p.jj[e]) ? 1.f : 0.f;
            sEP[tid][13] = __int_as_float(e);
            sEP[tid][14] = sc;
        }
    }
    for (int i = tid; i < 4 * kMaxDeg * kNV; i += kThreads) (&sAcc[0][0][0])[i] = 0.f;
    if (tid < 4) sFen[tid] = 0.f;

    // Gram wave mapping (fixed for all sub-tiles): this wave's pixel part and
    // tile pairs; lane (gi, gk) feeds row gi of a tile at pixel offset 4 gk
    const int wv = __builtin_amdgcn_readfirstlane(wave);
    const int gnrows = Rext * PD;
    const int gnT = gram_tiles(Rext, PD);
    const int gnp = gnT * (gnT + 1) / 2;
    const int gks = gram_ksplit(Rext, PD);
    const int gwpk = 4 / gks;
    const int gpart = wv / gwpk;
    const int gklen = TP / gks;
    const int gnslot = (gnp - wv % gwpk + gwpk - 1) / gwpk;
    const int gi = lane & 15, gk = lane >> 4;
    const RowMap rmap{has_self, R, has_self * PD + 6 * deg, has_self * PD + 6 * deg + 2};
    f32x4 gacc[kGramSlots];
#pragma unroll
    for (int j = 0; j < kGramSlots; j++) gacc[j] = f32x4{0.f, 0.f, 0.f, 0.f};
    float fen_acc = 0.f;
    __syncthreads();

    // weights (u, v) and targets (u, v) of edge slot s at this lane's pixel
    auto edge_in = [&](int s, int pix, bool valid, float4& o) {
        if (valid && s < deg) {
            const int64_t eb = (int64_t)__float_as_int(sEP[s][13]) * 2 * HW + pix;
            o = make_float4(a.weights[eb], a.weights[eb + HW], a.targets[eb], a.targets[eb + HW]);
        }
    };
    for (int sub = 0; sub < p.SUB; sub++) {
        const int pix = (tile * p.SUB + sub) * TP + tid;
        const bool valid = (tid < TP) && (pix < HW);
        const int py = valid ? pix / p.W : 0;
        const int px = valid ? pix - py * p.W : 0;
        float X0 = ((float)px - cx) / fx;
        float Y0 = ((float)py - cy) / fy;
        if (CAM) rig_iproj<CAM>(Kf, (float)px, (float)py, X0, Y0);
        const float d = valid ? a.disps[(int64_t)kf * HW + pix] : 0.f;
        const float ds = (!MOTION && valid) ? a.disps_sens[(int64_t)kf * HW + pix] : 0.f;
        const float et = (!MOTION && valid) ? a.eta[(int64_t)k * HW + pix] : 0.f;
        // the inputs of the next two edges are in flight while an edge is reduced
        float4 in1 = make_float4(0.f, 0.f, 0.f, 0.f), in2 = in1;
        edge_in(0, pix, valid, in1);
        edge_in(1, pix, valid, in2);
        float Csum = 0.f, wsum = 0.f, Ef = 0.f;
        float es[PD];
#pragma unroll
        for (int n = 0; n < PD; n++) es[n] = 0.f;
        for (int s = 0; s < deg; s++) {
            const float* Rm = sEP[s];
            const float* t = sEP[s] + 9;
            const float sc = sEP[s][14];
            const bool stereo = sEP[s][12] != 0.f;
            const float4 in = in1;
            in1 = in2;
            edge_in(s + 2, pix, valid, in2);
            float v[32], fj[6], er[PD];
            if constexpr (CAM != 0) {
#pragma unroll
                for (int n = 0; n < 6; n++) fj[n] = 0.f;
                edge_pixel_rig<MOTION, ROBUST, CAM>(sEP[s], in, valid, X0, Y0, d, act, has_self, a, v, er, Csum,
                                                    wsum, es);
            } else {
                edge_pixel<INTR, MOTION, ROBUST, PD>(Rm, t, sc, stereo, in, valid, X0, Y0, d, act, has_self, a, fx,
                                                     fy, cx, cy, v, fj, er, Csum, wsum, Ef, es);
            }
            if (!MOTION && tid < TP) {
#pragma unroll
                for (int n = 0; n < 6; n++) dyn[(has_self * PD + 6 * s + n) * ldr + tid] = er[n];
            }
            // this wave's sums of the edge's partials; its LDS slot is its own, and
            // a wavefront's LDS accesses complete in order, so no barrier is needed
            const float tot = wave_sum32(v);
            if (lane < 28) sAcc[wave][s][lane] += tot;
            if (INTR) {
                if (lane >= 28 && lane < 30) sAcc[wave][s][lane] += tot;
#pragma unroll
                for (int n = 0; n < 6; n++) {
                    const float f = wave_sum(fj[n]);
                    if (lane == 0) sAcc[wave][s][32 + n] += f;
                }
            }
        }
        if (!MOTION) {
            float Q = 0.f, wt = 0.f, Eft = 0.f;
            if (valid) {
                const float m = (ds > 0.f ? 1.f : 0.f) * act;
                const float C = act * (Csum + et + m * a.alpha) + (1.f - act);
                wt = act * (wsum - m * a.alpha * (d - ds));
                Q = 1.f / C;
                Eft = Ef * act;
                fen_acc += m * a.alpha * (d - ds) * (d - ds);
                w.Qb[(int64_t)k * HW + pix] = Q;
                w.Wb[(int64_t)k * HW + pix] = wt;
            }
            if (tid < TP) {
                if (has_self) {
#pragma unroll
                    for (int n = 0; n < PD; n++) dyn[n * ldr + tid] = es[n];
                }
                dyn[(rmap.pbase + 0) * ldr + tid] = wt;
                dyn[(rmap.pbase + 1) * ldr + tid] = Eft;
                dyn[rmap.zrow * ldr + tid] = 0.f;
                sQ[tid] = Q;
            }
            __syncthreads();
            // G[16ti + r][16tj + c] += sum_q Q_q E[16ti + r][q] E[16tj + c][q]
            // over this wave's pixel part, 16 pixels (four k-steps) per pass;
            // rows at or past Rext*6 feed zeros
            const int qb = gpart * gklen;
            switch (gnslot) {
#define VH_GRAM_CASE(S)                                                                                    \
    case S:                                                                                                \
        gram_pass<S, PD>(gacc, dyn, sQ, qb, qb + gklen, wv % gwpk, gwpk, gnT, gnrows, rmap, ldr, gi, gk);  \
        break;
                VH_GRAM_CASE(1) VH_GRAM_CASE(2) VH_GRAM_CASE(3) VH_GRAM_CASE(4) VH_GRAM_CASE(5) VH_GRAM_CASE(6)
                VH_GRAM_CASE(7) VH_GRAM_CASE(8) VH_GRAM_CASE(9) VH_GRAM_CASE(10) VH_GRAM_CASE(11) VH_GRAM_CASE(12)
#undef VH_GRAM_CASE
                default: break;
            }
            __syncthreads();
        }
    }

    // per-edge partials (the sum of the four waves' slots)
    __syncthreads();
    for (int i = tid; i < deg * kNV; i += kThreads) {
        const int s = i / kNV, v = i % kNV;
        const int e = p.fr_edges[e0 + s];
        w.edge_part[((int64_t)tile * p.E + e) * kNV + v] = (sAcc[0][s][v] + sAcc[1][s][v]) + (sAcc[2][s][v] + sAcc[3][s][v]);
    }
    if (!MOTION) {
        // depth-prior energy
        const float f = wave_sum(fen_acc);
        if (lane == 0) sFen[wave] = f;
        // Gram tiles of every pixel part into LDS [ks][np][lane][4] (the E rows are dead)
        float* red = dyn;
#pragma unroll
        for (int j = 0; j < kGramSlots; j++)
            if (j < gnslot)
                reinterpret_cast<f32x4*>(red)[(gpart * gnp + wv % gwpk + gwpk * j) * 64 + lane] = gacc[j];
        __syncthreads();
        if (tid == 0) w.fen_part[(int64_t)tile * p.K + k] = sFen[0] + sFen[1] + sFen[2] + sFen[3];
        // PD x PD blocks (ra <= rb): element (n, m) is G[PD ra + n][PD rb + m], read
        // as G[c][r] below the diagonal; D[r][c] of a tile sits at lane c + 16 (r / 4),
        // component r % 4
        constexpr int PD2 = PD * PD;
        const int nvals = NB * PD2;
        float* out = w.gram_part + (int64_t)tile * p.GT + (int64_t)p.gram_off[k] * PD2;
        for (int i = tid; i < nvals; i += kThreads) {
            const int b = i / PD2, v = i - PD2 * (i / PD2);
            int ra, rb;
            gram_rows(b, Rext, ra, rb);
            int r = PD * ra + v / PD, c = PD * rb + v % PD;
            if (r > c) {
                const int t = r;
                r = c;
                c = t;
            }
            const int ti = r >> 4, tj = c >> 4;
            const int pair = ti * gnT - ti * (ti - 1) / 2 + (tj - ti);
            const int off = pair * 256 + ((c & 15) + 16 * ((r & 15) >> 2)) * 4 + (r & 3);
            float s = 0.f;
            for (int kp = 0; kp < gks; kp++) s += red[kp * gnp * 256 + off];
            out[i] = s;
        }
    }
}


// ---------------------------------------------------------------------------
// 1b. wide-frame kernel: depth frames with more than kMaxDeg outgoing edges
// ---------------------------------------------------------------------------
// The same sums as the frame kernel for one 256-pixel tile of one wide frame,
// one pixel per thread.  Its edges are set up and reduced kWideGroup at a time
// (their per-edge partials are complete after one pass over the tile), and its
// stored E rows (the RowMap layout) go to the workspace instead of LDS.  The
// Schur Gram then runs over the tile's 256 pixels in chunks of 4 x kGramSlots
// tile pairs (each wave owns its pairs, so no cross-wave reduction), and each
// chunk's elements go straight to their PD x PD blocks.
template <bool INTR, bool MOTION, bool ROBUST, int PD, int CAM = 0>
__global__ __launch_bounds__(kThreads) void ba_frame_wide_kernel(DevPlan p, RunArgs a, Work w) {
    static_assert(CAM == 0 || (PD == 6 && !INTR), "rig plans: SE3 poses, fixed intrinsics");
    __shared__ float sEP[kWideGroup][CAM ? kRigEP : 16];  // R(9) t(3) stereo(12) edge(13) scale(14); rig: kRigEP
    __shared__ __align__(16) float sQ[kThreads];
    __shared__ float sAcc[4][kWideGroup][kNV];
    __shared__ float sFen[4];
    const int slot = blockIdx.y;
    const int k = p.wide[slot];
    const int tile = blockIdx.x;
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int kf = p.kx[k];
    const int e0 = p.fr_ptr[k];
    const int deg = p.fr_ptr[k + 1] - e0;
    const int has_self = p.fr_self[k];
    const int R = has_self + deg;
    const int Rext = R + 1;
    const int HW = p.HW;
    const float act = a.depth_active[kf];
    const float* Kf = CAM ? a.intrinsics + (kf % max(p.V, 1)) * a.kstride : a.intrinsics;
    const float fx = Kf[0], fy = Kf[1], cx = Kf[2], cy = Kf[3];
    const int pix = tile * kThreads + tid;  // 256-pixel tiles, as the frame kernel's TP x SUB
    const bool valid = pix < HW;
    const int py = valid ? pix / p.W : 0;
    const int px = valid ? pix - py * p.W : 0;
    float X0 = ((float)px - cx) / fx;
    float Y0 = ((float)py - cy) / fy;
    if (CAM) rig_iproj<CAM>(Kf, (float)px, (float)py, X0, Y0);
    const float d = valid ? a.disps[(int64_t)kf * HW + pix] : 0.f;
    const RowMap rmap{has_self, R, has_self * PD + 6 * deg, has_self * PD + 6 * deg + 2};
    float* Eg = w.wide_e + ((int64_t)slot * p.T + tile) * (int64_t)p.wide_rows * kThreads;
    float Csum = 0.f, wsum = 0.f, Ef = 0.f;
    float es[PD];
#pragma unroll
    for (int n = 0; n < PD; n++) es[n] = 0.f;

    auto edge_in = [&](int s, float4& o) {
        if (valid && s < deg) {
            const int64_t eb = (int64_t)p.fr_edges[e0 + s] * 2 * HW + pix;
            o = make_float4(a.weights[eb], a.weights[eb + HW], a.targets[eb], a.targets[eb + HW]);
        }
    };
    for (int g0 = 0; g0 < deg; g0 += kWideGroup) {
        const int gn = min(kWideGroup, deg - g0);
        if (tid < gn) {
            const int e = p.fr_edges[e0 + g0 + tid];
            if (CAM) {
                rig_edge_record(p, a, e, sEP[tid]);
            } else {
                float Rm[9], t[3], sc;
                edge_transform<PD>(a.poses, p.ii[e], p.jj[e], Rm, t, sc);
                for (int i = 0; i < 9; i++) sEP[tid][i] = Rm[i];
                for (int i = 0; i < 3; i++) sEP[tid][9 + i] = t[i];
                sEP[tid][12] = (p.ii[e] == p.jj[e]) ? 1.f : 0.f;
                sEP[tid][13] = __int_as_float(e);
                sEP[tid][14] = sc;
            }
        }
        for (int i = tid; i < 4 * kWideGroup * kNV; i += kThreads) (&sAcc[0][0][0])[i] = 0.f;
        float4 in1 = make_float4(0.f, 0.f, 0.f, 0.f), in2 = in1;
        edge_in(g0, in1);
        edge_in(g0 + 1, in2);
        __syncthreads();
        for (int s = 0; s < gn; s++) {
            const float4 in = in1;
            in1 = in2;
            edge_in(g0 + s + 2 < g0 + gn ? g0 + s + 2 : deg, in2);
            float v[32], fj[6], er[PD];
            if constexpr (CAM != 0) {
#pragma unroll
                for (int n = 0; n < 6; n++) fj[n] = 0.f;
                edge_pixel_rig<MOTION, ROBUST, CAM>(sEP[s], in, valid, X0, Y0, d, act, has_self, a, v, er, Csum,
                                                    wsum, es);
            } else {
                edge_pixel<INTR, MOTION, ROBUST, PD>(sEP[s], sEP[s] + 9, sEP[s][14], sEP[s][12] != 0.f, in, valid,
                                                     X0, Y0, d, act, has_self, a, fx, fy, cx, cy, v, fj, er, Csum,
                                                     wsum, Ef, es);
            }
            if (!MOTION) {
#pragma unroll
                for (int n = 0; n < 6; n++) Eg[(int64_t)(has_self * PD + 6 * (g0 + s) + n) * kThreads + tid] = er[n];
            }
            const float tot = wave_sum32(v);
            if (lane < 28) sAcc[wave][s][lane] += tot;
            if (INTR) {
                if (lane >= 28 && lane < 30) sAcc[wave][s][lane] += tot;
#pragma unroll
                for (int n = 0; n < 6; n++) {
                    const float f = wave_sum(fj[n]);
                    if (lane == 0) sAcc[wave][s][32 + n] += f;
                }
            }
        }
        __syncthreads();
        for (int i = tid; i < gn * kNV; i += kThreads) {
            const int sl = i / kNV, v = i % kNV;
            const int e = __float_as_int(sEP[sl][13]);
            w.edge_part[((int64_t)tile * p.E + e) * kNV + v] =
                (sAcc[0][sl][v] + sAcc[1][sl][v]) + (sAcc[2][sl][v] + sAcc[3][sl][v]);
        }
        __syncthreads();
    }
    if (MOTION) return;

    float Q = 0.f, wt = 0.f, Eft = 0.f, fen = 0.f;
    if (valid) {
        const float ds = a.disps_sens[(int64_t)kf * HW + pix];
        const float et = a.eta[(int64_t)k * HW + pix];
        const float m = (ds > 0.f ? 1.f : 0.f) * act;
        const float C = act * (Csum + et + m * a.alpha) + (1.f - act);
        wt = act * (wsum - m * a.alpha * (d - ds));
        Q = 1.f / C;
        Eft = Ef * act;
        fen = m * a.alpha * (d - ds) * (d - ds);
        w.Qb[(int64_t)k * HW + pix] = Q;
        w.Wb[(int64_t)k * HW + pix] = wt;
    }
    if (has_self) {
#pragma unroll
        for (int n = 0; n < PD; n++) Eg[(int64_t)n * kThreads + tid] = es[n];
    }
    Eg[(int64_t)rmap.pbase * kThreads + tid] = wt;
    Eg[(int64_t)(rmap.pbase + 1) * kThreads + tid] = Eft;
    Eg[(int64_t)rmap.zrow * kThreads + tid] = 0.f;
    sQ[tid] = Q;
    const float fsum = wave_sum(fen);
    if (lane == 0) sFen[wave] = fsum;
    __syncthreads();
    if (tid == 0) w.fen_part[(int64_t)tile * p.K + k] = sFen[0] + sFen[1] + sFen[2] + sFen[3];

    // Gram chunks: pair c0 + wave + 4 j of the upper tile pairs
    constexpr int PD2 = PD * PD;
    const int nrows = Rext * PD;
    const int nT = gram_tiles(Rext, PD);
    const int np = nT * (nT + 1) / 2;
    const int gi = lane & 15, gk = lane >> 4;
    float* out = w.gram_part + (int64_t)tile * p.GT + (int64_t)p.gram_off[k] * PD2;
    for (int c0 = 0; c0 < np; c0 += 4 * kGramSlots) {
        const int nslot = (min(np - c0, 4 * kGramSlots) - wave + 3) / 4;
        f32x4 acc[kGramSlots];
#pragma unroll
        for (int j = 0; j < kGramSlots; j++) acc[j] = f32x4{0.f, 0.f, 0.f, 0.f};
        switch (nslot) {
#define VH_GRAM_CASE(S)                                                                                      \
    case S:                                                                                                  \
        gram_pass<S, PD>(acc, Eg, sQ, 0, kThreads, c0 + wave, 4, nT, nrows, rmap, kThreads, gi, gk);        \
        break;
            VH_GRAM_CASE(1) VH_GRAM_CASE(2) VH_GRAM_CASE(3) VH_GRAM_CASE(4) VH_GRAM_CASE(5) VH_GRAM_CASE(6)
            VH_GRAM_CASE(7) VH_GRAM_CASE(8) VH_GRAM_CASE(9) VH_GRAM_CASE(10) VH_GRAM_CASE(11) VH_GRAM_CASE(12)
#undef VH_GRAM_CASE
            default: break;
        }
        // D[r][c] of a tile sits at lane c + 16 (r / 4), component r % 4; the
        // upper elements (r <= c) land in block (r / PD, c / PD), and a
        // diagonal block also takes their mirror
#pragma unroll
        for (int j = 0; j < kGramSlots; j++) {
            if (j >= nslot) break;
            int ti, tj;
            gram_rows(c0 + wave + 4 * j, nT, ti, tj);
            const int c = 16 * tj + (lane & 15);
#pragma unroll
            for (int m = 0; m < 4; m++) {
                const int r = 16 * ti + 4 * (lane >> 4) + m;
                if (r > c || c >= nrows) continue;
                const int ra = r / PD, rb = c / PD, n = r - PD * ra, q = c - PD * rb;
                const int blk = ra * Rext - ra * (ra - 1) / 2 + (rb - ra);
                out[blk * PD2 + n * PD + q] = acc[j][m];
                if (ra == rb && n != q) out[blk * PD2 + q * PD + n] = acc[j][m];
            }
        }
    }
}

// ---------------------------------------------------------------------------
// 2. finalize: tile sums in FP64, i-side edge blocks
// ---------------------------------------------------------------------------
template <int PD, int CAM = 0>
__global__ __launch_bounds__(kThreads) void ba_finalize_kernel(DevPlan p, RunArgs a, Work w, int motion) {
    using RC = Rec<PD>;
    constexpr int PD2 = PD * PD;
    __shared__ double s[64];
    __shared__ double Ad[PD2];
    __shared__ double Arig[CAM ? 36 : 1], Trig[CAM ? 36 : 1], Hrig[CAM ? 36 : 1], vrig[CAM ? 6 : 1];
    __shared__ double RT[13];  // R (row-major), t and s of the edge's relative pose
    __shared__ double Tm[PD2];  // Adj^T Hjj
    __shared__ double tpart[kThreads / kNV][kNV];  // per-slice tile sums of an edge
    const int tid = threadIdx.x;
    const int bid = blockIdx.x;
    if (bid < p.E) {
        const int e = bid;
        // thread (slice, v) sums tiles slice, slice + 6, ... of value v; the loads
        // of a slice are issued together (8 in flight), then the 6 slices are summed
        constexpr int kSl = kThreads / kNV;  // 6
        const int sl = tid / kNV, v = tid % kNV;
        if (sl < kSl) {
            double acc[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
            for (int t0 = sl; t0 < p.T; t0 += 8 * kSl) {
                float x[8];
#pragma unroll
                for (int u = 0; u < 8; u++)
                    x[u] = w.edge_part[((int64_t)min(t0 + u * kSl, p.T - 1) * p.E + e) * kNV + v];
#pragma unroll
                for (int u = 0; u < 8; u++) acc[u] += (t0 + u * kSl < p.T) ? (double)x[u] : 0.0;
            }
            tpart[sl][v] = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
        } else if (tid == kThreads - 1) {
            // the relative pose, while the other threads sum the tiles
            float Rm[9], t[3], sc = 1.f;
            if (CAM) rig_pose_transform(p, a, e, Rm, t);
            else edge_transform<PD>(a.poses, p.ii[e], p.jj[e], Rm, t, sc);
            for (int i = 0; i < 9; i++) RT[i] = Rm[i];
            for (int i = 0; i < 3; i++) RT[9 + i] = t[i];
            RT[12] = sc;
        } else if (CAM && tid == kThreads - 2) {
            rig_adj_inv(a.rig, p.qj[e], Arig);
        }
        __syncthreads();
        if (tid < kNV) {
            double a2 = 0.0;
#pragma unroll
            for (int u = 0; u < kSl; u++) a2 += tpart[u][tid];
            s[tid] = a2;
        }
        // Adj = [[R, [t]x R], [0, R]] (SE3) or [[sR, [t]x R, -t], [0, R, 0], [0, 0, 1]]
        // (Sim3), one entry per thread
        if (tid < PD2) {
            const int r = tid / PD, c = tid % PD;
            double v = 0.0;
            if (r < 3 && c < 3) {
                v = RT[r * 3 + c] * (PD == 7 ? RT[12] : 1.0);
            } else if (r >= 3 && r < 6 && c >= 3 && c < 6) {
                v = RT[(r - 3) * 3 + (c - 3)];
            } else if (r < 3 && c < 6) {  // ([t]x R)[r][c-3]
                const double* td = RT + 9;
                const double tx[3] = {r == 0 ? 0.0 : (r == 1 ? td[2] : -td[1]),
                                      r == 0 ? -td[2] : (r == 1 ? 0.0 : td[0]),
                                      r == 0 ? td[1] : (r == 1 ? -td[0] : 0.0)};
                for (int q = 0; q < 3; q++) v += tx[q] * RT[q * 3 + (c - 3)];
            } else if (r < 3 && c == 6) {
                v = -RT[9 + r];
            } else if (r == 6 && c == 6) {
                v = 1.0;
            }
            Ad[tid] = v;
        }
        __syncthreads();
        double* rec = w.ed + (int64_t)e * RC::SIZE;
        auto hsum = [&](int n, int m) -> double {
            const int hi = n > m ? n : m, lo = n > m ? m : n;
            return s[hi * (hi + 1) / 2 + lo];
        };
        if (CAM) {
            // rig plans: the partials are camera-frame; Hjj = A^T Hc A, vj = A^T vc
            if (tid < 36) {
                const int i = tid / 6, j = tid % 6;
                double t = 0.0;
                for (int q = 0; q < 6; q++) t += Arig[q * 6 + i] * hsum(q, j);
                Trig[tid] = t;
            } else if (tid >= 64 && tid < 70) {
                const int i = tid - 64;
                double t = 0.0;
                for (int q = 0; q < 6; q++) t += Arig[q * 6 + i] * s[21 + q];
                vrig[i] = t;
            }
            __syncthreads();
            if (tid < 36) {
                const int i = tid / 6, j = tid % 6;
                double h = 0.0;
                for (int q = 0; q < 6; q++) h += Trig[i * 6 + q] * Arig[q * 6 + j];
                Hrig[tid] = h;
            }
            __syncthreads();
        }
        // Hjj from the 6-dof partials (its sigma row and column are zero);
        // T = Adj^T Hjj (the i-j block is -T), then Hii = T Adj
        auto hjj = [&](int n, int m) -> double {
            if (n >= 6 || m >= 6) return 0.0;
            return CAM ? Hrig[n * 6 + m] : hsum(n, m);
        };
        auto vjv = [&](int q) -> double { return CAM ? vrig[q] : s[21 + q]; };
        if (tid < PD2) {
            const int i = tid / PD, j = tid % PD;
            rec[RC::HJJ + tid] = hjj(i, j);
            double aTh = 0.0;
            for (int q = 0; q < PD; q++) aTh += Ad[q * PD + i] * hjj(q, j);
            Tm[tid] = aTh;
            rec[RC::HIJ + tid] = -aTh;
        }
        __syncthreads();
        if (tid < PD2) {
            const int i = tid / PD, j = tid % PD;
            double aTha = 0.0;
            for (int q = 0; q < PD; q++) aTha += Tm[i * PD + q] * Ad[q * PD + j];
            rec[RC::HII + tid] = aTha;
        } else if (tid >= 64 && tid < 64 + PD) {
            const int i = tid - 64;
            double vi = 0.0, fi = 0.0;
            for (int q = 0; q < 6; q++) {
                vi += Ad[q * PD + i] * vjv(q);
                fi += Ad[q * PD + i] * s[32 + q];
            }
            rec[RC::VI + i] = -vi;
            rec[RC::VJ + i] = i < 6 ? vjv(i) : 0.0;
            rec[RC::FI + i] = -fi;
            rec[RC::FJ + i] = i < 6 ? s[32 + i] : 0.0;
        } else if (tid == 64 + PD) {
            rec[RC::HFF] = s[28];
            rec[RC::VF] = s[29];
            rec[RC::EN] = s[27];
        }
        return;
    }
    if (motion) return;
    // kGramSplit workgroups per depth frame, each a contiguous share of its Gram
    const int k = (bid - p.E) / kGramSplit, part = (bid - p.E) % kGramSplit;
    const int64_t off = (int64_t)p.gram_off[k] * PD2;
    const int nv = (p.gram_off[k + 1] - p.gram_off[k]) * PD2;
    const int per = (nv + kGramSplit - 1) / kGramSplit;
    const int i0 = part * per, i1 = min(nv, i0 + per);
    for (int i = i0 + tid; i < i1; i += kThreads) {
        // the loads of 32 tiles are issued together (a tile past T reads the
        // last tile and is selected away), 8 partial sums
        const float* g = w.gram_part + off + i;
        double acc[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        for (int t0 = 0; t0 < p.T; t0 += 32) {
            float v[32];
#pragma unroll
            for (int u = 0; u < 32; u++) v[u] = g[(int64_t)min(t0 + u, p.T - 1) * p.GT];
#pragma unroll
            for (int u = 0; u < 32; u++) acc[u & 7] += (t0 + u < p.T) ? (double)v[u] : 0.0;
        }
        w.gd[off + i] = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
    }
    if (part == 0 && tid == 0) {
        double acc = 0.0;
        for (int t0 = 0; t0 < p.T; t0 += 32) {
            float v[32];
#pragma unroll
            for (int u = 0; u < 32; u++) v[u] = w.fen_part[(int64_t)min(t0 + u, p.T - 1) * p.K + k];
#pragma unroll
            for (int u = 0; u < 32; u++) acc += (t0 + u < p.T) ? (double)v[u] : 0.0;
        }
        w.fen[k] = acc;
    }
}

// Rig plans: a depth frame's Gram blocks hold camera-frame E rows for its
// edge rows.  Block (ra, rb), ra <= rb, becomes A_a^T B A_b, where A is the
// edge's Adj(rig[qj]^-1) for an edge row and the identity for the self row and
// the pseudo row [w, Ef].  One workgroup per depth frame, one block per thread
// at a time, a row (then a column) of the block in registers.
__global__ __launch_bounds__(64) void ba_gram_rig_kernel(DevPlan p, RunArgs a, Work w) {
    const int k = blockIdx.x;
    const int has_self = p.fr_self[k];
    const int e0 = p.fr_ptr[k];
    const int R = has_self + p.fr_ptr[k + 1] - e0;
    const int Rext = R + 1;
    const int NB = Rext * (Rext + 1) / 2;
    for (int b = threadIdx.x; b < NB; b += 64) {
        int ra, rb;
        gram_rows(b, Rext, ra, rb);
        const bool ea = ra >= has_self && ra < R, eb = rb >= has_self && rb < R;
        double* B = w.gd + ((int64_t)p.gram_off[k] + b) * 36;
        double A[36];
        if (eb) {
            rig_adj_inv(a.rig, p.qj[p.fr_edges[e0 + rb - has_self]], A);
            for (int i = 0; i < 6; i++) {
                double row[6];
                for (int q = 0; q < 6; q++) row[q] = B[6 * i + q];
                for (int j = 0; j < 6; j++) {
                    double o = 0.0;
                    for (int q = 0; q < 6; q++) o += row[q] * A[6 * q + j];
                    B[6 * i + j] = o;
                }
            }
        }
        if (ea) {
            rig_adj_inv(a.rig, p.qj[p.fr_edges[e0 + ra - has_self]], A);
            for (int j = 0; j < 6; j++) {
                double col[6];
                for (int q = 0; q < 6; q++) col[q] = B[6 * q + j];
                for (int i = 0; i < 6; i++) {
                    double o = 0.0;
                    for (int q = 0; q < 6; q++) o += A[6 * q + i] * col[q];
                    B[6 * i + j] = o;
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// 3. assemble the dense lower triangle of H - S (+ intrinsics) and b - v
// ---------------------------------------------------------------------------
// Address and sign of one assembly-list item's contribution to entry (i, j) of
// its block (sign 0: no contribution; the address is then still a valid one),
// computed without branches so that the loads of several items can be in
// flight together.
template <int PD>
__device__ __forceinline__ const double* assembly_item(const Work& w, int item, int i, int j, int motion,
                                                      double& sign) {
    using RC = Rec<PD>;
    const int kind = item >> 27, idx = item & ((1 << 27) - 1);
    const bool tr = (kind == K_HIJT) || (kind == K_GT);
    const int ij = tr ? j * PD + i : i * PD + j;
    const bool gram = (kind == K_G) || (kind == K_GT);
    const int eoff = (kind == K_HII) ? RC::HII : (kind == K_HJJ) ? RC::HJJ : RC::HIJ;
    sign = gram ? (motion ? 0.0 : -1.0) : (kind <= K_HIJT ? 1.0 : 0.0);
    return gram ? w.gd + (int64_t)idx * (PD * PD) + ij : w.ed + (int64_t)idx * RC::SIZE + eoff + ij;
}

template <int PD>
__global__ __launch_bounds__(64) void ba_assemble_kernel(DevPlan p, RunArgs a, Work w, int motion) {
    using RC = Rec<PD>;
    constexpr int PD2 = PD * PD;
    const int tid = threadIdx.x;
    const int bid = blockIdx.x;
    const int ld = p.ld;
    const int n6 = PD * p.P;  // first row after the pose unknowns
    if (bid < p.nlb) {
        if (tid >= PD2) return;
        const int ab = p.blk_ab[2 * bid], bb = p.blk_ab[2 * bid + 1];
        const int i = tid / PD, j = tid % PD;
        // items in list order, 8 per round: their indices, then their values,
        // are loaded together.  Schur terms exist only when the depth system is
        // solved; in motion-only mode the Gram buffer is never written.
        double s = 0.0;
        const int beg = p.blk_ptr[bid], end = p.blk_ptr[bid + 1];
        for (int it0 = beg; it0 < end; it0 += 8) {
            int item[8];
#pragma unroll
            for (int u = 0; u < 8; u++) item[u] = p.blk_items[min(it0 + u, end - 1)];
            double val[8], sg[8];
#pragma unroll
            for (int u = 0; u < 8; u++) {
                val[u] = *assembly_item<PD>(w, item[u], i, j, motion, sg[u]);
                if (it0 + u >= end) sg[u] = 0.0;
            }
#pragma unroll
            for (int u = 0; u < 8; u++) s += (sg[u] != 0.0) ? sg[u] * val[u] : 0.0;
        }
        if (ab == bb && i == j) s += (double)a.ep + (double)a.lm * s;
        w.A[(int64_t)(PD * ab + i) * ld + PD * bb + j] = s;
        return;
    }
    const int r = bid - p.nlb;
    if (r < p.P) {
        if (tid >= PD) return;
        double sb = 0.0, sf = 0.0;
        const int beg = p.rhs_ptr[r], end = p.rhs_ptr[r + 1];
        for (int it0 = beg; it0 < end; it0 += 8) {
            int item[8];
#pragma unroll
            for (int u = 0; u < 8; u++) item[u] = p.rhs_items[min(it0 + u, end - 1)];
            double vb[8], vf[8], sg[8];
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const int kind = item[u] >> 27, idx = item[u] & ((1 << 27) - 1);
                const bool row = (kind == K_ROW);
                // Gram block (row r, pseudo row): element (tid, 0) is sum Q E w, (tid, 1) sum Q E Ef
                const double* pb = row ? w.gd + (int64_t)idx * PD2 + tid * PD
                                       : w.ed + (int64_t)idx * RC::SIZE + (kind == K_VI ? RC::VI : RC::VJ) + tid;
                const double* pf = row ? w.gd + (int64_t)idx * PD2 + tid * PD + 1
                                       : w.ed + (int64_t)idx * RC::SIZE + (kind == K_VI ? RC::FI : RC::FJ) + tid;
                vb[u] = *pb;
                vf[u] = *pf;
                sg[u] = row ? (motion ? 0.0 : -1.0) : ((kind == K_VI || kind == K_VJ) ? 1.0 : 0.0);
                if (it0 + u >= end) sg[u] = 0.0;
            }
#pragma unroll
            for (int u = 0; u < 8; u++) {
                sb += (sg[u] != 0.0) ? sg[u] * vb[u] : 0.0;
                sf += (sg[u] != 0.0) ? sg[u] * vf[u] : 0.0;
            }
        }
        w.b[PD * r + tid] = sb;
        if (p.intr) w.A[(int64_t)n6 * ld + PD * r + tid] = sf;
        return;
    }
    if (r == p.P) {
        // intrinsics scalar, energy, failure flag: sums split over the wave
        double hff = 0.0, vf = 0.0, en = 0.0;
        for (int e = tid; e < p.E; e += 64) {
            hff += w.ed[(int64_t)e * RC::SIZE + RC::HFF];
            vf += w.ed[(int64_t)e * RC::SIZE + RC::VF];
            en += w.ed[(int64_t)e * RC::SIZE + RC::EN];
        }
        if (!motion) {
            for (int k = tid; k < p.K; k += 64) {
                const int64_t po = (int64_t)p.pseudo_off[k] * PD2;
                hff -= w.gd[po + 1 * PD + 1];
                vf -= w.gd[po + 0 * PD + 1];
                en += w.fen[k];
            }
        }
        hff = wave_sum_f64(hff);
        vf = wave_sum_f64(vf);
        en = wave_sum_f64(en);
        if (tid == 0) {
            if (p.intr) {
                w.A[(int64_t)n6 * ld + n6] = hff + (double)a.iep + (double)a.ilm * hff;
                w.b[n6] = vf;
            }
            if (a.compute_energy) a.energy[a.itr] = (float)en;
            w.flag[0] = 0;
        }
        return;
    }
    // padding rows n..ld-1: identity
    const int prow = p.n + (r - p.P - 1);
    if (prow < ld) {
        for (int c = tid; c < ld; c += 64) w.A[(int64_t)prow * ld + c] = (c == prow) ? 1.0 : 0.0;
        if (tid == 0) w.b[prow] = 0.0;
    }
}

// ---------------------------------------------------------------------------
// 4. blocked FP64 Cholesky (lower), 64 x 64 blocks, right-looking
// ---------------------------------------------------------------------------
// Launches per solve: chol_factor_kernel for block 0, then chol_step_kernel for
// kb = 0 .. nb-2, then chol_backsolve_kernel.  Step kb has one workgroup per
// trailing lower block (i, j), kb < j <= i, which
//   forms L_ik = A_ik X_k and L_jk = A_jk X_k        (X_k = L_kk^-T)
//   updates A_ij -= L_ik L_jk^T
// and, on the diagonal (i == j), stores L_ik^T in the unused upper block (kb, i)
// and carries the right-hand side: y_k = X_k^T b'_k, b'_i -= L_ik y_k.  The
// workgroup of block (kb+1, kb+1) then factors its freshly updated block, so
// the next step's X is ready when the launch ends.  The forward substitution
// is done when the factorisation is; chol_backsolve_kernel runs L^T x = y.
// Entries above the diagonal of A are scratch: the assembly never writes a
// result there, and only the stored L_ik^T blocks are read back.

// 1/sqrt(d): hardware estimate and one Newton step, which squares its
// relative error (zero, negative and NaN pivots are flagged by the caller).
__device__ __forceinline__ double rsqrt_nr(double d) {
    const double y = __builtin_amdgcn_rsq(d);
    const double e = fma(-d * y, y, 1.0);  // 1 - d y^2
    return fma(0.5 * y, e, y);
}

constexpr int kSB = 16;  // sub-block of the diagonal-block inverse

struct FactorLds {
    double invd[kCB];
    double T[3][kSB][kSB + 1];
};

__device__ __forceinline__ double readlane_f64(double x, int lane) {
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

typedef double f64x4 __attribute__((ext_vector_type(4)));

// v_mfma_f64_16x16x4f64 takes A[i][k] from lane i + 16 k and B[k][j] from lane
// j + 16 k, and leaves D[i][j] in register i / 4 of lane j + 16 (i % 4).
// One wave: acc += P Q over k < K (a multiple of 4) for a 16 x 16 tile, with
// P[i][k] at p[i * ps + k] and Q[k][j] at q[k * qs + j], or at q[j * qs + k]
// when QT.
template <bool QT>
__device__ __forceinline__ f64x4 mfma_tile16(const double* p, int ps, const double* q, int qs, int K, int lane,
                                             f64x4 acc) {
    const int i = lane & 15, k = lane >> 4;
    for (int k0 = 0; k0 < K; k0 += 4) {
        const double a = p[i * ps + k0 + k];
        const double b = QT ? q[i * qs + k0 + k] : q[(k0 + k) * qs + i];
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
    }
    return acc;
}

// Wave 0: factor the 16-column panel at column o (rows o..63, row r in lane r).
// Every value a step needs from another row is read from that row's lane with
// readlane into scalar registers: the pivot, and the pivot column's entries
// that update the columns to its right.  Nothing goes through LDS, so the
// pivot chain is VALU latency only (about 4000 cycles per panel against 5500
// with an LDS column broadcast).  Lanes above the panel's diagonal compute
// values that are never read or stored.  Writes L's panel to S and 1/L_jj to
// invd.
__device__ __forceinline__ void factor_panel(double (*S)[kCB + 1], FactorLds& F, int o, int lane, bool& bad) {
    double v[kSB], invs[kSB];
#pragma unroll
    for (int m = 0; m < kSB; m++) v[m] = S[lane][o + m];
    double dn = readlane_f64(v[0], o);
#pragma unroll
    for (int j = 0; j < kSB; j++) {
        const double d = dn;
        bad |= (d <= 0.0);  // a NaN pivot propagates, like the reference's LLT
        const double inv = rsqrt_nr(d);
        const double sc = v[j] * (inv * inv);  // A'[r][o+j] / d
        if (j + 1 < kSB) {
            // the next pivot's column first: its diagonal entry is the next pivot
            v[j + 1] -= sc * readlane_f64(v[j], o + j + 1);
            dn = readlane_f64(v[j + 1], o + j + 1);
        }
#pragma unroll
        for (int c = j + 2; c < kSB; c++) {
            v[c] -= sc * readlane_f64(v[j], o + c);
            __builtin_amdgcn_sched_barrier(0);
        }
        v[j] *= inv;
        invs[j] = inv;
    }
    if (lane < kSB) {
        double iv = invs[0];
#pragma unroll
        for (int m = 1; m < kSB; m++) iv = (lane == m) ? invs[m] : iv;
        F.invd[o + lane] = iv;
    }
#pragma unroll
    for (int m = 0; m < kSB; m++)
        if (lane >= o + m) S[lane][o + m] = v[m];
}

// S[r][c] -= sum_{m in [p0, p0+16)} S[r][m] S[c][m] over 4 x 4 tiles with
// c in [c_lo, c_hi) and r in [c_lo, 64), tiles wholly above the diagonal
// skipped; thread t of nt.
__device__ __forceinline__ void panel_update(double (*S)[kCB + 1], int p0, int c_lo, int c_hi, int t, int nt) {
    const int ntc = (c_hi - c_lo) / 4, ntr = (kCB - c_lo) / 4;
    for (int task = t; task < ntc * ntr; task += nt) {
        const int tc = c_lo + 4 * (task % ntc), tr = c_lo + 4 * (task / ntc);
        if (tr < tc) continue;
        double acc[4][4];
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 4; j++) acc[i][j] = 0.0;
#pragma unroll
        for (int m = 0; m < kSB; m++) {
            double a[4], bb[4];
#pragma unroll
            for (int i = 0; i < 4; i++) {
                a[i] = S[tr + i][p0 + m];
                bb[i] = S[tc + i][p0 + m];
            }
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < 4; j++) acc[i][j] += a[i] * bb[j];
        }
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 4; j++) S[tr + i][tc + j] -= acc[i][j];
    }
}

// One wave: M_tt = L_tt^-1 for the 16 x 16 diagonal block at o.  Lane r < 16
// solves L_tt x = e_r by forward substitution (column r of M_tt).
__device__ __forceinline__ void diag_inverse(const double (*S)[kCB + 1], double (*M)[kCB + 1], const FactorLds& F,
                                             int o, int lane) {
    const int r = lane & (kSB - 1);
    double lq[kSB][kSB];  // L_tt[c][j], broadcast reads
    double x[kSB];
#pragma unroll
    for (int c = 0; c < kSB; c++) {
        x[c] = (c == r) ? 1.0 : 0.0;
#pragma unroll
        for (int j = 0; j < c; j++) lq[c][j] = S[o + c][o + j];
    }
#pragma unroll
    for (int j = 0; j < kSB; j++) {
        const double xj = x[j] * F.invd[o + j];
        x[j] = xj;
#pragma unroll
        for (int c = j + 1; c < kSB; c++) x[c] -= xj * lq[c][j];
    }
    if (lane < kSB) {
#pragma unroll
        for (int c = 0; c < kSB; c++) M[o + c][o + r] = x[c];
    }
}

// T_q = sum_{p=q}^{t-1} L_tp M_pq for q < t (16 x 16 blocks), thread u of nu.
__device__ __forceinline__ void inverse_row_partial(const double (*S)[kCB + 1], const double (*M)[kCB + 1],
                                                    FactorLds& F, int t, int u, int nu) {
    for (int e = u; e < t * kSB * kSB; e += nu) {
        const int q = e >> 8, i = (e >> 4) & 15, j = e & 15;
        double acc[4] = {0.0, 0.0, 0.0, 0.0};
        for (int k = kSB * q; k < kSB * t; k += 4) {
#pragma unroll
            for (int m = 0; m < 4; m++) acc[m] += S[kSB * t + i][k + m] * M[k + m][kSB * q + j];
        }
        F.T[q][i][j] = (acc[0] + acc[1]) + (acc[2] + acc[3]);
    }
}

// Factor the 64 x 64 block whose lower triangle is in S (256 threads) and
// write X = L^-T (row-major) to Xout.  S receives L; M is scratch.
// Four rounds, one per 16-column panel s (o = 16 s):
//   A: wave 0 factors panel s.  Meanwhile, for t = s - 1, wave 1 forms the
//      diagonal inverse M_tt, and waves 2-3 apply panel t's update to the
//      columns right of panel s and form T_q = sum_{p=q}^{t-1} L_tp M_pq.
//   B: one 16 x 16 matrix-core tile per wave: panel s's update of panel s+1's
//      columns (3 - s tiles), and block row t of M = L^-1: M_tq = -M_tt T_q.
// A last round does block row 3 of M, then X = M^T is stored.  Entries above
// the diagonal of S are scratch (the diagonal tiles update them too).  In
// every round, waves 1-3 then run hook(round, wave, lane): in round 0 the
// caller may use it to finish columns 16..63 of S while panel 0 is factored,
// later for work of its own.
struct NoHook {
    __device__ void operator()(int, int, int) const {}
};

template <typename Hook = NoHook>
__device__ __forceinline__ void factor_block(double (*S)[kCB + 1], double (*M)[kCB + 1], FactorLds& F,
                                             double* __restrict__ Xout, int* flag, const Hook& hook = Hook()) {
    constexpr int LS = kCB + 1, LT = kSB + 1;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform
    const int ti = lane & 15, tk = lane >> 4;                 // tile element (4 r + tk, ti) in register r
    const f64x4 zero = {0.0, 0.0, 0.0, 0.0};
    bool bad = false;
#pragma unroll
    for (int s = 0; s < kCB / kSB; s++) {
        const int o = s * kSB, t = s - 1;
        if (wv == 0) {
            factor_panel(S, F, o, lane, bad);
        } else {
            if (wv == 1) {
                if (t >= 0) diag_inverse(S, M, F, kSB * t, lane);
            } else {
                if (t >= 0 && o + kSB < kCB) panel_update(S, kSB * t, o + kSB, kCB, tid - 128, 128);
                if (t >= 1) inverse_row_partial(S, M, F, t, tid - 128, 128);
            }
            hook(s, wv, lane);
        }
        __syncthreads();
        const int npu = (s + 1 < kCB / kSB) ? kCB / kSB - 1 - s : 0;
        if (wv < npu) {
            const int r0 = o + kSB * (wv + 1), c0 = o + kSB;
            const f64x4 acc = mfma_tile16<true>(&S[r0][o], LS, &S[c0][o], LS, kSB, lane, zero);
#pragma unroll
            for (int r = 0; r < 4; r++) S[r0 + 4 * r + tk][c0 + ti] -= acc[r];
        } else if (wv - npu < t) {
            const int q = wv - npu;
            const f64x4 acc = mfma_tile16<false>(&M[kSB * t][kSB * t], LS, &F.T[q][0][0], LT, kSB, lane, zero);
#pragma unroll
            for (int r = 0; r < 4; r++) M[kSB * t + 4 * r + tk][kSB * q + ti] = -acc[r];
        }
        __syncthreads();
    }
    if (tid == 0 && bad) atomicExch(flag, 1);
    constexpr int tl = kCB - kSB;  // last block row
    if (wv == 1) {
        diag_inverse(S, M, F, tl, lane);
    } else {
        const int q = wv == 0 ? 0 : wv - 1;
        const f64x4 acc = mfma_tile16<false>(&S[tl][kSB * q], LS, &M[kSB * q][kSB * q], LS, tl - kSB * q, lane, zero);
#pragma unroll
        for (int r = 0; r < 4; r++) F.T[q][4 * r + tk][ti] = acc[r];
    }
    __syncthreads();
    if (wv < 3) {
        const f64x4 acc = mfma_tile16<false>(&M[tl][tl], LS, &F.T[wv][0][0], LT, kSB, lane, zero);
#pragma unroll
        for (int r = 0; r < 4; r++) M[tl + 4 * r + tk][kSB * wv + ti] = -acc[r];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 16; q++) {
        const int e = tid + q * 256;
        const int r = e >> 6, c = e & 63;
        Xout[e] = (c >= r) ? M[c][r] : 0.0;
    }
}

// Load a 64 x 64 FP64 block (row stride ld) into LDS, all 16 loads per thread
// in flight before the first store.  TRANS stores it transposed.
template <bool TRANS>
__device__ __forceinline__ void load_block(const double* __restrict__ src, int64_t ld, double (*dst)[kCB + 1],
                                           int tid) {
    double tmp[16];
#pragma unroll
    for (int q = 0; q < 16; q++) {
        const int e = tid + q * 256;
        tmp[q] = src[(int64_t)(e >> 6) * ld + (e & 63)];
    }
#pragma unroll
    for (int q = 0; q < 16; q++) {
        const int e = tid + q * 256;
        if (TRANS) dst[e & 63][e >> 6] = tmp[q];
        else dst[e >> 6][e & 63] = tmp[q];
    }
}

// 64 x 64 x 64 FP64 products on the matrix cores (layout at mfma_tile16).
// Wave w owns the 32 x 32 quadrant (w / 2, w % 2) of the result as 2 x 2 tiles.

// row / column of element r of tile (ta, tb) of quadrant (qr, qc) in this lane
__device__ __forceinline__ int quad_row(int qr, int ta, int r, int lane) { return 32 * qr + 16 * ta + 4 * r + (lane >> 4); }
__device__ __forceinline__ int quad_col(int qc, int tb, int lane) { return 32 * qc + 16 * tb + (lane & 15); }

// acc += P Q (BT false: Q[m][c]) or P Q^T (BT true: Q[c][m]) over the wave's
// quadrant, with P[r][m] and Q in LDS.
template <bool BT>
__device__ __forceinline__ void mfma_quad_64(const double (*P)[kCB + 1], const double (*Q)[kCB + 1], int qr, int qc,
                                             int lane, f64x4 (&acc)[2][2]) {
    const int i = lane & 15, k = lane >> 4;
#pragma unroll 4
    for (int m0 = 0; m0 < kCB; m0 += 4) {
        const double a0 = P[32 * qr + i][m0 + k], a1 = P[32 * qr + 16 + i][m0 + k];
        const double b0 = BT ? Q[32 * qc + i][m0 + k] : Q[m0 + k][32 * qc + i];
        const double b1 = BT ? Q[32 * qc + 16 + i][m0 + k] : Q[m0 + k][32 * qc + 16 + i];
        acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
        acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
        acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
        acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
    }
}

// acc[cb] += P X over the 16-row strip w of the result and its four column
// tiles cb, for an upper triangular X (X[m][c] = 0 for m > c): column tile cb
// needs m < 16 (cb + 1) only, 40 of the 64 k-steps, the same for every strip.
// Element r of acc[cb] in this lane is row 16 w + 4 r + lane / 16, column
// 16 cb + lane % 16.
__device__ __forceinline__ void mfma_strip_upper(const double (*P)[kCB + 1], const double (*X)[kCB + 1], int w,
                                                 int lane, f64x4 (&acc)[4]) {
    const int i = lane & 15, k = lane >> 4;
#pragma unroll
    for (int m0 = 0; m0 < kCB; m0 += 4) {
        const double a = P[16 * w + i][m0 + k];
#pragma unroll
        for (int cb = 0; cb < 4; cb++)
            if (m0 < 16 * (cb + 1)) acc[cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, X[m0 + k][16 * cb + i], acc[cb], 0, 0, 0);
    }
}

__global__ __launch_bounds__(256) void chol_factor_kernel(const double* __restrict__ A, double* __restrict__ LinvT,
                                                          int ld, int kb, int* flag) {
    __shared__ double S[kCB][kCB + 1];
    __shared__ double M[kCB][kCB + 1];
    __shared__ FactorLds F;
    load_block<false>(A + (int64_t)kb * kCB * ld + kb * kCB, ld, S, threadIdx.x);
    __syncthreads();
    factor_block(S, M, F, LinvT + (int64_t)kb * kCB * kCB, flag);
}

// R[16 a + i][16 c + j] -= sum_m L[16 a + i][m] L[16 c + j][m] for one 16 x 16
// tile (a, c) of a diagonal block (one wave; L and R in LDS).
__device__ __forceinline__ void syrk_tile(double (*R)[kCB + 1], const double (*L)[kCB + 1], int a, int c, int lane) {
    const f64x4 acc = mfma_tile16<true>(&L[16 * a][0], kCB + 1, &L[16 * c][0], kCB + 1, kCB, lane,
                                        f64x4{0.0, 0.0, 0.0, 0.0});
    const int ti = lane & 15, tk = lane >> 4;
#pragma unroll
    for (int r = 0; r < 4; r++) R[16 * a + 4 * r + tk][16 * c + ti] -= acc[r];
}

// lower 16 x 16 tiles of a 64 x 64 block, column 0 first: (a, c) = (kTileA[n], kTileC[n])
__constant__ const int kTileA[10] = {0, 1, 2, 3, 1, 2, 3, 2, 3, 3};
__constant__ const int kTileC[10] = {0, 0, 0, 0, 1, 1, 1, 2, 2, 3};

__global__ __launch_bounds__(256) void chol_step_kernel(double* __restrict__ A, double* __restrict__ LinvT,
                                                        double* __restrict__ b, double* __restrict__ y, int ld, int kb,
                                                        int* flag) {
    __shared__ double R0[kCB][kCB + 1];  // X_k (R0[m][c] = X[m][c]); then factorisation scratch
    __shared__ double R1[kCB][kCB + 1];  // A_ik, then L_ik
    __shared__ double R2[kCB][kCB + 1];  // A_jk, then L_jk (off-diagonal); A_ii, updated, factored (diagonal)
    __shared__ FactorLds F;
    __shared__ double yk[kCB];
    __shared__ double bk[kCB];
    int idx = blockIdx.x, row = 0;
    while (idx > row) {
        idx -= row + 1;
        row++;
    }
    const int bi = kb + 1 + row, bj = kb + 1 + idx;
    const bool diag = (bi == bj);
    const bool next = diag && bi == kb + 1;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6), qr = wv >> 1, qc = wv & 1;
    const int c0 = kb * kCB;
    double* Ab = A + (int64_t)bi * kCB * ld + bj * kCB;
    load_block<false>(LinvT + (int64_t)kb * kCB * kCB, kCB, R0, tid);
    load_block<false>(A + (int64_t)bi * kCB * ld + c0, ld, R1, tid);
    load_block<false>(diag ? Ab : A + (int64_t)bj * kCB * ld + c0, ld, R2, tid);
    if (diag && tid < kCB) bk[tid] = b[c0 + tid];
    double cur[2][2][4];
    if (!diag) {
#pragma unroll
        for (int ta = 0; ta < 2; ta++)
#pragma unroll
            for (int tb = 0; tb < 2; tb++)
#pragma unroll
                for (int r = 0; r < 4; r++)
                    cur[ta][tb][r] = Ab[(int64_t)quad_row(qr, ta, r, lane) * ld + quad_col(qc, tb, lane)];
    }
    __syncthreads();
    f64x4 li[4], lj[4];
#pragma unroll
    for (int cb = 0; cb < 4; cb++) li[cb] = lj[cb] = f64x4{0.0, 0.0, 0.0, 0.0};
    mfma_strip_upper(R1, R0, wv, lane, li);  // L_ik = A_ik X_k
    if (!diag) mfma_strip_upper(R2, R0, wv, lane, lj);
    if (diag && tid < kCB) {
        double s[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int m = 0; m < kCB; m++) s[m & 3] += bk[m] * R0[m][tid];
        yk[tid] = (s[0] + s[1]) + (s[2] + s[3]);
    }
    __syncthreads();
#pragma unroll
    for (int cb = 0; cb < 4; cb++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int rr = 16 * wv + 4 * r + (lane >> 4), cc = 16 * cb + (lane & 15);
            R1[rr][cc] = li[cb][r];
            if (!diag) R2[rr][cc] = lj[cb][r];
        }
    if (next && tid < kCB) y[c0 + tid] = yk[tid];
    __syncthreads();
    if (!diag) {
        f64x4 acc[2][2];
#pragma unroll
        for (int ta = 0; ta < 2; ta++)
#pragma unroll
            for (int tb = 0; tb < 2; tb++) acc[ta][tb] = f64x4{0.0, 0.0, 0.0, 0.0};
        mfma_quad_64<true>(R1, R2, qr, qc, lane, acc);  // L_ik L_jk^T
#pragma unroll
        for (int ta = 0; ta < 2; ta++)
#pragma unroll
            for (int tb = 0; tb < 2; tb++)
#pragma unroll
                for (int r = 0; r < 4; r++)
                    Ab[(int64_t)quad_row(qr, ta, r, lane) * ld + quad_col(qc, tb, lane)] = cur[ta][tb][r] - acc[ta][tb][r];
        return;
    }
    // diagonal block: L_ik^T goes into the upper block (kb, bi) (row c of it is
    // column c of L_ik) and the right-hand side b'_i -= L_ik y_k; thread u of nu
    auto store_lt_and_rhs = [&](int u, int nu) {
        for (int e = u; e < kCB * kCB; e += nu)
            A[(int64_t)(c0 + (e >> 6)) * ld + bi * kCB + (e & 63)] = R1[e & 63][e >> 6];
        if (u < kCB) {
            double s[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int m = 0; m < kCB; m++) s[m & 3] += R1[u][m] * yk[m];
            b[bi * kCB + u] -= (s[0] + s[1]) + (s[2] + s[3]);
        }
    };
    // A_ii -= L_ik L_ik^T on the lower tiles, column 0 first (one tile per wave)
    syrk_tile(R2, R1, wv, 0, lane);
    if (!next) {
        store_lt_and_rhs(tid, 256);
        for (int n = 4 + wv; n < 10; n += 4) syrk_tile(R2, R1, kTileA[n], kTileC[n], lane);
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 16; q++) {
            const int e = tid + q * 256;
            Ab[(int64_t)(e >> 6) * ld + (e & 63)] = R2[e >> 6][e & 63];
        }
        return;
    }
    // this workgroup owns the next diagonal block: factor it while waves 1-3
    // finish its remaining six tiles (columns 16..63) during panel 0, and the
    // L_ik^T store and the right-hand side during panel 1
    __syncthreads();
    factor_block(R2, R0, F, LinvT + (int64_t)(kb + 1) * kCB * kCB, flag, [&](int round, int w, int ln) {
        if (round == 0) {
            syrk_tile(R2, R1, kTileA[2 + 2 * w], kTileC[2 + 2 * w], ln);
            syrk_tile(R2, R1, kTileA[3 + 2 * w], kTileC[3 + 2 * w], ln);
        } else if (round == 1) {
            store_lt_and_rhs(64 * (w - 1) + ln, 192);
        }
    });
}

// Backward substitution L^T x = y, zeros when the factorisation failed.
//   last block:  y = X^T b'            x = X y
//   block k:     t = y_k - U_k x_>k    x_k = X_k t
// where row c of U_k is row kb*64+c of A right of the diagonal block (the
// stored L_ik^T blocks).  16 waves; wave w takes rows w + 16 u of U_k with its
// lanes along the columns (coalesced).  The loads of block k-1 (X, the first
// kPre column chunks of U, y) are issued before block k is computed, so only
// the first block waits for memory.
constexpr int kBT = 1024;           // threads of the backward substitution
constexpr int kBW = kBT / 64;       // its waves
constexpr int kBR = kCB / kBW;      // rows of a block per wave
constexpr int kPre = 3;             // 128-column chunks of a U row held in registers (ld <= 448)

// Operands of block kb of the backward substitution for thread (wv, lane), all
// as 16-byte pairs: X_k entries (2 tid + 2 kBT q, +1) in row-major order,
// U_k[wv + kBW u][2 lane + 128 c, +1] for c < kPre, and y_k of the row whose sum
// this lane ends up holding (back_row).
struct BackOps {
    double2 x[kBR / 2], u[kBR][kPre];
    double y;
};

// After the halving reduction, lane l holds row wv + kBW back_row(l) of the block.
__device__ __forceinline__ int back_row(int lane) { return 2 * ((lane >> 5) & 1) + ((lane >> 4) & 1); }

__device__ __forceinline__ void back_fetch(BackOps& o, const double* __restrict__ A, const double* __restrict__ LinvT,
                                           const double* __restrict__ y, int ld, int nb, int kb, int tid) {
    const int lane = tid & 63, wv = tid >> 6;
    const double2* X = reinterpret_cast<const double2*>(LinvT + (int64_t)kb * kCB * kCB);
#pragma unroll
    for (int q = 0; q < kBR / 2; q++) o.x[q] = X[tid + q * kBT];
    const int c0 = kb * kCB, r0 = c0 + kCB, nr = (nb - kb - 1) * kCB;
    o.y = y[c0 + wv + kBW * back_row(lane)];
#pragma unroll
    for (int u = 0; u < kBR; u++) {
        const double2* row = reinterpret_cast<const double2*>(A + (int64_t)(c0 + wv + kBW * u) * ld + r0);
#pragma unroll
        for (int c = 0; c < kPre; c++) {
            const int col = 2 * lane + 128 * c;
            o.u[u][c] = (col < nr) ? row[lane + 64 * c] : make_double2(0.0, 0.0);
        }
    }
}

// Sum of 4 per-lane values over the wave: a half-wave swap keeps rows {0,1}
// or {2,3}, a 16-lane-row swap one of those, and four DPP steps sum within the
// 16-lane row.  Lane l ends with the sum of row back_row(l).
__device__ __forceinline__ double back_reduce4(const double (&s)[4]) {
    // after the half-wave swap the lower half holds rows 0 / 1 and the upper
    // half rows 2 / 3; after the row swap, rows 0, 2 of the wave hold the
    // first of those and rows 1, 3 the second
    const double a0 = swap32_sum_f64(s[0], s[2]);
    const double a1 = swap32_sum_f64(s[1], s[3]);
    return row_sum16_f64(swap16_sum_f64(a0, a1));
}

__global__ __launch_bounds__(kBT) void chol_backsolve_kernel(const double* __restrict__ A,
                                                             const double* __restrict__ LinvT, int ld, int nb,
                                                             const double* __restrict__ b,
                                                             const double* __restrict__ y, double* __restrict__ x,
                                                             const int* flag, int npose, float* dx) {
    static_assert(kBR == 4, "back_reduce4 sums 4 rows per wave");
    static_assert(kBT * 2 * (kBR / 2) == kCB * kCB, "X_k pairs cover the block");
    extern __shared__ double xs[];  // [ld]
    __shared__ double Xk[kCB][kCB + 1];
    __shared__ double t[kCB];
    __shared__ double part[kBW][kCB];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wv = tid >> 6;
    if (flag[0] != 0) {
        for (int i = tid; i < ld; i += kBT) x[i] = 0.0;
        for (int i = tid; i < npose; i += kBT) dx[i] = 0.f;
        return;
    }
    // The operands of block kb-1 are requested as soon as those of block kb
    // are consumed, so their latency overlaps block kb's reductions.
    BackOps o;
    back_fetch(o, A, LinvT, y, ld, nb, nb - 1, tid);
    for (int kb = nb - 1; kb >= 0; kb--) {
        const int c0 = kb * kCB;
        const bool last = (kb == nb - 1);
        double s[kBR];
#pragma unroll
        for (int u = 0; u < kBR; u++) s[u] = 0.0;
        if (!last) {
            const int r0 = c0 + kCB, nr = (nb - kb - 1) * kCB;
#pragma unroll
            for (int c = 0; c < kPre; c++) {
                const int col = 2 * lane + 128 * c;
                const double x0 = (col < nr) ? xs[r0 + col] : 0.0, x1 = (col < nr) ? xs[r0 + col + 1] : 0.0;
#pragma unroll
                for (int u = 0; u < kBR; u++) s[u] += o.u[u][c].x * x0 + o.u[u][c].y * x1;
            }
            for (int col = lane + 128 * kPre; col < nr; col += 64) {
                const double xq = xs[r0 + col];
#pragma unroll
                for (int u = 0; u < kBR; u++) s[u] += A[(int64_t)(c0 + wv + kBW * u) * ld + r0 + col] * xq;
            }
        }
        const double yc = o.y;
#pragma unroll
        for (int q = 0; q < kBR / 2; q++) {
            const int e = 2 * (tid + q * kBT);
            Xk[e >> 6][e & 63] = o.x[q].x;
            Xk[e >> 6][(e & 63) + 1] = o.x[q].y;
        }
        if (kb > 0) back_fetch(o, A, LinvT, y, ld, nb, kb - 1, tid);
        if (!last) {
            const double r = back_reduce4(s);
            if ((lane & 15) == 0) t[wv + kBW * back_row(lane)] = yc - r;
        }
        __syncthreads();
        if (last) {
            // y_k = X^T b': wave wv takes rows m in [kBR wv, kBR wv + kBR) of column lane
            double a = 0.0;
#pragma unroll
            for (int u = 0; u < kBR; u++) a += b[c0 + kBR * wv + u] * Xk[kBR * wv + u][lane];
            part[wv][lane] = a;
            __syncthreads();
            if (wv == 0) {
                double a2 = 0.0;
#pragma unroll
                for (int u = 0; u < kBW; u++) a2 += part[u][lane];
                t[lane] = a2;
            }
            __syncthreads();
        }
        // x_k[r] = sum_c X[r][c] t[c]: wave wv takes c in [kBR wv, kBR wv + kBR)
        double s2 = 0.0;
#pragma unroll
        for (int u = 0; u < kBR; u++) s2 += Xk[lane][kBR * wv + u] * t[kBR * wv + u];
        part[wv][lane] = s2;
        __syncthreads();
        if (wv == 0) {
            double a = 0.0;
#pragma unroll
            for (int u = 0; u < kBW; u++) a += part[u][lane];
            xs[c0 + lane] = a;
        }
        __syncthreads();
    }
    for (int i = tid; i < ld; i += kBT) x[i] = xs[i];
    for (int i = tid; i < npose; i += kBT) dx[i] = (float)xs[i];
}

// ---------------------------------------------------------------------------
// 5. back-substitution: dz = Q (w - sum_r E_r . dx_r) - Q Ef df
// ---------------------------------------------------------------------------
template <bool INTR, int PD = 6, int CAM = 0>
__global__ __launch_bounds__(kThreads) void ba_backsub_kernel(DevPlan p, RunArgs a, Work w) {
    static_assert(CAM == 0 || (PD == 6 && !INTR), "rig plans: SE3 poses, fixed intrinsics");
    __shared__ float sEP[kMaxDeg][CAM ? kRigEP : 16];
    __shared__ float sU[kMaxDeg][6];  // the sigma entry of u meets a zero E entry
    const int k = blockIdx.y;
    const int tid = threadIdx.x;
    const int kf = p.kx[k];
    const int e0 = p.fr_ptr[k];
    const int deg = p.fr_ptr[k + 1] - e0;
    const int has_self = p.fr_self[k];
    const float act = a.depth_active[kf];
    const float* Kf = CAM ? a.intrinsics + (kf % max(p.V, 1)) * a.kstride : a.intrinsics;
    const float fx = Kf[0], fy = Kf[1], cx = Kf[2], cy = Kf[3];
    const int HW = p.HW;
    // edge slot tid of the group starting at edge g0: its relative pose and
    // u = dx[jj] (if 1 <= jj-t0 < P) - Adj(Tij) dx[k] (self row, if 1 <= k-t0 < P);
    // rig plans: u = A (dx[pj] - Adj(G) dx[pi]) with the same conditions, and
    // zero for an edge within one frame (its E rows are zero)
    auto setup = [&](int g0, int gn) {
        if (tid >= gn) return;
        const int e = p.fr_edges[e0 + g0 + tid];
        float Rm[9], t[3], sc = 1.f;
        if (CAM) {
            rig_edge_record(p, a, e, sEP[tid]);
            rig_pose_transform(p, a, e, Rm, t);
        } else {
            edge_transform<PD>(a.poses, p.ii[e], p.jj[e], Rm, t, sc);
            for (int i = 0; i < 9; i++) sEP[tid][i] = Rm[i];
            for (int i = 0; i < 3; i++) sEP[tid][9 + i] = t[i];
            sEP[tid][14] = sc;
        }
        double u[6] = {0, 0, 0, 0, 0, 0};
        const int lj = p.jj[e] - p.t0;
        if (lj >= 1 && lj < p.P)
            for (int n = 0; n < 6; n++) u[n] += w.x[PD * lj + n];
        const int lk = (CAM ? p.kpose[k] : kf) - p.t0;
        if (has_self && lk >= 1 && lk < p.P) {
            const double* v = w.x + PD * lk;
            double rv[3], rw[3];
            for (int r = 0; r < 3; r++) {
                rv[r] = Rm[3 * r] * v[0] + Rm[3 * r + 1] * v[1] + Rm[3 * r + 2] * v[2];
                rw[r] = Rm[3 * r] * v[3] + Rm[3 * r + 1] * v[4] + Rm[3 * r + 2] * v[5];
            }
            // top = s R v_tau + t x (R v_phi) - t v_sigma, bottom = R v_phi
            // (the sigma entry of Adj v is v_sigma, dotted with a zero E entry)
            const double sv = PD == 7 ? (double)sc : 1.0;
            const double vs = PD == 7 ? v[PD - 1] : 0.0;
            u[0] -= sv * rv[0] + (t[1] * rw[2] - t[2] * rw[1]) - t[0] * vs;
            u[1] -= sv * rv[1] + (t[2] * rw[0] - t[0] * rw[2]) - t[1] * vs;
            u[2] -= sv * rv[2] + (t[0] * rw[1] - t[1] * rw[0]) - t[2] * vs;
            u[3] -= rw[0];
            u[4] -= rw[1];
            u[5] -= rw[2];
        }
        if (CAM) {
            double A[36], ua[6];
            rig_adj_inv(a.rig, p.qj[e], A);
            const bool inframe = p.ii[e] == p.jj[e];
            for (int i = 0; i < 6; i++) {
                double o = 0.0;
                for (int q = 0; q < 6; q++) o += A[6 * i + q] * u[q];
                ua[i] = inframe ? 0.0 : o;
            }
            for (int n = 0; n < 6; n++) u[n] = ua[n];
        }
        for (int n = 0; n < 6; n++) sU[tid][n] = (float)u[n];
    };
    setup(0, min(deg, kMaxDeg));
    // this pixel's own inputs and the first two edges' are requested before the
    // barrier, so their latency overlaps the edge set-up above
    const int pix = blockIdx.x * kThreads + tid;
    const bool valid = pix < HW;
    const int64_t dpos = (int64_t)kf * HW + pix;
    const float d = valid ? a.disps[dpos] : 0.f;
    const float Q = valid ? w.Qb[(int64_t)k * HW + pix] : 0.f;
    const float wt = valid ? w.Wb[(int64_t)k * HW + pix] : 0.f;
    const bool rob = a.rk.type != 0;
    // weights (u, v) and, with a robust kernel, targets (u, v) of edge s (below `end`)
    auto edge_in = [&](int s, int end, float4& o) {
        if (valid && s < end) {
            const int64_t eb = (int64_t)p.fr_edges[e0 + s] * 2 * HW + pix;
            o.x = a.weights[eb];
            o.y = a.weights[eb + HW];
            if (rob) {
                o.z = a.targets[eb];
                o.w = a.targets[eb + HW];
            }
        }
    };
    const double df = INTR ? w.x[PD * p.P] : 0.0;
    const int py = valid ? pix / p.W : 0, px = valid ? pix - py * p.W : 0;
    float X0 = ((float)px - cx) / fx;
    float Y0 = ((float)py - cy) / fy;
    if (CAM) rig_iproj<CAM>(Kf, (float)px, (float)py, X0, Y0);
    float dot = 0.f, Ef = 0.f;
    // edges in groups of kMaxDeg (one group unless the frame is wide)
    for (int g0 = 0; g0 < deg; g0 += kMaxDeg) {
        const int gn = min(kMaxDeg, deg - g0);
        if (g0 > 0) {
            __syncthreads();  // the previous group's set-up is read
            setup(g0, gn);
        }
        float4 in1 = make_float4(0.f, 0.f, 0.f, 0.f), in2 = in1;
        edge_in(g0, g0 + gn, in1);
        edge_in(g0 + 1, g0 + gn, in2);
        __syncthreads();
        if (!valid) continue;
        for (int s = 0; s < gn; s++) {
            const float4 in = in1;
            in1 = in2;
            edge_in(g0 + s + 2, g0 + gn, in2);
            PixTerms T;
            if (rob) {
                // the frame kernel's robust weights: residuals at the same linearisation point
                if (CAM) rig_pixel_terms<CAM, true>(sEP[s], X0, Y0, d, in.x, in.y, in.z, in.w, T);
                else pixel_terms<INTR, true, PD>(sEP[s], sEP[s] + 9, sEP[s][14], X0, Y0, d, fx, fy, cx, cy, in.x,
                                                 in.y, in.z, in.w, a.iscale, T);
                T.wu *= robust_weight(T.ru, a.rk);
                T.wv *= robust_weight(T.rv, a.rk);
            } else {
                if (CAM) rig_pixel_terms<CAM, false>(sEP[s], X0, Y0, d, in.x, in.y, 0.f, 0.f, T);
                else pixel_terms<INTR, false, PD>(sEP[s], sEP[s] + 9, sEP[s][14], X0, Y0, d, fx, fy, cx, cy, in.x,
                                                  in.y, 0.f, 0.f, a.iscale, T);
            }
            const float wdu = T.wu * act, wdv = T.wv * act;
            const float cu = act * wdu * T.Jzu, cv = act * wdv * T.Jzv;
            float s6 = 0.f;
#pragma unroll
            for (int n = 0; n < 6; n++) s6 += (cu * T.Ju[n] + cv * T.Jv[n]) * sU[s][n];
            dot += s6;
            if (INTR) Ef += wdu * T.Jzu * T.Jfu + wdv * T.Jzv * T.Jfv;
        }
    }
    if (!valid) return;
    float dz = Q * (wt - dot);
    if (INTR) dz = dz - Q * (Ef * act) * (float)df;
    a.dz[(int64_t)k * HW + pix] = dz;
    a.disps[dpos] = d + dz;
}

// ---------------------------------------------------------------------------
// 6. retraction T <- Exp(dx) T in FP64 (expSE3/retrSE3 semantics,
//    geom_kernels.cu:116-177, :974-1023), intrinsics update (:1038-1041)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void ba_retract_kernel(float* poses, float* intrinsics, const double* x, int t0,
                                                        int P, int intr, float iscale) {
    const int i = blockIdx.x * 64 + threadIdx.x;
    if (intr && i == 0) {
        const double df = x[6 * P] * (double)iscale;
        intrinsics[0] = (float)((double)intrinsics[0] + df);
        intrinsics[1] = (float)((double)intrinsics[1] + df);
    }
    if (i >= P) return;
    const double* xi = x + 6 * i;
    const double tau[3] = {xi[0], xi[1], xi[2]};
    const double phi[3] = {xi[3], xi[4], xi[5]};
    const double th2 = phi[0] * phi[0] + phi[1] * phi[1] + phi[2] * phi[2];
    const double th = sqrt(th2);
    double imag, real;
    if (th2 < 1e-8) {
        imag = 0.5 - th2 / 48.0 + th2 * th2 / 3840.0;
        real = 1.0 - th2 / 8.0 + th2 * th2 / 384.0;
    } else {
        imag = sin(0.5 * th) / th;
        real = cos(0.5 * th);
    }
    const double dq[4] = {imag * phi[0], imag * phi[1], imag * phi[2], real};
    double dt[3] = {tau[0], tau[1], tau[2]};
    if (th > 1e-4) {
        const double ca = (1 - cos(th)) / th2;
        const double cb = (th - sin(th)) / (th * th2);
        double c1[3] = {phi[1] * tau[2] - phi[2] * tau[1], phi[2] * tau[0] - phi[0] * tau[2],
                        phi[0] * tau[1] - phi[1] * tau[0]};
        double c2[3] = {phi[1] * c1[2] - phi[2] * c1[1], phi[2] * c1[0] - phi[0] * c1[2],
                        phi[0] * c1[1] - phi[1] * c1[0]};
        for (int r = 0; r < 3; r++) dt[r] += ca * c1[r] + cb * c2[r];
    }
    float* T = poses + 7 * (t0 + i);
    const double t[3] = {T[0], T[1], T[2]};
    const double q[4] = {T[3], T[4], T[5], T[6]};
    double q1[4];
    q1[0] = dq[3] * q[0] + dq[0] * q[3] + dq[1] * q[2] - dq[2] * q[1];
    q1[1] = dq[3] * q[1] + dq[1] * q[3] + dq[2] * q[0] - dq[0] * q[2];
    q1[2] = dq[3] * q[2] + dq[2] * q[3] + dq[0] * q[1] - dq[1] * q[0];
    q1[3] = dq[3] * q[3] - dq[0] * q[0] - dq[1] * q[1] - dq[2] * q[2];
    // actSO3(dq, t) with the uv form
    double uv[3] = {2.0 * (dq[1] * t[2] - dq[2] * t[1]), 2.0 * (dq[2] * t[0] - dq[0] * t[2]),
                    2.0 * (dq[0] * t[1] - dq[1] * t[0])};
    double t1[3];
    t1[0] = t[0] + dq[3] * uv[0] + (dq[1] * uv[2] - dq[2] * uv[1]) + dt[0];
    t1[1] = t[1] + dq[3] * uv[1] + (dq[2] * uv[0] - dq[0] * uv[2]) + dt[1];
    t1[2] = t[2] + dq[3] * uv[2] + (dq[0] * uv[1] - dq[1] * uv[0]) + dt[2];
    for (int r = 0; r < 3; r++) T[r] = (float)t1[r];
    for (int r = 0; r < 4; r++) T[3 + r] = (float)q1[r];
}

// Sim3 retraction T <- Exp(dx) T in FP64 with lietorch's Exp and product
// (sim3.h Exp and operator*, rxso3.h calcW; quaternions normalised as the
// group's constructors do), i.e. Sim3.retr of groups.py:157-160.
__global__ __launch_bounds__(64) void ba_retract_sim3_kernel(float* poses, float* intrinsics, const double* x, int t0,
                                                             int P, int intr, float iscale) {
    const int i = blockIdx.x * 64 + threadIdx.x;
    if (intr && i == 0) {
        const double df = x[7 * P] * (double)iscale;
        intrinsics[0] = (float)((double)intrinsics[0] + df);
        intrinsics[1] = (float)((double)intrinsics[1] + df);
    }
    if (i >= P) return;
    float* T = poses + 8 * (t0 + i);
    double Td[8];
    for (int r = 0; r < 8; r++) Td[r] = T[r];
    lie::Elem<double> X, Ex, Y;
    lie::Sim3T<double>::load(Td, X);
    lie::Sim3T<double>::exp(x + 7 * i, Ex);
    lie::elem_mul(Ex, X, Y);
    lie::Sim3T<double>::store(Y, Td);
    for (int r = 0; r < 8; r++) T[r] = (float)Td[r];
}

// ---------------------------------------------------------------------------
// 7. implicit adjoint of one Gauss-Newton iteration (no counterpart in the
//    reference, whose BA has no backward)
// ---------------------------------------------------------------------------
// Forward of one iteration, per depth pixel z and residual row (e, c):
//   pose terms  H_pp += pw J^T J,  b_p += pw r J        pw = stereo ? 0 : w
//   depth terms C += a^2 w Jz^2, b_z += a^2 w r Jz, E_row += a^2 w Jz J
//   (w = 0.001 weight on valid depth, a = depth_active of the frame)
//   S = H_pp - E Q E^T damped by ep + lm diag;  y = S^-1 (b_p - E Q b_z)
//   dz = Q (b_z - E~^T y)   (E~: E without rows of local pose index <= 0)
// Given gy = dL/dy and gz = dL/dz the adjoint is
//   ybar = gy - E~ (Q gz),  mu = S^-1 ybar  (the forward's Cholesky factor)
//   dL/db_p = mu,  dL/db_z = Q (gz - E^T mu)
//   dL/dS_raw = -mu y^T - lm diag(mu * y)
//   dL/dQ = gz (b_z - E~^T y) - (E^T mu) b_z + (E^T mu)(E^T y) + lm sum_i mu_i y_i E_i^2
//   dL/dE = Q [-mu b_z + (E^T y) mu + (E^T mu) y + 2 lm (mu*y) E] - mask * y Q gz
// and per row dL/dtarget = dL/dr, dL/dweight = 0.001 valid dL/dw with
//   dL/dr = pw J.mu + a^2 w Jz dL/db_z
//   dL/dw = ns [r J.mu - (J.mu)(J.y) - lm sum_i mu_i y_i J_i^2]
//         + a^2 [Jz J.dL/dE + Jz^2 dL/dC + r Jz dL/db_z]
// where J.v applies the row to pose vectors: v[jj] - Adj(Tij) v[kf].
// The lm terms take the pose rows of one frame as distinct apart from stereo
// edges, which the host checks when it builds the plan.

// Zero a double array.
__global__ void zero_f64_kernel(double* __restrict__ p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

struct AdjArgs {
    const float* gdx;  // [P][6]
    const float* gdz;  // [K][HW]
    float* gt;         // [E][2][HW]
    float* gw;         // [E][2][HW]
    double* esum;      // [E][6]
    double* ybar;      // [ld]
    double* mu;        // [ld]
    float lm;
};

// E row (j side) of edge s at one pixel: a^2 (wu Jzu Ju + wv Jzv Jv).
__device__ __forceinline__ void e_row(const PixTerms& T, float a2, float* er) {
#pragma unroll
    for (int n = 0; n < 6; n++) er[n] = a2 * (T.wu * T.Jzu * T.Ju[n] + T.wv * T.Jzv * T.Jv[n]);
}

// Pass 1: esum[e] = sum_p er_e(p) Q(p) gz(p).
__global__ __launch_bounds__(kThreads) void ba_adj_rhs_kernel(DevPlan p, RunArgs a, Work w, AdjArgs g) {
    __shared__ float sEP[kMaxDeg][16];
    __shared__ float sRed[4][6];
    const int k = blockIdx.y;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int kf = p.kx[k];
    const int e0 = p.fr_ptr[k];
    const int deg = p.fr_ptr[k + 1] - e0;
    const float act = a.depth_active[kf];
    const float fx = a.intrinsics[0], fy = a.intrinsics[1], cx = a.intrinsics[2], cy = a.intrinsics[3];
    const int HW = p.HW;
    if (tid < deg) {
        const int e = p.fr_edges[e0 + tid];
        float Rm[9], t[3], sc;
        edge_transform<6>(a.poses, p.ii[e], p.jj[e], Rm, t, sc);
        for (int i = 0; i < 9; i++) sEP[tid][i] = Rm[i];
        for (int i = 0; i < 3; i++) sEP[tid][9 + i] = t[i];
    }
    __syncthreads();
    const int pix = blockIdx.x * kThreads + tid;
    const bool valid = pix < HW;
    const int py = valid ? pix / p.W : 0, px = valid ? pix - py * p.W : 0;
    const float X0 = ((float)px - cx) / fx, Y0 = ((float)py - cy) / fy;
    const float d = valid ? a.disps[(int64_t)kf * HW + pix] : 0.f;
    const float tz = valid ? w.Qb[(int64_t)k * HW + pix] * g.gdz[(int64_t)k * HW + pix] : 0.f;
    for (int s = 0; s < deg; s++) {
        const int e = p.fr_edges[e0 + s];
        float v[6] = {0, 0, 0, 0, 0, 0};
        if (valid) {
            const int64_t eb = (int64_t)e * 2 * HW + pix;
            PixTerms T;
            pixel_terms<false, false>(sEP[s], sEP[s] + 9, 1.f, X0, Y0, d, fx, fy, cx, cy, a.weights[eb], a.weights[eb + HW],
                                      0.f, 0.f, 1.f, T);
            float er[6];
            e_row(T, act * act, er);
#pragma unroll
            for (int n = 0; n < 6; n++) v[n] = er[n] * tz;
        }
#pragma unroll
        for (int n = 0; n < 6; n++) {
            const float sum = wave_sum(v[n]);
            if (lane == 0) sRed[wave][n] = sum;
        }
        __syncthreads();
        if (tid < 6)
            atomicAdd(g.esum + (int64_t)e * 6 + tid,
                      (double)sRed[0][tid] + (double)sRed[1][tid] + (double)sRed[2][tid] + (double)sRed[3][tid]);
        __syncthreads();
    }
}

// Pass 2: ybar = gy - E~ t on the pose entries (quirk mask: local index >= 1),
// zeros on the padding.  One wave per pose walks the pose's right-hand-side
// list (the edges whose ii or jj is that pose): an edge reaching it as jj adds
// its esum, as ii its kf-side row -Adj^T esum.  Blocks past P clear the padding.
__global__ __launch_bounds__(64) void ba_adj_pose_kernel(DevPlan p, RunArgs a, AdjArgs g, int motion) {
    const int l = blockIdx.x;
    const int lane = threadIdx.x;
    if (l >= p.P) {
        const int i = 6 * p.P + (l - p.P) * 64 + lane;
        if (i < p.ld) g.ybar[i] = 0.0;
        return;
    }
    double c[6] = {0, 0, 0, 0, 0, 0};
    if (!motion && l >= 1) {
        const int beg = p.rhs_ptr[l], end = p.rhs_ptr[l + 1];
        for (int it = beg + lane; it < end; it += 64) {
            const int item = p.rhs_items[it];
            const int kind = item >> 27, e = item & ((1 << 27) - 1);
            if (kind != K_VI && kind != K_VJ) continue;
            const double* es = g.esum + (int64_t)e * 6;
            if (kind == K_VJ) {
                for (int n = 0; n < 6; n++) c[n] += es[n];
            } else {
                // kf side: E_kf = -Adj^T er, so -E~ t adds Adj^T esum
                float Rm[9], t[3], sc;
                edge_transform<6>(a.poses, p.ii[e], p.jj[e], Rm, t, sc);
                float ef[6], o[6];
                for (int q = 0; q < 6; q++) ef[q] = (float)es[q];
                neg_adjT(Rm, t, 1.f, ef, o);
                for (int n = 0; n < 6; n++) c[n] += (double)o[n];
            }
        }
    }
#pragma unroll
    for (int n = 0; n < 6; n++) c[n] = wave_sum_f64(c[n]);
    double mine = c[0];
#pragma unroll
    for (int n = 1; n < 6; n++) mine = lane == n ? c[n] : mine;
    if (lane < 6) g.ybar[6 * l + lane] = (double)g.gdx[6 * l + lane] - mine;
}

// x = (L L^T)^-1 rhs with the stored factor (X_k = L_kk^-T in LinvT, L_ik^T in
// the upper blocks of A); zeros when the factorisation failed.
//   forward   z_k = X_k^T (rhs_k - sum_{j<k} L_kj z_j)
//   backward  x_k = X_k (z_k - sum_{i>k} L_ik^T x_i)
__global__ __launch_bounds__(1024) void chol_solve_full_kernel(const double* __restrict__ A,
                                                               const double* __restrict__ LinvT, int ld, int nb,
                                                               const double* __restrict__ rhs,
                                                               double* __restrict__ out, const int* flag) {
    // z of the forward sweep, overwritten block by block with x by the backward
    // sweep (block kb reads its z before it stores its x, and reads x only of
    // the blocks after it)
    extern __shared__ double zs[];  // [ld]
    double* xs = zs;
    __shared__ double Xk[kCB][kCB + 1];
    __shared__ double t[kCB];
    __shared__ double part[16][kCB];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wv = tid >> 6;
    if (flag[0] != 0) {
        for (int i = tid; i < ld; i += 1024) out[i] = 0.0;
        return;
    }
    for (int kb = 0; kb < nb; kb++) {
        const int c0 = kb * kCB;
        const double* X = LinvT + (int64_t)kb * kCB * kCB;
        double xv[4];
#pragma unroll
        for (int q = 0; q < 4; q++) xv[q] = X[tid + q * 1024];
        // s[r] = sum_{q < c0} A[q][c0 + r] z[q]: lanes along r, waves split q
        double s = 0.0;
        for (int q = wv; q < c0; q += 16) s += A[(int64_t)q * ld + c0 + lane] * zs[q];
        part[wv][lane] = s;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int e = tid + q * 1024;
            Xk[e >> 6][e & 63] = xv[q];
        }
        __syncthreads();
        if (wv == 0) {
            double acc = 0.0;
#pragma unroll
            for (int u = 0; u < 16; u++) acc += part[u][lane];
            t[lane] = rhs[c0 + lane] - acc;
        }
        __syncthreads();
        // z_k[c] = sum_m X[m][c] t[m]: wave wv takes m in [4 wv, 4 wv + 4)
        double s2 = 0.0;
#pragma unroll
        for (int u = 0; u < 4; u++) s2 += Xk[4 * wv + u][lane] * t[4 * wv + u];
        part[wv][lane] = s2;
        __syncthreads();
        if (wv == 0) {
            double acc = 0.0;
#pragma unroll
            for (int u = 0; u < 16; u++) acc += part[u][lane];
            zs[c0 + lane] = acc;
        }
        __syncthreads();
    }
    for (int kb = nb - 1; kb >= 0; kb--) {
        const int c0 = kb * kCB;
        const double* X = LinvT + (int64_t)kb * kCB * kCB;
        double xv[4];
#pragma unroll
        for (int q = 0; q < 4; q++) xv[q] = X[tid + q * 1024];
        const int r0 = c0 + kCB, nr = (nb - kb - 1) * kCB;
        double s[4] = {0.0, 0.0, 0.0, 0.0};
        for (int q = lane; q < nr; q += 64) {
            const double xq = xs[r0 + q];
#pragma unroll
            for (int u = 0; u < 4; u++) s[u] += A[(int64_t)(c0 + wv + 16 * u) * ld + r0 + q] * xq;
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            s[u] = wave_sum_f64(s[u]);
        }
        if (lane == 0) {
#pragma unroll
            for (int u = 0; u < 4; u++) t[wv + 16 * u] = zs[c0 + wv + 16 * u] - s[u];
        }
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int e = tid + q * 1024;
            Xk[e >> 6][e & 63] = xv[q];
        }
        __syncthreads();
        double s2 = 0.0;
#pragma unroll
        for (int u = 0; u < 4; u++) s2 += Xk[lane][4 * wv + u] * t[4 * wv + u];
        part[wv][lane] = s2;
        __syncthreads();
        if (wv == 0) {
            double acc = 0.0;
#pragma unroll
            for (int u = 0; u < 16; u++) acc += part[u][lane];
            xs[c0 + lane] = acc;
        }
        __syncthreads();
    }
    for (int i = tid; i < ld; i += 1024) out[i] = xs[i];
}

// Pass 3: per-pixel aggregates and per-row gradients.
template <bool MOTION>
__global__ __launch_bounds__(kThreads) void ba_adj_grad_kernel(DevPlan p, RunArgs a, Work w, AdjArgs g) {
    __shared__ float sEP[kMaxDeg][16];  // R(9) t(3) stereo(12)
    __shared__ float sUm[kMaxDeg][6], sUy[kMaxDeg][6], sUq[kMaxDeg][6], sMj[kMaxDeg][6];
    __shared__ float sMk[6];
    const int k = blockIdx.y;
    const int tid = threadIdx.x;
    const int kf = p.kx[k];
    const int e0 = p.fr_ptr[k];
    const int deg = p.fr_ptr[k + 1] - e0;
    const int has_self = p.fr_self[k];
    const int lk = kf - p.t0;
    const float act = a.depth_active[kf];
    const float a2 = act * act;
    const float fx = a.intrinsics[0], fy = a.intrinsics[1], cx = a.intrinsics[2], cy = a.intrinsics[3];
    const int HW = p.HW;
    const float lm = g.lm;
    const double* y = w.x;
    const double* mu = g.mu;
    if (tid < deg) {
        const int e = p.fr_edges[e0 + tid];
        float Rm[9], t[3], sc;
        edge_transform<6>(a.poses, p.ii[e], p.jj[e], Rm, t, sc);
        for (int i = 0; i < 9; i++) sEP[tid][i] = Rm[i];
        for (int i = 0; i < 3; i++) sEP[tid][9 + i] = t[i];
        const bool stereo = p.ii[e] == p.jj[e];
        sEP[tid][12] = stereo ? 1.f : 0.f;
        // u(v) = v[jj] - Adj(Tij) v[kf] over the window, and with the back-substitution mask
        const int lj = p.jj[e] - p.t0;
        const bool jv = lj >= 0 && lj < p.P, kv = has_self && lk >= 0 && lk < p.P;
        double um[6] = {0, 0, 0, 0, 0, 0}, uy[6] = {0, 0, 0, 0, 0, 0}, uq[6] = {0, 0, 0, 0, 0, 0};
        for (int n = 0; n < 6; n++) {
            if (jv) {
                um[n] = mu[6 * lj + n];
                uy[n] = y[6 * lj + n];
                if (lj >= 1) uq[n] = y[6 * lj + n];
            }
            sMj[tid][n] = (jv && !stereo) ? (float)(mu[6 * lj + n] * y[6 * lj + n]) : 0.f;
        }
        if (kv) {
            const double* vm = mu + 6 * lk;
            const double* vy = y + 6 * lk;
            double am[6], ay[6];
            for (int r = 0; r < 3; r++) {
                const double rvm = Rm[3 * r] * vm[0] + Rm[3 * r + 1] * vm[1] + Rm[3 * r + 2] * vm[2];
                const double rwm = Rm[3 * r] * vm[3] + Rm[3 * r + 1] * vm[4] + Rm[3 * r + 2] * vm[5];
                const double rvy = Rm[3 * r] * vy[0] + Rm[3 * r + 1] * vy[1] + Rm[3 * r + 2] * vy[2];
                const double rwy = Rm[3 * r] * vy[3] + Rm[3 * r + 1] * vy[4] + Rm[3 * r + 2] * vy[5];
                am[r] = rvm; am[3 + r] = rwm;
                ay[r] = rvy; ay[3 + r] = rwy;
            }
            // Adj v = [R v_tau + t x (R v_phi), R v_phi]
            double adm[6] = {am[0] + (t[1] * am[5] - t[2] * am[4]), am[1] + (t[2] * am[3] - t[0] * am[5]),
                             am[2] + (t[0] * am[4] - t[1] * am[3]), am[3], am[4], am[5]};
            double ady[6] = {ay[0] + (t[1] * ay[5] - t[2] * ay[4]), ay[1] + (t[2] * ay[3] - t[0] * ay[5]),
                             ay[2] + (t[0] * ay[4] - t[1] * ay[3]), ay[3], ay[4], ay[5]};
            for (int n = 0; n < 6; n++) {
                um[n] -= adm[n];
                uy[n] -= ady[n];
                if (lk >= 1) uq[n] -= ady[n];
            }
        }
        for (int n = 0; n < 6; n++) {
            sUm[tid][n] = (float)um[n];
            sUy[tid][n] = (float)uy[n];
            sUq[tid][n] = (float)uq[n];
        }
    }
    if (tid < 6) {
        const bool kv = has_self && lk >= 0 && lk < p.P;
        sMk[tid] = kv ? (float)(mu[6 * lk + tid] * y[6 * lk + tid]) : 0.f;
    }
    __syncthreads();
    const bool kv = has_self && lk >= 0 && lk < p.P;
    const int pix = blockIdx.x * kThreads + tid;
    if (pix >= HW) return;
    const int py = pix / p.W, px = pix - py * p.W;
    const float X0 = ((float)px - cx) / fx, Y0 = ((float)py - cy) / fy;
    const float d = a.disps[(int64_t)kf * HW + pix];
    float Q = 0.f, bz = 0.f, lz = 0.f, tz = 0.f, bbz = 0.f, Cbar = 0.f, Ey = 0.f, Emu = 0.f;
    float Ekf[6] = {0, 0, 0, 0, 0, 0};
    if (!MOTION) {
        Q = w.Qb[(int64_t)k * HW + pix];
        bz = w.Wb[(int64_t)k * HW + pix];
        lz = g.gdz[(int64_t)k * HW + pix];
        tz = Q * lz;
        float Eyq = 0.f, lmq = 0.f;
        for (int s = 0; s < deg; s++) {
            const int e = p.fr_edges[e0 + s];
            const int64_t eb = (int64_t)e * 2 * HW + pix;
            PixTerms T;
            pixel_terms<false, false>(sEP[s], sEP[s] + 9, 1.f, X0, Y0, d, fx, fy, cx, cy, a.weights[eb], a.weights[eb + HW],
                                      0.f, 0.f, 1.f, T);
            float er[6];
            e_row(T, a2, er);
            float o[6];
            neg_adjT(sEP[s], sEP[s] + 9, 1.f, er, o);
            const bool stereo = sEP[s][12] != 0.f;
#pragma unroll
            for (int n = 0; n < 6; n++) {
                Ey += er[n] * sUy[s][n];
                Emu += er[n] * sUm[s][n];
                Eyq += er[n] * sUq[s][n];
                lmq += sMj[s][n] * er[n] * er[n];
                Ekf[n] += (kv ? o[n] : 0.f) + (stereo && kv ? er[n] : 0.f);
            }
        }
#pragma unroll
        for (int n = 0; n < 6; n++) lmq += sMk[n] * Ekf[n] * Ekf[n];
        bbz = Q * (lz - Emu);
        const float Qbar = lz * (bz - Eyq) - Emu * bz + Emu * Ey + lm * lmq;
        Cbar = -Q * Q * Qbar;
    }
    for (int s = 0; s < deg; s++) {
        const int e = p.fr_edges[e0 + s];
        const int64_t eb = (int64_t)e * 2 * HW + pix;
        const float* Rm = sEP[s];
        const float* tt = sEP[s] + 9;
        const bool stereo = sEP[s][12] != 0.f;
        PixTerms T;
        pixel_terms<false, true>(Rm, tt, 1.f, X0, Y0, d, fx, fy, cx, cy, a.weights[eb], a.weights[eb + HW], a.targets[eb],
                                 a.targets[eb + HW], 1.f, T);
        float er[6];
        e_row(T, a2, er);
#pragma unroll
        for (int c = 0; c < 2; c++) {
            const float* J = c == 0 ? T.Ju : T.Jv;
            const float Jz = c == 0 ? T.Jzu : T.Jzv;
            const float wr = c == 0 ? T.wu : T.wv;
            const float r = c == 0 ? T.ru : T.rv;
            float Ji[6];
            neg_adjT(Rm, tt, 1.f, J, Ji);
            float Jm = 0.f, Jy = 0.f, Jq = 0.f, lmp = 0.f, lmr = 0.f;
#pragma unroll
            for (int n = 0; n < 6; n++) {
                Jm += J[n] * sUm[s][n];
                Jy += J[n] * sUy[s][n];
                Jq += J[n] * sUq[s][n];
                const float jk = (kv ? Ji[n] : 0.f) + (stereo && kv ? J[n] : 0.f);  // row at pose kf
                lmp += sMj[s][n] * J[n] * J[n] + sMk[n] * (kv ? Ji[n] * Ji[n] : 0.f);
                lmr += J[n] * sMj[s][n] * er[n] + jk * sMk[n] * Ekf[n];
            }
            const float pw = stereo ? 0.f : wr;
            float rbar = pw * Jm;
            float wbar = stereo ? 0.f : (r * Jm - Jm * Jy - lm * lmp);
            if (!MOTION) {
                rbar += a2 * wr * Jz * bbz;
                wbar += a2 * (Jz * Q * (Jm * Ey + Jy * Emu + 2.f * lm * lmr) - Jz * Jm * Q * bz - Jz * Jq * tz +
                              Jz * Jz * Cbar + r * Jz * bbz);
            }
            // w = 0.001 weight where the depth test passes (pixel_terms)
            g.gt[eb + c * HW] = rbar;
            g.gw[eb + c * HW] = T.ok ? 0.001f * wbar : 0.f;
        }
    }
}

}  // namespace ba
}  // namespace vh

// ---------------------------------------------------------------------------
// host: plan construction and the run loop
// ---------------------------------------------------------------------------
using namespace vh;
using namespace vh::ba;

struct vh_ba_plan {
    DevPlan d;
    std::vector<int64_t> kx;
    void* dev = nullptr;  // one allocation holding every index table
    int deg_max = 0;
    size_t ws_bytes = 0;
    // workspace offsets
    size_t o_edge, o_gram, o_fen, o_Q, o_W, o_ed, o_gd, o_fenk, o_A, o_b, o_x, o_y, o_Li, o_flag;
    size_t o_esum, o_ybar, o_mu;  // adjoint
    size_t o_wide;                // E rows of the wide frames
    int gram_blocks = 0;          // Gram blocks over all depth frames
    int TP7 = 64;                 // frame kernel pixel tile for Sim3 runs
    // plan variants: the system size depends on whether intrinsics are optimised
    int P;
    // every non-stereo edge of a depth frame reaches a different pose (the
    // adjoint's damping terms rely on it)
    bool adj_distinct = true;
};

namespace {

size_t align256(size_t v) { return (v + 255) & ~size_t(255); }

// Workspace of vh_ba_run_static: one array per GPU, held until the process
// ends.  Host code never uses its name as an address: the device address of
// each GPU's copy comes from hipGetSymbolAddress on that GPU's first call.
constexpr size_t kStaticScratchBytes = VH_BA_STATIC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;
__device__ __align__(256) char scratch[kStaticScratchBytes];
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of the current GPU's scratch array (nullptr and the error
// recorded on failure).
void* static_scratch() {
    int dev = -1;
    hipError_t err = hipGetDevice(&dev);
    if (err != hipSuccess) {
        hip_fail(err, "ba_run_static hipGetDevice");
        return nullptr;
    }
    if (dev < 0 || dev >= kMaxDevices) {
        fail(VH_ERR_UNSUPPORTED, "ba_run_static: device %d (at most %d devices)", dev, kMaxDevices);
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        err = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (err != hipSuccess || !ptr) {
            hip_fail(err != hipSuccess ? err : hipErrorInvalidSymbol, "ba_run_static hipGetSymbolAddress");
            return nullptr;
        }
        g_scratch_ptr[dev] = ptr;
    }
    return g_scratch_ptr[dev];
}

}  // namespace

__global__ __launch_bounds__(256) void robust_weight_kernel(RobustK k, const float* __restrict__ x,
                                                            float* __restrict__ w, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        w[i] = robust_weight(x[i], k);
}

static vh_ba_plan* plan_build(const int64_t* ii_host, const int64_t* jj_host, const int64_t* qi_host,
                              const int64_t* qj_host, int V, int E, int N, int H, int W, int t0, int t1,
                              int pixel_tile);

extern "C" {

vh_ba_plan* vh_ba_plan_create(const int64_t* ii_host, const int64_t* jj_host, int E, int N, int H, int W, int t0,
                              int t1) {
    return vh_ba_plan_create_tiled(ii_host, jj_host, E, N, H, W, t0, t1, 0);
}

vh_ba_plan* vh_ba_plan_create_tiled(const int64_t* ii_host, const int64_t* jj_host, int E, int N, int H, int W,
                                    int t0, int t1, int pixel_tile) {
    return plan_build(ii_host, jj_host, nullptr, nullptr, 0, E, N, H, W, t0, t1, pixel_tile);
}

vh_ba_plan* vh_ba_plan_create_rig(const int64_t* pi_host, const int64_t* pj_host, const int64_t* qi_host,
                                  const int64_t* qj_host, int E, int N, int V, int H, int W, int t0, int t1,
                                  int pixel_tile) {
    if (V < 1) {
        fail(VH_ERR_ARG, "ba_plan_create_rig: %d views", V);
        return nullptr;
    }
    if (E > 0 && (!pi_host || !pj_host || !qi_host || !qj_host)) {
        fail(VH_ERR_ARG, "ba_plan_create_rig: null index array");
        return nullptr;
    }
    return plan_build(pi_host, pj_host, qi_host, qj_host, V, E, N, H, W, t0, t1, pixel_tile);
}

}  // extern "C"

// Plan of a single-view graph (V = 0: depth frame id = pose id = ii) or of a
// rig graph (V >= 1: edge e from view qi of pose ii to view qj of pose jj;
// depth frame id ii V + qi over the flattened [N V, H, W] disparities).
static vh_ba_plan* plan_build(const int64_t* ii_host, const int64_t* jj_host, const int64_t* qi_host,
                              const int64_t* qj_host, int V, int E, int N, int H, int W, int t0, int t1,
                              int pixel_tile) {
    if (pixel_tile != 0 && pixel_tile != 64 && pixel_tile != 128 && pixel_tile != 256) {
        fail(VH_ERR_ARG, "ba_plan_create: pixel_tile %d is not 0, 64, 128 or 256", pixel_tile);
        return nullptr;
    }
    if (E < 0 || N <= 0 || H <= 0 || W <= 0 || t0 < 0 || t1 <= t0 || t1 > N || (int64_t)N * std::max(V, 1) >= (1 << 27)) {
        fail(VH_ERR_ARG, "ba_plan_create: bad sizes (E=%d N=%d V=%d H=%d W=%d t0=%d t1=%d)", E, N, V, H, W, t0, t1);
        return nullptr;
    }
    for (int e = 0; e < E; e++) {
        if (ii_host[e] < 0 || ii_host[e] >= N || jj_host[e] < 0 || jj_host[e] >= N) {
            fail(VH_ERR_ARG, "ba_plan_create: edge %d (%lld -> %lld) outside [0, %d)", e, (long long)ii_host[e],
                 (long long)jj_host[e], N);
            return nullptr;
        }
        if (V > 0) {
            if (qi_host[e] < 0 || qi_host[e] >= V || qj_host[e] < 0 || qj_host[e] >= V) {
                fail(VH_ERR_ARG, "ba_plan_create_rig: edge %d views (%lld -> %lld) outside [0, %d)", e,
                     (long long)qi_host[e], (long long)qj_host[e], V);
                return nullptr;
            }
            if (ii_host[e] == jj_host[e] && qi_host[e] == qj_host[e]) {
                fail(VH_ERR_ARG, "ba_plan_create_rig: edge %d maps view %lld of frame %lld onto itself", e,
                     (long long)qi_host[e], (long long)ii_host[e]);
                return nullptr;
            }
        }
    }
    vh_ba_plan* pl = new (std::nothrow) vh_ba_plan();
    if (!pl) {
        fail(VH_ERR_ARG, "ba_plan_create: out of host memory");
        return nullptr;
    }
    const int P = t1 - t0;
    const int HW = H * W;
    pl->P = P;
    if (P > kMaxPoses) {
        delete pl;
        fail(VH_ERR_UNSUPPORTED, "ba_plan_create: %d poses in the window (limit %d)", P, kMaxPoses);
        return nullptr;
    }
    // kx = unique(cat(arange(t0,t1), ii)), sorted (geom_kernels.cu:1648-1652);
    // rig plans: every view of the window's frames and the edges' source
    // depth frames ii V + qi
    const int VV = std::max(V, 1);
    std::vector<int64_t> dsrc(E);
    for (int e = 0; e < E; e++) dsrc[e] = V > 0 ? ii_host[e] * V + qi_host[e] : ii_host[e];
    std::vector<int64_t> cand;
    for (int64_t t = (int64_t)t0 * VV; t < (int64_t)t1 * VV; t++) cand.push_back(t);
    for (int e = 0; e < E; e++) cand.push_back(dsrc[e]);
    std::sort(cand.begin(), cand.end());
    cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
    pl->kx = cand;
    const int K = (int)cand.size();
    std::vector<int> slot((size_t)N * VV, -1);
    for (int k = 0; k < K; k++) slot[cand[k]] = k;
    auto kpose_of = [&](int k) { return (int)(cand[k] / VV); };
    // per-frame edge lists in edge order
    std::vector<int> fr_ptr(K + 1, 0), fr_edges(E), fr_self(K);
    for (int e = 0; e < E; e++) fr_ptr[slot[dsrc[e]] + 1]++;
    for (int k = 0; k < K; k++) fr_ptr[k + 1] += fr_ptr[k];
    {
        std::vector<int> fill(fr_ptr.begin(), fr_ptr.end() - 1);
        for (int e = 0; e < E; e++) fr_edges[fill[slot[dsrc[e]]]++] = e;
    }
    int deg_max = 0, Rext_max = 1, wide_rows = 0;
    std::vector<int> wide;
    for (int k = 0; k < K; k++) {
        std::vector<int64_t> js;
        for (int q = fr_ptr[k]; q < fr_ptr[k + 1]; q++) {
            const int e = fr_edges[q];
            if (ii_host[e] != jj_host[e]) js.push_back(jj_host[e]);
        }
        std::sort(js.begin(), js.end());
        if (std::adjacent_find(js.begin(), js.end()) != js.end()) pl->adj_distinct = false;
        fr_self[k] = (kpose_of(k) >= t0 && kpose_of(k) < t1) ? 1 : 0;
        const int deg = fr_ptr[k + 1] - fr_ptr[k];
        deg_max = std::max(deg_max, deg);
        if (deg > kMaxDeg) {
            // stored rows: self block (sized for 7-dof poses), 6 per edge, the pseudo pair and the zero row
            wide.push_back(k);
            wide_rows = std::max(wide_rows, fr_self[k] * 7 + 6 * deg + 3);
        } else {
            Rext_max = std::max(Rext_max, fr_self[k] + deg + 1);
        }
    }
    // Gram storage, in blocks of PD x PD: frame k holds Rext_k (Rext_k+1)/2 blocks
    std::vector<int> gram_off(K + 1, 0), pseudo_off(K);
    for (int k = 0; k < K; k++) {
        const int Rext = fr_self[k] + (fr_ptr[k + 1] - fr_ptr[k]) + 1;
        gram_off[k + 1] = gram_off[k] + Rext * (Rext + 1) / 2;
    }
    auto gblock = [&](int k, int ra, int rb) {  // ra <= rb
        const int Rext = fr_self[k] + (fr_ptr[k + 1] - fr_ptr[k]) + 1;
        const int b = ra * Rext - ra * (ra - 1) / 2 + (rb - ra);
        return gram_off[k] + b;
    };
    for (int k = 0; k < K; k++) {
        const int Rext = fr_self[k] + (fr_ptr[k + 1] - fr_ptr[k]) + 1;
        pseudo_off[k] = gblock(k, Rext - 1, Rext - 1);
    }
    // tiles: TP pixels per sub-tile, sized so the E rows fit the LDS budget
    // unless the caller chose TP (then it must fit the workgroup's LDS)
    // (sized for SE3 blocks; a Sim3 run checks its 7-row blocks against kLdsMax)
    int TP = 256;
    while (TP > 64 && frame_lds_bytes(Rext_max, TP, 6) > (size_t)kLdsBudget) TP /= 2;
    if (pixel_tile != 0) {
        if (frame_lds_bytes(Rext_max, pixel_tile, 6) > (size_t)kLdsMax) {
            delete pl;
            fail(VH_ERR_UNSUPPORTED, "ba_plan_create: %d-pixel tiles need %zu B of LDS for %d E rows (limit %d)",
                 pixel_tile, frame_lds_bytes(Rext_max, pixel_tile, 6), Rext_max * 6, kLdsMax);
            return nullptr;
        }
        TP = pixel_tile;
    }
    // Sim3 frames store one more self row: their pixel tile is chosen on its own
    int TP7 = 256;
    while (TP7 > 64 && frame_lds_bytes(Rext_max, TP7, 7) > (size_t)kLdsBudget) TP7 /= 2;
    if (pixel_tile != 0) TP7 = pixel_tile;
    pl->TP7 = TP7;
    const int SUB = std::max(1, 256 / TP);  // 256-pixel tiles
    const int T = (HW + TP * SUB - 1) / (TP * SUB);
    // lower pose blocks and their contribution lists (fixed order)
    std::vector<std::vector<int>> items((size_t)P * (P + 1) / 2);
    auto lidx = [&](int a, int b) { return a * (a + 1) / 2 + b; };  // a >= b
    for (int e = 0; e < E; e++) {
        const int li = (int)ii_host[e] - t0, lj = (int)jj_host[e] - t0;
        if (ii_host[e] == jj_host[e]) continue;  // stereo: no pose terms
        const bool iin = li >= 0 && li < P, jin = lj >= 0 && lj < P;
        if (iin) items[lidx(li, li)].push_back(pk(K_HII, e));
        if (jin) items[lidx(lj, lj)].push_back(pk(K_HJJ, e));
        if (iin && jin) {
            if (li > lj) items[lidx(li, lj)].push_back(pk(K_HIJ, e));
            else items[lidx(lj, li)].push_back(pk(K_HIJT, e));
        }
    }
    std::vector<std::vector<int>> ritems(P);
    for (int k = 0; k < K; k++) {
        const int deg = fr_ptr[k + 1] - fr_ptr[k];
        const int R = fr_self[k] + deg;
        std::vector<int> rpose(R);
        for (int r = 0; r < R; r++) {
            if (fr_self[k] && r == 0) rpose[r] = kpose_of(k) - t0;
            else rpose[r] = (int)jj_host[fr_edges[fr_ptr[k] + r - fr_self[k]]] - t0;
        }
        for (int ra = 0; ra < R; ra++) {
            const int pa = rpose[ra];
            if (pa < 0 || pa >= P) continue;
            for (int rb = ra; rb < R; rb++) {
                const int pb = rpose[rb];
                if (pb < 0 || pb >= P) continue;
                const int g = gblock(k, ra, rb);
                if (pa > pb) items[lidx(pa, pb)].push_back(pk(K_G, g));
                else if (pa < pb) items[lidx(pb, pa)].push_back(pk(K_GT, g));
                else {
                    items[lidx(pa, pa)].push_back(pk(K_G, g));
                    if (ra != rb) items[lidx(pa, pa)].push_back(pk(K_GT, g));
                }
            }
        }
    }
    // right-hand side / intrinsics column per pose, ordered by edge then by row
    for (int e = 0; e < E; e++) {
        const int li = (int)ii_host[e] - t0, lj = (int)jj_host[e] - t0;
        if (li >= 0 && li < P) ritems[li].push_back(pk(K_VI, e));
        if (lj >= 0 && lj < P) ritems[lj].push_back(pk(K_VJ, e));
    }
    for (int k = 0; k < K; k++) {
        const int deg = fr_ptr[k + 1] - fr_ptr[k];
        const int R = fr_self[k] + deg;
        for (int r = 0; r < R; r++) {
            int pr;
            if (fr_self[k] && r == 0) pr = kpose_of(k) - t0;
            else pr = (int)jj_host[fr_edges[fr_ptr[k] + r - fr_self[k]]] - t0;
            if (pr < 0 || pr >= P) continue;
            ritems[pr].push_back(pk(K_ROW, gblock(k, r, R)));
        }
    }
    const int nlb = P * (P + 1) / 2;
    std::vector<int> blk_ab(2 * nlb), blk_ptr(nlb + 1, 0), blk_items;
    for (int a = 0, q = 0; a < P; a++)
        for (int b = 0; b <= a; b++, q++) {
            blk_ab[2 * q] = a;
            blk_ab[2 * q + 1] = b;
        }
    for (int q = 0; q < nlb; q++) {
        blk_ptr[q + 1] = blk_ptr[q] + (int)items[q].size();
        blk_items.insert(blk_items.end(), items[q].begin(), items[q].end());
    }
    std::vector<int> rhs_ptr(P + 1, 0), rhs_items;
    for (int a = 0; a < P; a++) {
        rhs_ptr[a + 1] = rhs_ptr[a] + (int)ritems[a].size();
        rhs_items.insert(rhs_items.end(), ritems[a].begin(), ritems[a].end());
    }
    std::vector<int> kx32(K), ii32(E), jj32(E), qi32(E, 0), qj32(E, 0), kpose32(K);
    for (int k = 0; k < K; k++) {
        kx32[k] = (int)cand[k];
        kpose32[k] = kpose_of(k);
    }
    for (int e = 0; e < E; e++) {
        ii32[e] = (int)ii_host[e];
        jj32[e] = (int)jj_host[e];
        if (V > 0) {
            qi32[e] = (int)qi_host[e];
            qj32[e] = (int)qj_host[e];
        }
    }
    // upload all tables in one allocation
    std::vector<std::pair<const std::vector<int>*, const int**>> tabs = {
        {&kx32, &pl->d.kx}, {&fr_ptr, &pl->d.fr_ptr}, {&fr_edges, &pl->d.fr_edges}, {&fr_self, &pl->d.fr_self},
        {&gram_off, &pl->d.gram_off}, {&ii32, &pl->d.ii}, {&jj32, &pl->d.jj}, {&blk_ab, &pl->d.blk_ab},
        {&blk_ptr, &pl->d.blk_ptr}, {&blk_items, &pl->d.blk_items}, {&rhs_ptr, &pl->d.rhs_ptr},
        {&rhs_items, &pl->d.rhs_items}, {&pseudo_off, &pl->d.pseudo_off}, {&wide, &pl->d.wide},
        {&qi32, &pl->d.qi}, {&qj32, &pl->d.qj}, {&kpose32, &pl->d.kpose}};
    size_t total = 0;
    std::vector<size_t> offs;
    for (auto& t : tabs) {
        offs.push_back(total);
        total += align256(std::max<size_t>(1, t.first->size()) * sizeof(int));
    }
    std::vector<char> host(total, 0);
    for (size_t i = 0; i < tabs.size(); i++)
        if (!tabs[i].first->empty()) memcpy(host.data() + offs[i], tabs[i].first->data(), tabs[i].first->size() * 4);
    hipError_t err = hipMalloc(&pl->dev, total);
    if (err != hipSuccess) {
        hip_fail(err, "ba_plan_create hipMalloc");
        delete pl;
        return nullptr;
    }
    err = hipMemcpy(pl->dev, host.data(), total, hipMemcpyHostToDevice);
    if (err != hipSuccess) {
        hip_fail(err, "ba_plan_create hipMemcpy");
        (void)hipFree(pl->dev);
        delete pl;
        return nullptr;
    }
    for (size_t i = 0; i < tabs.size(); i++) *tabs[i].second = (const int*)((char*)pl->dev + offs[i]);

    DevPlan& d = pl->d;
    d.E = E; d.N = N; d.H = H; d.W = W; d.HW = HW; d.t0 = t0; d.t1 = t1; d.P = P; d.K = K;
    d.TP = TP; d.SUB = SUB; d.T = T; d.Rext_max = Rext_max;
    d.n_wide = (int)wide.size();
    d.wide_rows = wide_rows;
    d.V = V;
    d.nlb = nlb;
    pl->gram_blocks = gram_off[K];
    // SE3 sizes; ba_run_impl sets them for the run's pose group
    d.pd = 6;
    d.GT = pl->gram_blocks * 36;
    d.ld = (6 * P + 1 + kCB - 1) / kCB * kCB;
    d.nb = d.ld / kCB;
    d.n = 6 * P;
    d.intr = 0;
    pl->deg_max = deg_max;
    // workspace layout, sized for the larger (Sim3) blocks so one workspace serves both groups
    const int GTmax = std::max(pl->gram_blocks * 49, 1);
    const int ldmax = (7 * P + 1 + kCB - 1) / kCB * kCB;
    size_t o = 0;
    pl->o_edge = o; o += align256((size_t)T * E * kNV * 4);
    pl->o_gram = o; o += align256((size_t)T * GTmax * 4);
    pl->o_fen = o; o += align256((size_t)T * K * 4);
    pl->o_Q = o; o += align256((size_t)K * HW * 4);
    pl->o_W = o; o += align256((size_t)K * HW * 4);
    pl->o_ed = o; o += align256((size_t)std::max(E, 1) * kEDSMax * 8);
    pl->o_gd = o; o += align256((size_t)GTmax * 8);
    pl->o_fenk = o; o += align256((size_t)K * 8);
    pl->o_A = o; o += align256((size_t)ldmax * ldmax * 8);
    pl->o_b = o; o += align256((size_t)ldmax * 8);
    pl->o_x = o; o += align256((size_t)ldmax * 8);
    pl->o_Li = o; o += align256((size_t)ldmax * kCB * 8);
    pl->o_y = o; o += align256((size_t)ldmax * 8);
    pl->o_flag = o; o += 256;
    pl->o_esum = o; o += align256((size_t)std::max(E, 1) * 6 * 8);
    pl->o_ybar = o; o += align256((size_t)ldmax * 8);
    pl->o_mu = o; o += align256((size_t)ldmax * 8);
    pl->o_wide = o; o += align256((size_t)wide.size() * T * wide_rows * kThreads * 4);
    pl->ws_bytes = o;
    return pl;
}

extern "C" {

void vh_ba_plan_destroy(vh_ba_plan* plan) {
    if (!plan) return;
    if (plan->dev) (void)hipFree(plan->dev);
    delete plan;
}

int vh_ba_plan_num_depth_frames(const vh_ba_plan* plan) { return plan ? (int)plan->kx.size() : -1; }

int vh_ba_plan_depth_frames(const vh_ba_plan* plan, int64_t* kx_out) {
    if (!plan || !kx_out) return fail(VH_ERR_ARG, "ba_plan_depth_frames: null argument");
    for (size_t i = 0; i < plan->kx.size(); i++) kx_out[i] = plan->kx[i];
    return VH_OK;
}

size_t vh_ba_workspace_size(const vh_ba_plan* plan) { return plan ? plan->ws_bytes : 0; }

int vh_ba_plan_pixel_tile(const vh_ba_plan* plan) { return plan ? plan->d.TP : -1; }

}  // extern "C"

// The wide-frame kernel over the plan's wide frames, in the frame kernel's variant.
template <int PD, int CAM = 0>
static void launch_wide(const DevPlan& d, const RunArgs& a, const Work& w, bool motion, bool intr, bool robust,
                        hipStream_t s) {
    const dim3 g(d.T, d.n_wide), b(kThreads);
    if constexpr (CAM != 0) {
        // rig plans: fixed intrinsics
        if (robust) {
            if (motion) hipLaunchKernelGGL((ba_frame_wide_kernel<false, true, true, 6, CAM>), g, b, 0, s, d, a, w);
            else hipLaunchKernelGGL((ba_frame_wide_kernel<false, false, true, 6, CAM>), g, b, 0, s, d, a, w);
        } else {
            if (motion) hipLaunchKernelGGL((ba_frame_wide_kernel<false, true, false, 6, CAM>), g, b, 0, s, d, a, w);
            else hipLaunchKernelGGL((ba_frame_wide_kernel<false, false, false, 6, CAM>), g, b, 0, s, d, a, w);
        }
        return;
    }
#define VH_WIDE(I, M, R) hipLaunchKernelGGL((ba_frame_wide_kernel<I, M, R, PD>), g, b, 0, s, d, a, w)
    if (robust) {
        if (motion) {
            if (intr) VH_WIDE(true, true, true);
            else VH_WIDE(false, true, true);
        } else {
            if (intr) VH_WIDE(true, false, true);
            else VH_WIDE(false, false, true);
        }
    } else if (motion) {
        if (intr) VH_WIDE(true, true, false);
        else VH_WIDE(false, true, false);
    } else {
        if (intr) VH_WIDE(true, false, false);
        else VH_WIDE(false, false, false);
    }
#undef VH_WIDE
}

// The frame kernel of a rig plan (fixed intrinsics) in its variant.
template <int CAM>
static void launch_frame_rig(const DevPlan& d, const RunArgs& a, const Work& w, bool motion, bool robust, size_t lds,
                             hipStream_t s) {
    const dim3 g(d.T, d.K), b(kThreads);
    if (robust) {
        if (motion) hipLaunchKernelGGL((ba_frame_kernel<false, true, true, 6, CAM>), g, b, lds, s, d, a, w);
        else hipLaunchKernelGGL((ba_frame_kernel<false, false, true, 6, CAM>), g, b, lds, s, d, a, w);
    } else {
        if (motion) hipLaunchKernelGGL((ba_frame_kernel<false, true, false, 6, CAM>), g, b, lds, s, d, a, w);
        else hipLaunchKernelGGL((ba_frame_kernel<false, false, false, 6, CAM>), g, b, lds, s, d, a, w);
    }
}

// The launches of `iterations` Gauss-Newton iterations for pose blocks of PD
// (CAM 0: a single-view plan; 1 / 2: a rig plan with the pinhole / Mei model).
template <int PD, int CAM = 0>
static int ba_iterations(const DevPlan& d, RunArgs a, const Work& w, int iterations, bool motion, bool intr,
                         const RobustK& rk, float* poses, float* intrinsics, float* dx, float intrinsics_scale,
                         hipStream_t s) {
    const size_t lds = frame_lds_bytes(d.Rext_max, d.TP, PD);
    const dim3 fgrid(d.T, d.K);
    const int npad = d.ld - d.n;
    for (int itr = 0; itr < iterations; itr++) {
        a.itr = itr;
        if constexpr (CAM != 0) {
          if (d.K > 0 && (d.E > 0 || !motion)) {
            launch_frame_rig<CAM>(d, a, w, motion, rk.type != 0, lds, s);
            VH_LAUNCH_CHECK("ba_frame_kernel");
            if (d.n_wide > 0) {
                launch_wide<PD, CAM>(d, a, w, motion, false, rk.type != 0, s);
                VH_LAUNCH_CHECK("ba_frame_wide_kernel");
            }
            hipLaunchKernelGGL((ba_finalize_kernel<PD, CAM>), dim3(d.E + (motion ? 0 : d.K * kGramSplit)),
                               dim3(kThreads), 0, s, d, a, w, motion ? 1 : 0);
            VH_LAUNCH_CHECK("ba_finalize_kernel");
            if (!motion) {
                hipLaunchKernelGGL(ba_gram_rig_kernel, dim3(d.K), dim3(64), 0, s, d, a, w);
                VH_LAUNCH_CHECK("ba_gram_rig_kernel");
            }
          }  // motion only without edges: nothing to accumulate
        } else if (d.E > 0 && d.K > 0) {
            if (rk.type) {
                // robust weights: a separate instantiation, so plain least squares pays nothing
                if (motion) {
                    if (intr) hipLaunchKernelGGL((ba_frame_kernel<true, true, true, PD>), fgrid, dim3(kThreads), lds, s, d, a, w);
                    else hipLaunchKernelGGL((ba_frame_kernel<false, true, true, PD>), fgrid, dim3(kThreads), lds, s, d, a, w);
                } else {
                    if (intr) hipLaunchKernelGGL((ba_frame_kernel<true, false, true, PD>), fgrid, dim3(kThreads), lds, s, d, a, w);
                    else hipLaunchKernelGGL((ba_frame_kernel<false, false, true, PD>), fgrid, dim3(kThreads), lds, s, d, a, w);
                }
            } else if (motion) {
                if (intr) hipLaunchKernelGGL((ba_frame_kernel<true, true, false, PD>), fgrid, dim3(kThreads), lds, s, d, a, w);
                else hipLaunchKernelGGL((ba_frame_kernel<false, true, false, PD>), fgrid, dim3(kThreads), lds, s, d, a, w);
            } else {
                if (intr) hipLaunchKernelGGL((ba_frame_kernel<true, false, false, PD>), fgrid, dim3(kThreads), lds, s, d, a, w);
                else hipLaunchKernelGGL((ba_frame_kernel<false, false, false, PD>), fgrid, dim3(kThreads), lds, s, d, a, w);
            }
            VH_LAUNCH_CHECK("ba_frame_kernel");
            if (d.n_wide > 0) {
                launch_wide<PD>(d, a, w, motion, intr, rk.type != 0, s);
                VH_LAUNCH_CHECK("ba_frame_wide_kernel");
            }
            hipLaunchKernelGGL(ba_finalize_kernel<PD>, dim3(d.E + (motion ? 0 : d.K * kGramSplit)), dim3(kThreads), 0,
                               s, d, a, w, motion ? 1 : 0);
            VH_LAUNCH_CHECK("ba_finalize_kernel");
        } else if (!motion && d.K > 0) {
            // no edges: every depth frame still has its prior / damping terms
            hipLaunchKernelGGL((ba_frame_kernel<false, false, false, PD>), fgrid, dim3(kThreads), lds, s, d, a, w);
            VH_LAUNCH_CHECK("ba_frame_kernel");
            hipLaunchKernelGGL(ba_finalize_kernel<PD>, dim3(d.E + d.K * kGramSplit), dim3(kThreads), 0, s, d, a, w, 0);
            VH_LAUNCH_CHECK("ba_finalize_kernel");
        }
        hipLaunchKernelGGL(ba_assemble_kernel<PD>, dim3(d.nlb + d.P + 1 + npad), dim3(64), 0, s, d, a, w,
                           motion ? 1 : 0);
        VH_LAUNCH_CHECK("ba_assemble_kernel");
        hipLaunchKernelGGL(chol_factor_kernel, dim3(1), dim3(256), 0, s, w.A, w.Li, d.ld, 0, w.flag);
        VH_LAUNCH_CHECK("chol_factor_kernel");
        for (int kb = 0; kb + 1 < d.nb; kb++) {
            const int rem = d.nb - kb - 1;
            hipLaunchKernelGGL(chol_step_kernel, dim3(rem * (rem + 1) / 2), dim3(256), 0, s, w.A, w.Li, w.b, w.y, d.ld,
                               kb, w.flag);
            VH_LAUNCH_CHECK("chol_step_kernel");
        }
        hipLaunchKernelGGL(chol_backsolve_kernel, dim3(1), dim3(kBT), (size_t)d.ld * 8, s, w.A, w.Li, d.ld, d.nb, w.b,
                           w.y, w.x, w.flag, PD * d.P, dx);
        VH_LAUNCH_CHECK("chol_backsolve_kernel");
        if (!motion && d.K > 0) {
            const dim3 bgrid((d.HW + kThreads - 1) / kThreads, d.K);
            if constexpr (CAM != 0) {
                hipLaunchKernelGGL((ba_backsub_kernel<false, 6, CAM>), bgrid, dim3(kThreads), 0, s, d, a, w);
            } else {
                if (intr) hipLaunchKernelGGL((ba_backsub_kernel<true, PD>), bgrid, dim3(kThreads), 0, s, d, a, w);
                else hipLaunchKernelGGL((ba_backsub_kernel<false, PD>), bgrid, dim3(kThreads), 0, s, d, a, w);
            }
            VH_LAUNCH_CHECK("ba_backsub_kernel");
        }
        if (PD == 7)
            hipLaunchKernelGGL(ba_retract_sim3_kernel, dim3((d.P + 63) / 64), dim3(64), 0, s, poses, intrinsics, w.x,
                               d.t0, d.P, intr ? 1 : 0, intrinsics_scale);
        else
            hipLaunchKernelGGL(ba_retract_kernel, dim3((d.P + 63) / 64), dim3(64), 0, s, poses, intrinsics, w.x,
                               d.t0, d.P, intr ? 1 : 0, intrinsics_scale);
        VH_LAUNCH_CHECK("ba_retract_kernel");
    }
    return VH_OK;
}

extern "C" {

// camera < 0: a single-view plan; VH_CAM_PINHOLE / VH_CAM_MEI: a rig plan
// with per-view intrinsics [V][4 or 5] and rig [V][7]
static int ba_run_impl(const vh_ba_plan* plan, float* poses, float* disps, float* intrinsics,
                       const float* disps_sens, const float* targets, const float* weights, const float* eta,
                       const float* depth_active, int iterations, float lm, float ep, float alpha,
                       float intrinsics_lm, float intrinsics_ep, float intrinsics_scale, int flags, const RobustK& rk,
                       float* dx, float* dz, float* energy, void* workspace, void* stream, int camera = -1,
                       const float* rig = nullptr) {
    if (!plan) return fail(VH_ERR_ARG, "ba_run: null plan");
    if (plan->d.V > 0 && camera < 0)
        return fail(VH_ERR_ARG, "ba_run: a rig plan (vh_ba_plan_create_rig) runs through vh_ba_run_rig");
    if (plan->d.V == 0 && camera >= 0)
        return fail(VH_ERR_ARG, "ba_run_rig: the plan is a single-view plan (make it with vh_ba_plan_create_rig)");
    if (camera >= 0 && (flags & (VH_BA_SIM3 | VH_BA_OPTIMIZE_INTRINSICS)))
        return fail(VH_ERR_UNSUPPORTED, "ba_run_rig: SE3 poses with fixed intrinsics only");
    if (iterations < 0) return fail(VH_ERR_ARG, "ba_run: negative iterations");
    if (iterations == 0) return VH_OK;
    if (!poses || !disps || !intrinsics || !disps_sens || !targets || !weights || !depth_active || !dx ||
        !workspace)
        return fail(VH_ERR_ARG, "ba_run: null tensor pointer");
    const bool motion = (flags & VH_BA_MOTION_ONLY) != 0;
    const bool intr = (flags & VH_BA_OPTIMIZE_INTRINSICS) != 0;
    const bool cen = (flags & VH_BA_COMPUTE_ENERGY) != 0;
    const bool sim3 = (flags & VH_BA_SIM3) != 0;
    if (!motion && (!eta || !dz)) return fail(VH_ERR_ARG, "ba_run: eta and dz are required unless motion only");
    if (cen && !energy) return fail(VH_ERR_ARG, "ba_run: energy output required");
    hipStream_t s = (hipStream_t)stream;
    DevPlan d = plan->d;
    const int pd = sim3 ? 7 : 6;
    if (sim3) {
        // the plan was sized for SE3 blocks; 7-row blocks must still fit
        if (pd * d.P + 1 > kMaxLd)
            return fail(VH_ERR_UNSUPPORTED, "ba_run: %d Sim3 poses in the window (limit %d)", d.P, (kMaxLd - 1) / 7);
        if (!gram_fits(d.Rext_max, pd) || frame_lds_bytes(d.Rext_max, plan->TP7, pd) > (size_t)kLdsMax)
            return fail(VH_ERR_UNSUPPORTED, "ba_run: the frame kernel's 7-row blocks do not fit (%d E-row blocks)",
                        d.Rext_max);
    }
    if (sim3) {
        // 256-pixel tiles either way, so the tile count T (and every buffer) is unchanged
        d.TP = plan->TP7;
        d.SUB = std::max(1, 256 / d.TP);
    }
    d.pd = pd;
    d.GT = plan->gram_blocks * pd * pd;
    d.intr = intr ? 1 : 0;
    d.n = pd * d.P + (intr ? 1 : 0);
    d.ld = (pd * d.P + 1 + kCB - 1) / kCB * kCB;
    d.nb = d.ld / kCB;
    char* ws = (char*)workspace;
    Work w;
    w.edge_part = (float*)(ws + plan->o_edge);
    w.gram_part = (float*)(ws + plan->o_gram);
    w.fen_part = (float*)(ws + plan->o_fen);
    w.Qb = (float*)(ws + plan->o_Q);
    w.Wb = (float*)(ws + plan->o_W);
    w.ed = (double*)(ws + plan->o_ed);
    w.gd = (double*)(ws + plan->o_gd);
    w.fen = (double*)(ws + plan->o_fenk);
    w.A = (double*)(ws + plan->o_A);
    w.b = (double*)(ws + plan->o_b);
    w.x = (double*)(ws + plan->o_x);
    w.Li = (double*)(ws + plan->o_Li);
    w.y = (double*)(ws + plan->o_y);
    w.flag = (int*)(ws + plan->o_flag);
    w.wide_e = (float*)(ws + plan->o_wide);
    RunArgs a{};
    a.rk = rk;
    a.poses = poses; a.disps = disps; a.intrinsics = intrinsics; a.disps_sens = disps_sens;
    a.targets = targets; a.weights = weights; a.eta = eta; a.depth_active = depth_active;
    a.lm = lm; a.ep = ep; a.alpha = alpha; a.ilm = intrinsics_lm; a.iep = intrinsics_ep;
    a.iscale = intrinsics_scale; a.dx = dx; a.dz = dz; a.energy = energy; a.compute_energy = cen ? 1 : 0;
    if (camera >= 0) {
        if (camera != VH_CAM_PINHOLE && camera != VH_CAM_MEI)
            return fail(VH_ERR_ARG, "ba_run_rig: unknown camera model %d", camera);
        if (!rig) return fail(VH_ERR_ARG, "ba_run_rig: null rig");
        a.rig = rig;
        a.kstride = camera == VH_CAM_MEI ? 5 : 4;
        if (camera == VH_CAM_MEI)
            return ba_iterations<6, 2>(d, a, w, iterations, motion, false, rk, poses, intrinsics, dx, 1.f, s);
        return ba_iterations<6, 1>(d, a, w, iterations, motion, false, rk, poses, intrinsics, dx, 1.f, s);
    }
    if (sim3) return ba_iterations<7>(d, a, w, iterations, motion, intr, rk, poses, intrinsics, dx, intrinsics_scale, s);
    return ba_iterations<6>(d, a, w, iterations, motion, intr, rk, poses, intrinsics, dx, intrinsics_scale, s);
}

int vh_ba_run(const vh_ba_plan* plan, float* poses, float* disps, float* intrinsics, const float* disps_sens,
              const float* targets, const float* weights, const float* eta, const float* depth_active,
              int iterations, float lm, float ep, float alpha, float intrinsics_lm, float intrinsics_ep,
              float intrinsics_scale, int flags, float* dx, float* dz, float* energy, void* workspace,
              void* stream) {
    return ba_run_impl(plan, poses, disps, intrinsics, disps_sens, targets, weights, eta, depth_active, iterations,
                       lm, ep, alpha, intrinsics_lm, intrinsics_ep, intrinsics_scale, flags, RobustK{0, 0.f, 0.f, 0.f},
                       dx, dz, energy, workspace, stream);
}

int vh_ba_run_static(const vh_ba_plan* plan, float* poses, float* disps, float* intrinsics,
                     const float* disps_sens, const float* targets, const float* weights, const float* eta,
                     const float* depth_active, int iterations, float lm, float ep, float alpha, float intrinsics_lm,
                     float intrinsics_ep, float intrinsics_scale, int flags, float* dx, float* dz, float* energy,
                     void* stream) {
    if (!plan) return fail(VH_ERR_ARG, "ba_run_static: null plan");
    if (plan->ws_bytes > kStaticScratchBytes)
        return fail(VH_ERR_UNSUPPORTED, "ba_run_static: the plan needs %zu B of workspace (static limit %zu B)",
                    plan->ws_bytes, kStaticScratchBytes);
    // no kernel reads workspace memory it has not written in the same call,
    // so the array needs no clearing between calls
    void* ws = static_scratch();
    if (!ws) return VH_ERR_HIP;
    return ba_run_impl(plan, poses, disps, intrinsics, disps_sens, targets, weights, eta, depth_active, iterations,
                       lm, ep, alpha, intrinsics_lm, intrinsics_ep, intrinsics_scale, flags, RobustK{0, 0.f, 0.f, 0.f},
                       dx, dz, energy, ws, stream);
}

int vh_ba_run_robust(const vh_ba_plan* plan, float* poses, float* disps, float* intrinsics, const float* disps_sens,
                     const float* targets, const float* weights, const float* eta, const float* depth_active,
                     int iterations, float lm, float ep, float alpha, float intrinsics_lm, float intrinsics_ep,
                     float intrinsics_scale, int flags, int robust_type, float threshold, float gnc_mu,
                     float gnc_mu_max, float* dx, float* dz, float* energy, void* workspace, void* stream) {
    if (robust_type < VH_ROBUST_NONE || robust_type > VH_ROBUST_GNC_TLS)
        return fail(VH_ERR_ARG, "ba_run_robust: unknown robust kernel %d", robust_type);
    if (robust_type != VH_ROBUST_NONE && !(threshold > 0.f))
        return fail(VH_ERR_ARG, "ba_run_robust: threshold must be positive");
    if (robust_type == VH_ROBUST_GNC_TLS && !(gnc_mu > 0.f))
        return fail(VH_ERR_ARG, "ba_run_robust: GNC mu must be positive");
    return ba_run_impl(plan, poses, disps, intrinsics, disps_sens, targets, weights, eta, depth_active, iterations,
                       lm, ep, alpha, intrinsics_lm, intrinsics_ep, intrinsics_scale, flags,
                       RobustK{robust_type, threshold, gnc_mu, gnc_mu_max}, dx, dz, energy, workspace, stream);
}

int vh_ba_run_rig(const vh_ba_plan* plan, float* poses, float* disps, const float* intrinsics, int camera,
                  const float* rig, const float* disps_sens, const float* targets, const float* weights,
                  const float* eta, const float* depth_active, int iterations, float lm, float ep, float alpha,
                  int flags, int robust_type, float threshold, float gnc_mu, float gnc_mu_max, float* dx, float* dz,
                  float* energy, void* workspace, void* stream) {
    if (camera < 0) return fail(VH_ERR_ARG, "ba_run_rig: unknown camera model %d", camera);
    if (robust_type < VH_ROBUST_NONE || robust_type > VH_ROBUST_GNC_TLS)
        return fail(VH_ERR_ARG, "ba_run_rig: unknown robust kernel %d", robust_type);
    if (robust_type != VH_ROBUST_NONE && !(threshold > 0.f))
        return fail(VH_ERR_ARG, "ba_run_rig: threshold must be positive");
    if (robust_type == VH_ROBUST_GNC_TLS && !(gnc_mu > 0.f))
        return fail(VH_ERR_ARG, "ba_run_rig: GNC mu must be positive");
    return ba_run_impl(plan, poses, disps, const_cast<float*>(intrinsics), disps_sens, targets, weights, eta,
                       depth_active, iterations, lm, ep, alpha, 0.f, 0.f, 1.f, flags,
                       RobustK{robust_type, threshold, gnc_mu, gnc_mu_max}, dx, dz, energy, workspace, stream, camera,
                       rig);
}

int vh_robust_weights(int robust_type, float threshold, float gnc_mu, float gnc_mu_max, const float* x, float* w,
                      int64_t n, void* stream) {
    if (robust_type < VH_ROBUST_NONE || robust_type > VH_ROBUST_GNC_TLS)
        return fail(VH_ERR_ARG, "robust_weights: unknown robust kernel %d", robust_type);
    if (n < 0) return fail(VH_ERR_ARG, "robust_weights: negative size");
    if (n == 0) return VH_OK;
    if (!x || !w) return fail(VH_ERR_ARG, "robust_weights: null pointer");
    const RobustK k{robust_type, threshold, gnc_mu, gnc_mu_max};
    const int64_t blocks = (n + 255) / 256;
    hipLaunchKernelGGL(robust_weight_kernel, dim3((unsigned)std::min<int64_t>(blocks, 65536)), dim3(256), 0,
                       (hipStream_t)stream, k, x, w, n);
    VH_LAUNCH_CHECK("robust_weight_kernel");
    return VH_OK;
}


int vh_ba_adjoint(const vh_ba_plan* plan, const float* poses, const float* disps, const float* intrinsics,
                  const float* targets, const float* weights, const float* depth_active, float lm, int flags,
                  const float* grad_dx, const float* grad_dz, float* grad_targets, float* grad_weights,
                  void* workspace, void* stream) {
    if (!plan) return fail(VH_ERR_ARG, "ba_adjoint: null plan");
    if (plan->d.V > 0) return fail(VH_ERR_UNSUPPORTED, "ba_adjoint: rig plans (single-view plans only)");
    const bool motion = (flags & VH_BA_MOTION_ONLY) != 0;
    if (flags & VH_BA_OPTIMIZE_INTRINSICS) return fail(VH_ERR_UNSUPPORTED, "ba_adjoint: intrinsics optimisation");
    if (flags & VH_BA_SIM3) return fail(VH_ERR_UNSUPPORTED, "ba_adjoint: Sim3 poses (SE3 only)");
    if (plan->d.n_wide > 0)
        return fail(VH_ERR_UNSUPPORTED, "ba_adjoint: a depth frame has %d outgoing edges (limit %d)", plan->deg_max,
                    kMaxDeg);
    if (!poses || !disps || !intrinsics || !targets || !weights || !depth_active || !grad_dx || !workspace ||
        (!motion && !grad_dz) || (plan->d.E > 0 && (!grad_targets || !grad_weights)))
        return fail(VH_ERR_ARG, "ba_adjoint: null tensor pointer");
    if (lm != 0.f && !plan->adj_distinct)
        return fail(VH_ERR_UNSUPPORTED, "ba_adjoint: a depth frame has two edges to the same pose (needs lm = 0)");
    hipStream_t s = (hipStream_t)stream;
    DevPlan d = plan->d;
    d.intr = 0;
    d.n = 6 * d.P;
    char* ws = (char*)workspace;
    Work w{};
    w.Qb = (float*)(ws + plan->o_Q);
    w.Wb = (float*)(ws + plan->o_W);
    w.A = (double*)(ws + plan->o_A);
    w.x = (double*)(ws + plan->o_x);
    w.Li = (double*)(ws + plan->o_Li);
    w.flag = (int*)(ws + plan->o_flag);
    w.wide_e = (float*)(ws + plan->o_wide);
    RunArgs a{};
    a.poses = const_cast<float*>(poses);
    a.disps = const_cast<float*>(disps);
    a.intrinsics = const_cast<float*>(intrinsics);
    a.targets = targets;
    a.weights = weights;
    a.depth_active = depth_active;
    AdjArgs g;
    g.gdx = grad_dx;
    g.gdz = grad_dz;
    g.gt = grad_targets;
    g.gw = grad_weights;
    g.esum = (double*)(ws + plan->o_esum);
    g.ybar = (double*)(ws + plan->o_ybar);
    g.mu = (double*)(ws + plan->o_mu);
    g.lm = lm;
    const dim3 pgrid((d.HW + kThreads - 1) / kThreads, d.K);
    if (!motion && d.E > 0 && d.K > 0) {
        hipLaunchKernelGGL(zero_f64_kernel, dim3((d.E * 6 + 255) / 256), dim3(256), 0, s, g.esum, d.E * 6);
        VH_LAUNCH_CHECK("ba_adjoint zero");
        hipLaunchKernelGGL(ba_adj_rhs_kernel, pgrid, dim3(kThreads), 0, s, d, a, w, g);
        VH_LAUNCH_CHECK("ba_adj_rhs_kernel");
    }
    hipLaunchKernelGGL(ba_adj_pose_kernel, dim3(d.P + (d.ld - 6 * d.P + 63) / 64), dim3(64), 0, s, d, a, g,
                       (motion || d.E == 0) ? 1 : 0);
    VH_LAUNCH_CHECK("ba_adj_pose_kernel");
    hipLaunchKernelGGL(chol_solve_full_kernel, dim3(1), dim3(1024), (size_t)d.ld * 8, s, w.A, w.Li, d.ld, d.nb, g.ybar,
                       g.mu, w.flag);
    VH_LAUNCH_CHECK("chol_solve_full_kernel");
    if (d.E > 0 && d.K > 0) {
        if (motion) hipLaunchKernelGGL(ba_adj_grad_kernel<true>, pgrid, dim3(kThreads), 0, s, d, a, w, g);
        else hipLaunchKernelGGL(ba_adj_grad_kernel<false>, pgrid, dim3(kThreads), 0, s, d, a, w, g);
        VH_LAUNCH_CHECK("ba_adj_grad_kernel");
    }
    return VH_OK;
}

}  // extern "C"
