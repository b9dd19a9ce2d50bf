// Lie-group maths for SO3 / RxSO3 / SE3 / Sim3 on fixed-size arrays, usable
// from device kernels and from the host loop that serves CPU tensors.
//
// Semantics follow csrc/lietorch_ext/{so3,rxso3,se3,sim3}.h (quaternions
// normalised on construction, EPS = 1e-6 small-angle series, atan-based log
// with the +-pi/n branch, Sim3 left Jacobian without the Xi^5/720 term,
// truncated Sim3 inverse Jacobian).  Layouts: SO3 [q], RxSO3 [q,s], SE3 [t,q],
// Sim3 [t,q,s]; tangents SO3 [phi], RxSO3 [phi,sigma], SE3 [tau,phi],
// Sim3 [tau,phi,sigma].
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

#define VH_HD __host__ __device__ __forceinline__

namespace vh {
namespace lie {

constexpr double kEps = 1e-6;
constexpr double kPi = 3.14159265358979323846;

// Transcendentals with explicit float / double overloads (host and device).
VH_HD float m_sqrt(float x) { return sqrtf(x); }
VH_HD double m_sqrt(double x) { return ::sqrt(x); }
VH_HD float m_sin(float x) { return sinf(x); }
VH_HD double m_sin(double x) { return ::sin(x); }
VH_HD float m_cos(float x) { return cosf(x); }
VH_HD double m_cos(double x) { return ::cos(x); }
VH_HD float m_atan(float x) { return atanf(x); }
VH_HD double m_atan(double x) { return ::atan(x); }
VH_HD float m_fabs(float x) { return fabsf(x); }
VH_HD double m_fabs(double x) { return ::fabs(x); }
VH_HD float m_exp(float x) { return expf(x); }
VH_HD double m_exp(double x) { return ::exp(x); }
VH_HD float m_log(float x) { return logf(x); }
VH_HD double m_log(double x) { return ::log(x); }

// A group element in one uniform form: x = s * R(q) * p + t.
template <typename T>
struct Elem {
    T t[3];
    T q[4];  // x, y, z, w (unit)
    T s;
};

template <typename T>
VH_HD void hat(const T* v, T M[3][3]) {
    M[0][0] = 0;     M[0][1] = -v[2]; M[0][2] = v[1];
    M[1][0] = v[2];  M[1][1] = 0;     M[1][2] = -v[0];
    M[2][0] = -v[1]; M[2][1] = v[0];  M[2][2] = 0;
}

template <typename T>
VH_HD void matmul3(const T A[3][3], const T B[3][3], T C[3][3]) {
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) C[i][j] = A[i][0] * B[0][j] + A[i][1] * B[1][j] + A[i][2] * B[2][j];
}

template <typename T>
VH_HD void cross(const T* a, const T* b, T* c) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

template <typename T>
VH_HD void qnormalize(T* q) {
    T n = m_sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    for (int i = 0; i < 4; i++) q[i] /= n;
}

// Hamilton product c = a * b, [x,y,z,w] storage.
template <typename T>
VH_HD void qmul(const T* a, const T* b, T* c) {
    T x = a[3] * b[0] + b[3] * a[0] + a[1] * b[2] - a[2] * b[1];
    T y = a[3] * b[1] + b[3] * a[1] + a[2] * b[0] - a[0] * b[2];
    T z = a[3] * b[2] + b[3] * a[2] + a[0] * b[1] - a[1] * b[0];
    T w = a[3] * b[3] - a[0] * b[0] - a[1] * b[1] - a[2] * b[2];
    c[0] = x; c[1] = y; c[2] = z; c[3] = w;
}

// Rotation of a point with the uv form: p + w*uv + v x uv, uv = 2 v x p.
template <typename T>
VH_HD void qrot(const T* q, const T* p, T* o) {
    T uv[3], c[3];
    cross(q, p, uv);
    uv[0] += uv[0]; uv[1] += uv[1]; uv[2] += uv[2];
    cross(q, uv, c);
    o[0] = p[0] + q[3] * uv[0] + c[0];
    o[1] = p[1] + q[3] * uv[1] + c[1];
    o[2] = p[2] + q[3] * uv[2] + c[2];
}

template <typename T>
VH_HD void qmat(const T* q, T R[3][3]) {
    T x = q[0], y = q[1], z = q[2], w = q[3];
    R[0][0] = 1 - 2 * (y * y + z * z); R[0][1] = 2 * (x * y - w * z);     R[0][2] = 2 * (x * z + w * y);
    R[1][0] = 2 * (x * y + w * z);     R[1][1] = 1 - 2 * (x * x + z * z); R[1][2] = 2 * (y * z - w * x);
    R[2][0] = 2 * (x * z - w * y);     R[2][1] = 2 * (y * z + w * x);     R[2][2] = 1 - 2 * (x * x + y * y);
}

// SO3 exponential -> unit quaternion (so3.h:135-152).
template <typename T>
VH_HD void so3_exp(const T* phi, T* q) {
    T th2 = phi[0] * phi[0] + phi[1] * phi[1] + phi[2] * phi[2];
    T th = m_sqrt(th2);
    T imag, real;
    if (th < T(kEps)) {
        T th4 = th2 * th2;
        imag = T(0.5) - T(1.0 / 48.0) * th2 + T(1.0 / 3840.0) * th4;
        real = T(1) - T(1.0 / 8.0) * th2 + T(1.0 / 384.0) * th4;
    } else {
        imag = m_sin(T(0.5) * th) / th;
        real = m_cos(T(0.5) * th);
    }
    q[0] = imag * phi[0]; q[1] = imag * phi[1]; q[2] = imag * phi[2]; q[3] = real;
    qnormalize(q);
}

// SO3 logarithm (so3.h:97-133).
template <typename T>
VH_HD void so3_log(const T* q, T* phi) {
    T n2 = q[0] * q[0] + q[1] * q[1] + q[2] * q[2];
    T w = q[3];
    T f;
    if (n2 < T(kEps * kEps)) {
        f = T(2) / w - T(2.0 / 3.0) * n2 / (w * w * w);
    } else {
        T n = m_sqrt(n2);
        if (m_fabs(w) < T(kEps)) {
            f = (w > T(0)) ? T(kPi) / n : -T(kPi) / n;
        } else {
            f = T(2) * m_atan(n / w) / n;
        }
    }
    phi[0] = f * q[0]; phi[1] = f * q[1]; phi[2] = f * q[2];
}

template <typename T>
VH_HD void so3_left_jac(const T* phi, T J[3][3]) {
    T P[3][3], P2[3][3];
    hat(phi, P);
    matmul3(P, P, P2);
    T th2 = phi[0] * phi[0] + phi[1] * phi[1] + phi[2] * phi[2];
    T th = m_sqrt(th2);
    T c1 = (th < T(kEps)) ? T(0.5) - T(1.0 / 24.0) * th2 : (T(1) - m_cos(th)) / th2;
    T c2 = (th < T(kEps)) ? T(1.0 / 6.0) - T(1.0 / 120.0) * th2 : (th - m_sin(th)) / (th2 * th);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) J[i][j] = (i == j ? T(1) : T(0)) + c1 * P[i][j] + c2 * P2[i][j];
}

template <typename T>
VH_HD void so3_left_jac_inv(const T* phi, T J[3][3]) {
    T P[3][3], P2[3][3];
    hat(phi, P);
    matmul3(P, P, P2);
    T th = m_sqrt(phi[0] * phi[0] + phi[1] * phi[1] + phi[2] * phi[2]);
    T half = T(0.5) * th;
    T c2 = (th < T(kEps)) ? T(1.0 / 12.0) : (T(1) - th * m_cos(half) / (T(2) * m_sin(half))) / (th * th);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) J[i][j] = (i == j ? T(1) : T(0)) - T(0.5) * P[i][j] + c2 * P2[i][j];
}

// RxSO3::calcW (rxso3.h:175-215).
template <typename T>
VH_HD void calcW(const T* phi, T sigma, T Wm[3][3]) {
    T th = m_sqrt(phi[0] * phi[0] + phi[1] * phi[1] + phi[2] * phi[2]);
    T P[3][3], P2[3][3];
    hat(phi, P);
    matmul3(P, P, P2);
    T scale = m_exp(sigma);
    T A, B, C;
    const T one(1), half(0.5);
    if (m_fabs(sigma) < T(kEps)) {
        C = one;
        if (m_fabs(th) < T(kEps)) {
            A = half;
            B = T(1. / 6.);
        } else {
            T ts = th * th;
            A = (one - m_cos(th)) / ts;
            B = (th - m_sin(th)) / (ts * th);
        }
    } else {
        C = (scale - one) / sigma;
        if (m_fabs(th) < T(kEps)) {
            T ss = sigma * sigma;
            A = ((sigma - one) * scale + one) / ss;
            B = (scale * half * ss + scale - one - sigma * scale) / (ss * sigma);
        } else {
            T ts = th * th;
            T a = scale * m_sin(th);
            T b = scale * m_cos(th);
            T c = ts + sigma * sigma;
            A = (a * sigma + (one - b) * th) / (th * c);
            B = (C - ((b - one) * sigma + a * th) / c) * one / ts;
        }
    }
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) Wm[i][j] = A * P[i][j] + B * P2[i][j] + (i == j ? C : T(0));
}

template <typename T>
VH_HD bool inv3(const T M[3][3], T I[3][3]) {
    T c00 = M[1][1] * M[2][2] - M[1][2] * M[2][1];
    T c01 = M[1][2] * M[2][0] - M[1][0] * M[2][2];
    T c02 = M[1][0] * M[2][1] - M[1][1] * M[2][0];
    T det = M[0][0] * c00 + M[0][1] * c01 + M[0][2] * c02;
    T id = T(1) / det;
    I[0][0] = c00 * id;
    I[0][1] = (M[0][2] * M[2][1] - M[0][1] * M[2][2]) * id;
    I[0][2] = (M[0][1] * M[1][2] - M[0][2] * M[1][1]) * id;
    I[1][0] = c01 * id;
    I[1][1] = (M[0][0] * M[2][2] - M[0][2] * M[2][0]) * id;
    I[1][2] = (M[0][2] * M[1][0] - M[0][0] * M[1][2]) * id;
    I[2][0] = c02 * id;
    I[2][1] = (M[0][1] * M[2][0] - M[0][0] * M[2][1]) * id;
    I[2][2] = (M[0][0] * M[1][1] - M[0][1] * M[1][0]) * id;
    return det != T(0);
}

// SE3::calcQ (se3.h:135-160).
template <typename T>
VH_HD void se3_calcQ(const T* tau, const T* phi, T Q[3][3]) {
    T Ta[3][3], P[3][3];
    hat(tau, Ta);
    hat(phi, P);
    T th = m_sqrt(phi[0] * phi[0] + phi[1] * phi[1] + phi[2] * phi[2]);
    T t2 = th * th, t4 = t2 * t2;
    T c1 = (th < T(kEps)) ? T(1.0 / 6.0) - T(1.0 / 120.0) * t2 : (th - m_sin(th)) / (t2 * th);
    T c2 = (th < T(kEps)) ? T(1.0 / 24.0) - T(1.0 / 720.0) * t2 : (t2 + T(2) * m_cos(th) - T(2)) / (T(2) * t4);
    T c3 = (th < T(kEps)) ? T(1.0 / 120.0) - T(1.0 / 2520.0) * t2
                          : (T(2) * th - T(3) * m_sin(th) + th * m_cos(th)) / (T(2) * t4 * th);
    T PT[3][3], TP[3][3], PTP[3][3], PP[3][3], PPT[3][3], TPP[3][3], PTPP[3][3], PPTP[3][3];
    matmul3(P, Ta, PT);
    matmul3(Ta, P, TP);
    matmul3(PT, P, PTP);
    matmul3(P, P, PP);
    matmul3(PP, Ta, PPT);
    matmul3(TP, P, TPP);
    matmul3(PTP, P, PTPP);
    matmul3(PP, TP, PPTP);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            Q[i][j] = T(0.5) * Ta[i][j] + c1 * (PT[i][j] + TP[i][j] + PTP[i][j]) +
                      c2 * (PPT[i][j] + TPP[i][j] - T(3) * PTP[i][j]) + c3 * (PTPP[i][j] + PPTP[i][j]);
}

// ---------------------------------------------------------------------------
// Group traits.  Each provides N, K, load/store, exp, log, Adj, adj,
// left_jac, left_jac_inv, act_jac, act4_jac, projector.
// ---------------------------------------------------------------------------

template <typename T>
VH_HD void elem_identity(Elem<T>& X) {
    X.t[0] = X.t[1] = X.t[2] = T(0);
    X.q[0] = X.q[1] = X.q[2] = T(0);
    X.q[3] = T(1);
    X.s = T(1);
}

template <typename T>
VH_HD void elem_act(const Elem<T>& X, const T* p, T* o) {
    qrot(X.q, p, o);
    for (int i = 0; i < 3; i++) o[i] = X.s * o[i] + X.t[i];
}

template <typename T>
VH_HD void elem_act4(const Elem<T>& X, const T* p, T* o) {
    qrot(X.q, p, o);
    for (int i = 0; i < 3; i++) o[i] = X.s * o[i] + p[3] * X.t[i];
    o[3] = p[3];
}

template <typename T>
VH_HD void elem_mul(const Elem<T>& A, const Elem<T>& B, Elem<T>& C) {
    T rt[3];
    qrot(A.q, B.t, rt);
    T q[4];
    qmul(A.q, B.q, q);
    qnormalize(q);
    for (int i = 0; i < 3; i++) C.t[i] = A.t[i] + A.s * rt[i];
    for (int i = 0; i < 4; i++) C.q[i] = q[i];
    C.s = A.s * B.s;
}

template <typename T>
VH_HD void elem_inv(const Elem<T>& A, Elem<T>& B) {
    T qi[4] = {-A.q[0], -A.q[1], -A.q[2], A.q[3]};
    T si = T(1) / A.s;
    T rt[3];
    qrot(qi, A.t, rt);
    for (int i = 0; i < 3; i++) B.t[i] = -(si * rt[i]);
    for (int i = 0; i < 4; i++) B.q[i] = qi[i];
    B.s = si;
}

template <typename T>
VH_HD void elem_matrix4(const Elem<T>& X, T* M) {  // row-major 4x4
    T R[3][3];
    qmat(X.q, R);
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++) M[i * 4 + j] = X.s * R[i][j];
        M[i * 4 + 3] = X.t[i];
    }
    M[12] = M[13] = M[14] = T(0);
    M[15] = T(1);
}

template <typename T>
VH_HD void so3_projector4(const T* q, T* J, int ld, int r0, int c0) {
    // 0.5*(w I + hat(-v)) in rows r0..r0+2, cols c0..c0+2; row r0+3 = -0.5 v.
    T v[3] = {-q[0], -q[1], -q[2]};
    T H[3][3];
    hat(v, H);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) J[(r0 + i) * ld + c0 + j] = T(0.5) * ((i == j ? q[3] : T(0)) + H[i][j]);
    for (int j = 0; j < 3; j++) J[(r0 + 3) * ld + c0 + j] = T(0.5) * (-q[j]);
}

template <typename T>
struct SO3T {
    static constexpr int N = 4, K = 3;
    VH_HD static void load(const T* d, Elem<T>& X) {
        elem_identity(X);
        for (int i = 0; i < 4; i++) X.q[i] = d[i];
        qnormalize(X.q);
    }
    VH_HD static void store(const Elem<T>& X, T* d) {
        for (int i = 0; i < 4; i++) d[i] = X.q[i];
    }
    VH_HD static void exp(const T* a, Elem<T>& X) {
        elem_identity(X);
        so3_exp(a, X.q);
    }
    VH_HD static void log(const Elem<T>& X, T* a) { so3_log(X.q, a); }
    VH_HD static void Adj(const Elem<T>& X, T* A) {
        T R[3][3];
        qmat(X.q, R);
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) A[i * K + j] = R[i][j];
    }
    VH_HD static void adj(const T* a, T* A) {
        T H[3][3];
        hat(a, H);
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) A[i * K + j] = H[i][j];
    }
    VH_HD static void left_jac(const T* a, T* J) {
        T M[3][3];
        so3_left_jac(a, M);
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) J[i * K + j] = M[i][j];
    }
    VH_HD static void left_jac_inv(const T* a, T* J) {
        T M[3][3];
        so3_left_jac_inv(a, M);
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) J[i * K + j] = M[i][j];
    }
    VH_HD static void act_jac(const T* p, T* J) {  // 3 x K
        T v[3] = {-p[0], -p[1], -p[2]};
        T H[3][3];
        hat(v, H);
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) J[i * K + j] = H[i][j];
    }
    VH_HD static void act4_jac(const T* p, T* J) {  // 4 x K
        for (int i = 0; i < 4 * K; i++) J[i] = T(0);
        T v[3] = {-p[0], -p[1], -p[2]};
        T H[3][3];
        hat(v, H);
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) J[i * K + j] = H[i][j];
    }
    VH_HD static void projector(const Elem<T>& X, T* J) {  // N x N
        for (int i = 0; i < N * N; i++) J[i] = T(0);
        so3_projector4(X.q, J, N, 0, 0);
    }
};

template <typename T>
struct RxSO3T {
    static constexpr int N = 5, K = 4;
    VH_HD static void load(const T* d, Elem<T>& X) {
        elem_identity(X);
        for (int i = 0; i < 4; i++) X.q[i] = d[i];
        qnormalize(X.q);
        X.s = d[4];
    }
    VH_HD static void store(const Elem<T>& X, T* d) {
        for (int i = 0; i < 4; i++) d[i] = X.q[i];
        d[4] = X.s;
    }
    VH_HD static void exp(const T* a, Elem<T>& X) {
        elem_identity(X);
        so3_exp(a, X.q);
        X.s = m_exp(a[3]);
    }
    VH_HD static void log(const Elem<T>& X, T* a) {
        so3_log(X.q, a);
        a[3] = m_log(X.s);
    }
    VH_HD static void Adj(const Elem<T>& X, T* A) {
        T R[3][3];
        qmat(X.q, R);
        for (int i = 0; i < K * K; i++) A[i] = T(0);
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) A[i * K + j] = R[i][j];
        A[3 * K + 3] = T(1);
    }
    VH_HD static void adj(const T* a, T* A) {
        T H[3][3];
        hat(a, H);
        for (int i = 0; i < K * K; i++) A[i] = T(0);
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) A[i * K + j] = H[i][j];
    }
    VH_HD static void left_jac(const T* a, T* J) {
        T M[3][3];
        so3_left_jac(a, M);
        for (int i = 0; i < K * K; i++) J[i] = T(0);
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) J[i * K + j] = M[i][j];
        J[3 * K + 3] = T(1);
    }
    VH_HD static void left_jac_inv(const T* a, T* J) {
        T M[3][3];
        so3_left_jac_inv(a, M);
        for (int i = 0; i < K * K; i++) J[i] = T(0);
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) J[i * K + j] = M[i][j];
        J[3 * K + 3] = T(1);
    }
    VH_HD static void act_jac(const T* p, T* J) {  // [hat(-p), p]
        T v[3] = {-p[0], -p[1], -p[2]};
        T H[3][3];
        hat(v, H);
        for (int i = 0; i < 3; i++) {
            for (int j = 0; j < 3; j++) J[i * K + j] = H[i][j];
            J[i * K + 3] = p[i];
        }
    }
    VH_HD static void act4_jac(const T* p, T* J) {
        for (int i = 0; i < 4 * K; i++) J[i] = T(0);
        T v[3] = {-p[0], -p[1], -p[2]};
        T H[3][3];
        hat(v, H);
        for (int i = 0; i < 3; i++) {
            for (int j = 0; j < 3; j++) J[i * K + j] = H[i][j];
            J[i * K + 3] = p[i];
        }
    }
    VH_HD static void projector(const Elem<T>& X, T* J) {
        for (int i = 0; i < N * N; i++) J[i] = T(0);
        so3_projector4(X.q, J, N, 0, 0);
        J[4 * N + 3] = X.s;
    }
};

template <typename T>
struct SE3T {
    static constexpr int N = 7, K = 6;
    VH_HD static void load(const T* d, Elem<T>& X) {
        elem_identity(X);
        for (int i = 0; i < 3; i++) X.t[i] = d[i];
        for (int i = 0; i < 4; i++) X.q[i] = d[3 + i];
        qnormalize(X.q);
    }
    VH_HD static void store(const Elem<T>& X, T* d) {
        for (int i = 0; i < 3; i++) d[i] = X.t[i];
        for (int i = 0; i < 4; i++) d[3 + i] = X.q[i];
    }
    VH_HD static void exp(const T* a, Elem<T>& X) {
        elem_identity(X);
        so3_exp(a + 3, X.q);
        T J[3][3];
        so3_left_jac(a + 3, J);
        for (int i = 0; i < 3; i++) X.t[i] = J[i][0] * a[0] + J[i][1] * a[1] + J[i][2] * a[2];
    }
    VH_HD static void log(const Elem<T>& X, T* a) {
        so3_log(X.q, a + 3);
        T Vi[3][3];
        so3_left_jac_inv(a + 3, Vi);
        for (int i = 0; i < 3; i++) a[i] = Vi[i][0] * X.t[0] + Vi[i][1] * X.t[1] + Vi[i][2] * X.t[2];
    }
    VH_HD static void Adj(const Elem<T>& X, T* A) {
        T R[3][3], tx[3][3], tR[3][3];
        qmat(X.q, R);
        hat(X.t, tx);
        matmul3(tx, R, tR);
        for (int i = 0; i < K * K; i++) A[i] = T(0);
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                A[i * K + j] = R[i][j];
                A[i * K + 3 + j] = tR[i][j];
                A[(3 + i) * K + 3 + j] = R[i][j];
            }
    }
    VH_HD static void adj(const T* a, T* A) {
        T Ta[3][3], P[3][3];
        hat(a, Ta);
        hat(a + 3, P);
        for (int i = 0; i < K * K; i++) A[i] = T(0);
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                A[i * K + j] = P[i][j];
                A[i * K + 3 + j] = Ta[i][j];
                A[(3 + i) * K + 3 + j] = P[i][j];
            }
    }
    VH_HD static void left_jac(const T* a, T* Jo) {
        T J[3][3], Q[3][3];
        so3_left_jac(a + 3, J);
        se3_calcQ(a, a + 3, Q);
        for (int i = 0; i < K * K; i++) Jo[i] = T(0);
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                Jo[i * K + j] = J[i][j];
                Jo[i * K + 3 + j] = Q[i][j];
                Jo[(3 + i) * K + 3 + j] = J[i][j];
            }
    }
    VH_HD static void left_jac_inv(const T* a, T* Jo) {
        T Ji[3][3], Q[3][3], A1[3][3], A2[3][3];
        so3_left_jac_inv(a + 3, Ji);
        se3_calcQ(a, a + 3, Q);
        matmul3(Ji, Q, A1);
        matmul3(A1, Ji, A2);
        for (int i = 0; i < K * K; i++) Jo[i] = T(0);
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                Jo[i * K + j] = Ji[i][j];
                Jo[i * K + 3 + j] = -A2[i][j];
                Jo[(3 + i) * K + 3 + j] = Ji[i][j];
            }
    }
    VH_HD static void act_jac(const T* p, T* J) {  // [I, hat(-p)]
        T v[3] = {-p[0], -p[1], -p[2]};
        T H[3][3];
        hat(v, H);
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                J[i * K + j] = (i == j) ? T(1) : T(0);
                J[i * K + 3 + j] = H[i][j];
            }
    }
    VH_HD static void act4_jac(const T* p, T* J) {  // [[p3 I, hat(-p)], [0]]
        for (int i = 0; i < 4 * K; i++) J[i] = T(0);
        T v[3] = {-p[0], -p[1], -p[2]};
        T H[3][3];
        hat(v, H);
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                J[i * K + j] = (i == j) ? p[3] : T(0);
                J[i * K + 3 + j] = H[i][j];
            }
    }
    VH_HD static void projector(const Elem<T>& X, T* J) {
        for (int i = 0; i < N * N; i++) J[i] = T(0);
        T v[3] = {-X.t[0], -X.t[1], -X.t[2]};
        T H[3][3];
        hat(v, H);
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                J[i * N + j] = (i == j) ? T(1) : T(0);
                J[i * N + 3 + j] = H[i][j];
            }
        so3_projector4(X.q, J, N, 3, 3);
    }
};

template <typename T>
struct Sim3T {
    static constexpr int N = 8, K = 7;
    VH_HD static void load(const T* d, Elem<T>& X) {
        elem_identity(X);
        for (int i = 0; i < 3; i++) X.t[i] = d[i];
        for (int i = 0; i < 4; i++) X.q[i] = d[3 + i];
        qnormalize(X.q);
        X.s = d[7];
    }
    VH_HD static void store(const Elem<T>& X, T* d) {
        for (int i = 0; i < 3; i++) d[i] = X.t[i];
        for (int i = 0; i < 4; i++) d[3 + i] = X.q[i];
        d[7] = X.s;
    }
    VH_HD static void exp(const T* a, Elem<T>& X) {
        elem_identity(X);
        so3_exp(a + 3, X.q);
        X.s = m_exp(a[6]);
        T Wm[3][3];
        calcW(a + 3, a[6], Wm);
        for (int i = 0; i < 3; i++) X.t[i] = Wm[i][0] * a[0] + Wm[i][1] * a[1] + Wm[i][2] * a[2];
    }
    VH_HD static void log(const Elem<T>& X, T* a) {
        so3_log(X.q, a + 3);
        a[6] = m_log(X.s);
        T Wm[3][3], Wi[3][3];
        calcW(a + 3, a[6], Wm);
        inv3(Wm, Wi);
        for (int i = 0; i < 3; i++) a[i] = Wi[i][0] * X.t[0] + Wi[i][1] * X.t[1] + Wi[i][2] * X.t[2];
    }
    VH_HD static void Adj(const Elem<T>& X, T* A) {
        T R[3][3], tx[3][3], tR[3][3];
        qmat(X.q, R);
        hat(X.t, tx);
        matmul3(tx, R, tR);
        for (int i = 0; i < K * K; i++) A[i] = T(0);
        for (int i = 0; i < 3; i++) {
            for (int j = 0; j < 3; j++) {
                A[i * K + j] = X.s * R[i][j];
                A[i * K + 3 + j] = tR[i][j];
                A[(3 + i) * K + 3 + j] = R[i][j];
            }
            A[i * K + 6] = -X.t[i];
        }
        A[6 * K + 6] = T(1);
    }
    VH_HD static void adj(const T* a, T* A) {
        T Ta[3][3], P[3][3];
        hat(a, Ta);
        hat(a + 3, P);
        for (int i = 0; i < K * K; i++) A[i] = T(0);
        for (int i = 0; i < 3; i++) {
            for (int j = 0; j < 3; j++) {
                A[i * K + j] = P[i][j] + (i == j ? a[6] : T(0));
                A[i * K + 3 + j] = Ta[i][j];
                A[(3 + i) * K + 3 + j] = P[i][j];
            }
            A[i * K + 6] = -a[i];
        }
    }
    // Truncated series: I + X/2 + X^2/6 + X^3/24 + X^4/120 (no X^5 term, sim3.h:160-169).
    VH_HD static void left_jac(const T* a, T* J) {
        T X[K * K], X2[K * K], X3[K * K], X4[K * K];
        adj(a, X);
        mm(X, X, X2);
        mm(X, X2, X3);
        mm(X2, X2, X4);
        for (int i = 0; i < K * K; i++)
            J[i] = ((i % (K + 1)) == 0 ? T(1) : T(0)) + T(1.0 / 2.0) * X[i] + T(1.0 / 6.0) * X2[i] +
                   T(1.0 / 24.0) * X3[i] + T(1.0 / 120.0) * X4[i];
    }
    // I - X/2 + X^2/12 - X^4/720 (sim3.h:171-178).
    VH_HD static void left_jac_inv(const T* a, T* J) {
        T X[K * K], X2[K * K], X4[K * K];
        adj(a, X);
        mm(X, X, X2);
        mm(X2, X2, X4);
        for (int i = 0; i < K * K; i++)
            J[i] = ((i % (K + 1)) == 0 ? T(1) : T(0)) - T(1.0 / 2.0) * X[i] + T(1.0 / 12.0) * X2[i] -
                   T(1.0 / 720.0) * X4[i];
    }
    VH_HD static void mm(const T* A, const T* B, T* C) {
        for (int i = 0; i < K; i++)
            for (int j = 0; j < K; j++) {
                T s = 0;
                for (int k = 0; k < K; k++) s += A[i * K + k] * B[k * K + j];
                C[i * K + j] = s;
            }
    }
    VH_HD static void act_jac(const T* p, T* J) {  // [I, hat(-p), p]
        T v[3] = {-p[0], -p[1], -p[2]};
        T H[3][3];
        hat(v, H);
        for (int i = 0; i < 3; i++) {
            for (int j = 0; j < 3; j++) {
                J[i * K + j] = (i == j) ? T(1) : T(0);
                J[i * K + 3 + j] = H[i][j];
            }
            J[i * K + 6] = p[i];
        }
    }
    VH_HD static void act4_jac(const T* p, T* J) {
        for (int i = 0; i < 4 * K; i++) J[i] = T(0);
        T v[3] = {-p[0], -p[1], -p[2]};
        T H[3][3];
        hat(v, H);
        for (int i = 0; i < 3; i++) {
            for (int j = 0; j < 3; j++) {
                J[i * K + j] = (i == j) ? p[3] : T(0);
                J[i * K + 3 + j] = H[i][j];
            }
            J[i * K + 6] = p[i];
        }
    }
    VH_HD static void projector(const Elem<T>& X, T* J) {
        for (int i = 0; i < N * N; i++) J[i] = T(0);
        T v[3] = {-X.t[0], -X.t[1], -X.t[2]};
        T H[3][3];
        hat(v, H);
        for (int i = 0; i < 3; i++) {
            for (int j = 0; j < 3; j++) {
                J[i * N + j] = (i == j) ? T(1) : T(0);
                J[i * N + 3 + j] = H[i][j];
            }
            J[i * N + 6] = X.t[i];
        }
        so3_projector4(X.q, J, N, 3, 3);
        J[7 * N + 6] = X.s;
    }
};

}  // namespace lie
}  // namespace vh
