// lietorch_ext on gfx950: the 13 group ops (+ as_matrix, projector, Jinv) for
// SO3 / RxSO3 / SE3 / Sim3 in float and double.  One element per lane,
// grid-stride, 256-thread blocks; the same element functors serve CPU
// tensors through a host loop (the reference's lietorch_cpu.cpp path).
//
// Forward / backward formulas follow csrc/lietorch_ext/lietorch_gpu.cu:24-299.
#include "lie_math.hpp"
#include "vh_common.hpp"

namespace vh {
namespace {

using namespace lie;

// ---- element functors ------------------------------------------------------
// Each is a POD with device pointers and an operator()(i) for element i.

template <class G, typename T>
struct ExpF {
    const T* a; T* X;
    VH_HD void operator()(int64_t i) const {
        Elem<T> E; G::exp(a + i * G::K, E); G::store(E, X + i * G::N);
    }
};
template <class G, typename T>
struct ExpB {  // da = grad[:K] * J_l(a)
    const T* g; const T* a; T* da;
    VH_HD void operator()(int64_t i) const {
        T J[G::K * G::K];
        G::left_jac(a + i * G::K, J);
        const T* gi = g + i * G::N;
        for (int c = 0; c < G::K; c++) {
            T s = 0;
            for (int r = 0; r < G::K; r++) s += gi[r] * J[r * G::K + c];
            da[i * G::K + c] = s;
        }
    }
};
template <class G, typename T>
struct LogF {
    const T* X; T* a;
    VH_HD void operator()(int64_t i) const {
        Elem<T> E; G::load(X + i * G::N, E); G::log(E, a + i * G::K);
    }
};
template <class G, typename T>
struct LogB {  // dX[:K] = grad * J_l^-1(log X)
    const T* g; const T* X; T* dX;
    VH_HD void operator()(int64_t i) const {
        Elem<T> E; G::load(X + i * G::N, E);
        T a[G::K], J[G::K * G::K];
        G::log(E, a);
        G::left_jac_inv(a, J);
        const T* gi = g + i * G::K;
        for (int c = 0; c < G::K; c++) {
            T s = 0;
            for (int r = 0; r < G::K; r++) s += gi[r] * J[r * G::K + c];
            dX[i * G::N + c] = s;
        }
    }
};
template <class G, typename T>
struct InvF {
    const T* X; T* Y;
    VH_HD void operator()(int64_t i) const {
        Elem<T> E, F; G::load(X + i * G::N, E); elem_inv(E, F); G::store(F, Y + i * G::N);
    }
};
template <class G, typename T>
struct InvB {  // dX[:K] = -grad[:K] * Adj(X^-1)
    const T* g; const T* X; T* dX;
    VH_HD void operator()(int64_t i) const {
        Elem<T> E, F; G::load(X + i * G::N, E); elem_inv(E, F);
        T A[G::K * G::K]; G::Adj(F, A);
        const T* gi = g + i * G::N;
        for (int c = 0; c < G::K; c++) {
            T s = 0;
            for (int r = 0; r < G::K; r++) s += gi[r] * A[r * G::K + c];
            dX[i * G::N + c] = -s;
        }
    }
};
template <class G, typename T>
struct MulF {
    const T* X; const T* Y; T* Z;
    VH_HD void operator()(int64_t i) const {
        Elem<T> A, B, C; G::load(X + i * G::N, A); G::load(Y + i * G::N, B);
        elem_mul(A, B, C); G::store(C, Z + i * G::N);
    }
};
template <class G, typename T>
struct MulB {  // dX = grad ; dY = grad * Adj(X)
    const T* g; const T* X; T* dX; T* dY;
    VH_HD void operator()(int64_t i) const {
        Elem<T> A; G::load(X + i * G::N, A);
        T Ad[G::K * G::K]; G::Adj(A, Ad);
        const T* gi = g + i * G::N;
        for (int c = 0; c < G::K; c++) {
            dX[i * G::N + c] = gi[c];
            T s = 0;
            for (int r = 0; r < G::K; r++) s += gi[r] * Ad[r * G::K + c];
            dY[i * G::N + c] = s;
        }
    }
};
template <class G, typename T>
struct AdjF {
    const T* X; const T* a; T* b;
    VH_HD void operator()(int64_t i) const {
        Elem<T> A; G::load(X + i * G::N, A);
        T Ad[G::K * G::K]; G::Adj(A, Ad);
        for (int r = 0; r < G::K; r++) {
            T s = 0;
            for (int c = 0; c < G::K; c++) s += Ad[r * G::K + c] * a[i * G::K + c];
            b[i * G::K + r] = s;
        }
    }
};
template <class G, typename T>
struct AdjB {  // da = db * Adj(X); dX = -db * ad(Adj(X) a)
    const T* g; const T* X; const T* a; T* dX; T* da;
    VH_HD void operator()(int64_t i) const {
        Elem<T> A; G::load(X + i * G::N, A);
        T Ad[G::K * G::K], b[G::K], ad[G::K * G::K];
        G::Adj(A, Ad);
        for (int r = 0; r < G::K; r++) {
            T s = 0;
            for (int c = 0; c < G::K; c++) s += Ad[r * G::K + c] * a[i * G::K + c];
            b[r] = s;
        }
        G::adj(b, ad);
        const T* gi = g + i * G::K;
        for (int c = 0; c < G::K; c++) {
            T s1 = 0, s2 = 0;
            for (int r = 0; r < G::K; r++) {
                s1 += gi[r] * Ad[r * G::K + c];
                s2 += gi[r] * ad[r * G::K + c];
            }
            da[i * G::K + c] = s1;
            dX[i * G::N + c] = -s2;
        }
    }
};
template <class G, typename T>
struct AdjTF {
    const T* X; const T* a; T* b;
    VH_HD void operator()(int64_t i) const {
        Elem<T> A; G::load(X + i * G::N, A);
        T Ad[G::K * G::K]; G::Adj(A, Ad);
        for (int c = 0; c < G::K; c++) {
            T s = 0;
            for (int r = 0; r < G::K; r++) s += Ad[r * G::K + c] * a[i * G::K + r];
            b[i * G::K + c] = s;
        }
    }
};
template <class G, typename T>
struct AdjTB {  // da = Adj(X) db; dX = -a * ad(Adj(X) db)
    const T* g; const T* X; const T* a; T* dX; T* da;
    VH_HD void operator()(int64_t i) const {
        Elem<T> A; G::load(X + i * G::N, A);
        T Ad[G::K * G::K], v[G::K], ad[G::K * G::K];
        G::Adj(A, Ad);
        const T* gi = g + i * G::K;
        for (int r = 0; r < G::K; r++) {
            T s = 0;
            for (int c = 0; c < G::K; c++) s += Ad[r * G::K + c] * gi[c];
            v[r] = s;
            da[i * G::K + r] = s;
        }
        G::adj(v, ad);
        for (int c = 0; c < G::K; c++) {
            T s = 0;
            for (int r = 0; r < G::K; r++) s += a[i * G::K + r] * ad[r * G::K + c];
            dX[i * G::N + c] = -s;
        }
    }
};
template <class G, typename T>
struct ActF {
    const T* X; const T* p; T* q;
    VH_HD void operator()(int64_t i) const {
        Elem<T> A; G::load(X + i * G::N, A); elem_act(A, p + i * 3, q + i * 3);
    }
};
template <class G, typename T>
struct ActB {  // dp = dq * M[:3,:3]; dX = dq * act_jac(X p)
    const T* g; const T* X; const T* p; T* dX; T* dp;
    VH_HD void operator()(int64_t i) const {
        Elem<T> A; G::load(X + i * G::N, A);
        T M[16], qp[3], J[3 * G::K];
        elem_matrix4(A, M);
        elem_act(A, p + i * 3, qp);
        G::act_jac(qp, J);
        const T* gi = g + i * 3;
        for (int c = 0; c < 3; c++) dp[i * 3 + c] = gi[0] * M[c] + gi[1] * M[4 + c] + gi[2] * M[8 + c];
        for (int c = 0; c < G::K; c++)
            dX[i * G::N + c] = gi[0] * J[c] + gi[1] * J[G::K + c] + gi[2] * J[2 * G::K + c];
    }
};
template <class G, typename T>
struct Act4F {
    const T* X; const T* p; T* q;
    VH_HD void operator()(int64_t i) const {
        Elem<T> A; G::load(X + i * G::N, A); elem_act4(A, p + i * 4, q + i * 4);
    }
};
template <class G, typename T>
struct Act4B {  // dp = dq * M; dX = dq * act4_jac(X.act4(p))
    const T* g; const T* X; const T* p; T* dX; T* dp;
    VH_HD void operator()(int64_t i) const {
        Elem<T> A; G::load(X + i * G::N, A);
        T M[16], qp[4], J[4 * G::K];
        elem_matrix4(A, M);
        elem_act4(A, p + i * 4, qp);
        G::act4_jac(qp, J);
        const T* gi = g + i * 4;
        for (int c = 0; c < 4; c++)
            dp[i * 4 + c] = gi[0] * M[c] + gi[1] * M[4 + c] + gi[2] * M[8 + c] + gi[3] * M[12 + c];
        for (int c = 0; c < G::K; c++) {
            T s = 0;
            for (int r = 0; r < 4; r++) s += gi[r] * J[r * G::K + c];
            dX[i * G::N + c] = s;
        }
    }
};
template <class G, typename T>
struct AsMatrixF {
    const T* X; T* M;
    VH_HD void operator()(int64_t i) const {
        Elem<T> A; G::load(X + i * G::N, A); elem_matrix4(A, M + i * 16);
    }
};
template <class G, typename T>
struct ProjectorF {
    const T* X; T* Pm;
    VH_HD void operator()(int64_t i) const {
        Elem<T> A; G::load(X + i * G::N, A); G::projector(A, Pm + i * G::N * G::N);
    }
};
template <class G, typename T>
struct JinvF {  // b = J_l^-1(log X) a
    const T* X; const T* a; T* b;
    VH_HD void operator()(int64_t i) const {
        Elem<T> A; G::load(X + i * G::N, A);
        T t[G::K], J[G::K * G::K];
        G::log(A, t);
        G::left_jac_inv(t, J);
        for (int r = 0; r < G::K; r++) {
            T s = 0;
            for (int c = 0; c < G::K; c++) s += J[r * G::K + c] * a[i * G::K + c];
            b[i * G::K + r] = s;
        }
    }
};

template <class F>
__global__ __launch_bounds__(256) void lie_kernel(F f, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) f(i);
}

template <class F>
int run(const F& f, int64_t n, int place, void* stream, const char* name) {
    if (n < 0) return fail(VH_ERR_ARG, "%s: negative element count", name);
    if (n == 0) return VH_OK;
    if (place == VH_HOST) {
        for (int64_t i = 0; i < n; i++) f(i);
        return VH_OK;
    }
    if (place != VH_DEVICE) return fail(VH_ERR_ARG, "%s: bad placement %d", name, place);
    int64_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(lie_kernel<F>, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, f, n);
    VH_LAUNCH_CHECK(name);
    return VH_OK;
}

// Dispatch group id x dtype to a functor template; unknown ids/dtypes give
// VH_ERR_UNDEFINED like the reference's undefined tensor (dispatch.h:40-50).
#define VH_LIE_DISPATCH(NAME, BUILD)                                                          \
    switch (dtype) {                                                                          \
        case VH_F32: {                                                                        \
            using T = float;                                                                  \
            switch (group) {                                                                  \
                case 1: { using G = SO3T<T>; BUILD; }                                         \
                case 2: { using G = RxSO3T<T>; BUILD; }                                       \
                case 3: { using G = SE3T<T>; BUILD; }                                         \
                case 4: { using G = Sim3T<T>; BUILD; }                                        \
                default: break;                                                               \
            }                                                                                 \
            break;                                                                            \
        }                                                                                     \
        case VH_F64: {                                                                        \
            using T = double;                                                                 \
            switch (group) {                                                                  \
                case 1: { using G = SO3T<T>; BUILD; }                                         \
                case 2: { using G = RxSO3T<T>; BUILD; }                                       \
                case 3: { using G = SE3T<T>; BUILD; }                                         \
                case 4: { using G = Sim3T<T>; BUILD; }                                        \
                default: break;                                                               \
            }                                                                                 \
            break;                                                                            \
        }                                                                                     \
        default: break;                                                                       \
    }                                                                                         \
    return fail(VH_ERR_UNDEFINED, "%s: unsupported group id %d / dtype %d", NAME, group, dtype);

}  // namespace
}  // namespace vh

using namespace vh;
using namespace vh::lie;

extern "C" {

int vh_lie_expm(int group, int dtype, int64_t n, const void* a, void* X, int place, void* stream) {
    VH_LIE_DISPATCH("expm", return run(ExpF<G, T>{(const T*)a, (T*)X}, n, place, stream, "expm"))
}
int vh_lie_expm_backward(int group, int dtype, int64_t n, const void* grad, const void* a, void* da, int place,
                         void* stream) {
    VH_LIE_DISPATCH("expm_backward",
                    return run(ExpB<G, T>{(const T*)grad, (const T*)a, (T*)da}, n, place, stream, "expm_backward"))
}
int vh_lie_logm(int group, int dtype, int64_t n, const void* X, void* a, int place, void* stream) {
    VH_LIE_DISPATCH("logm", return run(LogF<G, T>{(const T*)X, (T*)a}, n, place, stream, "logm"))
}
int vh_lie_logm_backward(int group, int dtype, int64_t n, const void* grad, const void* X, void* dX, int place,
                         void* stream) {
    VH_LIE_DISPATCH("logm_backward",
                    return run(LogB<G, T>{(const T*)grad, (const T*)X, (T*)dX}, n, place, stream, "logm_backward"))
}
int vh_lie_inv(int group, int dtype, int64_t n, const void* X, void* Y, int place, void* stream) {
    VH_LIE_DISPATCH("inv", return run(InvF<G, T>{(const T*)X, (T*)Y}, n, place, stream, "inv"))
}
int vh_lie_inv_backward(int group, int dtype, int64_t n, const void* grad, const void* X, void* dX, int place,
                        void* stream) {
    VH_LIE_DISPATCH("inv_backward",
                    return run(InvB<G, T>{(const T*)grad, (const T*)X, (T*)dX}, n, place, stream, "inv_backward"))
}
int vh_lie_mul(int group, int dtype, int64_t n, const void* X, const void* Y, void* Z, int place, void* stream) {
    VH_LIE_DISPATCH("mul", return run(MulF<G, T>{(const T*)X, (const T*)Y, (T*)Z}, n, place, stream, "mul"))
}
int vh_lie_mul_backward(int group, int dtype, int64_t n, const void* grad, const void* X, const void* Y, void* dX,
                        void* dY, int place, void* stream) {
    (void)Y;
    VH_LIE_DISPATCH("mul_backward", return run(MulB<G, T>{(const T*)grad, (const T*)X, (T*)dX, (T*)dY}, n, place,
                                               stream, "mul_backward"))
}
int vh_lie_adj(int group, int dtype, int64_t n, const void* X, const void* a, void* b, int place, void* stream) {
    VH_LIE_DISPATCH("adj", return run(AdjF<G, T>{(const T*)X, (const T*)a, (T*)b}, n, place, stream, "adj"))
}
int vh_lie_adj_backward(int group, int dtype, int64_t n, const void* grad, const void* X, const void* a, void* dX,
                        void* da, int place, void* stream) {
    VH_LIE_DISPATCH("adj_backward", return run(AdjB<G, T>{(const T*)grad, (const T*)X, (const T*)a, (T*)dX, (T*)da},
                                               n, place, stream, "adj_backward"))
}
int vh_lie_adjT(int group, int dtype, int64_t n, const void* X, const void* a, void* b, int place, void* stream) {
    VH_LIE_DISPATCH("adjT", return run(AdjTF<G, T>{(const T*)X, (const T*)a, (T*)b}, n, place, stream, "adjT"))
}
int vh_lie_adjT_backward(int group, int dtype, int64_t n, const void* grad, const void* X, const void* a, void* dX,
                         void* da, int place, void* stream) {
    VH_LIE_DISPATCH("adjT_backward", return run(AdjTB<G, T>{(const T*)grad, (const T*)X, (const T*)a, (T*)dX,
                                                            (T*)da},
                                                n, place, stream, "adjT_backward"))
}
int vh_lie_act(int group, int dtype, int64_t n, const void* X, const void* p, void* q, int place, void* stream) {
    VH_LIE_DISPATCH("act", return run(ActF<G, T>{(const T*)X, (const T*)p, (T*)q}, n, place, stream, "act"))
}
int vh_lie_act_backward(int group, int dtype, int64_t n, const void* grad, const void* X, const void* p, void* dX,
                        void* dp, int place, void* stream) {
    VH_LIE_DISPATCH("act_backward", return run(ActB<G, T>{(const T*)grad, (const T*)X, (const T*)p, (T*)dX, (T*)dp},
                                               n, place, stream, "act_backward"))
}
int vh_lie_act4(int group, int dtype, int64_t n, const void* X, const void* p, void* q, int place, void* stream) {
    VH_LIE_DISPATCH("act4", return run(Act4F<G, T>{(const T*)X, (const T*)p, (T*)q}, n, place, stream, "act4"))
}
int vh_lie_act4_backward(int group, int dtype, int64_t n, const void* grad, const void* X, const void* p, void* dX,
                         void* dp, int place, void* stream) {
    VH_LIE_DISPATCH("act4_backward", return run(Act4B<G, T>{(const T*)grad, (const T*)X, (const T*)p, (T*)dX,
                                                            (T*)dp},
                                                n, place, stream, "act4_backward"))
}
int vh_lie_as_matrix(int group, int dtype, int64_t n, const void* X, void* Tm, int place, void* stream) {
    VH_LIE_DISPATCH("as_matrix",
                    return run(AsMatrixF<G, T>{(const T*)X, (T*)Tm}, n, place, stream, "as_matrix"))
}
int vh_lie_projector(int group, int dtype, int64_t n, const void* X, void* Pm, int place, void* stream) {
    VH_LIE_DISPATCH("projector",
                    return run(ProjectorF<G, T>{(const T*)X, (T*)Pm}, n, place, stream, "projector"))
}
int vh_lie_Jinv(int group, int dtype, int64_t n, const void* X, const void* a, void* b, int place, void* stream) {
    VH_LIE_DISPATCH("Jinv", return run(JinvF<G, T>{(const T*)X, (const T*)a, (T*)b}, n, place, stream, "Jinv"))
}

}  // extern "C"
