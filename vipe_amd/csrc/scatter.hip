// scatter_ext: scatter reductions along one axis (csrc/scatter_ext/scatter.cpp,
// cuda/scatter_cuda.cu:21-129, cuda/reducer.cuh).
//
// src is viewed as [B, E, K] around the scatter axis, out as [B, N, K]; element
// (b, e, k) of src goes to out[b, index(b, e, k), k], with index read at
// b*sB + e*sE + k*sK (int64, so broadcast indices with zero strides cost
// nothing).  Indices outside [0, N) are skipped.
//
// gfx950 design:
//  * sum: native float / double / int32 / int64 atomic adds; 8- and 16-bit
//    types update their 32-bit word by compare-and-swap.
//  * mul: compare-and-swap on the element (its 32-bit word below 32 bits).
//  * min / max: out holds order-preserving unsigned keys while the reduction
//    runs (floats: sign-magnitude flipped; signed integers: sign bit flipped),
//    so the reduction is one native unsigned atomic min / max for 32- and 64-bit
//    types and the result does not depend on the order of arrival.  NaN sources
//    are skipped (as a NaN-ignoring max / min would).  A finalize pass decodes
//    the keys; without a given out, entries equal to the reduction's identity
//    become 0.  arg_out is the smallest e whose value equals the result (E if
//    none): an atomic min over e, so ties resolve deterministically.
// Integer sums and products wrap (mod 2^bits) and are exact in any order;
// float sums and products are exact up to the order of the atomics.

#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cstdint>
#include <limits>
#include <type_traits>

#include "vh_common.hpp"

namespace vh {
namespace {

enum { kSum = 0, kMul = 1, kMin = 2, kMax = 3 };

// ---- value <-> float-or-integer arithmetic for each storage type -----------
template <typename T> struct Val {
    using C = T;  // compute type
    __device__ static C get(T v) { return v; }
    __device__ static T put(C v) { return v; }
};
template <> struct Val<__half> {
    using C = float;
    __device__ static C get(__half v) { return __half2float(v); }
    __device__ static __half put(C v) { return __float2half(v); }
};
template <> struct Val<__hip_bfloat16> {
    using C = float;
    __device__ static C get(__hip_bfloat16 v) { return __bfloat162float(v); }
    __device__ static __hip_bfloat16 put(C v) { return __float2bfloat16(v); }
};
template <> struct Val<uint8_t> {
    using C = uint32_t;
    __device__ static C get(uint8_t v) { return v; }
    __device__ static uint8_t put(C v) { return (uint8_t)v; }
};
template <> struct Val<int8_t> {
    using C = uint32_t;  // wrapping arithmetic in unsigned, truncated on store
    __device__ static C get(int8_t v) { return (uint32_t)(int32_t)v; }
    __device__ static int8_t put(C v) { return (int8_t)(uint8_t)v; }
};
template <> struct Val<int16_t> {
    using C = uint32_t;
    __device__ static C get(int16_t v) { return (uint32_t)(int32_t)v; }
    __device__ static int16_t put(C v) { return (int16_t)(uint16_t)v; }
};
template <> struct Val<int32_t> {
    using C = uint32_t;
    __device__ static C get(int32_t v) { return (uint32_t)v; }
    __device__ static int32_t put(C v) { return (int32_t)v; }
};
template <> struct Val<int64_t> {
    using C = uint64_t;
    __device__ static C get(int64_t v) { return (uint64_t)v; }
    __device__ static int64_t put(C v) { return (int64_t)v; }
};

template <typename T> constexpr bool is_float_v =
    std::is_same_v<T, float> || std::is_same_v<T, double> || std::is_same_v<T, __half> ||
    std::is_same_v<T, __hip_bfloat16>;

template <int BYTES> struct UIntOf;
template <> struct UIntOf<1> { using type = uint8_t; };
template <> struct UIntOf<2> { using type = uint16_t; };
template <> struct UIntOf<4> { using type = uint32_t; };
template <> struct UIntOf<8> { using type = uint64_t; };

template <typename T> using Bits = typename UIntOf<sizeof(T)>::type;

template <typename T> __device__ __forceinline__ Bits<T> bits_of(T v) { return __builtin_bit_cast(Bits<T>, v); }
template <typename T> __device__ __forceinline__ T from_bits(Bits<T> b) { return __builtin_bit_cast(T, b); }

template <typename T> __device__ __forceinline__ bool is_nan(T v) {
    if constexpr (is_float_v<T>) {
        const auto c = Val<T>::get(v);
        return c != c;
    } else {
        return false;
    }
}

// Order-preserving map to unsigned keys.
template <typename T> __device__ __forceinline__ Bits<T> encode(T v) {
    using U = Bits<T>;
    constexpr U sign = (U)((U)1 << (sizeof(T) * 8 - 1));
    const U b = bits_of(v);
    if constexpr (is_float_v<T>) return (b & sign) ? (U)~b : (U)(b | sign);
    else if constexpr (std::is_signed_v<T>) return (U)(b ^ sign);
    else return b;
}
template <typename T> __device__ __forceinline__ T decode(Bits<T> k) {
    using U = Bits<T>;
    constexpr U sign = (U)((U)1 << (sizeof(T) * 8 - 1));
    if constexpr (is_float_v<T>) return from_bits<T>((k & sign) ? (U)(k & (U)~sign) : (U)~k);
    else if constexpr (std::is_signed_v<T>) return from_bits<T>((U)(k ^ sign));
    else return from_bits<T>(k);
}

// Compare-and-swap update of one element: new = f(old).  Elements narrower
// than 32 bits update the aligned 32-bit word that holds them.
template <typename T, typename F> __device__ __forceinline__ void cas_update(T* addr, F f) {
    if constexpr (sizeof(T) >= 4) {
        using W = std::conditional_t<sizeof(T) == 4, unsigned int, unsigned long long>;
        W* w = reinterpret_cast<W*>(addr);
        W old = __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        while (true) {
            const W nv = (W)bits_of(f(from_bits<T>((Bits<T>)old)));
            if (nv == old) return;
            const W seen = atomicCAS(w, old, nv);
            if (seen == old) return;
            old = seen;
        }
    } else {
        const uintptr_t a = reinterpret_cast<uintptr_t>(addr);
        unsigned int* w = reinterpret_cast<unsigned int*>(a & ~(uintptr_t)3);
        const unsigned shift = (unsigned)(a & 3) * 8;
        constexpr unsigned mask = sizeof(T) == 1 ? 0xffu : 0xffffu;
        unsigned int old = __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        while (true) {
            const T cur = from_bits<T>((Bits<T>)((old >> shift) & mask));
            const unsigned nb = (unsigned)bits_of(f(cur)) & mask;
            const unsigned nv = (old & ~(mask << shift)) | (nb << shift);
            if (nv == old) return;
            const unsigned seen = atomicCAS(w, old, nv);
            if (seen == old) return;
            old = seen;
        }
    }
}

template <typename T> __device__ __forceinline__ void atomic_sum(T* p, T v) {
    if constexpr (std::is_same_v<T, float> || std::is_same_v<T, double>) {
        atomicAdd(p, v);
    } else if constexpr (std::is_same_v<T, int32_t>) {
        atomicAdd(reinterpret_cast<unsigned int*>(p), (unsigned int)v);
    } else if constexpr (std::is_same_v<T, int64_t>) {
        atomicAdd(reinterpret_cast<unsigned long long*>(p), (unsigned long long)v);
    } else {
        const auto c = Val<T>::get(v);
        cas_update(p, [c](T o) { return Val<T>::put(Val<T>::get(o) + c); });
    }
}

template <typename T> __device__ __forceinline__ void atomic_mul(T* p, T v) {
    const auto c = Val<T>::get(v);
    cas_update(p, [c](T o) { return Val<T>::put(Val<T>::get(o) * c); });
}

// min / max on keys already stored at p
template <typename T, bool MAX> __device__ __forceinline__ void atomic_key(T* p, Bits<T> key) {
    if constexpr (sizeof(T) == 4) {
        unsigned int* w = reinterpret_cast<unsigned int*>(p);
        if (MAX) atomicMax(w, (unsigned int)key); else atomicMin(w, (unsigned int)key);
    } else if constexpr (sizeof(T) == 8) {
        unsigned long long* w = reinterpret_cast<unsigned long long*>(p);
        if (MAX) atomicMax(w, (unsigned long long)key); else atomicMin(w, (unsigned long long)key);
    } else {
        using U = Bits<T>;
        cas_update(reinterpret_cast<U*>(p), [key](U o) { return MAX ? (o > key ? o : key) : (o < key ? o : key); });
    }
}

template <typename T, int OP> __device__ __forceinline__ T identity() {
    if constexpr (OP == kSum) return Val<T>::put(0);
    else if constexpr (OP == kMul) return Val<T>::put(1);
    else if constexpr (std::is_same_v<T, __half>) return from_bits<__half>(OP == kMin ? 0x7bff : 0xfbff);
    else if constexpr (std::is_same_v<T, __hip_bfloat16>) return from_bits<__hip_bfloat16>(OP == kMin ? 0x7f7f : 0xff7f);
    else return OP == kMin ? std::numeric_limits<T>::max() : std::numeric_limits<T>::lowest();
}

struct Geom {
    int64_t B, E, K, N;   // src [B,E,K], out [B,N,K]
    int64_t sB, sE, sK;   // index strides (elements)
};

// out (and arg_out) before the reduction: identity / keys / E.
template <typename T, int OP>
__global__ __launch_bounds__(256) void scatter_prepare_kernel(T* __restrict__ out, int64_t* __restrict__ arg,
                                                              int64_t n_out, int64_t E, int out_given) {
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n_out; t += (int64_t)gridDim.x * blockDim.x) {
        if constexpr (OP == kMin || OP == kMax) {
            T v = out_given ? out[t] : identity<T, OP>();
            Bits<T> k = encode(v);
            // a NaN already in out loses to every source, as in a NaN-ignoring min / max
            if (is_nan(v)) k = OP == kMax ? (Bits<T>)0 : (Bits<T>)~(Bits<T>)0;
            out[t] = from_bits<T>(k);
            arg[t] = E;
        } else if (!out_given) {
            out[t] = identity<T, OP>();
        }
    }
}

template <typename IDX>
__device__ __forceinline__ bool locate(const int64_t* __restrict__ index, const Geom& g, int64_t t, int64_t& e,
                                       int64_t& o) {
    const IDX K = (IDX)g.K, E = (IDX)g.E;
    const IDX tt = (IDX)t;
    const IDX k = tt % K;
    const IDX be = tt / K;
    const IDX ee = be % E;
    const IDX b = be / E;
    const int64_t n = index[(int64_t)b * g.sB + (int64_t)ee * g.sE + (int64_t)k * g.sK];
    e = ee;
    if (n < 0 || n >= g.N) return false;
    o = ((int64_t)b * g.N + n) * g.K + k;
    return true;
}

template <typename T, int OP, typename IDX>
__global__ __launch_bounds__(256) void scatter_reduce_kernel(const T* __restrict__ src, const int64_t* __restrict__ index,
                                                             T* __restrict__ out, Geom g, int64_t n_src) {
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n_src; t += (int64_t)gridDim.x * blockDim.x) {
        int64_t e, o;
        if (!locate<IDX>(index, g, t, e, o)) continue;
        const T v = src[t];
        if constexpr (OP == kSum) atomic_sum(out + o, v);
        else if constexpr (OP == kMul) atomic_mul(out + o, v);
        else {
            if (is_nan(v)) continue;
            atomic_key<T, OP == kMax>(out + o, encode(v));
        }
    }
}

template <typename T, int OP>
__global__ __launch_bounds__(256) void scatter_finalize_kernel(T* __restrict__ out, int64_t n_out, int out_given) {
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n_out; t += (int64_t)gridDim.x * blockDim.x) {
        T v = decode<T>(bits_of(out[t]));
        if (!out_given && bits_of(v) == bits_of(identity<T, OP>())) v = Val<T>::put(0);
        out[t] = v;
    }
}

template <typename T, typename IDX>
__global__ __launch_bounds__(256) void scatter_arg_kernel(const T* __restrict__ src, const int64_t* __restrict__ index,
                                                          const T* __restrict__ out, int64_t* __restrict__ arg, Geom g,
                                                          int64_t n_src) {
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n_src; t += (int64_t)gridDim.x * blockDim.x) {
        int64_t e, o;
        if (!locate<IDX>(index, g, t, e, o)) continue;
        if (Val<T>::get(src[t]) == Val<T>::get(out[o]))
            atomicMin(reinterpret_cast<unsigned long long*>(arg + o), (unsigned long long)e);
    }
}

unsigned grid_for(int64_t n) {
    const int64_t blocks = (n + 255) / 256;
    return (unsigned)(blocks < 65536 ? (blocks > 0 ? blocks : 1) : 65536);
}

template <typename T, int OP>
int run_op(const void* src, const int64_t* index, const Geom& g, void* out, int64_t* arg, int out_given,
           hipStream_t s) {
    const int64_t n_src = g.B * g.E * g.K, n_out = g.B * g.N * g.K;
    T* o = (T*)out;
    if (n_out > 0 && (OP == kMin || OP == kMax || !out_given)) {
        hipLaunchKernelGGL((scatter_prepare_kernel<T, OP>), dim3(grid_for(n_out)), dim3(256), 0, s, o, arg, n_out,
                           g.E, out_given);
        VH_LAUNCH_CHECK("scatter prepare");
    }
    if (n_src > 0 && n_out > 0) {
        if (n_src < ((int64_t)1 << 31))
            hipLaunchKernelGGL((scatter_reduce_kernel<T, OP, uint32_t>), dim3(grid_for(n_src)), dim3(256), 0, s,
                               (const T*)src, index, o, g, n_src);
        else
            hipLaunchKernelGGL((scatter_reduce_kernel<T, OP, int64_t>), dim3(grid_for(n_src)), dim3(256), 0, s,
                               (const T*)src, index, o, g, n_src);
        VH_LAUNCH_CHECK("scatter reduce");
    }
    if constexpr (OP == kMin || OP == kMax) {
        if (n_out > 0) {
            hipLaunchKernelGGL((scatter_finalize_kernel<T, OP>), dim3(grid_for(n_out)), dim3(256), 0, s, o, n_out,
                               out_given);
            VH_LAUNCH_CHECK("scatter finalize");
        }
        if (n_src > 0 && n_out > 0) {
            if (n_src < ((int64_t)1 << 31))
                hipLaunchKernelGGL((scatter_arg_kernel<T, uint32_t>), dim3(grid_for(n_src)), dim3(256), 0, s,
                                   (const T*)src, index, (const T*)o, arg, g, n_src);
            else
                hipLaunchKernelGGL((scatter_arg_kernel<T, int64_t>), dim3(grid_for(n_src)), dim3(256), 0, s,
                                   (const T*)src, index, (const T*)o, arg, g, n_src);
            VH_LAUNCH_CHECK("scatter arg");
        }
    }
    return VH_OK;
}

template <typename T>
int run_typed(int reduce, const void* src, const int64_t* index, const Geom& g, void* out, int64_t* arg,
              int out_given, hipStream_t s) {
    switch (reduce) {
        case kSum: return run_op<T, kSum>(src, index, g, out, arg, out_given, s);
        case kMul: return run_op<T, kMul>(src, index, g, out, arg, out_given, s);
        case kMin: return run_op<T, kMin>(src, index, g, out, arg, out_given, s);
        case kMax: return run_op<T, kMax>(src, index, g, out, arg, out_given, s);
        default: return fail(VH_ERR_ARG, "scatter: unknown reduction %d", reduce);
    }
}

}  // namespace
}  // namespace vh

using namespace vh;

extern "C" int vh_scatter(int reduce, int dtype, const void* src, const int64_t* index, int64_t B, int64_t E,
                          int64_t K, int64_t sB, int64_t sE, int64_t sK, int64_t N, void* out, int64_t* arg_out,
                          int out_given, void* stream) {
    if (B < 0 || E < 0 || K < 0 || N < 0) return fail(VH_ERR_ARG, "scatter: negative size");
    if (reduce < kSum || reduce > kMax) return fail(VH_ERR_ARG, "scatter: unknown reduction %d", reduce);
    const int64_t n_src = B * E * K, n_out = B * N * K;
    if ((n_out > 0 && !out) || (n_src > 0 && (!src || !index)) ||
        ((reduce == kMin || reduce == kMax) && n_out > 0 && !arg_out))
        return fail(VH_ERR_ARG, "scatter: null pointer");
    const Geom g{B, E, K, N, sB, sE, sK};
    hipStream_t s = (hipStream_t)stream;
    switch (dtype) {
        case VH_F32: return run_typed<float>(reduce, src, index, g, out, arg_out, out_given, s);
        case VH_F64: return run_typed<double>(reduce, src, index, g, out, arg_out, out_given, s);
        case VH_F16: return run_typed<__half>(reduce, src, index, g, out, arg_out, out_given, s);
        case VH_BF16: return run_typed<__hip_bfloat16>(reduce, src, index, g, out, arg_out, out_given, s);
        case VH_U8: return run_typed<uint8_t>(reduce, src, index, g, out, arg_out, out_given, s);
        case VH_I8: return run_typed<int8_t>(reduce, src, index, g, out, arg_out, out_given, s);
        case VH_I16: return run_typed<int16_t>(reduce, src, index, g, out, arg_out, out_given, s);
        case VH_I32: return run_typed<int32_t>(reduce, src, index, g, out, arg_out, out_given, s);
        case VH_I64: return run_typed<int64_t>(reduce, src, index, g, out, arg_out, out_given, s);
        default: return fail(VH_ERR_UNSUPPORTED, "scatter: dtype %d", dtype);
    }
}
