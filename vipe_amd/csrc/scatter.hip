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
//
// When the index depends on e alone (a 1-D index broadcast over the other
// axes, the common case) and K >= 32, a segmented path replaces the atomics:
// see scatter_segment_sort_kernel / scatter_segment_gather_kernel.

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

// ---- segmented path: index depends on e only --------------------------------
// A stable counting sort of the E indices gives, per output slot n, the list of
// source positions e in ascending order (start[n] .. start[n+1] into order[]).
// The gather kernel then reduces each output element over its list in that
// order: no atomics, each src element read once, each out element written once,
// the same result on every run, and 8 / 16-bit floats reduced in FP32 with one
// final rounding.
constexpr int kSegMaxN = 12000;   // LDS counters (47 KiB)
constexpr int kSegIdxCache = 4096;   // indices kept in LDS for the placement pass (16 KiB)
constexpr int kSegParE = 1024;       // up to this many sources, every thread places (O(E^2 / 256) reads)
constexpr int kSegMaxE = 1 << 16;   // the placement pass is one wave
constexpr int kSegMinK = 32;      // lanes along k share one list

__global__ __launch_bounds__(256) void scatter_segment_sort_kernel(const int64_t* __restrict__ index, int64_t sE,
                                                                   int E, int N, int* __restrict__ start,
                                                                   int* __restrict__ order) {
    __shared__ int cnt[kSegMaxN];
    __shared__ int sidx[kSegIdxCache];
    __shared__ int tsum[4];
    const int tid = threadIdx.x;
    for (int n = tid; n < N; n += 256) cnt[n] = 0;
    __syncthreads();
    for (int e = tid; e < E; e += 256) {
        const int64_t n = index[(int64_t)e * sE];
        const bool ok = n >= 0 && n < N;
        if (ok) atomicAdd(&cnt[n], 1);
        if (e < kSegIdxCache) sidx[e] = ok ? (int)n : -1;
    }
    __syncthreads();
    // exclusive scan: each thread sums a contiguous chunk; the 256 chunk sums
    // are scanned within each wave by shuffles, then across the four waves
    const int per = (N + 255) / 256;
    const int lo = min(N, tid * per), hi = min(N, lo + per);
    int sum = 0;
    for (int n = lo; n < hi; n++) sum += cnt[n];
    const int lane = tid & 63, wv = tid >> 6;
    int inc = sum;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int v = __shfl_up(inc, d, 64);
        if (lane >= d) inc += v;
    }
    if (lane == 63) tsum[wv] = inc;
    __syncthreads();
    int wbase = 0;
    for (int w = 0; w < wv; w++) wbase += tsum[w];
    if (tid == 255) start[N] = wbase + inc;
    int run = wbase + inc - sum;
    for (int n = lo; n < hi; n++) {
        const int c = cnt[n];
        start[n] = run;
        cnt[n] = run;  // becomes the placement cursor
        run += c;
    }
    __syncthreads();
    if (E <= kSegParE) {
        // all threads: the rank of e in its bucket is the number of earlier
        // positions with the same index (one broadcast LDS read per position)
        for (int e = tid; e < E; e += 256) {
            const int n = sidx[e];
            if (n < 0) continue;
            int rank = 0;
#pragma unroll 8
            for (int j = 0; j < e; j++) rank += sidx[j] == n;
            order[cnt[n] + rank] = e;   // cnt holds the bucket's start here
        }
        return;
    }
    // stable placement by one wave, 64 positions at a time: a lane's slot is its
    // bucket's cursor plus the number of lower lanes in the same bucket; the
    // highest such lane advances the cursor.  One wave's LDS reads and writes
    // complete in order.
    if (tid < 64) {
        for (int base = 0; base < E; base += 64) {
            const int e = base + lane;
            int n = -1;
            if (e < E && e < kSegIdxCache) {
                n = sidx[e];
            } else if (e < E) {
                const int64_t nn = index[(int64_t)e * sE];
                n = (nn >= 0 && nn < N) ? (int)nn : -1;
            }
            int rank = 0;
            bool higher = false;
            for (int j = 0; j < 64; j++) {
                const int nj = __shfl(n, j, 64);
                if (nj == n) {
                    rank += j < lane;
                    higher |= j > lane;
                }
            }
            if (n >= 0) {
                const int pos = cnt[n] + rank;
                order[pos] = e;
                if (!higher) cnt[n] = pos + 1;
            }
        }
    }
}

template <typename T, int OP> struct SegAcc {
    using Cv = typename Val<T>::C;
    Cv acc;
    __device__ void init(T v0) { acc = Val<T>::get(v0); }
    __device__ void add(T v, int) {
        if constexpr (OP == kSum) acc = acc + Val<T>::get(v); else acc = acc * Val<T>::get(v);
    }
    __device__ T finish(int, int64_t&, int64_t) const { return Val<T>::put(acc); }
};
// min / max on order-preserving keys: the same result as the atomic path; the
// arg is the smallest e whose value equals the running extremum (e ascends).
template <typename T, int OP> struct SegArg {
    using U = Bits<T>;
    static constexpr bool MAX = OP == kMax;
    U best;
    int64_t a;
    __device__ void init(T v0) {
        best = encode(v0);
        if (is_nan(v0)) best = MAX ? (U)0 : (U)~(U)0;
        a = -1;
    }
    __device__ void add(T v, int e) {
        if (is_nan(v)) return;
        const U key = encode(v);
        const bool beats = MAX ? key > best : key < best;
        const bool same = Val<T>::get(v) == Val<T>::get(decode<T>(best));
        if (beats) {
            if (!same || a < 0) a = e;
            best = key;
        } else if (same && a < 0) {
            a = e;
        }
    }
    __device__ T finish(int out_given, int64_t& arg, int64_t E) const {
        T r = decode<T>(best);
        arg = a < 0 ? E : a;
        if (!out_given && bits_of(r) == bits_of(identity<T, OP>())) {
            r = Val<T>::put(0);
            arg = E;
        }
        return r;
    }
};

template <typename T, int VEC> struct alignas(sizeof(T) * VEC) VecT {
    T v[VEC];
};

// One lane per VEC consecutive output elements along k; the segment's source
// rows are read four positions at a time so that four vector loads are in
// flight per lane.  VEC > 1 needs K % VEC == 0 and src / out aligned to
// VEC * sizeof(T) (checked at launch).
template <typename T, int OP, int VEC>
__global__ __launch_bounds__(256) void scatter_segment_gather_kernel(const T* __restrict__ src,
                                                                     const int* __restrict__ start,
                                                                     const int* __restrict__ order,
                                                                     T* __restrict__ out, int64_t* __restrict__ arg,
                                                                     Geom g, int out_given) {
    using State = std::conditional_t<OP == kMin || OP == kMax, SegArg<T, OP>, SegAcc<T, OP>>;
    using V = VecT<T, VEC>;
    const int64_t KV = g.K / VEC;
    const int64_t n_vec = g.B * g.N * KV;
    for (int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; o < n_vec; o += (int64_t)gridDim.x * blockDim.x) {
        const int64_t kv = o % KV;
        const int64_t bn = o / KV;
        const int n = (int)(bn % g.N);
        const int64_t b = bn / g.N;
        const T* sp = src + b * g.E * g.K + kv * VEC;
        const int64_t ob = bn * g.K + kv * VEC;
        const int s0 = start[n], s1 = start[n + 1];
        State st[VEC];
        if (out_given) {
            const V o0 = *reinterpret_cast<const V*>(out + ob);
#pragma unroll
            for (int i = 0; i < VEC; i++) st[i].init(o0.v[i]);
        } else {
#pragma unroll
            for (int i = 0; i < VEC; i++) st[i].init(identity<T, OP>());
        }
        int j = s0;
        for (; j + 4 <= s1; j += 4) {
            int e[4];
            V v[4];
#pragma unroll
            for (int u = 0; u < 4; u++) e[u] = order[j + u];
#pragma unroll
            for (int u = 0; u < 4; u++) v[u] = *reinterpret_cast<const V*>(sp + (int64_t)e[u] * g.K);
#pragma unroll
            for (int u = 0; u < 4; u++)
#pragma unroll
                for (int i = 0; i < VEC; i++) st[i].add(v[u].v[i], e[u]);
        }
        for (; j < s1; j++) {
            const int e = order[j];
            const V v = *reinterpret_cast<const V*>(sp + (int64_t)e * g.K);
#pragma unroll
            for (int i = 0; i < VEC; i++) st[i].add(v.v[i], e);
        }
        V r;
        int64_t a[VEC];
#pragma unroll
        for (int i = 0; i < VEC; i++) r.v[i] = st[i].finish(out_given, a[i], g.E);
        *reinterpret_cast<V*>(out + ob) = r;
        if constexpr (OP == kMin || OP == kMax) {
#pragma unroll
            for (int i = 0; i < VEC; i++) arg[ob + i] = a[i];
        }
    }
}

bool segment_path(const Geom& g) {
    return (g.B == 1 || g.sB == 0) && (g.K == 1 || g.sK == 0) && g.K >= kSegMinK && g.N >= 1 && g.N <= kSegMaxN &&
           g.E >= 1 && g.E <= kSegMaxE;
}

size_t segment_workspace(const Geom& g) {
    return segment_path(g) ? (size_t)(g.N + 1 + g.E) * sizeof(int) : 0;
}

template <typename T, int OP>
int run_segmented(const void* src, const int64_t* index, const Geom& g, void* out, int64_t* arg, int out_given,
                  void* ws, hipStream_t s) {
    int* start = (int*)ws;
    int* order = start + g.N + 1;
    hipLaunchKernelGGL(scatter_segment_sort_kernel, dim3(1), dim3(256), 0, s, index, g.sE, (int)g.E, (int)g.N,
                       start, order);
    VH_LAUNCH_CHECK("scatter segment sort");
    constexpr int VEC = 16 / sizeof(T) < 8 ? 16 / sizeof(T) : 8;   // 16-byte loads, at most 8 lanes of state
    const bool vec_ok = g.K % VEC == 0 && reinterpret_cast<uintptr_t>(src) % (VEC * sizeof(T)) == 0 &&
                        reinterpret_cast<uintptr_t>(out) % (VEC * sizeof(T)) == 0;
    if (vec_ok)
        hipLaunchKernelGGL((scatter_segment_gather_kernel<T, OP, VEC>), dim3(grid_for(g.B * g.N * (g.K / VEC))),
                           dim3(256), 0, s, (const T*)src, start, order, (T*)out, arg, g, out_given);
    else
        hipLaunchKernelGGL((scatter_segment_gather_kernel<T, OP, 1>), dim3(grid_for(g.B * g.N * g.K)), dim3(256), 0,
                           s, (const T*)src, start, order, (T*)out, arg, g, out_given);
    VH_LAUNCH_CHECK("scatter segment gather");
    return VH_OK;
}

template <typename T, int OP>
int run_op(const void* src, const int64_t* index, const Geom& g, void* out, int64_t* arg, int out_given, void* ws,
           size_t ws_bytes, hipStream_t s) {
    const int64_t n_src = g.B * g.E * g.K, n_out = g.B * g.N * g.K;
    if (n_src > 0 && n_out > 0 && segment_path(g) && ws && ws_bytes >= segment_workspace(g))
        return run_segmented<T, OP>(src, index, g, out, arg, out_given, ws, s);
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
              int out_given, void* ws, size_t ws_bytes, hipStream_t s) {
    switch (reduce) {
        case kSum: return run_op<T, kSum>(src, index, g, out, arg, out_given, ws, ws_bytes, s);
        case kMul: return run_op<T, kMul>(src, index, g, out, arg, out_given, ws, ws_bytes, s);
        case kMin: return run_op<T, kMin>(src, index, g, out, arg, out_given, ws, ws_bytes, s);
        case kMax: return run_op<T, kMax>(src, index, g, out, arg, out_given, ws, ws_bytes, s);
        default: return fail(VH_ERR_ARG, "scatter: unknown reduction %d", reduce);
    }
}

}  // namespace
}  // namespace vh

using namespace vh;

extern "C" size_t vh_scatter_workspace_size(int64_t B, int64_t E, int64_t K, int64_t sB, int64_t sE, int64_t sK,
                                             int64_t N) {
    if (B < 0 || E < 0 || K < 0 || N < 0) return 0;
    return segment_workspace(Geom{B, E, K, N, sB, sE, sK});
}

extern "C" int vh_scatter(int reduce, int dtype, const void* src, const int64_t* index, int64_t B, int64_t E,
                          int64_t K, int64_t sB, int64_t sE, int64_t sK, int64_t N, void* out, int64_t* arg_out,
                          int out_given, void* workspace, size_t workspace_bytes, void* stream) {
    if (B < 0 || E < 0 || K < 0 || N < 0) return fail(VH_ERR_ARG, "scatter: negative size");
    if (reduce < kSum || reduce > kMax) return fail(VH_ERR_ARG, "scatter: unknown reduction %d", reduce);
    const int64_t n_src = B * E * K, n_out = B * N * K;
    if ((n_out > 0 && !out) || (n_src > 0 && (!src || !index)) ||
        ((reduce == kMin || reduce == kMax) && n_out > 0 && !arg_out))
        return fail(VH_ERR_ARG, "scatter: null pointer");
    const Geom g{B, E, K, N, sB, sE, sK};
    hipStream_t s = (hipStream_t)stream;
    switch (dtype) {
        case VH_F32: return run_typed<float>(reduce, src, index, g, out, arg_out, out_given, workspace, workspace_bytes, s);
        case VH_F64: return run_typed<double>(reduce, src, index, g, out, arg_out, out_given, workspace, workspace_bytes, s);
        case VH_F16: return run_typed<__half>(reduce, src, index, g, out, arg_out, out_given, workspace, workspace_bytes, s);
        case VH_BF16: return run_typed<__hip_bfloat16>(reduce, src, index, g, out, arg_out, out_given, workspace, workspace_bytes, s);
        case VH_U8: return run_typed<uint8_t>(reduce, src, index, g, out, arg_out, out_given, workspace, workspace_bytes, s);
        case VH_I8: return run_typed<int8_t>(reduce, src, index, g, out, arg_out, out_given, workspace, workspace_bytes, s);
        case VH_I16: return run_typed<int16_t>(reduce, src, index, g, out, arg_out, out_given, workspace, workspace_bytes, s);
        case VH_I32: return run_typed<int32_t>(reduce, src, index, g, out, arg_out, out_given, workspace, workspace_bytes, s);
        case VH_I64: return run_typed<int64_t>(reduce, src, index, g, out, arg_out, out_given, workspace, workspace_bytes, s);
        default: return fail(VH_ERR_UNSUPPORTED, "scatter: dtype %d", dtype);
    }
}
