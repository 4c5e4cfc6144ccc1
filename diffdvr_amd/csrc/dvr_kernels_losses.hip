// Elementwise losses (renderer/losses_impl.cu:35-240, formulas of
// renderer/renderer_utils.cuh:180-298), float32 and float64:
//   mul_log(t)        = t log2 t, 0 below FLT_EPSILON
//   log_mse(x, y)     = log((e^x - e^y)^2) = 2 (log(1 - e^(b - a)) + a), a = max, b = min
//   log_l1(x, y)      = log|e^x - e^y|     =    log(1 - e^(b - a)) + a
// and their derivatives.  log(1 - e^u) for u <= 0 is evaluated as
// log(-expm1(u)) for u > -ln 2 and log1p(-e^u) otherwise (Maechler, "Accurately
// computing log(1 - exp(-|a|))", 2012), and 1 - e^u in the derivatives as
// -expm1(u): the same functions as the reference's log(1 - exp(u)) without its
// cancellation when x and y are close.
//
// Memory-bound streams: each lane moves 16 bytes per array (float4 / double2)
// when every pointer is 16-byte aligned, one vector per lane up to 2^20
// workgroups; outputs are written with non-temporal stores.
#include <hip/hip_runtime.h>

#include "dvr_launch.hpp"

namespace dvr {

namespace {

constexpr int kEwBlock = 256;

__device__ __forceinline__ float log2_(float v) { return log2f(v); }
__device__ __forceinline__ double log2_(double v) { return log2(v); }
__device__ __forceinline__ float expm1_(float v) { return expm1f(v); }
__device__ __forceinline__ double expm1_(double v) { return expm1(v); }

template <class T>
__device__ __forceinline__ T log1mexp(T u)
{
    return u > T(-0.693147180559945309417) ? log(-expm1(u)) : log1p(-exp(u));
}
__device__ __forceinline__ float log1mexp(float u)
{
    return u > -0.693147180559945309417f ? logf(-expm1f(u)) : log1pf(-expf(u));
}

template <int OP, class T>
__device__ __forceinline__ void ew_op(T a, T b, T g, T& o1, T& o2)
{
    constexpr T eps = T(1.192092896e-07);            // FLT_EPSILON for both types (losses_impl.cu:28-30,46)
    constexpr T invLn2 = T(1.44269504088896340736);  // 1 / ln 2
    if constexpr (OP == 0) {
        o1 = a < eps ? T(0) : a * log2_(a);
    } else if constexpr (OP == 1) {
        o1 = a < eps ? T(0) : g * (log2_(a) + invLn2);
    } else if constexpr (OP == 2 || OP == 4) {
        const T hi = a > b ? a : b, lo = a > b ? b : a;
        const T l = log1mexp(lo - hi) + hi;
        o1 = OP == 2 ? T(2) * l : l;
    } else {
        const T k = OP == 3 ? T(2) * g : g;
        o1 = k / -expm1_(b - a);
        o2 = k / -expm1_(a - b);
    }
}

template <int OP> constexpr bool uses_b() { return OP >= 2; }
template <int OP> constexpr bool uses_g() { return OP == 1 || OP == 3 || OP == 5; }
template <int OP> constexpr bool has_o2() { return OP == 3 || OP == 5; }

template <class T, int V> struct VecOf {
    typedef T type __attribute__((ext_vector_type(V)));
};
template <class T> struct VecOf<T, 1> {
    typedef T type;
};

template <int OP, class T, int V>
__global__ __launch_bounds__(kEwBlock) void k_elementwise(ElementwiseParams P)
{
    typedef typename VecOf<T, V>::type vec;
    const T* A = static_cast<const T*>(P.a);
    const T* Bp = static_cast<const T*>(P.b);
    const T* G = static_cast<const T*>(P.g);
    T* O1 = static_cast<T*>(P.o1);
    T* O2 = static_cast<T*>(P.o2);
    const long long nv = V > 1 ? P.n / V : 0;
    const long long stride = (long long)gridDim.x * kEwBlock;
    if constexpr (V == 1) {
        for (long long j = (long long)blockIdx.x * kEwBlock + threadIdx.x; j < P.n; j += stride) {
            T r1, r2 = T(0);
            ew_op<OP, T>(A[j], uses_b<OP>() ? Bp[j] : T(0), uses_g<OP>() ? G[j] : T(0), r1, r2);
            O1[j] = r1;
            if constexpr (has_o2<OP>()) O2[j] = r2;
        }
    } else {
        for (long long j = (long long)blockIdx.x * kEwBlock + threadIdx.x; j < nv; j += stride) {
            const vec a = reinterpret_cast<const vec*>(A)[j];
            vec b = a, g = a, o1, o2;
            if constexpr (uses_b<OP>()) b = reinterpret_cast<const vec*>(Bp)[j];
            if constexpr (uses_g<OP>()) g = reinterpret_cast<const vec*>(G)[j];
#pragma unroll
            for (int k = 0; k < V; ++k) {
                T r1, r2 = T(0);
                ew_op<OP, T>(a[k], b[k], g[k], r1, r2);
                o1[k] = r1;
                o2[k] = r2;
            }
            __builtin_nontemporal_store(o1, reinterpret_cast<vec*>(O1) + j);
            if constexpr (has_o2<OP>()) __builtin_nontemporal_store(o2, reinterpret_cast<vec*>(O2) + j);
        }
        if (blockIdx.x == 0) {  // the last n mod V elements
            for (long long j = nv * V + threadIdx.x; j < P.n; j += kEwBlock) {
                T r1, r2 = T(0);
                ew_op<OP, T>(A[j], uses_b<OP>() ? Bp[j] : T(0), uses_g<OP>() ? G[j] : T(0), r1, r2);
                O1[j] = r1;
                if constexpr (has_o2<OP>()) O2[j] = r2;
            }
        }
    }
}

bool aligned16(const ElementwiseParams& P)
{
    const uintptr_t m = reinterpret_cast<uintptr_t>(P.a) | reinterpret_cast<uintptr_t>(P.b) |
                        reinterpret_cast<uintptr_t>(P.g) | reinterpret_cast<uintptr_t>(P.o1) |
                        reinterpret_cast<uintptr_t>(P.o2);
    return (m & 15u) == 0;
}

template <int OP, class T>
hipError_t launch_op(const ElementwiseParams& P, hipStream_t s)
{
    constexpr int V = 16 / sizeof(T);
    const bool vec = aligned16(P);
    const long long work = vec ? P.n / V : P.n;
    long long blocks = (work + kEwBlock - 1) / kEwBlock;
    blocks = blocks < 1 ? 1 : (blocks > (1 << 20) ? (1 << 20) : blocks);
    if (vec) hipLaunchKernelGGL((k_elementwise<OP, T, V>), dim3(unsigned(blocks)), dim3(kEwBlock), 0, s, P);
    else hipLaunchKernelGGL((k_elementwise<OP, T, 1>), dim3(unsigned(blocks)), dim3(kEwBlock), 0, s, P);
    return hipGetLastError();
}

template <class T>
hipError_t launch_typed(const ElementwiseParams& P, int op, hipStream_t s)
{
    switch (op) {
    case 0: return launch_op<0, T>(P, s);
    case 1: return launch_op<1, T>(P, s);
    case 2: return launch_op<2, T>(P, s);
    case 3: return launch_op<3, T>(P, s);
    case 4: return launch_op<4, T>(P, s);
    case 5: return launch_op<5, T>(P, s);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace

hipError_t launch_elementwise(const ElementwiseParams& P, int op, bool dbl, hipStream_t s)
{
    if (P.n == 0) return hipSuccess;
    return dbl ? launch_typed<double>(P, op, s) : launch_typed<float>(P, op, s);
}

}  // namespace dvr
