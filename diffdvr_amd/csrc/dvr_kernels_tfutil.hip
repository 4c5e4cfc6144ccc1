// TF utilities on the GPU: pre-shading a density volume through a TF and the
// adjoint of that w.r.t. the density (renderer/tf_utils_impl.cu:17-190), the
// best-fit inversion of a colour volume by random search (:225-520), and
// importance sampling of volume positions (renderer/losses_impl.cu:254-421).
//
// The voxel-wise kernels give one voxel to a lane and stride over the
// X*Y*Z grid in the memory order of the main input (VoxelGrid); the TF table
// is staged in LDS once per workgroup.  GaussianLog (renderer_tf.cuh:653-814)
// stages a transformed table: log colours and 1/(2 sigma^2).
//
// Random numbers.  The reference seeds one 48-bit linear congruential stream
// per thread from 1024 random seeds and lets grid-stride iterations continue
// it, so its draws depend on the launch shape.  Here each voxel (best fit) and
// each sample (importance sampling) owns a stream derived from the caller's
// 64-bit seed and its logical index with SplitMix64 (Steele, Lea and Flood,
// OOPSLA 2014), so results do not depend on the launch.  Best-fit candidates
// then follow the reference's generator: the 48-bit LCG with multiplier
// 0x5DEECE66D and increment 11 (java.util.Random), 24 high bits per float.
#include <algorithm>

#include "dvr_device.hpp"
#include "dvr_launch.hpp"

namespace dvr {

namespace {

constexpr int kAuxBlock = 256;
constexpr unsigned long long kLcgMul = 0x5DEECE66Dull, kLcgAdd = 0xBull, kLcgMask = (1ull << 48) - 1ull;

__device__ __forceinline__ unsigned long long splitmix64(unsigned long long x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

__device__ __forceinline__ unsigned fast_div(unsigned n, FastDiv d) { return (__umulhi(n, d.m) + n) >> d.l; }

// logical (x, y, z) of flat index i
__device__ __forceinline__ void voxel_xyz(const VoxelGrid& G, int i, int& x, int& y, int& z)
{
    const unsigned q0 = fast_div(unsigned(i), G.div0);
    const unsigned q1 = fast_div(q0, G.div1);
    const int i0 = i - int(q0) * G.n[0];
    const int i1 = int(q0) - int(q1) * G.n[1];
    const int i2 = int(q1);
    x = G.axis[0] == 0 ? i0 : (G.axis[1] == 0 ? i1 : i2);
    y = G.axis[0] == 1 ? i0 : (G.axis[1] == 1 ? i1 : i2);
    z = G.axis[0] == 2 ? i0 : (G.axis[1] == 2 ? i1 : i2);
}

__device__ __forceinline__ int off3(const Strides3& s, int x, int y, int z) { return x * s.x + y * s.y + z * s.z; }

// Copies the TF into LDS as a dense R x C table; GaussianLog rows become
// (log r, log g, log b, log a, mu, 1/(2 sigma^2)).  Every thread of the block
// must call it.
template <int TFM>
__device__ __forceinline__ Tf stage_tf(const TfTable& t, float* lds)
{
    constexpr int C = TFM == 6 ? 6 : tf_cols<TFM>();  // the row stride is a compile-time constant
    const int n = t.R * C;
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        const int r = k / C, c = k - r * C;
        float v = t.p[r * t.sr + c * t.sc];
        if constexpr (TFM == 6) {
            if (c < 4) v = logf(v);
            else if (c == 5) v = 1.0f / (2.0f * v * v);
        }
        lds[k] = v;
    }
    __syncthreads();
    return Tf{lds, t.R, C, C, 1};
}

// GaussianLog: per channel log-sum-exp over i of log(c_i) - (d - mu_i)^2 / (2 sigma_i^2)
// (renderer_tf.cuh:676-688 with renderer_utils.cuh:165-178: maximum first, then the sum).
__device__ __forceinline__ f4 glog_max(const Tf& T, float d)
{
    const float ninf = -__builtin_inff();
    f4 m = mk4(ninf, ninf, ninf, ninf);
    for (int i = 0; i < T.R; ++i) {
        const float* r = T.p + i * T.sr;
        const float u = d - r[4];
        const float ln = -(u * u) * r[5];
        m.x = fmaxf(m.x, r[0] + ln); m.y = fmaxf(m.y, r[1] + ln);
        m.z = fmaxf(m.z, r[2] + ln); m.w = fmaxf(m.w, r[3] + ln);
    }
    return m;
}

__device__ __forceinline__ f4 glog_sum(const Tf& T, float d, const f4& m)
{
    f4 s = mk4(0.f, 0.f, 0.f, 0.f);
    for (int i = 0; i < T.R; ++i) {
        const float* r = T.p + i * T.sr;
        const float u = d - r[4];
        const float ln = -(u * u) * r[5];
        s.x += expf(r[0] + ln - m.x); s.y += expf(r[1] + ln - m.y);
        s.z += expf(r[2] + ln - m.z); s.w += expf(r[3] + ln - m.w);
    }
    return s;
}

__device__ __forceinline__ f4 glog_eval(const Tf& T, float d)
{
    const f4 m = glog_max(T, d);
    const f4 s = glog_sum(T, d, m);
    return mk4(m.x + logf(s.x), m.y + logf(s.y), m.z + logf(s.z), m.w + logf(s.w));
}

// d/d(density) of <adj, GaussianLog(density)> (renderer_tf.cuh:706-744): the
// reference's adj_t_i = adj / sum_j exp(t_j - t_i) is adj times the softmax
// weight exp(t_i - m) / s, so one sum per channel serves every i.
__device__ __forceinline__ float glog_adjoint(const Tf& T, float d, const f4& adj)
{
    const f4 m = glog_max(T, d);
    const f4 s = glog_sum(T, d, m);
    const f4 a = mk4(adj.x / s.x, adj.y / s.y, adj.z / s.z, adj.w / s.w);
    float out = 0.0f;
    for (int i = 0; i < T.R; ++i) {
        const float* r = T.p + i * T.sr;
        const float u = d - r[4];
        const float ln = -(u * u) * r[5];
        const float w = a.x * expf(r[0] + ln - m.x) + a.y * expf(r[1] + ln - m.y) + a.z * expf(r[2] + ln - m.z) +
                        a.w * expf(r[3] + ln - m.w);
        out += w * (-2.0f * r[5] * u);  // (mu - d) / sigma^2
    }
    return out;
}

template <int TFM>
__device__ __forceinline__ f4 tf_color(const Tf& T, float d)
{
    if constexpr (TFM == 6) return glog_eval(T, d);
    else return tf_eval<TFM>(T, d);
}

// ---------------------------------------------------------------------------
// pre-shading (renderer/tf_utils_impl.cu:17-32 forward, :96-121 adjoint).
// A lane takes V consecutive voxels along the innermost loop axis: V = 4 moves
// 16-byte vectors when that axis has unit stride in every tensor and every
// vector is 16-byte aligned (checked on the host), V = 1 otherwise.
// ---------------------------------------------------------------------------
typedef float v4f __attribute__((ext_vector_type(4)));

template <int V>
__device__ __forceinline__ void load_v(const float* p, float (&d)[V])
{
    if constexpr (V == 4) {
        const v4f v = *reinterpret_cast<const v4f*>(p);
        d[0] = v.x; d[1] = v.y; d[2] = v.z; d[3] = v.w;
    } else {
        d[0] = p[0];
    }
}

// outputs are streamed: non-temporal vector stores keep them from displacing
// the inputs in the caches (measured 5.3 -> 5.7 TB/s at 256^3)
template <int V>
__device__ __forceinline__ void store_v(float* p, const float (&d)[V])
{
    if constexpr (V == 4) {
        v4f v;
        v.x = d[0]; v.y = d[1]; v.z = d[2]; v.w = d[3];
        __builtin_nontemporal_store(v, reinterpret_cast<v4f*>(p));
    } else {
        __builtin_nontemporal_store(d[0], p);
    }
}

// Gaussian TF (renderer_tf.cuh:491-519, adjoint :562-614) for V voxels at once:
// each row is read from LDS once and its 1/(2 sigma^2), 1/sigma^2 formed once.
template <int V>
__device__ __forceinline__ void gauss_eval_v(const Tf& T, const float (&d)[V], float (&c)[4][V])
{
#pragma unroll
    for (int k = 0; k < V; ++k) c[0][k] = c[1][k] = c[2][k] = c[3][k] = 0.0f;
    for (int i = 0; i < T.R; ++i) {
        const float* r = T.p + i * T.sr;
        const float c0 = r[0], c1 = r[1], c2 = r[2], c3 = r[3], mu = r[4], sg = r[5];
        const float k2 = -1.0f / (2.0f * sg * sg);
#pragma unroll
        for (int k = 0; k < V; ++k) {
            const float u = d[k] - mu;
            const float n = expf(u * u * k2);
            c[0][k] += c0 * n; c[1][k] += c1 * n; c[2][k] += c2 * n; c[3][k] += c3 * n;
        }
    }
}

template <int V>
__device__ __forceinline__ void gauss_adjoint_v(const Tf& T, const float (&d)[V], const float (&a)[4][V],
                                                float (&g)[V])
{
#pragma unroll
    for (int k = 0; k < V; ++k) g[k] = 0.0f;
    for (int i = 0; i < T.R; ++i) {
        const float* r = T.p + i * T.sr;
        const float c0 = r[0], c1 = r[1], c2 = r[2], c3 = r[3], mu = r[4], sg = r[5];
        const float is2 = 1.0f / (sg * sg), k2 = -0.5f * is2;
#pragma unroll
        for (int k = 0; k < V; ++k) {
            const float u = d[k] - mu;
            const float n = expf(u * u * k2);
            const float ac = a[0][k] * c0 + a[1][k] * c1 + a[2][k] * c2 + a[3][k] * c3;
            g[k] += ac * (-u * is2) * n;  // (mu - d) / sigma^2 * n
        }
    }
}

template <int TFM, bool ADJ, int V>
__global__ __launch_bounds__(kAuxBlock) void k_preshade(PreshadeParams P)
{
    extern __shared__ float s_tf[];
    const Tf T = stage_tf<TFM>(P.tf, s_tf);
    const int groups = P.G.total / V;
    for (int j = blockIdx.x * kAuxBlock + threadIdx.x; j < groups; j += gridDim.x * kAuxBlock) {
        int x, y, z;
        voxel_xyz(P.G, j * V, x, y, z);
        float d[V];
        load_v<V>(P.dens + off3(P.ds, x, y, z), d);
        if constexpr (!ADJ) {
            float c[4][V];
            if constexpr (TFM == 3) {
                gauss_eval_v<V>(T, d, c);
            } else {
#pragma unroll
                for (int k = 0; k < V; ++k) {
                    const f4 v = tf_color<TFM>(T, d[k]);
                    c[0][k] = v.x; c[1][k] = v.y; c[2][k] = v.z; c[3][k] = v.w;
                }
            }
            float* o = P.col + off3(P.cs, x, y, z);
#pragma unroll
            for (int ch = 0; ch < 4; ++ch) store_v<V>(o + ch * P.cs.c, c[ch]);
        } else {
            const float* a = P.acol + off3(P.as, x, y, z);
            float av[4][V];
#pragma unroll
            for (int ch = 0; ch < 4; ++ch) load_v<V>(a + ch * P.as.c, av[ch]);
            float g[V];
            if constexpr (TFM == 3) {
                gauss_adjoint_v<V>(T, d, av, g);
            } else {
#pragma unroll
                for (int k = 0; k < V; ++k) {
                    const f4 adj = mk4(av[0][k], av[1][k], av[2][k], av[3][k]);
                    if constexpr (TFM == 6) {
                        g[k] = glog_adjoint(T, d[k], adj);
                    } else {
                        SinkNone sn;
                        g[k] = tf_adjoint<TFM, false>(T, d[k], adj, sn);
                    }
                }
            }
            store_v<V>(P.adens + off3(P.ads, x, y, z), g);
        }
    }
}

// ---------------------------------------------------------------------------
// best fit (renderer/tf_utils_impl.cu:225-272, :351-420): per voxel, the
// lowest-cost density among num_samples uniform candidates, cost =
// |rgb - rgb_tf|^2 + w_o log(1 + |a - a_tf|) [+ w_n sum over the existing
// axis neighbours of (d - previous)^2]; ties keep the first candidate.
// ---------------------------------------------------------------------------
template <int TFM, bool NEIGH>
__global__ __launch_bounds__(kAuxBlock) void k_best_fit(BestFitParams P)
{
    extern __shared__ float s_tf[];
    const Tf T = stage_tf<TFM>(P.tf, s_tf);
    const unsigned long long seed = splitmix64(P.seed);
    for (int i = blockIdx.x * kAuxBlock + threadIdx.x; i < P.G.total; i += gridDim.x * kAuxBlock) {
        int x, y, z;
        voxel_xyz(P.G, i, x, y, z);
        const float* c = P.col + off3(P.cs, x, y, z);
        const f4 target = mk4(c[0], c[P.cs.c], c[2 * P.cs.c], c[3 * P.cs.c]);
        float nb[6], nw[6];
        if constexpr (NEIGH) {
            // neighbours clamped into the grid (every load in bounds); a missing one weighs 0
            const bool has[6] = {x > 0, x < P.G.X - 1, y > 0, y < P.G.Y - 1, z > 0, z < P.G.Z - 1};
            nb[0] = P.prev[off3(P.ps, max(x - 1, 0), y, z)];
            nb[1] = P.prev[off3(P.ps, min(x + 1, P.G.X - 1), y, z)];
            nb[2] = P.prev[off3(P.ps, x, max(y - 1, 0), z)];
            nb[3] = P.prev[off3(P.ps, x, min(y + 1, P.G.Y - 1), z)];
            nb[4] = P.prev[off3(P.ps, x, y, max(z - 1, 0))];
            nb[5] = P.prev[off3(P.ps, x, y, min(z + 1, P.G.Z - 1))];
#pragma unroll
            for (int j = 0; j < 6; ++j) {
                nw[j] = has[j] ? P.neighborWeighting : 0.0f;
                nb[j] = has[j] ? nb[j] : 0.0f;
            }
        }
        const unsigned long long v = (unsigned long long)((long long)(x * P.G.Y + y) * P.G.Z + z);
        unsigned long long state = splitmix64(seed + v) & kLcgMask;
        float best = 3.402823466e38f, bestD = 0.0f;
        for (int k = 0; k < P.numSamples; ++k) {
            state = (state * kLcgMul + kLcgAdd) & kLcgMask;
            const float d = float(unsigned(state >> 24)) * (1.0f / 16777216.0f);
            const f4 cur = tf_color<TFM>(T, d);
            const float dx = target.x - cur.x, dy = target.y - cur.y, dz = target.z - cur.z;
            float cost = (dx * dx + dy * dy + dz * dz) + P.opacityWeighting * log1pf(fabsf(target.w - cur.w));
            if constexpr (NEIGH) {
#pragma unroll
                for (int j = 0; j < 6; ++j) cost += nw[j] * ((d - nb[j]) * (d - nb[j]));
            }
            if (cost < best) {
                best = cost;
                bestD = d;
            }
        }
        P.out[off3(P.os, x, y, z)] = bestD;
    }
}

// ---------------------------------------------------------------------------
// importance sampling (renderer/losses_impl.cu:254-292 maxima, :298-371 decision)
// ---------------------------------------------------------------------------
__global__ void k_clear3(float* p)
{
    if (threadIdx.x < 3) p[threadIdx.x] = 0.0f;
}

__device__ __forceinline__ float norm3(float a, float b, float c) { return sqrtf(a * a + b * b + c * c); }

template <int TFM>
__global__ __launch_bounds__(kAuxBlock) void k_importance_maxima(ImportanceParams P)
{
    extern __shared__ float s_tf[];
    const Tf T = stage_tf<TFM>(P.tf, s_tf);
    const int X = P.G.X, Y = P.G.Y, Z = P.G.Z;
    float m0 = 0.0f, m1 = 0.0f, m2 = 0.0f;
    for (int i = blockIdx.x * kAuxBlock + threadIdx.x; i < P.G.total; i += gridDim.x * kAuxBlock) {
        int x, y, z;
        voxel_xyz(P.G, i, x, y, z);
        const Strides3& s = P.ds;
        const float* c = P.dens + off3(s, x, y, z);
        const float dC = c[0];
        const float dXm = P.dens[off3(s, max(0, x - 1), y, z)], dXp = P.dens[off3(s, min(X - 1, x + 1), y, z)];
        const float dYm = P.dens[off3(s, x, max(0, y - 1), z)], dYp = P.dens[off3(s, x, min(Y - 1, y + 1), z)];
        const float dZm = P.dens[off3(s, x, y, max(0, z - 1))], dZp = P.dens[off3(s, x, y, min(Z - 1, z + 1))];
        const float oC = tf_absorption<TFM>(T, dC);
        const float oXm = tf_absorption<TFM>(T, dXm), oXp = tf_absorption<TFM>(T, dXp);
        const float oYm = tf_absorption<TFM>(T, dYm), oYp = tf_absorption<TFM>(T, dYp);
        const float oZm = tf_absorption<TFM>(T, dZm), oZp = tf_absorption<TFM>(T, dZp);
        m0 = fmaxf(m0, norm3((dXp - dXm) / 2.0f, (dYp - dYm) / 2.0f, (dZp - dZm) / 2.0f));
        m1 = fmaxf(m1, oC);
        m2 = fmaxf(m2, norm3((oXp - oXm) / 2.0f, (oYp - oYm) / 2.0f, (oZp - oZm) / 2.0f));
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        m0 = fmaxf(m0, __shfl_xor(m0, o, 64));
        m1 = fmaxf(m1, __shfl_xor(m1, o, 64));
        m2 = fmaxf(m2, __shfl_xor(m2, o, 64));
    }
    // workgroup maximum through LDS, then one atomic per value and workgroup;
    // non-negative floats order like their bit patterns
    __shared__ float s_max[3][kAuxBlock / 64];
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        s_max[0][w] = m0;
        s_max[1][w] = m1;
        s_max[2][w] = m2;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        float m = s_max[threadIdx.x][0];
        for (int k = 1; k < kAuxBlock / 64; ++k) m = fmaxf(m, s_max[threadIdx.x][k]);
        atomicMax(reinterpret_cast<unsigned*>(P.maxima) + threadIdx.x, __float_as_uint(m));
    }
}

// trilinear fetch at voxel position p (renderer_interpolation.cuh:145-177)
__device__ __forceinline__ float fetch_tri(const Vol& V, float px, float py, float pz)
{
    const int ix = int(px), iy = int(py), iz = int(pz);
    float d[8];
    cell_load(V, ix, iy, iz, d);
    return tri_lerp(d, px - float(ix), py - float(iy), pz - float(iz));
}

template <int TFM>
__global__ __launch_bounds__(kAuxBlock) void k_sample_importance(ImportanceParams P)
{
    extern __shared__ float s_tf[];
    const Tf T = stage_tf<TFM>(P.tf, s_tf);
    const Vol V{P.dens, P.G.X, P.G.Y, P.G.Z, P.ds.x, P.ds.y, P.ds.z};
    const float sx = float(P.G.X - 1), sy = float(P.G.Y - 1), sz = float(P.G.Z - 1);
    const float mx0 = P.maxima[0], mx1 = P.maxima[1], mx2 = P.maxima[2];
    const unsigned long long seed = splitmix64(P.seed);
    for (int i = blockIdx.x * kAuxBlock + threadIdx.x; i < P.N; i += gridDim.x * kAuxBlock) {
        const float* l = P.loc + i * P.lsN;
        const float px = l[0] * sx, py = l[P.lsC] * sy, pz = l[2 * P.lsC] * sz;
        // two uniform numbers in (0, 1]
        const unsigned long long h = splitmix64(seed + (unsigned long long)i);
        const float rndSample = float(unsigned(h >> 40) + 1u) * (1.0f / 16777216.0f);
        const float rndReject = float(unsigned((h >> 16) & 0xFFFFFFull) + 1u) * (1.0f / 16777216.0f);
        const float dC = fetch_tri(V, px, py, pz);
        const float dXm = fetch_tri(V, px - 1.0f, py, pz), dXp = fetch_tri(V, px + 1.0f, py, pz);
        const float dYm = fetch_tri(V, px, py - 1.0f, pz), dYp = fetch_tri(V, px, py + 1.0f, pz);
        const float dZm = fetch_tri(V, px, py, pz - 1.0f), dZp = fetch_tri(V, px, py, pz + 1.0f);
        bool sampled;
        if (rndSample <= P.wUniform) {
            sampled = true;
        } else if (rndSample <= P.wDensityGradient) {
            sampled = norm3((dXp - dXm) / 2.0f, (dYp - dYm) / 2.0f, (dZp - dZm) / 2.0f) / mx0 >= rndReject;
        } else if (rndSample <= P.wOpacity) {
            sampled = tf_absorption<TFM>(T, dC) / mx1 >= rndReject;
        } else {
            const float oXm = tf_absorption<TFM>(T, dXm), oXp = tf_absorption<TFM>(T, dXp);
            const float oYm = tf_absorption<TFM>(T, dYm), oYp = tf_absorption<TFM>(T, dYp);
            const float oZm = tf_absorption<TFM>(T, dZm), oZp = tf_absorption<TFM>(T, dZp);
            sampled = norm3((oXp - oXm) / 2.0f, (oYp - oYm) / 2.0f, (oZp - oZm) / 2.0f) / mx2 >= rndReject;
        }
        P.out[i * P.os] = sampled ? 1 : 0;
    }
}

// one work item per lane up to 2^20 workgroups (the grid-stride loops cover the rest)
int aux_grid(long long work)
{
    const long long want = (work + kAuxBlock - 1) / kAuxBlock;
    return int(want < (1 << 20) ? (want > 0 ? want : 1) : (1 << 20));
}

size_t tf_lds(const TfTable& t) { return size_t(t.R) * size_t(t.C) * sizeof(float); }

}  // namespace

#define DVR_AUX_TF_SWITCH(tfm, CALL)            \
    switch (tfm) {                              \
    case 0: CALL(0); break;                     \
    case 1: CALL(1); break;                     \
    case 2: CALL(2); break;                     \
    case 3: CALL(3); break;                     \
    case 6: CALL(6); break;                     \
    default: return hipErrorInvalidValue;       \
    }

namespace {
// 16-byte vectors along the innermost loop axis are possible for this tensor
bool vec4_ok(const VoxelGrid& G, const void* p, const Strides3& s, bool channels)
{
    if (p == nullptr) return true;
    if ((reinterpret_cast<uintptr_t>(p) & 15u) != 0) return false;
    const int n[3] = {G.X, G.Y, G.Z}, st[3] = {s.x, s.y, s.z};
    if (st[G.axis[0]] != 1) return false;
    for (int a = 0; a < 3; ++a)
        if (a != G.axis[0] && n[a] > 1 && st[a] % 4 != 0) return false;
    return !channels || s.c % 4 == 0;
}
}  // namespace

hipError_t launch_preshade(const PreshadeParams& P, int tfm, bool adjoint, hipStream_t s)
{
    if (P.G.total == 0) return hipSuccess;
    const bool v4 = P.G.n[0] % 4 == 0 && vec4_ok(P.G, P.dens, P.ds, false) &&
                    (adjoint ? vec4_ok(P.G, P.acol, P.as, true) && vec4_ok(P.G, P.adens, P.ads, false)
                             : vec4_ok(P.G, P.col, P.cs, true));
    const dim3 grid(aux_grid(v4 ? P.G.total / 4 : P.G.total)), block(kAuxBlock);
    const size_t lds = tf_lds(P.tf);
#define DVR_PRESHADE(T)                                                                              \
    if (adjoint && v4) hipLaunchKernelGGL((k_preshade<T, true, 4>), grid, block, lds, s, P);         \
    else if (adjoint) hipLaunchKernelGGL((k_preshade<T, true, 1>), grid, block, lds, s, P);          \
    else if (v4) hipLaunchKernelGGL((k_preshade<T, false, 4>), grid, block, lds, s, P);              \
    else hipLaunchKernelGGL((k_preshade<T, false, 1>), grid, block, lds, s, P)
    DVR_AUX_TF_SWITCH(tfm, DVR_PRESHADE)
#undef DVR_PRESHADE
    return hipGetLastError();
}

hipError_t launch_best_fit(const BestFitParams& P, int tfm, hipStream_t s)
{
    if (P.G.total == 0) return hipSuccess;
    if (tfm == 6) return hipErrorInvalidValue;  // not a best-fit mode in the reference
    const dim3 grid(aux_grid(P.G.total)), block(kAuxBlock);
    const size_t lds = tf_lds(P.tf);
    const bool neigh = P.prev != nullptr;
#define DVR_BEST_FIT(T)                                                                 \
    if (neigh) hipLaunchKernelGGL((k_best_fit<T, true>), grid, block, lds, s, P);       \
    else hipLaunchKernelGGL((k_best_fit<T, false>), grid, block, lds, s, P)
    switch (tfm) {
    case 0: DVR_BEST_FIT(0); break;
    case 1: DVR_BEST_FIT(1); break;
    case 2: DVR_BEST_FIT(2); break;
    case 3: DVR_BEST_FIT(3); break;
    default: return hipErrorInvalidValue;
    }
#undef DVR_BEST_FIT
    return hipGetLastError();
}

hipError_t launch_sample_importance(const ImportanceParams& P, int tfm, hipStream_t s)
{
    hipLaunchKernelGGL(k_clear3, dim3(1), dim3(64), 0, s, P.maxima);
    const size_t lds = tf_lds(P.tf);
    // maxima: at most 1024 workgroups (grid-stride), so few atomics meet on the 3 values
    const dim3 gm(std::min(aux_grid(P.G.total), 1024)), gs(aux_grid(P.N)), block(kAuxBlock);
#define DVR_IMPORTANCE(T)                                                                          \
    if (P.G.total > 0) hipLaunchKernelGGL((k_importance_maxima<T>), gm, block, lds, s, P);         \
    if (P.N > 0) hipLaunchKernelGGL((k_sample_importance<T>), gs, block, lds, s, P)
    switch (tfm) {
    case 0: DVR_IMPORTANCE(0); break;
    case 1: DVR_IMPORTANCE(1); break;
    case 2: DVR_IMPORTANCE(2); break;
    case 3: DVR_IMPORTANCE(3); break;
    default: return hipErrorInvalidValue;
    }
#undef DVR_IMPORTANCE
    return hipGetLastError();
}

}  // namespace dvr
