// Forward-mode differentiation (renderer/renderer_kernels.cuh:135-308,
// :504-598; renderer/renderer_compareToImage.cuh).  The reference carries
// dual numbers with D derivative slots through the whole march (cudAD
// fvar<real4, D>), which for a TF with R*C traced entries puts R*C*4 values
// per thread in registers.  Here each launch slice traces a chunk of kFwdChunk
// derivative indices (blockIdx.y selects the chunk), so the per-thread state is
// bounded for any D; the value part is the march of the tiled forward kernel
// (identical colour and termination).
//
// Traced variables (per pixel): step size, ray start and ray direction (the
// ray of the pixel after camera evaluation, renderer_camera.cuh:64-130), and
// TF entries through the index table d_tf.  tmin is not traced (the
// reference casts it to real).
#include "dvr_tiled_impl.hpp"

namespace dvr {

// Density and its gradient in voxel coordinates at p (trilinear: the
// derivative of the lerp form; nearest: zero).
template <int FILTER>
__device__ __forceinline__ float fetch_grad(const Vol& V, f3 p, Cell& c, f3& g)
{
    if constexpr (FILTER == 0) {
        g = mk3(0.f, 0.f, 0.f);
        return gld(V.p, unsigned(nearest_offset(V, p)));
    } else {
        const int ix = int(p.x), iy = int(p.y), iz = int(p.z);
        if (ix != c.ix || iy != c.iy || iz != c.iz) {
            c.ix = ix; c.iy = iy; c.iz = iz;
            cell_load(V, ix, iy, iz, c.d);
        }
        const float fx = p.x - float(ix), fy = p.y - float(iy), fz = p.z - float(iz);
        const float* d = c.d;
        const float gx = 1.0f - fx, gy = 1.0f - fy, gz = 1.0f - fz;
        g.x = (d[4] - d[0]) * (gy * gz) + (d[5] - d[1]) * (gy * fz) + (d[6] - d[2]) * (fy * gz) + (d[7] - d[3]) * (fy * fz);
        g.y = (d[2] - d[0]) * (gx * gz) + (d[3] - d[1]) * (gx * fz) + (d[6] - d[4]) * (fx * gz) + (d[7] - d[5]) * (fx * fz);
        g.z = (d[1] - d[0]) * (gx * gy) + (d[3] - d[2]) * (gx * fy) + (d[5] - d[4]) * (fx * gy) + (d[7] - d[6]) * (fx * fy);
        return tri_lerp(d, fx, fy, fz);
    }
}

// Derivatives of the TF output at `density`: w.r.t. the density (ds) and
// w.r.t. TF entry ent[k] = r*C + col (de[k]; ent[k] < 0: none).  Dual-number
// forms of renderer_tf.cuh:122-145 (identity), :256-287 (texture), :437-485
// (piecewise linear) and :522-560 (Gaussian); the value is tf_eval's.
// The clamps take the derivatives of the reference's adjoint, so that the
// forward Jacobian is exactly the transpose of the adjoint: the identity
// TF passes the density derivative strictly inside (0, 1) only
// (renderer_tf.cuh:103); the linear TF differentiates
// frac = (clamp(d, p0, p1) - p0) / (p1 - p0) with the clamped density held
// fixed and passes the density derivative 1 / (p1 - p0) unmasked
// (renderer_tf.cuh:337-357).
template <int TFM, int NE>
__device__ __forceinline__ void tf_deriv(const Tf& T, float density, f4& ds, const int (&ent)[NE], f4 (&de)[NE])
{
    constexpr int C = tf_cols<TFM>();
#pragma unroll
    for (int k = 0; k < NE; ++k) de[k] = mk4(0.f, 0.f, 0.f, 0.f);
    if constexpr (TFM == 0) {
        const float dc = fclamp(density, 0.0f, 1.0f);
        const float sa = T.p[0], sc = T.p[1];
        const float inside = (density > 0.0f && density < 1.0f) ? 1.0f : 0.0f;
        ds = mk4(sc * inside, sc * inside, sc * inside, sa * inside);
#pragma unroll
        for (int k = 0; k < NE; ++k) {
            if (ent[k] == 0) de[k].w = dc;
            else if (ent[k] == 1) de[k] = mk4(dc, dc, dc, 0.f);
        }
    } else if constexpr (TFM == 1) {
        const float d = density * float(T.R) - 0.5f;
        const int di = int(floorf(d));
        const float df = d - float(di);
        const int i0 = iclamp(di, 0, T.R - 1), i1 = iclamp(di + 1, 0, T.R - 1);
        const float* a = T.p + i0 * C;
        const float* c = T.p + i1 * C;
        const float R = float(T.R);
        ds = mk4((c[0] - a[0]) * R, (c[1] - a[1]) * R, (c[2] - a[2]) * R, (c[3] - a[3]) * R);
#pragma unroll
        for (int k = 0; k < NE; ++k) {
            if (ent[k] < 0) continue;
            const int r = ent[k] / C, col = ent[k] - r * C;
            const float w = (r == i0 ? 1.0f - df : 0.0f) + (r == i1 ? df : 0.0f);
            if (col == 0) de[k].x = w; else if (col == 1) de[k].y = w; else if (col == 2) de[k].z = w; else de[k].w = w;
        }
    } else if constexpr (TFM == 2) {
        int i;
        for (i = 0; i < T.R - 2; ++i)
            if (T.p[(i + 1) * C + 4] > density) break;
        const float* a = T.p + i * C;
        const float* c = a + C;
        const float p0 = a[4], p1 = c[4];
        const float d2 = fclamp(density, p0, p1);
        const float f = (d2 - p0) / (p1 - p0);
        const float dfd = 1.0f / (p1 - p0);
        const f4 dv = mk4(c[0] - a[0], c[1] - a[1], c[2] - a[2], c[3] - a[3]);
        ds = mk4(dv.x * dfd, dv.y * dfd, dv.z * dfd, dv.w * dfd);
        const float den = (p1 - p0) * (p1 - p0);
#pragma unroll
        for (int k = 0; k < NE; ++k) {
            if (ent[k] < 0) continue;
            const int r = ent[k] / C, col = ent[k] - r * C;
            if (col < 4) {
                const float w = (r == i ? 1.0f - f : 0.0f) + (r == i + 1 ? f : 0.0f);
                if (col == 0) de[k].x = w; else if (col == 1) de[k].y = w; else if (col == 2) de[k].z = w; else de[k].w = w;
            } else {
                const float dfp = (r == i ? (d2 - p1) / den : 0.0f) + (r == i + 1 ? (p0 - d2) / den : 0.0f);
                de[k] = mk4(dv.x * dfp, dv.y * dfp, dv.z * dfp, dv.w * dfp);
            }
        }
    } else {
        ds = mk4(0.f, 0.f, 0.f, 0.f);
        for (int i = 0; i < T.R; ++i) {
            const float* r = T.p + i * C;
            const float mu = r[4], sg = r[5];
            const float n = gauss_n<true>(density, mu, sg);  // as the colour (tf_eval<3, true>)
            const float dn = n * (mu - density) / (sg * sg);
            ds.x += r[0] * dn; ds.y += r[1] * dn; ds.z += r[2] * dn; ds.w += r[3] * dn;
#pragma unroll
            for (int k = 0; k < NE; ++k) {
                if (ent[k] < i * C || ent[k] >= (i + 1) * C) continue;
                const int col = ent[k] - i * C;
                if (col == 0) de[k].x = n;
                else if (col == 1) de[k].y = n;
                else if (col == 2) de[k].z = n;
                else if (col == 3) de[k].w = n;
                else {
                    const float dd = density - mu;
                    const float q = col == 4 ? n * dd / (sg * sg) : n * dd * dd / (sg * sg * sg);
                    de[k] = mk4(r[0] * q, r[1] * q, r[2] * q, r[3] * q);
                }
            }
        }
    }
}

template <int FILTER, int TFM, int BLEND>
__global__ __launch_bounds__(kBlock) void k_forward_gradients(RenderParams P, FwdGradParams G, int tilesX,
                                                              int tilesPerView, int nTiles)
{
    extern __shared__ float lds[];
    const TileXY tc = tile_xy(tilesX, tilesPerView, nTiles);
    stage_tf(P, tc.b, lds);
    __syncthreads();
    if (tc.x >= P.W || tc.y >= P.H) return;
    const int b = tc.b, x = tc.x, y = tc.y;
    const int d0 = blockIdx.y * kFwdChunk;
    const int nSlots = min(kFwdChunk, G.D - d0);
    // slot -> TF entry traced by that derivative index (uniform over the block)
    int ent[kFwdChunk];
#pragma unroll
    for (int k = 0; k < kFwdChunk; ++k)
        ent[k] = (G.tfEntry && k < nSlots) ? G.tfEntry[b * G.tfEntrySb + d0 + k] : -1;

    const Ray r = setup_ray(P, x, y, b);
    const Vol V = make_vol(P, b);
    constexpr int kC = tf_cols<TFM>();
    const Tf T{lds, P.R, kC, kC, 1};
    // derivative seeds of slot k (uniform): step, start (3), direction (3)
    float sStep[kFwdChunk];
    f3 sStart[kFwdChunk], sDir[kFwdChunk];
#pragma unroll
    for (int k = 0; k < kFwdChunk; ++k) {
        const int d = d0 + k;
        sStep[k] = (k < nSlots && d == G.dStep) ? 1.0f : 0.0f;
        sStart[k] = mk3(d == G.dStart[0] ? 1.f : 0.f, d == G.dStart[1] ? 1.f : 0.f, d == G.dStart[2] ? 1.f : 0.f);
        sDir[k] = mk3(d == G.dDir[0] ? 1.f : 0.f, d == G.dDir[1] ? 1.f : 0.f, d == G.dDir[2] ? 1.f : 0.f);
        if (k >= nSlots) { sStart[k] = mk3(0.f, 0.f, 0.f); sDir[k] = mk3(0.f, 0.f, 0.f); }
    }
    // value state and derivatives of (r, g, b, A, T)
    Acc acc{0.f, 0.f, 0.f, 0.f, 1.f};
    float dr[kFwdChunk], dg[kFwdChunk], db[kFwdChunk], dA[kFwdChunk], dT[kFwdChunk];
#pragma unroll
    for (int k = 0; k < kFwdChunk; ++k) { dr[k] = 0.f; dg[k] = 0.f; db[k] = 0.f; dA[k] = 0.f; dT[k] = 0.f; }
    const float tEarly = 1.0f - P.earlyOut;
    Cell cell;
    cell_invalidate(cell);
    int i = 0;
    for (; i < kMaxSamples; ++i) {
        const float t = r.t0 + float(i) * r.step;
        if (t > r.t1) break;
        if (acc.T < tEarly) break;
        const f3 p = sample_pos(r, t);
        f3 gd;
        const float dens = fetch_grad<FILTER>(V, p, cell, gd);
        // d density / d variable: step -> i * (grad . dp), start_a -> grad_a / voxel_a, dir_a -> t grad_a / voxel_a
        const f3 gv = mul3(gd, r.invVoxel);
        const float gStep = float(i) * dot3(gd, r.dp);
        f4 ds, de[kFwdChunk];
        const f4 s = tf_eval<TFM, true>(T, dens);
        tf_deriv<TFM, kFwdChunk>(T, dens, ds, ent, de);
        float alpha, keep;
        alpha_keep<BLEND>(s.w, r.step, alpha, keep);
        const float w = acc.T * alpha;
        const bool lin = BLEND == 0 || s.w * r.step < 1.0f;
#pragma unroll
        for (int k = 0; k < kFwdChunk; ++k) {
            const float dd = sStep[k] * gStep + dot3(sStart[k], gv) + t * dot3(sDir[k], gv);
            const f4 sp = mk4(ds.x * dd + de[k].x, ds.y * dd + de[k].y, ds.z * dd + de[k].z, ds.w * dd + de[k].w);
            // x = absorption * step
            const float dx = sp.w * r.step + s.w * sStep[k];
            float dalpha, dkeep;
            if constexpr (BLEND == 0) {
                dalpha = keep * dx;
                dkeep = -dalpha;
            } else {
                dalpha = lin ? dx : 0.0f;
                dkeep = -dalpha;
            }
            const float dw = dT[k] * alpha + acc.T * dalpha;
            dr[k] += dw * s.x + w * sp.x;
            dg[k] += dw * s.y + w * sp.y;
            db[k] += dw * s.z + w * sp.z;
            dA[k] += dw;
            dT[k] = dT[k] * keep + acc.T * dkeep;
        }
        acc.r += w * s.x;
        acc.g += w * s.y;
        acc.b += w * s.z;
        acc.A += w;
        acc.T *= keep;
    }
    if (blockIdx.y == 0) {
        float* o = P.color + b * P.colSb + y * P.colSy + x * P.colSx;
        o[0] = acc.r; o[P.colSc] = acc.g; o[2 * P.colSc] = acc.b; o[3 * P.colSc] = acc.A;
        P.term[b * P.termSb + y * P.termSy + x * P.termSx] = i;
    }
    float* g = G.grad + b * G.gSb + y * G.gSy + x * G.gSx + d0 * G.gSd;
#pragma unroll
    for (int k = 0; k < kFwdChunk; ++k) {
        if (k < nSlots) {
            float* q = g + k * G.gSd;
            q[0] = dr[k]; q[G.gSc] = dg[k]; q[2 * G.gSc] = db[k]; q[3 * G.gSc] = dA[k];
        }
    }
}

template <int FILTER, int TFM, int BLEND>
static hipError_t fwdgrad_t(const RenderParams& P, const FwdGradParams& G, hipStream_t s)
{
    const int tilesX = (P.W + kTileW - 1) / kTileW, tilesY = (P.H + kTileH - 1) / kTileH;
    const int nTiles = tilesX * tilesY * P.B;
    if (nTiles == 0 || G.D == 0) return hipSuccess;
    const int chunks = (G.D + kFwdChunk - 1) / kFwdChunk;
    const size_t lds = sizeof(float) * size_t(P.R) * P.C;
    hipLaunchKernelGGL((k_forward_gradients<FILTER, TFM, BLEND>), dim3(nTiles, chunks), dim3(kBlock), lds, s, P, G,
                       tilesX, tilesX * tilesY, nTiles);
    return hipGetLastError();
}

hipError_t launch_forward_gradients(const RenderParams& P, const FwdGradParams& G, int filter, int tfm, int blend,
                                    hipStream_t s)
{
    DVR_DISPATCH3(filter, tfm, blend, return fwdgrad_t, (P, G, s));
    return hipErrorInvalidValue;
}

// ---------------------------------------------------------------------------
// forward variables -> adjoint variables (renderer/renderer_kernels.cuh:504-598)
// ---------------------------------------------------------------------------
__device__ __forceinline__ float fv_dot(const FwdToAdjParams& F, const float* fp, f4 a, int d)
{
    const float* q = fp + d * F.fSd;
    return a.x * q[0] + a.y * q[F.fSc] + a.z * q[2 * F.fSc] + a.w * q[3 * F.fSc];
}

__device__ __forceinline__ void acc_out(float* p, float v, int atomic)
{
    if (atomic) atomicAdd(p, v);
    else *p += v;
}

// blockIdx.y = batch; per pixel: the camera / step adjoints, and the
// contraction sum_pixels dot(adj_color, fv[:, d]) for every d, reduced in LDS.
__global__ __launch_bounds__(256) void k_forward_to_adjoint(FwdToAdjParams F)
{
    extern __shared__ float s_sum[];
    const int b = blockIdx.y;
    if (F.sums)
        for (int d = threadIdx.x; d < F.D; d += 256) s_sum[d] = 0.0f;
    __syncthreads();
    const int pix = blockIdx.x * 256 + threadIdx.x;
    const bool valid = pix < F.H * F.W;
    const int y = valid ? pix / F.W : 0, x = valid ? pix - (pix / F.W) * F.W : 0;
    const float* fp = F.fv + b * F.fSb + y * F.fSy + x * F.fSx;
    f4 a = mk4(0.f, 0.f, 0.f, 0.f);
    if (valid) {
        const float* ap = F.ac + b * F.aSb + y * F.aSy + x * F.aSx;
        a = mk4(ap[0], ap[F.aSc], ap[2 * F.aSc], ap[3 * F.aSc]);
        if (F.ast && F.dStep >= 0)
            acc_out(F.ast + b * F.astSb + y * F.astSy + x * F.astSx, fv_dot(F, fp, a, F.dStep), F.astAtomic);
        for (int c = 0; c < 3; ++c) {
            if (F.ars && F.dStart[c] >= 0)
                acc_out(F.ars + b * F.arsSb + y * F.arsSy + x * F.arsSx + c * F.arsSc, fv_dot(F, fp, a, F.dStart[c]),
                        F.camAtomic);
            if (F.ard && F.dDir[c] >= 0)
                acc_out(F.ard + b * F.ardSb + y * F.ardSy + x * F.ardSx + c * F.ardSc, fv_dot(F, fp, a, F.dDir[c]),
                        F.camAtomic);
        }
    }
    if (F.sums) {
        const int lane = threadIdx.x & 63;
        for (int d = 0; d < F.D; ++d) {
            const float v = wave_sum(valid ? fv_dot(F, fp, a, d) : 0.0f);
            if (lane == 0 && v != 0.0f) atomicAdd(&s_sum[d], v);
        }
        __syncthreads();
        for (int d = threadIdx.x; d < F.D; d += 256)
            if (s_sum[d] != 0.0f) atomicAdd(F.sums + b * F.D + d, s_sum[d]);
    }
}

// adj_tf[b][r][c] += sums[b][d_tf[b][r][c]] for d_tf >= 0 (atomic: the
// reference accumulates, and broadcast batches sum over the forward batch)
__global__ __launch_bounds__(256) void k_forward_to_adjoint_tf(const float* sums, int D, int B, const int* dtf,
                                                               int dSb, int dSr, int dSc, float* atf, int tSb,
                                                               int tSr, int tSc, int R, int C)
{
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= B * R * C) return;
    const int b = e / (R * C), rc = e - b * R * C, r = rc / C, c = rc - r * C;
    const int d = dtf[b * dSb + r * dSr + c * dSc];
    if (d < 0 || d >= D) return;
    const float v = sums[b * D + d];
    if (v != 0.0f) atomicAdd(atf + b * tSb + r * tSr + c * tSc, v);
}

__global__ __launch_bounds__(256) void k_fill_zero(float* p, long long n, long long stride)
{
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i * stride] = 0.0f;
}

hipError_t launch_forward_to_adjoint(const FwdToAdjParams& F, const int* dtf, int dSb, int dSr, int dSc, float* atf,
                                     int tSb, int tSr, int tSc, int R, int C, hipStream_t s)
{
    if (F.B == 0 || F.H * F.W == 0) return hipSuccess;
    if (F.sums) {
        const long long n = (long long)F.B * F.D;
        hipLaunchKernelGGL(k_fill_zero, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, F.sums, n, 1ll);
    }
    const size_t lds = F.sums ? sizeof(float) * size_t(F.D) : 0;
    hipLaunchKernelGGL(k_forward_to_adjoint, dim3((F.H * F.W + 255) / 256, F.B), dim3(256), lds, s, F);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || !F.sums || !atf) return e;
    const int n = F.B * R * C;
    hipLaunchKernelGGL(k_forward_to_adjoint_tf, dim3((n + 255) / 256), dim3(256), 0, s, F.sums, F.D, F.B, dtf, dSb,
                       dSr, dSc, atf, tSb, tSr, tSc, R, C);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// compareToImage (renderer/renderer.cpp:562-658, renderer_compareToImage.cuh):
// per pixel L = |color - reference| (Euclidean norm over RGBA) and its forward
// derivatives dL/dθ_d = dot(color - reference, g_d) / L; reduce = mean over
// every pixel of the batch.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_compare_to_image(CompareParams Q)
{
    const int b = blockIdx.y;
    const int pix = blockIdx.x * 256 + threadIdx.x;
    const bool valid = pix < Q.H * Q.W;
    const int y = valid ? pix / Q.W : 0, x = valid ? pix - (pix / Q.W) * Q.W : 0;
    f4 df = mk4(0.f, 0.f, 0.f, 0.f);
    float L = 0.0f;
    if (valid) {
        const float* c = Q.col + b * Q.cSb + y * Q.cSy + x * Q.cSx;
        const float* r = Q.ref + b * Q.rSb + y * Q.rSy + x * Q.rSx;
        df = mk4(c[0] - r[0], c[Q.cSc] - r[Q.rSc], c[2 * Q.cSc] - r[2 * Q.rSc], c[3 * Q.cSc] - r[3 * Q.rSc]);
        L = sqrtf(df.x * df.x + df.y * df.y + df.z * df.z + df.w * df.w);
    }
    const float* gp = Q.gin + b * Q.gSb + y * Q.gSy + x * Q.gSx;
    const float invL = L > 0.0f ? 1.0f / L : 0.0f;
    if (!Q.reduce) {
        if (!valid) return;
        Q.dout[b * Q.oSb + y * Q.oSy + x * Q.oSx] = L;
        for (int d = 0; d < Q.D; ++d) {
            const float* g = gp + d * Q.gSd;
            Q.gout[b * Q.qSb + y * Q.qSy + x * Q.qSx + d * Q.qSd] =
                (df.x * g[0] + df.y * g[Q.gSc] + df.z * g[2 * Q.gSc] + df.w * g[3 * Q.gSc]) * invL;
        }
        return;
    }
    const int lane = threadIdx.x & 63;
    const float s = wave_sum(L * Q.norm);
    if (lane == 0 && s != 0.0f) atomicAdd(Q.dout, s);
    for (int d = 0; d < Q.D; ++d) {
        float v = 0.0f;
        if (valid) {
            const float* g = gp + d * Q.gSd;
            v = (df.x * g[0] + df.y * g[Q.gSc] + df.z * g[2 * Q.gSc] + df.w * g[3 * Q.gSc]) * invL * Q.norm;
        }
        v = wave_sum(v);
        if (lane == 0 && v != 0.0f) atomicAdd(Q.gout + d * Q.qSd, v);
    }
}

hipError_t launch_compare_to_image(const CompareParams& Q, hipStream_t s)
{
    if (Q.reduce) {
        // the reduced outputs are sums: clear them first (renderer.cpp:596-597)
        hipLaunchKernelGGL(k_fill_zero, dim3(1), dim3(256), 0, s, Q.dout, 1ll, 1ll);
        if (Q.D > 0)
            hipLaunchKernelGGL(k_fill_zero, dim3(unsigned((Q.D + 255) / 256)), dim3(256), 0, s, Q.gout, (long long)Q.D,
                               (long long)Q.qSd);
    }
    if (Q.B == 0 || Q.H * Q.W == 0) return hipGetLastError();
    hipLaunchKernelGGL(k_compare_to_image, dim3((Q.H * Q.W + 255) / 256, Q.B), dim3(256), 0, s, Q);
    return hipGetLastError();
}

}  // namespace dvr
