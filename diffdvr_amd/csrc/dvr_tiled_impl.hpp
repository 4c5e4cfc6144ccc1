// Tiled kernels for gfx950: each workgroup is one wave64 rendering an 8x8
// pixel tile of one view, so the wave's trilinear gathers touch a compact
// bundle of nearby cells, and a wave that finishes frees its slot at once.  Tiles are ordered so that each XCD owns a
// contiguous band of them (shared L2).  The transfer function is staged in
// LDS; its gradient is privatised in LDS and flushed once per workgroup.
//
// The adjoint is checkpointed: the forward march records the accumulated
// alpha every K samples; the reverse march recomputes each K-sample segment
// from its checkpoint into registers and walks it backwards.  No blend step is
// ever inverted (the reference inverts it, renderer/renderer_blending.cuh:75-77,
// which loses precision as the accumulated alpha approaches 1).
//
// Forward and reverse semantics: renderer/renderer_kernels.cuh:22-90 and
// :324-444 (step-size adjoint accumulated over all samples, see dvr_device.hpp).
#pragma once
#include <type_traits>
#include "dvr_device.hpp"
#include "dvr_launch.hpp"

namespace dvr {

constexpr int kInvalidCell = -0x40000000;

struct TileXY {
    int tile, b, x, y;
};

__device__ __forceinline__ TileXY tile_xy(int tilesX, int tilesPerView, int nTiles)
{
    TileXY t;
    t.tile = xcd_remap(blockIdx.x, nTiles);
    t.b = t.tile / tilesPerView;
    const int tl = t.tile - t.b * tilesPerView;
    const int ty = tl / tilesX, tx = tl - ty * tilesX;
    t.x = tx * kTileW + int(threadIdx.x & 7);
    t.y = ty * kTileH + int(threadIdx.x >> 3);
    return t;
}

__device__ __forceinline__ void stage_tf(const RenderParams& P, int b, float* s_tf)
{
    const float* g = P.tf + b * P.tfSb;
    const int RC = P.R * P.C;
    for (int e = threadIdx.x; e < RC; e += blockDim.x) {
        const int r = e / P.C, c = e - r * P.C;
        s_tf[e] = g[r * P.tfSr + c * P.tfSc];
    }
}

// ---------------------------------------------------------------------------
// Forward march of one ray (transmittance-tracking blend, dvr_device.hpp).
// frozen >= 0 replays exactly that many samples.  Writes the transmittance
// after every K-th sample to ck[c*ckS] for c < cap (ck may be NULL).
// Early termination tests the accumulated alpha as the reference does
// (renderer_kernels.cuh:64), evaluated as 1 - T.
// ---------------------------------------------------------------------------
template <int FILTER, int TFM, int BLEND>
__device__ __forceinline__ void march_forward(const Vol& V, const Tf& T, const Ray& r, float earlyOut, int frozen,
                                              Acc& acc, int& n, float* ck, int ckS, int K, int cap,
                                              float* dsw = nullptr, int dsS = 0, int dsCap = 0)
{
    acc = Acc{0.f, 0.f, 0.f, 0.f, 1.f};
    const float tEarly = 1.0f - earlyOut;
    int countdown = K, cidx = 0;
    int i = 0;
    float t = r.t0;
    f3 p = sample_pos(r, t);
    FetchPipe<FILTER> pipe;
    bool started = false;
    for (; i < kMaxSamples; ++i) {
        if (frozen >= 0) {
            if (i >= frozen) break;
        } else {
            if (t > r.t1) break;
            if (acc.T < tEarly) break;
        }
        if (!started) {
            pipe.start(V, p);
            started = true;
        }
        const float d = pipe.value(p);
        if (dsw) *(gfloat*)(dsw + min(i, dsCap - 1) * dsS) = d;  // past the capacity: never read back
        const float tn = r.t0 + float(i + 1) * r.step;
        const f3 pn = sample_pos(r, tn);
        pipe.request(V, pn);
        blend_t<BLEND>(acc, tf_eval<TFM, true>(T, d), r.step);
        if (ck) {
            if (--countdown == 0) {
                countdown = K;
                if (cidx < cap) ck[cidx * ckS] = acc.T;
                ++cidx;
            }
        }
        pipe.advance();
        t = tn;
        p = pn;
    }
    n = i;
}

// ---------------------------------------------------------------------------
// TF-gradient sink of the tiled adjoint.
//   identity : two registers per lane, wave-reduced at the end
//   texture / linear : per-lane cache of the two most recent rows (densities
//       change slowly along a ray), flushed into a workgroup LDS table with
//       LDS atomics only when a row leaves (a move by one texel keeps the
//       shared row); the batched kernel spreads the atomics of a wave over
//       kTfCopies copies of the table (neighbouring rays hit the same rows)
//   gaussian : every sample touches every entry; per-thread private LDS slices
//       (entry-major, conflict-free) when they fit, else LDS atomics
// ---------------------------------------------------------------------------
template <int TFM>
struct SinkTiled {
    float* tab;   // workgroup table (R*C) or private slice base (priv)
    float* tabc;  // this lane's copy of the table (copies: lanes of one wave spread their LDS atomics)
    int C;
    int nthreads; // threads of the workgroup (slice stride of the private slices)
    bool priv;
    int r0, r1;
    float a0[5], a1[5];
    float id0, id1;

    // copyStride > 0: kTfCopies tables copyStride floats apart, lane & (kTfCopies-1) adds to its own
    __device__ __forceinline__ void init(float* table, int C_, bool usePriv, int threads = kBlock, int copyStride = 0)
    {
        tab = table; C = C_; priv = usePriv; nthreads = threads;
        tabc = table + (copyStride > 0 ? int(threadIdx.x & (kTfCopies - 1)) * copyStride : 0);
        r0 = r1 = -1;
#pragma unroll
        for (int k = 0; k < 5; ++k) { a0[k] = 0.f; a1[k] = 0.f; }
        id0 = id1 = 0.f;
    }
    __device__ __forceinline__ void add(int r, int c, float v)
    {
        if constexpr (TFM == 0) {
            if (c == 0) id0 += v; else id1 += v;
        } else {
            if (priv) tab[(r * C + c) * nthreads + threadIdx.x] += v;
            else atomicAdd(&tabc[r * C + c], v);
        }
    }
    template <int NC>
    __device__ __forceinline__ void flush_row(int r, float (&a)[5])
    {
        if (r >= 0) {
#pragma unroll
            for (int c = 0; c < NC; ++c) atomicAdd(&tabc[r * C + c], a[c]);
        }
#pragma unroll
        for (int k = 0; k < 5; ++k) a[k] = 0.f;
    }
    template <bool POS>
    __device__ __forceinline__ void add_pair(int i0, float w0, int i1, float w1, const f4& adj, float p0, float p1)
    {
        constexpr int NC = POS ? 5 : 4;
        if (i0 != r0) {
            // a move by one texel keeps the accumulator of the row both pairs share
            if (i0 == r1 && r1 >= 0) {
                flush_row<NC>(r0, a0);
#pragma unroll
                for (int k = 0; k < 5; ++k) { a0[k] = a1[k]; a1[k] = 0.f; }
                r0 = i0; r1 = -1;
            } else if (i1 == r0 && i1 != i0) {
                flush_row<NC>(r1, a1);
#pragma unroll
                for (int k = 0; k < 5; ++k) { a1[k] = a0[k]; a0[k] = 0.f; }
                r1 = i1; r0 = -1;
            }
        }
        if (i0 != r0) { flush_row<NC>(r0, a0); r0 = i0; }
        a0[0] += adj.x * w0; a0[1] += adj.y * w0; a0[2] += adj.z * w0; a0[3] += adj.w * w0;
        if constexpr (POS) a0[4] += p0;
        if (i1 != r1) { flush_row<NC>(r1, a1); r1 = i1; }
        a1[0] += adj.x * w1; a1[1] += adj.y * w1; a1[2] += adj.z * w1; a1[3] += adj.w * w1;
        if constexpr (POS) a1[4] += p1;
    }
    // called by every thread of the workgroup (uniform control flow); NC
    // columns per row (5 with the control-point positions of a linear TF)
    template <int NC>
    __device__ __forceinline__ void finish()
    {
        if constexpr (TFM == 0) {
            const float s0 = wave_sum(id0), s1 = wave_sum(id1);
            if ((threadIdx.x & 63) == 0) {
                atomicAdd(&tab[0], s0);
                atomicAdd(&tab[1], s1);
            }
        } else if constexpr (TFM == 1 || TFM == 2) {
            flush_row<NC>(r0, a0);
            flush_row<NC>(r1, a1);
        }
    }
};

// ---------------------------------------------------------------------------
// Reverse-walk cell state of the volume and position adjoints.  Consecutive
// cells of a ray that differ by one step along one axis share the 4 corners
// of the face between them -- also where the reference clamps corners at the
// volume border, since the old cell's high corner and the new cell's low
// corner clamp to the same voxel.  Their gradient accumulators (and corner
// densities) are kept; only the 4 corners that leave are flushed with global
// atomics (and only the 4 new ones loaded), which halves the volume-gradient
// atomics of a ray crossing one face per step.  Corner bit order (x<<2)|(y<<1)|z.
// ---------------------------------------------------------------------------
__device__ __forceinline__ unsigned corner_off(const CellOffsets& o, int k)
{
    return o.base + ((k & 4) ? o.dx : 0u) + ((k & 2) ? o.dy : 0u) + ((k & 1) ? o.dz : 0u);
}

// ---------------------------------------------------------------------------
// Volume-gradient sinks of the reverse walk.  Both take element offsets from
// the gradient base; zero values are dropped.
//
// VolDirect: one device-memory float atomic per corner (the reference's
// flush, renderer_interpolation.cuh:299-324).
//
// VolAgg: a device-memory float atomic costs one 64-byte request whether one
// lane or sixteen lanes of a wave instruction address that 64-byte chunk
// (measured on MI355X: ~21 G requests/s for lanes on distinct chunks, so
// ~21 G floats/s scattered and ~330 G floats/s for 16-lane groups on one
// chunk).  The 64 rays of a wave, walked over a few samples, touch ~7 corner
// values per distinct 64-byte chunk, so the wave's corner gradients are first
// summed in an LDS table of chunks (open addressing on the chunk index:
// ds_cmpst claims a slot, ds_add_f32 adds the value) and the table is flushed
// with one 16-lane atomic group per occupied chunk.  The walk checks every four
// samples how many slots are taken, and the flush lists the occupied slots, by
// scanning the tags.  A chunk that finds no free
// slot within kAggProbe probes goes straight to device memory.  The table
// belongs to one wave (one-wave workgroups) and to one gradient tensor: it is
// flushed when it is half full, at the end of every tile and before its
// gradient base changes.  The flush runs on whichever lanes are active (the
// work is spread by rank among them), so it may be called from divergent code.
// ---------------------------------------------------------------------------
constexpr int kAggProbe = 8;
constexpr unsigned kAggEmpty = 0xFFFFFFFFu;

// Element offsets of a cell's 8 corners in a gradient layout: corner (a,b,c)
// (bits x<<2 | y<<1 | z) lies at x[a] + y[b] + z[c].  Corners past the border
// clamp to it, as the reference's trilinear fetch does (cell_offsets).
struct CornerIdx {
    unsigned x[2], y[2], z[2];
};
__device__ __forceinline__ unsigned corner_at(const CornerIdx& o, int k)
{
    return o.x[(k >> 2) & 1] + o.y[(k >> 1) & 1] + o.z[k & 1];
}
// the two corner coordinates of cell index i along an axis of n voxels
__device__ __forceinline__ void cell_axis(int i, int n, int& c0, int& c1)
{
    c0 = iclamp(i, 0, n - 1);
    c1 = unsigned(i) < unsigned(n - 1) ? i + 1 : c0;
}

// X x Y x Z gradient with element strides (the caller's tensor)
struct LinearLayout {
    int X, Y, Z, sx, sy, sz;
    __device__ __forceinline__ CornerIdx cell(int ix, int iy, int iz) const
    {
        const CellOffsets o = cell_offsets(ix, iy, iz, X, Y, Z, sx, sy, sz);
        const unsigned bx = unsigned(iclamp(ix, 0, X - 1)) * unsigned(sx);
        const unsigned by = unsigned(iclamp(iy, 0, Y - 1)) * unsigned(sy);
        const unsigned bz = unsigned(iclamp(iz, 0, Z - 1)) * unsigned(sz);
        return CornerIdx{{bx, bx + o.dx}, {by, by + o.dy}, {bz, bz + o.dz}};
    }
    __device__ __forceinline__ unsigned point(int x, int y, int z) const
    {
        return unsigned(x) * unsigned(sx) + unsigned(y) * unsigned(sy) + unsigned(z) * unsigned(sz);
    }
};

// Bricked gradient scratch: every 64-byte chunk holds a 2 x 2 x 4 brick of
// voxels (x, y, z; lane (x&1)<<3 | (y&1)<<2 | z&3), bricks z-fastest.  A ray's
// corners -- and those of its neighbours, in any direction -- then fall into
// fewer chunks than along 16-voxel z rows, so the chunk table flushes fewer
// device requests.  k_debrick adds the scratch into the caller's gradient.
struct BrickLayout {
    int X, Y, Z;
    unsigned sxb, syb;  // floats per brick column step in x (BY * BZ * 16) and y (BZ * 16)
    __device__ __forceinline__ void set(int X_, int Y_, int Z_)
    {
        X = X_; Y = Y_; Z = Z_;
        syb = unsigned((Z + 3) >> 2) * 16u;
        sxb = unsigned((Y + 1) >> 1) * syb;
    }
    __device__ __forceinline__ unsigned xp(int x) const { return unsigned(x >> 1) * sxb + (unsigned(x & 1) << 3); }
    __device__ __forceinline__ unsigned yp(int y) const { return unsigned(y >> 1) * syb + (unsigned(y & 1) << 2); }
    __device__ __forceinline__ unsigned zp(int z) const { return (unsigned(z >> 2) << 4) + unsigned(z & 3); }
    __device__ __forceinline__ CornerIdx cell(int ix, int iy, int iz) const
    {
        int x0, x1, y0, y1, z0, z1;
        cell_axis(ix, X, x0, x1);
        cell_axis(iy, Y, y0, y1);
        cell_axis(iz, Z, z0, z1);
        return CornerIdx{{xp(x0), xp(x1)}, {yp(y0), yp(y1)}, {zp(z0), zp(z1)}};
    }
    __device__ __forceinline__ unsigned point(int x, int y, int z) const { return xp(x) + yp(y) + zp(z); }
};

struct VolDirect {
    static constexpr bool kAgg = false;
    float* p;
    LinearLayout L;
    __device__ __forceinline__ void add(unsigned off, float v)
    {
        if (v != 0.f) gatomic_add(p, off, v);
    }
    __device__ __forceinline__ void add2(unsigned o0, float v0, unsigned o1, float v1)
    {
        add(o0, v0);
        add(o1, v1);
    }
    template <int N>
    __device__ __forceinline__ void add_n(const unsigned (&o)[N], const float (&v)[N])
    {
#pragma unroll
        for (int k = 0; k < N; ++k) add(o[k], v[k]);
    }
    __device__ __forceinline__ void maybe_flush() {}
};

template <class Layout>
struct VolAgg {
    static constexpr bool kAgg = true;
    unsigned lg;      // log2 of the slot count (uniform)
    Layout L;         // element offsets of the gradient the table adds into
    unsigned* tags;   // [slots] chunk index, kAggEmpty = free
    unsigned* list;   // [slots] occupied slots, listed by the flush
    float* vals;      // [slots][16]
    float* abase;     // gradient base rounded down to 64 bytes (never dereferenced below the base)
    unsigned boff;    // floats from abase to the gradient base

    __device__ __forceinline__ unsigned slots() const { return 1u << lg; }
    // every thread of the (one-wave) workgroup; lds is 16-byte aligned, agg_bytes(log2) long
    __device__ __forceinline__ void init(void* lds, int log2)
    {
        lg = unsigned(log2);
        const unsigned kSlots = slots();
        tags = (unsigned*)lds;
        list = tags + kSlots;
        vals = (float*)(list + kSlots + 4);  // 16 bytes of padding after the list
        for (int e = threadIdx.x; e < kSlots; e += blockDim.x) tags[e] = kAggEmpty;
        for (int e = threadIdx.x; e < kSlots * 16; e += blockDim.x) vals[e] = 0.f;
    }
    // the table must be empty (flushed) when the base changes
    __device__ __forceinline__ void bind(float* g)
    {
        const unsigned long long a = (unsigned long long)g;
        abase = (float*)(a & ~63ull);
        boff = unsigned(a & 63ull) >> 2;
    }
    __device__ __forceinline__ unsigned home(unsigned q) const { return (q * 0x9E3779B1u) >> (32u - lg); }
    // Linear probing, only for the lanes whose slot holds another chunk: nested
    // rather than a loop, so each level costs one exec-mask save and restore and
    // the levels below are skipped as soon as no lane needs them.
    template <int P>
    __device__ __forceinline__ void probe(unsigned q, unsigned& s, unsigned& old)
    {
        if constexpr (P < kAggProbe) {
            if (old != kAggEmpty && old != q) {
                s = (s + 1) & (slots() - 1u);
                old = atomicCAS(&tags[s], kAggEmpty, q);
                probe<P + 1>(q, s, old);
            }
        }
    }
    // The slot of chunk q: its home slot or the next free or matching one within
    // kAggProbe probes; kAggEmpty when there is none.
    __device__ __forceinline__ unsigned claim(unsigned q)
    {
        unsigned s = home(q);
        unsigned old = atomicCAS(&tags[s], kAggEmpty, q);
        probe<1>(q, s, old);
        return (old == kAggEmpty || old == q) ? s : kAggEmpty;
    }
    __device__ __forceinline__ void put(unsigned s, unsigned e, float v)
    {
        if (s != kAggEmpty) atomicAdd(&vals[s * 16 + (e & 15)], v);
        else gatomic_add(abase, e, v);
    }
    __device__ __forceinline__ void add(unsigned off, float v)
    {
        if (v == 0.f) return;
        const unsigned e = boff + off;
        put(claim(e >> 4), e, v);
    }
    // N values in pairs (2k, 2k+1), usually z-neighbours in one chunk (then one
    // claim per pair).  The pairs are claimed one after the other: issuing all
    // claims of a cell change before waiting for the first measured the same
    // time and held more registers.
    template <int N>
    __device__ __forceinline__ void add_n(const unsigned (&o)[N], const float (&v)[N])
    {
        static_assert(N % 2 == 0, "pairs");
#pragma unroll
        for (int p = 0; p < N / 2; ++p) {
            const unsigned e0 = boff + o[2 * p], e1 = boff + o[2 * p + 1];
            const bool same = (e0 >> 4) == (e1 >> 4);
            const bool n0 = v[2 * p] != 0.f, n1 = v[2 * p + 1] != 0.f;
            unsigned s0 = kAggEmpty, s1 = kAggEmpty;
            if (n0 || (n1 && same)) s0 = claim(e0 >> 4);
            if (n1 && !same) s1 = claim(e1 >> 4);
            if (same) s1 = s0;
            if (n0) put(s0, e0, v[2 * p]);
            if (n1) put(s1, e1, v[2 * p + 1]);
        }
    }
    __device__ __forceinline__ void add2(unsigned o0, float v0, unsigned o1, float v1)
    {
        const unsigned o[2] = {o0, o1};
        const float v[2] = {v0, v1};
        add_n<2>(o, v);
    }
    __device__ __forceinline__ void flush()
    {
        const unsigned long long act = __ballot(1);
        const unsigned cnt = unsigned(__popcll(act));
        const unsigned rank = __builtin_amdgcn_mbcnt_hi(unsigned(act >> 32), __builtin_amdgcn_mbcnt_lo(unsigned(act), 0u));
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        // list the occupied slots
        unsigned n = 0u;
        const unsigned kSlots = slots();
        for (unsigned t0 = 0; t0 < kSlots; t0 += cnt) {
            const unsigned t = t0 + rank;
            const bool occ = t < kSlots && tags[t] != kAggEmpty;
            const unsigned long long m = __ballot(occ);
            if (occ) {
                const unsigned below = __builtin_amdgcn_mbcnt_hi(unsigned(m >> 32), __builtin_amdgcn_mbcnt_lo(unsigned(m), 0u));
                list[n + below] = t;
            }
            n += unsigned(__popcll(m));
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        for (unsigned e = rank; e < n * 16u; e += cnt) {
            const unsigned s = list[e >> 4], j = e & 15u;
            const float v = vals[s * 16 + j];
            if (v != 0.f) gatomic_add(abase, tags[s] * 16u + j, v);
            vals[s * 16 + j] = 0.f;
        }
        for (unsigned e = rank; e < n; e += cnt) tags[list[e]] = kAggEmpty;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    }
    // flush when more than half of the slots are taken: the active lanes count the tags
    __device__ __forceinline__ void maybe_flush()
    {
        const unsigned long long act = __ballot(1);
        const unsigned cnt = unsigned(__popcll(act));
        const unsigned rank = __builtin_amdgcn_mbcnt_hi(unsigned(act >> 32), __builtin_amdgcn_mbcnt_lo(unsigned(act), 0u));
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        const unsigned kSlots = slots();
        unsigned n = 0u;
        for (unsigned t0 = 0; t0 < kSlots; t0 += cnt) {
            const unsigned t = t0 + rank;
            n += unsigned(__popcll(__ballot(t < kSlots && tags[t] != kAggEmpty)));
        }
        if (n > (kSlots >> 1)) flush();
    }
};

struct CornerCache {
    int cx, cy, cz;   // current cell, kInvalidCell = none
    float w[8];       // accumulated corner gradients (volume adjoint)
    float d[8];       // corner densities (position adjoint)
};

// Move across the face normal to axis bit B, towards +1 (UP) or -1: the four
// leaving corners' offsets and values go to lo/lv[0..3] (pairs along z unless
// B is z), the shared corners' state moves over, the fresh corners start empty.
template <int B, bool UP, bool VOL, bool POS>
__device__ __forceinline__ void corner_shift(CornerCache& c, const Vol& V, const CornerIdx& oldG,
                                             const CellOffsets& newV, unsigned (&lo)[8], float (&lv)[8])
{
    int m = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        if (k & B) continue;
        const int f = UP ? k : (k | B);   // leaves; takes the shared corner's state
        const int s = UP ? (k | B) : k;   // the new cell's fresh corner
        if constexpr (VOL) {
            lo[m] = corner_at(oldG, f);
            lv[m] = c.w[f];
            c.w[f] = c.w[s];
            c.w[s] = 0.f;
        }
        if constexpr (POS) {
            c.d[f] = c.d[s];
            c.d[s] = gld(V.p, corner_off(newV, s));
        }
        ++m;
    }
}

template <bool VOL, bool POS, class VS>
__device__ __forceinline__ void corner_flush_all(CornerCache& c, const Vol& V, VS& vs)
{
    if constexpr (VOL) {
        if (c.cx != kInvalidCell) {
            const CornerIdx o = vs.L.cell(c.cx, c.cy, c.cz);
            unsigned lo[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) lo[k] = corner_at(o, k);
            vs.add_n(lo, c.w);
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) c.w[k] = 0.f;
    }
}

// The leaving corners are collected by the (divergent) direction cases and
// handed to the sink once, after them: the sink's insertion code then runs
// once per cell change instead of once per direction case.
template <bool VOL, bool POS, class VS>
__device__ __forceinline__ void corner_move(CornerCache& c, const Vol& V, VS& vs, int ix, int iy, int iz)
{
    const int ddx = ix - c.cx, ddy = iy - c.cy, ddz = iz - c.cz;
    const bool valid = c.cx != kInvalidCell;
    const bool unit = valid && (abs(ddx) + abs(ddy) + abs(ddz)) == 1;
    // written by the case a lane takes before it reads them (the upper four only by the jumps)
    unsigned lo[8];
    float lv[8];
    if (unit) {
        CornerIdx oldG{};
        if constexpr (VOL) oldG = vs.L.cell(c.cx, c.cy, c.cz);
        const CellOffsets newV = cell_offsets(ix, iy, iz, V.X, V.Y, V.Z, V.sx, V.sy, V.sz);
        if (ddx > 0) corner_shift<4, true, VOL, POS>(c, V, oldG, newV, lo, lv);
        else if (ddx < 0) corner_shift<4, false, VOL, POS>(c, V, oldG, newV, lo, lv);
        else if (ddy > 0) corner_shift<2, true, VOL, POS>(c, V, oldG, newV, lo, lv);
        else if (ddy < 0) corner_shift<2, false, VOL, POS>(c, V, oldG, newV, lo, lv);
        else if (ddz > 0) corner_shift<1, true, VOL, POS>(c, V, oldG, newV, lo, lv);
        else corner_shift<1, false, VOL, POS>(c, V, oldG, newV, lo, lv);
    } else {
        if constexpr (VOL) {
            // a ray's first cell (no current cell yet) holds zero gradients, which claim and add
            // nothing; its corner offsets are formed from the clamped invalid index and never used
            const CornerIdx o = vs.L.cell(valid ? c.cx : 0, valid ? c.cy : 0, valid ? c.cz : 0);
#pragma unroll
            for (int k = 0; k < 8; ++k) { lo[k] = corner_at(o, k); lv[k] = c.w[k]; }
#pragma unroll
            for (int k = 0; k < 8; ++k) c.w[k] = 0.f;
        }
        if constexpr (POS) cell_load(V, ix, iy, iz, c.d);
    }
    if constexpr (VOL) {
        // the four leaving corners of a move across one face; all eight for a jump
        {
            const unsigned lo4[4] = {lo[0], lo[1], lo[2], lo[3]};
            const float lv4[4] = {lv[0], lv[1], lv[2], lv[3]};
            vs.add_n(lo4, lv4);
        }
        if (!unit) {
            const unsigned lo4[4] = {lo[4], lo[5], lo[6], lo[7]};
            const float lv4[4] = {lv[4], lv[5], lv[6], lv[7]};
            vs.add_n(lo4, lv4);
        }
    }
    c.cx = ix; c.cy = iy; c.cz = iz;
}

// ---------------------------------------------------------------------------
// Reverse march of one ray with transmittance checkpoints every K*ckEvery
// samples, walked back in register segments of K samples.  GM is the set of gradient
// kinds compiled in: bit0 TF, bit1 volume, bit2 position (camera / step);
// the runtime flags select among them.
// ---------------------------------------------------------------------------
template <int FILTER, int TFM, int BLEND, int K, int GM, class Sink, bool SPLIT = false, class VS = VolDirect>
__device__ __forceinline__ void march_reverse(const Vol& V, const Tf& T, const Ray& r, int n, f4 adj_acc,
                                              const float* ck, int ckS, int cap, int ckEvery, Sink& sink,
                                              bool wantTf,
                                              bool wantVol, bool wantPos, VS& vs,
                                              f3& adj_start, f3& adj_dir, float& adj_step,
                                              const float* dsr = nullptr, int dsS = 0)
{
    constexpr bool CT_TF = (GM & 1) != 0, CT_VOL = (GM & 2) != 0, CT_POS = (GM & 4) != 0;
    wantTf = CT_TF && wantTf;
    wantVol = CT_VOL && wantVol;
    wantPos = CT_POS && wantPos;
    // Beer-Lambert segments of up to 16 samples also keep each sample's (alpha,
    // keep) from the recompute, so the backward walk evaluates no exponential
    // (not in the volume-gradient instances without the position adjoint: they run more waves
    // per SIMD and need the registers)
    constexpr bool kVolOnly = CT_VOL && !CT_POS;
    constexpr bool kKeepAK = BLEND == 0 && K <= 16 && !kVolOnly;
    // a linear TF's interval row per sample, located once by the recompute and reused by the walk
    // (in the volume-gradient instances, whose walk is a rolled loop): two 16-bit rows per
    // register (a linear TF has at most kMaxTfLds / 5 < 65536 rows)
    constexpr bool kKeepRow = TFM == 2 && kVolOnly;
    static_assert(kMaxTfLds / 5 < 65536, "TF rows must fit 16 bits");
    float dv[K], tin[K];
    float al[kKeepAK ? K : 1], kp[kKeepAK ? K : 1];
    unsigned rows[kKeepRow ? (K + 1) / 2 : 1];
    // reverse-walk cell state for the volume / position adjoint
    CornerCache cc;
    cc.cx = kInvalidCell; cc.cy = 0; cc.cz = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) { cc.w[k] = 0.f; cc.d[k] = 0.f; }
    int nLast = -1;
    float nAcc = 0.0f;

    const int nseg = (n + K - 1) / K;
    for (int seg = nseg - 1; seg >= 0; --seg) {
        const int i0 = seg * K;
        const int cnt = min(K, n - i0);
        // transmittance before sample i0
        float Tr = 1.0f;
        {
            // checkpoints hold T after every K*ckEvery samples; recompute from the
            // last one before i0 (ckEvery > 1, or a ray longer than the capacity)
            const int c = min(seg / ckEvery, cap);
            if (c > 0) Tr = ck[(c - 1) * ckS];
            Cell cc;
            cell_invalidate(cc);
            for (int i = c * K * ckEvery; i < i0; ++i) {
                const float t = r.t0 + float(i) * r.step;
                const float d = dsr ? *(const gfloat*)(dsr + i * dsS) : fetch<FILTER>(V, sample_pos(r, t), cc);
                float alpha, keep;
                alpha_keep<BLEND>(tf_absorption<TFM, true>(T, d), r.step, alpha, keep);
                Tr *= keep;
            }
        }
        // recompute the segment forward: densities from the forward march's store
        // (all loads of the segment in flight at once), or re-interpolated with the
        // fetches pipelined one sample ahead; keep density and transmittance-before
        if (dsr) {
            // no per-sample predicate: indices past the ray's last sample re-read that sample
            // (in bounds), and the values formed for them are never used (the walk skips
            // j >= cnt; Tr restarts from a checkpoint every segment)
#pragma unroll
            for (int j = 0; j < K; ++j) dv[j] = *(const gfloat*)(dsr + min(i0 + j, n - 1) * dsS);
#pragma unroll
            for (int j = 0; j < K; ++j) {
                tin[j] = Tr;
                float alpha, keep;
                if constexpr (kKeepRow) {
                    const TfLoc L = tf_linear_locate<true>(T, dv[j]);
                    if (j & 1) rows[j >> 1] |= unsigned(L.i) << 16;
                    else rows[j >> 1] = unsigned(L.i);
                    alpha_keep<BLEND>(tf_absorption<TFM, true>(T, dv[j], &L), r.step, alpha, keep);
                } else {
                    alpha_keep<BLEND>(tf_absorption<TFM, true>(T, dv[j]), r.step, alpha, keep);
                }
                if constexpr (kKeepAK) { al[j] = alpha; kp[j] = keep; }
                Tr *= keep;
            }
        } else {
            FetchPipe<FILTER> pipe;
            f3 p = sample_pos(r, r.t0 + float(i0) * r.step);
            pipe.start(V, p);
#pragma unroll
            for (int j = 0; j < K; ++j) {
                if (j < cnt) {
                    dv[j] = pipe.value(p);
                    const f3 pn = sample_pos(r, r.t0 + float(i0 + j + 1) * r.step);
                    const bool more = j + 1 < cnt;
                    if (more) pipe.request(V, pn);
                    tin[j] = Tr;
                    float alpha, keep;
                    if constexpr (kKeepRow) {
                        const TfLoc L = tf_linear_locate<true>(T, dv[j]);
                        if (j & 1) rows[j >> 1] |= unsigned(L.i) << 16;
                        else rows[j >> 1] = unsigned(L.i);
                        alpha_keep<BLEND>(tf_absorption<TFM, true>(T, dv[j], &L), r.step, alpha, keep);
                    } else {
                        alpha_keep<BLEND>(tf_absorption<TFM, true>(T, dv[j]), r.step, alpha, keep);
                    }
                    if constexpr (kKeepAK) { al[j] = alpha; kp[j] = keep; }
                    Tr *= keep;
                    if (more) pipe.advance();
                    p = pn;
                }
            }
        }
        // walk the segment backwards; with SPLIT, a segment that is full on every active lane
        // of the wave (all but a ray's last) runs without a per-sample predicate (a second
        // copy of the unrolled walk)
        auto walk = [&](int j) __attribute__((always_inline)) {
            const int i = i0 + j;
            const float t = r.t0 + float(i) * r.step;
            // a linear TF's interval and the Gaussian weights are formed once for the value and the adjoint
            TfPre pre;
            if constexpr (kKeepRow) pre.L = tf_linear_row<true>(T, int((rows[j >> 1] >> ((j & 1) * 16)) & 0xffffu));
            else tf_prepare<TFM, true>(T, dv[j], pre);
            const f4 s = tf_eval<TFM, true>(T, dv[j], &pre);
            f4 adj_s;
            float adj_st;
            if constexpr (kKeepAK) blend_adjoint_ak<BLEND>(tin[j], s, r.step, al[j], kp[j], adj_acc, adj_s, adj_st);
            else blend_adjoint_t<BLEND>(tin[j], s, r.step, adj_acc, adj_s, adj_st);
            float adj_d;
            if (wantTf) adj_d = tf_adjoint<TFM, CT_TF, Sink, true>(T, dv[j], adj_s, sink, &pre);
            else {
                SinkNone none;
                adj_d = tf_adjoint<TFM, false, SinkNone, true>(T, dv[j], adj_s, none, &pre);
            }
            if constexpr (CT_VOL || CT_POS) {
                if (wantVol || wantPos) {
                    const f3 p = sample_pos(r, t);
                    f3 adj_p = mk3(0.f, 0.f, 0.f);
                    if constexpr (FILTER == 0) {
                        if (wantVol) {
                            int gx, gy, gz;
                            nearest_index(p, V.X, V.Y, V.Z, gx, gy, gz);
                            const int goff = int(vs.L.point(gx, gy, gz));
                            if (goff != nLast) {
                                if (nLast >= 0) vs.add(unsigned(nLast), nAcc);
                                nLast = goff;
                                nAcc = 0.f;
                            }
                            nAcc += adj_d;
                        }
                    } else {
                        const int ix = int(p.x), iy = int(p.y), iz = int(p.z);
                        if (ix != cc.cx || iy != cc.cy || iz != cc.cz) {
                            if (wantVol && wantPos) corner_move<true, true>(cc, V, vs, ix, iy, iz);
                            else if (wantVol) corner_move<true, false>(cc, V, vs, ix, iy, iz);
                            else corner_move<false, true>(cc, V, vs, ix, iy, iz);
                        }
                        const float fx = p.x - float(ix), fy = p.y - float(iy), fz = p.z - float(iz);
                        float w[8];
                        if (wantPos) tri_adjoint<true>(cc.d, fx, fy, fz, adj_d, w, adj_p);
                        else tri_adjoint<false>(cc.d, fx, fy, fz, adj_d, w, adj_p);
                        if (wantVol) {
#pragma unroll
                            for (int k = 0; k < 8; ++k) cc.w[k] += w[k];
                        }
                    }
                    if (wantPos) {
                        const f3 adj_w = mul3(adj_p, r.invVoxel);
                        adj_start = add3(adj_start, adj_w);
                        adj_dir = add3(adj_dir, scl3(adj_w, t));
                        adj_step += dot3(adj_w, r.dir) * float(i);
                    }
                }
            }
            adj_step += adj_st;
            if constexpr (VS::kAgg && CT_VOL) {
                if ((j & 3) == 0 && wantVol) vs.maybe_flush();
            }
        };
        // Volume-gradient instances (without the position adjoint) walk the segment as a loop:
        // their per-sample body (cell changes, chunk-table claims) is long, and unrolled K times
        // it made the C5 instance 143 KB of code; rolled it is 35 KB at the same speed (register
        // arrays indexed by j become selects) and needs fewer registers.
        if constexpr (kVolOnly) {
#pragma unroll 1
            for (int j = K - 1; j >= 0; --j)
                if (j < cnt) walk(j);
        } else {
            bool full = false;
            if constexpr (SPLIT) full = __all(cnt == K);
            if (full) {
#pragma unroll
                for (int j = K - 1; j >= 0; --j) walk(j);
            } else {
#pragma unroll
                for (int j = K - 1; j >= 0; --j)
                    if (j < cnt) walk(j);
            }
        }
    }
    if constexpr (CT_VOL) {
        if (wantVol) {
            if constexpr (FILTER == 0) {
                if (nLast >= 0) vs.add(unsigned(nLast), nAcc);
            } else {
                corner_flush_all<true, false>(cc, V, vs);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
template <int FILTER, int TFM, int BLEND>
__global__ __launch_bounds__(kBlock) void k_forward_tiled(RenderParams P, CkptParams CK, int tilesX, int tilesPerView,
                                                          int nTiles)
{
    extern __shared__ float lds[];
    const TileXY tc = tile_xy(tilesX, tilesPerView, nTiles);
    stage_tf(P, tc.b, lds);
    __syncthreads();
    if (tc.x >= P.W || tc.y >= P.H) return;
    const Ray r = setup_ray(P, tc.x, tc.y, tc.b);
    const Vol V = make_vol(P, tc.b);
    constexpr int kC = tf_cols<TFM>();
    const Tf T{lds, P.R, kC, kC, 1};
    float* ck = CK.buf ? CK.buf + (size_t(tc.tile) * CK.cap) * kBlock + threadIdx.x : nullptr;
    float* ds = CK.ds ? CK.ds + (size_t(tc.tile) * CK.dsCap) * kBlock + threadIdx.x : nullptr;
    Acc acc;
    int n;
    march_forward<FILTER, TFM, BLEND>(V, T, r, P.earlyOut, -1, acc, n, ck, kBlock, CK.K * CK.every, CK.cap, ds, kBlock,
                                      CK.dsCap);
    float* o = P.color + tc.b * P.colSb + tc.y * P.colSy + tc.x * P.colSx;
    o[0] = acc.r; o[P.colSc] = acc.g; o[2 * P.colSc] = acc.b; o[3 * P.colSc] = acc.A;
    P.term[tc.b * P.termSb + tc.y * P.termSy + tc.x * P.termSx] = n;
}

// Volume + TF gradient instances (GM 3) are compiled for four waves per SIMD
// (the walk is latency-bound, as in the batched kernel); the others keep the
// compiler's default register budget.
template <int FILTER, int TFM, int BLEND, int K, int GM>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(GM == 3 ? 4 : 1))) void k_adjoint_tiled(
    RenderParams P, AdjointParams A, CkptParams CK, LossParams L, int fused, int ckFromFwd, int usePriv, int tilesX,
    int tilesPerView, int nTiles, int aggLog2)
{
    extern __shared__ float lds[];
    const int RC = P.R * P.C;
    float* s_tf = lds;
    float* s_g = lds + RC;
    const TileXY tc = tile_xy(tilesX, tilesPerView, nTiles);
    const bool wantTf = A.atf != nullptr;
    const int ng = wantTf ? (usePriv ? RC * kBlock : RC) : 0;
    stage_tf(P, tc.b, s_tf);
    for (int e = threadIdx.x; e < ng; e += kBlock) s_g[e] = 0.0f;
    // volume gradient: the wave's corner values are summed per 64-byte chunk in LDS
    constexpr bool kVolInst = (GM & 2) != 0;
    using VS = std::conditional_t<kVolInst, VolAgg<LinearLayout>, VolDirect>;
    VS vs;
    vs.L = LinearLayout{P.X, P.Y, P.Z, A.avSx, A.avSy, A.avSz};
    const bool aggOn = kVolInst && A.avol != nullptr;
    if constexpr (kVolInst) {
        if (aggOn) {
            vs.init(lds + agg_offset(RC, ng), aggLog2);
            vs.bind(A.avol + tc.b * A.avSb);
        }
    }
    __syncthreads();

    SinkTiled<TFM> sink;
    sink.init(s_g, tf_cols<TFM>(), usePriv != 0);
    float lossLocal = 0.0f;
    const bool valid = tc.x < P.W && tc.y < P.H;
    if (valid) {
        const int b = tc.b, x = tc.x, y = tc.y;
        const Ray r = setup_ray(P, x, y, b);
        const Vol V = make_vol(P, b);
        constexpr int kC = tf_cols<TFM>();
        const Tf T{s_tf, P.R, kC, kC, 1};
        float* ck = CK.buf + (size_t(tc.tile) * CK.cap) * kBlock + threadIdx.x;
        // density store: written by this kernel's forward march (fused, or replayed without
        // checkpoints) or by the forward launch that wrote the checkpoints
        float* ds = CK.ds ? CK.ds + (size_t(tc.tile) * CK.dsCap) * kBlock + threadIdx.x : nullptr;
        int n;
        f4 adj_acc;
        if (fused) {
            Acc acc;
            march_forward<FILTER, TFM, BLEND>(V, T, r, P.earlyOut, -1, acc, n, ck, kBlock, K * CK.every, CK.cap, ds,
                                              kBlock, CK.dsCap);
            float* o = P.color + b * P.colSb + y * P.colSy + x * P.colSx;
            o[0] = acc.r; o[P.colSc] = acc.g; o[2 * P.colSc] = acc.b; o[3 * P.colSc] = acc.A;
            P.term[b * P.termSb + y * P.termSy + x * P.termSx] = n;
            const float* tg = L.target + b * L.tgSb + y * L.tgSy + x * L.tgSx;
            const float d0 = acc.r - tg[0], d1 = acc.g - tg[L.tgSc], d2 = acc.b - tg[2 * L.tgSc],
                        d3 = acc.A - tg[3 * L.tgSc];
            if (L.kind == DVR_LOSS_L1) {
                auto sgn = [](float v) { return v > 0.f ? 1.f : (v < 0.f ? -1.f : 0.f); };
                adj_acc = mk4(sgn(d0) * L.scale, sgn(d1) * L.scale, sgn(d2) * L.scale, sgn(d3) * L.scale);
                lossLocal = (fabsf(d0) + fabsf(d1) + fabsf(d2) + fabsf(d3)) * L.scale;
            } else {
                adj_acc = mk4(2.f * d0 * L.scale, 2.f * d1 * L.scale, 2.f * d2 * L.scale, 2.f * d3 * L.scale);
                lossLocal = (d0 * d0 + d1 * d1 + d2 * d2 + d3 * d3) * L.scale;
            }
        } else {
            n = min(A.fterm[b * A.ftSb + y * A.ftSy + x * A.ftSx], kMaxSamples);
            const float* ac = A.acol + b * A.acSb + y * A.acSy + x * A.acSx;
            adj_acc = mk4(ac[0], ac[A.acSc], ac[2 * A.acSc], ac[3 * A.acSc]);
            if (!ckFromFwd) {
                Acc acc;
                int n2;
                march_forward<FILTER, TFM, BLEND>(V, T, r, P.earlyOut, n, acc, n2, ck, kBlock, K * CK.every, CK.cap,
                                                  ds, kBlock, CK.dsCap);
            }
        }
        f3 adj_start = mk3(0.f, 0.f, 0.f), adj_dir = mk3(0.f, 0.f, 0.f);
        float adj_step = 0.0f;
        const bool wantVol = A.avol != nullptr;
        const bool wantPos = A.ars != nullptr || A.ast != nullptr;
        const bool stored = ds != nullptr && __all(n <= CK.dsCap);  // uniform over the wave
        // the split walk (a second, predicate-free copy for segments full on every lane) pays off
        // except for the arithmetic-bound Gaussian TF
        if constexpr (!kVolInst) vs.p = nullptr;
        march_reverse<FILTER, TFM, BLEND, K, GM, SinkTiled<TFM>, TFM != 3, VS>(V, T, r, n, adj_acc, ck, kBlock, CK.cap,
                                                                               CK.every, sink, wantTf, wantVol,
                                                 wantPos, vs, adj_start, adj_dir, adj_step, stored ? ds : nullptr,
                                                 kBlock);
        if (A.ast) {
            float* o = A.ast + b * A.astSb + y * A.astSy + x * A.astSx;
            if (A.stepBroadcast) atomicAdd(o, adj_step);
            else *o = adj_step;
        }
        if (A.ars) {
            float* os = A.ars + b * A.arsSb + y * A.arsSy + x * A.arsSx;
            float* od = A.ard + b * A.ardSb + y * A.ardSy + x * A.ardSx;
            if (A.camBroadcast) {
                atomicAdd(os, adj_start.x); atomicAdd(os + A.arsSc, adj_start.y); atomicAdd(os + 2 * A.arsSc, adj_start.z);
                atomicAdd(od, adj_dir.x); atomicAdd(od + A.ardSc, adj_dir.y); atomicAdd(od + 2 * A.ardSc, adj_dir.z);
            } else {
                os[0] = adj_start.x; os[A.arsSc] = adj_start.y; os[2 * A.arsSc] = adj_start.z;
                od[0] = adj_dir.x; od[A.ardSc] = adj_dir.y; od[2 * A.ardSc] = adj_dir.z;
            }
        }
    }
    if constexpr (kVolInst) {
        if (aggOn) vs.flush();
    }
    if (wantTf) sink.template finish<TFM == 2 ? 5 : 4>();
    if (fused && L.lossOut) {
        const float s = wave_sum(lossLocal);
        if ((threadIdx.x & 63) == 0) atomicAdd(L.lossOut, s);
    }
    if (wantTf) {
        __syncthreads();
        float* g = A.atf + tc.b * A.atfSb;
        for (int e = threadIdx.x; e < RC; e += kBlock) {
            float v;
            if (usePriv) {
                v = 0.0f;
                for (int t = 0; t < kBlock; ++t) v += s_g[e * kBlock + t];
            } else {
                v = s_g[e];
            }
            const int rr = e / P.C, cc = e - rr * P.C;
            if (v != 0.0f) atomicAdd(g + rr * A.atfSr + cc * A.atfSc, v);
        }
    }
}

// ---------------------------------------------------------------------------
// launch of one instance
// ---------------------------------------------------------------------------
template <int FILTER, int TFM, int BLEND, int K, int GM>
static hipError_t adj_tiled_k(const RenderParams& P, const AdjointParams& A, const CkptParams& CK, const LossParams& L,
                              bool fused, bool ckFromFwd, hipStream_t s)
{
    const int tilesX = (P.W + kTileW - 1) / kTileW, tilesY = (P.H + kTileH - 1) / kTileH;
    const int nTiles = tilesX * tilesY * P.B;
    if (nTiles == 0) return hipSuccess;
    const int RC = P.R * P.C;
    const bool usePriv = (TFM == 3) && A.atf && RC <= 48;  // private slices <= 48 KiB of LDS
    const int ng = A.atf ? (usePriv ? RC * kBlock : RC) : 0;
    int aggLog2 = kAggLog2;
    size_t lds;
    if ((GM & 2) && A.avol) {
        const size_t base = sizeof(float) * size_t(agg_offset(RC, ng));
        if (GM == 3) {
            // the largest chunk table beside the TF tables that leaves room for sixteen one-wave
            // workgroups per CU (the four waves per SIMD the instance's registers allow)
            while (aggLog2 > kAggMinLog2 && base + agg_bytes(aggLog2) > size_t(kLdsPerCu) / kTiledVolWgPerCu)
                --aggLog2;
            lds = base + agg_bytes(aggLog2);
        } else {
            lds = occupancy_lds(base + agg_bytes(aggLog2), tiled_wg_per_cu(TFM));
        }
    } else {
        lds = occupancy_lds(sizeof(float) * (size_t(RC) + size_t(ng)), tiled_wg_per_cu(TFM));
    }
    hipLaunchKernelGGL((k_adjoint_tiled<FILTER, TFM, BLEND, K, GM>), dim3(nTiles), dim3(kBlock), lds, s, P, A, CK, L,
                       fused ? 1 : 0, ckFromFwd ? 1 : 0, usePriv ? 1 : 0, tilesX, tilesX * tilesY, nTiles, aggLog2);
    return hipGetLastError();
}

// gradient kinds compiled into the instance: TF only, volume (+TF), position (+TF), or everything
// (camera / step gradients without volume gradients run an instance compiled
// without the volume-gradient code: fewer registers)
inline int tiled_grad_mask(const AdjointParams& A)
{
    const bool pos = A.ars || A.ast;
    return pos ? (A.avol ? 7 : 5) : (A.avol ? 3 : 1);
}

template <int FILTER, int BLEND, int K>
static hipError_t adj_tiled_fbk(const RenderParams& P, const AdjointParams& A, const CkptParams& CK,
                                const LossParams& L, bool fused, bool ckFromFwd, int tfm, hipStream_t s)
{
    const int gm = tiled_grad_mask(A);
#define DVR_GM_CASES(TV)                                                                              \
    if (gm == 1) return adj_tiled_k<FILTER, TV, BLEND, K, 1>(P, A, CK, L, fused, ckFromFwd, s);      \
    if (gm == 3) return adj_tiled_k<FILTER, TV, BLEND, K, 3>(P, A, CK, L, fused, ckFromFwd, s);      \
    if (gm == 5) return adj_tiled_k<FILTER, TV, BLEND, K, 5>(P, A, CK, L, fused, ckFromFwd, s);      \
    return adj_tiled_k<FILTER, TV, BLEND, K, 7>(P, A, CK, L, fused, ckFromFwd, s);
    switch (tfm) {
    case 0: { DVR_GM_CASES(0) }
    case 1: { DVR_GM_CASES(1) }
    case 2: { DVR_GM_CASES(2) }
    case 3: { DVR_GM_CASES(3) }
    default: return hipErrorInvalidValue;
    }
#undef DVR_GM_CASES
}

}  // namespace dvr
