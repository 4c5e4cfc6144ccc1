// Kernel parameter blocks shared by the host dispatch and the HIP kernels.
// Element strides are int32 (the reference's accessors are 32-bit too,
// renderer/renderer_tensor.cuh:122-269); the host validates that every
// offset fits.  A stride of 0 on a size-1 batch dimension broadcasts it.
#pragma once
#include <stdint.h>

namespace dvr {

constexpr int kTileW = 8;         // pixels per workgroup tile (x)
constexpr int kTileH = 8;         // pixels per workgroup tile (y)
constexpr int kBlock = 64;        // threads per workgroup = one wave of 8x8 rays
constexpr int kBatchBlock = 64;   // batched persistent kernel: one wave per workgroup, one wave tile per pop
constexpr int kTfCopies = 8;      // batched kernel: copies of the LDS TF-gradient table, lane & 7 selects one
constexpr int kNumXcd = 8;        // MI355X: 8 XCDs, workgroups dispatched round-robin
constexpr int kMaxTfLds = 4096;   // max R*C floats of a TF staged in LDS by the tiled kernels
// LDS table of the volume-gradient aggregation (one per wave): 2^log2 slots of
// one 64-byte chunk each, plus tags, slot list and 16 bytes of padding.  The per-call tiled
// adjoint uses kAggLog2 where it also carries the position adjoint; its volume +
// TF instances and a batched plan take the largest table (kAggMinLog2 ..
// kAggLog2) that keeps its instance at batched_vol_waves waves per SIMD.
constexpr int kAggLog2 = 8;
constexpr int kAggMinLog2 = 5;
constexpr size_t agg_bytes(int log2) { return (size_t(1) << log2) * (4 + 4 + 16 * 4) + 16; }
constexpr size_t kAggBytes = agg_bytes(kAggLog2);
// Batched volume-gradient instances: waves per SIMD the compiler must fit
// (registers).  Measured on C5 / C3 (linear / identity TF): the walk is bound by
// latency at two waves per SIMD; four waves ran C5 in 105 ms (64-slot table)
// against 110 ms at three (128 slots) and 136 ms at two (256 slots), C3 in
// 19.6 ms (128 slots) against 21.7 / 24.1 ms, and C3 with a 64-slot table in
// 27 ms: the table shrinks only as far as the TF tables beside it require.
// Instances that also carry the position adjoint (camera gradients) need more
// registers and keep two waves per SIMD.
constexpr int batched_vol_waves(int gm) { return (gm & 4) ? 2 : 4; }
// float offset of the table in dynamic LDS after used0 + used1 floats (16-byte aligned)
constexpr int agg_offset(int used0, int used1) { return (used0 + used1 + 3) & ~3; }
constexpr int kMaxSamples = 1 << 24;  // hard cap on samples per ray (termination guarantee)

struct RenderParams {
    int W, H, B;
    // volume (B) x X x Y x Z
    const float* vol; int X, Y, Z; int vsb, vsx, vsy, vsz;
    // bounding box; pointer NULL -> uniform value
    const float* bmin; int bminSb, bminSc; float bminU[3];
    const float* bsize; int bsizeSb, bsizeSc; float bsizeU[3];
    // camera
    int camMode;
    const float* rs; int rsSb, rsSy, rsSx, rsSc;
    const float* rd; int rdSb, rdSy, rdSx, rdSc;
    const float* cm; int cmSb, cmSr, cmSc; float cmU[16];
    float fovY;
    double tanX, tanY;   // tan(fovX/2), tan(fovY/2) for ReferenceFrame, evaluated on the host in double
    // step size (B) x (H) x (W); pointer NULL -> uniform value
    const float* st; int stSb, stSy, stSx; float stU;
    // transfer function (B) x R x C
    const float* tf; int R, C; int tfSb, tfSr, tfSc;
    float earlyOut;
    // forward outputs
    float* color; int colSb, colSy, colSx, colSc;
    int* term; int termSb, termSy, termSx;
};

struct AdjointParams {
    // forward outputs used as inputs
    const float* fcol; int fcSb, fcSy, fcSx, fcSc;
    const int* fterm; int ftSb, ftSy, ftSx;
    // adjoint of the output colour
    const float* acol; int acSb, acSy, acSx, acSc;
    // adjoint outputs (NULL = not requested)
    float* avol; int avSb, avSx, avSy, avSz;
    float* ast; int astSb, astSy, astSx; int stepBroadcast;
    float* ars; int arsSb, arsSy, arsSx, arsSc;
    float* ard; int ardSb, ardSy, ardSx, ardSc; int camBroadcast;
    float* atf; int atfSb, atfSr, atfSc;
    int tfDelayed;
};

// Checkpoints: ckpt c (0-based) of a ray holds the transmittance 1 - alpha
// accumulated after (c+1)*K*every samples.  Layout [tile][c][thread] so one
// wave stores 256 B.  The reverse march walks K-sample register segments.
struct CkptParams {
    float* buf;   // NULL = none
    int K;        // register segment length (8, 16, 32)
    int cap;      // checkpoints per ray
    int every;    // checkpoint interval in segments (1 or 2): interval = K*every
    float* ds;    // density store [tile][dsCap][kBlock] of the forward march, or NULL
    int dsCap;    // samples per ray it holds
};

// Per-pixel loss whose gradient the fused forward+adjoint kernel forms in
// registers between the forward and the reverse march.
#define DVR_LOSS_NONE 0   // adjoint colour read from AdjointParams::acol
#define DVR_LOSS_L1 1     // mean |color - target|   (torch.nn.L1Loss)
#define DVR_LOSS_L2 2     // mean (color - target)^2 (torch.nn.MSELoss)

struct LossParams {
    int kind;
    const float* target; int tgSb, tgSy, tgSx, tgSc;
    float scale;          // 1 / number of loss elements
    float* lossOut;       // optional scalar accumulator (atomic), may be NULL
};


// ---------------------------------------------------------------------------
// Batched persistent launch: many views (cameras) of many volumes and TFs in
// one kernel.  The views are cut into workgroup tiles; tiles are numbered
// view after view and split into one contiguous range per XCD.  Persistent
// workgroups pop tiles from their XCD's queue and steal from the others when
// it runs dry.  Volumes and TFs are contiguous (x slowest, z fastest; row
// major R x C).
// ---------------------------------------------------------------------------
struct BView {
    int vol, tf, W, H;
    int tileBegin, tilesX, camMode, pad0;
    double tanX, tanY;          // ReferenceFrame tangents (host, double)
    const float* cam;           // 3x3 viewport rows (eye, right, up) or 4x4 row-major matrix
    float step, lossScale;
    float* color;               // H x W x 4
    int* term;                  // H x W
    const float* target;        // H x W x 4, loss target (loss kinds L1/L2)
    const float* acol;          // H x W x 4, given adjoint colour (loss kind none)
    float* acam;                // 3 x 3 viewport adjoint (accumulated) or NULL
};

struct BVol {
    const float* p;
    float* adj;                 // gradient (accumulated) or NULL
    float* brick;               // with adj: bricked gradient scratch of the batched kernel (2x2x4-voxel chunks)
    int X, Y, Z, pad0;
    float bmin[3], bsize[3];
};

// Bricked scratch of one volume gradient, added into the caller's gradient after the batched kernel
struct DebrickSeg {
    const float* brick;
    float* adj;
    int X, Y, Z, pad0;
};

struct BTf {
    const float* p;
    float* adj;                 // gradient (accumulated) or NULL
    int R, C;
};

struct BatchParams {
    const BView* views;
    const BVol* vols;
    const BTf* tfs;
    unsigned* queues;           // kNumXcd tile counters, zeroed before every launch
    float* ck;                  // checkpoints [workgroup][cap][kBatchBlock]
    float* ds;                  // sample densities [workgroup][dsCap][kBatchBlock] of the forward march, or NULL
    int dsCap;                  // samples per ray the density store holds
    float* lossOut;             // scalar loss accumulator or NULL
    int nViews, nTiles, ckCap;
    int ckEvery;                // checkpoint interval = K * ckEvery samples
    int qBegin[kNumXcd + 1];    // tile range of each XCD's queue
    float earlyOut;
    int lossKind;               // DVR_LOSS_*
    int wantTf, wantVol, wantCam;
    int waveShift;              // log2 of the wave tile width; wave tile = (1<<s) x (64>>s) rays
    int tileW, tileH;           // workgroup tile = one wave tile
    int ldsTf;                  // floats of the LDS TF table (max rows x the staged row stride)
    int usePriv;                // Gaussian TF gradient in per-thread LDS slices
    int tfCopyStride;           // floats between the kTfCopies TF-gradient tables (0: one table)
    int ldsAgg;                 // float offset of the volume-gradient chunk table in LDS (0: none)
    int aggLog2;                // log2 of the chunk table's slots
};

// One segment of a zero-fill list (gradients, counters) cleared by a kernel
// (graph-capturable; no memset nodes).
struct ZeroSeg {
    float* p;
    long long n;
};

// One parameter block of the multi-block Adam step: rows x cols floats with a
// row stride shared by param, grad and both moment buffers.
struct AdamSeg {
    float* p;
    const float* grad;
    float* m;
    float* v;
    long long rows, cols, stride;
    float lo, hi;               // clamp after the update when lo <= hi
};

// ---------------------------------------------------------------------------
// Forward-mode differentiation (renderer/renderer_kernels.cuh:135-308, :504-598)
// ---------------------------------------------------------------------------
constexpr int kFwdChunk = 8;      // derivative indices traced per launch slice
constexpr int kMaxFwdSums = 16384; // D bound of the LDS pixel contraction

struct FwdGradParams {
    int D;
    int dStep;                 // derivative index of the step size or -1
    int dStart[3], dDir[3];    // derivative indices of ray start / direction components or -1
    const int* tfEntry;        // [batch][D]: r*C + c of the TF entry traced by index d, or -1 (NULL = none)
    int tfEntrySb;             // D, or 0 when one table serves every batch
    float* grad;               // B x H x W x D x 4
    int gSb, gSy, gSx, gSd, gSc;
};

// forward variables -> adjoint variables.  The per-pixel outputs are
// accumulated (+=); an output that broadcasts over the batch or image
// (stride 0) takes atomic adds.
struct FwdToAdjParams {
    int B, H, W, D;
    const float* fv; int fSb, fSy, fSx, fSd, fSc;      // B x H x W x D x 4
    const float* ac; int aSb, aSy, aSx, aSc;           // B x H x W x 4
    int dStep, dStart[3], dDir[3];
    float* ast; int astSb, astSy, astSx, astAtomic;    // step-size adjoint or NULL
    float* ars; int arsSb, arsSy, arsSx, arsSc;        // ray start adjoint or NULL
    float* ard; int ardSb, ardSy, ardSx, ardSc;        // ray direction adjoint or NULL
    int camAtomic;
    float* sums;                                       // [B][D] contraction over pixels (zeroed here), or NULL
};

// compareToImage (renderer/renderer.cpp:562-658)
struct CompareParams {
    int B, H, W, D;
    const float* col; int cSb, cSy, cSx, cSc;
    const float* gin; int gSb, gSy, gSx, gSd, gSc;
    const float* ref; int rSb, rSy, rSx, rSc;
    float* dout; int oSb, oSy, oSx;        // B x H x W x 1 (reduce: 1 element)
    float* gout; int qSb, qSy, qSx, qSd;   // B x H x W x D (reduce: D elements, stride qSd)
    int reduce;
    float norm;                            // 1 / (B*H*W)
};

// ---------------------------------------------------------------------------
// Voxel-wise kernels over an X x Y x Z grid (pre-shading, best fit,
// importance maxima).  A flat index i runs over loop levels n[0] (innermost),
// n[1], n[2]; level k is the logical axis axis[k] (0 = x, 1 = y, 2 = z).  The
// host puts the axis of the main input's smallest stride innermost, so
// neighbouring lanes touch neighbouring elements.  Divisions by n[0] and n[1]
// use multiply-high magic numbers (Granlund and Montgomery, PLDI 1994):
// q = (umulhi(n, m) + n) >> l for n < 2^31.
// ---------------------------------------------------------------------------
constexpr int kMaxAuxTf = 16384;  // max R*C floats of a TF staged in LDS by the voxel-wise kernels

struct FastDiv {
    unsigned m;
    int l;
};

struct VoxelGrid {
    int X, Y, Z;
    int n[3];
    int axis[3];
    FastDiv div0, div1;
    int total;
};

// a TF table (one batch entry) in global memory: row stride sr, column stride sc
struct TfTable {
    const float* p;
    int R, C, sr, sc;
};

// per-tensor element strides for the logical axes x, y, z (+ channel)
struct Strides3 {
    int x, y, z, c;
};

// TFUtils.preshade_volume and its adjoint (renderer/tf_utils_impl.cu:17-190)
struct PreshadeParams {
    VoxelGrid G;
    TfTable tf;
    const float* dens; Strides3 ds;   // 1 x X x Y x Z
    float* col; Strides3 cs;          // 4 x X x Y x Z (forward output)
    const float* acol; Strides3 as;   // 4 x X x Y x Z (adjoint input)
    float* adens; Strides3 ads;       // 1 x X x Y x Z (adjoint output, overwritten)
};

// TFUtils.find_best_fit (renderer/tf_utils_impl.cu:225-520)
struct BestFitParams {
    VoxelGrid G;
    TfTable tf;
    const float* col; Strides3 cs;    // 4 x X x Y x Z target colours
    const float* prev; Strides3 ps;   // previous densities or NULL (no neighbour term)
    float* out; Strides3 os;          // 1 x X x Y x Z
    int numSamples;
    float opacityWeighting, neighborWeighting;
    unsigned long long seed;
};

// sample_importance (renderer/losses_impl.cu:254-421)
struct ImportanceParams {
    VoxelGrid G;                       // maxima pass
    TfTable tf;
    const float* dens; Strides3 ds;   // 1 x X x Y x Z
    const float* loc; int lsN, lsC;   // N x 3 sample locations in [0,1]^3
    unsigned char* out; int os;       // N booleans
    int N;
    float* maxima;                     // 3 floats: max |grad density|, max opacity, max |grad opacity|
    float wUniform, wDensityGradient, wOpacity;  // cumulative, normalised weights
    unsigned long long seed;
};

// elementwise losses (renderer/losses_impl.cu:35-240): contiguous arrays of n elements
struct ElementwiseParams {
    const void* a; const void* b; const void* g;
    void* o1; void* o2;
    long long n;
};

}  // namespace dvr
