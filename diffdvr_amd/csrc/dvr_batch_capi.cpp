// C-ABI of the batched persistent launch (plans), its autotuner and the
// multi-block Adam step.  A plan validates the batch once, lays out the tile
// queues and uploads the descriptor tables; a run is two kernel launches
// (clear + persistent render) with no allocation, copy or memset, so that
// optimisation steps can be captured into HIP graphs.
#include <hip/hip_runtime.h>

#include <sys/stat.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/dvr.h"
#include "dvr_host_util.hpp"
#include "dvr_launch.hpp"
#include "dvr_params.hpp"

using dvr_host::check_hip;
using dvr_host::DeviceGuard;
using dvr_host::fail;
using dvr_host::guarded;

namespace {

constexpr double kMaxDensityStore = 2.0 * 1024 * 1024 * 1024;  // bytes

template <class T> T* dev_alloc(size_t n)
{
    void* p = nullptr;
    if (n == 0) n = 1;
    check_hip(hipMalloc(&p, n * sizeof(T)), "hipMalloc");
    return static_cast<T*>(p);
}

void dev_free(void* p)
{
    if (p) (void)hipFree(p);
}

int wave_shift_of(int ww)
{
    switch (ww) {
    case 4: return 2;
    case 8: return 3;
    case 16: return 4;
    case 32: return 5;
    default: return -1;
    }
}

int tf_cols(int tfm)
{
    switch (tfm) {
    case DVR_TF_IDENTITY: return 2;
    case DVR_TF_TEXTURE: return 4;
    case DVR_TF_LINEAR: return 5;
    case DVR_TF_GAUSSIAN: return 6;
    default: return -1;
    }
}

}  // namespace

struct dvr_batch_plan {
    int device = 0;
    int filter = 1, tfm = 0, blend = 0, gm = 0, lossKind = 0;
    bool volGrad = false;  // some volume gradient is requested (the chunk table is needed)
    float earlyOut = 1.0f - 1e-5f;
    float* lossOut = nullptr;
    std::vector<dvr::BView> views;  // sorted by (tf, volume); tileBegin/tilesX set by layout()
    std::vector<dvr::BVol> vols;
    std::vector<dvr::BTf> tfs;
    int maxRC = 0;
    bool usePriv = false;
    double maxSamples = 0;  // upper bound of samples per ray over the views
    // layout
    int K = 16, waveWidth = 8, gridReq = 0, grid = 0, nTiles = 0, cap = 1;
    bool reinterpolate = false;  // no density store
    int cus = 0, resident = 0;  // compute units, resident workgroups of the current instance
    int qBegin[dvr::kNumXcd + 1] = {};
    size_t lds = 0;
    int aggOffset = 0;  // float offset of the volume-gradient chunk table in LDS (0: none)
    int aggLog2 = dvr::kAggLog2;  // log2 of the chunk table's slots
    // device
    dvr::BView* dViews = nullptr;
    dvr::BVol* dVols = nullptr;
    dvr::BTf* dTfs = nullptr;
    unsigned* dQueues = nullptr;
    dvr::ZeroSeg* dSegsAll = nullptr;   // [queues, bricked scratch..., gradients...]
    int nSegsAll = 0;
    long long maxSegAll = 0;
    int nSegsAlways = 1;                // cleared by every run: queues and bricked scratch
    long long maxSegAlways = dvr::kNumXcd;
    float* dBricks = nullptr;           // bricked volume-gradient scratch of every gradient volume
    size_t brickFloats = 0;
    dvr::DebrickSeg* dDebrick = nullptr;
    int nDebrick = 0;
    long long maxBricks = 0;
    float* dCk = nullptr;
    size_t ckFloats = 0;
    float* dDs = nullptr;   // density store of the forward march (adjoint plans), or NULL
    size_t dsFloats = 0;
    int dsCap = 0;
    std::string signature;  // autotune cache key
    bool frozen = false;    // captured graphs hold the scratch pointers: no re-layout

    ~dvr_batch_plan()
    {
        dev_free(dViews); dev_free(dVols); dev_free(dTfs); dev_free(dQueues); dev_free(dSegsAll); dev_free(dCk);
        dev_free(dDs); dev_free(dBricks); dev_free(dDebrick);
    }

    // floats between the TF-gradient table copies (odd: the copies start in different LDS banks);
    // 0 (one table) when the copies would take more than 16 KiB of LDS per workgroup, which keeps
    // eight one-wave workgroups resident per CU
    int copy_stride() const
    {
        const int st = maxRC | 1;
        return size_t(dvr::kTfCopies) * st * sizeof(float) <= 16384 ? st : 0;
    }

    // floats of the LDS TF table (the kernel pads linear-TF rows to 8 floats, dvr::tf_lds_stride)
    int tf_lds() const { return tfm == 2 ? (maxRC / 5) * 8 : maxRC; }

    // register segment length of checkpoint interval K (64 = two 32-sample segments)
    int segK() const { return K > 32 ? 32 : K; }

    dvr::BatchParams params() const
    {
        dvr::BatchParams P{};
        P.views = dViews; P.vols = dVols; P.tfs = dTfs; P.queues = dQueues; P.ck = dCk; P.lossOut = lossOut;
        P.ds = dsCap > 0 ? dDs : nullptr; P.dsCap = dsCap;
        P.nViews = (int)views.size(); P.nTiles = nTiles; P.ckCap = cap;
        P.ckEvery = K > 32 ? K / 32 : 1;
        for (int i = 0; i <= dvr::kNumXcd; ++i) P.qBegin[i] = qBegin[i];
        P.earlyOut = earlyOut; P.lossKind = lossKind;
        P.wantTf = (gm & 1) ? 1 : 0; P.wantVol = (gm & 2) ? 1 : 0; P.wantCam = (gm & 4) ? 1 : 0;
        P.waveShift = wave_shift_of(waveWidth);
        P.tileW = waveWidth; P.tileH = 64 / waveWidth;
        P.ldsTf = tf_lds(); P.usePriv = usePriv ? 1 : 0;
        P.tfCopyStride = (gm && !usePriv) ? copy_stride() : 0;
        P.ldsAgg = aggOffset;
        P.aggLog2 = aggLog2;
        return P;
    }

    // tiles, queues, grid and checkpoint scratch for the current options
    void layout()
    {
        const int tileW = waveWidth, tileH = 64 / waveWidth;  // one wave tile per workgroup
        long long t = 0;
        for (auto& v : views) {
            v.tilesX = (v.W + tileW - 1) / tileW;
            const int tilesY = (v.H + tileH - 1) / tileH;
            v.tileBegin = (int)t;
            t += (long long)v.tilesX * tilesY;
        }
        if (t >= (1LL << 31)) fail(DVR_ERR_INVALID_ARGUMENT, "batch has too many tiles");
        nTiles = (int)t;
        for (int q = 0; q <= dvr::kNumXcd; ++q) qBegin[q] = (int)((long long)nTiles * q / dvr::kNumXcd);
        const size_t ldsG =
            gm ? (usePriv ? size_t(maxRC) * dvr::kBatchBlock
                          : (copy_stride() > 0 ? size_t(dvr::kTfCopies) * copy_stride() : size_t(maxRC)))
               : 0;
        lds = sizeof(float) * (size_t(tf_lds()) + ldsG);
        // volume gradients: the per-wave table of 64-byte chunks follows the TF tables (not
        // reserved for camera-only plans, which run the same instance without it)
        aggOffset = 0;
        int occ = 0;
        if ((gm & 2) && volGrad) {
            // the largest table that keeps the instance at the waves per SIMD its registers
            // were fitted for (workgroups per CU from the runtime's occupancy calculator)
            aggOffset = dvr::agg_offset(tf_lds(), (int)ldsG);
            const int target = 4 * dvr::batched_vol_waves(gm);
            for (int lg = dvr::kAggLog2; lg >= dvr::kAggMinLog2; --lg) {
                aggLog2 = lg;
                lds = sizeof(float) * size_t(aggOffset) + dvr::agg_bytes(lg);
                check_hip(dvr::launch_batched(params(), filter, tfm, blend, segK(), gm, 0, lds, nullptr, &occ),
                          "occupancy query");
                if (occ >= target) break;
            }
        } else {
            check_hip(dvr::launch_batched(params(), filter, tfm, blend, segK(), gm, 0, lds, nullptr, &occ),
                      "occupancy query");
        }
        hipDeviceProp_t prop;
        check_hip(hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties");
        cus = prop.multiProcessorCount;
        resident = std::max(1, occ) * cus;
        // default: two waves per SIMD (eight one-wave workgroups per CU; measured
        // best for gather-bound tiles on MI355X: fewer concurrent rays per CU keep
        // their cells in L1), within residency; volume-gradient instances are bound
        // by the latency of their walk and fill the CU (four waves per SIMD)
        int g = gridReq > 0 ? gridReq : (((gm & 2) && volGrad) ? resident : std::min(resident, 8 * cus));
        g = std::max(1, std::min(g, std::max(nTiles, 1)));
        grid = g;
        cap = std::max(1, std::min((int)std::ceil(maxSamples / K), 1 << 14));
        const size_t need = gm ? size_t(grid) * cap * dvr::kBatchBlock : 0;
        if (need > ckFloats) {
            dev_free(dCk);
            dCk = nullptr;
            ckFloats = 0;
            dCk = dev_alloc<float>(need);
            ckFloats = need;
        }
        // Density store: the forward march keeps every sample's density, so the reverse march
        // reads each segment back (one 4-byte load per sample, all in flight) instead of
        // re-interpolating it (eight gathers per new cell).  Sized for the longest possible ray;
        // plans whose store would exceed kMaxDensityStore bytes re-interpolate.
        const double dsWant = std::ceil(maxSamples) + 2.0;
        const double dsBytes = double(grid) * dsWant * dvr::kBatchBlock * sizeof(float);
        dsCap = (gm && !reinterpolate && dsBytes <= double(kMaxDensityStore)) ? (int)dsWant : 0;
        const size_t dsNeed = size_t(grid) * size_t(dsCap) * dvr::kBatchBlock;
        if (dsNeed > dsFloats) {
            dev_free(dDs);
            dDs = nullptr;
            dsFloats = 0;
            dDs = dev_alloc<float>(dsNeed);
            dsFloats = dsNeed;
        }
        check_hip(hipMemcpy(dViews, views.data(), views.size() * sizeof(dvr::BView), hipMemcpyHostToDevice),
                  "hipMemcpy views");
    }

    void zero(bool gradients, hipStream_t s) const
    {
        const int n = gradients ? nSegsAll : nSegsAlways;
        const long long mx = gradients ? maxSegAll : maxSegAlways;
        check_hip(dvr::launch_zero_segments(dSegsAll, n, mx, s), "zero launch");
    }

    // evBegin / evEnd (optional) are recorded around the persistent kernel alone
    void run(int flags, hipStream_t s, hipEvent_t evBegin = nullptr, hipEvent_t evEnd = nullptr) const
    {
        zero((flags & DVR_BATCH_ZERO_GRADIENTS) != 0, s);
        if (evBegin) check_hip(hipEventRecord(evBegin, s), "hipEventRecord");
        if (nTiles > 0)
            check_hip(dvr::launch_batched(params(), filter, tfm, blend, segK(), gm, grid, lds, s, nullptr),
                      "batched launch");
        if (evEnd) check_hip(hipEventRecord(evEnd, s), "hipEventRecord");
        if (nDebrick > 0) check_hip(dvr::launch_debrick(dDebrick, nDebrick, maxBricks, s), "debrick launch");
    }
};

namespace {

std::mutex g_tuneMu;
struct Tuned {
    int K, ww, grid;
};
std::map<std::string, Tuned> g_tuneCache;  // signature -> tuned options

// Autotune results persist across processes in a text file, one line per batch signature:
// "<64-bit FNV-1a of the signature, hex> <K> <wave width> <grid>".  The file is
// $DVR_AUTOTUNE_CACHE, else $XDG_CACHE_HOME/diffdvr_amd/autotune.txt, else
// $HOME/.cache/diffdvr_amd/autotune.txt; DVR_AUTOTUNE_CACHE set to an empty string turns it off.
// Lines are appended (a later line for the same signature wins); read and write failures only
// disable the persistence.
std::string tune_file()
{
    if (const char* e = std::getenv("DVR_AUTOTUNE_CACHE")) return e;
    std::string dir;
    if (const char* x = std::getenv("XDG_CACHE_HOME"); x && *x) dir = x;
    else if (const char* h = std::getenv("HOME"); h && *h) dir = std::string(h) + "/.cache";
    else return "";
    return dir + "/diffdvr_amd/autotune.txt";
}

uint64_t fnv1a(const std::string& s)
{
    uint64_t h = 1469598103934665603ull;
    for (unsigned char c : s) {
        h ^= c;
        h *= 1099511628211ull;
    }
    return h;
}

bool load_tuned(const std::string& key, Tuned& out)
{
    const std::string f = tune_file();
    if (f.empty()) return false;
    FILE* fp = std::fopen(f.c_str(), "r");
    if (!fp) return false;
    const unsigned long long want = fnv1a(key);
    unsigned long long h;
    int K, ww, grid;
    bool found = false;
    while (std::fscanf(fp, "%llx %d %d %d", &h, &K, &ww, &grid) == 4)
        if (h == want && (K == 4 || K == 8 || K == 16 || K == 32 || K == 64) && wave_shift_of(ww) >= 0 && grid > 0) {
            out = Tuned{K, ww, grid};
            found = true;
        }
    std::fclose(fp);
    return found;
}

void store_tuned(const std::string& key, const Tuned& t)
{
    const std::string f = tune_file();
    if (f.empty()) return;
    for (size_t i = 1; i < f.size(); ++i)  // create the directory chain (failures show at fopen)
        if (f[i] == '/') (void)mkdir(f.substr(0, i).c_str(), 0755);
    FILE* fp = std::fopen(f.c_str(), "a");
    if (!fp) return;
    std::fprintf(fp, "%016llx %d %d %d\n", (unsigned long long)fnv1a(key), t.K, t.ww, t.grid);
    std::fclose(fp);
}

void check_options(const dvr_batch_options* o, int& K, int& ww)
{
    K = o && o->checkpoint_interval ? o->checkpoint_interval : 16;
    ww = o && o->wave_width ? o->wave_width : 8;
    if (K != 4 && K != 8 && K != 16 && K != 32 && K != 64)
        fail(DVR_ERR_INVALID_ARGUMENT, "checkpoint_interval must be 4, 8, 16, 32 or 64");
    if (wave_shift_of(ww) < 0) fail(DVR_ERR_INVALID_ARGUMENT, "wave_width must be 4, 8, 16 or 32");
    if (o && o->grid < 0) fail(DVR_ERR_INVALID_ARGUMENT, "grid must be >= 0");
}

void set_options(dvr_batch_plan* p, const dvr_batch_options* o)
{
    int K, ww;
    check_options(o, K, ww);
    if (p->frozen) {
        const bool same = K == p->K && ww == p->waveWidth && (o ? o->grid : 0) == p->gridReq &&
                          (o && o->reinterpolate != 0) == p->reinterpolate;
        if (!same)
            fail(DVR_ERR_INVALID_ARGUMENT,
                 "the plan is frozen (a captured graph holds its scratch buffers and kernel arguments): "
                 "its options cannot change; create a new plan");
        return;
    }
    p->K = K;
    p->waveWidth = ww;
    p->gridReq = o ? o->grid : 0;
    p->reinterpolate = o && o->reinterpolate != 0;
    p->layout();
}

dvr_batch_plan* create(const dvr_batch* b, const dvr_batch_options* o)
{
    if (!b) fail(DVR_ERR_INVALID_ARGUMENT, "batch is NULL");
    if (b->num_views < 0 || b->num_volumes < 0 || b->num_tfs < 0) fail(DVR_ERR_INVALID_ARGUMENT, "negative count");
    if (b->num_views > 0 && (!b->views || !b->volumes || !b->tfs))
        fail(DVR_ERR_INVALID_ARGUMENT, "views, volumes and tfs must be given");
    if (b->device < 0) fail(DVR_ERR_UNSUPPORTED, "batch device is the host: this library renders on the GPU only");
    if (b->volume_filter_mode != 0 && b->volume_filter_mode != 1)
        fail(DVR_ERR_UNSUPPORTED, "volume_filter_mode %d is not supported (Nearest=0, Trilinear=1)", b->volume_filter_mode);
    const int C = tf_cols(b->tf_mode);
    if (C < 0) fail(DVR_ERR_UNSUPPORTED, "tf mode %d is not supported (Identity=0, Texture=1, Linear=2, Gaussian=3)", b->tf_mode);
    if (b->blend_mode != 0 && b->blend_mode != 1) fail(DVR_ERR_INVALID_ARGUMENT, "unknown blend enum value");
    if (b->loss != DVR_LOSS_NONE && b->loss != DVR_LOSS_L1 && b->loss != DVR_LOSS_L2)
        fail(DVR_ERR_INVALID_ARGUMENT, "loss must be DVR_LOSS_NONE, DVR_LOSS_L1 or DVR_LOSS_L2");
    {
        int K, ww;
        check_options(o, K, ww);
    }

    std::unique_ptr<dvr_batch_plan> p(new dvr_batch_plan());
    p->device = b->device;
    p->filter = b->volume_filter_mode;
    p->tfm = b->tf_mode;
    p->blend = b->blend_mode;
    p->earlyOut = b->blending_early_out > 0.0f ? b->blending_early_out : 1.0f - 1e-5f;
    p->lossKind = b->loss;
    p->lossOut = b->forward_only ? nullptr : b->loss_out;

    bool anyTfGrad = false, anyVolGrad = false, anyCamGrad = false;
    for (int i = 0; i < b->num_volumes; ++i) {
        const dvr_batch_volume& v = b->volumes[i];
        if (!v.data) fail(DVR_ERR_INVALID_ARGUMENT, "volume %d has no data", i);
        if (v.size_x < 2 || v.size_y < 2 || v.size_z < 2)
            fail(DVR_ERR_INVALID_ARGUMENT, "volume %d must have at least 2 voxels per dimension, got %dx%dx%d", i,
                 v.size_x, v.size_y, v.size_z);
        if ((long long)v.size_x * v.size_y * v.size_z > (1LL << 30))
            fail(DVR_ERR_UNSUPPORTED, "volume %d: one volume may hold at most 2^30 voxels", i);
        for (int k = 0; k < 3; ++k)
            if (!(v.box_size[k] > 0.0f) || !std::isfinite(v.box_size[k]) || !std::isfinite(v.box_min[k]))
                fail(DVR_ERR_INVALID_ARGUMENT, "volume %d: box size must be positive and finite", i);
        dvr::BVol d{};
        d.p = v.data; d.adj = b->forward_only ? nullptr : v.adj_data;
        d.X = v.size_x; d.Y = v.size_y; d.Z = v.size_z;
        for (int k = 0; k < 3; ++k) { d.bmin[k] = v.box_min[k]; d.bsize[k] = v.box_size[k]; }
        anyVolGrad |= d.adj != nullptr;
        p->vols.push_back(d);
    }
    for (int i = 0; i < b->num_tfs; ++i) {
        const dvr_batch_tf& t = b->tfs[i];
        if (!t.data) fail(DVR_ERR_INVALID_ARGUMENT, "tf %d has no data", i);
        if (t.cols != C) fail(DVR_ERR_INVALID_ARGUMENT, "tf %d must have %d entries at dimension 2, but has %d", i, C, t.cols);
        if (t.rows < 1 || (b->tf_mode == DVR_TF_IDENTITY && t.rows != 1))
            fail(DVR_ERR_INVALID_ARGUMENT, "tf %d has an invalid number of rows (%d)", i, t.rows);
        if (b->tf_mode == DVR_TF_LINEAR && t.rows < 2)
            fail(DVR_ERR_INVALID_ARGUMENT, "tensor 'tf' must have at least two control points in TFLinear-mode");
        if ((long long)t.rows * t.cols > dvr::kMaxTfLds)
            fail(DVR_ERR_UNSUPPORTED, "batched rendering needs R*C <= %d transfer-function entries", dvr::kMaxTfLds);
        dvr::BTf d{};
        d.p = t.data; d.adj = b->forward_only ? nullptr : t.adj_data; d.R = t.rows; d.C = t.cols;
        anyTfGrad |= d.adj != nullptr;
        p->maxRC = std::max(p->maxRC, t.rows * t.cols);
        p->tfs.push_back(d);
    }
    long long totalPix = 0;
    std::vector<dvr::BView> vs;
    for (int i = 0; i < b->num_views; ++i) {
        const dvr_batch_view& v = b->views[i];
        if (v.volume < 0 || v.volume >= b->num_volumes) fail(DVR_ERR_INVALID_ARGUMENT, "view %d: volume index out of range", i);
        if (v.tf < 0 || v.tf >= b->num_tfs) fail(DVR_ERR_INVALID_ARGUMENT, "view %d: tf index out of range", i);
        if (v.width <= 0 || v.height <= 0) fail(DVR_ERR_INVALID_ARGUMENT, "view %d: image size must be positive", i);
        if ((long long)v.width * v.height * 4 > (1LL << 30)) fail(DVR_ERR_INVALID_ARGUMENT, "view %d: image too large", i);
        if (!v.camera || !v.color || !v.termination_index)
            fail(DVR_ERR_INVALID_ARGUMENT, "view %d: camera, color and termination_index must be given", i);
        if (v.camera_mode != DVR_CAMERA_REFERENCE_FRAME && v.camera_mode != DVR_CAMERA_INVERSE_VIEW_MATRIX)
            fail(DVR_ERR_UNSUPPORTED, "view %d: batched views take ReferenceFrame or InverseViewMatrix cameras", i);
        if (!(v.step_size > 0.0f) || !std::isfinite(v.step_size))
            fail(DVR_ERR_INVALID_ARGUMENT, "view %d: step size must be positive and finite", i);
        if (!b->forward_only) {
            if (b->loss == DVR_LOSS_NONE && !v.adj_color) fail(DVR_ERR_INVALID_ARGUMENT, "view %d: adj_color is NULL", i);
            if (b->loss != DVR_LOSS_NONE && !v.target) fail(DVR_ERR_INVALID_ARGUMENT, "view %d: target is NULL", i);
            if (v.adj_camera && v.camera_mode != DVR_CAMERA_REFERENCE_FRAME)
                fail(DVR_ERR_UNSUPPORTED, "view %d: camera gradients need a ReferenceFrame camera", i);
        }
        dvr::BView d{};
        d.vol = v.volume; d.tf = v.tf; d.W = v.width; d.H = v.height;
        d.camMode = v.camera_mode;
        const double fovX = double(v.fov_y_radians) * double(v.width) / double(v.height);
        d.tanX = std::tan(fovX / 2.0);
        d.tanY = std::tan(double(v.fov_y_radians) / 2.0);
        d.cam = v.camera;
        d.step = v.step_size;
        d.lossScale = v.loss_scale;
        d.color = v.color; d.term = v.termination_index;
        d.target = v.target; d.acol = v.adj_color;
        d.acam = b->forward_only ? nullptr : v.adj_camera;
        anyCamGrad |= d.acam != nullptr;
        totalPix += (long long)v.width * v.height;
        const dvr::BVol& vol = p->vols[v.volume];
        const double diag = std::sqrt(double(vol.bsize[0]) * vol.bsize[0] + double(vol.bsize[1]) * vol.bsize[1] +
                                      double(vol.bsize[2]) * vol.bsize[2]);
        p->maxSamples = std::max(p->maxSamples, diag / double(v.step_size) + 2.0);
        vs.push_back(d);
    }
    if (!b->forward_only && b->loss != DVR_LOSS_NONE) {
        for (auto& v : vs)
            if (v.lossScale == 0.0f) v.lossScale = float(1.0 / (4.0 * double(std::max(totalPix, 1LL))));
    }
    // tiles of views sharing a TF (then a volume) are contiguous: a workgroup
    // restages its LDS TF table rarely and XCDs share volumes
    std::stable_sort(vs.begin(), vs.end(), [](const dvr::BView& a, const dvr::BView& c) {
        return a.tf != c.tf ? a.tf < c.tf : a.vol < c.vol;
    });
    p->views = vs;
    if (b->forward_only) p->gm = 0;
    else {
        if (!anyTfGrad && !anyVolGrad && !anyCamGrad && !p->lossOut)
            fail(DVR_ERR_NO_GRADIENT, "At least one adjoint variable must be activated!");
        p->gm = anyCamGrad ? (anyVolGrad ? 7 : 5) : (anyVolGrad ? 3 : 1);
        p->volGrad = anyVolGrad;
    }
    p->usePriv = p->gm && b->tf_mode == DVR_TF_GAUSSIAN && p->maxRC <= 48;

    // validation done: device set-up from here on
    DeviceGuard guard(b->device);
    // clear list: queue counters first, then every gradient buffer and the loss
    std::vector<dvr::ZeroSeg> segs;
    p->dQueues = dev_alloc<unsigned>(dvr::kNumXcd);
    segs.push_back({reinterpret_cast<float*>(p->dQueues), dvr::kNumXcd});
    // bricked volume-gradient scratch (2 x 2 x 4 voxels per 64-byte chunk), one range per gradient
    // volume, each 256-byte aligned; cleared by every run, added into the gradient after the kernel
    if (p->gm & 2) {
        std::vector<size_t> off;
        size_t total = 0;
        for (auto& v : p->vols) {
            off.push_back(total);
            if (!v.adj) continue;
            const size_t n = size_t((v.X + 1) / 2) * size_t((v.Y + 1) / 2) * size_t((v.Z + 3) / 4) * 16;
            total += (n + 63) / 64 * 64;
        }
        if (total > 0) {
            p->dBricks = dev_alloc<float>(total);
            p->brickFloats = total;
            std::vector<dvr::DebrickSeg> db;
            for (size_t k = 0; k < p->vols.size(); ++k) {
                dvr::BVol& v = p->vols[k];
                if (!v.adj) continue;
                v.brick = p->dBricks + off[k];
                const long long nb = (long long)((v.X + 1) / 2) * ((v.Y + 1) / 2) * ((v.Z + 3) / 4);
                segs.push_back({v.brick, nb * 16});
                p->maxSegAlways = std::max(p->maxSegAlways, nb * 16);
                db.push_back({v.brick, v.adj, v.X, v.Y, v.Z, 0});
                p->maxBricks = std::max(p->maxBricks, nb);
            }
            p->nDebrick = (int)db.size();
            p->dDebrick = dev_alloc<dvr::DebrickSeg>(db.size());
            check_hip(hipMemcpy(p->dDebrick, db.data(), db.size() * sizeof(dvr::DebrickSeg), hipMemcpyHostToDevice),
                      "hipMemcpy");
        }
    }
    p->nSegsAlways = (int)segs.size();
    std::vector<dvr::ZeroSeg> g;
    for (auto& v : p->vols) if (v.adj) g.push_back({v.adj, (long long)v.X * v.Y * v.Z});
    for (auto& t : p->tfs) if (t.adj) g.push_back({t.adj, (long long)t.R * t.C});
    for (auto& v : p->views) if (v.acam) g.push_back({v.acam, 9});
    if (p->lossOut) g.push_back({p->lossOut, 1});
    std::sort(g.begin(), g.end(), [](const dvr::ZeroSeg& a, const dvr::ZeroSeg& c) { return a.p < c.p; });
    for (auto& s : g) {  // merge duplicates and adjacent ranges (not into the always-cleared ranges)
        if ((int)segs.size() > p->nSegsAlways) {
            dvr::ZeroSeg& l = segs.back();
            if (s.p <= l.p + l.n) {
                const long long end = std::max<long long>(l.n, (long long)(s.p - l.p) + s.n);
                l.n = end;
                continue;
            }
        }
        segs.push_back(s);
    }
    if (segs.size() > 65535) fail(DVR_ERR_UNSUPPORTED, "too many separate gradient buffers (%zu)", segs.size());
    p->nSegsAll = (int)segs.size();
    p->maxSegAll = 0;
    for (auto& s : segs) p->maxSegAll = std::max(p->maxSegAll, s.n);
    p->dSegsAll = dev_alloc<dvr::ZeroSeg>(segs.size());
    check_hip(hipMemcpy(p->dSegsAll, segs.data(), segs.size() * sizeof(dvr::ZeroSeg), hipMemcpyHostToDevice), "hipMemcpy");

    p->dViews = dev_alloc<dvr::BView>(p->views.size());
    p->dVols = dev_alloc<dvr::BVol>(p->vols.size());
    p->dTfs = dev_alloc<dvr::BTf>(p->tfs.size());
    check_hip(hipMemcpy(p->dVols, p->vols.data(), p->vols.size() * sizeof(dvr::BVol), hipMemcpyHostToDevice), "hipMemcpy");
    check_hip(hipMemcpy(p->dTfs, p->tfs.data(), p->tfs.size() * sizeof(dvr::BTf), hipMemcpyHostToDevice), "hipMemcpy");

    // signature for the autotune cache: device, modes, gradients and the batch shape
    {
        hipDeviceProp_t prop;
        check_hip(hipGetDeviceProperties(&prop, p->device), "hipGetDeviceProperties");
        std::string s = std::string(prop.gcnArchName) + "|" + std::to_string(prop.multiProcessorCount);
        s += "|" + std::to_string(p->filter) + std::to_string(p->tfm) + std::to_string(p->blend) + "|" +
             std::to_string(p->gm) + "|" + std::to_string(p->lossKind);
        for (auto& v : p->vols) s += "|v" + std::to_string(v.X) + "x" + std::to_string(v.Y) + "x" + std::to_string(v.Z);
        for (auto& t : p->tfs) s += "|t" + std::to_string(t.R);
        for (auto& v : p->views) {
            char buf[96];
            snprintf(buf, sizeof(buf), "|%d,%d,%dx%d,%.6g", v.vol, v.tf, v.W, v.H, double(v.step));
            s += buf;
        }
        p->signature = s;
    }
    set_options(p.get(), o);
    return p.release();
}

}  // namespace

extern "C" {

dvr_status dvr_batch_plan_create(const dvr_batch* batch, const dvr_batch_options* options, dvr_batch_plan** plan)
{
    return guarded([&] {
        if (!plan) fail(DVR_ERR_INVALID_ARGUMENT, "plan is NULL");
        *plan = nullptr;
        *plan = create(batch, options);
    });
}

dvr_status dvr_batch_plan_destroy(dvr_batch_plan* plan)
{
    return guarded([&] {
        if (!plan) return;
        DeviceGuard g(plan->device);
        delete plan;
    });
}

dvr_status dvr_batch_plan_run(dvr_batch_plan* plan, int32_t flags, void* stream)
{
    return guarded([&] {
        if (!plan) fail(DVR_ERR_INVALID_ARGUMENT, "plan is NULL");
        DeviceGuard g(plan->device);
        plan->run(flags, (hipStream_t)stream);
    });
}

dvr_status dvr_batch_plan_time_kernel(dvr_batch_plan* plan, int32_t flags, void* stream, float* kernel_ms)
{
    return guarded([&] {
        if (!plan || !kernel_ms) fail(DVR_ERR_INVALID_ARGUMENT, "NULL argument");
        DeviceGuard g(plan->device);
        hipStream_t s = (hipStream_t)stream;
        hipEvent_t e0, e1;
        check_hip(hipEventCreate(&e0), "hipEventCreate");
        if (hipEventCreate(&e1) != hipSuccess) {
            (void)hipEventDestroy(e0);
            fail(DVR_ERR_HIP, "hipEventCreate failed");
        }
        try {
            plan->run(flags, s, e0, e1);
            check_hip(hipEventSynchronize(e1), "hipEventSynchronize");
            check_hip(hipEventElapsedTime(kernel_ms, e0, e1), "hipEventElapsedTime");
        } catch (...) {
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            throw;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    });
}

dvr_status dvr_batch_plan_freeze(dvr_batch_plan* plan)
{
    return guarded([&] {
        if (!plan) fail(DVR_ERR_INVALID_ARGUMENT, "plan is NULL");
        plan->frozen = true;
    });
}

dvr_status dvr_batch_plan_zero_gradients(dvr_batch_plan* plan, void* stream)
{
    return guarded([&] {
        if (!plan) fail(DVR_ERR_INVALID_ARGUMENT, "plan is NULL");
        DeviceGuard g(plan->device);
        plan->zero(true, (hipStream_t)stream);
    });
}

dvr_status dvr_batch_plan_get_options(const dvr_batch_plan* plan, dvr_batch_options* options)
{
    return guarded([&] {
        if (!plan || !options) fail(DVR_ERR_INVALID_ARGUMENT, "NULL argument");
        options->checkpoint_interval = plan->K;
        options->wave_width = plan->waveWidth;
        options->grid = plan->grid;
        options->reinterpolate = plan->reinterpolate ? 1 : 0;
    });
}

dvr_status dvr_batch_plan_set_options(dvr_batch_plan* plan, const dvr_batch_options* options)
{
    return guarded([&] {
        if (!plan) fail(DVR_ERR_INVALID_ARGUMENT, "plan is NULL");
        DeviceGuard g(plan->device);
        check_hip(hipDeviceSynchronize(), "hipDeviceSynchronize");
        set_options(plan, options);
    });
}

dvr_status dvr_batch_plan_autotune(dvr_batch_plan* plan, int32_t iterations, dvr_batch_options* chosen, void* stream)
{
    return guarded([&] {
        if (!plan) fail(DVR_ERR_INVALID_ARGUMENT, "plan is NULL");
        if (plan->frozen)
            fail(DVR_ERR_INVALID_ARGUMENT,
                 "the plan is frozen (a captured graph holds its scratch buffers and kernel arguments): "
                 "autotune before capturing");
        DeviceGuard g(plan->device);
        hipStream_t s = (hipStream_t)stream;
        const int gridReq = plan->gridReq;
        // the leading tag names the kernels' and search's revision: results of older ones are not reused
        const std::string tuneKey = "r2|" + plan->signature + (plan->reinterpolate ? "|reinterpolate" : "");
        Tuned best{0, 0, 0};
        {
            std::lock_guard<std::mutex> lk(g_tuneMu);
            auto it = g_tuneCache.find(tuneKey);
            if (it != g_tuneCache.end()) best = it->second;
            else if (load_tuned(tuneKey, best)) g_tuneCache[tuneKey] = best;
        }
        if (best.K == 0) {
            const int its = std::max(1, (int)iterations);
            hipEvent_t e0, e1;
            check_hip(hipEventCreate(&e0), "hipEventCreate");
            check_hip(hipEventCreate(&e1), "hipEventCreate");
            const int reint = plan->reinterpolate ? 1 : 0;
            auto time_it = [&](int K, int ww, int g) {
                check_hip(hipStreamSynchronize(s), "hipStreamSynchronize");
                dvr_batch_options o{K, ww, g, reint};
                set_options(plan, &o);
                plan->run(0, s);  // warm-up
                check_hip(hipEventRecord(e0, s), "hipEventRecord");
                for (int i = 0; i < its; ++i) plan->run(0, s);
                check_hip(hipEventRecord(e1, s), "hipEventRecord");
                check_hip(hipEventSynchronize(e1), "hipEventSynchronize");
                float ms = 0.0f;
                check_hip(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
                return ms;
            };
            try {
                // joint search over checkpoint interval and workgroups per CU (each interval is
                // its own instance, with its own registers and residency), then the wave tile shape
                Tuned cur{16, 8, 0};
                float bestMs = 1e30f;
                std::vector<int> Ks;
                if (plan->gm) Ks = {4, 8, 16, 32, 64};
                else Ks = {16};
                for (int K : Ks) {
                    {
                        dvr_batch_options o{K, 8, 0, reint};
                        set_options(plan, &o);
                    }
                    const int cus = plan->cus, resident = plan->resident;
                    std::vector<int> grids;
                    if (gridReq > 0) grids.push_back(gridReq);
                    else {
                        for (int k = 1; k <= 4; ++k)  // 1..4 waves per SIMD
                            if (4 * k * cus <= resident) grids.push_back(4 * k * cus);
                        if (grids.empty() || grids.back() < resident) grids.push_back(resident);
                    }
                    for (int g : grids) {
                        const float ms = time_it(K, 8, g);
                        if (ms < bestMs) { bestMs = ms; cur = Tuned{K, 8, g}; }
                    }
                }
                for (int ww : {16, 4, 32}) {
                    const float ms = time_it(cur.K, ww, cur.grid);
                    if (ms < bestMs) { bestMs = ms; cur.ww = ww; }
                }
                best = cur;
            } catch (...) {
                (void)hipEventDestroy(e0);
                (void)hipEventDestroy(e1);
                throw;
            }
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            std::lock_guard<std::mutex> lk(g_tuneMu);
            g_tuneCache[tuneKey] = best;
            store_tuned(tuneKey, best);
        }
        check_hip(hipStreamSynchronize(s), "hipStreamSynchronize");
        dvr_batch_options o{best.K, best.ww, best.grid, plan->reinterpolate ? 1 : 0};
        set_options(plan, &o);
        plan->zero(true, s);
        if (chosen) {
            chosen->checkpoint_interval = plan->K;
            chosen->wave_width = plan->waveWidth;
            chosen->grid = plan->grid;
            chosen->reinterpolate = plan->reinterpolate ? 1 : 0;
        }
    });
}

int32_t dvr_batch_plan_chunk_slots(const dvr_batch_plan* plan)
{
    if (!plan) return -1;
    return plan->aggOffset > 0 ? (int32_t(1) << plan->aggLog2) : 0;
}

int64_t dvr_batch_plan_workspace_bytes(const dvr_batch_plan* plan)
{
    if (!plan) return -1;
    return (int64_t)((plan->ckFloats + plan->dsFloats + plan->brickFloats) * sizeof(float) +
                     plan->nDebrick * sizeof(dvr::DebrickSeg) + plan->views.size() * sizeof(dvr::BView) +
                     plan->vols.size() * sizeof(dvr::BVol) + plan->tfs.size() * sizeof(dvr::BTf) +
                     plan->nSegsAll * sizeof(dvr::ZeroSeg) + dvr::kNumXcd * sizeof(unsigned));
}

// ---------------------------------------------------------------------------
// Adam
// ---------------------------------------------------------------------------
struct dvr_adam_group {
    int device = 0;
    dvr::AdamSeg* dSegs = nullptr;
    long long* dStep = nullptr;  // step counter, advanced by every step launch
    int n = 0;
    long long maxN = 0;
    ~dvr_adam_group()
    {
        dev_free(dSegs);
        dev_free(dStep);
    }
};

dvr_status dvr_adam_group_create(const dvr_adam_block* blocks, int32_t num_blocks, int32_t device,
                                 dvr_adam_group** group)
{
    return guarded([&] {
        if (!group) fail(DVR_ERR_INVALID_ARGUMENT, "group is NULL");
        *group = nullptr;
        if (num_blocks < 0 || (num_blocks > 0 && !blocks)) fail(DVR_ERR_INVALID_ARGUMENT, "invalid blocks");
        if (num_blocks > 65535) fail(DVR_ERR_UNSUPPORTED, "too many parameter blocks");
        if (device < 0) fail(DVR_ERR_UNSUPPORTED, "Adam runs on the GPU only");
        std::vector<dvr::AdamSeg> segs;
        long long mx = 0;
        for (int i = 0; i < num_blocks; ++i) {
            const dvr_adam_block& b = blocks[i];
            if (!b.param || !b.grad || !b.exp_avg || !b.exp_avg_sq) fail(DVR_ERR_INVALID_ARGUMENT, "block %d: NULL buffer", i);
            if (b.rows < 0 || b.cols < 0 || b.stride < b.cols) fail(DVR_ERR_INVALID_ARGUMENT, "block %d: invalid shape", i);
            segs.push_back({b.param, b.grad, b.exp_avg, b.exp_avg_sq, b.rows, b.cols, b.stride, b.clamp_min, b.clamp_max});
            mx = std::max(mx, (long long)(b.rows * b.cols));
        }
        DeviceGuard g(device);
        std::unique_ptr<dvr_adam_group> a(new dvr_adam_group());
        a->device = device;
        a->n = (int)segs.size();
        a->maxN = mx;
        a->dSegs = dev_alloc<dvr::AdamSeg>(segs.size());
        a->dStep = dev_alloc<long long>(1);
        const long long zero = 0;
        check_hip(hipMemcpy(a->dStep, &zero, sizeof(zero), hipMemcpyHostToDevice), "hipMemcpy");
        if (!segs.empty())
            check_hip(hipMemcpy(a->dSegs, segs.data(), segs.size() * sizeof(dvr::AdamSeg), hipMemcpyHostToDevice),
                      "hipMemcpy");
        *group = a.release();
    });
}

dvr_status dvr_adam_group_destroy(dvr_adam_group* group)
{
    return guarded([&] {
        if (!group) return;
        DeviceGuard g(group->device);
        delete group;
    });
}

dvr_status dvr_adam_group_step(dvr_adam_group* group, double lr, double beta1, double beta2, double eps,
                               void* stream)
{
    return guarded([&] {
        if (!group) fail(DVR_ERR_INVALID_ARGUMENT, "group is NULL");
        DeviceGuard g(group->device);
        check_hip(dvr::launch_adam(group->dSegs, group->n, group->maxN, group->dStep, lr, beta1, beta2, eps,
                                   (hipStream_t)stream),
                  "adam launch");
    });
}

dvr_status dvr_adam_group_set_step(dvr_adam_group* group, int64_t step)
{
    return guarded([&] {
        if (!group) fail(DVR_ERR_INVALID_ARGUMENT, "group is NULL");
        if (step < 0) fail(DVR_ERR_INVALID_ARGUMENT, "step must be >= 0");
        DeviceGuard g(group->device);
        const long long v = step;
        check_hip(hipMemcpy(group->dStep, &v, sizeof(v), hipMemcpyHostToDevice), "hipMemcpy");
    });
}

dvr_status dvr_adam_group_get_step(dvr_adam_group* group, int64_t* step)
{
    return guarded([&] {
        if (!group || !step) fail(DVR_ERR_INVALID_ARGUMENT, "NULL argument");
        DeviceGuard g(group->device);
        long long v = 0;
        check_hip(hipMemcpy(&v, group->dStep, sizeof(v), hipMemcpyDeviceToHost), "hipMemcpy");
        *step = v;
    });
}

}  // extern "C"
