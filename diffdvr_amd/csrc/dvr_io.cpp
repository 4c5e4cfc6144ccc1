// Host-side data formats around the renderer: the .cvol volume file, the
// implicit (synthetic) datasets, the density histogram and the transfer-function
// builders of the TF editor.
//
//   .cvol            renderer/volume.cpp:59-285 (64-byte header, uchar / ushort /
//                    float voxels, x fastest)
//   implicit sets    renderer/volume_implicit.cpp:13-236, renderer/volume.cpp:369-387
//   histogram        renderer/volume.cpp:287-356 (512 bins over the non-zero voxels)
//   TF builders      renderer/tf_utils.cpp:12-172, colour space renderer/tf_utils.h:20-90
//
// Voxel buffers crossing this interface are dense float32 host arrays with x
// fastest (index x + X*(y + Y*z)), the layout of the reference's Volume blob.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <limits>
#include <string>
#include <vector>

#include "../../include/dvr.h"
#include "dvr_host_util.hpp"

using dvr_host::fail;
using dvr_host::guarded;

namespace {

constexpr char kMagic[4] = {'c', 'v', 'o', 'l'};
constexpr int kBytesPerType[3] = {1, 2, 4};

void read_header(std::ifstream& f, const char* path, dvr_cvol_header* h)
{
    char magic[4];
    f.read(magic, 4);
    if (!f || std::memcmp(magic, kMagic, 4) != 0) fail(DVR_ERR_INVALID_ARGUMENT, "%s: Illegal magic number", path);
    uint64_t sz[3];
    double vs[3];
    int32_t type = 0;
    char compressed = 0;
    f.read(reinterpret_cast<char*>(sz), 24);
    f.read(reinterpret_cast<char*>(vs), 24);
    f.read(reinterpret_cast<char*>(&type), 4);
    f.read(&compressed, 1);
    f.ignore(7);
    if (!f) fail(DVR_ERR_INVALID_ARGUMENT, "%s: truncated header", path);
    if (type < DVR_CVOL_UCHAR || type > DVR_CVOL_FLOAT) fail(DVR_ERR_INVALID_ARGUMENT, "%s: unknown voxel type %d", path, type);
    for (int d = 0; d < 3; ++d)
        if (sz[d] == 0) fail(DVR_ERR_INVALID_ARGUMENT, "%s: empty volume", path);
    if (double(sz[0]) * double(sz[1]) * double(sz[2]) > double(std::numeric_limits<int>::max()))
        fail(DVR_ERR_INVALID_ARGUMENT, "Dataset is too big, must fit into an 32-bit signed int");
    for (int d = 0; d < 3; ++d) {
        h->size[d] = int64_t(sz[d]);
        h->voxel_size[d] = vs[d];
    }
    h->type = type;
    h->compressed = compressed ? 1 : 0;
}

// ---------------------------------------------------------------------------
// implicit datasets: f(x, y, z) on [boxMin, boxMax]^3 (renderer/volume_implicit.cpp)
// ---------------------------------------------------------------------------
struct Params {
    int n;
    const char* const* names;
    const float* values;
    float get(const char* key, float def) const
    {
        for (int i = 0; i < n; ++i)
            if (names[i] && std::strcmp(names[i], key) == 0) return values[i];
        return def;
    }
};

inline float sq(float v) { return v * v; }
inline float cube(float v) { return v * v * v; }
// zero level set -> density 0.5, clamped to [0, 1]
inline float level_to_density(float i) { return std::max(0.0f, std::min(1.0f, 0.5f - i)); }

void implicit_box(int eq, float& lo, float& hi)
{
    switch (eq) {
    case DVR_IMPLICIT_DING_DONG: case DVR_IMPLICIT_ENDRASS: case DVR_IMPLICIT_KUMMER: case DVR_IMPLICIT_BLOBBY:
        lo = -2.0f; hi = 2.0f; return;
    case DVR_IMPLICIT_BARTH: lo = -1.5f; hi = 1.5f; return;
    case DVR_IMPLICIT_KLEINE: lo = -5.0f; hi = 5.0f; return;
    case DVR_IMPLICIT_STEINER: lo = -0.5f; hi = 0.5f; return;
    default: lo = -1.0f; hi = 1.0f; return;
    }
}

// The expressions are the reference's, including its use of `z * 2` where a
// square z*z is written in the published surfaces (Barth, Heart, Kleine,
// Steiner, Cross-cap, Kummer, Blobby): the datasets must be the same.
float implicit_eval(int eq, float x, float y, float z, const Params& p)
{
    const float s2 = std::sqrt(2.0f);
    switch (eq) {
    case DVR_IMPLICIT_MARSCHNER_LOBB: {
        const double fM = p.get("fM", 6.0f), alpha = p.get("alpha", 0.25f);
        const double r = std::sqrt(x * x + y * y);  // float radius, as the reference evaluates it
        const double rho = std::cos(2.0 * M_PI * fM * std::cos(M_PI * r / 2.0));
        return float(((1.0 - std::sin(M_PI * z / 2.0)) + alpha * (1.0 + rho)) / (2.0 * (1.0 + alpha)));
    }
    case DVR_IMPLICIT_CUBE: {
        const float e = p.get("scale", 0.5f);
        const float ox = std::max(0.0f, std::fabs(x) - e), oy = std::max(0.0f, std::fabs(y) - e),
                    oz = std::max(0.0f, std::fabs(z) - e);
        return 1.0f - std::sqrt(sq(ox) + sq(oy) + sq(oz));
    }
    case DVR_IMPLICIT_SPHERE: return 1.0f - std::sqrt(sq(x) + sq(y) + sq(z));
    case DVR_IMPLICIT_INVERSE_SPHERE: return std::sqrt(sq(x) + sq(y) + sq(z));
    case DVR_IMPLICIT_DING_DONG: return level_to_density(x * x + y * y - z * (1.0f - z * z));
    case DVR_IMPLICIT_ENDRASS: {
        const float a = sq(x + y) - 2.0f, b = sq(x - y) - 2.0f;
        const float c = -4.0f * (1.0f - s2);
        const float d = 8.0f * (2.0f - s2) * z * z + 2.0f * (2.0f - 7.0f * s2) * (x * x + y * y);
        const float e = -16.0f * sq(sq(z)) + 8.0f * (1.0f + 2.0f * s2) * sq(z) - 1.0f + 12.0f * s2;
        return 0.5f + (64.0f * (x * x - 1.0f) * (y * y - 1.0f) * a * b - sq(c + d + e));
    }
    case DVR_IMPLICIT_BARTH: {
        const float zz = z + 0.5f;
        const float phi = (1.0f + std::sqrt(5.0f)) / 2.0f;
        const float x2 = x * x, y2 = y * y, z2 = zz * 2.0f, p2 = phi * phi;
        return 0.5f + (4.0f * (p2 * x2 - y2) * (p2 * y2 - z2) * (p2 * z2 - x2) -
                       (1.0f + 2.0f * phi) * sq(x2 + y2 + z2 - 1.0f));
    }
    case DVR_IMPLICIT_HEART: {
        const float x2 = x * x, y2 = y * y, z2 = z * 2.0f;
        return level_to_density(cube(2.0f * x2 + 2.0f * y2 + z2 - 1.0f) - 0.1f * x2 * z2 * z - y2 * z2 * z);
    }
    case DVR_IMPLICIT_KLEINE: {
        const float x2 = x * x, y2 = y * y, z2 = z * 2.0f;
        const float q = x2 + y2 + z2;
        return 0.5f + -((q + 2.0f * y - 1.0f) * sq(q - 2.0f * y - 1.0f) - 8.0f * z2 +
                        16.0f * x * y * (q - 2.0f * y - 1.0f));
    }
    case DVR_IMPLICIT_CASSINI: {
        const float a = p.get("a", 0.25f);
        return level_to_density((sq(x + a) + y * y) * (sq(x - a) + y * y) - z * z);
    }
    case DVR_IMPLICIT_STEINER: {
        const float x2 = x * x, y2 = y * y, z2 = z * 2.0f;
        return level_to_density(x2 * y2 + x2 * z2 + y2 * z2 - 2.0f * x * y * z);
    }
    case DVR_IMPLICIT_CROSS_CAP: {
        const float x2 = x * x, y2 = y * y, z2 = z * 2.0f;
        return level_to_density(4.0f * x2 * (x2 + y2 * z2 + z) + y2 * (y2 + z2 - 1.0f));
    }
    case DVR_IMPLICIT_KUMMER: {
        const float x2 = x * x, y2 = y * y, z2 = z * 2.0f;
        return level_to_density(x2 * x2 + y2 * y2 + z2 * z2 - x2 - y2 - z2 - x2 * y2 - y2 * z2 - z2 * x2 + 1.0f);
    }
    case DVR_IMPLICIT_BLOBBY: {
        const float x2 = x * x, y2 = y * y, z2 = z * 2.0f;
        return level_to_density(x2 + y2 * z2 + std::sin(4.0f * x) - std::cos(4.0f * y) + std::sin(4.0f * z) - 1.0f);
    }
    default: {  // DVR_IMPLICIT_TUBE
        const float r = std::sqrt(y * y + z * z);
        return ((1.0f - (r * cube(0.9f - 0.5f * std::cos(7.0f * x)))) - 0.9f) * 10.0f;
    }
    }
}

// ---------------------------------------------------------------------------
// colour space: sRGB (D65) <-> CIE L*a*b*, channels of sRGB in [0, 1]
// ---------------------------------------------------------------------------
constexpr float kWhite[3] = {95.047f, 100.0f, 108.883f};  // D65 reference white, Y = 100

inline float srgb_to_linear(float c) { return c > 0.04045f ? std::pow((c + 0.055f) / 1.055f, 2.4f) : c / 12.92f; }
inline float linear_to_srgb(float c)
{
    return c > 0.0031308f ? 1.055f * std::pow(c, 1.0f / 2.4f) - 0.055f : 12.92f * c;
}

void rgb_to_lab(const float rgb[3], float lab[3])
{
    const float r = srgb_to_linear(rgb[0]) * 100.0f, g = srgb_to_linear(rgb[1]) * 100.0f,
                b = srgb_to_linear(rgb[2]) * 100.0f;
    const float xyz[3] = {r * 0.4124564f + g * 0.3575761f + b * 0.1804375f,
                          r * 0.2126729f + g * 0.7151522f + b * 0.0721750f,
                          r * 0.0193339f + g * 0.1191920f + b * 0.9503041f};
    float f[3];
    for (int k = 0; k < 3; ++k) {
        const float t = xyz[k] / kWhite[k];
        f[k] = t > 0.008856f ? std::cbrt(t) : 7.787f * t + 16.0f / 116.0f;
    }
    lab[0] = 116.0f * f[1] - 16.0f;
    lab[1] = 500.0f * (f[0] - f[1]);
    lab[2] = 200.0f * (f[1] - f[2]);
}

void lab_to_rgb(const float lab[3], float rgb[3])
{
    const float fy = (lab[0] + 16.0f) / 116.0f;
    const float fx = lab[1] / 500.0f + fy;
    const float fz = fy - lab[2] / 200.0f;
    const float f[3] = {fx, fy, fz};
    float xyz[3];
    for (int k = 0; k < 3; ++k) {
        const float c = f[k] * f[k] * f[k];
        xyz[k] = (c > 0.008856f ? c : (f[k] - 16.0f / 116.0f) / 7.787f) * kWhite[k] / 100.0f;
    }
    const float r = xyz[0] * 3.2404542f + xyz[1] * -1.5371385f + xyz[2] * -0.4985314f;
    const float g = xyz[0] * -0.9692660f + xyz[1] * 1.8760108f + xyz[2] * 0.0415560f;
    const float b = xyz[0] * 0.0556434f + xyz[1] * -0.2040259f + xyz[2] * 1.0572252f;
    rgb[0] = linear_to_srgb(r);
    rgb[1] = linear_to_srgb(g);
    rgb[2] = linear_to_srgb(b);
}

}  // namespace

extern "C" {

dvr_status dvr_cvol_read_header(const char* path, dvr_cvol_header* header)
{
    return guarded([&] {
        if (!path || !header) fail(DVR_ERR_INVALID_ARGUMENT, "path or header is NULL");
        std::ifstream f(path, std::ios::binary);
        if (!f) fail(DVR_ERR_INVALID_ARGUMENT, "cannot open %s", path);
        read_header(f, path, header);
    });
}

dvr_status dvr_cvol_read(const char* path, float* data, int64_t count)
{
    return guarded([&] {
        if (!path || !data) fail(DVR_ERR_INVALID_ARGUMENT, "path or data is NULL");
        std::ifstream f(path, std::ios::binary);
        if (!f) fail(DVR_ERR_INVALID_ARGUMENT, "cannot open %s", path);
        dvr_cvol_header h;
        read_header(f, path, &h);
        const int64_t n = h.size[0] * h.size[1] * h.size[2];
        if (count != n) fail(DVR_ERR_INVALID_ARGUMENT, "%s holds %lld voxels, buffer has %lld", path, (long long)n,
                             (long long)count);
        if (h.compressed)
            fail(DVR_ERR_UNSUPPORTED,
                 "%s: LZ4-compressed .cvol files are not supported (the chunk framing of the reference's lz4 "
                 "wrapper is not available); store the volume uncompressed", path);
        const size_t bytes = size_t(n) * kBytesPerType[h.type];
        std::vector<unsigned char> raw(bytes);
        f.read(reinterpret_cast<char*>(raw.data()), std::streamsize(bytes));
        if (size_t(f.gcount()) != bytes) fail(DVR_ERR_INVALID_ARGUMENT, "%s: truncated voxel data", path);
        if (h.type == DVR_CVOL_FLOAT) {
            std::memcpy(data, raw.data(), bytes);
        } else if (h.type == DVR_CVOL_UCHAR) {
            for (int64_t i = 0; i < n; ++i) data[i] = raw[size_t(i)] / 255.0f;
        } else {
            const uint16_t* s = reinterpret_cast<const uint16_t*>(raw.data());
            for (int64_t i = 0; i < n; ++i) data[i] = s[i] / 65535.0f;
        }
    });
}

dvr_status dvr_cvol_write(const char* path, const float* data, const int64_t size[3], const double world_size[3],
                          int32_t compression)
{
    return guarded([&] {
        if (!path || !data || !size || !world_size) fail(DVR_ERR_INVALID_ARGUMENT, "NULL argument");
        if (compression < 0 || compression > 9) fail(DVR_ERR_INVALID_ARGUMENT, "Illegal compression factor");
        if (compression > 0)
            fail(DVR_ERR_UNSUPPORTED, "LZ4 compression of .cvol files is not supported; use compression 0");
        for (int d = 0; d < 3; ++d)
            if (size[d] <= 0) fail(DVR_ERR_INVALID_ARGUMENT, "volume size must be positive");
        std::ofstream f(path, std::ios::binary);
        if (!f) fail(DVR_ERR_INVALID_ARGUMENT, "cannot create %s", path);
        f.write(kMagic, 4);
        for (int d = 0; d < 3; ++d) {
            const uint64_t s = uint64_t(size[d]);
            f.write(reinterpret_cast<const char*>(&s), 8);
        }
        for (int d = 0; d < 3; ++d) {
            const double v = world_size[d] / double(size[d]);  // renderer/volume.cpp:104-106
            f.write(reinterpret_cast<const char*>(&v), 8);
        }
        const int32_t type = DVR_CVOL_FLOAT;
        f.write(reinterpret_cast<const char*>(&type), 4);
        const char pad[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        f.write(pad, 8);  // compression flag 0 + 7 bytes padding
        f.write(reinterpret_cast<const char*>(data), std::streamsize(sizeof(float) * size[0] * size[1] * size[2]));
        if (!f) fail(DVR_ERR_INVALID_ARGUMENT, "writing %s failed", path);
    });
}

dvr_status dvr_volume_implicit(int32_t equation, int32_t resolution, int32_t nparams, const char* const* names,
                               const float* values, float* data)
{
    return guarded([&] {
        if (equation < 0 || equation > DVR_IMPLICIT_TUBE) fail(DVR_ERR_INVALID_ARGUMENT, "unknown implicit equation %d", equation);
        if (resolution < 2) fail(DVR_ERR_INVALID_ARGUMENT, "resolution must be at least 2");
        if (!data || (nparams > 0 && (!names || !values))) fail(DVR_ERR_INVALID_ARGUMENT, "NULL argument");
        const Params p{nparams, names, values};
        float lo, hi;
        implicit_box(equation, lo, hi);
        const float scale = (hi - lo) / float(resolution - 1);
        const int64_t R = resolution;
        for (int64_t z = 0; z < R; ++z)
            for (int64_t y = 0; y < R; ++y)
                for (int64_t x = 0; x < R; ++x)
                    data[x + R * (y + R * z)] = implicit_eval(equation, lo + float(x) * scale, lo + float(y) * scale,
                                                              lo + float(z) * scale, p);
    });
}

dvr_status dvr_volume_histogram(const float* data, const int64_t size[3], const int64_t strides[3], float* bins,
                                float* min_density, float* max_density, uint32_t* nonzero)
{
    return guarded([&] {
        if (!data || !size || !strides || !bins || !min_density || !max_density || !nonzero)
            fail(DVR_ERR_INVALID_ARGUMENT, "NULL argument");
        float mn = FLT_MAX, mx = 0.0f;
        uint32_t nnz = 0;
        for (int64_t x = 0; x < size[0]; ++x)
            for (int64_t y = 0; y < size[1]; ++y)
                for (int64_t z = 0; z < size[2]; ++z) {
                    const float d = data[x * strides[0] + y * strides[1] + z * strides[2]];
                    mn = std::fmin(mn, d);
                    mx = std::fmax(mx, d);
                    if (d > 0) ++nnz;
                }
        for (int i = 0; i < DVR_HISTOGRAM_BINS; ++i) bins[i] = 0.0f;
        const float frac = 1.0f / float(nnz);
        const float width = (mx - mn) / float(DVR_HISTOGRAM_BINS);
        for (int64_t x = 0; x < size[0]; ++x)
            for (int64_t y = 0; y < size[1]; ++y)
                for (int64_t z = 0; z < size[2]; ++z) {
                    const float d = data[x * strides[0] + y * strides[1] + z * strides[2]];
                    if (d > 0) {
                        const int b = std::min(int((d - mn) / width), DVR_HISTOGRAM_BINS - 1);
                        bins[b] += frac;
                    }
                }
        *min_density = mn;
        *max_density = mx;
        *nonzero = nnz;
    });
}

void dvr_rgb_to_lab(const float rgb[3], float lab[3]) { rgb_to_lab(rgb, lab); }
void dvr_lab_to_rgb(const float lab[3], float rgb[3]) { lab_to_rgb(lab, rgb); }

dvr_status dvr_tf_assemble_from_settings(const float* color_lab, const float* color_pos, int32_t nc,
                                         const float* opacity, const float* opacity_pos, int32_t no,
                                         float min_density, float max_density, float opacity_scaling,
                                         int32_t purge_zero_regions, dvr_tf_point* points, int32_t capacity,
                                         int32_t* count)
{
    return guarded([&] {
        if (nc < 1 || no < 1) fail(DVR_ERR_INVALID_ARGUMENT, "at least one colour and one opacity control point");
        if (!color_lab || !color_pos || !opacity || !opacity_pos || !count) fail(DVR_ERR_INVALID_ARGUMENT, "NULL argument");
        struct C3 { float v[3]; };
        std::vector<C3> cv(static_cast<size_t>(nc));
        std::vector<float> cp(color_pos, color_pos + nc), ov(opacity, opacity + no), op(opacity_pos, opacity_pos + no);
        for (int i = 0; i < nc; ++i) for (int k = 0; k < 3; ++k) cv[size_t(i)].v[k] = color_lab[3 * i + k];
        // editor positions in [0, 1] -> densities
        for (float& d : cp) d = min_density + d * (max_density - min_density);
        for (float& d : op) d = min_density + d * (max_density - min_density);
        // sentinel points outside [0, 1] (at -1 and 2) when the axes do not reach the ends
        if (cp.front() > 0) { cp.insert(cp.begin(), -1.0f); cv.insert(cv.begin(), cv.front()); }
        if (op.front() > 0) { op.insert(op.begin(), -1.0f); ov.insert(ov.begin(), ov.front()); }
        if (cp.back() < 1) { cp.push_back(2.0f); cv.push_back(cv.back()); }
        if (op.back() < 1) { op.push_back(2.0f); ov.push_back(ov.back()); }
        auto rgb_of = [](const float lab[3], float out[3]) { lab_to_rgb(lab, out); };
        std::vector<dvr_tf_point> pts;
        {
            dvr_tf_point p;
            p.pos = cp[0];
            rgb_of(cv[0].v, p.val);
            p.val[3] = ov[0];
            pts.push_back(p);
        }
        // merge the two axes in density order; each new point interpolates the other axis
        size_t io = 0, ic = 0;
        while (io + 1 < op.size() && ic + 1 < cp.size()) {
            dvr_tf_point p;
            if (op[io + 1] < cp[ic + 1]) {
                const float t = (op[io + 1] - cp[ic]) / (cp[ic + 1] - cp[ic]);
                float lab[3];
                for (int k = 0; k < 3; ++k) lab[k] = cv[ic].v[k] + t * (cv[ic + 1].v[k] - cv[ic].v[k]);
                p.pos = op[io + 1];
                rgb_of(lab, p.val);
                p.val[3] = ov[io + 1];
                ++io;
            } else {
                const float t = (cp[ic + 1] - op[io]) / (op[io + 1] - op[io]);
                p.pos = cp[ic + 1];
                rgb_of(cv[ic + 1].v, p.val);
                p.val[3] = ov[io] + t * (ov[io + 1] - ov[io]);
                ++ic;
            }
            pts.push_back(p);
        }
        if (purge_zero_regions) {
            // drop a point between two others when all three are transparent, and points
            // closer than 1e-7 to their predecessor
            const float eps = 1e-7f;
            for (int64_t i = 0; i < int64_t(pts.size()) - 2;) {
                const bool transparent = pts[size_t(i)].val[3] < eps && pts[size_t(i + 1)].val[3] < eps &&
                                         pts[size_t(i + 2)].val[3] < eps;
                if (transparent || pts[size_t(i + 1)].pos - pts[size_t(i)].pos < eps) pts.erase(pts.begin() + (i + 1));
                else ++i;
            }
        }
        for (auto& p : pts) {
            for (int k = 0; k < 4; ++k) p.val[k] = std::min(std::max(p.val[k], 0.0f), 1.0f - FLT_EPSILON);
            p.val[3] *= opacity_scaling;
        }
        *count = int32_t(pts.size());
        if (points) {
            if (capacity < int32_t(pts.size()))
                fail(DVR_ERR_INVALID_ARGUMENT, "points needs room for %d control points", int(pts.size()));
            std::copy(pts.begin(), pts.end(), points);
        }
    });
}

dvr_status dvr_tf_texture_from_points(const dvr_tf_point* points, int32_t n, int32_t resolution, float* texture)
{
    return guarded([&] {
        if (!points || !texture) fail(DVR_ERR_INVALID_ARGUMENT, "NULL argument");
        if (n < 2) fail(DVR_ERR_INVALID_ARGUMENT, "at least two control points");
        if (resolution < 1) fail(DVR_ERR_INVALID_ARGUMENT, "resolution must be positive");
        for (int i = 0; i < resolution; ++i) {
            const float density = (float(i) + 0.5f) / float(resolution);  // texel centre
            int k = 0;
            while (k < n - 2 && !(points[k + 1].pos > density)) ++k;
            const float lo = points[k].pos, hi = points[k + 1].pos;
            const float f = (density - lo) / (hi - lo);
            for (int c = 0; c < 4; ++c) texture[4 * i + c] = (1.0f - f) * points[k].val[c] + f * points[k + 1].val[c];
        }
    });
}

}  // extern "C"
