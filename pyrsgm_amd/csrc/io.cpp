// io.cpp -- PFM files, MiddEval3 calib.txt and disparity metrics on the host
// (pyrsgm_io.h).
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <fstream>
#include <string>

#include "pyrsgm/pyrsgm_io.h"

namespace {

void copyErr(const std::string& what, char* err, int errlen)
{
    if (err && errlen > 0) snprintf(err, (size_t)errlen, "%s", what.c_str());
}

template <class T>
const T* rowAt(const T* base, size_t pitch, int y)
{
    return reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + (size_t)y * pitch);
}

uint32_t byteSwap(uint32_t u)
{
    return ((u & 0xffu) << 24) | ((u & 0xff00u) << 8) | ((u >> 8) & 0xff00u) | (u >> 24);
}

} // namespace

extern "C" {

// Reference: bench/middlebury.h:22-53 (header tokens, one whitespace byte,
// bottom row first, byte order from the scale's sign).
int pyrsgm_read_pfm(const char* path, float** data, int* width, int* height, char* err,
                    int errlen)
{
    if (!path || !data || !width || !height) {
        copyErr("pyrsgm: null argument", err, errlen);
        return -1;
    }
    *data = nullptr;
    std::ifstream f(path, std::ios::binary);
    if (!f) {
        copyErr(std::string("cannot open ") + path, err, errlen);
        return -1;
    }
    std::string magic;
    f >> magic;
    if (magic != "Pf") {
        copyErr(std::string(path) + ": only grayscale PFM supported", err, errlen);
        return -1;
    }
    long w = 0, h = 0;
    double scale = 0.0;
    f >> w >> h >> scale;
    f.get();
    if (!f || w <= 0 || h <= 0 || w > (1 << 20) || h > (1 << 20)) {
        copyErr(std::string(path) + ": bad PFM header", err, errlen);
        return -1;
    }
    const bool little = scale < 0.0;
    float* out = static_cast<float*>(malloc((size_t)w * (size_t)h * sizeof(float)));
    if (!out) {
        copyErr("pyrsgm: out of host memory", err, errlen);
        return -1;
    }
    for (long y = h - 1; y >= 0; --y) {
        float* row = out + (size_t)y * (size_t)w;
        f.read(reinterpret_cast<char*>(row), (std::streamsize)((size_t)w * sizeof(float)));
        if (!f) {
            free(out);
            copyErr(std::string(path) + ": truncated PFM", err, errlen);
            return -1;
        }
        if (!little) {
            for (long x = 0; x < w; ++x) {
                uint32_t u;
                memcpy(&u, &row[x], 4);
                u = byteSwap(u);
                memcpy(&row[x], &u, 4);
            }
        }
    }
    *data = out;
    *width = (int)w;
    *height = (int)h;
    return 0;
}

// Reference: examples/stereo_pair.cpp:38-45.
int pyrsgm_write_pfm(const char* path, const float* data, size_t pitch_bytes, int width,
                     int height)
{
    if (!path || !data || width <= 0 || height <= 0 || pitch_bytes < (size_t)width * sizeof(float))
        return -1;
    std::ofstream f(path, std::ios::binary);
    if (!f) return -1;
    f << "Pf\n" << width << " " << height << "\n-1.0\n";
    for (int y = height - 1; y >= 0; --y)
        f.write(reinterpret_cast<const char*>(rowAt(data, pitch_bytes, y)),
                (std::streamsize)((size_t)width * sizeof(float)));
    return f ? 0 : -1;
}

void pyrsgm_free(void* p) { free(p); }

// Reference: bench/middlebury.h:56-64.
int pyrsgm_read_ndisp(const char* path)
{
    if (!path) return 0;
    std::ifstream f(path);
    std::string line;
    while (std::getline(f, line))
        if (line.rfind("ndisp=", 0) == 0) return atoi(line.c_str() + 6);
    return 0;
}

// Reference: bench/pyrsgm_bench.cpp:443-453.
void pyrsgm_disp16_to_float(const short* disp16, size_t disp16_pitch, float* out,
                            size_t out_pitch, int width, int height)
{
    if (!disp16 || !out) return;
    for (int y = 0; y < height; ++y) {
        const short* s = rowAt(disp16, disp16_pitch, y);
        float* d = const_cast<float*>(rowAt(out, out_pitch, y));
        for (int x = 0; x < width; ++x) d[x] = s[x] > 0 ? (float)s[x] / 16.0f : 0.0f;
    }
}

// Reference: bench/middlebury.h:80-120 (double accumulation, strict > thresholds).
int pyrsgm_evaluate(const float* disp, size_t disp_pitch, const float* gt, size_t gt_pitch,
                    const unsigned char* mask, size_t mask_pitch, int width, int height,
                    float max_disp, pyrsgm_metrics_c* out)
{
    if (!disp || !gt || !out || width <= 0 || height <= 0) return -1;
    pyrsgm_metrics_c m;
    memset(&m, 0, sizeof(m));
    double sum = 0.0, sq = 0.0;
    long b05 = 0, b1 = 0, b2 = 0, b4 = 0;
    for (int y = 0; y < height; ++y) {
        const float* g = rowAt(gt, gt_pitch, y);
        const float* d = rowAt(disp, disp_pitch, y);
        const unsigned char* mk = mask ? rowAt(mask, mask_pitch, y) : nullptr;
        for (int x = 0; x < width; ++x) {
            if (!isfinite(g[x]) || g[x] <= 0.0f || g[x] > max_disp) continue;
            if (mk && mk[x] != 255) continue;
            ++m.evaluated;
            if (!(d[x] > 0.0f)) continue;
            ++m.valid;
            const double e = fabs((double)d[x] - (double)g[x]);
            sum += e;
            sq += e * e;
            b05 += e > 0.5;
            b1 += e > 1.0;
            b2 += e > 2.0;
            b4 += e > 4.0;
        }
    }
    if (m.valid) {
        m.avg_err = sum / m.valid;
        m.rms_err = sqrt(sq / m.valid);
        m.bad05 = 100.0 * (double)b05 / m.valid;
        m.bad1 = 100.0 * (double)b1 / m.valid;
        m.bad2 = 100.0 * (double)b2 / m.valid;
        m.bad4 = 100.0 * (double)b4 / m.valid;
    }
    if (m.evaluated) {
        m.density = 100.0 * m.valid / m.evaluated;
        m.invalid = 100.0 - m.density;
        m.bad1_all = 100.0 * (double)(b1 + (m.evaluated - m.valid)) / m.evaluated;
        m.bad2_all = 100.0 * (double)(b2 + (m.evaluated - m.valid)) / m.evaluated;
    }
    *out = m;
    return 0;
}

} // extern "C"
