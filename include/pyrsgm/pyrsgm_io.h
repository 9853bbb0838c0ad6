/*
 * pyrsgm_io.h -- on-disk formats and disparity metrics of the reference's
 * benchmark and examples, as host C functions.
 *
 * Counterparts: mb::readPFM, mb::readNdisp, mb::Metrics, mb::evaluate
 * (reference bench/middlebury.h:22-120), writePfm (examples/stereo_pair.cpp:38-45)
 * and the bench's fixed-point to float conversion (bench/pyrsgm_bench.cpp:443-453).
 * The reference implements them on cv::Mat; these take plain host pointers and
 * row pitches in bytes, so that the library needs no OpenCV.  Everything here
 * runs on the host.
 */
#ifndef PYRSGM_IO_H
#define PYRSGM_IO_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* mb::Metrics: percentages are of the valid (bad*) or evaluated (density,
 * invalid, bad*_all) pixels. */
typedef struct {
    int evaluated, valid;
    double avg_err, rms_err;
    double bad05, bad1, bad2, bad4;
    double density, invalid;
    double bad1_all, bad2_all;
} pyrsgm_metrics_c;

/* Grayscale PFM ("Pf") into a malloc'd row-major float array (free it with
 * pyrsgm_free).  Big- and little-endian (scale sign) and the bottom-up row
 * order are handled; inf (unknown) stays inf.  0, or -1 with a message. */
int pyrsgm_read_pfm(const char* path, float** data, int* width, int* height, char* err,
                    int errlen);
/* Little-endian grayscale PFM (scale -1.0), bottom row first.  0 / -1. */
int pyrsgm_write_pfm(const char* path, const float* data, size_t pitch_bytes, int width,
                     int height);
void pyrsgm_free(void* p);
/* ndisp of a MiddEval3 calib.txt ("ndisp=" line); 0 if absent or unreadable. */
int pyrsgm_read_ndisp(const char* path);
/* disp32[x] = disp16[x] > 0 ? disp16[x] / 16 : 0 (values <= 0 are invalid). */
void pyrsgm_disp16_to_float(const short* disp16, size_t disp16_pitch, float* out,
                            size_t out_pitch, int width, int height);
/* mb::evaluate: disp <= 0 is invalid; gt inf, <= 0 or > max_disp is unknown;
 * mask (optional, 255 = evaluate) selects non-occluded pixels. */
int pyrsgm_evaluate(const float* disp, size_t disp_pitch, const float* gt, size_t gt_pitch,
                    const unsigned char* mask, size_t mask_pitch, int width, int height,
                    float max_disp, pyrsgm_metrics_c* out);

#ifdef __cplusplus
}
#endif
#endif
