// Builds against the installed headers and library; with a GPU and any
// argument it also matches one small synthetic pair (prints the level count and
// the centre disparity).
#include <pyrsgm/pyrsgm.h>

#include <cstdio>
#include <vector>

#define CHECK(x)                                                               \
    do {                                                                       \
        if ((x) != hipSuccess) {                                               \
            std::fprintf(stderr, "%s failed\n", #x);                           \
            return 1;                                                          \
        }                                                                      \
    } while (0)

int main(int argc, char** argv)
{
    (void)argv;
    std::printf("pyrsgm %s\n", pyrsgm::version());
    if (argc < 2) return 0;  // link check only
    const int W = 320, H = 240;
    pyrsgm::Params p;
    p.num_disparities = 64;
    p.pyramid_levels = 1;
    pyrsgm::StereoMatcher sgm(W, H, p);
    std::vector<uint8_t> img(W * H);
    for (int i = 0; i < W * H; ++i) img[i] = (uint8_t)((i * 2654435761u) >> 24);
    uint8_t *dl = nullptr, *dr = nullptr;
    int16_t* d16 = nullptr;
    CHECK(hipMalloc(&dl, W * H));
    CHECK(hipMalloc(&dr, W * H));
    CHECK(hipMalloc(&d16, W * H * 2));
    CHECK(hipMemcpy(dl, img.data(), W * H, hipMemcpyHostToDevice));
    CHECK(hipMemcpy(dr, img.data(), W * H, hipMemcpyHostToDevice));
    sgm.compute(dl, W, dr, W, d16, W * 2, nullptr, 0, nullptr, 0);
    std::vector<int16_t> out(W * H);
    CHECK(hipMemcpy(out.data(), d16, W * H * 2, hipMemcpyDeviceToHost));
    std::printf("levels %d, disparity at centre %d\n", sgm.numLevels(), out[(H / 2) * W + W / 2]);
    CHECK(hipFree(dl));
    CHECK(hipFree(dr));
    CHECK(hipFree(d16));
    return 0;
}
