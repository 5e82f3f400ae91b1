// kf_gfusion_example.cpp -- the frame block of kinect.cpp:654-695 written
// against the C++ gFusion drop-in, fed by the seeded synthetic sequence.
//
//   kf_gfusion_example [frames] [volume] [storage 0|1|2]
//
// Prints one line per frame: frame, tracked, translation, error vs ground
// truth (mm).  Exit status 0 when every frame after the first tracked.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "gfusion.hpp"

extern "C" int kfs_trajectory(uint64_t seed, int frames, float *poses);
extern "C" int kfs_render(uint64_t seed, const float *poses, int frames, int first_frame, const float cam[4], int w,
                          int h, int noise, uint16_t *out);

int main(int argc, char **argv)
{
    const int frames = argc > 1 ? std::atoi(argv[1]) : 30;
    const int volume = argc > 2 ? std::atoi(argv[2]) : 256;
    const int storage = argc > 3 ? std::atoi(argv[3]) : KF_STORAGE_F32;
    const int W = 512, H = 424;
    const float cam[4] = {365.5f, 365.5f, 255.5f, 211.5f};
    const uint64_t seed = 1000;

    std::vector<float> gt(16 * (size_t)frames);
    std::vector<uint16_t> depth((size_t)W * H * frames);
    if (kfs_trajectory(seed, frames, gt.data()) || kfs_render(seed, gt.data(), frames, 0, cam, W, H, 1, depth.data())) {
        std::fprintf(stderr, "synthetic sequence failed\n");
        return 2;
    }

    gFusion gfusion;
    gFusionConfig gconfig;
    for (int k = 0; k < 3; ++k) {
        gconfig.volumeSize[k] = (float)volume;
        gconfig.volumeDimensions[k] = 2.0f;
    }
    gconfig.mu = 3.0f * gconfig.stepSize();
    gconfig.iterations = {4, 5, 10};
    gconfig.storage = storage;
    try {
        // gFusionInit (kinect.cpp:23-46) + "Reset Volume" (kinect.cpp:446-477)
        gfusion.setConfig(gconfig);
        gfusion.setCameraParams(Vec4{cam[0], cam[1], cam[2], cam[3]});
        Mat4 initPose = Mat4::translate(gconfig.volumeDimensions[0] / 2.0f, gconfig.volumeDimensions[1] / 2.0f, 0.0f);
        gfusion.setPose(initPose);
        gfusion.initTextures();
        gfusion.initVolume();
        gfusion.allocateBuffers();
        gfusion.Reset(initPose, false);
        bool reset = true;
        int counter = 0;
        gfusion.raycast();

        int untracked = 0;
        std::vector<float> depthArray((size_t)W * H);
        for (int f = 0; f < frames; ++f) {
            const uint16_t *d = depth.data() + (size_t)f * W * H;
            for (size_t i = 0; i < depthArray.size(); ++i) depthArray[i] = (float)d[i];
            gfusion.depthToVertex(depthArray.data());
            gfusion.vertexToNormal();
            bool tracked = gfusion.Track();
            if (tracked || reset) {
                gfusion.integrate();
                if (counter > 2) reset = false;
            }
            if (!tracked && f > 0) ++untracked;
            if (!tracked) gfusion.recoverPose();
            counter++;
            gfusion.raycast();
            std::vector<float> t = gfusion.getTransPose();
            const float *g = gt.data() + 16 * (size_t)f;
            const double err = std::sqrt((t[0] - g[12]) * (t[0] - g[12]) + (t[1] - g[13]) * (t[1] - g[13]) +
                                         (t[2] - g[14]) * (t[2] - g[14]));
            std::printf("%d %d %.5f %.5f %.5f %.2f\n", f, tracked ? 1 : 0, t[0], t[1], t[2], err * 1000.0);
        }
        return untracked == 0 ? 0 : 1;
    } catch (const gfusion_error &e) {
        std::fprintf(stderr, "%s (status %d)\n", e.what(), e.code);
        return 3;
    }
}
