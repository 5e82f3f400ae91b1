// gfusion.hpp -- C++ drop-in for `class gFusion` (include/gFusion.h:124-276 of
// the reference) on top of the C ABI in include/kfusion.h.
//
// Same method names, argument meaning and call order as the reference, so the
// frame block of kinect.cpp:654-695 compiles against it unchanged apart from
// the matrix type (the reference uses glm::mat4; here Mat4 is 16 floats in
// glm's column-major layout, m[c*4 + r]).  GL texture ids become device
// buffers of a kfusion session; the getters that returned texture ids copy
// the maps to host memory instead.  Errors: the reference exits on shader
// build failure and has no GPU error checking; here every failing call throws
// gfusion_error carrying the kfusion status and message.
#pragma once

#include <algorithm>
#include <array>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "kfusion.h"

struct Mat4 {
    std::array<float, 16> m{1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    float *data() { return m.data(); }
    const float *data() const { return m.data(); }
    float &operator()(int col, int row) { return m[col * 4 + row]; }
    float operator()(int col, int row) const { return m[col * 4 + row]; }
    static Mat4 translate(float x, float y, float z)
    {
        Mat4 t;
        t(3, 0) = x;
        t(3, 1) = y;
        t(3, 2) = z;
        return t;
    }
};

struct Vec4 {
    float x, y, z, w;
};

class gfusion_error : public std::runtime_error {
public:
    gfusion_error(int code, const std::string &what) : std::runtime_error(what), code(code) {}
    int code;
};

// gFusionConfig: include/gFusion.h:42-82
struct gFusionConfig {
    float volumeSize[3] = {256, 256, 256};
    float volumeDimensions[3] = {1.0f, 1.0f, 1.0f};
    float mu = 0.05f;
    float maxWeight = 100.0f;
    float nearPlane = 0.1f;
    float farPlane = 3.0f;
    float dist_threshold = 0.05f;
    float normal_threshold = 0.8f;
    std::vector<int> iterations{5, 5, 5};
    float track_threshold = 0.5f;
    int storage = KF_STORAGE_RG16I;  // beyond the reference: TSDF format
    int reduce_rule = KF_RULE_LAST_ROW;

    float stepSize() const
    {
        float minxyz = std::min(std::min(volumeDimensions[0], volumeDimensions[1]), volumeDimensions[2]);
        float maxxyz = std::max(std::max(volumeSize[0], volumeSize[1]), volumeSize[2]);
        return minxyz / maxxyz;
    }
};

class gFusion {
public:
    gFusion() = default;
    ~gFusion()
    {
        if (s_) kf_session_destroy(s_);
    }
    gFusion(const gFusion &) = delete;
    gFusion &operator=(const gFusion &) = delete;

    gFusionConfig configuration;

    // -- configuration ----------------------------------------------------
    void setConfig(const gFusionConfig &config) { configuration = config; }
    void setCameraParams(Vec4 camPams, Vec4 /*camPamsColor*/ = Vec4{0, 0, 0, 0})
    {
        cam_[0] = camPams.x;
        cam_[1] = camPams.y;
        cam_[2] = camPams.z;
        cam_[3] = camPams.w;
    }
    void setDepthSize(int w, int h)
    {
        width_ = w;
        height_ = h;
    }
    void setStream(void *hip_stream) { stream_ = hip_stream; }
    void setPose(const Mat4 &pose)
    {
        pose_ = pose;
        if (s_) check(kf_set_pose(s_, pose.data()));
    }
    Mat4 getPose()
    {
        if (!s_) return pose_;
        Mat4 p;
        check(kf_get_pose(s_, p.data()));
        return p;
    }
    std::vector<float> getTransPose()
    {
        Mat4 p = getPose();
        return {p(3, 0), p(3, 1), p(3, 2)};
    }

    // -- allocation (gFusion.cpp:200-330) -----------------------------------
    void queryDeviceLimits() {}
    void compileAndLinkShader() {}
    void setLocations() {}
    void initTextures() {}
    void initVolume() {}
    void allocateBuffers()
    {
        if (s_) {
            kf_session_destroy(s_);
            s_ = nullptr;
        }
        kf_config c;
        check(kf_config_default(&c));
        for (int k = 0; k < 3; ++k) {
            c.volume_size[k] = (int)configuration.volumeSize[k];
            c.volume_dim[k] = configuration.volumeDimensions[k];
            c.iterations[k] = k < (int)configuration.iterations.size() ? configuration.iterations[k] : 0;
        }
        c.mu = configuration.mu;
        c.max_weight = configuration.maxWeight;
        c.near_plane = configuration.nearPlane;
        c.far_plane = configuration.farPlane;
        c.dist_threshold = configuration.dist_threshold;
        c.normal_threshold = configuration.normal_threshold;
        for (int k = 0; k < 4; ++k) c.cam[k] = cam_[k];
        c.width = width_;
        c.height = height_;
        c.storage = configuration.storage;
        c.reduce_rule = configuration.reduce_rule;
        check(kf_session_create(&c, stream_, &s_));
        check(kf_set_pose(s_, pose_.data()));
    }

    // -- reset (gFusion.cpp:226-287) ---------------------------------------
    void Reset(const Mat4 &pose, bool deleteFlag)
    {
        if (deleteFlag) allocateBuffers();
        pose_ = pose;
        check(kf_reset(session(), pose.data()));
    }
    void resetVolume() { check(kf_reset_volume(session())); }
    void resetPose(const Mat4 &pose) { setPose(pose); }

    // -- per frame (kinect.cpp:654-695) --------------------------------------
    void depthToVertex(const float *depthArray)
    {
        if (depthArray) check(kf_depth_to_vertex_f32(session(), depthArray));
    }
    void depthToVertex(const uint16_t *depthArray)
    {
        if (depthArray) check(kf_depth_to_vertex_u16(session(), depthArray));
    }
    void vertexToNormal() {}  // produced by the depthToVertex launch
    bool Track()
    {
        int tracked = 0;
        check(kf_track(session(), &tracked));
        return tracked != 0;
    }
    // TrackSDF (gFusion.cpp:835-1038) and recoverPose (:1616-1653)
    bool TrackSDF()
    {
        int tracked = 0;
        check(kf_track_sdf(session(), &tracked));
        return tracked != 0;
    }
    bool recoverPose()
    {
        int found = 0;
        check(kf_recover_pose(session(), &found));
        return found != 0;
    }
    std::vector<float> getCumTwist()
    {
        std::vector<float> t(6);
        check(kf_get_cum_twist(session(), t.data()));
        return t;
    }
    void track(int layer) { check(kf_icp_iteration(session(), layer, 1, 1)); }
    void reduce(int /*layer*/) {}  // fused into track()
    void getReduction(std::vector<float> &b, std::vector<float> &C, float &alignmentEnergy)
    {
        float parts[256];
        check(kf_read_reduction(session(), parts, &alignmentEnergy));
        std::vector<float> row(32, 0.0f);
        for (int k = 0; k < 32; ++k) {
            if (configuration.reduce_rule == KF_RULE_LAST_ROW) {
                row[k] = parts[7 * 32 + k];
            } else {
                for (int g = 0; g < 8; ++g) row[k] += parts[g * 32 + k];
            }
        }
        b.assign(row.begin() + 1, row.begin() + 7);
        C.assign(36, 0.0f);
        int q = 7;
        for (int i = 0; i < 6; ++i)
            for (int j = i; j < 6; ++j, ++q) C[i * 6 + j] = C[j * 6 + i] = row[q];
    }
    void intensityProjection() { check(kf_intensity_projection(session())); }
    std::vector<float> getIntensityProjection() { return readMap(KF_MAP_INTENSITY, 0, 4); }
    void integrate() { check(kf_integrate(session(), KF_INTEGRATE_ALWAYS)); }
    void raycast() { check(kf_raycast(session())); }
    // The whole frame block without host synchronisation (device-side gate).
    void frame(const float *depthArray) { check(kf_frame_f32(session(), depthArray)); }

    // -- read-back -----------------------------------------------------------
    std::vector<float> getVerts() { return readMap(KF_MAP_REF_VERTEX, 0, 4); }
    std::vector<float> getNorms() { return readMap(KF_MAP_REF_NORMAL, 0, 4); }
    std::vector<float> getDepthImage() { return readMap(KF_MAP_DEPTH, 0, 1); }
    std::vector<float> readMap(int which, int level, int comps)
    {
        std::vector<float> out((size_t)(width_ >> level) * (height_ >> level) * comps);
        check(kf_read_map(session(), which, level, out.data(), out.size() * sizeof(float)));
        return out;
    }
    kf_session *session()
    {
        if (!s_) throw gfusion_error(KF_ERR_STATE, "allocateBuffers() has not been called");
        return s_;
    }

private:
    static void check(int rc)
    {
        if (rc != KF_OK) throw gfusion_error(rc, std::string("kfusion: ") + kf_last_error());
    }
    kf_session *s_ = nullptr;
    void *stream_ = nullptr;
    float cam_[4] = {365.5f, 365.5f, 255.5f, 211.5f};
    int width_ = 512, height_ = 424;
    Mat4 pose_;
};

// mCubeConfig (include/mcubes.h:25-43)
struct mCubeConfig {
    float gridSize[3] = {128, 128, 128};
    unsigned numVoxels = 128 * 128 * 128;
    float isoValue = 1650.0f;
    unsigned maxVerts = 128 * 128 * 128;
};

// C++ drop-in for `class MCubes` (include/mcubes.h:46-170) over kf_mesh_*:
// the volume is the gFusion session's TSDF.
class MCubes {
public:
    void setVolumeTexture(gFusion &g) { g_ = &g; }
    void setConfig(const mCubeConfig &c) { config_ = c; }
    void init() {}
    void setIsolevel(float iso) { iso_ = iso; }
    void generateMarchingCubes()
    {
        uint64_t n = 0;
        check(kf_mesh_generate(g_->session(), iso_, &n));
        triangles_ = n;
    }
    // m_totalSum: the vertex count (mcubes.h:76-79)
    int getNumberTriangles() const { return (int)(3 * triangles_); }
    std::vector<float> getPosBuffer()
    {
        std::vector<float> v((size_t)triangles_ * 12);
        if (triangles_) check(kf_mesh_read(g_->session(), v.data(), triangles_));
        return v;
    }
    void exportMesh(const std::string &path = "data/meshes/marchingCubesBin.stl")
    {
        check(kf_mesh_export_stl(g_->session(), path.c_str()));
    }
    void exportPointCloud(const std::string &path = "data/meshes/marchingCubesBin.ply")
    {
        check(kf_mesh_export_ply(g_->session(), path.c_str()));
    }

private:
    static void check(int rc)
    {
        if (rc != KF_OK) throw gfusion_error(rc, std::string("kfusion: ") + kf_last_error());
    }
    gFusion *g_ = nullptr;
    mCubeConfig config_;
    float iso_ = 0.0f;
    uint64_t triangles_ = 0;
};

// loadCalibration (kinect.cpp:78-161): 1 on success, 0 for an extension other
// than .xml / .json; throws gfusion_error for a file it cannot read or parse.
inline int loadCalibration(const std::string &calibFile, kf_cam_params &out)
{
    const auto ends = [&](const char *e) {
        const size_t n = std::string(e).size();
        return calibFile.size() >= n && calibFile.compare(calibFile.size() - n, n, e) == 0;
    };
    if (!ends(".xml") && !ends(".json")) return 0;
    const int rc = kf_load_calibration(calibFile.c_str(), &out);
    if (rc != KF_OK) throw gfusion_error(rc, std::string("kfusion: ") + kf_last_error());
    return 1;
}
