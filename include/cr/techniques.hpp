// C++ host-side mirror of the reference's technique plugin interface
// (CloudRendering-Vulkan/RenderTechnique.h:15-66 and its subclasses
// RenderTechniquePT.h:11, RenderTechniqueSV.h:10, RenderTechniquePPB.h:10),
// header-only over the C-ABI of include/cr/abi.h.
//
// Differences that follow from dropping Vulkan:
//  * descriptor infos become the uniform structs themselves (by value, as
//    the reference copies them into mapped buffers, VulkanBuffer.cpp:33-39);
//  * RecordDrawCommands issues the work on the context's HIP stream instead
//    of recording into a VkCommandBuffer; the swapchain blit is dropped;
//  * errors surface as cr::Error (the reference aborts on a Vulkan error,
//    Validation.cpp:4-10); nothing throws across the C-ABI itself.
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

#include "abi.h"

namespace cr {

class Error : public std::runtime_error {
public:
    Error(cr_status s, const std::string& msg) : std::runtime_error(msg), status(s) {}
    cr_status status;
};

inline void check(cr_context* ctx, cr_status s, const char* what)
{
    if (s != CR_OK)
        throw Error(s, std::string(what) + ": " + cr_status_string(s) + ": " + (ctx ? cr_last_error(ctx) : ""));
}

// Owns the device, its stream and every resident buffer (the role of
// main.cpp's globals and InitializeVulkan, main.cpp:765-910).
class Device {
public:
    explicit Device(int device = 0)
    {
        check(nullptr, cr_create(device, &ctx_), "cr_create");
    }
    ~Device() { cr_destroy(ctx_); }
    Device(const Device&) = delete;
    Device& operator=(const Device&) = delete;
    cr_context* get() const { return ctx_; }

    // Grid3D<float>::Load + SetCloudProperties + UpdateCloudData (main.cpp:166-190, :266-335)
    void LoadCloudFile(const std::string& path, int storage = CR_DTYPE_F32)
    {
        check(ctx_, cr_grid_load_xyz(ctx_, path.c_str(), storage), "cr_grid_load_xyz");
    }
    void GenerateCloud(const cr_fbm_params& p, int storage = CR_DTYPE_F32)
    {
        check(ctx_, cr_grid_generate_fbm(ctx_, &p, storage), "cr_grid_generate_fbm");
    }
    cr_cloud_properties CloudProperties() const
    {
        cr_cloud_properties c;
        check(ctx_, cr_get_cloud(ctx_, &c), "cr_get_cloud");
        return c;
    }
    void Synchronize() { check(ctx_, cr_synchronize(ctx_), "cr_synchronize"); }

private:
    cr_context* ctx_ = nullptr;
};

// RenderTechnique.h:15-66: QueueUpdate* record bindings, UpdateDescriptorSets
// commits them, RecordDrawCommands issues the dispatches.
class RenderTechnique {
public:
    RenderTechnique(Device* device, cr_push_constants* pushConstants) : dev_(device), push_(pushConstants) {}
    virtual ~RenderTechnique() = default;

    void QueueUpdateCloudData(const cr_cloud_properties& c, unsigned int /*imageIdx*/ = 0) { cloud_ = c; hasCloud_ = true; }
    void QueueUpdateCameraProperties(const cr_camera_properties& c, unsigned int = 0) { cam_ = c; hasCam_ = true; }
    void QueueUpdateParameters(const cr_parameters& p, unsigned int = 0) { params_ = p; hasParams_ = true; }
    void QueueUpdateShadowVolume(const cr_shadow_volume_properties& s, unsigned int = 0) { sv_ = s; hasSv_ = true; }
    // The samplers bind the context's grid and shadow volume; nothing to queue.
    void QueueUpdateCloudDataSampler(unsigned int = 0) {}
    void QueueUpdateShadowVolumeSampler(unsigned int = 0) {}

    void UpdateDescriptorSets()
    {
        cr_context* c = dev_->get();
        if (hasCloud_) check(c, cr_set_cloud(c, &cloud_), "cr_set_cloud");
        if (hasCam_) check(c, cr_set_camera(c, &cam_), "cr_set_camera");
        if (hasParams_) check(c, cr_set_parameters(c, &params_), "cr_set_parameters");
        if (hasSv_) check(c, cr_set_shadow_volume(c, &sv_), "cr_set_shadow_volume");
        hasCloud_ = hasCam_ = hasParams_ = hasSv_ = false;
    }
    virtual uint32_t GetRequiredSetCount() const { return 1; }
    virtual void RecordDrawCommands(unsigned int imageIndex = 0) = 0;

protected:
    Device* dev_;
    cr_push_constants* push_;
    cr_cloud_properties cloud_{};
    cr_camera_properties cam_{};
    cr_parameters params_{};
    cr_shadow_volume_properties sv_{};
    bool hasCloud_ = false, hasCam_ = false, hasParams_ = false, hasSv_ = false;
};

// RenderTechniqueSV: shaders/ShadowVolume.comp (RenderTechniqueSV.cpp:102-113).
class RenderTechniqueSV : public RenderTechnique {
public:
    RenderTechniqueSV(Device* d, const cr_shadow_volume_properties* svp, cr_push_constants* pc) : RenderTechnique(d, pc)
    {
        if (svp) QueueUpdateShadowVolume(*svp);
    }
    void RecordDrawCommands(unsigned int = 0) override
    {
        UpdateDescriptorSets();
        check(dev_->get(), cr_sv_build(dev_->get()), "cr_sv_build");
    }
};

// RenderTechniquePT: shaders/PathTracer.comp (RenderTechniquePT.cpp:129-168).
// One call renders the frame in *pushConstants; RecordFrames renders many in
// one persistent launch.
class RenderTechniquePT : public RenderTechnique {
public:
    RenderTechniquePT(Device* d, const cr_camera_properties* cam, cr_push_constants* pc) : RenderTechnique(d, pc)
    {
        if (cam) QueueUpdateCameraProperties(*cam);
    }
    void SetExtensions(const cr_render_ext& e) { check(dev_->get(), cr_set_render_ext(dev_->get(), &e), "cr_set_render_ext"); }
    void RecordDrawCommands(unsigned int = 0) override
    {
        UpdateDescriptorSets();
        check(dev_->get(), cr_pt_render(dev_->get(), push_, 1), "cr_pt_render");
    }
    void RecordFrames(const std::vector<cr_push_constants>& frames)
    {
        UpdateDescriptorSets();
        check(dev_->get(), cr_pt_render(dev_->get(), frames.data(), (uint32_t)frames.size()), "cr_pt_render");
    }
    void Clear() { check(dev_->get(), cr_pt_clear(dev_->get()), "cr_pt_clear"); }
    std::vector<float> ReadImage(int width, int height)
    {
        std::vector<float> rgba((size_t)width * height * 4);
        check(dev_->get(), cr_pt_download(dev_->get(), rgba.data()), "cr_pt_download");
        return rgba;
    }
};

// RenderTechniquePPB (RenderTechniquePPB.h:7-129): beam tracing, LBVH build
// and gather, every stage dispatched (the reference dispatches only the first
// two, RenderTechniquePPB.cpp:436-519).
class RenderTechniquePPB : public RenderTechnique {
public:
    RenderTechniquePPB(Device* d, cr_push_constants* pc, const cr_camera_properties* cam, float initialRadius)
        : RenderTechnique(d, pc), initialRadius_(initialRadius), radius_(initialRadius)
    {
        if (cam) QueueUpdateCameraProperties(*cam);
        cr_beam_trace_params_default(&trace_);
        trace_.radius = initialRadius;
        cr_gather_ext_default(&gather_);
    }
    // RenderTechniquePPB.cpp:282-355: every beam, sort and tree buffer, up front
    void AllocateResources(uint32_t maxBeams = 1u << 22)
    {
        maxBeams_ = maxBeams;
        check(dev_->get(), cr_ppb_allocate(dev_->get(), maxBeams), "cr_ppb_allocate");
    }
    void FreeResources() { check(dev_->get(), cr_ppb_free(dev_->get()), "cr_ppb_free"); }
    void SetTraceParams(const cr_beam_trace_params& p) { trace_ = p; }
    void SetGatherExtensions(const cr_gather_ext& g) { gather_ = g; }
    // RenderTechniquePPB.cpp:530-546; RecordDrawCommands calls it with the frame's frameCount (:439).
    void UpdateRadius(unsigned int frameNumber)
    {
        radius_ = cr_ppb_update_radius(frameNumber, CR_PPB_BEAMS_PER_PASS, CR_PPB_ALPHA, initialRadius_, radius_);
    }
    float Radius() const { return radius_; }
    void RecordDrawCommands(unsigned int = 0) override
    {
        UpdateDescriptorSets();
        cr_context* c = dev_->get();
        uint32_t nb = 0, nd = 0;
        UpdateRadius(push_->frameCount);
        trace_.radius = radius_;
        trace_.seed = push_->seed;
        check(c, cr_ppb_trace_beams(c, &trace_, maxBeams_, &nb, &nd), "cr_ppb_trace_beams");
        check(c, cr_lbvh_build(c), "cr_lbvh_build");
        check(c, cr_ppb_gather(c, push_, 1, &gather_), "cr_ppb_gather");
    }

private:
    float initialRadius_;
    float radius_;
    uint32_t maxBeams_ = 1u << 22;
    cr_beam_trace_params trace_{};
    cr_gather_ext gather_{};
};

// RenderTechniquePPM (RenderTechniquePPM.h:8-67): per frame, UpdateRadius,
// photon tracing into the photon grid with the frame's seed, then the
// ray-marched estimate blended into the accumulation image
// (RenderTechniquePPM.cpp:229-300).
class RenderTechniquePPM : public RenderTechnique {
public:
    RenderTechniquePPM(Device* d, const cr_camera_properties* cam, const cr_photon_map_properties* pm,
                       cr_push_constants* pc, float initialRadius)
        : RenderTechnique(d, pc), initialRadius_(initialRadius)
    {
        if (cam) QueueUpdateCameraProperties(*cam);
        if (pm) UpdatePhotonMapProperties(*pm);
        pc->pmRadius = initialRadius;
        cr_ppm_params_default(&params_);
    }
    void UpdatePhotonMapProperties(const cr_photon_map_properties& pm)
    {
        check(dev_->get(), cr_set_photon_map(dev_->get(), &pm), "cr_set_photon_map");
    }
    void SetParams(const cr_ppm_params& p) { params_ = p; }
    // RenderTechniquePPM.cpp:302-312
    void UpdateRadius(unsigned int frameNumber)
    {
        push_->pmRadius = cr_ppm_update_radius(frameNumber, CR_PPM_ALPHA, initialRadius_, push_->pmRadius);
    }
    void RecordDrawCommands(unsigned int = 0) override
    {
        UpdateDescriptorSets();
        cr_context* c = dev_->get();
        UpdateRadius(push_->frameCount);
        check(c, cr_ppm_trace(c, push_->seed, &params_, nullptr), "cr_ppm_trace");
        check(c, cr_ppm_estimate(c, push_, 1, &params_), "cr_ppm_estimate");
    }

private:
    float initialRadius_;
    cr_ppm_params params_{};
};

} // namespace cr
