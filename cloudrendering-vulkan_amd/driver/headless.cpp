// Headless driver: the reference's StartSimulation / RenderLoop
// (main.cpp:927-944, :714-736) without a window.  It builds the cloud (an
// .xyz file given on the command line, or the seeded fBm cumulus), binds the
// uniforms in the order main.cpp's UpdateCloudData / UpdateShadowVolume use,
// renders N frames with seeds from std::rand() after std::srand(0) and
// writes the accumulated image as a PFM (RGB float, bottom row first) through
// the library's progressive loop and image writer.
//
//   headless [frames] [width] [height] [cloud.xyz] [out.pfm]
#include <cstdio>
#include <cstdlib>
#include <string>

#include "../../include/cr/techniques.hpp"

int main(int argc, char** argv)
{
    const uint32_t nFrames = argc > 1 ? (uint32_t)atoi(argv[1]) : 16;
    const int W = argc > 2 ? atoi(argv[2]) : 256, H = argc > 3 ? atoi(argv[3]) : 256;
    const std::string cloudFile = argc > 4 ? argv[4] : "";
    const std::string out = argc > 5 ? argv[5] : "headless.pfm";
    try {
        cr::Device dev(0);
        cr_cloud_properties cloud;
        cr_cloud_default(&cloud);
        cloud.baseScaling = 1.0f;
        cloud.densityScaling = 64.0f;
        cr::check(dev.get(), cr_set_cloud(dev.get(), &cloud), "cr_set_cloud");
        if (!cloudFile.empty()) {
            dev.LoadCloudFile(cloudFile);
        } else {
            cr_fbm_params p;
            cr_fbm_params_default(&p);
            p.thresholdLo = 0.16f;
            p.thresholdHi = 0.36f;
            dev.GenerateCloud(p);
        }
        cloud = dev.CloudProperties();

        cr_camera_properties cam;
        cr_camera_default(&cam);
        cam.position[2] = -0.2f;
        cam.nearPlane = 0.05f;
        cr_camera_set_resolution(&cam, W, H);
        float fov = 90.0f;
        cr_camera_set_fov(&cam, &fov);
        cr_parameters params;
        cr_parameters_default(&params);
        cr_parameters_set_phase_g(&params, 0.85f);
        cr_shadow_volume_properties svp;
        cr_shadow_volume_default(&svp);
        svp.voxelAxisCount = 256;
        const float light[3] = {1.0f, -1.0f, 0.0f};
        cr_shadow_volume_set_light(&svp, light, cloud.bounds[0], cloud.bounds[1]);

        cr_push_constants push{};
        cr::RenderTechniqueSV sv(&dev, &svp, &push);
        sv.RecordDrawCommands();
        cr::RenderTechniquePT pt(&dev, &cam, &push);
        pt.QueueUpdateParameters(params);
        pt.UpdateDescriptorSets();   // commits the camera, which sizes the accumulation image
        pt.Clear();
        // the render loop: std::rand() seeds after std::srand(0); frameCount from 0 for the exact mean
        cr::check(dev.get(), cr_pt_progressive_reset(dev.get(), 0, 0, 1), "cr_pt_progressive_reset");
        cr_progressive_state ps;
        cr::check(dev.get(), cr_pt_progressive(dev.get(), nFrames, &ps), "cr_pt_progressive");
        cr::check(dev.get(), cr_pt_write_pfm(dev.get(), out.c_str()), "cr_pt_write_pfm");
        cr_pt_stats s;
        cr_pt_stats_get(dev.get(), &s);
        printf("%u frames %dx%d: %llu paths, %llu vertices, %.3f ms -> %s\n", nFrames, W, H,
               (unsigned long long)s.paths, (unsigned long long)s.vertices, s.totalKernelMs, out.c_str());
    } catch (const cr::Error& e) {
        fprintf(stderr, "error: %s\n", e.what());
        return 2;
    }
    return 0;
}
