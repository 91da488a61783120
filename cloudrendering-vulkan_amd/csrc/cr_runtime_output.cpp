// Progressive accumulation and image output: the reference's render loop and
// swapchain blit, headless.
//
//   RenderLoop (main.cpp:714-731): per frame seed = std::rand() (after
//   std::srand(0), main.cpp:949), draw, frameCount++.
//   UpdateShadowVolume (main.cpp:226-264), which every parameter, camera or
//   cloud change goes through (main.cpp:560-590): frameCount = 1, the image is
//   not cleared, so the next frame blends half of the old image.
//   RenderTechniquePT::RecordDrawCommands (RenderTechniquePT.cpp:143-167): blit
//   of the accumulation image to the swapchain -- here a PFM file.
#include <cstdio>
#include <vector>

#include "cr_runtime.h"

using namespace cr;

#define CR_CHECK_CTX(ctx) \
    do { if (!(ctx)) return CR_ERR_INVALID_ARGUMENT; } while (0)
#define CR_HIP(ctx, call, what) \
    do { hipError_t _e = (call); if (_e != hipSuccess) return hip_fail((ctx), _e, (what)); } while (0)

extern "C" {

cr_status cr_pt_progressive_reset(cr_context* ctx, uint32_t srand_seed, uint32_t first_frame_count,
                                  int clear_on_restart)
{
    CR_CHECK_CTX(ctx);
    ctx->prog.randState = srand_seed;
    ctx->prog.firstFrameCount = first_frame_count;
    ctx->prog.clearOnRestart = clear_on_restart != 0;
    ctx->prog.started = false;
    ctx->prog.restarts = 0;
    ctx->prog.framesRendered = 0;
    ctx->prog.frameCount = first_frame_count;
    return CR_OK;
}

cr_status cr_pt_progressive(cr_context* ctx, uint32_t n_frames, cr_progressive_state* state_out)
{
    CR_CHECK_CTX(ctx);
    Progressive& P = ctx->prog;
    if (!P.started || P.sceneVersion != ctx->sceneVersion) {
        if (P.started) P.restarts++;
        P.frameCount = P.firstFrameCount;
        P.framesRendered = 0;
        if (P.clearOnRestart && ctx->image) {
            cr_status st = cr_pt_clear(ctx);
            if (st != CR_OK) return st;
        }
        P.started = true;
        P.sceneVersion = ctx->sceneVersion;
    }
    if (n_frames) {
        std::vector<cr_push_constants> fr(n_frames);
        uint32_t rs = P.randState;
        for (uint32_t i = 0; i < n_frames; i++) {
            rs = rs * 214013u + 2531011u;                  // MSVC std::rand
            fr[i].time = 0.0;
            fr[i].seed = (int32_t)((rs >> 16) & 0x7fffu);
            fr[i].frameCount = P.frameCount + i;
            fr[i].pmRadius = 0.0f;
            fr[i].currentBuffer = 0;
        }
        cr_status st = cr_pt_render(ctx, fr.data(), n_frames);
        if (st != CR_OK) return st;
        P.randState = rs;
        P.frameCount += n_frames;
        P.framesRendered += n_frames;
    }
    if (state_out) {
        state_out->frameCount = P.frameCount;
        state_out->framesRendered = P.framesRendered;
        state_out->restarts = P.restarts;
        state_out->randState = P.randState;
    }
    return CR_OK;
}

// Portable float map: "PF\n<W> <H>\n-1.0\n" (negative scale: little-endian),
// then RGB float32 rows from the bottom row up.
cr_status cr_pt_write_pfm(cr_context* ctx, const char* path)
{
    CR_CHECK_CTX(ctx);
    if (!path) return fail(ctx, CR_ERR_INVALID_ARGUMENT, "null path");
    if (!ctx->image) return fail(ctx, CR_ERR_STATE, "no camera set");
    const int W = ctx->W, H = ctx->H;
    std::vector<float> rgba((size_t)W * H * 4), rgb((size_t)W * H * 3);
    cr_status st = cr_pt_download(ctx, rgba.data());
    if (st != CR_OK) return st;
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++)
            for (int c = 0; c < 3; c++)
                rgb[3 * ((size_t)(H - 1 - y) * W + x) + c] = rgba[4 * ((size_t)y * W + x) + c];
    FILE* f = fopen(path, "wb");
    if (!f) return fail(ctx, CR_ERR_IO, std::string("cannot create ") + path);
    bool ok = fprintf(f, "PF\n%d %d\n-1.0\n", W, H) > 0 && fwrite(rgb.data(), 4, rgb.size(), f) == rgb.size();
    ok = (fclose(f) == 0) && ok;
    return ok ? CR_OK : fail(ctx, CR_ERR_IO, "write failed");
}

} // extern "C"
