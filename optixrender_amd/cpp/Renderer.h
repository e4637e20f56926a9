// Renderer with the reference's interface (core/include/Renderer.h:10-25,
// 139-143): Renderer(const Scene*), Render(), Resize(vec2i),
// DownloadPixels(uint32_t[]), SetCamera(Camera) and the public fields
// denoiserOn, progressive and launchParams (numPixelSamples, maxBounce,
// filterType, frame.frameID, frame.size).  OptiX is replaced by the HIP
// library behind include/oxr.h; errors throw std::runtime_error.
#pragma once
#include <cstdint>

#include "CUDABuffer.h"
#include "Scene.h"

struct LaunchParams {
    int numPixelSamples = 1;
    int maxBounce = 25;
    int rrStart = 0;  // first bounce with Russian roulette (the reference: 0)
    struct {
        int frameID = 0;
        vec2i size;
    } frame;
    FilterType filterType = FilterType::Gaussian;
};

class Renderer {
public:
    Renderer(const Scene* scene);
    ~Renderer();
    Renderer(const Renderer&) = delete;
    Renderer& operator=(const Renderer&) = delete;

    void Render();
    void Resize(const vec2i& newSize);
    void DownloadPixels(uint32_t h_pixels[]);
    void SetCamera(const Camera& camera);

    // statistics of the last Render() (device time, rays, per-stage times)
    oxr_render_stats LastStats() const;
    float BvhBuildMs() const { return buildMs; }

    bool denoiserOn = false;
    // Loads the KPCNN weights (OXR_DENOISER_PARAMS raw floats, the layout
    // oxr_denoiser_load takes).  Render() calls it on the first denoised
    // frame with $OXR_DENOISER_WEIGHTS, or denoiser_kpcnn.f32 next to the
    // executable.
    void LoadDenoiser(const std::string& path);
    bool progressive = false;
    // Streamed readback for a viewer: with presentOn, Render() copies each
    // frame to page-locked memory on a copy stream (Renderer.cpp:787-792
    // asks for streams and double-buffering), DownloadPixels reads from there,
    // and PresentPixels() hands out the page-locked frame itself (valid
    // through the next Render()).
    bool presentOn = false;
    const uint32_t* PresentPixels();
    LaunchParams launchParams;

protected:
    const Scene* scene;
    uint32_t sceneId = 0;
    float buildMs = 0.0f;
    Camera lastSetCamera;
    bool cameraSet = false;
    bool denoiserLoaded = false;
};
