// pbrt-v4 scene subset for the "PBRT Scene" node: what merian's PBRTScene loads for the pt path
// (src/merian-shaders/scene/pbrt_scene.cpp), parsed into this library's scene description.
//
// Supported: LookAt / Camera "perspective" (fov on the shorter image axis, :970-1000) / Film
// resolution; transforms (Translate, Scale, Rotate, Transform, ConcatTransform, Identity) and
// AttributeBegin/End; Material, MakeNamedMaterial, NamedMaterial with diffuse, coateddiffuse,
// conductor (reflectance or eta/k), dielectric and mix mapped onto OpenPBR (:365-487, roughness
// remap :332-364); AreaLightSource "diffuse" (emission); LightSource "infinite" with "rgb L"
// or a PFM image (equal-area octahedral, clamp to edge, :924-968); Shape "trianglemesh" with
// P, indices, N, uv (normals computed by area weighting when absent, :55-70, :533-575),
// "loopsubdiv" (unsubdivided), "sphere" and "disk" (tessellated, :692-765), "plymesh" (ascii or
// binary PLY, optionally .gz, :577-690); Include / Import. A file without a camera gets the
// default camera framing the scene's bounds (:994-1003). Geometry is visible from both sides unless it is a two-sided emitter
// (:802-815); ReverseOrientation xor a mirroring transform flips facing (:817).
// Anything else is skipped with a warning.
#pragma once

#include <string>
#include <vector>

#include "../../include/merian_amd.h"

namespace mrt::pbrt {

// The scene's camera: the file's, or, for a file without one, a default camera framing the
// scene's bounds (pbrt_scene.cpp:994-1003, camera.cpp:373-423).
struct Camera {
    bool present = false;
    float eye[3], forward[3], up[3];
    float vfov = 0.f; // radians, vertical
    uint32_t width = 1920, height = 1080;
    bool has_resolution = true; // false: the default camera, the image size is the GBuffer node's
    float aspect = 0.f;         // the default camera's fixed aspect ratio (0: that of the image)
};

struct Scene {
    std::vector<mrt_vertex> vertices;
    std::vector<uint32_t> indices;
    std::vector<mrt_geometry> geometries;
    std::vector<mrt_material> materials;
    std::vector<float> env_rgba; // RGBA32F texels
    uint32_t env_width = 0, env_height = 0;
    float env_scale = 1.f;
    float env_to_local[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    Camera camera;
    std::vector<std::string> warnings;

    // A scene description pointing into this scene's arrays.
    mrt_scene_desc desc() const;
};

// Throws std::runtime_error when the file cannot be read or parsed.
Scene load(const std::string& path);

// Reads a PFM image (PF / Pf) into RGBA32F, top row first.
bool read_pfm(const std::string& path, std::vector<float>& rgba, uint32_t& width, uint32_t& height);
// Writes RGB (of RGBA32F input, top row first) as a little-endian PF file.
bool write_pfm(const std::string& path, const float* rgba, uint32_t width, uint32_t height);

} // namespace mrt::pbrt
