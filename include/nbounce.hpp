/* nbounce.hpp — C++17 host-side mirror of the reference's `pathtracing` renderer surface over
 * the C ABI of nbounce.h (libnbounce.so).  Header only.
 *
 * Names, arguments and behaviour follow the reference's Rust types:
 *   Context         WGPUContext (src/common/wgpu_context.rs:13-79): one HIP device and stream
 *   Scene           Scene (src/pathtracing/scene.rs:100-217) with parse_gltf
 *   build_scene     SceneBuffers::from_scene without the upload (scene.rs:280-307)
 *   SceneBuffers    SceneBuffers (scene.rs:271-307), device-resident
 *   CameraController  camera state and CameraBuffer (src/common/camera.rs:14-52, 121-142)
 *   EnvMap          EnvMap (src/pathtracing/envmap.rs:5-26), load() from a cube DDS file
 *   Pathtracer      Pathtracer (src/pathtracing/pathtracer.rs:15-239): the Sobol table, the
 *                   output texture rounded down to multiples of 8, dispatch() adding one
 *                   sample per pixel with weight 1/sample up to 1024 samples, invalidate(),
 *                   update(), resize()
 * The window is replaced by an explicit output size.  Where the reference panics (wgpu
 * errors) or returns io::Error / MeshError, these throw nbounce::Error carrying the nb_status and
 * nb_last_error()'s text (MeshError's texts for rejected glTF files).  Objects own their handles
 * (move-only).  A Context is used from one thread at a time; rendering is asynchronous on it, and
 * output_texture() waits. */
#ifndef NBOUNCE_HPP
#define NBOUNCE_HPP

#include "nbounce.h"

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace nbounce {

class Error : public std::runtime_error {
public:
    Error(nb_status status, const std::string &message)
        : std::runtime_error("nb_status " + std::to_string(static_cast<int>(status)) + ": " + message),
          status_(status) {}
    nb_status status() const { return status_; }

private:
    nb_status status_;
};

inline void check(nb_status s) {
    if (s != NB_OK) throw Error(s, nb_last_error());
}

namespace detail {
template <typename T, nb_status (*Destroy)(T *)>
struct Deleter {
    void operator()(T *p) const { Destroy(p); }
};
template <typename T, nb_status (*Destroy)(T *)>
using Handle = std::unique_ptr<T, Deleter<T, Destroy>>;
}  // namespace detail

// ---- Context ------------------------------------------------------------------------------
class Context {
public:
    explicit Context(int device = 0) {
        nb_context *c = nullptr;
        check(nb_context_create(device, &c));
        h_.reset(c);
    }
    nb_context *get() const { return h_.get(); }
    void synchronize() const { check(nb_context_synchronize(get())); }
    nb_counters counters() const {
        nb_counters c{};
        check(nb_counters_get(get(), &c));
        return c;
    }
    void enable_counters(bool on) const { check(nb_counters_enable(get(), on ? 1u : 0u)); }
    void reset_counters() const { check(nb_counters_reset(get())); }
    // Native option (no counterpart in the reference): Russian roulette from this bounce on
    // (0: the reference's rule, pathtracing.wgsl:227-236).
    void set_render_options(uint32_t roulette_first_bounce) const {
        nb_render_options o{roulette_first_bounce};
        check(nb_render_options_set(get(), &o));
    }

private:
    detail::Handle<nb_context, nb_context_destroy> h_;
};

// ---- Scene (scene.rs:100-217) ------------------------------------------------------------
struct Scene {
    std::vector<nb_vertex> vertices;
    std::vector<uint32_t> indices;
    std::vector<nb_primitive> primitives;

    // Scene::parse_gltf (scene.rs:108-217): appends the file's mesh primitives, one primitive
    // per mesh primitive of each mesh node, its indices and index ranges continuing after what
    // the scene already holds (scene.rs:135-136).  A rejected file throws (NB_ERR_FORMAT with the
    // MeshError text) and leaves the scene unchanged.  share_geometry: mesh primitives naming the
    // same accessors share one copy of the geometry (use with build_scene(…, true)).
    void parse_gltf(const std::string &path, bool share_geometry = false) {
        nb_host_scene *raw = nullptr;
        check(nb_host_scene_create(&raw));
        detail::Handle<nb_host_scene, nb_host_scene_destroy> hs(raw);
        check(nb_host_scene_parse_gltf(raw, path.c_str(), share_geometry ? 1u : 0u));
        uint32_t nv = 0, ni = 0, np = 0;
        check(nb_host_scene_counts(raw, &nv, &ni, &np));
        std::vector<nb_vertex> v(nv);
        std::vector<uint32_t> i(ni);
        std::vector<nb_primitive> p(np);
        check(nb_host_scene_copy(raw, v.data(), i.data(), p.data()));
        const uint32_t v0 = static_cast<uint32_t>(vertices.size()), i0 = static_cast<uint32_t>(indices.size());
        for (uint32_t &x : i) x += v0;
        for (nb_primitive &x : p) { x.index_start += i0; x.index_end += i0; }
        vertices.insert(vertices.end(), v.begin(), v.end());
        indices.insert(indices.end(), i.begin(), i.end());
        primitives.insert(primitives.end(), p.begin(), p.end());
    }
};

// The reference-layout arrays SceneBuffers::from_scene uploads (scene.rs:280-307).
struct SceneArrays {
    std::vector<nb_bvh_node> blas, tlas;
    std::vector<nb_instance> instances;
};

// SceneBuffers::from_scene without the upload: builds the BLASes, instance bounds and TLAS and
// permutes scene.indices in place, bit-exact with the reference builder.  share_blas: build each
// index range once (NB_BUILD_SHARE_BLAS) instead of once per primitive as the reference does.
inline SceneArrays build_scene(Scene &scene, bool share_blas = false) {
    const uint32_t flags = share_blas ? NB_BUILD_SHARE_BLAS : 0u;
    const uint32_t np = static_cast<uint32_t>(scene.primitives.size());
    uint32_t bcap = 0, tcap = 0;
    check(nb_build_scene_capacity(static_cast<uint32_t>(scene.indices.size()), scene.primitives.data(), np, flags,
                                  &bcap, &tcap));
    SceneArrays a;
    a.blas.resize(bcap > 0 ? bcap : 1);
    a.tlas.resize(tcap > 0 ? tcap : 1);
    a.instances.resize(np);
    uint32_t nb = 0, nt = 0;
    check(nb_build_scene(scene.vertices.data(), static_cast<uint32_t>(scene.vertices.size()), scene.indices.data(),
                         static_cast<uint32_t>(scene.indices.size()), scene.primitives.data(), np, flags,
                         a.blas.data(), static_cast<uint32_t>(a.blas.size()), &nb, a.tlas.data(),
                         static_cast<uint32_t>(a.tlas.size()), &nt, a.instances.data()));
    a.blas.resize(nb);
    a.tlas.resize(nt);
    return a;
}

// ---- SceneBuffers (scene.rs:271-307) -------------------------------------------------------
class SceneBuffers {
public:
    SceneBuffers(const Context &ctx, const SceneArrays &a, const std::vector<nb_vertex> &vertices,
                 const std::vector<uint32_t> &indices) {
        nb_scene *s = nullptr;
        check(nb_scene_create(ctx.get(), a.blas.data(), static_cast<uint32_t>(a.blas.size()), a.tlas.data(),
                              static_cast<uint32_t>(a.tlas.size()), a.instances.data(),
                              static_cast<uint32_t>(a.instances.size()), vertices.data(),
                              static_cast<uint32_t>(vertices.size()), indices.data(),
                              static_cast<uint32_t>(indices.size()), &s));
        h_.reset(s);
    }
    // SceneBuffers::from_scene (scene.rs:280): build on the host (permuting scene.indices as the
    // reference does), then upload.
    static SceneBuffers from_scene(const Context &ctx, Scene &scene, bool share_blas = false) {
        const SceneArrays a = build_scene(scene, share_blas);
        return SceneBuffers(ctx, a, scene.vertices, scene.indices);
    }
    nb_scene *get() const { return h_.get(); }
    uint64_t device_bytes() const {
        uint64_t b = 0;
        check(nb_scene_device_bytes(get(), &b));
        return b;
    }

private:
    detail::Handle<nb_scene, nb_scene_destroy> h_;
};

// ---- CameraController (camera.rs:14-52, 121-142; no input handling) ------------------------
struct CameraController {
    float world_position[3] = {5.0f, 0.0f, 0.0f};
    float target[3] = {0.0f, 0.0f, 0.0f};
    float up[3] = {0.0f, 1.0f, 0.0f};
    float fov = 1.0471975511965976f;   // pi / 3
    float aspect_ratio = 1.0f;
    float near = 0.1f;

    CameraController() = default;
    CameraController(const float eye[3], const float look_at[3], float aspect) : aspect_ratio(aspect) {
        for (int k = 0; k < 3; ++k) { world_position[k] = eye[k]; target[k] = look_at[k]; }
    }
    void resize(float aspect) { aspect_ratio = aspect; }
    // CameraBuffer (camera.rs:7-12): world_to_clip and its inverse.
    nb_camera data() const {
        nb_camera c{};
        check(nb_camera_look_at(world_position, target, up, fov, aspect_ratio, near, &c));
        return c;
    }
};

// ---- EnvMap (envmap.rs:5-26) ---------------------------------------------------------------
class EnvMap {
public:
    // Six RGBA32F faces of face_size^2 texels, +X, -X, +Y, -Y, +Z, -Z.
    EnvMap(const Context &ctx, const std::vector<float> &faces, uint32_t face_size) {
        if (faces.size() != static_cast<size_t>(6) * face_size * face_size * 4)
            throw Error(NB_ERR_INVALID_ARGUMENT, "cube faces: expected 6 x face_size^2 RGBA32F texels");
        nb_envmap *e = nullptr;
        check(nb_envmap_create_cube(ctx.get(), faces.data(), face_size, &e));
        h_.reset(e);
    }
    // EnvMap::load (envmap.rs:12): a BC6H_UF16 (or RGBA16F / RGBA32F) cube DDS file.
    static EnvMap load(const Context &ctx, const std::string &path) {
        nb_envmap *e = nullptr;
        check(nb_envmap_load(ctx.get(), path.c_str(), &e));
        return EnvMap(e);
    }
    // Native option (no counterpart in the reference): an equirectangular RGBA32F map, row 0 at
    // the top (+y).
    static EnvMap latlong(const Context &ctx, const std::vector<float> &texels, uint32_t width, uint32_t height) {
        if (texels.size() != static_cast<size_t>(width) * height * 4)
            throw Error(NB_ERR_INVALID_ARGUMENT, "lat-long map: expected width x height RGBA32F texels");
        nb_envmap *e = nullptr;
        check(nb_envmap_create_latlong(ctx.get(), texels.data(), width, height, &e));
        return EnvMap(e);
    }
    nb_envmap *get() const { return h_.get(); }

private:
    explicit EnvMap(nb_envmap *e) { h_.reset(e); }
    detail::Handle<nb_envmap, nb_envmap_destroy> h_;
};

// ---- Pathtracer (pathtracer.rs:15-239) -----------------------------------------------------
class Pathtracer {
public:
    static constexpr uint32_t COMPUTE_SIZE = 8;     // pathtracer.rs:48
    static constexpr uint32_t LDS_PER_BOUNCE = 2;   // pathtracer.rs:49: Sobol sets per bounce

    // Pathtracer::new (pathtracer.rs:51-111): the Sobol table of max_sample_count rows and
    // bounces * 2 + 1 sets (the default 8 bounces: 17, the kernel's row stride) and the output
    // texture of (window * resolution_factor) rounded down to multiples of 8 (:192-202).  The
    // context, scene and environment are held by reference and must outlive the tracer; the
    // camera's buffer is copied here and by update().
    Pathtracer(const Context &ctx, const SceneBuffers &scene, const CameraController &camera, const EnvMap &envmap,
               uint32_t window_width, uint32_t window_height, float resolution_factor = 0.3f)
        : resolution_factor(resolution_factor), ctx_(&ctx), scene_(&scene), envmap_(&envmap),
          camera_(camera.data()), window_{window_width, window_height} {
        const uint32_t dims = globals.bounces * LDS_PER_BOUNCE + 1u;
        std::vector<float> table(static_cast<size_t>(max_sample_count) * dims * 4);
        check(nb_sobol_generate(max_sample_count, dims, 0u, table.data()));
        check(nb_sobol_set(ctx.get(), table.data(), max_sample_count, dims, dims));
        create_output_texture();
    }

    // Replaces the Sobol table (for example the exact table a reference run used); stride: the
    // float4 entries per sample row the kernel steps by (17 reproduces the reference).
    void set_sobol_table(const std::vector<float> &table, uint32_t rows, uint32_t dims, uint32_t stride) {
        check(nb_sobol_set(ctx_->get(), table.data(), rows, dims, stride));
    }

    // The reference's public fields (pathtracer.rs:21-23; Globals::default, :35-44).
    nb_globals globals{0u, 0.0f, 8u, 4.0f};
    float resolution_factor;
    uint32_t max_sample_count = 1024;

    uint32_t width() const { return width_; }
    uint32_t height() const { return height_; }
    uint32_t sample_count() const { return globals.sample; }   // pathtracer.rs:217
    void invalidate() { globals.sample = 0; }                   // pathtracer.rs:221

    // pathtracer.rs:212: new camera and environment, accumulation restarts.
    void update(const CameraController &camera, const EnvMap &envmap) {
        camera_ = camera.data();
        envmap_ = &envmap;
        invalidate();
    }
    // pathtracer.rs:208: a new output texture for the new window size.
    void resize(uint32_t window_width, uint32_t window_height) {
        window_[0] = window_width;
        window_[1] = window_height;
        create_output_texture();
    }

    // pathtracer.rs:225-239: one more sample per pixel, mixed in with weight 1/sample; nothing
    // once max_sample_count samples are in.
    void dispatch() {
        if (globals.sample >= max_sample_count) return;
        globals.sample += 1;
        globals.weight = 1.0f / static_cast<float>(globals.sample);
        check(nb_render(ctx_->get(), scene_->get(), target_.get(), &camera_, &globals, envmap_->get()));
    }
    // n dispatch() calls in one launch (the same image bit for bit); returns the frames added.
    uint32_t dispatch_frames(uint32_t n) {
        if (globals.sample >= max_sample_count) return 0;
        if (n > max_sample_count - globals.sample) n = max_sample_count - globals.sample;
        if (n == 0) return 0;
        nb_globals first = globals;
        first.sample = globals.sample + 1;
        check(nb_render_frames(ctx_->get(), scene_->get(), target_.get(), &camera_, &first, n, envmap_->get()));
        globals.sample += n;
        globals.weight = 1.0f / static_cast<float>(globals.sample);
        return n;
    }

    // The RGBA32F accumulation image, row-major, height x width x 4 (waits for the GPU).
    std::vector<float> output_texture() const {
        std::vector<float> out(static_cast<size_t>(width_) * height_ * 4);
        check(nb_target_read(target_.get(), out.data()));
        return out;
    }
    nb_target *target() const { return target_.get(); }

private:
    void create_output_texture() {
        // (window as f32 * resolution_factor) as u32, then down to a multiple of COMPUTE_SIZE
        const uint32_t w = static_cast<uint32_t>(static_cast<float>(window_[0]) * resolution_factor);
        const uint32_t h = static_cast<uint32_t>(static_cast<float>(window_[1]) * resolution_factor);
        const uint32_t tw = w / COMPUTE_SIZE * COMPUTE_SIZE, th = h / COMPUTE_SIZE * COMPUTE_SIZE;
        if (tw == 0 || th == 0) throw Error(NB_ERR_INVALID_ARGUMENT, "output size is empty");
        nb_target *t = nullptr;
        check(nb_target_create(ctx_->get(), tw, th, &t));
        target_.reset(t);
        width_ = tw;
        height_ = th;
    }

    const Context *ctx_;
    const SceneBuffers *scene_;
    const EnvMap *envmap_;
    nb_camera camera_;
    uint32_t window_[2];
    uint32_t width_ = 0, height_ = 0;
    detail::Handle<nb_target, nb_target_destroy> target_;
};

// ---- BlitRenderer (src/pathtracing/blit_renderer.rs, blit.wgsl:27-38) ----------------------
// Draws a Pathtracer's output texture onto an RGBA16F surface of its own size (the surface format
// of wgpu_context.rs:62): bilinear, clamp-to-edge, surface row 0 at the top.
class BlitRenderer {
public:
    BlitRenderer(const Context &ctx, uint32_t width, uint32_t height) : width_(width), height_(height) {
        nb_display *d = nullptr;
        check(nb_display_create(ctx.get(), width, height, &d));
        h_.reset(d);
    }
    // Enqueues the blit of the tracer's output texture (asynchronous on the context stream).
    void render(const Pathtracer &pathtracer) const { check(nb_display_blit(h_.get(), pathtracer.target())); }
    // The surface as height x width x 4 half-float bit patterns (waits for the blit).
    std::vector<uint16_t> surface() const {
        std::vector<uint16_t> out(static_cast<size_t>(width_) * height_ * 4);
        check(nb_display_read(h_.get(), out.data()));
        return out;
    }

private:
    uint32_t width_, height_;
    detail::Handle<nb_display, nb_display_destroy> h_;
};

// Image files of an output_texture() image (target row order): Portable Float Map (RGB) and
// OpenEXR (RGBA, FLOAT, or HALF when half is set).
inline void write_pfm(const std::string &path, const std::vector<float> &rgba, uint32_t width, uint32_t height) {
    if (rgba.size() != static_cast<size_t>(width) * height * 4) throw Error(NB_ERR_INVALID_ARGUMENT, "image size");
    check(nb_image_write_pfm(path.c_str(), rgba.data(), width, height));
}
inline void write_exr(const std::string &path, const std::vector<float> &rgba, uint32_t width, uint32_t height,
                      bool half = false) {
    if (rgba.size() != static_cast<size_t>(width) * height * 4) throw Error(NB_ERR_INVALID_ARGUMENT, "image size");
    check(nb_image_write_exr(path.c_str(), rgba.data(), width, height, half ? NB_EXR_HALF : 0u));
}

}  // namespace nbounce

#endif  // NBOUNCE_HPP
