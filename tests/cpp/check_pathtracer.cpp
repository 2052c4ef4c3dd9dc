// Drives the C++ host mirror (include/nbounce.hpp) end to end through libnbounce.so, for the
// tests in tests/test_cpp_host.py (host mode, no GPU) and tests/test_gpu_cpp_host.py (render).
//
//   check_pathtracer host <scene.glb> <out_dir>
//       Scene::parse_gltf of the file twice (the second copy appended after the first), the
//       reference build (build_scene), and a camera buffer, written as raw arrays to out_dir
//       (vertices, indices, primitives, blas, tlas, instances, camera .bin); then the error
//       behaviour: a missing file throws NB_ERR_IO, a file that is not glTF NB_ERR_FORMAT, and the
//       scene is unchanged after either.
//   check_pathtracer render <scene.glb> <cube.dds> <width> <height> <frames> <ex> <ey> <ez>
//                           <tx> <ty> <tz> <out.raw>
//       SceneBuffers::from_scene, EnvMap::load, Pathtracer at resolution factor 1: one
//       dispatch(), a dispatch_frames() of frames - 2 and one more dispatch(); then with
//       max_sample_count lowered to the samples taken, one more dispatch() must change nothing.
//       Writes the RGBA32F image (height x width x 4) to out.raw, the BlitRenderer surface of it
//       at (width + 24) x (height + 8) to out.raw.surface (RGBA16F bits), and the image as PFM and
//       OpenEXR files to out.raw.pfm and out.raw.exr.
// Prints one line per check and exits nonzero when one fails.
#include "nbounce.hpp"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <functional>
#include <string>
#include <vector>

namespace {

int failures = 0;

void expect(bool ok, const char *what) {
    std::printf("%s %s\n", ok ? "ok" : "FAIL", what);
    if (!ok) ++failures;
}

template <typename T>
void write_raw(const std::string &path, const std::vector<T> &v) {
    std::ofstream f(path, std::ios::binary);
    f.write(reinterpret_cast<const char *>(v.data()), static_cast<std::streamsize>(v.size() * sizeof(T)));
    if (!f) throw std::runtime_error("cannot write " + path);
}

nb_status status_of(const std::function<void()> &fn) {
    try {
        fn();
    } catch (const nbounce::Error &e) {
        return e.status();
    }
    return NB_OK;
}

int host_mode(const std::string &glb, const std::string &dir) {
    nbounce::Scene scene;
    scene.parse_gltf(glb);
    const size_t nv = scene.vertices.size(), ni = scene.indices.size(), np = scene.primitives.size();
    scene.parse_gltf(glb);
    expect(scene.vertices.size() == 2 * nv && scene.indices.size() == 2 * ni && scene.primitives.size() == 2 * np,
           "parse_gltf appends a second copy");
    bool ranges = true;
    for (size_t k = 0; k < np; ++k) {
        const nb_primitive &a = scene.primitives[k], &b = scene.primitives[np + k];
        ranges = ranges && b.index_start == a.index_start + ni && b.index_end == a.index_end + ni;
    }
    for (size_t k = 0; k < ni; ++k) ranges = ranges && scene.indices[ni + k] == scene.indices[k] + nv;
    expect(ranges, "appended indices and index ranges continue after the first copy");
    write_raw(dir + "/vertices.bin", scene.vertices);
    write_raw(dir + "/primitives.bin", scene.primitives);
    const nbounce::SceneArrays a = nbounce::build_scene(scene);
    write_raw(dir + "/indices.bin", scene.indices);   // permuted by the build
    write_raw(dir + "/blas.bin", a.blas);
    write_raw(dir + "/tlas.bin", a.tlas);
    write_raw(dir + "/instances.bin", a.instances);
    const float eye[3] = {0.0f, 1.0f, 0.9f}, target[3] = {0.0f, 0.0f, -0.5f};
    const nb_camera cam = nbounce::CameraController(eye, target, 1.5f).data();
    write_raw(dir + "/camera.bin", std::vector<nb_camera>{cam});

    const size_t before = scene.vertices.size();
    expect(status_of([&] { scene.parse_gltf(dir + "/no_such_file.glb"); }) == NB_ERR_IO, "missing file: NB_ERR_IO");
    expect(status_of([&] { scene.parse_gltf(dir + "/camera.bin"); }) == NB_ERR_FORMAT, "not glTF: NB_ERR_FORMAT");
    expect(scene.vertices.size() == before, "scene unchanged after rejected files");
    return failures == 0 ? 0 : 1;
}

int render_mode(char **argv) {
    const std::string glb = argv[2], dds = argv[3];
    const uint32_t w = static_cast<uint32_t>(std::atoi(argv[4])), h = static_cast<uint32_t>(std::atoi(argv[5]));
    const uint32_t frames = static_cast<uint32_t>(std::atoi(argv[6]));
    const float eye[3] = {std::strtof(argv[7], nullptr), std::strtof(argv[8], nullptr), std::strtof(argv[9], nullptr)};
    const float target[3] = {std::strtof(argv[10], nullptr), std::strtof(argv[11], nullptr),
                             std::strtof(argv[12], nullptr)};
    const std::string out = argv[13];
    if (frames < 2) throw std::runtime_error("frames must be at least 2");

    nbounce::Context ctx(0);
    nbounce::Scene scene;
    scene.parse_gltf(glb);
    const nbounce::SceneBuffers sb = nbounce::SceneBuffers::from_scene(ctx, scene);
    const nbounce::EnvMap env = nbounce::EnvMap::load(ctx, dds);
    const nbounce::CameraController cam(eye, target, static_cast<float>(w) / static_cast<float>(h));
    nbounce::Pathtracer pt(ctx, sb, cam, env, w, h, 1.0f);
    expect(pt.width() == w && pt.height() == h, "output texture of the window size at resolution factor 1");
    pt.dispatch();
    expect(pt.dispatch_frames(frames - 2) == frames - 2, "dispatch_frames adds its frames");
    pt.dispatch();
    expect(pt.sample_count() == frames, "sample_count counts every frame");
    const std::vector<float> img = pt.output_texture();
    pt.max_sample_count = frames;
    pt.dispatch();
    expect(pt.sample_count() == frames && pt.output_texture() == img, "no dispatch beyond max_sample_count");
    write_raw(out, img);
    // the display and the image files of the same image, beside it
    const nbounce::BlitRenderer blit(ctx, w + 24, h + 8);
    blit.render(pt);
    write_raw(out + ".surface", blit.surface());
    nbounce::write_pfm(out + ".pfm", img, w, h);
    nbounce::write_exr(out + ".exr", img, w, h);
    expect(true, "display and image files written");
    return failures == 0 ? 0 : 1;
}

}  // namespace

int main(int argc, char **argv) {
    try {
        if (argc == 4 && std::strcmp(argv[1], "host") == 0) return host_mode(argv[2], argv[3]);
        if (argc == 14 && std::strcmp(argv[1], "render") == 0) return render_mode(argv);
        std::fprintf(stderr, "usage: check_pathtracer host <scene.glb> <out_dir>\n"
                             "       check_pathtracer render <scene.glb> <cube.dds> <w> <h> <frames> "
                             "<ex> <ey> <ez> <tx> <ty> <tz> <out.raw>\n");
        return 2;
    } catch (const std::exception &e) {
        std::printf("FAIL exception: %s\n", e.what());
        return 1;
    }
}
