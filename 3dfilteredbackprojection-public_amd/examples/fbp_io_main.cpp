// fbp_io_main.cpp -- the formats around the reference's pipeline in
// Mains/fbp_cone_fdk_main.cpp: scanner frames in through
// FBP3D::read_projections (:19), slices out through normalize_histogram +
// write_volume (:26-27), on the drop-in FBP3D.
//
//   fbp_io_main read <dir> <file_w> <file_h> <channels> <bytes> <w> <h> <n> <out.raw> [sino]
//       read_projections (or read_projections_as_sinograms with "sino") and
//       dump the projections (sinograms) as raw float32
//   fbp_io_main write <volume.raw> <X> <Y> <Z> <dir> [normalize|textures]
//       load a raw float32 (Z, Y, X) volume, optionally normalize_histogram()
//       ("textures": through compute_{min,max}_value_of_volume and the
//       two-texture overload), then write_volume(dir)
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "FBP3D.h"

using namespace ctrecon;

static int dump(const TextureStorage &t, const char *path)
{
    std::vector<float> host(t.texel_count() * t.channels());
    t.download(host.data());
    std::FILE *f = std::fopen(path, "wb");
    if (!f) return 3;
    std::fwrite(host.data(), sizeof(float), host.size(), f);
    std::fclose(f);
    return 0;
}

int main(int argc, char **argv)
{
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s read|write ...\n", argv[0]);
        return 2;
    }
    try {
        auto solver = std::make_shared<FBP3D>();
        if (!std::strcmp(argv[1], "read") && argc >= 11) {
            const bool sino = argc > 11 && !std::strcmp(argv[11], "sino");
            const int a[7] = {std::atoi(argv[3]), std::atoi(argv[4]), std::atoi(argv[5]),
                              std::atoi(argv[6]), std::atoi(argv[7]), std::atoi(argv[8]),
                              std::atoi(argv[9])};
            if (sino) {
                solver->read_projections_as_sinograms(argv[2], a[0], a[1], a[2], a[3], a[4], a[5],
                                                      a[6]);
                return dump(*solver->sinograms, argv[10]);
            }
            solver->read_projections(argv[2], a[0], a[1], a[2], a[3], a[4], a[5], a[6]);
            return dump(*solver->projections, argv[10]);
        }
        if (!std::strcmp(argv[1], "write") && argc >= 7) {
            const int X = std::atoi(argv[3]), Y = std::atoi(argv[4]), Z = std::atoi(argv[5]);
            solver->generate_blank_volume(X, Y, Z);
            std::vector<float> host((size_t)X * Y * Z);
            std::FILE *f = std::fopen(argv[2], "rb");
            if (!f || std::fread(host.data(), sizeof(float), host.size(), f) != host.size())
                return 3;
            std::fclose(f);
            solver->volume->upload(host.data());
            const std::string mode = argc > 7 ? argv[7] : "";
            if (mode == "normalize") {
                solver->normalize_histogram();
            } else if (mode == "textures") {
                Texture1D lo(1, ColorTextureFormat::R32UI), hi(1, ColorTextureFormat::R32UI);
                solver->compute_min_value_of_volume(lo);
                solver->compute_max_value_of_volume(hi);
                solver->normalize_histogram(lo, hi);
            }
            solver->write_volume(argv[6]);
            return 0;
        }
        std::fprintf(stderr, "bad arguments\n");
        return 2;
    } catch (const Error &e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
}
