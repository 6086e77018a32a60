// fbp3d_copies_main.cpp -- exercises the slice and view copies of the
// drop-in FBP3D (FBP3D.cpp:639-778): load/store_volume_slice_{x,y,z},
// load/store_projection, load/store_sinogram and
// load_projection_from_sinograms / store_projection_to_sinograms, each with
// a texture of the exact size and one of a different size (the copied
// extent is the min of both; the rest of the target keeps its value).
// Inputs and outputs are raw float32 files in <dir>:
//   in:  volume.raw (Z, Y, X), projections.raw (N, nv, nu),
//        sinograms.raw (nv, N, nu), patch.raw (4096 floats)
//   out: <method>_<exact|odd>.raw for every load, and volume_out.raw,
//        projections_out.raw, sinograms_out.raw after all stores
//
//   fbp3d_copies_main <dir> X Y Z nu nv N
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "FBP3D.h"

using namespace ctrecon;

static std::string g_dir;

static void dump(const TextureStorage &t, const std::string &name)
{
    std::vector<float> host(t.texel_count() * t.channels());
    t.download(host.data());
    std::FILE *f = std::fopen((g_dir + "/" + name + ".raw").c_str(), "wb");
    if (!f) raise_error("[CTAnalyzer Error] cannot write " + name);
    std::fwrite(host.data(), sizeof(float), host.size(), f);
    std::fclose(f);
}

static void load(TextureStorage &t, const std::string &name)
{
    std::vector<float> host(t.texel_count());
    std::FILE *f = std::fopen((g_dir + "/" + name + ".raw").c_str(), "rb");
    if (!f || std::fread(host.data(), sizeof(float), host.size(), f) != host.size())
        raise_error("[CTAnalyzer Error] cannot read " + name);
    std::fclose(f);
    t.upload(host.data());
}

// A w x h texture cleared to -1, or filled from the first w*h patch values.
static std::shared_ptr<Texture2D> texture(int w, int h, const std::vector<float> *patch)
{
    auto t = std::make_shared<Texture2D>(w, h, Texture2D::ColorTextureFormat::R32F);
    if (patch)
        t->upload(patch->data());
    else
        t->clear(-1.0f);
    return t;
}

int main(int argc, char **argv)
{
    if (argc < 8) {
        std::fprintf(stderr, "usage: %s dir X Y Z nu nv N\n", argv[0]);
        return 2;
    }
    g_dir = argv[1];
    const int X = std::atoi(argv[2]), Y = std::atoi(argv[3]), Z = std::atoi(argv[4]);
    const int nu = std::atoi(argv[5]), nv = std::atoi(argv[6]), N = std::atoi(argv[7]);
    try {
        FBP3D solver;
        solver.generate_blank_volume(X, Y, Z);
        solver.generate_blank_projections(nu, nv, N);
        solver.generate_blank_sinograms(nu, nv, N);
        load(*solver.volume, "volume");
        load(*solver.projections, "projections");
        load(*solver.sinograms, "sinograms");
        std::vector<float> patch(4096);
        {
            Texture1D p(4096, Texture1D::ColorTextureFormat::R32F);
            load(p, "patch");
            p.download(patch.data());
        }
        // odd-sized targets: 3 wider and 2 shorter than the exact extent
        struct Case {
            const char *name;
            int w, h, index;
        };
        const Case loads[] = {{"slice_x", Y, Z, X / 3},   {"slice_y", X, Z, Y / 2},
                              {"slice_z", X, Y, Z - 1},   {"projection", nu, nv, N / 2},
                              {"sinogram", nu, N, nv / 3}, {"projection_from_sinograms", nu, nv, 1}};
        for (const Case &c : loads)
            for (int odd = 0; odd < 2; odd++) {
                auto t = texture(c.w + 3 * odd, c.h - 2 * odd, nullptr);
                const std::string n = c.name;
                if (n == "slice_x") solver.load_volume_slice_x(c.index, *t);
                if (n == "slice_y") solver.load_volume_slice_y(c.index, *t);
                if (n == "slice_z") solver.load_volume_slice_z(c.index, *t);
                if (n == "projection") solver.load_projection(c.index, *t);
                if (n == "sinogram") solver.load_sinogram(c.index, *t);
                if (n == "projection_from_sinograms")
                    solver.load_projection_from_sinograms(c.index, *t);
                dump(*t, n + (odd ? "_odd" : "_exact"));
            }
        // stores: exact-size patches at one index, odd-size ones at another
        solver.store_volume_slice_x(*texture(Y, Z, &patch), 1);
        solver.store_volume_slice_x(*texture(Y - 2, Z + 1, &patch), X - 1);
        solver.store_volume_slice_y(*texture(X, Z, &patch), 0);
        solver.store_volume_slice_y(*texture(X + 2, Z - 3, &patch), Y - 2);
        solver.store_volume_slice_z(*texture(X, Y, &patch), 2);
        solver.store_volume_slice_z(*texture(X - 1, Y + 2, &patch), Z / 2);
        solver.store_projection(*texture(nu, nv, &patch), 0);
        solver.store_projection(*texture(nu + 1, nv - 1, &patch), N - 1);
        solver.store_sinogram(*texture(nu, N, &patch), 0);
        solver.store_sinogram(*texture(nu - 2, N + 1, &patch), nv - 1);
        solver.store_projection_to_sinograms(*texture(nu, nv, &patch), 2);
        solver.store_projection_to_sinograms(*texture(nu + 2, nv - 2, &patch), N - 2);
        dump(*solver.volume, "volume_out");
        dump(*solver.projections, "projections_out");
        dump(*solver.sinograms, "sinograms_out");
        // an out-of-range index is the reference's error
        solver.load_projection(N, *texture(nu, nv, nullptr));
    } catch (const Error &e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    return 0;
}
