// fbp3d_passes_main.cpp -- the detector and volume passes of the drop-in
// FBP3D (FBP3D.cpp:333-458): log_normalize_projections / _sinograms,
// apply_fdk_weights_to_projections / _sinograms, apply_filter_to_projections
// / _sinograms and clip_negatives_of_volume, on data read from raw float32
// files in <dir> (projections.raw (N, nv, nu) of intensities, sinograms.raw
// (nv, N, nu), volume.raw (Z, Y, X)); each stage is written back as raw
// float32 so a test can compare it with the oracle.
//
//   fbp3d_passes_main <dir> nu nv N X Y Z base sod odd width
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "FBP3D.h"

using namespace ctrecon;

static std::string g_dir;

static void dump(const TextureStorage &t, const std::string &name)
{
    std::vector<float> host(t.texel_count());
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

int main(int argc, char **argv)
{
    if (argc < 12) {
        std::fprintf(stderr, "usage: %s dir nu nv N X Y Z base sod odd width\n", argv[0]);
        return 2;
    }
    g_dir = argv[1];
    const int nu = std::atoi(argv[2]), nv = std::atoi(argv[3]), N = std::atoi(argv[4]);
    const int X = std::atoi(argv[5]), Y = std::atoi(argv[6]), Z = std::atoi(argv[7]);
    const float base = (float)std::atof(argv[8]), sod = (float)std::atof(argv[9]);
    const float odd = (float)std::atof(argv[10]), width = (float)std::atof(argv[11]);
    try {
        FBP3D solver;
        solver.generate_blank_projections(nu, nv, N);
        solver.generate_blank_sinograms(nu, nv, N);
        solver.generate_blank_volume(X, Y, Z);
        load(*solver.projections, "projections");
        load(*solver.sinograms, "sinograms");
        load(*solver.volume, "volume");

        solver.log_normalize_projections(base);
        dump(*solver.projections, "projections_log");
        solver.apply_fdk_weights_to_projections(sod, odd, width);
        dump(*solver.projections, "projections_fdk");
        solver.apply_filter_to_projections(FBP2D::FilterType::SHEPP_LOGAN);
        dump(*solver.projections, "projections_filtered");

        solver.log_normalize_sinograms(base);
        dump(*solver.sinograms, "sinograms_log");
        solver.apply_fdk_weights_to_sinograms(sod, odd, width);
        dump(*solver.sinograms, "sinograms_fdk");
        solver.apply_filter_to_sinograms(FBP2D::FilterType::COSINE);
        dump(*solver.sinograms, "sinograms_filtered");

        solver.clip_negatives_of_volume();
        dump(*solver.volume, "volume_clipped");
    } catch (const Error &e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    return 0;
}
