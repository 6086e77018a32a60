// fbp_cone_fdk_main.cpp -- the call sequence of the reference's
// Mains/fbp_cone_fdk_main.cpp:19-32 on the drop-in FBP3D (a Shepp-Logan
// volume is projected first, as in its commented-out lines :21-22, since
// there are no scanner files here).  Writes the reconstruction as raw
// float32 (Z, Y, X) so a test can compare it with the C-ABI pipeline.
//
//   fbp_cone_fdk_main <size> <views> <out.raw> [fft_length] [staged_dtype]
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "FBP3D.h"

using namespace ctrecon;

int main(int argc, char **argv)
{
    if (argc < 4) {
        std::fprintf(stderr, "usage: %s size views out.raw [fft_length] [staged_dtype]\n", argv[0]);
        return 2;
    }
    const int size = std::atoi(argv[1]);
    const int projection_count = std::atoi(argv[2]);
    try {
        auto solver = std::make_shared<FBP3D>();
        if (argc > 4) solver->fft_length = std::atoi(argv[4]);
        if (argc > 5) solver->staged_dtype = std::atoi(argv[5]);

        solver->generate_shepplogan(size, size, size);
        solver->project_forward_cone_to_projections(730.87f, 669.04f, 409.60f, 213.84f, 1,
                                                    projection_count, 0);
        solver->apply_fdk_weights_to_projections(730.87f, 669.04f, 409.60f);
        solver->apply_filter_to_projections(FBP2D::FilterType::SHEPP_LOGAN);
        solver->project_backward_cone_fdk_from_projections(730.87f, 669.04f, 409.60f, 213.84f,
                                                           1, size, size, 0);

        std::vector<float> host(solver->volume->texel_count());
        solver->volume->download(host.data());
        std::FILE *f = std::fopen(argv[3], "wb");
        if (!f) return 3;
        std::fwrite(host.data(), sizeof(float), host.size(), f);
        std::fclose(f);
        std::printf("fbp_cone_fdk_main: %d^3 from %d views -> %s\n", size, projection_count,
                    argv[3]);
    } catch (const Error &e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    return 0;
}
