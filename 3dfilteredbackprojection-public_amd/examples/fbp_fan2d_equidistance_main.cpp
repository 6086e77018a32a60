// fbp_fan2d_equidistance_main.cpp -- the call sequence of the reference's
// Mains/fbp_fan2d_equidistance_main.cpp:32-49 on the drop-in FBP2D and FFFT:
// a W x W image is projected by the parallel projector and by the
// flat-detector fan projector (SOD 730.87, detector distance 669.04,
// detector 409.60, field 213.84 wide, W views over one turn), the fan
// sinogram is copied and rebinned to parallel rays
// (parallelize_fan_beam_sinogram), filtered through the same Fourier-space
// chain as the parallel main and backprojected by FBP2D::project_backward.
// Input and every stage are raw float32 files, as in fbp_parallel2d_main.
//
//   fbp_fan2d_equidistance_main <image.raw> <W> <outdir> [filter]
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "FBP2D.h"
#include "FastFourierTransformGPU.h"

using namespace ctrecon;

static void dump(const TextureStorage &t, const std::string &path)
{
    std::vector<float> host(t.texel_count() * t.channels());
    t.download(host.data());
    std::FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) raise_error("[CTAnalyzer Error] cannot write " + path);
    std::fwrite(host.data(), sizeof(float), host.size(), f);
    std::fclose(f);
}

static std::vector<float> read_raw(const char *path, size_t count)
{
    std::vector<float> v(count);
    std::FILE *f = std::fopen(path, "rb");
    if (!f || std::fread(v.data(), sizeof(float), count, f) != count)
        raise_error(std::string("[CTAnalyzer Error] cannot read ") + path);
    std::fclose(f);
    return v;
}

int main(int argc, char **argv)
{
    if (argc < 4) {
        std::fprintf(stderr, "usage: %s image.raw W outdir [filter]\n", argv[0]);
        return 2;
    }
    const int projection_width = std::atoi(argv[2]);
    const std::string out = argv[3];
    const int filter = argc > 4 ? std::atoi(argv[4]) : FBP2D::FilterType::RAM_LAK;
    try {
        FBP2D fbp_solver;
        FFFT fft_solver;

        std::vector<float> image = read_raw(argv[1], (size_t)projection_width * projection_width);
        auto source_texture = std::make_shared<Texture2D>(projection_width, projection_width,
                                                          Texture2D::ColorTextureFormat::R32F);
        source_texture->load_data(image.data(), Texture2D::ColorFormat::RED, Texture2D::Type::FLOAT);

        std::shared_ptr<Texture2D> sinogram_parallel =
            fbp_solver.project_forward(*source_texture, projection_width, 1);
        dump(*sinogram_parallel, out + "/sinogram_parallel.raw");
        std::shared_ptr<Texture2D> sinogram_original = fbp_solver.project_forward_equiangular(
            *source_texture, projection_width, 730.87f, 669.04f, 409.60f, 213.84f, 1);
        std::shared_ptr<Texture2D> sinogram_fan =
            sinogram_original->create_texture_with_same_parameters();
        std::shared_ptr<Texture2D> sinogram = sinogram_original->create_texture_with_same_parameters();
        fbp_solver._texture_blit_float1_to_float1(*sinogram_original, *sinogram_fan);
        dump(*sinogram_fan, out + "/sinogram_fan.raw");
        fbp_solver.parallelize_fan_beam_sinogram(*sinogram_fan, *sinogram, 730.87f, 669.04f,
                                                 409.60f, 213.84f, 1);
        dump(*sinogram, out + "/sinogram_rebinned.raw");

        std::shared_ptr<Texture2D> sinogram_complex =
            fft_solver.create_complex_texture_from_real(*sinogram);
        fft_solver.fft_radix2(*sinogram_complex);
        fft_solver.fft_shift(*sinogram_complex);

        fbp_solver.apply_filter(*sinogram_complex, (FBP2D::FilterType)filter);

        fft_solver.inverse_fft_shift(*sinogram_complex);
        fft_solver.inverse_fft_radix2(*sinogram_complex);
        fft_solver.blit_texture_complex_to_real_r(*sinogram_complex, *sinogram);
        dump(*sinogram, out + "/filtered.raw");

        std::shared_ptr<Texture2D> reconstruction = fbp_solver.project_backward(*sinogram);
        dump(*reconstruction, out + "/reconstruction.raw");
        std::printf("fbp_fan2d_equidistance_main: %d^2, %d views -> %s\n", projection_width,
                    projection_width, out.c_str());
    } catch (const Error &e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    return 0;
}
