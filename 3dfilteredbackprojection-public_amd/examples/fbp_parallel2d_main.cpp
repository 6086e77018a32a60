// fbp_parallel2d_main.cpp -- the call sequence of the reference's
// Mains/fbp_parallel2d_main.cpp:29-41 on the drop-in FBP2D and FFFT:
// parallel forward projection of a W x W image over one turn with W views,
// then the Fourier-space chain create_complex_texture_from_real ->
// fft_radix2 -> fft_shift -> FBP2D::apply_filter -> inverse_fft_shift ->
// inverse_fft_radix2 -> blit_texture_complex_to_real_r, then
// FBP2D::project_backward and the RGBA copy for display.  The image comes
// from a raw float32 file (the reference loads a PNG through GraphicsCortex,
// which is absent); every stage is written as raw float32 so a test can
// compare each one with the oracle.
//
//   fbp_parallel2d_main <image.raw> <W> <outdir> [filter]
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

        std::shared_ptr<Texture2D> sinogram =
            fbp_solver.project_forward(*source_texture, projection_width, 1);
        dump(*sinogram, out + "/sinogram.raw");

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

        auto reconstruction_white = std::make_shared<Texture2D>(
            source_texture->get_size().x, source_texture->get_size().y,
            Texture2D::ColorTextureFormat::RGBA32F);
        fbp_solver._texture_blit_float1_to_float4(*reconstruction, *reconstruction_white);
        dump(*reconstruction_white, out + "/reconstruction_rgba.raw");
        std::printf("fbp_parallel2d_main: %d^2, %d views -> %s\n", projection_width,
                    projection_width, out.c_str());
    } catch (const Error &e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    return 0;
}
