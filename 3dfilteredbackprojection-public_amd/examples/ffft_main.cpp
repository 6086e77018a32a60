// ffft_main.cpp -- every public method of the drop-in FFFT
// (Source/FastFourierTransformGPU/FastFourierTransformGPU.h:10-66) on
// textures read from raw files, each result written back as raw float32 so
// a test can compare it with the oracle:
//   in:  rows.raw      H rows of W complex values (W need not be a power of 2)
//        rows_p2.raw   H rows of P complex values (P a power of 2)
//   out: fft_padded      fft(src): padded to 2^ceil(log2 W), then FFT
//        inverse_cropped inverse_fft(fft_padded, W): IFFT, first W columns
//        inverse_same    inverse_fft(src, dst) with equal widths
//        shift_out, shift_in_place, unshift_in_place, shift_by_3
//        stages          reverse_bit_swap + fft_single_step 0 .. log2(P) - 1
//        real_r, real_i, complex_from_real
//        conj_mul_div    conjugate, x 2.5, / 4
//   stdout: floor_log2 / ceil_log2 of a few sizes and the error lines of the
//   reference's precondition checks (non-power-of-two FFT, wrong pad size)
//
//   ffft_main <dir> W P H
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "FastFourierTransformGPU.h"

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

static std::shared_ptr<Texture2D> load(const std::string &name, int w, int h)
{
    auto t = std::make_shared<Texture2D>(w, h, Texture2D::ColorTextureFormat::RG32F);
    std::vector<float> host(t->texel_count() * 2);
    std::FILE *f = std::fopen((g_dir + "/" + name + ".raw").c_str(), "rb");
    if (!f || std::fread(host.data(), sizeof(float), host.size(), f) != host.size())
        raise_error("[CTAnalyzer Error] cannot read " + name);
    std::fclose(f);
    t->upload(host.data());
    return t;
}

template <typename F>
static void expect_error(const char *what, F f)
{
    try {
        f();
        std::printf("no error: %s\n", what);
    } catch (const Error &e) {
        std::printf("error as expected: %s\n", what);
    }
}

int main(int argc, char **argv)
{
    if (argc < 5) {
        std::fprintf(stderr, "usage: %s dir W P H\n", argv[0]);
        return 2;
    }
    g_dir = argv[1];
    const int W = std::atoi(argv[2]), P = std::atoi(argv[3]), H = std::atoi(argv[4]);
    try {
        FFFT fft;
        auto src = load("rows", W, H);
        auto src_p2 = load("rows_p2", P, H);

        std::shared_ptr<Texture2D> padded = fft.fft(*src);
        dump(*padded, "fft_padded");
        dump(*fft.inverse_fft(*padded, W), "inverse_cropped");
        auto same = src_p2->create_texture_with_same_parameters();
        fft.inverse_fft(*src_p2, *same);
        dump(*same, "inverse_same");

        auto shifted = src_p2->create_texture_with_same_parameters();
        fft.fft_shift(*src_p2, *shifted);
        dump(*shifted, "shift_out");
        auto t = load("rows_p2", P, H);
        fft.fft_shift(*t);
        dump(*t, "shift_in_place");
        fft.inverse_fft_shift(*t);
        dump(*t, "unshift_in_place");
        fft.fft_shift(*t, 3);
        dump(*t, "shift_by_3");

        auto a = load("rows_p2", P, H);
        auto b = a->create_texture_with_same_parameters();
        fft.reverse_bit_swap(*a);
        for (int s = 0; s < fft.floor_log2(P); s++) {
            fft.fft_single_step(*a, *b, s);
            std::swap(a, b);
        }
        dump(*a, "stages");

        dump(*fft.create_real_texture_from_complex_r(*src), "real_r");
        dump(*fft.create_real_texture_from_complex_i(*src), "real_i");
        auto real = fft.create_real_texture_from_complex_r(*src);
        dump(*fft.create_complex_texture_from_real(*real), "complex_from_real");

        auto c = load("rows", W, H);
        fft.conjugate_complex_texture(*c);
        fft.multiply_complex_texture(*c, 2.5f);
        fft.divide_complex_texture(*c, 4.0f);
        dump(*c, "conj_mul_div");

        for (int n : {1, 2, 3, 300, 512, 513})
            std::printf("log2 %d: floor %d ceil %d\n", n, fft.floor_log2(n), fft.ceil_log2(n));
        expect_error("fft_radix2 of a non-power-of-two width", [&] { fft.fft_radix2(*src); });
        expect_error("fft into a target of the wrong width", [&] { fft.fft(*src, *src_p2); });
    } catch (const Error &e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    return 0;
}
