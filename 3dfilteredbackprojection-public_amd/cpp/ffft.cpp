// ffft.cpp -- FFFT methods (reference: Source/FastFourierTransformGPU/
// FastFourierTransformGPU.cpp) on top of the C ABI.
#include <algorithm>
#include <string>

#include "FastFourierTransformGPU.h"
#include "ctrecon.h"

namespace ctrecon {

static void blit(int mode, Texture2D &src, Texture2D &dst, const char *what)
{
    glm::ivec2 a = src.get_size(), b = dst.get_size();
    check_status(ctr_blit(get_stream(), mode, src.data(), a.x, a.y, dst.data(), b.x, b.y), what);
}

// FastFourierTransformGPU.cpp:3-28
void FFFT::fft_radix2(Texture2D &complex_texture)
{
    int size = complex_texture.get_size().x;
    int log2_size = floor_log2(size);
    if ((1 << log2_size) != size)
        raise_error("[CTAnalyzer Error] FFFT::fft_radix2() is called but texture's width is "
                    "not a power of 2");
    check_status(ctr_fft_radix2(get_stream(), complex_texture.data(), size,
                                complex_texture.get_size().y),
                 "FFFT::fft_radix2");
}

// FastFourierTransformGPU.cpp:30-36
void FFFT::inverse_fft_radix2(Texture2D &complex_texture)
{
    int size = complex_texture.get_size().x;
    int log2_size = floor_log2(size);
    if ((1 << log2_size) != size)
        raise_error("[CTAnalyzer Error] FFFT::fft_radix2() is called but texture's width is "
                    "not a power of 2");
    check_status(ctr_inverse_fft_radix2(get_stream(), complex_texture.data(), size,
                                        complex_texture.get_size().y),
                 "FFFT::inverse_fft_radix2");
}

// FastFourierTransformGPU.cpp:38-52
void FFFT::fft(Texture2D &source_complex_texture, Texture2D &target_fourier_texture)
{
    int size = source_complex_texture.get_size().x;
    int logsize = 1 << ceil_log2(size);
    if (target_fourier_texture.get_size().x != logsize)
        raise_error("[CTAnalyzer Error] FFFT::fft(Texture2D& source_complex_texture, Texture2D& "
                    "target_fourier_texture) was called but target_fourier_texture's x "
                    "dimention is now radix-2, it had to be " + std::to_string(logsize));
    target_fourier_texture.clear(glm::vec2(0, 0), 0);
    blit_texture_complex_to_complex(source_complex_texture, target_fourier_texture);
    fft_radix2(target_fourier_texture);
}

// FastFourierTransformGPU.cpp:54-64
std::shared_ptr<Texture2D> FFFT::fft(Texture2D &source_complex_texture)
{
    std::shared_ptr<Texture2D> target = created_padded_complex_texture(source_complex_texture);
    fft(source_complex_texture, *target);
    return target;
}

// FastFourierTransformGPU.cpp:66-81
void FFFT::inverse_fft(Texture2D &source_fourier_texture, Texture2D &target_texture)
{
    if (source_fourier_texture.get_size().x == target_texture.get_size().x) {
        blit_texture_complex_to_complex(source_fourier_texture, target_texture);
        inverse_fft_radix2(target_texture);
    } else {
        std::shared_ptr<Texture2D> copy = source_fourier_texture.create_texture_with_same_parameters();
        blit_texture_complex_to_complex(source_fourier_texture, *copy);
        inverse_fft_radix2(*copy);
        blit_texture_complex_to_complex(*copy, target_texture);
    }
}

// FastFourierTransformGPU.cpp:83-91
std::shared_ptr<Texture2D> FFFT::inverse_fft(Texture2D &source_fourier_texture,
                                             int texture_size_before_padding)
{
    auto target = std::make_shared<Texture2D>(texture_size_before_padding,
                                              source_fourier_texture.get_size().y,
                                              ColorTextureFormat::RG32F);
    inverse_fft(source_fourier_texture, *target);
    return target;
}

// FastFourierTransformGPU.cpp:93-142
void FFFT::fft_shift(Texture2D &source_complex_texture, Texture2D &target_complex_texture)
{
    fft_shift(source_complex_texture, target_complex_texture,
              source_complex_texture.get_size().x / 2);
}

void FFFT::fft_shift(Texture2D &complex_texture)
{
    fft_shift(complex_texture, complex_texture.get_size().x / 2);
}

void FFFT::fft_shift(Texture2D &source_complex_texture, Texture2D &target_complex_texture,
                     int shift_amount)
{
    glm::ivec2 s = source_complex_texture.get_size();
    if (target_complex_texture.get_size() != s)
        raise_error("[CTAnalyzer Error] FFFT::fft_shift() source and target sizes differ");
    check_status(ctr_fft_shift(get_stream(), source_complex_texture.data(),
                               target_complex_texture.data(), s.x, s.y, shift_amount),
                 "FFFT::fft_shift");
}

// The reference shifts into a new texture and blits it back (:113-119); the
// C ABI shifts in place without a temporary.
void FFFT::fft_shift(Texture2D &complex_texture, int shift_amount)
{
    fft_shift(complex_texture, complex_texture, shift_amount);
}

void FFFT::inverse_fft_shift(Texture2D &source_complex_texture, Texture2D &target_complex_texture)
{
    inverse_fft_shift(source_complex_texture, target_complex_texture,
                      source_complex_texture.get_size().x / 2);
}

void FFFT::inverse_fft_shift(Texture2D &complex_texture)
{
    inverse_fft_shift(complex_texture, complex_texture.get_size().x / 2);
}

void FFFT::inverse_fft_shift(Texture2D &source_complex_texture, Texture2D &target_complex_texture,
                             int shift_amount)
{
    fft_shift(source_complex_texture, target_complex_texture, -shift_amount);
}

void FFFT::inverse_fft_shift(Texture2D &complex_texture, int shift_amount)
{
    fft_shift(complex_texture, complex_texture, -shift_amount);
}

// FastFourierTransformGPU.cpp:238-244
void FFFT::reverse_bit_swap(Texture2D &texture)
{
    glm::ivec2 s = texture.get_size();
    check_status(ctr_fft_bit_reverse(get_stream(), texture.data(), s.x, s.y),
                 "FFFT::reverse_bit_swap");
}

// FastFourierTransformGPU.cpp:246-256
void FFFT::fft_single_step(Texture2D &read_texture, Texture2D &write_texture, int step_index)
{
    glm::ivec2 s = read_texture.get_size();
    check_status(ctr_fft_single_stage(get_stream(), read_texture.data(), write_texture.data(), s.x,
                                      s.y, step_index),
                 "FFFT::fft_single_step");
}

// FastFourierTransformGPU.cpp:258-267
std::shared_ptr<Texture2D> FFFT::created_padded_complex_texture(Texture2D &source_complex_texture)
{
    int padded_width = 1 << ceil_log2(source_complex_texture.get_size().x);
    auto padded = std::make_shared<Texture2D>(padded_width, source_complex_texture.get_size().y,
                                              ColorTextureFormat::RG32F);
    blit_texture_complex_to_complex(source_complex_texture, *padded);
    return padded;
}

// FastFourierTransformGPU.cpp:144-186
void FFFT::blit_texture_complex_to_complex(Texture2D &s, Texture2D &t)
{
    blit(CTR_BLIT_COMPLEX_TO_COMPLEX, s, t, "FFFT::blit_texture_complex_to_complex");
}
void FFFT::blit_texture_real_to_complex(Texture2D &s, Texture2D &t)
{
    blit(CTR_BLIT_REAL_TO_COMPLEX, s, t, "FFFT::blit_texture_real_to_complex");
}
void FFFT::blit_texture_complex_to_real_r(Texture2D &s, Texture2D &t)
{
    blit(CTR_BLIT_COMPLEX_TO_REAL_R, s, t, "FFFT::blit_texture_complex_to_real_r");
}
void FFFT::blit_texture_complex_to_real_i(Texture2D &s, Texture2D &t)
{
    blit(CTR_BLIT_COMPLEX_TO_REAL_I, s, t, "FFFT::blit_texture_complex_to_real_i");
}

// FastFourierTransformGPU.cpp:188-210
std::shared_ptr<Texture2D> FFFT::create_complex_texture_from_real(Texture2D &real_texture)
{
    auto c = std::make_shared<Texture2D>(real_texture.get_size().x, real_texture.get_size().y,
                                         ColorTextureFormat::RG32F);
    blit_texture_real_to_complex(real_texture, *c);
    return c;
}
std::shared_ptr<Texture2D> FFFT::create_real_texture_from_complex_r(Texture2D &complex_texture)
{
    auto r = std::make_shared<Texture2D>(complex_texture.get_size().x,
                                         complex_texture.get_size().y, ColorTextureFormat::R32F);
    blit_texture_complex_to_real_r(complex_texture, *r);
    return r;
}
std::shared_ptr<Texture2D> FFFT::create_real_texture_from_complex_i(Texture2D &complex_texture)
{
    auto r = std::make_shared<Texture2D>(complex_texture.get_size().x,
                                         complex_texture.get_size().y, ColorTextureFormat::R32F);
    blit_texture_complex_to_real_i(complex_texture, *r);
    return r;
}

// FastFourierTransformGPU.cpp:212-236
void FFFT::conjugate_complex_texture(Texture2D &texture)
{
    check_status(ctr_conjugate(get_stream(), texture.data(), texture.texel_count()),
                 "FFFT::conjugate_complex_texture");
}
void FFFT::multiply_complex_texture(Texture2D &texture, float multiplier)
{
    check_status(ctr_scale_complex(get_stream(), texture.data(), texture.texel_count(), multiplier),
                 "FFFT::multiply_complex_texture");
}
void FFFT::divide_complex_texture(Texture2D &texture, float divisor)
{
    check_status(ctr_scale_complex(get_stream(), texture.data(), texture.texel_count(),
                                   1.0f / divisor),
                 "FFFT::divide_complex_texture");
}

// FastFourierTransformGPU.cpp:269-280
int FFFT::floor_log2(int n)
{
    int log2_size = 0;
    for (unsigned long t = (unsigned long)n; t > 1UL; t >>= 1) log2_size++;
    return log2_size;
}
int FFFT::ceil_log2(int n) { return floor_log2(2 * n - 1); }

}  // namespace ctrecon
