// FastFourierTransformGPU.h -- drop-in FFFT over the gfx950 library.
//
// Same class name, public methods, parameter order and error behaviour as
// the reference's Source/FastFourierTransformGPU/FastFourierTransformGPU.h
// (class FFFT, :10-66).  Textures are the HIP device-memory stand-ins of
// Texture.h; complex textures are RG32F rows.  The reference's
// ComputeProgram members (cp_*, :48-62) have no counterpart: every method
// calls the C ABI in include/ctrecon.h instead of dispatching GLSL.
#pragma once

#include <memory>

#include "Texture.h"

namespace ctrecon {

class FFFT {
public:
    void fft_radix2(Texture2D &complex_texture);
    void inverse_fft_radix2(Texture2D &complex_texture);

    void fft(Texture2D &source_complex_texture, Texture2D &target_fourier_texture);
    std::shared_ptr<Texture2D> fft(Texture2D &source_complex_texture);

    void inverse_fft(Texture2D &source_fourier_texture, Texture2D &target_texture);
    std::shared_ptr<Texture2D> inverse_fft(Texture2D &source_fourier_texture,
                                           int texture_size_before_padding);

    void fft_shift(Texture2D &source_complex_texture, Texture2D &target_complex_texture);
    void fft_shift(Texture2D &complex_texture);
    void fft_shift(Texture2D &source_complex_texture, Texture2D &target_complex_texture,
                   int shift_amount);
    void fft_shift(Texture2D &complex_texture, int shift_amount);
    void inverse_fft_shift(Texture2D &source_complex_texture, Texture2D &target_complex_texture);
    void inverse_fft_shift(Texture2D &complex_texture);
    void inverse_fft_shift(Texture2D &source_complex_texture, Texture2D &target_complex_texture,
                           int shift_amount);
    void inverse_fft_shift(Texture2D &complex_texture, int shift_amount);

    // helper functions
    void reverse_bit_swap(Texture2D &texture);
    void fft_single_step(Texture2D &read_texture, Texture2D &write_texture, int step_index);

    std::shared_ptr<Texture2D> created_padded_complex_texture(Texture2D &source_complex_texture);

    void blit_texture_complex_to_complex(Texture2D &source_complex_texture,
                                         Texture2D &target_complex_texture);
    void blit_texture_real_to_complex(Texture2D &source_real_texture,
                                      Texture2D &target_complex_texture);
    void blit_texture_complex_to_real_r(Texture2D &source_complex_texture,
                                        Texture2D &target_real_texture);
    void blit_texture_complex_to_real_i(Texture2D &source_complex_texture,
                                        Texture2D &target_real_texture);

    std::shared_ptr<Texture2D> create_complex_texture_from_real(Texture2D &real_texture);
    std::shared_ptr<Texture2D> create_real_texture_from_complex_r(Texture2D &complex_texture);
    std::shared_ptr<Texture2D> create_real_texture_from_complex_i(Texture2D &complex_texture);

    void conjugate_complex_texture(Texture2D &texture);
    void multiply_complex_texture(Texture2D &texture, float multiplier);
    void divide_complex_texture(Texture2D &texture, float divisor);

    int floor_log2(int n);
    int ceil_log2(int n);
};

}  // namespace ctrecon
