// FBP2D.h -- drop-in 2D projector / backprojector / Fourier filter over the
// gfx950 library.  Same class, enum, methods and default arguments as the
// reference's Source/FilteredBackProjection/FBP2D.h (:8-72).  The
// ComputeProgram members (:44-70) have no counterpart; the legacy integer
// pipeline helpers (_int_texture_add/_mult, declared but never defined in the
// reference, :39-40) are omitted.
#pragma once

#include <memory>

#include "Texture.h"

namespace ctrecon {

class FBP2D {
public:
    enum FilterType {
        RAM_LAK = 0,
        SHEPP_LOGAN = 1,
        COSINE = 2,
    };

    // main functions
    void project_forward(Texture2D &source_image, Texture2D &target_image,
                         float total_rotation_count);
    void project_forward_equiangular(Texture2D &source_image, Texture2D &target_image,
                                     float xray_source_distance, float detector_panel_distance,
                                     float detector_width, float volume_width,
                                     float total_rotation_count);
    void project_backward(Texture2D &sinogram, Texture2D &out_texture);

    std::shared_ptr<Texture2D> project_forward(Texture2D &source_image, int projection_count,
                                               float total_rotation_count);
    std::shared_ptr<Texture2D> project_forward_equiangular(
        Texture2D &source_image, int projection_count, float xray_source_distance,
        float detector_panel_distance, float detector_width, float volume_width,
        float total_rotation_count);
    std::shared_ptr<Texture2D> project_backward(Texture2D &sinogram);

    void parallelize_fan_beam_sinogram(Texture2D &source_sinogram, Texture2D &target_sinogram,
                                       float xray_source_distance, float detector_panel_distance,
                                       float detector_width, float volume_width,
                                       float total_rotation_count);

    void apply_filter(Texture2D &sinogram, FilterType filter_type = FilterType::RAM_LAK);

    // utility functions
    void _texture_blit_float1_to_float4(Texture2D &source_float_texture,
                                        Texture2D &target_float_texture);
    void _texture_blit_float1_to_float1(Texture2D &source_float_texture,
                                        Texture2D &target_float_texture);
    void _int_texture_div(Texture2D &int_texture, float denominator);

private:
    DeviceBuffer staged_;
};

}  // namespace ctrecon
