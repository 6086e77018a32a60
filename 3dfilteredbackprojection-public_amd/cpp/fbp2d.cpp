// fbp2d.cpp -- FBP2D methods (reference: Source/FilteredBackProjection/
// FBP2D.cpp) on top of the C ABI.  A 2D image (W, H) is passed to the 3D
// entry points as a volume (W, 1, H); a sinogram (nu, N) as projections with
// nv = 1.  Pixel pitch is 1 (the reference's volume_width = image width,
// FBP2D.cpp:67,111).
#include <cmath>

#include "FBP2D.h"
#include "ctrecon.h"

namespace ctrecon {

// FBP2D.cpp:5-11
std::shared_ptr<Texture2D> FBP2D::project_forward(Texture2D &source_image, int projection_count,
                                                  float total_rotation_count)
{
    auto sinogram = std::make_shared<Texture2D>(source_image.get_size().x, projection_count,
                                                ColorTextureFormat::R32F);
    project_forward(source_image, *sinogram, total_rotation_count);
    return sinogram;
}

// FBP2D.cpp:13-19
std::shared_ptr<Texture2D> FBP2D::project_forward_equiangular(
    Texture2D &source_image, int projection_count, float xray_source_distance,
    float detector_panel_distance, float detector_width, float volume_width,
    float total_rotation_count)
{
    auto sinogram = std::make_shared<Texture2D>(source_image.get_size().x, projection_count,
                                                ColorTextureFormat::R32F);
    project_forward_equiangular(source_image, *sinogram, xray_source_distance,
                                detector_panel_distance, detector_width, volume_width,
                                total_rotation_count);
    return sinogram;
}

// FBP2D.cpp:21-29
std::shared_ptr<Texture2D> FBP2D::project_backward(Texture2D &sinogram)
{
    glm::ivec2 s = sinogram.get_size();
    auto slice = std::make_shared<Texture2D>(s.x, s.x, ColorTextureFormat::R32F);
    project_backward(sinogram, *slice);
    return slice;
}

// FBP2D.cpp:31-52: rays over a parallel detector of width volume_width*sqrt(2)
void FBP2D::parallelize_fan_beam_sinogram(Texture2D &source_sinogram, Texture2D &target_sinogram,
                                          float xray_source_distance,
                                          float detector_panel_distance, float detector_width,
                                          float volume_width, float total_rotation_count)
{
    glm::ivec2 s = source_sinogram.get_size(), t = target_sinogram.get_size();
    if (s.y != t.y)
        raise_error("[CTAnalyzer Error] FBP2D::parallelize_fan_beam_sinogram() source and "
                    "target sinograms need the same projection count");
    check_status(ctr_rebin_fan_to_parallel(get_stream(), source_sinogram.data(), s.x, s.y,
                                           xray_source_distance, detector_panel_distance,
                                           detector_width, total_rotation_count,
                                           target_sinogram.data(), t.x,
                                           volume_width * std::sqrt(2.0)),
                 "FBP2D::parallelize_fan_beam_sinogram");
}

// FBP2D.cpp:54-72: parallel rays, plane width W*sqrt(2), offset 0
void FBP2D::project_forward(Texture2D &source_image, Texture2D &target_image,
                            float total_rotation_count)
{
    glm::ivec2 img = source_image.get_size(), sino = target_image.get_size();
    check_status(ctr_forward_project(get_stream(), nullptr, CTR_GEOMETRY_PARALLEL, source_image.data(),
                                     img.x, 1, img.y, 0.0, 0.0, img.x * std::sqrt(2.0),
                                     (double)img.x, total_rotation_count, 0.0,
                                     target_image.data(), CTR_LAYOUT_PROJECTIONS, sino.x, 1,
                                     sino.y, 1.0),
                 "FBP2D::project_forward");
}

// FBP2D.cpp:74-94: flat-detector fan rays (the reference's demo calls this
// path "equidistance", Mains/fbp_fan2d_equidistance_main.cpp:33)
void FBP2D::project_forward_equiangular(Texture2D &source_image, Texture2D &target_image,
                                        float xray_source_distance,
                                        float detector_panel_distance, float detector_width,
                                        float volume_width, float total_rotation_count)
{
    glm::ivec2 img = source_image.get_size(), sino = target_image.get_size();
    check_status(ctr_forward_project(get_stream(), nullptr, CTR_GEOMETRY_CONE, source_image.data(), img.x,
                                     1, img.y, xray_source_distance, detector_panel_distance,
                                     detector_width, volume_width, total_rotation_count, 0.0,
                                     target_image.data(), CTR_LAYOUT_PROJECTIONS, sino.x, 1,
                                     sino.y, 1.0),
                 "FBP2D::project_forward_equiangular");
}

// FBP2D.cpp:96-122: pixel-driven parallel backprojection over one full turn
// with a 0.006 rad offset and plane width W*sqrt(2), summed then / N.
void FBP2D::project_backward(Texture2D &sinogram, Texture2D &out_texture)
{
    glm::ivec2 s = sinogram.get_size(), o = out_texture.get_size();
    size_t bytes = 0;
    check_status(ctr_staged_layout(s.x, 1, s.y, CTR_F32, nullptr, nullptr, &bytes),
                 "FBP2D::project_backward");
    if (staged_.bytes() < bytes) staged_.resize(bytes);
    check_status(ctr_stage_projections(get_stream(), nullptr, sinogram.data(),
                                       CTR_LAYOUT_PROJECTIONS, s.x, 1, s.y, staged_.get(), CTR_F32,
                                       nullptr),
                 "FBP2D::project_backward");
    check_status(ctr_backproject(get_stream(), nullptr, CTR_GEOMETRY_PARALLEL, staged_.get(),
                                 CTR_F32, nullptr, s.x, 1, s.y, 0.0, 0.0, o.x * std::sqrt(2.0),
                                 (double)o.x, 1.0, 0.006, out_texture.data(), o.x, 1, o.y,
                                 1.0f / (float)s.y, 0, 0),
                 "FBP2D::project_backward");
}

// FBP2D.cpp:124-144
void FBP2D::apply_filter(Texture2D &sinogram, FilterType filter_type)
{
    if (filter_type != RAM_LAK && filter_type != SHEPP_LOGAN && filter_type != COSINE)
        raise_error("[CTAnalyzer Error] FBP2D::apply_filter() unknown filter type");
    glm::ivec2 s = sinogram.get_size();
    check_status(ctr_apply_filter(get_stream(), sinogram.data(), s.x, s.y, (int)filter_type),
                 "FBP2D::apply_filter");
}

// FBP2D.cpp:168-173
void FBP2D::_texture_blit_float1_to_float1(Texture2D &source_float_texture,
                                           Texture2D &target_float_texture)
{
    glm::ivec2 a = source_float_texture.get_size(), b = target_float_texture.get_size();
    copy_block(target_float_texture.data(), (size_t)b.x, source_float_texture.data(),
               (size_t)a.x, (size_t)std::min(a.x, b.x), (size_t)std::min(a.y, b.y));
}

// FBP2D.cpp:160-166: replicate the red channel into RGBA
void FBP2D::_texture_blit_float1_to_float4(Texture2D &source_float_texture,
                                           Texture2D &target_float_texture)
{
    glm::ivec2 a = source_float_texture.get_size(), b = target_float_texture.get_size();
    int w = std::min(a.x, b.x), h = std::min(a.y, b.y);
    int ch = target_float_texture.channels();
    for (int c = 0; c < ch; c++)
        for (int y = 0; y < h; y++)
            copy_block(target_float_texture.data() + ((size_t)y * b.x) * ch + c, (size_t)ch,
                       source_float_texture.data() + (size_t)y * a.x, 1, 1, (size_t)w);
}

// FBP2D.cpp:189-195
void FBP2D::_int_texture_div(Texture2D &int_texture, float denominator)
{
    check_status(ctr_scale_real(get_stream(), int_texture.data(),
                                int_texture.texel_count() * int_texture.channels(),
                                1.0f / denominator),
                 "FBP2D::_int_texture_div");
}

}  // namespace ctrecon
