// fbp3d.cpp -- FBP3D methods (reference: Source/FilteredBackProjection/
// FBP3D.cpp) on top of the C ABI.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <string>

#include "FBP3D.h"
#include "ctrecon.h"

namespace ctrecon {

static int ceil_pow2_i(int n)
{
    int p = 1;
    while (p < n) p <<= 1;
    return p;
}

FBP3D::FBP3D() {}

// FBP3D.cpp:16-28: reuse and clear when the size matches, else reallocate
void FBP3D::generate_blank_volume(int x_size, int y_size, int z_size)
{
    if (volume != nullptr && volume->get_size() == glm::ivec3(x_size, y_size, z_size)) {
        volume->force_allocation();
        volume->clear(0.0f, 0);
        return;
    }
    volume = std::make_shared<Texture3D>(x_size, y_size, z_size, ColorTextureFormat::R32F);
}

// FBP3D.cpp:30-42
void FBP3D::generate_blank_projections(int x_size, int y_size, int projection_amount)
{
    if (projections != nullptr &&
        projections->get_size() == glm::ivec3(x_size, y_size, projection_amount)) {
        projections->force_allocation();
        projections->clear(0.0f, 0);
        return;
    }
    projections = std::make_shared<Texture2DArray>(x_size, y_size, projection_amount,
                                                   ColorTextureFormat::R32F);
}

// FBP3D.cpp:44-56: sinograms are (nu, N, nv)
void FBP3D::generate_blank_sinograms(int x_size, int y_size, int projection_amount)
{
    if (sinograms != nullptr &&
        sinograms->get_size() == glm::ivec3(x_size, projection_amount, y_size)) {
        sinograms->force_allocation();
        sinograms->clear(0.0f, 0);
        return;
    }
    sinograms = std::make_shared<Texture3D>(x_size, projection_amount, y_size,
                                            ColorTextureFormat::R32F);
}

static void require(bool ok, const std::string &what)
{
    if (!ok) raise_error("[CTAnalyzer Error] " + what);
}

// FBP3D.cpp:609-628: centre (c/2 + 1/2) * size and semi-axes (a/2) * size
// in voxel units; rendered with voxel pitch 1.
void FBP3D::render_ellipsoid(glm::vec3 center, glm::vec3 axis_lengths,
                             glm::vec3 rotation_angles_radian, float value)
{
    require(volume != nullptr, "FBP3D tried to render_ellipsoid() but FBP3D.volume was nullptr");
    glm::ivec3 s = volume->get_size();
    double e[10] = {center.x / 2.0 * s.x, center.y / 2.0 * s.y, center.z / 2.0 * s.z,
                    axis_lengths.x / 2.0 * s.x, axis_lengths.y / 2.0 * s.y,
                    axis_lengths.z / 2.0 * s.z, rotation_angles_radian.x,
                    rotation_angles_radian.y, rotation_angles_radian.z, value};
    check_status(ctr_render_ellipsoids(get_stream(), nullptr, e, 1, volume->data(), s.x, s.y, s.z,
                                       (double)s.x),
                 "FBP3D::render_ellipsoid");
}

// FBP3D.cpp:595-607
void FBP3D::render_sphere(glm::vec3 center, float radius, float value)
{
    require(volume != nullptr, "FBP3D tried to render_sphere() but FBP3D.volume was nullptr");
    glm::ivec3 s = volume->get_size();
    double rv = radius / 2.0 * s.x;
    double e[10] = {center.x / 2.0 * s.x, center.y / 2.0 * s.y, center.z / 2.0 * s.z,
                    rv, rv, rv, 0.0, 0.0, 0.0, value};
    check_status(ctr_render_ellipsoids(get_stream(), nullptr, e, 1, volume->data(), s.x, s.y, s.z,
                                       (double)s.x),
                 "FBP3D::render_sphere");
}

// FBP3D.cpp:58-72: the 3D Shepp-Logan table of the reference (degrees
// converted with 3.14/180 as there).
void FBP3D::generate_shepplogan(int x_size, int y_size, int z_size)
{
    generate_blank_volume(x_size, y_size, z_size);
    const float k = 3.14f / 180.0f;
    struct Row { float c[3], a[3], deg, v; };
    const Row rows[10] = {
        {{0, 0, 0}, {0.69f, 0.9f, 0.92f}, 0, 2.0f / 3.0f},
        {{0, 0, 0}, {0.6624f, 0.88f, 0.874f}, 0, -0.98f / 3.0f},
        {{-0.22f, -0.25f, 0}, {0.41f, 0.21f, 0.16f}, 108, -0.02f * 75 / 3.0f},
        {{0.22f, -0.25f, 0}, {0.31f, 0.22f, 0.11f}, 72, -0.02f * 75 / 3.0f},
        {{0, -0.25f, 0.4f}, {0.23f, 0.46f, 0.23f}, 0, 0.02f * 75 / 3.0f},
        {{0, -0.25f, 0.1f}, {0.046f, 0.046f, 0.046f}, 0, 0.02f * 75 / 3.0f},
        {{-0.8f, -0.25f, -0.65f}, {0.046f, 0.02f, 0.023f}, 0, 0.01f * 75 / 3.0f},
        {{0.06f, -0.25f, -0.065f}, {0.046f, 0.02f, 0.023f}, 90, 0.01f * 75 / 3.0f},
        {{0.06f, 0.625f, -0.105f}, {0.56f, 0.1f, 0.04f}, 90, 0.02f * 75 / 3.0f},
        {{0, -0.625f, 0.1f}, {0.056f, 0.1f, 0.056f}, 0, -0.02f * 75 / 3.0f},
    };
    for (const Row &r : rows)
        render_ellipsoid(glm::vec3(r.c[0], r.c[1], r.c[2]), glm::vec3(r.a[0], r.a[1], r.a[2]),
                         glm::vec3(0.0f, r.deg * k, 0.0f), r.v);
}

// FBP3D.cpp:197-210
void FBP3D::free_volume() { volume = nullptr; }
void FBP3D::free_projections() { projections = nullptr; }
void FBP3D::free_sinograms() { sinograms = nullptr; }

// FBP3D.cpp:212-240: detector (X, Y), pitch one voxel
void FBP3D::project_forward_parallel_to_projections(float full_rotation_count,
                                                    int projection_count,
                                                    float rotation_offset_radian)
{
    require(volume != nullptr,
            "FBP3D tried to project_forward_parallel() but FBP3D.volume was nullptr");
    glm::ivec3 v = volume->get_size();
    generate_blank_projections(v.x, v.y, projection_count);
    check_status(ctr_forward_project(get_stream(), nullptr, CTR_GEOMETRY_PARALLEL, volume->data(), v.x,
                                     v.y, v.z, 0.0, 0.0, (double)v.x, (double)v.x,
                                     full_rotation_count, rotation_offset_radian,
                                     projections->data(), CTR_LAYOUT_PROJECTIONS, v.x, v.y,
                                     projection_count, 1.0),
                 "FBP3D::project_forward_parallel_to_projections");
}

// FBP3D.cpp:242-267
void FBP3D::project_forward_parallel_to_sinogram(float full_rotation_count, int projection_count,
                                                 float rotation_offset_radian)
{
    require(volume != nullptr,
            "FBP3D tried to project_forward_parallel_to_sinogram() but FBP3D.volume was nullptr");
    glm::ivec3 v = volume->get_size();
    generate_blank_sinograms(v.x, v.y, projection_count);
    check_status(ctr_forward_project(get_stream(), nullptr, CTR_GEOMETRY_PARALLEL, volume->data(), v.x,
                                     v.y, v.z, 0.0, 0.0, (double)v.x, (double)v.x,
                                     full_rotation_count, rotation_offset_radian,
                                     sinograms->data(), CTR_LAYOUT_SINOGRAMS, v.x, v.y,
                                     projection_count, 1.0),
                 "FBP3D::project_forward_parallel_to_sinogram");
}

// FBP3D.cpp:269-299 (detector defaults to the volume's (X, Y) as at :276)
void FBP3D::project_forward_cone_to_projections(float xray_source_distance,
                                                float detector_panel_distance,
                                                float detector_width, float volume_width,
                                                float full_rotation_count, int projection_count,
                                                float rotation_offset_radian,
                                                int detector_resolution_u,
                                                int detector_resolution_v)
{
    require(volume != nullptr, "FBP3D tried to project_forward_cone_to_projections() but "
                               "FBP3D.volume was nullptr");
    glm::ivec3 v = volume->get_size();
    int nu = detector_resolution_u > 0 ? detector_resolution_u : v.x;
    int nv = detector_resolution_v > 0 ? detector_resolution_v : v.y;
    generate_blank_projections(nu, nv, projection_count);
    check_status(ctr_forward_project(get_stream(), nullptr, CTR_GEOMETRY_CONE, volume->data(), v.x, v.y,
                                     v.z, xray_source_distance, detector_panel_distance,
                                     detector_width, volume_width, full_rotation_count,
                                     rotation_offset_radian, projections->data(),
                                     CTR_LAYOUT_PROJECTIONS, nu, nv, projection_count, 1.0),
                 "FBP3D::project_forward_cone_to_projections");
}

// FBP3D.cpp:301-330
void FBP3D::project_forward_cone_to_sinogram(float xray_source_distance,
                                             float detector_panel_distance, float detector_width,
                                             float volume_width, float full_rotation_count,
                                             int projection_count, float rotation_offset_radian,
                                             int detector_resolution_u, int detector_resolution_v)
{
    require(volume != nullptr, "FBP3D tried to project_forward_cone_to_sinogram() but "
                               "FBP3D.volume was nullptr");
    glm::ivec3 v = volume->get_size();
    int nu = detector_resolution_u > 0 ? detector_resolution_u : v.x;
    int nv = detector_resolution_v > 0 ? detector_resolution_v : v.y;
    generate_blank_sinograms(nu, nv, projection_count);
    check_status(ctr_forward_project(get_stream(), nullptr, CTR_GEOMETRY_CONE, volume->data(), v.x, v.y,
                                     v.z, xray_source_distance, detector_panel_distance,
                                     detector_width, volume_width, full_rotation_count,
                                     rotation_offset_radian, sinograms->data(),
                                     CTR_LAYOUT_SINOGRAMS, nu, nv, projection_count, 1.0),
                 "FBP3D::project_forward_cone_to_sinogram");
}

// FBP3D.cpp:333-349: p = -ln(I / base); intensities below FLT_MIN clamp there
void FBP3D::log_normalize_projections(float base_value)
{
    require(projections != nullptr, "FBP3D tried to log_normalize_projections() but "
                                    "FBP3D.projections was nullptr");
    check_status(ctr_log_normalize(get_stream(), projections->data(), projections->texel_count(),
                                   base_value, FLT_MIN),
                 "FBP3D::log_normalize_projections");
}

// FBP3D.cpp:351-367 (the reference sized this by `projections`; here it
// normalises the sinograms themselves)
void FBP3D::log_normalize_sinograms(float base_value)
{
    require(sinograms != nullptr,
            "FBP3D tried to log_normalize_sinograms() but FBP3D.sinograms was nullptr");
    check_status(ctr_log_normalize(get_stream(), sinograms->data(), sinograms->texel_count(),
                                   base_value, FLT_MIN),
                 "FBP3D::log_normalize_sinograms");
}

// FBP3D.cpp:369-387
void FBP3D::apply_fdk_weights_to_projections(float xray_source_distance,
                                             float detector_panel_distance, float detector_width)
{
    require(projections != nullptr, "FBP3D tried to apply_fdk_weights_to_projections() but "
                                    "FBP3D.projections was nullptr");
    glm::ivec3 p = projections->get_size();
    check_status(ctr_fdk_weight(get_stream(), projections->data(), CTR_LAYOUT_PROJECTIONS, p.x,
                                p.y, p.z, xray_source_distance, detector_panel_distance,
                                detector_width),
                 "FBP3D::apply_fdk_weights_to_projections");
}

// declared at FBP3D.h:53 but never defined in the reference
void FBP3D::apply_fdk_weights_to_sinograms(float xray_source_distance,
                                           float detector_panel_distance, float detector_width)
{
    require(sinograms != nullptr, "FBP3D tried to apply_fdk_weights_to_sinograms() but "
                                  "FBP3D.sinograms was nullptr");
    glm::ivec3 s = sinograms->get_size();  // (nu, N, nv)
    check_status(ctr_fdk_weight(get_stream(), sinograms->data(), CTR_LAYOUT_SINOGRAMS, s.x, s.z,
                                s.y, xray_source_distance, detector_panel_distance,
                                detector_width),
                 "FBP3D::apply_fdk_weights_to_sinograms");
}

// FBP3D.cpp:389-414: every detector row, in place
void FBP3D::apply_filter_to_projections(FBP2D::FilterType filter_type)
{
    require(projections != nullptr, "FBP3D tried to apply_filter_to_projections() but "
                                    "FBP3D.projections was nullptr");
    glm::ivec3 p = projections->get_size();
    check_status(ctr_filter_rows(get_stream(), projections->data(), projections->data(), p.x,
                                 p.y * p.z, fft_length, (int)filter_type),
                 "FBP3D::apply_filter_to_projections");
    filtered_fft_length_ = fft_length ? fft_length : ceil_pow2_i(p.x);
}

// FBP3D.cpp:417-442
void FBP3D::apply_filter_to_sinograms(FBP2D::FilterType filter_type)
{
    require(sinograms != nullptr,
            "FBP3D tried to apply_filter_to_sinograms() but FBP3D.sinograms was nullptr");
    glm::ivec3 s = sinograms->get_size();
    check_status(ctr_filter_rows(get_stream(), sinograms->data(), sinograms->data(), s.x,
                                 s.y * s.z, fft_length, (int)filter_type),
                 "FBP3D::apply_filter_to_sinograms");
    filtered_fft_length_ = fft_length ? fft_length : ceil_pow2_i(s.x);
}

// FBP3D.cpp:444-458
void FBP3D::clip_negatives_of_volume()
{
    require(volume != nullptr,
            "FBP3D tried to clip_negatives_of_volume() but FBP3D.volume was nullptr");
    check_status(ctr_clip_negative(get_stream(), volume->data(), volume->texel_count()),
                 "FBP3D::clip_negatives_of_volume");
}

// Shared body of the four 3D backprojections.  Scale: (pi / N) for parallel
// and (pi / N) * SDD / SOD for cone (Feldkamp's discretisation of the
// angular integral), divided by fft_length * detector pitch, which undoes
// the factor n in the filter gain n * g(i/n - 1/2) and converts the
// per-sample ramp to physical units.
void FBP3D::backproject(int geometry, const float *data, int layout, int nu, int nv, int n,
                        float sod, float odd, float detector_width, float volume_width,
                        float rotation, float offset, int H, int V)
{
    generate_blank_volume(H, V, H);
    size_t bytes = 0;
    check_status(ctr_staged_layout(nu, nv, n, staged_dtype, nullptr, nullptr, &bytes),
                 "FBP3D backprojection");
    if (staged_.bytes() < bytes) staged_.resize(bytes);
    if (view_scale_.bytes() < sizeof(float) * n) view_scale_.resize(sizeof(float) * n);
    float *vs = staged_dtype == CTR_F32 ? nullptr : (float *)view_scale_.get();
    check_status(ctr_stage_projections(get_stream(), nullptr, data, layout, nu, nv, n,
                                       staged_.get(), staged_dtype, vs),
                 "FBP3D backprojection");
    int nfft = filtered_fft_length_ ? filtered_fft_length_ : ceil_pow2_i(nu);
    double delta = (double)detector_width / nu;
    double scale = M_PI / n / (nfft * delta);
    if (geometry == CTR_GEOMETRY_CONE) scale *= ((double)sod + odd) / sod;
    check_status(ctr_backproject(get_stream(), nullptr, geometry, staged_.get(), staged_dtype, vs,
                                 nu, nv, n, sod, odd, detector_width, volume_width, rotation,
                                 offset, volume->data(), H, V, H, (float)scale, 0, 1),
                 "FBP3D backprojection");
}

// FBP3D.cpp:460-490: detector plane width = X voxels, volume width = X
void FBP3D::project_backward_parallel_from_projections(float full_rotation_count,
                                                       int volume_resolution_horizontal,
                                                       int volume_resolution_vertical,
                                                       float rotation_offset_radian)
{
    require(projections != nullptr, "FBP3D tried to project_backward_parallel_from_projections() "
                                    "but FBP3D.projections was nullptr");
    glm::ivec3 p = projections->get_size();
    float w = (float)volume_resolution_horizontal;
    backproject(CTR_GEOMETRY_PARALLEL, projections->data(), CTR_LAYOUT_PROJECTIONS, p.x, p.y, p.z,
                0.0f, 0.0f, w, w, full_rotation_count, rotation_offset_radian,
                volume_resolution_horizontal, volume_resolution_vertical);
}

// FBP3D.cpp:492-522
void FBP3D::project_backward_parallel_from_sinograms(float full_rotation_count,
                                                     int volume_resolution_horizontal,
                                                     int volume_resolution_vertical,
                                                     float rotation_offset_radian)
{
    require(sinograms != nullptr, "FBP3D tried to project_backward_parallel_from_sinograms() but "
                                  "FBP3D.sinograms was nullptr");
    glm::ivec3 s = sinograms->get_size();  // (nu, N, nv)
    float w = (float)volume_resolution_horizontal;
    backproject(CTR_GEOMETRY_PARALLEL, sinograms->data(), CTR_LAYOUT_SINOGRAMS, s.x, s.z, s.y,
                0.0f, 0.0f, w, w, full_rotation_count, rotation_offset_radian,
                volume_resolution_horizontal, volume_resolution_vertical);
}

// FBP3D.cpp:524-559
void FBP3D::project_backward_cone_fdk_from_projections(
    float xray_source_distance, float detector_panel_distance, float detector_width,
    float volume_width, float full_rotation_count, int volume_resolution_horizontal,
    int volume_resolution_vertical, float rotation_offset_radian)
{
    require(projections != nullptr, "FBP3D tried to project_backward_cone_fdk() but "
                                    "FBP3D.projections was nullptr");
    glm::ivec3 p = projections->get_size();
    backproject(CTR_GEOMETRY_CONE, projections->data(), CTR_LAYOUT_PROJECTIONS, p.x, p.y, p.z,
                xray_source_distance, detector_panel_distance, detector_width, volume_width,
                full_rotation_count, rotation_offset_radian, volume_resolution_horizontal,
                volume_resolution_vertical);
}

// FBP3D.cpp:561-593
void FBP3D::project_backward_cone_fdk_from_sinograms(
    float xray_source_distance, float detector_panel_distance, float detector_width,
    float volume_width, float full_rotation_count, int volume_resolution_horizontal,
    int volume_resolution_vertical, float rotation_offset_radian)
{
    require(sinograms != nullptr, "FBP3D tried to project_backward_cone_fdk_from_sinograms() but "
                                  "FBP3D.sinograms was nullptr");
    glm::ivec3 s = sinograms->get_size();
    backproject(CTR_GEOMETRY_CONE, sinograms->data(), CTR_LAYOUT_SINOGRAMS, s.x, s.z, s.y,
                xray_source_distance, detector_panel_distance, detector_width, volume_width,
                full_rotation_count, rotation_offset_radian, volume_resolution_horizontal,
                volume_resolution_vertical);
}

// ---- slice and view copies (FBP3D.cpp:639-778) ----------------------------

void FBP3D::load_volume_slice_x(int x_index, Texture2D &t)
{
    glm::ivec3 v = volume->get_size();
    glm::ivec2 s = t.get_size();
    require(x_index >= 0 && x_index < v.x, "load_volume_slice_x: index out of range");
    // rows j of all z-planes, one element each: texel (j, k)
    if (s.x == v.y && s.y == v.z)
        copy_block(t.data(), 1, volume->data() + x_index, (size_t)v.x, 1, (size_t)v.y * v.z);
    else
        for (int k = 0; k < std::min(s.y, v.z); k++)
            copy_block(t.data() + (size_t)k * s.x, 1,
                       volume->data() + (size_t)k * v.x * v.y + x_index, (size_t)v.x, 1,
                       (size_t)std::min(s.x, v.y));
}

void FBP3D::store_volume_slice_x(Texture2D &t, int x_index)
{
    glm::ivec3 v = volume->get_size();
    glm::ivec2 s = t.get_size();
    require(x_index >= 0 && x_index < v.x, "store_volume_slice_x: index out of range");
    for (int k = 0; k < std::min(s.y, v.z); k++)
        copy_block(volume->data() + (size_t)k * v.x * v.y + x_index, (size_t)v.x,
                   t.data() + (size_t)k * s.x, 1, 1, (size_t)std::min(s.x, v.y));
}

void FBP3D::load_volume_slice_y(int y_index, Texture2D &t)
{
    glm::ivec3 v = volume->get_size();
    glm::ivec2 s = t.get_size();
    require(y_index >= 0 && y_index < v.y, "load_volume_slice_y: index out of range");
    copy_block(t.data(), (size_t)s.x, volume->data() + (size_t)y_index * v.x,
               (size_t)v.x * v.y, (size_t)std::min(s.x, v.x), (size_t)std::min(s.y, v.z));
}

void FBP3D::store_volume_slice_y(Texture2D &t, int y_index)
{
    glm::ivec3 v = volume->get_size();
    glm::ivec2 s = t.get_size();
    require(y_index >= 0 && y_index < v.y, "store_volume_slice_y: index out of range");
    copy_block(volume->data() + (size_t)y_index * v.x, (size_t)v.x * v.y, t.data(), (size_t)s.x,
               (size_t)std::min(s.x, v.x), (size_t)std::min(s.y, v.z));
}

void FBP3D::load_volume_slice_z(int z_index, Texture2D &t)
{
    glm::ivec3 v = volume->get_size();
    glm::ivec2 s = t.get_size();
    require(z_index >= 0 && z_index < v.z, "load_volume_slice_z: index out of range");
    copy_block(t.data(), (size_t)s.x, volume->data() + (size_t)z_index * v.x * v.y, (size_t)v.x,
               (size_t)std::min(s.x, v.x), (size_t)std::min(s.y, v.y));
}

void FBP3D::store_volume_slice_z(Texture2D &t, int z_index)
{
    glm::ivec3 v = volume->get_size();
    glm::ivec2 s = t.get_size();
    require(z_index >= 0 && z_index < v.z, "store_volume_slice_z: index out of range");
    copy_block(volume->data() + (size_t)z_index * v.x * v.y, (size_t)v.x, t.data(), (size_t)s.x,
               (size_t)std::min(s.x, v.x), (size_t)std::min(s.y, v.y));
}

void FBP3D::load_projection(int projection_index, Texture2D &t)
{
    glm::ivec3 p = projections->get_size();
    glm::ivec2 s = t.get_size();
    require(projection_index >= 0 && projection_index < p.z, "load_projection: index out of range");
    copy_block(t.data(), (size_t)s.x, projections->data() + (size_t)projection_index * p.x * p.y,
               (size_t)p.x, (size_t)std::min(s.x, p.x), (size_t)std::min(s.y, p.y));
}

void FBP3D::store_projection(Texture2D &t, int projection_index)
{
    glm::ivec3 p = projections->get_size();
    glm::ivec2 s = t.get_size();
    require(projection_index >= 0 && projection_index < p.z, "store_projection: index out of range");
    copy_block(projections->data() + (size_t)projection_index * p.x * p.y, (size_t)p.x, t.data(),
               (size_t)s.x, (size_t)std::min(s.x, p.x), (size_t)std::min(s.y, p.y));
}

// sinogram `y_coord` = detector row: a contiguous (nu, N) plane
void FBP3D::load_sinogram(int y_coord, Texture2D &t)
{
    glm::ivec3 g = sinograms->get_size();
    glm::ivec2 s = t.get_size();
    require(y_coord >= 0 && y_coord < g.z, "load_sinogram: index out of range");
    copy_block(t.data(), (size_t)s.x, sinograms->data() + (size_t)y_coord * g.x * g.y,
               (size_t)g.x, (size_t)std::min(s.x, g.x), (size_t)std::min(s.y, g.y));
}

void FBP3D::store_sinogram(Texture2D &t, int y_coord)
{
    glm::ivec3 g = sinograms->get_size();
    glm::ivec2 s = t.get_size();
    require(y_coord >= 0 && y_coord < g.z, "store_sinogram: index out of range");
    copy_block(sinograms->data() + (size_t)y_coord * g.x * g.y, (size_t)g.x, t.data(), (size_t)s.x,
               (size_t)std::min(s.x, g.x), (size_t)std::min(s.y, g.y));
}

// view n of the sinogram layout: rows of nu at stride nu * N
void FBP3D::load_projection_from_sinograms(int projection_index, Texture2D &t)
{
    glm::ivec3 g = sinograms->get_size();
    glm::ivec2 s = t.get_size();
    require(projection_index >= 0 && projection_index < g.y,
            "load_projection_from_sinograms: index out of range");
    copy_block(t.data(), (size_t)s.x, sinograms->data() + (size_t)projection_index * g.x,
               (size_t)g.x * g.y, (size_t)std::min(s.x, g.x), (size_t)std::min(s.y, g.z));
}

void FBP3D::store_projection_to_sinograms(Texture2D &t, int projection_index)
{
    glm::ivec3 g = sinograms->get_size();
    glm::ivec2 s = t.get_size();
    require(projection_index >= 0 && projection_index < g.y,
            "store_projection_to_sinograms: index out of range");
    copy_block(sinograms->data() + (size_t)projection_index * g.x, (size_t)g.x * g.y, t.data(),
               (size_t)s.x, (size_t)std::min(s.x, g.x), (size_t)std::min(s.y, g.z));
}

}  // namespace ctrecon
