// FBP3D.h -- drop-in 3D CT driver over the gfx950 library.
//
// Same class, public methods, parameter order, default arguments and public
// data members (fbp_solver, fft_solver, volume, projections, sinograms) as
// the reference's Source/FilteredBackProjection/FBP3D.h (:16-170); copy
// construction is deleted as there (:19).  Layouts follow the reference:
// volume Texture3D (X, Y, Z) with Y the rotation axis, projections
// Texture2DArray (nu, nv, N), sinograms Texture3D (nu, N, nv)
// (FBP3D.cpp:26,40,54).
//
// Differences, all additive: every 3D backprojection runs as ONE launch over
// all views instead of one dispatch per view; `fft_length` (default 0 = the
// reference's 2^ceil(log2 nu)) lets callers pick the zero-padded FFT length;
// `staged_dtype` selects FP32 or FP16/BF16 filtered-projection storage for
// the backprojector; the cone forward projector takes an optional explicit
// detector size.  Projection ingest, slice output and the volume range
// helpers (FBP3D.cpp:74-195,780-826) decode / convert on the device (see
// fbp3d_io.cpp); blur_volume, the read_volume / write_projections /
// read_sinograms / write_sinograms declarations the reference never defines
// and the ComputeProgram members have no counterpart.
#pragma once

#include <memory>
#include <string>

#include "FBP2D.h"
#include "FastFourierTransformGPU.h"
#include "Texture.h"

namespace ctrecon {

class FBP3D {
public:
    FBP3D(const FBP3D &other) = delete;
    FBP3D();

    void generate_blank_volume(int x_size, int y_size, int z_size);
    void generate_blank_projections(int x_size, int y_size, int projection_amount);
    void generate_blank_sinograms(int x_size, int y_size, int projection_amount);

    void render_sphere(glm::vec3 center, float radius, float value);
    void render_ellipsoid(glm::vec3 center, glm::vec3 axis_lengths,
                          glm::vec3 rotation_angles_radian, float value);
    void generate_shepplogan(int x_size, int y_size, int z_size);

    // FBP3D.cpp:74-176: frames named "...NNNN.ext" (NNNN = projection index),
    // channel 0 of file_channel_count interleaved channels of
    // byte_per_channel (1 or 2) bytes, resampled to the target size.
    void read_projections(const std::string &upper_diretory, int file_width, int file_height,
                          int file_channel_count, int byte_per_channel,
                          int target_projection_width, int target_projection_height,
                          int projection_count);
    void read_projections_as_sinograms(const std::string &upper_diretory, int file_width,
                                       int file_height, int file_channel_count,
                                       int byte_per_channel, int target_projection_width,
                                       int target_projection_height, int projection_count);
    // FBP3D.cpp:178-195: one raw unorm16 file per y-slice.
    void write_volume(std::string upper_directory);

    void free_volume();
    void free_projections();
    void free_sinograms();

    void project_forward_parallel_to_projections(float full_rotation_count, int projection_count,
                                                 float rotation_offset_radian = 0.0f);
    void project_forward_cone_to_projections(float xray_source_distance,
                                             float detector_panel_distance, float detector_width,
                                             float volume_width, float full_rotation_count,
                                             int projection_count,
                                             float rotation_offset_radian = 0.0f,
                                             int detector_resolution_u = 0,
                                             int detector_resolution_v = 0);
    void project_forward_parallel_to_sinogram(float full_rotation_count, int projection_count,
                                              float rotation_offset_radian = 0.0f);
    void project_forward_cone_to_sinogram(float xray_source_distance,
                                          float detector_panel_distance, float detector_width,
                                          float volume_width, float full_rotation_count,
                                          int projection_count,
                                          float rotation_offset_radian = 0.0f,
                                          int detector_resolution_u = 0,
                                          int detector_resolution_v = 0);

    void log_normalize_projections(float base_value);
    void apply_fdk_weights_to_projections(float xray_source_distance,
                                          float detector_panel_distance, float detector_width);
    void apply_filter_to_projections(FBP2D::FilterType filter_type);
    void log_normalize_sinograms(float base_value);
    void apply_fdk_weights_to_sinograms(float xray_source_distance, float detector_panel_distance,
                                        float detector_width);
    void apply_filter_to_sinograms(FBP2D::FilterType filter_type);
    void clip_negatives_of_volume();

    void project_backward_parallel_from_projections(float full_rotation_count,
                                                    int volume_resolution_horizontal,
                                                    int volume_resolution_vertical,
                                                    float rotation_offset_radian = 0.0f);
    void project_backward_cone_fdk_from_projections(float xray_source_distance,
                                                    float detector_panel_distance,
                                                    float detector_width, float volume_width,
                                                    float full_rotation_count,
                                                    int volume_resolution_horizontal,
                                                    int volume_resolution_vertical,
                                                    float rotation_offset_radian = 0.0f);
    void project_backward_parallel_from_sinograms(float full_rotation_count,
                                                  int volume_resolution_horizontal,
                                                  int volume_resolution_vertical,
                                                  float rotation_offset_radian = 0.0f);
    void project_backward_cone_fdk_from_sinograms(float xray_source_distance,
                                                  float detector_panel_distance,
                                                  float detector_width, float volume_width,
                                                  float full_rotation_count,
                                                  int volume_resolution_horizontal,
                                                  int volume_resolution_vertical,
                                                  float rotation_offset_radian = 0.0f);

    // Slice copies (FBP3D.cpp:639-704).  An x slice is laid out (Y, Z), a y
    // slice (X, Z), a z slice (X, Y); the copied extent is the min of both.
    void load_volume_slice_x(int x_index, Texture2D &target_texture_out);
    void load_volume_slice_y(int y_index, Texture2D &target_texture_out);
    void load_volume_slice_z(int z_index, Texture2D &target_texture_out);
    void store_volume_slice_x(Texture2D &source_texture, int x_index);
    void store_volume_slice_y(Texture2D &source_texture, int y_index);
    void store_volume_slice_z(Texture2D &source_texture, int z_index);

    void load_projection(int projection_index, Texture2D &target_texture_out);
    void store_projection(Texture2D &source_texture, int projection_index);

    void load_sinogram(int projection_index, Texture2D &target_texture_out);
    void store_sinogram(Texture2D &source_texture, int projection_index);

    void load_projection_from_sinograms(int projection_index, Texture2D &target_texture_out);
    void store_projection_to_sinograms(Texture2D &soruce_texture, int projection_index);

    // FBP3D.cpp:780-826: 1-texel textures carry the value's float bits.
    void compute_max_value_of_volume(Texture1D &target_int_texture);
    void compute_min_value_of_volume(Texture1D &target_int_texture);
    void normalize_histogram();
    void normalize_histogram(Texture1D &source_min_int_texture,
                             Texture1D &source_max_int_texture);

    std::shared_ptr<FBP2D> fbp_solver = std::make_shared<FBP2D>();
    std::shared_ptr<FFFT> fft_solver = std::make_shared<FFFT>();

    std::shared_ptr<Texture3D> volume = nullptr;
    std::shared_ptr<Texture2DArray> projections = nullptr;
    std::shared_ptr<Texture3D> sinograms = nullptr;

    // Additions (no reference counterpart).
    int fft_length = 0;            // zero-padded FFT length, 0 = 2^ceil(log2 nu)
    int staged_dtype = 0;          // CTR_F32 / CTR_F16 / CTR_BF16 backprojector input

private:
    void backproject(int geometry, const float *data, int layout, int nu, int nv, int n,
                     float sod, float odd, float detector_width, float volume_width,
                     float rotation, float offset, int H, int V);
    int filtered_fft_length_ = 0;  // FFT length of the last apply_filter_* call
    DeviceBuffer staged_;
    DeviceBuffer view_scale_;
    DeviceBuffer minmax_;
    static constexpr size_t kSliceIndexDigits = 6;  // FBP3D.h:161
};

}  // namespace ctrecon
