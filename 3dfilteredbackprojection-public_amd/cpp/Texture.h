// Texture.h -- HIP device-memory stand-ins for the GraphicsCortex texture
// types the reference's FBP2D / FBP3D / FFFT classes are written against
// (Texture1D, Texture2D, Texture2DArray, Texture3D; used throughout
// Source/FilteredBackProjection/FBP3D.cpp and
// Source/FastFourierTransformGPU/FastFourierTransformGPU.cpp).
//
// They keep the member names those classes call -- get_size(),
// create_texture_with_same_parameters(), clear(), force_allocation(),
// is_bindless, load_data(), get_image() -- so call sites compile unchanged.
// Storage is linear device memory, x fastest, `channels` floats (or uint32
// for R32UI) per texel.  There is no GL, no sampler and no mipmapping.
#pragma once

#include <cstddef>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "glm_lite.h"

namespace ctrecon {

// "[CTAnalyzer Error] ..." + ASSERT(false) of the reference become a console
// line and this exception.
struct Error : std::runtime_error {
    explicit Error(const std::string &m) : std::runtime_error(m) {}
};
[[noreturn]] void raise_error(const std::string &message);
// Checks a ctr_* status code; raises with ctr_last_error() text on failure.
void check_status(int rc, const char *what);

// Stream the calling thread's classes enqueue on (hipStream_t as void*,
// NULL = default); each thread keeps its own.
void set_stream(void *stream);
void *get_stream();

class DeviceBuffer {
public:
    DeviceBuffer() = default;
    explicit DeviceBuffer(size_t bytes);
    ~DeviceBuffer();
    DeviceBuffer(const DeviceBuffer &) = delete;
    DeviceBuffer &operator=(const DeviceBuffer &) = delete;
    void resize(size_t bytes);  // contents are not preserved
    void *get() const { return ptr_; }
    size_t bytes() const { return bytes_; }

private:
    void *ptr_ = nullptr;
    size_t bytes_ = 0;
};

enum class ColorTextureFormat { R32F, RG32F, RGBA32F, R16, R32UI };
enum class ColorFormat { RED, RG, RGB, RGBA };
enum class PixelType { FLOAT, UNSIGNED_BYTE, UNSIGNED_SHORT, UNSIGNED_INT };

int channels_of(ColorTextureFormat f);

// Host-side image, the stand-in for GraphicsCortex `Image` in get_image().
struct HostImage {
    int width = 0, height = 0, channels = 1;
    std::vector<float> data;
};

class TextureStorage {
public:
    bool is_bindless = false;
    float *data() const { return static_cast<float *>(buf_.get()); }
    size_t texel_count() const { return texels_; }
    int channels() const { return channels_; }
    ColorTextureFormat format() const { return format_; }
    void force_allocation() {}
    // Upload / download the whole texture as float channels.
    void upload(const float *host);
    void download(float *host) const;

protected:
    void allocate(size_t texels, ColorTextureFormat f);
    void clear_all(float v);
    DeviceBuffer buf_;
    size_t texels_ = 0;
    int channels_ = 1;
    ColorTextureFormat format_ = ColorTextureFormat::R32F;
};

class Texture1D : public TextureStorage {
public:
    using ColorTextureFormat = ctrecon::ColorTextureFormat;
    Texture1D(int width, ColorTextureFormat f, int mipmap_levels = 1, float anisotropy = 0.0f);
    int get_size() const { return width_; }
    void clear(float v, int mip = 0) { (void)mip; clear_all(v); }

private:
    int width_;
};

class Texture2D : public TextureStorage {
public:
    using ColorTextureFormat = ctrecon::ColorTextureFormat;
    using ColorFormat = ctrecon::ColorFormat;
    using Type = ctrecon::PixelType;
    Texture2D(int width, int height, ColorTextureFormat f, int mipmap_levels = 1,
              float anisotropy = 0.0f, int multisample = 0);
    glm::ivec2 get_size() const { return size_; }
    std::shared_ptr<Texture2D> create_texture_with_same_parameters() const;
    void clear(float v, int mip = 0) { (void)mip; clear_all(v); }
    void clear(glm::vec2 v, int mip = 0) { (void)mip; clear_all(v.x); }
    void clear(glm::vec4 v, int mip = 0) { (void)mip; clear_all(v.x); }
    // Texture2D::load_data(image, format, type, mip): host texels are
    // converted to float (unsigned types normalised to [0, 1], as GL does).
    void load_data(const void *host, ColorFormat fmt, PixelType type, int mip = 0);
    std::shared_ptr<HostImage> get_image(ColorFormat fmt, PixelType type, int mip = 0) const;

private:
    glm::ivec2 size_;
};

class Texture2DArray : public TextureStorage {
public:
    using ColorTextureFormat = ctrecon::ColorTextureFormat;
    Texture2DArray(int width, int height, int layers, ColorTextureFormat f,
                   int mipmap_levels = 1, float anisotropy = 0.0f, int multisample = 0);
    glm::ivec3 get_size() const { return size_; }
    void clear(float v, int mip = 0) { (void)mip; clear_all(v); }

private:
    glm::ivec3 size_;
};

class Texture3D : public TextureStorage {
public:
    using ColorTextureFormat = ctrecon::ColorTextureFormat;
    Texture3D(int width, int height, int depth, ColorTextureFormat f, int mipmap_levels = 1,
              float anisotropy = 0.0f);
    glm::ivec3 get_size() const { return size_; }
    void clear(float v, int mip = 0) { (void)mip; clear_all(v); }

private:
    glm::ivec3 size_;
};

// Copies a w x h float block between row-pitched device arrays (elements).
void copy_block(float *dst, size_t dst_pitch, const float *src, size_t src_pitch, size_t w,
                size_t h);

}  // namespace ctrecon
