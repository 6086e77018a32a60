// textures.cpp -- device-memory texture stand-ins (see Texture.h).
#include <hip/hip_runtime.h>

#include <cstring>
#include <iostream>

#include "Texture.h"
#include "ctrecon.h"

namespace ctrecon {

// Per thread, like the reference's one GL context per thread: each thread's
// classes enqueue on the stream that thread set.
static thread_local void *g_stream = nullptr;

void set_stream(void *stream) { g_stream = stream; }
void *get_stream() { return g_stream; }

void raise_error(const std::string &message)
{
    std::cout << message << std::endl;
    throw Error(message);
}

void check_status(int rc, const char *what)
{
    if (rc != CTR_OK)
        raise_error(std::string("[CTAnalyzer Error] ") + what + ": " + ctr_last_error());
}

static void hip_check(hipError_t e, const char *what)
{
    if (e != hipSuccess)
        raise_error(std::string("[CTAnalyzer Error] ") + what + ": " + hipGetErrorString(e));
}

DeviceBuffer::DeviceBuffer(size_t bytes) { resize(bytes); }

DeviceBuffer::~DeviceBuffer()
{
    if (ptr_) (void)hipFree(ptr_);
}

void DeviceBuffer::resize(size_t bytes)
{
    if (bytes == bytes_) return;
    if (ptr_) {
        hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
        hip_check(hipFree(ptr_), "hipFree");
        ptr_ = nullptr;
    }
    bytes_ = 0;
    if (bytes) {
        hip_check(hipMalloc(&ptr_, bytes), "hipMalloc");
        bytes_ = bytes;
    }
}

int channels_of(ColorTextureFormat f)
{
    switch (f) {
    case ColorTextureFormat::RG32F: return 2;
    case ColorTextureFormat::RGBA32F: return 4;
    default: return 1;
    }
}

void TextureStorage::allocate(size_t texels, ColorTextureFormat f)
{
    format_ = f;
    channels_ = channels_of(f);
    texels_ = texels;
    buf_.resize(texels * channels_ * sizeof(float));
    clear_all(0.0f);
}

void TextureStorage::clear_all(float v)
{
    check_status(ctr_fill(g_stream, data(), texels_ * channels_, v), "Texture::clear");
}

void TextureStorage::upload(const float *host)
{
    hip_check(hipMemcpyAsync(data(), host, texels_ * channels_ * sizeof(float),
                             hipMemcpyHostToDevice, (hipStream_t)g_stream),
              "texture upload");
    hip_check(hipStreamSynchronize((hipStream_t)g_stream), "texture upload");
}

void TextureStorage::download(float *host) const
{
    hip_check(hipMemcpyAsync(host, data(), texels_ * channels_ * sizeof(float),
                             hipMemcpyDeviceToHost, (hipStream_t)g_stream),
              "texture download");
    hip_check(hipStreamSynchronize((hipStream_t)g_stream), "texture download");
}

Texture1D::Texture1D(int width, ColorTextureFormat f, int, float) : width_(width)
{
    allocate((size_t)width, f);
}

Texture2D::Texture2D(int width, int height, ColorTextureFormat f, int, float, int)
    : size_(width, height)
{
    allocate((size_t)width * height, f);
}

std::shared_ptr<Texture2D> Texture2D::create_texture_with_same_parameters() const
{
    return std::make_shared<Texture2D>(size_.x, size_.y, format_);
}

void Texture2D::load_data(const void *host, ColorFormat fmt, PixelType type, int)
{
    int src_ch = fmt == ColorFormat::RED ? 1 : fmt == ColorFormat::RG ? 2
                 : fmt == ColorFormat::RGB ? 3 : 4;
    std::vector<float> t(texels_ * channels_, 0.0f);
    for (size_t p = 0; p < texels_; p++)
        for (int c = 0; c < channels_ && c < src_ch; c++) {
            size_t s = p * src_ch + c;
            float v;
            switch (type) {
            case PixelType::UNSIGNED_BYTE: v = ((const uint8_t *)host)[s] / 255.0f; break;
            case PixelType::UNSIGNED_SHORT: v = ((const uint16_t *)host)[s] / 65535.0f; break;
            case PixelType::UNSIGNED_INT: v = (float)(((const uint32_t *)host)[s] / 4294967295.0); break;
            default: v = ((const float *)host)[s]; break;
            }
            t[p * channels_ + c] = v;
        }
    upload(t.data());
}

std::shared_ptr<HostImage> Texture2D::get_image(ColorFormat fmt, PixelType, int) const
{
    auto img = std::make_shared<HostImage>();
    img->width = size_.x;
    img->height = size_.y;
    img->channels = fmt == ColorFormat::RED ? 1 : channels_;
    std::vector<float> t(texels_ * channels_);
    download(t.data());
    img->data.resize(texels_ * img->channels);
    for (size_t p = 0; p < texels_; p++)
        for (int c = 0; c < img->channels; c++) img->data[p * img->channels + c] = t[p * channels_ + c];
    return img;
}

Texture2DArray::Texture2DArray(int width, int height, int layers, ColorTextureFormat f, int,
                               float, int)
    : size_(width, height, layers)
{
    allocate((size_t)width * height * layers, f);
}

Texture3D::Texture3D(int width, int height, int depth, ColorTextureFormat f, int, float)
    : size_(width, height, depth)
{
    allocate((size_t)width * height * depth, f);
}

void copy_block(float *dst, size_t dst_pitch, const float *src, size_t src_pitch, size_t w,
                size_t h)
{
    if (w == 0 || h == 0) return;
    hip_check(hipMemcpy2DAsync(dst, dst_pitch * sizeof(float), src, src_pitch * sizeof(float),
                               w * sizeof(float), h, hipMemcpyDeviceToDevice,
                               (hipStream_t)g_stream),
              "copy_block");
}

}  // namespace ctrecon
