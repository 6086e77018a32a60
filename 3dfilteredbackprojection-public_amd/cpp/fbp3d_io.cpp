// fbp3d_io.cpp -- FBP3D's file formats either side of the reconstruction:
// projection ingest (FBP3D::read_projections, read_projections_as_sinograms,
// FBP3D.cpp:74-176), slice output (FBP3D::write_volume, FBP3D.cpp:178-195)
// and the volume range helpers (compute_{min,max}_value_of_volume,
// normalize_histogram, FBP3D.cpp:780-826).
//
// The host only moves bytes: files are read by a pool of 16 threads (the
// reference's OpenMP team size, FBP3D.cpp:107) in batches of 128 frames into
// pinned memory, copied to the device once per batch and decoded + resampled
// there by ctr_decode_frames; slices are converted to unorm16 on the device
// (ctr_slices_unorm16) and written one file per y-slice.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <filesystem>
#include <iostream>
#include <string>
#include <thread>
#include <vector>

#include "FBP3D.h"
#include "ctrecon.h"

namespace ctrecon {

namespace {

void hip_ok(hipError_t e, const char *what)
{
    if (e != hipSuccess)
        raise_error(std::string("[CTAnalyzer Error] ") + what + ": " + hipGetErrorString(e));
}

struct PinnedBuffer {
    void *ptr = nullptr;
    explicit PinnedBuffer(size_t bytes)
    {
        hip_ok(hipHostMalloc(&ptr, bytes, hipHostMallocDefault), "hipHostMalloc");
    }
    ~PinnedBuffer()
    {
        if (ptr) (void)hipHostFree(ptr);
    }
    PinnedBuffer(const PinnedBuffer &) = delete;
    PinnedBuffer &operator=(const PinnedBuffer &) = delete;
};

// FBP3D.cpp:94-104: file i is the one whose name has i in the 4 characters
// before its 4-character extension ("....0012.raw"); indices outside
// [0, count) are ignored.  Names whose 4 characters are not a number are
// skipped (the reference's std::stoi would throw).
std::vector<std::string> index_files(const std::string &dir, int count)
{
    std::vector<std::string> files(count);
    for (const auto &entry : std::filesystem::directory_iterator(dir)) {
        const std::string path = entry.path().string();
        if (path.size() < 8) continue;
        int index = -1;
        try {
            index = std::stoi(path.substr(path.size() - 8, 4));
        } catch (const std::exception &) {
            continue;
        }
        if (index < 0 || index >= count) continue;
        files[index] = path;
    }
    return files;
}

constexpr int kReadThreads = 16;
constexpr int kBatch = 128;

// Reads `count` frames of `frame_bytes` bytes (the first frame_bytes of each
// file) into `dst`, kReadThreads files at a time.
void read_files(const std::vector<std::string> &files, int first, int count, size_t frame_bytes,
                unsigned char *dst)
{
    std::atomic<int> next{0};
    std::vector<std::string> errors(kReadThreads);
    auto worker = [&](int t) {
        for (int i = next++; i < count; i = next++) {
            const std::string &path = files[first + i];
            std::FILE *f = std::fopen(path.c_str(), "rb");
            if (!f) {
                errors[t] = "cannot open \"" + path + "\"";
                return;
            }
            const size_t got = std::fread(dst + (size_t)i * frame_bytes, 1, frame_bytes, f);
            std::fclose(f);
            if (got != frame_bytes) {
                errors[t] = "\"" + path + "\" holds " + std::to_string(got) + " of the " +
                            std::to_string(frame_bytes) + " bytes of a frame";
                return;
            }
        }
    };
    std::vector<std::thread> pool;
    for (int t = 0; t < kReadThreads; ++t) pool.emplace_back(worker, t);
    for (auto &th : pool) th.join();
    for (const auto &e : errors)
        if (!e.empty()) raise_error("[CTAnalyzer Error] FBP3D read_projections(): " + e);
}

}  // namespace

// Shared body of read_projections / read_projections_as_sinograms.
static void load_frames(const std::string &dir, int file_width, int file_height,
                        int file_channel_count, int byte_per_channel, int target_w, int target_h,
                        int count, int layout, float *out, const char *what)
{
    std::cout << "[FBP3D Info] loading from \"" << dir << "\"" << std::endl;
    const auto begin = std::chrono::system_clock::now();
    if (!std::filesystem::exists(dir))
        raise_error(std::string("[CTAnalyzer Error] FBP3D tried to ") + what + "() but \"" + dir +
                    "\" doesn't exist");
    if (file_width <= 0 || file_height <= 0 || file_channel_count <= 0 ||
        (byte_per_channel != 1 && byte_per_channel != 2))
        raise_error("[CTAnalyzer Error] FBP3D read_projections(): unsupported frame format");
    const std::vector<std::string> files = index_files(dir, count);
    for (int i = 0; i < count; ++i)
        if (files[i].empty())
            raise_error("[CTAnalyzer Error] FBP3D read_projections(): no file for projection " +
                        std::to_string(i) + " in \"" + dir + "\"");
    const size_t frame_bytes =
        (size_t)file_width * file_height * file_channel_count * byte_per_channel;
    const int batch = std::min(kBatch, count);
    PinnedBuffer host(frame_bytes * batch);
    DeviceBuffer dev(frame_bytes * batch);
    hipStream_t s = static_cast<hipStream_t>(get_stream());
    for (int first = 0; first < count; first += batch) {
        const int n = std::min(batch, count - first);
        read_files(files, first, n, frame_bytes, static_cast<unsigned char *>(host.ptr));
        hip_ok(hipMemcpyAsync(dev.get(), host.ptr, frame_bytes * n, hipMemcpyHostToDevice, s),
               "hipMemcpyAsync");
        check_status(ctr_decode_frames(s, dev.get(), file_width, file_height, file_channel_count,
                                       byte_per_channel, n, out, layout, target_w, target_h,
                                       count, first),
                     what);
        hip_ok(hipStreamSynchronize(s), "hipStreamSynchronize");
    }
    const auto end = std::chrono::system_clock::now();
    std::cout << "[FBP3D Info] load finished in "
              << std::chrono::duration<double>(end - begin).count() << " seconds" << std::endl;
}

void FBP3D::read_projections(const std::string &upper_diretory, int file_width, int file_height,
                             int file_channel_count, int byte_per_channel,
                             int target_projection_width, int target_projection_height,
                             int projection_count)
{
    generate_blank_projections(target_projection_width, target_projection_height,
                               projection_count);
    load_frames(upper_diretory, file_width, file_height, file_channel_count, byte_per_channel,
                target_projection_width, target_projection_height, projection_count,
                CTR_LAYOUT_PROJECTIONS, projections->data(), "load_projections");
}

void FBP3D::read_projections_as_sinograms(const std::string &upper_diretory, int file_width,
                                          int file_height, int file_channel_count,
                                          int byte_per_channel, int target_projection_width,
                                          int target_projection_height, int projection_count)
{
    generate_blank_sinograms(target_projection_width, target_projection_height,
                             projection_count);
    load_frames(upper_diretory, file_width, file_height, file_channel_count, byte_per_channel,
                target_projection_width, target_projection_height, projection_count,
                CTR_LAYOUT_SINOGRAMS, sinograms->data(), "load_sinograms");
}

// FBP3D.cpp:178-195: one file per y-slice, "CTAnalyzer_Slice#" + the slice
// index zero-padded to 6 digits + ".raw", raw little-endian unorm16 of the
// (X, Z) slice, x fastest.  The directory gets a trailing '/' unless it
// ends in '/' or '\\' (the reference appends one unconditionally), and is
// created when missing.
void FBP3D::write_volume(std::string upper_directory)
{
    if (volume == nullptr)
        raise_error("[CTAnalyzer Error] FBP3D tried to write_volume() but FBP3D.volume was nullptr");
    if (upper_directory.empty() ||
        (upper_directory.back() != '/' && upper_directory.back() != '\\'))
        upper_directory.push_back('/');
    std::filesystem::create_directories(upper_directory);
    const glm::ivec3 size = volume->get_size();
    const int X = size.x, Y = size.y, Z = size.z;
    const size_t slice_values = (size_t)X * Z;
    const int batch = std::max(1, std::min(Y, (int)((256u << 20) / (slice_values * 2))));
    DeviceBuffer dev(slice_values * 2 * batch);
    PinnedBuffer host(slice_values * 2 * batch);
    hipStream_t s = static_cast<hipStream_t>(get_stream());
    for (int y0 = 0; y0 < Y; y0 += batch) {
        const int n = std::min(batch, Y - y0);
        check_status(ctr_slices_unorm16(s, volume->data(), X, Y, Z, y0, n,
                                        static_cast<unsigned short *>(dev.get())),
                     "FBP3D::write_volume");
        hip_ok(hipMemcpyAsync(host.ptr, dev.get(), slice_values * 2 * n, hipMemcpyDeviceToHost,
                              s),
               "hipMemcpyAsync");
        hip_ok(hipStreamSynchronize(s), "hipStreamSynchronize");
        for (int k = 0; k < n; ++k) {
            std::string index = std::to_string(y0 + k);
            index.insert(index.begin(), index.size() < kSliceIndexDigits
                                            ? kSliceIndexDigits - index.size()
                                            : 0,
                         '0');
            const std::string path = upper_directory + "CTAnalyzer_Slice#" + index + ".raw";
            std::FILE *f = std::fopen(path.c_str(), "wb");
            if (!f) raise_error("[CTAnalyzer Error] FBP3D::write_volume cannot create " + path);
            const size_t put = std::fwrite(
                static_cast<const unsigned char *>(host.ptr) + slice_values * 2 * k, 1,
                slice_values * 2, f);
            std::fclose(f);
            if (put != slice_values * 2)
                raise_error("[CTAnalyzer Error] FBP3D::write_volume short write to " + path);
        }
    }
}

// FBP3D.cpp:780-800.  The 1-texel target holds the value's float bits (the
// reference's R32UI textures); normalize_histogram(min, max) reads them back.
void FBP3D::compute_max_value_of_volume(Texture1D &target_int_texture)
{
    if (volume == nullptr)
        raise_error("[CTAnalyzer Error] FBP3D tried to compute_max_value_of_volume() but "
                    "FBP3D.volume was nullptr");
    minmax_.resize(2 * sizeof(float));
    float *mm = static_cast<float *>(minmax_.get());
    check_status(ctr_min_max(get_stream(), nullptr, volume->data(), volume->texel_count(), mm),
                 "FBP3D::compute_max_value_of_volume");
    hip_ok(hipMemcpyAsync(target_int_texture.data(), mm + 1, sizeof(float),
                          hipMemcpyDeviceToDevice, static_cast<hipStream_t>(get_stream())),
           "hipMemcpyAsync");
}

void FBP3D::compute_min_value_of_volume(Texture1D &target_int_texture)
{
    if (volume == nullptr)
        raise_error("[CTAnalyzer Error] FBP3D tried to compute_min_value_of_volume() but "
                    "FBP3D.volume was nullptr");
    minmax_.resize(2 * sizeof(float));
    float *mm = static_cast<float *>(minmax_.get());
    check_status(ctr_min_max(get_stream(), nullptr, volume->data(), volume->texel_count(), mm),
                 "FBP3D::compute_min_value_of_volume");
    hip_ok(hipMemcpyAsync(target_int_texture.data(), mm, sizeof(float), hipMemcpyDeviceToDevice,
                          static_cast<hipStream_t>(get_stream())),
           "hipMemcpyAsync");
}

// FBP3D.cpp:802-812: volume = (volume - min) / (max - min)
void FBP3D::normalize_histogram()
{
    if (volume == nullptr)
        raise_error("[CTAnalyzer Error] FBP3D tried to normalize_histogram() but FBP3D.volume "
                    "was nullptr");
    minmax_.resize(2 * sizeof(float));
    float *mm = static_cast<float *>(minmax_.get());
    check_status(ctr_min_max(get_stream(), nullptr, volume->data(), volume->texel_count(), mm),
                 "FBP3D::normalize_histogram");
    check_status(ctr_normalize_range(get_stream(), volume->data(), volume->texel_count(), mm),
                 "FBP3D::normalize_histogram");
}

// FBP3D.cpp:814-826
void FBP3D::normalize_histogram(Texture1D &source_min_int_texture,
                                Texture1D &source_max_int_texture)
{
    if (volume == nullptr)
        raise_error("[CTAnalyzer Error] FBP3D tried to normalize_histogram() but FBP3D.volume "
                    "was nullptr");
    minmax_.resize(2 * sizeof(float));
    float *mm = static_cast<float *>(minmax_.get());
    hipStream_t s = static_cast<hipStream_t>(get_stream());
    hip_ok(hipMemcpyAsync(mm, source_min_int_texture.data(), sizeof(float),
                          hipMemcpyDeviceToDevice, s),
           "hipMemcpyAsync");
    hip_ok(hipMemcpyAsync(mm + 1, source_max_int_texture.data(), sizeof(float),
                          hipMemcpyDeviceToDevice, s),
           "hipMemcpyAsync");
    check_status(ctr_normalize_range(s, volume->data(), volume->texel_count(), mm),
                 "FBP3D::normalize_histogram");
}

}  // namespace ctrecon
