// Output drivers: the film as a Radiance RGBE file (builtin/driver/outputhdr.go:30-57,
// image/hdr/writer.go:34-129, image/hdr/hdr.go:24-48) or as raw float32 (builtin/driver/outputfloat.go:30-42).
//
// The RGBE encoding runs on the GPU, one thread per pixel, writing the file's scanline order (the
// bottom film row first, as writer.go:96-106 walks k = Height - j - 1).  Every value follows the
// reference's float32 arithmetic on amd64, so the bytes are the reference's:
//   d = max(r, g, b) by "greater than" tests (a NaN channel never becomes the max);
//   d < 1e-6 -> 0 0 0 0; else Frexp(d) = (n, e) ((d, 0) for an infinite or NaN d),
//   df = n * 255.999 / d (float32, left to right);
//   each byte is byte(x * df): Go converts through a truncating int32 conversion (CVTTSS2SL), whose
//   NaN or out-of-range result is 0x80000000, then keeps the low byte; the exponent byte is e + 128
//   modulo 256.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <vector>

#include "scene_internal.h"

namespace {

__device__ __forceinline__ uint8_t go_byte(float x) {
    // float32 -> int32 truncation with the x86 "integer indefinite" result, then the low byte
    const int32_t i = (x >= -2147483648.0f && x < 2147483648.0f) ? (int32_t)x : (int32_t)0x80000000u;
    return (uint8_t)(uint32_t)i;
}

__global__ __launch_bounds__(256) void k_rgbe(const float* __restrict__ film, uint32_t w, uint32_t h,
                                             uchar4* __restrict__ out) {
    const uint64_t np = (uint64_t)w * h;
    for (uint64_t q = (uint64_t)blockIdx.x * 256 + threadIdx.x; q < np; q += (uint64_t)gridDim.x * 256) {
        // output position q = j * w + i holds film row k = h - j - 1, column i
        const uint32_t j = (uint32_t)(q / w), i = (uint32_t)(q % w);
        const uint64_t p = ((uint64_t)(h - j - 1) * w + i) * 3;
        const float r = film[p], g = film[p + 1], b = film[p + 2];
        float d = r;
        if (g > d) d = g;
        if (b > d) d = b;
        uchar4 o = make_uchar4(0, 0, 0, 0);
        if (!(d < 0.000001f)) {
            // math.Frexp returns (d, 0) for an infinite or NaN d; the quotient below is then NaN
            int e = 0;
            const float n = isfinite(d) ? (float)frexp((double)d, &e) : d;
            const float df = n * 255.999f / d;
            o = make_uchar4(go_byte(r * df), go_byte(g * df), go_byte(b * df), (uint8_t)(e + 128));
        }
        out[q] = o;
    }
}

const char kHeaderFmt[] = "#?RADIANCE\n# Created by Vermeer Light Tools (http://www.vermeerlt.com)\nFORMAT=32-bit_rle_rgbe\n\n+Y %u +X %u\n";

int write_file(const char* path, const std::vector<char>& head, const void* body, size_t bytes) {
    FILE* f = std::fopen(path, "wb");
    if (!f) return vgpu::fail(VGPU_ERR_ARG, "cannot create %s", path);
    const bool ok = std::fwrite(head.data(), 1, head.size(), f) == head.size() &&
                    (bytes == 0 || std::fwrite(body, 1, bytes, f) == bytes);
    const bool closed = std::fclose(f) == 0;
    if (!ok || !closed) return vgpu::fail(VGPU_ERR_ARG, "writing %s failed", path);
    return VGPU_OK;
}

// Device copy of the film (device pointers are used as they are).
struct FilmOnDevice {
    const float* ptr = nullptr;
    float* owned = nullptr;
    ~FilmOnDevice() { if (owned) (void)hipFree(owned); }
};

int film_on_device(uint32_t flags, const float* film, size_t floats, hipStream_t st, FilmOnDevice& out) {
    if (flags & VGPU_TRACE_DEVICE_PTRS) { out.ptr = film; return VGPU_OK; }
    HIP_TRY(hipMalloc((void**)&out.owned, floats * sizeof(float)));
    HIP_TRY(hipMemcpyAsync(out.owned, film, floats * sizeof(float), hipMemcpyHostToDevice, st));
    out.ptr = out.owned;
    return VGPU_OK;
}

int check_film_args(uint32_t flags, const float* film, uint32_t w, uint32_t h) {
    if (!film) return vgpu::fail(VGPU_ERR_ARG, "film is NULL");
    if (w == 0 || h == 0) return vgpu::fail(VGPU_ERR_ARG, "empty image %ux%u", w, h);
    if ((flags & ~VGPU_TRACE_DEVICE_PTRS) != 0) return vgpu::fail(VGPU_ERR_ARG, "unknown flags 0x%x", flags);
    return VGPU_OK;
}

}  // namespace

extern "C" {

int vgpu_film_to_rgbe(int device, uint32_t flags, const float* film, uint32_t width, uint32_t height, uint8_t* rgbe,
                      void* stream) {
    int rc = check_film_args(flags, film, width, height);
    if (rc) return rc;
    if (!rgbe) return vgpu::fail(VGPU_ERR_ARG, "rgbe is NULL");
    vgpu_detail::DeviceGuard dg(device);
    hipStream_t st = (hipStream_t)stream;
    const size_t np = (size_t)width * height;
    FilmOnDevice f;
    if ((rc = film_on_device(flags, film, np * 3, st, f))) return rc;
    uchar4* d_out = (uchar4*)rgbe;
    uint8_t* owned = nullptr;
    if (!(flags & VGPU_TRACE_DEVICE_PTRS)) {
        HIP_TRY(hipMalloc((void**)&owned, np * 4));
        d_out = (uchar4*)owned;
    }
    int cus = 256;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
    const uint64_t want = (np + 255) / 256, cap = (uint64_t)cus * 16;
    const unsigned grid = (unsigned)(want < cap ? want : cap);
    hipLaunchKernelGGL(k_rgbe, dim3(grid), dim3(256), 0, st, f.ptr, width, height, d_out);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess && owned) {
        e = hipMemcpyAsync(rgbe, owned, np * 4, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
    } else if (e == hipSuccess && !(flags & VGPU_TRACE_DEVICE_PTRS)) {
        e = hipStreamSynchronize(st);
    }
    if (owned) (void)hipFree(owned);
    if (e != hipSuccess) return vgpu::fail(VGPU_ERR_HIP, "RGBE encoding: %s", hipGetErrorString(e));
    return VGPU_OK;
}

int vgpu_write_hdr(int device, uint32_t flags, const float* film, uint32_t width, uint32_t height, const char* path,
                   void* stream) {
    int rc = check_film_args(flags, film, width, height);
    if (rc) return rc;
    if (!path) return vgpu::fail(VGPU_ERR_ARG, "path is NULL");
    std::vector<uint8_t> body((size_t)width * height * 4);
    if (flags & VGPU_TRACE_DEVICE_PTRS) {
        // device film: encode into a device buffer, then copy the bytes to the host
        vgpu_detail::DeviceGuard dg(device);
        uint8_t* d_out = nullptr;
        HIP_TRY(hipMalloc((void**)&d_out, body.size()));
        rc = vgpu_film_to_rgbe(device, flags, film, width, height, d_out, stream);
        hipError_t e = rc ? hipSuccess : hipMemcpyAsync(body.data(), d_out, body.size(), hipMemcpyDeviceToHost,
                                                         (hipStream_t)stream);
        if (!rc && e == hipSuccess) e = hipStreamSynchronize((hipStream_t)stream);
        (void)hipFree(d_out);
        if (rc) return rc;
        if (e != hipSuccess) return vgpu::fail(VGPU_ERR_HIP, "RGBE copy: %s", hipGetErrorString(e));
    } else if ((rc = vgpu_film_to_rgbe(device, flags, film, width, height, body.data(), stream))) {
        return rc;
    }
    std::vector<char> head(256);
    const int n = std::snprintf(head.data(), head.size(), kHeaderFmt, height, width);
    head.resize((size_t)n);
    return write_file(path, head, body.data(), body.size());
}

int vgpu_write_float(int device, uint32_t flags, const float* film, uint32_t width, uint32_t height, const char* path,
                     void* stream) {
    int rc = check_film_args(flags, film, width, height);
    if (rc) return rc;
    if (!path) return vgpu::fail(VGPU_ERR_ARG, "path is NULL");
    const size_t floats = (size_t)width * height * 3;
    std::vector<float> host;
    const float* src = film;
    if (flags & VGPU_TRACE_DEVICE_PTRS) {
        vgpu_detail::DeviceGuard dg(device);
        host.resize(floats);
        HIP_TRY(hipMemcpyAsync(host.data(), film, floats * sizeof(float), hipMemcpyDeviceToHost, (hipStream_t)stream));
        HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
        src = host.data();
    }
    // binary.Write(fp, binary.LittleEndian, FrameBuf()): the float32 values in film order, little-endian
    // (the host byte order here)
    return write_file(path, std::vector<char>(), src, floats * sizeof(float));
}

}  // extern "C"
