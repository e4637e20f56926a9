// CUDABuffer-compatible device buffer over the C ABI
// (reference: common/CUDABuffer.h:26-81).  Same method set and meaning:
// d_pointer, resize, alloc, free, alloc_and_upload, upload, download, size,
// public sizeInBytes / d_ptr.  Size mismatches are asserts and device errors
// throw std::runtime_error, as CUDA_CHECK does (common/optix7.h:25-35).
#pragma once
#include <cassert>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/oxr.h"

namespace oxr_cpp {
// The process-wide context used by buffers and renderers (device 0).
oxr_ctx* default_context();
inline void check(int rc, const char* what) {
    if (rc != OXR_OK)
        throw std::runtime_error(std::string(what) + " failed: " + oxr_last_error(default_context()));
}
}  // namespace oxr_cpp

#define OXR_CHECK(call) oxr_cpp::check((call), #call)

typedef uintptr_t CUdeviceptr;

struct CUDABuffer {
    inline CUdeviceptr d_pointer() const { return (CUdeviceptr)d_ptr; }

    void resize(size_t size) {
        if (d_ptr) free();
        alloc(size);
    }
    void alloc(size_t size) {
        assert(d_ptr == nullptr);
        this->sizeInBytes = size;
        OXR_CHECK(oxr_malloc(oxr_cpp::default_context(), sizeInBytes, &d_ptr));
    }
    void free() {
        OXR_CHECK(oxr_free(oxr_cpp::default_context(), d_ptr));
        d_ptr = nullptr;
        sizeInBytes = 0;
    }
    template <typename T>
    void alloc_and_upload(const std::vector<T>& vt) {
        alloc(vt.size() * sizeof(T));
        upload((const T*)vt.data(), vt.size());
    }
    template <typename T>
    void upload(const T* t, size_t count) {
        assert(d_ptr != nullptr);
        assert(sizeInBytes == count * sizeof(T));
        OXR_CHECK(oxr_memcpy_h2d(oxr_cpp::default_context(), d_ptr, (const void*)t, count * sizeof(T)));
    }
    template <typename T>
    void download(T* t, size_t count) {
        assert(d_ptr != nullptr);
        assert(sizeInBytes == count * sizeof(T));
        OXR_CHECK(oxr_memcpy_d2h(oxr_cpp::default_context(), (void*)t, d_ptr, count * sizeof(T)));
    }
    inline size_t size() const { return sizeInBytes; }
    size_t sizeInBytes{0};
    void* d_ptr{nullptr};
};
