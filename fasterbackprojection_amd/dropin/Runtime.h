// Process-wide device context of the drop-in classes (the reference's
// CudaContext singleton, FB/CudaContext.cpp:22-58) and the error
// convention of CudaHelper::checkError (FB/CudaHelper.cpp:66-69): any
// non-zero fbp_status becomes a std::runtime_error.
#pragma once

#include <stdexcept>
#include <string>

#include "fbp.h"

namespace fbp {

class Runtime {
public:
    // Device: FBP_DEVICE environment variable, else the reference's pick
    // (largest clock x multiprocessor count).
    static Runtime& instance();
    fbp_context context() const { return _ctx; }
    ~Runtime();

private:
    Runtime();
    fbp_context _ctx = nullptr;
};

inline void check(fbp_status st, const char* what) {
    if (st != FBP_OK) {
        std::string msg = std::string(what) + ": " + fbp_status_string(st);
        const char* detail = fbp_context_last_error(Runtime::instance().context());
        if (detail && *detail) msg += std::string(" (") + detail + ")";
        throw std::runtime_error(msg);
    }
}

// Device memory helpers with the reference's CudaHelper semantics.
void* deviceAlloc(size_t bytes);
void deviceFree(void* p);
void upload(void* dst, const void* src, size_t bytes);
void download(void* dst, const void* src, size_t bytes);
void synchronize(const char* what);

}  // namespace fbp
