#include "Runtime.h"

#include <hip/hip_runtime_api.h>

#include <cstdlib>

namespace fbp {

Runtime::Runtime() {
    int dev = -1;
    if (const char* e = std::getenv("FBP_DEVICE")) dev = std::atoi(e);
    fbp_status st = fbp_context_create(dev, &_ctx);
    if (st != FBP_OK)
        throw std::runtime_error(std::string("fbp_context_create: ") + fbp_status_string(st));
}

Runtime::~Runtime() {
    if (_ctx) fbp_context_destroy(_ctx);
}

Runtime& Runtime::instance() {
    static Runtime rt;
    return rt;
}

static void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

void* deviceAlloc(size_t bytes) {
    Runtime::instance();
    void* p = nullptr;
    hip_check(hipMalloc(&p, bytes ? bytes : 1), "hipMalloc");
    return p;
}

void deviceFree(void* p) {
    if (p) hipFree(p);
}

void upload(void* dst, const void* src, size_t bytes) {
    hip_check(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice), "hipMemcpy H2D");
}

void download(void* dst, const void* src, size_t bytes) {
    hip_check(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), "hipMemcpy D2H");
}

void synchronize(const char* what) { hip_check(hipDeviceSynchronize(), what); }

}  // namespace fbp
