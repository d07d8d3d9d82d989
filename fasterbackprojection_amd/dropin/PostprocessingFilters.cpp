#include "PostprocessingFilters.h"

#include <utility>

#include "Runtime.h"

namespace {
fbp_uvec3 abi(const fbp::uvec3& s) { return fbp_uvec3{s.x, s.y, s.z}; }
size_t voxels(const fbp::uvec3& s) { return size_t(s.x) * s.y * s.z; }
}  // namespace

void Laplacian::compute(float*& input, const fbp::uvec3& size,
                        const TransientParameters& transientParameters) const {
    float* out = static_cast<float*>(fbp::deviceAlloc(voxels(size) * sizeof(float)));
    fbp::check(fbp_filter_laplacian(fbp::Runtime::instance().context(), input, out, abi(size),
                                    transientParameters._kernelSize, nullptr),
               "laplacianFilter");
    fbp::synchronize("laplacianFilter");
    std::swap(input, out);
    fbp::deviceFree(out);
}

std::vector<float> LoG::calculateLaplacianKernel(int size, float std1) {
    std::vector<float> w(size_t(size) * size * size);
    fbp::check(fbp_log_kernel_host(size, std1, w.data()), "calculateLaplacianKernel");
    return w;
}

void LoG::compute(float*& input, const fbp::uvec3& size,
                  const TransientParameters& transientParameters) const {
    float* out = static_cast<float*>(fbp::deviceAlloc(voxels(size) * sizeof(float)));
    fbp::check(fbp_filter_log(fbp::Runtime::instance().context(), input, out, abi(size),
                              transientParameters._kernelSize, transientParameters._sigma, nullptr),
               "LoGFilter");
    fbp::synchronize("LoGFilter");
    std::swap(input, out);
    fbp::deviceFree(out);
}

void LoGFFT::compute(float*& input, const fbp::uvec3& size,
                     const TransientParameters& transientParameters) const {
    fbp::check(fbp_filter_log_fft(fbp::Runtime::instance().context(), input, abi(size),
                                  transientParameters._sigma, nullptr),
               "LoGFFT");
    fbp::synchronize("LoGFFT");
}
