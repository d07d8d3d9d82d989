// Post-processing filters: FB/PostprocessingFilters.h:5-39.  compute() takes
// a device volume by reference; Laplacian and LoG allocate the output, swap
// it into `input` and free the old buffer; LoGFFT works in place; None is a
// no-op (FB/PostprocessingFilters.cu:29-31, 96-97).
#pragma once

#include <vector>

#include "GpuStructs.h"
#include "TransientParameters.h"

class PostprocessingFilters {
public:
    virtual ~PostprocessingFilters() = default;
    virtual void compute(float*& input, const fbp::uvec3& size,
                         const TransientParameters& transientParameters) const = 0;
};

class None : public PostprocessingFilters {
public:
    void compute(float*&, const fbp::uvec3&, const TransientParameters&) const override {}
};

class Laplacian : public PostprocessingFilters {
public:
    void compute(float*& input, const fbp::uvec3& size,
                 const TransientParameters& transientParameters) const override;
};

class LoG : public PostprocessingFilters {
public:
    // The k^3 weights as the reference builds them (calculateLaplacianKernel).
    static std::vector<float> calculateLaplacianKernel(int size, float std1);
    void compute(float*& input, const fbp::uvec3& size,
                 const TransientParameters& transientParameters) const override;
};

class LoGFFT : public PostprocessingFilters {
public:
    void compute(float*& input, const fbp::uvec3& size,
                 const TransientParameters& transientParameters) const override;
};
