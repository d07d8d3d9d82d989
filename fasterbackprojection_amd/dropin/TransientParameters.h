// Reconstruction parameters: FB/TransientParameters.h:5-44, same members
// and defaults, plus the LCT parameters the reference's python/lct.py keeps
// as locals (new, defaulted to lct.py's values).
#pragma once

#include <string>

#include "GpuStructs.h"

enum PostprocessingFilterType { NONE, LAPLACIAN, LOG, LOG_FFT, NUM_POSTPROCESSING_FILTERS };

class TransientParameters {
public:
    bool _useFourierFilter;
    bool _reconstructAABB;
    uint32_t _numReconstructionDepths;
    fbp::uvec3 _voxelResolution;

    PostprocessingFilterType _postprocessingFilterType;
    int _kernelSize;
    float _sigma;

    std::string _outputFolder;
    bool _saveReconstructedBoundingBox;

    // LCT (python/lct.py:29-33)
    float _lctSnr;
    bool _lctDiffuse;
    uint32_t _lctZTrim;
    // K: time bins summed pairwise K times before the LCT (python/lct.py:31,
    // 98-105).  lct.py sets 1 for its 4 ps captures; 0 keeps the capture's
    // own bins, as _lctZTrim = 0 keeps all of them.
    uint32_t _lctDownsample;

    TransientParameters()
        : _useFourierFilter(true),
          _reconstructAABB(true),
          _numReconstructionDepths(200),
          _voxelResolution(256),
          _postprocessingFilterType(PostprocessingFilterType::NONE),
          _kernelSize(5),
          _sigma(1.0f),
          _outputFolder("output/"),
          _saveReconstructedBoundingBox(false),
          _lctSnr(0.8f),
          _lctDiffuse(false),
          _lctZTrim(0),
          _lctDownsample(0) {}
};
