// NLOS capture container: FB/NLosData.h:8-62.  Public fields, the
// filename constructor with the .nlos binary cache (FB/NLosData.cpp:15-127,
// 320-405) and toGpu (:131-159).  The y-tal HDF5 reader needs HighFive/HDF5,
// which this build does not link: a filename whose .nlos cache is missing
// raises std::runtime_error.  Synthetic captures are built in memory.
#pragma once

#include <string>
#include <vector>

#include "GpuStructs.h"
#include "TransientParameters.h"

// FB/AABB.h: {vptr, _max, _min} in memory; only the two corners are data.
class AABB {
public:
    explicit AABB(const fbp::vec3& mn = fbp::vec3(3.402823466e+38f),
                  const fbp::vec3& mx = fbp::vec3(-3.402823466e+38f))
        : _max(mx), _min(mn) {}
    fbp::vec3 maxPoint() const { return _max; }
    fbp::vec3 minPoint() const { return _min; }
    fbp::vec3 size() const { return _max - _min; }

    fbp::vec3 _max, _min;
};

class NLosData {
public:
    std::vector<float> _data;  // confocal [y][x][t]; exhaustive [lY][lX][cY][cX][t]
    std::vector<size_t> _dims;

    std::vector<fbp::vec3> _cameraGridPositions;
    std::vector<fbp::vec3> _cameraGridNormals;
    fbp::vec3 _cameraPosition;
    fbp::vec2 _cameraGridSize;

    std::vector<fbp::vec3> _laserGridPositions;
    std::vector<fbp::vec3> _laserGridNormals;
    fbp::vec3 _laserPosition;
    fbp::vec2 _laserGridSize;

    AABB _hiddenGeometry;

    uint32_t _temporalResolution = 0;
    float _deltaT = 0.f, _t0 = 0.f;

    bool _isConfocal = true;

    // The reference's toGpu always writes 0 (FB/NLosData.cpp:147); captures
    // without first/last bounces (all synthetic ones) set 1.
    uint32_t _discardFirstLastBounces = 0;

public:
    NLosData() = default;
    NLosData(const std::string& filename, bool saveBinary = true, bool useBinary = true);
    virtual ~NLosData() = default;

    // Uploads intensity and target positions (device buffers the caller frees
    // with fbp::deviceFree) and fills recInfo (FB/NLosData.cpp:131-159).
    void toGpu(ReconstructionInfo& recInfo, ReconstructionBuffers& recBuffers,
               const TransientParameters& transientParameters);

    bool loadBinaryFile(const std::string& filename);
    bool saveBinaryFile(const std::string& filename) const;

    size_t numTimeSeries() const { return _temporalResolution ? _data.size() / _temporalResolution : 0; }
};
