#include "NLosData.h"

#include <fstream>
#include <stdexcept>

#include "Runtime.h"

NLosData::NLosData(const std::string& filename, bool saveBinary, bool useBinary) {
    // FB/NLosData.cpp:15-127: the .nlos cache beside the file is tried first.
    const std::string binaryFile = filename.substr(0, filename.find_last_of('.')) + ".nlos";
    if (useBinary && loadBinaryFile(binaryFile)) return;
    if (filename.size() >= 5 && filename.compare(filename.size() - 5, 5, ".nlos") == 0 &&
        loadBinaryFile(filename))
        return;
    (void)saveBinary;
    throw std::runtime_error("NLosData: no readable .nlos cache for '" + filename +
                             "' (HDF5 / .mat readers are not part of this build)");
}

void NLosData::toGpu(ReconstructionInfo& recInfo, ReconstructionBuffers& recBuffers,
                     const TransientParameters& transientParameters) {
    if (_cameraGridPositions.empty() || _laserGridPositions.empty() || _data.empty())
        throw std::runtime_error("NLosData::toGpu: empty capture");
    recBuffers._intensity = static_cast<float*>(fbp::deviceAlloc(_data.size() * sizeof(float)));
    fbp::upload(recBuffers._intensity, _data.data(), _data.size() * sizeof(float));
    recBuffers._sensorTargets =
        static_cast<fbp::vec3*>(fbp::deviceAlloc(_cameraGridPositions.size() * sizeof(fbp::vec3)));
    fbp::upload(recBuffers._sensorTargets, _cameraGridPositions.data(),
                _cameraGridPositions.size() * sizeof(fbp::vec3));
    recBuffers._laserTargets =
        static_cast<fbp::vec3*>(fbp::deviceAlloc(_laserGridPositions.size() * sizeof(fbp::vec3)));
    fbp::upload(recBuffers._laserTargets, _laserGridPositions.data(),
                _laserGridPositions.size() * sizeof(fbp::vec3));

    recInfo._sensorPosition = _cameraPosition;
    recInfo._numSensorTargets = uint32_t(_cameraGridPositions.size());
    recInfo._laserPosition = _laserPosition;
    recInfo._numLaserTargets = uint32_t(_laserGridPositions.size());
    recInfo._numTimeBins = _temporalResolution;
    recInfo._timeStep = _deltaT;
    recInfo._timeOffset = _t0;
    recInfo._captureSystem = _isConfocal ? CaptureSystem::Confocal : CaptureSystem::Exhaustive;
    recInfo._discardFirstLastBounces = _discardFirstLastBounces;
    recInfo._relayWallNormal = _cameraGridNormals.empty() ? fbp::vec3(0.f, 1.f, 0.f) : _cameraGridNormals.front();
    recInfo._relayWallMinPosition = fbp::vec3(-_cameraGridSize.x / 2.0f, 0.f, -_cameraGridSize.y / 2.0f);
    recInfo._relayWallSize = fbp::vec3(_cameraGridSize.x, 0.f, _cameraGridSize.y);
    recInfo._hiddenVolumeMin = _hiddenGeometry.minPoint();
    recInfo._hiddenVolumeMax = _hiddenGeometry.maxPoint();
    recInfo._hiddenVolumeSize = _hiddenGeometry.size();
    recInfo._voxelResolution = transientParameters._voxelResolution;
    const fbp::vec3 res(float(recInfo._voxelResolution.x), float(recInfo._voxelResolution.y),
                        float(recInfo._voxelResolution.z));
    recInfo._hiddenVolumeVoxelSize = _hiddenGeometry.size() / res;
}

// .nlos layout (FB/NLosData.cpp:320-405): vec3 cameraPosition, vec3
// laserPosition, uint temporalResolution, float deltaT, float t0, bool
// isConfocal, AABB (8-byte vptr slot, _max, _min), size_t numDims + dims,
// size_t n + cameraGridPositions + cameraGridNormals, vec2 cameraGridSize,
// the same for the laser grid, size_t numValues + float data.
bool NLosData::loadBinaryFile(const std::string& filename) {
    std::ifstream f(filename, std::ios::binary);
    if (!f.is_open()) return false;
    auto rd = [&](void* p, size_t n) { f.read(static_cast<char*>(p), std::streamsize(n)); return bool(f); };
    bool ok = rd(&_cameraPosition, 12) && rd(&_laserPosition, 12) && rd(&_temporalResolution, 4) &&
              rd(&_deltaT, 4) && rd(&_t0, 4);
    unsigned char confocal = 0;
    ok = ok && rd(&confocal, 1);
    _isConfocal = confocal != 0;
    unsigned char vptr[8];
    ok = ok && rd(vptr, 8) && rd(&_hiddenGeometry._max, 12) && rd(&_hiddenGeometry._min, 12);
    uint64_t n = 0;
    ok = ok && rd(&n, 8) && n < 16;
    if (!ok) return false;
    _dims.resize(n);
    std::vector<uint64_t> d(n);
    ok = rd(d.data(), n * 8);
    for (size_t i = 0; i < n; ++i) _dims[i] = size_t(d[i]);
    ok = ok && rd(&n, 8) && n < (uint64_t(1) << 32);
    if (!ok) return false;
    _cameraGridPositions.resize(n);
    _cameraGridNormals.resize(n);
    ok = rd(_cameraGridPositions.data(), n * 12) && rd(_cameraGridNormals.data(), n * 12) &&
         rd(&_cameraGridSize, 8) && rd(&n, 8) && n < (uint64_t(1) << 32);
    if (!ok) return false;
    _laserGridPositions.resize(n);
    _laserGridNormals.resize(n);
    ok = rd(_laserGridPositions.data(), n * 12) && rd(_laserGridNormals.data(), n * 12) &&
         rd(&_laserGridSize, 8) && rd(&n, 8) && n < (uint64_t(1) << 40);
    if (!ok) return false;
    _data.resize(n);
    return rd(_data.data(), n * 4);
}

bool NLosData::saveBinaryFile(const std::string& filename) const {
    std::ofstream f(filename, std::ios::binary);
    if (!f.is_open()) return false;
    auto wr = [&](const void* p, size_t n) { f.write(static_cast<const char*>(p), std::streamsize(n)); };
    wr(&_cameraPosition, 12);
    wr(&_laserPosition, 12);
    wr(&_temporalResolution, 4);
    wr(&_deltaT, 4);
    wr(&_t0, 4);
    const unsigned char confocal = _isConfocal ? 1 : 0;
    wr(&confocal, 1);
    const unsigned char vptr[8] = {0};
    wr(vptr, 8);
    wr(&_hiddenGeometry._max, 12);
    wr(&_hiddenGeometry._min, 12);
    uint64_t n = _dims.size();
    wr(&n, 8);
    for (size_t v : _dims) {
        const uint64_t u = v;
        wr(&u, 8);
    }
    n = _cameraGridPositions.size();
    wr(&n, 8);
    wr(_cameraGridPositions.data(), n * 12);
    std::vector<fbp::vec3> normals = _cameraGridNormals;
    normals.resize(n, fbp::vec3(0.f, 1.f, 0.f));
    wr(normals.data(), n * 12);
    wr(&_cameraGridSize, 8);
    n = _laserGridPositions.size();
    wr(&n, 8);
    wr(_laserGridPositions.data(), n * 12);
    normals = _laserGridNormals;
    normals.resize(n, fbp::vec3(0.f, 1.f, 0.f));
    wr(normals.data(), n * 12);
    wr(&_laserGridSize, 8);
    n = _data.size();
    wr(&n, 8);
    wr(_data.data(), n * 4);
    return bool(f);
}
