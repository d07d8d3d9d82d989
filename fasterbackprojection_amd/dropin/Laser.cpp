#include "Laser.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <random>
#include <stdexcept>
#include <thread>

#include "Runtime.h"

const PostprocessingFilters* Laser::_postprocessingFilters[NUM_POSTPROCESSING_FILTERS] = {
    new None(), new Laplacian(), new LoG(), new LoGFFT()};

Laser::Laser(NLosData* nlosData) : _nlosData(nlosData) {
    if (!nlosData) throw std::runtime_error("Laser: null NLosData");
}

Laser::~Laser() = default;

std::vector<double> Laser::linearSpace(double minValue, double maxValue, int n) {
    // FB/Laser.cu:69-78: n samples, endpoint excluded.
    std::vector<double> v(size_t(n > 0 ? n : 0));
    const double step = (maxValue - minValue) / static_cast<double>(n);
    for (int i = 0; i < n; ++i) v[size_t(i)] = minValue + step * i;
    return v;
}

void Laser::filter_H(float* deviceData, float wl_mean, float wl_sigma, const std::string& border) {
    // FB/Laser.cu:128-240 on the device copy; the filtered series are copied
    // back into NLosData::_data as the reference leaves them there.
    const int b = (border == "edge") ? 1 : 0;
    fbp::check(fbp_filter_h(fbp::Runtime::instance().context(), deviceData,
                            _nlosData->numTimeSeries(), _nlosData->_temporalResolution,
                            _nlosData->_deltaT, wl_mean, wl_sigma, b, nullptr),
               "filter_H");
    fbp::download(_nlosData->_data.data(), deviceData, _nlosData->_data.size() * sizeof(float));
}

void Laser::reconstructShape(const TransientParameters& transientParams) {
    // FB/Laser.cu:40-67
    if (transientParams._postprocessingFilterType < NONE ||
        transientParams._postprocessingFilterType >= NUM_POSTPROCESSING_FILTERS)
        throw std::runtime_error("Laser: invalid post-processing filter type");
    ReconstructionInfo recInfo;
    ReconstructionBuffers recBuffers;
    _nlosData->toGpu(recInfo, recBuffers, transientParams);
    try {
        if (transientParams._useFourierFilter)
            filter_H(recBuffers._intensity, _nlosData->_deltaT * 10.0f, 0.0f);
        std::cout << "Reconstructing shape...\n";
        if (transientParams._reconstructAABB)
            reconstructShapeAABB(recInfo, recBuffers, transientParams);
        else
            reconstructShapeDepths(recInfo, recBuffers, transientParams);
    } catch (...) {
        fbp::deviceFree(recBuffers._laserTargets);
        fbp::deviceFree(recBuffers._sensorTargets);
        fbp::deviceFree(recBuffers._intensity);
        throw;
    }
    fbp::deviceFree(recBuffers._laserTargets);
    fbp::deviceFree(recBuffers._sensorTargets);
    fbp::deviceFree(recBuffers._intensity);
}

void Laser::reconstructShapeAABB(const ReconstructionInfo& recInfo,
                                 const ReconstructionBuffers& recBuffers,
                                 const TransientParameters& transientParams) {
    // FB/Laser.cu:341-363
    fbp_context ctx = fbp::Runtime::instance().context();
    const fbp::uvec3 res = recInfo._voxelResolution;
    const size_t numVoxels = size_t(res.x) * res.y * res.z;
    float* volume = static_cast<float*>(fbp::deviceAlloc(numVoxels * sizeof(float)));
    const auto t0 = std::chrono::steady_clock::now();
    fbp_transient tr{recBuffers._intensity, FBP_STORAGE_F32, FBP_LAYOUT_TIME_FASTEST, nullptr};
    try {
        if (recInfo._captureSystem == CaptureSystem::Confocal)
            fbp::check(fbp_backproject_confocal_voxel(
                           ctx, recInfo.abi(), &tr,
                           reinterpret_cast<const fbp_vec3*>(recBuffers._laserTargets), volume, nullptr),
                       "backprojectConfocalVoxel");
        else
            fbp::check(fbp_backproject_exhaustive_voxel(
                           ctx, recInfo.abi(), &tr,
                           reinterpret_cast<const fbp_vec3*>(recBuffers._laserTargets),
                           reinterpret_cast<const fbp_vec3*>(recBuffers._sensorTargets), volume,
                           nullptr),
                       "backprojectExhaustiveVoxel");
        _postprocessingFilters[transientParams._postprocessingFilterType]->compute(volume, res,
                                                                                   transientParams);
        fbp::check(fbp_normalize_minmax(ctx, volume, numVoxels, nullptr), "normalizeMatrix");
        fbp::synchronize("reconstructShapeAABB");
    } catch (...) {
        fbp::deviceFree(volume);
        throw;
    }
    _elapsedMs = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    std::cout << "Reconstruction finished in " << _elapsedMs << " milliseconds.\n";
    _volume.resize(numVoxels);
    fbp::download(_volume.data(), volume, numVoxels * sizeof(float));
    fbp::deviceFree(volume);
    // saveReconstructedAABB("output/aabb.cube") (FB/Laser.cu:518-524): raw float32.
    std::filesystem::create_directories(transientParams._outputFolder);
    std::ofstream f(transientParams._outputFolder + "aabb.cube", std::ios::binary);
    if (f) f.write(reinterpret_cast<const char*>(_volume.data()), std::streamsize(numVoxels * 4));
}

void Laser::reconstructShapeDepths(const ReconstructionInfo& recInfo,
                                   const ReconstructionBuffers& recBuffers,
                                   const TransientParameters& transientParams) {
    // FB/Laser.cu:365-480: planes at focus +- 0.1, unweighted sums, min-max
    // normalised, one PNG image per depth.  The activation is also kept in
    // memory and written raw to <outputFolder>depths.raw ([L][D] float32).
    fbp_context ctx = fbp::Runtime::instance().context();
    const float focus = recInfo.getFocusDepth();
    std::vector<double> depths = linearSpace(focus - 0.1f, focus + 0.1f,
                                             int(transientParams._numReconstructionDepths));
    if (depths.empty()) throw std::runtime_error("Laser: no reconstruction depths");
    const uint32_t D = uint32_t(depths.size());
    const size_t n = size_t(recInfo._numLaserTargets) * D;
    double* dDepths = static_cast<double*>(fbp::deviceAlloc(D * sizeof(double)));
    float* act = static_cast<float*>(fbp::deviceAlloc(n * sizeof(float)));
    fbp::upload(dDepths, depths.data(), D * sizeof(double));
    fbp_transient tr{recBuffers._intensity, FBP_STORAGE_F32, FBP_LAYOUT_TIME_FASTEST, nullptr};
    const auto t0 = std::chrono::steady_clock::now();
    fbp_status st;
    if (recInfo._captureSystem == CaptureSystem::Confocal)
        st = fbp_backproject_depth_confocal(ctx, recInfo.abi(), &tr,
                                            reinterpret_cast<const fbp_vec3*>(recBuffers._laserTargets),
                                            dDepths, D, act, nullptr);
    else
        st = fbp_backproject_depth_exhaustive(
            ctx, recInfo.abi(), &tr, reinterpret_cast<const fbp_vec3*>(recBuffers._laserTargets),
            reinterpret_cast<const fbp_vec3*>(recBuffers._sensorTargets), dDepths, D, act, nullptr);
    if (st == FBP_OK) st = fbp_normalize_minmax(ctx, act, n, nullptr);
    if (st != FBP_OK) {
        fbp::deviceFree(dDepths);
        fbp::deviceFree(act);
        fbp::check(st, "reconstructShapeDepths");
    }
    fbp::synchronize("reconstructShapeDepths");
    _elapsedMs = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    std::cout << "Reconstruction finished in " << _elapsedMs << " milliseconds.\n";
    _activation.resize(n);
    fbp::download(_activation.data(), act, n * sizeof(float));
    std::filesystem::create_directories(transientParams._outputFolder);
    {
        std::ofstream f(transientParams._outputFolder + "depths.raw", std::ios::binary);
        if (f) f.write(reinterpret_cast<const char*>(_activation.data()), std::streamsize(n * 4));
    }
    // One image per depth, <outputFolder>transient_<d>.png (FB/Laser.cu:410-424,
    // 463-476): res = floor(sqrt(S)) pixels a side, enlarged to
    // res * (res < 256 ? 256 / res : 1) (confocal) or 4 res (exhaustive).
    uint32_t res = uint32_t(std::sqrt(double(recInfo._numSensorTargets)));
    while (uint64_t(res) * res > recInfo._numSensorTargets) --res;
    if (res == 0 || uint64_t(res) * res > recInfo._numLaserTargets) {
        fbp::deviceFree(dDepths);
        fbp::deviceFree(act);
        throw std::runtime_error("Laser: depth images need at least one sensor target per pixel");
    }
    const bool confocal = recInfo._captureSystem == CaptureSystem::Confocal;
    const uint32_t side = confocal ? res * (res < 256u ? 256u / res : 1u) : res * 4u;
    const size_t imageBytes = size_t(side) * side * 4;
    uint8_t* dRgba = static_cast<uint8_t*>(fbp::deviceAlloc(imageBytes * D));
    st = fbp_depth_slices_rgba(ctx, act, res, res, D, side, side, dRgba, nullptr);
    std::vector<uint8_t> rgba;
    if (st == FBP_OK) {
        rgba.resize(imageBytes * D);
        fbp::download(rgba.data(), dRgba, rgba.size());
    }
    fbp::deviceFree(dRgba);
    fbp::deviceFree(dDepths);
    fbp::deviceFree(act);
    fbp::check(st, "reconstructShapeDepths: slice images");
    std::vector<std::thread> workers;
    std::atomic<uint32_t> next{0}, failed{0};
    const unsigned nw = std::max(1u, std::min(8u, std::thread::hardware_concurrency()));
    for (unsigned w = 0; w < nw; ++w)
        workers.emplace_back([&] {
            for (uint32_t d = next++; d < D; d = next++) {
                const std::string name =
                    transientParams._outputFolder + "transient_" + std::to_string(d) + ".png";
                if (fbp_write_png_rgba(name.c_str(), rgba.data() + d * imageBytes, side, side) !=
                    FBP_OK)
                    ++failed;
            }
        });
    for (auto& t : workers) t.join();
    if (failed) throw std::runtime_error("Laser: could not write the depth images");
}

void Laser::reconstructAABBConfocalMIS(float* volume, const ReconstructionInfo& recInfo) const {
    // Noise buffer of numVoxels uniform floats from a default-seeded mt19937,
    // as RandomUtilities::getUniformRandom produces (FB/RandomUtilities.h:38-44).
    const size_t numVoxels =
        size_t(recInfo._voxelResolution.x) * recInfo._voxelResolution.y * recInfo._voxelResolution.z;
    std::mt19937 gen;
    std::uniform_real_distribution<float> dist(0.0f, 1.0f);
    std::vector<float> noise(numVoxels);
    for (float& v : noise) v = dist(gen);
    NLosData& d = *_nlosData;
    float* dI = static_cast<float*>(fbp::deviceAlloc(d._data.size() * sizeof(float)));
    fbp::vec3* dL = static_cast<fbp::vec3*>(fbp::deviceAlloc(d._laserGridPositions.size() * sizeof(fbp::vec3)));
    float* dN = static_cast<float*>(fbp::deviceAlloc(numVoxels * sizeof(float)));
    fbp::upload(dI, d._data.data(), d._data.size() * sizeof(float));
    fbp::upload(dL, d._laserGridPositions.data(), d._laserGridPositions.size() * sizeof(fbp::vec3));
    fbp::upload(dN, noise.data(), numVoxels * sizeof(float));
    fbp_transient tr{dI, FBP_STORAGE_F32, FBP_LAYOUT_TIME_FASTEST, nullptr};
    fbp_mis_params mp{recInfo._numLaserTargets / 4, FBP_MIS_REFERENCE, dN, uint32_t(numVoxels)};
    fbp_status st = fbp_backproject_confocal_mis(fbp::Runtime::instance().context(), recInfo.abi(), &tr,
                                                 reinterpret_cast<const fbp_vec3*>(dL), &mp, volume,
                                                 nullptr);
    fbp::deviceFree(dI);
    fbp::deviceFree(dL);
    fbp::deviceFree(dN);
    fbp::check(st, "backprojectConfocalVoxelMIS");
}
