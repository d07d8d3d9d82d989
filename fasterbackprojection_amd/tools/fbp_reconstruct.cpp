// Command-line driver of the drop-in C++ classes: what the reference's
// SceneContent::buildScenario does (FB/SceneContent.cpp:14-32) with the
// capture and parameters taken from the command line.
//
//   fbp_reconstruct info <capture.nlos>
//   fbp_reconstruct aabb <capture.nlos> <out/> <res> <filter 0..3> <fourier 0|1> <discard 0|1>
//   fbp_reconstruct depths <capture.nlos> <out/> <num_depths> <fourier 0|1> <discard 0|1>
//   fbp_reconstruct lct <capture.nlos> <out/> <snr> <z_trim> [downsample K]
//   fbp_reconstruct mis <capture.nlos> <out/> <res> <discard 0|1>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>

#include "LCT.h"
#include "Laser.h"
#include "NLosData.h"
#include "Runtime.h"

static int usage() {
    std::cerr << "usage: fbp_reconstruct info|aabb|depths|lct|mis <capture.nlos> ...\n";
    return 2;
}

static void write_raw(const std::string& path, const std::vector<float>& v) {
    std::ofstream f(path, std::ios::binary);
    f.write(reinterpret_cast<const char*>(v.data()), std::streamsize(v.size() * 4));
}

int main(int argc, char** argv) {
    if (argc < 3) return usage();
    const std::string mode = argv[1];
    try {
        NLosData data(argv[2]);
        if (mode == "info") {
            double sum = 0.0;
            for (float v : data._data) sum += v;
            std::printf("{\"dims\": [");
            for (size_t i = 0; i < data._dims.size(); ++i) std::printf("%s%zu", i ? ", " : "", data._dims[i]);
            std::printf("], \"t\": %u, \"deltaT\": %.9g, \"t0\": %.9g, \"confocal\": %d, "
                        "\"camera\": %zu, \"laser\": %zu, \"values\": %zu, \"sum\": %.17g, "
                        "\"hmin\": [%.9g, %.9g, %.9g], \"hmax\": [%.9g, %.9g, %.9g]}\n",
                        data._temporalResolution, data._deltaT, data._t0, int(data._isConfocal),
                        data._cameraGridPositions.size(), data._laserGridPositions.size(),
                        data._data.size(), sum, data._hiddenGeometry._min.x, data._hiddenGeometry._min.y,
                        data._hiddenGeometry._min.z, data._hiddenGeometry._max.x,
                        data._hiddenGeometry._max.y, data._hiddenGeometry._max.z);
            return 0;
        }
        if (argc < 4) return usage();
        TransientParameters p;
        p._outputFolder = argv[3];
        if (!p._outputFolder.empty() && p._outputFolder.back() != '/') p._outputFolder += '/';
        if (mode == "aabb" && argc >= 9) {
            p._voxelResolution = fbp::uvec3(uint32_t(std::atoi(argv[4])));
            p._postprocessingFilterType = PostprocessingFilterType(std::atoi(argv[5]));
            p._useFourierFilter = std::atoi(argv[6]) != 0;
            data._discardFirstLastBounces = uint32_t(std::atoi(argv[7]));
            (void)argv[8];
            Laser laser(&data);
            laser.reconstructShape(p);
            return 0;
        }
        if (mode == "depths" && argc >= 7) {
            p._reconstructAABB = false;
            p._numReconstructionDepths = uint32_t(std::atoi(argv[4]));
            p._useFourierFilter = std::atoi(argv[5]) != 0;
            data._discardFirstLastBounces = uint32_t(std::atoi(argv[6]));
            Laser laser(&data);
            laser.reconstructShape(p);
            return 0;
        }
        if (mode == "lct" && argc >= 6) {
            p._lctSnr = float(std::atof(argv[4]));
            p._lctZTrim = uint32_t(std::atoi(argv[5]));
            if (argc >= 7) p._lctDownsample = uint32_t(std::atoi(argv[6]));
            ReconstructionInfo info;
            ReconstructionBuffers buf;
            data.toGpu(info, buf, p);
            LCT lct(info);
            lct.reconstructAABB(info, buf, p);
            fbp::deviceFree(buf._intensity);
            fbp::deviceFree(buf._laserTargets);
            fbp::deviceFree(buf._sensorTargets);
            write_raw(p._outputFolder + "lct.raw", lct.volume());
            return 0;
        }
        if (mode == "mis" && argc >= 6) {
            p._voxelResolution = fbp::uvec3(uint32_t(std::atoi(argv[4])));
            data._discardFirstLastBounces = uint32_t(std::atoi(argv[5]));
            ReconstructionInfo info;
            ReconstructionBuffers buf;
            data.toGpu(info, buf, p);
            const size_t nv = size_t(info._voxelResolution.x) * info._voxelResolution.y * info._voxelResolution.z;
            float* vol = static_cast<float*>(fbp::deviceAlloc(nv * 4));
            Laser laser(&data);
            laser.reconstructAABBConfocalMIS(vol, info);
            std::vector<float> h(nv);
            fbp::download(h.data(), vol, nv * 4);
            fbp::deviceFree(vol);
            fbp::deviceFree(buf._intensity);
            fbp::deviceFree(buf._laserTargets);
            fbp::deviceFree(buf._sensorTargets);
            write_raw(p._outputFolder + "mis.raw", h);
            return 0;
        }
        return usage();
    } catch (const std::exception& e) {
        std::cerr << "error: " << e.what() << "\n";
        return 1;
    }
}
