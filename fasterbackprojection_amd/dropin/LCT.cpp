#include "LCT.h"

#include <cmath>
#include <stdexcept>

#include "Runtime.h"

LCT::LCT(const ReconstructionInfo& recInfo) {
    const uint32_t n = uint32_t(std::lround(std::sqrt(double(recInfo._numLaserTargets))));
    _size = fbp::uvec3(n, n, recInfo._numTimeBins);
}

void LCT::reconstructAABB(const ReconstructionInfo& recInfo, const ReconstructionBuffers& recBuffers,
                          const TransientParameters& transientParams) {
    const uint32_t N = _size.x, T = recInfo._numTimeBins, K = transientParams._lctDownsample;
    if (size_t(N) * N != recInfo._numLaserTargets || recInfo._captureSystem != CaptureSystem::Confocal)
        throw std::runtime_error("LCT: needs a square confocal scan grid");
    if (K > 16 || (T >> K) < 2) throw std::runtime_error("LCT: downsampling leaves fewer than 2 bins");
    const uint32_t M = T >> K;
    fbp_lct_params p{};
    p.n = N;
    p.m = T;
    p.width = recInfo._relayWallSize.x * 0.5f;
    p.range_max = float(T) * recInfo._timeStep;  // M * c * bin_resolution before downsampling
    p.snr = transientParams._lctSnr;
    p.is_diffuse = transientParams._lctDiffuse ? 1 : 0;
    p.is_backprop = 0;
    p.z_trim = transientParams._lctZTrim;
    p.downsample = K;
    const size_t n = size_t(N) * N * M;
    float* out = static_cast<float*>(fbp::deviceAlloc(n * sizeof(float)));
    fbp_status st = fbp_lct(fbp::Runtime::instance().context(), recBuffers._intensity, &p, out, nullptr);
    if (st == FBP_OK) {
        fbp::synchronize("LCT");
        _volume.resize(n);
        fbp::download(_volume.data(), out, n * sizeof(float));
    }
    fbp::deviceFree(out);
    fbp::check(st, "LCT::reconstructAABB");
    _size = fbp::uvec3(N, N, M);
}
