// Reconstruction orchestration: FB/Laser.cuh:14-48 / FB/Laser.cu.
// reconstructShape runs filter_H (when enabled), uploads the capture,
// backprojects into the AABB volume (or depth planes), applies the selected
// post-filter, min-max normalises and writes <outputFolder>aabb.cube (raw
// float32) like Laser::reconstructShapeAABB.
#pragma once

#include <string>
#include <vector>

#include "NLosData.h"
#include "PostprocessingFilters.h"
#include "TransientParameters.h"

class Laser {
private:
    NLosData* _nlosData = nullptr;
    static const PostprocessingFilters* _postprocessingFilters[NUM_POSTPROCESSING_FILTERS];

    std::vector<float> _volume;       // last AABB reconstruction, [Z][Y][X]
    std::vector<float> _activation;   // last depth-plane reconstruction, [L][D]
    double _elapsedMs = 0.0;

    static std::vector<double> linearSpace(double minValue, double maxValue, int n);
    void filter_H(float* deviceData, float wl_mean, float wl_sigma = .0f,
                  const std::string& border = "zero");
    void reconstructShapeAABB(const ReconstructionInfo& recInfo,
                              const ReconstructionBuffers& recBuffers,
                              const TransientParameters& transientParams);
    void reconstructShapeDepths(const ReconstructionInfo& recInfo,
                                const ReconstructionBuffers& recBuffers,
                                const TransientParameters& transientParams);

public:
    // MIS-sampled confocal backprojection into a device volume
    // (FB/Laser.cu:526-611); uploads the capture itself.
    void reconstructAABBConfocalMIS(float* volume, const ReconstructionInfo& recInfo) const;

    explicit Laser(NLosData* nlosData);
    virtual ~Laser();

    void reconstructShape(const TransientParameters& transientParams);

    const std::vector<float>& volume() const { return _volume; }
    const std::vector<float>& activation() const { return _activation; }
    double elapsedMilliseconds() const { return _elapsedMs; }
};
