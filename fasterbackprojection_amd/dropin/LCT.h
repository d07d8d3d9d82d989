// Confocal LCT: the reference's LCT class (FB/LCT.h:6-12) has a private
// constructor and empty bodies (FB/LCT.cpp:8-19); this one is public and
// implements python/lct.py:cnlos_reconstruction (steps 1-4) over the C ABI.
// The scan grid must be square (N x N, N a power of two) with the confocal
// transient in [row][col][t] order and T a power of two.  The result is the
// (M, N, N) volume of lct.py, on the host.
#pragma once

#include <vector>

#include "GpuStructs.h"
#include "TransientParameters.h"

class LCT {
public:
    explicit LCT(const ReconstructionInfo& recInfo);
    void reconstructAABB(const ReconstructionInfo& recInfo, const ReconstructionBuffers& recBuffers,
                         const TransientParameters& transientParams);

    const std::vector<float>& volume() const { return _volume; }
    fbp::uvec3 volumeSize() const { return _size; }  // (N, N, M) as (x, y, z)

private:
    std::vector<float> _volume;
    fbp::uvec3 _size;
};
