// Ghost exchange between the sub-volumes of one rank: X, then Y, then Z, so edges and corners are filled by the
// sequencing.
//
// Reference: infrastructure/haloManager3D_gpu.h:14-60 (interface), infrastructure/haloManager3D_gpu.cu:5-58
// (constructor), :133-205 (local exchanges captured as graphs), infrastructure/haloManager3D_gpu_kernels.cuh:14-472
// (copy_2way_X/Y/Z face extents). On MI355X each axis is ONE batched copy launch over a validated descriptor table
// (tti_copy_plan) instead of a graph of one kernel per sub-volume pair and field; ghosts on the global boundary are
// not touched.
#pragma once

#include <array>
#include <vector>

#include "decomp.h"
#include "grid.h"
#include "tti_grid.h"

class haloManager3D_gpu
{
public:
    // data[field][sub]: volumes of every field; wSz[axis][side][field][sub]: ghost width received on that side
    // (side 0 = negative). includeEdges / includeCorners are accepted and, as in the reference, treated as true.
    haloManager3D_gpu(decompositionManager3D* decomp, std::vector<cart_volume<float>**>& data,
                      std::vector<std::vector<int>> wSz[3][2], std::vector<std::vector<bool>>& includeEdges,
                      std::vector<std::vector<bool>>& includeCorners);
    ~haloManager3D_gpu();
    haloManager3D_gpu(const haloManager3D_gpu&) = delete;
    haloManager3D_gpu& operator=(const haloManager3D_gpu&) = delete;

    void start_update() { start_update(nullptr); }
    void finish_update() { finish_update(nullptr); }
    // The exchange is ordered on `stream`; finish_update only keeps the reference's pairing (the copies of one rank
    // need no host wait).
    void start_update(hipStream_t stream);
    void finish_update(hipStream_t stream);

    int exchange_count(int axis) const { return counts_[axis]; }

private:
    std::array<tti_copy_plan*, 3> plans_{ nullptr, nullptr, nullptr };
    std::array<int, 3> counts_{ 0, 0, 0 };
};
