// Zone decomposition of the global grid on one rank: every non-empty zone becomes a sub-volume.
//
// Reference: infrastructure/decompositionManager3D.cpp:12-89 (constructor, partition), :340-404 (sub-domain creation),
// infrastructure/decompositionManager3D.h:110-177 (accessors). With one rank per axis the weighted 1-D partition gives
// the whole axis to the rank, so the sub-domains are the zones themselves, numbered x fastest.
#pragma once

#include <vector>

#include "comm.h"

// Neighbour slots of a sub-domain, x fastest: slot = (sz + 1)*9 + (sy + 1)*3 + (sx + 1) for offsets s in {-1, 0, 1}
// (the numbering of utils/mpi_const.h:37-66).
enum NH_ID_3D
{
    _NH_ID_3D_NEGX_NEGY_NEGZ, _NH_ID_3D_CENX_NEGY_NEGZ, _NH_ID_3D_POSX_NEGY_NEGZ,
    _NH_ID_3D_NEGX_CENY_NEGZ, _NH_ID_3D_CENX_CENY_NEGZ, _NH_ID_3D_POSX_CENY_NEGZ,
    _NH_ID_3D_NEGX_POSY_NEGZ, _NH_ID_3D_CENX_POSY_NEGZ, _NH_ID_3D_POSX_POSY_NEGZ,
    _NH_ID_3D_NEGX_NEGY_CENZ, _NH_ID_3D_CENX_NEGY_CENZ, _NH_ID_3D_POSX_NEGY_CENZ,
    _NH_ID_3D_NEGX_CENY_CENZ, _NH_ID_3D_CENX_CENY_CENZ, _NH_ID_3D_POSX_CENY_CENZ,
    _NH_ID_3D_NEGX_POSY_CENZ, _NH_ID_3D_CENX_POSY_CENZ, _NH_ID_3D_POSX_POSY_CENZ,
    _NH_ID_3D_NEGX_NEGY_POSZ, _NH_ID_3D_CENX_NEGY_POSZ, _NH_ID_3D_POSX_NEGY_POSZ,
    _NH_ID_3D_NEGX_CENY_POSZ, _NH_ID_3D_CENX_CENY_POSZ, _NH_ID_3D_POSX_CENY_POSZ,
    _NH_ID_3D_NEGX_POSY_POSZ, _NH_ID_3D_CENX_POSY_POSZ, _NH_ID_3D_POSX_POSY_POSZ,
    _NH_ID_3D_TOTAL
};

class decompositionManager3D
{
public:
    // numPts[d][zone]: points of each zone; costs and stencil widths are checked (>= 0) and otherwise unused on one
    // rank. npe must be {1, 1, 1}.
    decompositionManager3D(int numZones[3], int* numPts[3], float*** costs, int*** sw[3][2], int npe[3],
                           MPI_Comm comm, int hack_first_interior_z_rank_max_nz = 0);

    void getSplitNumLocalSubDom(int n[3]) const
    {
        for (int d = 0; d < 3; ++d)
            n[d] = (int)split_[d].size();
    }
    void getSplitLocalSubDomID(int id, int s[3]) const
    {
        const int nx = (int)split_[0].size(), ny = (int)split_[1].size();
        s[0] = id % nx;
        s[1] = (id % (nx * ny)) / nx;
        s[2] = id / (nx * ny);
    }
    int getSubDomIDfromSplit(int ix, int iy, int iz) const
    {
        return (iz * (int)split_[1].size() + iy) * (int)split_[0].size() + ix;
    }
    int getNumPtsSplit(int split, int d) const { return npts_[d][split_[d][split]]; }
    int getOffset(int split, int d) const { return off_[d][split]; }
    void getStartAndEndOffsets(int id, int st[3], int end[3]) const;
    int getTotalNumLocalSubDom() const
    {
        return (int)(split_[0].size() * split_[1].size() * split_[2].size());
    }
    int getNumPts(int d) const { return gN_[d]; }
    int getGlobalNumPts(int d) const { return gN_[d]; }
    // local sub-domain in each of the 27 neighbour slots (-1 when there is none; the centre slot is id itself)
    void getLocalNeighborID(int id, int nh[_NH_ID_3D_TOTAL]) const;
    MPI_Comm get_comm() const { return comm_; }
    int get_rank() const { return 0; }
    int get_npes() const { return 1; }

private:
    MPI_Comm comm_;
    int gN_[3] = { 0, 0, 0 };
    std::vector<int> npts_[3];  // points per zone
    std::vector<int> split_[3]; // zone index of each local split sub-domain
    std::vector<int> off_[3];   // global offset of each split sub-domain
};
