#include "decomp.h"

#include <stdexcept>

decompositionManager3D::decompositionManager3D(int numZones[3], int* numPts[3], float*** costs, int*** sw[3][2],
                                               int npe[3], MPI_Comm comm, int)
    : comm_(comm)
{
    for (int d = 0; d < 3; ++d) {
        if (numZones[d] < 1)
            throw std::runtime_error("decompositionManager3D: at least one zone per axis");
        if (npe[d] != 1)
            throw std::runtime_error("decompositionManager3D: one rank per axis (this build runs one GPU per process)");
    }
    // basic checks on the inputs (infrastructure/decompositionManager3D.cpp:16-33)
    for (int k = 0; k < numZones[2]; ++k)
        for (int j = 0; j < numZones[1]; ++j)
            for (int i = 0; i < numZones[0]; ++i) {
                if (costs && costs[k][j][i] < 0.0f)
                    throw std::runtime_error("decompositionManager3D: negative cost");
                for (int a = 0; a < 3; ++a)
                    for (int s = 0; s < 2; ++s)
                        if (sw && sw[a][s] && sw[a][s][k][j][i] < 0)
                            throw std::runtime_error("decompositionManager3D: negative stencil width");
            }
    for (int d = 0; d < 3; ++d) {
        int off = 0;
        for (int z = 0; z < numZones[d]; ++z) {
            const int n = numPts[d][z];
            if (n < 0)
                throw std::runtime_error("decompositionManager3D: negative zone size");
            npts_[d].push_back(n);
            if (n > 0) {
                split_[d].push_back(z);
                off_[d].push_back(off);
            }
            off += n;
        }
        gN_[d] = off;
        if (split_[d].empty())
            throw std::runtime_error("decompositionManager3D: empty axis");
    }
}

void
decompositionManager3D::getStartAndEndOffsets(int id, int st[3], int end[3]) const
{
    // end offsets count the points after the sub-domain (infrastructure/decompositionManager3D.h:166-174)
    int s[3];
    getSplitLocalSubDomID(id, s);
    for (int d = 0; d < 3; ++d) {
        st[d] = off_[d][s[d]];
        end[d] = gN_[d] - (st[d] + getNumPtsSplit(s[d], d));
    }
}

void
decompositionManager3D::getLocalNeighborID(int id, int nh[_NH_ID_3D_TOTAL]) const
{
    int s[3];
    getSplitLocalSubDomID(id, s);
    for (int slot = 0; slot < _NH_ID_3D_TOTAL; ++slot) {
        const int t[3] = { s[0] + slot % 3 - 1, s[1] + (slot / 3) % 3 - 1, s[2] + slot / 9 - 1 };
        bool inside = true;
        for (int d = 0; d < 3; ++d)
            inside = inside && t[d] >= 0 && t[d] < (int)split_[d].size();
        nh[slot] = inside ? getSubDomIDfromSplit(t[0], t[1], t[2]) : -1;
    }
}
