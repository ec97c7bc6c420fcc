// One-rank communicator: the type and the few calls the reference's driver and infrastructure constructors make
// (MPI_Comm arguments of simulator2SetUp, decompositionManager3D, file_snap_factory::create). Every rank query answers
// rank 0 of 1. When a real MPI header is included first, its definitions are used instead.
#pragma once

#ifndef MPI_VERSION
typedef int MPI_Comm;
#define MPI_COMM_WORLD ((MPI_Comm)0)
#define MPI_COMM_NULL ((MPI_Comm)-1)

inline int
MPI_Comm_rank(MPI_Comm, int* rank)
{
    *rank = 0;
    return 0;
}

inline int
MPI_Comm_size(MPI_Comm, int* size)
{
    *size = 1;
    return 0;
}
#endif
