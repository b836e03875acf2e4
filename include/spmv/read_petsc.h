// spmv::read_petsc_binary / read_petsc_binary_vector (the interface of
// spmv/read_petsc.h:10-31), backed by the spmvx C ABI.
//
// One process drives one rank: with a one-process communicator the file is
// read whole; the overloads taking (nranks, rank) read one rank of an
// nranks-way row split (owner_ranges, read_petsc.cpp:13-30), with the
// exchange plan of that rank in the column map.
#pragma once

#include <string>
#include <vector>

#include "Matrix.h"

namespace spmv
{

/// Binary PETSc matrix file (32-bit indices, big-endian).  Prints
/// "Read file: <name>: <rows>x<cols> = <nnz>" as the reference does.
Matrix<double> read_petsc_binary(MPI_Comm comm, std::string filename);
Matrix<double> read_petsc_binary(MPI_Comm comm, std::string filename, int nranks, int rank);

/// Binary PETSc vector file: this rank's part (host values; the reference
/// returns an Eigen::VectorXd).
std::vector<double> read_petsc_binary_vector(MPI_Comm comm, std::string filename);
std::vector<double> read_petsc_binary_vector(MPI_Comm comm, std::string filename, int nranks,
                                             int rank);

} // namespace spmv
