// Size limit of the DIIS history shared by the device solvers.
//
// DIIS (reference SFNewton::DIIS / Ax, src/sfnewton.cpp:793-880) keeps the
// last m error vectors and solves A c = (1,...,1) for their Gram matrix A.
// On the device that small system is solved by a wavefront-parallel pivoted
// elimination (wave_diis_coefficients in scf_device.h); kMaxDIIS bounds m so
// the Gram matrix, its column sums and the coefficients fit in fixed-size
// LDS and register arrays.
#pragma once

namespace nscf {

constexpr int kMaxDIIS = 16;

}  // namespace nscf
