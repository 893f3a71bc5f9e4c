// Local coordinate systems from atomic positions and their position derivatives (host).
#pragma once

#include <vector>

#include "hc_internal.hpp"

namespace discamb_amd {

// direction types (LcsDirectionType, LocalCoordinateSystem.h:17)
constexpr int DH_LCS_ANY_ORTHOGONAL_ = 0;
constexpr int DH_LCS_AVERAGE_DIRECTION_ = 1;
constexpr int DH_LCS_AVERAGE_POSITION_ = 2;
// atoms defining one frame at most (central + reference atoms of both directions)
constexpr int kLcsMaxAtoms = 10;

// an atom of the asymmetric unit under a symmetry operation (AtomInCrystalID)
struct LcsRef {
    int atom = 0;
    int rot[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};   // fractional rotation, row-major
    double trans[3] = {0, 0, 0};                // fractional translation
};

struct LcsDef {
    LcsRef central;
    int dir1 = DH_LCS_ANY_ORTHOGONAL_, dir2 = DH_LCS_ANY_ORTHOGONAL_;
    std::vector<LcsRef> ref1, ref2, chirality;
    int c1 = 0, c2 = 1;     // axis along direction 1, axis in the plane of directions 1 and 2
    bool is_r = true;
};

void lcs_validate(const LcsDef& d, int n_atoms);
// m9: row-major 3x3 with the local axes as columns, Cartesian (f2c: fractional -> Cartesian)
void lcs_matrix(const LcsDef& d, const double* xyz_frac, const double* f2c, double* m9);
// sum_k dT_dL[k] dL[k] / dx_frac[atom][c] for the atoms defining the frame: writes their indices to
// atoms[] and the values to grad[3 * slot + c] (both kLcsMaxAtoms long); returns the atom count
int lcs_chain(const LcsDef& d, const double* xyz_frac, const double* f2c, const double* dT_dL, int* atoms,
              double* grad);

}  // namespace discamb_amd
