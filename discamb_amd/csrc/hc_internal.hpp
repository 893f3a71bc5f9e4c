// Internal host-side types of the discamb_amd library (not part of the C ABI).
#pragma once

#include <array>
#include <complex>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace discamb_amd {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// ---------------------------------------------------------------- crystal conventions
// Orthogonalisation with a || x and c* || z (International Tables vol. B 3.3), the convention of
// DiSCaMB's UnitCell::calculateTransformationMatrices (src/CrystalStructure/UnitCell.cpp:183-212).
struct Cell {
    double par[6] = {1, 1, 1, 90, 90, 90};
    double f2c[9] = {0}, c2f[9] = {0};  // row-major
    double rec_len[3] = {0};            // |a*|, |b*|, |c*|
    void set(const double* p6);
    // U_cart = A U_cif A^T with A = f2c diag(|a*|,|b*|,|c*|)
    void u_cif_to_cart(const double* ucif6, double* ucart6) const;
    // chain rule dT/dU_cif from dT/dU_cart (6-component parameter convention: off-diagonal
    // parameters enter the symmetric tensor twice)
    void d_ucart_to_d_ucif(const double* d_cart6, double* d_cif6) const;
    void u_cart_to_cif(const double* ucart6, double* ucif6) const;
};

struct SymOp {
    int r[9];        // fractional rotation, row-major
    double t[3];     // fractional translation
};

// Space group split into an explicit subset of coset representatives, lattice centering and an
// optional inversion centre - the representation DiSCaMB's SpaceGroup exposes through
// nSymmetryOperationsInSubset / centering / inversionCenterTranslation (SpaceGroup.cpp:388-420).
struct SpaceGroupInfo {
    int order = 1;                 // |G|
    std::vector<SymOp> subset;     // explicit operations
    bool centro = false;
    double inv_t[3] = {0, 0, 0};   // translation part of the inversion operation (fractional)
    char centering = 'P';
    bool obverse = true;
    std::vector<std::array<double, 3>> centering_vectors;
};
SpaceGroupInfo analyse_space_group(const std::vector<SymOp>& ops);

// Per-reflection centering multipliers (scattering_utilities.cpp:103-170), including that
// routine's treatment of reverse R centering.
double centering_multiplier(char centering, bool obverse, const int* hkl);

// ---------------------------------------------------------------- Hansen-Coppens model
struct SlaterOrbital {
    int n = 1, l = 0;
    std::vector<double> coeff, expo;
    std::vector<int> pow;
};

struct WfnType {
    std::string label;
    std::vector<SlaterOrbital> orbitals;
    std::vector<int> occupancy;
    std::vector<int> core, valence;       // orbital indices
    double def_val_exponent = 1.0;
    std::vector<int> def_val_power;       // n_l, size 0..5
    std::complex<double> anomalous{0, 0};
};

struct AtomType {
    int n_l = 0;                 // size of p_lm
    double plm[25] = {0};        // [l*l + l + m], density normalised, local frame
    double p_val = 0, kappa = 1, kappa_prime = 1;
};

// Spherical Slater density sum_k c_k r^p_k exp(-e_k r) of one wavefunction type.
struct SlaterDensity {
    std::vector<double> c, e;
    std::vector<int> p;
};

// Flattened, engine-ready Hansen-Coppens model (HansenCoppensStructureFactorCalculator::setModel,
// .cpp:157-290, and Engine3::pre_hkl_loop_sf_calc, Engine3.cpp:195-256).
struct EngineModel {
    int n_wfn = 0, n_types = 0;
    std::vector<SlaterDensity> core, valence;     // per wfn type; valence / valence occupancy
    std::vector<double> dv_zeta;                  // per wfn
    std::vector<int> dv_n;                        // per wfn, 5
    std::vector<int> dv_nl;                       // per wfn
    std::vector<double> anom;                     // per wfn, 2
    std::vector<AtomType> types;
    std::vector<int> type_wfn;                    // type -> wfn (last atom wins, as Engine3)
    std::vector<int> type_max_l;                  // min(4, n_dv - 1, highest non-zero l)
    // Engine3 takes the core term from the atom's own wavefunction but the valence and
    // deformation terms from the wavefunction of the type's last atom.  Atoms whose wavefunction
    // differs from that one get an engine type of their own (appended after the model's types,
    // same parameters) over an engine wavefunction combining their core with the type's valence.
    int n_model_types = 0;
    std::vector<int> type_origin;                 // engine type -> model type
    std::vector<int> wfn_origin;                  // engine wfn -> model wfn of its core, f', f'', Z
};

void build_sphere_density(const WfnType& w, bool valence, SlaterDensity& out);

// N-Gaussian atomic form factors (generated tables, ngauss_tables.h)
struct NGaussEntry;
const NGaussEntry* ngauss_lookup(const std::string& label, const std::string& table, int* n_gauss);
double ngauss_value(const NGaussEntry& e, int n, double h);
std::string element_symbol(int z);
// engine_atom_type: the engine type of every atom (the model type unless split as above)
EngineModel build_engine_model(const std::vector<WfnType>& wfn, const std::vector<AtomType>& types,
                               const std::vector<int>& atom_wfn, const std::vector<int>& atom_type,
                               std::vector<int>* engine_atom_type = nullptr);
// Independent-atom model (N-Gaussians of the named table per wavefunction label, Waasmaier-Kirfel
// by default); atom_type is set to the atom-to-wavefunction map.
EngineModel build_iam_model(const std::vector<std::string>& labels, const std::vector<std::complex<double>>& anom,
                            const std::vector<int>& atom_wfn, std::vector<int>& atom_type,
                            const std::string& table = "Waasmeier-Kirfel");

}  // namespace discamb_amd
