// Check of include/discamb_amd/discamb_binding.hpp against DiSCaMB's own types: builds a
// discamb::Crystal and a discamb::HC_ModelParameters the way a DiSCaMB program does (space group
// from operation strings, wavefunctions from the built-in Clementi-Roetti bank with the UBDB
// deformation-valence defaults), flattens them with the binding and prints every array of the
// resulting dh_model_desc and dh_atoms, so that a test can compare them with the Python path's.
// Built against the reference's headers and the reference sources compiled in oracle/_ref
// (oracle/Makefile.ref); it runs on the host only.
//
// Input (stdin, whitespace separated):
//   cell(6) xyz_convention adp_convention            (fractional|cartesian, U_cif|U_cart)
//   n_ops, then one operation string per line
//   n_wfn, then one bank label per line
//   n_types, then per type: n_l p_val kappa kappa' and P_lm for l < n_l (2l + 1 each)
//   n_atoms, then per atom: wfn type multiplicity occupancy x y z n_adp adp(n_adp) lcs(9, row-major)
//   n_hkl, then per reflection: h k l Re(dT/dF) Im(dT/dF)
#include <cstdio>
#include <iostream>
#include <string>
#include <vector>

#include "discamb/HC_Model/ClementiRoettiData.h"
#include "discamb/HC_Model/DeformationValenceParameters.h"
#include "discamb_amd/discamb_binding.hpp"

using namespace discamb;

template <typename T>
static void dump(const char* name, const T* p, size_t n)
{
    std::printf("%s %zu", name, n);
    for (size_t i = 0; i < n; i++) std::printf(" %.17g", (double)p[i]);
    std::printf("\n");
}

int main()
{
    Crystal crystal;
    double cell[6];
    std::string xc, ac;
    std::cin >> cell[0] >> cell[1] >> cell[2] >> cell[3] >> cell[4] >> cell[5] >> xc >> ac;
    crystal.unitCell.set(cell[0], cell[1], cell[2], cell[3], cell[4], cell[5]);
    crystal.xyzCoordinateSystem = xc == "cartesian" ? structural_parameters_convention::XyzCoordinateSystem::cartesian
                                                    : structural_parameters_convention::XyzCoordinateSystem::fractional;
    crystal.adpConvention = ac == "U_cart" ? structural_parameters_convention::AdpConvention::U_cart
                                           : structural_parameters_convention::AdpConvention::U_cif;
    int n_ops;
    std::cin >> n_ops;
    std::string line;
    std::getline(std::cin, line);
    std::vector<SpaceGroupOperation> ops;
    for (int s = 0; s < n_ops; s++) {
        std::getline(std::cin, line);
        ops.push_back(SpaceGroupOperation(line));
    }
    crystal.spaceGroup.set(ops);

    HC_ModelParameters p;
    int n_wfn;
    std::cin >> n_wfn;
    ClementiRoettiData bank;
    DeformationValenceParameters dv(DeformationValenceParameters::ParametersType::UBDB_CR);
    for (int w = 0; w < n_wfn; w++) {
        std::string label;
        std::cin >> label;
        bool found = false;
        HC_WfnType t;
        t = bank.getEntry(label, found);
        if (!found) {
            std::fprintf(stderr, "no bank entry %s\n", label.c_str());
            return 2;
        }
        if (!dv.getParameters(label, t.deformation_valence_exponent, t.deformation_valence_power)) {
            std::fprintf(stderr, "no deformation-valence parameters for %s\n", label.c_str());
            return 2;
        }
        t.anomalous_scattering = 0.0;
        p.wfn_parameters.push_back(t);
    }
    int n_types;
    std::cin >> n_types;
    for (int k = 0; k < n_types; k++) {
        HC_AtomTypeParameters t;
        int n_l;
        std::cin >> n_l >> t.p_val >> t.kappa_spherical_valence >> t.kappa_deformation_valence;
        t.p_lm.resize(n_l);
        for (int l = 0; l < n_l; l++) {
            t.p_lm[l].resize(2 * l + 1);
            for (int m = 0; m < 2 * l + 1; m++) std::cin >> t.p_lm[l][m];
        }
        p.type_parameters.push_back(t);
    }
    int n_atoms;
    std::cin >> n_atoms;
    std::vector<Matrix3d> frames(n_atoms);
    for (int i = 0; i < n_atoms; i++) {
        AtomInCrystal a;
        int w, t, na;
        std::cin >> w >> t >> a.multiplicity >> a.occupancy >> a.coordinates[0] >> a.coordinates[1] >> a.coordinates[2] >> na;
        a.adp.resize(na);
        for (int j = 0; j < na; j++) std::cin >> a.adp[j];
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 3; c++) std::cin >> frames[i](r, c);
        crystal.atoms.push_back(a);
        p.atom_to_wfn_map.push_back(w);
        p.atom_to_type_map.push_back(t);
    }
    int n_hkl;
    std::cin >> n_hkl;
    std::vector<Vector3i> hkl(n_hkl);
    std::vector<std::complex<double>> dtdf(n_hkl);
    for (int k = 0; k < n_hkl; k++) {
        double re, im;
        std::cin >> hkl[k][0] >> hkl[k][1] >> hkl[k][2] >> re >> im;
        dtdf[k] = std::complex<double>(re, im);
    }
    if (!std::cin) {
        std::fprintf(stderr, "malformed input\n");
        return 2;
    }

    discamb_amd::binding::FlatModel m(crystal, p);
    const dh_model_desc& d = m.desc;
    dump("cell", d.cell, 6);
    dump("symop_rot", d.symop_rot, 9 * (size_t)d.n_symops);
    dump("symop_trans", d.symop_trans, 3 * (size_t)d.n_symops);
    const int conv[4] = {d.xyz_convention, d.adp_convention, d.d_xyz_convention, d.d_adp_convention};
    dump("conventions", conv, 4);
    const size_t n_orb = d.wfn_orbital_offset[d.n_wfn];
    const size_t n_terms = d.orbital_term_offset[n_orb];
    dump("wfn_orbital_offset", d.wfn_orbital_offset, d.n_wfn + 1);
    dump("orbital_term_offset", d.orbital_term_offset, n_orb + 1);
    dump("orbital_nl", d.orbital_nl, 2 * n_orb);
    dump("term_coeff", d.term_coeff, n_terms);
    dump("term_exp", d.term_exp, n_terms);
    dump("term_pow", d.term_pow, n_terms);
    dump("orbital_occupancy", d.orbital_occupancy, n_orb);
    dump("orbital_kind", d.orbital_kind, n_orb);
    dump("def_val_exponent", d.def_val_exponent, d.n_wfn);
    dump("def_val_pow_offset", d.def_val_pow_offset, d.n_wfn + 1);
    dump("def_val_pow", d.def_val_pow, d.def_val_pow_offset[d.n_wfn]);
    dump("anomalous", d.anomalous, 2 * (size_t)d.n_wfn);
    dump("wfn_atomic_number", d.wfn_atomic_number, d.n_wfn);
    std::printf("wfn_label %d", d.n_wfn);
    for (int w = 0; w < d.n_wfn; w++) std::printf(" %s", d.wfn_label[w]);
    std::printf("\n");
    dump("plm_n_l", d.plm_n_l, d.n_types);
    dump("plm", d.plm, 25 * (size_t)d.n_types);
    dump("p_val", d.p_val, d.n_types);
    dump("kappa", d.kappa, d.n_types);
    dump("kappa_prime", d.kappa_prime, d.n_types);
    dump("atom_to_wfn", d.atom_to_wfn, d.n_atoms);
    dump("atom_to_type", d.atom_to_type, d.n_atoms);
    dump("atom_multiplicity", d.atom_multiplicity, d.n_atoms);

    discamb_amd::binding::FlatCall call(crystal.atoms, frames, hkl, dtdf, std::vector<bool>());
    const dh_atoms& a = call.atoms;
    dump("xyz", a.xyz, 3 * (size_t)a.n_atoms);
    dump("adp_size", a.adp_size, a.n_atoms);
    dump("adp", a.adp, 6 * (size_t)a.n_atoms);
    dump("occupancy", a.occupancy, a.n_atoms);
    dump("lcs", a.lcs, 9 * (size_t)a.n_atoms);
    dump("include", a.include, a.n_atoms);
    dump("hkl", call.hkl.data(), call.hkl.size());
    dump("dtdf", call.dtdf.data(), call.dtdf.size());
    return 0;
}
