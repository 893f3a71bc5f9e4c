// Reference-side binding: DiSCaMB's own model and structure types flattened into the C ABI of
// include/discamb_amd.h.  A DiSCaMB build that adds the "HIP" engine includes this header (it needs
// DiSCaMB's include directory) inside HansenCoppensStructureFactorCalculator
// (src/Scattering/HansenCoppensStructureFactorCalculator.cpp, engine switch at :417-502):
//
//   discamb_amd::binding::FlatModel model(crystal, hcParameters);      // after setModel (.cpp:157-290)
//   dh_hc_create(&model.desc, device, &hip);
//   discamb_amd::binding::FlatCall call(atoms, localCoordinateSystems, hkl, dTarget_df, countAtomContribution);
//   dh_hc_sf_and_derivs(hip, &call.atoms, call.n_hkl, call.hkl.data(), call.dtdf.data(), DH_D_ALL,
//                       f2, d_xyz, d_adp, d_occ);
//
// Flattened from discamb::Crystal (Crystal.h: unit cell, space group, atoms, conventions) and
// discamb::HC_ModelParameters (HC_ModelParameters.h:22-28: wfn_parameters, type_parameters,
// atom_to_wfn_map, atom_to_type_map), with HC_WfnType (HC_WfnData.h:126-163) and
// HC_AtomTypeParameters (HC_AtomTypeParameters.h:19-50).  The flat structs own every array the
// descriptors point into; keep them alive while the descriptors are used.
#pragma once

#include <complex>
#include <stdexcept>
#include <string>
#include <vector>

#include "discamb/CrystalStructure/Crystal.h"
#include "discamb/HC_Model/HC_ModelParameters.h"
#include "discamb/MathUtilities/Matrix3.h"
#include "discamb/MathUtilities/Vector3.h"

#include "../discamb_amd.h"

namespace discamb_amd {
namespace binding {

inline int xyz_convention(discamb::structural_parameters_convention::XyzCoordinateSystem c)
{
    return c == discamb::structural_parameters_convention::XyzCoordinateSystem::cartesian ? DH_XYZ_CARTESIAN
                                                                                         : DH_XYZ_FRACTIONAL;
}

inline int adp_convention(discamb::structural_parameters_convention::AdpConvention c)
{
    if (c == discamb::structural_parameters_convention::AdpConvention::U_cart) return DH_ADP_U_CART;
    if (c == discamb::structural_parameters_convention::AdpConvention::U_cif) return DH_ADP_U_CIF;
    throw std::invalid_argument("discamb_amd: ADPs as U* are not supported; convert to U_cif or U_cart");
}

// Crystal + HC_ModelParameters -> dh_model_desc.  Derivative conventions default to the crystal's
// input conventions, as the reference calculator's constructor sets them (.cpp:57-73).
struct FlatModel {
    dh_model_desc desc{};
    std::vector<int> rot, wfn_orbital_offset, orbital_term_offset, orbital_nl, term_pow, orbital_occupancy,
        orbital_kind, def_val_pow_offset, def_val_pow, plm_n_l, atom_to_wfn, atom_to_type, wfn_z;
    std::vector<double> trans, term_coeff, term_exp, def_val_exponent, anomalous, plm, p_val, kappa, kappa_prime,
        multiplicity;
    std::vector<std::string> labels;
    std::vector<const char*> label_ptr;

    FlatModel(const discamb::Crystal& crystal, const discamb::HC_ModelParameters& p)
        : FlatModel(crystal, p, xyz_convention(crystal.xyzCoordinateSystem), adp_convention(crystal.adpConvention))
    {
    }

    FlatModel(const discamb::Crystal& crystal, const discamb::HC_ModelParameters& p, int d_xyz, int d_adp)
    {
        crystal.unitCell.get(desc.cell[0], desc.cell[1], desc.cell[2], desc.cell[3], desc.cell[4], desc.cell[5]);
        const int n_ops = crystal.spaceGroup.nSymmetryOperations();
        for (int s = 0; s < n_ops; s++) {
            discamb::Matrix3<int> r;
            discamb::Vector3d t;
            discamb::Matrix3d rd;
            crystal.spaceGroup.getSpaceGroupOperation(s).getRotation(r);
            crystal.spaceGroup.getSpaceGroupOperation(s).get(rd, t);
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) rot.push_back(r(i, j));
            for (int i = 0; i < 3; i++) trans.push_back(t[i]);
        }
        // wavefunction types: orbitals, Slater terms, core / valence split, def-val, f' + i f''
        wfn_orbital_offset.push_back(0);
        orbital_term_offset.push_back(0);
        def_val_pow_offset.push_back(0);
        for (const discamb::HC_WfnType& w : p.wfn_parameters) {
            for (size_t o = 0; o < w.orbitals.size(); o++) {
                const discamb::SlaterTypeAtomicOrbitalRdf& orb = w.orbitals[o];
                orbital_nl.push_back(orb.principal_number);
                orbital_nl.push_back(orb.azimuthal_number);
                for (size_t k = 0; k < orb.coefficients.size(); k++) {
                    term_coeff.push_back(orb.coefficients[k]);
                    term_exp.push_back(orb.exponents[k]);
                    term_pow.push_back(orb.power_r[k]);
                }
                orbital_term_offset.push_back((int)term_coeff.size());
                orbital_occupancy.push_back(o < w.orbital_occupancy.size() ? w.orbital_occupancy[o] : 0);
                int kind = -1;
                for (int c : w.core_orbitals_indices)
                    if (c == (int)o) kind = 0;
                for (int v : w.valence_orbitals_indices)
                    if (v == (int)o) kind = 1;
                orbital_kind.push_back(kind);
            }
            wfn_orbital_offset.push_back((int)orbital_nl.size() / 2);
            def_val_exponent.push_back(w.deformation_valence_exponent);
            for (int n : w.deformation_valence_power) def_val_pow.push_back(n);
            def_val_pow_offset.push_back((int)def_val_pow.size());
            anomalous.push_back(w.anomalous_scattering.real());
            anomalous.push_back(w.anomalous_scattering.imag());
            labels.push_back(w.wfnAtomTypeName);
            wfn_z.push_back(w.atomic_number);
        }
        for (const std::string& l : labels) label_ptr.push_back(l.c_str());
        // atom types: P_lm [l][l + m] -> l*l + l + m (density normalised, local frame)
        for (const discamb::HC_AtomTypeParameters& t : p.type_parameters) {
            if (t.p_lm.size() > 5) throw std::invalid_argument("discamb_amd: multipoles beyond l = 4");
            plm_n_l.push_back((int)t.p_lm.size());
            std::vector<double> row(25, 0.0);
            for (size_t l = 0; l < t.p_lm.size(); l++)
                for (size_t j = 0; j < t.p_lm[l].size() && j <= 2 * l; j++) row[l * l + j] = t.p_lm[l][j];
            plm.insert(plm.end(), row.begin(), row.end());
            p_val.push_back(t.p_val);
            kappa.push_back(t.kappa_spherical_valence);
            kappa_prime.push_back(t.kappa_deformation_valence);
        }
        atom_to_wfn = p.atom_to_wfn_map;
        atom_to_type = p.atom_to_type_map;
        for (const discamb::AtomInCrystal& a : crystal.atoms) multiplicity.push_back(a.multiplicity);
        if (multiplicity.size() != atom_to_type.size())
            throw std::invalid_argument("discamb_amd: the crystal's atoms and the model's atom-to-type map differ");
        desc.n_symops = n_ops;
        desc.symop_rot = rot.data();
        desc.symop_trans = trans.data();
        desc.xyz_convention = xyz_convention(crystal.xyzCoordinateSystem);
        desc.adp_convention = adp_convention(crystal.adpConvention);
        desc.d_xyz_convention = d_xyz;
        desc.d_adp_convention = d_adp;
        desc.n_wfn = (int)p.wfn_parameters.size();
        desc.wfn_orbital_offset = wfn_orbital_offset.data();
        desc.orbital_term_offset = orbital_term_offset.data();
        desc.orbital_nl = orbital_nl.data();
        desc.term_coeff = term_coeff.data();
        desc.term_exp = term_exp.data();
        desc.term_pow = term_pow.data();
        desc.orbital_occupancy = orbital_occupancy.data();
        desc.orbital_kind = orbital_kind.data();
        desc.def_val_exponent = def_val_exponent.data();
        desc.def_val_pow_offset = def_val_pow_offset.data();
        desc.def_val_pow = def_val_pow.data();
        desc.anomalous = anomalous.data();
        desc.n_types = (int)p.type_parameters.size();
        desc.plm_n_l = plm_n_l.data();
        desc.plm = plm.data();
        desc.p_val = p_val.data();
        desc.kappa = kappa.data();
        desc.kappa_prime = kappa_prime.data();
        desc.n_atoms = (int)atom_to_type.size();
        desc.atom_to_wfn = atom_to_wfn.data();
        desc.atom_to_type = atom_to_type.data();
        desc.atom_multiplicity = multiplicity.data();
        desc.wfn_label = label_ptr.data();
        desc.wfn_atomic_number = wfn_z.data();
    }
    FlatModel(const FlatModel&) = delete;
    FlatModel& operator=(const FlatModel&) = delete;
};

// One calculateStructureFactorsAndDerivatives call's inputs (.h:148-165): vector<AtomInCrystal>,
// the local frames (vector<Matrix3d>, columns = local axes), vector<Vector3i> hkl,
// vector<complex<double>> dTarget/dF and vector<bool> countAtomContribution.
struct FlatCall {
    dh_atoms atoms{};
    int n_hkl = 0;
    std::vector<double> xyz, adp, occ, lcs, dtdf;
    std::vector<int> adp_size, hkl;
    std::vector<unsigned char> include;

    FlatCall(const std::vector<discamb::AtomInCrystal>& at, const std::vector<discamb::Matrix3d>& frames,
             const std::vector<discamb::Vector3i>& h, const std::vector<std::complex<double>>& dt_df,
             const std::vector<bool>& count_atom)
    {
        const size_t n = at.size();
        if (frames.size() != n || (!count_atom.empty() && count_atom.size() != n))
            throw std::invalid_argument("discamb_amd: per-atom arrays of different sizes");
        adp.assign(6 * n, 0.0);
        for (size_t i = 0; i < n; i++) {
            for (int j = 0; j < 3; j++) xyz.push_back(at[i].coordinates[j]);
            const size_t na = at[i].adp.size();
            if (na != 0 && na != 1 && na != 6) throw std::invalid_argument("discamb_amd: ADPs with 0, 1 or 6 components");
            adp_size.push_back((int)na);
            for (size_t j = 0; j < na; j++) adp[6 * i + j] = at[i].adp[j];
            occ.push_back(at[i].occupancy);
            for (int r = 0; r < 3; r++)
                for (int c = 0; c < 3; c++) lcs.push_back(frames[i](r, c));
            include.push_back(count_atom.empty() || count_atom[i] ? 1 : 0);
        }
        n_hkl = (int)h.size();
        for (const discamb::Vector3i& v : h)
            for (int j = 0; j < 3; j++) hkl.push_back(v[j]);
        for (const std::complex<double>& g : dt_df) {
            dtdf.push_back(g.real());
            dtdf.push_back(g.imag());
        }
        if ((int)dt_df.size() != n_hkl && !dt_df.empty())
            throw std::invalid_argument("discamb_amd: dT/dF and hkl of different sizes");
        atoms.n_atoms = (int)n;
        atoms.xyz = xyz.data();
        atoms.adp_size = adp_size.data();
        atoms.adp = adp.data();
        atoms.occupancy = occ.data();
        atoms.lcs = lcs.data();
        atoms.anomalous = nullptr;
        atoms.include = include.data();
    }
    FlatCall(const FlatCall&) = delete;
    FlatCall& operator=(const FlatCall&) = delete;
};

}  // namespace binding
}  // namespace discamb_amd
