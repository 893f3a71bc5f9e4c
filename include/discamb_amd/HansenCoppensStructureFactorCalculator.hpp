// C++ drop-in for DiSCaMB's HansenCoppensStructureFactorCalculator, on the discamb_amd C ABI.
//
// Mirrors include/discamb/Scattering/HansenCoppensStructureFactorCalculator.h:36-375 (constructor
// from Crystal + HC_ModelParameters, setCalculationMode, calculateStructureFactorsAndDerivatives
// with and without DerivativesSelector, calculateStructureFactors) and the value types it takes
// (Crystal.h, AtomInCrystal, HC_ModelParameters.h:22-28, HC_WfnData.h:126-163,
// HC_AtomTypeParameters.h:19-50, SF_CalcDataTypes.h:22-56), including the per-hkl template
// overload that hands each reflection's SfDerivativesAtHkl to a caller-supplied processor.  Errors throw
// discamb_amd::Exception, as the reference throws discamb::Exception.  Link with
// -ldiscamb_amd (discamb_amd/libdiscamb_amd.so).
#pragma once

#include <array>
#include <cctype>
#include <complex>
#include <cstdlib>
#include <exception>
#include <stdexcept>
#include <string>
#include <vector>

#include "../discamb_amd.h"

namespace discamb_amd {

struct Exception : std::runtime_error {
    int code;
    Exception(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

struct Vector3i {
    int x = 0, y = 0, z = 0;
    Vector3i() = default;
    Vector3i(int a, int b, int c) : x(a), y(b), z(c) {}
    int operator[](int i) const { return i == 0 ? x : i == 1 ? y : z; }
};

template <typename T>
struct Vector3 {
    T x{}, y{}, z{};
    Vector3() = default;
    Vector3(T a, T b, T c) : x(a), y(b), z(c) {}
    T operator[](int i) const { return i == 0 ? x : i == 1 ? y : z; }
    T& operator[](int i) { return i == 0 ? x : i == 1 ? y : z; }
};
using Vector3d = Vector3<double>;

// 3x3 matrix; element (i, j) is row i, column j.  Local coordinate systems hold the local x, y, z
// axes (in Cartesian coordinates) as columns.
struct Matrix3d {
    std::array<double, 9> m{{1, 0, 0, 0, 1, 0, 0, 0, 1}};
    double& operator()(int i, int j) { return m[3 * i + j]; }
    double operator()(int i, int j) const { return m[3 * i + j]; }
};

namespace structural_parameters_convention {
enum class XyzCoordinateSystem { fractional, cartesian };
enum class AdpConvention { U_cif, U_cart };
}  // namespace structural_parameters_convention

struct UnitCell {
    double a = 1, b = 1, c = 1, alpha = 90, beta = 90, gamma = 90;
};

struct AtomInCrystal {
    std::string label, type;
    Vector3d coordinates;
    std::vector<double> adp;        // 1 (U_iso) or 6 (U11 U22 U33 U12 U13 U23) or empty
    double occupancy = 1.0;
    double multiplicity = 1.0;
};

struct Crystal {
    UnitCell unitCell;
    std::vector<std::string> spaceGroupOperations;   // full group, e.g. "-x,y+1/2,-z+1/2"
    std::vector<AtomInCrystal> atoms;
    structural_parameters_convention::XyzCoordinateSystem xyzCoordinateSystem =
        structural_parameters_convention::XyzCoordinateSystem::fractional;
    structural_parameters_convention::AdpConvention adpConvention =
        structural_parameters_convention::AdpConvention::U_cif;
};

struct SlaterTypeAtomicOrbitalRdf {
    int principal_number = 1, azimuthal_number = 0;
    std::vector<int> power_r;
    std::vector<double> exponents, coefficients;
};

struct HC_WfnType {
    std::string wfnAtomTypeName;
    int atomic_number = 0, charge = 0;
    std::vector<SlaterTypeAtomicOrbitalRdf> orbitals;
    std::vector<int> orbital_occupancy, core_orbitals_indices, valence_orbitals_indices;
    double deformation_valence_exponent = 1.0;
    std::vector<int> deformation_valence_power;
    std::complex<double> anomalous_scattering{0, 0};
};

struct HC_AtomTypeParameters {
    std::vector<std::vector<double>> p_lm;   // p_lm[l][l + m]
    double p_val = 0, kappa_spherical_valence = 1, kappa_deformation_valence = 1;
};

struct HC_ModelParameters {
    std::vector<HC_WfnType> wfn_parameters;
    std::vector<HC_AtomTypeParameters> type_parameters;
    std::vector<int> atom_to_wfn_map, atom_to_type_map;
};

struct TargetFunctionAtomicParamDerivatives {
    std::vector<double> adp_derivatives;
    Vector3d atomic_position_derivatives;
    double occupancy_derivatives = 0;
};

// dF/dp of one reflection for every atom (SF_CalcDataTypes.h:49-55); the field names keep the
// reference's spelling.
struct SfDerivativesAtHkl {
    std::vector<Vector3<std::complex<double>>> atomicPostionDerivatives;
    std::vector<std::vector<std::complex<double>>> adpDerivatives;
    std::vector<std::complex<double>> occupancyDerivatives;
    std::vector<std::complex<double>> anomalousScatteringDerivatives;
};

struct DerivativesSelector {
    bool d_xyz = true, d_adp = true, d_occ = true, d_anom = false;
};

// An atom of the asymmetric unit under a symmetry operation (AtomInCrystalID).
struct AtomInCrystalID {
    int index = 0;
    std::string operation = "x,y,z";
};

enum class LcsDirectionType { ANY_ORTHOGONAL = DH_LCS_ANY_ORTHOGONAL, AVERAGE_DIRECTION = DH_LCS_AVERAGE_DIRECTION,
                              AVERAGE_POSITION = DH_LCS_AVERAGE_POSITION };

// Definition of one atom's local coordinate system from atomic positions
// (LocalCoordinateSystem<AtomInCrystalID>, AtomTyping/LocalCoordinateSystem.h:14-45): X, Y, Z =
// coordinate 0, 1, 2; coordinate_1 along direction 1, coordinate_2 in the plane of both
// directions; chirality empty or three atoms.
struct LocalCoordinateSystem {
    bool isR = true;
    LcsDirectionType direction1_type = LcsDirectionType::ANY_ORTHOGONAL;
    LcsDirectionType direction2_type = LcsDirectionType::ANY_ORTHOGONAL;
    AtomInCrystalID centralAtom;
    std::vector<AtomInCrystalID> refPoint_1, refPoint_2;
    int coordinate_1 = 0, coordinate_2 = 1;
    std::vector<AtomInCrystalID> chirality;
};

namespace detail {
inline void check(int rc)
{
    if (rc != DH_OK) throw Exception(rc, dh_last_error());
}

// "x", "-y+1/2", "z-1/4" ... -> row of the rotation and translation
inline void parse_component(const std::string& s, int* row, double* t)
{
    row[0] = row[1] = row[2] = 0;
    *t = 0;
    size_t i = 0;
    while (i < s.size()) {
        int sign = 1;
        while (i < s.size() && (s[i] == '+' || s[i] == '-' || s[i] == ' ')) {
            if (s[i] == '-') sign = -sign;
            i++;
        }
        size_t j = i;
        while (j < s.size() && s[j] != '+' && s[j] != '-') j++;
        std::string term = s.substr(i, j - i);
        while (!term.empty() && term.back() == ' ') term.pop_back();
        if (!term.empty()) {
            char c = (char)std::tolower(term.back());
            if (c == 'x' || c == 'y' || c == 'z') {
                int k = c - 'x';
                std::string coef = term.substr(0, term.size() - 1);
                row[k] += sign * (coef.empty() || coef == "*" ? 1 : std::atoi(coef.c_str()));
            } else {
                size_t sl = term.find('/');
                double v = sl == std::string::npos ? std::atof(term.c_str())
                                                   : std::atof(term.substr(0, sl).c_str()) /
                                                         std::atof(term.substr(sl + 1).c_str());
                *t += sign * v;
            }
        }
        i = j;
    }
}
// "-x+1/2,y,-z" -> fractional rotation (row-major) and translation
inline void parse_operation(const std::string& op, int* rot, double* tr)
{
    std::string parts[3];
    int k = 0;
    for (char ch : op) {
        if (ch == ',') {
            if (++k > 2) throw Exception(DH_E_MODEL, "invalid symmetry operation " + op);
        } else
            parts[k] += ch;
    }
    if (k != 2) throw Exception(DH_E_MODEL, "invalid symmetry operation " + op);
    for (int r = 0; r < 3; r++) parse_component(parts[r], rot + 3 * r, tr + r);
}

// AtomInCrystal values, local frames and countAtomContribution flattened for the C ABI.
struct PackedAtoms {
    std::vector<double> xyz, adp, occ, lcs;
    std::vector<int> adp_n;
    std::vector<unsigned char> inc;
    dh_atoms da{};
    PackedAtoms(const std::vector<AtomInCrystal>& atoms, const std::vector<Matrix3d>& localCoordinateSystems,
                const std::vector<bool>& countAtomContribution)
    {
        const size_t n = atoms.size();
        if (localCoordinateSystems.size() != n || countAtomContribution.size() != n)
            throw Exception(DH_E_ARGUMENT, "inconsistent number of atoms");
        xyz.resize(3 * n + 3);
        adp.assign(6 * n + 6, 0.0);
        occ.resize(n + 1);
        lcs.resize(9 * n + 9);
        adp_n.resize(n + 1);
        inc.resize(n + 1);
        for (size_t i = 0; i < n; i++) {
            const AtomInCrystal& a = atoms[i];
            for (int j = 0; j < 3; j++) xyz[3 * i + j] = a.coordinates[j];
            adp_n[i] = (int)a.adp.size();
            for (size_t j = 0; j < a.adp.size() && j < 6; j++) adp[6 * i + j] = a.adp[j];
            occ[i] = a.occupancy;
            for (int j = 0; j < 9; j++) lcs[9 * i + j] = localCoordinateSystems[i].m[j];
            inc[i] = countAtomContribution[i] ? 1 : 0;
        }
        da.n_atoms = (int)n;
        da.xyz = xyz.data();
        da.adp_size = adp_n.data();
        da.adp = adp.data();
        da.occupancy = occ.data();
        da.lcs = lcs.data();
        da.anomalous = nullptr;
        da.include = inc.data();
    }
};

inline std::vector<int> flat_hkl(const std::vector<Vector3i>& hkl)
{
    std::vector<int> h(3 * hkl.size() + 3);
    for (size_t k = 0; k < hkl.size(); k++) {
        h[3 * k] = hkl[k].x;
        h[3 * k + 1] = hkl[k].y;
        h[3 * k + 2] = hkl[k].z;
    }
    return h;
}

// Adapts a singleHklResultsProcessor to the C callback of dh_hc_per_hkl.
template <typename T>
struct PerHklContext {
    T* processor;
    SfDerivativesAtHkl d;
    size_t n;
    std::exception_ptr error;

    static int call(void* user, int k, const double* f, const double* row)
    {
        PerHklContext& c = *static_cast<PerHklContext*>(user);
        try {
            const size_t n = c.n;
            auto z = [&](int plane, size_t a) {
                const double* v = row + 2 * (plane * n + a);
                return std::complex<double>(v[0], v[1]);
            };
            for (size_t a = 0; a < n; a++) {
                c.d.atomicPostionDerivatives[a] = Vector3<std::complex<double>>(z(0, a), z(1, a), z(2, a));
                std::vector<std::complex<double>>& ad = c.d.adpDerivatives[a];
                for (size_t j = 0; j < ad.size(); j++) ad[j] = z(3 + (int)j, a);
                c.d.occupancyDerivatives[a] = z(9, a);
            }
            std::complex<double> F(f[0], f[1]);
            c.processor->onPerHklCalculation(k, F, c.d);
            return 0;
        } catch (...) {
            c.error = std::current_exception();
            return 1;
        }
    }
};
}  // namespace detail

class HansenCoppensStructureFactorCalculator {
public:
    HansenCoppensStructureFactorCalculator(const Crystal& crystal, const HC_ModelParameters& parameters,
                                           int device = 0)
        : mAtomCount(crystal.atoms.size())
    {
        namespace spc = structural_parameters_convention;
        const UnitCell& uc = crystal.unitCell;
        std::vector<int> rot;
        std::vector<double> tr;
        for (const std::string& op : crystal.spaceGroupOperations) {
            int r9[9];
            double t3[3];
            detail::parse_operation(op, r9, t3);
            rot.insert(rot.end(), r9, r9 + 9);
            tr.insert(tr.end(), t3, t3 + 3);
        }
        // flatten the model
        std::vector<int> wo{0}, to{0}, nl, tp, occ, kind, dvo{0}, dvp;
        std::vector<double> cc, ce, dve, an;
        for (const HC_WfnType& w : parameters.wfn_parameters) {
            for (size_t i = 0; i < w.orbitals.size(); i++) {
                const auto& o = w.orbitals[i];
                nl.push_back(o.principal_number);
                nl.push_back(o.azimuthal_number);
                cc.insert(cc.end(), o.coefficients.begin(), o.coefficients.end());
                ce.insert(ce.end(), o.exponents.begin(), o.exponents.end());
                tp.insert(tp.end(), o.power_r.begin(), o.power_r.end());
                to.push_back((int)cc.size());
                occ.push_back(i < w.orbital_occupancy.size() ? w.orbital_occupancy[i] : 0);
                int kd = -1;
                for (int c : w.core_orbitals_indices)
                    if (c == (int)i) kd = 0;
                for (int v : w.valence_orbitals_indices)
                    if (v == (int)i) kd = 1;
                kind.push_back(kd);
            }
            wo.push_back((int)occ.size());
            dve.push_back(w.deformation_valence_exponent);
            dvp.insert(dvp.end(), w.deformation_valence_power.begin(), w.deformation_valence_power.end());
            dvo.push_back((int)dvp.size());
            an.push_back(w.anomalous_scattering.real());
            an.push_back(w.anomalous_scattering.imag());
        }
        std::vector<int> plm_nl;
        std::vector<double> plm, pval, kap, kapp, mult;
        for (const HC_AtomTypeParameters& t : parameters.type_parameters) {
            plm_nl.push_back((int)t.p_lm.size());
            double p[25] = {0};
            for (size_t l = 0; l < t.p_lm.size() && l < 5; l++)
                for (size_t i = 0; i < t.p_lm[l].size() && i < 2 * l + 1; i++) p[l * l + i] = t.p_lm[l][i];
            plm.insert(plm.end(), p, p + 25);
            pval.push_back(t.p_val);
            kap.push_back(t.kappa_spherical_valence);
            kapp.push_back(t.kappa_deformation_valence);
        }
        for (const AtomInCrystal& a : crystal.atoms) mult.push_back(a.multiplicity);
        // the C ABI copies everything it needs during dh_hc_create
        auto pad_i = [](std::vector<int>& v) { v.push_back(0); };
        auto pad_d = [](std::vector<double>& v) { v.push_back(0.0); };
        pad_i(nl); pad_i(tp); pad_i(occ); pad_i(kind); pad_i(dvp); pad_i(plm_nl);
        pad_d(cc); pad_d(ce); pad_d(dve); pad_d(an); pad_d(an); pad_d(pval); pad_d(kap); pad_d(kapp); pad_d(mult);
        for (int i = 0; i < 25; i++) plm.push_back(0.0);
        std::vector<int> a2w(parameters.atom_to_wfn_map), a2t(parameters.atom_to_type_map);
        a2w.push_back(0);
        a2t.push_back(0);
        rot.push_back(0);
        tr.push_back(0.0);
        dh_model_desc d{};
        d.cell[0] = uc.a; d.cell[1] = uc.b; d.cell[2] = uc.c;
        d.cell[3] = uc.alpha; d.cell[4] = uc.beta; d.cell[5] = uc.gamma;
        d.n_symops = (int)crystal.spaceGroupOperations.size();
        d.symop_rot = rot.data();
        d.symop_trans = tr.data();
        const bool cart = crystal.xyzCoordinateSystem == spc::XyzCoordinateSystem::cartesian;
        const bool ucart = crystal.adpConvention == spc::AdpConvention::U_cart;
        // derivative conventions follow the crystal's (reference constructor, .cpp:57-73)
        d.xyz_convention = d.d_xyz_convention = cart ? DH_XYZ_CARTESIAN : DH_XYZ_FRACTIONAL;
        d.adp_convention = d.d_adp_convention = ucart ? DH_ADP_U_CART : DH_ADP_U_CIF;
        d.n_wfn = (int)parameters.wfn_parameters.size();
        d.wfn_orbital_offset = wo.data();
        d.orbital_term_offset = to.data();
        d.orbital_nl = nl.data();
        d.term_coeff = cc.data();
        d.term_exp = ce.data();
        d.term_pow = tp.data();
        d.orbital_occupancy = occ.data();
        d.orbital_kind = kind.data();
        d.def_val_exponent = dve.data();
        d.def_val_pow_offset = dvo.data();
        d.def_val_pow = dvp.data();
        d.anomalous = an.data();
        d.n_types = (int)parameters.type_parameters.size();
        d.plm_n_l = plm_nl.data();
        d.plm = plm.data();
        d.p_val = pval.data();
        d.kappa = kap.data();
        d.kappa_prime = kapp.data();
        d.n_atoms = (int)parameters.atom_to_type_map.size();
        d.atom_to_wfn = a2w.data();
        d.atom_to_type = a2t.data();
        d.atom_multiplicity = mult.data();
        std::vector<const char*> labels;
        std::vector<int> z;
        for (const HC_WfnType& w : parameters.wfn_parameters) {
            labels.push_back(w.wfnAtomTypeName.c_str());
            z.push_back(w.atomic_number);
        }
        labels.push_back(nullptr);
        z.push_back(0);
        d.wfn_label = labels.data();
        d.wfn_atomic_number = z.data();
        detail::check(dh_hc_create(&d, device, &mCalc));
    }

    ~HansenCoppensStructureFactorCalculator() { dh_hc_destroy(mCalc); }
    HansenCoppensStructureFactorCalculator(const HansenCoppensStructureFactorCalculator&) = delete;
    HansenCoppensStructureFactorCalculator& operator=(const HansenCoppensStructureFactorCalculator&) = delete;

    // "HIP" (FP64 drop-in; "GPU" / "CPU" / "CPU_v4" accepted as aliases), "HIP_MIXED", "CPU_IAM"
    // (independent-atom model; "HIP_IAM_MIXED" in mixed precision)
    void setCalculationMode(const std::string& mode) { detail::check(dh_hc_set_calculation_mode(mCalc, mode.c_str())); }
    void setElectronScattering(bool electronScattering)
    {
        detail::check(dh_hc_set_electron_scattering(mCalc, electronScattering ? 1 : 0));
    }
    // N-Gaussian table of the CPU_IAM mode: "Waasmeier-Kirfel" (default), "IT92", "electron-IT",
    // "electron-cctbx"
    void setIamTable(const std::string& table) { detail::check(dh_hc_set_iam_table(mCalc, table.c_str())); }

    // One definition per atom (empty: detach).  Every later call recomputes the frames from the
    // positions it is given and ignores its localCoordinateSystems argument (AnyHcCalculator
    // without "frozen lcs"); chainRule adds dT/dL : dL/dxyz to the position derivatives.
    void setLocalCoordinateSystems(const std::vector<LocalCoordinateSystem>& lcs, bool chainRule = false)
    {
        std::vector<std::vector<dh_atom_ref>> refs;
        refs.reserve(3 * lcs.size() + 1);
        auto conv = [](const AtomInCrystalID& a) {
            dh_atom_ref r{};
            r.atom = a.index;
            detail::parse_operation(a.operation, r.rot, r.trans);
            return r;
        };
        auto list = [&](const std::vector<AtomInCrystalID>& v) {
            refs.emplace_back();
            for (const AtomInCrystalID& a : v) refs.back().push_back(conv(a));
            refs.back().push_back(dh_atom_ref{});
            return refs.back().data();
        };
        std::vector<dh_lcs_def> defs(lcs.size() + 1);
        for (size_t i = 0; i < lcs.size(); i++) {
            const LocalCoordinateSystem& l = lcs[i];
            dh_lcs_def& d = defs[i];
            d.central = conv(l.centralAtom);
            d.direction1_type = (int)l.direction1_type;
            d.direction2_type = (int)l.direction2_type;
            d.n_ref1 = (int)l.refPoint_1.size();
            d.n_ref2 = (int)l.refPoint_2.size();
            d.n_chirality = (int)l.chirality.size();
            d.ref1 = list(l.refPoint_1);
            d.ref2 = list(l.refPoint_2);
            d.chirality = list(l.chirality);
            d.coordinate_1 = l.coordinate_1;
            d.coordinate_2 = l.coordinate_2;
            d.is_r = l.isR ? 1 : 0;
        }
        detail::check(dh_hc_set_lcs_definitions(mCalc, (int)lcs.size(), defs.data(), chainRule ? 1 : 0));
    }
    // the frames the last call used
    std::vector<Matrix3d> lastLocalCoordinateSystems() const
    {
        std::vector<double> m(9 * mAtomCount + 9);
        detail::check(dh_hc_last_lcs(mCalc, m.data()));
        std::vector<Matrix3d> out(mAtomCount);
        for (size_t i = 0; i < mAtomCount; i++)
            for (int j = 0; j < 9; j++) out[i].m[j] = m[9 * i + j];
        return out;
    }

    void calculateStructureFactorsAndDerivatives(const std::vector<AtomInCrystal>& atoms,
                                                 const std::vector<Matrix3d>& localCoordinateSystems,
                                                 const std::vector<Vector3i>& hkl,
                                                 std::vector<std::complex<double>>& f,
                                                 std::vector<TargetFunctionAtomicParamDerivatives>& dTarget_dparam,
                                                 const std::vector<std::complex<double>>& dTarget_df,
                                                 const std::vector<bool>& countAtomContribution)
    {
        calculateStructureFactorsAndDerivatives(atoms, localCoordinateSystems, hkl, f, dTarget_dparam, dTarget_df,
                                                countAtomContribution, DerivativesSelector());
    }

    void calculateStructureFactorsAndDerivatives(const std::vector<AtomInCrystal>& atoms,
                                                 const std::vector<Matrix3d>& localCoordinateSystems,
                                                 const std::vector<Vector3i>& hkl,
                                                 std::vector<std::complex<double>>& f,
                                                 std::vector<TargetFunctionAtomicParamDerivatives>& dTarget_dparam,
                                                 const std::vector<std::complex<double>>& dTarget_df,
                                                 const std::vector<bool>& countAtomContribution,
                                                 const DerivativesSelector& selector)
    {
        const size_t n = atoms.size(), m = hkl.size();
        if (dTarget_df.size() != m)
            throw Exception(DH_E_ARGUMENT, "Inconsistent size of hkl vectors set and (d target/ d F) table when "
                                           "calculating multipolar structure factors.");
        detail::PackedAtoms pa(atoms, localCoordinateSystems, countAtomContribution);
        std::vector<int> h = detail::flat_hkl(hkl);
        std::vector<double> dd(2 * m + 1), fo(2 * m + 2), dx(3 * n + 3), du(6 * n + 6), dov(n + 1);
        for (size_t k = 0; k < m; k++) {
            dd[2 * k] = dTarget_df[k].real();
            dd[2 * k + 1] = dTarget_df[k].imag();
        }
        const std::vector<int>& adp_n = pa.adp_n;
        unsigned sel = (selector.d_xyz ? DH_D_XYZ : 0u) | (selector.d_adp ? DH_D_ADP : 0u) |
                       (selector.d_occ ? DH_D_OCC : 0u);
        detail::check(dh_hc_sf_and_derivs(mCalc, &pa.da, (int)m, h.data(), dd.data(), sel, fo.data(), dx.data(),
                                          du.data(), dov.data()));
        f.resize(m);
        for (size_t k = 0; k < m; k++) f[k] = {fo[2 * k], fo[2 * k + 1]};
        dTarget_dparam.resize(n);
        for (size_t i = 0; i < n; i++) {
            auto& d = dTarget_dparam[i];
            d.atomic_position_derivatives = Vector3d(dx[3 * i], dx[3 * i + 1], dx[3 * i + 2]);
            d.adp_derivatives.assign(du.begin() + 6 * i, du.begin() + 6 * i + adp_n[i]);
            d.occupancy_derivatives = dov[i];
        }
    }

    // Per-hkl overloads (HansenCoppensStructureFactorCalculator.h:206-242, 384-545): for every
    // reflection, in order, singleHklResultsProcessor.onPerHklCalculation(hkl_idx, F, derivatives)
    // with dF/dp of every atom in the calculator's derivative conventions.  An exception thrown by
    // the processor stops the calculation and is rethrown here.
    template <typename T>
    void calculateStructureFactorsAndDerivatives(const std::vector<AtomInCrystal>& atoms,
                                                 const std::vector<Matrix3d>& localCoordinateSystems,
                                                 const std::vector<Vector3i>& hkl, T& singleHklResultsProcessor) const
    {
        calculateStructureFactorsAndDerivatives(atoms, localCoordinateSystems, hkl, singleHklResultsProcessor,
                                                std::vector<bool>(atoms.size(), true));
    }

    template <typename T>
    void calculateStructureFactorsAndDerivatives(const std::vector<AtomInCrystal>& atoms,
                                                 const std::vector<Matrix3d>& localCoordinateSystems,
                                                 const std::vector<Vector3i>& hkl, T& singleHklResultsProcessor,
                                                 const std::vector<bool>& countAtomContribution) const
    {
        detail::PackedAtoms pa(atoms, localCoordinateSystems, countAtomContribution);
        std::vector<int> h = detail::flat_hkl(hkl);
        const size_t n = atoms.size();
        detail::PerHklContext<T> ctx{&singleHklResultsProcessor, SfDerivativesAtHkl(), n, nullptr};
        ctx.d.atomicPostionDerivatives.resize(n);
        ctx.d.adpDerivatives.resize(n);
        ctx.d.occupancyDerivatives.resize(n);
        for (size_t i = 0; i < n; i++) ctx.d.adpDerivatives[i].resize(pa.adp_n[i]);
        const int rc = dh_hc_per_hkl(mCalc, &pa.da, (int)hkl.size(), h.data(), &detail::PerHklContext<T>::call, &ctx);
        if (ctx.error) std::rethrow_exception(ctx.error);
        detail::check(rc);
    }

    // Form-factor route (HansenCoppensStructureFactorCalculator.h:171-181): atomic form factors of
    // every atom for one reflection (Miller indices or Cartesian 1/Angstrom), as the reference, and
    // for many reflections in one call (f[k * n_atoms + atom]).
    void calculateFormFactors(const std::vector<Matrix3d>& localCoordinateSystems, const Vector3i& hkl,
                              std::vector<std::complex<double>>& f, const std::vector<bool>& countAtom) const
    {
        calculateFormFactors(localCoordinateSystems, std::vector<Vector3i>(1, hkl), f, countAtom);
    }

    void calculateFormFactorsCart(const std::vector<Matrix3d>& localCoordinateSystems, const Vector3d& h_cart,
                                  std::vector<std::complex<double>>& f, const std::vector<bool>& countAtom) const
    {
        calculateFormFactorsCart(localCoordinateSystems, std::vector<Vector3d>(1, h_cart), f, countAtom);
    }

    void calculateFormFactors(const std::vector<Matrix3d>& localCoordinateSystems, const std::vector<Vector3i>& hkl,
                              std::vector<std::complex<double>>& f, const std::vector<bool>& countAtom) const
    {
        std::vector<int> h = detail::flat_hkl(hkl);
        form_factors(localCoordinateSystems, countAtom, hkl.size(), f,
                     [&](const double* l, const unsigned char* inc, double* out) {
                         return dh_hc_form_factors(mCalc, l, inc, (int)hkl.size(), h.data(), out);
                     });
    }

    void calculateFormFactorsCart(const std::vector<Matrix3d>& localCoordinateSystems,
                                  const std::vector<Vector3d>& h_cart, std::vector<std::complex<double>>& f,
                                  const std::vector<bool>& countAtom) const
    {
        std::vector<double> h(3 * h_cart.size() + 3);
        for (size_t k = 0; k < h_cart.size(); k++)
            for (int j = 0; j < 3; j++) h[3 * k + j] = h_cart[k][j];
        form_factors(localCoordinateSystems, countAtom, h_cart.size(), f,
                     [&](const double* l, const unsigned char* inc, double* out) {
                         return dh_hc_form_factors_cart(mCalc, l, inc, (int)h_cart.size(), h.data(), out);
                     });
    }

    void calculateStructureFactors(const std::vector<AtomInCrystal>& atoms,
                                   const std::vector<Matrix3d>& localCoordinateSystems,
                                   const std::vector<Vector3i>& hkl, std::vector<std::complex<double>>& f,
                                   const std::vector<bool>& countAtomContribution)
    {
        std::vector<TargetFunctionAtomicParamDerivatives> d;
        DerivativesSelector none;
        none.d_xyz = none.d_adp = none.d_occ = false;
        calculateStructureFactorsAndDerivatives(atoms, localCoordinateSystems, hkl, f, d,
                                                std::vector<std::complex<double>>(hkl.size(), 1.0),
                                                countAtomContribution, none);
    }

    void calculateStructureFactors(const std::vector<AtomInCrystal>& atoms,
                                   const std::vector<Matrix3d>& localCoordinateSystems,
                                   const std::vector<Vector3i>& hkl, std::vector<std::complex<double>>& f)
    {
        calculateStructureFactors(atoms, localCoordinateSystems, hkl, f, std::vector<bool>(atoms.size(), true));
    }

private:
    template <typename Call>
    void form_factors(const std::vector<Matrix3d>& lcs, const std::vector<bool>& countAtom, size_t n_h,
                      std::vector<std::complex<double>>& f, Call call) const
    {
        const size_t n = mAtomCount;
        if (lcs.size() != n || countAtom.size() != n) throw Exception(DH_E_ARGUMENT, "inconsistent number of atoms");
        std::vector<double> l(9 * n + 9), out(2 * n * n_h + 2);
        std::vector<unsigned char> inc(n + 1);
        for (size_t i = 0; i < n; i++) {
            for (int j = 0; j < 9; j++) l[9 * i + j] = lcs[i].m[j];
            inc[i] = countAtom[i] ? 1 : 0;
        }
        detail::check(call(l.data(), inc.data(), out.data()));
        f.resize(n * n_h);
        for (size_t i = 0; i < n * n_h; i++) f[i] = {out[2 * i], out[2 * i + 1]};
    }

    dh_calculator* mCalc = nullptr;
    size_t mAtomCount;
};

}  // namespace discamb_amd
