// Test-only C shim over the parts of the DiSCaMB reference that compile from their own source
// files (math kernels, wavefunction bank, crystal conventions).  It is built by
// oracle/Makefile.ref into oracle/_ref/libdiscamb_ref.so and is loaded only by tests/ to pin the
// oracle restatement.  The reference's HansenCoppensStructureFactorCalculator / Engine3 include
// the CMake-generated header discamb/discamb_config.h (via BasicUtilities/Timer.h) and are
// therefore not built here.
#include "discamb/Scattering/SlaterTypeOrbitalScattering.h"
#include "discamb/MathUtilities/real_spherical_harmonics.h"
#include "discamb/MathUtilities/SphConverter.h"
#include "discamb/MathUtilities/NaturalCubicSpline.h"
#include "discamb/Scattering/NGaussianFormFactorsTable.h"
#include "discamb/AtomTyping/LocalCoordinateSystemCalculator.h"
#include "discamb/HC_Model/HC_WfnData.h"
#include "discamb/HC_Model/ClementiRoettiData.h"
#include "discamb/HC_Model/DeformationValenceParameters.h"
#include "discamb/CrystalStructure/UnitCell.h"
#include "discamb/CrystalStructure/StructuralParametersConverter.h"
#include "discamb/CrystalStructure/SpaceGroup.h"
#include "discamb/CrystalStructure/SpaceGroupOperation.h"

#include <complex>
#include <cstring>
#include <string>
#include <vector>

using namespace discamb;
namespace spc = structural_parameters_convention;

extern "C" {

// NaturalCubicSpline(values, step, start) evaluated at m points   (NaturalCubicSpline.cpp)
void ref_spline(const double* values, int n, double step, double start, const double* x, int m, double* out)
{
    NaturalCubicSpline s(std::vector<double>(values, values + n), step, start);
    for (int i = 0; i < m; i++) out[i] = s(x[i]);
}

// n_gaussian_form_factors_table::getFormFactor(label, table).calculate_h(h) at m values of |h|;
// returns 0 if the table has no entry for the label (hasFormFactor)
int ref_ngauss(const char* label, const char* table, const double* h, int m, double* out)
{
    if (!n_gaussian_form_factors_table::hasFormFactor(label, table)) return 0;
    NGaussianFormFactor f = n_gaussian_form_factors_table::getFormFactor(label, table);
    for (int i = 0; i < m; i++) out[i] = f.calculate_h(h[i]);
    return 1;
}

// LocalCoordinateSystemCalculator::calculate (LocalCoordinateSystemCalculator.cpp:194-238) for one
// definition: atoms at fractional xyz; reference atoms are (index, operation string) pairs; direction
// types 0 ANY_ORTHOGONAL, 1 AVERAGE_DIRECTION, 2 AVERAGE_POSITION; m9 = local axes as columns.
int ref_lcs(const double* cell6, int n_atoms, const double* xyz, int central, const char* central_op, int t1,
            int n1, const int* r1, const char* const* op1, int t2, int n2, const int* r2, const char* const* op2,
            int c1, int c2, int is_r, int nch, const int* ch, const char* const* opch, double* m9)
{
    Crystal c;
    c.unitCell.set(cell6[0], cell6[1], cell6[2], cell6[3], cell6[4], cell6[5]);
    c.atoms.resize(n_atoms);
    for (int i = 0; i < n_atoms; i++) c.atoms[i].coordinates = Vector3d(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
    const LcsDirectionType types[3] = {LcsDirectionType::ANY_ORTHOGONAL, LcsDirectionType::AVERAGE_DIRECTION,
                                       LcsDirectionType::AVERAGE_POSITION};
    LocalCoordinateSystem<AtomInCrystalID> lcs;
    lcs.centralAtom = AtomInCrystalID(central, SpaceGroupOperation(std::string(central_op)));
    lcs.direction1_type = types[t1];
    lcs.direction2_type = types[t2];
    for (int i = 0; i < n1; i++) lcs.refPoint_1.push_back(AtomInCrystalID(r1[i], SpaceGroupOperation(std::string(op1[i]))));
    for (int i = 0; i < n2; i++) lcs.refPoint_2.push_back(AtomInCrystalID(r2[i], SpaceGroupOperation(std::string(op2[i]))));
    for (int i = 0; i < nch; i++) lcs.chirality.push_back(AtomInCrystalID(ch[i], SpaceGroupOperation(std::string(opch[i]))));
    lcs.coordinate_1 = c1;
    lcs.coordinate_2 = c2;
    lcs.isR = is_r != 0;
    LocalCoordinateSystemCalculator calc(lcs, c);
    Matrix3d m;
    bool same = true;
    calc.calculate(m, c, same);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) m9[3 * i + j] = m(i, j);
    return same ? 1 : 0;
}

// sto_scattering::gFunction<l>(n, h, Z)   (SlaterTypeOrbitalScattering.h:51-183)
double ref_g_function(int l, int n, double h, double z) { return sto_scattering::gFunction(l, n, h, z); }

// real_spherical_harmonics::getDensityNormalized<4>  (real_spherical_harmonics.h:102-115)
void ref_density_normalized_ylm(const double* u, double* out25)
{
    std::vector<std::vector<double>> v(5);
    for (int l = 0; l < 5; l++) v[l].resize(2 * l + 1);
    real_spherical_harmonics::getDensityNormalized<4>(Vector3d(u[0], u[1], u[2]), v);
    int k = 0;
    for (int l = 0; l < 5; l++)
        for (int i = 0; i < 2 * l + 1; i++) out25[k++] = v[l][i];
}

// real_spherical_harmonics::getDensityToWfnMultipliers  (real_spherical_harmonics.cpp:95-103)
void ref_density_to_wfn(double* out25)
{
    std::vector<std::vector<double>> v;
    real_spherical_harmonics::getDensityToWfnMultipliers(4, v);
    for (int l = 0; l < 5; l++)
        for (int m = 0; m < 5; m++) out25[5 * l + m] = m <= l ? v[l][m] : 0.0;
}

// Global-frame P_lm exactly as HansenCoppens_SF_Engine3::calculateGlobalCoordinatesPlm
// (Engine3.cpp:594-654) composes it: density->wfn normalisation, SphConverter matrix, back.
// lcs is row-major 3x3 (lcs[3*i+j] = Matrix3d(i,j)); plm_local and plm_global are [l][l+m], 25.
void ref_global_plm(const double* lcs, const double* plm_local, double* plm_global)
{
    std::vector<std::vector<double>> den2wfn;
    real_spherical_harmonics::getDensityToWfnMultipliers(4, den2wfn);
    SphConverter conv;
    conv.setMaxL(4);
    std::vector<std::vector<double>> oldc(3, std::vector<double>(3)), newc{{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) oldc[i][j] = lcs[3 * i + j];
    std::vector<std::vector<std::vector<double>>> mats;
    conv.convert(oldc, newc, mats);
    int off = 0;
    for (int l = 0; l <= 4; l++) {
        for (int i = 0; i < 2 * l + 1; i++) {
            double s = 0;
            for (int j = 0; j < 2 * l + 1; j++)
                s += mats[l][i][j] * plm_local[off + j] * den2wfn[l][std::abs(l - j)];
            plm_global[off + i] = s / den2wfn[l][std::abs(l - i)];
        }
        off += 2 * l + 1;
    }
}

// Clementi-Roetti bank entry -> spherical core/valence density expansions
// (SlaterOrbitalWfnData + sto_atomic_wfn::atomicStoWfnToSphericalDensity, HC_WfnData.cpp:11-73),
// valence coefficients divided by valence occupancy as in
// HansenCoppensStructureFactorCalculator::setModel (.cpp:198-236).  Returns the number of terms
// written (core first, then valence) or -1.
int ref_wfn_density(const char* label, int kind, int max_terms, double* coeff, double* expo, int* powr)
{
    ClementiRoettiData cr;
    bool has = false;
    HC_WfnBankEntry e = cr.getEntry(label, has);
    if (!has) return -1;
    std::vector<double> c, x;
    std::vector<int> p;
    sto_atomic_wfn::atomicStoWfnToSphericalDensity(
        e, c, x, p, kind == 0 ? sto_atomic_wfn::CORE_DENSITY : sto_atomic_wfn::VALENCE_DENSITY);
    if (kind == 1) {
        double occ = 0;
        for (int i : e.valence_orbitals_indices) occ += e.orbital_occupancy[i];
        for (auto& v : c) v /= occ;
    }
    int n = (int)c.size();
    if (n > max_terms) return -1;
    for (int i = 0; i < n; i++) { coeff[i] = c[i]; expo[i] = x[i]; powr[i] = p[i]; }
    return n;
}

// DeformationValenceParameters (UBDB_CR when ubdb != 0, else STANDARD_CR)
int ref_def_val(const char* label, int ubdb, double* zeta, int* powers5)
{
    DeformationValenceParameters dv(ubdb ? DeformationValenceParameters::ParametersType::UBDB_CR
                                         : DeformationValenceParameters::ParametersType::STANDARD_CR);
    std::vector<int> p;
    if (!dv.getParameters(label, *zeta, p)) return -1;
    for (int i = 0; i < 5 && i < (int)p.size(); i++) powers5[i] = p[i];
    return (int)p.size();
}

// UnitCell fractional->Cartesian and Cartesian->fractional matrices, row-major.
void ref_unit_cell(const double* par6, double* f2c, double* c2f)
{
    UnitCell uc(par6[0], par6[1], par6[2], par6[3], par6[4], par6[5]);
    const Matrix3d& a = uc.getFractionalToCartesianMatrix();
    const Matrix3d& b = uc.getCartesianToFractionalMatrix();
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) { f2c[3 * i + j] = a(i, j); c2f[3 * i + j] = b(i, j); }
}

// StructuralParametersConverter: U_cif -> U_cart and the derivative conversions back
// (U_cart -> U_cif, Cartesian -> fractional xyz), StructuralParametersConverter.cpp:384-473.
void ref_convert(const double* par6, const double* u_cif6, double* u_cart6, const double* d_ucart6,
                 double* d_ucif6, const double* d_xyz_cart3, double* d_xyz_frac3)
{
    UnitCell uc(par6[0], par6[1], par6[2], par6[3], par6[4], par6[5]);
    StructuralParametersConverter c(uc);
    std::vector<double> in(u_cif6, u_cif6 + 6), out(6, 0.0);
    c.convertADP(in, out, spc::AdpConvention::U_cif, spc::AdpConvention::U_cart);
    for (int i = 0; i < 6; i++) u_cart6[i] = out[i];
    std::vector<std::complex<double>> din(6), dout(6);
    for (int i = 0; i < 6; i++) din[i] = d_ucart6[i];
    c.convertDerivativesADP(din, dout, spc::AdpConvention::U_cart, spc::AdpConvention::U_cif);
    for (int i = 0; i < 6; i++) d_ucif6[i] = dout[i].real();
    Vector3<std::complex<double>> xin(d_xyz_cart3[0], d_xyz_cart3[1], d_xyz_cart3[2]), xout;
    c.convertDerivativesXyz(xin, xout, spc::XyzCoordinateSystem::cartesian, spc::XyzCoordinateSystem::fractional);
    for (int i = 0; i < 3; i++) d_xyz_frac3[i] = xout[i].real();
}

// SpaceGroup built from operation strings separated by ';'.  Outputs the order |G|, the size of
// the explicit subset, centrosymmetry flag, inversion-centre translation (fractional), centering
// letter and the subset's rotations (int, row-major) and translations (fractional).
int ref_space_group(const char* ops, int* n_ops, int* n_subset, int* centro, double* inv_t3,
                    char* centering, int* obverse, int* rot9n, double* trans3n, int max_ops)
{
    std::vector<SpaceGroupOperation> v;
    std::string s(ops), tok;
    size_t pos = 0;
    while (pos <= s.size()) {
        size_t e = s.find(';', pos);
        if (e == std::string::npos) e = s.size();
        tok = s.substr(pos, e - pos);
        if (!tok.empty()) v.push_back(SpaceGroupOperation(tok));
        pos = e + 1;
    }
    SpaceGroup sg;
    sg.set(v);
    *n_ops = sg.nSymmetryOperations();
    *n_subset = sg.nSymmetryOperationsInSubset();
    *centro = sg.isCentrosymmetric() ? 1 : 0;
    Vector3<CrystallographicRational> t;
    if (*centro) {
        sg.inversionCenterTranslation(t);
        for (int i = 0; i < 3; i++) inv_t3[i] = t[i].get();
    } else
        inv_t3[0] = inv_t3[1] = inv_t3[2] = 0;
    bool obv = false;
    *centering = sg.centering(obv);
    *obverse = obv ? 1 : 0;
    if (*n_subset > max_ops) return -1;
    for (int k = 0; k < *n_subset; k++) {
        Matrix3d r;
        Vector3d tr;
        sg.getSpaceGroupOperation(0, 0, k).get(r, tr);
        for (int i = 0; i < 3; i++) {
            for (int j = 0; j < 3; j++) rot9n[9 * k + 3 * i + j] = (int)r(i, j);
            trans3n[3 * k + i] = tr[i];
        }
    }
    return 0;
}

}  // extern "C"
