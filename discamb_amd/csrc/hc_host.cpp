// Host-side model preparation and crystallographic conventions for the Hansen-Coppens path.
#include "hc_internal.hpp"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstring>
#include <map>

#include "ngauss_tables.h"

namespace discamb_amd {

static const double kPi = 3.14159265358979323846;

static void mat_mul3(const double* a, const double* b, double* c)
{
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double s = 0;
            for (int k = 0; k < 3; k++) s += a[3 * i + k] * b[3 * k + j];
            c[3 * i + j] = s;
        }
}

// UnitCell.cpp:183-212 (ITC vol. B 3.3 orthogonalisation)
void Cell::set(const double* p6)
{
    std::memcpy(par, p6, sizeof(par));
    const double f = kPi / 180.0;
    double ca = std::cos(par[3] * f), cb = std::cos(par[4] * f), cg = std::cos(par[5] * f);
    double sg = std::sin(par[5] * f);
    double phi = std::sqrt(1 - ca * ca - cb * cb - cg * cg + 2 * ca * cb * cg);
    double a = par[0], b = par[1], c = par[2];
    std::memset(f2c, 0, sizeof(f2c));
    std::memset(c2f, 0, sizeof(c2f));
    f2c[0] = a;
    f2c[1] = b * cg;
    f2c[2] = c * cb;
    f2c[4] = b * sg;
    f2c[5] = (c / sg) * (ca - cb * cg);
    f2c[8] = c * phi / sg;
    c2f[0] = 1.0 / a;
    c2f[1] = -cg / (a * sg);
    c2f[2] = (ca * cg - cb) / (a * phi * sg);
    c2f[4] = 1.0 / (b * sg);
    c2f[5] = (cb * cg - ca) / (b * phi * sg);
    c2f[8] = sg / (c * phi);
    // reciprocal basis vectors in the Cartesian frame are the rows of c2f
    for (int i = 0; i < 3; i++)
        rec_len[i] = std::sqrt(c2f[3 * i] * c2f[3 * i] + c2f[3 * i + 1] * c2f[3 * i + 1] +
                               c2f[3 * i + 2] * c2f[3 * i + 2]);
}

static void sym6_to_mat(const double* u, double* m)
{
    m[0] = u[0]; m[4] = u[1]; m[8] = u[2];
    m[1] = m[3] = u[3];
    m[2] = m[6] = u[4];
    m[5] = m[7] = u[5];
}

static void mat_to_sym6(const double* m, double* u)
{
    u[0] = m[0]; u[1] = m[4]; u[2] = m[8];
    u[3] = m[1]; u[4] = m[2]; u[5] = m[5];
}

void Cell::u_cif_to_cart(const double* ucif6, double* ucart6) const
{
    double A[9], At[9], U[9], T[9], R[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) A[3 * i + j] = f2c[3 * i + j] * rec_len[j];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) At[3 * i + j] = A[3 * j + i];
    sym6_to_mat(ucif6, U);
    mat_mul3(A, U, T);
    mat_mul3(T, At, R);
    mat_to_sym6(R, ucart6);
}

void Cell::u_cart_to_cif(const double* ucart6, double* ucif6) const
{
    // U_cif = B U_cart B^T, B = diag(1/|a*|..) c2f
    double B[9], Bt[9], U[9], T[9], R[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) B[3 * i + j] = c2f[3 * i + j] / rec_len[i];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) Bt[3 * i + j] = B[3 * j + i];
    sym6_to_mat(ucart6, U);
    mat_mul3(B, U, T);
    mat_mul3(T, Bt, R);
    mat_to_sym6(R, ucif6);
}

void Cell::d_ucart_to_d_ucif(const double* d_cart6, double* d_cif6) const
{
    // G (symmetric, dT = sum_ij G_ij dU_ij) from parameter derivatives; G_cif = A^T G A.
    double A[9], At[9], G[9], T[9], R[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) A[3 * i + j] = f2c[3 * i + j] * rec_len[j];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) At[3 * i + j] = A[3 * j + i];
    G[0] = d_cart6[0]; G[4] = d_cart6[1]; G[8] = d_cart6[2];
    G[1] = G[3] = 0.5 * d_cart6[3];
    G[2] = G[6] = 0.5 * d_cart6[4];
    G[5] = G[7] = 0.5 * d_cart6[5];
    mat_mul3(At, G, T);
    mat_mul3(T, A, R);
    d_cif6[0] = R[0]; d_cif6[1] = R[4]; d_cif6[2] = R[8];
    d_cif6[3] = 2 * R[1]; d_cif6[4] = 2 * R[2]; d_cif6[5] = 2 * R[5];
}

// ---------------------------------------------------------------- space group analysis

static double wrap01(double x)
{
    double r = x - std::floor(x);
    if (std::fabs(r - 1.0) < 1e-9 || std::fabs(r) < 1e-9) r = 0.0;
    return r;
}

static bool same_mod1(const double* a, const double* b)
{
    for (int i = 0; i < 3; i++) {
        double d = a[i] - b[i];
        d -= std::round(d);
        if (std::fabs(d) > 1e-6) return false;
    }
    return true;
}

SpaceGroupInfo analyse_space_group(const std::vector<SymOp>& ops_in)
{
    SpaceGroupInfo g;
    if (ops_in.empty()) throw Error(2, "space group has no symmetry operations");
    std::vector<SymOp> ops = ops_in;
    for (auto& o : ops)
        for (int i = 0; i < 3; i++) o.t[i] = wrap01(o.t[i]);
    g.order = (int)ops.size();
    static const int I[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    // centering translations: operations with identity rotation
    for (auto& o : ops)
        if (std::equal(o.r, o.r + 9, I)) g.centering_vectors.push_back({o.t[0], o.t[1], o.t[2]});
    if (g.centering_vectors.empty()) throw Error(2, "space group lacks the identity operation");
    // inversion
    for (auto& o : ops) {
        bool inv = true;
        for (int i = 0; i < 9; i++)
            if (o.r[i] != -I[i]) inv = false;
        if (inv) {
            g.centro = true;
            for (int i = 0; i < 3; i++) g.inv_t[i] = o.t[i];
            break;
        }
    }
    // coset representatives modulo centering and inversion
    for (auto& o : ops) {
        bool dup = false;
        for (auto& s : g.subset) {
            bool same_r = std::equal(o.r, o.r + 9, s.r);
            bool neg_r = true;
            for (int i = 0; i < 9; i++)
                if (o.r[i] != -s.r[i]) neg_r = false;
            for (auto& c : g.centering_vectors) {
                if (same_r) {
                    double t[3] = {s.t[0] + c[0], s.t[1] + c[1], s.t[2] + c[2]};
                    if (same_mod1(o.t, t)) dup = true;
                }
                if (g.centro && neg_r) {
                    double t[3] = {-s.t[0] + g.inv_t[0] + c[0], -s.t[1] + g.inv_t[1] + c[1],
                                   -s.t[2] + g.inv_t[2] + c[2]};
                    if (same_mod1(o.t, t)) dup = true;
                }
            }
            if (dup) break;
        }
        if (!dup) g.subset.push_back(o);
    }
    // centering letter
    auto has = [&](double a, double b, double c) {
        double t[3] = {a, b, c};
        for (auto& v : g.centering_vectors)
            if (same_mod1(v.data(), t)) return true;
        return false;
    };
    const double h = 0.5, t1 = 1.0 / 3.0, t2 = 2.0 / 3.0;
    int nc = (int)g.centering_vectors.size();
    g.centering = 'P';
    if (nc == 4 && has(0, h, h) && has(h, 0, h) && has(h, h, 0)) g.centering = 'F';
    else if (nc == 2 && has(0, h, h)) g.centering = 'A';
    else if (nc == 2 && has(h, 0, h)) g.centering = 'B';
    else if (nc == 2 && has(h, h, 0)) g.centering = 'C';
    else if (nc == 2 && has(h, h, h)) g.centering = 'I';
    else if (nc == 3 && has(t2, t1, t1)) { g.centering = 'R'; g.obverse = true; }
    else if (nc == 3 && has(t1, t2, t1)) { g.centering = 'R'; g.obverse = false; }
    else if (nc != 1) throw Error(2, "unsupported lattice centering");
    if (g.order % (nc * (g.centro ? 2 : 1)) != 0 ||
        (int)g.subset.size() * nc * (g.centro ? 2 : 1) != g.order)
        throw Error(2, "symmetry operations do not form a complete space group");
    return g;
}

double centering_multiplier(char c, bool obverse, const int* hkl)
{
    int h = hkl[0], k = hkl[1], l = hkl[2];
    switch (c) {
    case 'P': return 1.0;
    case 'R':
        if (obverse) return std::abs(-h + k + l) % 3 == 0 ? 3.0 : 0.0;
        return std::abs(h - k + l) % 2 == 0 ? 3.0 : 0.0;   // as scattering_utilities.cpp:130
    case 'F': {
        int a = std::abs(h % 2), b = std::abs(k % 2), d = std::abs(l % 2);
        return (a == b && a == d) ? 4.0 : 0.0;
    }
    case 'I': return (h + k + l) % 2 == 0 ? 2.0 : 0.0;
    case 'A': return (k + l) % 2 == 0 ? 2.0 : 0.0;
    case 'B': return (h + l) % 2 == 0 ? 2.0 : 0.0;
    case 'C': return (h + k) % 2 == 0 ? 2.0 : 0.0;
    default: throw Error(2, "reflection conditions for this centering are not implemented");
    }
}

// ---------------------------------------------------------------- model preparation

// sto_atomic_wfn::atomicStoWfnToSphericalDensity (HC_WfnData.cpp:11-73): products of the
// orbital's Slater terms, weighted by occupancy / 4 pi.
void build_sphere_density(const WfnType& w, bool valence, SlaterDensity& out)
{
    out.c.clear(); out.e.clear(); out.p.clear();
    const std::vector<int>& idx = valence ? w.valence : w.core;
    for (int oi : idx) {
        if (oi < 0 || oi >= (int)w.orbitals.size()) throw Error(2, "orbital index out of range");
        double occ = (double)w.occupancy.at(oi);
        if (occ == 0.0) continue;
        const SlaterOrbital& o = w.orbitals[oi];
        for (size_t i = 0; i < o.coeff.size(); i++)
            for (size_t j = 0; j <= i; j++) {
                double c = o.coeff[i] * o.coeff[j] * occ / (4 * kPi);
                if (i != j) c *= 2;
                out.c.push_back(c);
                out.e.push_back(o.expo[i] + o.expo[j]);
                out.p.push_back(o.pow[i] + o.pow[j]);
            }
    }
}

EngineModel build_engine_model(const std::vector<WfnType>& wfn, const std::vector<AtomType>& types,
                               const std::vector<int>& atom_wfn, const std::vector<int>& atom_type,
                               std::vector<int>* engine_atom_type)
{
    EngineModel m;
    m.n_wfn = (int)wfn.size();
    m.n_types = (int)types.size();
    m.core.resize(m.n_wfn);
    m.valence.resize(m.n_wfn);
    for (int w = 0; w < m.n_wfn; w++) {
        build_sphere_density(wfn[w], false, m.core[w]);
        build_sphere_density(wfn[w], true, m.valence[w]);
        double vocc = 0;
        for (int oi : wfn[w].valence) vocc += wfn[w].occupancy.at(oi);
        // HansenCoppensStructureFactorCalculator.cpp:228-230 (division by the valence occupancy)
        for (auto& c : m.valence[w].c) c /= vocc;
        for (auto e : m.core[w].e)
            if (!(e > 0)) throw Error(2, "non-positive Slater exponent");
        for (size_t k = 0; k < m.core[w].p.size(); k++)
            if (m.core[w].p[k] + 2 > 12) throw Error(2, "Slater power too high (max r^10)");
        for (size_t k = 0; k < m.valence[w].p.size(); k++)
            if (m.valence[w].p[k] + 2 > 12) throw Error(2, "Slater power too high (max r^10)");
        m.dv_zeta.push_back(wfn[w].def_val_exponent);
        int nl = (int)wfn[w].def_val_power.size();
        if (nl > 5) nl = 5;
        m.dv_nl.push_back(nl);
        for (int l = 0; l < 5; l++) {
            int n = l < nl ? wfn[w].def_val_power[l] : 0;
            if (l < nl && (n + 2 < l + 1 || n + 2 > 12))
                throw Error(2, "deformation-valence power out of range for g-function");
            m.dv_n.push_back(n);
        }
        m.anom.push_back(wfn[w].anomalous.real());
        m.anom.push_back(wfn[w].anomalous.imag());
    }
    m.types = types;
    m.type_wfn.assign(m.n_types, 0);
    if (atom_wfn.size() != atom_type.size()) throw Error(1, "atom-to-wfn and atom-to-type maps differ in size");
    for (size_t a = 0; a < atom_wfn.size(); a++) {
        if (atom_type[a] < 0 || atom_type[a] >= m.n_types || atom_wfn[a] < 0 || atom_wfn[a] >= m.n_wfn)
            throw Error(1, "atom type or wavefunction index out of range");
        m.type_wfn[atom_type[a]] = atom_wfn[a];
    }
    m.n_model_types = m.n_types;
    for (int t = 0; t < m.n_types; t++) m.type_origin.push_back(t);
    for (int w = 0; w < m.n_wfn; w++) m.wfn_origin.push_back(w);
    // core of the atom's wavefunction, valence and deformation terms of the type's
    // (HansenCoppens_SF_Engine3.cpp:218-224, 1295-1300)
    std::map<std::pair<int, int>, int> pair_wfn, pair_type;
    std::vector<int> eat(atom_type);
    for (size_t a = 0; a < atom_wfn.size(); a++) {
        const int t = atom_type[a], cw = atom_wfn[a], vw = m.type_wfn[t];
        if (cw == vw) continue;
        auto pw = pair_wfn.find({cw, vw});
        int ew;
        if (pw == pair_wfn.end()) {
            ew = m.n_wfn++;
            pair_wfn[{cw, vw}] = ew;
            m.core.push_back(m.core[cw]);
            m.valence.push_back(m.valence[vw]);
            m.dv_zeta.push_back(m.dv_zeta[vw]);
            m.dv_nl.push_back(m.dv_nl[vw]);
            for (int l = 0; l < 5; l++) m.dv_n.push_back(m.dv_n[5 * vw + l]);
            m.anom.push_back(m.anom[2 * cw]);
            m.anom.push_back(m.anom[2 * cw + 1]);
            m.wfn_origin.push_back(cw);
        } else {
            ew = pw->second;
        }
        auto pt = pair_type.find({t, cw});
        if (pt == pair_type.end()) {
            const int et = m.n_types++;
            pair_type[{t, cw}] = et;
            m.types.push_back(m.types[t]);
            m.type_wfn.push_back(ew);
            m.type_origin.push_back(t);
            eat[a] = et;
        } else {
            eat[a] = pt->second;
        }
    }
    m.type_max_l.resize(m.n_types);
    for (int t = 0; t < m.n_types; t++) {
        const AtomType& ty = m.types[t];
        int lp = -1;
        for (int l = 0; l < ty.n_l && l < 5; l++)
            for (int i = l * l; i < (l + 1) * (l + 1); i++)
                if (ty.plm[i] != 0.0) lp = l;
        int lw = std::min(4, m.dv_nl[m.type_wfn[t]] - 1);
        m.type_max_l[t] = std::min(lw, lp);
    }
    if (engine_atom_type) *engine_atom_type = eat;
    return m;
}

// ---------------------------------------------------------------- N-Gaussian form factors

template <size_t N>
static const NGaussEntry* find_entry(const NGaussEntry (&t)[N], const std::string& label)
{
    for (size_t i = 0; i < N; i++)
        if (label == t[i].label) return &t[i];
    return nullptr;
}

// n_gaussian_form_factors_table::getFormFactor (NGaussianFormFactorsTable.cpp:3176-3217): exact
// label match in the named table, else the alternative spelling of an ion ("O-" -> "O1-").
const NGaussEntry* ngauss_lookup(const std::string& label, const std::string& table, int* n_gauss)
{
    const NGaussEntry* e = nullptr;
    int n = 5;
    if (table == "Waasmeier-Kirfel") e = find_entry(kNGauss_Waasmeier_Kirfel, label);
    else if (table == "IT92") {
        e = find_entry(kNGauss_X_Ray_It95, label);
        n = 4;
    } else if (table == "electron-IT") e = find_entry(kNGauss_Electron_IT, label);
    else if (table == "electron-cctbx") e = find_entry(kNGauss_Electron_cctbx, label);
    else throw Error(1, "unknown form-factor table " + table);
    if (!e && label.size() > 1 && (label.back() == '-' || label.back() == '+') &&
        !std::isdigit((unsigned char)label[label.size() - 2]))
        return ngauss_lookup(label.substr(0, label.size() - 1) + "1" + label.back(), table, n_gauss);
    if (n_gauss) *n_gauss = n;
    return e;
}

double ngauss_value(const NGaussEntry& e, int n, double h)
{
    double r = 0.0;
    const double x2 = h * h / 4;
    for (int i = 0; i < n; i++) r += e.a[i] * std::exp(-e.b[i] * x2);
    return r + e.c;
}

static const char* const kElements[] = {
    "H", "He", "Li", "Be", "B", "C", "N", "O", "F", "Ne", "Na", "Mg", "Al", "Si", "P", "S", "Cl", "Ar", "K", "Ca",
    "Sc", "Ti", "V", "Cr", "Mn", "Fe", "Co", "Ni", "Cu", "Zn", "Ga", "Ge", "As", "Se", "Br", "Kr", "Rb", "Sr", "Y",
    "Zr", "Nb", "Mo", "Tc", "Ru", "Rh", "Pd", "Ag", "Cd", "In", "Sn", "Sb", "Te", "I", "Xe", "Cs", "Ba", "La", "Ce",
    "Pr", "Nd", "Pm", "Sm", "Eu", "Gd", "Tb", "Dy", "Ho", "Er", "Tm", "Yb", "Lu", "Hf", "Ta", "W", "Re", "Os", "Ir",
    "Pt", "Au", "Hg", "Tl", "Pb", "Bi", "Po", "At", "Rn", "Fr", "Ra", "Ac", "Th", "Pa", "U", "Np", "Pu", "Am", "Cm",
    "Bk", "Cf"};

std::string element_symbol(int z)
{
    if (z < 1 || z > (int)(sizeof(kElements) / sizeof(kElements[0])))
        throw Error(2, "atomic number out of range: " + std::to_string(z));
    return kElements[z - 1];
}

// The independent-atom model of HansenCoppens_SF_Engine::calculateSF_IAM
// (HansenCoppens_SF_Engine.cpp:229-278): per wavefunction type the Waasmaier-Kirfel form factor
// of its label plus its anomalous term, no valence or deformation terms.  Each wavefunction gets
// its own atom type (so per-type radial records follow the element); the Gaussians are stored as
// core "terms" with power -1.
EngineModel build_iam_model(const std::vector<std::string>& labels, const std::vector<std::complex<double>>& anom,
                            const std::vector<int>& atom_wfn, std::vector<int>& atom_type, const std::string& table)
{
    EngineModel m;
    m.n_wfn = (int)labels.size();
    m.n_types = m.n_wfn;
    m.core.resize(m.n_wfn);
    m.valence.resize(m.n_wfn);
    for (int w = 0; w < m.n_wfn; w++) {
        int n = 0;
        const NGaussEntry* e = ngauss_lookup(labels[w], table, &n);
        if (!e)
            throw Error(2, "request for Gaussian type atomic form factor parameter for unknown atom type: " + labels[w]);
        for (int i = 0; i < n; i++) {
            m.core[w].c.push_back(e->a[i]);
            m.core[w].e.push_back(e->b[i]);
            m.core[w].p.push_back(-1);
        }
        m.core[w].c.push_back(e->c);
        m.core[w].e.push_back(0.0);
        m.core[w].p.push_back(-1);
        m.dv_zeta.push_back(1.0);
        m.dv_nl.push_back(0);
        for (int l = 0; l < 5; l++) m.dv_n.push_back(0);
        m.anom.push_back(anom[w].real());
        m.anom.push_back(anom[w].imag());
    }
    m.types.assign(m.n_types, AtomType());
    m.type_wfn.resize(m.n_types);
    for (int t = 0; t < m.n_types; t++) m.type_wfn[t] = t;
    m.type_max_l.assign(m.n_types, -1);
    m.n_model_types = m.n_types;
    for (int t = 0; t < m.n_types; t++) {
        m.type_origin.push_back(t);
        m.wfn_origin.push_back(t);
    }
    atom_type = atom_wfn;
    for (int w : atom_wfn)
        if (w < 0 || w >= m.n_wfn) throw Error(1, "atom wavefunction index out of range");
    return m;
}

}  // namespace discamb_amd
