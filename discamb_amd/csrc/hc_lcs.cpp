// Local coordinate systems from atomic positions (LocalCoordinateSystemCalculator,
// src/AtomTyping/LocalCoordinateSystemCalculator.cpp:54-238, with CrossProductLcs,
// src/MathUtilities/CrossProductLcs.cpp:10-62) and their derivatives with respect to the
// fractional coordinates of the atoms that define them (forward-mode dual numbers), for the
// per-step recomputation of the frames and the chain rule dT/dxyz += dT/dL : dL/dxyz.
#include "hc_lcs.hpp"

#include <cmath>

namespace discamb_amd {

namespace {

// value and derivatives along N tangent directions (3 per atom defining the frame)
template <int N>
struct Dual {
    double v = 0.0;
    double d[N] = {};
    Dual() = default;
    Dual(double x) : v(x) {}
};

template <int N>
inline Dual<N> operator+(const Dual<N>& a, const Dual<N>& b)
{
    Dual<N> r(a.v + b.v);
    for (int i = 0; i < N; i++) r.d[i] = a.d[i] + b.d[i];
    return r;
}
template <int N>
inline Dual<N> operator-(const Dual<N>& a, const Dual<N>& b)
{
    Dual<N> r(a.v - b.v);
    for (int i = 0; i < N; i++) r.d[i] = a.d[i] - b.d[i];
    return r;
}
template <int N>
inline Dual<N> operator*(const Dual<N>& a, const Dual<N>& b)
{
    Dual<N> r(a.v * b.v);
    for (int i = 0; i < N; i++) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
    return r;
}
template <int N>
inline Dual<N> operator*(double s, const Dual<N>& a)
{
    Dual<N> r(s * a.v);
    for (int i = 0; i < N; i++) r.d[i] = s * a.d[i];
    return r;
}
template <int N>
inline Dual<N> operator/(const Dual<N>& a, const Dual<N>& b)
{
    Dual<N> r(a.v / b.v);
    const double ib = 1.0 / b.v;
    for (int i = 0; i < N; i++) r.d[i] = (a.d[i] - r.v * b.d[i]) * ib;
    return r;
}
template <int N>
inline Dual<N> sqrt(const Dual<N>& a)
{
    Dual<N> r(std::sqrt(a.v));
    const double h = 0.5 / r.v;
    for (int i = 0; i < N; i++) r.d[i] = a.d[i] * h;
    return r;
}
template <int N>
inline double value(const Dual<N>& a) { return a.v; }
inline double value(double a) { return a; }

template <typename T>
struct V3 {
    T x, y, z;
};
template <typename T> V3<T> operator+(const V3<T>& a, const V3<T>& b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <typename T> V3<T> operator-(const V3<T>& a, const V3<T>& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <typename T> V3<T> scale(double s, const V3<T>& a) { return {s * a.x, s * a.y, s * a.z}; }
template <typename T> V3<T> divide(const V3<T>& a, const T& s) { return {a.x / s, a.y / s, a.z / s}; }
template <typename T> T dot(const V3<T>& a, const V3<T>& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <typename T> V3<T> cross(const V3<T>& a, const V3<T>& b)
{
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <typename T> V3<T> normalized(const V3<T>& a)
{
    using std::sqrt;
    return divide(a, sqrt(dot(a, a)));
}

V3<double> cart_position(const LcsRef& r, const double* xyz_frac, const double* f2c)
{
    const double* x = xyz_frac + 3 * (size_t)r.atom;
    double u[3];
    for (int i = 0; i < 3; i++) u[i] = r.rot[3 * i] * x[0] + r.rot[3 * i + 1] * x[1] + r.rot[3 * i + 2] * x[2] + r.trans[i];
    double c[3];
    for (int i = 0; i < 3; i++) c[i] = f2c[3 * i] * u[0] + f2c[3 * i + 1] * u[1] + f2c[3 * i + 2] * u[2];
    return {c[0], c[1], c[2]};
}

// Cartesian position of a reference atom: f2c (R x + t); with duals, the tangents of atom slot s
// are d/dx_frac of that atom (columns of f2c R).
template <typename T>
struct Position {
    static V3<double> at(const LcsRef& r, const double* xyz_frac, const double* f2c, const std::vector<int>&)
    {
        return cart_position(r, xyz_frac, f2c);
    }
};

template <int N>
struct Position<Dual<N>> {
    static V3<Dual<N>> at(const LcsRef& r, const double* xyz_frac, const double* f2c, const std::vector<int>& slots)
    {
        const V3<double> p = cart_position(r, xyz_frac, f2c);
        V3<Dual<N>> out{Dual<N>(p.x), Dual<N>(p.y), Dual<N>(p.z)};
        int s = -1;
        for (size_t k = 0; k < slots.size(); k++)
            if (slots[k] == r.atom) s = (int)k;
        if (s < 0) return out;
        // d(f2c R x)/dx_c = sum_k f2c[i][k] R[k][c]
        Dual<N>* comp[3] = {&out.x, &out.y, &out.z};
        for (int i = 0; i < 3; i++)
            for (int c = 0; c < 3; c++) {
                double v = 0.0;
                for (int k = 0; k < 3; k++) v += f2c[3 * i + k] * r.rot[3 * k + c];
                comp[i]->d[3 * s + c] = v;
            }
        return out;
    }
};

template <typename T>
V3<T> position(const LcsRef& r, const double* xyz_frac, const double* f2c, const std::vector<int>& slots)
{
    return Position<T>::at(r, xyz_frac, f2c, slots);
}

template <typename T>
V3<T> direction(const std::vector<LcsRef>& refs, int kind, const V3<T>& r0, const double* xyz, const double* f2c,
                const std::vector<int>& slots)
{
    V3<T> v{T(0.0), T(0.0), T(0.0)};
    if (refs.empty()) throw Error(1, "problem when calculating atomic local coordinate system");
    if (kind == DH_LCS_AVERAGE_DIRECTION_) {
        for (const LcsRef& r : refs) v = v + normalized(position<T>(r, xyz, f2c, slots) - r0);
    } else if (kind == DH_LCS_AVERAGE_POSITION_) {
        for (const LcsRef& r : refs) v = v + position<T>(r, xyz, f2c, slots);
        v = scale(1.0 / (double)refs.size(), v) - r0;
    } else {
        throw Error(1, "problem when calculating atomic local coordinate system");
    }
    return normalized(v);
}

// calculateAnyOrthogonal: the larger of r0 x X and r0 x Y, normalised
template <typename T>
V3<T> any_orthogonal(const V3<T>& r0)
{
    const V3<T> ex{T(1.0), T(0.0), T(0.0)}, ey{T(0.0), T(1.0), T(0.0)};
    const V3<T> rx = cross(r0, ex), ry = cross(r0, ey);
    return normalized(value(dot(rx, rx)) > value(dot(ry, ry)) ? rx : ry);
}

// axes[c] = local axis c (x, y, z); returns false if the frame is the identity by definition
template <typename T>
void frame(const LcsDef& d, const double* xyz, const double* f2c, const std::vector<int>& slots, V3<T> (&axes)[3])
{
    if (d.dir1 == DH_LCS_ANY_ORTHOGONAL_ && d.dir2 == DH_LCS_ANY_ORTHOGONAL_) {
        axes[0] = {T(1.0), T(0.0), T(0.0)};
        axes[1] = {T(0.0), T(1.0), T(0.0)};
        axes[2] = {T(0.0), T(0.0), T(1.0)};
        return;
    }
    const V3<T> r0 = position<T>(d.central, xyz, f2c, slots);
    V3<T> r1, r2;
    if (d.dir1 == DH_LCS_ANY_ORTHOGONAL_) {
        r2 = direction<T>(d.ref2, d.dir2, r0, xyz, f2c, slots);
        r1 = any_orthogonal(r2);
    } else {
        r1 = direction<T>(d.ref1, d.dir1, r0, xyz, f2c, slots);
        r2 = d.dir2 == DH_LCS_ANY_ORTHOGONAL_ ? any_orthogonal(r1) : direction<T>(d.ref2, d.dir2, r0, xyz, f2c, slots);
    }
    // CrossProductLcs with positive axis labels: c1 along r1, c2 in the (r1, r2) plane
    const int c1 = d.c1 + 1, c2 = d.c2 + 1, c3 = 6 - c1 - c2;
    const double s = (1 - 2 * ((2 + c2 - c1) % 3)) * (d.is_r ? 1.0 : -1.0);
    const V3<T> a0 = normalized(r1);
    const V3<T> a2 = normalized(scale(s, cross(a0, r2)));
    const V3<T> a1 = scale(s, cross(a2, a0));
    axes[c1 - 1] = a0;
    axes[c2 - 1] = a1;
    axes[c3 - 1] = a2;
    // takeChiralityIntoAccount: a negative triple product flips the third axis
    if (d.chirality.size() == 3) {
        const V3<double> p0 = position<double>(d.central, xyz, f2c, slots);
        V3<double> q[3];
        for (int k = 0; k < 3; k++) q[k] = position<double>(d.chirality[k], xyz, f2c, slots) - p0;
        if (dot(cross(q[0], q[1]), q[2]) < 0) {
            const int third = 3 - d.c1 - d.c2;
            axes[third] = scale(-1.0, axes[third]);
        }
    }
}

std::vector<int> involved_atoms(const LcsDef& d)
{
    std::vector<int> a;
    auto add = [&](const LcsRef& r) {
        for (int x : a)
            if (x == r.atom) return;
        a.push_back(r.atom);
    };
    add(d.central);
    for (const LcsRef& r : d.ref1) add(r);
    for (const LcsRef& r : d.ref2) add(r);
    return a;
}

}  // namespace

void lcs_validate(const LcsDef& d, int n_atoms)
{
    auto ok = [&](const LcsRef& r) { return r.atom >= 0 && r.atom < n_atoms; };
    if (!ok(d.central)) throw Error(1, "local coordinate system: atom index out of range");
    for (const auto* v : {&d.ref1, &d.ref2, &d.chirality})
        for (const LcsRef& r : *v)
            if (!ok(r)) throw Error(1, "local coordinate system: atom index out of range");
    if (d.c1 < 0 || d.c1 > 2 || d.c2 < 0 || d.c2 > 2 || d.c1 == d.c2)
        throw Error(1, "local coordinate system: invalid axis pair");
    if (!d.chirality.empty() && d.chirality.size() != 3)
        throw Error(1, "local coordinate system: chirality needs three atoms");
    for (int t : {d.dir1, d.dir2})
        if (t < 0 || t > 2) throw Error(1, "local coordinate system: invalid direction type");
    if ((d.dir1 != DH_LCS_ANY_ORTHOGONAL_ && d.ref1.empty()) || (d.dir2 != DH_LCS_ANY_ORTHOGONAL_ && d.ref2.empty()))
        throw Error(1, "local coordinate system: a direction needs reference atoms");
    if ((int)involved_atoms(d).size() > kLcsMaxAtoms)
        throw Error(1, "local coordinate system: too many defining atoms");
}

void lcs_matrix(const LcsDef& d, const double* xyz_frac, const double* f2c, double* m9)
{
    V3<double> ax[3];
    frame<double>(d, xyz_frac, f2c, {}, ax);
    for (int c = 0; c < 3; c++) {
        m9[c] = ax[c].x;
        m9[3 + c] = ax[c].y;
        m9[6 + c] = ax[c].z;
    }
}

namespace {

template <int N>
void chain_n(const LcsDef& d, const double* xyz_frac, const double* f2c, const std::vector<int>& slots,
             const double* dT_dL, double* grad)
{
    V3<Dual<N>> ax[3];
    frame<Dual<N>>(d, xyz_frac, f2c, slots, ax);
    // L[i][c] = component i of axis c
    const Dual<N>* L[9] = {&ax[0].x, &ax[1].x, &ax[2].x, &ax[0].y, &ax[1].y, &ax[2].y, &ax[0].z, &ax[1].z, &ax[2].z};
    for (size_t s = 0; s < slots.size(); s++)
        for (int c = 0; c < 3; c++) {
            double acc = 0.0;
            for (int k = 0; k < 9; k++) acc += dT_dL[k] * L[k]->d[3 * s + c];
            grad[3 * s + c] = acc;
        }
}

}  // namespace

int lcs_chain(const LcsDef& d, const double* xyz_frac, const double* f2c, const double* dT_dL, int* atoms,
              double* grad)
{
    if (d.dir1 == DH_LCS_ANY_ORTHOGONAL_ && d.dir2 == DH_LCS_ANY_ORTHOGONAL_) return 0;
    const std::vector<int> slots = involved_atoms(d);
    switch (slots.size()) {   // the dual width follows the number of defining atoms
    case 1: chain_n<3>(d, xyz_frac, f2c, slots, dT_dL, grad); break;
    case 2: chain_n<6>(d, xyz_frac, f2c, slots, dT_dL, grad); break;
    case 3: chain_n<9>(d, xyz_frac, f2c, slots, dT_dL, grad); break;
    case 4: chain_n<12>(d, xyz_frac, f2c, slots, dT_dL, grad); break;
    case 5: chain_n<15>(d, xyz_frac, f2c, slots, dT_dL, grad); break;
    case 6: chain_n<18>(d, xyz_frac, f2c, slots, dT_dL, grad); break;
    case 7: chain_n<21>(d, xyz_frac, f2c, slots, dT_dL, grad); break;
    case 8: chain_n<24>(d, xyz_frac, f2c, slots, dT_dL, grad); break;
    case 9: chain_n<27>(d, xyz_frac, f2c, slots, dT_dL, grad); break;
    case 10: chain_n<30>(d, xyz_frac, f2c, slots, dT_dL, grad); break;
    default: throw Error(1, "local coordinate system: too many defining atoms");
    }
    for (size_t s = 0; s < slots.size(); s++) atoms[s] = slots[s];
    return (int)slots.size();
}

}  // namespace discamb_amd
