// HIP engine for the Hansen-Coppens structure factor F(h) and the adjoint dT/dp (xyz, U, occ).
//
// Work decomposition (one GPU):
//   * prep kernels build, per call, (a) per (reflection, symmetry op) records - density-normalised
//     real harmonics of the rotated unit vector, Debye-Waller multipliers and the integer
//     fractional phase vector - (b) per (reflection, atom type) radial records - core +
//     Pval * valence Slater transforms and the five deformation-valence Fourier-Bessel terms - and
//     (c) per-atom records with P_lm rotated into the global frame;
//   * the main kernel is atom-stationary: each lane owns one atom (kept in registers), a wave
//     streams a chunk of reflections whose records are wave-uniform (scalar loads), evaluates
//     X = f T exp(i phi) per (atom, op, reflection), reduces w * sum_s X across the wave for
//     F(h) and accumulates the adjoint in registers - no atomics anywhere;
//   * two atomic-free stage-2 kernels sum the per-wave F partials and per-chunk derivative
//     partials in FP64 in a fixed order (bitwise reproducible).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <type_traits>

#include "../../include/discamb_amd.h"
#include "hc_device_math.hpp"
#include "hc_fp64_trig.hpp"
#include "hc_engine.hpp"

namespace discamb_amd {


static constexpr int kHklopRec = 36;    // doubles / floats per (reflection, op) record
static constexpr int kRadRec = 8;       // per (reflection, type) radial record
static constexpr int kWave = 64;
static constexpr int kWavesPerBlock = 4;
// Workgroups are dealt to the 8 XCDs of an MI355X round-robin by workgroup id; each XCD has its
// own L2.  The main kernel maps ids so that all atom blocks of one reflection chunk share an XCD.
static constexpr int kXcds = 8;


struct TypeRec {
    int wfn, nl, max_l, lp;     // lp = min(4, n_l - 1, size(p_lm) - 1): l range of the multipole adjoint
    int n[5], pad2[3];
    double pval, kappa, kappa_p, zeta;
    double inv_kappa, inv_kappa_p;
    double gscale[5];           // 4 pi (i^l sign) N_l: STO normalisation zeta^(n+3)/(n+2)!
    double plm[25];
};

struct WfnRec {
    int core_off, core_n, val_off, val_n;
    int n[5], pad;              // deformation-valence STO powers n_l
    double zeta;                // deformation-valence exponent
    double gscale[5];           // 4 pi (i^l sign) N_l: STO normalisation zeta^(n+3)/(n+2)!
};

struct TermRec {
    double c, e;
    int p, pad;
};

struct AtomIn {
    double xf[3];
    double u[6];
    double w;
    double anom[2];
    double lcs[9];
    int type;
    int orig;
};

struct AtomRec {
    double xf[3];
    double u[6];
    double w;
    double anom[2];
    double plm[25];
    int type;
    int orig;
    double pval, inv_kappa, inv_kappa_p;    // type parameters for on-the-fly radial functions
    int wfn, pad;
};

// Set while an allocation-free call runs on this thread: a buffer that would have to grow is an
// error instead of an allocation.
static thread_local bool g_no_alloc = false;

static void refuse_allocation(size_t n)
{
    throw Error(4, "allocation-free call: " + std::to_string(n) +
                       " bytes of device memory would have to be allocated (a buffer of the reflection records "
                       "for another calculation mode, or partial sums beyond the static scratch); prepare with "
                       "dh_hc_set_reflections in the current calculation mode");
}

void DeviceBuffer::ensure(size_t n)
{
    if (n <= bytes && p) return;
    if (g_no_alloc) refuse_allocation(n);
    release();
    size_t want = std::max<size_t>(n, 256);
    if (hipMalloc(&p, want) != hipSuccess) {
        p = nullptr;
        bytes = 0;
        throw Error(4, "device allocation failed");
    }
    bytes = want;
}

void DeviceBuffer::release()
{
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
}

void HostBuffer::ensure(size_t n)
{
    if (n <= bytes && p) return;
    if (g_no_alloc) refuse_allocation(n);
    release();
    size_t want = std::max<size_t>(n, 4096);
    if (hipHostMalloc(&p, want, hipHostMallocDefault) != hipSuccess) {
        p = nullptr;
        bytes = 0;
        throw Error(4, "pinned host allocation failed");
    }
    bytes = want;
}

void HostBuffer::release()
{
    if (p) (void)hipHostFree(p);
    p = nullptr;
    bytes = 0;
}

// ------------------------------------------------------------------ prep kernels

struct OpDev {
    int rf[9];
    double tf[3];
    double rc[9];
    double anom_sign;
};

__device__ __forceinline__ void hkl_prep_item(int k, const int* __restrict__ hkl, const OpDev* __restrict__ ops,
                                              int n_ops, const double* __restrict__ c2f, double* __restrict__ rec64,
                                              float* __restrict__ rec32, double* __restrict__ phase32,
                                              double* __restrict__ hlen)
{
    const int h0 = hkl[3 * k], h1 = hkl[3 * k + 1], h2 = hkl[3 * k + 2];
    double hc[3];
    for (int j = 0; j < 3; j++) hc[j] = h0 * c2f[j] + h1 * c2f[3 + j] + h2 * c2f[6 + j];
    double hl = sqrt(hc[0] * hc[0] + hc[1] * hc[1] + hc[2] * hc[2]);
    hlen[k] = hl;
    bool h000 = hl < 1e-10;
    for (int s = 0; s < n_ops; s++) {
        const OpDev& o = ops[s];
        double hs[3], hf[3];
        for (int j = 0; j < 3; j++) {
            hs[j] = hc[0] * o.rc[j] + hc[1] * o.rc[3 + j] + hc[2] * o.rc[6 + j];
            hf[j] = double(h0 * o.rf[j] + h1 * o.rf[3 + j] + h2 * o.rf[6 + j]);
        }
        double ht = h0 * o.tf[0] + h1 * o.tf[1] + h2 * o.tf[2];
        double y[25];
        if (h000) {
            for (int i = 0; i < 25; i++) y[i] = 0.0;
        } else {
            double inv = 1.0 / hl;
            ylm_density<double>(hs[0] * inv, hs[1] * inv, hs[2] * inv, y);
        }
        double m[6] = {hs[0] * hs[0], hs[1] * hs[1], hs[2] * hs[2], 2 * hs[0] * hs[1], 2 * hs[0] * hs[2],
                       2 * hs[1] * hs[2]};
        size_t base = ((size_t)k * n_ops + s) * kHklopRec;
        if (rec64) {
            double* r = rec64 + base;
            for (int i = 0; i < 25; i++) r[i] = y[i];
            for (int i = 0; i < 6; i++) r[25 + i] = m[i];
            for (int i = 0; i < 3; i++) r[31 + i] = hf[i];
            r[34] = ht;
            r[35] = o.anom_sign;
        }
        if (rec32) {
            float* r = rec32 + base;
            for (int i = 0; i < 25; i++) r[i] = (float)y[i];
            // Debye-Waller multipliers pre-scaled by log2(e) for exp2
            for (int i = 0; i < 6; i++) r[25 + i] = (float)(m[i] * 1.4426950408889634074);
            for (int i = 0; i < 3; i++) r[31 + i] = (float)hf[i];
            r[34] = 0.f;
            r[35] = (float)o.anom_sign;
            double* p = phase32 + ((size_t)k * n_ops + s) * 4;
            p[0] = hf[0]; p[1] = hf[1]; p[2] = hf[2]; p[3] = ht;
        }
    }
}

// Diagonal groups (mixed precision): one FP32 record per reflection [d_j(h) x 25, m(h) log2(e) x 6,
// h x 3, pad x 2], an FP64 [h, 0] phase record, h.t_s per operation, and |h|.
__device__ __forceinline__ void hkl_prep_diag_item(int k, const int* __restrict__ hkl,
                                                   const OpDev* __restrict__ ops, int n_ops,
                                                   const double* __restrict__ c2f, float* __restrict__ rec32,
                                                   double* __restrict__ hd, double* __restrict__ ht,
                                                   double* __restrict__ hlen)
{
    const int h0 = hkl[3 * k], h1 = hkl[3 * k + 1], h2 = hkl[3 * k + 2];
    double hc[3];
    for (int j = 0; j < 3; j++) hc[j] = h0 * c2f[j] + h1 * c2f[3 + j] + h2 * c2f[6 + j];
    const double hl = sqrt(hc[0] * hc[0] + hc[1] * hc[1] + hc[2] * hc[2]);
    hlen[k] = hl;
    double y[25];
    if (hl < 1e-10) {
        for (int i = 0; i < 25; i++) y[i] = 0.0;
    } else {
        const double inv = 1.0 / hl;
        ylm_density<double>(hc[0] * inv, hc[1] * inv, hc[2] * inv, y);
    }
    const double m[6] = {hc[0] * hc[0], hc[1] * hc[1], hc[2] * hc[2], 2 * hc[0] * hc[1], 2 * hc[0] * hc[2],
                         2 * hc[1] * hc[2]};
    float* r = rec32 + (size_t)k * kHklopRec;
    for (int i = 0; i < 25; i++) r[i] = (float)y[i];
    for (int i = 0; i < 6; i++) r[25 + i] = (float)(m[i] * 1.4426950408889634074);
    r[31] = (float)h0; r[32] = (float)h1; r[33] = (float)h2;
    r[34] = 0.f; r[35] = 0.f;
    double* p = hd + (size_t)k * 4;
    p[0] = h0; p[1] = h1; p[2] = h2; p[3] = 0.0;
    for (int s = 0; s < n_ops; s++)
        ht[(size_t)k * n_ops + s] = h0 * ops[s].tf[0] + h1 * ops[s].tf[1] + h2 * ops[s].tf[2];
}

// DPP lane sums (quad_perm xor 1, xor 2, row_half_mirror): total of eight adjacent lanes in each.
template <int CTRL>
__device__ __forceinline__ float dpp_add(float v)
{
    return v + __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xF, 0xF,
                                                                      false));
}

template <int CTRL>
__device__ __forceinline__ double dpp_add(double v)
{
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)b, CTRL, 0xF, 0xF, false);
    const unsigned hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), CTRL, 0xF, 0xF, false);
    return v + __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// sum over the eight lanes 8j .. 8j+7 (quad_perm [1,0,3,2], [2,3,0,1], row_half_mirror)
template <typename T>
__device__ __forceinline__ T octet_sum(T v)
{
    v = dpp_add<0xB1>(v);
    v = dpp_add<0x4E>(v);
    return dpp_add<0x141>(v);
}

// Radial functions of the model are smooth in s and shared by every atom type of a wavefunction:
// core and valence form factors 4 pi sum_k c_k g_0(p_k + 2, s, e_k) and the deformation-valence
// Fourier-Bessel transforms G_l(s) = 4 pi (i^l sign) N_l g_l(n_l + 2, s, zeta), l = 0..4.  They are
// tabulated per wavefunction on a uniform grid (value and d/ds at every node) and read back by
// cubic Hermite interpolation: core at s = |h|, valence at |h| / kappa, G_l at |h| / kappa'.  With
// a step of kTabStep = 5e-4 / A the interpolation error is below 1e-13 of f(0) for H..S (measured
// against the closed forms), far inside the FP64 parity bound.  Node i of wavefunction w sits at
// s = i * kTabStep and holds 16 values {core f, f', valence f, f', G_0 f, f', ..., G_4 f, f', 0, 0}
// (FP64 table).  The mixed-precision kernels read an FP32 table on a grid kLinRatio times finer
// by linear interpolation instead: node i holds {f(s_i), f(s_i+1) - f(s_i)} for the same seven
// functions, so a lookup is one FMA per function.  At 1.25e-4 / A the interpolation error for
// H..S was measured at <= 3e-8 of f(0) for the spherical terms and <= 1e-7 of max |G_l| for the
// deformation terms (against the closed forms for s <= 3.2 / A); linear interpolation error grows
// with the square of the step, so the 2.5e-4 / A grid used here stays within 1.2e-7 and 4e-7.  The
// finer grid's table did not fit the L1 and L2 caches of the on-the-fly kernels as well: on the
// 256-structure batch the step went from 9.33 ms to 8.76 ms while max |dF| / max |F| against FP64
// moved from 2.52e-7 to 2.60e-7.
static constexpr double kTabStep = 5e-4;
static constexpr int kTabNode = 16;
static constexpr int kLinRatio = 2;
static constexpr double kLinStep = kTabStep / kLinRatio;

struct TabRec {
    int off, n;        // first node, node count
};

// Node j of the tables, one 64-lane wave: lanes 8c .. 8c+7 build component c (0 core, 1 valence,
// 2..6 G_0..G_4, 7 padding), value and d/ds; the Slater terms of a component are spread over its
// eight lanes and summed with a fixed-order DPP reduction.
__device__ __forceinline__ void tab_build_node(int j, int lane, const TabRec* __restrict__ tabs, int n_wfn,
                                               const WfnRec* __restrict__ wfns, const TermRec* __restrict__ core,
                                               const TermRec* __restrict__ val, double* __restrict__ tab)
{
    const int c = lane >> 3, q = lane & 7;
    int w = 0;
    while (w + 1 < n_wfn && tabs[w + 1].off <= j) w++;
    const double sv = (j - tabs[w].off) * kTabStep;
    const WfnRec& wf = wfns[w];
    double f = 0.0, d = 0.0;
    if (c < 2) {
        const TermRec* terms = c == 0 ? core + wf.core_off : val + wf.val_off;
        const int nt = c == 0 ? wf.core_n : wf.val_n;
        double fg = 0.0, dgs = 0.0;
        for (int i = q; i < nt; i += 8) {
            const TermRec tr = terms[i];
            if (tr.p < 0) {
                // independent-atom model: Gaussian term a exp(-b s^2 / 4) of an N-Gaussian fit
                const double g = tr.c * exp(-0.25 * tr.e * sv * sv);
                fg += g;
                dgs += g * (-0.5 * tr.e * sv);
            } else {
                double dg;
                f += tr.c * g_function_dh(0, tr.p + 2, sv, tr.e, &dg);
                d += tr.c * dg;
            }
        }
        const double four_pi = 12.566370614359172953850;
        f = f * four_pi + fg;
        d = d * four_pi + dgs;
    } else if (c < 7 && q == 0) {
        const int l = c - 2;
        if (wf.n[l] >= 0) {
            double dg;
            f = wf.gscale[l] * g_function_dh(l, wf.n[l] + 2, sv, wf.zeta, &dg);
            d = wf.gscale[l] * dg;
        }
    }
    f = octet_sum(f);
    d = octet_sum(d);
    if (q == 0) {
        tab[(size_t)j * kTabNode + 2 * c] = f;
        tab[(size_t)j * kTabNode + 2 * c + 1] = d;
    }
}

// Hermite basis at x = s / kTabStep: node index i (clamped to the table) and weights for
// (f_i, f_{i+1}, f'_i, f'_{i+1}) with the step folded into the derivative weights.
template <typename R>
__device__ __forceinline__ int hermite(R x, int n, R (&wt)[4])
{
    const int i = min((int)x, n - 2);
    const R t = x - (R)i;
    const R u = R(1) - t, tt = t * t;
    wt[1] = tt * (R(3) - R(2) * t);
    wt[0] = R(1) - wt[1];
    wt[2] = R(kTabStep) * t * u * u;
    wt[3] = -R(kTabStep) * tt * u;
    return i;
}

// component pair (c, c + 1) of the table at x = s / kTabStep
template <typename R>
__device__ __forceinline__ R tab_at(const R* __restrict__ tab, TabRec tr, R x, int c)
{
    R wt[4];
    const int i = hermite<R>(x, tr.n, wt);
    const R* a = tab + (size_t)(tr.off + i) * kTabNode + c;
    return wt[0] * a[0] + wt[1] * a[kTabNode] + wt[2] * a[1] + wt[3] * a[kTabNode + 1];
}

// On-the-fly radial factors of one atom at |h| = h: S = f_core + Pval f_val + f' and G_0..G_L.
template <int L, typename R>
__device__ __forceinline__ void radial_otf(const R* __restrict__ tab, TabRec tr, R h, R pval, R inv_kappa,
                                           R inv_kappa_p, R anr, R& S, R (&G)[5])
{
    if constexpr (sizeof(R) == sizeof(float)) {
        // FP32 linear tables (k_lin_build): value + t * difference
        constexpr float inv_lstep = float(1.0 / kLinStep);
        const int nl = kLinRatio * tr.n;
        const R* base = tab + (size_t)kLinRatio * tr.off * kTabNode;
        auto at = [&](R x, int c) {
            const int i = min((int)x, nl - 2);
            const R t = x - (R)i;
            const R* a = base + (size_t)i * kTabNode + c;
            return a[0] + t * a[1];
        };
        S = at(h * inv_lstep, 0) + pval * at(h * (inv_kappa * inv_lstep), 2) + anr;
        const R x = h * (inv_kappa_p * inv_lstep);
        const int i = min((int)x, nl - 2);
        const R t = x - (R)i;
        const R* a = base + (size_t)i * kTabNode + 4;
#pragma unroll
        for (int l = 0; l < 5; l++) G[l] = l <= L ? a[2 * l] + t * a[2 * l + 1] : R(0);
        return;
    }
    constexpr R inv_step = R(1.0 / kTabStep);
    S = tab_at<R>(tab, tr, h * inv_step, 0) + pval * tab_at<R>(tab, tr, h * (inv_kappa * inv_step), 2) + anr;
    R wt[4];
    const int i = hermite<R>(h * (inv_kappa_p * inv_step), tr.n, wt);
    const R* a = tab + (size_t)(tr.off + i) * kTabNode + 4;
#pragma unroll
    for (int l = 0; l < 5; l++) {
        G[l] = R(0);
        if (l <= L)
            G[l] = wt[0] * a[2 * l] + wt[1] * a[kTabNode + 2 * l] + wt[2] * a[2 * l + 1] +
                   wt[3] * a[kTabNode + 2 * l + 1];
    }
}

// radial_otf on the FP32 linear tables with each lane's last core, valence and deformation
// intervals kept in registers.  Reflections in |h| order fall into one interval (kLinStep wide)
// for several consecutive reflections, so a lane reloads an interval only when its index
// changes: a divergent load, issued for the lanes whose interval moved.  Without it every lane
// gathers two or three table lines per reflection, a working set (64 lanes x 16 waves per CU)
// that overflows L1.  The values are radial_otf's, from the same table entries.
template <int L>
struct LinTabCache {
    int ic = -1, iv = -1, id = -1;
    float c0 = 0.f, c1 = 0.f, v0 = 0.f, v1 = 0.f;
    float d[2 * (L + 1)];

    __device__ __forceinline__ void eval(const float* __restrict__ tab, TabRec tr, float h, float pval,
                                         float inv_kappa, float inv_kappa_p, float anr, float& S, float (&G)[5])
    {
        constexpr float inv_lstep = float(1.0 / kLinStep);
        const int nl = kLinRatio * tr.n;
        const float* base = tab + (size_t)kLinRatio * tr.off * kTabNode;
        const float xc = h * inv_lstep;
        const int jc = min((int)xc, nl - 2);
        if (jc != ic) {
            ic = jc;
            c0 = base[(size_t)jc * kTabNode];
            c1 = base[(size_t)jc * kTabNode + 1];
        }
        const float xv = h * (inv_kappa * inv_lstep);
        const int jv = min((int)xv, nl - 2);
        if (jv != iv) {
            iv = jv;
            v0 = base[(size_t)jv * kTabNode + 2];
            v1 = base[(size_t)jv * kTabNode + 3];
        }
        S = (c0 + (xc - (float)jc) * c1) + pval * (v0 + (xv - (float)jv) * v1) + anr;
        const float xd = h * (inv_kappa_p * inv_lstep);
        const int jd = min((int)xd, nl - 2);
        if (jd != id) {
            id = jd;
            const float* a = base + (size_t)jd * kTabNode + 4;
#pragma unroll
            for (int j = 0; j < 2 * (L + 1); j++) d[j] = a[j];
        }
        const float t = xd - (float)jd;
#pragma unroll
        for (int l = 0; l < 5; l++) G[l] = l <= L ? d[2 * l] + t * d[2 * l + 1] : 0.f;
    }
};

// Per (reflection, type): S = f_core + Pval f_val(h/kappa) and
// G_l = 4 pi (i^l sign) N_l g_l(n_l + 2, h/kappa', zeta), l <= lp.  With radx set (multipole
// adjoint) f_val and its derivative are summed in closed form and the record
// [f_val, dS/dkappa, dG_0..4/dkappa', 0] is written as well.
// Electron scattering (Mott-Bethe, HansenCoppens_SF_Engine3.cpp:1387-1398): the X-ray form factor
// without its anomalous term becomes 0.023934 (Z - f_x) / (h^2 / 4) and, at h = 0, the type's
// electron form factor at zero angle; ez holds (Z, f_e(0)) per type.  The anomalous term, added in
// the main kernels, is unchanged.
__device__ __forceinline__ void radial_prep_item(int idx, const double* __restrict__ hlen,
                                                 const double* __restrict__ tab, const TabRec* __restrict__ tabs,
                                                 const TypeRec* __restrict__ types, const WfnRec* __restrict__ wfns,
                                                 const TermRec* __restrict__ val, int n_types, int n_hkl,
                                                 double* __restrict__ rad64, float* __restrict__ rad32,
                                                 double* __restrict__ radx64, float* __restrict__ radx32,
                                                 const double* __restrict__ ez)
{
    int t = idx / n_hkl, k = idx - t * n_hkl;       // type-major: a wave shares one type
    const double four_pi = 12.566370614359172953850;
    const double h = hlen[k];
    const TypeRec& ty = types[t];
    const TabRec tr = tabs[ty.wfn];
    const bool want_d = radx64 || radx32;
    const double hk = h * ty.inv_kappa;
    // S and G_l come from the tables with or without the derivative records, so the records of a
    // multipole-adjoint prep equal a plain one's bit for bit (and serve plain calls)
    const double fv = tab_at<double>(tab, tr, hk * (1.0 / kTabStep), 2);
    double r[kRadRec] = {0, 0, 0, 0, 0, 0, 0, 0};
    double x[kRadRec] = {0, 0, 0, 0, 0, 0, 0, 0};
    r[0] = tab_at<double>(tab, tr, h * (1.0 / kTabStep), 0) + ty.pval * fv;
    const double hp = h * ty.inv_kappa_p;
    {
        // G_l from the wavefunction's table (same Hermite basis for every l)
        double wt[4];
        const int i = hermite<double>(hp * (1.0 / kTabStep), tr.n, wt);
        const double* a = tab + (size_t)(tr.off + i) * kTabNode + 4;
        for (int l = 0; l <= ty.lp; l++)
            r[1 + l] = wt[0] * a[2 * l] + wt[1] * a[kTabNode + 2 * l] + wt[2] * a[2 * l + 1] +
                       wt[3] * a[kTabNode + 2 * l + 1];
    }
    if (want_d) {
        // the derivative records from the closed forms: f_val and its derivative summed over the
        // valence terms, dG_l / dkappa' from the g-function derivatives
        const WfnRec wf = wfns[ty.wfn];
        double fvc = 0.0, dfv = 0.0;
        for (int i = 0; i < wf.val_n; i++) {
            const TermRec tv = val[wf.val_off + i];
            double dg;
            fvc += tv.c * g_function_dh(0, tv.p + 2, hk, tv.e, &dg);
            dfv += tv.c * dg;
        }
        x[0] = fvc * four_pi;
        // d f_val(h/kappa) / d kappa = f_val'(h/kappa) * (-h / kappa^2)
        x[1] = ty.pval * (dfv * four_pi) * (-hk * ty.inv_kappa);
        for (int l = 0; l <= ty.lp; l++) {
            double dg;
            (void)g_function_dh(l, ty.n[l] + 2, hp, ty.zeta, &dg);
            x[2 + l] = ty.gscale[l] * dg * (-hp * ty.inv_kappa_p);
        }
    }
    if (ez) {
        if (h < 1e-10) {
            for (int i = 0; i < kRadRec; i++) r[i] = x[i] = 0.0;
            r[0] = ez[2 * t + 1];
        } else {
            const double b = -0.023934 / (0.25 * h * h);
            r[0] = -b * ez[2 * t] + b * r[0];
            for (int i = 1; i < kRadRec; i++) r[i] *= b;
            for (int i = 0; i < kRadRec; i++) x[i] *= b;
        }
    }
    size_t base = ((size_t)k * n_types + t) * kRadRec;
    if (rad64)
        for (int i = 0; i < kRadRec; i++) rad64[base + i] = r[i];
    if (rad32)
        for (int i = 0; i < kRadRec; i++) rad32[base + i] = (float)r[i];
    if (radx64)
        for (int i = 0; i < kRadRec; i++) radx64[base + i] = x[i];
    if (radx32)
        for (int i = 0; i < kRadRec; i++) radx32[base + i] = (float)x[i];
}

// One 64-lane workgroup per atom slot: lanes 0..49 evaluate the local-frame multipole expansion
// f(u_p) = sum_j P_j d_j(L^T u_p) at the projection nodes, lanes 0..24 project it back onto the
// global-frame harmonics.
// one 64-lane workgroup (lane p) per atom slot i; fsh: 64 doubles of LDS
__device__ __forceinline__ void atom_prep_item(int i, int p, double* fsh, const AtomIn* __restrict__ in,
                                               const TypeRec* __restrict__ types, AtomRec* __restrict__ out)
{
    const AtomIn& a = in[i];
    AtomRec& r = out[i];
    const bool real_atom = a.orig >= 0;
    int max_l = -1;
    if (real_atom) max_l = types[a.type].max_l;
    double f = 0.0;
    if (max_l >= 0 && p < 50) {
        const double* P = types[a.type].plm;
        const double u0 = c_proj_dir[p][0], u1 = c_proj_dir[p][1], u2 = c_proj_dir[p][2];
        // local coordinates: columns of the LCS are the local axes
        const double x = a.lcs[0] * u0 + a.lcs[3] * u1 + a.lcs[6] * u2;
        const double y = a.lcs[1] * u0 + a.lcs[4] * u1 + a.lcs[7] * u2;
        const double z = a.lcs[2] * u0 + a.lcs[5] * u1 + a.lcs[8] * u2;
        double d[25];
        ylm_density<double>(x, y, z, d);
#pragma unroll
        for (int j = 0; j < 25; j++) f += P[j] * d[j];
    }
    fsh[p] = f;
    __syncthreads();
    if (p < 25) {
        double s = 0.0;
        const int l = p < 1 ? 0 : p < 4 ? 1 : p < 9 ? 2 : p < 16 ? 3 : 4;
        if (l <= max_l)
            for (int q = 0; q < 50; q++) s += c_proj_w[q][p] * fsh[q];
        r.plm[p] = s;
    } else if (p < 28) {
        r.xf[p - 25] = a.xf[p - 25];
    } else if (p < 34) {
        r.u[p - 28] = a.u[p - 28];
    } else if (p == 34) {
        r.w = a.w;
    } else if (p < 37) {
        r.anom[p - 35] = a.anom[p - 35];
    } else if (p == 37) {
        r.type = a.type;
        r.orig = a.orig;
    } else if (p == 38) {
        const TypeRec& ty = types[a.type];
        r.pval = ty.pval;
        r.inv_kappa = ty.inv_kappa;
        r.inv_kappa_p = ty.inv_kappa_p;
        r.wfn = ty.wfn;
        r.pad = 0;
    }
}

__global__ void k_hkl_prep(const int* __restrict__ hkl, const OpDev* __restrict__ ops, int n_ops, int n_hkl,
                           const double* __restrict__ c2f, double* __restrict__ rec64, float* __restrict__ rec32,
                           double* __restrict__ phase32, double* __restrict__ hlen)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n_hkl) hkl_prep_item(k, hkl, ops, n_ops, c2f, rec64, rec32, phase32, hlen);
}

__global__ void k_hkl_prep_diag(const int* __restrict__ hkl, const OpDev* __restrict__ ops, int n_ops, int n_hkl,
                                const double* __restrict__ c2f, float* __restrict__ rec32, double* __restrict__ hd,
                                double* __restrict__ ht, double* __restrict__ hlen)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n_hkl) hkl_prep_diag_item(k, hkl, ops, n_ops, c2f, rec32, hd, ht, hlen);
}

// FP64 diagonal-group line kernels: one record per reflection, [d_j(h) x 25, m(h) x 6, h x 3,
// h.t_s for s < n_ops (<= 4), pad] (kDiag64Rec doubles), and |h|.
__device__ __forceinline__ void hkl_prep_diag64_item(int k, const int* __restrict__ hkl,
                                                     const OpDev* __restrict__ ops, int n_ops,
                                                     const double* __restrict__ c2f, double* __restrict__ rec,
                                                     double* __restrict__ hlen)
{
    const int h0 = hkl[3 * k], h1 = hkl[3 * k + 1], h2 = hkl[3 * k + 2];
    double hc[3];
    for (int j = 0; j < 3; j++) hc[j] = h0 * c2f[j] + h1 * c2f[3 + j] + h2 * c2f[6 + j];
    const double hl = sqrt(hc[0] * hc[0] + hc[1] * hc[1] + hc[2] * hc[2]);
    hlen[k] = hl;
    double y[25];
    if (hl < 1e-10) {
        for (int i = 0; i < 25; i++) y[i] = 0.0;
    } else {
        const double inv = 1.0 / hl;
        ylm_density<double>(hc[0] * inv, hc[1] * inv, hc[2] * inv, y);
    }
    double* r = rec + (size_t)k * 40;
    for (int i = 0; i < 25; i++) r[i] = y[i];
    r[25] = hc[0] * hc[0];
    r[26] = hc[1] * hc[1];
    r[27] = hc[2] * hc[2];
    r[28] = 2 * hc[0] * hc[1];
    r[29] = 2 * hc[0] * hc[2];
    r[30] = 2 * hc[1] * hc[2];
    r[31] = h0;
    r[32] = h1;
    r[33] = h2;
    for (int s = 0; s < 6; s++) {
        const OpDev* o = s < n_ops ? ops + s : nullptr;
        r[34 + s] = o ? h0 * o->tf[0] + h1 * o->tf[1] + h2 * o->tf[2] : 0.0;
    }
}

__global__ void k_hkl_prep_diag64(const int* __restrict__ hkl, const OpDev* __restrict__ ops, int n_ops, int n_hkl,
                                  const double* __restrict__ c2f, double* __restrict__ rec, double* __restrict__ hlen)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n_hkl) hkl_prep_diag64_item(k, hkl, ops, n_ops, c2f, rec, hlen);
}

__global__ __launch_bounds__(64) void k_tab_build(const TabRec* __restrict__ tabs, int n_wfn,
                                                  const WfnRec* __restrict__ wfns,
                            const TermRec* __restrict__ core, const TermRec* __restrict__ val, int n_nodes,
                            double* __restrict__ tab)
{
    tab_build_node(blockIdx.x, threadIdx.x, tabs, n_wfn, wfns, core, val, tab);
}

// The FP32 linear tables from the FP64 Hermite tables: one thread per (fine node, function);
// wavefunction w's fine nodes are [kLinRatio * off, kLinRatio * (off + n)).
__global__ void k_lin_build(const TabRec* __restrict__ tabs, int n_wfn, const double* __restrict__ tab, int n_nodes,
                            float* __restrict__ lin)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n_nodes * kLinRatio * 8) return;
    const int j = idx >> 3, c = idx & 7;
    int w = 0;
    while (w + 1 < n_wfn && kLinRatio * tabs[w + 1].off <= j) w++;
    const TabRec tr = tabs[w];
    float* out = lin + (size_t)j * kTabNode + 2 * c;
    if (c == 7) {
        out[0] = out[1] = 0.0f;
        return;
    }
    const double x = (double)(j - kLinRatio * tr.off) / kLinRatio;       // in coarse-node units
    const double f0 = tab_at<double>(tab, tr, x, 2 * c);
    const double f1 = tab_at<double>(tab, tr, x + 1.0 / kLinRatio, 2 * c);
    out[0] = (float)f0;
    out[1] = (float)(f1 - f0);
}

__global__ void k_radial_prep(const double* __restrict__ hlen, const double* __restrict__ tab,
                              const TabRec* __restrict__ tabs, const TypeRec* __restrict__ types,
                              const WfnRec* __restrict__ wfns, const TermRec* __restrict__ val, int n_types,
                              int n_hkl, double* __restrict__ rad64, float* __restrict__ rad32,
                              double* __restrict__ radx64, float* __restrict__ radx32, const double* __restrict__ ez)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < n_hkl * n_types)
        radial_prep_item(idx, hlen, tab, tabs, types, wfns, val, n_types, n_hkl, rad64, rad32, radx64, radx32, ez);
}

__global__ __launch_bounds__(64) void k_atom_prep(const AtomIn* __restrict__ in, const TypeRec* __restrict__ types,
                                                  AtomRec* __restrict__ out, int n_slots)
{
    __shared__ double fsh[64];
    atom_prep_item(blockIdx.x, threadIdx.x, fsh, in, types, out);
}

// ------------------------------------------------------------------ main kernels

struct MainArgs {
    const void* hklop;        // double[36] or float[36] per (k, s)
    const double* phase;      // mixed path: double[4] per (k, s)
    const void* radial;       // [k][t][8]
    const void* radx;         // multipole adjoint: [k][t][8] = f_val, dS/dkappa, dG_l/dkappa'
    const double* dtdf;       // [k][2]
    const AtomRec* atoms;     // sorted slots
    const int* wave_l;        // multipole class (2 or 4) per wave
    double* fpart;            // [wave][k][2]
    double* dpart;            // [chunk][slot][nd]
    int n_hkl, n_ops, n_types, n_slots, n_waves, chunk;
    int grid_x;               // atom blocks per reflection chunk
    int n_blocks;             // grid_x * number of chunks
    int per_xcd;              // ceil(n_blocks / kXcds)
    int diag_mask;            // diagonal groups: bit p set when an operation has sign pattern p
    int diag_op[8];           // diagonal groups: operation index of sign pattern p (bit i: axis i flips)
    float diag_anom[8];       // diagonal groups: anomalous sign of the operation with pattern p
    const double* ht;         // diagonal groups: h.t_s per (reflection, operation)
    const void* tab;          // on-the-fly radial functions: per-wavefunction table (FP32 mixed / FP64)
    const TabRec* tabs;       // node range of each wavefunction in tab
    const double* hlen;       // |h| per reflection
    const int* line_cont;     // FP64 line kernels: 1 where reflection k continues the line of k - 1
    const double* line_ops;   // FP64 line kernels: per operation [Delta R_f (3), Delta.t, m(Delta_s) (6)]
};

// Pointers read out of argument structs in memory (the batched kernels) are generic to the
// compiler, which then emits flat loads with full waits and no scalar loads.  Every buffer here is
// device global memory: gmem() re-types a pointer as global (address space 1) and cmem() a
// pointer to data no kernel of the step writes while the main kernel runs as constant (address
// space 4), so uniform reads become scalar loads again.  The integer round trip keeps the
// compiler from folding the casts away.
template <typename T>
__device__ __forceinline__ T* gmem(T* p)
{
    return (T*)((__attribute__((address_space(1))) T*)(unsigned long long)p);
}

template <typename T>
__device__ __forceinline__ const T* cmem(const T* p)
{
    return (const T*)((const __attribute__((address_space(4))) T*)(unsigned long long)p);
}

// per-atom partial layout: 0-2 d/dx_frac, 3-8 d/dU, 9 d/docc, then with the multipole adjoint
// 10-34 d/dP_lm (global frame), 35 d/dPval, 36 d/dkappa, 37 d/dkappa'
static constexpr int kNdBasic = 10;
static constexpr int kNdMultipole = 38;

// F staging.  Each lane deposits its weighted contribution to reflection kk of the current block
// of kStage; the flush sums the 64 atoms of every staged reflection in a fixed order
// (deterministic) and writes one FP64 partial per (wave, reflection).  In the flush lane
// kk*8 + part sums atoms part*8 .. part*8+7 of reflection kk from LDS, then the eight parts of a
// reflection (eight adjacent lanes) are combined with DPP.  Each group of eight atoms is padded so
// that the 128-bit column reads of the flush fall on distinct LDS banks.  LDS accesses of one wave
// complete in order; the fences keep the compiler from moving them across the staging boundary.
static constexpr int kStage = 8;                               // reflections per staging block
static constexpr int kGrpF32 = 10;                             // float2 slots per 8 atoms (2 pad)
static constexpr int kGrpF64 = 9;                              // double2 slots per 8 atoms (1 pad)
static constexpr int kStageBytesF32 = kStage * 8 * kGrpF32 * 8;    // per wave, mixed path
static constexpr int kStageBytesF64 = kStage * 8 * kGrpF64 * 16;   // per wave, FP64 path

__device__ __forceinline__ void lds_wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ int stage_slot(int kk, int lane, int grp)
{
    return kk * 8 * grp + lane + (lane >> 3) * (grp - 8);
}

__device__ __forceinline__ void stage_f32(float2* buf, int kk, int lane, float re, float im)
{
    buf[stage_slot(kk, lane, kGrpF32)] = make_float2(re, im);
}

// The mixed-precision partials are FP32 sums of FP32 terms: stored as FP32 (the stage-2 reduction
// converts them exactly), half the bytes of FP64 partials.
__device__ __forceinline__ void flush_f32(const float2* buf, int nk, int lane, float* __restrict__ fp)
{
    lds_wave_sync();
    const int kk = lane >> 3, part = lane & 7;
    const float4* seg = reinterpret_cast<const float4*>(buf + kk * 8 * kGrpF32 + part * kGrpF32);
    float re = 0.f, im = 0.f;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const float4 v = seg[i];
        re += v.x;
        im += v.y;
        re += v.z;
        im += v.w;
    }
    re = octet_sum(re);
    im = octet_sum(im);
    if (part == 0 && kk < nk) {
        fp[2 * kk] = re;
        fp[2 * kk + 1] = im;
    }
    lds_wave_sync();
}

__device__ __forceinline__ void stage_f64(double2* buf, int kk, int lane, double re, double im)
{
    buf[stage_slot(kk, lane, kGrpF64)] = make_double2(re, im);
}

__device__ __forceinline__ void flush_f64(const double2* buf, int nk, int lane, double* __restrict__ fp)
{
    lds_wave_sync();
    const int kk = lane >> 3, part = lane & 7;
    const double2* seg = buf + kk * 8 * kGrpF64 + part * kGrpF64;
    double re = 0.0, im = 0.0;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        const double2 v = seg[i];
        re += v.x;
        im += v.y;
    }
    re = octet_sum(re);
    im = octet_sum(im);
    if (part == 0 && kk < nk) {
        fp[2 * kk] = re;
        fp[2 * kk + 1] = im;
    }
    lds_wave_sync();
}

// Wave-uniform FP64 record entries are fetched as 32-bit words so that type-based alias analysis
// can separate them from the FP64 partial-sum stores and the loads stay on the scalar unit.
__device__ __forceinline__ double ld_f64(const unsigned* __restrict__ p, int i)
{
    unsigned lo = p[2 * i], hi = p[2 * i + 1];
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// One (atom, op, reflection) term in working precision R: the def-val contraction, phase,
// Debye-Waller factor and X = f T exp(i phi).  Y: the 25 harmonics, m: 6 DW multipliers.
template <int L, typename R>
struct Term {
    R A[5];
    R Er, Ei;      // T exp(i phi)
    R Xr, Xi;      // f T exp(i phi)
};

template <int L, typename R>
__device__ __forceinline__ void contract(const R* P, const R* Y, R* A)
{
    constexpr int NLM = (L + 1) * (L + 1);
    A[0] = P[0] * Y[0];
    A[1] = P[1] * Y[1] + P[2] * Y[2] + P[3] * Y[3];
    R a2 = 0;
#pragma unroll
    for (int j = 4; j < 9; j++) a2 += P[j] * Y[j];
    A[2] = a2;
    A[3] = A[4] = R(0);
    if (L == 4) {
        R a3 = 0, a4 = 0;
#pragma unroll
        for (int j = 9; j < 16; j++) a3 += P[j < NLM ? j : 0] * Y[j];
#pragma unroll
        for (int j = 16; j < 25; j++) a4 += P[j < NLM ? j : 0] * Y[j];
        A[3] = a3;
        A[4] = a4;
    }
}

template <int L, bool DERIV, int MP, bool OTF>
__device__ __forceinline__ void body_fp64(const MainArgs& a, int wave, int lane, int chunk_id, int k0, int k1,
                                          const unsigned* __restrict__ hop, const double* __restrict__ rad,
                                          const double* __restrict__ radx, const unsigned* __restrict__ dtdf32,
                                          double2* __restrict__ stage, double* __restrict__ fpart,
                                          double* __restrict__ dpart)
{
    constexpr int NLM = (L + 1) * (L + 1);
    constexpr bool SC = MP == 1;    // multipole pass 1: F, xyz/U/occ and the Pval, kappa, kappa' sums
    constexpr bool BO = MP == 2;    // multipole pass 2: the global-frame B_lm sums only
    const int slot = wave * kWave + lane;
    const AtomRec& at = cmem(a.atoms)[slot];
    const double x0 = at.xf[0], x1 = at.xf[1], x2 = at.xf[2];
    double u[6];
#pragma unroll
    for (int j = 0; j < 6; j++) u[j] = at.u[j];
    double P[NLM];
#pragma unroll
    for (int j = 0; j < NLM; j++) P[j] = at.plm[j];
    const double w = at.w, anr = at.anom[0], ani = at.anom[1];
    const int t = at.type;
    TabRec tr{0, 2};
    double pv = 0, ik = 0, ikp = 0;
    if (OTF) {
        tr = cmem(a.tabs)[at.wfn];
        pv = at.pval;
        ik = at.inv_kappa;
        ikp = at.inv_kappa_p;
    }
    const double* tab = cmem(static_cast<const double*>(a.tab));
    const unsigned* hl32 = cmem(reinterpret_cast<const unsigned*>(a.hlen));
    double dx[3] = {0, 0, 0}, du[6] = {0, 0, 0, 0, 0, 0}, docc = 0;
    double B[BO ? NLM : 1];
#pragma unroll
    for (int j = 0; j < (BO ? NLM : 1); j++) B[j] = 0;
    double dpv = 0, dka = 0, dkp = 0;
    for (int kb = k0; kb < k1; kb += kStage) {
    const int nk = min(kStage, k1 - kb);
    for (int kk = 0; kk < nk; ++kk) {
        const int k = kb + kk;
        double S, G0, G1, G2, G3, G4;
        if (OTF) {
            double G[5];
            radial_otf<L, double>(tab, tr, ld_f64(hl32, k), pv, ik, ikp, anr, S, G);
            G0 = G[0]; G1 = G[1]; G2 = G[2]; G3 = G[3]; G4 = G[4];
        } else {
            const double* rr = rad + ((size_t)k * a.n_types + t) * kRadRec;
            S = rr[0] + anr; G0 = rr[1]; G1 = rr[2]; G2 = rr[3]; G3 = rr[4]; G4 = rr[5];
        }
        double fval = 0, dSk = 0, dG[5] = {0, 0, 0, 0, 0};
        if (SC) {
            const double* rx = radx + ((size_t)k * a.n_types + t) * kRadRec;
            fval = rx[0];
            dSk = rx[1];
#pragma unroll
            for (int l = 0; l < 5; l++) dG[l] = rx[2 + l];
        }
        const double cr = ld_f64(dtdf32, 2 * k), ci = ld_f64(dtdf32, 2 * k + 1);
        double sr = 0, si = 0, Rh = 0;
        for (int s = 0; s < a.n_ops; ++s) {
            const unsigned* rw = hop + ((size_t)k * a.n_ops + s) * (2 * kHklopRec);
            double r[kHklopRec];
#pragma unroll
            for (int j = 0; j < kHklopRec; j++) r[j] = ld_f64(rw, j);
            double A[5] = {0, 0, 0, 0, 0};
            double fr = 0, fi = 0;
            if (!BO) {
                contract<L, double>(P, r, A);
                fr = S + G0 * A[0] + G2 * A[2];
                fi = ani * r[35] + G1 * A[1];
                if (L == 4) {
                    fr += G4 * A[4];
                    fi += G3 * A[3];
                }
            }
            // h.(R x + t) reduced to [0, 1) revolutions
            const double ph = __builtin_amdgcn_fract(fma(r[33], x2, fma(r[32], x1, fma(r[31], x0, r[34]))));
            double sn, cs;
            sincos_rev(ph, &sn, &cs);
            double e = r[25] * u[0] + r[26] * u[1] + r[27] * u[2] + r[28] * u[3] + r[29] * u[4] + r[30] * u[5];
            double T = exp_neg(e);
            double Er = T * cs, Ei = T * sn;
            if (!BO) {
                double Xr = fr * Er - fi * Ei, Xi = fr * Ei + fi * Er;
                sr += Xr;
                si += Xi;
                if (DERIV) {
                    double zr = cr * Xr - ci * Xi, zi = cr * Xi + ci * Xr;
#pragma unroll
                    for (int j = 0; j < 6; j++) du[j] -= r[25 + j] * zr;
#pragma unroll
                    for (int j = 0; j < 3; j++) dx[j] -= r[31 + j] * zi;
                }
            }
            if (SC || BO) {
                double cer = cr * Er - ci * Ei, cei = cr * Ei + ci * Er;
                if (SC) {
                    Rh += cer;
                    dkp += (dG[0] * A[0] + dG[2] * A[2] + dG[4] * A[4]) * cer - (dG[1] * A[1] + dG[3] * A[3]) * cei;
                } else {
                    double q[5] = {G0 * cer, -G1 * cei, G2 * cer, -G3 * cei, G4 * cer};
#pragma unroll
                    for (int j = 0; j < NLM; j++) {
                        const int l = j < 1 ? 0 : j < 4 ? 1 : j < 9 ? 2 : j < 16 ? 3 : 4;
                        B[j] += q[l] * r[j];
                    }
                }
            }
        }
        if (DERIV && !BO) docc += cr * sr - ci * si;
        if (SC) {
            dpv += fval * Rh;
            dka += dSk * Rh;
        }
        if (!BO) stage_f64(stage, kk, lane, w * sr, w * si);
    }
    if (!BO) flush_f64(stage, nk, lane, fpart + ((size_t)wave * a.n_hkl + kb) * 2);
    }
    if (DERIV) {
        constexpr int ND = MP ? kNdMultipole : kNdBasic;
        double* dp = dpart + ((size_t)chunk_id * a.n_slots + slot) * ND;
        if (!BO) {
            dp[0] = dx[0]; dp[1] = dx[1]; dp[2] = dx[2];
#pragma unroll
            for (int j = 0; j < 6; j++) dp[3 + j] = du[j];
            dp[9] = docc;
        }
        if (BO) {
#pragma unroll
            for (int j = 0; j < 25; j++) dp[10 + j] = j < NLM ? B[j < NLM ? j : 0] : 0.0;
        }
        if (SC) {
            dp[35] = dpv;
            dp[36] = dka;
            dp[37] = dkp;
        }
        if (MP == 3) dp[35] = dp[36] = dp[37] = 0.0;
    }
}

// FP64 with the phase and Debye-Waller factors carried along reflection lines (the reference's
// line recurrences, scattering_utilities.cpp:520-756, used at HansenCoppens_SF_Engine3.cpp:
// 1258-1288).  Reflections are in line order: consecutive reflections k-1, k of one line differ
// by Delta = one step along a fixed hkl axis (line_cont[k] = 1).  For operation s with h_s = h R_s,
// x the atom's fractional position and U its (2 pi^2) Cartesian ADP quadratic form:
//   phase:  exp(i phi(h + Delta)) = exp(i phi(h)) M_s,   M_s = exp(2 pi i (Delta R_f . x + Delta . t_s))
//   DW:     T(h + Delta) = T(h) q,   q <- q c_s,   c_s = exp(-2 m(Delta_s) . u)
// so a term continuing a line costs one complex and two real products for E = T exp(i phi) instead
// of sin, cos and exp.  Lines and work chunks start with direct evaluation (q from the next
// reflection's Debye-Waller multipliers); the recurrence error grows by a few ulp per step.  The
// lane keeps E, q, M, c of every operation in registers, hence the compile-time operation count.
template <int L, bool DERIV, bool OTF, int NOPS>
__device__ __forceinline__ void body_fp64_line(const MainArgs& a, int wave, int lane, int chunk_id, int k0, int k1,
                                               const unsigned* __restrict__ hop, const double* __restrict__ rad,
                                               const unsigned* __restrict__ dtdf32, double2* __restrict__ stage,
                                               double* __restrict__ fpart, double* __restrict__ dpart)
{
    constexpr int NLM = (L + 1) * (L + 1);
    const int slot = wave * kWave + lane;
    const AtomRec& at = cmem(a.atoms)[slot];
    const double x0 = at.xf[0], x1 = at.xf[1], x2 = at.xf[2];
    double u[6];
#pragma unroll
    for (int j = 0; j < 6; j++) u[j] = at.u[j];
    double P[NLM];
#pragma unroll
    for (int j = 0; j < NLM; j++) P[j] = at.plm[j];
    const double w = at.w, anr = at.anom[0], ani = at.anom[1];
    const int t = at.type;
    TabRec tr{0, 2};
    double pv = 0, ik = 0, ikp = 0;
    if (OTF) {
        tr = cmem(a.tabs)[at.wfn];
        pv = at.pval;
        ik = at.inv_kappa;
        ikp = at.inv_kappa_p;
    }
    const double* tab = cmem(static_cast<const double*>(a.tab));
    const unsigned* hl32 = cmem(reinterpret_cast<const unsigned*>(a.hlen));
    const int* cont = cmem(a.line_cont);
    const unsigned* lop = cmem(reinterpret_cast<const unsigned*>(a.line_ops));
    // per-(atom, operation) line multipliers
    double Mr[NOPS], Mi[NOPS], Cq[NOPS], Er[NOPS], Ei[NOPS], Q[NOPS];
#pragma unroll
    for (int s = 0; s < NOPS; s++) {
        const unsigned* o = lop + s * 20;
        const double ph = __builtin_amdgcn_fract(fma(ld_f64(o, 2), x2, fma(ld_f64(o, 1), x1, fma(ld_f64(o, 0), x0, ld_f64(o, 3)))));
        sincos_rev(ph, &Mi[s], &Mr[s]);
        double e2 = 0.0;
#pragma unroll
        for (int j = 0; j < 6; j++) e2 = fma(ld_f64(o, 4 + j), u[j], e2);
        Cq[s] = exp_neg(2.0 * e2);
        Er[s] = Ei[s] = Q[s] = 0.0;
    }
    double dx[3] = {0, 0, 0}, du[6] = {0, 0, 0, 0, 0, 0}, docc = 0;
    for (int kb = k0; kb < k1; kb += kStage) {
    const int nk = min(kStage, k1 - kb);
    for (int kk = 0; kk < nk; ++kk) {
        const int k = kb + kk;
        // wave-uniform: every lane is on reflection k
        const bool along = k > k0 && cont[k] != 0;
        if (along) {
#pragma unroll
            for (int s = 0; s < NOPS; s++) {
                const double qr = Q[s] * Mr[s], qi = Q[s] * Mi[s];
                const double er = Er[s] * qr - Ei[s] * qi;
                const double ei = Er[s] * qi + Ei[s] * qr;
                Er[s] = er;
                Ei[s] = ei;
                Q[s] *= Cq[s];
            }
        } else {
            const bool next = k + 1 < k1 && cont[k + 1] != 0;
#pragma unroll
            for (int s = 0; s < NOPS; s++) {
                const unsigned* rw = hop + ((size_t)k * NOPS + s) * (2 * kHklopRec);
                // h.(R x + t) reduced to [0, 1) revolutions
                const double ph = __builtin_amdgcn_fract(
                    fma(ld_f64(rw, 33), x2, fma(ld_f64(rw, 32), x1, fma(ld_f64(rw, 31), x0, ld_f64(rw, 34)))));
                double sn, cs;
                sincos_rev(ph, &sn, &cs);
                double e = 0.0;
#pragma unroll
                for (int j = 0; j < 6; j++) e = fma(ld_f64(rw, 25 + j), u[j], e);
                const double T = exp_neg(e);
                Er[s] = T * cs;
                Ei[s] = T * sn;
                if (next) {
                    const unsigned* rn = hop + ((size_t)(k + 1) * NOPS + s) * (2 * kHklopRec);
                    double e1 = 0.0;
#pragma unroll
                    for (int j = 0; j < 6; j++) e1 = fma(ld_f64(rn, 25 + j), u[j], e1);
                    Q[s] = exp_neg(e1 - e);
                }
            }
        }
        double S, G0, G1, G2, G3, G4;
        if (OTF) {
            double G[5];
            radial_otf<L, double>(tab, tr, ld_f64(hl32, k), pv, ik, ikp, anr, S, G);
            G0 = G[0]; G1 = G[1]; G2 = G[2]; G3 = G[3]; G4 = G[4];
        } else {
            const double* rr = rad + ((size_t)k * a.n_types + t) * kRadRec;
            S = rr[0] + anr; G0 = rr[1]; G1 = rr[2]; G2 = rr[3]; G3 = rr[4]; G4 = rr[5];
        }
        const double cr = ld_f64(dtdf32, 2 * k), ci = ld_f64(dtdf32, 2 * k + 1);
        double sr = 0, si = 0;
#pragma unroll
        for (int s = 0; s < NOPS; s++) {
            const unsigned* rw = hop + ((size_t)k * NOPS + s) * (2 * kHklopRec);
            double r[kHklopRec];
#pragma unroll
            for (int j = 0; j < kHklopRec; j++) r[j] = ld_f64(rw, j);
            double A[5] = {0, 0, 0, 0, 0};
            contract<L, double>(P, r, A);
            double fr = S + G0 * A[0] + G2 * A[2];
            double fi = ani * r[35] + G1 * A[1];
            if (L == 4) {
                fr += G4 * A[4];
                fi += G3 * A[3];
            }
            const double Xr = fr * Er[s] - fi * Ei[s], Xi = fr * Ei[s] + fi * Er[s];
            sr += Xr;
            si += Xi;
            if (DERIV) {
                const double zr = cr * Xr - ci * Xi, zi = cr * Xi + ci * Xr;
#pragma unroll
                for (int j = 0; j < 6; j++) du[j] -= r[25 + j] * zr;
#pragma unroll
                for (int j = 0; j < 3; j++) dx[j] -= r[31 + j] * zi;
            }
        }
        if (DERIV) docc += cr * sr - ci * si;
        stage_f64(stage, kk, lane, w * sr, w * si);
    }
    flush_f64(stage, nk, lane, fpart + ((size_t)wave * a.n_hkl + kb) * 2);
    }
    if (DERIV) {
        double* dp = dpart + ((size_t)chunk_id * a.n_slots + slot) * kNdBasic;
        dp[0] = dx[0]; dp[1] = dx[1]; dp[2] = dx[2];
#pragma unroll
        for (int j = 0; j < 6; j++) dp[3 + j] = du[j];
        dp[9] = docc;
    }
}

// Mixed precision: FP64 phase range reduction and FP64 accumulation across reflections, FP32
// harmonics / radial / Debye-Waller / sincos / exp per term.
template <int L, bool DERIV, int MP, bool OTF>
__device__ __forceinline__ void body_mixed(const MainArgs& a, int wave, int lane, int chunk_id, int k0, int k1,
                                           const float* __restrict__ hop, const float* __restrict__ rad,
                                           const float* __restrict__ radx, const unsigned* __restrict__ phs32,
                                           const unsigned* __restrict__ dtdf32, float2* __restrict__ stage,
                                           double* __restrict__ fpart, double* __restrict__ dpart)
{
    constexpr int NLM = (L + 1) * (L + 1);
    constexpr int kFlush = kStage;    // reflections accumulated in FP32 before an FP64 flush
    constexpr bool SC = MP == 1;      // multipole pass 1: F, xyz/U/occ and the Pval, kappa, kappa' sums
    constexpr bool BO = MP == 2;      // multipole pass 2: the global-frame B_lm sums only
    const int slot = wave * kWave + lane;
    const AtomRec& at = cmem(a.atoms)[slot];
    const double x0 = at.xf[0], x1 = at.xf[1], x2 = at.xf[2];
    float u[6];
#pragma unroll
    for (int j = 0; j < 6; j++) u[j] = (float)at.u[j];
    float P[NLM];
#pragma unroll
    for (int j = 0; j < NLM; j++) P[j] = (float)at.plm[j];
    const float w = (float)at.w, anr = (float)at.anom[0], ani = (float)at.anom[1];
    const int t = at.type;
    TabRec tr{0, 2};
    float pv = 0, ik = 0, ikp = 0;
    if (OTF) {
        tr = cmem(a.tabs)[at.wfn];
        pv = (float)at.pval;
        ik = (float)at.inv_kappa;
        ikp = (float)at.inv_kappa_p;
    }
    const float* tab = cmem(static_cast<const float*>(a.tab));
    LinTabCache<L> rtc;    // the lane's last table intervals (reflections in |h| order)
    const unsigned* hl32 = cmem(reinterpret_cast<const unsigned*>(a.hlen));
    double dx[3] = {0, 0, 0}, du[6] = {0, 0, 0, 0, 0, 0}, docc = 0;
    double B[BO ? NLM : 1];
#pragma unroll
    for (int j = 0; j < (BO ? NLM : 1); j++) B[j] = 0;
    double dpv = 0, dka = 0, dkp = 0;
    for (int kb = k0; kb < k1; kb += kFlush) {
        const int kend = min(kb + kFlush, k1);
        float fx[3] = {0, 0, 0}, fu[6] = {0, 0, 0, 0, 0, 0}, focc = 0;
        float fB[BO ? NLM : 1];
#pragma unroll
        for (int j = 0; j < (BO ? NLM : 1); j++) fB[j] = 0;
        float fpv = 0, fka = 0, fkp = 0;
        for (int k = kb; k < kend; ++k) {
            const int kk = k - kb;
            float S, G0, G1, G2, G3, G4;
            if (OTF) {
                float G[5];
                rtc.eval(tab, tr, (float)ld_f64(hl32, k), pv, ik, ikp, anr, S, G);
                G0 = G[0]; G1 = G[1]; G2 = G[2]; G3 = G[3]; G4 = G[4];
            } else {
                const float* rr = rad + ((size_t)k * a.n_types + t) * kRadRec;
                S = rr[0] + anr; G0 = rr[1]; G1 = rr[2]; G2 = rr[3]; G3 = rr[4]; G4 = rr[5];
            }
            float fval = 0, dSk = 0, dG[5] = {0, 0, 0, 0, 0};
            if (SC) {
                const float* rx = radx + ((size_t)k * a.n_types + t) * kRadRec;
                fval = rx[0];
                dSk = rx[1];
#pragma unroll
                for (int l = 0; l < 5; l++) dG[l] = rx[2 + l];
            }
            const float cr = (float)ld_f64(dtdf32, 2 * k), ci = (float)ld_f64(dtdf32, 2 * k + 1);
            float sr = 0, si = 0, Rh = 0;
            for (int s = 0; s < a.n_ops; ++s) {
                const float* r = hop + ((size_t)k * a.n_ops + s) * kHklopRec;
                const unsigned* pr = phs32 + ((size_t)k * a.n_ops + s) * 8;
                float A[5] = {0, 0, 0, 0, 0};
                float fr = 0, fi = 0;
                if (!BO) {
                    contract<L, float>(P, r, A);
                    fr = S + G0 * A[0] + G2 * A[2];
                    fi = ani * r[35] + G1 * A[1];
                    if (L == 4) {
                        fr += G4 * A[4];
                        fi += G3 * A[3];
                    }
                }
                // FP64 range reduction of h.(R x + t) to [0, 1) revolutions
                const double ph = fma(ld_f64(pr, 2), x2, fma(ld_f64(pr, 1), x1, fma(ld_f64(pr, 0), x0, ld_f64(pr, 3))));
                const float phf = (float)__builtin_amdgcn_fract(ph);
                float sn = __builtin_amdgcn_sinf(phf), cs = __builtin_amdgcn_cosf(phf);
                float e = r[25] * u[0] + r[26] * u[1] + r[27] * u[2] + r[28] * u[3] + r[29] * u[4] + r[30] * u[5];
                float T = __builtin_amdgcn_exp2f(-e);
                float Er = T * cs, Ei = T * sn;
                if (!BO) {
                    float Xr = fr * Er - fi * Ei, Xi = fr * Ei + fi * Er;
                    sr += Xr;
                    si += Xi;
                    if (DERIV) {
                        float zr = cr * Xr - ci * Xi, zi = cr * Xi + ci * Xr;
#pragma unroll
                        for (int j = 0; j < 6; j++) fu[j] -= r[25 + j] * zr;
#pragma unroll
                        for (int j = 0; j < 3; j++) fx[j] -= r[31 + j] * zi;
                    }
                }
                if (SC || BO) {
                    float cer = cr * Er - ci * Ei, cei = cr * Ei + ci * Er;
                    if (SC) {
                        Rh += cer;
                        fkp += (dG[0] * A[0] + dG[2] * A[2] + dG[4] * A[4]) * cer -
                               (dG[1] * A[1] + dG[3] * A[3]) * cei;
                    } else {
                        float q[5] = {G0 * cer, -G1 * cei, G2 * cer, -G3 * cei, G4 * cer};
#pragma unroll
                        for (int j = 0; j < NLM; j++) {
                            const int l = j < 1 ? 0 : j < 4 ? 1 : j < 9 ? 2 : j < 16 ? 3 : 4;
                            fB[j] += q[l] * r[j];
                        }
                    }
                }
            }
            if (DERIV && !BO) focc += cr * sr - ci * si;
            if (SC) {
                fpv += fval * Rh;
                fka += dSk * Rh;
            }
            if (!BO) stage_f32(stage, kk, lane, w * sr, w * si);
        }
        if (DERIV && !BO) {
#pragma unroll
            for (int j = 0; j < 3; j++) dx[j] += fx[j];
#pragma unroll
            for (int j = 0; j < 6; j++) du[j] += fu[j];
            docc += focc;
        }
        if (BO) {
#pragma unroll
            for (int j = 0; j < NLM; j++) B[j] += fB[j];
        }
        if (SC) {
            dpv += fpv;
            dka += fka;
            dkp += fkp;
        }
        if (!BO) flush_f32(stage, kend - kb, lane, reinterpret_cast<float*>(fpart) + ((size_t)wave * a.n_hkl + kb) * 2);
    }
    if (DERIV) {
        // U multipliers carried log2(e); undo for the derivative
        const double inv_log2e = 0.69314718055994530942;
        constexpr int ND = MP ? kNdMultipole : kNdBasic;
        double* dp = dpart + ((size_t)chunk_id * a.n_slots + slot) * ND;
        if (!BO) {
            dp[0] = dx[0]; dp[1] = dx[1]; dp[2] = dx[2];
#pragma unroll
            for (int j = 0; j < 6; j++) dp[3 + j] = du[j] * inv_log2e;
            dp[9] = docc;
        }
        if (BO) {
#pragma unroll
            for (int j = 0; j < 25; j++) dp[10 + j] = j < NLM ? B[j < NLM ? j : 0] : 0.0;
        }
        if (SC) {
            dp[35] = dpv;
            dp[36] = dka;
            dp[37] = dkp;
        }
        if (MP == 3) dp[35] = dp[36] = dp[37] = 0.0;
    }
}

// ------------------------------------------------------------------ diagonal point groups
// When every explicit operation has a diagonal rotation (R = diag(sx, sy, sz), the same signs in
// fractional and Cartesian axes: triclinic, monoclinic and orthorhombic groups) the rotated
// reflection is h_s = (sx h, sy k, sz l).  Each density-normalised harmonic has a definite parity
// in x, y and z, so d_j(h_s) = sign_j(s) d_j(h): the atom's P.d contraction is done once per
// reflection as parity-class sums and every operation combines them with its own signs
// (compile-time add/sub, one code path per sign pattern).  The Debye-Waller exponent splits the
// same way into a diagonal part and three signed cross terms.  Records are per reflection:
// [d_j(h) x 25, DW multipliers of h x 6, h x 3, pad] plus h.t_s per operation.  Mixed precision
// only: in FP64 the class sums double the register footprint and cost more occupancy than the
// saved contraction work.

// parity bits of harmonic j: 1 = odd in x, 2 = odd in y, 4 = odd in z
__host__ __device__ constexpr int harmonic_parity(int j)
{
    constexpr int t[25] = {0, 2, 4, 1, 3, 6, 0, 5, 0, 2, 7, 2, 4, 1, 4, 1, 3, 6, 3, 6, 0, 5, 0, 5, 0};
    return t[j];
}

__host__ __device__ constexpr int harmonic_l(int j) { return j < 1 ? 0 : j < 4 ? 1 : j < 9 ? 2 : j < 16 ? 3 : 4; }

__host__ __device__ constexpr bool class_present(int l, int key)
{
    for (int j = l * l; j < (l + 1) * (l + 1); j++)
        if (harmonic_parity(j) == key) return true;
    return false;
}

template <int L>
__host__ __device__ constexpr bool class_used(int key)
{
    for (int l = 0; l <= L; l++)
        if (class_present(l, key)) return true;
    return false;
}

__host__ __device__ constexpr bool odd_under(int key, int pat)
{
    return (__builtin_popcount((unsigned)(key & pat)) & 1) != 0;
}

template <int L, typename R>
__device__ __forceinline__ void class_sums(const R* P, const R* Y, R (&C)[5][8])
{
    constexpr int NLM = (L + 1) * (L + 1);
#pragma unroll
    for (int l = 0; l < 5; l++)
#pragma unroll
        for (int c = 0; c < 8; c++) C[l][c] = R(0);
#pragma unroll
    for (int j = 0; j < NLM; j++) C[harmonic_l(j)][harmonic_parity(j)] += P[j] * Y[j];
}

template <int L, int PAT, typename R>
__device__ __forceinline__ void signed_contraction(const R (&C)[5][8], R* A)
{
#pragma unroll
    for (int l = 0; l < 5; l++) {
        R acc = R(0);
        if (l <= L) {
#pragma unroll
            for (int c = 0; c < 8; c++)
                if (class_present(l, c)) acc = odd_under(c, PAT) ? acc - C[l][c] : acc + C[l][c];
        }
        A[l] = acc;
    }
}

// per-reflection values shared by all operations of one atom (working precision R)
// FP64 line kernel for diagonal groups (every rotation a sign matrix, at most four explicit
// operations): one record per reflection, [d_j(h) x 25, m(h) x 6, h x 3, h.t_s per operation,
// pad], instead of one per (reflection, operation).  Operation s has sign pattern
// (PATS >> 3 s) & 7 (bit i: axis i flips), so d_j(h_s) = +-d_j(h), the cross Debye-Waller
// multipliers and h_s flip sign by parity and the flips are compile-time negations of the
// record's scalar operands.  Phases and Debye-Waller factors are carried along the lines exactly
// as in body_fp64_line.
static constexpr int kDiag64Rec = 40;
// Sign-pattern sets with compiled kernels (bit 12 marks the code; operation s has pattern bits
// 3s..3s+2, bit i of a pattern: axis i flips).  The engine lists a diagonal group's operations in
// increasing pattern order, so every set is stored in that order: the identity first, then a
// 2-fold along y (P2_1/c, C2/c, P2_1), along z or along x, an inversion partner listed explicitly
// (a centre off the origin), a mirror normal to x, y (Pc, Cc) or z; the orthorhombic 222 set
// (identity, 2z, 2y, 2x) and mm2 with the 2-fold along z (Pna2_1, Pca2_1: identity, m_x, m_y, 2z),
// along y or along x.
static constexpr int kDiag64Pats0 = 4096 | (5 << 3);
static constexpr int kDiag64Pats1 = 4096 | (3 << 3);
static constexpr int kDiag64Pats2 = 4096 | (6 << 3);
static constexpr int kDiag64Pats3 = 4096 | (7 << 3);
static constexpr int kDiag64Pats4 = 4096 | (3 << 3) | (5 << 6) | (6 << 9);
static constexpr int kDiag64Pats5 = 4096 | (1 << 3);
static constexpr int kDiag64Pats6 = 4096 | (2 << 3);
static constexpr int kDiag64Pats7 = 4096 | (4 << 3);
static constexpr int kDiag64Pats8 = 4096 | (1 << 3) | (2 << 6) | (3 << 9);
static constexpr int kDiag64Pats9 = 4096 | (1 << 3) | (4 << 6) | (5 << 9);
static constexpr int kDiag64Pats10 = 4096 | (2 << 3) | (4 << 6) | (6 << 9);
// X(code, number of operations) for every compiled set
#define DH_FOR_EACH_PATTERN_SET(X)                                                                    \
    X(kDiag64Pats0, 2) X(kDiag64Pats1, 2) X(kDiag64Pats2, 2) X(kDiag64Pats3, 2) X(kDiag64Pats5, 2)     \
    X(kDiag64Pats6, 2) X(kDiag64Pats7, 2) X(kDiag64Pats4, 4) X(kDiag64Pats8, 4) X(kDiag64Pats9, 4)     \
    X(kDiag64Pats10, 4)
// added to a sign-pattern code: the mixed parity-folded kernel of that set without line recurrences
static constexpr int kDiagDirect = 8192;

template <int PAT>
__device__ __forceinline__ constexpr double psign(int key)
{
    return odd_under(key, PAT) ? -1.0 : 1.0;
}

template <int L, int PAT>
__device__ __forceinline__ void contract_signed(const double* P, const double* Y, double* A)
{
    constexpr int NLM = (L + 1) * (L + 1);
    A[0] = P[0] * Y[0];
    A[1] = fma(P[3], psign<PAT>(harmonic_parity(3)) * Y[3],
               fma(P[2], psign<PAT>(harmonic_parity(2)) * Y[2], P[1] * (psign<PAT>(harmonic_parity(1)) * Y[1])));
    double a2 = 0;
#pragma unroll
    for (int j = 4; j < 9; j++) a2 = fma(P[j], psign<PAT>(harmonic_parity(j)) * Y[j], a2);
    A[2] = a2;
    A[3] = A[4] = 0.0;
    if (L == 4) {
        double a3 = 0, a4 = 0;
#pragma unroll
        for (int j = 9; j < 16; j++) a3 = fma(P[j < NLM ? j : 0], psign<PAT>(harmonic_parity(j)) * Y[j], a3);
#pragma unroll
        for (int j = 16; j < 25; j++) a4 = fma(P[j < NLM ? j : 0], psign<PAT>(harmonic_parity(j)) * Y[j], a4);
        A[3] = a3;
        A[4] = a4;
    }
}

// Debye-Waller exponent m(h_s) . u with the cross terms signed by the pattern
template <int PAT>
__device__ __forceinline__ double dw_signed(const double* m, const double* u)
{
    double e = m[0] * u[0];
    e = fma(m[1], u[1], e);
    e = fma(m[2], u[2], e);
    e = fma(psign<PAT>(3) * m[3], u[3], e);
    e = fma(psign<PAT>(5) * m[4], u[4], e);
    e = fma(psign<PAT>(6) * m[5], u[5], e);
    return e;
}

// Form factors of every operation of a diagonal sign-pattern set at once.  Under operation s
// the harmonics of parity class c change sign when odd_under(c, pattern s); the classes with the
// same signs under every operation (the same sign key) enter every f_s together, so
//   f_s = S + sum_l G_l A_l(s) = sum_k sign_s(k) B[k],  B[k] = sum_l G_l sum_{j in l, key(j) = k} P_j d_j
// (real parts from even l and S, imaginary parts from odd l): one pass over the harmonics and
// the radial factors per reflection, then NOPS signed sums of the few keys, instead of a signed
// contraction and an assembly per operation.
template <int NOPS, int PATS>
__host__ __device__ constexpr int sign_key(int c)
{
    int k = 0;
    for (int s = 0; s < NOPS; s++)
        if (odd_under(c, (PATS >> (3 * s)) & 7)) k |= 1 << s;
    return k;
}

template <int L, int NOPS, int PATS>
__host__ __device__ constexpr bool key_bin_used(int l, int k)
{
    if (l > L) return false;
    for (int j = l * l; j < (l + 1) * (l + 1); j++)
        if (sign_key<NOPS, PATS>(harmonic_parity(j)) == k) return true;
    return false;
}

template <int L, int NOPS, int PATS>
__host__ __device__ constexpr bool key_used(int k, bool odd)
{
    for (int l = odd ? 1 : 0; l <= L; l += 2)
        if (key_bin_used<L, NOPS, PATS>(l, k)) return true;
    return !odd && k == 0;    // S enters the all-plus key
}

// Compile-time loop: f(std::integral_constant<int, i>) for i = B .. E - 1, so that tables indexed
// by i (harmonic parity, sign keys) are evaluated by the compiler, not at run time.
template <int B, int E, typename F>
__device__ __forceinline__ void static_for(F&& f)
{
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}

// The sign keys are linear in the class bits over GF(2) (odd_under(c1 ^ c2, p) = odd_under(c1, p) ^
// odd_under(c2, p)), so the keys of all classes span the subspace generated by the keys of x, y
// and z.  key_dim: its dimension D; key_of(alpha): the key with coordinates alpha in the basis;
// op_vector(s): operation s's signs on the basis keys (bit i: basis key i is odd under s).  For the
// compiled sets the NOPS operations take the 2^D sign vectors once each, so sums over operations
// signed by key are Walsh-Hadamard transforms (D * 2^D additions instead of NOPS * (NOPS - 1)).
template <int NOPS, int PATS>
__host__ __device__ constexpr int key_basis(int i)
{
    int v[3] = {0, 0, 0}, d = 0;
    const int c3[3] = {1, 2, 4};
    for (int q = 0; q < 3; q++) {
        const int k = sign_key<NOPS, PATS>(c3[q]);
        bool in = false;
        for (int a = 0; a < (1 << d); a++) {
            int x = 0;
            for (int j = 0; j < d; j++)
                if ((a >> j) & 1) x ^= v[j];
            if (x == k) in = true;
        }
        if (!in) v[d++] = k;
    }
    return i < 0 ? d : v[i];
}

template <int NOPS, int PATS>
__host__ __device__ constexpr int key_dim()
{
    return key_basis<NOPS, PATS>(-1);
}

template <int NOPS, int PATS>
__host__ __device__ constexpr int key_of(int alpha)
{
    int x = 0;
    for (int j = 0; j < key_dim<NOPS, PATS>(); j++)
        if ((alpha >> j) & 1) x ^= key_basis<NOPS, PATS>(j);
    return x;
}

template <int NOPS, int PATS>
__host__ __device__ constexpr int key_coords(int k)
{
    for (int a = 0; a < (1 << key_dim<NOPS, PATS>()); a++)
        if (key_of<NOPS, PATS>(a) == k) return a;
    return -1;
}

template <int NOPS, int PATS>
__host__ __device__ constexpr int op_vector(int s)
{
    int b = 0;
    for (int j = 0; j < key_dim<NOPS, PATS>(); j++)
        if ((key_basis<NOPS, PATS>(j) >> s) & 1) b |= 1 << j;
    return b;
}

template <int NOPS, int PATS>
__host__ __device__ constexpr bool ops_span_keys()
{
    if (NOPS != (1 << key_dim<NOPS, PATS>())) return false;
    for (int s = 0; s < NOPS; s++)
        for (int t = 0; t < s; t++)
            if (op_vector<NOPS, PATS>(s) == op_vector<NOPS, PATS>(t)) return false;
    return true;
}

// in-place Walsh-Hadamard transform: x[b] <- sum_a (-1)^popcount(a & b) x[a]
template <int D, typename R>
__device__ __forceinline__ void wht(R (&x)[1 << D])
{
    static_for<0, D>([&](auto Ii) {
        constexpr int i = decltype(Ii)::value;
        static_for<0, (1 << D)>([&](auto Ai) {
            constexpr int a = decltype(Ai)::value;
            if constexpr (((a >> i) & 1) == 0) {
                const R p = x[a], q = x[a | (1 << i)];
                x[a] = p + q;
                x[a | (1 << i)] = p - q;
            }
        });
    });
}

// Harmonic sums binned by (l, sign key): C[l][k] = sum_{j in l, key(j) = k} P_j d_j.
template <int L, int NOPS, int PATS, typename R>
__device__ __forceinline__ void keyed_bins(const R* P, const R* Y, R (&C)[5][1 << NOPS])
{
    static_assert(ops_span_keys<NOPS, PATS>(), "the operations must take every sign vector of the key basis once");
    constexpr int NK = 1 << NOPS;
    constexpr int NLM = (L + 1) * (L + 1);
    static_for<0, 5>([&](auto Li) {
        static_for<0, NK>([&](auto Ki) { C[decltype(Li)::value][decltype(Ki)::value] = R(0); });
    });
    static_for<0, NLM>([&](auto Ji) {
        constexpr int j = decltype(Ji)::value;
        constexpr int l = harmonic_l(j), k = sign_key<NOPS, PATS>(harmonic_parity(j));
        C[l][k] = fma(P[j], Y[j], C[l][k]);
    });
}

// f_s = S + sum_l G_l A_l(s) for every operation from the binned sums (real parts from even l
// and S, imaginary parts from odd l)
template <int L, int NOPS, int PATS, typename R>
__device__ __forceinline__ void keyed_weights(const R (&C)[5][1 << NOPS], R S, R G0, R G1, R G2, R G3, R G4,
                                              R (&fr)[NOPS], R (&fi)[NOPS])
{
    constexpr int NK = 1 << NOPS;
    const R G[5] = {G0, G1, G2, G3, G4};
    R Br[NK], Bi[NK];
    static_for<0, NK>([&](auto Ki) {
        constexpr int k = decltype(Ki)::value;
        Br[k] = k == 0 ? S : R(0);
        Bi[k] = R(0);
        static_for<0, L + 1>([&](auto Li) {
            constexpr int l = decltype(Li)::value;
            if constexpr (key_bin_used<L, NOPS, PATS>(l, k)) {
                if constexpr (l % 2 == 0) Br[k] = fma(G[l], C[l][k], Br[k]);
                else Bi[k] = fma(G[l], C[l][k], Bi[k]);
            }
        });
    });
    // f_s = sum_k sign_s(k) B[k]: a Walsh-Hadamard transform over the key basis
    constexpr int D = key_dim<NOPS, PATS>();
    R ar[1 << D], ai[1 << D];
    static_for<0, (1 << D)>([&](auto Ai) {
        constexpr int al = decltype(Ai)::value;
        ar[al] = Br[key_of<NOPS, PATS>(al)];
        ai[al] = Bi[key_of<NOPS, PATS>(al)];
    });
    wht<D>(ar);
    wht<D>(ai);
    static_for<0, NOPS>([&](auto Si) {
        constexpr int s = decltype(Si)::value;
        fr[s] = ar[op_vector<NOPS, PATS>(s)];
        fi[s] = ai[op_vector<NOPS, PATS>(s)];
    });
}

// Sums over the operations signed like each parity class: W[alpha] = sum_s sign x_s, class c's
// sum at W[key_coords(c)]; one Walsh-Hadamard transform of the operations' values.
template <int NOPS, int PATS, typename R>
__device__ __forceinline__ void class_sums_all(const R (&x)[NOPS], R (&w)[1 << key_dim<NOPS, PATS>()])
{
    constexpr int D = key_dim<NOPS, PATS>();
    static_for<0, NOPS>([&](auto Si) {
        constexpr int s = decltype(Si)::value;
        w[op_vector<NOPS, PATS>(s)] = x[s];
    });
    wht<D>(w);
}

template <int NOPS, int PATS>
__host__ __device__ constexpr int class_coords(int c)
{
    return key_coords<NOPS, PATS>(sign_key<NOPS, PATS>(c));
}

// MP: the passes of the multipole adjoint as in body_fp64 (1: F, xyz/U/occupancy and the Pval,
// kappa, kappa' sums; 2: the B_lm sums only, the class-signed sums of c E_s; 3: the plain pass in
// the multipole row layout).
template <int L, bool DERIV, bool OTF, int NOPS, int PATS, int MP = 0, bool LINES = true>
__device__ __forceinline__ void body_fp64_line_diag(const MainArgs& a, int wave, int lane, int chunk_id, int k0,
                                                    int k1, const unsigned* __restrict__ hop,
                                                    const double* __restrict__ rad, const double* __restrict__ radx,
                                                    const unsigned* __restrict__ dtdf32,
                                                    double2* __restrict__ stage, double* __restrict__ fpart,
                                                    double* __restrict__ dpart)
{
    constexpr int NLM = (L + 1) * (L + 1);
    constexpr bool SC = MP == 1;
    constexpr bool BO = MP == 2;
    static_assert(!(MP && OTF), "the multipole adjoint uses radial records");
    const int slot = wave * kWave + lane;
    const AtomRec& at = cmem(a.atoms)[slot];
    const double x[3] = {at.xf[0], at.xf[1], at.xf[2]};
    double u[6];
#pragma unroll
    for (int j = 0; j < 6; j++) u[j] = at.u[j];
    double P[NLM];
#pragma unroll
    for (int j = 0; j < NLM; j++) P[j] = at.plm[j];
    const double w = at.w, anr = at.anom[0], ani = at.anom[1];
    const int t = at.type;
    TabRec tr{0, 2};
    double pv = 0, ik = 0, ikp = 0;
    if (OTF) {
        tr = cmem(a.tabs)[at.wfn];
        pv = at.pval;
        ik = at.inv_kappa;
        ikp = at.inv_kappa_p;
    }
    const double* tab = cmem(static_cast<const double*>(a.tab));
    const unsigned* hl32 = cmem(reinterpret_cast<const unsigned*>(a.hlen));
    const int* cont = cmem(a.line_cont);
    const unsigned* lop = cmem(reinterpret_cast<const unsigned*>(a.line_ops));
    // With diagonal rotations the step of every operation is +-Delta, so the Debye-Waller step
    // ratio C = exp(-2 Delta_s^T U Delta_s) is the same for all of them.  LINES false (reflections
    // in |h| order for the table lookups): every term direct, no step factors.
    double Mr[NOPS], Mi[NOPS], Er[NOPS], Ei[NOPS], Q[NOPS], an[NOPS];
    double Cq = 0.0;
    if constexpr (LINES) {
        double e2 = 0.0;
#pragma unroll
        for (int j = 0; j < 6; j++) e2 = fma(ld_f64(lop, 4 + j), u[j], e2);
        Cq = exp_neg(2.0 * e2);
    }
#pragma unroll
    for (int s = 0; s < NOPS; s++) {
        Mr[s] = Mi[s] = 0.0;
        if constexpr (LINES) {
            const unsigned* o = lop + s * 20;
            const double ph = __builtin_amdgcn_fract(
                fma(ld_f64(o, 2), x[2], fma(ld_f64(o, 1), x[1], fma(ld_f64(o, 0), x[0], ld_f64(o, 3)))));
            sincos_rev(ph, &Mi[s], &Mr[s]);
        }
        Er[s] = Ei[s] = Q[s] = 0.0;
        an[s] = (double)a.diag_anom[(PATS >> (3 * s)) & 7] * ani;
    }
    double dx[3] = {0, 0, 0}, du[6] = {0, 0, 0, 0, 0, 0}, docc = 0;
    double B[BO ? NLM : 1];
#pragma unroll
    for (int j = 0; j < (BO ? NLM : 1); j++) B[j] = 0;
    double dpv = 0, dka = 0, dkp = 0;
    for (int kb = k0; kb < k1; kb += kStage) {
    const int nk = min(kStage, k1 - kb);
    for (int kk = 0; kk < nk; ++kk) {
        const int k = kb + kk;
        const unsigned* rw = hop + (size_t)k * (2 * kDiag64Rec);
        double m[6], hf[3];
#pragma unroll
        for (int j = 0; j < 6; j++) m[j] = ld_f64(rw, 25 + j);
#pragma unroll
        for (int j = 0; j < 3; j++) hf[j] = ld_f64(rw, 31 + j);
        const bool along = LINES && k > k0 && cont[k] != 0;
        if (along) {
#pragma unroll
            for (int s = 0; s < NOPS; s++) {
                const double qr = Q[s] * Mr[s], qi = Q[s] * Mi[s];
                const double er = Er[s] * qr - Ei[s] * qi;
                const double ei = Er[s] * qi + Ei[s] * qr;
                Er[s] = er;
                Ei[s] = ei;
                Q[s] *= Cq;
            }
        } else {
            const bool next = LINES && k + 1 < k1 && cont[k + 1] != 0;
            double mn[6];
            if (next) {
                const unsigned* rn = hop + (size_t)(k + 1) * (2 * kDiag64Rec);
#pragma unroll
                for (int j = 0; j < 6; j++) mn[j] = ld_f64(rn, 25 + j);
            }
#pragma unroll
            for (int s = 0; s < NOPS; s++) {
                double sn, cs, e, e1 = 0.0;
                const double hts = ld_f64(rw, 34 + s);
                switch ((PATS >> (3 * s)) & 7) {
#define DIAG64_DIRECT(PT)                                                                                       \
    case PT: {                                                                                                \
        const double ph = __builtin_amdgcn_fract(                                                             \
            fma(psign<PT>(4) * hf[2], x[2], fma(psign<PT>(2) * hf[1], x[1], fma(psign<PT>(1) * hf[0], x[0], hts)))); \
        sincos_rev(ph, &sn, &cs);                                                                             \
        e = dw_signed<PT>(m, u);                                                                              \
        if (next) e1 = dw_signed<PT>(mn, u);                                                                  \
        break;                                                                                                \
    }
                    DIAG64_DIRECT(0) DIAG64_DIRECT(1) DIAG64_DIRECT(2) DIAG64_DIRECT(3)
                    DIAG64_DIRECT(4) DIAG64_DIRECT(5) DIAG64_DIRECT(6) DIAG64_DIRECT(7)
#undef DIAG64_DIRECT
                }
                const double T = exp_neg(e);
                Er[s] = T * cs;
                Ei[s] = T * sn;
                if (next) Q[s] = exp_neg(e1 - e);
            }
        }
        double S, G0, G1, G2, G3, G4;
        if (OTF) {
            double G[5];
            radial_otf<L, double>(tab, tr, ld_f64(hl32, k), pv, ik, ikp, anr, S, G);
            G0 = G[0]; G1 = G[1]; G2 = G[2]; G3 = G[3]; G4 = G[4];
        } else {
            const double* rr = rad + ((size_t)k * a.n_types + t) * kRadRec;
            S = rr[0] + anr; G0 = rr[1]; G1 = rr[2]; G2 = rr[3]; G3 = rr[4]; G4 = rr[5];
        }
        const double cr = ld_f64(dtdf32, 2 * k), ci = ld_f64(dtdf32, 2 * k + 1);
        double Y[25];
#pragma unroll
        for (int j = 0; j < 25; j++) Y[j] = j < NLM ? ld_f64(rw, j) : 0.0;
        if constexpr (BO) {
            // B_j += G'_l d_j(h) V[class(j)], V the class-signed sums of Re (even classes) or Im
            // (odd classes) of c E_s
            double cer[NOPS], cei[NOPS];
#pragma unroll
            for (int s = 0; s < NOPS; s++) {
                cer[s] = cr * Er[s] - ci * Ei[s];
                cei[s] = cr * Ei[s] + ci * Er[s];
            }
            double Vr[1 << key_dim<NOPS, PATS>()], Vi[1 << key_dim<NOPS, PATS>()];
            class_sums_all<NOPS, PATS>(cer, Vr);
            class_sums_all<NOPS, PATS>(cei, Vi);
            const double Gs[5] = {G0, -G1, G2, -G3, G4};
            static_for<0, NLM>([&](auto Ji) {
                constexpr int j = decltype(Ji)::value;
                constexpr int c = harmonic_parity(j), q = class_coords<NOPS, PATS>(c);
                const double v = (__builtin_popcount((unsigned)c) & 1) ? Vi[q] : Vr[q];
                B[j] = fma(Gs[harmonic_l(j)] * Y[j], v, B[j]);
            });
            continue;
        }
        double dG[5] = {0, 0, 0, 0, 0}, fval = 0, dSk = 0, Rh = 0;
        if constexpr (SC) {
            const double* rx = radx + ((size_t)k * a.n_types + t) * kRadRec;
            fval = rx[0];
            dSk = rx[1];
#pragma unroll
            for (int l = 0; l < 5; l++) dG[l] = rx[2 + l];
        }
        double Xr[NOPS], Xi[NOPS];
#pragma unroll
        for (int s = 0; s < NOPS; s++) {
            double A[5] = {0, 0, 0, 0, 0};
            switch ((PATS >> (3 * s)) & 7) {
#define DIAG64_TERM(PT)                                                                                       \
    case PT: {                                                                                                \
        contract_signed<L, PT>(P, Y, A);                                                                      \
        break;                                                                                                \
    }
                DIAG64_TERM(0) DIAG64_TERM(1) DIAG64_TERM(2) DIAG64_TERM(3)
                DIAG64_TERM(4) DIAG64_TERM(5) DIAG64_TERM(6) DIAG64_TERM(7)
#undef DIAG64_TERM
            }
            double fr = S + G0 * A[0] + G2 * A[2];
            double fi = an[s] + G1 * A[1];
            if (L == 4) {
                fr += G4 * A[4];
                fi += G3 * A[3];
            }
            Xr[s] = fr * Er[s] - fi * Ei[s];
            Xi[s] = fr * Ei[s] + fi * Er[s];
            if constexpr (SC) {
                const double cer = cr * Er[s] - ci * Ei[s], cei = cr * Ei[s] + ci * Er[s];
                Rh += cer;
                dkp += (dG[0] * A[0] + dG[2] * A[2] + dG[4] * A[4]) * cer - (dG[1] * A[1] + dG[3] * A[3]) * cei;
            }
        }
        if constexpr (SC) {
            dpv += fval * Rh;
            dka += dSk * Rh;
        }
        // sums over the operations signed by parity class (as in the mixed line kernel)
        double Wr[1 << key_dim<NOPS, PATS>()], Wi[1 << key_dim<NOPS, PATS>()];
        class_sums_all<NOPS, PATS>(Xr, Wr);
        class_sums_all<NOPS, PATS>(Xi, Wi);
        const double sr = Wr[class_coords<NOPS, PATS>(0)], si = Wi[class_coords<NOPS, PATS>(0)];
        if (DERIV) {
            const double z0 = cr * sr - ci * si;
            docc += z0;
            du[0] -= m[0] * z0;
            du[1] -= m[1] * z0;
            du[2] -= m[2] * z0;
            static_for<0, 3>([&](auto Ji) {
                constexpr int j = decltype(Ji)::value;
                constexpr int cu = class_coords<NOPS, PATS>(j == 0 ? 3 : j == 1 ? 5 : 6);    // xy, xz, yz
                du[3 + j] -= m[3 + j] * (cr * Wr[cu] - ci * Wi[cu]);
                constexpr int cx = class_coords<NOPS, PATS>(1 << j);                         // x, y, z
                dx[j] -= hf[j] * (cr * Wi[cx] + ci * Wr[cx]);
            });
        }
        stage_f64(stage, kk, lane, w * sr, w * si);
    }
    if (!BO) flush_f64(stage, nk, lane, fpart + ((size_t)wave * a.n_hkl + kb) * 2);
    }
    if (DERIV) {
        constexpr int ND = MP ? kNdMultipole : kNdBasic;
        double* dp = dpart + ((size_t)chunk_id * a.n_slots + slot) * ND;
        if (!BO) {
            dp[0] = dx[0]; dp[1] = dx[1]; dp[2] = dx[2];
#pragma unroll
            for (int j = 0; j < 6; j++) dp[3 + j] = du[j];
            dp[9] = docc;
        }
        if (BO) {
#pragma unroll
            for (int j = 0; j < 25; j++) dp[10 + j] = j < NLM ? B[j < NLM ? j : 0] : 0.0;
        }
        if (SC) {
            dp[35] = dpv;
            dp[36] = dka;
            dp[37] = dkp;
        }
        if (MP == 3) dp[35] = dp[36] = dp[37] = 0.0;
    }
}

template <typename R>
struct DiagRefl {
    R C[5][8];             // parity-class sums of P_j d_j(h)
    R D, O3, O4, O5;       // DW exponent: diagonal part and the xy, xz, yz cross terms
    R S, G0, G1, G2, G3, G4, cr, ci;
    R dG[5];               // d G_l / d kappa' (first multipole pass only)
};

// Mixed precision: one operation with sign pattern PAT.  SC adds the sums of the first
// multipole pass: Rh = sum Re(c E) and the kappa' sum.
template <int L, bool DERIV, int PAT, bool SC = false>
__device__ __forceinline__ void diag_term_mixed(const DiagRefl<float>& q, const float* __restrict__ r, double h0,
                                                double h1, double h2, double ht, float an, float ani, double x0,
                                                double x1, double x2, float& sr, float& si, float (&fx)[3],
                                                float (&fu)[6], float& Rh, float& fkp)
{
    constexpr double sx = (PAT & 1) ? -1.0 : 1.0, sy = (PAT & 2) ? -1.0 : 1.0, sz = (PAT & 4) ? -1.0 : 1.0;
    constexpr float fxy = float(sx * sy), fxz = float(sx * sz), fyz = float(sy * sz);
    float A[5];
    signed_contraction<L, PAT>(q.C, A);
    float fr = q.S + q.G0 * A[0] + q.G2 * A[2];
    float fi = ani * an + q.G1 * A[1];
    if (L == 4) {
        fr += q.G4 * A[4];
        fi += q.G3 * A[3];
    }
    const double ph = __builtin_amdgcn_fract(fma(sz * h2, x2, fma(sy * h1, x1, fma(sx * h0, x0, ht))));
    const float phf = (float)ph;
    const float sn = __builtin_amdgcn_sinf(phf), cs = __builtin_amdgcn_cosf(phf);
    const float e = q.D + fxy * q.O3 + fxz * q.O4 + fyz * q.O5;
    const float T = __builtin_amdgcn_exp2f(-e);
    const float Er = T * cs, Ei = T * sn;
    const float Xr = fr * Er - fi * Ei, Xi = fr * Ei + fi * Er;
    sr += Xr;
    si += Xi;
    if (DERIV) {
        const float zr = q.cr * Xr - q.ci * Xi, zi = q.cr * Xi + q.ci * Xr;
        fu[0] -= r[25] * zr;
        fu[1] -= r[26] * zr;
        fu[2] -= r[27] * zr;
        fu[3] -= (fxy * r[28]) * zr;
        fu[4] -= (fxz * r[29]) * zr;
        fu[5] -= (fyz * r[30]) * zr;
        fx[0] -= (float(sx) * r[31]) * zi;
        fx[1] -= (float(sy) * r[32]) * zi;
        fx[2] -= (float(sz) * r[33]) * zi;
    }
    if (SC) {
        const float cer = q.cr * Er - q.ci * Ei, cei = q.cr * Ei + q.ci * Er;
        Rh += cer;
        fkp += (q.dG[0] * A[0] + q.dG[2] * A[2] + q.dG[4] * A[4]) * cer - (q.dG[1] * A[1] + q.dG[3] * A[3]) * cei;
    }
}

// Second multipole pass, one operation with sign pattern PAT: V[c] += sign_c(PAT) Re(c E) for the
// even parity classes (even l) and sign_c(PAT) Im(c E) for the odd ones, so that afterwards
// B_j += G'_l d_j(h) V[class(j)] sums every operation at once.
template <int L, int PAT>
__device__ __forceinline__ void diag_b_term(const DiagRefl<float>& q, double h0, double h1, double h2, double ht,
                                            double x0, double x1, double x2, float (&V)[8])
{
    constexpr double sx = (PAT & 1) ? -1.0 : 1.0, sy = (PAT & 2) ? -1.0 : 1.0, sz = (PAT & 4) ? -1.0 : 1.0;
    constexpr float fxy = float(sx * sy), fxz = float(sx * sz), fyz = float(sy * sz);
    const double ph = __builtin_amdgcn_fract(fma(sz * h2, x2, fma(sy * h1, x1, fma(sx * h0, x0, ht))));
    const float phf = (float)ph;
    const float sn = __builtin_amdgcn_sinf(phf), cs = __builtin_amdgcn_cosf(phf);
    const float e = q.D + fxy * q.O3 + fxz * q.O4 + fyz * q.O5;
    const float T = __builtin_amdgcn_exp2f(-e);
    const float Er = T * cs, Ei = T * sn;
    const float cer = q.cr * Er - q.ci * Ei, cei = q.cr * Ei + q.ci * Er;
#pragma unroll
    for (int c = 0; c < 8; c++) {
        if (!class_used<L>(c)) continue;
        const float v = (__builtin_popcount((unsigned)c) & 1) ? cei : cer;
        V[c] = odd_under(c, PAT) ? V[c] - v : V[c] + v;
    }
}

template <int L, bool DERIV, int MP, bool OTF>
__device__ __forceinline__ void body_mixed_diag(const MainArgs& a, int wave, int lane, int chunk_id, int k0, int k1,
                                                const float* __restrict__ hrec, const float* __restrict__ rad,
                                                const float* __restrict__ radx, const unsigned* __restrict__ hd32,
                                                const unsigned* __restrict__ ht32, const unsigned* __restrict__ dtdf32,
                                                float2* __restrict__ stage, double* __restrict__ fpart,
                                                double* __restrict__ dpart)
{
    constexpr int NLM = (L + 1) * (L + 1);
    constexpr bool SC = MP == 1;    // multipole pass 1: F, xyz/U/occ and the Pval, kappa, kappa' sums
    constexpr bool BO = MP == 2;    // multipole pass 2: the global-frame B_lm sums only
    const int slot = wave * kWave + lane;
    const AtomRec& at = cmem(a.atoms)[slot];
    const double x0 = at.xf[0], x1 = at.xf[1], x2 = at.xf[2];
    float u[6];
#pragma unroll
    for (int j = 0; j < 6; j++) u[j] = (float)at.u[j];
    float P[NLM];
#pragma unroll
    for (int j = 0; j < NLM; j++) P[j] = (float)at.plm[j];
    const float w = (float)at.w, anr = (float)at.anom[0], ani = (float)at.anom[1];
    const int t = at.type;
    TabRec tr{0, 2};
    float pv = 0, ik = 0, ikp = 0;
    if (OTF) {
        tr = cmem(a.tabs)[at.wfn];
        pv = (float)at.pval;
        ik = (float)at.inv_kappa;
        ikp = (float)at.inv_kappa_p;
    }
    const float* tab = cmem(static_cast<const float*>(a.tab));
    LinTabCache<L> rtc;    // the lane's last table intervals (reflections in |h| order)
    const unsigned* hl32 = cmem(reinterpret_cast<const unsigned*>(a.hlen));
    double dx[3] = {0, 0, 0}, du[6] = {0, 0, 0, 0, 0, 0}, docc = 0;
    double B[BO ? NLM : 1];
#pragma unroll
    for (int j = 0; j < (BO ? NLM : 1); j++) B[j] = 0;
    double dpv = 0, dka = 0, dkp = 0;
    for (int kb = k0; kb < k1; kb += kStage) {
        const int kend = min(kb + kStage, k1);
        float fx[3] = {0, 0, 0}, fu[6] = {0, 0, 0, 0, 0, 0}, focc = 0;
        float fB[BO ? NLM : 1];
#pragma unroll
        for (int j = 0; j < (BO ? NLM : 1); j++) fB[j] = 0;
        float fpv = 0, fka = 0, fkp = 0;
        for (int k = kb; k < kend; ++k) {
            const float* r = hrec + (size_t)k * kHklopRec;
            DiagRefl<float> q;
            if (OTF) {
                float G[5];
                rtc.eval(tab, tr, (float)ld_f64(hl32, k), pv, ik, ikp, anr, q.S, G);
                q.G0 = G[0]; q.G1 = G[1]; q.G2 = G[2]; q.G3 = G[3]; q.G4 = G[4];
            } else {
                const float* rr = rad + ((size_t)k * a.n_types + t) * kRadRec;
                q.S = rr[0] + anr;
                q.G0 = rr[1];
                q.G1 = rr[2];
                q.G2 = rr[3];
                q.G3 = rr[4];
                q.G4 = rr[5];
            }
            float fval = 0, dSk = 0;
#pragma unroll
            for (int l = 0; l < 5; l++) q.dG[l] = 0;
            if (SC) {
                const float* rx = radx + ((size_t)k * a.n_types + t) * kRadRec;
                fval = rx[0];
                dSk = rx[1];
#pragma unroll
                for (int l = 0; l < 5; l++) q.dG[l] = rx[2 + l];
            }
            q.cr = (float)ld_f64(dtdf32, 2 * k);
            q.ci = (float)ld_f64(dtdf32, 2 * k + 1);
            if (!BO) class_sums<L, float>(P, r, q.C);
            q.D = r[25] * u[0] + r[26] * u[1] + r[27] * u[2];
            q.O3 = r[28] * u[3];
            q.O4 = r[29] * u[4];
            q.O5 = r[30] * u[5];
            const unsigned* hd = hd32 + (size_t)k * 8;
            const double h0 = ld_f64(hd, 0), h1 = ld_f64(hd, 1), h2 = ld_f64(hd, 2);
            const unsigned* htk = ht32 + (size_t)k * a.n_ops * 2;
            if (BO) {
                float V[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#define DIAG_OP(PAT)                                                                                          \
    if (a.diag_mask & (1 << PAT)) diag_b_term<L, PAT>(q, h0, h1, h2, ld_f64(htk, a.diag_op[PAT]), x0, x1, x2, V);
                DIAG_OP(0) DIAG_OP(1) DIAG_OP(2) DIAG_OP(3) DIAG_OP(4) DIAG_OP(5) DIAG_OP(6) DIAG_OP(7)
#undef DIAG_OP
                const float Gs[5] = {q.G0, -q.G1, q.G2, -q.G3, q.G4};
#pragma unroll
                for (int j = 0; j < NLM; j++) fB[j] += (Gs[harmonic_l(j)] * r[j]) * V[harmonic_parity(j)];
            } else {
                float sr = 0, si = 0, Rh = 0;
                // one uniform block per sign pattern present in the group (each at most once)
#define DIAG_OP(PAT)                                                                                          \
    if (a.diag_mask & (1 << PAT))                                                                             \
        diag_term_mixed<L, DERIV, PAT, SC>(q, r, h0, h1, h2, ld_f64(htk, a.diag_op[PAT]), a.diag_anom[PAT], ani, \
                                           x0, x1, x2, sr, si, fx, fu, Rh, fkp);
                DIAG_OP(0) DIAG_OP(1) DIAG_OP(2) DIAG_OP(3) DIAG_OP(4) DIAG_OP(5) DIAG_OP(6) DIAG_OP(7)
#undef DIAG_OP
                if (DERIV) focc += q.cr * sr - q.ci * si;
                if (SC) {
                    fpv += fval * Rh;
                    fka += dSk * Rh;
                }
                stage_f32(stage, k - kb, lane, w * sr, w * si);
            }
        }
        if (DERIV && !BO) {
#pragma unroll
            for (int j = 0; j < 3; j++) dx[j] += fx[j];
#pragma unroll
            for (int j = 0; j < 6; j++) du[j] += fu[j];
            docc += focc;
        }
        if (BO) {
#pragma unroll
            for (int j = 0; j < NLM; j++) B[j] += fB[j];
        }
        if (SC) {
            dpv += fpv;
            dka += fka;
            dkp += fkp;
        }
        if (!BO) flush_f32(stage, kend - kb, lane, reinterpret_cast<float*>(fpart) + ((size_t)wave * a.n_hkl + kb) * 2);
    }
    if (DERIV) {
        const double inv_log2e = 0.69314718055994530942;
        constexpr int ND = MP ? kNdMultipole : kNdBasic;
        double* dp = dpart + ((size_t)chunk_id * a.n_slots + slot) * ND;
        if (!BO) {
            dp[0] = dx[0]; dp[1] = dx[1]; dp[2] = dx[2];
#pragma unroll
            for (int j = 0; j < 6; j++) dp[3 + j] = du[j] * inv_log2e;
            dp[9] = docc;
        }
        if (BO) {
#pragma unroll
            for (int j = 0; j < 25; j++) dp[10 + j] = j < NLM ? B[j < NLM ? j : 0] : 0.0;
        }
        if (SC) {
            dp[35] = dpv;
            dp[36] = dka;
            dp[37] = dkp;
        }
        if (MP == 3) dp[35] = dp[36] = dp[37] = 0.0;
    }
}

// ------------------------------------------------------------------ mixed precision along hkl lines
// Reflections ordered along hkl lines (line_cont[k] = 1 when k continues the line of k - 1 by one
// step Delta along the line axis).  The FP32 factor E_s = T_s exp(i phi_s) of each operation is
// evaluated directly (FP64 range reduction of the phase, FP32 sin, cos and exp2) at every
// kLineRestart-th reflection of a staging block and wherever a line starts, and carried to the
// next reflection of the same line by E <- E (Q M), Q <- Q C: M = exp(2 pi i (Delta R_s x + Delta t_s))
// and C = exp(-2 Delta_s^T U Delta_s) are the atom's per-operation step factors (computed in FP64,
// rounded once), Q = T(h + Delta) / T(h) comes from the two direct exponents.  The roundings of M,
// Q, C and of each product compound along a chain (the error grows about quadratically with its
// length); chains of at most kLineRestart - 1 = 3 steps keep max|dF|/max|F| within about twice
// the direct FP32 evaluation's (8 steps: about four times).  Replaces the phase reduction, two
// sines/cosines and the exponential of three terms in four by one complex and three real products.
static constexpr int kLineRestart = 4;
static_assert(kStage % kLineRestart == 0 && (kLineRestart & (kLineRestart - 1)) == 0,
              "restarts fall on staging-block boundaries");
template <int L, bool DERIV, bool OTF, int NOPS>
__device__ __forceinline__ void line_steps_f32(const MainArgs& a, const double* xf, const double* u64, float (&Mr)[NOPS],
                                               float (&Mi)[NOPS], float (&Cq)[NOPS])
{
    const unsigned* lop = cmem(reinterpret_cast<const unsigned*>(a.line_ops));
#pragma unroll
    for (int s = 0; s < NOPS; s++) {
        const unsigned* o = lop + s * 20;
        const double ph = __builtin_amdgcn_fract(
            fma(ld_f64(o, 2), xf[2], fma(ld_f64(o, 1), xf[1], fma(ld_f64(o, 0), xf[0], ld_f64(o, 3)))));
        double sn, cs;
        sincos_rev(ph, &sn, &cs);
        Mr[s] = (float)cs;
        Mi[s] = (float)sn;
        double e2 = 0.0;
#pragma unroll
        for (int j = 0; j < 6; j++) e2 = fma(ld_f64(o, 4 + j), u64[j], e2);
        Cq[s] = (float)exp_neg(2.0 * e2);
    }
}

// Diagonal groups (compile-time sign patterns PATS as in the FP64 diagonal line kernel): form
// factors from sign-keyed harmonic sums, derivative sums per parity class.  LINES false
// evaluates every term directly (reflections in any order: the batched step, whose per-atom
// table lookups want |h| order).  MP: the passes of the multipole adjoint as in body_mixed_diag
// (1: F, xyz/U/occupancy and the Pval, kappa, kappa' sums; 2: the B_lm sums only; 3: the plain
// pass in the multipole row layout).
template <int L, bool DERIV, bool OTF, int NOPS, int PATS, bool LINES = true, int MP = 0>
__device__ __forceinline__ void body_mixed_line_diag(const MainArgs& a, int wave, int lane, int chunk_id, int k0,
                                                     int k1, const float* __restrict__ hrec,
                                                     const float* __restrict__ rad, const float* __restrict__ radx,
                                                     const unsigned* __restrict__ hd32,
                                                     const unsigned* __restrict__ ht32,
                                                     const unsigned* __restrict__ dtdf32, float2* __restrict__ stage,
                                                     double* __restrict__ fpart, double* __restrict__ dpart)
{
    constexpr int NLM = (L + 1) * (L + 1);
    constexpr bool SC = MP == 1;    // multipole pass 1: F, xyz/U/occ and the Pval, kappa, kappa' sums
    constexpr bool BO = MP == 2;    // multipole pass 2: the global-frame B_lm sums only
    static_assert(!(MP && OTF), "the multipole adjoint uses radial records");
    const int slot = wave * kWave + lane;
    const AtomRec& at = cmem(a.atoms)[slot];
    const double x0 = at.xf[0], x1 = at.xf[1], x2 = at.xf[2];
    float u[6];
#pragma unroll
    for (int j = 0; j < 6; j++) u[j] = (float)at.u[j];
    float P[NLM];
#pragma unroll
    for (int j = 0; j < NLM; j++) P[j] = (float)at.plm[j];
    const float w = (float)at.w, anr = (float)at.anom[0], ani = (float)at.anom[1];
    const int t = at.type;
    TabRec tr{0, 2};
    float pv = 0, ik = 0, ikp = 0;
    if (OTF) {
        tr = cmem(a.tabs)[at.wfn];
        pv = (float)at.pval;
        ik = (float)at.inv_kappa;
        ikp = (float)at.inv_kappa_p;
    }
    const float* tab = cmem(static_cast<const float*>(a.tab));
    LinTabCache<L> rtc;    // the lane's last table intervals (reflections in |h| order)
    const unsigned* hl32 = cmem(reinterpret_cast<const unsigned*>(a.hlen));
    const int* cont = LINES ? cmem(a.line_cont) : nullptr;
    float Mr[NOPS], Mi[NOPS], Cq[NOPS], Er[NOPS], Ei[NOPS], Q[NOPS], an[NOPS];
    if (LINES) {
        const double xf[3] = {x0, x1, x2};
        double u64[6];
#pragma unroll
        for (int j = 0; j < 6; j++) u64[j] = at.u[j];
        line_steps_f32<L, DERIV, OTF, NOPS>(a, xf, u64, Mr, Mi, Cq);
    }
#pragma unroll
    for (int s = 0; s < NOPS; s++) {
        if (!LINES) Mr[s] = Mi[s] = Cq[s] = 0.f;
        // diagonal rotations step every operation by +-Delta: one Debye-Waller step ratio for all
        if (s > 0) Cq[s] = Cq[0];
        Er[s] = Ei[s] = Q[s] = 0.f;
        an[s] = a.diag_anom[(PATS >> (3 * s)) & 7] * ani;
    }
    double dx[3] = {0, 0, 0}, du[6] = {0, 0, 0, 0, 0, 0}, docc = 0;
    double B[BO ? NLM : 1];
#pragma unroll
    for (int j = 0; j < (BO ? NLM : 1); j++) B[j] = 0;
    double dpv = 0, dka = 0, dkp = 0;
    for (int kb = k0; kb < k1; kb += kStage) {
        const int kend = min(kb + kStage, k1);
        float fx[3] = {0, 0, 0}, fu[6] = {0, 0, 0, 0, 0, 0}, focc = 0;
        float fB[BO ? NLM : 1];
#pragma unroll
        for (int j = 0; j < (BO ? NLM : 1); j++) fB[j] = 0;
        float fpv = 0, fka = 0, fkp = 0;
        for (int k = kb; k < kend; ++k) {
            const float* r = hrec + (size_t)k * kHklopRec;
            // wave-uniform: every lane is on reflection k
            const bool along = LINES && ((k - kb) & (kLineRestart - 1)) != 0 && cont[k] != 0;
            if (along) {
#pragma unroll
                for (int s = 0; s < NOPS; s++) {
                    const float qr = Q[s] * Mr[s], qi = Q[s] * Mi[s];
                    const float er = Er[s] * qr - Ei[s] * qi;
                    const float ei = Er[s] * qi + Ei[s] * qr;
                    Er[s] = er;
                    Ei[s] = ei;
                    Q[s] *= Cq[s];
                }
            } else {
                const bool next =
                    LINES && k + 1 < kend && ((k + 1 - kb) & (kLineRestart - 1)) != 0 && cont[k + 1] != 0;
                const unsigned* hd = hd32 + (size_t)k * 8;
                const double h0 = ld_f64(hd, 0), h1 = ld_f64(hd, 1), h2 = ld_f64(hd, 2);
                const unsigned* htk = ht32 + (size_t)k * a.n_ops * 2;
                const float D = r[25] * u[0] + r[26] * u[1] + r[27] * u[2];
                const float O3 = r[28] * u[3], O4 = r[29] * u[4], O5 = r[30] * u[5];
                float Dn = 0.f, O3n = 0.f, O4n = 0.f, O5n = 0.f;
                if (next) {
                    const float* rn = r + kHklopRec;
                    Dn = rn[25] * u[0] + rn[26] * u[1] + rn[27] * u[2];
                    O3n = rn[28] * u[3];
                    O4n = rn[29] * u[4];
                    O5n = rn[30] * u[5];
                }
#pragma unroll
                for (int s = 0; s < NOPS; s++) {
                    const double hts = ld_f64(htk, s);
                    switch ((PATS >> (3 * s)) & 7) {
#define MLINE_DIRECT(PT)                                                                                      \
    case PT: {                                                                                                \
        constexpr float fxy = float(psign<PT>(3)), fxz = float(psign<PT>(5)), fyz = float(psign<PT>(6));      \
        const double ph = __builtin_amdgcn_fract(                                                             \
            fma(psign<PT>(4) * h2, x2, fma(psign<PT>(2) * h1, x1, fma(psign<PT>(1) * h0, x0, hts))));         \
        const float phf = (float)ph;                                                                          \
        const float sn = __builtin_amdgcn_sinf(phf), cs = __builtin_amdgcn_cosf(phf);                         \
        const float e = D + fxy * O3 + fxz * O4 + fyz * O5;                                                   \
        const float T = __builtin_amdgcn_exp2f(-e);                                                           \
        Er[s] = T * cs;                                                                                       \
        Ei[s] = T * sn;                                                                                       \
        if (next) Q[s] = __builtin_amdgcn_exp2f(e - (Dn + fxy * O3n + fxz * O4n + fyz * O5n));                \
        break;                                                                                                \
    }
                        MLINE_DIRECT(0) MLINE_DIRECT(1) MLINE_DIRECT(2) MLINE_DIRECT(3)
                        MLINE_DIRECT(4) MLINE_DIRECT(5) MLINE_DIRECT(6) MLINE_DIRECT(7)
#undef MLINE_DIRECT
                    }
                }
            }
            float S, G0, G1, G2, G3, G4;
            if (OTF) {
                float G[5];
                rtc.eval(tab, tr, (float)ld_f64(hl32, k), pv, ik, ikp, anr, S, G);
                G0 = G[0]; G1 = G[1]; G2 = G[2]; G3 = G[3]; G4 = G[4];
            } else {
                const float* rr = rad + ((size_t)k * a.n_types + t) * kRadRec;
                S = rr[0] + anr; G0 = rr[1]; G1 = rr[2]; G2 = rr[3]; G3 = rr[4]; G4 = rr[5];
            }
            const float cr = (float)ld_f64(dtdf32, 2 * k), ci = (float)ld_f64(dtdf32, 2 * k + 1);
            if constexpr (BO) {
                // B_j += G'_l d_j(h) V[class(j)], V the class-signed sums of Re (even classes) or
                // Im (odd classes) of c E_s
                float cer[NOPS], cei[NOPS];
#pragma unroll
                for (int s = 0; s < NOPS; s++) {
                    cer[s] = cr * Er[s] - ci * Ei[s];
                    cei[s] = cr * Ei[s] + ci * Er[s];
                }
                float Vr[1 << key_dim<NOPS, PATS>()], Vi[1 << key_dim<NOPS, PATS>()];
                class_sums_all<NOPS, PATS>(cer, Vr);
                class_sums_all<NOPS, PATS>(cei, Vi);
                const float Gs[5] = {G0, -G1, G2, -G3, G4};
                static_for<0, NLM>([&](auto Ji) {
                    constexpr int j = decltype(Ji)::value;
                    constexpr int c = harmonic_parity(j), q = class_coords<NOPS, PATS>(c);
                    const float v = (__builtin_popcount((unsigned)c) & 1) ? Vi[q] : Vr[q];
                    fB[j] += (Gs[harmonic_l(j)] * r[j]) * v;
                });
                continue;
            }
            float C[5][1 << NOPS];
            keyed_bins<L, NOPS, PATS, float>(P, r, C);
            float frs[NOPS], fis[NOPS];
            keyed_weights<L, NOPS, PATS, float>(C, S, G0, G1, G2, G3, G4, frs, fis);
            if constexpr (SC) {
                // kappa' sums: the form factors with dG_l / dkappa' (no spherical or anomalous
                // term) against c E_s; Pval and kappa sums: f_val and dS / dkappa against sum Re(c E_s)
                const float* rx = radx + ((size_t)k * a.n_types + t) * kRadRec;
                float dfr[NOPS], dfi[NOPS];
                keyed_weights<L, NOPS, PATS, float>(C, 0.f, rx[2], rx[3], rx[4], rx[5], rx[6], dfr, dfi);
                float Rh = 0.f;
#pragma unroll
                for (int s = 0; s < NOPS; s++) {
                    const float cer = cr * Er[s] - ci * Ei[s], cei = cr * Ei[s] + ci * Er[s];
                    Rh += cer;
                    fkp += dfr[s] * cer - dfi[s] * cei;
                }
                fpv += rx[0] * Rh;
                fka += rx[1] * Rh;
            }
            // X_s = f_s E_s; the sums over operations that F and the derivatives need are the
            // terms signed by a parity class (the ADP cross terms xy, xz, yz and the x, y, z
            // components change sign with the operation like the harmonics of that class), so
            // they are formed once per reflection and multiplied by the adjoint c once
            float Xr[NOPS], Xi[NOPS];
#pragma unroll
            for (int s = 0; s < NOPS; s++) {
                const float fr = frs[s], fi = an[s] + fis[s];
                Xr[s] = fr * Er[s] - fi * Ei[s];
                Xi[s] = fr * Ei[s] + fi * Er[s];
            }
            float Wr[1 << key_dim<NOPS, PATS>()], Wi[1 << key_dim<NOPS, PATS>()];
            class_sums_all<NOPS, PATS>(Xr, Wr);
            class_sums_all<NOPS, PATS>(Xi, Wi);
            const float sr = Wr[class_coords<NOPS, PATS>(0)], si = Wi[class_coords<NOPS, PATS>(0)];
            if (DERIV) {
                const float z0 = cr * sr - ci * si;
                focc += z0;
                fu[0] -= r[25] * z0;
                fu[1] -= r[26] * z0;
                fu[2] -= r[27] * z0;
                static_for<0, 3>([&](auto Ji) {
                    constexpr int j = decltype(Ji)::value;
                    constexpr int cu = class_coords<NOPS, PATS>(j == 0 ? 3 : j == 1 ? 5 : 6);    // xy, xz, yz
                    fu[3 + j] -= r[28 + j] * (cr * Wr[cu] - ci * Wi[cu]);
                    constexpr int cx = class_coords<NOPS, PATS>(1 << j);                         // x, y, z
                    fx[j] -= r[31 + j] * (cr * Wi[cx] + ci * Wr[cx]);
                });
            }
            stage_f32(stage, k - kb, lane, w * sr, w * si);
        }
        if (DERIV && !BO) {
#pragma unroll
            for (int j = 0; j < 3; j++) dx[j] += fx[j];
#pragma unroll
            for (int j = 0; j < 6; j++) du[j] += fu[j];
            docc += focc;
        }
        if (BO) {
#pragma unroll
            for (int j = 0; j < (BO ? NLM : 1); j++) B[j] += fB[j];
        }
        if (SC) {
            dpv += fpv;
            dka += fka;
            dkp += fkp;
        }
        if (!BO) flush_f32(stage, kend - kb, lane, reinterpret_cast<float*>(fpart) + ((size_t)wave * a.n_hkl + kb) * 2);
    }
    if (DERIV) {
        const double inv_log2e = 0.69314718055994530942;
        constexpr int ND = MP ? kNdMultipole : kNdBasic;
        double* dp = dpart + ((size_t)chunk_id * a.n_slots + slot) * ND;
        if (!BO) {
            dp[0] = dx[0]; dp[1] = dx[1]; dp[2] = dx[2];
#pragma unroll
            for (int j = 0; j < 6; j++) dp[3 + j] = du[j] * inv_log2e;
            dp[9] = docc;
        }
        if (BO) {
#pragma unroll
            for (int j = 0; j < 25; j++) dp[10 + j] = j < NLM ? B[j < NLM ? j : 0] : 0.0;
        }
        if (SC) {
            dp[35] = dpv;
            dp[36] = dka;
            dp[37] = dkp;
        }
        if (MP == 3) dp[35] = dp[36] = dp[37] = 0.0;
    }
}

// General rotations (per-(reflection, operation) records and FP64 phase records of body_mixed),
// NOPS explicit operations.
template <int L, bool DERIV, bool OTF, int NOPS>
__device__ __forceinline__ void body_mixed_line(const MainArgs& a, int wave, int lane, int chunk_id, int k0, int k1,
                                                const float* __restrict__ hop, const float* __restrict__ rad,
                                                const unsigned* __restrict__ phs32,
                                                const unsigned* __restrict__ dtdf32, float2* __restrict__ stage,
                                                double* __restrict__ fpart, double* __restrict__ dpart)
{
    constexpr int NLM = (L + 1) * (L + 1);
    const int slot = wave * kWave + lane;
    const AtomRec& at = cmem(a.atoms)[slot];
    const double x0 = at.xf[0], x1 = at.xf[1], x2 = at.xf[2];
    float u[6];
#pragma unroll
    for (int j = 0; j < 6; j++) u[j] = (float)at.u[j];
    float P[NLM];
#pragma unroll
    for (int j = 0; j < NLM; j++) P[j] = (float)at.plm[j];
    const float w = (float)at.w, anr = (float)at.anom[0], ani = (float)at.anom[1];
    const int t = at.type;
    TabRec tr{0, 2};
    float pv = 0, ik = 0, ikp = 0;
    if (OTF) {
        tr = cmem(a.tabs)[at.wfn];
        pv = (float)at.pval;
        ik = (float)at.inv_kappa;
        ikp = (float)at.inv_kappa_p;
    }
    const float* tab = cmem(static_cast<const float*>(a.tab));
    LinTabCache<L> rtc;    // the lane's last table intervals (reflections in |h| order)
    const unsigned* hl32 = cmem(reinterpret_cast<const unsigned*>(a.hlen));
    const int* cont = cmem(a.line_cont);
    float Mr[NOPS], Mi[NOPS], Cq[NOPS], Er[NOPS], Ei[NOPS], Q[NOPS];
    {
        const double xf[3] = {x0, x1, x2};
        double u64[6];
#pragma unroll
        for (int j = 0; j < 6; j++) u64[j] = at.u[j];
        line_steps_f32<L, DERIV, OTF, NOPS>(a, xf, u64, Mr, Mi, Cq);
    }
#pragma unroll
    for (int s = 0; s < NOPS; s++) Er[s] = Ei[s] = Q[s] = 0.f;
    double dx[3] = {0, 0, 0}, du[6] = {0, 0, 0, 0, 0, 0}, docc = 0;
    for (int kb = k0; kb < k1; kb += kStage) {
        const int kend = min(kb + kStage, k1);
        float fx[3] = {0, 0, 0}, fu[6] = {0, 0, 0, 0, 0, 0}, focc = 0;
        for (int k = kb; k < kend; ++k) {
            const bool along = ((k - kb) & (kLineRestart - 1)) != 0 && cont[k] != 0;
            if (along) {
#pragma unroll
                for (int s = 0; s < NOPS; s++) {
                    const float qr = Q[s] * Mr[s], qi = Q[s] * Mi[s];
                    const float er = Er[s] * qr - Ei[s] * qi;
                    const float ei = Er[s] * qi + Ei[s] * qr;
                    Er[s] = er;
                    Ei[s] = ei;
                    Q[s] *= Cq[s];
                }
            } else {
                const bool next = k + 1 < kend && ((k + 1 - kb) & (kLineRestart - 1)) != 0 && cont[k + 1] != 0;
#pragma unroll
                for (int s = 0; s < NOPS; s++) {
                    const float* r = hop + ((size_t)k * NOPS + s) * kHklopRec;
                    const unsigned* pr = phs32 + ((size_t)k * NOPS + s) * 8;
                    const double ph =
                        fma(ld_f64(pr, 2), x2, fma(ld_f64(pr, 1), x1, fma(ld_f64(pr, 0), x0, ld_f64(pr, 3))));
                    const float phf = (float)__builtin_amdgcn_fract(ph);
                    const float sn = __builtin_amdgcn_sinf(phf), cs = __builtin_amdgcn_cosf(phf);
                    float e = 0.f;
#pragma unroll
                    for (int j = 0; j < 6; j++) e = fmaf(r[25 + j], u[j], e);
                    const float T = __builtin_amdgcn_exp2f(-e);
                    Er[s] = T * cs;
                    Ei[s] = T * sn;
                    if (next) {
                        const float* rn = hop + ((size_t)(k + 1) * NOPS + s) * kHklopRec;
                        float en = 0.f;
#pragma unroll
                        for (int j = 0; j < 6; j++) en = fmaf(rn[25 + j], u[j], en);
                        Q[s] = __builtin_amdgcn_exp2f(e - en);
                    }
                }
            }
            float S, G0, G1, G2, G3, G4;
            if (OTF) {
                float G[5];
                rtc.eval(tab, tr, (float)ld_f64(hl32, k), pv, ik, ikp, anr, S, G);
                G0 = G[0]; G1 = G[1]; G2 = G[2]; G3 = G[3]; G4 = G[4];
            } else {
                const float* rr = rad + ((size_t)k * a.n_types + t) * kRadRec;
                S = rr[0] + anr; G0 = rr[1]; G1 = rr[2]; G2 = rr[3]; G3 = rr[4]; G4 = rr[5];
            }
            const float cr = (float)ld_f64(dtdf32, 2 * k), ci = (float)ld_f64(dtdf32, 2 * k + 1);
            float sr = 0, si = 0;
#pragma unroll
            for (int s = 0; s < NOPS; s++) {
                const float* r = hop + ((size_t)k * NOPS + s) * kHklopRec;
                float A[5] = {0, 0, 0, 0, 0};
                contract<L, float>(P, r, A);
                float fr = S + G0 * A[0] + G2 * A[2];
                float fi = ani * r[35] + G1 * A[1];
                if (L == 4) {
                    fr += G4 * A[4];
                    fi += G3 * A[3];
                }
                const float Xr = fr * Er[s] - fi * Ei[s], Xi = fr * Ei[s] + fi * Er[s];
                sr += Xr;
                si += Xi;
                if (DERIV) {
                    const float zr = cr * Xr - ci * Xi, zi = cr * Xi + ci * Xr;
#pragma unroll
                    for (int j = 0; j < 6; j++) fu[j] -= r[25 + j] * zr;
#pragma unroll
                    for (int j = 0; j < 3; j++) fx[j] -= r[31 + j] * zi;
                }
            }
            if (DERIV) focc += cr * sr - ci * si;
            stage_f32(stage, k - kb, lane, w * sr, w * si);
        }
        if (DERIV) {
#pragma unroll
            for (int j = 0; j < 3; j++) dx[j] += fx[j];
#pragma unroll
            for (int j = 0; j < 6; j++) du[j] += fu[j];
            docc += focc;
        }
        flush_f32(stage, kend - kb, lane, reinterpret_cast<float*>(fpart) + ((size_t)wave * a.n_hkl + kb) * 2);
    }
    if (DERIV) {
        const double inv_log2e = 0.69314718055994530942;
        double* dp = dpart + ((size_t)chunk_id * a.n_slots + slot) * kNdBasic;
        dp[0] = dx[0]; dp[1] = dx[1]; dp[2] = dx[2];
#pragma unroll
        for (int j = 0; j < 6; j++) dp[3 + j] = du[j] * inv_log2e;
        dp[9] = docc;
    }
}

// One (atom block, reflection chunk) work unit: wave wib of the block runs atom wave
// ablock * kWavesPerBlock + wib over the reflections of chunk chunk_id.
template <bool MIXED, bool DERIV, int MP, bool DIAG, bool OTF, int NOPS = 0, int DPATS = 0>
__device__ __forceinline__ void main_unit(const MainArgs& a, int chunk_id, int ablock, int wib, int lane, void* stage)
{
    static_assert(MIXED || !DIAG, "the parity-folded kernels are mixed precision only");
    static_assert(!(MP && OTF), "the multipole adjoint uses radial records");
    static_assert(NOPS == 0 || MP == 0 || (MIXED && DIAG) || (!MIXED && DPATS > 0),
                  "multipole passes along lines: parity-folded (mixed) and diagonal-record (FP64) kernels only");
    static_assert(!(MIXED && DIAG && NOPS > 0 && DPATS == 0), "a mixed diagonal line kernel needs its sign patterns");
    const int wave = ablock * kWavesPerBlock + wib;
    if (wave >= a.n_waves) return;
    const int k0 = chunk_id * a.chunk;
    const int k1 = min(k0 + a.chunk, a.n_hkl);
    const int cls = __builtin_amdgcn_readfirstlane(cmem(a.wave_l)[wave]);
    const unsigned* dtdf32 = cmem(reinterpret_cast<const unsigned*>(a.dtdf));
    const unsigned* ht32 = cmem(reinterpret_cast<const unsigned*>(a.ht));
    double* fpart = gmem(a.fpart);
    double* dpart = gmem(a.dpart);
    if constexpr (MIXED) {
        const float* hop = cmem(static_cast<const float*>(a.hklop));
        const float* rad = cmem(static_cast<const float*>(a.radial));
        const float* radx = cmem(static_cast<const float*>(a.radx));
        const unsigned* phs = cmem(reinterpret_cast<const unsigned*>(a.phase));
        float2* st = static_cast<float2*>(stage);
        if constexpr (NOPS > 0 && DIAG) {
            constexpr int N = NOPS > 0 ? NOPS : 1;
            constexpr int PS = DPATS & 4095;
            constexpr bool LN = (DPATS & kDiagDirect) == 0;
            if (cls <= 2)
                body_mixed_line_diag<2, DERIV, OTF, N, PS, LN, MP>(a, wave, lane, chunk_id, k0, k1, hop, rad, radx,
                                                                   phs, ht32, dtdf32, st, fpart, dpart);
            else
                body_mixed_line_diag<4, DERIV, OTF, N, PS, LN, MP>(a, wave, lane, chunk_id, k0, k1, hop, rad, radx,
                                                                   phs, ht32, dtdf32, st, fpart, dpart);
        } else if constexpr (NOPS > 0) {
            constexpr int N = NOPS > 0 ? NOPS : 1;
            if (cls <= 2)
                body_mixed_line<2, DERIV, OTF, N>(a, wave, lane, chunk_id, k0, k1, hop, rad, phs, dtdf32, st, fpart,
                                                  dpart);
            else
                body_mixed_line<4, DERIV, OTF, N>(a, wave, lane, chunk_id, k0, k1, hop, rad, phs, dtdf32, st, fpart,
                                                  dpart);
        } else if constexpr (DIAG) {
            if (cls <= 2)
                body_mixed_diag<2, DERIV, MP, OTF>(a, wave, lane, chunk_id, k0, k1, hop, rad, radx, phs, ht32, dtdf32,
                                                   st, fpart, dpart);
            else
                body_mixed_diag<4, DERIV, MP, OTF>(a, wave, lane, chunk_id, k0, k1, hop, rad, radx, phs, ht32, dtdf32,
                                                   st, fpart, dpart);
        } else if (cls <= 2) {
            body_mixed<2, DERIV, MP, OTF>(a, wave, lane, chunk_id, k0, k1, hop, rad, radx, phs, dtdf32, st, fpart, dpart);
        } else {
            body_mixed<4, DERIV, MP, OTF>(a, wave, lane, chunk_id, k0, k1, hop, rad, radx, phs, dtdf32, st, fpart, dpart);
        }
    } else {
        const unsigned* hop = cmem(static_cast<const unsigned*>(a.hklop));
        const double* rad = cmem(static_cast<const double*>(a.radial));
        const double* radx = cmem(static_cast<const double*>(a.radx));
        double2* st = static_cast<double2*>(stage);
        if constexpr (NOPS > 0 && DPATS > 0) {
            constexpr int N = NOPS > 0 ? NOPS : 1;
            constexpr int PS = DPATS & 4095;
            constexpr bool LN = (DPATS & kDiagDirect) == 0;
            if (cls <= 2)
                body_fp64_line_diag<2, DERIV, OTF, N, PS, MP, LN>(a, wave, lane, chunk_id, k0, k1, hop, rad, radx,
                                                                  dtdf32, st, fpart, dpart);
            else
                body_fp64_line_diag<4, DERIV, OTF, N, PS, MP, LN>(a, wave, lane, chunk_id, k0, k1, hop, rad, radx,
                                                                  dtdf32, st, fpart, dpart);
        } else if constexpr (NOPS > 0) {
            constexpr int N = NOPS > 0 ? NOPS : 1;
            if (cls <= 2)
                body_fp64_line<2, DERIV, OTF, N>(a, wave, lane, chunk_id, k0, k1, hop, rad, dtdf32, st, fpart, dpart);
            else
                body_fp64_line<4, DERIV, OTF, N>(a, wave, lane, chunk_id, k0, k1, hop, rad, dtdf32, st, fpart, dpart);
        } else if (cls <= 2) {
            body_fp64<2, DERIV, MP, OTF>(a, wave, lane, chunk_id, k0, k1, hop, rad, radx, dtdf32, st, fpart, dpart);
        } else {
            body_fp64<4, DERIV, MP, OTF>(a, wave, lane, chunk_id, k0, k1, hop, rad, radx, dtdf32, st, fpart, dpart);
        }
    }
}

// One workgroup = kWavesPerBlock atom waves x one reflection chunk, over a 1-D grid of
// kXcds * per_xcd workgroups: hardware id b runs on XCD b % kXcds and is mapped to logical block
// (b % kXcds) * per_xcd + b / kXcds, so each XCD walks a contiguous range of chunks and every
// record of a chunk is fetched into one L2 only.
// MP: 0 no multipole adjoint; 1 and 2 its two passes over the same work units (1: F, the
// xyz/U/occupancy derivatives and the Pval, kappa, kappa' sums; 2: the 25 global-frame B_lm sums),
// split so that neither pass holds both the atom's P_lm and the B_lm accumulators; 3 replaces 1
// when only the B_lm are wanted (the chain rule through the local frames): the plain pass in the
// multipole row layout.  Waves per SIMD the register allocation must allow: mixed plain passes
// four (five for the general-rotation kernel); mixed multipole passes along lines four, except
// pass 1 of a four-operation set (it would spill 39 VGPRs at four); FP64 pass 1 along lines two,
// FP64 pass 2 three; the others uncapped.
template <bool MIXED, bool DERIV, int MP, bool DIAG, bool OTF, int NOPS = 0, int DPATS = 0>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(MIXED && (MP == 0 || MP == 3) ? (DIAG || OTF || NOPS > 0 ? 4 : 5) : MIXED && NOPS > 0 && (MP == 2 || NOPS == 2) ? 4 : (!MIXED && MP == 2 ? 3 : (!MIXED && MP == 1 && NOPS > 0 ? 2 : 1))))) void k_main(MainArgs a)
{
    constexpr int kStageBytes = MIXED ? kStageBytesF32 : kStageBytesF64;
    __shared__ __align__(16) unsigned char stage_lds[kWavesPerBlock * kStageBytes];
    const int b = blockIdx.x;
    const int logical = (b % kXcds) * a.per_xcd + b / kXcds;
    if (logical >= a.n_blocks) return;
    const int chunk_id = logical / a.grid_x;
    const int wib = (int)(threadIdx.x / kWave);
    main_unit<MIXED, DERIV, MP, DIAG, OTF, NOPS, DPATS>(a, chunk_id, logical - chunk_id * a.grid_x, wib,
                                                 threadIdx.x & (kWave - 1), stage_lds + wib * kStageBytes);
}

// ------------------------------------------------------------------ per-reflection derivatives
//
// The per-hkl path (HansenCoppensStructureFactorCalculator.h:384-545 over
// HansenCoppens_SF_Engine.h:512-780): for every reflection k of a block and every atom a the
// kernel writes dF_k/dp_a for the atom's 10 parameters (xyz, 6 ADP, occupancy) as complex FP64
// values, already scaled and converted to the caller's conventions, and F_k through the same
// per-wave staging and stage-2 reduction as the summed path.  The output is ~160 bytes per
// (atom, reflection) against ~130 FP64 flops per (atom, operation, reflection) term, so the
// kernel sits near the HBM-write roof; the mixed-precision variant (FP32 terms, FP64 phase
// reduction and rows) takes the term work off that path.
//
// Work unit = (atom wave, reflection chunk of the block); each of the four waves of a workgroup
// takes its own unit, atom wave fastest, so the waves of a workgroup share reflection records
// and write neighbouring atoms of the same rows.

struct JacAtom {
    double mult_w;     // m / |G|: occupancy-derivative weight
    int adp_kind;      // 0 (no ADP), 1 (U_iso), 6
    int pad;
};

struct JacArgs {
    double* rows;          // block rows: [k - k0][plane][n_atoms] complex (re, im)
    double* f;             // block F: [k - k0] complex
    const JacAtom* atoms;  // caller atom order
    const double* jrec;    // per (reflection, operation): kJacRec doubles, see k_jac_rec
    int n_atoms, k0, k1, chunk, n_units;
};

static constexpr int kJacPlanes = 10;
static constexpr int kJacRec = 12;

// one complex entry of a per-reflection row; the rows are written once and not read back by
// this kernel, so the stores are marked non-temporal (streamed past the caches' retention)
__device__ __forceinline__ void put_row(double* p, double re, double im)
{
    typedef double d2v __attribute__((ext_vector_type(2)));
    const d2v v = {re, im};
    __builtin_nontemporal_store(v, reinterpret_cast<d2v*>(p));
}

// Per (reflection, operation): the multipliers of the xyz and ADP derivatives already carried into
// the caller's conventions, so that the main kernel needs no conversion:
//   [0..2] xyz_m . h_s (fractional h_s; xyz_m = identity for fractional derivatives, c2f^T for
//          Cartesian ones),  [3..8] adp_m . m(h_s) (m = Debye-Waller multipliers of U_cart;
//          adp_m = identity for U_cart derivatives, the U_cart -> U_cif chain rule otherwise).
__global__ void k_jac_rec(const double* __restrict__ rec64, int n, const double* __restrict__ conv,
                          double* __restrict__ jrec)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double* r = rec64 + (size_t)i * kHklopRec;
    double* o = jrec + (size_t)i * kJacRec;
    for (int a = 0; a < 3; a++) o[a] = conv[3 * a] * r[31] + conv[3 * a + 1] * r[32] + conv[3 * a + 2] * r[33];
    for (int a = 0; a < 6; a++) {
        double v = 0.0;
        for (int b = 0; b < 6; b++) v += conv[9 + 6 * a + b] * r[25 + b];
        o[3 + a] = v;
    }
    o[9] = o[10] = o[11] = 0.0;
}

// The same from the mixed-precision records (Debye-Waller multipliers carry log2(e)).
__global__ void k_jac_rec32(const float* __restrict__ rec32, int n, const double* __restrict__ conv,
                            double* __restrict__ jrec)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* r = rec32 + (size_t)i * kHklopRec;
    double* o = jrec + (size_t)i * kJacRec;
    const double inv_log2e = 0.69314718055994530942;
    for (int a = 0; a < 3; a++)
        o[a] = conv[3 * a] * (double)r[31] + conv[3 * a + 1] * (double)r[32] + conv[3 * a + 2] * (double)r[33];
    for (int a = 0; a < 6; a++) {
        double v = 0.0;
        for (int b = 0; b < 6; b++) v += conv[9 + 6 * a + b] * ((double)r[25 + b] * inv_log2e);
        o[3 + a] = v;
    }
    o[9] = o[10] = o[11] = 0.0;
}

// CENTRIC: centre of symmetry at the origin with only the coset representatives explicit; the
// inversion partner is folded in per operation (f'' kept unconjugated, as the per-hkl engine does):
//   F part   X = 2 T [(f0 + f') cos phi + Re(f_def e^{i phi})] + 2 i T f'' cos phi
//   xyz part Y = -2 T f'' sin phi + 2 i T [(f0 + f') sin phi + Im(f_def e^{i phi})]
// Otherwise every explicit operation (inversion partners of a shifted centre included) adds
// X = Y = f T e^{i phi} with f = f0 + f' + f_def + i f''.
template <int L, bool CENTRIC, bool OTF>
__device__ __forceinline__ void body_jac(const MainArgs& a, const JacArgs& jb, int wave, int lane, int k0, int k1,
                                         const unsigned* __restrict__ hop, const unsigned* __restrict__ jrec,
                                         const double* __restrict__ rad, const unsigned* __restrict__ mult32,
                                         double2* __restrict__ stage, double* __restrict__ fpart,
                                         double* __restrict__ rows)
{
    constexpr int NLM = (L + 1) * (L + 1);
    const int slot = wave * kWave + lane;
    const AtomRec& at = cmem(a.atoms)[slot];
    const double x0 = at.xf[0], x1 = at.xf[1], x2 = at.xf[2];
    double u[6];
#pragma unroll
    for (int j = 0; j < 6; j++) u[j] = at.u[j];
    double P[NLM];
#pragma unroll
    for (int j = 0; j < NLM; j++) P[j] = at.plm[j];
    const double w = at.w, anr = at.anom[0], ani = at.anom[1];
    const int t = at.type;
    const int orig = at.orig;
    double mw = 0.0;
    int kind = 0;
    if (orig >= 0) {
        const JacAtom ja = gmem(jb.atoms)[orig];
        mw = ja.mult_w;
        kind = ja.adp_kind;
    }
    TabRec tr{0, 2};
    double pv = 0, ik = 0, ikp = 0;
    if (OTF) {
        tr = cmem(a.tabs)[at.wfn];
        pv = at.pval;
        ik = at.inv_kappa;
        ikp = at.inv_kappa_p;
    }
    const double* tab = cmem(static_cast<const double*>(a.tab));
    const unsigned* hl32 = cmem(reinterpret_cast<const unsigned*>(a.hlen));
    const size_t plane = (size_t)jb.n_atoms * 2;
    for (int kb = k0; kb < k1; kb += kStage) {
    const int nk = min(kStage, k1 - kb);
    for (int kk = 0; kk < nk; ++kk) {
        const int k = kb + kk;
        const double hl = ld_f64(hl32, k);
        double S, G0, G1, G2, G3, G4;
        if (OTF) {
            double G[5];
            radial_otf<L, double>(tab, tr, hl, pv, ik, ikp, anr, S, G);
            G0 = G[0]; G1 = G[1]; G2 = G[2]; G3 = G[3]; G4 = G[4];
        } else {
            const double* rr = rad + ((size_t)k * a.n_types + t) * kRadRec;
            S = rr[0] + anr; G0 = rr[1]; G1 = rr[2]; G2 = rr[3]; G3 = rr[4]; G4 = rr[5];
        }
        double s0r = 0, s0i = 0, sxr[3] = {0, 0, 0}, sxi[3] = {0, 0, 0};
        double sur[6] = {0, 0, 0, 0, 0, 0}, sui[6] = {0, 0, 0, 0, 0, 0};
        for (int s = 0; s < a.n_ops; ++s) {
            const unsigned* rw = hop + ((size_t)k * a.n_ops + s) * (2 * kHklopRec);
            double r[kHklopRec];
#pragma unroll
            for (int j = 0; j < kHklopRec; j++) r[j] = ld_f64(rw, j);
            const unsigned* qw = jrec + ((size_t)k * a.n_ops + s) * (2 * kJacRec);
            double q[9];
#pragma unroll
            for (int j = 0; j < 9; j++) q[j] = ld_f64(qw, j);
            double A[5];
            contract<L, double>(P, r, A);
            double fr = S + G0 * A[0] + G2 * A[2];
            double fi = ani + G1 * A[1];
            if (L == 4) {
                fr += G4 * A[4];
                fi += G3 * A[3];
            }
            const double ph = __builtin_amdgcn_fract(fma(r[33], x2, fma(r[32], x1, fma(r[31], x0, r[34]))));
            double sn, cs;
            sincos_rev(ph, &sn, &cs);
            const double e = r[25] * u[0] + r[26] * u[1] + r[27] * u[2] + r[28] * u[3] + r[29] * u[4] + r[30] * u[5];
            const double T = exp_neg(e);
            const double Er = T * cs, Ei = T * sn;
            const double Xr = fr * Er - fi * Ei, Xi = fr * Ei + fi * Er;
            double gr, gi, yr, yi;
            if (CENTRIC) {
                gr = 2.0 * (Xr + ani * Ei);
                gi = 2.0 * ani * Er;
                yr = -2.0 * ani * Ei;
                yi = 2.0 * (Xi - ani * Er);
            } else {
                gr = Xr;
                gi = Xi;
                yr = Xr;
                yi = Xi;
            }
            s0r += gr;
            s0i += gi;
#pragma unroll
            for (int j = 0; j < 3; j++) {
                sxr[j] += q[j] * yr;
                sxi[j] += q[j] * yi;
            }
#pragma unroll
            for (int j = 0; j < 6; j++) {
                sur[j] -= q[3 + j] * gr;
                sui[j] -= q[3 + j] * gi;
            }
        }
        stage_f64(stage, kk, lane, w * s0r, w * s0i);
        if (orig >= 0) {
            const double c = ld_f64(mult32, k);
            const double cx = 6.283185307179586477 * w * c, cu = 19.739208802178717238 * w * c, co = mw * c;
            double* o = rows + ((size_t)(k - jb.k0) * kJacPlanes * jb.n_atoms + orig) * 2;
            // xyz: 2 pi i w c sum h'_s Y_s
#pragma unroll
            for (int i = 0; i < 3; i++) put_row(o + i * plane, -cx * sxi[i], cx * sxr[i]);
            // ADP: 2 pi^2 w c sum (-m'(h_s)) X_s; U_iso: 2 pi^2 w c (-|h|^2) sum X_s
            double vr[6], vi[6];
#pragma unroll
            for (int i = 0; i < 6; i++) {
                vr[i] = kind == 6 ? cu * sur[i] : 0.0;
                vi[i] = kind == 6 ? cu * sui[i] : 0.0;
            }
            if (kind == 1) {
                vr[0] = -cu * hl * hl * s0r;
                vi[0] = -cu * hl * hl * s0i;
            }
#pragma unroll
            for (int i = 0; i < 6; i++) put_row(o + (3 + i) * plane, vr[i], vi[i]);
            put_row(o + 9 * plane, co * s0r, co * s0i);
        }
    }
    flush_f64(stage, nk, lane, fpart + ((size_t)wave * a.n_hkl + kb) * 2);
    }
}

// Mixed precision (HIP_MIXED): the terms as in body_mixed (FP64 phase reduction, FP32 harmonics,
// radial factors, Debye-Waller, sincos and the sums over operations), the rows' scale factors and
// the stored rows in FP64.
template <int L, bool CENTRIC, bool OTF>
__device__ __forceinline__ void body_jac_mixed(const MainArgs& a, const JacArgs& jb, int wave, int lane, int k0,
                                               int k1, const float* __restrict__ hop,
                                               const unsigned* __restrict__ phs32, const unsigned* __restrict__ jrec,
                                               const float* __restrict__ rad, const unsigned* __restrict__ mult32,
                                               float2* __restrict__ stage, double* __restrict__ fpart,
                                               double* __restrict__ rows)
{
    constexpr int NLM = (L + 1) * (L + 1);
    const int slot = wave * kWave + lane;
    const AtomRec& at = cmem(a.atoms)[slot];
    const double x0 = at.xf[0], x1 = at.xf[1], x2 = at.xf[2];
    float u[6];
#pragma unroll
    for (int j = 0; j < 6; j++) u[j] = (float)at.u[j];
    float P[NLM];
#pragma unroll
    for (int j = 0; j < NLM; j++) P[j] = (float)at.plm[j];
    const double w = at.w;
    const float anr = (float)at.anom[0], ani = (float)at.anom[1];
    const int t = at.type;
    const int orig = at.orig;
    double mw = 0.0;
    int kind = 0;
    if (orig >= 0) {
        const JacAtom ja = gmem(jb.atoms)[orig];
        mw = ja.mult_w;
        kind = ja.adp_kind;
    }
    TabRec tr{0, 2};
    float pv = 0, ik = 0, ikp = 0;
    if (OTF) {
        tr = cmem(a.tabs)[at.wfn];
        pv = (float)at.pval;
        ik = (float)at.inv_kappa;
        ikp = (float)at.inv_kappa_p;
    }
    const float* tab = cmem(static_cast<const float*>(a.tab));
    const unsigned* hl32 = cmem(reinterpret_cast<const unsigned*>(a.hlen));
    const size_t plane = (size_t)jb.n_atoms * 2;
    for (int kb = k0; kb < k1; kb += kStage) {
        const int nk = min(kStage, k1 - kb);
        for (int kk = 0; kk < nk; ++kk) {
            const int k = kb + kk;
            const double hl = ld_f64(hl32, k);
            float S, G0, G1, G2, G3, G4;
            if (OTF) {
                float G[5];
                radial_otf<L, float>(tab, tr, (float)hl, pv, ik, ikp, anr, S, G);
                G0 = G[0]; G1 = G[1]; G2 = G[2]; G3 = G[3]; G4 = G[4];
            } else {
                const float* rr = rad + ((size_t)k * a.n_types + t) * kRadRec;
                S = rr[0] + anr; G0 = rr[1]; G1 = rr[2]; G2 = rr[3]; G3 = rr[4]; G4 = rr[5];
            }
            float s0r = 0, s0i = 0, sxr[3] = {0, 0, 0}, sxi[3] = {0, 0, 0};
            float sur[6] = {0, 0, 0, 0, 0, 0}, sui[6] = {0, 0, 0, 0, 0, 0};
            for (int s = 0; s < a.n_ops; ++s) {
                const float* r = hop + ((size_t)k * a.n_ops + s) * kHklopRec;
                const unsigned* pr = phs32 + ((size_t)k * a.n_ops + s) * 8;
                const unsigned* qw = jrec + ((size_t)k * a.n_ops + s) * (2 * kJacRec);
                float q[9];
#pragma unroll
                for (int j = 0; j < 9; j++) q[j] = (float)ld_f64(qw, j);
                float A[5];
                contract<L, float>(P, r, A);
                float fr = S + G0 * A[0] + G2 * A[2];
                float fi = ani + G1 * A[1];
                if (L == 4) {
                    fr += G4 * A[4];
                    fi += G3 * A[3];
                }
                // FP64 range reduction of h.(R x + t) to [0, 1) revolutions
                const double ph = fma(ld_f64(pr, 2), x2, fma(ld_f64(pr, 1), x1, fma(ld_f64(pr, 0), x0, ld_f64(pr, 3))));
                const float phf = (float)__builtin_amdgcn_fract(ph);
                const float sn = __builtin_amdgcn_sinf(phf), cs = __builtin_amdgcn_cosf(phf);
                const float e = r[25] * u[0] + r[26] * u[1] + r[27] * u[2] + r[28] * u[3] + r[29] * u[4] + r[30] * u[5];
                const float T = __builtin_amdgcn_exp2f(-e);
                const float Er = T * cs, Ei = T * sn;
                const float Xr = fr * Er - fi * Ei, Xi = fr * Ei + fi * Er;
                float gr, gi, yr, yi;
                if (CENTRIC) {
                    gr = 2.0f * (Xr + ani * Ei);
                    gi = 2.0f * ani * Er;
                    yr = -2.0f * ani * Ei;
                    yi = 2.0f * (Xi - ani * Er);
                } else {
                    gr = Xr;
                    gi = Xi;
                    yr = Xr;
                    yi = Xi;
                }
                s0r += gr;
                s0i += gi;
#pragma unroll
                for (int j = 0; j < 3; j++) {
                    sxr[j] += q[j] * yr;
                    sxi[j] += q[j] * yi;
                }
#pragma unroll
                for (int j = 0; j < 6; j++) {
                    sur[j] -= q[3 + j] * gr;
                    sui[j] -= q[3 + j] * gi;
                }
            }
            stage_f32(stage, kk, lane, (float)w * s0r, (float)w * s0i);
            if (orig >= 0) {
                const double c = ld_f64(mult32, k);
                const double cx = 6.283185307179586477 * w * c, cu = 19.739208802178717238 * w * c, co = mw * c;
                double* o = rows + ((size_t)(k - jb.k0) * kJacPlanes * jb.n_atoms + orig) * 2;
#pragma unroll
                for (int i = 0; i < 3; i++) put_row(o + i * plane, -cx * sxi[i], cx * sxr[i]);
                double vr[6], vi[6];
#pragma unroll
                for (int i = 0; i < 6; i++) {
                    vr[i] = kind == 6 ? cu * sur[i] : 0.0;
                    vi[i] = kind == 6 ? cu * sui[i] : 0.0;
                }
                if (kind == 1) {
                    vr[0] = -cu * hl * hl * s0r;
                    vi[0] = -cu * hl * hl * s0i;
                }
#pragma unroll
                for (int i = 0; i < 6; i++) put_row(o + (3 + i) * plane, vr[i], vi[i]);
                put_row(o + 9 * plane, co * s0r, co * s0i);
            }
        }
        flush_f32(stage, nk, lane, reinterpret_cast<float*>(fpart) + ((size_t)wave * a.n_hkl + kb) * 2);
    }
}

template <bool CENTRIC, bool OTF, bool MIXED>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(MIXED ? 4 : 2))) void k_jac(MainArgs a, JacArgs jb)
{
    __shared__ __align__(16) unsigned char stage_lds[kWavesPerBlock * (MIXED ? kStageBytesF32 : kStageBytesF64)];
    // wave-uniform (the compiler cannot see that threadIdx.x / 64 is), so that the reflection
    // range and every record address derived from it stay in scalar registers
    const int wib = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / kWave));
    const int unit = blockIdx.x * kWavesPerBlock + wib;
    if (unit >= jb.n_units) return;
    const int wave = unit % a.n_waves;
    const int chunk_id = unit / a.n_waves;
    const int k0 = jb.k0 + chunk_id * jb.chunk;
    const int k1 = min(k0 + jb.chunk, jb.k1);
    const int lane = threadIdx.x & (kWave - 1);
    const int cls = __builtin_amdgcn_readfirstlane(cmem(a.wave_l)[wave]);
    const unsigned* jrec = cmem(reinterpret_cast<const unsigned*>(jb.jrec));
    const unsigned* mult32 = cmem(reinterpret_cast<const unsigned*>(a.dtdf + 2 * (size_t)a.n_hkl));
    if (MIXED) {
        const float* hop = cmem(static_cast<const float*>(a.hklop));
        const unsigned* phs = cmem(reinterpret_cast<const unsigned*>(a.phase));
        const float* rad = cmem(static_cast<const float*>(a.radial));
        float2* st = reinterpret_cast<float2*>(stage_lds + wib * kStageBytesF32);
        if (cls <= 2)
            body_jac_mixed<2, CENTRIC, OTF>(a, jb, wave, lane, k0, k1, hop, phs, jrec, rad, mult32, st,
                                            gmem(a.fpart), gmem(jb.rows));
        else
            body_jac_mixed<4, CENTRIC, OTF>(a, jb, wave, lane, k0, k1, hop, phs, jrec, rad, mult32, st,
                                            gmem(a.fpart), gmem(jb.rows));
        return;
    }
    const unsigned* hop = cmem(static_cast<const unsigned*>(a.hklop));
    const double* rad = cmem(static_cast<const double*>(a.radial));
    double2* st = reinterpret_cast<double2*>(stage_lds + wib * kStageBytesF64);
    if (cls <= 2)
        body_jac<2, CENTRIC, OTF>(a, jb, wave, lane, k0, k1, hop, jrec, rad, mult32, st, gmem(a.fpart),
                                  gmem(jb.rows));
    else
        body_jac<4, CENTRIC, OTF>(a, jb, wave, lane, k0, k1, hop, jrec, rad, mult32, st, gmem(a.fpart),
                                  gmem(jb.rows));
}

// F of a block of reflections [k0, k0 + n): mult(h) * sum_waves partial, into f[k - k0].
template <typename P2>
__global__ void k_reduce_f_rows(const P2* __restrict__ fpart, const double* __restrict__ mult, int n_waves,
                                int n_hkl, int k0, int n, double* __restrict__ f)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int k = k0 + i;
    double re = 0, im = 0;
    for (int w = 0; w < n_waves; w++) {
        const P2 v = fpart[(size_t)w * n_hkl + k];
        re += v.x;
        im += v.y;
    }
    f[2 * i] = re * mult[k];
    f[2 * i + 1] = im * mult[k];
}

// ------------------------------------------------------------------ form-factor route
//
// Atomic form factors f_a(h) = f_core(|h|) + Pval f_val(|h| / kappa) + f_def(h) for a list of
// Cartesian h (HansenCoppensStructureFactorCalculator::calculateFormFactors(Cart),
// .cpp:1274-1333, over HansenCoppens_SF_Engine::calculateFormFactors, HansenCoppens_SF_Engine.cpp:
// 383-495).  The spherical terms come, as in the reference, from cubic splines on a 0.001 A^-1
// grid of 10 000 nodes (setDataForCpuEngineCallsForSingleHkl, .cpp:1244-1272); f_def uses the
// closed-form g-functions.  No anomalous term.

static constexpr double kSplineStep = 0.001;
static constexpr int kSplineNodes = 10000;

// Grid values: spline s < n_wfn is the core density of wavefunction s, spline n_wfn + t the
// valence term Pval_t f_val(h / kappa_t) of type t, both at h = i * step.
__global__ void k_spline_values(const TypeRec* __restrict__ types, const WfnRec* __restrict__ wfns,
                                const TermRec* __restrict__ core, const TermRec* __restrict__ val, int n_wfn,
                                int n_types, double* __restrict__ out)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (n_wfn + n_types) * kSplineNodes) return;
    const int sp = idx / kSplineNodes, i = idx - sp * kSplineNodes;
    const double h = i * kSplineStep;
    double v = 0.0;
    if (sp < n_wfn) {
        const WfnRec w = wfns[sp];
        for (int q = 0; q < w.core_n; q++) {
            const TermRec tc = core[w.core_off + q];
            v += tc.c * g_function(0, tc.p + 2, h, tc.e);
        }
        v *= 12.566370614359172953850;
    } else {
        const TypeRec& ty = types[sp - n_wfn];
        const WfnRec w = wfns[ty.wfn];
        const double hk = h / ty.kappa;
        for (int q = 0; q < w.val_n; q++) {
            const TermRec tv = val[w.val_off + q];
            v += tv.c * g_function(0, tv.p + 2, hk, tv.e);
        }
        v *= 12.566370614359172953850 * ty.pval;
    }
    out[idx] = v;
}

// NaturalCubicSpline::evaluate (NaturalCubicSpline.cpp): zero outside the grid and on its last
// node; c holds 4 coefficients per segment.
__device__ __forceinline__ double spline_at(const double* __restrict__ c, double x)
{
    const double u = x / kSplineStep;
    if (u < 0.0) return 0.0;
    const int i = (int)u;
    if (i >= kSplineNodes) return 0.0;
    const double t = (x - i * kSplineStep) / kSplineStep;
    const double* q = c + 4 * (size_t)i;
    return q[0] + t * q[1] + t * t * q[2] + t * t * t * q[3];
}

// Per (h, type): [S = core + valence spline, G_0..G_4 (4 pi i^l-sign N_l g_l(n_l + 2, |h| / kappa',
// zeta))]; per h: the 25 harmonics of h/|h| (zero for |h| < 1e-10).
__global__ void k_ff_prep(const double* __restrict__ hc, int n_h, const TypeRec* __restrict__ types, int n_types,
                          int n_wfn, const double* __restrict__ spl, double* __restrict__ rec,
                          double* __restrict__ hrec)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n_h * n_types) return;
    const int k = idx / n_types, t = idx - k * n_types;
    const double h0 = hc[3 * k], h1 = hc[3 * k + 1], h2 = hc[3 * k + 2];
    const double hl = sqrt(h0 * h0 + h1 * h1 + h2 * h2);
    if (t == 0) {
        double y[25];
        if (hl < 1e-10) {
            for (int i = 0; i < 25; i++) y[i] = 0.0;
        } else {
            ylm_density<double>(h0 / hl, h1 / hl, h2 / hl, y);
        }
        for (int i = 0; i < 25; i++) hrec[(size_t)k * 26 + i] = y[i];
        hrec[(size_t)k * 26 + 25] = hl;
    }
    const TypeRec& ty = types[t];
    double r[kRadRec] = {0, 0, 0, 0, 0, 0, 0, 0};
    r[0] = spline_at(spl + (size_t)ty.wfn * kSplineNodes * 4, hl) +
           spline_at(spl + (size_t)(n_wfn + t) * kSplineNodes * 4, hl);
    const double hp = hl / ty.kappa_p;
    for (int l = 0; l <= ty.lp; l++) r[1 + l] = ty.gscale[l] * g_function(l, ty.n[l] + 2, hp, ty.zeta);
    for (int i = 0; i < kRadRec; i++) rec[((size_t)k * n_types + t) * kRadRec + i] = r[i];
}

// lane = atom (caller order), work unit = (atom wave, chunk of h); out[(k - k0) * n_atoms + atom]
__global__ __launch_bounds__(256) void k_ff(const AtomRec* __restrict__ atoms, int n_waves, int n_atoms,
                                            const double* __restrict__ hrec, const double* __restrict__ rec,
                                            int n_types, int k0, int k1, int chunk, int n_units,
                                            double* __restrict__ out)
{
    const int wib = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / kWave));
    const int unit = blockIdx.x * kWavesPerBlock + wib;
    if (unit >= n_units) return;
    const int wave = unit % n_waves;
    const int ka = k0 + (unit / n_waves) * chunk, kb = min(ka + chunk, k1);
    const int slot = wave * kWave + (int)(threadIdx.x & (kWave - 1));
    const AtomRec& at = atoms[slot];
    double P[25];
#pragma unroll
    for (int j = 0; j < 25; j++) P[j] = at.plm[j];
    const int t = at.type, orig = at.orig;
    const unsigned* h32 = cmem(reinterpret_cast<const unsigned*>(hrec));
    for (int k = ka; k < kb; k++) {
        double y[25];
#pragma unroll
        for (int j = 0; j < 25; j++) y[j] = ld_f64(h32 + (size_t)k * 52, j);
        double A[5];
        contract<4, double>(P, y, A);
        const double* rr = rec + ((size_t)k * n_types + t) * kRadRec;
        const double fr = rr[0] + rr[1] * A[0] + rr[3] * A[2] + rr[5] * A[4];
        const double fi = rr[2] * A[1] + rr[4] * A[3];
        if (orig >= 0) put_row(out + ((size_t)(k - k0) * n_atoms + orig) * 2, fr, fi);
    }
}

// F(h) = mult(h) * sum_waves partial  (centric at origin: 2 Re, 0).  Loads are issued eight at a
// time and added in wave order.
template <typename P2>
__device__ __forceinline__ void reduce_f_item(int k, const P2* __restrict__ fpart,
                                              const double* __restrict__ mult, int n_waves, int n_hkl,
                                              int centric_origin, double* __restrict__ f)
{
    double re = 0, im = 0;
    int w = 0;
    for (; w + 8 <= n_waves; w += 8) {
        P2 v[8];
#pragma unroll
        for (int q = 0; q < 8; q++) v[q] = fpart[(size_t)(w + q) * n_hkl + k];
#pragma unroll
        for (int q = 0; q < 8; q++) {
            re += v[q].x;
            im += v[q].y;
        }
    }
    for (; w < n_waves; w++) {
        const P2 v = fpart[(size_t)w * n_hkl + k];
        re += v.x;
        im += v.y;
    }
    if (centric_origin) {
        re *= 2;
        im = 0;
    }
    f[2 * k] = re * mult[k];
    f[2 * k + 1] = im * mult[k];
}

template <typename P2>
__global__ void k_reduce_f(const P2* __restrict__ fpart, const double* __restrict__ mult, int n_waves,
                           int n_hkl, int centric_origin, double* __restrict__ f)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n_hkl) reduce_f_item(k, fpart, mult, n_waves, n_hkl, centric_origin, f);
}

// d[orig][j] = sum over reflection chunks of the per-(chunk, slot) partials; one thread per
// (slot, j), chunks added in order, loads issued eight at a time.
__device__ __forceinline__ void reduce_d_item(int idx, const double* __restrict__ dpart,
                                              const AtomRec* __restrict__ atoms, int n_chunks, int n_slots, int nd,
                                              double* __restrict__ d)
{
    const int i = idx / nd;
    const int orig = atoms[i].orig;
    if (orig < 0) return;
    const size_t stride = (size_t)n_slots * nd;
    const double* p = dpart + idx;
    double s = 0.0;
    int c = 0;
    for (; c + 8 <= n_chunks; c += 8) {
        double v[8];
#pragma unroll
        for (int q = 0; q < 8; q++) v[q] = p[(size_t)(c + q) * stride];
#pragma unroll
        for (int q = 0; q < 8; q++) s += v[q];
    }
    for (; c < n_chunks; c++) s += p[(size_t)c * stride];
    d[(size_t)orig * nd + (idx - i * nd)] = s;
}

__global__ void k_reduce_d(const double* __restrict__ dpart, const AtomRec* __restrict__ atoms, int n_chunks,
                           int n_slots, int nd, double* __restrict__ d)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < n_slots * nd) reduce_d_item(idx, dpart, atoms, n_chunks, n_slots, nd, d);
}

// dT/dP_lm in the atom's local frame from the global-frame derivative: P_glob = D P_loc with
// D_ij = sum_p w_p d_i(u_p) d_j(L^T u_p) / int d_i^2, hence dT/dP_loc = D^T dT/dP_glob.
__global__ __launch_bounds__(64) void k_plm_back(const AtomIn* __restrict__ in, int n_slots, int nd,
                                                 double* __restrict__ d)
{
    const int i = blockIdx.x;
    const int p = threadIdx.x;
    __shared__ double v[25];
    __shared__ double sp[64];
    const AtomIn& a = in[i];
    if (a.orig < 0) return;               // whole block exits together
    double* row = d + (size_t)a.orig * nd + 10;
    if (p < 25) v[p] = row[p];
    __syncthreads();
    double s = 0.0;
    if (p < 50)
        for (int q = 0; q < 25; q++) s += c_proj_w[p][q] * v[q];
    sp[p] = s;
    __syncthreads();
    if (p < 25) {
        double acc = 0.0;
        for (int q = 0; q < 50; q++) {
            const double u0 = c_proj_dir[q][0], u1 = c_proj_dir[q][1], u2 = c_proj_dir[q][2];
            const double x = a.lcs[0] * u0 + a.lcs[3] * u1 + a.lcs[6] * u2;
            const double y = a.lcs[1] * u0 + a.lcs[4] * u1 + a.lcs[7] * u2;
            const double z = a.lcs[2] * u0 + a.lcs[5] * u1 + a.lcs[8] * u2;
            double dl[25];
            ylm_density<double>(x, y, z, dl);
            double dj = 0.0;
#pragma unroll
            for (int j = 0; j < 25; j++) dj = (j == p) ? dl[j] : dj;
            acc += dj * sp[q];
        }
        row[p] = acc;
    }
}

// dT/dL for every atom's local frame L (columns = local axes), from the global-frame multipole
// adjoint B = dT/dP_glob (weight excluded) before k_plm_back maps it to the local frame:
// P_glob,i = sum_q W_qi f(L^T u_q) with f(v) = sum_j P_loc,j d_j(v) (the projection of
// atom_prep_item), hence dT/dL_ab = sum_q s_q u_q,a (grad f)_b(L^T u_q), s_q = sum_{i, l <= max_l}
// W_qi B_i.  One 64-lane workgroup per atom slot, lane q = quadrature node; lane sums in order.
__global__ __launch_bounds__(64) void k_lcs_grad(const AtomIn* __restrict__ in, const TypeRec* __restrict__ types,
                                                 int nd, const double* __restrict__ d, double* __restrict__ g)
{
    const int i = blockIdx.x;
    const int p = threadIdx.x;
    __shared__ double part[9][64];
    const AtomIn& a = in[i];
    if (a.orig < 0) return;               // whole block exits together
    const TypeRec& ty = types[a.type];
    double G[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (p < 50 && ty.max_l >= 0) {
        const double* B = d + (size_t)a.orig * nd + 10;
        double s = 0.0;
        for (int j = 0; j < 25; j++) {
            const int l = j < 1 ? 0 : j < 4 ? 1 : j < 9 ? 2 : j < 16 ? 3 : 4;
            if (l <= ty.max_l) s += c_proj_w[p][j] * B[j];
        }
        const double u0 = c_proj_dir[p][0], u1 = c_proj_dir[p][1], u2 = c_proj_dir[p][2];
        const double x = a.lcs[0] * u0 + a.lcs[3] * u1 + a.lcs[6] * u2;
        const double y = a.lcs[1] * u0 + a.lcs[4] * u1 + a.lcs[7] * u2;
        const double z = a.lcs[2] * u0 + a.lcs[5] * u1 + a.lcs[8] * u2;
        double gy[25][3];
        ylm_density_grad(x, y, z, gy);
        double gf[3] = {0, 0, 0};
        for (int j = 0; j < 25; j++) {
            const int l = j < 1 ? 0 : j < 4 ? 1 : j < 9 ? 2 : j < 16 ? 3 : 4;
            if (l <= ty.max_l)
                for (int c = 0; c < 3; c++) gf[c] += ty.plm[j] * gy[j][c];
        }
        const double u[3] = {u0, u1, u2};
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 3; c++) G[3 * r + c] = s * u[r] * gf[c];
    }
    for (int k = 0; k < 9; k++) part[k][p] = G[k];
    __syncthreads();
    if (p < 9) {
        double acc = 0.0;
        for (int q = 0; q < 50; q++) acc += part[p][q];
        g[(size_t)a.orig * 9 + p] = acc;
    }
}

// ------------------------------------------------------------------ host side

static bool g_constants_ready[64] = {false};

static void legendre_nodes5(double* x, double* w)
{
    const int n = 5;
    for (int i = 0; i < n; i++) {
        double z = std::cos(3.14159265358979323846 * (i + 0.75) / (n + 0.5));
        for (int it = 0; it < 100; it++) {
            double p1 = 1.0, p2 = 0.0;
            for (int j = 0; j < n; j++) {
                double p3 = p2;
                p2 = p1;
                p1 = ((2.0 * j + 1.0) * z * p2 - j * p3) / (j + 1);
            }
            double pp = n * (z * p1 - p2) / (z * z - 1.0);
            double z1 = z;
            z = z1 - p1 / pp;
            if (std::fabs(z - z1) < 1e-16) {
                w[i] = 2.0 / ((1.0 - z * z) * pp * pp);
                break;
            }
            w[i] = 2.0 / ((1.0 - z * z) * pp * pp);
        }
        x[i] = z;
    }
}

static void host_ylm(const double* u, double* d)
{
    double x = u[0], y = u[1], z = u[2], xx = x * x, yy = y * y, zz = z * z;
    double v[25] = {1.0, y, z, x, x * y, y * z, 2 * zz - xx - yy, x * z, xx - yy, (3 * xx - yy) * y, x * y * z,
                    y * (4 * zz - xx - yy), z * (2 * zz - 3 * xx - 3 * yy), x * (4 * zz - xx - yy), (xx - yy) * z,
                    (xx - 3 * yy) * x, x * y * (xx - yy), (3 * xx - yy) * y * z, x * y * (7 * zz - 1),
                    y * z * (7 * zz - 3), 35 * zz * zz - 30 * zz + 3, x * z * (7 * zz - 3), (xx - yy) * (7 * zz - 1),
                    (xx - 3 * yy) * x * z, xx * (xx - 3 * yy) - yy * (3 * xx - yy)};
    for (int i = 0; i < 25; i++) d[i] = v[i] * kHcDensityNorm[i];
}

// ------------------------------------------------------------------ static scratch
// Memory of the allocation-free device calls (dh_hc_sf_and_derivs_dev_noalloc): every per-call
// buffer of such a call (atom records, sorted dT/dF, partial sums, sorted F, raw derivatives) is
// carved out of this array, so the call makes no device allocation.  Each GPU the library runs on
// holds its own copy, at its own address, until the process ends; all calculators on that GPU
// share it, so allocation-free calls that overlap on one GPU race.  Host code never passes the
// array's name to a kernel (in host code it is a stand-in, not a device address): it takes the
// device address once per device with hipGetSymbolAddress and passes that.
static constexpr size_t kScratchBytes = DH_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

static std::mutex g_scratch_mu;
static void* g_scratch_addr[64] = {};

// Called with `device` current (engine set-up); once per device.
static void* scratch_address(int device)
{
    if (device < 0 || device >= 64) throw Error(5, "HIP device index out of range for the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_addr[device]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            throw Error(3, std::string("HIP error ") + hipGetErrorString(e) + " taking the scratch array's address");
        g_scratch_addr[device] = ptr;
    }
    return g_scratch_addr[device];
}

size_t Engine::scratch_bytes() { return kScratchBytes; }

static void upload_constants(int device)
{
    if (device >= 0 && device < 64 && g_constants_ready[device]) return;
    HIP_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(c_density_norm), kHcDensityNorm, sizeof(kHcDensityNorm)));
    static_assert(sizeof(kHcGCoef) <= 400 * sizeof(double), "g-function table too large");
    HIP_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(c_gcoef), kHcGCoef, sizeof(kHcGCoef)));
    HIP_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(c_goff), kHcGOffset, sizeof(kHcGOffset)));
    double gx[5], gw[5];
    legendre_nodes5(gx, gw);
    double dir[50][3], wq[50];
    int p = 0;
    for (int a = 0; a < 5; a++)
        for (int b = 0; b < 10; b++) {
            double phi = 2 * 3.14159265358979323846 * b / 10.0;
            double z = gx[a], rho = std::sqrt(1 - z * z);
            dir[p][0] = rho * std::cos(phi);
            dir[p][1] = rho * std::sin(phi);
            dir[p][2] = z;
            wq[p] = gw[a] * 2 * 3.14159265358979323846 / 10.0;
            p++;
        }
    double norm[25] = {0}, dv[50][25];
    for (int q = 0; q < 50; q++) {
        host_ylm(dir[q], dv[q]);
        for (int i = 0; i < 25; i++) norm[i] += wq[q] * dv[q][i] * dv[q][i];
    }
    double pw[50][25];
    for (int q = 0; q < 50; q++)
        for (int i = 0; i < 25; i++) pw[q][i] = wq[q] * dv[q][i] / norm[i];
    HIP_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(c_proj_dir), dir, sizeof(dir)));
    HIP_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(c_proj_w), pw, sizeof(pw)));
    if (device >= 0 && device < 64) g_constants_ready[device] = true;
}

// 4 pi (i^l sign) N for the deformation-valence STO r^n exp(-zeta r): N = zeta^(n+3) / (n+2)!
static double def_val_scale(double zeta, int n, int l)
{
    double norm = 1.0;
    for (int q = 0; q < n + 3; q++) norm *= zeta;
    for (int q = 2; q <= n + 2; q++) norm /= q;
    const double sgn = (l == 2 || l == 3) ? -1.0 : 1.0;   // i^l: +, +i, -, -i, +
    return sgn * 12.566370614359172953850 * norm;
}

static std::vector<TypeRec> type_records(const EngineModel& model)
{
    std::vector<TypeRec> tr(model.n_types);
    for (int t = 0; t < model.n_types; t++) {
        TypeRec& r = tr[t];
        std::memset(&r, 0, sizeof(r));
        r.wfn = model.type_wfn[t];
        r.nl = model.dv_nl[r.wfn];
        r.max_l = model.type_max_l[t];
        r.lp = std::min(std::min(4, r.nl - 1), model.types[t].n_l - 1);
        if (r.lp < r.max_l) r.lp = r.max_l;
        for (int l = 0; l < 5; l++) r.n[l] = model.dv_n[5 * r.wfn + l];
        r.pval = model.types[t].p_val;
        r.kappa = model.types[t].kappa;
        r.kappa_p = model.types[t].kappa_prime;
        r.zeta = model.dv_zeta[r.wfn];
        r.inv_kappa = 1.0 / r.kappa;
        r.inv_kappa_p = 1.0 / r.kappa_p;
        for (int l = 0; l < 5; l++) r.gscale[l] = def_val_scale(r.zeta, r.n[l], l);
        for (int l = 0; l <= r.max_l; l++)
            for (int i = l * l; i < (l + 1) * (l + 1); i++) r.plm[i] = model.types[t].plm[i];
    }
    return tr;
}

int Engine::type_class(int t) const
{
    const int lp = std::max(model_.type_max_l[t],
                            std::min(std::min(4, model_.dv_nl[model_.type_wfn[t]] - 1), model_.types[t].n_l - 1));
    return lp <= 2 ? 2 : 4;
}

void Engine::set_types(const std::vector<AtomType>& types)
{
    HIP_CHECK(hipSetDevice(device_));
    if ((int)types.size() != model_.n_types) throw Error(1, "number of atom types differs from the model");
    for (const AtomType& t : types)
        if (!(t.kappa > 0) || !(t.kappa_prime > 0)) throw Error(2, "kappa must be positive");
    model_.types = types;
    for (int t = 0; t < model_.n_types; t++) {
        int lp = -1;
        for (int l = 0; l < types[t].n_l && l < 5; l++)
            for (int i = l * l; i < (l + 1) * (l + 1); i++)
                if (types[t].plm[i] != 0.0) lp = l;
        int lw = std::min(4, model_.dv_nl[model_.type_wfn[t]] - 1);
        model_.type_max_l[t] = std::min(lw, lp);
    }
    std::vector<TypeRec> tr = type_records(model_);
    if (!tr.empty())
        HIP_CHECK(hipMemcpyAsync(d_type_.p, tr.data(), tr.size() * sizeof(TypeRec), hipMemcpyHostToDevice, stream_));
    HIP_CHECK(hipStreamSynchronize(stream_));
    plan_tables();
    jac_prepped_ = false;      // radial records depend on the type parameters
    prep_valid_ = false;
    splines_ready_ = false;    // the valence splines carry Pval and kappa
}

// Node counts of the per-wavefunction form-factor tables: core at s = |h|, valence at
// s = |h| / kappa for every type of the wavefunction, so the grid runs to h_max / min(1, kappa).
void Engine::plan_tables()
{
    const int nw = std::max(model_.n_wfn, 1);
    std::vector<double> kmin(nw, 1.0);
    for (int t = 0; t < model_.n_types; t++) {
        const int w = model_.type_wfn[t];
        kmin[w] = std::min(kmin[w], std::min(model_.types[t].kappa, model_.types[t].kappa_prime));
    }
    std::vector<TabRec> tr(nw);
    int nodes = 0;
    for (int w = 0; w < nw; w++) {
        const double smax = h_max_ / kmin[w];
        tr[w].off = nodes;
        tr[w].n = (int)std::ceil(smax / kTabStep) + 3;
        nodes += tr[w].n;
    }
    n_tab_nodes_ = nodes;
    tab_nodes_.resize(nw);
    for (int w = 0; w < nw; w++) tab_nodes_[w] = tr[w].n;
    d_tab_.ensure((size_t)nodes * kTabNode * sizeof(double));
    d_tab32_.ensure((size_t)nodes * kLinRatio * kTabNode * sizeof(float));
    d_tabrec_.ensure(tr.size() * sizeof(TabRec));
    HIP_CHECK(hipMemcpy(d_tabrec_.p, tr.data(), tr.size() * sizeof(TabRec), hipMemcpyHostToDevice));
}

// Sign pattern of an operation whose fractional and Cartesian rotations are the same diagonal sign
// matrix (bit i: axis i flips), -1 otherwise.
static int diagonal_pattern(const EngineOp& op)
{
    int pat = 0;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            const int f = op.rf[3 * i + j];
            const double c = op.rc[3 * i + j];
            if (i != j) {
                if (f != 0 || std::fabs(c) > 1e-12) return -1;
            } else {
                if ((f != 1 && f != -1) || std::fabs(c - f) > 1e-12) return -1;
                if (f == -1) pat |= 1 << i;
            }
        }
    return pat;
}

// A diagonal group's operations in increasing sign-pattern order, so that every listing of a set
// with compiled kernels reaches them (F and the derivatives are sums over the operations; only the
// order of those sums' roundings changes).  Other groups keep the caller's order.
static std::vector<EngineOp> canonical_ops(const std::vector<EngineOp>& ops)
{
    std::vector<int> pat(ops.size());
    for (size_t s = 0; s < ops.size(); s++) {
        pat[s] = diagonal_pattern(ops[s]);
        if (pat[s] < 0) return ops;
    }
    std::vector<size_t> idx(ops.size());
    std::iota(idx.begin(), idx.end(), size_t(0));
    std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) { return pat[a] < pat[b]; });
    std::vector<EngineOp> out;
    for (size_t i : idx) out.push_back(ops[i]);
    return out;
}

Engine::Engine(int device, const EngineModel& model, const double* cell6, const std::vector<EngineOp>& ops_in,
               bool centric_origin)
    : device_(device), model_(model), ops_(canonical_ops(ops_in)), centric_origin_(centric_origin)
{
    const std::vector<EngineOp>& ops = ops_;
    int n_dev = 0;
    if (hipGetDeviceCount(&n_dev) != hipSuccess || n_dev == 0) throw Error(5, "no HIP device available");
    if (device < 0 || device >= n_dev) throw Error(5, "HIP device index out of range");
    HIP_CHECK(hipSetDevice(device));
    upload_constants(device);
    scratch_ = scratch_address(device);
    HIP_CHECK(hipStreamCreate(&stream_));
    HIP_CHECK(hipEventCreate(&ev0_));
    HIP_CHECK(hipEventCreate(&ev1_));
    HIP_CHECK(hipEventCreateWithFlags(&ev_up_, hipEventDisableTiming));
    Cell c;
    c.set(cell6);
    std::memcpy(f2c_, c.f2c, sizeof(f2c_));
    std::memcpy(c2f_, c.c2f, sizeof(c2f_));
    // model tables
    std::vector<TermRec> core, val;
    std::vector<WfnRec> wr(model.n_wfn);
    for (int w = 0; w < model.n_wfn; w++) {
        wr[w].core_off = (int)core.size();
        wr[w].core_n = (int)model.core[w].c.size();
        for (size_t i = 0; i < model.core[w].c.size(); i++)
            core.push_back({model.core[w].c[i], model.core[w].e[i], model.core[w].p[i], 0});
        wr[w].val_off = (int)val.size();
        wr[w].val_n = (int)model.valence[w].c.size();
        wr[w].zeta = model.dv_zeta[w];
        wr[w].pad = 0;
        for (int l = 0; l < 5; l++) {
            const bool has = l < model.dv_nl[w];
            wr[w].n[l] = has ? model.dv_n[5 * w + l] : -1;
            wr[w].gscale[l] = has ? def_val_scale(model.dv_zeta[w], model.dv_n[5 * w + l], l) : 0.0;
        }
        for (size_t i = 0; i < model.valence[w].c.size(); i++)
            val.push_back({model.valence[w].c[i], model.valence[w].e[i], model.valence[w].p[i], 0});
    }
    std::vector<TypeRec> tr = type_records(model);
    size_t bytes_ct = std::max<size_t>(1, core.size()) * sizeof(TermRec);
    size_t bytes_vt = std::max<size_t>(1, val.size()) * sizeof(TermRec);
    d_core_.ensure(bytes_ct);
    d_val_.ensure(bytes_vt);
    d_type_.ensure(std::max<size_t>(1, tr.size()) * sizeof(TypeRec) + std::max<size_t>(1, wr.size()) * sizeof(WfnRec));
    if (!core.empty()) HIP_CHECK(hipMemcpy(d_core_.p, core.data(), core.size() * sizeof(TermRec), hipMemcpyHostToDevice));
    if (!val.empty()) HIP_CHECK(hipMemcpy(d_val_.p, val.data(), val.size() * sizeof(TermRec), hipMemcpyHostToDevice));
    if (!tr.empty()) HIP_CHECK(hipMemcpy(d_type_.p, tr.data(), tr.size() * sizeof(TypeRec), hipMemcpyHostToDevice));
    if (!wr.empty())
        HIP_CHECK(hipMemcpy(static_cast<char*>(d_type_.p) + tr.size() * sizeof(TypeRec), wr.data(),
                            wr.size() * sizeof(WfnRec), hipMemcpyHostToDevice));
    std::vector<OpDev> od(ops.size());
    for (size_t s = 0; s < ops.size(); s++) {
        std::memcpy(od[s].rf, ops[s].rf, sizeof(od[s].rf));
        std::memcpy(od[s].tf, ops[s].tf, sizeof(od[s].tf));
        std::memcpy(od[s].rc, ops[s].rc, sizeof(od[s].rc));
        od[s].anom_sign = ops[s].anom_sign;
    }
    // every operation diagonal with equal fractional / Cartesian signs -> parity-class path
    // (with a single operation the direct contraction is cheaper than the class sums)
    diag_ = ops.size() >= 2 && ops.size() <= 8;
    for (size_t s = 0; s < ops.size() && diag_; s++) {
        const int pat = diagonal_pattern(ops[s]);
        if (pat < 0) {
            diag_ = false;
            break;
        }
        for (size_t q = 0; q < s && diag_; q++)
            if (diag_pat_[q] == pat) diag_ = false;    // each sign pattern at most once
        diag_pat_[s] = pat;
        diag_anom_[s] = (float)ops[s].anom_sign;
    }
    d_ops_.ensure(od.size() * sizeof(OpDev) + 9 * sizeof(double));
    HIP_CHECK(hipMemcpy(d_ops_.p, od.data(), od.size() * sizeof(OpDev), hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(static_cast<char*>(d_ops_.p) + od.size() * sizeof(OpDev), c2f_, 9 * sizeof(double),
                        hipMemcpyHostToDevice));
}

Engine::~Engine()
{
    if (ev0_) (void)hipEventDestroy(ev0_);
    if (ev1_) (void)hipEventDestroy(ev1_);
    if (ev_up_) (void)hipEventDestroy(ev_up_);
    if (stream_) (void)hipStreamDestroy(stream_);
}

void Engine::synchronize() { HIP_CHECK(hipStreamSynchronize(stream_)); }

void Engine::upload(const EngineAtoms& at, int n_hkl, const int* hkl, const double* dtdf, bool rows_layout)
{
    // No result of an earlier run survives a new upload, and nothing runs until an upload has
    // completed: a failed upload leaves the engine without inputs (and without reusable prep).
    uploaded_ = false;
    ran_ = false;
    jac_ok_ = false;
    jac_prepped_ = false;
    HIP_CHECK(hipSetDevice(device_));
    if (n_hkl < 0 || at.n < 0) throw Error(1, "negative size");
    for (int i = 0; i < at.n; i++)
        if (at.type[i] < 0 || at.type[i] >= model_.n_types) throw Error(1, "atom type index out of range");
    try {
        upload_checked(at, n_hkl, hkl, dtdf, rows_layout);
    } catch (...) {
        prep_valid_ = false;
        hkl_gen_++;
        last_hkl_.clear();
        hkl_perm_.clear();
        n_atoms_ = n_hkl_ = n_sorted_ = n_included_ = 0;
        order_.clear();
        throw;
    }
    uploaded_ = true;
}

int Engine::diag64_code() const
{
    if (!diag_ || (ops_.size() != 2 && ops_.size() != 4)) return 0;
    int code = 4096;
    for (size_t q = 0; q < ops_.size(); q++) code |= diag_pat_[q] << (3 * q);
#define DH_SET_MATCH(C, N)                                                                            \
    if ((int)ops_.size() == N && code == C) return C;
    DH_FOR_EACH_PATTERN_SET(DH_SET_MATCH)
#undef DH_SET_MATCH
    return 0;
}

// Reflections along hkl lines for the FP64 line kernels: the axis with the fewest lines (as the
// reference chooses, scattering_utilities.cpp:371-435), lines in lexicographic order of the other
// two indices, each by increasing index along the axis.  Fills line_cont (reflection continues the
// previous one's line) and the per-operation step constants, and reorders key (caller index in
// .second) into that order.
void Engine::line_sort(int n_hkl, const int* hkl, std::vector<std::pair<double, int>>& key)
{
    int best = 2;
    size_t best_lines = (size_t)-1;
    for (int ax = 0; ax < 3; ax++) {
        const int o1 = (ax + 1) % 3, o2 = (ax + 2) % 3;
        std::vector<long long> ids(n_hkl);
        for (int k = 0; k < n_hkl; k++)
            ids[k] = ((long long)hkl[3 * k + o1] << 32) ^ (long long)(unsigned)hkl[3 * k + o2];
        std::sort(ids.begin(), ids.end());
        const size_t lines = std::unique(ids.begin(), ids.end()) - ids.begin();
        if (lines < best_lines) {
            best_lines = lines;
            best = ax;
        }
    }
    line_axis_ = best;
    const int ax = best, o1 = (ax + 1) % 3, o2 = (ax + 2) % 3;
    std::stable_sort(key.begin(), key.end(), [&](const std::pair<double, int>& x, const std::pair<double, int>& y) {
        const int* a = hkl + 3 * x.second;
        const int* b = hkl + 3 * y.second;
        if (a[o1] != b[o1]) return a[o1] < b[o1];
        if (a[o2] != b[o2]) return a[o2] < b[o2];
        return a[ax] < b[ax];
    });
    line_cont_host_.assign(std::max(n_hkl, 1), 0);
    for (int i = 1; i < n_hkl; i++) {
        const int* a = hkl + 3 * key[i - 1].second;
        const int* b = hkl + 3 * key[i].second;
        line_cont_host_[i] = a[o1] == b[o1] && a[o2] == b[o2] && b[ax] == a[ax] + 1;
    }
    d_line_cont_.ensure(line_cont_host_.size() * sizeof(int));
    HIP_CHECK(hipMemcpyAsync(d_line_cont_.p, line_cont_host_.data(), line_cont_host_.size() * sizeof(int),
                             hipMemcpyHostToDevice, stream_));
    // per operation: Delta R_f (the phase step in revolutions is Delta R_f . x + Delta . t), Delta . t,
    // and m(Delta_s) with Delta_s = Delta in Cartesian reciprocal coordinates times R_s
    line_ops_host_.assign(ops_.size() * 10, 0.0);
    double dc[3];
    for (int j = 0; j < 3; j++) dc[j] = c2f_[3 * ax + j];
    for (size_t s = 0; s < ops_.size(); s++) {
        const EngineOp& o = ops_[s];
        double* q = &line_ops_host_[10 * s];
        for (int j = 0; j < 3; j++) q[j] = o.rf[3 * ax + j];
        q[3] = o.tf[ax];
        double ds[3];
        for (int j = 0; j < 3; j++) ds[j] = dc[0] * o.rc[j] + dc[1] * o.rc[3 + j] + dc[2] * o.rc[6 + j];
        const double m[6] = {ds[0] * ds[0], ds[1] * ds[1], ds[2] * ds[2], 2 * ds[0] * ds[1], 2 * ds[0] * ds[2],
                             2 * ds[1] * ds[2]};
        for (int j = 0; j < 6; j++) q[4 + j] = m[j];
    }
    d_line_ops_.ensure(std::max<size_t>(line_ops_host_.size(), 1) * sizeof(double));
    HIP_CHECK(hipMemcpyAsync(d_line_ops_.p, line_ops_host_.data(), line_ops_host_.size() * sizeof(double),
                             hipMemcpyHostToDevice, stream_));
}

void Engine::upload_checked(const EngineAtoms& at, int n_hkl, const int* hkl, const double* dtdf, bool rows_layout)
{
    // the copies of the previous upload read the page-locked staging buffer: they must be done
    // before it is rewritten (or reallocated)
    if (up_pending_) {
        HIP_CHECK(hipEventSynchronize(ev_up_));
        up_pending_ = false;
    }
    dev_ready_ = false;
    n_atoms_ = at.n;
    n_hkl_ = n_hkl;
    rows_layout_ = rows_layout;
    // The atom order depends only on the atoms' types, the include mask and the layout: a
    // refinement's calls keep the order of the first.
    const bool same_atoms = !order_.empty() && rows_layout == last_rows_layout_ &&
                            at.type == last_atom_type_ && at.include == last_include_;
    if (!same_atoms) {
    last_rows_layout_ = rows_layout;
    last_atom_type_ = at.type;
    last_include_ = at.include;
    // sort included atoms by multipole class, then type; pad each class to whole waves
    std::vector<int> idx;
    for (int i = 0; i < at.n; i++)
        if (at.include[i]) idx.push_back(i);
    auto cls_of = [&](int i) { return type_class(at.type[i]); };
    order_.clear();
    std::vector<int>& wave_cls = wave_cls_;
    wave_cls.clear();
    if (rows_layout) {
        // caller order (included atoms), so that a wave's 64 atoms are adjacent in every
        // per-reflection row and its stores are contiguous; a wave takes the larger multipole
        // class of its atoms (the l > 2 coefficients of a class-2 type are zero)
        for (int i = 0; i < at.n; i++)
            if (at.include[i]) order_.push_back(i);
        while (order_.size() % kWave) order_.push_back(-1);
        for (size_t wv = 0; wv < order_.size() / kWave; wv++) {
            int c = 2;
            for (int l = 0; l < kWave; l++) {
                const int i = order_[wv * kWave + l];
                if (i >= 0) c = std::max(c, cls_of(i));
            }
            wave_cls.push_back(c);
        }
    } else {
        // within a class by wavefunction, then type: the lanes of a wave then share table rows
        std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) {
            int cx = cls_of(x), cy = cls_of(y);
            if (cx != cy) return cx < cy;
            const int wx = model_.type_wfn[at.type[x]], wy = model_.type_wfn[at.type[y]];
            if (wx != wy) return wx < wy;
            return at.type[x] < at.type[y];
        });
        size_t n2 = 0;
        for (int i : idx) n2 += cls_of(i) == 2 ? 1 : 0;
        const size_t n4 = idx.size() - n2;
        auto waves = [](size_t n) { return (n + kWave - 1) / kWave; };
        if (waves(n2 + n4) < waves(n2) + waves(n4)) {
            // one padded run when that saves a wave (small structures): a wave takes the larger
            // class of its atoms; the l > 2 multipole and radial terms of a class-2 type are zero
            for (int i : idx) order_.push_back(i);
            while (order_.size() % kWave) order_.push_back(-1);
            for (size_t wv = 0; wv < order_.size() / kWave; wv++) {
                int c = 2;
                for (int l = 0; l < kWave; l++) {
                    const int i = order_[wv * kWave + l];
                    if (i >= 0) c = std::max(c, cls_of(i));
                }
                wave_cls.push_back(c);
            }
        } else {
            for (int pass = 2; pass <= 4; pass += 2) {
                size_t start = order_.size();
                for (int i : idx)
                    if (cls_of(i) == pass) order_.push_back(i);
                size_t cnt = order_.size() - start;
                size_t padded = (cnt + kWave - 1) / kWave * kWave;
                while (order_.size() < start + padded) order_.push_back(-1);
                for (size_t wv = 0; wv < padded / kWave; wv++) wave_cls.push_back(pass);
            }
        }
    }
    n_included_ = (int)idx.size();
    }
    const std::vector<int>& wave_cls = wave_cls_;
    n_sorted_ = (int)order_.size();
    // reflection order: caller's (rows); along hkl lines when the line kernels apply and the
    // radial factors come from per-(reflection, type) records; by |h| otherwise, so that per-atom
    // table lookups of neighbouring reflections fall on the same table nodes (along a line |h|
    // jumps by a whole reflection step and the lookups of a chunk spread over the whole table)
    const int order_mode =
        rows_layout ? 0 : (line_order && line_ops_ok() && !tune.direct_phases && !use_otf(false) ? 2 : 1);
    // page-locked staging: [atoms | wave classes | dT/dF and multipliers | sorted reflections]
    const size_t n_ain = std::max(n_sorted_, 1);
    const size_t off_cls = n_ain * sizeof(AtomIn);
    const size_t in_bytes = off_cls + wave_cls.size() * sizeof(int);
    const size_t off_dtdf = (in_bytes + 255) / 256 * 256;
    const size_t off_hkl = off_dtdf + 3 * (size_t)std::max(n_hkl, 1) * sizeof(double);
    h_up_.ensure(off_hkl + 3 * (size_t)std::max(n_hkl, 1) * sizeof(int));
    AtomIn* ain = h_up_.at<AtomIn>(0);
    std::memset(ain, 0, sizeof(AtomIn));
    for (int s = 0; s < n_sorted_; s++) {
        AtomIn& r = ain[s];
        std::memset(&r, 0, sizeof(r));
        int i = order_[s];
        r.orig = i;
        if (i < 0) {
            r.type = 0;
            r.lcs[0] = r.lcs[4] = r.lcs[8] = 1.0;
            continue;
        }
        for (int j = 0; j < 3; j++) r.xf[j] = at.xyz_frac[3 * i + j];
        for (int j = 0; j < 6; j++) r.u[j] = at.u6[6 * i + j];
        r.w = at.weight[i];
        r.anom[0] = at.anom[2 * i];
        r.anom[1] = at.anom[2 * i + 1];
        for (int j = 0; j < 9; j++) r.lcs[j] = at.lcs[9 * i + j];
        r.type = at.type[i];
    }
    if (!wave_cls.empty()) std::memcpy(h_up_.at<int>(off_cls), wave_cls.data(), wave_cls.size() * sizeof(int));
    d_atoms_in_.ensure(in_bytes + 64);
    HIP_CHECK(hipMemcpyAsync(d_atoms_in_.p, h_up_.p, in_bytes, hipMemcpyHostToDevice, stream_));
    d_atoms_.ensure(n_ain * sizeof(AtomRec));
    // Reflections are processed in order of increasing |h| (stable; hkl_perm_[i] = caller index of
    // position i): consecutive reflections of a wave then read neighbouring nodes of the radial
    // tables.  |h| range for the form-factor tables.  A refinement passes the same list every
    // call: then the order, the tables' range and the device copy of the indices are kept.
    const bool same_hkl = order_mode == last_order_ && (size_t)3 * n_hkl == last_hkl_.size() &&
                          (n_hkl == 0 || std::memcmp(last_hkl_.data(), hkl, (size_t)3 * n_hkl * sizeof(int)) == 0);
    d_hkl_.ensure(std::max(n_hkl, 1) * 3 * sizeof(int));
    d_dtdf_.ensure(std::max(n_hkl, 1) * 3 * sizeof(double));   // dtdf (2) + mult (1)
    int* hkl_s = h_up_.at<int>(off_hkl);
    if (!same_hkl) {
        prep_valid_ = false;
        hkl_gen_++;
        std::vector<std::pair<double, int>> key(n_hkl);
        for (int k = 0; k < n_hkl; k++) {
            double v = 0.0;
            for (int j = 0; j < 3; j++) {
                const double hc = hkl[3 * k] * c2f_[j] + hkl[3 * k + 1] * c2f_[3 + j] + hkl[3 * k + 2] * c2f_[6 + j];
                v += hc * hc;
            }
            key[k] = {v, k};
        }
        h_max_ = 0.0;
        for (int k = 0; k < n_hkl; k++) h_max_ = std::max(h_max_, std::sqrt(key[k].first));
        if (order_mode == 1) std::sort(key.begin(), key.end());     // (|h|^2, index): the stable order
        if (order_mode == 2) line_sort(n_hkl, hkl, key);
        hkl_perm_.resize(n_hkl);
        for (int k = 0; k < n_hkl; k++) hkl_perm_[k] = key[k].second;
        plan_tables();
        for (int i = 0; i < n_hkl; i++)
            for (int j = 0; j < 3; j++) hkl_s[3 * i + j] = hkl[3 * hkl_perm_[i] + j];
        last_hkl_.assign(hkl, hkl + 3 * (size_t)n_hkl);
        last_order_ = order_mode;
        if (n_hkl > 0)
            HIP_CHECK(hipMemcpyAsync(d_hkl_.p, hkl_s, (size_t)n_hkl * 3 * sizeof(int), hipMemcpyHostToDevice,
                                     stream_));
    }
    double* dtdf_s = h_up_.at<double>(off_dtdf);
    for (int i = 0; i < n_hkl; i++) {
        const int k = hkl_perm_[i];
        dtdf_s[2 * i] = dtdf[2 * k];
        dtdf_s[2 * i + 1] = dtdf[2 * k + 1];
        dtdf_s[2 * (size_t)n_hkl + i] = dtdf[2 * (size_t)n_hkl + k];
    }
    if (n_hkl > 0)
        HIP_CHECK(hipMemcpyAsync(d_dtdf_.p, dtdf_s, (size_t)n_hkl * 3 * sizeof(double), hipMemcpyHostToDevice,
                                 stream_));
    // the copies complete on the stream, ahead of the kernels of run(); the next upload waits
    HIP_CHECK(hipEventRecord(ev_up_, stream_));
    up_pending_ = true;
}

__global__ void k_fill(double* __restrict__ p, size_t n, double v)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

void Engine::fill(DeviceBuffer& b, size_t n, double v) { fill(b.as<double>(), n, v); }

void Engine::fill(double* p, size_t n, double v)
{
    if (n == 0) return;
    const int T = 256;
    hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + T - 1) / T)), dim3(T), 0, stream_, p, n, v);
    HIP_CHECK(hipGetLastError());
}

// every per-call table of the prep kernels (|h| lives in d_radial32_)
void Engine::ensure_prep_buffers(Precision prec, bool mp)
{
    const int n_types = std::max(model_.n_types, 1);
    const int n_hkl = n_hkl_;
    const bool mixed = prec == kMixed;
    const size_t n_rec = (size_t)n_hkl * ops_.size();
    if (!use_otf(mp)) d_radial_.ensure((size_t)n_hkl * n_types * kRadRec * (mixed ? sizeof(float) : sizeof(double)));
    d_hklop_.ensure(std::max(n_rec * kHklopRec, (size_t)std::max(n_hkl, 1) * kDiag64Rec) *
                    (mixed ? sizeof(float) : sizeof(double)));
    if (mixed) d_phase_.ensure(n_rec * 4 * sizeof(double));
    d_radial32_.ensure((size_t)std::max(n_hkl, 1) * sizeof(double));
    if (use_diag(mixed, mp)) d_ht_.ensure(std::max<size_t>(n_rec, 1) * sizeof(double));
    if (mp) d_radx_.ensure((size_t)n_hkl * n_types * kRadRec * (mixed ? sizeof(float) : sizeof(double)));
}

void Engine::run_prep(Precision prec, bool mp, bool electron, bool general)
{
    jac_prepped_ = false;
    ensure_prep_buffers(prec, mp);
    // the reflection records, tables and radial records depend on the reflections, the types and
    // the path only: kept from the previous call with the same key until one of them changes
    const bool mixed_k = prec == kMixed;
    // per-reflection FP64 diagonal records: the FP64 line kernels and the FP64 direct kernels of
    // the compiled sign-pattern sets (table lookups)
    const bool diag64 = (use_diag64(mixed_k, mp) || (!mixed_k && (main_code(mixed_k, mp) & kDiagDirect) != 0)) &&
                        !general;
    const int key = (mixed_k ? 1 : 0) | (mp ? 2 : 0) | (electron && electron_ ? 4 : 0) |
                    (use_diag(mixed_k, mp) && !general ? 8 : 0) | (use_otf(mp) ? 16 : 0) | (diag64 ? 32 : 0);
    // the records of a multipole-adjoint prep (key | 2) are those of the plain prep plus the
    // derivative records: they serve plain calls too (the torch step alternates the two)
    const bool reuse = prep_valid_ && (prep_key_ == key || prep_key_ == (key | 2));
    prep_valid_ = true;
    if (!reuse) prep_key_ = key;
    const int n_ops = (int)ops_.size();
    const int n_types = std::max(model_.n_types, 1);
    const int n_hkl = n_hkl_;
    const int n_slots = n_sorted_;
    const bool mixed = prec == kMixed;
    DeviceBuffer& hlen = d_radial32_;   // |h| per reflection
    const OpDev* ops = d_ops_.as<OpDev>();
    const double* c2f = reinterpret_cast<const double*>(static_cast<const char*>(d_ops_.p) + n_ops * sizeof(OpDev));
    const TypeRec* types = d_type_.as<TypeRec>();
    const WfnRec* wfns = reinterpret_cast<const WfnRec*>(static_cast<const char*>(d_type_.p) +
                                                         model_.n_types * sizeof(TypeRec));
    const AtomIn* ain = d_atoms_in_.as<AtomIn>();
    const int T = 256;
    if (!reuse) {
        if (use_diag(mixed, mp) && !general) {
            hipLaunchKernelGGL(k_hkl_prep_diag, dim3((n_hkl + T - 1) / T), dim3(T), 0, stream_, d_hkl_.as<int>(), ops,
                               n_ops, n_hkl, c2f, d_hklop_.as<float>(), d_phase_.as<double>(), d_ht_.as<double>(),
                               hlen.as<double>());
        } else if (diag64) {
            hipLaunchKernelGGL(k_hkl_prep_diag64, dim3((n_hkl + T - 1) / T), dim3(T), 0, stream_, d_hkl_.as<int>(),
                               ops, n_ops, n_hkl, c2f, d_hklop_.as<double>(), hlen.as<double>());
        } else {
            hipLaunchKernelGGL(k_hkl_prep, dim3((n_hkl + T - 1) / T), dim3(T), 0, stream_, d_hkl_.as<int>(), ops, n_ops,
                               n_hkl, c2f, mixed ? nullptr : d_hklop_.as<double>(), mixed ? d_hklop_.as<float>() : nullptr,
                               mixed ? d_phase_.as<double>() : nullptr, hlen.as<double>());
        }
        HIP_CHECK(hipGetLastError());
        if (model_.n_types > 0) {
            hipLaunchKernelGGL(k_tab_build, dim3(n_tab_nodes_), dim3(64), 0, stream_,
                               d_tabrec_.as<TabRec>(), model_.n_wfn, wfns, d_core_.as<TermRec>(), d_val_.as<TermRec>(),
                               n_tab_nodes_, d_tab_.as<double>());
            HIP_CHECK(hipGetLastError());
            if (mixed && use_otf(mp)) {
                const int nl = n_tab_nodes_ * kLinRatio * 8;
                hipLaunchKernelGGL(k_lin_build, dim3((nl + T - 1) / T), dim3(T), 0, stream_, d_tabrec_.as<TabRec>(),
                                   model_.n_wfn, d_tab_.as<double>(), n_tab_nodes_, d_tab32_.as<float>());
                HIP_CHECK(hipGetLastError());
            }
        }
        if (model_.n_types > 0 && !use_otf(mp)) {
            int nt = n_hkl * model_.n_types;
            hipLaunchKernelGGL(k_radial_prep, dim3((nt + T - 1) / T), dim3(T), 0, stream_, hlen.as<double>(),
                               d_tab_.as<double>(), d_tabrec_.as<TabRec>(), types, wfns, d_val_.as<TermRec>(),
                               model_.n_types, n_hkl,
                               mixed ? nullptr : d_radial_.as<double>(), mixed ? d_radial_.as<float>() : nullptr,
                               (mp && !mixed) ? d_radx_.as<double>() : nullptr,
                               (mp && mixed) ? d_radx_.as<float>() : nullptr,
                               electron && electron_ ? d_ez_.as<double>() : nullptr);
            HIP_CHECK(hipGetLastError());
        }
    }
    hipLaunchKernelGGL(k_atom_prep, dim3(n_slots), dim3(64), 0, stream_, ain, types, d_atoms_.as<AtomRec>(),
                       n_slots);
    HIP_CHECK(hipGetLastError());
}

// mp_pass: 1 or 2 selects the pass of the multipole adjoint (0 without it); line_ops > 0: the FP64
// line kernel for that many operations
static const void* main_kernel(bool mixed, bool derivs, int mp_pass, bool diag, bool otf, int line_ops = 0,
                               int diag64 = 0)
{
#define KM(M, D, P, G, O) (const void*)k_main<M, D, P, G, O>
#define KL(M, N) (derivs ? (const void*)k_main<M, true, 0, false, false, N> : (const void*)k_main<M, false, 0, false, false, N>)
#define KD(M, N, C) (derivs ? (const void*)k_main<M, true, 0, M, false, N, C> : (const void*)k_main<M, false, 0, M, false, N, C>)
    // line kernels (radial records only: reflections are ordered along lines only on that path):
    // FP64 diagonal groups take per-reflection FP64 records, mixed diagonal groups the
    // parity-folded FP32 records; other groups take the general records of their precision.
    // The multipole passes run along lines in the mixed parity-folded kernels only.
    if (diag64 > 0 && mp_pass > 0 && !otf && (!mixed || diag) && (diag64 & kDiagDirect) == 0) {
#define KPM(M, N, C) (mp_pass == 1 ? (const void*)k_main<M, true, 1, M, false, N, C>                              \
                                   : mp_pass == 2 ? (const void*)k_main<M, true, 2, M, false, N, C>               \
                                                  : (const void*)k_main<M, true, 3, M, false, N, C>)
#define KP(N, C) (mixed ? KPM(true, N, C) : KPM(false, N, C))
#define DH_CASE(C, N)                                                                                 \
    case C: return KP(N, C);
        switch (diag64) {
            DH_FOR_EACH_PATTERN_SET(DH_CASE)
        default: break;
        }
#undef DH_CASE
#undef KP
#undef KPM
    }
    if (diag64 > 0 && mp_pass == 0 && !otf && (!mixed || diag)) {
#define DH_CASE(C, N)                                                                                 \
    case C: return mixed ? KD(true, N, C) : KD(false, N, C);
        switch (diag64) {
            DH_FOR_EACH_PATTERN_SET(DH_CASE)
        default: break;
        }
#undef DH_CASE
    }
#undef KD
    // FP64 diagonal-record kernel with a compiled sign-pattern set, every term direct (table lookups)
    if ((diag64 & kDiagDirect) && !mixed && otf && mp_pass == 0) {
#define DH_CASE(C, N)                                                                                 \
    case C:                                                                                           \
        return derivs ? (const void*)k_main<false, true, 0, false, true, N, C | kDiagDirect>           \
                      : (const void*)k_main<false, false, 0, false, true, N, C | kDiagDirect>;
        switch (diag64 & ~kDiagDirect) {
            DH_FOR_EACH_PATTERN_SET(DH_CASE)
        default: break;
        }
#undef DH_CASE
    }
    // mixed parity-folded kernel with a compiled sign-pattern set, every term direct (table lookups)
    if ((diag64 & kDiagDirect) && mixed && diag && otf && mp_pass == 0) {
#define DH_CASE(C, N)                                                                                 \
    case C:                                                                                           \
        return derivs ? (const void*)k_main<true, true, 0, true, true, N, C | kDiagDirect>             \
                      : (const void*)k_main<true, false, 0, true, true, N, C | kDiagDirect>;
        switch (diag64 & ~kDiagDirect) {
            DH_FOR_EACH_PATTERN_SET(DH_CASE)
        default: break;
        }
#undef DH_CASE
    }
    if (line_ops > 0 && mp_pass == 0 && !otf && !(mixed && diag)) {
        switch (line_ops) {
        case 1: return mixed ? KL(true, 1) : KL(false, 1);
        case 2: return mixed ? KL(true, 2) : KL(false, 2);
        case 3: return mixed ? KL(true, 3) : KL(false, 3);
        case 4: return mixed ? KL(true, 4) : KL(false, 4);
        default: break;
        }
    }
#undef KL
    if (mp_pass == 1) return mixed ? (diag ? KM(true, true, 1, true, false) : KM(true, true, 1, false, false))
                                   : KM(false, true, 1, false, false);
    if (mp_pass == 2) return mixed ? (diag ? KM(true, true, 2, true, false) : KM(true, true, 2, false, false))
                                   : KM(false, true, 2, false, false);
    if (mp_pass == 3) return mixed ? (diag ? KM(true, true, 3, true, false) : KM(true, true, 3, false, false))
                                   : KM(false, true, 3, false, false);
    if (mixed) {
        if (diag) {
            if (otf) return derivs ? KM(true, true, 0, true, true) : KM(true, false, 0, true, true);
            return derivs ? KM(true, true, 0, true, false) : KM(true, false, 0, true, false);
        }
        if (otf) return derivs ? KM(true, true, 0, false, true) : KM(true, false, 0, false, true);
        return derivs ? KM(true, true, 0, false, false) : KM(true, false, 0, false, false);
    }
    if (otf) return derivs ? KM(false, true, 0, false, true) : KM(false, false, 0, false, true);
    return derivs ? KM(false, true, 0, false, false) : KM(false, false, 0, false, false);
#undef KM
}

int Engine::main_code(bool mixed, bool mp) const
{
    const int d64 = use_diag64(mixed, mp);
    if (d64) return d64;
    if (!mp && use_otf(mp) && (mixed ? use_diag(mixed, mp) : diag_ && !tune.general_path)) {
        const int c = diag64_code();
        if (c != 0) return c | kDiagDirect;
    }
    return 0;
}

int Engine::resident_waves(bool mixed, bool derivs, bool mp)
{
    const int line = use_line(mixed, mp) ? (use_diag64(mixed, mp) ? 5 : (int)ops_.size())
                                         : ((main_code(mixed, mp) & kDiagDirect) ? 6 : 0);
    const int key = line * 32 + (use_otf(mp) ? 16 : 0) + (use_diag(mixed, mp) ? 8 : 0) + (mixed ? 4 : 0) +
                    (derivs ? 2 : 0) + (mp ? 1 : 0);
    if (resident_waves_[key] > 0) return resident_waves_[key];
    int blocks = 0;
    HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &blocks, main_kernel(mixed, derivs, mp ? 1 : 0, use_diag(mixed, mp), use_otf(mp),
                             use_line(mixed, mp) ? (int)ops_.size() : 0, main_code(mixed, mp)),
        256, 0));
    hipDeviceProp_t prop;
    HIP_CHECK(hipGetDeviceProperties(&prop, device_));
    resident_waves_[key] = std::max(1, blocks) * kWavesPerBlock * prop.multiProcessorCount;
    return resident_waves_[key];
}

int Engine::default_chunk(bool mixed, bool derivs, bool mp)
{
    // about four resident-capacity rounds of (atom wave, reflection chunk) work units
    const int n_waves = n_sorted_ / kWave;
    const double target = 4.0 * resident_waves(mixed, derivs, mp);
    int n_chunks = std::max(1, (int)std::lround(target / std::max(n_waves, 1)));
    return std::max(8, (n_hkl_ + n_chunks - 1) / n_chunks);
}

// Arguments of the main kernel for this structure (partial-sum buffers sized here).
void Engine::main_args(Precision prec, bool derivs, bool mp, int chunk, MainArgs* out)
{
    const bool mixed = prec == kMixed;
    const int n_slots = n_sorted_;
    const int n_waves = n_slots / kWave;
    const int n_hkl = n_hkl_;
    const int n_chunks = (n_hkl + chunk - 1) / chunk;
    d_fpart_.ensure(std::max<size_t>((size_t)n_waves * n_hkl * 2 * (mixed ? sizeof(float) : sizeof(double)), 1));
    part_f32_ = mixed;
    const int nd = mp ? kNdMultipole : kNdBasic;
    if (derivs || mp) d_dpart_.ensure((size_t)n_chunks * n_slots * nd * sizeof(double));
    MainArgs& a = *out;
    std::memset(&a, 0, sizeof(a));
    a.hklop = d_hklop_.p;
    a.phase = mixed ? d_phase_.as<double>() : nullptr;
    a.radial = d_radial_.p;
    a.radx = mp ? d_radx_.p : nullptr;
    a.dtdf = d_dtdf_.as<double>();    // [re, im] pairs followed by the centering multipliers
    a.atoms = d_atoms_.as<AtomRec>();
    a.wave_l = reinterpret_cast<const int*>(static_cast<const char*>(d_atoms_in_.p) +
                                            std::max(n_sorted_, 1) * sizeof(AtomIn));
    a.fpart = d_fpart_.as<double>();
    a.dpart = d_dpart_.as<double>();
    a.n_hkl = n_hkl;
    a.n_ops = (int)ops_.size();
    a.n_types = std::max(model_.n_types, 1);
    a.n_slots = n_slots;
    a.n_waves = n_waves;
    a.chunk = chunk;
    a.grid_x = (n_waves + kWavesPerBlock - 1) / kWavesPerBlock;
    a.n_blocks = a.grid_x * n_chunks;
    a.per_xcd = (a.n_blocks + kXcds - 1) / kXcds;
    const bool diag = use_diag(mixed, mp) || use_diag64(mixed, mp) || (main_code(mixed, mp) & kDiagDirect) != 0;
    a.ht = use_diag(mixed, mp) ? d_ht_.as<double>() : nullptr;
    a.tab = mixed ? d_tab32_.p : d_tab_.p;
    a.tabs = d_tabrec_.as<TabRec>();
    a.hlen = d_radial32_.as<double>();
    a.line_cont = d_line_cont_.as<int>();
    a.line_ops = d_line_ops_.as<double>();
    a.diag_mask = 0;
    for (int p = 0; p < 8; p++) {
        a.diag_op[p] = 0;
        a.diag_anom[p] = 0.f;
    }
    if (diag)
        for (int s = 0; s < (int)ops_.size(); s++) {
            a.diag_mask |= 1 << diag_pat_[s];
            a.diag_op[diag_pat_[s]] = s;
            a.diag_anom[diag_pat_[s]] = diag_anom_[s];
        }
}

void Engine::launch_main(Precision prec, bool derivs, bool mp, int chunk)
{
    MainArgs a;
    main_args(prec, derivs, mp, chunk, &a);
    dim3 grid(a.per_xcd * kXcds);
    last_grid_x = a.grid_x;
    last_grid_y = a.n_blocks / std::max(a.grid_x, 1);
    last_chunk = chunk;
    void* args[] = {&a};
    const int passes[2] = {mp ? (mp_scalars ? 1 : 3) : 0, 2};
    const int line = use_line(prec == kMixed, mp) ? (int)ops_.size() : 0;
    const int d64 = main_code(prec == kMixed, mp);
    for (int p = 0; p < (mp ? 2 : 1); p++) {
        const void* kern =
            main_kernel(prec == kMixed, derivs, passes[p], use_diag(prec == kMixed, mp), use_otf(mp), line, d64);
        HIP_CHECK(hipLaunchKernel(kern, grid, dim3(256), args, 0, stream_));
    }
}

int Engine::chunk_for(Precision prec, bool derivs, bool mp)
{
    if (chunk_override_ > 0) return chunk_override_;
    if (tune.hkl_chunk > 0) return tune.hkl_chunk;
    auto it = tuned_.find(shape_key(prec, derivs, mp));
    if (it != tuned_.end()) return it->second;
    return default_chunk(prec == kMixed, derivs, mp);
}

long long Engine::shape_key(Precision prec, bool derivs, bool mp) const
{
    return ((((long long)n_sorted_ * 1000003LL + n_hkl_) * 64 + (long long)ops_.size()) * 3 +
            (use_line(prec == kMixed, mp) ? 1 : 0) + (use_diag64(prec == kMixed, mp) ? 1 : 0)) * 32 +
           (use_otf(mp) ? 16 : 0) + (use_diag(prec == kMixed, mp) ? 8 : 0) + (prec == kMixed ? 4 : 0) +
           (derivs ? 2 : 0) + (mp ? 1 : 0);
}

void Engine::run(Precision prec, bool derivs, bool mp)
{
    HIP_CHECK(hipSetDevice(device_));
    if (!uploaded_) throw Error(1, "no inputs: upload them before running the calculation");
    ran_ = true;
    if (mp) derivs = true;
    last_nd_ = mp ? kNdMultipole : kNdBasic;
    const int n_hkl = n_hkl_;
    const int n_slots = n_sorted_;
    const int n_waves = n_slots / kWave;
    d_f_.ensure(std::max(n_hkl, 1) * 2 * sizeof(double));
    d_d_.ensure((size_t)std::max(n_atoms_, 1) * last_nd_ * sizeof(double));
    fill(d_d_, (size_t)std::max(n_atoms_, 1) * last_nd_, 0.0);
    if (n_hkl == 0) return;
    if (n_waves == 0) {
        fill(d_f_, (size_t)n_hkl * 2, 0.0);
        return;
    }
    run_prep(prec, mp);
    const int chunk = chunk_for(prec, derivs, mp);
    if (timing) HIP_CHECK(hipEventRecord(ev0_, stream_));
    launch_main(prec, derivs, mp, chunk);
    if (timing) HIP_CHECK(hipEventRecord(ev1_, stream_));
    reduce(derivs, mp, chunk);
    if (timing) {
        HIP_CHECK(hipEventSynchronize(ev1_));
        HIP_CHECK(hipEventElapsedTime(&last_main_ms, ev0_, ev1_));
    }
}

// Stage 2: F from the per-wave partials, derivatives from the per-chunk partials.
void Engine::reduce(bool derivs, bool mp, int chunk)
{
    const int n_hkl = n_hkl_;
    const int n_slots = n_sorted_;
    const int n_waves = n_slots / kWave;
    const int n_chunks = (n_hkl + chunk - 1) / chunk;
    const int nd = mp ? kNdMultipole : kNdBasic;
    // outputs sized here: the autotuner reaches this before any run() has allocated them
    d_f_.ensure((size_t)std::max(n_hkl, 1) * 2 * sizeof(double));
    if (derivs) d_d_.ensure((size_t)std::max(n_atoms_, 1) * nd * sizeof(double));
    if (d_fpart_.bytes < (size_t)n_waves * n_hkl * 2 * (part_f32_ ? sizeof(float) : sizeof(double)) ||
        (derivs && d_dpart_.bytes < (size_t)n_chunks * n_slots * nd * sizeof(double)))
        throw Error(1, "internal: stage-2 partial buffers smaller than the launch geometry");
    const int T = 64;
    if (part_f32_)
        hipLaunchKernelGGL(k_reduce_f<float2>, dim3((n_hkl + T - 1) / T), dim3(T), 0, stream_, d_fpart_.as<float2>(),
                           d_dtdf_.as<double>() + 2 * (size_t)n_hkl, n_waves, n_hkl, centric_origin_ ? 1 : 0,
                           d_f_.as<double>());
    else
        hipLaunchKernelGGL(k_reduce_f<double2>, dim3((n_hkl + T - 1) / T), dim3(T), 0, stream_, d_fpart_.as<double2>(),
                           d_dtdf_.as<double>() + 2 * (size_t)n_hkl, n_waves, n_hkl, centric_origin_ ? 1 : 0,
                           d_f_.as<double>());
    HIP_CHECK(hipGetLastError());
    if (!derivs) return;
    hipLaunchKernelGGL(k_reduce_d, dim3((n_slots * nd + T - 1) / T), dim3(T), 0, stream_, d_dpart_.as<double>(),
                       d_atoms_.as<AtomRec>(), n_chunks, n_slots, nd, d_d_.as<double>());
    HIP_CHECK(hipGetLastError());
    if (mp) {
        if (lcs_grad) {
            d_lcsg_.ensure((size_t)std::max(n_atoms_, 1) * 9 * sizeof(double));
            fill(d_lcsg_, (size_t)std::max(n_atoms_, 1) * 9, 0.0);
            hipLaunchKernelGGL(k_lcs_grad, dim3(n_slots), dim3(64), 0, stream_, d_atoms_in_.as<AtomIn>(),
                               d_type_.as<TypeRec>(), nd, d_d_.as<double>(), d_lcsg_.as<double>());
            HIP_CHECK(hipGetLastError());
        }
        hipLaunchKernelGGL(k_plm_back, dim3(n_slots), dim3(64), 0, stream_, d_atoms_in_.as<AtomIn>(), n_slots, nd,
                           d_d_.as<double>());
        HIP_CHECK(hipGetLastError());
    }
}

void Engine::download_lcs_grad(double* g)
{
    HIP_CHECK(hipSetDevice(device_));
    if (!ran_ || !lcs_grad || last_nd_ < kNdMultipole) throw Error(1, "internal: no local-frame gradient computed");
    HIP_CHECK(hipMemcpyAsync(g, d_lcsg_.p, (size_t)n_atoms_ * 9 * sizeof(double), hipMemcpyDeviceToHost, stream_));
    HIP_CHECK(hipStreamSynchronize(stream_));
}

int Engine::autotune(Precision prec, bool derivs, bool mp)
{
    HIP_CHECK(hipSetDevice(device_));
    if (!uploaded_) throw Error(1, "no inputs: upload them before tuning");
    if (mp) derivs = true;
    const int n_waves = n_sorted_ / kWave;
    if (n_hkl_ == 0 || n_waves == 0) return 0;
    run_prep(prec, mp);
    const int cap = resident_waves(prec == kMixed, derivs, mp);
    std::vector<int> cands;
    for (double rounds : {1.0, 2.0, 3.0, 4.0, 6.0, 8.0, 12.0, 16.0, 24.0}) {
        int n_chunks = std::max(1, (int)std::lround(rounds * cap / n_waves));
        int c = std::max(8, (n_hkl_ + n_chunks - 1) / n_chunks);
        if (std::find(cands.begin(), cands.end(), c) == cands.end()) cands.push_back(c);
    }
    auto time_once = [&](int c) {
        HIP_CHECK(hipEventRecord(ev0_, stream_));
        launch_main(prec, derivs, mp, c);
        reduce(derivs, mp, c);
        HIP_CHECK(hipEventRecord(ev1_, stream_));
        HIP_CHECK(hipEventSynchronize(ev1_));
        float ms = 0;
        HIP_CHECK(hipEventElapsedTime(&ms, ev0_, ev1_));
        return ms;
    };
    // one untimed launch first (clocks and caches), then the best of three per candidate
    (void)time_once(cands.front());
    std::vector<float> ms_min(cands.size(), 1e30f);
    for (size_t i = 0; i < cands.size(); i++)
        for (int rep = 0; rep < 3; rep++) ms_min[i] = std::min(ms_min[i], time_once(cands[i]));
    std::vector<size_t> rank(cands.size());
    for (size_t i = 0; i < rank.size(); i++) rank[i] = i;
    std::sort(rank.begin(), rank.end(), [&](size_t x, size_t y) { return ms_min[x] < ms_min[y]; });
    int best = cands[rank[0]];
    if (rank.size() > 1) {
        // the two leaders again, interleaved, best of five more each (they often differ by noise)
        const size_t a = rank[0], b = rank[1];
        for (int rep = 0; rep < 5; rep++) {
            ms_min[a] = std::min(ms_min[a], time_once(cands[a]));
            ms_min[b] = std::min(ms_min[b], time_once(cands[b]));
        }
        best = ms_min[b] < ms_min[a] ? cands[b] : cands[a];
    }
    tuned_[shape_key(prec, derivs, mp)] = best;
    ran_ = false;     // the outputs hold a tuning launch, not a run() result
    return best;
}

void Engine::download(double* f, double* d10)
{
    HIP_CHECK(hipSetDevice(device_));
    if (!ran_) throw Error(1, "no results: run the calculation before fetching");
    // through page-locked staging: [F (sorted order) | per-atom derivative sums]
    const size_t f_bytes = (size_t)n_hkl_ * 2 * sizeof(double);
    const size_t d_bytes = (size_t)n_atoms_ * last_nd_ * sizeof(double);
    const size_t off_d = (f_bytes + 255) / 256 * 256;
    h_down_.ensure(off_d + d_bytes);
    const double* fs = h_down_.at<double>(0);
    if (f && n_hkl_ > 0)
        HIP_CHECK(hipMemcpyAsync(h_down_.p, d_f_.p, f_bytes, hipMemcpyDeviceToHost, stream_));
    if (d10 && n_atoms_ > 0)
        HIP_CHECK(hipMemcpyAsync(h_down_.at<char>(off_d), d_d_.p, d_bytes, hipMemcpyDeviceToHost, stream_));
    HIP_CHECK(hipStreamSynchronize(stream_));
    if (d10 && n_atoms_ > 0) std::memcpy(d10, h_down_.at<char>(off_d), d_bytes);
    if (f)
        for (int i = 0; i < n_hkl_; i++) {
            f[2 * (size_t)hkl_perm_[i]] = fs[2 * (size_t)i];
            f[2 * (size_t)hkl_perm_[i] + 1] = fs[2 * (size_t)i + 1];
        }
}

// ------------------------------------------------------------------ device-resident calls

struct DevAtomArgs {
    const double* xyz;
    const int* adp_size;
    const double* adp;
    const double* occ;
    const double* lcs;
    const double* anom;
    const unsigned char* include;
    const double* mult_w;        // m / |G| per atom (per image with coset representatives)
    const double* anom_default;  // f', f'' per atom from its wavefunction type (per image)
    int n_slots, n_waves;
    // coset representatives (set_dev_cosets): engine atom i is image i % K of caller atom i / K,
    // representative k = reps[21 k ..]: fractional rotation (9), translation (3), Cartesian
    // rotation (9).  K = 1: no images.
    int K;
    const double* reps;
};

// caller atom of engine atom (image) i
__device__ __forceinline__ int dev_atom(const DevAtomArgs& a, int i) { return a.K > 1 ? i / a.K : i; }

// AtomIn of every slot from the caller's arrays: the slot layout (atom index, type, padding)
// comes from the skeleton prepare_dev uploaded, the values are converted to the engine's
// conventions (fractional xyz, 2 pi^2 U_cart, weight occ * m/|G| or 0 for excluded atoms).  The
// multipole class of every wave is copied with it.  skel and out may be the same array.
__global__ void k_dev_atoms(const AtomIn* __restrict__ skel, const int* __restrict__ skel_cls, AtomIn* out,
                            int* out_cls, DevAtomArgs a, DevConv c)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < a.n_waves) out_cls[s] = skel_cls[s];
    if (s >= a.n_slots) return;
    AtomIn r = skel[s];
    const int img = r.orig;
    if (img >= 0) {
        const int i = dev_atom(a, img);    // the caller's atom
        const double* x = a.xyz + 3 * (size_t)i;
        if (c.xyz_identity) {
            for (int j = 0; j < 3; j++) r.xf[j] = x[j];
        } else {
            for (int j = 0; j < 3; j++) r.xf[j] = c.c2f[3 * j] * x[0] + c.c2f[3 * j + 1] * x[1] + c.c2f[3 * j + 2] * x[2];
        }
        const int na = a.adp_size[i];
        const double* u = a.adp + 6 * (size_t)i;
        for (int j = 0; j < 6; j++) r.u[j] = 0.0;
        if (na == 6) {
            for (int j = 0; j < 6; j++) {
                double v = 0.0;
                for (int k = 0; k < 6; k++) v += c.u2c[6 * j + k] * u[k];
                r.u[j] = v;
            }
        } else if (na == 1) {
            const double two_pi2 = 2.0 * M_PI * M_PI;
            for (int j = 0; j < 3; j++) r.u[j] = u[0] * two_pi2;
        }
        const bool inc = a.include ? a.include[i] != 0 : true;
        r.w = inc ? a.occ[i] * a.mult_w[img] : 0.0;
        const double* an = a.anom ? a.anom + 2 * (size_t)i : a.anom_default + 2 * (size_t)img;
        r.anom[0] = an[0];
        r.anom[1] = an[1];
        for (int j = 0; j < 9; j++) r.lcs[j] = a.lcs ? a.lcs[9 * (size_t)i + j] : (j % 4 == 0 ? 1.0 : 0.0);
        if (a.K > 1) {
            // image under representative k: position R x + t, U -> R U R^T, frame R L
            const double* g = a.reps + 21 * (size_t)(img % a.K);
            double xf[3], L[9];
            for (int p = 0; p < 3; p++) xf[p] = g[9 + p] + g[3 * p] * r.xf[0] + g[3 * p + 1] * r.xf[1] + g[3 * p + 2] * r.xf[2];
            for (int p = 0; p < 3; p++) r.xf[p] = xf[p];
            const double* R = g + 12;
            if (na == 6) {
                const double U[9] = {r.u[0], r.u[3], r.u[4], r.u[3], r.u[1], r.u[5], r.u[4], r.u[5], r.u[2]};
                double RU[9], V[9];
                for (int p = 0; p < 3; p++)
                    for (int q = 0; q < 3; q++)
                        RU[3 * p + q] = R[3 * p] * U[q] + R[3 * p + 1] * U[3 + q] + R[3 * p + 2] * U[6 + q];
                for (int p = 0; p < 3; p++)
                    for (int q = 0; q < 3; q++)
                        V[3 * p + q] = RU[3 * p] * R[3 * q] + RU[3 * p + 1] * R[3 * q + 1] + RU[3 * p + 2] * R[3 * q + 2];
                r.u[0] = V[0];
                r.u[1] = V[4];
                r.u[2] = V[8];
                r.u[3] = 0.5 * (V[1] + V[3]);
                r.u[4] = 0.5 * (V[2] + V[6]);
                r.u[5] = 0.5 * (V[5] + V[7]);
            }
            for (int p = 0; p < 3; p++)
                for (int q = 0; q < 3; q++)
                    L[3 * p + q] = R[3 * p] * r.lcs[q] + R[3 * p + 1] * r.lcs[3 + q] + R[3 * p + 2] * r.lcs[6 + q];
            for (int j = 0; j < 9; j++) r.lcs[j] = L[j];
        }
    }
    out[s] = r;
}

// dT/dF in the engine's reflection order: conj(dT/dF) * centring multiplier (2 Re for a centre at
// the origin), followed by the multipliers themselves (what upload() stages on the host).
__global__ void k_dev_dtdf(const double* __restrict__ dtdf, const int* __restrict__ perm,
                           const double* __restrict__ mult, int n_hkl, int centric_origin, double* __restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_hkl) return;
    const int k = perm[i];
    const double m = mult[k];
    // F only (dtdf null): dT/dF is not used, the multipliers still scale F
    double re = dtdf ? dtdf[2 * (size_t)k] * m : 0.0, im = dtdf ? -dtdf[2 * (size_t)k + 1] * m : 0.0;
    if (centric_origin) {
        re *= 2.0;
        im = 0.0;
    }
    out[2 * (size_t)i] = re;
    out[2 * (size_t)i + 1] = im;
    out[2 * (size_t)n_hkl + i] = m;
}

// F back in caller order and the per-atom derivatives in the caller's conventions (what the host
// fetch does: x 2 pi w, x 2 pi^2 w, x m/|G|, the U_iso sum and the convention matrices).
__global__ void k_dev_out(const double* __restrict__ fs, const int* __restrict__ perm, const double* __restrict__ draw,
                          int nd, int n_hkl, int n_atoms, DevAtomArgs a, DevConv c, unsigned sel,
                          double* __restrict__ f, double* __restrict__ dxyz, double* __restrict__ dadp,
                          double* __restrict__ docc)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < n_hkl && f) {
        const int k = perm[idx];
        f[2 * (size_t)k] = fs[2 * (size_t)idx];
        f[2 * (size_t)k + 1] = fs[2 * (size_t)idx + 1];
    }
    if (idx >= n_atoms) return;
    const int i = idx;
    double r[10];
    if (a.K > 1) {
        // the images' rows folded back: xyz through R^T (fractional), the symmetric U gradient
        // through R^T G R (Cartesian), occupancy added
        for (int j = 0; j < 10; j++) r[j] = 0.0;
        for (int k = 0; k < a.K; k++) {
            const double* q = draw + (size_t)nd * ((size_t)i * a.K + k);
            const double* g = a.reps + 21 * (size_t)k;
            const double* R = g + 12;
            for (int p = 0; p < 3; p++) r[p] += g[p] * q[0] + g[3 + p] * q[1] + g[6 + p] * q[2];
            const double G[9] = {q[3], 0.5 * q[6], 0.5 * q[7], 0.5 * q[6], q[4], 0.5 * q[8], 0.5 * q[7], 0.5 * q[8], q[5]};
            double M[9];
            for (int p = 0; p < 3; p++)
                for (int t = 0; t < 3; t++) {
                    double v = 0.0;
                    for (int m = 0; m < 3; m++)
                        v += R[3 * m + p] * (G[3 * m] * R[t] + G[3 * m + 1] * R[3 + t] + G[3 * m + 2] * R[6 + t]);
                    M[3 * p + t] = v;
                }
            r[3] += M[0];
            r[4] += M[4];
            r[5] += M[8];
            r[6] += M[1] + M[3];
            r[7] += M[2] + M[6];
            r[8] += M[5] + M[7];
            r[9] += q[9];
        }
    } else {
        for (int j = 0; j < 10; j++) r[j] = draw[(size_t)nd * i + j];
    }
    const double mw = a.mult_w[(size_t)i * a.K];
    const bool inc = a.include ? a.include[i] != 0 : true;
    const double w = inc ? a.occ[i] * mw : 0.0;
    if (dxyz) {
        double* o = dxyz + 3 * (size_t)i;
        if ((sel & 1u) && inc) {
            const double two_pi = 2.0 * M_PI;
            const double g[3] = {two_pi * w * r[0], two_pi * w * r[1], two_pi * w * r[2]};
            for (int j = 0; j < 3; j++)
                o[j] = c.dxyz_identity ? g[j] : c.dxyz_m[3 * j] * g[0] + c.dxyz_m[3 * j + 1] * g[1] + c.dxyz_m[3 * j + 2] * g[2];
        } else {
            o[0] = o[1] = o[2] = 0.0;
        }
    }
    if (dadp) {
        double* o = dadp + 6 * (size_t)i;
        double out6[6] = {0, 0, 0, 0, 0, 0};
        const int na = a.adp_size[i];
        if ((sel & 2u) && inc && (na == 1 || na == 6)) {
            const double two_pi2 = 2.0 * M_PI * M_PI;
            double du[6];
            for (int j = 0; j < 6; j++) du[j] = two_pi2 * w * r[3 + j];
            if (na == 1) {
                out6[0] = du[0] + du[1] + du[2];
            } else {
                for (int j = 0; j < 6; j++) {
                    double v = 0.0;
                    for (int k = 0; k < 6; k++) v += c.dadp_m[6 * j + k] * du[k];
                    out6[j] = v;
                }
            }
        }
        for (int j = 0; j < 6; j++) o[j] = out6[j];
    }
    if (docc) docc[i] = ((sel & 4u) && inc) ? mw * r[9] : 0.0;
}

// Per-type sums of a device call's multipole derivatives, as the host call forms them:
// dP[t][j] = sum over the atoms a of type t of w_a dT/dP_j (local frame, rows 10..34) and the
// Pval, kappa, kappa' sums from rows 35..37, w_a = occupancy x m/|G| (0 for excluded atoms);
// P_lm entries with l >= n_l[t] are not parameters and read 0.  The atoms of type t are
// list[off[t] .. off[t + 1]) (ascending).  One workgroup per type: thread (column c, lane q) sums
// the list entries q, q + 8, ..., then lane 0 adds the eight lane sums in a fixed order (no
// atomics: the sums are reproducible bit for bit).
__global__ __launch_bounds__(256) void k_dev_type_sums(const double* __restrict__ draw, int nd,
                                                       const int* __restrict__ list, const int* __restrict__ off,
                                                       const int* __restrict__ n_l, DevAtomArgs a,
                                                       double* __restrict__ dplm, double* __restrict__ dpval,
                                                       double* __restrict__ dkappa, double* __restrict__ dkp)
{
    __shared__ double part[8][32];
    const int t = blockIdx.x;
    const int c = threadIdx.x & 31, q = threadIdx.x >> 5;
    double s = 0.0;
    if (c < 28)
        for (int e = off[t] + q; e < off[t + 1]; e += 8) {
            const int img = list[e], i = dev_atom(a, img);
            if (a.include ? a.include[i] != 0 : true) s += a.occ[i] * a.mult_w[img] * draw[(size_t)nd * img + 10 + c];
        }
    part[q][c] = s;
    __syncthreads();
    if (q != 0 || c >= 28) return;
    double v = 0.0;
    for (int r = 0; r < 8; r++) v += part[r][c];
    if (c < 25) {
        const int l = c < 1 ? 0 : c < 4 ? 1 : c < 9 ? 2 : c < 16 ? 3 : 4;
        if (dplm) dplm[25 * (size_t)t + c] = l < n_l[t] ? v : 0.0;
    } else if (c == 25) {
        if (dpval) dpval[t] = v;
    } else if (c == 26) {
        if (dkappa) dkappa[t] = v;
    } else if (dkp) {
        dkp[t] = v;
    }
}

void Engine::set_dev_cosets(const std::vector<EngineOp>& reps)
{
    HIP_CHECK(hipSetDevice(device_));
    const int K = std::max<int>(1, (int)reps.size());
    if (n_atoms_ % K != 0) throw Error(1, "internal: images of another atom count");
    std::vector<double> r(21 * (size_t)K, 0.0);
    for (size_t k = 0; k < reps.size(); k++) {
        for (int j = 0; j < 9; j++) r[21 * k + j] = reps[k].rf[j];
        for (int j = 0; j < 3; j++) r[21 * k + 9 + j] = reps[k].tf[j];
        for (int j = 0; j < 9; j++) r[21 * k + 12 + j] = reps[k].rc[j];
    }
    // earlier calls on any stream may still read the representatives: wait before rewriting them
    HIP_CHECK(hipDeviceSynchronize());
    d_dev_reps_.ensure(r.size() * sizeof(double));
    HIP_CHECK(hipMemcpy(d_dev_reps_.p, r.data(), r.size() * sizeof(double), hipMemcpyHostToDevice));
    dev_K_ = K;
}

void Engine::set_type_sum_map(const std::vector<int>& atom_type, const std::vector<int>& n_l)
{
    HIP_CHECK(hipSetDevice(device_));
    if ((int)atom_type.size() != n_atoms_) throw Error(1, "internal: type map of another atom count");
    // [atoms in type order (ascending within a type), type offsets (n_types + 1), n_l (n_types)]
    const int nt = (int)n_l.size();
    std::vector<int> off(nt + 1, 0);
    for (int t : atom_type) {
        if (t < 0 || t >= nt) throw Error(1, "internal: atom type out of range");
        off[t + 1]++;
    }
    for (int t = 0; t < nt; t++) off[t + 1] += off[t];
    std::vector<int> m(atom_type.size()), fill_at(off.begin(), off.end() - 1);
    for (size_t i = 0; i < atom_type.size(); i++) m[fill_at[atom_type[i]]++] = (int)i;
    m.insert(m.end(), off.begin(), off.end());
    m.insert(m.end(), n_l.begin(), n_l.end());
    // earlier calls on any stream may still read the map: wait for them before rewriting it
    HIP_CHECK(hipDeviceSynchronize());
    d_tsum_map_.ensure(std::max<size_t>(m.size(), 1) * sizeof(int));
    HIP_CHECK(hipMemcpy(d_tsum_map_.p, m.data(), m.size() * sizeof(int), hipMemcpyHostToDevice));
    tsum_types_ = (int)n_l.size();
}

void Engine::prepare_dev(const EngineAtoms& layout, int n_hkl, const int* hkl, const double* mult,
                         const std::vector<double>& mult_w, const std::vector<double>& anom_default,
                         const DevConv& conv, Precision prec)
{
    HIP_CHECK(hipSetDevice(device_));
    if ((int)mult_w.size() != layout.n || (int)anom_default.size() != 2 * layout.n)
        throw Error(1, "internal: per-atom constants of the device layout");
    for (int i = 0; i < layout.n; i++)
        if (!layout.include[i]) throw Error(1, "internal: the device layout holds every atom");
    // dT/dF zero, multipliers in the third part: the layout upload itself stages the reflections
    std::vector<double> dd(3 * (size_t)std::max(n_hkl, 0), 0.0);
    for (int k = 0; k < n_hkl; k++) dd[2 * (size_t)n_hkl + k] = mult[k];
    upload(layout, n_hkl, hkl, dd.data(), false);
    d_dev_perm_.ensure((size_t)std::max(n_hkl, 1) * sizeof(int));
    d_dev_mult_.ensure((size_t)std::max(n_hkl, 1) * sizeof(double));
    d_dev_atomc_.ensure((size_t)std::max(layout.n, 1) * 3 * sizeof(double));
    std::vector<double> atomc(3 * (size_t)layout.n);
    for (int i = 0; i < layout.n; i++) {
        atomc[i] = mult_w[i];
        atomc[layout.n + 2 * (size_t)i] = anom_default[2 * (size_t)i];
        atomc[layout.n + 2 * (size_t)i + 1] = anom_default[2 * (size_t)i + 1];
    }
    if (n_hkl > 0) {
        HIP_CHECK(hipMemcpyAsync(d_dev_perm_.p, hkl_perm_.data(), (size_t)n_hkl * sizeof(int), hipMemcpyHostToDevice,
                                 stream_));
        HIP_CHECK(hipMemcpyAsync(d_dev_mult_.p, mult, (size_t)n_hkl * sizeof(double), hipMemcpyHostToDevice, stream_));
    }
    if (layout.n > 0)
        HIP_CHECK(hipMemcpyAsync(d_dev_atomc_.p, atomc.data(), atomc.size() * sizeof(double), hipMemcpyHostToDevice,
                                 stream_));
    dev_conv_ = conv;
    // reflection records for this mode now, and the launch-shape queries, so that later calls
    // (including allocation-free ones and captured ones) find them in place
    if (n_hkl > 0 && n_sorted_ > 0) {
        run_prep(prec, false);
        (void)chunk_for(prec, true, false);
        (void)chunk_for(prec, false, false);
        const int n_waves = n_sorted_ / kWave;
        d_f_.ensure((size_t)n_hkl * 2 * sizeof(double));
        d_fpart_.ensure((size_t)n_waves * n_hkl * 2 * sizeof(double));
    }
    HIP_CHECK(hipStreamSynchronize(stream_));
    dev_ready_ = true;
    ran_ = false;
}

namespace {
// For one device-resident call: the engine's stream replaced by the caller's and, with a scratch
// area, the per-call buffers replaced by views into it while allocation is refused.  Everything
// is put back on exit, also when the call throws.
struct DevScope {
    struct Saved {
        DeviceBuffer* b;
        void* p;
        size_t bytes;
    };
    Saved saved[8];
    int n_saved = 0;
    hipStream_t* stream_slot;
    hipStream_t old_stream;
    bool* timing_slot;
    bool old_timing;
    int* chunk_slot;
    bool* ran_slot;
    DevScope(hipStream_t* slot, hipStream_t st, bool* timing, int* chunk, bool* ran)
        : stream_slot(slot), old_stream(*slot), timing_slot(timing), old_timing(*timing), chunk_slot(chunk),
          ran_slot(ran)
    {
        *slot = st;
        *timing = false;     // no host synchronisation inside a device-resident call
    }
    void view(DeviceBuffer& b, void* p, size_t bytes)
    {
        saved[n_saved++] = {&b, b.p, b.bytes};
        b.p = p;
        b.bytes = bytes;
    }
    ~DevScope()
    {
        for (int i = n_saved - 1; i >= 0; i--) {
            saved[i].b->p = saved[i].p;
            saved[i].b->bytes = saved[i].bytes;
        }
        *stream_slot = old_stream;
        *timing_slot = old_timing;
        *chunk_slot = 0;
        *ran_slot = false;     // results went to the caller's memory (or the call failed)
        g_no_alloc = false;
    }
};

size_t align256(size_t n) { return (n + 255) / 256 * 256; }
}  // namespace

void Engine::run_dev(const DevCall& call, Precision prec, hipStream_t stream, void* scratch_mem, size_t scratch_size)
{
    HIP_CHECK(hipSetDevice(device_));
    if (!dev_ready_ || !uploaded_)
        throw Error(1, "device-resident call: set the reflections first (dh_hc_set_reflections)");
    // with coset representatives (set_dev_cosets) the engine's atoms are the K images of each of
    // the caller's n atoms
    const int n_img = n_atoms_, n = n_atoms_ / dev_K_, n_hkl = n_hkl_;
    if (n > 0 && (!call.xyz || !call.adp_size || !call.adp || !call.occ)) throw Error(1, "null atom array");
    if (n_hkl > 0 && !call.f) throw Error(1, "null structure-factor output");
    if (n_hkl > 0 && call.selector && !call.dtdf) throw Error(1, "null dT/dF");
    const bool mp = call.multipole;
    const bool derivs = call.selector != 0 || mp;
    if (mp && scratch_mem) throw Error(1, "allocation-free calls take no multipole derivatives");
    if (mp && tsum_types_ == 0) throw Error(1, "internal: no atom-type map for the multipole sums");
    if (mp && n_hkl > 0 && !call.dtdf) throw Error(1, "null dT/dF");
    const int n_ain = std::max(n_sorted_, 1);
    const int n_waves = n_sorted_ / kWave;
    AtomIn* skel = d_atoms_in_.as<AtomIn>();
    const int* skel_cls = reinterpret_cast<const int*>(static_cast<const char*>(d_atoms_in_.p) + n_ain * sizeof(AtomIn));
    DevScope scope(&stream_, stream, &timing, &chunk_override_, &ran_);
    if (scratch_mem) {
        // carve the per-call buffers; the reflection chunk is lengthened if the derivative partials
        // of the usual launch shape do not fit
        const size_t b_ain = align256(n_ain * sizeof(AtomIn) + n_waves * sizeof(int) + 64);
        const size_t b_atoms = align256(n_ain * sizeof(AtomRec));
        const size_t b_dtdf = align256((size_t)std::max(n_hkl, 1) * 3 * sizeof(double));
        const size_t b_f = align256((size_t)std::max(n_hkl, 1) * 2 * sizeof(double));
        const size_t b_d = align256((size_t)std::max(n_img, 1) * kNdBasic * sizeof(double));
        const size_t b_fpart = align256((size_t)n_waves * n_hkl * 2 * (prec == kMixed ? sizeof(float) : sizeof(double)));
        const size_t fixed = b_ain + b_atoms + b_dtdf + b_f + b_d + b_fpart;
        const size_t per_chunk = (size_t)n_sorted_ * kNdBasic * sizeof(double);
        if (fixed + (derivs ? per_chunk : 0) > scratch_size)
            throw Error(4, "allocation-free call: " + std::to_string(n) + " atoms x " + std::to_string(n_hkl) +
                               " reflections need more than the " + std::to_string(scratch_size >> 20) +
                               " MiB static scratch");
        size_t b_dpart = 256;
        if (derivs && n_hkl > 0 && n_waves > 0) {
            int chunk = chunk_for(prec, true, false);
            int n_chunks = (n_hkl + chunk - 1) / chunk;
            const size_t max_chunks = (scratch_size - fixed) / per_chunk;
            if ((size_t)n_chunks > max_chunks) {
                n_chunks = (int)max_chunks;
                chunk = (n_hkl + n_chunks - 1) / n_chunks;
                n_chunks = (n_hkl + chunk - 1) / chunk;
            }
            chunk_override_ = chunk;
            b_dpart = align256((size_t)n_chunks * per_chunk);
        }
        char* base = static_cast<char*>(scratch_mem);
        size_t off = 0;
        auto carve = [&](DeviceBuffer& b, size_t bytes) {
            scope.view(b, base + off, bytes);
            off += bytes;
        };
        carve(d_atoms_in_, b_ain);
        carve(d_atoms_, b_atoms);
        carve(d_dtdf_, b_dtdf);
        carve(d_f_, b_f);
        carve(d_d_, b_d);
        carve(d_fpart_, b_fpart);
        carve(d_dpart_, b_dpart);
        g_no_alloc = true;
    }
    DevAtomArgs aa;
    aa.xyz = call.xyz;
    aa.adp_size = call.adp_size;
    aa.adp = call.adp;
    aa.occ = call.occ;
    aa.lcs = call.lcs;
    aa.anom = call.anom;
    aa.include = call.include;
    aa.mult_w = d_dev_atomc_.as<double>();
    aa.anom_default = d_dev_atomc_.as<double>() + n_img;   // [m/|G| | f', f''] per engine atom
    aa.n_slots = n_sorted_;
    aa.n_waves = n_waves;
    aa.K = dev_K_;
    aa.reps = d_dev_reps_.as<double>();
    const int T = 256;
    if (n_sorted_ > 0) {
        int* out_cls = reinterpret_cast<int*>(static_cast<char*>(d_atoms_in_.p) + n_ain * sizeof(AtomIn));
        hipLaunchKernelGGL(k_dev_atoms, dim3((n_sorted_ + T - 1) / T), dim3(T), 0, stream_, skel, skel_cls,
                           d_atoms_in_.as<AtomIn>(), out_cls, aa, dev_conv_);
        HIP_CHECK(hipGetLastError());
    }
    if (n_hkl > 0) {
        hipLaunchKernelGGL(k_dev_dtdf, dim3((n_hkl + T - 1) / T), dim3(T), 0, stream_, derivs ? call.dtdf : nullptr,
                           d_dev_perm_.as<int>(), d_dev_mult_.as<double>(), n_hkl, dev_conv_.centric_origin,
                           d_dtdf_.as<double>());
        HIP_CHECK(hipGetLastError());
    }
    if (mp) {
        mp_scalars = true;       // pass 1 with the Pval, kappa, kappa' sums
        lcs_grad = false;
    }
    run(prec, derivs, mp);
    const int n_out = std::max(n_hkl, n);
    if (n_out > 0) {
        hipLaunchKernelGGL(k_dev_out, dim3((n_out + T - 1) / T), dim3(T), 0, stream_, d_f_.as<double>(),
                           d_dev_perm_.as<int>(), d_d_.as<double>(), last_nd_, n_hkl, n, aa, dev_conv_, call.selector,
                           call.f, call.d_xyz, call.d_adp, call.d_occ);
        HIP_CHECK(hipGetLastError());
    }
    if (mp) {
        const int* list = d_tsum_map_.as<int>();
        const int* off = list + n_img;
        hipLaunchKernelGGL(k_dev_type_sums, dim3(tsum_types_), dim3(256), 0, stream_, d_d_.as<double>(), last_nd_,
                           list, off, off + tsum_types_ + 1, aa, call.d_plm, call.d_pval, call.d_kappa,
                           call.d_kappa_prime);
        HIP_CHECK(hipGetLastError());
    }
}

// ------------------------------------------------------------------ per-reflection rows

void Engine::set_electron(bool on, const std::vector<double>& z_f0)
{
    HIP_CHECK(hipSetDevice(device_));
    if (on) {
        if ((int)z_f0.size() != 2 * model_.n_types) throw Error(1, "internal: electron data per type");
        d_ez_.ensure(std::max<size_t>(z_f0.size(), 1) * sizeof(double));
        HIP_CHECK(hipMemcpy(d_ez_.p, z_f0.data(), z_f0.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    electron_ = on;
    jac_prepped_ = false;
    prep_valid_ = false;
}

void Engine::set_jacobian_conventions(const JacConv& conv)
{
    HIP_CHECK(hipSetDevice(device_));
    if ((int)conv.mult_w.size() != n_atoms_ || (int)conv.adp_kind.size() != n_atoms_)
        throw Error(1, "internal: per-atom data of the per-reflection rows do not match the upload");
    if (!rows_layout_) throw Error(1, "internal: per-reflection rows need the caller's reflection and atom order");
    std::vector<double> m(45);
    std::memcpy(m.data(), conv.xyz_m, 9 * sizeof(double));
    std::memcpy(m.data() + 9, conv.adp_m, 36 * sizeof(double));
    d_jconv_.ensure(m.size() * sizeof(double));
    HIP_CHECK(hipMemcpyAsync(d_jconv_.p, m.data(), m.size() * sizeof(double), hipMemcpyHostToDevice, stream_));
    std::vector<JacAtom> ja(std::max(n_atoms_, 1));
    for (int i = 0; i < n_atoms_; i++) {
        ja[i].mult_w = conv.mult_w[i];
        ja[i].adp_kind = conv.adp_kind[i];
        ja[i].pad = 0;
    }
    d_jatoms_.ensure(ja.size() * sizeof(JacAtom));
    HIP_CHECK(hipMemcpyAsync(d_jatoms_.p, ja.data(), ja.size() * sizeof(JacAtom), hipMemcpyHostToDevice, stream_));
    HIP_CHECK(hipStreamSynchronize(stream_));
    jac_ok_ = true;
}

static const void* jac_kernel(bool centric, bool otf, bool mixed)
{
    if (mixed) {
        if (centric) return otf ? (const void*)k_jac<true, true, true> : (const void*)k_jac<true, false, true>;
        return otf ? (const void*)k_jac<false, true, true> : (const void*)k_jac<false, false, true>;
    }
    if (centric) return otf ? (const void*)k_jac<true, true, false> : (const void*)k_jac<true, false, false>;
    return otf ? (const void*)k_jac<false, true, false> : (const void*)k_jac<false, false, false>;
}

// Reflections per work unit: one staging block (8), or fewer when that would leave resident wave
// slots idle.  Longer chunks lose more to the last, partly filled round of units than they save
// in per-unit set-up (measured on 60- and 5000-atom structures).
int Engine::jac_chunk(int n_rows, Precision prec)
{
    if (tune.hkl_chunk > 0) return tune.hkl_chunk;
    const int m = prec == kMixed ? 1 : 0;
    if (jac_resident_[m] == 0) {
        int blocks = 0;
        HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &blocks, jac_kernel(centric_origin_, use_otf(false), prec == kMixed), 256, 0));
        hipDeviceProp_t prop;
        HIP_CHECK(hipGetDeviceProperties(&prop, device_));
        jac_resident_[m] = std::max(1, blocks) * kWavesPerBlock * prop.multiProcessorCount;
    }
    const long long work = (long long)n_rows * std::max(n_sorted_ / kWave, 1);
    return (int)std::max(1LL, std::min((long long)kStage, work / jac_resident_[m]));
}

void Engine::jacobian(int k0, int k1, double* f_dev, double* rows_dev, Precision prec)
{
    HIP_CHECK(hipSetDevice(device_));
    if (!jac_ok_) throw Error(1, "per-reflection derivatives: no upload for them (dh_hc_jacobian_upload)");
    if (k0 < 0 || k1 > n_hkl_ || k0 > k1) throw Error(1, "reflection range outside the uploaded reflections");
    const int n = k1 - k0;
    if (n == 0) return;
    if (!f_dev || !rows_dev) throw Error(1, "null output pointer");
    const int n_waves = n_sorted_ / kWave;
    if (n_waves == 0 || n_included_ < n_atoms_) fill(rows_dev, (size_t)n * kJacPlanes * n_atoms_ * 2, 0.0);
    if (n_waves == 0) {
        fill(f_dev, (size_t)n * 2, 0.0);
        return;
    }
    const bool mixed = prec == kMixed;
    if (!jac_prepped_ || jac_prec_ != prec) {
        run_prep(prec, false, false, true);     // the per-hkl engine has no electron-scattering mode
        const int n_rec = n_hkl_ * (int)ops_.size();
        d_jrec_.ensure((size_t)std::max(n_rec, 1) * kJacRec * sizeof(double));
        if (mixed)
            hipLaunchKernelGGL(k_jac_rec32, dim3((n_rec + 255) / 256), dim3(256), 0, stream_, d_hklop_.as<float>(),
                               n_rec, d_jconv_.as<double>(), d_jrec_.as<double>());
        else
            hipLaunchKernelGGL(k_jac_rec, dim3((n_rec + 255) / 256), dim3(256), 0, stream_, d_hklop_.as<double>(),
                               n_rec, d_jconv_.as<double>(), d_jrec_.as<double>());
        HIP_CHECK(hipGetLastError());
        jac_prepped_ = true;
        jac_prec_ = prec;
    }
    const int chunk = jac_chunk(n, prec);
    MainArgs a;
    main_args(prec, false, false, chunk, &a);
    JacArgs jb;
    std::memset(&jb, 0, sizeof(jb));
    jb.rows = rows_dev;
    jb.f = f_dev;
    jb.atoms = d_jatoms_.as<JacAtom>();
    jb.n_atoms = n_atoms_;
    jb.k0 = k0;
    jb.k1 = k1;
    jb.jrec = d_jrec_.as<double>();
    jb.chunk = chunk;
    jb.n_units = n_waves * ((n + chunk - 1) / chunk);
    last_chunk = chunk;
    last_grid_x = (jb.n_units + kWavesPerBlock - 1) / kWavesPerBlock;
    last_grid_y = 1;
    void* args[] = {&a, &jb};
    if (timing) HIP_CHECK(hipEventRecord(ev0_, stream_));
    HIP_CHECK(hipLaunchKernel(jac_kernel(centric_origin_, use_otf(false), mixed), dim3(last_grid_x), dim3(256), args,
                              0, stream_));
    if (timing) HIP_CHECK(hipEventRecord(ev1_, stream_));
    const int T = 64;
    if (mixed)
        hipLaunchKernelGGL(k_reduce_f_rows<float2>, dim3((n + T - 1) / T), dim3(T), 0, stream_, d_fpart_.as<float2>(),
                           d_dtdf_.as<double>() + 2 * (size_t)n_hkl_, n_waves, n_hkl_, k0, n, f_dev);
    else
        hipLaunchKernelGGL(k_reduce_f_rows<double2>, dim3((n + T - 1) / T), dim3(T), 0, stream_,
                           d_fpart_.as<double2>(), d_dtdf_.as<double>() + 2 * (size_t)n_hkl_, n_waves, n_hkl_, k0, n,
                           f_dev);
    HIP_CHECK(hipGetLastError());
    if (timing) {
        HIP_CHECK(hipEventSynchronize(ev1_));
        HIP_CHECK(hipEventElapsedTime(&last_main_ms, ev0_, ev1_));
    }
}

// ------------------------------------------------------------------ form-factor route (host)

// NaturalCubicSpline::set (NaturalCubicSpline.cpp) on the grid values of every spline: the
// tridiagonal system of the derivatives x_i (rows 1 4 1, first and last rows with 4 on the
// diagonal as in the reference), solved by forward elimination and back substitution, then
// segment coefficients [y_i, x_i, 3(y_{i+1} - y_i) - 2 x_i - x_{i+1}, 2(y_i - y_{i+1}) + x_i + x_{i+1}];
// the last node's segment stays zero.
static void spline_coefficients(const double* y, int n, double* coeff)
{
    std::vector<double> b(n, 4.0), d(n), x(n);
    d[0] = 3.0 * (y[1] - y[0]);
    for (int i = 1; i < n - 1; i++) d[i] = 3.0 * (y[i + 1] - y[i - 1]);
    d[n - 1] = 3.0 * (y[n - 1] - y[n - 2]);
    for (int i = 1; i < n; i++) {
        const double w = 1.0 / b[i - 1];
        b[i] -= w;
        d[i] -= w * d[i - 1];
    }
    x[n - 1] = d[n - 1] / b[n - 1];
    for (int i = n - 2; i >= 0; i--) x[i] = (d[i] - x[i + 1]) / b[i];
    for (int i = 0; i < n - 1; i++) {
        double* c = coeff + 4 * (size_t)i;
        c[0] = y[i];
        c[1] = x[i];
        c[2] = 3.0 * (y[i + 1] - y[i]) - 2.0 * x[i] - x[i + 1];
        c[3] = 2.0 * (y[i] - y[i + 1]) + x[i] + x[i + 1];
    }
    for (int j = 0; j < 4; j++) coeff[4 * (size_t)(n - 1) + j] = 0.0;
}

void Engine::build_splines()
{
    const int n_spl = model_.n_wfn + model_.n_types;
    const size_t nv = (size_t)n_spl * kSplineNodes;
    DeviceBuffer vals;
    vals.ensure(nv * sizeof(double));
    const TypeRec* types = d_type_.as<TypeRec>();
    const WfnRec* wfns = reinterpret_cast<const WfnRec*>(static_cast<const char*>(d_type_.p) +
                                                         model_.n_types * sizeof(TypeRec));
    hipLaunchKernelGGL(k_spline_values, dim3((unsigned)((nv + 255) / 256)), dim3(256), 0, stream_, types, wfns,
                       d_core_.as<TermRec>(), d_val_.as<TermRec>(), model_.n_wfn, model_.n_types, vals.as<double>());
    HIP_CHECK(hipGetLastError());
    std::vector<double> y(nv), c(nv * 4);
    HIP_CHECK(hipMemcpyAsync(y.data(), vals.p, nv * sizeof(double), hipMemcpyDeviceToHost, stream_));
    HIP_CHECK(hipStreamSynchronize(stream_));
    for (int sp = 0; sp < n_spl; sp++)
        spline_coefficients(y.data() + (size_t)sp * kSplineNodes, kSplineNodes, c.data() + (size_t)sp * kSplineNodes * 4);
    d_spline_.ensure(c.size() * sizeof(double));
    HIP_CHECK(hipMemcpyAsync(d_spline_.p, c.data(), c.size() * sizeof(double), hipMemcpyHostToDevice, stream_));
    HIP_CHECK(hipStreamSynchronize(stream_));
    splines_ready_ = true;
}

void Engine::form_factors(int n_atoms, const int* type, const double* lcs, int n_h, const double* h_cart,
                          double* f)
{
    HIP_CHECK(hipSetDevice(device_));
    if (n_atoms <= 0 || n_h <= 0) return;
    if (model_.n_types <= 0) throw Error(1, "form factors need atom types");
    for (int i = 0; i < n_atoms; i++)
        if (type[i] < 0 || type[i] >= model_.n_types) throw Error(1, "atom type index out of range");
    if (!splines_ready_) build_splines();
    // atoms in caller order, padded to whole waves; P_lm to the global frame by k_atom_prep
    const int n_slots = (n_atoms + kWave - 1) / kWave * kWave, n_waves = n_slots / kWave;
    std::vector<AtomIn> ain(n_slots);
    for (int s = 0; s < n_slots; s++) {
        AtomIn& r = ain[s];
        std::memset(&r, 0, sizeof(r));
        if (s < n_atoms) {
            r.orig = s;
            r.type = type[s];
            for (int j = 0; j < 9; j++) r.lcs[j] = lcs[9 * (size_t)s + j];
        } else {
            r.orig = -1;
            r.lcs[0] = r.lcs[4] = r.lcs[8] = 1.0;
        }
    }
    d_ff_ain_.ensure(ain.size() * sizeof(AtomIn));
    d_ff_atoms_.ensure(ain.size() * sizeof(AtomRec));
    HIP_CHECK(hipMemcpyAsync(d_ff_ain_.p, ain.data(), ain.size() * sizeof(AtomIn), hipMemcpyHostToDevice, stream_));
    hipLaunchKernelGGL(k_atom_prep, dim3(n_slots), dim3(64), 0, stream_, d_ff_ain_.as<AtomIn>(), d_type_.as<TypeRec>(),
                       d_ff_atoms_.as<AtomRec>(), n_slots);
    HIP_CHECK(hipGetLastError());
    // blocks of reflections bounded by a 256 MB output buffer
    const size_t row = (size_t)n_atoms * 2 * sizeof(double);
    const int B = (int)std::max<size_t>(1, std::min<size_t>((size_t)n_h, (256u << 20) / row));
    const int nt = model_.n_types;
    d_ff_h_.ensure((size_t)B * 3 * sizeof(double));
    d_ff_rec_.ensure((size_t)B * nt * kRadRec * sizeof(double));
    d_ff_hrec_.ensure((size_t)B * 26 * sizeof(double));
    d_ff_out_.ensure((size_t)B * row);
    float ff_ms = 0.f;
    for (int b0 = 0; b0 < n_h; b0 += B) {
        const int n = std::min(B, n_h - b0);
        HIP_CHECK(hipMemcpyAsync(d_ff_h_.p, h_cart + 3 * (size_t)b0, (size_t)n * 3 * sizeof(double),
                                 hipMemcpyHostToDevice, stream_));
        hipLaunchKernelGGL(k_ff_prep, dim3((n * nt + 255) / 256), dim3(256), 0, stream_, d_ff_h_.as<double>(), n,
                           d_type_.as<TypeRec>(), nt, model_.n_wfn, d_spline_.as<double>(), d_ff_rec_.as<double>(),
                           d_ff_hrec_.as<double>());
        HIP_CHECK(hipGetLastError());
        const int chunk = 8;
        const int n_units = n_waves * ((n + chunk - 1) / chunk);
        if (timing) HIP_CHECK(hipEventRecord(ev0_, stream_));
        hipLaunchKernelGGL(k_ff, dim3((n_units + kWavesPerBlock - 1) / kWavesPerBlock), dim3(256), 0, stream_,
                           d_ff_atoms_.as<AtomRec>(), n_waves, n_atoms, d_ff_hrec_.as<double>(), d_ff_rec_.as<double>(),
                           nt, 0, n, chunk, n_units, d_ff_out_.as<double>());
        HIP_CHECK(hipGetLastError());
        if (timing) HIP_CHECK(hipEventRecord(ev1_, stream_));
        HIP_CHECK(hipMemcpyAsync(f + (size_t)b0 * n_atoms * 2, d_ff_out_.p, (size_t)n * row, hipMemcpyDeviceToHost,
                                 stream_));
        HIP_CHECK(hipStreamSynchronize(stream_));
        if (timing) {
            float ms = 0.f;
            HIP_CHECK(hipEventElapsedTime(&ms, ev0_, ev1_));
            ff_ms += ms;
        }
    }
    if (timing) {
        last_main_ms = ff_ms;
        last_chunk = 8;
        last_grid_x = (n_waves * ((B + 7) / 8) + kWavesPerBlock - 1) / kWavesPerBlock;
        last_grid_y = (n_h + B - 1) / B;     // blocks of reflections
    }
}

// ------------------------------------------------------------------ batched structures

struct BatchItem {
    // prep kernels
    const int* hkl;
    const OpDev* ops;
    const double* c2f;
    double* rec64;
    float* rec32;
    double* phase;
    double* ht;
    double* hlen;
    const WfnRec* wfns;
    const TermRec* core;
    const TermRec* val;
    const TypeRec* types;
    const TabRec* tabs;
    double* tab;
    float* tab32;
    double* rad64;
    float* rad32;
    const AtomIn* ain;
    AtomRec* atoms;
    // stage 2
    const double* mult;
    double* f;
    double* d;
    int n_ops, n_hkl, n_wfn, n_types, n_slots, n_atoms, nd, n_chunks, centric_origin, diag, n_tab;
    int diag64;       // FP64: per-reflection diagonal records (kDiag64Rec) for a compiled sign-pattern body
    int part_f32;     // F partials stored as FP32 (mixed precision)
    int body;         // kernel body of the batched main kernel (BatchBody)
    MainArgs main;
};

// Kernel bodies of the batched step.  Each runs as its own persistent launch over the structures
// that use it (contiguous in the batch order): one kernel holding every body would be allocated
// the registers of the largest and spill in all of them.  The launches run concurrently (forked
// streams), so their tails overlap.  Mixed precision uses every body; FP64 the general body and,
// for 222 structures, the FP64 parity-folded direct body.
enum BatchBody {
    kBodyGeneral = 0,    // general rotations (per-(reflection, operation) records)
    kBodyDiag,           // parity-folded records, sign patterns read at run time
    kBodyPats0,          // parity-folded records, 2/m (b axis) set compiled in
    kBodyPats4,          // ... 222 set
    kNumBodies
};
static_assert(kNumBodies <= 4 && kNumBodies * kXcds * sizeof(int) <= 256,
              "the batch's per-body grids and queue counters are sized for at most four bodies");

// index spaces of the batched launches; offsets are prefix sums over structures.  The main
// kernel's (structure, atom wave, reflection chunk) units follow in kXcds queues, queue x holding
// the chunks c with c % kXcds == x of every structure (kSpUnit + x).
enum BatchSpace { kSpFill = 0, kSpHkl, kSpTab, kSpSlot, kSpRedD, kSpUnit, kNumSpaces = kSpUnit + kXcds };

// structure b with off[b] <= i < off[b + 1] (empty structures are skipped)
__device__ __forceinline__ int find_item(const int* __restrict__ off, int nb, int i)
{
    int lo = 0, hi = nb;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (off[mid] <= i) lo = mid;
        else hi = mid;
    }
    return lo;
}

__global__ void k_batch_fill(const BatchItem* __restrict__ it, const int* __restrict__ off, int nb, int total,
                             int* __restrict__ counter)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < kXcds * kNumBodies) counter[i] = 0;    // work counters of the persistent main kernels' queues
    if (i >= total) return;
    const int b = find_item(off, nb, i);
    it[b].d[i - off[b]] = 0.0;
}

__global__ void k_batch_hkl_prep(const BatchItem* __restrict__ it, const int* __restrict__ off, int nb, int total)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const int b = find_item(off, nb, i);
    const BatchItem& q = it[b];
    const int k = i - off[b];
    if (q.diag) hkl_prep_diag_item(k, q.hkl, q.ops, q.n_ops, q.c2f, q.rec32, q.phase, q.ht, q.hlen);
    else if (q.diag64) hkl_prep_diag64_item(k, q.hkl, q.ops, q.n_ops, q.c2f, q.rec64, q.hlen);
    else hkl_prep_item(k, q.hkl, q.ops, q.n_ops, q.c2f, q.rec64, q.rec32, q.phase, q.hlen);
}



__global__ __launch_bounds__(64) void k_batch_atom_prep(const BatchItem* __restrict__ it, const int* __restrict__ off,
                                                        int nb)
{
    __shared__ double fsh[64];
    const int b = find_item(off, nb, blockIdx.x);
    const BatchItem& q = it[b];
    atom_prep_item(blockIdx.x - off[b], threadIdx.x, fsh, q.ain, q.types, q.atoms);
}

// Persistent: every wave takes (structure, atom wave, reflection chunk) units from the device
// counters until none is left, so uneven structures share the machine without a tail per
// structure and a structure of fewer than four atom waves leaves no wave of a workgroup idle.
// Workgroup b runs on XCD b % kXcds and drains that XCD's queue first: all atom waves of a chunk
// then run on one XCD and its L2 holds the chunk's reflection records once.  A wave whose queue
// is empty takes units from the other queues in turn.  One launch per kernel body (BODY) covers
// the nb structures that use it: `it` and `uoff` start at the first of them, queue x's prefix
// offsets are uoff + x * stride.  In each queue units are ordered costliest structure first, and
// within a structure chunk by chunk.  Radial factors are evaluated on the fly from the
// per-wavefunction tables (no per-(reflection, type) records); reflections are in |h| order, so
// every term is evaluated directly, and the mixed bodies keep each lane's table intervals in
// registers (LinTabCache).  Mixed bodies run at four waves per SIMD (128 VGPRs), which hides the
// table gathers' latency better than three waves without the spills; the body with run-time
// sign patterns would spill at four and runs at three.
template <bool MIXED, bool DERIV, int BODY>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(MIXED ? (BODY == kBodyDiag ? 3 : 4) : 2))) void k_batch_main(
    const BatchItem* __restrict__ it, const int* __restrict__ uoff, int stride, int nb, int* __restrict__ counter)
{
    constexpr int kStageBytes = MIXED ? kStageBytesF32 : kStageBytesF64;
    __shared__ __align__(16) unsigned char stage_lds[kWavesPerBlock * kStageBytes];
    const int wib = (int)(threadIdx.x / kWave);
    const int lane = threadIdx.x & (kWave - 1);
    void* stage = stage_lds + wib * kStageBytes;
    const int home = (int)(blockIdx.x % kXcds);
    for (int qi = 0; qi < kXcds; qi++) {
        const int x = (home + qi) % kXcds;
        const int* off = uoff + (size_t)x * stride;
        const int base = off[0], total = off[nb];
        for (;;) {
            int got = 0;
            if (lane == 0) got = atomicAdd(counter + x, 1);
            const int u = base + __builtin_amdgcn_readfirstlane(__shfl(got, 0));
            if (u >= total) break;
            const int b = __builtin_amdgcn_readfirstlane(find_item(off, nb, u));
            const BatchItem& q = it[b];
            const MainArgs& a = q.main;
            const int local = u - off[b];
            const int c = local / a.n_waves;
            const int wave = local - c * a.n_waves;
            const int chunk_id = c * kXcds + x;
            const int ab = wave / kWavesPerBlock, wb = wave % kWavesPerBlock;
            if constexpr (MIXED && BODY == kBodyDiag)
                main_unit<true, DERIV, 0, true, true>(a, chunk_id, ab, wb, lane, stage);
            else if constexpr (MIXED && BODY == kBodyPats0)
                main_unit<true, DERIV, 0, true, true, 2, kDiag64Pats0 | kDiagDirect>(a, chunk_id, ab, wb, lane, stage);
            else if constexpr (MIXED && BODY == kBodyPats4)
                main_unit<true, DERIV, 0, true, true, 4, kDiag64Pats4 | kDiagDirect>(a, chunk_id, ab, wb, lane, stage);
            else if constexpr (!MIXED && BODY == kBodyPats4)
                main_unit<false, DERIV, 0, false, true, 4, kDiag64Pats4 | kDiagDirect>(a, chunk_id, ab, wb, lane, stage);
            else
                main_unit<MIXED, DERIV, 0, false, true>(a, chunk_id, ab, wb, lane, stage);
        }
    }
}

__global__ void k_batch_reduce_f(const BatchItem* __restrict__ it, const int* __restrict__ off, int nb, int total)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const int b = find_item(off, nb, i);
    const BatchItem& q = it[b];
    if (q.part_f32)
        reduce_f_item(i - off[b], reinterpret_cast<const float2*>(q.main.fpart), q.mult, q.main.n_waves, q.n_hkl,
                      q.centric_origin, q.f);
    else
        reduce_f_item(i - off[b], reinterpret_cast<const double2*>(q.main.fpart), q.mult, q.main.n_waves, q.n_hkl,
                      q.centric_origin, q.f);
}

__global__ void k_batch_reduce_d(const BatchItem* __restrict__ it, const int* __restrict__ off, int nb, int total)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const int b = find_item(off, nb, i);
    const BatchItem& q = it[b];
    reduce_d_item(i - off[b], q.main.dpart, q.atoms, q.n_chunks, q.n_slots, q.nd, q.d);
}

void Engine::batch_item(Precision prec, bool derivs, int chunk, BatchItem* out)
{
    HIP_CHECK(hipSetDevice(device_));
    if (!uploaded_) throw Error(1, "batched step: a structure has no uploaded inputs");
    const bool mixed = prec == kMixed;
    jac_prepped_ = false;     // the batched step rewrites this engine's atom records
    prep_valid_ = false;
    BatchItem& q = *out;
    std::memset(&q, 0, sizeof(q));
    main_args(prec, derivs, false, chunk, &q.main);
    const int n_ops = (int)ops_.size();
    q.hkl = d_hkl_.as<int>();
    q.ops = d_ops_.as<OpDev>();
    q.c2f = reinterpret_cast<const double*>(static_cast<const char*>(d_ops_.p) + n_ops * sizeof(OpDev));
    q.diag = use_diag(mixed, false) ? 1 : 0;
    q.part_f32 = mixed ? 1 : 0;
    // reflection records (rec64/rec32, phase, ht, hlen) live in the batch's own buffer
    q.types = d_type_.as<TypeRec>();
    q.wfns = reinterpret_cast<const WfnRec*>(static_cast<const char*>(d_type_.p) + model_.n_types * sizeof(TypeRec));
    q.core = d_core_.as<TermRec>();
    q.val = d_val_.as<TermRec>();
    q.tabs = d_tabrec_.as<TabRec>();
    q.tab = d_tab_.as<double>();
    q.tab32 = d_tab32_.as<float>();
    q.n_tab = model_.n_types > 0 ? n_tab_nodes_ : 0;
    q.ain = d_atoms_in_.as<AtomIn>();
    q.atoms = d_atoms_.as<AtomRec>();
    q.nd = kNdBasic;
    d_f_.ensure((size_t)std::max(n_hkl_, 1) * 2 * sizeof(double));
    d_d_.ensure((size_t)std::max(n_atoms_, 1) * q.nd * sizeof(double));
    q.f = d_f_.as<double>();
    q.d = d_d_.as<double>();
    q.mult = d_dtdf_.as<double>() + 2 * (size_t)n_hkl_;
    q.n_ops = n_ops;
    q.n_hkl = n_hkl_;
    q.n_wfn = model_.n_types > 0 ? model_.n_wfn : 0;
    q.n_types = model_.n_types;
    q.n_slots = n_sorted_;
    q.n_atoms = n_atoms_;
    q.n_chunks = n_hkl_ > 0 ? (n_hkl_ + chunk - 1) / chunk : 0;
    q.centric_origin = centric_origin_ ? 1 : 0;
    q.body = kBodyGeneral;
    if (q.diag) {
        const int c = diag64_code();
        q.body = c == kDiag64Pats4 ? kBodyPats4 : c == kDiag64Pats0 ? kBodyPats0 : kBodyDiag;
    }
    q.diag64 = 0;
    if (!mixed && diag_ && !tune.general_path) {
        // FP64 222 structures: the parity-folded direct body on per-reflection records (2/m
        // structures measured faster on the general body in the batched step)
        const int c = diag64_code();
        if (c == kDiag64Pats4) {
            q.diag64 = 1;
            q.body = kBodyPats4;
            MainArgs& a = q.main;
            a.diag_mask = 0;
            for (int s = 0; s < n_ops; s++) {
                a.diag_mask |= 1 << diag_pat_[s];
                a.diag_op[diag_pat_[s]] = s;
                a.diag_anom[diag_pat_[s]] = diag_anom_[s];
            }
        }
    }
}

void Engine::mark_batched_run(bool derivs)
{
    ran_ = true;
    last_nd_ = kNdBasic;
    (void)derivs;
}

BatchEngine::BatchEngine(const std::vector<Engine*>& engines) : eng_(engines)
{
    if (eng_.empty()) throw Error(1, "empty batch");
    device_ = eng_[0]->device();
    for (Engine* e : eng_)
        if (!e || e->device() != device_) throw Error(1, "every structure of a batch must use the same device");
    HIP_CHECK(hipSetDevice(device_));
    HIP_CHECK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    HIP_CHECK(hipEventCreate(&ev0_));
    HIP_CHECK(hipEventCreate(&ev1_));
    HIP_CHECK(hipEventCreateWithFlags(&ev_fork_, hipEventDisableTiming));
    for (int i = 0; i < 3; i++) {
        HIP_CHECK(hipStreamCreateWithFlags(&side_[i], hipStreamNonBlocking));
        HIP_CHECK(hipEventCreateWithFlags(&ev_join_[i], hipEventDisableTiming));
    }
    d_counter_.ensure(256);
}

BatchEngine::~BatchEngine()
{
    drop_graph();
    if (ev0_) (void)hipEventDestroy(ev0_);
    if (ev1_) (void)hipEventDestroy(ev1_);
    if (ev_fork_) (void)hipEventDestroy(ev_fork_);
    for (int i = 0; i < 3; i++) {
        if (ev_join_[i]) (void)hipEventDestroy(ev_join_[i]);
        if (side_[i]) (void)hipStreamDestroy(side_[i]);
    }
    if (stream_) (void)hipStreamDestroy(stream_);
}

void BatchEngine::drop_graph()
{
    if (exec_) (void)hipGraphExecDestroy(exec_);
    exec_ = nullptr;
    graph_key_.clear();
}

static const void* batch_main_kernel(bool mixed, bool derivs, int body)
{
#define KB(M, B) (derivs ? (const void*)k_batch_main<M, true, B> : (const void*)k_batch_main<M, false, B>)
    if (!mixed) return body == kBodyPats4 ? KB(false, kBodyPats4) : KB(false, kBodyGeneral);
    switch (body) {
        case kBodyDiag: return KB(true, kBodyDiag);
        case kBodyPats0: return KB(true, kBodyPats0);
        case kBodyPats4: return KB(true, kBodyPats4);
        default: return KB(true, kBodyGeneral);
    }
#undef KB
}

int BatchEngine::chunk_for(Precision prec, bool derivs)
{
    if (hkl_chunk > 0) return hkl_chunk;
    auto it = tuned_.find((prec == kMixed ? 2 : 0) + (derivs ? 1 : 0));
    if (it != tuned_.end()) return it->second;
    return 64;
}

// Build the per-structure descriptors and the prefix offsets of every index space; structures
// are ordered by decreasing main-kernel work so the persistent kernel starts the long ones first.
void BatchEngine::plan(Precision prec, bool derivs, int chunk)
{
    const int nb = (int)eng_.size();
    std::vector<BatchItem> items(nb);
    for (int b = 0; b < nb; b++) eng_[b]->batch_item(prec, derivs, chunk, &items[b]);
    std::vector<int> order(nb);
    for (int b = 0; b < nb; b++) order[b] = b;
    // work of a structure: its atom-wave lanes times reflections times (operations + 1), the 1 for
    // the radial factors evaluated once per (atom, reflection)
    auto work = [&](int b) {
        const BatchItem& q = items[b];
        return (double)q.n_slots * (q.n_ops + 1) * q.n_hkl;
    };
    // structures of one kernel body together (the body with the most work first), costliest
    // structure first within a body
    std::vector<double> body_work(kNumBodies, 0.0);
    for (int b = 0; b < nb; b++) body_work[items[b].body] += work(b);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        const int bx = items[x].body, by = items[y].body;
        if (bx != by) return body_work[bx] != body_work[by] ? body_work[bx] > body_work[by] : bx < by;
        return work(x) > work(y);
    });
    std::vector<BatchItem> sorted(nb);
    for (int b = 0; b < nb; b++) sorted[b] = items[order[b]];
    share_tables(prec, order, sorted);
    // reflection records: one 256-byte aligned slice per structure of the batch's own buffer
    const bool mixed = prec == kMixed;
    auto align = [](size_t x) { return (x + 255) & ~(size_t)255; };
    std::vector<size_t> sz(4 * (size_t)nb);
    size_t rec_total = 0;
    for (int b = 0; b < nb; b++) {
        const BatchItem& q = sorted[b];
        const size_t n_hkl = (size_t)q.n_hkl, n_rec = q.diag ? n_hkl : n_hkl * q.n_ops;
        size_t* z = &sz[4 * (size_t)b];
        z[0] = q.diag64 ? align(n_hkl * kDiag64Rec * sizeof(double))                   // records
                        : align(n_rec * kHklopRec * (mixed ? sizeof(float) : sizeof(double)));
        z[1] = mixed ? align(n_rec * 4 * sizeof(double)) : 0;                           // phases / h
        z[2] = q.diag ? align(n_hkl * q.n_ops * sizeof(double)) : 0;                   // h.t per op
        z[3] = align(n_hkl * sizeof(double));                                           // |h|
        rec_total += z[0] + z[1] + z[2] + z[3];
    }
    d_rec_.ensure(std::max<size_t>(rec_total, 256));
    char* base = static_cast<char*>(d_rec_.p);
    rec_key_.assign({(unsigned long long)(mixed ? 1 : 0), (unsigned long long)(uintptr_t)d_rec_.p,
                     (unsigned long long)nb});
    for (int b = 0; b < nb; b++) {
        BatchItem& q = sorted[b];
        const size_t* z = &sz[4 * (size_t)b];
        q.rec64 = mixed ? nullptr : reinterpret_cast<double*>(base);
        q.rec32 = mixed ? reinterpret_cast<float*>(base) : nullptr;
        q.main.hklop = base;
        base += z[0];
        q.phase = mixed ? reinterpret_cast<double*>(base) : nullptr;
        q.main.phase = q.phase;
        base += z[1];
        q.ht = q.diag ? reinterpret_cast<double*>(base) : nullptr;
        q.main.ht = q.ht;
        base += z[2];
        q.hlen = reinterpret_cast<double*>(base);
        q.main.hlen = q.hlen;
        base += z[3];
        const Engine* e = eng_[order[b]];
        rec_key_.push_back((unsigned long long)(uintptr_t)e);
        rec_key_.push_back(e->hkl_generation());
        rec_key_.push_back((unsigned long long)q.n_hkl);
    }
    off_.assign((size_t)kNumSpaces * (nb + 1), 0);
    total_.assign(kNumSpaces, 0);
    for (int sp = 0; sp < kNumSpaces; sp++) {
        long long acc = 0;
        for (int b = 0; b < nb; b++) {
            off_[(size_t)sp * (nb + 1) + b] = (int)acc;
            const BatchItem& q = sorted[b];
            long long n = 0;
            switch (sp) {
                case kSpFill: n = (long long)q.n_atoms * q.nd; break;
                case kSpHkl: n = q.n_hkl; break;
                case kSpTab: n = q.n_tab; break;
                case kSpSlot: n = q.n_slots; break;
                case kSpRedD: n = derivs ? (long long)q.n_slots * q.nd : 0; break;
                default: {    // queue x = sp - kSpUnit: chunks x, x + kXcds, ...
                    const int x = sp - kSpUnit;
                    const long long nc = q.n_chunks > x ? (q.n_chunks - 1 - x) / kXcds + 1 : 0;
                    n = q.main.n_waves > 0 ? (long long)q.main.n_waves * nc : 0;
                }
            }
            acc += n;
        }
        if (acc > 0x7fffffffLL) throw Error(1, "batch too large for 32-bit index spaces");
        off_[(size_t)sp * (nb + 1) + nb] = (int)acc;
        total_[sp] = (int)acc;
    }
    items_.resize(sizeof(BatchItem) * nb);
    std::memcpy(items_.data(), sorted.data(), items_.size());
    d_items_.ensure(items_.size());
    d_off_.ensure(off_.size() * sizeof(int));
    HIP_CHECK(hipMemcpyAsync(d_items_.p, items_.data(), items_.size(), hipMemcpyHostToDevice, stream_));
    HIP_CHECK(hipMemcpyAsync(d_off_.p, off_.data(), off_.size() * sizeof(int), hipMemcpyHostToDevice, stream_));
    HIP_CHECK(hipStreamSynchronize(stream_));
    last_chunk = chunk;
    last_units = 0;
    for (int x = 0; x < kXcds; x++) last_units += total_[kSpUnit + x];
    // one persistent launch per kernel body over its contiguous range of structures
    groups_.clear();
    for (int b = 0; b < nb; b++) {
        const int bd = sorted[b].body;
        if (groups_.empty() || groups_.back().body != bd) groups_.push_back({bd, b, b, 0});
        BodyGroup& g = groups_.back();
        g.b1 = b + 1;
        for (int x = 0; x < kXcds; x++)
            g.units += off_[(size_t)(kSpUnit + x) * (nb + 1) + b + 1] - off_[(size_t)(kSpUnit + x) * (nb + 1) + b];
    }
}

// Radial tables depend only on the wavefunction data (Slater terms, deformation-valence exponent
// and powers), not on the structure: one table per distinct wavefunction of the batch, as long as
// the longest grid any structure needs, replaces the per-structure tables.  Each structure gets
// its own small TabRec array mapping its wavefunction indices into the shared table.
void BatchEngine::share_tables(Precision prec, const std::vector<int>& order, std::vector<BatchItem>& items)
{
    const int nb = (int)eng_.size();
    std::map<std::vector<double>, int> index;
    std::vector<WfnRec> wr;
    std::vector<TermRec> core, val;
    std::vector<int> nodes;
    std::vector<std::vector<int>> local(nb);
    for (int b = 0; b < nb; b++) {
        const Engine& e = *eng_[order[b]];
        const EngineModel& m = e.model();
        const int nw = e.model().n_types > 0 ? m.n_wfn : 0;
        for (int w = 0; w < nw; w++) {
            std::vector<double> key;
            key.push_back(m.dv_zeta[w]);
            for (int l = 0; l < 5; l++) key.push_back(l < m.dv_nl[w] ? m.dv_n[5 * w + l] : -1);
            for (const SlaterDensity* d : {&m.core[w], &m.valence[w]}) {
                key.push_back((double)d->c.size());
                for (size_t i = 0; i < d->c.size(); i++) {
                    key.push_back(d->c[i]);
                    key.push_back(d->e[i]);
                    key.push_back(d->p[i]);
                }
            }
            auto it = index.find(key);
            int g;
            if (it == index.end()) {
                g = (int)wr.size();
                index[key] = g;
                WfnRec r;
                std::memset(&r, 0, sizeof(r));
                r.core_off = (int)core.size();
                r.core_n = (int)m.core[w].c.size();
                for (size_t i = 0; i < m.core[w].c.size(); i++)
                    core.push_back({m.core[w].c[i], m.core[w].e[i], m.core[w].p[i], 0});
                r.val_off = (int)val.size();
                r.val_n = (int)m.valence[w].c.size();
                for (size_t i = 0; i < m.valence[w].c.size(); i++)
                    val.push_back({m.valence[w].c[i], m.valence[w].e[i], m.valence[w].p[i], 0});
                r.zeta = m.dv_zeta[w];
                for (int l = 0; l < 5; l++) {
                    const bool has = l < m.dv_nl[w];
                    r.n[l] = has ? m.dv_n[5 * w + l] : -1;
                    r.gscale[l] = has ? def_val_scale(m.dv_zeta[w], m.dv_n[5 * w + l], l) : 0.0;
                }
                wr.push_back(r);
                nodes.push_back(0);
            } else {
                g = it->second;
            }
            nodes[g] = std::max(nodes[g], e.tab_nodes(w));
            local[b].push_back(g);
        }
    }
    std::vector<TabRec> gtab(wr.size());
    int total = 0;
    for (size_t g = 0; g < wr.size(); g++) {
        gtab[g].off = total;
        gtab[g].n = nodes[g];
        total += nodes[g];
    }
    n_shared_nodes_ = total;
    n_shared_wfn_ = (int)wr.size();
    std::vector<TabRec> per;      // every structure's map, concatenated
    std::vector<size_t> start(nb);
    for (int b = 0; b < nb; b++) {
        start[b] = per.size();
        for (int g : local[b]) per.push_back(gtab[g]);
        if (local[b].empty()) per.push_back(TabRec{0, 2});
    }
    d_gwfn_.ensure(std::max<size_t>(wr.size(), 1) * sizeof(WfnRec));
    d_gcore_.ensure(std::max<size_t>(core.size(), 1) * sizeof(TermRec));
    d_gval_.ensure(std::max<size_t>(val.size(), 1) * sizeof(TermRec));
    d_gtabs_.ensure(std::max<size_t>(gtab.size(), 1) * sizeof(TabRec));
    d_ltabs_.ensure(std::max<size_t>(per.size(), 1) * sizeof(TabRec));
    d_gtab_.ensure((size_t)std::max(total, 1) * kTabNode * sizeof(double));
    d_gtab32_.ensure((size_t)std::max(total, 1) * kLinRatio * kTabNode * sizeof(float));
    auto up = [&](DeviceBuffer& d, const void* src, size_t bytes) {
        if (bytes) HIP_CHECK(hipMemcpyAsync(d.p, src, bytes, hipMemcpyHostToDevice, stream_));
    };
    up(d_gwfn_, wr.data(), wr.size() * sizeof(WfnRec));
    up(d_gcore_, core.data(), core.size() * sizeof(TermRec));
    up(d_gval_, val.data(), val.size() * sizeof(TermRec));
    up(d_gtabs_, gtab.data(), gtab.size() * sizeof(TabRec));
    up(d_ltabs_, per.data(), per.size() * sizeof(TabRec));
    HIP_CHECK(hipStreamSynchronize(stream_));
    for (int b = 0; b < nb; b++) {
        items[b].main.tabs = d_ltabs_.as<TabRec>() + start[b];
        items[b].main.tab = prec == kMixed ? d_gtab32_.p : d_gtab_.p;
        items[b].n_tab = 0;        // no per-structure table build
    }
}

void BatchEngine::launch(Precision prec, bool derivs)
{
    const bool mixed = prec == kMixed;
    const int nb = (int)eng_.size();
    const BatchItem* it = d_items_.as<BatchItem>();
    const int* off = d_off_.as<int>();
    auto o = [&](int sp) { return off + (size_t)sp * (nb + 1); };
    auto blocks = [](int n, int t) { return dim3((unsigned)std::max(1, (n + t - 1) / t)); };
    const int T = 256;
    hipLaunchKernelGGL(k_batch_fill, blocks(total_[kSpFill], T), dim3(T), 0, stream_, it, o(kSpFill), nb,
                       total_[kSpFill], d_counter_.as<int>());
    if (n_shared_nodes_ > 0) {
        hipLaunchKernelGGL(k_tab_build, dim3(n_shared_nodes_), dim3(64), 0, stream_, d_gtabs_.as<TabRec>(),
                           n_shared_wfn_, d_gwfn_.as<WfnRec>(), d_gcore_.as<TermRec>(), d_gval_.as<TermRec>(),
                           n_shared_nodes_, d_gtab_.as<double>());
        if (prec == kMixed)
            hipLaunchKernelGGL(k_lin_build, blocks(n_shared_nodes_ * kLinRatio * 8, T), dim3(T), 0, stream_,
                               d_gtabs_.as<TabRec>(), n_shared_wfn_, d_gtab_.as<double>(), n_shared_nodes_,
                               d_gtab32_.as<float>());
    }
    if (total_[kSpSlot] > 0)
        hipLaunchKernelGGL(k_batch_atom_prep, dim3(total_[kSpSlot]), dim3(64), 0, stream_, it, o(kSpSlot), nb);
    HIP_CHECK(hipGetLastError());
    last_grid = 0;
    // one persistent launch per kernel body: the first on stream_, the others on side streams
    // forked after the prep, so that a body's workgroups start as those of the previous body
    // retire (the launches' tails overlap); joined back before the reductions
    if (groups_.size() > 4) throw Error(1, "internal: more kernel bodies than side streams");
    HIP_CHECK(hipEventRecord(ev_fork_, stream_));
    bool joined[3] = {false, false, false};
    for (size_t gi = 0; gi < groups_.size(); gi++) {
        const BodyGroup& g = groups_[gi];
        if (g.units == 0) continue;
        hipStream_t s = stream_;
        if (gi > 0) {
            s = side_[gi - 1];
            HIP_CHECK(hipStreamWaitEvent(s, ev_fork_, 0));
        }
        const int grid = std::min(grid_[(mixed ? 2 : 0) + (derivs ? 1 : 0)][g.body], std::max(g.units, kXcds));
        last_grid = std::max(last_grid, grid);
        const BatchItem* itg = it + g.b0;
        const int* uo = o(kSpUnit) + g.b0;
        const int stride = nb + 1, n = g.b1 - g.b0;
        int* cnt = d_counter_.as<int>() + (size_t)g.body * kXcds;
        void* args[] = {(void*)&itg, (void*)&uo, (void*)&stride, (void*)&n, (void*)&cnt};
        HIP_CHECK(hipLaunchKernel(batch_main_kernel(mixed, derivs, g.body), dim3(grid), dim3(256), args, 0, s));
        if (gi > 0) {
            HIP_CHECK(hipEventRecord(ev_join_[gi - 1], s));
            joined[gi - 1] = true;
        }
    }
    for (int i = 0; i < 3; i++)
        if (joined[i]) HIP_CHECK(hipStreamWaitEvent(stream_, ev_join_[i], 0));
    const int R = 64;
    if (total_[kSpHkl] > 0)
        hipLaunchKernelGGL(k_batch_reduce_f, blocks(total_[kSpHkl], R), dim3(R), 0, stream_, it, o(kSpHkl), nb,
                           total_[kSpHkl]);
    if (derivs && total_[kSpRedD] > 0)
        hipLaunchKernelGGL(k_batch_reduce_d, blocks(total_[kSpRedD], R), dim3(R), 0, stream_, it, o(kSpRedD), nb,
                           total_[kSpRedD]);
    HIP_CHECK(hipGetLastError());
}

// The per-reflection records depend on the reflection lists, the operations and the cell only: a
// refinement's steps reuse them, so they are built outside the step's graph when rec_key_ changes.
void BatchEngine::launch_hkl_prep()
{
    const int nb = (int)eng_.size();
    if (total_[kSpHkl] > 0)
        hipLaunchKernelGGL(k_batch_hkl_prep, dim3((unsigned)((total_[kSpHkl] + 255) / 256)), dim3(256), 0, stream_,
                           d_items_.as<BatchItem>(), d_off_.as<int>() + (size_t)kSpHkl * (nb + 1), nb,
                           total_[kSpHkl]);
    HIP_CHECK(hipGetLastError());
}

void BatchEngine::run(Precision prec, bool derivs, bool graph)
{
    HIP_CHECK(hipSetDevice(device_));
    const int chunk = chunk_for(prec, derivs);
    plan(prec, derivs, chunk);
    for (const BodyGroup& bg : groups_) {
        int& g = grid_[(prec == kMixed ? 2 : 0) + (derivs ? 1 : 0)][bg.body];
        if (g == 0) {
            // persistent grid: every workgroup the device holds at once
            int per_cu = 0;
            HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &per_cu, batch_main_kernel(prec == kMixed, derivs, bg.body), 256, 0));
            hipDeviceProp_t prop;
            HIP_CHECK(hipGetDeviceProperties(&prop, device_));
            g = std::max(1, per_cu) * prop.multiProcessorCount;
        }
    }
    std::vector<unsigned char> key(items_);
    key.insert(key.end(), reinterpret_cast<const unsigned char*>(off_.data()),
               reinterpret_cast<const unsigned char*>(off_.data() + off_.size()));
    key.push_back((unsigned char)(prec == kMixed ? 1 : 0));
    key.push_back((unsigned char)(derivs ? 1 : 0));
    if (graph) {
        if (!exec_ || key != graph_key_) {
            drop_graph();
            hipGraph_t g = nullptr;
            HIP_CHECK(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
            launch(prec, derivs);
            HIP_CHECK(hipStreamEndCapture(stream_, &g));
            HIP_CHECK(hipGraphInstantiate(&exec_, g, nullptr, nullptr, 0));
            (void)hipGraphDestroy(g);
            graph_key_ = key;
        }
        HIP_CHECK(hipEventRecord(ev0_, stream_));
        if (rec_key_ != last_rec_key_) launch_hkl_prep();
        HIP_CHECK(hipGraphLaunch(exec_, stream_));
        HIP_CHECK(hipEventRecord(ev1_, stream_));
    } else {
        HIP_CHECK(hipEventRecord(ev0_, stream_));
        if (rec_key_ != last_rec_key_) launch_hkl_prep();
        launch(prec, derivs);
        HIP_CHECK(hipEventRecord(ev1_, stream_));
    }
    last_rec_key_ = rec_key_;
    HIP_CHECK(hipEventSynchronize(ev1_));
    HIP_CHECK(hipEventElapsedTime(&last_ms, ev0_, ev1_));
    for (Engine* e : eng_) e->mark_batched_run(derivs);
}

int BatchEngine::autotune(Precision prec, bool derivs)
{
    HIP_CHECK(hipSetDevice(device_));
    const int saved = hkl_chunk;
    int best = 64;
    float best_ms = 1e30f;
    for (int c : {16, 32, 64, 96, 128, 192, 256, 384, 512}) {
        hkl_chunk = c;
        float ms_min = 1e30f;
        for (int rep = 0; rep < 2; rep++) {
            run(prec, derivs, false);
            ms_min = std::min(ms_min, last_ms);
        }
        if (ms_min < best_ms) {
            best_ms = ms_min;
            best = c;
        }
    }
    hkl_chunk = saved;
    tuned_[(prec == kMixed ? 2 : 0) + (derivs ? 1 : 0)] = best;
    return best;
}

}  // namespace discamb_amd
