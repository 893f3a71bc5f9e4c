/*
 * TEST INFRASTRUCTURE ONLY - FP64 CPU restatement of DiSCaMB's Hansen-Coppens structure-factor
 * and derivative engine.  Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg
 * may load this; the product library never does.
 *
 * It follows HansenCoppens_SF_Engine3::calculateSF (src/Scattering/HansenCoppens_SF_Engine3.cpp:
 * 678-950) and its three serial kernels:
 *   acentric                        calculateSF_SerialAcentric              (.cpp:1030-1509)
 *   centrosymmetric, centre at 0    calculateSF_SerialCentrosymmetric_ordered_hkl2 (.cpp:1511-1992)
 *   centrosymmetric, centre != 0    calculateSF_SerialSymmetryCenterNotAtOrigin (.cpp:2649-2896)
 * with the per-reflection tables of pre_atom_loop_sf_calc (.cpp:319-418) and pre_hkl_loop_sf_calc
 * (.cpp:195-256).  Differences that only change rounding: every phase factor and Debye-Waller
 * factor is evaluated directly (sin/cos/exp per term) instead of by the hkl-line recurrences of
 * scattering_utilities.cpp:520-756, and the deformation-valence term is evaluated in each atom's
 * local frame (calculateDeformationValence with the LCS, Engine3.cpp:153-193) instead of with
 * globally rotated P_lm - the two are equal by rotation invariance of the multipole expansion.
 *
 * Units / conventions of the inputs are those Engine3 receives: Cartesian positions (Angstrom),
 * ADPs multiplied by 2*pi^2 (6 U_cart components or 1 U_iso), Cartesian reciprocal vectors,
 * Cartesian symmetry operations, dT/dF already conjugated and multiplied by the centering factor.
 * Outputs: F (complex, interleaved), dT/dxyz (Cartesian), dT/dU (U_cart, 6 or 1), dT/docc.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../discamb_amd/csrc/hc_tables.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

typedef struct {
    int n_wfn;
    const int *core_off, *val_off;       /* n_wfn+1 offsets into the term arrays */
    const double *core_c, *core_e;
    const int *core_p;
    const double *val_c, *val_e;
    const int *val_p;
    const double *dv_zeta;               /* n_wfn */
    const int *dv_n;                     /* n_wfn*5 */
    const int *dv_nl;                    /* n_wfn: number of deformation-valence powers */
    const double *anom;                  /* n_wfn*2 (f', f'') */
    int n_types;
    const double *pval, *kappa, *kappa_p;
    const double *plm;                   /* n_types*25, density normalised, local frame */
    const int *plm_nl;                   /* n_types: size of p_lm (number of l values) */
} oracle_model;

/* Optional replacements used by the independent-atom and electron-scattering restatements:
 *   core override  - f_core of wavefunction w at reflection k taken from core_f[k * n_wfn + w]
 *                    (the N-Gaussian form factors of HansenCoppens_SF_Engine::calculateSF_IAM);
 *   electron       - Mott-Bethe transform of every term's atomic form factor
 *                    (HansenCoppens_SF_Engine3.cpp:1387-1398) with per-atom Z and f_e(0). */
static const double *g_core_f = NULL;
static int g_core_nwfn = 0;
static const double *g_el_z = NULL, *g_el_f0 = NULL;

void hc_oracle_set_core_override(const double *core_f, int n_wfn) { g_core_f = core_f; g_core_nwfn = n_wfn; }
void hc_oracle_set_electron(const double *z, const double *f0) { g_el_z = z; g_el_f0 = f0; }

static void electron_ff(int a, double hl2, int h000, double anr, double ani, double *ar, double *ai)
{
    if (!g_el_z) return;
    if (h000) {
        *ar = g_el_f0[a] + anr;
        *ai = ani;
        return;
    }
    double s2 = 0.25 * hl2;
    double nuc = 0.023934 * g_el_z[a] / s2;
    double er = -0.023934 * (*ar - anr) / s2, ei = -0.023934 * (*ai - ani) / s2;
    *ar = nuc + er + anr;
    *ai = ei + ani;
}

/* int_0^inf r^n exp(-Z r) j_l(K r) dr, K = 2 pi h */
static double g_fun(int l, int n, double h, double z)
{
    double k = 2.0 * M_PI * h;
    int off = kHcGOffset[l][n];
    double d = k * k + z * z, num = 0.0, kp = 1.0;
    double zp[16];
    zp[0] = 1.0;
    for (int i = 1; i < n; i++) zp[i] = zp[i - 1] * z;
    for (int i = 0; i < n; i++) {
        num += kHcGCoef[off + i] * kp * zp[n - 1 - i];
        kp *= k;
    }
    return num / pow(d, n);
}

static double sph_density_sf(const double *c, const double *e, const int *p, int n, double h)
{
    double s = 0.0;
    for (int k = 0; k < n; k++) s += c[k] * g_fun(0, p[k] + 2, h, e[k]);
    return s * 4.0 * M_PI;
}

static void ylm_density(const double *u, double *d)
{
    double x = u[0], y = u[1], z = u[2];
    double xx = x * x, yy = y * y, zz = z * z;
    double w[25];
    w[0] = 1.0;
    w[1] = y; w[2] = z; w[3] = x;
    w[4] = x * y; w[5] = y * z; w[6] = 2 * zz - xx - yy; w[7] = x * z; w[8] = xx - yy;
    w[9] = (3 * xx - yy) * y; w[10] = x * y * z; w[11] = y * (4 * zz - xx - yy);
    w[12] = z * (2 * zz - 3 * xx - 3 * yy); w[13] = x * (4 * zz - xx - yy); w[14] = (xx - yy) * z;
    w[15] = (xx - 3 * yy) * x;
    w[16] = x * y * (xx - yy); w[17] = (3 * xx - yy) * y * z; w[18] = x * y * (7 * zz - 1);
    w[19] = y * z * (7 * zz - 3); w[20] = 35 * zz * zz - 30 * zz + 3; w[21] = x * z * (7 * zz - 3);
    w[22] = (xx - yy) * (7 * zz - 1); w[23] = (xx - 3 * yy) * x * z;
    w[24] = xx * (xx - 3 * yy) - yy * (3 * xx - yy);
    for (int i = 0; i < 25; i++) d[i] = w[i] * kHcDensityNorm[i];
}

/* 4 pi [ g0 A0 - g2 A2 + g4 A4 + i (g1 A1 - g3 A3) ],  A_l = sum_m P_lm d_lm(u_local)
   (combine_multipolar_terms, HansenCoppens_SF_Engine3.h:986-1074) */
static void def_val(const double *plm, int max_l, const double *g, const double *u_local,
                    double *re, double *im)
{
    double d[25], a[5] = {0, 0, 0, 0, 0};
    *re = *im = 0.0;
    if (max_l < 0) return;
    ylm_density(u_local, d);
    for (int l = 0; l <= max_l; l++)
        for (int i = l * l; i < (l + 1) * (l + 1); i++) a[l] += plm[i] * d[i];
    double r = g[0] * a[0], m = 0.0;
    if (max_l >= 1) m += g[1] * a[1];
    if (max_l >= 2) r -= g[2] * a[2];
    if (max_l >= 3) m -= g[3] * a[3];
    if (max_l >= 4) r += g[4] * a[4];
    *re = 4.0 * M_PI * r;
    *im = 4.0 * M_PI * m;
}

/* returns 0 on success */
int hc_oracle_engine(
    const oracle_model *mdl,
    int n_atoms, const int *atom_wfn, const int *atom_type, const double *xyz, const int *adp_n,
    const double *adp, const double *occ, const double *mult_w, const double *lcs,
    const double *anom_atom, const unsigned char *include,
    int n_ops, const double *rot, const double *trans, int centro, const double *inv_t,
    int n_hkl, const double *hv, const double *dtdf,
    double *f_out, double *d_xyz, double *d_adp, double *d_occ, int n_threads)
{
    const double two_pi = 2.0 * M_PI, two_pi2 = 2.0 * M_PI * M_PI;
    /* type -> wfn map and max L (pre_hkl_loop_sf_calc, Engine3.cpp:195-256) */
    int *type_wfn = (int *)calloc(mdl->n_types > 0 ? mdl->n_types : 1, sizeof(int));
    int *type_max_l = (int *)calloc(mdl->n_types > 0 ? mdl->n_types : 1, sizeof(int));
    for (int a = 0; a < n_atoms; a++) type_wfn[atom_type[a]] = atom_wfn[a];
    for (int t = 0; t < mdl->n_types; t++) {
        int lp = -1;
        for (int l = 0; l < mdl->plm_nl[t]; l++)
            for (int i = l * l; i < (l + 1) * (l + 1); i++)
                if (mdl->plm[25 * t + i] != 0.0) lp = l;
        int lw = mdl->dv_nl[type_wfn[t]] - 1;
        int ml = lw < 4 ? lw : 4;
        type_max_l[t] = ml < lp ? ml : lp;
    }
    int centric_origin = centro && inv_t[0] == 0.0 && inv_t[1] == 0.0 && inv_t[2] == 0.0;
    int centric_shift = centro && !centric_origin;
    int n_adp_tot = 0;
    int *adp_off = (int *)malloc(sizeof(int) * (n_atoms + 1));
    for (int a = 0; a < n_atoms; a++) { adp_off[a] = n_adp_tot; n_adp_tot += adp_n[a]; }
    adp_off[n_atoms] = n_adp_tot;
    if (n_threads < 1) n_threads = 1;
#ifndef _OPENMP
    n_threads = 1;
#endif
    /* per-thread derivative accumulators: [thread][atom][3 + 6 + 1] */
    double *acc = (double *)calloc((size_t)n_threads * n_atoms * 10, sizeof(double));

#pragma omp parallel num_threads(n_threads)
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        double *my = acc + (size_t)tid * n_atoms * 10;
        double *fcore = (double *)malloc(sizeof(double) * (mdl->n_wfn + 1));
        double *fval = (double *)malloc(sizeof(double) * (mdl->n_types + 1));
        double *gl = (double *)malloc(sizeof(double) * 5 * (mdl->n_types + 1));
#pragma omp for schedule(dynamic, 16)
        for (int k = 0; k < n_hkl; k++) {
            const double *h = hv + 3 * k;
            double hl2 = h[0] * h[0] + h[1] * h[1] + h[2] * h[2];
            double hl = sqrt(hl2);
            int h000 = hl < 1e-10;
            double cr = dtdf[2 * k], ci = dtdf[2 * k + 1];
            for (int w = 0; w < mdl->n_wfn; w++)
                fcore[w] = g_core_f ? g_core_f[(size_t)k * g_core_nwfn + w]
                                    : sph_density_sf(mdl->core_c + mdl->core_off[w], mdl->core_e + mdl->core_off[w],
                                                     mdl->core_p + mdl->core_off[w],
                                                     mdl->core_off[w + 1] - mdl->core_off[w], hl);
            for (int t = 0; t < mdl->n_types; t++) {
                int w = type_wfn[t];
                fval[t] = sph_density_sf(mdl->val_c + mdl->val_off[w], mdl->val_e + mdl->val_off[w],
                                         mdl->val_p + mdl->val_off[w],
                                         mdl->val_off[w + 1] - mdl->val_off[w], hl / mdl->kappa[t]);
                for (int l = 0; l < 5; l++) {
                    gl[5 * t + l] = 0.0;
                    if (l < mdl->dv_nl[w]) {
                        int nn = mdl->dv_n[5 * w + l];
                        double zeta = mdl->dv_zeta[w];
                        double norm = pow(zeta, nn + 3);
                        for (int q = 2; q <= nn + 2; q++) norm /= q;
                        gl[5 * t + l] = norm * g_fun(l, nn + 2, hl / mdl->kappa_p[t], zeta);
                    }
                }
            }
            double fr_tot = 0.0, fi_tot = 0.0;
            double inv_re = 1.0, inv_im = 0.0;
            if (centric_shift) {
                double ph = two_pi * (h[0] * inv_t[0] + h[1] * inv_t[1] + h[2] * inv_t[2]);
                inv_re = cos(ph);
                inv_im = sin(ph);
            }
            for (int a = 0; a < n_atoms; a++) {
                if (!include[a]) continue;
                int w = atom_wfn[a], t = atom_type[a];
                double weight = occ[a] * mult_w[a];
                double f_re = fcore[w] + fval[t] * mdl->pval[t];
                double f_im = 0.0;
                if (anom_atom) { f_re += anom_atom[2 * a]; f_im += anom_atom[2 * a + 1]; }
                else { f_re += mdl->anom[2 * w]; f_im += mdl->anom[2 * w + 1]; }
                const double *x = xyz + 3 * a;
                const double *L = lcs + 9 * a;
                const double *U = adp + adp_off[a];
                int nu = adp_n[a];
                double sr = 0, si = 0;                 /* sum over ops of X */
                double xr[3] = {0, 0, 0}, xi[3] = {0, 0, 0};
                double ur[6] = {0, 0, 0, 0, 0, 0}, ui[6] = {0, 0, 0, 0, 0, 0};
                for (int s = 0; s < n_ops; s++) {
                    const double *R = rot + 9 * s, *tt = trans + 3 * s;
                    double hs[3];
                    for (int j = 0; j < 3; j++) hs[j] = h[0] * R[j] + h[1] * R[3 + j] + h[2] * R[6 + j];
                    double phase = two_pi * (hs[0] * x[0] + hs[1] * x[1] + hs[2] * x[2] +
                                             h[0] * tt[0] + h[1] * tt[1] + h[2] * tt[2]);
                    double pr = cos(phase), pi_ = sin(phase);
                    double dre = 0.0, dim = 0.0;
                    if (!h000) {
                        double u[3] = {hs[0] / hl, hs[1] / hl, hs[2] / hl};
                        double ul[3];
                        for (int j = 0; j < 3; j++) ul[j] = L[j] * u[0] + L[3 + j] * u[1] + L[6 + j] * u[2];
                        def_val(mdl->plm + 25 * t, type_max_l[t], gl + 5 * t, ul, &dre, &dim);
                    }
                    double ar = f_re + dre, ai = f_im + dim;
                    electron_ff(a, hl2, h000, anom_atom ? anom_atom[2 * a] : mdl->anom[2 * w],
                                anom_atom ? anom_atom[2 * a + 1] : mdl->anom[2 * w + 1], &ar, &ai);
                    double m6[6] = {hs[0] * hs[0], hs[1] * hs[1], hs[2] * hs[2],
                                    2 * hs[0] * hs[1], 2 * hs[0] * hs[2], 2 * hs[1] * hs[2]};
                    double T = 1.0;
                    if (nu == 6 || centric_shift) {
                        if (nu == 6)
                            T = exp(-(m6[0] * U[0] + m6[1] * U[1] + m6[2] * U[2] + m6[3] * U[3] + m6[4] * U[4] +
                                      m6[5] * U[5]));
                        else if (nu == 1)
                            T = exp(-hl2 * U[0]);
                    }
                    if (!centric_shift) {
                        /* X = f T e^{i phi} */
                        double Xr = T * (ar * pr - ai * pi_), Xi = T * (ar * pi_ + ai * pr);
                        sr += Xr; si += Xi;
                        if (nu == 6)
                            for (int j = 0; j < 6; j++) { ur[j] -= m6[j] * Xr; ui[j] -= m6[j] * Xi; }
                        for (int j = 0; j < 3; j++) { xr[j] += hs[j] * Xr; xi[j] += hs[j] * Xi; }
                    } else {
                        /* term1 = f e^{i phi}; term2 = conj(f) e^{-i phi} e^{2 pi i h.t_inv} */
                        double t1r = ar * pr - ai * pi_, t1i = ar * pi_ + ai * pr;
                        double cfr = ar * pr - ai * pi_, cfi = -(ar * pi_ + ai * pr); /* conj(f e^{i phi}) */
                        double t2r = cfr * inv_re - cfi * inv_im, t2i = cfr * inv_im + cfi * inv_re;
                        double Xr = T * (t1r + t2r), Xi = T * (t1i + t2i);
                        sr += Xr; si += Xi;
                        /* T (term1 - term2) 2 pi i */
                        double dr = T * (t1r - t2r), di = T * (t1i - t2i);
                        double mr = -two_pi * di, mi = two_pi * dr;
                        for (int j = 0; j < 3; j++) { xr[j] += hs[j] * mr; xi[j] += hs[j] * mi; }
                        if (nu == 6)
                            for (int j = 0; j < 6; j++) { ur[j] -= m6[j] * Xr; ui[j] -= m6[j] * Xi; }
                    }
                }
                double *da = my + 10 * (size_t)a;
                if (!centric_shift) {
                    double T_iso = 1.0;
                    if (nu == 1) {
                        T_iso = exp(-hl2 * U[0]);
                        sr *= T_iso; si *= T_iso;
                    }
                    double auxr = cr, auxi = ci;
                    if (nu == 1) { auxr *= T_iso; auxi *= T_iso; }
                    if (!centro) {
                        if (nu == 1) da[3] -= hl2 * (cr * sr - ci * si);
                        else if (nu == 6)
                            for (int j = 0; j < 6; j++) da[3 + j] += cr * ur[j] - ci * ui[j];
                        da[9] += sr * cr - si * ci;
                        for (int j = 0; j < 3; j++) da[j] -= auxr * xi[j] + auxi * xr[j];
                        fr_tot += sr * weight;
                        fi_tot += si * weight;
                    } else {
                        if (nu == 1) da[3] -= 2 * hl2 * (cr * sr);
                        else if (nu == 6)
                            for (int j = 0; j < 6; j++) da[3 + j] += 2 * cr * ur[j];
                        da[9] += 2 * sr * cr;
                        for (int j = 0; j < 3; j++) da[j] -= 2 * auxr * xi[j];
                        fr_tot += sr * weight;
                    }
                } else {
                    double aur = cr * weight, aui = ci * weight;
                    da[9] += (sr * cr - si * ci) * mult_w[a];
                    fr_tot += sr * weight;
                    fi_tot += si * weight;
                    if (nu == 1) da[3] += -hl2 * two_pi2 * (sr * aur - si * aui);
                    else if (nu == 6)
                        for (int j = 0; j < 6; j++) da[3 + j] += two_pi2 * (aur * ur[j] - aui * ui[j]);
                    for (int j = 0; j < 3; j++) da[j] += aur * xr[j] - aui * xi[j];
                }
            }
            if (centric_origin) {
                f_out[2 * k] = 2 * fr_tot;
                f_out[2 * k + 1] = 0.0;
            } else {
                f_out[2 * k] = fr_tot;
                f_out[2 * k + 1] = fi_tot;
            }
        }
        free(fcore);
        free(fval);
        free(gl);
    }
    /* merge threads, final scaling (Engine3.cpp:1488-1507 / 1968-1991) */
    for (int a = 0; a < n_atoms; a++) {
        double s[10] = {0};
        for (int t = 0; t < n_threads; t++)
            for (int j = 0; j < 10; j++) s[j] += acc[((size_t)t * n_atoms + a) * 10 + j];
        double weight = occ[a] * mult_w[a];
        if (!centric_shift) {
            s[9] *= mult_w[a];
            for (int j = 3; j < 9; j++) s[j] *= two_pi2 * weight;
            for (int j = 0; j < 3; j++) s[j] *= two_pi * weight;
        }
        for (int j = 0; j < 3; j++) d_xyz[3 * a + j] = s[j];
        for (int j = 0; j < adp_n[a]; j++) d_adp[adp_off[a] + j] = s[3 + j];
        d_occ[a] = s[9];
    }
    free(acc);
    free(adp_off);
    free(type_wfn);
    free(type_max_l);
    return 0;
}

/* Timing restatement of Engine3's acentric path with its hkl-line recurrences, the CPU baseline of
 * the benchmark (its results are checked against hc_oracle_engine by the tests).  It follows the
 * reference's structure rather than the oracle's direct evaluation:
 *   - P_lm rotated once per atom to the global frame (calculateGlobalCoordinatesPlm,
 *     HansenCoppens_SF_Engine3.cpp:594-654; here by an exact projection quadrature: 5 Gauss-Legendre
 *     nodes in cos(theta) x 10 in phi integrate degree-8 products of harmonics exactly);
 *   - per reflection and operation the harmonics, Debye-Waller multipliers and per type the radial
 *     terms once (pre_atom_loop_sf_calc, .cpp:319-418);
 *   - reflections along lines with the phase factor and Debye-Waller factor carried by products
 *     (init_line_multipliers / calculate_line_phase_factors / calculate_line_temperature_factors,
 *     scattering_utilities.cpp:520-756, used at Engine3.cpp:1258-1288), direct evaluation at each
 *     line start; OpenMP over lines.
 * Reflections arrive in line order: line i is [line_off[i], line_off[i+1]) and consecutive
 * reflections of a line differ by the Cartesian step dh.  Acentric groups only (no inversion
 * folding), no electron or core override. */
/* g_fun with the power of the denominator by repeated multiplication (n <= 12) */
static double g_fun_ipow(int l, int n, double h, double z)
{
    const double k = 2.0 * M_PI * h;
    const int off = kHcGOffset[l][n];
    const double d = k * k + z * z;
    double num = 0.0, kp = 1.0, zp[16], dn = 1.0;
    zp[0] = 1.0;
    for (int i = 1; i < n; i++) zp[i] = zp[i - 1] * z;
    for (int i = 0; i < n; i++) {
        num += kHcGCoef[off + i] * kp * zp[n - 1 - i];
        kp *= k;
        dn *= d;
    }
    return num / dn;
}

static double sph_density_sf_ipow(const double *c, const double *e, const int *p, int n, double h)
{
    double s = 0.0;
    for (int k = 0; k < n; k++) s += c[k] * g_fun_ipow(0, p[k] + 2, h, e[k]);
    return s * 4.0 * M_PI;
}

static void legendre5(double *x, double *w)
{
    const double a = 0.5384693101056831, b = 0.9061798459386640;
    const double wa = 0.4786286704993665, wb = 0.2369268850561891, w0 = 0.5688888888888889;
    x[0] = -b; x[1] = -a; x[2] = 0.0; x[3] = a; x[4] = b;
    w[0] = wb; w[1] = wa; w[2] = w0; w[3] = wa; w[4] = wb;
}

/* P_glob = D P_loc, D_ij = sum_q w_q d_i(u_q) d_j(L^T u_q) / sum_q w_q d_i(u_q)^2 (same l only) */
static void global_plm(const double *plm, int max_l, const double *L, double *pg)
{
    double x[5], wx[5];
    legendre5(x, wx);
    double num[25] = {0}, den[25] = {0};
    for (int it = 0; it < 5; it++)
        for (int ip = 0; ip < 10; ip++) {
            const double ct = x[it], st = sqrt(1.0 - ct * ct), ph = 2.0 * M_PI * ip / 10.0;
            const double u[3] = {st * cos(ph), st * sin(ph), ct};
            double ul[3], dg[25], dl[25];
            for (int j = 0; j < 3; j++) ul[j] = L[j] * u[0] + L[3 + j] * u[1] + L[6 + j] * u[2];
            ylm_density(u, dg);
            ylm_density(ul, dl);
            double rho[5] = {0, 0, 0, 0, 0};
            for (int l = 0; l <= max_l; l++)
                for (int i = l * l; i < (l + 1) * (l + 1); i++) rho[l] += plm[i] * dl[i];
            for (int l = 0; l <= 4; l++)
                for (int i = l * l; i < (l + 1) * (l + 1); i++) {
                    num[i] += wx[it] * dg[i] * rho[l];
                    den[i] += wx[it] * dg[i] * dg[i];
                }
        }
    for (int i = 0; i < 25; i++) pg[i] = num[i] / den[i];
}

int hc_oracle_engine_lines(
    const oracle_model *mdl,
    int n_atoms, const int *atom_wfn, const int *atom_type, const double *xyz, const int *adp_n,
    const double *adp, const double *occ, const double *mult_w, const double *lcs,
    const double *anom_atom, const unsigned char *include,
    int n_ops, const double *rot, const double *trans, int centro, const double *inv_t,
    int n_hkl, const double *hv, int n_lines, const int *line_off, const double *dh, const double *dtdf,
    double *f_out, double *d_xyz, double *d_adp, double *d_occ, int n_threads)
{
    (void)inv_t;
    if (centro || g_core_f || g_el_z) return 1;
    const double two_pi = 2.0 * M_PI, two_pi2 = 2.0 * M_PI * M_PI, four_pi = 4.0 * M_PI;
    int *type_wfn = (int *)calloc(mdl->n_types > 0 ? mdl->n_types : 1, sizeof(int));
    int *type_max_l = (int *)calloc(mdl->n_types > 0 ? mdl->n_types : 1, sizeof(int));
    for (int a = 0; a < n_atoms; a++) type_wfn[atom_type[a]] = atom_wfn[a];
    for (int t = 0; t < mdl->n_types; t++) {
        int lp = -1;
        for (int l = 0; l < mdl->plm_nl[t]; l++)
            for (int i = l * l; i < (l + 1) * (l + 1); i++)
                if (mdl->plm[25 * t + i] != 0.0) lp = l;
        int lw = mdl->dv_nl[type_wfn[t]] - 1;
        int ml = lw < 4 ? lw : 4;
        type_max_l[t] = ml < lp ? ml : lp;
    }
    /* per atom: global P_lm, 6-component ADP (U_iso as U I), line constants per operation */
    double *pg = (double *)calloc((size_t)n_atoms * 25 + 1, sizeof(double));
    double *u6 = (double *)calloc((size_t)n_atoms * 6 + 1, sizeof(double));
    int off = 0;
    for (int a = 0; a < n_atoms; a++) {
        const int t = atom_type[a];
        if (type_max_l[t] >= 0) global_plm(mdl->plm + 25 * t, type_max_l[t], lcs + 9 * a, pg + 25 * a);
        if (adp_n[a] == 6)
            for (int j = 0; j < 6; j++) u6[6 * a + j] = adp[off + j];
        else if (adp_n[a] == 1)
            for (int j = 0; j < 3; j++) u6[6 * a + j] = adp[off];
        off += adp_n[a];
    }
    /* Delta in each operation's frame and m(Delta_s) */
    double *dhs = (double *)malloc(sizeof(double) * 3 * (n_ops + 1)), *dm = (double *)malloc(sizeof(double) * 6 * (n_ops + 1));
    for (int s = 0; s < n_ops; s++) {
        const double *R = rot + 9 * s;
        double *d = dhs + 3 * s;
        for (int j = 0; j < 3; j++) d[j] = dh[0] * R[j] + dh[1] * R[3 + j] + dh[2] * R[6 + j];
        double *m = dm + 6 * s;
        m[0] = d[0] * d[0]; m[1] = d[1] * d[1]; m[2] = d[2] * d[2];
        m[3] = 2 * d[0] * d[1]; m[4] = 2 * d[0] * d[2]; m[5] = 2 * d[1] * d[2];
    }
    int max_len = 1;
    for (int i = 0; i < n_lines; i++)
        if (line_off[i + 1] - line_off[i] > max_len) max_len = line_off[i + 1] - line_off[i];
    if (n_threads < 1) n_threads = 1;
#ifndef _OPENMP
    n_threads = 1;
#endif
    double *acc = (double *)calloc((size_t)n_threads * n_atoms * 10 + 1, sizeof(double));
#pragma omp parallel num_threads(n_threads)
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        double *my = acc + (size_t)tid * n_atoms * 10;
        /* per-line tables: [k][s] harmonics (25), m (6), h_s (3); [k][w] core; [k][t] valence, g_l */
        double *yt = (double *)malloc(sizeof(double) * (size_t)max_len * n_ops * 34);
        double *fc = (double *)malloc(sizeof(double) * (size_t)max_len * (mdl->n_wfn + 1));
        double *fv = (double *)malloc(sizeof(double) * (size_t)max_len * (mdl->n_types + 1));
        double *gt = (double *)malloc(sizeof(double) * (size_t)max_len * 5 * (mdl->n_types + 1));
        double *fl = (double *)malloc(sizeof(double) * (size_t)max_len * 2);
#pragma omp for schedule(dynamic, 4)
        for (int li = 0; li < n_lines; li++) {
            const int k0 = line_off[li], len = line_off[li + 1] - k0;
            for (int q = 0; q < len; q++) {
                const double *h = hv + 3 * (k0 + q);
                const double hl = sqrt(h[0] * h[0] + h[1] * h[1] + h[2] * h[2]);
                const int h000 = hl < 1e-10;
                for (int s = 0; s < n_ops; s++) {
                    const double *R = rot + 9 * s;
                    double *r = yt + ((size_t)q * n_ops + s) * 34;
                    double hs[3];
                    for (int j = 0; j < 3; j++) hs[j] = h[0] * R[j] + h[1] * R[3 + j] + h[2] * R[6 + j];
                    if (h000) {
                        for (int i = 0; i < 25; i++) r[i] = 0.0;
                    } else {
                        const double u[3] = {hs[0] / hl, hs[1] / hl, hs[2] / hl};
                        ylm_density(u, r);
                    }
                    r[25] = hs[0] * hs[0]; r[26] = hs[1] * hs[1]; r[27] = hs[2] * hs[2];
                    r[28] = 2 * hs[0] * hs[1]; r[29] = 2 * hs[0] * hs[2]; r[30] = 2 * hs[1] * hs[2];
                    r[31] = hs[0]; r[32] = hs[1]; r[33] = hs[2];
                }
                for (int w = 0; w < mdl->n_wfn; w++)
                    fc[(size_t)q * mdl->n_wfn + w] =
                        sph_density_sf_ipow(mdl->core_c + mdl->core_off[w], mdl->core_e + mdl->core_off[w],
                                       mdl->core_p + mdl->core_off[w], mdl->core_off[w + 1] - mdl->core_off[w], hl);
                for (int t = 0; t < mdl->n_types; t++) {
                    const int w = type_wfn[t];
                    fv[(size_t)q * mdl->n_types + t] =
                        sph_density_sf_ipow(mdl->val_c + mdl->val_off[w], mdl->val_e + mdl->val_off[w],
                                       mdl->val_p + mdl->val_off[w], mdl->val_off[w + 1] - mdl->val_off[w],
                                       hl / mdl->kappa[t]);
                    for (int l = 0; l < 5; l++) {
                        double g = 0.0;
                        if (l < mdl->dv_nl[w] && !h000) {
                            const int nn = mdl->dv_n[5 * w + l];
                            const double zeta = mdl->dv_zeta[w];
                            double norm = pow(zeta, nn + 3);
                            for (int qq = 2; qq <= nn + 2; qq++) norm /= qq;
                            g = norm * g_fun_ipow(l, nn + 2, hl / mdl->kappa_p[t], zeta);
                        }
                        gt[((size_t)q * mdl->n_types + t) * 5 + l] = four_pi * g;
                    }
                }
                fl[2 * q] = fl[2 * q + 1] = 0.0;
            }
            for (int a = 0; a < n_atoms; a++) {
                if (!include[a]) continue;
                const int w = atom_wfn[a], t = atom_type[a], ml = type_max_l[t];
                const double weight = occ[a] * mult_w[a];
                const double *x = xyz + 3 * a, *U = u6 + 6 * a, *P = pg + 25 * a;
                const double anr = anom_atom ? anom_atom[2 * a] : mdl->anom[2 * w];
                const double ani = anom_atom ? anom_atom[2 * a + 1] : mdl->anom[2 * w + 1];
                double *da = my + 10 * (size_t)a;
                for (int s = 0; s < n_ops; s++) {
                    const double *tt = trans + 3 * s, *d = dhs + 3 * s, *mD = dm + 6 * s;
                    const double *h0 = hv + 3 * k0;
                    const double *r0 = yt + (size_t)s * 34;
                    /* line start: direct phase factor, Debye-Waller factor and its first ratio */
                    const double ph0 = two_pi * (r0[31] * x[0] + r0[32] * x[1] + r0[33] * x[2] + h0[0] * tt[0] +
                                                 h0[1] * tt[1] + h0[2] * tt[2]);
                    const double e0 = r0[25] * U[0] + r0[26] * U[1] + r0[27] * U[2] + r0[28] * U[3] + r0[29] * U[4] +
                                      r0[30] * U[5];
                    const double T0 = exp(-e0);
                    double er = T0 * cos(ph0), ei = T0 * sin(ph0), q = 1.0;
                    if (len > 1) {
                        const double *r1 = yt + ((size_t)n_ops + s) * 34;
                        const double e1 = r1[25] * U[0] + r1[26] * U[1] + r1[27] * U[2] + r1[28] * U[3] +
                                          r1[29] * U[4] + r1[30] * U[5];
                        q = exp(-(e1 - e0));
                    }
                    const double phm = two_pi * (d[0] * x[0] + d[1] * x[1] + d[2] * x[2] + dh[0] * tt[0] + dh[1] * tt[1] +
                                                 dh[2] * tt[2]);
                    const double mr = cos(phm), mi = sin(phm);
                    const double c = exp(-2.0 * (mD[0] * U[0] + mD[1] * U[1] + mD[2] * U[2] + mD[3] * U[3] +
                                                 mD[4] * U[4] + mD[5] * U[5]));
                    for (int k = 0; k < len; k++) {
                        const double *r = yt + ((size_t)k * n_ops + s) * 34;
                        const double *g = gt + ((size_t)k * mdl->n_types + t) * 5;
                        double A[5] = {0, 0, 0, 0, 0};
                        for (int l = 0; l <= ml; l++)
                            for (int i = l * l; i < (l + 1) * (l + 1); i++) A[l] += P[i] * r[i];
                        const double fr = fc[(size_t)k * mdl->n_wfn + w] + mdl->pval[t] * fv[(size_t)k * mdl->n_types + t] +
                                          anr + g[0] * A[0] - g[2] * A[2] + g[4] * A[4];
                        const double fi = ani + g[1] * A[1] - g[3] * A[3];
                        const double Xr = fr * er - fi * ei, Xi = fr * ei + fi * er;
                        fl[2 * k] += weight * Xr;
                        fl[2 * k + 1] += weight * Xi;
                        const double cr = dtdf[2 * (k0 + k)], ci = dtdf[2 * (k0 + k) + 1];
                        const double zr = cr * Xr - ci * Xi, zi = cr * Xi + ci * Xr;
                        for (int j = 0; j < 3; j++) da[j] -= r[31 + j] * zi;
                        for (int j = 0; j < 6; j++) da[3 + j] -= r[25 + j] * zr;
                        da[9] += zr;
                        /* next reflection of the line */
                        const double qr = q * mr, qi = q * mi;
                        const double nr = er * qr - ei * qi, ni = er * qi + ei * qr;
                        er = nr;
                        ei = ni;
                        q *= c;
                    }
                }
            }
            for (int k = 0; k < len; k++) {
                f_out[2 * (k0 + k)] = fl[2 * k];
                f_out[2 * (k0 + k) + 1] = fl[2 * k + 1];
            }
        }
        free(yt); free(fc); free(fv); free(gt); free(fl);
    }
    off = 0;
    for (int a = 0; a < n_atoms; a++) {
        double sm[10] = {0};
        for (int t = 0; t < n_threads; t++)
            for (int j = 0; j < 10; j++) sm[j] += acc[((size_t)t * n_atoms + a) * 10 + j];
        const double weight = occ[a] * mult_w[a];
        for (int j = 0; j < 3; j++) d_xyz[3 * a + j] = two_pi * weight * sm[j];
        if (adp_n[a] == 6)
            for (int j = 0; j < 6; j++) d_adp[off + j] = two_pi2 * weight * sm[3 + j];
        else if (adp_n[a] == 1)
            d_adp[off] = two_pi2 * weight * (sm[3] + sm[4] + sm[5]);
        off += adp_n[a];
        d_occ[a] = mult_w[a] * sm[9];
    }
    free(acc); free(pg); free(u6); free(dhs); free(dm); free(type_wfn); free(type_max_l);
    return 0;
}

/* Per-reflection structure factors and derivatives: the per-hkl engine that the reference's
 * template calculateStructureFactorsAndDerivatives<T>(atoms, lcs, hkl, processor[, include])
 * dispatches to (HansenCoppens_SF_Engine.h:512-780, Engine v1, not Engine3).  For every
 * reflection k it returns F_k and, per atom a, the complex derivatives
 *   rows[((k * 10 + plane) * n_atoms + a) * 2 + {re, im}]
 *   plane 0-2: dF/dx_cart   = w c 2 pi i sum_s h_s Y_s            (.h:767-770)
 *   plane 3-8: dF/dU_cart   = 2 pi^2 w c sum_s (-m_j(h_s)) X_s     (.h:728-738, 755-762)
 *              U_iso atoms: plane 3 = 2 pi^2 w c sum_s (-|h|^2) X_s; atoms without ADP: 0
 *   plane 9:   dF/docc      = (m/|G|) c sum_s X_s                 (.h:741)
 * with w = occ m/|G| and c the centering multiplier (.h:637), X_s the transformed atomic form
 * factor of operation s and Y_s its position-derivative multiplier:
 *   acentric              X = T (f + f_def) e^{i phi},  Y = X                             (.h:690-699)
 *   centre at t_inv != 0  t1 = (f + f_def) e^{i phi}, t2 = (f + conj f_def) e^{-i phi} e^{2 pi i h.t_inv},
 *                         X = T (t1 + t2),  Y = T (t1 - t2)                               (.h:704-714)
 *   centre at the origin  X = 2 T (f cos phi + Re(f_def e^{i phi})),
 *                         Y = 2 T i (f sin phi + Im(f_def e^{i phi}))                      (.h:716-724)
 * where f = f_core + Pval f_val + f' + i f'' is NOT conjugated for the inversion partner (Engine3
 * conjugates it - the two engines differ for centrosymmetric groups with f'' != 0).  The reference
 * divides by |h| for h = 0 here; this restatement sets the deformation term of h = 0 to zero, as
 * Engine3 does.  Excluded atoms keep zero derivatives and do not enter F (.h:622-623).
 * Inputs as hc_oracle_engine, plus mult = the centering multiplier of every reflection. */
int hc_oracle_per_hkl(
    const oracle_model *mdl,
    int n_atoms, const int *atom_wfn, const int *atom_type, const double *xyz, const int *adp_n,
    const double *adp, const double *occ, const double *mult_w, const double *lcs,
    const double *anom_atom, const unsigned char *include,
    int n_ops, const double *rot, const double *trans, int centro, const double *inv_t,
    int n_hkl, const double *hv, const double *mult,
    double *f_out, double *rows, int n_threads)
{
    const double two_pi = 2.0 * M_PI, two_pi2 = 2.0 * M_PI * M_PI;
    int *type_wfn = (int *)calloc(mdl->n_types > 0 ? mdl->n_types : 1, sizeof(int));
    int *type_max_l = (int *)calloc(mdl->n_types > 0 ? mdl->n_types : 1, sizeof(int));
    for (int a = 0; a < n_atoms; a++) type_wfn[atom_type[a]] = atom_wfn[a];
    /* pre_hkl_loop_sf_calc (HansenCoppens_SF_Engine.cpp:160-221) */
    for (int t = 0; t < mdl->n_types; t++) {
        int lp = -1;
        for (int l = 0; l < mdl->plm_nl[t]; l++)
            for (int i = l * l; i < (l + 1) * (l + 1); i++)
                if (mdl->plm[25 * t + i] != 0.0) lp = l;
        int lw = mdl->dv_nl[type_wfn[t]] - 1;
        int ml = lw < 4 ? lw : 4;
        type_max_l[t] = ml < lp ? ml : lp;
    }
    int centric_origin = centro && inv_t[0] == 0.0 && inv_t[1] == 0.0 && inv_t[2] == 0.0;
    int centric_shift = centro && !centric_origin;
    int *adp_off = (int *)malloc(sizeof(int) * (n_atoms + 1));
    int n_adp_tot = 0;
    for (int a = 0; a < n_atoms; a++) { adp_off[a] = n_adp_tot; n_adp_tot += adp_n[a]; }
    if (n_threads < 1) n_threads = 1;
#ifndef _OPENMP
    n_threads = 1;
#endif
#pragma omp parallel num_threads(n_threads)
    {
        double *fcore = (double *)malloc(sizeof(double) * (mdl->n_wfn + 1));
        double *fval = (double *)malloc(sizeof(double) * (mdl->n_types + 1));
        double *gl = (double *)malloc(sizeof(double) * 5 * (mdl->n_types + 1));
#pragma omp for schedule(dynamic, 16)
        for (int k = 0; k < n_hkl; k++) {
            const double *h = hv + 3 * k;
            double hl2 = h[0] * h[0] + h[1] * h[1] + h[2] * h[2];
            double hl = sqrt(hl2);
            int h000 = hl < 1e-10;
            double *row = rows + (size_t)k * 10 * n_atoms * 2;
            memset(row, 0, sizeof(double) * 10 * (size_t)n_atoms * 2);
            /* pre_atom_loop_sf_calc (HansenCoppens_SF_Engine.cpp:281-377) */
            for (int w = 0; w < mdl->n_wfn; w++)
                fcore[w] = g_core_f ? g_core_f[(size_t)k * g_core_nwfn + w]
                                    : sph_density_sf(mdl->core_c + mdl->core_off[w], mdl->core_e + mdl->core_off[w],
                                                     mdl->core_p + mdl->core_off[w],
                                                     mdl->core_off[w + 1] - mdl->core_off[w], hl);
            for (int t = 0; t < mdl->n_types; t++) {
                int w = type_wfn[t];
                fval[t] = sph_density_sf(mdl->val_c + mdl->val_off[w], mdl->val_e + mdl->val_off[w],
                                         mdl->val_p + mdl->val_off[w],
                                         mdl->val_off[w + 1] - mdl->val_off[w], hl / mdl->kappa[t]);
                for (int l = 0; l < 5; l++) {
                    gl[5 * t + l] = 0.0;
                    if (l < mdl->dv_nl[w]) {
                        int nn = mdl->dv_n[5 * w + l];
                        double zeta = mdl->dv_zeta[w];
                        double norm = pow(zeta, nn + 3);
                        for (int q = 2; q <= nn + 2; q++) norm /= q;
                        gl[5 * t + l] = norm * g_fun(l, nn + 2, hl / mdl->kappa_p[t], zeta);
                    }
                }
            }
            double inv_re = 1.0, inv_im = 0.0;
            if (centric_shift) {
                double ph = two_pi * (h[0] * inv_t[0] + h[1] * inv_t[1] + h[2] * inv_t[2]);
                inv_re = cos(ph);
                inv_im = sin(ph);
            }
            const double c = mult[k];
            double fr_tot = 0.0, fi_tot = 0.0;
            for (int a = 0; a < n_atoms; a++) {
                if (!include[a]) continue;
                int w = atom_wfn[a], t = atom_type[a];
                double weight = occ[a] * mult_w[a] * c;
                double f_re = fcore[w] + fval[t] * mdl->pval[t];
                double f_im = 0.0;
                if (anom_atom) { f_re += anom_atom[2 * a]; f_im += anom_atom[2 * a + 1]; }
                else { f_re += mdl->anom[2 * w]; f_im += mdl->anom[2 * w + 1]; }
                const double *x = xyz + 3 * a;
                const double *L = lcs + 9 * a;
                const double *U = adp + adp_off[a];
                int nu = adp_n[a];
                double sr = 0, si = 0, xr[3] = {0, 0, 0}, xi[3] = {0, 0, 0};
                double ur[6] = {0, 0, 0, 0, 0, 0}, ui[6] = {0, 0, 0, 0, 0, 0};
                for (int s = 0; s < n_ops; s++) {
                    const double *R = rot + 9 * s, *tt = trans + 3 * s;
                    double hs[3];
                    for (int j = 0; j < 3; j++) hs[j] = h[0] * R[j] + h[1] * R[3 + j] + h[2] * R[6 + j];
                    double phase = two_pi * (hs[0] * x[0] + hs[1] * x[1] + hs[2] * x[2] +
                                             h[0] * tt[0] + h[1] * tt[1] + h[2] * tt[2]);
                    double pc = cos(phase), ps = sin(phase);
                    double dre = 0.0, dim = 0.0;
                    if (!h000) {
                        double u[3] = {hs[0] / hl, hs[1] / hl, hs[2] / hl};
                        double ul[3];
                        for (int j = 0; j < 3; j++) ul[j] = L[j] * u[0] + L[3 + j] * u[1] + L[6 + j] * u[2];
                        def_val(mdl->plm + 25 * t, type_max_l[t], gl + 5 * t, ul, &dre, &dim);
                    }
                    double m6[6] = {hs[0] * hs[0], hs[1] * hs[1], hs[2] * hs[2],
                                    2 * hs[0] * hs[1], 2 * hs[0] * hs[2], 2 * hs[1] * hs[2]};
                    double T = 1.0;
                    if (nu == 6)
                        T = exp(-(m6[0] * U[0] + m6[1] * U[1] + m6[2] * U[2] + m6[3] * U[3] + m6[4] * U[4] +
                                  m6[5] * U[5]));
                    else if (nu == 1)
                        T = exp(-hl2 * U[0]);
                    /* (f + f_def) e^{i phi} and f_def e^{i phi} */
                    double er = (f_re + dre) * pc - (f_im + dim) * ps, ei = (f_re + dre) * ps + (f_im + dim) * pc;
                    double Xr, Xi, Yr, Yi;
                    if (!centro) {
                        Xr = T * er; Xi = T * ei; Yr = Xr; Yi = Xi;
                    } else if (centric_shift) {
                        /* (f + conj f_def) e^{-i phi} e^{2 pi i h.t_inv} */
                        double br = (f_re + dre) * pc + (f_im - dim) * ps, bi = (f_im - dim) * pc - (f_re + dre) * ps;
                        double t2r = br * inv_re - bi * inv_im, t2i = br * inv_im + bi * inv_re;
                        Xr = T * (er + t2r); Xi = T * (ei + t2i);
                        Yr = T * (er - t2r); Yi = T * (ei - t2i);
                    } else {
                        double dpr = dre * pc - dim * ps, dpi = dre * ps + dim * pc;   /* f_def e^{i phi} */
                        Xr = 2 * T * (f_re * pc + dpr); Xi = 2 * T * f_im * pc;
                        /* 2 T i (f sin phi + i Im(f_def e^{i phi}) / i) */
                        Yr = -2 * T * f_im * ps; Yi = 2 * T * (f_re * ps + dpi);
                    }
                    sr += Xr; si += Xi;
                    for (int j = 0; j < 3; j++) { xr[j] += hs[j] * Yr; xi[j] += hs[j] * Yi; }
                    if (nu == 1) { ur[0] -= hl2 * Xr; ui[0] -= hl2 * Xi; }
                    else if (nu == 6)
                        for (int j = 0; j < 6; j++) { ur[j] -= m6[j] * Xr; ui[j] -= m6[j] * Xi; }
                }
                fr_tot += weight * sr;
                fi_tot += weight * si;
                /* xyz: w 2 pi i sum h_s Y_s */
                for (int j = 0; j < 3; j++) {
                    double *o = row + ((size_t)j * n_atoms + a) * 2;
                    o[0] = -two_pi * weight * xi[j];
                    o[1] = two_pi * weight * xr[j];
                }
                for (int j = 0; j < nu; j++) {
                    double *o = row + ((size_t)(3 + j) * n_atoms + a) * 2;
                    o[0] = two_pi2 * weight * ur[j];
                    o[1] = two_pi2 * weight * ui[j];
                }
                double *o = row + ((size_t)9 * n_atoms + a) * 2;
                o[0] = sr * mult_w[a] * c;
                o[1] = si * mult_w[a] * c;
            }
            f_out[2 * k] = fr_tot;
            f_out[2 * k + 1] = fi_tot;
        }
        free(fcore);
        free(fval);
        free(gl);
    }
    free(adp_off);
    free(type_wfn);
    free(type_max_l);
    return 0;
}

/* Spherical Slater density form factor sum_k c_k 4 pi g(0, p_k + 2, h, e_k) at m values of h
 * (scatteringSphericalDensity, SlaterTypeOrbitalScattering.cpp:8-21). */
void hc_oracle_sph_density(const double *c, const double *e, const int *p, int n_terms, const double *h, int m,
                           double *out)
{
    for (int i = 0; i < m; i++) out[i] = sph_density_sf(c, e, p, n_terms, h[i]);
}

/* Deformation-valence part of the atomic form factors of the form-factor route
 * (HansenCoppens_SF_Engine::calculateFormFactors, HansenCoppens_SF_Engine.cpp:383-495): for
 * every Cartesian h and atom, 4 pi sum_l i^l N_l g_l(h / kappa') sum_m P_lm d_lm(L^T h/|h|),
 * zero at |h| < 1e-10 and for excluded atoms.  out: [k][atom] complex. */
void hc_oracle_def_val_ff(const oracle_model *mdl, int n_atoms, const int *atom_wfn, const int *atom_type,
                          const double *lcs, const unsigned char *include, int n_h, const double *hv, double *out)
{
    int *type_wfn = (int *)calloc(mdl->n_types > 0 ? mdl->n_types : 1, sizeof(int));
    int *type_max_l = (int *)calloc(mdl->n_types > 0 ? mdl->n_types : 1, sizeof(int));
    for (int a = 0; a < n_atoms; a++) type_wfn[atom_type[a]] = atom_wfn[a];
    for (int t = 0; t < mdl->n_types; t++) {
        int lp = -1;
        for (int l = 0; l < mdl->plm_nl[t]; l++)
            for (int i = l * l; i < (l + 1) * (l + 1); i++)
                if (mdl->plm[25 * t + i] != 0.0) lp = l;
        int lw = mdl->dv_nl[type_wfn[t]] - 1;
        int ml = lw < 4 ? lw : 4;
        type_max_l[t] = ml < lp ? ml : lp;
    }
    double *gl = (double *)malloc(sizeof(double) * 5 * (mdl->n_types + 1));
    for (int k = 0; k < n_h; k++) {
        const double *h = hv + 3 * k;
        double hl = sqrt(h[0] * h[0] + h[1] * h[1] + h[2] * h[2]);
        for (int t = 0; t < mdl->n_types; t++) {
            int w = type_wfn[t];
            for (int l = 0; l < 5; l++) {
                gl[5 * t + l] = 0.0;
                if (l < mdl->dv_nl[w]) {
                    int nn = mdl->dv_n[5 * w + l];
                    double zeta = mdl->dv_zeta[w];
                    double norm = pow(zeta, nn + 3);
                    for (int q = 2; q <= nn + 2; q++) norm /= q;
                    gl[5 * t + l] = norm * g_fun(l, nn + 2, hl / mdl->kappa_p[t], zeta);
                }
            }
        }
        for (int a = 0; a < n_atoms; a++) {
            double *o = out + ((size_t)k * n_atoms + a) * 2;
            o[0] = o[1] = 0.0;
            if (!include[a] || hl < 1e-10) continue;
            const double *L = lcs + 9 * a;
            double u[3] = {h[0] / hl, h[1] / hl, h[2] / hl}, ul[3];
            for (int j = 0; j < 3; j++) ul[j] = L[j] * u[0] + L[3 + j] * u[1] + L[6 + j] * u[2];
            int t = atom_type[a];
            def_val(mdl->plm + 25 * t, type_max_l[t], gl + 5 * t, ul, &o[0], &o[1]);
        }
    }
    free(gl);
    free(type_wfn);
    free(type_max_l);
}

/* test hooks: the oracle's own g-function and harmonics */
double hc_oracle_g(int l, int n, double h, double z) { return g_fun(l, n, h, z); }
void hc_oracle_ylm(const double *u, double *d25) { ylm_density(u, d25); }
