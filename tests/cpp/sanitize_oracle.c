/* Sanitizer driver for the oracle's C restatement (oracle/hc_oracle.c, compiled into this file):
 * built with AddressSanitizer + UndefinedBehaviorSanitizer, or ThreadSanitizer with LLVM's OpenMP
 * and its Archer tool, by tests/test_sanitizers.py on the host only.  A small acentric model
 * (three Slater wavefunctions, multipoles to l = 4, one isotropic atom, anomalous terms, an
 * excluded atom) goes through the direct engine, the line-recurrence restatement and the per-hkl
 * engine on four OpenMP threads; the two summed engines must agree.  Prints "ok". */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "../../oracle/hc_oracle.c"

#define NA 9
#define NW 3
#define NT 4

int main(void)
{
    /* wavefunctions: core and valence Slater terms, deformation-valence parameters */
    const int core_off[NW + 1] = {0, 2, 5, 7}, val_off[NW + 1] = {0, 2, 4, 7};
    const double core_c[7] = {20.0, 3.0, 40.0, 6.0, 1.0, 25.0, 4.0}, core_e[7] = {10.0, 6.0, 14.0, 8.0, 5.0, 12.0, 7.0};
    const int core_p[7] = {0, 1, 0, 1, 2, 0, 1};
    const double val_c[7] = {0.5, 0.3, 0.8, 0.2, 0.4, 0.3, 0.1}, val_e[7] = {3.0, 2.0, 4.0, 2.5, 3.5, 2.2, 1.8};
    const int val_p[7] = {2, 2, 2, 3, 2, 2, 3};
    const double dv_zeta[NW] = {3.0, 4.0, 3.5};
    const int dv_n[NW * 5] = {2, 2, 2, 3, 4, 2, 2, 2, 3, 4, 4, 4, 4, 4, 4}, dv_nl[NW] = {5, 5, 5};
    const double anom[NW * 2] = {0.1, 0.05, 0.0, 0.0, -0.2, 0.3};
    double pval[NT], kap[NT], kapp[NT], plm[NT * 25];
    int plm_nl[NT];
    for (int t = 0; t < NT; t++) {
        pval[t] = 4.0 + 0.3 * t;
        kap[t] = 0.97 + 0.02 * t;
        kapp[t] = 0.9 + 0.05 * t;
        plm_nl[t] = t == 3 ? 3 : 5;
        for (int i = 0; i < 25; i++) plm[25 * t + i] = (i < plm_nl[t] * plm_nl[t]) ? 0.05 * sin(1.0 + i + 3 * t) : 0.0;
    }
    oracle_model m = {NW, core_off, val_off, core_c, core_e, core_p, val_c, val_e, val_p, dv_zeta, dv_n, dv_nl, anom,
                      NT, pval, kap, kapp, plm, plm_nl};
    /* atoms in an orthorhombic 7 x 8 x 9 A cell, P2_1 2_1 2_1 (Cartesian operations) */
    const double cell[3] = {7.0, 8.0, 9.0};
    int a_wfn[NA], a_type[NA], adp_n[NA];
    double xyz[3 * NA], adp[6 * NA], occ[NA], mult_w[NA], lcs[9 * NA];
    unsigned char inc[NA];
    int off = 0;
    for (int a = 0; a < NA; a++) {
        a_wfn[a] = a % NW;
        a_type[a] = a % NT;
        for (int j = 0; j < 3; j++) xyz[3 * a + j] = cell[j] * fmod(0.137 * (a + 1) * (j + 2), 1.0);
        adp_n[a] = a == 4 ? 1 : 6;
        if (a == 4) {
            adp[off++] = 2 * M_PI * M_PI * 0.03;
        } else {
            const double u[6] = {0.02, 0.025, 0.03, 0.002, -0.001, 0.0015};
            for (int j = 0; j < 6; j++) adp[off++] = 2 * M_PI * M_PI * u[j] * (1.0 + 0.05 * a);
        }
        occ[a] = 1.0 - 0.03 * a;
        mult_w[a] = 1.0;
        const double c = cos(0.3 * a), s = sin(0.3 * a);
        const double L[9] = {c, -s, 0, s, c, 0, 0, 0, 1};
        for (int j = 0; j < 9; j++) lcs[9 * a + j] = L[j];
        inc[a] = a != 6;
    }
    const double rot[4 * 9] = {1, 0, 0, 0, 1, 0, 0, 0, 1, -1, 0, 0, 0, -1, 0, 0, 0, 1,
                               -1, 0, 0, 0, 1, 0, 0, 0, -1, 1, 0, 0, 0, -1, 0, 0, 0, -1};
    const double tf[4 * 3] = {0, 0, 0, 0.5, 0, 0.5, 0, 0.5, 0.5, 0.5, 0.5, 0};
    double trans[12];
    for (int s = 0; s < 4; s++)
        for (int j = 0; j < 3; j++) trans[3 * s + j] = tf[3 * s + j] * cell[j];
    const double inv_t[3] = {0, 0, 0};
    /* reflections in line order along l, lines over (h, k) */
    int n_hkl = 0, n_lines = 0, line_off[200];
    double hv[3 * 2000], dtdf[2 * 2000];
    for (int h = 0; h <= 4; h++)
        for (int k = 0; k <= 4; k++) {
            line_off[n_lines++] = n_hkl;
            for (int l = 0; l <= 6; l++) {
                hv[3 * n_hkl] = h / cell[0];
                hv[3 * n_hkl + 1] = k / cell[1];
                hv[3 * n_hkl + 2] = l / cell[2];
                dtdf[2 * n_hkl] = cos(0.7 * n_hkl);
                dtdf[2 * n_hkl + 1] = sin(1.3 * n_hkl);
                n_hkl++;
            }
        }
    line_off[n_lines] = n_hkl;
    const double dh[3] = {0, 0, 1.0 / cell[2]};
    double *f1 = calloc(2 * n_hkl, sizeof(double)), *f2 = calloc(2 * n_hkl, sizeof(double));
    double dx1[3 * NA], du1[6 * NA], do1[NA], dx2[3 * NA], du2[6 * NA], do2[NA];
    int rc = hc_oracle_engine(&m, NA, a_wfn, a_type, xyz, adp_n, adp, occ, mult_w, lcs, NULL, inc, 4, rot, trans, 0,
                              inv_t, n_hkl, hv, dtdf, f1, dx1, du1, do1, 4);
    rc |= hc_oracle_engine_lines(&m, NA, a_wfn, a_type, xyz, adp_n, adp, occ, mult_w, lcs, NULL, inc, 4, rot, trans, 0,
                                 inv_t, n_hkl, hv, n_lines, line_off, dh, dtdf, f2, dx2, du2, do2, 4);
    double fmax = 0, df = 0, dmax = 0, dd = 0;
    for (int i = 0; i < 2 * n_hkl; i++) {
        fmax = fmax > fabs(f1[i]) ? fmax : fabs(f1[i]);
        df = df > fabs(f1[i] - f2[i]) ? df : fabs(f1[i] - f2[i]);
    }
    for (int i = 0; i < 3 * NA; i++) {
        dmax = dmax > fabs(dx1[i]) ? dmax : fabs(dx1[i]);
        dd = dd > fabs(dx1[i] - dx2[i]) ? dd : fabs(dx1[i] - dx2[i]);
    }
    double *fr = calloc(2 * n_hkl, sizeof(double)), *rows = calloc((size_t)n_hkl * 10 * NA * 2, sizeof(double));
    double mult[2000];
    for (int k = 0; k < n_hkl; k++) mult[k] = 1.0;
    double anom_atom[2 * NA];
    for (int a = 0; a < NA; a++) {
        anom_atom[2 * a] = 0.1 * cos(a);
        anom_atom[2 * a + 1] = 0.05 * (a % 3);
    }
    rc |= hc_oracle_per_hkl(&m, NA, a_wfn, a_type, xyz, adp_n, adp, occ, mult_w, lcs, anom_atom, inc, 4, rot, trans, 0,
                            inv_t, n_hkl, hv, mult, fr, rows, 4);
    free(f1); free(f2); free(fr); free(rows);
    if (rc != 0 || !(df <= 1e-12 * fmax) || !(dd <= 1e-12 * dmax)) {
        fprintf(stderr, "mismatch rc=%d dF=%g (max %g) dxyz=%g (max %g)\n", rc, df, fmax, dd, dmax);
        return 1;
    }
    printf("ok\n");
    return 0;
}
