#!/usr/bin/env python3
"""Generate hc_tables.h: constant tables shared by the host library, the HIP kernels and the
FP64 oracle.

1. Density-normalised real spherical harmonics d_lm = N_lm * w_lm(x, y, z) for l <= 4, where
   w_lm are the standard integer-coefficient real-harmonic polynomials in the order m = -l..l
   (l = 1 ordered y, z, x) used by DiSCaMB (real_spherical_harmonics.h:136-280) and N_lm makes
   the integral of |d_lm| over the unit sphere equal to 1 (l = 0) or 2 (l > 0), the Hansen-Coppens
   density normalisation.  N_lm is computed here by exact splitting of |w_lm| at its nodes and
   adaptive quadrature in mpmath (30 digits).

2. Numerator polynomials of the Slater Fourier-Bessel transform
       g(l, n, K, Z) = int_0^inf r^n exp(-Z r) j_l(K r) dr = P_ln(K, Z) / (K^2 + Z^2)^n,
   derived from g(l, l+1) = (2K)^l l! / (K^2+Z^2)^(l+1) and g(l, n+1) = -d g(l, n) / dZ.
   P_ln is homogeneous of degree n-1; it is stored as coefficients of K^i Z^(n-1-i).
   (DiSCaMB tabulates the same closed forms in SlaterTypeOrbitalScattering.h:51-183.)
"""
import os

import mpmath as mp

mp.mp.dps = 30
HERE = os.path.dirname(os.path.abspath(__file__))
LMAX = 4
NMAX = 12


def w_poly(l, i, x, y, z):
    """Real-harmonic polynomial w_l,(i-l) evaluated at (x, y, z); i = l + m."""
    xx, yy, zz = x * x, y * y, z * z
    if l == 0:
        return 1
    if l == 1:
        return (y, z, x)[i]
    if l == 2:
        return (x * y, y * z, 2 * zz - xx - yy, x * z, xx - yy)[i]
    if l == 3:
        return ((3 * xx - yy) * y, x * y * z, y * (4 * zz - xx - yy), z * (2 * zz - 3 * xx - 3 * yy),
                x * (4 * zz - xx - yy), (xx - yy) * z, (xx - 3 * yy) * x)[i]
    return (x * y * (xx - yy), (3 * xx - yy) * y * z, x * y * (7 * zz - 1), y * z * (7 * zz - 3),
            35 * zz * zz - 30 * zz + 3, x * z * (7 * zz - 3), (xx - yy) * (7 * zz - 1),
            (xx - 3 * yy) * x * z, xx * (xx - 3 * yy) - yy * (3 * xx - yy))[i]


def abs_integral(l, i):
    """Integral of |w_l,i| over the unit sphere."""
    m = i - l
    am = abs(m)
    if m == 0:
        phi0 = mp.mpf(0)
        ang = 2 * mp.pi
    elif m > 0:
        phi0 = mp.mpf(0)          # cos(m phi) factor equals 1
        ang = mp.mpf(4)           # int_0^{2pi} |cos(m phi)| dphi
    else:
        phi0 = mp.pi / (2 * am)   # sin(m phi) factor equals 1
        ang = mp.mpf(4)

    def q(z):
        rho = mp.sqrt(1 - z * z)
        return w_poly(l, i, rho * mp.cos(phi0), rho * mp.sin(phi0), z) / rho ** am

    # nodes of q(z) on (-1, 1): sample, then refine sign changes
    pts = [mp.mpf(-1)]
    grid = [mp.mpf(-1) + 2 * mp.mpf(k) / 4000 for k in range(1, 4000)]
    prev = q(grid[0])
    for a, b in zip(grid[:-1], grid[1:]):
        qb = q(b)
        if prev == 0:
            pts.append(a)
        elif prev * qb < 0:
            pts.append(mp.findroot(q, (a, b), solver="bisect"))
        prev = qb
    pts.append(mp.mpf(1))
    f = lambda z: abs(q(z)) * mp.sqrt(1 - z * z) ** am
    return ang * mp.quad(f, pts)


def density_norms():
    out = []
    for l in range(LMAX + 1):
        target = 1 if l == 0 else 2
        for i in range(2 * l + 1):
            out.append(target / abs_integral(l, i))
    return out


def g_polys():
    """P[l][n] = dict {(i, j): coeff} for K^i Z^j."""
    polys = {}
    for l in range(LMAX + 1):
        # g(l, l+1) = (2K)^l l! / d^(l+1)
        fact = 1
        for k in range(2, l + 1):
            fact *= k
        cur = {(l, 0): (2 ** l) * fact}
        m = l + 1
        polys[(l, m)] = cur
        while m < NMAX:
            # -(d/dZ)(P / d^m) = (2 m Z P - d P_Z) / d^(m+1), d = K^2 + Z^2
            nxt = {}
            for (i, j), c in cur.items():
                nxt[(i, j + 1)] = nxt.get((i, j + 1), 0) + 2 * m * c
                if j > 0:
                    dz = c * j
                    nxt[(i + 2, j - 1)] = nxt.get((i + 2, j - 1), 0) - dz
                    nxt[(i, j + 1)] = nxt.get((i, j + 1), 0) - dz
            cur = {k: v for k, v in nxt.items() if v != 0}
            m += 1
            polys[(l, m)] = cur
    return polys


def main():
    norms = density_norms()
    polys = g_polys()
    lines = ["// Generated by gen_tables.py - do not edit.",
             "#pragma once",
             "",
             "// Density normalisation N_lm of real spherical harmonics, [l*l + l + m], l <= 4.",
             "static const double kHcDensityNorm[25] = {"]
    lines.append(",\n".join("    %s" % mp.nstr(v, 20) for v in norms))
    lines.append("};")
    lines.append("")
    lines.append("// g(l, n, K, Z) numerator: sum_i kHcGCoef[kHcGOffset[l][n] + i] * K^i * Z^(n-1-i),")
    lines.append("// denominator (K^2+Z^2)^n; valid for l+1 <= n <= %d." % NMAX)
    offs = [[-1] * (NMAX + 1) for _ in range(LMAX + 1)]
    coefs = []
    for l in range(LMAX + 1):
        for n in range(l + 1, NMAX + 1):
            offs[l][n] = len(coefs)
            p = polys[(l, n)]
            for i in range(n):
                coefs.append(p.get((i, n - 1 - i), 0))
    lines.append("static const int kHcGMaxN = %d;" % NMAX)
    lines.append("static const int kHcGOffset[%d][%d] = {" % (LMAX + 1, NMAX + 1))
    lines.append(",\n".join("    {" + ", ".join(str(v) for v in row) + "}" for row in offs))
    lines.append("};")
    lines.append("static const double kHcGCoef[%d] = {" % len(coefs))
    for k in range(0, len(coefs), 6):
        lines.append("    " + ", ".join("%d.0" % v for v in coefs[k:k + 6]) + ",")
    lines.append("};")
    with open(os.path.join(HERE, "hc_tables.h"), "w") as f:
        f.write("\n".join(lines) + "\n")
    print("norms:", [mp.nstr(v, 16) for v in norms])


if __name__ == "__main__":
    main()
