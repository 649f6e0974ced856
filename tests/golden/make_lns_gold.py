"""Write the LNS coefficient and transform golden fixtures.

The reference pins its pointwise physics with two Fortran 2008 gold
generators (tests/CPlusPlus/generate_lns_coefficients_gold.f90 and
generate_lns_transform_gold.f90) whose output the C++ tests compare at
5e-13 * max(1, |x|).  gfortran is not available here, so this script restates
those two Fortran programs -- their inputs, formulas, 1-based loops and output
order -- in plain Python floats and complex numbers, and writes the same text
the Fortran would write ('(ES25.17E3)' per value, '(2ES25.17E3)' per complex
value).  The fixtures are data: they are compared with the C oracle and the
GPU kernels, neither of which this script imports.

    python tests/golden/make_lns_gold.py
"""
from __future__ import annotations

from math import sqrt
from pathlib import Path

HERE = Path(__file__).resolve().parent


def points():
    """The two base-flow points of generate_lns_coefficients_gold.f90:18-35.

    Returns q[point][tuple][component] with tuples 0..9 = value, x, y, z, xx,
    yy, zz, xy, xz, yz (the Fortran's second index 1..10).
    """
    p1 = [[1.2, 2.3, 0.14, -0.08, 1.4],
          [0.01, 0.02, -0.03, 0.04, 0.05],
          [-0.02, 0.07, 0.04, -0.05, 0.06],
          [0.03, -0.02, 0.08, 0.09, -0.04],
          [0.004, -0.006, 0.007, -0.008, 0.009],
          [-0.005, 0.011, -0.012, 0.013, -0.014],
          [0.006, -0.015, 0.016, -0.017, 0.018],
          [-0.007, 0.019, -0.020, 0.021, -0.022],
          [0.008, -0.023, 0.024, -0.025, 0.026],
          [-0.009, 0.027, -0.028, 0.029, -0.030]]
    z = [0.0] * 5
    p2 = [[0.95, 3.1, -0.12, 0.06, 1.15], z,
          [0.015, -0.045, 0.025, 0.035, -0.055],
          [-0.012, 0.038, -0.028, 0.042, 0.048], z,
          [0.003, -0.009, 0.008, -0.007, 0.011],
          [-0.004, 0.010, -0.012, 0.014, -0.013], z, z,
          [0.005, -0.016, 0.017, -0.018, 0.019]]
    return [p1, [list(r) for r in p2]]


def get_unadorned(q):
    """Eleven 5x5 matrices, 1-based [row][col], Fortran formulas verbatim."""
    Re, Ma, Pr = 32500.0, 6.0, 0.72
    gamma, Te, C1 = 1.4, 63.334, 110.4
    d1d3, d2d3, d4d3 = 1.0 / 3.0, 2.0 / 3.0, 4.0 / 3.0

    def mat():
        return [[0.0] * 6 for _ in range(6)]

    G, A, B, C, D = mat(), mat(), mat(), mat(), mat()
    Vxxm, Vyym, Vzzm, Vxym, Vxzm, Vyzm = mat(), mat(), mat(), mat(), mat(), mat()
    rho, U, V, W, T = q[0]
    rhox, Ux, Vx, Wx, Tx = q[1]
    rhoy, Uy, Vy, Wy, Ty = q[2]
    rhoz, Uz, Vz, Wz, Tz = q[3]
    _, Uxx, nVxx, Wxx, Txx = q[4]
    _, Uyy, nVyy, Wyy, Tyy = q[5]
    _, Uzz, nVzz, Wzz, Tzz = q[6]
    _, Uxy, nVxy, Wxy, _ = q[7]
    _, Uxz, nVxz, Wxz, _ = q[8]
    _, Uyz, nVyz, Wyz, _ = q[9]

    Pe = 1.0 / (gamma * Ma * Ma)
    g2 = 1.0 / ((gamma - 1.0) * Ma * Ma)
    cm = C1 / Te
    mu = T * sqrt(T) * (1.0 + cm) / (T + cm)
    muT = mu * (1.5 / T - 1.0 / (T + cm))
    muTT = muT * (1.5 / T - 1.0 / (T + cm)) - mu * (1.5 / T ** 2 - 1.0 / (T + cm) ** 2)
    mux, muy, muz = muT * Tx, muT * Ty, muT * Tz

    G[1][1] = 1; G[2][2] = rho; G[3][3] = rho; G[4][4] = rho
    G[5][1] = -Pe * T; G[5][5] = rho * g2 - Pe * rho

    A[1][1] = U; A[1][2] = rho; A[2][1] = Pe * T
    A[2][2] = rho * U - d4d3 * mux / Re
    A[2][3] = -muy / Re; A[2][4] = -muz / Re
    A[2][5] = Pe * rho - muT / Re * (d4d3 * Ux - d2d3 * Vy - d2d3 * Wz)
    A[3][2] = d2d3 * muy / Re; A[3][3] = rho * U - mux / Re
    A[3][5] = -muT * (Uy + Vx) / Re
    A[4][2] = d2d3 * muz / Re; A[4][4] = rho * U - mux / Re
    A[4][5] = -muT * (Wx + Uz) / Re
    A[5][1] = -Pe * U * T
    A[5][2] = -2.0 * mu * (d4d3 * Ux - d2d3 * Vy - d2d3 * Wz) / Re
    A[5][3] = -2.0 * mu * (Uy + Vx) / Re
    A[5][4] = -2.0 * mu * (Wx + Uz) / Re
    A[5][5] = rho * U * g2 - rho * U * Pe - 2.0 * Tx * muT / Re / Pr * g2

    B[1][1] = V; B[1][3] = rho
    B[2][2] = rho * V - muy / Re; B[2][3] = d2d3 * mux / Re
    B[2][5] = -muT * (Uy + Vx) / Re
    B[3][1] = Pe * T; B[3][2] = -mux / Re
    B[3][3] = rho * V - d4d3 * muy / Re; B[3][4] = -muz / Re
    B[3][5] = Pe * rho - muT * (-d2d3 * Ux + d4d3 * Vy - d2d3 * Wz) / Re
    B[4][3] = d2d3 * muz / Re; B[4][4] = rho * V - muy / Re
    B[4][5] = -muT * (Vz + Wy) / Re
    B[5][1] = -Pe * V * T; B[5][2] = -2.0 * mu * (Uy + Vx) / Re
    B[5][3] = -2.0 * mu * (-d2d3 * Ux + d4d3 * Vy - d2d3 * Wz) / Re
    B[5][4] = -2.0 * mu * (Vz + Wy) / Re
    B[5][5] = rho * V * g2 - rho * V * Pe - 2.0 * Ty * muT * g2 / Re / Pr

    C[1][1] = W; C[1][4] = rho
    C[2][2] = rho * W - muz / Re; C[2][4] = d2d3 * mux / Re
    C[2][5] = -muT * (Wx + Uz) / Re
    C[3][3] = rho * W - muz / Re; C[3][4] = d2d3 * muy / Re
    C[3][5] = -muT * (Vz + Wy) / Re
    C[4][1] = Pe * T; C[4][2] = -mux / Re; C[4][3] = -muy / Re
    C[4][4] = rho * W - d4d3 * muz / Re
    C[4][5] = Pe * rho - muT * (-d2d3 * Ux - d2d3 * Vy + d4d3 * Wz) / Re
    C[5][1] = -Pe * W * T; C[5][2] = -2.0 * mu * (Wx + Uz) / Re
    C[5][3] = -2.0 * mu * (Vz + Wy) / Re
    C[5][4] = -2.0 * mu * (-d2d3 * Ux - d2d3 * Vy + d4d3 * Wz) / Re
    C[5][5] = rho * W * g2 - rho * W * Pe - 2.0 * Tz * muT * g2 / Re / Pr

    D[1][1] = Ux + Vy + Wz; D[1][2] = rhox; D[1][3] = rhoy; D[1][4] = rhoz
    D[2][1] = U * Ux + V * Uy + W * Uz + Pe * Tx
    D[2][2] = rho * Ux; D[2][3] = rho * Uy; D[2][4] = rho * Uz
    D[2][5] = Pe * rhox - (muTT * Tx * d2d3 * (2.0 * Ux - Vy - Wz)
                           + muTT * Ty * (Uy + Vx) + muTT * Tz * (Wx + Uz)
                           + muT * (d4d3 * Uxx + Uyy + Uzz + d1d3 * nVxy + d1d3 * Wxz)) / Re
    D[3][1] = U * Vx + V * Vy + W * Vz + Pe * Ty
    D[3][2] = rho * Vx; D[3][3] = rho * Vy; D[3][4] = rho * Vz
    D[3][5] = Pe * rhoy - (muTT * Tx * (Uy + Vx)
                           + muTT * Ty * d2d3 * (-Ux + 2.0 * Vy - Wz) + muTT * Tz * (Vz + Wy)
                           + muT * (nVxx + d4d3 * nVyy + nVzz + d1d3 * Uxy + d1d3 * Wyz)) / Re
    D[4][1] = U * Wx + V * Wy + W * Wz + Pe * Tz
    D[4][2] = rho * Wx; D[4][3] = rho * Wy; D[4][4] = rho * Wz
    D[4][5] = Pe * rhoz - (muTT * Tx * (Wx + Uz) + muTT * Ty * (Vz + Wy)
                           + muTT * Tz * d2d3 * (-Ux - Vy + 2.0 * Wz)
                           + muT * (Wxx + Wyy + d4d3 * Wzz + d1d3 * Uxz + d1d3 * nVyz)) / Re
    D[5][1] = (g2 - Pe) * (U * Tx + V * Ty + W * Tz)
    D[5][2] = rho * Tx * g2 - Pe * (rho * Tx + T * rhox)
    D[5][3] = rho * Ty * g2 - Pe * (rho * Ty + T * rhoy)
    D[5][4] = rho * Tz * g2 - Pe * (rho * Tz + T * rhoz)
    D[5][5] = (-Pe * (U * rhox + V * rhoy + W * rhoz)
               - (Txx + Tyy + Tzz) * muT * g2 / Re / Pr
               - (Tx * Tx + Ty * Ty + Tz * Tz) * muTT * g2 / Re / Pr
               - muT * (Uy * Uy + Vx * Vx + 2.0 * Uy * Vx + Uz * Uz + Wx * Wx
                        + 2.0 * Uz * Wx + Vz * Vz + Wy * Wy + 2.0 * Vz * Wy) / Re
               - muT * d4d3 * (Ux * Ux + Vy * Vy + Wz * Wz - Ux * Wz - Ux * Vy - Vy * Wz) / Re)

    Vxxm[2][2] = d4d3 * mu / Re; Vxxm[3][3] = mu / Re
    Vxxm[4][4] = mu / Re; Vxxm[5][5] = mu * g2 / Re / Pr
    Vyym[2][2] = mu / Re; Vyym[3][3] = d4d3 * mu / Re
    Vyym[4][4] = mu / Re; Vyym[5][5] = mu * g2 / Re / Pr
    Vzzm[2][2] = mu / Re; Vzzm[3][3] = mu / Re
    Vzzm[4][4] = d4d3 * mu / Re; Vzzm[5][5] = mu * g2 / Re / Pr
    Vxym[2][3] = d1d3 * mu / Re; Vxym[3][2] = d1d3 * mu / Re
    Vxzm[2][4] = d1d3 * mu / Re; Vxzm[4][2] = d1d3 * mu / Re
    Vyzm[3][4] = d1d3 * mu / Re; Vyzm[4][3] = d1d3 * mu / Re
    return [G, A, B, C, D, Vxxm, Vxym, Vxzm, Vyym, Vyzm, Vzzm]


def coefficient_lines():
    out = []
    for q in points():
        mats = get_unadorned(q)
        for m in mats:
            for r in range(1, 6):
                for c in range(1, 6):
                    out.append(f"{m[r][c]:25.17E}")
    return out


def transform_lines():
    ci = complex(0.0, 1.0)
    out = []
    for point in (1, 2):
        phys = []
        for matrix in range(1, 12):
            m = [[0j] * 6 for _ in range(6)]
            for r in range(1, 6):
                for c in range(1, 6):
                    v = 0.13 * matrix + 0.017 * r + 0.003 * c + 0.0007 * point
                    if (matrix + r + c + point) % 2 == 0:
                        v = -v
                    m[r][c] = complex(v, 0.0)
            phys.append(m)
        if point == 1:
            alpha = complex(2.43, 0.17)
            met = [0.81, -0.23, 0.19, 1.07, -0.031, 0.024, -0.017, 0.028, -0.022, 0.013]
        else:
            alpha = complex(-0.64, 0.29)
            met = [1.14, 0.37, -0.26, 0.72, 0.041, -0.036, 0.025, -0.033, 0.019, -0.027]
        four = [[[0j] * 6 for _ in range(6)] for _ in range(7)]
        for r in range(1, 6):
            for c in range(1, 6):
                four[0][r][c] = phys[0][r][c]
                four[1][r][c] = phys[4][r][c] + ci * alpha * phys[1][r][c] + alpha * alpha * phys[5][r][c]
                four[2][r][c] = phys[2][r][c] - ci * alpha * phys[6][r][c]
                four[3][r][c] = phys[3][r][c] - ci * alpha * phys[7][r][c]
                four[4][r][c] = phys[8][r][c]
                four[5][r][c] = phys[9][r][c]
                four[6][r][c] = phys[10][r][c]
        xi_y, xi_z, eta_y, eta_z, xi_yy, xi_zz, xi_yz, eta_yy, eta_zz, eta_yz = met
        curv = [[[0j] * 6 for _ in range(6)] for _ in range(7)]
        for r in range(1, 6):
            for c in range(1, 6):
                f = [four[k][r][c] for k in range(7)]
                curv[0][r][c] = f[0]
                curv[1][r][c] = f[1]
                curv[2][r][c] = xi_y * f[2] + xi_z * f[3] - xi_yy * f[4] - xi_zz * f[6] - xi_yz * f[5]
                curv[3][r][c] = eta_y * f[2] + eta_z * f[3] - eta_yy * f[4] - eta_zz * f[6] - eta_yz * f[5]
                curv[4][r][c] = xi_y * xi_y * f[4] + xi_z * xi_z * f[6] + xi_y * xi_z * f[5]
                curv[5][r][c] = (2.0 * xi_y * eta_y * f[4] + 2.0 * xi_z * eta_z * f[6]
                                 + (xi_y * eta_z + eta_y * xi_z) * f[5])
                curv[6][r][c] = eta_y * eta_y * f[4] + eta_z * eta_z * f[6] + eta_y * eta_z * f[5]
        for blocks in (four, curv):
            for m in blocks:
                for r in range(1, 6):
                    for c in range(1, 6):
                        out.append(f"{m[r][c].real:25.17E}{m[r][c].imag:25.17E}")
    return out


if __name__ == "__main__":
    (HERE / "lns_coefficients_gold.txt").write_text("\n".join(coefficient_lines()) + "\n")
    (HERE / "lns_transform_gold.txt").write_text("\n".join(transform_lines()) + "\n")
    print("wrote", HERE / "lns_coefficients_gold.txt", HERE / "lns_transform_gold.txt")
