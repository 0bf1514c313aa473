"""McMurchie-Davidson electron-repulsion integrals, an independent check.

Test helper only.  The oracle and the product evaluate integrals with the
Obara-Saika vertical and Head-Gordon-Pople horizontal recurrences; this
module uses the Hermite-Gaussian (McMurchie-Davidson) expansion instead,
so agreement pins the recurrences, their direction choices and index
bookkeeping.  The auxiliary function is passed in (Boys F_n for the
Coulomb operator, G_n(R,T) for the truncated one): in both cases
R^n_000 = (-2 rho)^n g_n(T) and the Hermite recursion is unchanged,
because only T-derivatives at fixed R = R_cut sqrt(rho) occur.
"""
import math
import numpy as np


def cart_list(l):
    """cartesian components of shell l, descending x then descending y"""
    out = []
    for lx in range(l, -1, -1):
        for ly in range(l - lx, -1, -1):
            out.append((lx, ly, l - lx - ly))
    return out


def _hermite_E(la, lb, a, b, XA, XB):
    """E[i][j][t] along one axis for i<=la, j<=lb"""
    p = a + b
    mu = a * b / p
    XAB = XA - XB
    XPA = (a * XA + b * XB) / p - XA
    XPB = (a * XA + b * XB) / p - XB
    E = np.zeros((la + 1, lb + 1, la + lb + 2))
    E[0, 0, 0] = math.exp(-mu * XAB * XAB)
    for i in range(la + 1):
        for j in range(lb + 1):
            if i == 0 and j == 0:
                continue
            for t in range(i + j + 1):
                if i > 0:
                    v = XPA * E[i - 1, j, t]
                    if t > 0:
                        v += E[i - 1, j, t - 1] / (2 * p)
                    v += (t + 1) * E[i - 1, j, t + 1]
                else:
                    v = XPB * E[i, j - 1, t]
                    if t > 0:
                        v += E[i, j - 1, t - 1] / (2 * p)
                    v += (t + 1) * E[i, j - 1, t + 1]
                E[i, j, t] = v
    return E


def _hermite_R(L, rho, PQ, g):
    """R[t,u,v] = R^0_tuv for t+u+v <= L; g[n] = auxiliary function values"""
    X, Y, Z = PQ
    Rn = np.zeros((L + 1, L + 1, L + 1, L + 1))
    for n in range(L + 1):
        Rn[n, 0, 0, 0] = (-2 * rho) ** n * g[n]
    for tot in range(1, L + 1):
        for n in range(L - tot + 1):
            for t in range(tot + 1):
                for u in range(tot - t + 1):
                    v = tot - t - u
                    if t > 0:
                        val = X * Rn[n + 1, t - 1, u, v]
                        if t > 1:
                            val += (t - 1) * Rn[n + 1, t - 2, u, v]
                    elif u > 0:
                        val = Y * Rn[n + 1, t, u - 1, v]
                        if u > 1:
                            val += (u - 1) * Rn[n + 1, t, u - 2, v]
                    else:
                        val = Z * Rn[n + 1, t, u, v - 1]
                        if v > 1:
                            val += (v - 1) * Rn[n + 1, t, u, v - 2]
                    Rn[n, t, u, v] = val
    return Rn[0]


def eri_prim(la, lb, lc, ld, a, b, c, d, A, B, C, D, aux):
    """primitive cartesian (ab|cd), unnormalised x^l y^m z^n e^{-a r^2}
    functions; aux(rho, T, L) -> array of g_0..g_L.  Returns array
    [na, nb, nc, nd] in the library's component order."""
    p, q = a + b, c + d
    P = (a * np.asarray(A) + b * np.asarray(B)) / p
    Q = (c * np.asarray(C) + d * np.asarray(D)) / q
    rho = p * q / (p + q)
    PQ = P - Q
    T = rho * float(PQ @ PQ)
    L = la + lb + lc + ld
    g = aux(rho, T, L)
    Rt = _hermite_R(L, rho, PQ, g)
    Eab = [_hermite_E(la, lb, a, b, A[k], B[k]) for k in range(3)]
    Ecd = [_hermite_E(lc, ld, c, d, C[k], D[k]) for k in range(3)]
    ca, cb, cc, cd = cart_list(la), cart_list(lb), cart_list(lc), cart_list(ld)
    Lab, Lcd = la + lb, lc + ld
    tuv = [(t, u, L_ - t - u) for L_ in range(Lab + 1) for t in range(L_ + 1) for u in range(L_ - t + 1)]
    tnp = [(t, u, L_ - t - u) for L_ in range(Lcd + 1) for t in range(L_ + 1) for u in range(L_ - t + 1)]
    Mab = np.zeros((len(ca) * len(cb), len(tuv)))
    for i, va in enumerate(ca):
        for j, vb in enumerate(cb):
            for k, (t, u, v) in enumerate(tuv):
                Mab[i * len(cb) + j, k] = Eab[0][va[0], vb[0], t] * Eab[1][va[1], vb[1], u] * Eab[2][va[2], vb[2], v]
    Mcd = np.zeros((len(cc) * len(cd), len(tnp)))
    for i, vc in enumerate(cc):
        for j, vd in enumerate(cd):
            for k, (t, u, v) in enumerate(tnp):
                Mcd[i * len(cd) + j, k] = ((-1) ** (t + u + v)) * Ecd[0][vc[0], vd[0], t] * Ecd[1][vc[1], vd[1], u] * Ecd[2][vc[2], vd[2], v]
    Rm = np.zeros((len(tuv), len(tnp)))
    for i, (t, u, v) in enumerate(tuv):
        for j, (t2, u2, v2) in enumerate(tnp):
            Rm[i, j] = Rt[t + t2, u + u2, v + v2]
    pref = 2 * math.pi ** 2.5 / (p * q * math.sqrt(p + q))
    out = pref * (Mab @ Rm @ Mcd.T)
    return out.reshape(len(ca), len(cb), len(cc), len(cd))


def boys_aux(rho, T, L):
    """Coulomb: F_n(T) by the confluent hypergeometric series"""
    import mpmath as mp
    return np.array([float(mp.hyp1f1(n + 0.5, n + 1.5, -T) / (2 * n + 1)) for n in range(L + 1)])
