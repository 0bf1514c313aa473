"""Writes tcg_golden.json: the truncated-Coulomb auxiliary function
G_m(R,T), m = 0..12, at points that cover the low-R patches, the high-R
patches, the plain-Boys region and the zero region, evaluated with mpmath
at 40 digits from the closed form (independent of the patch tables)

  G_0(R,T) = sqrt(pi)/(4 sqrt(T)) [2 erf(sqrt T) + erf(R - sqrt T) - erf(R + sqrt T)]
  G_m(R,T) = (-d/dT)^m G_0(R,T) at fixed R   (every point has T > 0)

and F_m(T) = 1F1(m+1/2; m+3/2; -T) / (2m+1) for reference.  Test data;
run once: python tests/golden/make_tcg_golden.py
"""
import json
import os

import mpmath as mp

mp.mp.dps = 40
HERE = os.path.dirname(os.path.abspath(__file__))


def g0(R, T):
    R = mp.mpf(R)
    T = mp.mpf(T)
    s = mp.sqrt(T)
    return mp.sqrt(mp.pi) / (4 * s) * (2 * mp.erf(s) + mp.erf(R - s) - mp.erf(R + s))


def gm(R, T, M):
    # derivatives in T at fixed R; T > 0 only (points are chosen so)
    out = []
    for m in range(M + 1):
        d = mp.diff(lambda x: g0(R, x), mp.mpf(T), m)
        out.append(float((-1) ** m * d))
    return out


def fm(T, M):
    return [float(mp.hyp1f1(m + mp.mpf(1) / 2, m + mp.mpf(3) / 2, -mp.mpf(T)) / (2 * m + 1)) for m in range(M + 1)]


def main():
    pts = []
    # low R (R <= 11): T spans [0, R^2 + 11 R + 50]
    for R in (0.3, 1.7, 4.0, 7.5, 10.9):
        up = R * R + 11 * R + 50
        for f in (0.003, 0.07, 0.31, 0.64, 0.97):
            pts.append((R, f * up))
    # high R (R > 11): patches for T in [R^2 - 11 R, R^2 + 11 R + 50], Boys below
    for R in (11.4, 18.0, 35.0):
        lo, up = R * R - 11 * R, R * R + 11 * R + 50
        for f in (0.05, 0.4, 0.9):
            pts.append((R, lo + f * (up - lo)))
        pts.append((R, 0.5 * lo))   # plain Boys region
    # beyond the truncation cut
    for R in (2.0, 9.0, 20.0):
        pts.append((R, R * R + 11 * R + 50 + 5.0))
    out = []
    for R, T in pts:
        out.append({"R": R, "T": T, "G": gm(R, T, 12), "F": fm(T, 12)})
        print(R, T, out[-1]["G"][0], flush=True)
    with open(os.path.join(HERE, "tcg_golden.json"), "w") as f:
        json.dump({"dps": 40, "points": out}, f, indent=1)


if __name__ == "__main__":
    main()
