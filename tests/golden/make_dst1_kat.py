"""Writes dst1_kat.json: known-answer vectors for the 1-D DST-I
    Y_k = 2 sum_{j=1}^{N-1} x_j sin(pi j k / N)
evaluated term by term with mpmath at 40 significant digits from exactly
representable inputs (small dyadic rationals).  The reference's DST is FFTW
RODFT00 with this definition (src/myfft.f90:41); it ships no vectors."""
import json
import os

import mpmath


def main():
    mpmath.mp.dps = 40
    out = []
    for n in (3, 7, 9, 15, 19, 31, 39, 63, 99):
        N = n + 1
        x = [((7 * j * j + 3 * j) % 17 - 8) / 16.0 for j in range(1, N)]
        y = []
        for k in range(1, N):
            s = mpmath.mpf(0)
            for j in range(1, N):
                s += mpmath.mpf(x[j - 1]) * mpmath.sin(mpmath.pi * j * k / N)
            y.append(float(2 * s))
        out.append({"n": n, "x": x, "y": y})
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "dst1_kat.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=0)


if __name__ == "__main__":
    main()
