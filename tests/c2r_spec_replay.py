"""An independent replay of the project c2r definition in double precision:
Python floats are IEEE doubles, each fused multiply-add is evaluated exactly
with rationals and rounded once, and the twiddles are libm cos and sin called
separately.  Slow (pure Python); used on small sizes to pin that the oracle
follows the written operation sequence, not only that it is accurate."""
import math
from fractions import Fraction

import numpy as np

def fma(a, b, c):
    return float(Fraction(a) * Fraction(b) + Fraction(c))

def tw(j, N):
    a = 2.0 * math.pi * float(j) / float(N)
    return (math.cos(a), math.sin(a))

def cmul(y, w):
    return (fma(y[0], w[0], -(y[1] * w[1])), fma(y[0], w[1], y[1] * w[0]))

def spf(N):
    if N % 2 == 0: return 2
    p = 3
    while p * p <= N:
        if N % p == 0: return p
        p += 2
    return N

def pfa_n1(N):
    p = spf(N); n1 = 1
    while N % (n1 * p) == 0: n1 *= p
    return n1

def radix(N):
    return 4 if N % 4 == 0 else (2 if N % 2 == 0 else spf(N))

def modinv(a, m):
    for x in range(1, m):
        if (a % m) * x % m == 1: return x
    return 0

def bfly(u, R):
    if R == 2:
        return [(u[0][0] + u[1][0], u[0][1] + u[1][1]), (u[0][0] - u[1][0], u[0][1] - u[1][1])]
    if R == 4:
        t0 = (u[0][0] + u[2][0], u[0][1] + u[2][1]); t1 = (u[0][0] - u[2][0], u[0][1] - u[2][1])
        t2 = (u[1][0] + u[3][0], u[1][1] + u[3][1]); t3 = (u[1][0] - u[3][0], u[1][1] - u[3][1])
        return [(t0[0] + t2[0], t0[1] + t2[1]), (t1[0] - t3[1], t1[1] + t3[0]),
                (t0[0] - t2[0], t0[1] - t2[1]), (t1[0] + t3[1], t1[1] - t3[0])]
    hh = (R - 1) // 2
    sp = [None] * (hh + 1); sm = [None] * (hh + 1)
    for r in range(1, hh + 1):
        sp[r] = (u[r][0] + u[R - r][0], u[r][1] + u[R - r][1])
        sm[r] = (u[r][0] - u[R - r][0], u[r][1] - u[R - r][1])
    v = [None] * R
    x0 = u[0]
    for r in range(1, hh + 1): x0 = (x0[0] + sp[r][0], x0[1] + sp[r][1])
    v[0] = x0
    for s in range(1, hh + 1):
        A = u[0]; B = None
        for r in range(1, hh + 1):
            c, sn = tw((r * s) % R, R)
            A = (fma(sp[r][0], c, A[0]), fma(sp[r][1], c, A[1]))
            B = (sm[r][0] * sn, sm[r][1] * sn) if r == 1 else (fma(sm[r][0], sn, B[0]), fma(sm[r][1], sn, B[1]))
        v[s] = (A[0] - B[1], A[1] + B[0]); v[R - s] = (A[0] + B[1], A[1] - B[0])
    return v

def idft(x):
    N = len(x)
    if N == 1: return [x[0]]
    N1 = pfa_n1(N)
    if N1 < N:
        N2 = N // N1
        E1 = N2 * modinv(N2, N1); E2 = N1 * modinv(N1, N2)
        T = [idft([x[(N2 * n1 + N1 * n2) % N] for n1 in range(N1)]) for n2 in range(N2)]
        X = [None] * N
        for k1 in range(N1):
            o = idft([T[n2][k1] for n2 in range(N2)])
            for k2 in range(N2): X[(k1 * E1 + k2 * E2) % N] = o[k2]
        return X
    R = radix(N); M = N // R
    Y = [idft([x[r + j * R] for j in range(M)]) for r in range(R)]
    X = [None] * N
    for k in range(M):
        u = [Y[0][k]] + [Y[r][k] if r * k == 0 else cmul(Y[r][k], tw(r * k, N)) for r in range(1, R)]
        v = bfly(u, R)
        for s in range(R): X[k + M * s] = v[s]
    return X

def c2r(X):
    H = len(X) - 1
    Z = [(X[0][0] + X[H][0], X[0][0] - X[H][0])]
    for k in range(1, H):
        a = X[k]; b = X[H - k]
        er, ei = a[0] + b[0], a[1] - b[1]
        o = cmul((a[0] - b[0], a[1] + b[1]), tw(k, 2 * H))
        Z.append((er - o[1], ei + o[0]))
    z = idft(Z)
    out = []
    for s in range(H): out += [z[s][0], z[s][1]]
    return np.array(out)
