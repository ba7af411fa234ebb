"""CPU restatement of the plane-wave expansion, the cluster far field and the
cross sections (test infrastructure only: the product never imports this).

Conventions (those of the reference's Mie code):
  * pi_mn = m P_n^m(cos t) / sin t, tau_mn = d P_n^m(cos t) / dt, with P_n^m (m >= 0)
    without the Condon-Shortley phase and P_n^{-m} = (-1)^m (n-m)!/(n+m)! P_n^m
    (lib_math_legendre.f90:128-271, 834-964; the net sign convention of
    lib_mie_illumination.f90:714-726 at alpha = 0: pi_{n,+-1} = n(n+1)/2, 1/2 and
    tau_{n,+-1} = n(n+1)/2, -1/2);
  * plane wave travelling along (alpha, beta), amplitude 1, at sphere centre d:
        p_mn = e^{i k khat.d} tau_mn(cos alpha) e^{-i m beta} / (n(n+1))
        q_mn = e^{i k khat.d} pi_mn(cos alpha)  e^{-i m beta} / (n(n+1))
    (lib_mie_illumination.f90:636-763, Xu 1995 eq. 21 with its erratum);
  * scattered field E_s = sum_j sum_mn i E_mn [a^j_mn N^(3)_mn + b^j_mn M^(3)_mn],
    E_mn = i^n (2n+1)(n-m)!/(n+m)! (Xu 1995), so that far from the cluster
        E_s ~ e^{ikr}/(kr) F(theta, phi),
        F_theta =  i sum_j e^{-ik rhat.d_j} sum_mn c_mn e^{i m phi} (a tau_mn + b pi_mn)
        F_phi   = -  sum_j e^{-ik rhat.d_j} sum_mn c_mn e^{i m phi} (a pi_mn + b tau_mn)
    with c_mn = (2n+1)(n-m)!/(n+m)!;
  * cross sections (lib_mie_ss_helper_functions.f90:1380-1460 per sphere, Xu eq. 57):
        C_ext = 4 pi/k^2 sum_j sum_mn w_mn Re(conj(p) a + conj(q) b)
        C_abs = 4 pi/k^2 sum_j sum_mn w_mn (|a|^2 (Re(1/a_n) - 1) + |b|^2 (Re(1/b_n) - 1))
        C_sca = C_ext - C_abs,      w_mn = n(n+1)(2n+1)(n-m)!/(n+m)!
    (the absorption of sphere j follows from its exciting field a^j_mn / a_n^j);
  * amplitude matrix (Bohren & Huffman 3.12) from the far fields of an x- and a
    y-polarised wave along +z, and the Mueller matrix (Bohren & Huffman 3.16).

Pinned by tests/test_farfield_oracle.py: closed forms of pi, tau at alpha = 0 and a
brute-force mpmath evaluation; Lorenz-Mie S1, S2 of one sphere (textbook pi_n, tau_n
recurrences); the optical theorem and energy conservation for clusters.
"""
import math

import numpy as np


def _fact_ratio(n, m):
    """(n-m)!/(n+m)! for |m| <= n."""
    return math.exp(math.lgamma(n - m + 1) - math.lgamma(n + m + 1))


def pitau(nmax, theta):
    """pi[n][m+n], tau[n][m+n] for n = 1..nmax (index n-1), reference convention."""
    x = math.cos(theta)
    s = math.sin(theta)
    pi = [np.zeros(2 * n + 1) for n in range(1, nmax + 1)]
    tau = [np.zeros(2 * n + 1) for n in range(1, nmax + 1)]
    # u[n] = P_n^m / sin(theta) for fixed m >= 1 (no CS phase), upward in n
    for m in range(1, nmax + 1):
        u = {}
        df = 1.0
        for i in range(1, 2 * m, 2):
            df *= i
        u[m - 1] = 0.0
        u[m] = df * (s ** (m - 1) if m > 1 else 1.0)
        for n in range(m, nmax):
            u[n + 1] = ((2 * n + 1) * x * u[n] - (n + m) * u.get(n - 1, 0.0)) / (n - m + 1)
        for n in range(m, nmax + 1):
            p = m * u[n]
            t = n * x * u[n] - (n + m) * u[n - 1]
            f = _fact_ratio(n, m)
            pi[n - 1][m + n] = p
            tau[n - 1][m + n] = t
            pi[n - 1][-m + n] = (-1) ** (m + 1) * f * p
            tau[n - 1][-m + n] = (-1) ** m * f * t
            if m == 1:
                tau[n - 1][n] = -s * u[n]  # tau_n0 = -P_n^1
    return pi, tau


def _blocks(N):
    return N * (N + 2)


def plane_wave(pos, N, k, alpha=0.0, beta=0.0):
    """Incident coefficients (p block, q block per sphere) in the flat layout."""
    pos = np.asarray(pos, dtype=np.float64).reshape(-1, 3)
    khat = np.array([math.sin(alpha) * math.cos(beta), math.sin(alpha) * math.sin(beta), math.cos(alpha)])
    pi, tau = pitau(N, alpha)
    blk = _blocks(N)
    out = np.zeros(len(pos) * 2 * blk, dtype=np.complex128)
    for j, d in enumerate(pos):
        ph = np.exp(1j * k * float(khat @ d))
        base = j * 2 * blk
        for n in range(1, N + 1):
            for m in range(-n, n + 1):
                e = ph * np.exp(-1j * m * beta) / (n * (n + 1))
                i = n * n - 1 + m + n
                out[base + i] = tau[n - 1][m + n] * e
                out[base + blk + i] = pi[n - 1][m + n] * e
    return out


def far_field(coef, pos, N, k, theta, phi):
    """F_theta, F_phi at the directions (theta[i], phi[i])."""
    pos = np.asarray(pos, dtype=np.float64).reshape(-1, 3)
    blk = _blocks(N)
    theta = np.atleast_1d(theta)
    phi = np.atleast_1d(phi)
    Ft = np.zeros(len(theta), dtype=np.complex128)
    Fp = np.zeros(len(theta), dtype=np.complex128)
    ns = np.array([n for n in range(1, N + 1) for m in range(-n, n + 1)])
    ms = np.array([m for n in range(1, N + 1) for m in range(-n, n + 1)])
    c = np.array([(2 * n + 1) * _fact_ratio(n, m) for n, m in zip(ns, ms)])
    for i, (t, f) in enumerate(zip(theta, phi)):
        pi, tau = pitau(N, t)
        piv = np.concatenate(pi)
        tauv = np.concatenate(tau)
        rhat = np.array([math.sin(t) * math.cos(f), math.sin(t) * math.sin(f), math.cos(t)])
        em = np.exp(1j * ms * f) * c
        for j, d in enumerate(pos):
            a = coef[j * 2 * blk:j * 2 * blk + blk]
            b = coef[j * 2 * blk + blk:(j + 1) * 2 * blk]
            ph = np.exp(-1j * k * float(rhat @ d))
            Ft[i] += 1j * ph * np.sum(em * (a * tauv + b * piv))
            Fp[i] += -ph * np.sum(em * (a * piv + b * tauv))
    return Ft, Fp


def cross_sections(coef, inc, an, bn, N, k):
    """(C_sca, C_ext, C_abs) of the cluster; an, bn: [ns, N] Mie coefficients."""
    blk = _blocks(N)
    ns = len(coef) // (2 * blk)
    nn = np.array([n for n in range(1, N + 1) for m in range(-n, n + 1)])
    mm = np.array([m for n in range(1, N + 1) for m in range(-n, n + 1)])
    w = np.array([n * (n + 1) * (2 * n + 1) * _fact_ratio(n, m) for n, m in zip(nn, mm)])
    ext = 0.0
    ab = 0.0
    for j in range(ns):
        a = coef[j * 2 * blk:j * 2 * blk + blk]
        b = coef[j * 2 * blk + blk:(j + 1) * 2 * blk]
        p = inc[j * 2 * blk:j * 2 * blk + blk]
        q = inc[j * 2 * blk + blk:(j + 1) * 2 * blk]
        ext += np.sum(w * np.real(np.conj(p) * a + np.conj(q) * b))
        aj = an[j][nn - 1]
        bj = bn[j][nn - 1]
        with np.errstate(divide="ignore", invalid="ignore"):
            ra = np.where(aj != 0, np.real(1.0 / np.where(aj != 0, aj, 1.0)) - 1.0, 0.0)
            rb = np.where(bj != 0, np.real(1.0 / np.where(bj != 0, bj, 1.0)) - 1.0, 0.0)
        ab += np.sum(w * (np.abs(a) ** 2 * ra + np.abs(b) ** 2 * rb))
    f = 4.0 * math.pi / k ** 2
    return f * (ext - ab), f * ext, f * ab


def amplitude_matrix(Fx, Fy, phi):
    """S1..S4 from the far fields (F_theta, F_phi) of x- and y-polarised waves along +z."""
    (xt, xp), (yt, yp) = Fx, Fy
    c = np.cos(phi)
    s = np.sin(phi)
    S2 = -1j * (xt * c + yt * s)
    S3 = -1j * (xt * s - yt * c)
    S4 = 1j * (xp * c + yp * s)
    S1 = 1j * (xp * s - yp * c)
    return S1, S2, S3, S4


def mueller(S1, S2, S3, S4):
    """Mueller matrix M[..., 4, 4] (Bohren & Huffman 3.16)."""
    a1, a2, a3, a4 = (np.abs(v) ** 2 for v in (S1, S2, S3, S4))
    M = np.zeros(np.shape(S1) + (4, 4))
    M[..., 0, 0] = 0.5 * (a2 + a3 + a4 + a1)
    M[..., 0, 1] = 0.5 * (a2 - a3 + a4 - a1)
    M[..., 0, 2] = np.real(S2 * np.conj(S3) + S1 * np.conj(S4))
    M[..., 0, 3] = np.imag(S2 * np.conj(S3) - S1 * np.conj(S4))
    M[..., 1, 0] = 0.5 * (a2 - a4 + a3 - a1)
    M[..., 1, 1] = 0.5 * (a2 - a3 - a4 + a1)
    M[..., 1, 2] = np.real(S2 * np.conj(S3) - S1 * np.conj(S4))
    M[..., 1, 3] = np.imag(S2 * np.conj(S3) + S1 * np.conj(S4))
    M[..., 2, 0] = np.real(S2 * np.conj(S4) + S1 * np.conj(S3))
    M[..., 2, 1] = np.real(S2 * np.conj(S4) - S1 * np.conj(S3))
    M[..., 2, 2] = np.real(np.conj(S1) * S2 + np.conj(S3) * S4)
    M[..., 2, 3] = np.imag(S2 * np.conj(S1) + S4 * np.conj(S3))
    M[..., 3, 0] = np.imag(np.conj(S2) * S4 + np.conj(S3) * S1)
    M[..., 3, 1] = np.imag(np.conj(S2) * S4 - np.conj(S3) * S1)
    M[..., 3, 2] = np.imag(S1 * np.conj(S2) - S3 * np.conj(S4))
    M[..., 3, 3] = np.real(S1 * np.conj(S2) - S3 * np.conj(S4))
    return M
