"""Free-space spectral Ewald Stokeslet sum restated in numpy, with a direct
(non-uniform) Fourier sum in place of spreading, FFT and gathering.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the library's
free-space Ewald path; never by the product package.

The reference has no free-space Ewald sum (it calls an FMM,
FMM/src/StokesSLP/stokesSLP.f); what is restated here is the mathematics the
library's path implements, so that each piece can be pinned on its own:

* bhat_trunc: the Fourier transform of the biharmonic kernel B(r) = r^2 (ln r -
  1)/(8 pi), less the quadratic matching its value and slope at R, truncated
  to r < R, closed form and power series; pinned against numerical quadrature
  (tests/test_oracle_free_ewald.py).
* slp_free_grid: the same sum the way the library evaluates it -- the
  oracle's periodic real-space cell-list sum and its Gaussian spreading and
  gathering (oracle/se2p_oracle.c) in the periodic box of se2p_free_params,
  numpy FFTs, and the free-space filter (K^2 I - k k^T) B_R(K) (1 + K^2/(4
  xi^2)) exp(-(1 - eta) K^2/(4 xi^2)) -- to pin the library's parameter choice
  against the direct sum on the CPU.
* slp_free: u = (real + k + const)/(4 pi) with the real-space Hasimoto sum of
  the periodic path (E1 and Gaussian terms, all pairs, no cutoff), the k-space
  sum (1/(Lx Ly)) sum_k 4 pi (K^2 I - k k^T) B~_R(K) (1 + K^2/(4 xi^2))
  exp(-K^2/(4 xi^2)) f^(k) e^{i k.x} over the wavenumbers of a Lx x Ly box, the
  self term -(1 + gamma/2 + ln xi) f_i and the constant (1 - ln R) sum(f); pinned
  against the direct free-space sum (the FMM test's reference,
  FMM/tests/stokes_SLP_FMM_test.m:29-51).
"""
import numpy as np
from scipy import special

EULER_GAMMA = 0.5772156649015329


def bhat_trunc(K, R):
    """Transform of the biharmonic kernel B(r) = r^2 (ln r - 1)/(8 pi) minus the
    quadratic that matches its value and slope at R, truncated to r < R:
    B~(r) = B(r) - B(R) - beta (r^2 - R^2), beta = B'(R)/(2 R) = (2 ln R - 1)/(16 pi),
      int_{|r| < R} B~(r) e^{-i k.r} dr = (1 - J0(K R))/K^4 - R J1(K R)/(2 K^3)
    (the closed form of 1/4 int_0^R r^3 (ln r - 1) J0 minus the quadratic's
    transform; the R^3 J1/K and R^2 J2/K^2 terms cancel), and for K R < 2 its
    power series R^4/16 sum_{m>=2} (-1)^m (m - 1) (K R/2)^{2m-4}/(m!)^2.
    (d_i d_j - delta_ij Lap) of the quadratic is the constant -2 beta delta_ij,
    which the sum's constant term takes back."""
    K = np.asarray(K, dtype=np.float64)
    out = np.empty_like(K)
    x = K * R
    sm = x < 2
    t = (x[sm] / 2) ** 2
    s = np.zeros_like(t)
    c = np.ones_like(t)
    fac = 1.0
    for m in range(2, 24):
        fac *= m
        s += (-1) ** m * (m - 1) * c / fac ** 2
        c = c * t
    out[sm] = R ** 4 / 16 * s
    Kb, xb = K[~sm], x[~sm]
    out[~sm] = (1 - special.j0(xb)) / Kb ** 4 - R * special.j1(xb) / (2 * Kb ** 3)
    return out


def free_constant(R):
    """Constant term of the free-space sum per unit total density (u units):
    1/(8 pi) from the Stokeslet's own constant and -2 beta from the quadratic
    taken out of B: (1 - ln R)/(4 pi)."""
    return (1 - np.log(R)) / (4 * np.pi)


def slp_free_direct(p, f):
    """sum_{j != i} (-1/2 log r^2 f_j + (r.f_j) r / r^2)/(4 pi) at the sources."""
    u = np.zeros_like(f)
    for i in range(len(p)):
        r = p[i] - p
        r2 = (r ** 2).sum(1)
        m = r2 > 0
        rf = (r[m] * f[m]).sum(1)
        u[i] = (-0.5 * np.log(r2[m])[:, None] * f[m] + (rf / r2[m])[:, None] * r[m]).sum(0)
    return u / (4 * np.pi)


def slp_free(p, f, xi, delta=None, kcut=40.0):
    """The free-space Ewald sum at the sources p (N x 2, any position) for a
    density f, Ewald parameter xi; delta (default (sqrt(kcut) + 1)/xi) the margin
    of the truncation radius R = diameter + delta and of the box; wavenumbers
    with K^2/(4 xi^2) <= kcut."""
    lo, hi = p.min(0), p.max(0)
    wx, wy = hi - lo
    if delta is None:
        delta = (np.sqrt(kcut) + 1) / xi
    R = np.hypot(wx, wy) + delta
    Lx, Ly = wx + R + delta, wy + R + delta
    kmax = 2 * xi * np.sqrt(kcut)
    mx = int(np.ceil(kmax * Lx / (2 * np.pi)))
    my = int(np.ceil(kmax * Ly / (2 * np.pi)))
    kx = 2 * np.pi / Lx * np.arange(-mx, mx + 1)
    ky = 2 * np.pi / Ly * np.arange(-my, my + 1)
    KX, KY = np.meshgrid(kx, ky, indexing="ij")
    KX, KY = KX.ravel(), KY.ravel()
    K2 = KX ** 2 + KY ** 2
    a = 0.25 / xi ** 2
    H = (1 + a * K2) * np.exp(-a * K2)
    B = np.zeros_like(K2)
    nz = K2 > 0
    B[nz] = bhat_trunc(np.sqrt(K2[nz]), R)
    # real space: all pairs, Hasimoto kernel (4 pi scaled)
    ur = np.zeros_like(f)
    for i in range(len(p)):
        r = p[i] - p
        r2 = (r ** 2).sum(1)
        m = r2 > 0
        x = xi ** 2 * r2[m]
        E, e = special.exp1(x), np.exp(-x)
        rf = (r[m] * f[m]).sum(1)
        ur[i] = (0.5 * E[:, None] * f[m] + e[:, None] * ((rf / r2[m])[:, None] * r[m] - f[m])).sum(0)
    # k space by direct Fourier sums
    c = p - lo
    Ex = np.exp(-1j * (np.outer(KX, c[:, 0]) + np.outer(KY, c[:, 1])))
    fh1, fh2 = Ex @ f[:, 0], Ex @ f[:, 1]
    kd = KX * fh1 + KY * fh2
    g1 = 4 * np.pi * (K2 * fh1 - KX * kd) * B * H
    g2 = 4 * np.pi * (K2 * fh2 - KY * kd) * B * H
    uk = np.stack([(np.conj(Ex).T @ g1).real, (np.conj(Ex).T @ g2).real], 1) / (Lx * Ly)
    self = -(1 + EULER_GAMMA / 2 + np.log(xi)) * f
    return (ur + uk + self) / (4 * np.pi) + free_constant(R) * f.sum(0)


def _wavenum(M, L):
    j = np.arange(M)
    return 2 * np.pi / L * np.where(j <= M // 2, j, j - M)


def slp_free_grid(p, f, prm, Lx, Ly, O):
    """The free-space sum at the sources p (N x 2) with the parameters of
    se2p_free_params (prm dict, box Lx x Ly; p is centred here), O the oracle
    module (oracle/oracle.py)."""
    lo, hi = p.min(0), p.max(0)
    pc = np.ascontiguousarray(p - 0.5 * (lo + hi))
    xi, eta, w, P = prm["xi"], prm["eta"], prm["w"], prm["P"]
    Mx, My, R = prm["Mx"], prm["My"], prm["free_R"]
    ur = O.slp_real(pc, pc, f, xi, prm["nside_x"], prm["nside_y"], Lx, Ly)
    H1, H2 = O.spread(pc, f, Lx, Ly, xi, w, eta, P, Mx, My)
    F1, F2 = np.fft.fft2(H1), np.fft.fft2(H2)
    k1 = _wavenum(Mx, Lx)[:, None]
    k2 = _wavenum(My, Ly)[None, :]
    K2 = k1 ** 2 + k2 ** 2
    K2[0, 0] = 1.0
    a = 0.25 / xi ** 2
    g = bhat_trunc(np.sqrt(K2).ravel(), R).reshape(K2.shape) * (1 + a * K2) * \
        np.exp(-(1 - eta) * a * K2)
    g[0, 0] = 0.0
    d = k1 * F1 + k2 * F2
    h1 = np.ascontiguousarray(np.fft.ifft2((K2 * F1 - k1 * d) * g).real)
    h2 = np.ascontiguousarray(np.fft.ifft2((K2 * F2 - k2 * d) * g).real)
    uk = np.stack([O.gather(h1, pc, Lx, Ly, xi, w, eta, P),
                   O.gather(h2, pc, Lx, Ly, xi, w, eta, P)], 1)
    return ur + uk + free_constant(R) * f.sum(0)
