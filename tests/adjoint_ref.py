"""CPU restatement of the builder-defined reconstruction loss and its gradient (test
infrastructure): numpy for the spectra and the sensitivity formulas, the oracle for the
reference-path stages (calcdefp's right-hand side, the multigrid solve, nsmooth, defp2delta).

Forward map at def = 0 (the first outer iteration of mmh/main.f90:95-98):
  rhs   = calcdefp's right-hand side from rho                         mmh/limiter.f90:134-392
  phi   = mean-free solution of L(rho) phi = rhs - mean(rhs)           multigrid, solved to tol
  defp  = nsmooth^2(phi)                                               mmh/limiter.f90:421-423
  delta = defp2delta(defp)                                             mmh/reconstruct.f90:1-41
  P_b   = sum over DFT modes k in bin b of |delta_k|^2 / (N_b N^3)     (numpy fftn; 16 log bins)
  loss  = sum_b w_b P_b
The binning here uses numpy's full-grid FFT, independent of the NR packing the device uses."""
import numpy as np

import refbind as R


def log_edges(n, nb):
    k0, k1 = 2 * np.pi / n, np.pi
    e = k0 * np.exp(np.log(k1 / k0) * np.arange(nb + 1) / nb)
    e[0], e[-1] = k0, k1
    return e


def mode_bins(n, nb):
    """Bin index of every DFT mode of an n^3 grid (Fortran order of the field), -1 outside."""
    m = np.fft.fftfreq(n, 1.0 / n)
    mx, my, mz = np.meshgrid(m, m, m, indexing="ij")
    kk = 2 * np.pi * np.sqrt(mx ** 2 + my ** 2 + mz ** 2) / n
    e = log_edges(n, nb)
    b = np.searchsorted(e[1:nb], kk, side="right")
    b[(kk < e[0] * (1 - 1e-12)) | (kk > e[nb] * (1 + 1e-12))] = -1
    return b


def fortran3(x, n):
    return x.reshape((n, n, n), order="F")


def power(delta, n, nb):
    F = np.fft.fftn(fortran3(delta, n))
    b = mode_bins(n, nb)
    pk, nm = np.zeros(nb), np.zeros(nb)
    for q in range(nb):
        sel = b == q
        nm[q] = sel.sum()
        pk[q] = (np.abs(F[sel]) ** 2).sum() / (nm[q] * n ** 3) if nm[q] else 0.0
    return pk, nm


def aos(rho):
    u = np.zeros((rho.size, 5))
    u[:, 0] = rho
    return u.ravel()


def solve(rhs, rho, n, tol, guess=None):
    phi = np.zeros(n ** 3) if guess is None else guess.copy()
    b = rhs.copy()
    rc, rel = R.or_solve(phi, b, np.zeros(n ** 3), aos(rho), 5, n, tol=tol, max_cycles=250, schedule=0)
    # FP64 V-cycles level off near 1e-14 (16^3) to 2e-13 (32^3); that floor counts as converged
    assert rel[-1] <= max(tol, 3e-13), (rc, rel[-3:])
    return phi


def smooth(x, n, times):
    y = x.copy()
    for _ in range(times):
        R.or_nsmooth(y, n)
    return y


def forward(rho, n, w, tol=1e-13):
    rhs = R.or_limiter_rhs(np.zeros(n ** 3), aos(rho), n)
    phi = solve(rhs, rho, n, tol)
    delta = R.or_delta(smooth(phi, n, 2), n)
    pk, nm = power(delta, n, len(w))
    return float(np.dot(w, pk)), pk, nm, phi, delta


def gradient(rho, n, w, tol=1e-13):
    loss, pk, nm, phi, delta = forward(rho, n, w, tol)
    b = mode_bins(n, len(w))
    W = np.zeros(b.shape)
    for q in range(len(w)):
        if nm[q]:
            W[b == q] = w[q] / (nm[q] * n ** 3)
    # d loss / d delta_x = 2 sum_k W_k Re(F_k e^{ikx}) = 2 N^3 ifftn(W F)_x (W even in k)
    g_delta = (2 * n ** 3 * np.fft.ifftn(W * np.fft.fftn(fortran3(delta, n))).real).ravel(order="F")
    g_phi = smooth(R.or_delta(g_delta, n), n, 2)
    lam = solve(g_phi, rho, n, tol)
    L, P = fortran3(lam, n), fortran3(phi, n)
    g_bar = np.zeros_like(L)
    for ax in range(3):
        for s in (1, -1):
            g_bar += (np.roll(L, -s, ax) - L) * (np.roll(P, -s, ax) - P)
    g_bar *= 0.5
    g_op = np.zeros_like(g_bar)
    for a in (-1, 0, 1):
        for bb in (-1, 0, 1):
            for c in (-1, 0, 1):
                g_op += np.roll(g_bar, (a, bb, c), (0, 1, 2))
    g_op /= 27.0
    sm = smooth(lam, n, 4)
    g_rhs = np.where(np.abs(1.0 - rho) > 2.0, 0.0, -(sm / 30.0))
    return loss, g_op.ravel(order="F") + g_rhs


def fd_central(rho, n, w, cells, eps, tol=1e-14):
    """Central differences of the loss at the given cells, evaluated without cancellation: each
    perturbed solve starts from the unperturbed solution, and loss(+) - loss(-) is formed as
    sum_k W_k Re(d_k conj(s_k)) from the difference d and sum s of the two reconstructed
    densities (the loss is a quadratic form in delta)."""
    loss, pk, nm, phi0, _ = forward(rho, n, w, tol)
    b = mode_bins(n, len(w))
    W = np.zeros(b.shape)
    for q in range(len(w)):
        if nm[q]:
            W[b == q] = w[q] / (nm[q] * n ** 3)
    out = []
    for x in cells:
        deltas = []
        for sgn in (1, -1):
            r = rho.copy()
            r[x] += sgn * eps
            rhs = R.or_limiter_rhs(np.zeros(n ** 3), aos(r), n)
            deltas.append(solve(rhs, r, n, tol, guess=phi0))
        dphi = deltas[0] - deltas[1]
        sphi = deltas[0] + deltas[1]
        dd = np.fft.fftn(fortran3(R.or_delta(smooth(dphi, n, 2), n), n))
        sd = np.fft.fftn(fortran3(R.or_delta(smooth(sphi, n, 2), n), n))
        out.append(float((W * (dd * np.conj(sd)).real).sum()) / (2 * eps))
    return np.array(out)
