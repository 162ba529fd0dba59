"""Independent full-matrix lattice routines in numpy, written from the
geometry (coordinates, periodic shifts, 2x2 complex matrices) rather than
from the reference's offset arithmetic.  Used to pin the checker's staples,
boundary rules and observables."""
import numpy as np


def quat_to_mat(a):
    """[[a0 + i a1, a2 + i a3], [-a2 + i a3, a0 - i a1]] (functions/su2.h:98-100)."""
    a = np.asarray(a, dtype=np.float64)
    m = np.empty(a.shape[:-1] + (2, 2), dtype=np.complex128)
    m[..., 0, 0] = a[..., 0] + 1j * a[..., 1]
    m[..., 0, 1] = a[..., 2] + 1j * a[..., 3]
    m[..., 1, 0] = -a[..., 2] + 1j * a[..., 3]
    m[..., 1, 1] = a[..., 0] - 1j * a[..., 1]
    return m


def mat_to_quat(m):
    a = np.empty(m.shape[:-2] + (4,), dtype=np.float64)
    a[..., 0] = m[..., 0, 0].real
    a[..., 1] = m[..., 0, 0].imag
    a[..., 2] = m[..., 0, 1].real
    a[..., 3] = m[..., 0, 1].imag
    return a


def site_index(Lt, Lx, Ly):
    """index[t, x, y] of the site in the reference layout: parity block
    p = (t + x + y) % 2, half index y//2 + (Ly/2)(x + Lx t)."""
    t, x, y = np.meshgrid(np.arange(Lt), np.arange(Lx), np.arange(Ly), indexing="ij")
    p = (t + x + y) % 2
    h = y // 2 + (Ly // 2) * (x + Lx * t)
    return p * (Lt * Lx * Ly // 2) + h


def field_to_links(phi, Lt, Lx, Ly):
    """U[mu, t, x, y] as 2x2 complex matrices; mu 0 = y, 1 = x, 2 = t."""
    idx = site_index(Lt, Lx, Ly)
    q = np.asarray(phi).reshape(-1, 3, 4)[idx]  # [t, x, y, mu, 4]
    return quat_to_mat(np.moveaxis(q, 3, 0))


def links_to_field(U, Lt, Lx, Ly):
    idx = site_index(Lt, Lx, Ly)
    phi = np.zeros((Lt * Lx * Ly, 3, 4))
    phi[idx] = np.moveaxis(mat_to_quat(U), 0, 3)
    return phi.reshape(-1)


AXIS = {0: 2, 1: 1, 2: 0}  # direction -> array axis of [t, x, y]


def sh(M, mu, s):
    """Shift a field of matrices [t, x, y, 2, 2]."""
    return np.roll(M, -s, axis=AXIS[mu])


def dag(M):
    return np.conj(np.swapaxes(M, -1, -2))


def half_retr(M):
    return 0.5 * np.real(M[..., 0, 0] + M[..., 1, 1])


def plaquette(U, mu, nu):
    """1/2 Re tr U_mu(n) U_nu(n+mu) U_mu(n+nu)^dag U_nu(n)^dag at every n."""
    return half_retr(U[mu] @ sh(U[nu], mu, 1) @ dag(sh(U[mu], nu, 1)) @ dag(U[nu]))


def observables(U, pmc=False):
    """(E, P) of en_av: E = mean -(1 - P_xt); P = mean (P_yx + P_yt + P_xt)/3,
    with PMC the y-x and y-t plaquettes of y = Ly-1 counted as 1."""
    pyx, pyt, pxt = plaquette(U, 0, 1), plaquette(U, 0, 2), plaquette(U, 1, 2)
    if pmc:
        pyx = pyx.copy()
        pyt = pyt.copy()
        pyx[:, :, -1] = 1.0
        pyt[:, :, -1] = 1.0
    return float(np.mean(-(1 - pxt))), float(np.mean((pyx + pyt + pxt) / 3))


def action(U, beta, pmc=False):
    """Wilson action beta * sum (1 - 1/2 tr P) over all plaquettes; with PMC
    the y-plaquettes between y = Ly-1 and 0 are absent."""
    s = 0.0
    for mu, nu in ((0, 1), (0, 2), (1, 2)):
        p = plaquette(U, mu, nu)
        if pmc and mu == 0:
            p = p[:, :, :-1]
        s += np.sum(1 - p)
    return beta * s


def random_su2(shape, rng):
    a = rng.normal(size=shape + (4,))
    a /= np.linalg.norm(a, axis=-1, keepdims=True)
    return quat_to_mat(a)


def gauge_transform(U, g):
    """U_mu(n) -> g(n) U_mu(n) g(n+mu)^dag."""
    return np.stack([g @ U[mu] @ dag(sh(g, mu, 1)) for mu in range(3)])
