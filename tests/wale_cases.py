"""States with a known velocity gradient, shared by the WALE tests.

``synthetic_populations`` builds f = feq(rho, u) + f_neq(S) where
f_neq_q = -(w_q rho tau / cs^2) (c_q c_q - cs^2 I) : S.  Fourth-order
isotropy of D3Q19/D3Q27 (sum_q w_q c_a c_b c_c c_d = cs^4 (d_ab d_cd +
d_ac d_bd + d_ad d_bc)) gives sum_q c c f_neq = -2 cs^2 rho tau S, while
f_neq carries no mass or momentum, so the stress-based strain of such a
state is S exactly and its cached moments are (rho, u).
"""
import numpy as np

from oracle import oracle as O

CS2 = 1.0 / 3.0


def synthetic_populations(Q, rho, u, S, tau):
    """rho (N,), u (N, 3), S (N, 3, 3) symmetric -> f (Q, N), deviation form."""
    c, w, _, _ = O.lattice(Q)
    c = c.astype(np.float64)
    cu = u @ c.T
    usq = (u * u).sum(1)[:, None]
    feq = w[None, :] * ((rho - 1.0)[:, None] + rho[:, None] * (3.0 * cu + 4.5 * cu * cu - 1.5 * usq))
    qab = np.einsum("qa,qb->qab", c, c) - CS2 * np.eye(3)[None]
    fneq = -(w[None, :] * rho[:, None] * tau / CS2) * np.einsum("qab,nab->nq", qab, S)
    return np.ascontiguousarray((feq + fneq).T)


def coords(dims):
    nx, ny, nz = dims
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    return np.stack([x.reshape(-1), y.reshape(-1), z.reshape(-1)], 1).astype(np.float64)


def linear_state(Q, dims, G, tau, seed=0, rho_amp=0.01):
    """u = G (x - x0) over the box, S = sym(G); returns (f, rho, u)."""
    rng = np.random.default_rng(seed)
    X = coords(dims)
    x0 = (np.array(dims, np.float64) - 1.0) / 2.0
    u = (X - x0) @ np.asarray(G, np.float64).T
    n = X.shape[0]
    rho = 1.0 + rho_amp * (rng.random(n) - 0.5)
    S = np.broadcast_to(0.5 * (G + np.asarray(G).T), (n, 3, 3))
    return synthetic_populations(Q, rho, u, S, tau), rho, u


def channel_state(Q, dims, U, A, tau, with_strain):
    """Walls at y = 0 and y = ny-1 (half-way at 0.5 and ny-1.5), the top one
    moving with (U, 0, 0); u_x(y) = U (y - 1/2)/H + A (y - 1/2)(ny - 3/2 - y),
    H = ny - 2.  Returns (f, solid, wall_u, dudy) with dudy the exact
    derivative per cell; the populations carry S_xy = dudy/2 when asked."""
    nx, ny, nz = dims
    X = coords(dims)
    y = X[:, 1]
    H = ny - 2.0
    ux = U * (y - 0.5) / H + A * (y - 0.5) * (ny - 1.5 - y)
    dudy = U / H + A * (ny - 1.0 - 2.0 * y)
    n = X.shape[0]
    solid = ((y == 0) | (y == ny - 1)).astype(np.uint8)
    wall_u = np.zeros((n, 3))
    wall_u[y == ny - 1, 0] = U
    u = np.zeros((n, 3))
    u[:, 0] = np.where(solid == 1, 0.0, ux)
    S = np.zeros((n, 3, 3))
    if with_strain:
        S[:, 0, 1] = S[:, 1, 0] = 0.5 * dudy
    S[solid == 1] = 0.0
    f = synthetic_populations(Q, np.ones(n), u, S, tau)
    return f, solid, wall_u, dudy


def wale_closed_form(G, cw=O.WALE_CW):
    """nu_t of one constant gradient from the definition (les.rs:113-118)
    evaluated with numpy linear algebra."""
    G = np.asarray(G, np.float64)
    S = 0.5 * (G + G.T)
    G2 = G @ G
    Sd = 0.5 * (G2 + G2.T) - np.trace(G2) / 3.0 * np.eye(3)
    ss, sdsd = (S * S).sum(), (Sd * Sd).sum()
    den = ss ** 2.5 + sdsd ** 1.25
    return 0.0 if den == 0 else cw * cw * sdsd ** 1.5 / den
