"""FP64 numpy restatement of the learned per-atom energy model (the
checker of surrogate.hip): radial-Gaussian descriptors with a cosine cutoff,
per-species 2-hidden-layer SiLU MLP, energy and analytic forces."""
import numpy as np

NB, H = 32, 128


def constants(rc):
    mu0 = 0.5
    dmu = (rc - 0.5) / (NB - 1)
    eta = 0.5 / dmu ** 2
    return mu0 + dmu * np.arange(NB), eta


def pairs_within(R, L, rc, mol):
    """All ordered pairs (i, j), i != j, closer than rc under minimum image,
    excluding pairs of the same molecule (rigid body)."""
    d = R[:, None, :] - R[None, :, :]
    d -= L * np.round(d / L)
    r = np.sqrt((d * d).sum(-1))
    m = (r < rc) & (mol[:, None] != mol[None, :])
    i, j = np.nonzero(m)
    return i, j, d[i, j], r[i, j]


def descriptors(R, types, L, rc, mol, nt):
    mu, eta = constants(rc)
    i, j, d, r = pairs_within(R, L, rc, mol)
    fc = 0.5 * (np.cos(np.pi * r / rc) + 1.0)
    G = np.exp(-eta * (r[:, None] - mu[None, :]) ** 2) * fc[:, None]
    D = np.zeros((len(R), nt * NB))
    for t in range(nt):
        sel = types[j] == t
        np.add.at(D[:, t * NB:(t + 1) * NB], i[sel], G[sel])
    return D


def unpack(p, din):
    W1 = p[:din * H].reshape(din, H)
    o = din * H
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    w3 = p[o:o + H]; o += H
    b3 = p[o]
    return W1, b1, W2, b2, w3, b3


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def mlp(params, norm, D, types, nt):
    """Atom energies e and dE/dD (E = sum e)."""
    din = nt * NB
    P = len(params) // nt
    xm, xs = norm[:din], norm[din:2 * din]
    ym, ys = norm[2 * din:2 * din + nt], norm[2 * din + nt:2 * din + 2 * nt]
    X = (D - xm) / xs
    e = np.zeros(len(D))
    dD = np.zeros_like(D)
    for t in range(nt):
        W1, b1, W2, b2, w3, b3 = unpack(params[t * P:(t + 1) * P].astype(np.float64), din)
        sel = types == t
        Z1 = X[sel] @ W1 + b1
        H1 = silu(Z1)
        Z2 = H1 @ W2 + b2
        H2 = silu(Z2)
        out = H2 @ w3 + b3
        e[sel] = ys[t] * out + ym[t]
        G2 = ys[t] * w3[None, :] * dsilu(Z2)
        G1 = (G2 @ W2.T) * dsilu(Z1)
        dD[sel] = (G1 @ W1.T) / xs
    return e, dD


def energy_forces(params, norm, R, types, L, rc, mol, nt):
    mu, eta = constants(rc)
    D = descriptors(R, types, L, rc, mol, nt)
    e, dD = mlp(params, norm, D, types, nt)
    i, j, d, r = pairs_within(R, L, rc, mol)
    fc = 0.5 * (np.cos(np.pi * r / rc) + 1.0)
    dfc = -0.5 * np.pi / rc * np.sin(np.pi * r / rc)
    u = r[:, None] - mu[None, :]
    g = np.exp(-eta * u * u)
    dG = g * (dfc[:, None] - 2.0 * eta * u * fc[:, None])
    # E depends on r_ij through D_i (species of j) and D_j (species of i)
    gi = np.take_along_axis(dD[i].reshape(len(i), nt, NB), types[j][:, None, None], axis=1)[:, 0, :]
    gj = np.take_along_axis(dD[j].reshape(len(j), nt, NB), types[i][:, None, None], axis=1)[:, 0, :]
    dEdr = 0.5 * ((gi + gj) * dG).sum(1)   # each unordered pair appears twice
    F = np.zeros_like(R)
    np.add.at(F, i, -(dEdr / r)[:, None] * d)
    np.add.at(F, j, (dEdr / r)[:, None] * d)
    return e.sum(), F, e, D


def random_params(nt, seed, scale=1.0):
    rng = np.random.default_rng(seed)
    din = nt * NB
    P = din * H + 3 * H + H * H + 1
    p = np.zeros(nt * P, dtype=np.float32)
    for t in range(nt):
        q = p[t * P:(t + 1) * P]
        q[:din * H] = rng.uniform(-1, 1, din * H) * scale * np.sqrt(6.0 / (din + H))
        q[din * H:din * H + H] = rng.uniform(-0.1, 0.1, H)
        o = din * H + H
        q[o:o + H * H] = rng.uniform(-1, 1, H * H) * scale * np.sqrt(6.0 / (2 * H))
        o += H * H
        q[o:o + H] = rng.uniform(-0.1, 0.1, H)
        o += H
        q[o:o + H] = rng.uniform(-1, 1, H) * np.sqrt(6.0 / (H + 1))
        q[o + H] = 0.3
    return p


def loss_grad(params, norm, D, types, y, nt):
    """Per-species gradient of the per-atom MSE of normalised labels,
    loss_t = mean over atoms of species t of (out - (y - ym_t)/ys_t)^2, in
    the parameter layout of params."""
    din = nt * NB
    P = len(params) // nt
    xm, xs = norm[:din], norm[din:2 * din]
    ym, ys = norm[2 * din:2 * din + nt], norm[2 * din + nt:2 * din + 2 * nt]
    X = (D - xm) / xs
    g = np.zeros(len(params))
    for t in range(nt):
        W1, b1, W2, b2, w3, b3 = unpack(params[t * P:(t + 1) * P].astype(np.float64), din)
        sel = types == t
        Xs = X[sel]
        Z1 = Xs @ W1 + b1
        H1 = silu(Z1)
        Z2 = H1 @ W2 + b2
        H2 = silu(Z2)
        out = H2 @ w3 + b3
        yn = (y[sel] - ym[t]) / ys[t]
        s = 2.0 * (out - yn) / sel.sum()
        G2 = s[:, None] * w3[None, :] * dsilu(Z2)
        G1 = (G2 @ W2.T) * dsilu(Z1)
        o = t * P
        g[o:o + din * H] = (Xs.T @ G1).ravel(); o += din * H
        g[o:o + H] = G1.sum(0); o += H
        g[o:o + H * H] = (H1.T @ G2).ravel(); o += H * H
        g[o:o + H] = G2.sum(0); o += H
        g[o:o + H] = H2.T @ s; o += H
        g[o] = s.sum()
    return g


# ---- the exact long-range channel: reciprocal Ewald sum over |k| <= kcut ----

def lowk_vectors(L, kcut):
    """Half-space integer vectors n (first nonzero component positive) with
    |2 pi n / L| <= kcut, n3 fastest (the enumeration of kspace_ewald.f90:115-138)."""
    unit = 2.0 * np.pi / L
    nmax = max(1, int(np.ceil(kcut / unit)))
    out = []
    for n1 in range(0, nmax + 1):
        for n2 in range(-nmax, nmax + 1):
            for n3 in range(-nmax, nmax + 1):
                upper = n1 > 0 or (n1 == 0 and (n2 > 0 or (n2 == 0 and n3 > 0)))
                if upper and unit * unit * n1 * n1 + unit * unit * n2 * n2 + unit * unit * n3 * n3 <= kcut * kcut:
                    out.append((n1, n2, n3))
    return np.array(out, dtype=np.int64).reshape(-1, 3)


def lowk_energy_forces(R, q, L, alpha, kcut, kcoul):
    """E = kcoul sum_k prefac |S(k)|^2 over the half space, prefac =
    4 pi/V exp(-k^2/(4 alpha^2))/k^2 (kspace_ewald.f90:174-180, 287), forces,
    atom shares Re(q_j e^{ik.r_j} conj(kcoul S)) prefac, and the virial
    sum_k E_k (1 - k^2/(2 alpha^2))."""
    n = lowk_vectors(L, kcut)
    k = 2.0 * np.pi * n / L
    k2 = (k * k).sum(1)
    pref = 4.0 * np.pi / L ** 3 * np.exp(-k2 / (4.0 * alpha ** 2)) / k2
    ph = R @ k.T
    c, s = np.cos(ph), np.sin(ph)
    Sc, Ss = q @ c, q @ s
    Ek = kcoul * pref * (Sc * Sc + Ss * Ss)
    F = 2.0 * kcoul * q[:, None] * ((s * Sc - c * Ss) * pref) @ k
    e = kcoul * q * ((c * Sc + s * Ss) * pref).sum(1)
    W = np.sum(Ek * (1.0 - k2 / (2.0 * alpha ** 2)))
    return Ek.sum(), F, e, W
