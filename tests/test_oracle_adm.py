"""Pins the oracle's point algebra (oracle/oracle.c:or_point_compute, a
restatement of BSSN::set_bd_values + ev_* in components/bssn/bssn.cc and the
macros of bssn_macros.h) against brute-force ADM general relativity.

The reference's BSSN translation unit cannot be built here (it includes
<fftw3.h>), so no reference run pins the right-hand side.  Instead, analytic
smooth fields are differentiated exactly with torch autodiff (float64):

  * physical metric g_ij = e^{4 phi} gt_ij, extrinsic curvature
    K_ij = e^{4 phi} (At_ij + gt_ij K/3);
  * Christoffels, Ricci tensor, D_i D_j alpha of g by brute force;
  * ADM evolution  d_t g_ij = -2 alpha K_ij,
    d_t K_ij = -D_iD_j alpha + alpha (R_ij - 2 K_ik K^k_j + K K_ij)
               - 8 pi alpha (S_ij - g_ij (S - rho)/2);
  * the BSSN variables' time derivatives by the chain rule, plus the two
    constraint additions that define BSSN: d_t K gets -alpha H, and
    d_t Gt^i gets 2 alpha e^{4 phi} M^i (H, M^i the physical Hamiltonian and
    momentum constraints, also by brute force).

The only term the expected values take from the reference text rather than
from GR is the omission of DIFFalpha K_FRW^2/3 in ev_DIFFK (bssn.cc:778-795
writes alpha/3 DIFFK (DIFFK + 2 K_FRW) where GR has (alpha K^2 - K_FRW^2)/3);
it is added explicitly below.  A dropped term, a wrong sign or a transposed
index anywhere in the oracle's algebra fails this test.
"""
import math

import numpy as np
import pytest

torch = pytest.importorskip("torch")
from torch.func import jacfwd  # noqa: E402

torch.set_default_dtype(torch.float64)
PI = math.pi
SYM = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]


def sym_to_mat(v):
    return torch.stack([torch.stack([v[0], v[1], v[2]]), torch.stack([v[1], v[3], v[4]]),
                        torch.stack([v[2], v[4], v[5]])])


def mat_to_sym(m):
    return torch.stack([m[a, b] for a, b in SYM])


# Explicit 3x3 determinant and inverse: torch.func's forward-over-forward
# derivative of torch.det / torch.linalg.* is not exact, elementary ops are.
def det3(m):
    return (m[0, 0] * (m[1, 1] * m[2, 2] - m[1, 2] * m[2, 1]) - m[0, 1] * (m[1, 0] * m[2, 2] - m[1, 2] * m[2, 0])
            + m[0, 2] * (m[1, 0] * m[2, 1] - m[1, 1] * m[2, 0]))


def inv3(m):
    c = torch.stack([
        torch.stack([m[1, 1] * m[2, 2] - m[1, 2] * m[2, 1], m[0, 2] * m[2, 1] - m[0, 1] * m[2, 2],
                     m[0, 1] * m[1, 2] - m[0, 2] * m[1, 1]]),
        torch.stack([m[1, 2] * m[2, 0] - m[1, 0] * m[2, 2], m[0, 0] * m[2, 2] - m[0, 2] * m[2, 0],
                     m[0, 2] * m[1, 0] - m[0, 0] * m[1, 2]]),
        torch.stack([m[1, 0] * m[2, 1] - m[1, 1] * m[2, 0], m[0, 1] * m[2, 0] - m[0, 0] * m[2, 1],
                     m[0, 0] * m[1, 1] - m[0, 1] * m[1, 0]])])
    return c / det3(m)


class Data:
    """Analytic test data; amplitude `amp` sets the nonlinearity."""

    def __init__(self, seed, amp=0.05, dgam_amp=0.0):
        rng = np.random.default_rng(seed)
        self.k = torch.tensor(rng.integers(1, 3, size=(40, 3)) * rng.choice([-1, 1], size=(40, 3)),
                              dtype=torch.float64)
        self.ph = torch.tensor(rng.uniform(0, 2 * PI, 40))
        self.c = torch.tensor(rng.standard_normal(40)) * amp
        self.dgam_amp = dgam_amp
        self.phi_F, self.K_F, self.rho_F, self.S_F = 0.031, -2.9, 0.12, 0.013

    def w(self, x, n):
        """n-th smooth scalar function."""
        return self.c[n] * torch.sin(2 * PI * (self.k[n] @ x) + self.ph[n]) + \
            0.5 * self.c[n + 20] * torch.cos(2 * PI * (self.k[n + 20] @ x) + self.ph[n + 20])

    def gt(self, x):  # unit-determinant conformal metric
        m = torch.eye(3) + sym_to_mat(torch.stack([self.w(x, q) for q in range(6)]))
        return m / det3(m) ** (1.0 / 3.0)

    def phi(self, x):
        return self.phi_F + self.w(x, 6)

    def K(self, x):
        return self.K_F + self.w(x, 7)

    def alpha(self, x):
        return 1.0 + self.w(x, 8)

    def At(self, x):
        raw = sym_to_mat(torch.stack([self.w(x, 9 + q) for q in range(6)]))
        g = self.gt(x)
        return raw - g * torch.trace(inv3(g) @ raw) / 3.0

    def Gam(self, x):  # Gt^i = -d_j gt^{ij}, plus an optional constraint violation
        J = jacfwd(lambda y: inv3(self.gt(y)))(x)  # [i][j][d]
        G = -torch.stack([J[i, :, :].diagonal().sum() for i in range(3)])
        if self.dgam_amp:
            G = G + self.dgam_amp * torch.stack([self.w(x, 15 + i) for i in range(3)])
        return G

    def beta(self, x):  # shift vector beta^i
        return torch.stack([0.6 * self.w(x, 15 + q) + 0.3 * self.w(x, 4 + q) for q in range(3)])

    def expN(self, x):
        return 0.3 * self.w(x, 19)

    # matter (physical, lower indices)
    def rho(self, x):
        return self.rho_F + 0.2 * self.w(x, 18)

    def S(self, x):
        return self.S_F + 0.2 * self.w(x, 19)

    def Si(self, x):
        return torch.stack([0.1 * self.w(x, 3 + q) * self.w(x, 12 + q) + 0.05 * self.w(x, 16 + q) for q in range(3)])

    def STF(self, x):  # trace-free w.r.t. the physical metric
        raw = sym_to_mat(torch.stack([0.3 * self.w(x, 10 + q) for q in range(6)]))
        g = self.g(x)
        return raw - g * torch.trace(inv3(g) @ raw) / 3.0

    # physical ADM variables
    def g(self, x):
        return torch.exp(4 * self.phi(x)) * self.gt(x)

    def Kij(self, x):
        return torch.exp(4 * self.phi(x)) * (self.At(x) + self.gt(x) * self.K(x) / 3.0)


def christoffel(gfun, x):
    g = gfun(x)
    gi = inv3(g)
    dg = jacfwd(gfun)(x)  # [a][b][c] = d_c g_ab
    t = dg.permute(2, 0, 1)  # [c][a][b]
    # Gamma^i_jk = 1/2 g^{im} (d_j g_mk + d_k g_mj - d_m g_jk)
    low = 0.5 * (t.permute(1, 0, 2) + t.permute(1, 2, 0) - t)  # [m][j][k]
    return torch.einsum("im,mjk->ijk", gi, low)


def ricci(gfun, x):
    Gm = christoffel(gfun, x)
    dG = jacfwd(lambda y: christoffel(gfun, y))(x)  # [k][i][j][d]
    R = torch.einsum("kijk->ij", dG) - torch.einsum("kikj->ij", dG) \
        + torch.einsum("kkm,mij->ij", Gm, Gm) - torch.einsum("kjm,mik->ij", Gm, Gm)
    return R


def adm_rhs(d, x, shift=False):
    """(d_t g_ij, d_t K_ij) of ADM with lapse alpha and zero shift, or with
    shift beta^i the Lie derivatives L_beta g_ij and L_beta K_ij added."""
    g, K, al = d.g(x), d.Kij(x), d.alpha(x)
    gi = inv3(g)
    Gm = christoffel(d.g, x)
    dal = jacfwd(d.alpha)(x)
    ddal = jacfwd(jacfwd(d.alpha))(x)
    DDal = ddal - torch.einsum("kij,k->ij", Gm, dal)
    R = ricci(d.g, x)
    trK = torch.einsum("ij,ij->", gi, K)
    KK = K @ gi @ K
    Sij = d.STF(x) + g * d.S(x) / 3.0
    dtg = -2 * al * K
    dtK = -DDal + al * (R - 2 * KK + trK * K) - 8 * PI * al * (Sij - 0.5 * g * (d.S(x) - d.rho(x)))
    if shift:
        b = d.beta(x)
        db = jacfwd(d.beta)(x)  # [k][i] = d_i beta^k
        for T, dT in ((g, jacfwd(d.g)(x)), (K, jacfwd(d.Kij)(x))):  # dT[a][b][k] = d_k T_ab
            lie = torch.einsum("k,abk->ab", b, dT) + torch.einsum("kb,ka->ab", T, db) + torch.einsum("ak,kb->ab", T, db)
            if T is g:
                dtg = dtg + lie
            else:
                dtK = dtK + lie
    return dtg, dtK


def constraints(d, x):
    """Physical H = R + K^2 - K_ij K^ij - 16 pi rho and M^i = D_j K^ij - g^ij d_j K - 8 pi S^i."""
    g, K = d.g(x), d.Kij(x)
    gi = inv3(g)
    R = torch.einsum("ij,ij->", gi, ricci(d.g, x))
    trK = torch.einsum("ij,ij->", gi, K)
    H = R + trK**2 - torch.einsum("ij,ij->", gi @ K @ gi, K) - 16 * PI * d.rho(x)
    Kup = lambda y: inv3(d.g(y)) @ d.Kij(y) @ inv3(d.g(y))  # noqa: E731
    dKup = jacfwd(Kup)(x)  # [i][j][d]
    Gm = christoffel(d.g, x)
    KU = Kup(x)
    divK = torch.einsum("ijj->i", dKup) + torch.einsum("ijm,mj->i", Gm, KU) + torch.einsum("jjm,im->i", Gm, KU)
    dtrK = jacfwd(lambda y: torch.einsum("ij,ij->", inv3(d.g(y)), d.Kij(y)))(x)
    M = divK - gi @ dtrK - 8 * PI * gi @ d.Si(x)
    return H, M, R


def oracle_point(po, params, d, x, shift=False):
    pt = po.Point()
    gt, At = d.gt(x), d.At(x)
    for q, (a, b) in enumerate(SYM):
        pt.v[q] = float(gt[a, b] - (1.0 if a == b else 0.0))
        pt.v[7 + q] = float(At[a, b])
    pt.v[6] = float(d.phi(x)) - d.phi_F
    pt.v[13] = float(d.K(x)) - d.K_F
    G = d.Gam(x)
    for i in range(3):
        pt.v[14 + i] = float(G[i])
    pt.v[17] = float(d.alpha(x)) - 1.0
    src = [float(d.rho(x)) - d.rho_F, float(d.S(x)) - d.S_F] + [float(v) for v in d.Si(x)] + \
        [float(d.STF(x)[a, b]) for a, b in SYM]
    for q, v in enumerate(src):
        pt.src[q] = v
    dg = jacfwd(d.gt)(x)                  # [a][b][k]
    ddg = jacfwd(jacfwd(d.gt))(x)         # [a][b][k][l]
    dphi = jacfwd(d.phi)(x)
    ddphi = jacfwd(jacfwd(d.phi))(x)
    dal = jacfwd(d.alpha)(x)
    ddal = jacfwd(jacfwd(d.alpha))(x)
    dK = jacfwd(d.K)(x)
    dG = jacfwd(d.Gam)(x)                 # [k][j]
    dA = jacfwd(d.At)(x)                  # [a][b][j]
    for k in range(3):
        for q, (a, b) in enumerate(SYM):
            pt.dg[k][q] = float(dg[a, b, k])
            pt.dA[k][q] = float(dA[a, b, k])
        pt.dphi[k] = float(dphi[k])
        pt.dalpha[k] = float(dal[k])
        pt.dK[k] = float(dK[k])
        for j in range(3):
            pt.dGam[j][k] = float(dG[k, j])
    for s, (k, l) in enumerate(SYM):
        for q, (a, b) in enumerate(SYM):
            pt.ddg[s][q] = float(ddg[a, b, k, l])
        pt.ddphi[s] = float(ddphi[k, l])
        pt.ddalpha[s] = float(ddal[k, l])
    pt.phi_F, pt.K_F, pt.rho_F, pt.S_F = d.phi_F, d.K_F, d.rho_F, d.S_F
    if shift:
        # shift inputs: beta^i, d_j beta^i, d_k d_l beta^i, d expN, and the
        # advection term beta^k d_k f of every field (exact derivatives where
        # the grid driver uses upwind stencils)
        b = d.beta(x)
        db = jacfwd(d.beta)(x)               # [i][j] = d_j beta^i
        ddb = jacfwd(jacfwd(d.beta))(x)      # [i][k][l]
        for i in range(3):
            pt.xv[i] = float(b[i])
            for j in range(3):
                pt.dbeta[j][i] = float(db[i, j])
            for s_, (k, l) in enumerate(SYM):
                pt.ddbeta[i][s_] = float(ddb[i, k, l])
        pt.xv[3] = float(d.expN(x))
        dN = jacfwd(d.expN)(x)
        for k in range(3):
            pt.dexpN[k] = float(dN[k])
        for q, (a, c) in enumerate(SYM):
            pt.adv[q] = float(b @ dg[a, c])
            pt.adv[7 + q] = float(b @ dA[a, c])
        pt.adv[6] = float(b @ dphi)
        pt.adv[13] = float(b @ dK)
        for i in range(3):
            pt.adv[14 + i] = float(b @ dG[i])
        pt.adv[17] = float(b @ dal)
    po.point_compute(params, pt)
    return pt


def expected(d, x, lapse="Static", shift=False):
    """BSSN right-hand sides from ADM + chain rule (+ constraint additions)."""
    phi = d.phi(x)
    e4 = torch.exp(4 * phi)
    al = d.alpha(x)
    dtg, dtK = adm_rhs(d, x, shift)
    g, K = d.g(x), d.Kij(x)
    gi = inv3(g)
    trK = torch.einsum("ij,ij->", gi, K)
    dtgi = -gi @ dtg @ gi
    dt_trK_adm = torch.einsum("ij,ij->", gi, dtK) + torch.einsum("ij,ij->", dtgi, K)
    dtphi = torch.einsum("ij,ij->", gi, dtg) / 12.0
    dtgt = (dtg - 4 * dtphi * g) / e4
    At = (K - g * trK / 3) / e4
    dtAt = -4 * dtphi * At + (dtK - dtg * trK / 3 - g * dt_trK_adm / 3) / e4
    H, M, R = constraints(d, x)
    dt_trK = dt_trK_adm - al * H
    # d_t Gt^i = -d_j (d_t gt^{ij}) + 2 alpha e^{4 phi} M^i
    def dtgt_up(y):
        dtg_y, _ = adm_rhs(d, y, shift)
        g_y = d.g(y)
        gi_y = inv3(g_y)
        dtphi_y = torch.einsum("ij,ij->", gi_y, dtg_y) / 12.0
        e4y = torch.exp(4 * d.phi(y))
        dtgt_y = (dtg_y - 4 * dtphi_y * g_y) / e4y
        gti = inv3(d.gt(y))
        return -gti @ dtgt_y @ gti
    J = jacfwd(dtgt_up)(x)  # [i][j][d]
    dtGam = -torch.einsum("ijj->i", J) + 2 * al * e4 * M
    # DIFF variables: subtract the FRW background (FRW.h:93-129) and add the
    # reference's omission of DIFFalpha K_FRW^2/3 in ev_DIFFK
    dtK_frw = d.K_F**2 / 3 + 4 * PI * (d.rho_F + d.S_F)
    out = np.zeros(18)
    for q, (a, b) in enumerate(SYM):
        out[q] = float(dtgt[a, b])
        out[7 + q] = float(dtAt[a, b])
    out[6] = float(dtphi) - (-d.K_F / 6)
    out[13] = float(dt_trK) - dtK_frw - (float(al) - 1.0) * d.K_F**2 / 3
    for i in range(3):
        out[14 + i] = float(dtGam[i])
    out[17] = 0.0 if lapse == "Static" else float(-al * trK)
    if shift:
        # the lapse is advected too (bssn.cc:822-831)
        out[17] += float(d.beta(x) @ jacfwd(d.alpha)(x))
    return out, float(H), M, float(R)


@pytest.mark.parametrize("seed", [1, 2, 3])
@pytest.mark.parametrize("lapse", ["Static", "Harmonic"])
def test_point_rhs_matches_adm(oracle_mod, seed, lapse):
    d = Data(seed)
    x = torch.tensor(np.random.default_rng(100 + seed).uniform(0, 1, 3))
    params = oracle_mod.make_params(16, order=4, lapse=lapse)
    pt = oracle_point(oracle_mod, params, d, x)
    want, H, M, R = expected(d, x, lapse)
    got = np.array(pt.rhs[:])
    scale = np.abs(want).max()
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-11 * scale)
    assert abs(pt.ricci - R) <= 1e-11 * max(1.0, abs(R))
    # Hamiltonian constraint as the reference reports it (bssn.cc:1081-1083)
    phi = float(d.phi(x))
    assert abs(pt.H - (-math.exp(5 * phi) / 8 * H)) <= 1e-11 * max(1.0, abs(H))
    # momentum constraint lowered with the physical metric, times e^{6 phi}
    Mlow = (d.g(x) @ M).detach().numpy() * math.exp(6 * phi)
    np.testing.assert_allclose(oracle_mod.point_momentum(pt), Mlow, rtol=0, atol=1e-11 * max(1, np.abs(Mlow).max()))


def test_point_rhs_gamma_constraint_violation(oracle_mod):
    """With Gt^i off its constraint surface, the conformal Ricci differs from
    the metric's by gt_{k(i} d_{j)} dGt^k only (TERM2 of
    BSSN_CALCULATE_RICCI_UNITARY, bssn_macros.h:251-252, uses the evolved
    Gt^i; TERM3 uses the one computed from the metric)."""
    d0, d1 = Data(5), Data(5, dgam_amp=0.02)
    x = torch.tensor([0.31, 0.77, 0.12])
    params = oracle_mod.make_params(16, order=4)
    p0 = oracle_point(oracle_mod, params, d0, x)
    p1 = oracle_point(oracle_mod, params, d1, x)
    dd = jacfwd(lambda y: d1.Gam(y) - d0.Gam(y))(x)  # [k][j]
    gt = d0.gt(x)
    dR = 0.5 * (torch.einsum("ki,kj->ij", gt, dd) + torch.einsum("kj,ki->ij", gt, dd))
    e4 = math.exp(4 * float(d0.phi(x)))
    gti = inv3(gt)
    dRtf = dR - gt * torch.einsum("ij,ij->", gti, dR) / 3
    al = float(d0.alpha(x))
    for q, (a, b) in enumerate(SYM):
        assert abs((p1.rhs[7 + q] - p0.rhs[7 + q]) - al * float(dRtf[a, b]) / e4) < 1e-11
    assert abs((p1.ricci - p0.ricci) - float(torch.einsum("ij,ij->", gti, dR)) / e4) < 1e-11


def test_point_damping_terms(oracle_mod):
    """a_adj_amp / k_damping_amp add H-proportional terms (bssn.cc:748-820)."""
    d = Data(4)
    x = torch.tensor([0.2, 0.4, 0.9])
    p0 = oracle_point(oracle_mod, oracle_mod.make_params(16, order=4), d, x)
    par = oracle_mod.make_params(16, order=4, a_adj=0.7, k_damp=0.3)
    p1 = oracle_point(oracle_mod, par, d, x)
    H, dt = p0.H, par.dt
    for q in range(6):
        assert abs((p1.rhs[q] - p0.rhs[q]) - 0.5 * 0.7 * dt * H * p0.v[q]) < 1e-15
        assert abs((p1.rhs[7 + q] - p0.rhs[7 + q]) + 0.7 * dt * p0.v[7 + q] * H) < 1e-15
    assert abs((p1.rhs[13] - p0.rhs[13]) + 0.3 * H * math.exp(-5 * p0.phi)) < 1e-13
    assert abs((p1.rhs[6] - p0.rhs[6]) - 0.1 * 0.7 * dt * H) < 1e-15


@pytest.mark.parametrize("order,expect", [(4, 16.0), (6, 64.0)])  # measured: 14.2, 47.2
def test_grid_rhs_converges_to_point_algebra(oracle_mod, order, expect):
    """The oracle's grid driver (stencils wired into the point algebra, as in
    set_bd_values, bssn.cc:361-421) converges at the stencil order to the
    algebra evaluated with exact derivatives: a derivative fed to the wrong
    slot, axis or field leaves an O(1) error and fails the ratio."""
    from torch.func import vmap
    d = Data(7, amp=0.03)
    errs = []
    for n in (16, 32):
        ax = torch.arange(n, dtype=torch.float64) / n
        X = torch.stack(torch.meshgrid(ax, ax, ax, indexing="ij"), -1).reshape(-1, 3)
        gt = vmap(d.gt)(X)
        At = vmap(d.At)(X)
        G = vmap(d.Gam)(X)
        s = oracle_mod.Sim(oracle_mod.make_params(n, order=order, lapse="Harmonic"))
        for q, (a, b) in enumerate(SYM):
            s.field(f"DIFFgamma{a + 1}{b + 1}_a")[:] = (gt[:, a, b] - (1.0 if a == b else 0.0)).reshape(n, n, n).numpy()
            s.field(f"A{a + 1}{b + 1}_a")[:] = At[:, a, b].reshape(n, n, n).numpy()
            s.field(f"STF{a + 1}{b + 1}_a")[:] = vmap(d.STF)(X)[:, a, b].reshape(n, n, n).numpy()
        s.field("DIFFphi_a")[:] = (vmap(d.phi)(X) - d.phi_F).reshape(n, n, n).numpy()
        s.field("DIFFK_a")[:] = (vmap(d.K)(X) - d.K_F).reshape(n, n, n).numpy()
        s.field("DIFFalpha_a")[:] = (vmap(d.alpha)(X) - 1.0).reshape(n, n, n).numpy()
        for i in range(3):
            s.field(f"Gamma{i + 1}_a")[:] = G[:, i].reshape(n, n, n).numpy()
            s.field(f"S{i + 1}_a")[:] = vmap(d.Si)(X)[:, i].reshape(n, n, n).numpy()
        s.field("DIFFr_a")[:] = (vmap(d.rho)(X) - d.rho_F).reshape(n, n, n).numpy()
        s.field("DIFFS_a")[:] = (vmap(d.S)(X) - d.S_F).reshape(n, n, n).numpy()
        s.frw_set(d.phi_F, d.K_F)
        s.frw_add_fluid(d.rho_F, d.S_F / (3 * d.rho_F))
        s.RKEvolve()
        # compare at a few grid points
        e = 0.0
        for (i, j, k) in [(3, 5, 7), (0, n // 2, n - 1), (n // 4, 1, n // 3)]:
            x = torch.tensor([i / n, j / n, k / n])
            pt = oracle_point(oracle_mod, oracle_mod.make_params(n, order=order, lapse="Harmonic"), d, x)
            got = np.array([s.field(f + "_c")[i, j, k] for f in oracle_mod.FIELDS])
            e = max(e, float(np.abs(got - np.array(pt.rhs[:])).max()))
        errs.append(e)
    ratio = errs[0] / errs[1]
    assert ratio > 0.6 * expect, (errs, ratio)


@pytest.mark.parametrize("seed", [1, 2])
@pytest.mark.parametrize("lapse", ["Static", "Harmonic"])
def test_point_rhs_with_shift_matches_adm(oracle_mod, seed, lapse):
    """The USE_BSSN_SHIFT build's right-hand side (bssn_macros.h:292-385,
    bssn.cc:786-828): ADM with the Lie derivatives L_beta g_ij and L_beta K_ij,
    carried through the same chain rule, gives every BSSN field's shift
    terms (advection, d_j beta^i couplings, the second derivatives of beta in
    d_t Gt^i and the divergence in d_t phi).  The expansion e-fold count
    advances by -alpha K/3 plus its advection (bssn.cc:871-875)."""
    d = Data(seed)
    x = torch.tensor(np.random.default_rng(200 + seed).uniform(0, 1, 3))
    params = oracle_mod.make_params(16, order=4, lapse=lapse, shift="Static")
    pt = oracle_point(oracle_mod, params, d, x, shift=True)
    want, H, _, _ = expected(d, x, lapse, shift=True)
    got = np.array(pt.rhs[:])
    scale = np.abs(want).max()
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-11 * scale)
    # the shift terms are not negligible here
    p0 = oracle_point(oracle_mod, oracle_mod.make_params(16, order=4, lapse=lapse), d, x)
    assert np.abs(got - np.array(p0.rhs[:])).max() > 1e-3 * scale
    trK = float(torch.einsum("ij,ij->", inv3(d.g(x)), d.Kij(x)))
    dN = float(d.beta(x) @ jacfwd(d.expN)(x))
    assert abs(pt.rhsx[3] - (dN - float(d.alpha(x)) * trK / 3)) < 1e-12 * max(1.0, abs(trK))
    # Static shift: d_t beta^i = 0 (minus KO, zero here)
    assert all(pt.rhsx[i] == 0.0 for i in range(3))
