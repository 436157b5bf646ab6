"""FP64 reference gradient of one forward stage in a flat periodic box: the
stage restated in torch (CPU, float64) on the periodic interior arrays, so
that torch's autograd differentiates it.

Test infrastructure.  The restatement follows the oracle's stage (oracle_flux,
oracle_emf, oracle_ct_update, oracle_cons of oracle/hamr_oracle.c, which follow
GPU_interp.h:2-111, GPU_program1.cu:791-1068, GPU_phys_mhd.h:432-623,
GPU_const_trans.h:63-530, GPU_fixup.h:75-217) in the flat metric
(g = diag(-1, 1, 1, 1), gdet = 1, lapse 1, no shift, no source); the limiter
and upwind selections enter as masks, so autograd takes the derivative of the
branch the forward takes.  Its forward is checked against the oracle's.

Arrays: P (9, n1, n2, n3) cell-centred primitives, PS (3, n1, n2, n3) lower-face
fields; periodic in every direction.
"""
import torch

RHO, UU, U1, U2, U3, B1, B2, B3, KTOT = range(9)
SMALL = 1e-20


def sh(x, d, k):
    """x shifted along grid axis d (0, 1, 2) so that the result at i is x[i - k]."""
    return torch.roll(x, shifts=k, dims=x.dim() - 3 + d)


def mc(ym, y0, yp):
    Dqm = 2.0 * (y0 - ym)
    Dqp = 2.0 * (yp - y0)
    Dqc = 0.5 * (yp - ym)
    m = torch.minimum(Dqm.abs(), Dqp.abs())
    v = torch.where(Dqc.abs() < m, Dqc.abs(), m)
    v = torch.where(Dqc > 0, v, -v)
    return v * (Dqm * Dqp > 0).double()


def para(x1, x2, x3, x4, x5):
    dq1, dq2, dq3 = mc(x1, x2, x3), mc(x2, x3, x4), mc(x3, x4, x5)
    l = 0.5 * (x3 + x2) - (dq2 - dq1) / 6.0
    r = 0.5 * (x4 + x3) - (dq3 - dq2) / 6.0
    qa = (r - x3) * (x3 - l)
    qd = r - l
    qe = 6.0 * (x3 - 0.5 * (l + r))
    flat = qa <= 0
    l = torch.where(flat, x3, l)
    r = torch.where(flat, x3, r)
    c1 = qd * (qd - qe) < 0
    c2 = (~c1) & (qd * (qd + qe) < 0)
    l2 = torch.where(c1, 3.0 * x3 - 2.0 * r, l)
    r2 = torch.where(c2, 3.0 * x3 - 2.0 * l, r)
    return l2, r2


def state(pr):
    """ucon, ucov, bcon, bcov (lists of 4 tensors) in the flat metric."""
    U = [pr[U1], pr[U2], pr[U3]]
    gam = torch.sqrt(1.0 + U[0] * U[0] + U[1] * U[1] + U[2] * U[2])
    ucon = [gam, U[0], U[1], U[2]]
    ucov = [-gam, U[0], U[1], U[2]]
    b0 = pr[B1] * ucov[1] + pr[B2] * ucov[2] + pr[B3] * ucov[3]
    bcon = [b0] + [(pr[B1 + j] + b0 * ucon[j + 1]) / gam for j in range(3)]
    bcov = [-bcon[0], bcon[1], bcon[2], bcon[3]]
    return ucon, ucov, bcon, bcov


def primtoflux(pr, q, d, gam):
    ucon, ucov, bcon, bcov = q
    u = pr[UU]
    P = (gam - 1.0) * u
    w = pr[RHO] + P + u
    bsq = sum(bcon[m] * bcov[m] for m in range(4))
    eta = w + bsq
    ptot = P + 0.5 * bsq
    f = [None] * 9
    f[RHO] = pr[RHO] * ucon[d]
    mhd = [eta * ucon[d] * ucov[j] + (ptot if j == d else 0.0) - bcon[d] * bcov[j] for j in range(4)]
    f[UU] = mhd[0] + f[RHO]
    f[U1], f[U2], f[U3] = mhd[1], mhd[2], mhd[3]
    for j in range(3):
        f[B1 + j] = bcon[j + 1] * ucon[d] - bcon[d] * ucon[j + 1]
    f[KTOT] = f[RHO] * ((gam - 1.0) * u * torch.pow(pr[RHO], -gam))
    if d == 0:
        return f, None, None
    cs2 = gam * (gam - 1.0) * u / w
    va2 = bsq / eta
    cms2 = cs2 + va2 - cs2 * va2
    cms2 = torch.where(cms2 < 0, torch.full_like(cms2, SMALL), cms2)
    cms2 = torch.where(cms2 > 1, torch.ones_like(cms2), cms2)
    Asq, Bsq, AB = 1.0, -1.0, 0.0
    Au, Bu = ucon[d], ucon[0]
    A = Bu * Bu - (Bsq + Bu * Bu) * cms2
    B = 2.0 * (Au * Bu - (AB + Au * Bu) * cms2)
    C = Au * Au - (Asq + Au * Au) * cms2
    discr = B * B - 4.0 * A * C
    pos = discr > 0
    sd = torch.sqrt(torch.where(pos, discr, torch.ones_like(discr))) * pos.double()
    vp = -(-B + sd) / (2.0 * A)
    vm = -(-B - sd) / (2.0 * A)
    vmax = torch.where(vp > vm, vp, vm)
    vmin = torch.where(vp < vm, vp, vm)
    return f, vmax, vmin


def fluxes(P, PS, gam):
    F = []
    for d in range(3):
        pl, pr = [], []
        for k in range(9):
            x = P[k]
            l_m1, r_m1 = para(sh(x, d, 3), sh(x, d, 2), sh(x, d, 1), x, sh(x, d, -1))
            l_0, r_0 = para(sh(x, d, 2), sh(x, d, 1), x, sh(x, d, -1), sh(x, d, -2))
            pl.append(r_m1)
            pr.append(l_0)
        pl[B1 + d] = PS[d]
        pr[B1 + d] = PS[d]
        qL, qR = state(pl), state(pr)
        FL, vxl, vnl = primtoflux(pl, qL, d + 1, gam)
        UL, _, _ = primtoflux(pl, qL, 0, gam)
        FR, vxr, vnr = primtoflux(pr, qR, d + 1, gam)
        UR, _, _ = primtoflux(pr, qR, 0, gam)
        zero = torch.zeros_like(vxl)
        m1 = torch.where(0 > vxl, zero, vxl)
        cmax = torch.where(m1 > vxr, m1, vxr).abs()
        n1 = torch.where(0 > -vnl, zero, -vnl)
        cmin = torch.where(n1 > -vnr, n1, -vnr).abs()
        F.append([(cmax * FL[k] + cmin * FR[k] - cmax * cmin * (UR[k] - UL[k])) / (cmax + cmin + SMALL)
                  for k in range(9)])
    return F


def emf(P, F):
    """Corner EMFs (E1, E2, E3) with the upwind selections of oracle_emf."""
    ucon, _, bcon, _ = state(P)
    E1c = -(ucon[2] * bcon[3] - ucon[3] * bcon[2])
    E2c = -(ucon[3] * bcon[1] - ucon[1] * bcon[3])
    E3c = -(ucon[1] * bcon[2] - ucon[2] * bcon[1])
    F1, F2, F3 = F
    F1R, F2R, F3R = F1[RHO].detach(), F2[RHO].detach(), F3[RHO].detach()
    F1B2, F1B3, F2B1, F2B3, F3B1, F3B2 = F1[B2], F1[B3], F2[B1], F2[B3], F3[B1], F3[B2]
    L = lambda x: (x <= 0).double()
    G = lambda x: (x > 0).double()
    i, j, z = 0, 1, 2
    m = lambda x, *dk: x if not dk else m(sh(x, dk[0], 1), *dk[1:])  # x at g - e_d for each d in dk
    dEL13_1 = E1c - F3B2
    dEL13_2 = m(E1c, j) - m(F3B2, j)
    dER13_1 = F3B2 - m(E1c, z)
    dER13_2 = m(F3B2, j) - m(E1c, j, z)
    dEL12_1 = E1c + F2B3
    dEL12_2 = m(E1c, z) + m(F2B3, z)
    dER12_1 = -F2B3 - m(E1c, j)
    dER12_2 = -m(F2B3, z) - m(E1c, j, z)
    dEL21_1 = E2c - F1B3
    dEL21_2 = m(E2c, z) - m(F1B3, z)
    dER21_1 = F1B3 - m(E2c, i)
    dER21_2 = m(F1B3, z) - m(E2c, i, z)
    dEL23_1 = E2c + F3B1
    dEL23_2 = m(E2c, i) + m(F3B1, i)
    dER23_1 = -F3B1 - m(E2c, z)
    dER23_2 = -m(F3B1, i) - m(E2c, i, z)
    dEL31_1 = E3c + F1B2
    dEL31_2 = m(E3c, j) + m(F1B2, j)
    dER31_1 = -F1B2 - m(E3c, i)
    dER31_2 = -m(F1B2, j) - m(E3c, i, j)
    dEL32_1 = E3c - F2B1
    dEL32_2 = m(E3c, i) - m(F2B1, i)
    dER32_1 = F2B1 - m(E3c, j)
    dER32_2 = m(F2B1, i) - m(E3c, i, j)
    e1 = 0.25 * ((-F2B3 - (dEL13_1 * L(F2R) + dEL13_2 * G(F2R)))
                 + (-m(F2B3, z) + (dER13_1 * L(m(F2R, z)) + dER13_2 * G(m(F2R, z))))
                 + (F3B2 - (dEL12_1 * L(F3R) + dEL12_2 * G(F3R)))
                 + (m(F3B2, j) + (dER12_1 * L(m(F3R, j)) + dER12_2 * G(m(F3R, j)))))
    e2 = 0.25 * ((-F3B1 - (dEL21_1 * L(F3R) + dEL21_2 * G(F3R)))
                 + (-m(F3B1, i) + (dER21_1 * L(m(F3R, i)) + dER21_2 * G(m(F3R, i))))
                 + (F1B3 - (dEL23_1 * L(F1R) + dEL23_2 * G(F1R)))
                 + (m(F1B3, z) + (dER23_1 * L(m(F1R, z)) + dER23_2 * G(m(F1R, z)))))
    e3 = 0.25 * ((F2B1 - (dEL31_1 * L(F2R) + dEL31_2 * G(F2R)))
                 + (m(F2B1, i) + (dER31_1 * L(m(F2R, i)) + dER31_2 * G(m(F2R, i))))
                 + (-F1B2 - (dEL32_1 * L(F1R) + dEL32_2 * G(F1R)))
                 + (-m(F1B2, j) + (dER32_1 * L(m(F1R, j)) + dER32_2 * G(m(F1R, j)))))
    return e1, e2, e3


def stage(P, PS, Dt, dx, gam):
    """U' (9, n1, n2, n3) of one forward stage."""
    F = fluxes(P, PS, gam)
    E1, E2, E3 = emf(P, F)
    up = lambda x, d: sh(x, d, -1)  # x at g + e_d
    psf0 = PS[0] + (-Dt / dx[1]) * (up(E3, 1) - E3) + (Dt / dx[2]) * (up(E2, 2) - E2)
    psf2 = PS[2] + (-Dt / dx[0]) * (up(E2, 0) - E2) + (Dt / dx[1]) * (up(E1, 1) - E1)
    psf1 = PS[1] + (Dt / dx[0]) * (up(E3, 0) - E3) + (-Dt / dx[2]) * (up(E1, 2) - E1)
    U, _, _ = primtoflux(P, state(P), 0, gam)
    out = []
    for k in range(9):
        v = U[k]
        v = v - Dt * (up(F[0][k], 0) - F[0][k]) / dx[0]
        v = v - Dt * (up(F[1][k], 1) - F[1][k]) / dx[1]
        v = v - Dt * (up(F[2][k], 2) - F[2][k]) / dx[2]
        out.append(v)
    out[B1] = (psf0 + up(psf0, 0)) / 2.0
    out[B2] = (psf1 + up(psf1, 1)) / 2.0
    out[B3] = (psf2 + up(psf2, 2)) / 2.0
    return torch.stack(out)
