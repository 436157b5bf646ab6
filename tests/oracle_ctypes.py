"""ctypes bindings of the FP64 CPU oracle (oracle/liboracle.so).

Test infrastructure: the HIP step is tested against the oracle, which is never the thing
measured or shipped.  Arrays follow the layout documented in oracle/hamr_oracle.h
(structure of arrays over the ghost-inclusive global grid, z fastest).
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "liboracle.so")

NPR = 9
NG = 3
BC_PERIODIC, BC_OUTFLOW, BC_POLE = 0, 1, 2

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class Params(C.Structure):
    _fields_ = [
        ("n1", C.c_int), ("n2", C.c_int), ("n3", C.c_int),
        ("bs1", C.c_int), ("bs2", C.c_int), ("bs3", C.c_int),
        ("bc", C.c_int * 3),
        ("gam", C.c_double), ("cour", C.c_double),
        ("dx", C.c_double * 3),
        ("rho_floor", C.c_double), ("u_floor", C.c_double),
        ("riemann", C.c_int),
    ]


class Metric(C.Structure):
    _fields_ = [("gcov", _dp), ("gcon", _dp), ("gdet", _dp), ("conn", _dp), ("rcen", _dp)]


def _ptr(a, t=_dp):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(t)


class Oracle:
    def __init__(self, lib):
        self.lib = lib
        L = lib
        L.oracle_get_state.argtypes = [_dp] * 7
        L.oracle_primtoflux.argtypes = [_dp, _dp, _dp, C.c_double, C.c_double, C.c_int, _dp, _dp, _dp]
        L.oracle_source.argtypes = [_dp, _dp, _dp, C.c_double, _dp, C.c_double, _dp]
        L.oracle_para.argtypes = [_dp, _dp, _dp]
        L.oracle_utoprim.argtypes = [C.c_int, _dp, _dp, _dp, C.c_double, C.c_double, _dp]
        L.oracle_utoprim.restype = C.c_int
        P = C.POINTER(Params)
        M = C.POINTER(Metric)
        L.oracle_flux.argtypes = [P, C.c_int, _dp, _dp, M, _dp, _dp]
        L.oracle_emf.argtypes = [P, _dp, _dp, _dp, _dp, M, C.c_double, _dp, _dp]
        L.oracle_ct_update.argtypes = [P, C.c_double, _dp, _dp, _dp, M]
        L.oracle_update.argtypes = [P, C.c_int, C.c_double, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, M,
                                    _ip, C.POINTER(C.c_longlong)]
        L.oracle_cons.argtypes = [P, C.c_double, _dp, _dp, _dp, _dp, _dp, _dp, _dp, M]
        L.oracle_invert.argtypes = [P, C.c_int, _dp, _dp, _dp, _dp, M, _ip, C.POINTER(C.c_longlong)]
        L.oracle_bound.argtypes = [P, _dp, _dp, M]
        L.oracle_block_dt.argtypes = [P, _dp, _dp, _dp]
        L.oracle_block_dt.restype = C.c_double
        L.oracle_step.argtypes = [P, _dp, _dp, M, C.c_double, _ip, C.POINTER(C.c_longlong)]
        L.oracle_step.restype = C.c_double
        L.oracle_tetrad.argtypes = [_dp, _dp, C.c_int, _dp, _dp]
        L.oracle_hllc_face.argtypes = [_dp, _dp, _dp, _dp, C.c_double, C.c_double, C.c_int, _dp]
        L.oracle_hllc_face.restype = C.c_double
        L.oracle_rdump_block_write.argtypes = [_ip, _dp, _dp, _dp, _dp]
        L.oracle_dump_block.argtypes = [_ip, _dp, _dp, _dp, C.POINTER(C.c_float)]
        L.oracle_rdump_block_read.argtypes = [_ip, _ip, C.c_int, C.c_int, _dp, _dp, _dp, _dp]

    # ---- pointwise ----
    def get_state(self, pr, gcov, gcon):
        out = [np.zeros(4) for _ in range(4)]
        self.lib.oracle_get_state(_ptr(np.ascontiguousarray(pr, float)), _ptr(np.ascontiguousarray(gcov, float)),
                                  _ptr(np.ascontiguousarray(gcon, float)), *[_ptr(o) for o in out])
        return out

    def primtoflux(self, pr, gcov, gcon, g, gam, d):
        f = np.zeros(NPR)
        vx = np.zeros(1)
        vn = np.zeros(1)
        self.lib.oracle_primtoflux(_ptr(np.ascontiguousarray(pr, float)), _ptr(np.ascontiguousarray(gcov, float)),
                                   _ptr(np.ascontiguousarray(gcon, float)), g, gam, d, _ptr(f), _ptr(vx), _ptr(vn))
        return f, vx[0], vn[0]

    def source(self, pr, gcov, gcon, g, conn40, gam):
        dU = np.zeros(NPR)
        self.lib.oracle_source(_ptr(np.ascontiguousarray(pr, float)), _ptr(np.ascontiguousarray(gcov, float)),
                               _ptr(np.ascontiguousarray(gcon, float)), g,
                               _ptr(np.ascontiguousarray(conn40, float)), gam, _ptr(dU))
        return dU

    def para(self, x5):
        l = np.zeros(1)
        r = np.zeros(1)
        self.lib.oracle_para(_ptr(np.ascontiguousarray(x5, float)), _ptr(l), _ptr(r))
        return l[0], r[0]

    def utoprim(self, which, U, gcov, gcon, g, gam, guess):
        prim = np.array(guess, dtype=float, copy=True)
        ret = self.lib.oracle_utoprim(which, _ptr(np.ascontiguousarray(U, float)),
                                      _ptr(np.ascontiguousarray(gcov, float)),
                                      _ptr(np.ascontiguousarray(gcon, float)), g, gam, _ptr(prim))
        return ret, prim


    # ---- HLLC ----
    def tetrad(self, gcov, gcon, d):
        mud, mudi = np.zeros(16), np.zeros(16)
        self.lib.oracle_tetrad(_ptr(np.ascontiguousarray(gcov, float)), _ptr(np.ascontiguousarray(gcon, float)), d,
                               _ptr(mud), _ptr(mudi))
        return mud.reshape(4, 4), mudi.reshape(4, 4)

    def hllc_face(self, pl, pr, gcov, gcon, g, gam, d):
        """Coordinate HLLC flux (gdet included) through one face and ctop."""
        F = np.zeros(NPR)
        ctop = self.lib.oracle_hllc_face(_ptr(np.ascontiguousarray(pl, float)), _ptr(np.ascontiguousarray(pr, float)),
                                         _ptr(np.ascontiguousarray(gcov, float)),
                                         _ptr(np.ascontiguousarray(gcon, float)), g, gam, d, _ptr(F))
        return F, ctop

    # ---- restart record of one block ----
    def rdump_write(self, bs, p, ps, gdet):
        """Record of one block (p: 9 x ncell, ps: 3 x ncell, gdet: 5 x ncol)."""
        D = [b + 2 * NG for b in bs]
        rec = np.zeros(D[0] * D[1] * D[2] * (NPR + 4))
        b = np.array(bs, dtype=np.int32)
        self.lib.oracle_rdump_block_write(_ptr(b, _ip), _ptr(np.ascontiguousarray(p, float)),
                                          _ptr(np.ascontiguousarray(ps, float)),
                                          _ptr(np.ascontiguousarray(gdet, float)), _ptr(rec))
        return rec

    def dump_block(self, bs, p, gcov_cent, gcon_cent):
        """Dump record (9 floats per interior cell); gcov/gcon_cent: (ncol, 10)."""
        rec = np.zeros(int(np.prod(bs)) * 9, dtype=np.float32)
        b = np.array(bs, dtype=np.int32)
        self.lib.oracle_dump_block(_ptr(b, _ip), _ptr(np.ascontiguousarray(p, float)),
                                   _ptr(np.ascontiguousarray(gcov_cent, float)),
                                   _ptr(np.ascontiguousarray(gcon_cent, float)),
                                   rec.ctypes.data_as(C.POINTER(C.c_float)))
        return rec

    def rdump_read(self, bs, red, spherical, n2, rec, gdet, p, ps):
        """Reads a record (of a block red times finer) into copies of p, ps."""
        p = np.array(p, dtype=float, copy=True)
        ps = np.array(ps, dtype=float, copy=True)
        b = np.array(bs, dtype=np.int32)
        r = np.array(red, dtype=np.int32)
        self.lib.oracle_rdump_block_read(_ptr(b, _ip), _ptr(r, _ip), int(spherical), int(n2),
                                         _ptr(np.ascontiguousarray(rec, float)),
                                         _ptr(np.ascontiguousarray(gdet, float)), _ptr(p), _ptr(ps))
        return p, ps


class Grid:
    """A global ghost-inclusive grid stepped by the oracle."""

    def __init__(self, oracle, n, bs, bc, gam, cour, dx, metric, rho_floor=0.0, u_floor=0.0, riemann=0):
        self.o = oracle
        self.par = Params()
        self.par.riemann = riemann
        self.par.n1, self.par.n2, self.par.n3 = n
        self.par.bs1, self.par.bs2, self.par.bs3 = bs
        for d in range(3):
            self.par.bc[d] = bc[d]
            self.par.dx[d] = dx[d]
        self.par.gam = gam
        self.par.cour = cour
        self.par.rho_floor = rho_floor
        self.par.u_floor = u_floor
        self.D = tuple(x + 2 * NG for x in n)
        self.ncell = self.D[0] * self.D[1] * self.D[2]
        # metric arrays kept alive here
        self._m = {k: np.ascontiguousarray(v, dtype=float) for k, v in metric.items()}
        self.metric = Metric(_ptr(self._m["gcov"]), _ptr(self._m["gcon"]), _ptr(self._m["gdet"]),
                             _ptr(self._m["conn"]), _ptr(self._m["rcen"]))
        self.failcount = (C.c_longlong * 4)()

    def zeros(self, ncomp):
        return np.zeros((ncomp,) + self.D)

    def step(self, p, ps, dt):
        pflag = np.zeros(self.D, dtype=np.int32)
        ndt = self.o.lib.oracle_step(C.byref(self.par), _ptr(p), _ptr(ps), C.byref(self.metric), dt,
                                     _ptr(pflag, _ip), self.failcount)
        return ndt, pflag

    def bound(self, p, ps):
        self.o.lib.oracle_bound(C.byref(self.par), _ptr(p), _ptr(ps), C.byref(self.metric))

    def flux(self, d, p, ps):
        F = self.zeros(NPR)
        dtf = np.zeros(self.D)
        self.o.lib.oracle_flux(C.byref(self.par), d, _ptr(p), _ptr(ps), C.byref(self.metric), _ptr(F), _ptr(dtf))
        return F, dtf

    def emf(self, pb, F1, F2, F3, dt, with_centre=False):
        E = self.zeros(3)
        Ec = self.zeros(3)
        self.o.lib.oracle_emf(C.byref(self.par), _ptr(pb), _ptr(F1), _ptr(F2), _ptr(F3), C.byref(self.metric),
                              dt, _ptr(E), _ptr(Ec))
        return (E, Ec) if with_centre else E

    def ct_update(self, Dt, psi, psf, E):
        self.o.lib.oracle_ct_update(C.byref(self.par), Dt, _ptr(psi), _ptr(psf), _ptr(E), C.byref(self.metric))

    def update(self, full, Dt, pi, pb, pf, U0, psf, F1, F2, F3):
        pflag = np.zeros(self.D, dtype=np.int32)
        self.o.lib.oracle_update(C.byref(self.par), full, Dt, _ptr(pi), _ptr(pb), _ptr(pf), _ptr(U0), _ptr(psf),
                                 _ptr(F1), _ptr(F2), _ptr(F3), C.byref(self.metric), _ptr(pflag, _ip),
                                 self.failcount)
        return pflag

    def invert(self, full, guess, Unew):
        """The inversion cascade, floors and flags over the interior on Unew
        from the guess: (primitives, pflag); failure counts added to
        self.failcount."""
        pf = np.array(guess, dtype=float, copy=True)
        pflag = np.zeros(self.D, dtype=np.int32)
        g = np.ascontiguousarray(guess, dtype=float)
        self.o.lib.oracle_invert(C.byref(self.par), full, _ptr(g), _ptr(g), _ptr(pf),
                                 _ptr(np.ascontiguousarray(Unew, dtype=float)), C.byref(self.metric),
                                 _ptr(pflag, _ip), self.failcount)
        return pf, pflag

    def cons(self, Dt, pi, psf, F1, F2, F3):
        U0 = self.zeros(NPR)
        Un = self.zeros(NPR)
        self.o.lib.oracle_cons(C.byref(self.par), Dt, _ptr(pi), _ptr(U0), _ptr(Un), _ptr(psf), _ptr(F1), _ptr(F2),
                               _ptr(F3), C.byref(self.metric))
        return Un

    def stage(self, Dt, P, PS):
        """One forward stage (the predictor with step Dt) from a state with
        filled ghosts: U' over the interior and the updated face fields."""
        F = [self.flux(d, P, PS)[0] for d in (1, 2, 3)]
        E = self.emf(P, F[0], F[1], F[2], Dt)
        PSF = PS.copy()
        self.ct_update(Dt, PS, PSF, E)
        return self.cons(Dt, P, PSF, F[0], F[1], F[2]), PSF

    def block_dt(self, d1, d2, d3):
        return self.o.lib.oracle_block_dt(C.byref(self.par), _ptr(d1), _ptr(d2), _ptr(d3))


def load():
    if not os.path.exists(LIB):
        raise RuntimeError("oracle/liboracle.so is not built (run __graft_entry__.build())")
    return Oracle(C.CDLL(LIB))


def load_cpu_arith():
    """The oracle built with the inversions in the arithmetic of the
    reference's CPU file (ORACLE_CPU_ARITH): for pinning against the golden
    vectors of that file to the last bit."""
    path = os.path.join(os.path.dirname(LIB), "liboracle_cpuarith.so")
    if not os.path.exists(path):
        raise RuntimeError("oracle/liboracle_cpuarith.so is not built (run __graft_entry__.build())")
    return Oracle(C.CDLL(path))


# ---- multi-level block sets ----
class AmrBlock(C.Structure):
    _fields_ = [("level", C.c_int), ("o", C.c_int * 3), ("bs", C.c_int * 3)]


class AmrGridC(C.Structure):
    _fields_ = [("nblk", C.c_int), ("nlev", C.c_int), ("N", (C.c_int * 3) * 3), ("blk", C.POINTER(AmrBlock)),
                ("m", C.POINTER(Metric)), ("gam", C.c_double), ("cour", C.c_double), ("dx0", C.c_double * 3),
                ("rho_floor", C.c_double), ("u_floor", C.c_double), ("reflux", C.c_int), ("riemann", C.c_int)]


def _setup_amr(lib):
    if getattr(lib, "_amr_ready", False):
        return
    G = C.POINTER(AmrGridC)
    pp = C.POINTER(_dp)
    ipp = C.POINTER(_ip)
    lib.oracle_amr_bound.argtypes = [G, pp, pp, ipp]
    lib.oracle_amr_exchange.argtypes = [G, C.c_int, pp, pp, ipp]
    lib.oracle_amr_step.argtypes = [G, pp, pp, C.c_double, ipp, C.POINTER(C.c_longlong)]
    lib.oracle_amr_step.restype = C.c_double
    lib.oracle_amr_reflux.argtypes = [G, pp, pp, pp, C.c_double]
    lib.oracle_amr_fluxes.argtypes = [G, pp, pp, C.c_double, pp, pp, pp]
    lib.oracle_amr_ct.argtypes = [G, pp, pp, C.c_double, pp]
    lib.oracle_amr_las_cycle.argtypes = [G, _ip, C.c_int, pp, pp, C.c_double, ipp, C.POINTER(C.c_longlong), _dp,
                                         pp, pp]
    lib.oracle_amr_las_cycle.restype = C.c_double
    lib._amr_ready = True


class AmrGrid:
    """A multi-level block set stepped by the oracle.  blocks: rows (level, o1,
    o2, o3, bs1, bs2, bs3); N[l] cells of level l; metrics: per block, the
    oracle layout (as Grid's metric)."""

    def __init__(self, oracle, blocks, N, metrics, gam, cour, dx0, rho_floor=0.0, u_floor=0.0, reflux=True,
                 riemann=0):
        self.o = oracle
        _setup_amr(oracle.lib)
        nb = len(blocks)
        self._blk = (AmrBlock * nb)()
        for b, r in enumerate(blocks):
            self._blk[b].level = int(r[0])
            for k in range(3):
                self._blk[b].o[k] = int(r[1 + k])
                self._blk[b].bs[k] = int(r[4 + k])
        self.D = [tuple(int(r[4 + k]) + 2 * NG for k in range(3)) for r in blocks]
        self._m = [{k: np.ascontiguousarray(v, dtype=float) for k, v in m.items()} for m in metrics]
        self._mc = (Metric * nb)()
        for b, m in enumerate(self._m):
            self._mc[b] = Metric(_ptr(m["gcov"]), _ptr(m["gcon"]), _ptr(m["gdet"]), _ptr(m["conn"]), _ptr(m["rcen"]))
        self.g = AmrGridC()
        self.g.nblk = nb
        self.g.nlev = len(N)
        for l, n in enumerate(N):
            for k in range(3):
                self.g.N[l][k] = int(n[k])
        self.g.blk = self._blk
        self.g.m = self._mc
        self.g.gam, self.g.cour = gam, cour
        for k in range(3):
            self.g.dx0[k] = dx0[k]
        self.g.rho_floor, self.g.u_floor = rho_floor, u_floor
        self.g.reflux = 1 if reflux else 0
        self.g.riemann = riemann
        self.failcount = (C.c_longlong * 4)()

    @staticmethod
    def _pp(arrs, t=_dp):
        for a in arrs:
            assert a.flags["C_CONTIGUOUS"]
        return (C.POINTER(t._type_) * len(arrs))(*[a.ctypes.data_as(t) for a in arrs])

    def bound(self, p, ps, pflag):
        self.o.lib.oracle_amr_bound(C.byref(self.g), self._pp(p), self._pp(ps), self._pp(pflag, _ip))

    def exchange(self, stage, p, ps, pflag):
        self.o.lib.oracle_amr_exchange(C.byref(self.g), stage, self._pp(p), self._pp(ps), self._pp(pflag, _ip))

    def fluxes(self, p, ps, dt):
        """The corrector's fluxes of every block from (p, ps) (ghosts filled),
        corrected at level jumps when the grid refluxes: [F1, F2, F3] lists."""
        F = [[np.zeros_like(a) for a in p] for _ in range(3)]
        self.o.lib.oracle_amr_fluxes(C.byref(self.g), self._pp(p), self._pp(ps), dt, self._pp(F[0]), self._pp(F[1]),
                                     self._pp(F[2]))
        return F

    def ct(self, p, ps, dt):
        """The corrector's face-field update (fluxes, EMFs, polar average and
        the corrections at level jumps when the grid refluxes) from (p, ps):
        the updated face fields, ghosts not filled."""
        psf = [np.zeros_like(a) for a in ps]
        self.o.lib.oracle_amr_ct(C.byref(self.g), self._pp(p), self._pp(ps), dt, self._pp(psf))
        return psf

    def step(self, p, ps, pflag, dt):
        return self.o.lib.oracle_amr_step(C.byref(self.g), self._pp(p), self._pp(ps), dt, self._pp(pflag, _ip),
                                          self.failcount)

    def las_cycle(self, tl, maxtl, p, ps, pflag, dt, fint=False):
        """One local-adaptive-time-stepping cycle (oracle_amr_las_cycle);
        returns (next dt, per-block Courant dt) and with fint also the
        per-block [F1, F2, F3] and edge-EMF sums of Dt F and Dt E over each
        block's correctors."""
        tl = np.ascontiguousarray(tl, dtype=np.int32)
        bdt = np.zeros(len(p))
        F = [[np.zeros_like(a) for _ in range(3)] for a in p] if fint else None
        E = [np.zeros((3,) + a.shape[1:]) for a in p] if fint else None
        Fp = self._pp([x for blk in F for x in blk]) if fint else None
        Ep = self._pp(E) if fint else None
        ndt = self.o.lib.oracle_amr_las_cycle(C.byref(self.g), _ptr(tl, _ip), int(maxtl), self._pp(p), self._pp(ps), dt,
                                              self._pp(pflag, _ip), self.failcount, _ptr(bdt), Fp, Ep)
        if ndt < 0:
            raise RuntimeError("oracle_amr_las_cycle: bad time levels or allocation failure")
        return (ndt, bdt, F, E) if fint else (ndt, bdt)
