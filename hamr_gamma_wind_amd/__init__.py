"""MI355X-native GRMHD fluid step of H-AMR (tp2872/HAMR_gamma_wind).

Python host side of libhamr_hip.so (include/hamr_hip.h): a ctypes binding of
the C ABI plus `advance_step`, the per-block call sequence of the reference's
advance_GPU / GPU_step_ch (step_ch.c:916-1336), so a driver written against the
reference's launch layer reads the same here.  There is no CPU fallback: every
device call goes through the HIP library and raises if it is missing.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhamr_hip.so")
HOST_LIB_PATH = os.path.join(HERE, "libhamr_host.so")

NPR = 9
NG = 3
RHO, UU, U1, U2, U3, B1, B2, B3, KTOT = range(9)
FACE1, FACE2, CORN, CENT, FACE3 = range(5)
(HAMR_P, HAMR_PS, HAMR_PH, HAMR_PSH, HAMR_F1, HAMR_F2, HAMR_F3, HAMR_EMF, HAMR_ECENT, HAMR_U0,
 HAMR_UNEW) = range(11)
NCOMP = {HAMR_P: 9, HAMR_PS: 3, HAMR_PH: 9, HAMR_PSH: 3, HAMR_F1: 9, HAMR_F2: 9, HAMR_F3: 9, HAMR_EMF: 3,
         HAMR_ECENT: 3, HAMR_U0: 9, HAMR_UNEW: 9}
HAMR_PREC_FP64, HAMR_PREC_MIXED = 0, 1
HAMR_RIEMANN_HLL, HAMR_RIEMANN_HLLC = 0, 1


class HamrError(RuntimeError):
    pass


class Desc(C.Structure):
    _fields_ = [("bs", C.c_int * 3), ("nb", C.c_int * 3), ("gam", C.c_double), ("spin", C.c_double),
                ("cour", C.c_double), ("rin", C.c_double), ("rout", C.c_double), ("rho_floor", C.c_double),
                ("u_floor", C.c_double), ("minkowski", C.c_int), ("device", C.c_int)]


class TorusDesc(C.Structure):
    _fields_ = [("rin", C.c_double), ("rmax", C.c_double), ("kappa", C.c_double), ("beta_min", C.c_double),
                ("aphi_cut", C.c_double), ("perturb", C.c_double), ("seed", C.c_ulonglong)]


class AmrDescC(C.Structure):
    _fields_ = [("nlev", C.c_int), ("box", ((C.c_int * 2) * 3) * 3), ("seed", C.c_ulonglong), ("p_big", C.c_double),
                ("p_mid", C.c_double)]


def make_amr_desc(boxes, seed=1, p_big=0.35, p_mid=0.35):
    """boxes[l-1] = ((lo1, hi1), (lo2, hi2), (lo3, hi3)) in cells of level l, for l = 1 .. nlev-1."""
    a = AmrDescC()
    a.nlev = len(boxes) + 1
    for l, b in enumerate(boxes, start=1):
        for k in range(3):
            a.box[l][k][0], a.box[l][k][1] = int(b[k][0]), int(b[k][1])
    a.seed, a.p_big, a.p_mid = seed, p_big, p_mid
    return a


HAMR_TUNE_NAMES = ("flux1", "flux2", "flux3", "emf_centre", "emf_corner", "ct_update", "cons", "invert",
                   "fallback", "bound1", "bound2", "bound3", "exchange")


def make_desc(bs, nb, gam=13.0 / 9.0, spin=0.9375, cour=0.8, rin=0.0, rout=1000.0, rho_floor=1e-5,
              u_floor=1e-7, minkowski=0, device=0):
    d = Desc()
    for k in range(3):
        d.bs[k] = int(bs[k])
        d.nb[k] = int(nb[k])
    d.gam, d.spin, d.cour, d.rin, d.rout = gam, spin, cour, rin, rout
    d.rho_floor, d.u_floor, d.minkowski, d.device = rho_floor, u_floor, minkowski, device
    return d


def torus_desc(rin=6.0, rmax=12.0, kappa=1e-3, beta_min=100.0, aphi_cut=0.2, perturb=0.04, seed=1):
    t = TorusDesc()
    t.rin, t.rmax, t.kappa, t.beta_min, t.aphi_cut, t.perturb, t.seed = rin, rmax, kappa, beta_min, aphi_cut, perturb, seed
    return t


_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_llp = C.POINTER(C.c_longlong)
_lib = None


def _ptr(a, t=_dp):
    if a is None:
        return None
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("array must be C-contiguous")
    return a.ctypes.data_as(t)


_host = None


def load_host_library():
    """libhamr_host.so: the reference's per-block host step in C over the C ABI."""
    global _host
    if _host is not None:
        return _host
    load_library()
    if not os.path.exists(HOST_LIB_PATH):
        raise HamrError("libhamr_host.so is not built: run __graft_entry__.build()")
    L = C.CDLL(HOST_LIB_PATH)
    L.hamr_host_advance.argtypes = [C.c_void_p, C.c_int, C.c_double, _dp]
    L.hamr_host_advance.restype = C.c_int
    _host = L
    return L


def host_advance(ctx, dt):
    """One step through hamr_host_advance (the C host's per-block sequence); returns the next dt."""
    v = C.c_double()
    r = load_host_library().hamr_host_advance(ctx.h, ctx.nblk, dt, C.byref(v))
    if r:
        raise HamrError("hamr_host_advance: %s" % ctx.L.hamr_last_error(ctx.h).decode())
    return v.value


def load_library(path=None):
    """Load libhamr_hip.so (built in-tree by build.py); raises if it is missing."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise HamrError("libhamr_hip.so is not built: run __graft_entry__.build() "
                        "(hamr_gamma_wind_amd/build.py)")
    # torch ships its own HIP runtime.  Loading it before the library makes the
    # library bind to that same runtime, so that torch device memory can be
    # handed to the library (two runtimes in one process do not share a device).
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(p)
    D = C.POINTER(Desc)
    ctxp = C.c_void_p
    sig = {
        "hamr_desc_layout": [D, _ip, _llp, _llp, _llp, _llp, _ip, _ip],
        "hamr_desc_block_table": [D, _ip, _ip],
        "hamr_desc_coords": [D, _dp, _dp],
        "hamr_desc_metric_at": [D, _dp, _dp, _dp, _dp, _dp, _dp],
        "hamr_desc_geometry": [D, C.c_int, _dp, _dp, _dp, _dp, _dp],
        "hamr_desc_torus": [D, C.POINTER(TorusDesc), _dp, _dp],
        "hamr_create": [D, C.POINTER(ctxp)],
        "hamr_destroy": [ctxp],
        "hamr_layout": [ctxp, _ip, _llp, _llp, _ip, _ip],
        "hamr_block_row": [ctxp, C.c_int, _ip],
        "hamr_block_order": [ctxp, _ip],
        "hamr_set_grid": [ctxp],
        "hamr_metric_sizes": [ctxp, _llp, _llp],
        "hamr_upload_metric": [ctxp, C.c_int, _dp, _dp, _dp, _dp, _dp],
        "hamr_download_metric": [ctxp, C.c_int, _dp, _dp, _dp, _dp, _dp],
        "hamr_init_torus": [ctxp, C.POINTER(TorusDesc)],
        "hamr_upload": [ctxp, C.c_int, _dp, _dp],
        "hamr_download": [ctxp, C.c_int, _dp, _dp],
        "hamr_set_array": [ctxp, C.c_int, _dp],
        "hamr_get_array": [ctxp, C.c_int, _dp],
        "hamr_fluxcalc": [ctxp, C.c_int, C.c_int, C.c_int],
        "hamr_block_dt": [ctxp, C.c_int, C.c_int, _dp],
        "hamr_flux_correct": [ctxp, C.c_int, C.c_double],
        "hamr_emf_correct": [ctxp, C.c_int, C.c_double],
        "hamr_ct_emf": [ctxp, C.c_int, C.c_int],
        "hamr_ct_bound": [ctxp, C.c_double],
        "hamr_ct_update": [ctxp, C.c_int, C.c_double, C.c_int],
        "hamr_update": [ctxp, C.c_int, C.c_int, C.c_double],
        "hamr_boundprim": [ctxp, C.c_int],
        "hamr_invert": [ctxp, C.c_int],
        "hamr_flags": [ctxp, C.c_int, _ip, _ip],
        "hamr_clear_flags": [ctxp],
        "hamr_estimate_dt": [ctxp, _dp],
        "hamr_step": [ctxp, C.c_int, _dp],
        "hamr_step_graph": [ctxp, C.c_int, _dp],
        "hamr_sync": [ctxp],
        "hamr_time_stage": [ctxp, C.c_int, C.c_int, _dp],
        "hamr_copy_probe": [ctxp, C.c_longlong, C.c_int, _dp],
        "hamr_set_precision": [ctxp, C.c_int],
        "hamr_get_precision": [ctxp, _ip],
        "hamr_set_array_dev": [ctxp, C.c_int, C.c_void_p],
        "hamr_get_array_dev": [ctxp, C.c_int, C.c_void_p],
        "hamr_stage_forward": [ctxp, C.c_double],
        "hamr_stage_adjoint": [ctxp, C.c_double, _dp, _dp, _dp],
        "hamr_stage_adjoint_dev": [ctxp, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p],
        "hamr_stage_adjoint_dev_noalloc": [ctxp, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p],
        "hamr_create_amr": [D, C.POINTER(AmrDescC), C.POINTER(ctxp)],
        "hamr_desc_amr": [D, C.POINTER(AmrDescC), _llp, _ip, _ip, _ip],
        "hamr_desc_amr_geometry": [D, C.POINTER(AmrDescC), C.c_int, _dp, _dp, _dp, _dp, _dp],
        "hamr_desc_amr_torus": [D, C.POINTER(AmrDescC), C.POINTER(TorusDesc), _dp, _dp],
        "hamr_amr_info": [ctxp, _llp],
        "hamr_block_layout": [ctxp, C.c_int, _ip],
        "hamr_geometry_info": [ctxp, C.c_int, _llp],
        "hamr_autotune": [ctxp, C.c_int],
        "hamr_tune_get": [ctxp, C.c_int, _ip, _dp],
        "hamr_tune_set": [ctxp, C.c_int, _ip],
        "hamr_time_kernel": [ctxp, C.c_int, C.c_int, _dp],
        "hamr_set_riemann": [ctxp, C.c_int],
        "hamr_get_riemann": [ctxp, _ip],
        "hamr_face_tetrad": [_dp, _dp, C.c_int, _dp, _dp],
        "hamr_download_tetrads": [ctxp, C.c_int, _dp],
        "hamr_rdump_block_size": [ctxp, C.c_int, C.c_int, C.c_int, C.c_int, _llp],
        "hamr_rdump_block_write": [ctxp, C.c_int, _dp],
        "hamr_rdump_block_read": [ctxp, C.c_int, _dp, C.c_int, C.c_int, C.c_int],
        "hamr_rdump_offsets": [ctxp, _llp],
        "hamr_rdump_write_dev": [ctxp, C.c_void_p],
        "hamr_rdump_read_dev": [ctxp, C.c_void_p],
        "hamr_dump_block_size": [ctxp, C.c_int, _llp],
        "hamr_dump_block_write": [ctxp, C.c_int, C.POINTER(C.c_float)],
        "hamr_dump_offsets": [ctxp, _llp],
        "hamr_dump_write_dev": [ctxp, C.c_void_p],
        "hamr_set_timelevels": [ctxp, _ip],
        "hamr_get_timelevels": [ctxp, _ip],
        "hamr_block_dts": [ctxp, _dp],
        "hamr_timelevels_from_dt": [ctxp, C.c_int, _ip],
        "hamr_las_rule": [C.c_int, _dp, C.c_double, C.c_int, C.c_int, C.c_int, _ip, C.c_int, _ip, _ip],
        "hamr_step_las": [ctxp, C.c_int, C.c_int, _dp],
        "hamr_step_las_graph": [ctxp, C.c_int, C.c_int, _dp],
        "hamr_las_info": [ctxp, _llp],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = C.c_int
    L.hamr_last_error.argtypes = [ctxp]
    L.hamr_last_error.restype = C.c_char_p
    if path is None:
        _lib = L
    return L


# ---------------------------------------------------------------------------
# host-only set-up
def face_tetrad(gcov10, gcon10, d):
    """set_Mud (hllc.c:20-178) at one face of direction d from the packed metric
    there: (mud, mudi) as 4x4 arrays, mud[mu, a] = e_(a)^mu, mudi[a, mu] = e^(a)_mu."""
    L = load_library()
    mud, mudi = np.zeros(16), np.zeros(16)
    r = L.hamr_face_tetrad(_ptr(np.ascontiguousarray(gcov10, float)), _ptr(np.ascontiguousarray(gcon10, float)),
                           int(d), _ptr(mud), _ptr(mudi))
    if r:
        raise HamrError("hamr_face_tetrad: bad arguments")
    return mud.reshape(4, 4), mudi.reshape(4, 4)


def desc_layout(desc):
    L = load_library()
    D = (C.c_int * 3)()
    ncell, ksize, ncol, kcol = C.c_longlong(), C.c_longlong(), C.c_longlong(), C.c_longlong()
    nblk, ngeo = C.c_int(), C.c_int()
    if L.hamr_desc_layout(C.byref(desc), D, C.byref(ncell), C.byref(ksize), C.byref(ncol), C.byref(kcol),
                          C.byref(nblk), C.byref(ngeo)):
        raise HamrError("hamr_desc_layout: invalid descriptor")
    return dict(D=tuple(D), ncell=ncell.value, ksize=ksize.value, ncol=ncol.value, kcol=kcol.value,
                nblk=nblk.value, ngeo=ngeo.value)


def desc_block_table(desc):
    L = load_library()
    lay = desc_layout(desc)
    rows = np.zeros((lay["nblk"], 17), dtype=np.int32)
    order = np.zeros(lay["nblk"], dtype=np.int32)
    if L.hamr_desc_block_table(C.byref(desc), _ptr(rows, _ip), _ptr(order, _ip)):
        raise HamrError("hamr_desc_block_table failed")
    return rows, order


def desc_coords(desc):
    L = load_library()
    sx = np.zeros(3)
    dx = np.zeros(3)
    if L.hamr_desc_coords(C.byref(desc), _ptr(sx), _ptr(dx)):
        raise HamrError("hamr_desc_coords failed")
    return sx, dx


def desc_metric_at(desc, X):
    L = load_library()
    gcov, gcon, gdet, conn, rth = np.zeros(10), np.zeros(10), np.zeros(1), np.zeros(40), np.zeros(2)
    X = np.ascontiguousarray(X, dtype=float)
    if L.hamr_desc_metric_at(C.byref(desc), _ptr(X), _ptr(gcov), _ptr(gcon), _ptr(gdet), _ptr(conn), _ptr(rth)):
        raise HamrError("hamr_desc_metric_at failed")
    return gcov, gcon, gdet[0], conn, rth


def desc_geometry(desc, geo):
    L = load_library()
    lay = desc_layout(desc)
    kc = lay["kcol"]
    gcov, gcon = np.zeros(5 * 10 * kc), np.zeros(5 * 10 * kc)
    gdet, conn, rcen = np.zeros(5 * kc), np.zeros(40 * kc), np.zeros(kc)
    if L.hamr_desc_geometry(C.byref(desc), geo, _ptr(gcov), _ptr(gcon), _ptr(gdet), _ptr(conn), _ptr(rcen)):
        raise HamrError("hamr_desc_geometry failed")
    return dict(gcov=gcov.reshape(5, 10, kc), gcon=gcon.reshape(5, 10, kc), gdet=gdet.reshape(5, kc),
                conn=conn.reshape(40, kc), rcen=rcen)


def desc_torus(desc, tdesc):
    L = load_library()
    lay = desc_layout(desc)
    p = np.zeros((lay["nblk"], NPR) + lay["D"])
    ps = np.zeros((lay["nblk"], 3) + lay["D"])
    r = L.hamr_desc_torus(C.byref(desc), C.byref(tdesc), _ptr(p), _ptr(ps))
    if r:
        raise HamrError("hamr_desc_torus failed (%d)" % r)
    return p, ps


AMR_INFO_KEYS = ("blocks", "geometries", "cells", "zones", "ex1", "ex2", "ex3", "pro1", "pro2", "pro3", "rings", "amr",
                 "reflux1", "reflux2", "reflux3", "emf", "push3")


def desc_amr(desc, amr, tables=False):
    """The multi-level block set (host only): info, block rows (level, o, bs,
    geo, pole, face kinds) and, with tables=True, the exchange entries per stage."""
    L = load_library()
    info = np.zeros(len(AMR_INFO_KEYS), dtype=np.int64)
    r = L.hamr_desc_amr(C.byref(desc), C.byref(amr), _ptr(info, _llp), None, None, None)
    if r:
        raise HamrError("hamr_desc_amr: description rejected (%d)" % r)
    nb = int(info[0])
    rows = np.zeros((nb, 15), dtype=np.int32)
    nex = [int(info[4 + s]) for s in range(3)]
    npro = [int(info[7 + s]) for s in range(3)]
    ex = np.zeros((sum(nex), 4), dtype=np.int32) if tables else None
    pro = np.zeros((sum(npro), 16), dtype=np.int32) if tables else None
    L.hamr_desc_amr(C.byref(desc), C.byref(amr), _ptr(info, _llp), _ptr(rows, _ip),
                    _ptr(ex, _ip) if tables else None, _ptr(pro, _ip) if tables else None)
    out = dict(info={k: int(v) for k, v in zip(AMR_INFO_KEYS, info)}, rows=rows)
    if tables:
        out["ex"] = np.split(ex, np.cumsum(nex)[:-1])
        out["pro"] = np.split(pro, np.cumsum(npro)[:-1])
    return out


def amr_block_dims(rows):
    return [tuple(int(b) + 2 * NG for b in r[4:7]) for r in rows]


def desc_amr_geometry(desc, amr, geo, D1, D2):
    """Geometry slab geo of the block set (device layout; D1 x D2 columns)."""
    L = load_library()
    kc = (D1 * D2 + 63) // 64 * 64
    gcov, gcon = np.zeros(5 * 10 * kc), np.zeros(5 * 10 * kc)
    gdet, conn, rcen = np.zeros(5 * kc), np.zeros(40 * kc), np.zeros(kc)
    if L.hamr_desc_amr_geometry(C.byref(desc), C.byref(amr), geo, _ptr(gcov), _ptr(gcon), _ptr(gdet), _ptr(conn),
                                _ptr(rcen)):
        raise HamrError("hamr_desc_amr_geometry failed")
    return dict(gcov=gcov.reshape(5, 10, kc), gcon=gcon.reshape(5, 10, kc), gdet=gdet.reshape(5, kc),
                conn=conn.reshape(40, kc), rcen=rcen)


def desc_amr_torus(desc, amr, tdesc):
    """Torus initial state of the block set: lists of per-block p (9,)+D and ps (3,)+D."""
    L = load_library()
    dims = amr_block_dims(desc_amr(desc, amr)["rows"])
    n = sum(int(np.prod(D)) for D in dims)
    p, ps = np.zeros(NPR * n), np.zeros(3 * n)
    r = L.hamr_desc_amr_torus(C.byref(desc), C.byref(amr), C.byref(tdesc), _ptr(p), _ptr(ps))
    if r:
        raise HamrError("hamr_desc_amr_torus failed (%d)" % r)
    P, S, a, b = [], [], 0, 0
    for D in dims:
        m = int(np.prod(D))
        P.append(p[a:a + NPR * m].reshape((NPR,) + D))
        S.append(ps[b:b + 3 * m].reshape((3,) + D))
        a += NPR * m
        b += 3 * m
    return P, S


# ---------------------------------------------------------------------------
class Context:
    """A resident block set on one MI355X (hamr_ctx)."""

    def __init__(self, desc):
        self.L = load_library()
        self.desc = desc
        self.h = C.c_void_p()
        r = self.L.hamr_create(C.byref(desc), C.byref(self.h))
        if r:
            msg = self.L.hamr_last_error(self.h).decode() if self.h.value else "hamr_create failed"
            if self.h.value:
                self.L.hamr_destroy(self.h)
            self.h = C.c_void_p()
            raise HamrError(msg)
        lay = desc_layout(desc)
        self.D = lay["D"]
        self.ncell = lay["ncell"]
        self.nblk = lay["nblk"]
        self.ngeo = lay["ngeo"]
        self.kcol = lay["kcol"]

    def close(self):
        if self.h.value:
            self.L.hamr_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, r, what):
        if r:
            raise HamrError("%s: %s" % (what, self.L.hamr_last_error(self.h).decode()))

    # set-up
    def set_grid(self):
        self._chk(self.L.hamr_set_grid(self.h), "hamr_set_grid")

    def init_torus(self, tdesc):
        self._chk(self.L.hamr_init_torus(self.h, C.byref(tdesc)), "hamr_init_torus")

    def upload_metric(self, geo, m):
        arrs = [np.ascontiguousarray(m[k], dtype=float).ravel() for k in ("gcov", "gcon", "gdet", "conn", "rcen")]
        self._chk(self.L.hamr_upload_metric(self.h, geo, *[_ptr(a) for a in arrs]), "hamr_upload_metric")

    def download_metric(self, geo):
        kc = self.kcol
        gcov, gcon = np.zeros(5 * 10 * kc), np.zeros(5 * 10 * kc)
        gdet, conn, rcen = np.zeros(5 * kc), np.zeros(40 * kc), np.zeros(kc)
        self._chk(self.L.hamr_download_metric(self.h, geo, _ptr(gcov), _ptr(gcon), _ptr(gdet), _ptr(conn),
                                              _ptr(rcen)), "hamr_download_metric")
        return dict(gcov=gcov.reshape(5, 10, kc), gcon=gcon.reshape(5, 10, kc), gdet=gdet.reshape(5, kc),
                    conn=conn.reshape(40, kc), rcen=rcen)

    def block_row(self, n):
        row = np.zeros(17, dtype=np.int32)
        self._chk(self.L.hamr_block_row(self.h, n, _ptr(row, _ip)), "hamr_block_row")
        return row

    # arrays
    def get(self, which):
        a = np.zeros((self.nblk, NCOMP[which]) + self.D)
        self._chk(self.L.hamr_get_array(self.h, which, _ptr(a)), "hamr_get_array")
        return a

    def set(self, which, a):
        a = np.ascontiguousarray(a, dtype=float)
        assert a.shape == (self.nblk, NCOMP[which]) + self.D
        self._chk(self.L.hamr_set_array(self.h, which, _ptr(a)), "hamr_set_array")

    def upload(self, n, p_aos, ps_aos):
        self._chk(self.L.hamr_upload(self.h, n, _ptr(np.ascontiguousarray(p_aos, float)),
                                     _ptr(np.ascontiguousarray(ps_aos, float))), "hamr_upload")

    def download(self, n):
        p = np.zeros(self.D + (NPR,))
        ps = np.zeros(self.D + (4,))
        self._chk(self.L.hamr_download(self.h, n, _ptr(p), _ptr(ps)), "hamr_download")
        return p, ps

    def flags(self, n):
        pf = np.zeros(self.D, dtype=np.int32)
        fi = np.zeros(4, dtype=np.int32)
        self._chk(self.L.hamr_flags(self.h, n, _ptr(pf, _ip), _ptr(fi, _ip)), "hamr_flags")
        return pf, fi

    def all_flags(self):
        pf = np.zeros((self.nblk,) + self.D, dtype=np.int32)
        fi = np.zeros((self.nblk, 4), dtype=np.int32)
        for n in range(self.nblk):
            pf[n], fi[n] = self.flags(n)
        return pf, fi

    def clear_flags(self):
        self._chk(self.L.hamr_clear_flags(self.h), "hamr_clear_flags")

    # stages (reference names in the docstrings of include/hamr_hip.h)
    def fluxcalc(self, d, flag, n):
        self._chk(self.L.hamr_fluxcalc(self.h, d, flag, n), "hamr_fluxcalc")

    def block_dt(self, n, d):
        v = C.c_double()
        self._chk(self.L.hamr_block_dt(self.h, n, d, C.byref(v)), "hamr_block_dt")
        return v.value

    def flux_correct(self, flag, dt):
        self._chk(self.L.hamr_flux_correct(self.h, flag, dt), "hamr_flux_correct")

    def emf_correct(self, flag, dt):
        self._chk(self.L.hamr_emf_correct(self.h, flag, dt), "hamr_emf_correct")

    def ct_emf(self, flag, n):
        self._chk(self.L.hamr_ct_emf(self.h, flag, n), "hamr_ct_emf")

    def ct_bound(self, dt):
        self._chk(self.L.hamr_ct_bound(self.h, dt), "hamr_ct_bound")

    def ct_update(self, flag, Dt, n):
        self._chk(self.L.hamr_ct_update(self.h, flag, Dt, n), "hamr_ct_update")

    def update(self, flag, n, Dt):
        self._chk(self.L.hamr_update(self.h, flag, n, Dt), "hamr_update")

    def invert(self, flag):
        """hamr_invert: the inversion cascade on HAMR_UNEW (flag 0: guess p ->
        ph; flag 1: guess ph -> p)."""
        self._chk(self.L.hamr_invert(self.h, flag), "hamr_invert")

    def boundprim(self, flag):
        self._chk(self.L.hamr_boundprim(self.h, flag), "hamr_boundprim")

    def sync(self):
        self._chk(self.L.hamr_sync(self.h), "hamr_sync")

    # device-pointer access (torch tensors or any device buffer of this layout)
    def set_dev(self, which, ptr):
        self._chk(self.L.hamr_set_array_dev(self.h, which, C.c_void_p(ptr)), "hamr_set_array_dev")

    def get_dev(self, which, ptr):
        self._chk(self.L.hamr_get_array_dev(self.h, which, C.c_void_p(ptr)), "hamr_get_array_dev")

    # restart records: the rdump files of restart.c (rdump_block_write :52,
    # rdump_block_read :178); NPR + 4 doubles per ghost-inclusive cell
    def rdump_size(self, n, red=(1, 1, 1)):
        v = C.c_longlong()
        self._chk(self.L.hamr_rdump_block_size(self.h, n, *[int(r) for r in red], C.byref(v)),
                  "hamr_rdump_block_size")
        return v.value

    def rdump_block_write(self, n):
        rec = np.zeros(self.rdump_size(n))
        self._chk(self.L.hamr_rdump_block_write(self.h, n, _ptr(rec)), "hamr_rdump_block_write")
        return rec

    def rdump_block_read(self, n, rec, red=(1, 1, 1)):
        rec = np.ascontiguousarray(rec, dtype=float).reshape(-1)
        if rec.size != self.rdump_size(n, red):
            raise HamrError("rdump record of block %d: %d doubles, expected %d" % (n, rec.size,
                                                                                    self.rdump_size(n, red)))
        self._chk(self.L.hamr_rdump_block_read(self.h, n, _ptr(rec), *[int(r) for r in red]),
                  "hamr_rdump_block_read")

    def rdump_offsets(self):
        off = np.zeros(self.nblk + 1, dtype=np.int64)
        self._chk(self.L.hamr_rdump_offsets(self.h, _ptr(off, _llp)), "hamr_rdump_offsets")
        return off

    def rdump_write_dev(self, ptr):
        self._chk(self.L.hamr_rdump_write_dev(self.h, C.c_void_p(ptr)), "hamr_rdump_write_dev")

    def rdump_read_dev(self, ptr):
        self._chk(self.L.hamr_rdump_read_dev(self.h, C.c_void_p(ptr)), "hamr_rdump_read_dev")

    # dump records: the new_dump files of dump.c (dump_block :276)
    def dump_block_write(self, n):
        v = C.c_longlong()
        self._chk(self.L.hamr_dump_block_size(self.h, n, C.byref(v)), "hamr_dump_block_size")
        rec = np.zeros(v.value, dtype=np.float32)
        self._chk(self.L.hamr_dump_block_write(self.h, n, rec.ctypes.data_as(C.POINTER(C.c_float))),
                  "hamr_dump_block_write")
        return rec

    def dump_offsets(self):
        off = np.zeros(self.nblk + 1, dtype=np.int64)
        self._chk(self.L.hamr_dump_offsets(self.h, _ptr(off, _llp)), "hamr_dump_offsets")
        return off

    def dump_write_dev(self, ptr):
        self._chk(self.L.hamr_dump_write_dev(self.h, C.c_void_p(ptr)), "hamr_dump_write_dev")

    def dump_write(self, path, order=None):
        """dump_new (dump.c:4-66) for one rank holding every block: the records
        of the blocks in `order` (default: block index order) into one file."""
        with open(path, "wb") as fh:
            for n in (range(self.nblk) if order is None else order):
                self.dump_block_write(int(n)).tofile(fh)

    def block_numbers(self):
        """The reference's number of each block (AMR_set_coord, AMR.c:168-275):
        number[n] of this context's block n (the inverse of hamr_block_order)."""
        order = np.zeros(self.nblk, dtype=np.int32)
        self._chk(self.L.hamr_block_order(self.h, _ptr(order, _ip)), "hamr_block_order")
        number = np.empty_like(order)
        number[order] = np.arange(self.nblk, dtype=np.int32)
        return number

    def rdump_write(self, directory):
        """restart_write's block files (restart.c:40-45): directory/rdump<k>
        holds the record, raw doubles, of the block the reference numbers k."""
        os.makedirs(directory, exist_ok=True)
        number = self.block_numbers()
        for n in range(self.nblk):
            self.rdump_block_write(n).tofile(os.path.join(directory, "rdump%d" % number[n]))

    def rdump_read(self, directory, red=(1, 1, 1)):
        """restart_read (restart.c:94-175): every block from directory/rdump<k>,
        k its reference number (records of blocks red times finer), then the
        ghost fill."""
        number = self.block_numbers()
        for n in range(self.nblk):
            path = os.path.join(directory, "rdump%d" % number[n])
            if not os.path.exists(path):
                raise HamrError("cannot open restart file %s" % path)
            self.rdump_block_read(n, np.fromfile(path, dtype=np.float64), red)
        self.boundprim(1)

    # one forward stage and its adjoint
    def stage_forward(self, Dt):
        self._chk(self.L.hamr_stage_forward(self.h, Dt), "hamr_stage_forward")

    def stage_adjoint(self, Dt, w):
        """dL/dP and dL/d(face field) of L = sum w . U' (host arrays)."""
        w = np.ascontiguousarray(w, dtype=float)
        assert w.shape == (self.nblk, 9) + self.D
        gp = np.zeros((self.nblk, 9) + self.D)
        gps = np.zeros((self.nblk, 3) + self.D)
        self._chk(self.L.hamr_stage_adjoint(self.h, Dt, _ptr(w), _ptr(gp), _ptr(gps)), "hamr_stage_adjoint")
        return gp, gps

    def stage_adjoint_dev(self, Dt, w_ptr, gp_ptr, gps_ptr):
        self._chk(self.L.hamr_stage_adjoint_dev(self.h, Dt, C.c_void_p(w_ptr), C.c_void_p(gp_ptr),
                                                C.c_void_p(gps_ptr)), "hamr_stage_adjoint_dev")

    def stage_adjoint_dev_noalloc(self, Dt, w_ptr, gp_ptr, gps_ptr):
        """stage_adjoint_dev with the work arrays in the library's static device
        scratch (no allocation; boxes up to the scratch's size)."""
        self._chk(self.L.hamr_stage_adjoint_dev_noalloc(self.h, Dt, C.c_void_p(w_ptr), C.c_void_p(gp_ptr),
                                                        C.c_void_p(gps_ptr)), "hamr_stage_adjoint_dev_noalloc")

    def set_precision(self, mode):
        """HAMR_PREC_FP64 or HAMR_PREC_MIXED (FP32 reconstruction, fluxes, EMFs)."""
        self._chk(self.L.hamr_set_precision(self.h, mode), "hamr_set_precision")

    def set_riemann(self, solver):
        self._chk(self.L.hamr_set_riemann(self.h, solver), "hamr_set_riemann")

    def riemann(self):
        v = C.c_int()
        self._chk(self.L.hamr_get_riemann(self.h, C.byref(v)), "hamr_get_riemann")
        return v.value

    def download_tetrads(self, geo, kcol=None):
        kc = kcol if kcol is not None else self.kcol
        t = np.zeros(96 * kc)
        self._chk(self.L.hamr_download_tetrads(self.h, geo, _ptr(t)), "hamr_download_tetrads")
        return t.reshape(3, 32, kc)

    def precision(self):
        v = C.c_int()
        self._chk(self.L.hamr_get_precision(self.h, C.byref(v)), "hamr_get_precision")
        return v.value

    def time_stage(self, stage, nrep=10):
        v = C.c_double()
        self._chk(self.L.hamr_time_stage(self.h, stage, nrep, C.byref(v)), "hamr_time_stage")
        return v.value

    def copy_probe(self, nbytes, nrep=1):
        v = C.c_double()
        self._chk(self.L.hamr_copy_probe(self.h, nbytes, nrep, C.byref(v)), "hamr_copy_probe")
        return v.value

    # batched
    def estimate_dt(self):
        v = C.c_double()
        self._chk(self.L.hamr_estimate_dt(self.h, C.byref(v)), "hamr_estimate_dt")
        return v.value

    def step(self, nsteps, dt=None, graph=False):
        v = C.c_double(dt if dt is not None else -1.0)
        f = self.L.hamr_step_graph if graph else self.L.hamr_step
        self._chk(f(self.h, nsteps, C.byref(v)), "hamr_step")
        return v.value

    # launch autotuner
    def autotune(self, nrep=3):
        self._chk(self.L.hamr_autotune(self.h, nrep), "hamr_autotune")
        return self.tune_table()

    def tune_table(self):
        out = {}
        for k, name in enumerate(HAMR_TUNE_NAMES):
            cfg = np.zeros(3, dtype=np.int32)
            ms = C.c_double()
            self._chk(self.L.hamr_tune_get(self.h, k, _ptr(cfg, _ip), C.byref(ms)), "hamr_tune_get")
            out[name] = dict(threads=int(cfg[0]), grid=int(cfg[1]), xcd_order=int(cfg[2]) & 1,
                             second_form=(int(cfg[2]) >> 1) & 1, ms=ms.value)
        return out

    def time_kernel(self, slot, nrep=5):
        v = C.c_double()
        self._chk(self.L.hamr_time_kernel(self.h, slot, nrep, C.byref(v)), "hamr_time_kernel")
        return v.value

    def tune_set(self, slot, threads, grid, xcd_order, second_form=0):
        """Launch shape of kernel `slot`; second_form = 1 selects the slot's
        other form (flux slots: line-staged tiles of 256 threads instead of one
        thread per face; cons: k_cons + k_invert instead of k_update)."""
        cfg = np.array([threads, grid, (xcd_order & 1) | (second_form << 1)], dtype=np.int32)
        self._chk(self.L.hamr_tune_set(self.h, slot, _ptr(cfg, _ip)), "hamr_tune_set")


class AmrContext(Context):
    """A multi-level block set of uneven block sizes (hamr_create_amr).  Whole
    arrays are lists of per-block arrays, each (ncomp,) + that block's D."""

    def __init__(self, desc, amr):
        self.L = load_library()
        self.desc = desc
        self.amr = amr
        self.h = C.c_void_p()
        r = self.L.hamr_create_amr(C.byref(desc), C.byref(amr), C.byref(self.h))
        if r:
            msg = self.L.hamr_last_error(self.h).decode() if self.h.value else "hamr_create_amr failed"
            if self.h.value:
                self.L.hamr_destroy(self.h)
            self.h = C.c_void_p()
            raise HamrError(msg)
        info = self.info()
        self.nblk = info["blocks"]
        self.ngeo = info["geometries"]
        self.layout = np.zeros((self.nblk, 15), dtype=np.int32)
        for n in range(self.nblk):
            self._chk(self.L.hamr_block_layout(self.h, n, _ptr(self.layout[n], _ip)), "hamr_block_layout")
        self.Ds = [tuple(int(b) + 2 * NG for b in self.layout[n, 4:7]) for n in range(self.nblk)]

    def info(self):
        v = np.zeros(len(AMR_INFO_KEYS), dtype=np.int64)
        self._chk(self.L.hamr_amr_info(self.h, _ptr(v, _llp)), "hamr_amr_info")
        return {k: int(x) for k, x in zip(AMR_INFO_KEYS, v)}

    def geometry_info(self, geo):
        v = np.zeros(6, dtype=np.int64)
        self._chk(self.L.hamr_geometry_info(self.h, geo, _ptr(v, _llp)), "hamr_geometry_info")
        return dict(level=int(v[0]), o1=int(v[1]), o2=int(v[2]), D1=int(v[3]), D2=int(v[4]), kcol=int(v[5]))

    def download_metric(self, geo):
        kc = self.geometry_info(geo)["kcol"]
        gcov, gcon = np.zeros(5 * 10 * kc), np.zeros(5 * 10 * kc)
        gdet, conn, rcen = np.zeros(5 * kc), np.zeros(40 * kc), np.zeros(kc)
        self._chk(self.L.hamr_download_metric(self.h, geo, _ptr(gcov), _ptr(gcon), _ptr(gdet), _ptr(conn),
                                              _ptr(rcen)), "hamr_download_metric")
        return dict(gcov=gcov.reshape(5, 10, kc), gcon=gcon.reshape(5, 10, kc), gdet=gdet.reshape(5, kc),
                    conn=conn.reshape(40, kc), rcen=rcen)

    def get(self, which):
        nc = NCOMP[which]
        flat = np.zeros(sum(nc * int(np.prod(D)) for D in self.Ds))
        self._chk(self.L.hamr_get_array(self.h, which, _ptr(flat)), "hamr_get_array")
        out, pos = [], 0
        for D in self.Ds:
            m = nc * int(np.prod(D))
            out.append(flat[pos:pos + m].reshape((nc,) + D))
            pos += m
        return out

    def set(self, which, arrays):
        nc = NCOMP[which]
        assert len(arrays) == self.nblk
        flat = np.concatenate([np.ascontiguousarray(a, dtype=float).reshape(-1) for a in arrays])
        assert flat.size == sum(nc * int(np.prod(D)) for D in self.Ds)
        self._chk(self.L.hamr_set_array(self.h, which, _ptr(flat)), "hamr_set_array")

    def flags(self, n):
        pf = np.zeros(self.Ds[n], dtype=np.int32)
        fi = np.zeros(4, dtype=np.int32)
        self._chk(self.L.hamr_flags(self.h, n, _ptr(pf, _ip), _ptr(fi, _ip)), "hamr_flags")
        return pf, fi

    def all_flags(self):
        return [self.flags(n) for n in range(self.nblk)]

    # local adaptive time stepping
    def set_timelevels(self, tl):
        v = np.ascontiguousarray(tl, dtype=np.int32)
        assert v.size == self.nblk
        self._chk(self.L.hamr_set_timelevels(self.h, _ptr(v, _ip)), "hamr_set_timelevels")

    def timelevels(self):
        v = np.zeros(self.nblk, dtype=np.int32)
        self._chk(self.L.hamr_get_timelevels(self.h, _ptr(v, _ip)), "hamr_get_timelevels")
        return v

    def block_dts(self):
        v = np.zeros(self.nblk)
        self._chk(self.L.hamr_block_dts(self.h, _ptr(v)), "hamr_block_dts")
        return v

    def timelevels_from_dt(self, maxtl):
        v = np.zeros(self.nblk, dtype=np.int32)
        self._chk(self.L.hamr_timelevels_from_dt(self.h, maxtl, _ptr(v, _ip)), "hamr_timelevels_from_dt")
        return v

    def step_las(self, maxtl, ncycles=1, dt=None, graph=False):
        """ncycles cycles of 2 maxtl sub-iterations (each advances every block
        by maxtl dt), launched one by one or replayed from a captured graph;
        returns the next dt."""
        v = C.c_double(dt if dt is not None else -1.0)
        f = self.L.hamr_step_las_graph if graph else self.L.hamr_step_las
        self._chk(f(self.h, maxtl, ncycles, C.byref(v)), "hamr_step_las")
        return v.value

    def las_info(self):
        v = np.zeros(6, dtype=np.int64)
        self._chk(self.L.hamr_las_info(self.h, _ptr(v, _llp)), "hamr_las_info")
        return dict(flux1=int(v[0]), flux2=int(v[1]), flux3=int(v[2]), emf_same=int(v[3]), emf_jump=int(v[4]),
                    tlmax=int(v[5]))


def las_rule(bdt, dt, maxtl, rings=(), pairs=()):
    """set_timelevel's rule (hamr_las_rule) on plain arrays: rings a list of
    equal-length block-id lists, pairs (coarse, finer) block pairs."""
    L = load_library()
    bdt = np.ascontiguousarray(bdt, dtype=float)
    n = bdt.size
    r = np.ascontiguousarray(np.array(rings, dtype=np.int32).reshape(len(rings), -1) if len(rings) else
                             np.zeros((0, 0), dtype=np.int32))
    pr = np.ascontiguousarray(np.array(pairs, dtype=np.int32).reshape(-1, 2) if len(pairs) else
                              np.zeros((0, 2), dtype=np.int32))
    tl = np.zeros(n, dtype=np.int32)
    rc = L.hamr_las_rule(n, _ptr(bdt), float(dt), int(maxtl), r.shape[0], r.shape[1] if r.size else 0,
                         _ptr(r, _ip), pr.shape[0], _ptr(pr, _ip), _ptr(tl, _ip))
    if rc:
        raise HamrError("hamr_las_rule: bad arguments")
    return tl


def advance_step(ctx, dt):
    """One predictor + corrector step through the per-block stage API, in the
    order of advance_GPU (step_ch.c:997-1336) with every block at timelevel 1:
    fluxes dir 3, 2, 1; flux correction at level jumps (full step); CT EMF;
    polar E1 average; EMF correction at level jumps (full step); face-B update;
    conserved update + inversion; ghost fill.
    Returns the next dt (per-block harmonic of
    the corrector's Courant minima, minimum over blocks, clamped to 1.3 dt)."""
    nb = ctx.nblk
    for flag, Dt in ((0, 0.5 * dt), (1, dt)):
        for d in (3, 2, 1):
            for n in range(nb):
                ctx.fluxcalc(d, flag, n)
        if flag == 1:
            bdt = [[ctx.block_dt(n, d) for d in (1, 2, 3)] for n in range(nb)]
        else:
            for n in range(nb):
                for d in (1, 2, 3):
                    ctx.block_dt(n, d)
        ctx.flux_correct(flag, dt)
        for n in range(nb):
            ctx.ct_emf(flag, n)
        ctx.ct_bound(dt)
        ctx.emf_correct(flag, dt)
        for n in range(nb):
            ctx.ct_update(flag, Dt, n)
        for n in range(nb):
            ctx.update(flag, n, Dt)
        ctx.boundprim(1 if flag else 0)
    ndt = min(1.0 / (1.0 / a + 1.0 / b + 1.0 / c) for a, b, c in bdt)
    return min(ndt, 1.3 * dt)
