"""ctypes binding of the CPU oracle (oracle/_build/libmpm_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, where it is the checker. The product path
(mpmsnow_amd) never imports it.

Layouts follow the reference's host types: Vector3f arrays are (n, 3),
Matrix3f arrays are (n, 3, 3) in natural [row, col] order (the C side stores
them column-major like Eigen; this module converts).
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libmpm_oracle.so")

_lib = None

VEC3 = {"p", "v"}
MAT3 = {"F", "Fp", "FinvTrans", "R", "Ginv"}
SCAL = {"m", "volume", "J", "mu", "lambda", "mu0", "lambda0", "xi"}

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_lp = C.POINTER(C.c_long)
_bp = C.POINTER(C.c_byte)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            from oracle import build as _b  # noqa: F401  (lazy build when run from the repo root)
            _b.build()
        L = C.CDLL(LIB_PATH)
        sig = {
            "orc_last_error": (C.c_char_p, []),
            "orc_free": (None, [C.c_void_p]),
            "orc_pd_create": (C.c_void_p, [C.c_int, C.c_long]),
            "orc_pd_set": (C.c_int, [C.c_void_p, C.c_char_p, _dp, C.c_long]),
            "orc_pd_get": (C.c_int, [C.c_void_p, C.c_char_p, _dp, C.c_long]),
            "orc_pd_size": (C.c_long, [C.c_void_p]),
            "orc_pd_num_vars": (C.c_int, [C.c_void_p]),
            "orc_model_snow": (C.c_void_p, [C.c_int] + [C.c_double] * 5),
            "orc_model_dummy": (C.c_void_p, [C.c_int]),
            "orc_model_lame": (C.c_int, [C.c_void_p, _dp]),
            "orc_model_set_particles": (C.c_int, [C.c_void_p, C.c_void_p]),
            "orc_model_update": (C.c_int, [C.c_void_p]),
            "orc_model_energy": (C.c_int, [C.c_void_p, C.c_long, _dp]),
            "orc_model_dEdF": (C.c_int, [C.c_void_p, C.c_long, _dp]),
            "orc_model_dEdF_diff": (C.c_int, [C.c_void_p, _dp, C.c_long, _dp]),
            "orc_svd3": (None, [_dp, _dp, _dp, _dp]),
            "orc_bspline_w": (C.c_double, [C.c_int, C.c_double]),
            "orc_bspline_dw": (C.c_double, [C.c_int, C.c_double]),
            "orc_cols_create": (C.c_void_p, [C.c_int]),
            "orc_cols_add_plane": (C.c_int, [C.c_void_p, _dp, _dp, C.c_double, C.c_int]),
            "orc_cols_add_sphere": (C.c_int, [C.c_void_p, _dp, C.c_double, _dp, C.c_double, C.c_int]),
            "orc_cols_add_volume": (C.c_int, [C.c_void_p, _dp, C.POINTER(C.c_int), _dp, C.c_double, _dp, _dp,
                                              C.c_double, C.c_int]),
            "orc_cols_collide": (C.c_int, [C.c_void_p, _dp, _dp, _dp, C.c_int]),
            "orc_fields_create": (C.c_void_p, [C.c_int]),
            "orc_fields_add_gravity": (C.c_int, [C.c_void_p, _dp]),
            "orc_grid_create": (C.c_void_p, [C.c_void_p, _ip, C.c_long, C.c_double, _dp, C.c_int]),
            "orc_grid_info": (C.c_int, [C.c_void_p, _ip, _dp]),
            "orc_grid_get": (C.c_int, [C.c_void_p, C.c_char_p, _dp]),
            "orc_grid_set": (C.c_int, [C.c_void_p, C.c_char_p, _dp]),
            "orc_grid_inds": (C.c_long, [C.c_void_p, _ip]),
            "orc_grid_partition": (C.c_long, [C.c_void_p, C.c_int, _lp]),
            "orc_grid_volumes": (C.c_int, [C.c_void_p]),
            "orc_grid_stencil": (C.c_int, [C.c_void_p, _dp, _ip, _dp, _dp]),
            "orc_grid_forces": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, _dp]),
            "orc_grid_force_diff": (C.c_int, [C.c_void_p, C.c_void_p, _dp, _dp]),
            "orc_grid_dfidxi": (C.c_int, [C.c_void_p, C.c_void_p, _dp]),
            "orc_grid_explicit_momenta": (C.c_int, [C.c_void_p, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p, _dp, _bp]),
            "orc_grid_collision_velocities": (C.c_int, [C.c_void_p, C.c_void_p, _bp, _dp]),
            "orc_grid_update_velocities": (C.c_int, [C.c_void_p, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p,
                                                     C.c_int, C.c_double, _dp, _dp, C.c_int, _ip, _dp, _dp]),
            "orc_grid_update_velocities_iterates": (C.c_int, [C.c_void_p, C.c_double, C.c_void_p, C.c_void_p,
                                                              C.c_void_p, C.c_int, C.c_double, _dp, C.c_int, _ip]),
            "orc_grid_update_F": (C.c_int, [C.c_void_p, C.c_double]),
            "orc_grid_update_pv": (C.c_int, [C.c_void_p]),
            "orc_voxel_sort": (C.c_int, [C.c_int, _ip, C.c_long, C.c_double, _dp]),
            "orc_cr_dense": (C.c_int, [C.c_int, C.c_int, _dp, _dp, _dp, _dp, C.c_int, C.c_double, _dp, _dp, _dp,
                                       C.c_int, _ip]),
            "orc_sim_create": (C.c_void_p, [C.c_int, C.c_long, _dp, _dp, C.c_double, C.c_void_p, C.c_void_p,
                                            C.c_void_p, C.c_int, _dp, _dp, _dp]),
            "orc_sim_pd": (C.c_void_p, [C.c_void_p]),
            "orc_sim_advance": (C.c_int, [C.c_void_p, C.c_double, C.c_int, C.c_double]),
            "orc_sim_advance_cancel": (C.c_int, [C.c_void_p, C.c_double, C.c_int, C.c_double, C.c_int,
                                                 C.POINTER(C.c_int)]),
            "orc_sim_body_stats": (C.c_int, [C.c_void_p, C.c_long, _dp, _dp, C.c_int]),
            "orc_sim_num_stats": (C.c_long, [C.c_void_p]),
            "orc_sim_num_bodies": (C.c_long, [C.c_void_p]),
            "orc_sim_body": (C.c_long, [C.c_void_p, C.c_long, _ip]),
            "orc_sim_ballistic": (C.c_long, [C.c_void_p, _ip]),
            "orc_num_threads": (C.c_int, []),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


class OracleError(RuntimeError):
    pass


def _check(rc):
    if rc != 0:
        raise OracleError(lib().orc_last_error().decode())


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_dp)


def _i(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(_ip)


def _prec(p):
    return {"float": 0, "f32": 0, 0: 0, "double": 1, "f64": 1, 1: 1}[p]


def to_colmajor(M):
    """(n,3,3) row/col -> flat column-major (Eigen Matrix3f memory order)."""
    M = np.asarray(M, dtype=np.float64).reshape(-1, 3, 3)
    return np.ascontiguousarray(M.transpose(0, 2, 1)).reshape(-1)


def from_colmajor(a):
    return np.asarray(a, dtype=np.float64).reshape(-1, 3, 3).transpose(0, 2, 1).copy()


class _H:
    def __init__(self, ptr, owner=True):
        if not ptr:
            raise OracleError(lib().orc_last_error().decode())
        self.ptr = ptr
        self._owner = owner

    def __del__(self):
        if getattr(self, "_owner", False) and self.ptr and _lib is not None:
            _lib.orc_free(self.ptr)
            self.ptr = None


class ParticleData(_H):
    """Reference MaterialPointData (include/MpmSim/MaterialPointData.h)."""

    @classmethod
    def create(cls, n, prec="float"):
        return cls(lib().orc_pd_create(_prec(prec), n))

    def size(self):
        return lib().orc_pd_size(self.ptr)

    def num_variables(self):
        return lib().orc_pd_num_vars(self.ptr)

    def set(self, name, arr):
        n = self.size()
        if name in MAT3:
            flat = to_colmajor(arr)
        else:
            flat = np.asarray(arr, dtype=np.float64).reshape(-1)
        flat, p = _d(flat)
        _check(lib().orc_pd_set(self.ptr, name.encode(), p, flat.size))

    def get(self, name):
        n = self.size()
        cnt = n * (9 if name in MAT3 else 3 if name in VEC3 else 1)
        out = np.zeros(cnt)
        _check(lib().orc_pd_get(self.ptr, name.encode(), out.ctypes.data_as(_dp), cnt))
        if name in MAT3:
            return from_colmajor(out)
        if name in VEC3:
            return out.reshape(n, 3)
        return out


class Model(_H):
    @classmethod
    def snow(cls, E, nu, xi, theta_c, theta_s, prec="float"):
        return cls(lib().orc_model_snow(_prec(prec), E, nu, xi, theta_c, theta_s))

    @classmethod
    def dummy(cls, prec="float"):
        return cls(lib().orc_model_dummy(_prec(prec)))

    def lame(self):
        out = np.zeros(2)
        _check(lib().orc_model_lame(self.ptr, out.ctypes.data_as(_dp)))
        return out

    def set_particles(self, pd):
        self._pd = pd
        _check(lib().orc_model_set_particles(self.ptr, pd.ptr))

    def update_particle_data(self):
        _check(lib().orc_model_update(self.ptr))

    def energy_density(self, p):
        out = C.c_double()
        _check(lib().orc_model_energy(self.ptr, p, C.byref(out)))
        return out.value

    def dEdF(self, p):
        out = np.zeros(9)
        _check(lib().orc_model_dEdF(self.ptr, p, out.ctypes.data_as(_dp)))
        return from_colmajor(out)[0]

    def dEdF_differential(self, dF, p):
        a, pa = _d(to_colmajor(dF))
        out = np.zeros(9)
        _check(lib().orc_model_dEdF_diff(self.ptr, pa, p, out.ctypes.data_as(_dp)))
        return from_colmajor(out)[0]


class Colliders(_H):
    @classmethod
    def create(cls, prec="float"):
        return cls(lib().orc_cols_create(_prec(prec)))

    def add_plane(self, coeffs, v=(0, 0, 0), friction=0.5, sticky=False):
        c, pc = _d(coeffs)
        vv, pv = _d(v)
        _check(lib().orc_cols_add_plane(self.ptr, pc, pv, friction, int(sticky)))

    def add_sphere(self, center, radius, v=(0, 0, 0), friction=0.5, sticky=False):
        c, pc = _d(center)
        vv, pv = _d(v)
        _check(lib().orc_cols_add_sphere(self.ptr, pc, radius, pv, friction, int(sticky)))

    def add_volume(self, phi, origin, voxel, v=(0, 0, 0), friction=0.5, sticky=False):
        """Level-set volume: phi[k, j, i] at origin + (i, j, k) voxel; v a 3-vector (uniform) or an
        array of shape phi.shape + (3,)."""
        ph = np.ascontiguousarray(phi, dtype=np.float64)
        dims = (C.c_int * 3)(ph.shape[2], ph.shape[1], ph.shape[0])
        oo, po = _d(origin)
        va = np.asarray(v, dtype=np.float64)
        if va.shape == (3,):
            uv, pu = _d(va)
            pvel = None
        else:
            vf = np.ascontiguousarray(va.reshape(ph.shape + (3,)))
            pvel = vf.ctypes.data_as(_dp)
            uv, pu = _d(np.zeros(3))
        _check(lib().orc_cols_add_volume(self.ptr, ph.ctypes.data_as(_dp), dims, po, float(voxel), pvel, pu,
                                         friction, int(sticky)))

    def collide(self, v, x, frame_v=(0, 0, 0), add_collision_velocity=False):
        vv, pv = _d(np.array(v, dtype=np.float64))
        xx, px = _d(x)
        ff, pf = _d(frame_v)
        flag = lib().orc_cols_collide(self.ptr, pv, px, pf, int(add_collision_velocity))
        return flag, vv


class Fields(_H):
    @classmethod
    def create(cls, prec="float"):
        return cls(lib().orc_fields_create(_prec(prec)))

    def add_gravity(self, g):
        gg, pg = _d(g)
        _check(lib().orc_fields_add_gravity(self.ptr, pg))


class Grid(_H):
    """Reference Grid (include/MpmSim/Grid.h): construction does the P2G splat."""

    def __init__(self, pd, inds, h, frame_v=(0, 0, 0), dim=3):
        ii, pi = _i(inds)
        ff, pf = _d(frame_v)
        super().__init__(lib().orc_grid_create(pd.ptr, pi, ii.size, h, pf, dim))
        self.pd = pd
        n = np.zeros(3, dtype=np.int32)
        mn = np.zeros(3)
        _check(lib().orc_grid_info(self.ptr, n.ctypes.data_as(_ip), mn.ctypes.data_as(_dp)))
        self.n = n
        self.min = mn
        self.h = h
        self.num_nodes = int(np.prod(n.astype(np.int64)))

    def get(self, name):
        cnt = self.num_nodes * (3 if "elocities" in name else 1)
        out = np.zeros(cnt)
        _check(lib().orc_grid_get(self.ptr, name.encode(), out.ctypes.data_as(_dp)))
        return out

    def set(self, name, arr):
        a, p = _d(arr)
        _check(lib().orc_grid_set(self.ptr, name.encode(), p))

    def inds(self):
        n = lib().orc_grid_inds(self.ptr, None)
        out = np.zeros(n, dtype=np.int32)
        lib().orc_grid_inds(self.ptr, out.ctypes.data_as(_ip))
        return out

    def partition(self, colour):
        n = lib().orc_grid_partition(self.ptr, colour, None)
        out = np.zeros(2 * n, dtype=np.int64)
        lib().orc_grid_partition(self.ptr, colour, out.ctypes.data_as(_lp))
        return out.reshape(n, 2)

    def stencil(self, x):
        xx, px = _d(x)
        nodes = np.zeros(64 * 3, dtype=np.int32)
        w = np.zeros(64)
        dw = np.zeros(64 * 3)
        _check(lib().orc_grid_stencil(self.ptr, px, nodes.ctypes.data_as(_ip), w.ctypes.data_as(_dp),
                                      dw.ctypes.data_as(_dp)))
        return nodes.reshape(64, 3), w, dw.reshape(64, 3)

    def compute_particle_volumes(self):
        _check(lib().orc_grid_volumes(self.ptr))

    def calculate_forces(self, model, fields):
        out = np.zeros(3 * self.num_nodes)
        _check(lib().orc_grid_forces(self.ptr, model.ptr, fields.ptr, out.ctypes.data_as(_dp)))
        return out

    def force_differentials(self, model, dx):
        d, pd_ = _d(dx)
        out = np.zeros(3 * self.num_nodes)
        _check(lib().orc_grid_force_diff(self.ptr, model.ptr, pd_, out.ctypes.data_as(_dp)))
        return out

    def dforce_dxi(self, model):
        out = np.zeros(3 * self.num_nodes)
        _check(lib().orc_grid_dfidxi(self.ptr, model.ptr, out.ctypes.data_as(_dp)))
        return out

    def explicit_momenta(self, dt, model, cols, fields):
        em = np.zeros(3 * self.num_nodes)
        col = np.zeros(self.num_nodes, dtype=np.int8)
        _check(lib().orc_grid_explicit_momenta(self.ptr, dt, model.ptr, cols.ptr, fields.ptr,
                                               em.ctypes.data_as(_dp), col.ctypes.data_as(_bp)))
        return em, col

    def collision_velocities(self, cols, collided):
        c = np.ascontiguousarray(collided, dtype=np.int8)
        out = np.zeros(3 * self.num_nodes)
        _check(lib().orc_grid_collision_velocities(self.ptr, cols.ptr, c.ctypes.data_as(_bp), out.ctypes.data_as(_dp)))
        return out

    def update_grid_velocities(self, dt, model, cols, fields, max_iters, tol, want_rhs=False, hist_cap=10000):
        stats = np.zeros(3)
        hist = np.zeros(hist_cap)
        hl = C.c_int()
        rhs = np.zeros(3 * self.num_nodes) if want_rhs else None
        guess = np.zeros(3)
        _check(lib().orc_grid_update_velocities(self.ptr, dt, model.ptr, cols.ptr, fields.ptr, max_iters, tol,
                                                stats.ctypes.data_as(_dp), hist.ctypes.data_as(_dp), hist_cap,
                                                C.byref(hl), rhs.ctypes.data_as(_dp) if want_rhs else None,
                                                guess.ctypes.data_as(_dp)))
        out = {"iterations": int(stats[0]), "bnorm2": stats[1], "early_exit": bool(stats[2]),
               "residual_norm2": hist[:min(hl.value, hist_cap)].copy(), "guess": guess}
        if want_rhs:
            out["rhs"] = rhs
        return out

    def update_grid_velocities_iterates(self, dt, model, cols, fields, max_iters, tol, cap=512):
        """The solve of update_grid_velocities with the reference's debug hook: the CR iterates x
        (relative to the colliders, one per x update), as a (calls, 3 n) array."""
        its = np.zeros((cap, 3 * self.num_nodes))
        cnt = C.c_int()
        _check(lib().orc_grid_update_velocities_iterates(self.ptr, dt, model.ptr, cols.ptr, fields.ptr, max_iters,
                                                         tol, its.ctypes.data_as(_dp), cap, C.byref(cnt)))
        return its[:min(cnt.value, cap)].copy()

    def update_deformation_gradients(self, dt):
        _check(lib().orc_grid_update_F(self.ptr, dt))

    def update_particle_velocities(self):
        _check(lib().orc_grid_update_pv(self.ptr))


def voxel_sort(inds, voxel_size, x, prec="float"):
    ii = np.ascontiguousarray(inds, dtype=np.int32).copy()
    xx, px = _d(np.asarray(x).reshape(-1))
    _check(lib().orc_voxel_sort(_prec(prec), ii.ctypes.data_as(_ip), ii.size, voxel_size, px))
    return ii


def cr_dense(A, b, x0=None, diag=None, max_iters=100, tol=0.0, prec="float", log_cap=64):
    A = np.asarray(A, dtype=np.float64)
    N = A.shape[0]
    Acm, pA = _d(A.T.reshape(-1))  # column-major
    bb, pb = _d(b)
    x = np.zeros(N) if x0 is None else np.array(x0, dtype=np.float64)
    x = np.ascontiguousarray(x)
    dd = None
    pdg = None
    if diag is not None:
        dd, pdg = _d(diag)
    stats = np.zeros(3)
    logR = np.zeros(log_cap * N)
    logP = np.zeros(log_cap * N)
    ll = C.c_int()
    _check(lib().orc_cr_dense(_prec(prec), N, pA, pb, x.ctypes.data_as(_dp), pdg, max_iters, tol,
                              stats.ctypes.data_as(_dp), logR.ctypes.data_as(_dp), logP.ctypes.data_as(_dp),
                              log_cap, C.byref(ll)))
    L = min(ll.value, log_cap)
    return {"x": x, "iterations": int(stats[0]), "bnorm2": stats[1], "early_exit": bool(stats[2]),
            "residuals": logR[:L * N].reshape(L, N), "directions": logP[:L * N].reshape(L, N)}


def svd3(A):
    A = np.asarray(A, dtype=np.float64)
    a = np.ascontiguousarray(A.T.reshape(-1))
    U = np.zeros(9)
    s = np.zeros(3)
    V = np.zeros(9)
    lib().orc_svd3(a.ctypes.data_as(_dp), U.ctypes.data_as(_dp), s.ctypes.data_as(_dp), V.ctypes.data_as(_dp))
    return U.reshape(3, 3).T.copy(), s, V.reshape(3, 3).T.copy()


def bspline_w(x, prec="float"):
    return lib().orc_bspline_w(_prec(prec), float(x))


def bspline_dw(x, prec="float"):
    return lib().orc_bspline_dw(_prec(prec), float(x))


class Sim(_H):
    """Reference Sim (include/MpmSim/Sim.h)."""

    def __init__(self, x, m, h, model, cols, fields, dim=3, mu0=None, lambda0=None, xi=None, prec="float"):
        xx, px = _d(np.asarray(x).reshape(-1))
        mm, pm = _d(m)
        extra = [None, None, None]
        keep = []
        if mu0 is not None:
            for k, a in enumerate((mu0, lambda0, xi)):
                arr, p = _d(a)
                keep.append(arr)
                extra[k] = p
        super().__init__(lib().orc_sim_create(_prec(prec), mm.size, px, pm, h, model.ptr, cols.ptr, fields.ptr, dim,
                                              *extra))
        self.model, self.cols, self.fields = model, cols, fields
        self.pd = ParticleData(lib().orc_sim_pd(self.ptr), owner=False)
        self.pd._sim = self

    def advance(self, dt, max_iters, tol, cancel_on_call=None):
        """One Sim::advance; cancel_on_call = n makes the termination's cancelled() true from its n-th
        call on (one call after each body's solve); returns whether the step was cancelled."""
        if cancel_on_call is None:
            _check(lib().orc_sim_advance(self.ptr, dt, max_iters, tol))
            return False
        was = C.c_int(0)
        _check(lib().orc_sim_advance_cancel(self.ptr, dt, max_iters, tol, int(cancel_on_call), C.byref(was)))
        return bool(was.value)

    def bodies(self):
        nb = lib().orc_sim_num_bodies(self.ptr)
        out = []
        for b in range(nb):
            n = lib().orc_sim_body(self.ptr, b, None)
            a = np.zeros(n, dtype=np.int32)
            lib().orc_sim_body(self.ptr, b, a.ctypes.data_as(_ip))
            out.append(a)
        return out

    def ballistic(self):
        n = lib().orc_sim_ballistic(self.ptr, None)
        a = np.zeros(n, dtype=np.int32)
        lib().orc_sim_ballistic(self.ptr, a.ctypes.data_as(_ip))
        return a

    def last_stats(self, cap=10000):
        out = []
        for b in range(lib().orc_sim_num_stats(self.ptr)):
            s = np.zeros(10)
            hist = np.zeros(cap)
            L = lib().orc_sim_body_stats(self.ptr, b, s.ctypes.data_as(_dp), hist.ctypes.data_as(_dp), cap)
            out.append({"iterations": int(s[0]), "bnorm2": s[1], "early_exit": bool(s[2]), "frame_velocity": s[3:6],
                        "n": s[6:9].astype(int), "particles": int(s[9]), "residual_norm2": hist[:L].copy()})
        return out


def num_threads():
    return lib().orc_num_threads()
