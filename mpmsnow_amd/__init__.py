"""mpmsnow_amd: MI355X (gfx950) implicit MPM snow step behind a C-ABI.

This module binds include/mpmsnow_gpu.h through ctypes and mirrors the
reference MpmSim classes (Sim, Grid, SnowConstitutiveModel, ...) closely
enough that parity tests read like the reference's own tests. All compute
runs in mpmsnow_amd/_build/libmpmsnow_hip.so; there is no CPU fallback: if
the library or a GPU is missing, every entry point raises.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libmpmsnow_hip.so")

MPM_OK = 0
MPM_ERR_INVALID = 1

VEC3 = {"p", "v"}
MAT3 = {"F", "Fp", "FinvTrans", "R", "Ginv"}

_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int32)
_bp = C.POINTER(C.c_int8)
_dp = C.POINTER(C.c_double)

TERM_CB = C.CFUNCTYPE(C.c_int, C.c_void_p, _fp, C.c_int64, C.c_int)
DEBUG_CB = C.CFUNCTYPE(None, C.c_void_p, _fp, C.c_int64)
CANCEL_CB = C.CFUNCTYPE(C.c_int, C.c_void_p)


class Collider(C.Structure):
    _fields_ = [("type", C.c_int), ("coeffs", C.c_float * 4), ("center", C.c_float * 3), ("radius", C.c_float),
                ("velocity", C.c_float * 3), ("friction", C.c_float), ("sticky", C.c_int), ("dims", C.c_int * 3),
                ("origin", C.c_float * 3), ("voxel", C.c_float), ("phi", _fp), ("phi_velocity", _fp)]


class Collider64(C.Structure):
    _fields_ = [("type", C.c_int), ("coeffs", C.c_double * 4), ("center", C.c_double * 3), ("radius", C.c_double),
                ("velocity", C.c_double * 3), ("friction", C.c_double), ("sticky", C.c_int), ("dims", C.c_int * 3),
                ("origin", C.c_double * 3), ("voxel", C.c_double), ("phi", _dp), ("phi_velocity", _dp)]


class SolverParams(C.Structure):
    _fields_ = [("max_iters", C.c_int), ("tolerance", C.c_float), ("termination", TERM_CB), ("debug", DEBUG_CB),
                ("cancelled", CANCEL_CB), ("user", C.c_void_p)]


class SolveStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("early_exit", C.c_int), ("rhs_norm2", C.c_double),
                ("residual_norm2", C.c_double), ("history_length", C.c_int), ("history_capacity", C.c_int),
                ("history", _dp), ("particles", C.c_int), ("active_nodes", C.c_int), ("fallbacks", C.c_int),
                ("restarts", C.c_int)]


MATVEC_CB = C.CFUNCTYPE(None, C.c_void_p, _fp, _fp, C.c_int64)
PROJECT_CB = C.CFUNCTYPE(None, C.c_void_p, _fp, C.c_int64)


class ProceduralMatrix(C.Structure):
    _fields_ = [("mult", MATVEC_CB), ("mult_inverse", MATVEC_CB), ("project", PROJECT_CB), ("user", C.c_void_p),
                ("device_pointers", C.c_int)]


class StepStats(C.Structure):
    _fields_ = [("num_bodies", C.c_int), ("num_ballistic", C.c_int), ("total_iterations", C.c_int),
                ("max_iterations", C.c_int), ("cancelled", C.c_int), ("num_solved", C.c_int)]


class MpmError(RuntimeError):
    pass


_lib = None

# every symbol include/mpmsnow_gpu.h declares, with its ctypes signature
SIGNATURES = {
    "mpm_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int]),
    "mpm_destroy": (C.c_int, [C.c_void_p]),
    "mpm_last_error": (C.c_char_p, [C.c_void_p]),
    "mpm_version": (C.c_int, []),
    "mpm_particles_create": (C.c_int, [C.c_void_p, C.c_int64]),
    "mpm_particles_count": (C.c_int64, [C.c_void_p]),
    "mpm_particles_num_variables": (C.c_int, [C.c_void_p]),
    "mpm_particles_set": (C.c_int, [C.c_void_p, C.c_char_p, _fp, C.c_int64]),
    "mpm_particles_get": (C.c_int, [C.c_void_p, C.c_char_p, _fp, C.c_int64]),
    "mpm_model_snow": (C.c_int, [C.c_void_p] + [C.c_float] * 5),
    "mpm_model_update_particle_data": (C.c_int, [C.c_void_p]),
    "mpm_model_energy_density": (C.c_int, [C.c_void_p, _ip, C.c_int64, _fp]),
    "mpm_model_stress": (C.c_int, [C.c_void_p, _ip, C.c_int64, _fp]),
    "mpm_model_stress_differential": (C.c_int, [C.c_void_p, _ip, _fp, C.c_int64, _fp]),
    # device pointers (particles, dF, out)
    "mpm_model_stress_differential_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64,
                                                        C.c_void_p]),
    "mpm_set_colliders": (C.c_int, [C.c_void_p, C.POINTER(Collider), C.c_int]),
    "mpm_set_gravity": (C.c_int, [C.c_void_p, _fp, C.c_int]),
    "mpm_grid_create": (C.c_int, [C.c_void_p, _ip, C.c_int64, C.c_float, _fp, C.c_int, C.POINTER(C.c_void_p)]),
    "mpm_grid_destroy": (C.c_int, [C.c_void_p]),
    "mpm_grid_dims": (C.c_int, [C.c_void_p, _ip, _fp]),
    "mpm_grid_get": (C.c_int, [C.c_void_p, C.c_char_p, C.c_void_p]),
    "mpm_grid_set": (C.c_int, [C.c_void_p, C.c_char_p, C.c_void_p]),
    "mpm_grid_compute_particle_volumes": (C.c_int, [C.c_void_p]),
    "mpm_grid_calculate_forces": (C.c_int, [C.c_void_p, _fp]),
    "mpm_grid_force_differentials": (C.c_int, [C.c_void_p, _fp, _fp]),
    "mpm_grid_dforce_dxi": (C.c_int, [C.c_void_p, _fp]),
    "mpm_grid_explicit_momenta": (C.c_int, [C.c_void_p, C.c_float, _fp, _bp]),
    "mpm_grid_collision_velocities": (C.c_int, [C.c_void_p, _bp, _fp]),
    "mpm_grid_update_grid_velocities": (C.c_int, [C.c_void_p, C.c_float, C.POINTER(SolverParams),
                                                  C.POINTER(SolveStats)]),
    "mpm_grid_update_particle_velocities": (C.c_int, [C.c_void_p]),
    "mpm_grid_update_deformation_gradients": (C.c_int, [C.c_void_p, C.c_float]),
    "mpm_conjugate_residuals": (C.c_int, [C.c_void_p, C.POINTER(ProceduralMatrix), C.POINTER(ProceduralMatrix),
                                          _fp, _fp, C.c_int64, C.POINTER(SolverParams), C.POINTER(SolveStats),
                                          _fp, _fp, C.c_int, _ip]),
    "mpm_sim_init": (C.c_int, [C.c_void_p, C.c_float, C.c_int]),
    "mpm_sim_advance": (C.c_int, [C.c_void_p, C.c_float, C.POINTER(SolverParams), C.POINTER(StepStats)]),
    "mpm_sim_num_bodies": (C.c_int, [C.c_void_p]),
    "mpm_sim_grid_layout": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64)]),
    "mpm_sim_body": (C.c_int64, [C.c_void_p, C.c_int, _ip]),
    "mpm_sim_ballistic": (C.c_int64, [C.c_void_p, _ip]),
    "mpm_sim_body_solve_stats": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(SolveStats)]),
    "mpm_lpc_num_params": (C.c_int64, []),
    "mpm_lpc_init": (C.c_int, [C.c_void_p, C.c_uint64]),
    "mpm_lpc_set_params": (C.c_int, [C.c_void_p, _fp, C.c_int64]),
    "mpm_lpc_get_params": (C.c_int, [C.c_void_p, _fp, C.c_int64]),
    "mpm_lpc_set_mode": (C.c_int, [C.c_void_p, C.c_int]),
    "mpm_lpc_set_fallback": (C.c_int, [C.c_void_p, C.c_double]),
    "mpm_lpc_set_directions": (C.c_int, [C.c_void_p, C.c_int]),
    "mpm_lpc_set_handoff": (C.c_int, [C.c_void_p, C.c_int]),
    "mpm_lpc_set_capture_iterations": (C.c_int, [C.c_void_p, _ip, C.c_int]),
    "mpm_lpc_train_next": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_float]),
    "mpm_lpc_set_training_loss": (C.c_int, [C.c_void_p, C.c_int]),
    "mpm_lpc_set_strikes": (C.c_int, [C.c_void_p, C.c_int]),
    "mpm_lpc_last_pairs": (C.c_int, [C.c_void_p]),
    "mpm_lpc_last_loss": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                    C.POINTER(C.c_double)]),
    "mpm_grid_lpc_apply": (C.c_int, [C.c_void_p, _fp, _fp, C.c_float, _fp]),
    "mpm_grid_lpc_gradients": (C.c_int, [C.c_void_p, _fp, _fp, _fp, C.c_float, C.c_float, _fp,
                                         C.POINTER(C.c_double)]),
    "mpm_sim64_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int, C.c_int64]),
    "mpm_sim64_destroy": (C.c_int, [C.c_void_p]),
    "mpm_sim64_last_error": (C.c_char_p, [C.c_void_p]),
    "mpm_sim64_set": (C.c_int, [C.c_void_p, C.c_char_p, _dp, C.c_int64]),
    "mpm_sim64_get": (C.c_int, [C.c_void_p, C.c_char_p, _dp, C.c_int64]),
    "mpm_sim64_model_snow": (C.c_int, [C.c_void_p] + [C.c_double] * 5),
    "mpm_sim64_set_colliders": (C.c_int, [C.c_void_p, C.POINTER(Collider64), C.c_int]),
    "mpm_sim64_set_gravity": (C.c_int, [C.c_void_p, _dp, C.c_int]),
    "mpm_sim64_init": (C.c_int, [C.c_void_p, C.c_double]),
    "mpm_sim64_advance": (C.c_int, [C.c_void_p, C.c_double, C.c_int, C.c_double, C.POINTER(StepStats)]),
    "mpm_sim64_num_bodies": (C.c_int, [C.c_void_p]),
    "mpm_sim64_body_solve_stats": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(SolveStats)]),
    "mpm_set_small_bodies": (C.c_int, [C.c_void_p, C.c_int]),
    "mpm_profile_enable": (C.c_int, [C.c_void_p, C.c_int]),
    "mpm_profile_period": (C.c_int, [C.c_void_p, C.c_int]),
    "mpm_profile_query": (C.c_int, [C.c_void_p, C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_int64),
                                    C.POINTER(C.c_double)]),
}

PROFILE_KERNELS = ("p2g_mass", "p2g_vel", "forces", "hvp", "diag", "g2p", "cr_update_x", "cr_ar", "cr_update_px",
                   "lpc_forward", "cr_small")


def lib():
    """Load the gfx950 library; raises if it was not built (no fallback path exists)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise MpmError("libmpmsnow_hip.so is not built: run __graft_entry__.build() or python mpmsnow_amd/build.py")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _f(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a, a.ctypes.data_as(_fp)


def _i(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(_ip)


def to_colmajor(M):
    M = np.asarray(M, dtype=np.float32).reshape(-1, 3, 3)
    return np.ascontiguousarray(M.transpose(0, 2, 1)).reshape(-1)


def from_colmajor(a):
    return np.asarray(a).reshape(-1, 3, 3).transpose(0, 2, 1).copy()


def make_params(max_iters, tol, termination=None, debug=None, cancelled=None):
    """SolverParams for SquareMagnitudeTermination(max_iters, tol) plus optional Python hooks."""
    p = SolverParams()
    p.max_iters = int(max_iters)
    p.tolerance = float(tol)
    keep = []
    if termination is not None:
        def _t(user, r, n, it):
            return int(bool(termination(np.ctypeslib.as_array(r, shape=(n,)).copy(), it)))
        cb = TERM_CB(_t)
        keep.append(cb)
        p.termination = cb
    if debug is not None:
        def _d(user, x, n):
            debug(np.ctypeslib.as_array(x, shape=(n,)).copy())
        cb = DEBUG_CB(_d)
        keep.append(cb)
        p.debug = cb
    if cancelled is not None:
        cb = CANCEL_CB(lambda user: int(bool(cancelled())))
        keep.append(cb)
        p.cancelled = cb
    p._keep = keep
    return p


def _stats_dict(st, hist):
    return {"iterations": st.iterations, "early_exit": bool(st.early_exit), "bnorm2": st.rhs_norm2,
            "residual_norm2_last": st.residual_norm2,
            "residual_norm2": hist[:st.history_length].copy() if hist is not None else None,
            "particles": st.particles, "active_nodes": st.active_nodes, "fallbacks": st.fallbacks,
            "restarts": st.restarts}


def _procedural(mult, mult_inverse=None, project=None):
    """A host-pointer ProceduralMatrix over numpy callables: mult(x) -> M x, mult_inverse(x) -> M^-1 x,
    project(x) -> P x (ProceduralMatrix.h:8-19)."""
    keep = []
    m = ProceduralMatrix()

    def _wrap(f):
        def _cb(user, x, out, n):
            res = np.asarray(f(np.ctypeslib.as_array(x, shape=(n,)).copy()), dtype=np.float32).reshape(n)
            np.ctypeslib.as_array(out, shape=(n,))[:] = res
        cb = MATVEC_CB(_cb)
        keep.append(cb)
        return cb

    m.mult = _wrap(mult)
    if mult_inverse is not None:
        m.mult_inverse = _wrap(mult_inverse)
    if project is not None:
        def _p(user, x, n):
            v = np.ctypeslib.as_array(x, shape=(n,))
            v[:] = np.asarray(project(v.copy()), dtype=np.float32).reshape(n)
        cb = PROJECT_CB(_p)
        keep.append(cb)
        m.project = cb
    m.device_pointers = 0
    m._keep = keep
    return m


def conjugate_residuals(ctx, mult, b, x0, max_iters, tol, preconditioner=None, project=None, log=0,
                        termination=None, debug=None, hist_cap=4096):
    """ConjugateResiduals(termination, preconditioner, log)(mat, rhs, x, debug)
    (ConjugateResiduals.cpp:19-160) on the GPU over a matrix given by the callable mult (and project,
    its subspaceProject). preconditioner: (mult, mult_inverse) callables or None. log: capacity of the
    residual / search-direction logs (0: none). Returns x, the stats and the logs."""
    bb = np.ascontiguousarray(b, dtype=np.float32).ravel()
    x = np.ascontiguousarray(x0, dtype=np.float32).ravel().copy()
    n = bb.size
    A = _procedural(mult, project=project)
    P = _procedural(*preconditioner) if preconditioner is not None else None
    prm = make_params(max_iters, tol, termination, debug)
    hist = np.zeros(hist_cap)
    st = SolveStats()
    st.history_capacity = hist_cap
    st.history = hist.ctypes.data_as(_dp)
    res = np.zeros((max(log, 1), n), dtype=np.float32)
    dirs = np.zeros((max(log, 1), n), dtype=np.float32)
    logged = C.c_int32(0)
    ctx.check(lib().mpm_conjugate_residuals(ctx.ptr, C.byref(A), C.byref(P) if P is not None else None,
                                            bb.ctypes.data_as(_fp), x.ctypes.data_as(_fp), n, C.byref(prm),
                                            C.byref(st), res.ctypes.data_as(_fp), dirs.ctypes.data_as(_fp),
                                            int(log), C.byref(logged)))
    k = logged.value
    return x, _stats_dict(st, hist), res[:k].copy(), dirs[:k].copy()


class Context:
    """One device context: the particle store (MaterialPointData), the snow model, colliders and fields."""

    def __init__(self, device=0):
        h = C.c_void_p()
        rc = lib().mpm_create(C.byref(h), device)
        if rc != MPM_OK:
            raise MpmError("mpm_create failed (code %d): no usable GPU" % rc)
        self.ptr = h.value
        self.n = 0

    def close(self):
        if getattr(self, "ptr", None) and _lib is not None:
            _lib.mpm_destroy(self.ptr)
            self.ptr = None

    __del__ = close

    def check(self, rc):
        if rc != MPM_OK:
            raise MpmError(lib().mpm_last_error(self.ptr).decode())

    # MaterialPointData
    def particles_create(self, n):
        self.check(lib().mpm_particles_create(self.ptr, int(n)))
        self.n = int(n)

    def num_variables(self):
        return lib().mpm_particles_num_variables(self.ptr)

    def set(self, name, arr):
        flat = to_colmajor(arr) if name in MAT3 else np.asarray(arr, dtype=np.float32).reshape(-1)
        a, p = _f(flat)
        self.check(lib().mpm_particles_set(self.ptr, name.encode(), p, a.size))

    def get(self, name):
        w = 9 if name in MAT3 else 3 if name in VEC3 else 1
        out = np.zeros(self.n * w, dtype=np.float32)
        self.check(lib().mpm_particles_get(self.ptr, name.encode(), out.ctypes.data_as(_fp), out.size))
        if name in MAT3:
            return from_colmajor(out)
        if name in VEC3:
            return out.reshape(self.n, 3)
        return out

    # SnowConstitutiveModel
    def model_snow(self, E, nu, xi, theta_c, theta_s):
        self.check(lib().mpm_model_snow(self.ptr, E, nu, xi, theta_c, theta_s))

    def update_particle_data(self):
        self.check(lib().mpm_model_update_particle_data(self.ptr))

    def energy_density(self, particles):
        ii, pi = _i(particles)
        out = np.zeros(ii.size, dtype=np.float32)
        self.check(lib().mpm_model_energy_density(self.ptr, pi, ii.size, out.ctypes.data_as(_fp)))
        return out

    def stress(self, particles):
        ii, pi = _i(particles)
        out = np.zeros(9 * ii.size, dtype=np.float32)
        self.check(lib().mpm_model_stress(self.ptr, pi, ii.size, out.ctypes.data_as(_fp)))
        return from_colmajor(out)

    def stress_differential(self, particles, dF):
        ii, pi = _i(particles)
        d, pd = _f(to_colmajor(dF))
        out = np.zeros(9 * ii.size, dtype=np.float32)
        self.check(lib().mpm_model_stress_differential(self.ptr, pi, pd, ii.size, out.ctypes.data_as(_fp)))
        return from_colmajor(out)

    # CollisionObjectSet / ForceFieldSet
    def set_colliders(self, planes=(), spheres=(), volumes=()):
        """planes: (coeffs4, velocity3, friction, sticky); spheres: (center3, radius, velocity3, friction, sticky);
        volumes (level sets, the reference's VDBCollisionObject): (phi, origin3, voxel, velocity, friction,
        sticky) with phi[k, j, i] the value at origin + (i, j, k) voxel and velocity a 3-vector or an array
        of shape phi.shape + (3,). Order in the set: planes, spheres, volumes."""
        items = []
        keep = []
        for coeffs, v, mu, sticky in planes:
            c = Collider()
            c.type = 0
            c.coeffs[:] = list(map(float, coeffs))
            c.velocity[:] = list(map(float, v))
            c.friction = mu
            c.sticky = int(sticky)
            items.append(c)
        for center, r, v, mu, sticky in spheres:
            c = Collider()
            c.type = 1
            c.center[:] = list(map(float, center))
            c.radius = r
            c.velocity[:] = list(map(float, v))
            c.friction = mu
            c.sticky = int(sticky)
            items.append(c)
        for phi, origin, voxel, v, mu, sticky in volumes:
            ph = np.ascontiguousarray(phi, dtype=np.float32)
            if ph.ndim != 3:
                raise ValueError("phi must be a 3-D array indexed [k, j, i]")
            c = Collider()
            c.type = 2
            c.dims[:] = [ph.shape[2], ph.shape[1], ph.shape[0]]
            c.origin[:] = list(map(float, origin))
            c.voxel = float(voxel)
            c.phi = ph.ctypes.data_as(_fp)
            keep.append(ph)
            va = np.asarray(v, dtype=np.float32)
            if va.shape == (3,):
                c.velocity[:] = list(map(float, va))
            else:
                va = np.ascontiguousarray(va.reshape(ph.shape + (3,)))
                c.phi_velocity = va.ctypes.data_as(_fp)
                keep.append(va)
            c.friction = mu
            c.sticky = int(sticky)
            items.append(c)
        arr = (Collider * max(1, len(items)))(*items)
        self.check(lib().mpm_set_colliders(self.ptr, arr, len(items)))

    def set_gravity(self, gs):
        g = np.asarray(gs, dtype=np.float32).reshape(-1, 3)
        a, p = _f(g.reshape(-1))
        self.check(lib().mpm_set_gravity(self.ptr, p, len(g)))

    # ---- learned preconditioner
    def lpc_init(self, seed=1):
        self.check(lib().mpm_lpc_init(self.ptr, seed))

    def lpc_params(self):
        n = lib().mpm_lpc_num_params()
        out = np.zeros(n, dtype=np.float32)
        self.check(lib().mpm_lpc_get_params(self.ptr, out.ctypes.data_as(_fp), n))
        return out

    def lpc_set_params(self, params):
        a, p = _f(params)
        self.check(lib().mpm_lpc_set_params(self.ptr, p, len(a)))

    def lpc_mode(self, mode):
        self.check(lib().mpm_lpc_set_mode(self.ptr, int(mode)))

    def lpc_fallback(self, ratio):
        self.check(lib().mpm_lpc_set_fallback(self.ptr, float(ratio)))

    def lpc_directions(self, m):
        self.check(lib().mpm_lpc_set_directions(self.ptr, int(m)))

    def lpc_handoff(self, on):
        self.check(lib().mpm_lpc_set_handoff(self.ptr, int(on)))

    def lpc_capture_iterations(self, iters):
        a = np.ascontiguousarray(iters, dtype=np.int32)
        self.check(lib().mpm_lpc_set_capture_iterations(self.ptr, a.ctypes.data_as(_ip), len(a)))

    def small_bodies(self, max_particles):
        """Bodies of at most 8 node blocks and this many particles are solved in one workgroup
        each, concurrently with the batched loop (0: off). Same iterates either way."""
        self.check(lib().mpm_set_small_bodies(self.ptr, int(max_particles)))

    def lpc_last_pairs(self):
        return lib().mpm_lpc_last_pairs(self.ptr)

    def lpc_strikes(self, k):
        self.check(lib().mpm_lpc_set_strikes(self.ptr, int(k)))

    def lpc_training_loss(self, mode):
        """0: optimal-step residual loss; 1: exact-correction energy-norm loss."""
        self.check(lib().mpm_lpc_set_training_loss(self.ptr, int(mode)))

    def lpc_train_next(self, samples, epochs, lr):
        """The next advance trains the network on `samples` residuals of its solve."""
        self.check(lib().mpm_lpc_train_next(self.ptr, int(samples), int(epochs), float(lr)))

    def lpc_last_loss(self):
        """(first epoch, last epoch, diagonal step alone) mean residual loss of the last training."""
        a, b, d = C.c_double(), C.c_double(), C.c_double()
        self.check(lib().mpm_lpc_last_loss(self.ptr, C.byref(a), C.byref(b), C.byref(d)))
        return a.value, b.value, d.value


class Grid:
    """Grid of one body (reference include/MpmSim/Grid.h); construction splats mass and velocity."""

    def __init__(self, ctx, inds, h, frame_v=(0.0, 0.0, 0.0), dim=3):
        self.ctx = ctx
        ii, pi = _i(inds)
        fv, pf = _f(frame_v)
        g = C.c_void_p()
        ctx.check(lib().mpm_grid_create(ctx.ptr, pi, ii.size, h, pf, dim, C.byref(g)))
        self.ptr = g.value
        self.num_particles = int(ii.size)
        n = np.zeros(3, dtype=np.int32)
        mn = np.zeros(3, dtype=np.float32)
        lib().mpm_grid_dims(self.ptr, n.ctypes.data_as(_ip), mn.ctypes.data_as(_fp))
        self.n = n
        self.min = mn
        self.h = np.float32(h)
        self.num_nodes = int(np.prod(n.astype(np.int64)))

    def close(self):
        if getattr(self, "ptr", None) and _lib is not None:
            _lib.mpm_grid_destroy(self.ptr)
            self.ptr = None

    __del__ = close

    def get(self, name):
        """Grid variables in the reference's dense node order; "base_cells" gives each particle's
        stencil base node (the binning kernel's, Grid.cpp:1105-1136) as an (n, 3) int array in the
        order of the particle list the grid was made from."""
        if name == "base_cells":
            out = np.zeros((self.num_particles, 3), dtype=np.int32)
            self.ctx.check(lib().mpm_grid_get(self.ptr, name.encode(), out.ctypes.data))
            return out
        if name == "node_collided":
            out = np.zeros(self.num_nodes, dtype=np.int8)
        else:
            out = np.zeros(self.num_nodes * (3 if "velocities" in name else 1), dtype=np.float32)
        self.ctx.check(lib().mpm_grid_get(self.ptr, name.encode(), out.ctypes.data))
        return out

    def set(self, name, arr):
        a = np.ascontiguousarray(arr, dtype=np.int8 if name == "node_collided" else np.float32)
        self.ctx.check(lib().mpm_grid_set(self.ptr, name.encode(), a.ctypes.data))

    def compute_particle_volumes(self):
        self.ctx.check(lib().mpm_grid_compute_particle_volumes(self.ptr))

    def calculate_forces(self):
        out = np.zeros(3 * self.num_nodes, dtype=np.float32)
        self.ctx.check(lib().mpm_grid_calculate_forces(self.ptr, out.ctypes.data_as(_fp)))
        return out

    def force_differentials(self, dx):
        d, pd = _f(dx)
        out = np.zeros(3 * self.num_nodes, dtype=np.float32)
        self.ctx.check(lib().mpm_grid_force_differentials(self.ptr, pd, out.ctypes.data_as(_fp)))
        return out

    def dforce_dxi(self):
        out = np.zeros(3 * self.num_nodes, dtype=np.float32)
        self.ctx.check(lib().mpm_grid_dforce_dxi(self.ptr, out.ctypes.data_as(_fp)))
        return out

    def explicit_momenta(self, dt):
        em = np.zeros(3 * self.num_nodes, dtype=np.float32)
        col = np.zeros(self.num_nodes, dtype=np.int8)
        self.ctx.check(lib().mpm_grid_explicit_momenta(self.ptr, dt, em.ctypes.data_as(_fp), col.ctypes.data_as(_bp)))
        return em, col

    def collision_velocities(self, collided):
        c = np.ascontiguousarray(collided, dtype=np.int8)
        out = np.zeros(3 * self.num_nodes, dtype=np.float32)
        self.ctx.check(lib().mpm_grid_collision_velocities(self.ptr, c.ctypes.data_as(_bp), out.ctypes.data_as(_fp)))
        return out

    def update_grid_velocities(self, dt, max_iters, tol, termination=None, debug=None, hist_cap=4096):
        prm = make_params(max_iters, tol, termination, debug)
        hist = np.zeros(hist_cap)
        st = SolveStats()
        st.history_capacity = hist_cap
        st.history = hist.ctypes.data_as(_dp)
        self.ctx.check(lib().mpm_grid_update_grid_velocities(self.ptr, dt, C.byref(prm), C.byref(st)))
        return _stats_dict(st, hist)

    def update_particle_velocities(self):
        self.ctx.check(lib().mpm_grid_update_particle_velocities(self.ptr))

    def update_deformation_gradients(self, dt):
        self.ctx.check(lib().mpm_grid_update_deformation_gradients(self.ptr, dt))

    def lpc_apply(self, r, diag, scale):
        """Learned-preconditioner network output (dense (n, 3)) for residual r and diagonal diag."""
        ra, rp = _f(r)
        da, dp = _f(diag)
        out = np.zeros(ra.size, dtype=np.float32)
        self.ctx.check(lib().mpm_grid_lpc_apply(self.ptr, rp, dp, float(scale), out.ctypes.data_as(_fp)))
        return out.reshape(-1, 3)

    def lpc_gradients(self, r, diag, e, scale, weight):
        """Loss weight |out - e / scale|^2 and its gradient with respect to the network parameters."""
        ra, rp = _f(r)
        da, dp = _f(diag)
        ea, ep = _f(e)
        g = np.zeros(lib().mpm_lpc_num_params(), dtype=np.float32)
        loss = C.c_double()
        self.ctx.check(lib().mpm_grid_lpc_gradients(self.ptr, rp, dp, ep, float(scale), float(weight),
                                                    g.ctypes.data_as(_fp), C.byref(loss)))
        return loss.value, g


class Sim:
    """Reference Sim (include/MpmSim/Sim.h) on the GPU: particles, snow model, colliders, gravity."""

    def __init__(self, x, m, h, material=None, planes=(), spheres=(), gravity=(), device=0, mu0=None, lambda0=None,
                 xi=None, v=None, volumes=()):
        self.ctx = Context(device)
        x = np.asarray(x, dtype=np.float32).reshape(-1, 3)
        self.ctx.particles_create(len(x))
        self.ctx.set("p", x)
        self.ctx.set("m", np.asarray(m, dtype=np.float32))
        if v is not None:
            self.ctx.set("v", v)
        if material is not None:
            mat = dict(material)
            self.ctx.model_snow(mat["E"], mat["nu"], mat["xi"], mat["theta_c"], mat["theta_s"])
            if mu0 is not None:
                self.ctx.set("mu0", mu0)
                self.ctx.set("lambda0", lambda0)
                self.ctx.set("xi", xi)
        self.ctx.set_colliders(planes, spheres, volumes)
        if len(gravity):
            self.ctx.set_gravity(gravity)
        self.ctx.check(lib().mpm_sim_init(self.ctx.ptr, h, 3))
        self.h = h

    def advance(self, dt, max_iters, tol, cancelled=None, termination=None, debug=None):
        prm = make_params(max_iters, tol, termination, debug, cancelled)
        st = StepStats()
        self.ctx.check(lib().mpm_sim_advance(self.ctx.ptr, dt, C.byref(prm), C.byref(st)))
        return {"num_bodies": st.num_bodies, "num_ballistic": st.num_ballistic,
                "total_iterations": st.total_iterations, "max_iterations": st.max_iterations,
                "cancelled": bool(st.cancelled), "num_solved": st.num_solved}

    def grid_layout(self):
        """Layout of the last step's batched grids (measurement)."""
        a = (C.c_int64 * 7)()
        self.ctx.check(lib().mpm_sim_grid_layout(self.ctx.ptr, a))
        keys = ("bodies", "tiles", "chunks", "node_slots", "particles", "loop_tiles", "loop_chunks")
        return dict(zip(keys, [int(v) for v in a]))

    def bodies(self):
        out = []
        for b in range(lib().mpm_sim_num_bodies(self.ctx.ptr)):
            n = lib().mpm_sim_body(self.ctx.ptr, b, None)
            a = np.zeros(n, dtype=np.int32)
            lib().mpm_sim_body(self.ctx.ptr, b, a.ctypes.data_as(_ip))
            out.append(a)
        return out

    def ballistic(self):
        n = lib().mpm_sim_ballistic(self.ctx.ptr, None)
        a = np.zeros(n, dtype=np.int32)
        lib().mpm_sim_ballistic(self.ctx.ptr, a.ctypes.data_as(_ip))
        return a

    def body_stats(self, b, hist_cap=4096):
        """Solve statistics of body b in the last advance; hist_cap = 0 skips the residual history."""
        hist = np.zeros(hist_cap) if hist_cap > 0 else None
        st = SolveStats()
        st.history_capacity = hist_cap
        st.history = hist.ctypes.data_as(_dp) if hist is not None else None
        rc = lib().mpm_sim_body_solve_stats(self.ctx.ptr, b, C.byref(st))
        if rc != MPM_OK:
            raise MpmError("no solve statistics for body %d" % b)
        return _stats_dict(st, hist)

    def get(self, name):
        return self.ctx.get(name)

    def profile(self, on=True, period=1):
        """Time kernel classes with HIP events: True = all, False = off, or an iterable of names;
        one launch in `period` of each class is timed."""
        if on is True:
            mask = -1
        elif not on:
            mask = 0
        else:
            mask = 0
            for k in on:
                mask |= 1 << PROFILE_KERNELS.index(k)
        self.ctx.check(lib().mpm_profile_enable(self.ctx.ptr, mask))
        self.ctx.check(lib().mpm_profile_period(self.ctx.ptr, int(period)))

    def profile_query(self):
        out = {}
        for k in PROFILE_KERNELS:
            ms = C.c_double()
            n = C.c_int64()
            by = C.c_double()
            self.ctx.check(lib().mpm_profile_query(self.ctx.ptr, k.encode(), C.byref(ms), C.byref(n), C.byref(by)))
            out[k] = {"ms": ms.value, "launches": n.value, "bytes": by.value}
        return out


class Sim64:
    """The full-FP64 step on the GPU: Sim (include/MpmSim/Sim.h) with every scalar in double, the FP64
    run that precision studies compare the FP32-bulk Sim against. Same constructor arguments as Sim
    (colliders: planes and spheres); arrays in and out are float64."""

    def __init__(self, x, m, h, material, planes=(), spheres=(), gravity=(), device=0, mu0=None, lambda0=None,
                 xi=None, v=None):
        x = np.asarray(x, dtype=np.float64).reshape(-1, 3)
        self.n = len(x)
        ptr = C.c_void_p()
        rc = lib().mpm_sim64_create(C.byref(ptr), device, self.n)
        if rc != MPM_OK:
            raise MpmError("mpm_sim64_create failed (code %d)" % rc)
        self.ptr = ptr.value
        self.set("p", x)
        self.set("m", m)
        if v is not None:
            self.set("v", v)
        mat = dict(material)
        self.check(lib().mpm_sim64_model_snow(self.ptr, float(mat["E"]), float(mat["nu"]), float(mat["xi"]),
                                              float(mat["theta_c"]), float(mat["theta_s"])))
        if mu0 is not None:
            self.set("mu0", mu0)
            self.set("lambda0", lambda0)
            self.set("xi", xi)
        items = []
        for coeffs, vel, mu, sticky in planes:
            c = Collider64()
            c.type = 0
            c.coeffs[:] = list(map(float, coeffs))
            c.velocity[:] = list(map(float, vel))
            c.friction = float(mu)
            c.sticky = int(sticky)
            items.append(c)
        for center, r, vel, mu, sticky in spheres:
            c = Collider64()
            c.type = 1
            c.center[:] = list(map(float, center))
            c.radius = float(r)
            c.velocity[:] = list(map(float, vel))
            c.friction = float(mu)
            c.sticky = int(sticky)
            items.append(c)
        arr = (Collider64 * max(1, len(items)))(*items)
        self.check(lib().mpm_sim64_set_colliders(self.ptr, arr, len(items)))
        if len(gravity):
            g = np.ascontiguousarray(np.asarray(gravity, dtype=np.float64).reshape(-1, 3))
            self.check(lib().mpm_sim64_set_gravity(self.ptr, g.ctypes.data_as(_dp), len(g)))
        self.check(lib().mpm_sim64_init(self.ptr, float(h)))
        self.h = h

    def close(self):
        if getattr(self, "ptr", None) and _lib is not None:
            _lib.mpm_sim64_destroy(self.ptr)
            self.ptr = None

    __del__ = close

    def check(self, rc):
        if rc != MPM_OK:
            raise MpmError(lib().mpm_sim64_last_error(self.ptr).decode())

    def set(self, name, arr):
        if name in MAT3:
            flat = np.ascontiguousarray(np.asarray(arr, dtype=np.float64).reshape(-1, 3, 3).transpose(0, 2, 1))
        else:
            flat = np.ascontiguousarray(np.asarray(arr, dtype=np.float64))
        flat = flat.reshape(-1)
        self.check(lib().mpm_sim64_set(self.ptr, name.encode(), flat.ctypes.data_as(_dp), flat.size))

    def get(self, name):
        w = 9 if name in MAT3 else 3 if name in VEC3 else 1
        out = np.zeros(self.n * w, dtype=np.float64)
        self.check(lib().mpm_sim64_get(self.ptr, name.encode(), out.ctypes.data_as(_dp), out.size))
        if name in MAT3:
            return out.reshape(-1, 3, 3).transpose(0, 2, 1).copy()
        if name in VEC3:
            return out.reshape(self.n, 3)
        return out

    def advance(self, dt, max_iters, tol):
        st = StepStats()
        self.check(lib().mpm_sim64_advance(self.ptr, float(dt), int(max_iters), float(tol), C.byref(st)))
        return {"num_bodies": st.num_bodies, "num_ballistic": st.num_ballistic,
                "total_iterations": st.total_iterations, "max_iterations": st.max_iterations,
                "num_solved": st.num_solved}

    def num_bodies(self):
        return lib().mpm_sim64_num_bodies(self.ptr)

    def body_stats(self, b, hist_cap=4096):
        hist = np.zeros(hist_cap) if hist_cap > 0 else None
        st = SolveStats()
        st.history_capacity = hist_cap
        st.history = hist.ctypes.data_as(_dp) if hist is not None else None
        rc = lib().mpm_sim64_body_solve_stats(self.ptr, b, C.byref(st))
        if rc != MPM_OK:
            raise MpmError("no solve statistics for body %d" % b)
        return _stats_dict(st, hist)
