"""The reference's own grid tests with their own assertions, run on the GPU library through the C-ABI:
TestGrid::testDeformationGradients (TestGrid.cpp:216-278), TestGrid::testForces (TestGrid.cpp:281-396),
TestSnowConstitutiveModel (TestSnowConstitutiveModel.cpp:17-158) and the invariants of
TestGrid::testImplicitUpdate (TestGrid.cpp:456-708): collided nodes move only tangentially to their
collider, multi-contact nodes are at rest, and the solved velocities satisfy the projected system
P(M(v - vc) - dt^2 DF v) = p*, within the reference test's own bound (1e-6, absolute max-norm,
TestGrid.cpp:698-707), computed in FP64 from the GPU's FP32 outputs."""
import numpy as np
import pytest

import mpmsnow_amd as G
from test_gpu_parity import implicit_pair, relerr, say

pytestmark = pytest.mark.gpu


def test_deformation_gradients_linear_field():
    """Linear velocity field (x, -y, 0.5 z), h = 0.2, dt = 0.01, 7 steps: F of the particle at the
    origin stays diag(1 + 0.01 i, 1 - 0.01 i, 1 + 0.005 i) within 1e-3."""
    say("reference testDeformationGradients: 1001 particles, 7 steps")
    pts = [[0, 0, 0]] + [[i / 9 - 0.5, j / 9 - 0.5, k / 9 - 0.5] for i in range(10) for j in range(10) for k in range(10)]
    x = np.array(pts, dtype=np.float32)
    n = len(x)
    ctx = G.Context()
    ctx.particles_create(n)
    ctx.set("p", x)
    ctx.set("m", np.ones(n, np.float32))
    ctx.set("volume", np.ones(n, np.float32))
    v = np.stack([x[:, 0], -x[:, 1], 0.5 * x[:, 2]], 1).astype(np.float32)
    v[0] = 0
    ctx.set("v", v)
    for i in range(1, 8):
        g = G.Grid(ctx, np.arange(n), 0.2)
        g.update_deformation_gradients(0.01)
        g.close()
        x = (x + v * np.float32(0.01)).astype(np.float32)
        ctx.set("p", x)
        F0 = ctx.get("F")[0]
        assert abs(F0[0, 0] - (1 + 0.01 * i)) < 1e-3
        assert abs(F0[1, 1] - (1 - 0.01 * i)) < 1e-3
        assert abs(F0[2, 2] - (1 + 0.005 * i)) < 1e-3


def test_implicit_update_solves_projected_system(oracle):
    say("reference testImplicitUpdate invariants: 64 particles, two moving planes, maxIters=40 tol=0")
    P, h, cols = implicit_pair(oracle)
    dt = 0.002
    g = G.Grid(P.ctx, np.arange(P.n), h)
    g.compute_particle_volumes()
    em, col = g.explicit_momenta(dt)
    g.update_grid_velocities(dt, 40, 0.0)
    v = g.get("velocities").astype(np.float64).reshape(-1, 3)
    col = np.asarray(col)
    vc = g.collision_velocities(col).astype(np.float64).reshape(-1, 3)
    normals = [np.array([-1, -1, -1.0]) / np.sqrt(3), np.array([-1.0, 0, 0])]
    objv = [np.array([-0.1, -0.2, -0.3]), np.array([-0.2, -0.1, -0.4])]
    assert (col >= 0).any() and (col == -1).any()
    for idx in np.nonzero(col >= 0)[0]:
        o = col[idx]
        assert np.abs(vc[idx] - objv[o]).max() < 1e-7
        assert abs(normals[o] @ (v[idx] - objv[o])) < 1e-5
    for idx in np.nonzero(col == -2)[0]:
        assert np.linalg.norm(v[idx]) == 0
    df = g.force_differentials(g.get("velocities")).astype(np.float64).reshape(-1, 3)
    M = g.get("masses").astype(np.float64)
    lhs = M[:, None] * (v - vc) - df * dt * dt
    for idx in range(len(M)):
        o = col[idx]
        if o >= 0:
            lhs[idx] -= (normals[o] @ lhs[idx]) * normals[o]
        elif o == -2:
            lhs[idx] = 0
    err = np.linalg.norm(lhs - em.astype(np.float64).reshape(-1, 3), axis=1).max()
    say("max-norm of P(M(v - vc) - dt^2 DF v) - p*: %.3e (reference test bound 1e-6)" % err)
    assert err < 1e-6


def test_forces_and_hvp_against_finite_differences():
    """TestGrid::testForces (TestGrid.cpp:281-396) on the GPU: one particle (volume 0.5, E = 1.4e3,
    h = 1): zero force at a pure rotation (1e-4); then F = rotation + 0.01 random: each force
    component against -dE/dx by central differences of the energy when the grid moves by +-0.01
    along it (1%), and the HVP against the central difference of the forces (5%: the HVP keeps the
    force's F^T factor fixed while the difference recomputes it, as the reference's own comment on
    its 2.5% bound says; the oracle pins the HVP's assembly exactly)."""
    from test_oracle_pins import ROT, rand_mat

    say("reference testForces: single particle, FD of energy and forces")
    rng = np.random.default_rng(23)
    ctx = G.Context()
    ctx.particles_create(1)
    ctx.set("p", np.zeros((1, 3), np.float32))
    ctx.set("m", np.ones(1, np.float32))
    ctx.set("volume", np.full(1, 0.5, np.float32))
    ctx.set("F", ROT.reshape(1, 3, 3).astype(np.float32))
    ctx.model_snow(1.4e3, 0.2, 0.0, 1e5, 1e5)
    ctx.update_particle_data()
    g = G.Grid(ctx, np.arange(1), 1.0)
    f = g.calculate_forces()
    assert np.abs(f).max() < 1e-4
    Forig = (ROT + 0.01 * rand_mat(rng)).reshape(1, 3, 3).astype(np.float32)

    def reset():
        ctx.set("F", Forig)
        ctx.update_particle_data()

    reset()
    f = g.calculate_forces()
    dx = 0.01
    nv = g.num_nodes * 3
    worst_f = worst_df = 0.0
    for i in range(nv):
        if f[i] == 0:
            continue
        reset()
        vel = np.zeros(nv, np.float32)
        vel[i] = dx
        df = g.force_differentials(vel).astype(np.float64)
        ends = []
        for sgn in (1.0, -1.0):
            reset()
            vel[i] = sgn * dx
            g.set("velocities", vel)
            g.update_deformation_gradients(1.0)
            ctx.update_particle_data()
            ends.append((float(ctx.energy_density(np.arange(1))[0]) * 0.5, g.calculate_forces().astype(np.float64)))
        reset()
        (ep, fp), (em, fm) = ends
        worst_f = max(worst_f, abs((f[i] + (ep - em) / (2 * dx)) / f[i]))
        worst_df = max(worst_df, np.linalg.norm(0.5 * (fp - fm) - df) / np.linalg.norm(df))
    say("worst relative force error %.3e, worst HVP error %.3e" % (worst_f, worst_df))
    assert worst_f < 0.01
    assert worst_df < 0.05


def gpu_model(F, Fp=None, E=1.0, nu=0.25, xi=10.0, thc=0.1, ths=0.2):
    """One particle with the reference test's snow model (TestSnowConstitutiveModel.cpp:30-36)."""
    ctx = G.Context()
    ctx.particles_create(1)
    ctx.set("m", np.ones(1, np.float32))
    ctx.model_snow(E, nu, xi, thc, ths)
    ctx.set("F", np.asarray(F, np.float32).reshape(1, 3, 3))
    if Fp is not None:
        ctx.set("Fp", np.asarray(Fp, np.float32).reshape(1, 3, 3))
    ctx.update_particle_data()
    return ctx


def test_snow_constitutive_model_reference_assertions():
    """TestSnowConstitutiveModel (TestSnowConstitutiveModel.cpp:17-158) on the GPU, with its own
    bounds: updateParticleData keeps F Fp and gives F^-T and J = det F (1e-4); the clamp known
    answer sigma 1.3 -> 1.2, 0.8 -> 0.9; zero energy and stress at the identity and at a pure
    rotation (1e-6); dEnergyDensitydF against central differences of the energy and
    dEdFDifferential against the change of dEnergyDensitydF (1e-5)."""
    from test_oracle_pins import ROT, rand_mat

    say("reference TestSnowConstitutiveModel on the GPU")
    rng = np.random.default_rng(11)
    P = rand_mat(rng)
    P = P @ P
    Mm = rand_mat(rng)
    Mm = Mm @ Mm
    ctx = gpu_model(Mm, P)
    F = ctx.get("F")[0].astype(np.float64)
    Fp = ctx.get("Fp")[0].astype(np.float64)
    assert np.linalg.norm(F @ Fp - Mm @ P) < 1e-4 * max(1, np.linalg.norm(Mm @ P))
    assert np.linalg.norm(np.linalg.inv(F).T - ctx.get("FinvTrans")[0]) < 1e-3 * max(1, np.linalg.norm(np.linalg.inv(F)))
    assert abs(ctx.get("J")[0] - abs(np.linalg.det(F))) < 1e-3
    ctx = gpu_model(np.diag([1.3, 0.8, 1.0]), np.eye(3))
    F = ctx.get("F")[0]
    assert abs(F[0, 0] - 1.2) < 1e-6 and abs(F[1, 1] - 0.9) < 1e-6
    assert np.abs(F.astype(np.float64) @ ctx.get("Fp")[0] - np.diag([1.3, 0.8, 1.0])).max() < 1e-6
    for R in (np.eye(3), ROT):
        ctx = gpu_model(R, np.eye(3))
        assert abs(ctx.energy_density(np.arange(1))[0]) < 1e-6
        assert np.linalg.norm(ctx.stress(np.arange(1))[0]) < 1e-6
    # stress against central differences of the energy
    F0 = np.diag([1.02, 0.95, 1.0]) @ ROT
    ctx = gpu_model(F0, np.eye(3))
    Pst = ctx.stress(np.arange(1))[0].astype(np.float64)
    delta = 0.01
    fd = np.zeros((3, 3))
    for i in range(3):
        for j in range(3):
            e = []
            for sgn in (1.0, -1.0):
                Fq = F0.copy()
                Fq[i, j] += sgn * delta
                ctx.set("F", Fq.reshape(1, 3, 3).astype(np.float32))
                ctx.update_particle_data()
                e.append(float(ctx.energy_density(np.arange(1))[0]))
            fd[i, j] = (e[0] - e[1]) / (2 * delta)
    say("stress vs FD of the energy: %.3e" % np.abs(Pst - fd).max())
    assert np.abs(Pst - fd).max() < 1e-5
    # differential against the change of the stress
    rng = np.random.default_rng(21)
    ctx = gpu_model(F0, np.eye(3))
    P0 = ctx.stress(np.arange(1))[0].astype(np.float64)
    dF = 0.001 * rand_mat(rng)
    dP = ctx.stress_differential(np.arange(1), dF.reshape(1, 3, 3))[0].astype(np.float64)
    ctx.set("F", (F0 + dF).reshape(1, 3, 3).astype(np.float32))
    ctx.update_particle_data()
    fd = ctx.stress(np.arange(1))[0].astype(np.float64) - P0
    say("stress differential vs change of the stress: %.3e" % np.abs(fd - dP).max())
    assert np.abs(fd - dP).max() < 1e-5


def test_sim_cancellation_against_oracle(oracle):
    """Sim::advance cancelled after the second body's solve (Sim.cpp:107-111), on the GPU against
    the FP64 oracle with the same interrupt: cancelled() is called once per body until it returns
    true; no particle moves; body 0 gets its G2P, F update and plasticity; bodies 1 and 2 keep their
    velocities and F bit for bit; the ballistic particles keep their gravity update. The next,
    uncancelled step then matches the oracle (positions within 1e-4 h, F within 1e-5)."""
    from test_oracle_pins import cancel_oracle_sim, cancel_scene

    say("Sim cancellation on the second poll: 3 snowballs + 5 ballistic particles")
    x, m, v, h = cancel_scene()
    planes = [((0.0, 1.0, 0.0, 0.0), (0.0, 0.0, 0.0), 0.2, False)]
    mat = dict(E=1.4e5, nu=0.2, xi=10.0, theta_c=2.5e-2, theta_s=7.5e-3)
    sim = G.Sim(x, m, h, material=mat, planes=planes, gravity=[(0.0, -9.81, 0.0)], v=v)
    o = cancel_oracle_sim(oracle, "double", x, m, v, h)
    F0 = sim.get("F").copy()
    calls = []

    def cancelled():
        calls.append(1)
        return len(calls) >= 2

    st = sim.advance(1e-3, 1000, 1e-5, cancelled=cancelled)
    assert o.advance(1e-3, 1000, 1e-5, cancel_on_call=2)
    assert st["cancelled"] and len(calls) == 2
    assert np.array_equal(sim.get("p"), x)
    vg, vo = sim.get("v"), o.pd.get("v")
    assert relerr(vg[:2500], vo[:2500]) < 1e-5
    assert np.array_equal(vg[2500:7500], v[2500:7500])
    assert np.abs(vg[7500:] - vo[7500:]).max() < 1e-7
    Fg = sim.get("F")
    assert np.abs(Fg[:2500] - o.pd.get("F")[:2500]).max() < 1e-5
    assert np.array_equal(Fg[2500:], F0[2500:])
    sim.advance(1e-3, 1000, 1e-5)
    o.advance(1e-3, 1000, 1e-5)
    dx = np.abs(sim.get("p") - o.pd.get("p")).max()
    say("after the next step: max |x - x_oracle| = %.3e h" % (dx / h))
    assert dx < 1e-4 * h
    assert np.abs(sim.get("F") - o.pd.get("F")).max() < 1e-5


def lattice_ctx(oracle, v):
    """6^3 particles at spacing 0.05 (TestGrid-style lattice) with F = I (stress-free), on the GPU and
    in the FP64 oracle."""
    pos = np.array([[0.05 * i, 0.05 * j - 0.02, 0.05 * k] for i in range(6) for j in range(6) for k in range(6)],
                   dtype=np.float32)
    n = len(pos)
    ctx = G.Context()
    ctx.particles_create(n)
    ctx.set("p", pos)
    ctx.set("v", np.tile(np.asarray(v, np.float32), (n, 1)))
    ctx.set("m", np.ones(n, np.float32))
    ctx.set("volume", np.full(n, 1e-4, np.float32))
    ctx.model_snow(1.4e5, 0.2, 10.0, 2.5e-2, 7.5e-3)
    ctx.set_gravity([(0.0, -9.81, 0.0)])
    pd = oracle.ParticleData.create(n, "double")
    pd.set("p", pos)
    pd.set("v", np.tile(np.asarray(v, np.float64), (n, 1)))
    pd.set("m", np.ones(n))
    pd.set("volume", np.full(n, 1e-4, np.float32))
    model = oracle.Model.snow(1.4e5, 0.2, 10.0, 2.5e-2, 7.5e-3, "double")
    model.set_particles(pd)
    fields = oracle.Fields.create("double")
    fields.add_gravity([0, -9.81, 0])
    return ctx, pd, model, fields, n


def test_explicit_momenta_sticky_plane_sphere_and_multi_contact(oracle):
    """calculateExplicitMomenta and collisionVelocities (Grid.cpp:634-724) with the ground plane
    (friction 0.2), a moving sticky wall and a moving analytic sphere (friction 0.5), against the FP64
    oracle: collision flags exact (including multi-contact -2 and sticky nodes), explicit momenta
    within 1e-5, collision velocities within 1e-7 (src/CollisionObject.cpp:19-96)."""
    say("explicit momenta: ground plane, sticky moving wall, moving sphere")
    ctx, pd, model, fields, n = lattice_ctx(oracle, [0.5, -1.0, 0.3])
    planes = [((0.0, 1.0, 0.0, 0.0), (0.0, 0.0, 0.0), 0.2, False),
              ((-1.0, 0.0, 0.0, 0.2), (0.1, 0.0, 0.0), 0.0, True)]
    spheres = [((0.05, 0.1, 0.1), 0.06, (0.0, 0.2, 0.0), 0.5, False)]
    ctx.set_colliders(planes, spheres)
    cols = oracle.Colliders.create("double")
    for c, vv, mu, st in planes:
        cols.add_plane(c, v=vv, friction=mu, sticky=st)
    for c, r, vv, mu, st in spheres:
        cols.add_sphere(c, r, v=vv, friction=mu, sticky=st)
    h, dt = 0.1, 1e-3
    g = G.Grid(ctx, np.arange(n), h)
    og = oracle.Grid(pd, np.arange(n), h)
    em, col = g.explicit_momenta(dt)
    emo, colo = og.explicit_momenta(dt, model, cols, fields)
    col, colo = np.asarray(col), np.asarray(colo)
    say("flags: %s" % {int(k): int((colo == k).sum()) for k in np.unique(colo)})
    for k in (-2, 0, 1, 2):
        assert (colo == k).any(), k
    assert (col == colo).all()
    assert relerr(em, emo) < 1e-5
    vc = g.collision_velocities(col)
    vco = og.collision_velocities(cols, colo)
    assert np.abs(vc - vco).max() < 1e-7


def test_explicit_momenta_nan_error_from_frictionless_contact(oracle):
    """The reference's zero-friction quirk (src/CollisionObject.cpp:70-84): a node pressed straight
    into a frictionless plane has zero tangential velocity, 0 * (x / 0) is NaN, and the explicit
    momentum check raises "nan in explicit momenta!" (Grid.cpp:677-687); the GPU raises the same."""
    say("explicit momenta: frictionless plane, normal approach -> nan error")
    ctx, pd, model, fields, n = lattice_ctx(oracle, [0.0, -1.0, 0.0])
    ctx.set_colliders([((0.0, 1.0, 0.0, 0.0), (0.0, 0.0, 0.0), 0.0, False)])
    g = G.Grid(ctx, np.arange(n), 0.1)
    with pytest.raises(G.MpmError, match="nan in explicit momenta!"):
        g.explicit_momenta(1e-3)
