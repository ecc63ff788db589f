"""Closed-form and invariant pins for the oracle's contact physics and integration.

Each check is fixed by mathematics rather than by re-typing the oracle's expressions:
static Hertz overlap force, the Coulomb cap and its stored-spring sign flip, the EBPM axial
increment k1 * u * dt, bond breakage at zero strength, free flight under gravity (exact for the
update rule), momentum conservation, the quaternion recursion without dt (a composition of
rotations by 2*atan(w/2)), and the coefficient of restitution of a head-on and a wall impact.
"""
import math

import numpy as np
import pytest

from helpers import packed_sample
from oracle import oracle as O
from taichi_dem_julia_port_amd import DEMConfig
from taichi_dem_julia_port_amd.types import grains_in

RHO = 2500.0


def two_spheres(r1, r2, k1, k2, v1=(0, 0, 0), v2=(0, 0, 0), **cfg):
    V = [4 / 3 * math.pi * r1 ** 3, 4 / 3 * math.pi * r2 ** 3]
    g = grains_in([1, 2], V, [RHO * V[0], RHO * V[1]], np.array([k1, k2], dtype=float),
                  np.array([v1, v2], dtype=float))
    base = dict(domain_min=[-1.0, -1.0, -1.0], domain_max=[1.0, 1.0, 1.0],
                gravity=[0.0, 0.0, 0.0], wall_normal=[0.0, 0.0, -1.0], wall_distance=0.9,
                particle_elastic_modulus=1e8, particle_poisson_ratio=0.3)
    base.update(cfg)
    c = DEMConfig(**base)
    return c, g


def hertz_kn(E, nu, r1, r2, delta):
    Ys = 1.0 / ((1 - nu ** 2) / E + (1 - nu ** 2) / E)
    Rs = 1.0 / (1 / r1 + 1 / r2)
    return 4.0 / 3.0 * Ys * math.sqrt(Rs * delta)


def hertz_kt(E, nu, r1, r2, delta):
    Gs = 1.0 / (2 * (2 - nu) * (1 + nu) / E * 2)
    Rs = 1.0 / (1 / r1 + 1 / r2)
    return 8.0 * Gs * math.sqrt(Rs * delta)


def test_static_overlap_hertz_force():
    r, delta = 0.01, 1e-4
    cfg, g = two_spheres(r, r, (0, 0, 0), (2 * r - delta, 0, 0), dt=1e-7)
    o = O.OracleScene(cfg, g)
    o.step(1)
    c, b = o.read_contacts()
    assert len(c) == 1 and not b[0]
    F1 = hertz_kn(1e8, 0.3, r, r, delta) * delta  # Hertz: 4/3 Y* sqrt(R*) delta^(3/2)
    assert c[0]["F"][0] == pytest.approx(F1, rel=1e-13)
    assert abs(c[0]["F"][1]) < 1e-20 and abs(c[0]["F"][2]) < 1e-20
    F, T = o.read_forces()
    assert F[0] == pytest.approx([-F1, 0, 0], rel=1e-13, abs=1e-20)
    assert F[1] == pytest.approx([F1, 0, 0], rel=1e-13, abs=1e-20)
    assert np.abs(T).max() < 1e-20
    # the contact point sits on the first sphere's surface moved inward by the overlap
    assert c[0]["k"] == pytest.approx([r - delta, 0, 0], rel=1e-13, abs=1e-18)


def test_coulomb_cap_and_spring_sign_flip():
    r, delta, mu, dt = 0.01, 1e-4, 0.5, 1e-3
    cfg, g = two_spheres(r, r, (0, 0, 0), (2 * r - delta, 0, 0), v2=(0, 1.0, 0), dt=dt,
                         pp_coefficient_friction=mu)
    o = O.OracleScene(cfg, g)
    o.step(1)
    c, _ = o.read_contacts()
    F1 = hertz_kn(1e8, 0.3, r, r, delta) * delta
    kt = hertz_kt(1e8, 0.3, r, r, delta)
    assert kt * dt > mu * F1  # the trial spring force exceeds the cap: sliding branch
    # |F_t| = mu F1 along -s, and the stored spring becomes F_t / k_t (sign flipped)
    assert c[0]["F"][0] == pytest.approx(F1, rel=1e-12)
    assert c[0]["F"][1] == pytest.approx(-mu * F1, rel=1e-12)
    assert c[0]["s"][1] == pytest.approx(-mu * F1 / kt, rel=1e-12)
    assert c[0]["s"][1] < 0 < dt * 1.0


def test_no_slip_tangential_spring():
    r, delta, mu, dt = 0.01, 1e-4, 0.5, 1e-9
    cfg, g = two_spheres(r, r, (0, 0, 0), (2 * r - delta, 0, 0), v2=(0, 1.0, 0), dt=dt,
                         pp_coefficient_friction=mu)
    o = O.OracleScene(cfg, g)
    o.step(1)
    c, _ = o.read_contacts()
    assert c[0]["s"][1] == pytest.approx(1.0 * dt, rel=1e-12)  # s += v_t dt
    kt = hertz_kt(1e8, 0.3, r, r, delta)
    assert c[0]["F"][1] < -kt * dt * 0.999  # spring plus (positive-direction) damping


def test_ebpm_axial_increment():
    r, u, dt = 0.01, 0.3, 1e-6
    Eb, rho_b = 28e9, 0.5
    cfg, g = two_spheres(r, r, (0, 0, 0), (2 * r, 0, 0), v1=(-u, 0, 0), v2=(u, 0, 0), dt=dt,
                         bond_elastic_modulus=Eb, bond_radius_ratio=rho_b)
    o = O.OracleScene(cfg, g)
    o.initialize()
    assert o.counters()["bonded_contacts"] == 1
    o.step(1)
    c, b = o.read_contacts()
    assert b[0]
    k1 = Eb * math.pi * (rho_b * r) ** 2 / (2 * r)
    assert c[0]["F"][0] == pytest.approx(k1 * (-u - u) * dt, rel=1e-12)
    F, _ = o.read_forces()
    assert F[0][0] == pytest.approx(2 * k1 * u * dt, rel=1e-12)  # pulls i towards j
    assert c[0]["k"] == pytest.approx([r, 0, 0], rel=1e-13)  # midpoint


def test_bond_breaks_at_zero_strength():
    cfg, g, w = packed_sample(300, scale=1.0, vmax=0.0)
    cfg = cfg.replace(bond_tensile_strength=0.0)
    o = O.OracleScene(cfg, g, w)
    o.initialize()
    nb = o.counters()["bonded_contacts"]
    assert nb > 0
    o.step(1)
    cnt = o.counters()
    assert cnt["bonds_broken_total"] == nb  # sigma_t = 0 >= 0 for every bond
    assert cnt["bonded_contacts"] == 0


def test_free_flight_closed_form():
    dt, n = 1e-4, 500
    g0, v0, k0 = np.array([0.3, -0.2, -9.81]), np.array([1.0, 2.0, 3.0]), np.array([0.0, 0.1, 0.2])
    V = 4 / 3 * math.pi * 0.01 ** 3
    gr = grains_in([1], [V], [RHO * V], k0[None], v0[None])
    cfg = DEMConfig(domain_min=[-10.0] * 3, domain_max=[10.0] * 3, gravity=list(g0), dt=dt,
                    wall_distance=9.0)
    o = O.OracleScene(cfg, gr)
    o.step(n)
    t = n * dt
    out = o.read_grains()[0]
    assert out["k"] == pytest.approx(k0 + v0 * t + 0.5 * g0 * t * t, rel=1e-11)
    assert out["v"] == pytest.approx(v0 + g0 * t, rel=1e-12)


def test_momentum_conserved_without_walls_and_gravity():
    cfg, g, w = packed_sample(1500, scale=1.2, vmax=0.05)
    cfg = cfg.replace(gravity=[0.0, 0.0, 0.0])
    o = O.OracleScene(cfg, g, w[:0])  # no walls
    p0 = (g["m"][:, None] * g["v"]).sum(0)
    o.step(20)
    gr = o.read_grains()
    assert o.counters()["active_contacts"] > 0
    p1 = (gr["m"][:, None] * gr["v"]).sum(0)
    scale = (g["m"][:, None] * np.abs(g["v"])).sum()
    assert np.abs(p1 - p0).max() <= 1e-12 * scale


def test_quaternion_recursion_without_dt():
    w, n = 0.3, 12
    V = 4 / 3 * math.pi * 0.01 ** 3
    gr = grains_in([1], [V], [RHO * V], np.zeros((1, 3)), np.zeros((1, 3)))
    cfg = DEMConfig(domain_min=[-1.0] * 3, domain_max=[1.0] * 3, gravity=[0.0, 0.0, 0.0],
                    dt=1e-6, wall_distance=0.9)
    o = O.OracleScene(cfg, gr)
    s = o.read_grains()
    s[0]["w"] = [0.0, 0.0, w]
    o.write_grains(s)
    o.step(n)
    q = o.read_grains()[0]["q"]
    # q <- normalize(q + 0.5 (0, w) q): a rotation by theta with tan(theta / 2) = w / 2 per step
    theta = 2 * math.atan(w / 2)
    assert q == pytest.approx([math.cos(n * theta / 2), 0, 0, math.sin(n * theta / 2)], abs=1e-13)


def _relative_normal_speed(o):
    g = o.read_grains()
    d = g["k"][1] - g["k"][0]
    return float(np.dot(g["v"][1] - g["v"][0], d / np.linalg.norm(d)))


def test_head_on_restitution():
    r, v, e = 0.01, 0.5, 0.8
    # a small tangential velocity keeps |F_t| > 0: with F_t exactly zero the reference's slip
    # test 0 >= mu F1 (F1 < 0 late in the rebound) divides by zero (kernels.jl:287-290)
    cfg, g = two_spheres(r, r, (0, 0, 0), (2 * r + 1e-5, 0, 0), v1=(v, 1e-3, 0), v2=(-v, 0, 0),
                         dt=2e-7, pp_coefficient_restitution=e)
    o = O.OracleScene(cfg, g)
    before = _relative_normal_speed(o)
    for _ in range(400):
        o.step(25)
        if o.counters()["active_contacts"] == 0 and _relative_normal_speed(o) > 0:
            break
    after = _relative_normal_speed(o)
    assert before == pytest.approx(-2 * v, rel=1e-5)
    assert after / -before == pytest.approx(e, rel=0.05)


def test_wall_rebound_restitution():
    r, v, e = 0.01, 0.5, 0.7
    V = 4 / 3 * math.pi * r ** 3
    gr = grains_in([1], [V], [RHO * V], np.array([[0.0, 0.0, r + 1e-5]]),
                   np.array([[1e-3, 0, -v]]))
    cfg = DEMConfig(domain_min=[-1.0] * 3, domain_max=[1.0] * 3, gravity=[0.0, 0.0, 0.0],
                    dt=2e-7, wall_normal=[0.0, 0.0, -1.0], wall_distance=0.0,
                    pw_coefficient_restitution=e, particle_elastic_modulus=1e8)
    o = O.OracleScene(cfg, gr)
    for _ in range(400):
        o.step(25)
        g = o.read_grains()[0]
        if g["v"][2] > 0 and g["k"][2] > r:
            break
    assert o.read_grains()[0]["v"][2] / v == pytest.approx(e, rel=0.05)


def test_dropped_candidate_keeps_contact_active():
    """An unbonded active contact whose pair leaves the cp_list is not deactivated (kernels.jl:
    97-119) and is evaluated with gap > 0 (kernels.jl:241-245)."""
    r, delta = 0.01, 1e-5
    cfg, g = two_spheres(r, r, (0, 0, 0), (2 * r - delta, 0, 0), dt=1e-7)
    o = O.OracleScene(cfg, g)
    o.step(1)
    assert o.counters()["active_contacts"] == 1
    s = o.read_grains()
    s[1]["k"] = [0.5, 0.5, 0.5]  # far outside the 8-bucket window of particle 1
    o.write_grains(s)
    o.step(1)
    cnt = o.counters()
    assert cnt["active_contacts"] == 1
    assert cnt["fatal_gap"] >= 1


def test_exactly_normal_impact_reproduces_reference_nan():
    """Zero tangential velocity: once damping drives F1 negative the reference's slip branch
    computes ratio = mu F1 / 0 and F2 = -inf * 0 = NaN (kernels.jl:287-290); kept as is."""
    r, v = 0.01, 0.5
    cfg, g = two_spheres(r, r, (0, 0, 0), (2 * r + 1e-5, 0, 0), v1=(v, 0, 0), v2=(-v, 0, 0),
                         dt=2e-7, pp_coefficient_restitution=0.8)
    o = O.OracleScene(cfg, g)
    o.step(10000)
    assert np.isnan(o.read_grains()["k"]).any()
