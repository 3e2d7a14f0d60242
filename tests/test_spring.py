"""The cage spring simulator against the reference's own dynamics tests
(src/deformation/dynamics.rs:328-450), restated on the CPU, plus the step's
guards (dt clamp to MAX_DT, no motion for dt <= 0) and its closed form."""
import math

import numpy as np

from cindergs_amd.cages import CageSpringSimulator, CollisionPlane

F32 = np.float32


def test_spring_convergence_to_rest():
    """dynamics.rs:332-352: +2 in x relaxes to < 0.05 after 120 steps of 16 ms."""
    sim = CageSpringSimulator([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0]])
    sim.current_positions[0, 0] = 2.0
    for _ in range(120):
        sim.step(0.016)
    assert np.linalg.norm(sim.current_positions[0] - sim.rest_positions[0]) < 0.05


def test_timestep_clamping_prevents_explosion():
    """dynamics.rs:354-366: a 2 s hitch is clamped to MAX_DT = 0.033 s."""
    sim = CageSpringSimulator([[0.0, 0.0, 0.0]])
    sim.current_positions[0, 0] = 1.0
    sim.step(2.0)
    x = float(sim.current_positions[0, 0])
    assert not math.isnan(x) and abs(x) < 10.0
    # exactly one clamped step: v = -k x dt gamma, x' = x + v dt
    ref = CageSpringSimulator([[0.0, 0.0, 0.0]])
    ref.current_positions[0, 0] = 1.0
    ref.step(0.033)
    assert sim.current_positions[0, 0] == ref.current_positions[0, 0]
    dt = F32(0.033)
    gamma = math.exp(math.log(0.92) * 60.0 * 0.033)
    v = -180.0 * 1.0 * 0.033 * gamma
    assert abs(float(ref.velocities[0, 0]) - v) < 1e-5
    assert abs(x - (1.0 + v * float(dt))) < 1e-5


def test_non_positive_dt_does_nothing():
    sim = CageSpringSimulator([[0.0, 0.0, 0.0]])
    sim.current_positions[0] = [0.5, -0.25, 1.0]
    sim.velocities[0] = [1.0, 2.0, 3.0]
    before = (sim.current_positions.copy(), sim.velocities.copy())
    sim.step(0.0)
    sim.step(-0.1)
    np.testing.assert_array_equal(sim.current_positions, before[0])
    np.testing.assert_array_equal(sim.velocities, before[1])


def test_pinned_vertex_remains_static():
    """dynamics.rs:368-378"""
    sim = CageSpringSimulator([[0.0, 0.0, 0.0]])
    sim.current_positions[0] = [5.0, 5.0, 5.0]
    sim.set_pinned(0)
    sim.step(0.016)
    np.testing.assert_array_equal(sim.current_positions[0], [5.0, 5.0, 5.0])
    np.testing.assert_array_equal(sim.velocities[0], [0.0, 0.0, 0.0])


def test_damping_is_frame_rate_independent():
    """dynamics.rs:380-414: one simulated second at 30, 60 and 120 Hz ends
    within 2e-2 of each other."""
    sims = {}
    for hz in (30, 60, 120):
        s = CageSpringSimulator([[0.0, 0.0, 0.0]])
        s.current_positions[0, 0] = 1.0
        for _ in range(hz):
            s.step(1.0 / hz)
        sims[hz] = float(s.current_positions[0, 0])
    assert abs(sims[120] - sims[60]) < 2e-2
    assert abs(sims[30] - sims[60]) < 2e-2


def test_damping_coefficient_matches_reference_frame():
    """dynamics.rs:416-427: exp(-c * 1 s) = 0.92^60 to 1e-5."""
    c = CageSpringSimulator.damping_coefficient(0.92)
    assert abs(math.exp(-float(c)) - 0.92 ** 60) < 1e-5
    # the cage workload's spring: retention exp(-4/60) per 1/60 s frame is c = 4 /s
    c4 = CageSpringSimulator.damping_coefficient(math.exp(-4.0 / 60.0))
    assert abs(float(c4) - 4.0) < 1e-4


def test_multi_handle_pinning_and_collision_plane():
    """dynamics.rs:429-449"""
    sim = CageSpringSimulator([[0.0, 1.0, 0.0], [1.0, 1.0, 0.0], [2.0, 1.0, 0.0]])
    sim.pin_vertex(0)
    assert sim.toggle_pin_vertex(2)
    assert sim.pinned_count() == 2
    sim.set_pinned_position(0, [0.0, 2.5, 0.0])
    sim.current_positions[1] = [1.0, -1.0, 0.0]
    sim.velocities[1] = [0.0, -4.0, 0.0]
    sim.add_collision_plane(CollisionPlane.ground_y(0.25))
    sim.step(0.016)
    assert abs(float(sim.current_positions[0, 1]) - 2.5) < 1e-6
    assert float(sim.current_positions[1, 1]) >= 0.25 - 1e-5
    assert sim.is_vertex_pinned(2)
    # the contact reflected the normal velocity with restitution 0.25 and
    # damped the tangential part by friction 0.2 (resolve_contact, 47-60)
    assert float(sim.velocities[1, 1]) > 0.0
    sim.clear_pins()
    assert sim.pinned_count() == 0


def test_semi_implicit_euler_step_in_f32():
    """One step on many vertices equals the per-vertex f32 update of
    step (249-277): v = (v + (x_rest - x) (k / m) dt) gamma, x += v dt."""
    rng = np.random.default_rng(0)
    rest = rng.normal(size=(50, 3)).astype(F32)
    sim = CageSpringSimulator(rest, stiffness=40.0, damping=math.exp(-4.0 / 60.0), mass=1.5)
    sim.current_positions[:] = rest + rng.normal(scale=0.1, size=rest.shape).astype(F32)
    sim.velocities[:] = rng.normal(scale=0.2, size=rest.shape).astype(F32)
    x0, v0 = sim.current_positions.copy(), sim.velocities.copy()
    sim.step(1.0 / 60.0)
    dt = F32(1.0 / 60.0)
    kinv = F32(F32(40.0) * F32(F32(1.0) / F32(1.5)))
    c = F32(-np.log(F32(math.exp(-4.0 / 60.0)), dtype=F32) * F32(60.0))
    gamma = np.exp(F32(-c * dt), dtype=F32)
    for i in range(rest.shape[0]):
        for k in range(3):
            a = F32(F32(rest[i, k] - x0[i, k]) * kinv)
            v = F32(F32(v0[i, k] + F32(a * dt)) * gamma)
            assert sim.velocities[i, k] == v
            assert sim.current_positions[i, k] == F32(x0[i, k] + F32(v * dt))


def test_jiggle_impulse_is_radial():
    """trigger_jiggle_impulse (280-306): unit radial directions from the rest
    centroid times the magnitude; pinned vertices are untouched."""
    rest = np.array([[1, 0, 0], [-1, 0, 0], [0, 2, 0], [0, 0, 0]], F32)
    sim = CageSpringSimulator(rest)
    sim.pin_vertex(1)
    sim.trigger_jiggle_impulse(0.5)
    c = rest.mean(0)
    for i in (0, 2):
        d = (rest[i] - c) / np.linalg.norm(rest[i] - c)
        np.testing.assert_allclose(sim.velocities[i], 0.5 * d, atol=1e-6)
    np.testing.assert_array_equal(sim.velocities[1], [0, 0, 0])
    d3 = (rest[3] - c) / np.linalg.norm(rest[3] - c)
    np.testing.assert_allclose(sim.velocities[3], 0.5 * d3, atol=1e-6)
