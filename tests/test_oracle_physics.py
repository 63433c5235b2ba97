"""The oracle's rigid-body restatement (physics/) against closed forms.

Pins: exact free fall under gravity with power-of-two steps; the constant
spin of a free body against 2 atan(w dt / 2) per step; the sleep rule's
exact motion sequence with bias 1/2; the spring force and torque against a
binary64 evaluation; world-space inertia for a diagonal rotation; powf
against binary64 pow rounded to binary32; object transforms and inverses
against numpy products and inverses.
"""
import ctypes

import numpy as np
import pytest

import oracle

f32 = np.float32
FLT_MAX = np.finfo(np.float32).max


def default_body():
    """RigidBody::default() (rigid_body.rs:88-112)."""
    b = np.zeros(1, oracle.BODY_DTYPE)
    b["inverse_mass"] = FLT_MAX
    b["inverse_inertia_tensor"][0] = np.eye(3, dtype=f32).reshape(9)
    b["inverse_inertia_tensor_world"][0] = np.eye(3, dtype=f32).reshape(9)
    b["linear_drag"] = b["angular_drag"] = 1.0
    b["orientation"][0] = (0, 0, 0, 1)
    b["can_sleep"] = 1
    b["transform"][0] = np.eye(4, dtype=f32).reshape(16)
    return b


def gravity(body, g):
    f = np.zeros(1, oracle.FORCE_DTYPE)
    f["kind"], f["body"] = 1, body
    f["v0"][0] = g
    return f


def test_powf_matches_binary64_pow_rounded():
    rng = np.random.default_rng(3)
    x = rng.uniform(0, 4, 20000).astype(f32)
    y = rng.uniform(-10, 10, 20000).astype(f32)
    x[:500] = rng.uniform(0.9, 1.0, 500).astype(f32)
    y[:500] = rng.uniform(0, 0.05, 500).astype(f32)
    got = np.array([oracle.powf(a, b) for a, b in zip(x, y)], f32)
    ref = (x.astype(np.float64) ** y.astype(np.float64)).astype(f32)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    # exact special values
    assert oracle.powf(1.0, 0.37) == 1.0 and oracle.powf(0.37, 0.0) == 1.0
    assert oracle.powf(0.5, 1.0) == 0.5 and oracle.powf(0.5, 3.0) == 0.125 and oracle.powf(2.0, 10.0) == 1024.0
    assert oracle.powf(0.0, 2.0) == 0.0 and np.isinf(oracle.powf(0.0, -1.0)) and np.isnan(oracle.powf(-2.0, 0.5))


def test_powf_agrees_with_the_host_libm_almost_everywhere():
    """f32::powf is the platform libm's powf, which need not be correctly
    rounded; the restatement is.  Record how often they differ."""
    libm = ctypes.CDLL("libm.so.6")
    libm.powf.restype = ctypes.c_float
    libm.powf.argtypes = [ctypes.c_float, ctypes.c_float]
    rng = np.random.default_rng(4)
    x = rng.uniform(0.5, 1.0, 5000).astype(f32)
    y = rng.uniform(0, 1, 5000).astype(f32)
    g = np.array([libm.powf(float(a), float(b)) for a, b in zip(x, y)], f32)
    o = np.array([oracle.powf(a, b) for a, b in zip(x, y)], f32)
    diff = g.view(np.uint32).astype(np.int64) - o.view(np.uint32).astype(np.int64)
    assert np.abs(diff).max() <= 1 and (diff != 0).mean() < 0.01


def test_free_fall_is_exact_with_power_of_two_steps():
    b = default_body()
    b["inverse_mass"] = 1.0  # set_mass(1)
    b["is_awake"], b["can_sleep"] = 1, 0
    forces = gravity(0, (0, -8, 0))
    for n in range(1, 11):
        oracle.world_run_physics(b, forces, 0.5)
        assert b["velocity"][0].tolist() == [0.0, -4.0 * n, 0.0]
        assert b["position"][0].tolist() == [0.0, -float(n * (n + 1)), 0.0]
        assert b["force_accumulator"][0].tolist() == [0.0, 0.0, 0.0]
    # orientation (0, 0, 0, 1) is a half turn about z
    t = b["transform"][0].reshape(4, 4)  # [c][r]
    assert np.array_equal(t[:3, :3], np.diag([-1.0, -1.0, 1.0]).astype(f32))
    assert t[3].tolist() == [0.0, -110.0, 0.0, 1.0]
    assert np.array_equal(b["inverse_inertia_tensor_world"][0], np.eye(3, dtype=f32).reshape(9))


def test_default_mass_gravity_and_asleep_bodies():
    """RigidBody::default() has inverse mass f32::MAX: Gravity applies
    g * (1 / f32::MAX), a subnormal mass, and wakes the body; a body without
    forces that is asleep does not move."""
    b = np.concatenate([default_body(), default_body()])
    b["velocity"][1] = (1, 2, 3)
    oracle.world_run_physics(b, gravity(0, (0, -9.81, 0)), 0.25)
    mass = f32(1.0) / FLT_MAX
    assert mass > 0 and mass < np.finfo(np.float32).tiny
    lfa = f32(0) + (f32(-9.81) * mass) * FLT_MAX
    assert b["last_frame_acceleration"][0][1] == lfa
    assert b["is_awake"][0] == 1 and b["is_awake"][1] == 0
    assert b["velocity"][1].tolist() == [1.0, 2.0, 3.0] and b["position"][1].tolist() == [0.0, 0.0, 0.0]


def test_constant_spin_turns_by_the_closed_form_angle():
    b = default_body()
    b["orientation"][0] = (1, 0, 0, 0)
    b["rotation"][0] = (0, 0, 2)
    b["is_awake"], b["can_sleep"] = 1, 0
    dt = 1.0 / 64
    n = 200
    for _ in range(n):
        oracle.world_run_physics(b, np.zeros(0, oracle.FORCE_DTYPE), dt)
    q = b["orientation"][0].astype(np.float64)
    assert abs(np.linalg.norm(q) - 1) < 1e-6
    # each step q' = normalize(q + (0, 0, 0, w dt / 2) q) turns by 2 atan(w dt / 2)
    angle = n * 2 * np.arctan(2 * dt / 2)
    assert abs(2 * np.arctan2(q[3], q[0]) - angle) < 2e-5
    t = b["transform"][0].reshape(4, 4).astype(np.float64)
    assert np.allclose(t[0, :2], [np.cos(angle), np.sin(angle)], atol=2e-5)
    assert np.allclose(t[1, :2], [-np.sin(angle), np.cos(angle)], atol=2e-5)
    assert b["rotation"][0].tolist() == [0.0, 0.0, 2.0]


def test_sleep_rule_motion_sequence():
    """duration 1 makes the bias powf(0.5, 1) = 0.5: motion starts at
    0.3 * 2 (set_awake) and moves halfway to |v|^2 each step; the body falls
    asleep (velocity zeroed) once motion < 0.3, and motion is capped at
    10 * 0.3."""
    b = default_body()
    b["is_awake"], b["motion"] = 1, f32(0.3) * f32(2)
    b["velocity"][0] = (0.5, 0, 0)
    m = f32(0.3) * f32(2)
    for step in range(3):
        oracle.world_run_physics(b, np.zeros(0, oracle.FORCE_DTYPE), 1.0)
        m = f32(0.5) * m + (f32(1) - f32(0.5)) * f32(0.25)
        if step < 2:
            assert b["motion"][0] == m and b["is_awake"][0] == 1
    assert m < f32(0.3) and b["is_awake"][0] == 0 and b["velocity"][0].tolist() == [0, 0, 0]
    fast = default_body()
    fast["is_awake"] = 1
    fast["velocity"][0] = (10, 0, 0)
    oracle.world_run_physics(fast, np.zeros(0, oracle.FORCE_DTYPE), 1.0)
    assert fast["motion"][0] == f32(10) * f32(0.3)


def test_spring_force_and_torque_against_binary64():
    b = np.concatenate([default_body(), default_body()])
    b["orientation"][:] = (1, 0, 0, 0)
    b["position"][0], b["position"][1] = (0, 0, 0), (3, 1, 0)
    oracle.world_start_frame(b)  # transforms from position and orientation
    f = np.zeros(1, oracle.FORCE_DTYPE)
    f["kind"], f["body"], f["other"] = 2, 0, 1
    f["v0"][0], f["v1"][0] = (0, 1, 0), (0.5, 0, 0)
    f["spring_constant"], f["rest_length"] = 2.0, 1.0
    b["inverse_mass"] = 1.0
    # a body woken by a force keeps motion 0 (apply_force does not set it):
    # with small motion it falls asleep within the same integrate
    woken = b.copy()
    oracle.world_run_physics(woken, f, 1e-3)
    assert woken["is_awake"][0] == 0 and not woken["velocity"][0].any() and not woken["rotation"][0].any()
    b["can_sleep"] = 0
    b0 = b.copy()
    oracle.world_run_physics(b, f, 1e-3)
    lws = np.array([0, 1, 0.0])
    ows = np.array([3.5, 1, 0.0])
    d = lws - ows
    force = -(d / np.linalg.norm(d)) * 2.0 * abs(np.linalg.norm(d) - 1.0)
    torque = np.cross(lws - b0["position"][0], force)
    # the accumulators are cleared by integrate: read the acceleration it used
    assert np.allclose(b["last_frame_acceleration"][0], force, rtol=1e-6)
    assert b["is_awake"][0] == 1 and b["is_awake"][1] == 0
    # angular velocity after one step = I_w^-1 torque dt with identity inertia
    assert np.allclose(b["rotation"][0], torque * 1e-3, rtol=1e-5, atol=1e-9)


def test_objects_from_bodies_against_numpy():
    rng = np.random.default_rng(5)
    n = 20
    b = np.concatenate([default_body() for _ in range(n)])
    q = rng.normal(size=(n, 4)).astype(f32)
    b["orientation"] = q
    b["position"] = rng.uniform(-5, 5, (n, 3)).astype(f32)
    b["inverse_inertia_tensor"] = rng.uniform(0.5, 2, (n, 9)).astype(f32)
    oracle.world_start_frame(b)
    init = np.stack([np.diag(rng.uniform(0.5, 2, 4)).astype(f32) for _ in range(n)])
    init[:, 3, :3] = rng.uniform(-1, 1, (n, 3))
    init[:, 3, 3] = 1
    init = init.reshape(n, 16)
    object_body = np.arange(n, dtype=np.uint32)[::-1].copy()
    object_body[3] = 0xFFFFFFFF
    m, minv = oracle.objects_from_bodies(b, object_body, init)
    for i in range(n):
        if object_body[i] == 0xFFFFFFFF:
            assert not m[i].any()
            continue
        A = b["transform"][object_body[i]].reshape(4, 4).T.astype(np.float64)  # [r][c]
        B = init[i].reshape(4, 4).T.astype(np.float64)
        M = m[i].reshape(4, 4).T.astype(np.float64)
        assert np.allclose(M, A @ B, rtol=1e-6, atol=1e-6)
        assert np.allclose(minv[i].reshape(4, 4).T, np.linalg.inv(M), rtol=1e-5, atol=1e-6)
    # world-space inertia is R I R^T
    for i in range(n):
        R = b["transform"][i].reshape(4, 4).T[:3, :3].astype(np.float64)
        Ib = b["inverse_inertia_tensor"][i].reshape(3, 3).T.astype(np.float64)
        Iw = b["inverse_inertia_tensor_world"][i].reshape(3, 3).T.astype(np.float64)
        assert np.allclose(Iw, R @ Ib @ R.T, rtol=1e-5, atol=1e-5)
    # exact for a scaled translation of an identity body
    e = default_body()
    e["orientation"][0] = (1, 0, 0, 0)
    oracle.world_start_frame(e)
    s = np.diag([2.0, 2.0, 2.0, 1.0]).astype(f32)
    s[3, :3] = (-1, -1, -1)
    m1, mi1 = oracle.objects_from_bodies(e, np.zeros(1, np.uint32), s.reshape(1, 16))
    assert np.array_equal(m1[0], s.reshape(16))
    expect = np.diag([0.5, 0.5, 0.5, 1.0]).astype(f32)
    expect[3, :3] = 0.5
    assert np.array_equal(mi1[0], expect.reshape(16))
