"""Host mirror of the reference's rigid-body physics (bvhtracer/src/physics/)
over the GPU library.

RigidBody (rigid_body.rs:66-463) keeps the reference's record field for field
(BODY_DTYPE, binary32) with its setters and getters; World (world.rs:14-61)
registers bodies and force generators (NullForce, Gravity, Spring,
force_generator.rs) and runs ForceRegistry::apply_forces + RigidBody::integrate
on the device (bvht_world_run_physics).  Host-side setters do their
arithmetic in binary32 in the reference's order (set_orientation normalises,
apply_force accumulates), so a body built here is the body the reference
would build.

A RigidBodyInstance is the shared handle (the reference's Rc<RefCell<..>>):
``borrow()`` returns a copy of the body's current state and ``borrow_mut()``
is a context manager that writes the body back on exit.
"""
import contextlib

import numpy as np

from . import _lib as L
from .api import default_context

f32 = np.float32
FLT_MAX = np.finfo(np.float32).max


def _v3(v):
    return np.asarray(v, f32).reshape(3)


class RigidBody:
    """RigidBody (rigid_body.rs:66-463).  ``RigidBody()`` is
    RigidBody::default() (rigid_body.rs:88-112): inverse mass f32::MAX,
    identity inertia tensors, drags 1, orientation Quaternion::unit_z =
    (s, x, y, z) = (0, 0, 0, 1), asleep, can sleep, identity transform."""

    def __init__(self, record=None):
        if record is not None:
            self.rec = np.array(record, L.BODY_DTYPE).reshape(1)
            return
        r = np.zeros(1, L.BODY_DTYPE)
        r["inverse_mass"] = FLT_MAX
        r["inverse_inertia_tensor"][0] = np.eye(3, dtype=f32).reshape(9)
        r["inverse_inertia_tensor_world"][0] = np.eye(3, dtype=f32).reshape(9)
        r["linear_drag"] = 1.0
        r["angular_drag"] = 1.0
        r["orientation"][0] = (0.0, 0.0, 0.0, 1.0)
        r["can_sleep"] = 1
        r["transform"][0] = np.eye(4, dtype=f32).reshape(16)
        self.rec = r

    @classmethod
    def default(cls):
        return cls()

    def _f(self, name):
        return self.rec[name][0]

    # mass (rigid_body.rs:205-233)
    def set_mass(self, mass):
        mass = f32(mass)
        if mass == 0:
            raise ValueError("mass must be non-zero")
        self.rec["inverse_mass"] = f32(1.0) / mass

    def get_mass(self):
        im = self.rec["inverse_mass"][0]
        return FLT_MAX if im == 0 else f32(1.0) / im

    def set_inverse_mass(self, inverse_mass):
        self.rec["inverse_mass"] = f32(inverse_mass)

    def get_inverse_mass(self):
        return self.rec["inverse_mass"][0]

    def has_finite_mass(self):
        return bool(self.rec["inverse_mass"][0] >= 0)

    # inertia (rigid_body.rs:235-273); matrices column-major [c][r]
    def set_inverse_inertia_tensor(self, inverse_inertia_tensor):
        m = np.asarray(inverse_inertia_tensor, f32).reshape(9)
        if not np.all(np.isfinite(m)):
            raise ValueError("inverse inertia tensor must be finite")
        self.rec["inverse_inertia_tensor"][0] = m

    def set_inertia_tensor(self, inertia_tensor):
        """Stores the inverse (binary64 inverse, rounded; cglinalg's inverse
        is not vendored)."""
        m = np.asarray(inertia_tensor, np.float64).reshape(3, 3)  # [c][r]
        inv = np.linalg.inv(m.T).T
        self.set_inverse_inertia_tensor(inv.astype(f32).reshape(9))

    def get_inverse_inertia_tensor(self):
        return self.rec["inverse_inertia_tensor"][0].reshape(3, 3).copy()

    def get_inverse_inertia_tensor_world(self):
        return self.rec["inverse_inertia_tensor_world"][0].reshape(3, 3).copy()

    # drag (rigid_body.rs:275-300)
    def set_drag(self, linear_drag, angular_drag):
        self.rec["linear_drag"] = f32(linear_drag)
        self.rec["angular_drag"] = f32(angular_drag)

    def set_linear_drag(self, linear_drag):
        self.rec["linear_drag"] = f32(linear_drag)

    def get_linear_drag(self):
        return self.rec["linear_drag"][0]

    def set_angular_drag(self, angular_drag):
        self.rec["angular_drag"] = f32(angular_drag)

    def get_angular_drag(self):
        return self.rec["angular_drag"][0]

    # state (rigid_body.rs:302-376)
    def set_position(self, position):
        self.rec["position"][0] = _v3(position)

    def get_position(self):
        return self.rec["position"][0].copy()

    def set_orientation(self, orientation):
        """Stores orientation.normalize(): q / sqrt(s*s + ((x*x + y*y) + z*z))."""
        q = np.asarray(orientation, f32).reshape(4)
        s, x, y, z = q
        mag = np.sqrt(s * s + ((x * x + y * y) + z * z), dtype=f32)
        self.rec["orientation"][0] = np.array([s / mag, x / mag, y / mag, z / mag], f32)

    def get_orientation(self):
        return self.rec["orientation"][0].copy()

    def get_transform(self):
        """Body to world, column-major 16 floats."""
        return self.rec["transform"][0].copy()

    def get_point_in_world_space(self, point_body):
        m = self.rec["transform"][0]
        p = _v3(point_body)
        return np.array([m[0 + r] * p[0] + m[4 + r] * p[1] + m[8 + r] * p[2] + m[12 + r] * f32(1) for r in range(3)],
                        f32)

    def get_direction_in_world_space(self, direction_body):
        m = self.rec["transform"][0]
        v = _v3(direction_body)
        return np.array([m[0 + r] * v[0] + m[4 + r] * v[1] + m[8 + r] * v[2] + m[12 + r] * f32(0) for r in range(3)],
                        f32)

    def set_velocity(self, velocity):
        self.rec["velocity"][0] = _v3(velocity)

    def get_velocity(self):
        return self.rec["velocity"][0].copy()

    def add_velocity(self, delta):
        self.rec["velocity"][0] = self.rec["velocity"][0] + _v3(delta)

    def set_rotation(self, rotation):
        self.rec["rotation"][0] = _v3(rotation)

    def get_rotation(self):
        return self.rec["rotation"][0].copy()

    def add_rotation(self, delta):
        self.rec["rotation"][0] = self.rec["rotation"][0] + _v3(delta)

    # sleep (rigid_body.rs:378-420)
    def get_awake(self):
        return bool(self.rec["is_awake"][0])

    def set_awake(self, awake):
        if awake:
            self.rec["is_awake"] = 1
            self.rec["motion"] = f32(0.3) * f32(2.0)
        else:
            self.rec["is_awake"] = 0
            self.rec["velocity"][0] = 0
            self.rec["rotation"][0] = 0

    def get_can_sleep(self):
        return bool(self.rec["can_sleep"][0])

    def set_can_sleep(self, can_sleep):
        self.rec["can_sleep"] = 1 if can_sleep else 0
        if not can_sleep and not self.get_awake():
            self.set_awake(True)

    # forces (rigid_body.rs:422-462)
    def clear_accumulators(self):
        self.rec["force_accumulator"][0] = 0
        self.rec["torque_accumulator"][0] = 0

    def apply_force(self, force):
        self.rec["force_accumulator"][0] = self.rec["force_accumulator"][0] + _v3(force)
        self.rec["is_awake"] = 1

    def apply_force_at_point(self, force_world, point_world):
        f = _v3(force_world)
        p = _v3(point_world) - self.rec["position"][0]
        self.rec["force_accumulator"][0] = self.rec["force_accumulator"][0] + f
        torque = np.array([p[1] * f[2] - p[2] * f[1], p[2] * f[0] - p[0] * f[2], p[0] * f[1] - p[1] * f[0]], f32)
        self.rec["torque_accumulator"][0] = self.rec["torque_accumulator"][0] + torque
        self.rec["is_awake"] = 1

    def apply_force_at_body_point(self, force_world, point_body):
        self.apply_force_at_point(force_world, self.get_point_in_world_space(point_body))

    def apply_torque(self, torque):
        self.rec["torque_accumulator"][0] = self.rec["torque_accumulator"][0] + _v3(torque)
        self.rec["is_awake"] = 1

    def set_acceleration(self, acceleration):
        self.rec["acceleration"][0] = _v3(acceleration)

    def get_acceleration_world(self):
        return self.rec["acceleration"][0].copy()


class RigidBodyInstance:
    """Shared handle to a registered body (force_registry.rs:15-28)."""

    def __init__(self, world, index):
        self.world = world
        self.index = index

    def borrow(self):
        """A copy of the body's current state."""
        return RigidBody(self.world._body_record(self.index))

    @contextlib.contextmanager
    def borrow_mut(self):
        body = self.borrow()
        yield body
        self.world._set_body_record(self.index, body.rec)


class NullForce:
    """NullForce (force_generator.rs:9-31)."""

    def record(self, body):
        r = np.zeros(1, L.FORCE_DTYPE)
        r["kind"], r["body"] = L.FORCE_NULL, body
        return r


class Gravity:
    """Gravity (force_generator.rs:34-56): force g * mass on finite-mass bodies."""

    def __init__(self, gravity):
        self.gravity = _v3(gravity)

    def record(self, body):
        r = np.zeros(1, L.FORCE_DTYPE)
        r["kind"], r["body"] = L.FORCE_GRAVITY, body
        r["v0"][0] = self.gravity
        return r


class Spring:
    """Spring (force_generator.rs:58-121) between the body it is registered
    for (connection point in that body's coordinates) and ``object``."""

    def __init__(self, object, spring_connection_point_body, object_connection_point_body_object, spring_constant,
                 rest_length):
        self.object = object
        self.v0 = _v3(spring_connection_point_body)
        self.v1 = _v3(object_connection_point_body_object)
        self.spring_constant = f32(spring_constant)
        self.rest_length = f32(rest_length)

    def record(self, body):
        r = np.zeros(1, L.FORCE_DTYPE)
        r["kind"], r["body"], r["other"] = L.FORCE_SPRING, body, self.object.index
        r["v0"][0], r["v1"][0] = self.v0, self.v1
        r["spring_constant"], r["rest_length"] = self.spring_constant, self.rest_length
        return r


class World:
    """World (world.rs:14-61).  Bodies and the force registry live on the
    host until the first device call, then on the device."""

    def __init__(self, ctx=None):
        self.ctx = ctx or default_context()
        self._bodies = []
        self._forces = []
        self.handle = None

    @classmethod
    def new(cls, ctx=None):
        return cls(ctx)

    def register_body(self, body):
        if self.handle is not None:
            raise RuntimeError("bodies are registered before the world first runs")
        self._bodies.append(np.array(body.rec, L.BODY_DTYPE))
        return RigidBodyInstance(self, len(self._bodies) - 1)

    def register_force_generator(self, body, generator):
        if self.handle is not None:
            raise RuntimeError("force generators are registered before the world first runs")
        if body.world is not self:
            raise ValueError("the body belongs to another world")
        self._forces.append(generator.record(body.index))

    def __len__(self):
        return len(self._bodies)

    def device(self):
        """Create the device world (bvht_world_create) on first use."""
        if self.handle is None:
            import ctypes
            if not self._bodies:
                raise ValueError("a world needs at least one body")
            bodies = np.concatenate(self._bodies)
            forces = np.concatenate(self._forces) if self._forces else np.zeros(0, L.FORCE_DTYPE)
            h = ctypes.c_void_p()
            L.check(L.lib().bvht_world_create(self.ctx.handle, L.ptr(bodies), bodies.shape[0],
                                              L.ptr(forces) if forces.shape[0] else None, forces.shape[0],
                                              ctypes.byref(h)))
            self.handle = h.value
            self.ctx.adopt(self)
        return self.handle

    def _body_record(self, i):
        if self.handle is None:
            return self._bodies[i].copy()
        out = np.zeros(1, L.BODY_DTYPE)
        L.check(L.lib().bvht_world_get_bodies(self.ctx.handle, self.handle, L.ptr(out), i, 1))
        return out

    def _set_body_record(self, i, rec):
        rec = np.ascontiguousarray(rec, L.BODY_DTYPE).reshape(1)
        if self.handle is None:
            self._bodies[i] = rec.copy()
        else:
            L.check(L.lib().bvht_world_set_bodies(self.ctx.handle, self.handle, L.ptr(rec), i, 1))

    def bodies(self):
        """Every body's current state as a BODY_DTYPE array."""
        if self.handle is None:
            return np.concatenate(self._bodies)
        out = np.zeros(len(self._bodies), L.BODY_DTYPE)
        L.check(L.lib().bvht_world_get_bodies(self.ctx.handle, self.handle, L.ptr(out), 0, out.shape[0]))
        return out

    def start_frame(self):
        """World::start_frame (world.rs:44-49)."""
        L.check(L.lib().bvht_world_start_frame(self.ctx.handle, self.device()))

    def run_physics(self, duration):
        """World::run_physics (world.rs:57-60) on the device."""
        L.check(L.lib().bvht_world_run_physics(self.ctx.handle, self.device(), f32(duration)))

    def close(self):
        if self.handle is not None:
            L.lib().bvht_world_destroy(self.ctx.handle, self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
