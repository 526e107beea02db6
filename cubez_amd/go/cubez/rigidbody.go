package cubez

import (
	m "github.com/tbogdala/cubez/math"
)

// The reference's body constants (rigidbody.go:11-19).
const (
	defaultDamping = 0.95
	sleepEpsilon   = 0.3
)

var defaultAcceleration = m.Vector3{0.0, -9.78, 0.0}

// RigidBody has the exported fields of the reference's RigidBody
// (rigidbody.go:23-101), so a caller's field accesses compile and mean the
// same.  Its state lives on the device between flushes: the Go fields hold
// the values of the last download, plus whatever the caller assigned since
// (uploaded at the next flush).
type RigidBody struct {
	LinearDamping        m.Real
	AngularDamping       m.Real
	Position             m.Vector3
	Orientation          m.Quat
	Velocity             m.Vector3
	Acceleration         m.Vector3
	Rotation             m.Vector3
	InverseInertiaTensor m.Matrix3
	IsAwake              bool
	CanSleep             bool

	inverseInertiaTensorWorld m.Matrix3
	inverseMass               m.Real
	mass                      m.Real
	transform                 m.Matrix3x4
	forceAccum                m.Vector3
	torqueAccum               m.Vector3
	lastFrameAccelleration    m.Vector3
	motion                    m.Real

	// the device slot of this body in the engine (-1: not yet registered)
	slot int
	// the collider the body belongs to (nil: none; it then never collides)
	collider Collider
}

// NewRigidBody returns a body with the reference's defaults
// (rigidbody.go:104-114): identity orientation, damping 0.95 and 0.95,
// gravity (0, -9.78, 0), identity world inverse inertia, allowed to sleep
// and awake.  It is registered with the engine.
func NewRigidBody() *RigidBody {
	body := &RigidBody{slot: -1}
	body.Orientation = m.Quat{1, 0, 0, 0}
	body.LinearDamping = defaultDamping
	body.AngularDamping = defaultDamping
	body.Acceleration = defaultAcceleration
	body.inverseInertiaTensorWorld = m.Matrix3{1, 0, 0, 0, 1, 0, 0, 0, 1}
	body.CanSleep = true
	body.SetAwake(true)
	theEngine().register(body)
	return body
}

// Clone returns a new registered body with this body's current data
// (rigidbody.go:117-121).
func (body *RigidBody) Clone() *RigidBody {
	theEngine().sync()
	nb := NewRigidBody()
	slot := nb.slot
	*nb = *body
	nb.slot = slot
	nb.collider = nil
	return nb
}

// SetMass sets the mass and its inverse (rigidbody.go:124-127).
func (body *RigidBody) SetMass(mass m.Real) {
	body.mass = mass
	body.inverseMass = 1.0 / mass
}

// SetInfiniteMass makes the body immovable: inverse mass 0
// (rigidbody.go:131-134).
func (body *RigidBody) SetInfiniteMass() {
	body.mass = 0.0
	body.inverseMass = 0.0
}

// HasFiniteMass reports a positive inverse mass (rigidbody.go:137-142).
func (body *RigidBody) HasFiniteMass() bool { return body.inverseMass > 0.0 }

// GetMass is the mass, or m.MaxValue for an infinite one
// (rigidbody.go:145-150).
func (body *RigidBody) GetMass() m.Real {
	if body.inverseMass == 0.0 {
		return m.MaxValue
	}
	return body.mass
}

// GetInverseMass is the inverse mass (rigidbody.go:153-155).
func (body *RigidBody) GetInverseMass() m.Real { return body.inverseMass }

// GetTransform is the body-to-world transform of the last flush
// (rigidbody.go:158-160).
func (body *RigidBody) GetTransform() m.Matrix3x4 {
	theEngine().syncIfPending()
	return body.transform
}

// GetLastFrameAccelleration is the linear acceleration of the last
// integration (rigidbody.go:164-166).
func (body *RigidBody) GetLastFrameAccelleration() m.Vector3 {
	theEngine().syncIfPending()
	return body.lastFrameAccelleration
}

// GetInverseInertiaTensorWorld is the world-space inverse inertia of the
// last flush (rigidbody.go:170-172).
func (body *RigidBody) GetInverseInertiaTensorWorld() m.Matrix3 {
	theEngine().syncIfPending()
	return body.inverseInertiaTensorWorld
}

// SetInertiaTensor stores the inverse of the given inertia tensor
// (rigidbody.go:176-178).
func (body *RigidBody) SetInertiaTensor(t *m.Matrix3) {
	body.InverseInertiaTensor = t.Invert()
}

// SetAwake wakes the body with some motion, or puts it to sleep with its
// velocities cleared; it ignores CanSleep (rigidbody.go:182-192).
func (body *RigidBody) SetAwake(awake bool) {
	body.IsAwake = awake
	if awake {
		body.motion = sleepEpsilon * 2.0
		return
	}
	body.Velocity = m.Vector3{}
	body.Rotation = m.Vector3{}
}

// AddVelocity adds v to Velocity (rigidbody.go:195-197).
func (body *RigidBody) AddVelocity(v *m.Vector3) { body.Velocity.Add(v) }

// AddRotation adds v to Rotation (rigidbody.go:200-202).
func (body *RigidBody) AddRotation(v *m.Vector3) { body.Rotation.Add(v) }

// ClearAccumulators zeroes the force and torque accumulators
// (rigidbody.go:206-209).
func (body *RigidBody) ClearAccumulators() {
	body.forceAccum = m.Vector3{}
	body.torqueAccum = m.Vector3{}
}

// Integrate advances the body by duration (rigidbody.go:213-259, which
// ends with CalculateDerivedData).  The device integrates every body
// recorded in the frame at the frame's first collision check, resolve or
// Sync.
func (body *RigidBody) Integrate(duration m.Real) { theEngine().noteIntegrate(body, duration) }

// CalculateDerivedData recomputes the body's transform and world inverse
// inertia from its position and orientation (rigidbody.go:268-272), at the
// next flush.
func (body *RigidBody) CalculateDerivedData() { theEngine().noteDerive(body) }
