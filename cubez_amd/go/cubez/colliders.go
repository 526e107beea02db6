package cubez

import (
	m "github.com/tbogdala/cubez/math"
)

// Collider is the reference's collision primitive interface
// (colliders.go:15-23), unchanged.
type Collider interface {
	Clone() Collider
	CalculateDerivedData()
	GetBody() *RigidBody
	GetTransform() m.Matrix3x4
	CheckAgainstHalfSpace(plane *CollisionPlane, existingContacts []*Contact) (bool, []*Contact)
	CheckAgainstSphere(sphere *CollisionSphere, existingContacts []*Contact) (bool, []*Contact)
	CheckAgainstCube(secondCube *CollisionCube, existingContacts []*Contact) (bool, []*Contact)
}

// CollisionPlane is an infinite half-space without a body (colliders.go:29-35).
// The device knows it by value, at the first check that names it.
type CollisionPlane struct {
	Normal m.Vector3
	Offset m.Real
}

// CollisionCube is a box collider on a body (colliders.go:39-53).
type CollisionCube struct {
	Body      *RigidBody
	Offset    m.Matrix3x4
	transform m.Matrix3x4
	HalfSize  m.Vector3
}

// CollisionSphere is a sphere collider on a body (colliders.go:57-71).
type CollisionSphere struct {
	Body      *RigidBody
	Offset    m.Matrix3x4
	transform m.Matrix3x4
	Radius    m.Real
}

var identity3x4 = m.Matrix3x4{1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0}

// NewCollisionPlane returns the plane n . x = o (colliders.go:81-86).
func NewCollisionPlane(n m.Vector3, o m.Real) *CollisionPlane {
	return &CollisionPlane{Normal: n, Offset: o}
}

// Clone returns a plane with the same normal and offset
// (colliders.go:89-92).
func (p *CollisionPlane) Clone() Collider { return NewCollisionPlane(p.Normal, p.Offset) }

// CalculateDerivedData does nothing for a plane (colliders.go:95-96).
func (p *CollisionPlane) CalculateDerivedData() {}

// GetTransform is the identity: a plane has no transform (colliders.go:99-103).
func (p *CollisionPlane) GetTransform() m.Matrix3x4 { return identity3x4 }

// GetBody is nil: a plane has no body (colliders.go:106-108).
func (p *CollisionPlane) GetBody() *RigidBody { return nil }

// CheckAgainstHalfSpace: two planes never collide (colliders.go:111-113).
func (p *CollisionPlane) CheckAgainstHalfSpace(plane *CollisionPlane, existing []*Contact) (bool, []*Contact) {
	return false, existing
}

// CheckAgainstSphere is the sphere's test against this plane
// (colliders.go:116-118).
func (p *CollisionPlane) CheckAgainstSphere(sphere *CollisionSphere, existing []*Contact) (bool, []*Contact) {
	return sphere.CheckAgainstHalfSpace(p, existing)
}

// CheckAgainstCube is the cube's test against this plane
// (colliders.go:122-124).
func (p *CollisionPlane) CheckAgainstCube(cube *CollisionCube, existing []*Contact) (bool, []*Contact) {
	return cube.CheckAgainstHalfSpace(p, existing)
}

// NewCollisionSphere puts a sphere collider on the body (a new one when
// optBody is nil) (colliders.go:136-145).
func NewCollisionSphere(optBody *RigidBody, radius m.Real) *CollisionSphere {
	s := &CollisionSphere{Body: optBody, Offset: identity3x4, Radius: radius}
	if s.Body == nil {
		s.Body = NewRigidBody()
	}
	theEngine().attach(s.Body, s)
	return s
}

// Clone copies the sphere onto a clone of its body (colliders.go:148-157).
func (s *CollisionSphere) Clone() Collider {
	var b *RigidBody
	if s.Body != nil {
		b = s.Body.Clone()
	}
	ns := NewCollisionSphere(b, s.Radius)
	ns.Offset = s.Offset
	ns.transform = s.transform
	return ns
}

// GetTransform is the collider transform of the last flush (colliders.go:160-162).
func (s *CollisionSphere) GetTransform() m.Matrix3x4 {
	theEngine().syncIfPending()
	return s.transform
}

// GetBody is the sphere's body (colliders.go:165-167).
func (s *CollisionSphere) GetBody() *RigidBody { return s.Body }

// CalculateDerivedData: the collider transform, body transform times Offset
// (colliders.go:173-176), at the next flush.
func (s *CollisionSphere) CalculateDerivedData() { theEngine().noteColliderDerive(s.Body) }

// CheckAgainstHalfSpace (colliders.go:180-207).
func (s *CollisionSphere) CheckAgainstHalfSpace(plane *CollisionPlane, existing []*Contact) (bool, []*Contact) {
	return theEngine().check(s, plane, existing)
}

// CheckAgainstCube is the cube's test against this sphere (colliders.go:210-213).
func (s *CollisionSphere) CheckAgainstCube(cube *CollisionCube, existing []*Contact) (bool, []*Contact) {
	return cube.CheckAgainstSphere(s, existing)
}

// CheckAgainstSphere (colliders.go:216-251).
func (s *CollisionSphere) CheckAgainstSphere(second *CollisionSphere, existing []*Contact) (bool, []*Contact) {
	return theEngine().check(s, second, existing)
}

// NewCollisionCube puts a box collider of the given half-sizes on the body
// (a new one when optBody is nil) (colliders.go:265-274).
func NewCollisionCube(optBody *RigidBody, halfSize m.Vector3) *CollisionCube {
	c := &CollisionCube{Body: optBody, Offset: identity3x4, HalfSize: halfSize}
	if c.Body == nil {
		c.Body = NewRigidBody()
	}
	theEngine().attach(c.Body, c)
	return c
}

// Clone copies the cube onto a clone of its body (colliders.go:277-286).
func (cube *CollisionCube) Clone() Collider {
	var b *RigidBody
	if cube.Body != nil {
		b = cube.Body.Clone()
	}
	nc := NewCollisionCube(b, cube.HalfSize)
	nc.Offset = cube.Offset
	nc.transform = cube.transform
	return nc
}

// GetTransform is the collider transform of the last flush (colliders.go:289-291).
func (cube *CollisionCube) GetTransform() m.Matrix3x4 {
	theEngine().syncIfPending()
	return cube.transform
}

// GetBody is the cube's body (colliders.go:294-296).
func (cube *CollisionCube) GetBody() *RigidBody { return cube.Body }

// CalculateDerivedData: the collider transform, body transform times Offset
// (colliders.go:302-304), at the next flush.
func (cube *CollisionCube) CalculateDerivedData() { theEngine().noteColliderDerive(cube.Body) }

// CheckAgainstHalfSpace (colliders.go:308-366).
func (cube *CollisionCube) CheckAgainstHalfSpace(plane *CollisionPlane, existing []*Contact) (bool, []*Contact) {
	return theEngine().check(cube, plane, existing)
}

// CheckAgainstSphere (colliders.go:369-442).
func (cube *CollisionCube) CheckAgainstSphere(sphere *CollisionSphere, existing []*Contact) (bool, []*Contact) {
	return theEngine().check(cube, sphere, existing)
}

// CheckAgainstCube: the 15-axis separating-axis test and its contact
// (colliders.go:576-712).
func (cube *CollisionCube) CheckAgainstCube(second *CollisionCube, existing []*Contact) (bool, []*Contact) {
	return theEngine().check(cube, second, existing)
}

// CheckForCollisions tests one collider against another and appends the
// contacts found; the bool says whether this call appended any
// (colliders.go:720-747).  The dispatch is on the second collider's type,
// as in the reference.
func CheckForCollisions(one Collider, two Collider, existing []*Contact) (bool, []*Contact) {
	switch t := two.(type) {
	case *CollisionSphere:
		return one.CheckAgainstSphere(t, existing)
	case *CollisionCube:
		return one.CheckAgainstCube(t, existing)
	case *CollisionPlane:
		return one.CheckAgainstHalfSpace(t, existing)
	}
	return false, existing
}
