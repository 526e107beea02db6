package cubez

import (
	m "github.com/tbogdala/cubez/math"
)

// Contact has the exported fields of the reference's Contact
// (contact.go:17-51).  Its resolver internals (contact basis, relative
// positions, closing velocity, desired velocity change) live on the device.
type Contact struct {
	Bodies        [2]*RigidBody
	Friction      m.Real
	Restitution   m.Real
	ContactPoint  m.Vector3
	ContactNormal m.Vector3
	Penetration   m.Real
}

// NewContact returns a zero contact (contact.go:54-57).
func NewContact() *Contact { return new(Contact) }

// ResolveContacts resolves the contacts' interpenetration and closing
// velocities, worst first, for at most maxIterations iterations per phase
// (contact.go:208-222), on the device.  The bodies' state and the contacts'
// Penetration are then the reference's, in the Go values.
func ResolveContacts(maxIterations int, contacts []*Contact, duration m.Real) {
	if duration <= 0.0 || len(contacts) == 0 {
		return
	}
	theEngine().resolve(maxIterations, contacts, duration)
}
