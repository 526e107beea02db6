package cubez

// The drop-in against the reference package itself, call for call: the
// step loops of examples/cubedrop.go (d1) and examples/ballistic.go (d2),
// each body's state compared to the bit every frame, and every
// CheckForCollisions call's bool and appended contacts.  Needs a HIP
// device and libcubez_amd.so (built by __graft_entry__.build()).

import (
	"testing"

	ref "github.com/tbogdala/cubez"
	m "github.com/tbogdala/cubez/math"
)

const dt = m.Real(1.0 / 60.0)

func sameBody(t *testing.T, frame, i int, a *RigidBody, b *ref.RigidBody) {
	t.Helper()
	if a.Position != b.Position || a.Orientation != b.Orientation || a.Velocity != b.Velocity ||
		a.Rotation != b.Rotation || a.IsAwake != b.IsAwake {
		t.Fatalf("frame %d body %d: drop-in %v %v %v %v %v, reference %v %v %v %v %v", frame, i,
			a.Position, a.Orientation, a.Velocity, a.Rotation, a.IsAwake,
			b.Position, b.Orientation, b.Velocity, b.Rotation, b.IsAwake)
	}
}

func sameCall(t *testing.T, what string, fa bool, ca []*Contact, fb bool, cb []*ref.Contact) {
	t.Helper()
	if fa != fb || len(ca) != len(cb) {
		t.Fatalf("%s: drop-in (%v, %d contacts), reference (%v, %d)", what, fa, len(ca), fb, len(cb))
	}
	for k := range ca {
		a, b := ca[k], cb[k]
		if a.ContactPoint != b.ContactPoint || a.ContactNormal != b.ContactNormal || a.Penetration != b.Penetration ||
			a.Friction != b.Friction || a.Restitution != b.Restitution {
			t.Fatalf("%s: contact %d differs: %+v vs %+v", what, k, *a, *b)
		}
	}
}

// d1: cubes dropped in a row onto the ground (examples/cubedrop.go:140-175)
// and stepped with its updateCallback (:28-76).
func TestCubeDropLoopMatchesReference(t *testing.T) {
	resetEngine()
	const n = 10
	var amd []*CollisionCube
	var rc []*ref.CollisionCube
	for i := 0; i < n; i++ {
		pos := m.Vector3{m.Real(i*2-n/2) - 0.5 + m.Real(i%3)*0.3, 10.0 + m.Real(i%4)*0.7, m.Real(i%2) * 0.4}
		half := m.Vector3{0.5, 0.5, 0.5}
		var inertia m.Matrix3
		inertia.SetBlockInertiaTensor(&half, 8.0)

		a := NewCollisionCube(nil, half)
		a.Body.Position = pos
		a.Body.SetMass(8.0)
		a.Body.SetInertiaTensor(&inertia)
		a.Body.CalculateDerivedData()
		a.CalculateDerivedData()
		amd = append(amd, a)

		b := ref.NewCollisionCube(nil, half)
		b.Body.Position = pos
		b.Body.SetMass(8.0)
		b.Body.SetInertiaTensor(&inertia)
		b.Body.CalculateDerivedData()
		b.CalculateDerivedData()
		rc = append(rc, b)
	}
	groundA := NewCollisionPlane(m.Vector3{0, 1, 0}, 0)
	groundB := ref.NewCollisionPlane(m.Vector3{0, 1, 0}, 0)

	for frame := 0; frame < 400; frame++ {
		for i := range amd {
			amd[i].Body.Integrate(dt)
			amd[i].CalculateDerivedData()
			rc[i].Body.Integrate(dt)
			rc[i].CalculateDerivedData()
		}
		var ca []*Contact
		var cb []*ref.Contact
		anyA, anyB := false, false
		for i := range amd {
			la, lb := len(ca), len(cb)
			var fa, fb bool
			fa, ca = amd[i].CheckAgainstHalfSpace(groundA, ca)
			fb, cb = rc[i].CheckAgainstHalfSpace(groundB, cb)
			sameCall(t, "ground", fa, ca[la:], fb, cb[lb:])
			anyA, anyB = anyA || fa, anyB || fb
			for j := range amd {
				if i == j {
					continue
				}
				la, lb = len(ca), len(cb)
				fa, ca = CheckForCollisions(amd[i], amd[j], ca)
				fb, cb = ref.CheckForCollisions(rc[i], rc[j], cb)
				sameCall(t, "cube pair", fa, ca[la:], fb, cb[lb:])
				anyA, anyB = anyA || fa, anyB || fb
			}
		}
		if anyA != anyB {
			t.Fatalf("frame %d: found %v vs %v", frame, anyA, anyB)
		}
		if anyA {
			ResolveContacts(len(ca)*8, ca, dt)
			ref.ResolveContacts(len(cb)*8, cb, dt)
			for k := range ca {
				if ca[k].Penetration != cb[k].Penetration {
					t.Fatalf("frame %d contact %d: penetration %v vs %v", frame, k, ca[k].Penetration, cb[k].Penetration)
				}
			}
		}
		for i := range amd {
			sameBody(t, frame, i, amd[i].Body, rc[i].Body)
		}
	}
}

// d2: a cube, a static infinite-mass backboard and bullets
// (examples/ballistic.go:100-140), stepped with its updateCallback
// (:25-96): the backboard is never integrated, and the ground plane is a
// new CollisionPlane every frame.
func TestBallisticLoopMatchesReference(t *testing.T) {
	resetEngine()
	mk := func(pos m.Vector3, half m.Vector3, mass m.Real) (*CollisionCube, *ref.CollisionCube) {
		var inertia m.Matrix3
		inertia.SetBlockInertiaTensor(&half, mass)
		a := NewCollisionCube(nil, half)
		b := ref.NewCollisionCube(nil, half)
		a.Body.Position, b.Body.Position = pos, pos
		if mass > 0 {
			a.Body.SetMass(mass)
			b.Body.SetMass(mass)
			a.Body.SetInertiaTensor(&inertia)
			b.Body.SetInertiaTensor(&inertia)
		} else {
			a.Body.SetInfiniteMass()
			b.Body.SetInfiniteMass()
			a.Body.InverseInertiaTensor = m.Matrix3{}
			b.Body.InverseInertiaTensor = m.Matrix3{}
		}
		a.Body.CalculateDerivedData()
		a.CalculateDerivedData()
		b.Body.CalculateDerivedData()
		b.CalculateDerivedData()
		return a, b
	}
	cubeA, cubeB := mk(m.Vector3{0, 5, 0}, m.Vector3{0.5, 0.5, 0.5}, 8)
	boardA, boardB := mk(m.Vector3{0, 2, -10}, m.Vector3{0.5, 2, 0.25}, 0)
	var bulletsA []*CollisionSphere
	var bulletsB []*ref.CollisionSphere
	for k := 0; k < 8; k++ {
		pos := m.Vector3{m.Real(k%5-2) * 0.45, 2.0, 20 + 2.5*m.Real(k)}
		a := NewCollisionSphere(nil, 0.2)
		b := ref.NewCollisionSphere(nil, 0.2)
		a.Body.Position, b.Body.Position = pos, pos
		a.Body.SetMass(1.5)
		b.Body.SetMass(1.5)
		a.Body.Velocity, b.Body.Velocity = m.Vector3{0, 0, -40}, m.Vector3{0, 0, -40}
		a.Body.Acceleration, b.Body.Acceleration = m.Vector3{0, -2.5, 0}, m.Vector3{0, -2.5, 0}
		var inertia m.Matrix3
		inertia.SetInertiaTensorCoeffs(0.4*1.5*0.04, 0.4*1.5*0.04, 0.4*1.5*0.04, 0, 0, 0)
		a.Body.SetInertiaTensor(&inertia)
		b.Body.SetInertiaTensor(&inertia)
		a.Body.CalculateDerivedData()
		a.CalculateDerivedData()
		b.Body.CalculateDerivedData()
		b.CalculateDerivedData()
		bulletsA, bulletsB = append(bulletsA, a), append(bulletsB, b)
	}
	for frame := 0; frame < 200; frame++ {
		cubeA.Body.Integrate(dt)
		cubeA.CalculateDerivedData()
		cubeB.Body.Integrate(dt)
		cubeB.CalculateDerivedData()
		for k := range bulletsA {
			bulletsA[k].Body.Integrate(dt)
			bulletsA[k].CalculateDerivedData()
			bulletsB[k].Body.Integrate(dt)
			bulletsB[k].CalculateDerivedData()
		}
		groundA := NewCollisionPlane(m.Vector3{0, 1, 0}, 0)
		groundB := ref.NewCollisionPlane(m.Vector3{0, 1, 0}, 0)
		var ca []*Contact
		var cb []*ref.Contact
		anyA, anyB := false, false
		call := func(what string, fa bool, na []*Contact, fb bool, nb []*ref.Contact) {
			la, lb := len(ca), len(cb)
			ca, cb = na, nb
			sameCall(t, what, fa, ca[la:], fb, cb[lb:])
			anyA, anyB = anyA || fa, anyB || fb
		}
		{
			fa, na := CheckForCollisions(cubeA, groundA, ca)
			fb, nb := ref.CheckForCollisions(cubeB, groundB, cb)
			call("cube-ground", fa, na, fb, nb)
			fa, na = CheckForCollisions(cubeA, boardA, ca)
			fb, nb = ref.CheckForCollisions(cubeB, boardB, cb)
			call("cube-board", fa, na, fb, nb)
		}
		for k := range bulletsA {
			fa, na := CheckForCollisions(bulletsA[k], groundA, ca)
			fb, nb := ref.CheckForCollisions(bulletsB[k], groundB, cb)
			call("bullet-ground", fa, na, fb, nb)
			fa, na = CheckForCollisions(cubeA, bulletsA[k], ca)
			fb, nb = ref.CheckForCollisions(cubeB, bulletsB[k], cb)
			call("cube-bullet", fa, na, fb, nb)
			fa, na = CheckForCollisions(boardA, bulletsA[k], ca)
			fb, nb = ref.CheckForCollisions(boardB, bulletsB[k], cb)
			call("board-bullet", fa, na, fb, nb)
			for j := range bulletsA {
				if j == k {
					continue
				}
				fa, na = CheckForCollisions(bulletsA[j], bulletsA[k], ca)
				fb, nb = ref.CheckForCollisions(bulletsB[j], bulletsB[k], cb)
				call("bullet-bullet", fa, na, fb, nb)
			}
		}
		if anyA {
			ResolveContacts(len(ca)*8, ca, dt)
			ref.ResolveContacts(len(cb)*8, cb, dt)
		}
		sameBody(t, frame, 0, cubeA.Body, cubeB.Body)
		sameBody(t, frame, 1, boardA.Body, boardB.Body)
		for k := range bulletsA {
			sameBody(t, frame, 2+k, bulletsA[k].Body, bulletsB[k].Body)
		}
	}
}

// Clone, GetMass and HasFiniteMass keep the reference's meaning
// (rigidbody.go:117-150).
func TestBodyAccessorsMatchReference(t *testing.T) {
	resetEngine()
	a := NewRigidBody()
	b := ref.NewRigidBody()
	a.SetMass(4)
	b.SetMass(4)
	if a.GetMass() != b.GetMass() || a.GetInverseMass() != b.GetInverseMass() || a.HasFiniteMass() != b.HasFiniteMass() {
		t.Fatal("finite mass accessors differ")
	}
	a.SetInfiniteMass()
	b.SetInfiniteMass()
	if a.GetMass() != b.GetMass() || a.HasFiniteMass() != b.HasFiniteMass() {
		t.Fatal("infinite mass accessors differ")
	}
	a.Position = m.Vector3{1, 2, 3}
	c := a.Clone()
	if c == a || c.Position != a.Position || c.GetMass() != a.GetMass() {
		t.Fatal("Clone")
	}
}
