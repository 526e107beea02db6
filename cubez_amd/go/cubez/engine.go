package cubez

import (
	"sync"
	"unsafe"

	m "github.com/tbogdala/cubez/math"
)

// Sync runs whatever the frame has recorded (the integrations) on the
// device and downloads the bodies' state into their Go values.  A caller
// that reads a body's fields right after Integrate, before its collision
// checks, calls it first.
func Sync() { theEngine().sync() }

type phase int

const (
	phaseIdle      phase = iota // the Go values are the device's
	phaseIntegrate              // integrations recorded, not yet run
	phaseCollide                // the frame's integrations ran; checks are answered
)

type pair struct{ one, two int32 }

// a plane by value: planes carry no state, so every CollisionPlane of the
// same normal and offset is one device plane (a caller may build a new one
// every frame, as examples/ballistic.go:49 does)
type planeKey [4]m.Real

// engine batches the calls of a frame into device launches (see the
// package documentation).
type engine struct {
	dev        *device
	bodies     []*RigidBody
	planes     []planeKey
	planeSlot  map[planeKey]int
	sceneDirty bool

	// the values the device holds per field, as of the last upload or
	// download: a Go value that differs was assigned by the caller
	shadow      map[int][]float64
	shadowFlags []uint8

	phase        phase
	integ        []bool
	dt           m.Real
	haveDT       bool
	deriveBefore bool // CalculateDerivedData before the frame's integrations
	deriveAfter  bool // ... after them

	calls       []pair // this frame's checks so far
	pred        []pair // the previous frame's checks: the batch list
	batch       bool   // the device holds the batch's contacts and every check so far matched it
	batchCounts []int32
	batchOff    []int
	batchRec    contactRecords
	handed      []*Contact // contacts the frame's checks appended, in order
	handedVals  []Contact  // their values when appended
}

var (
	engineOnce sync.Once
	engineInst *engine
)

func theEngine() *engine {
	engineOnce.Do(func() {
		engineInst = &engine{dev: newDevice(), shadow: map[int][]float64{}, planeSlot: map[planeKey]int{}}
	})
	return engineInst
}

// resetEngine forgets every body, collider and plane (a new empty scene on
// the same device context); the tests start each scene with it.
func resetEngine() {
	e := theEngine()
	dev := e.dev
	*e = engine{dev: dev, shadow: map[int][]float64{}, planeSlot: map[planeKey]int{}}
}

// --- registration -----------------------------------------------------------

func (e *engine) register(b *RigidBody) {
	b.slot = len(e.bodies)
	e.bodies = append(e.bodies, b)
	e.integ = append(e.integ, false)
	e.sceneDirty = true
}

func (e *engine) attach(b *RigidBody, c Collider) {
	b.collider = c
	e.sceneDirty = true
}

// planeIndex is the device plane of the plane's current normal and
// offset, added to the scene when new.
func (e *engine) planeIndex(p *CollisionPlane) int {
	k := planeKey{p.Normal[0], p.Normal[1], p.Normal[2], p.Offset}
	if s, ok := e.planeSlot[k]; ok {
		return s
	}
	s := len(e.planes)
	e.planes = append(e.planes, k)
	e.planeSlot[k] = s
	e.sceneDirty = true
	return s
}

// index of a collider in a device pair: body slot, or -(p+1) for plane p
func (e *engine) index(c Collider) int32 {
	if p, ok := c.(*CollisionPlane); ok {
		return -int32(e.planeIndex(p)) - 1
	}
	return int32(c.GetBody().slot)
}

// --- the frame --------------------------------------------------------------

func (e *engine) noteIntegrate(b *RigidBody, dt m.Real) {
	if e.phase == phaseCollide {
		e.endFrame() // a new frame without a ResolveContacts
	}
	if e.haveDT && dt != e.dt {
		panic("cubez: every body integrated in a frame must use the same duration")
	}
	e.phase = phaseIntegrate
	e.integ[b.slot] = true
	e.dt, e.haveDT = dt, true
}

func (e *engine) noteDerive(b *RigidBody) {
	if e.phase == phaseIntegrate && e.integ[b.slot] {
		e.deriveAfter = true
	} else {
		e.deriveBefore = true
	}
	if e.phase == phaseIdle {
		e.phase = phaseIntegrate
	}
}

// The collider transform of a body integrated in the frame comes with the
// integration; any other body's needs the derived-data pass.
func (e *engine) noteColliderDerive(b *RigidBody) {
	if e.phase == phaseIntegrate && e.integ[b.slot] {
		return
	}
	e.noteDerive(b)
}

func (e *engine) syncIfPending() {
	if e.phase == phaseIntegrate || e.sceneDirty {
		e.flushIntegration()
	}
}

func (e *engine) sync() { e.syncIfPending() }

func (e *engine) endFrame() {
	e.pred = e.calls
	e.calls = nil
	e.handed, e.handedVals = nil, nil
	e.batch = false
	e.phase = phaseIdle
	for i := range e.integ {
		e.integ[i] = false
	}
	e.haveDT = false
}

// flushIntegration uploads the caller's edits (or the whole scene when
// bodies, colliders or planes were added), runs the recorded derived-data
// passes and integrations, and downloads the result.
func (e *engine) flushIntegration() {
	if e.phase == phaseCollide {
		e.ensureScene()
		return
	}
	if e.sceneDirty {
		e.rebuild()
	} else {
		e.uploadEdits()
	}
	if e.deriveBefore {
		e.dev.calculateDerived()
	}
	integrating := false
	for _, v := range e.integ {
		integrating = integrating || v
	}
	if integrating {
		e.dev.integrate(float64(e.dt))
	}
	if e.deriveAfter {
		e.dev.calculateDerived()
	}
	e.deriveBefore, e.deriveAfter = false, false
	e.download()
	e.phase = phaseCollide
}

// ensureScene rebuilds the device scene after the frame's integrations
// when a check named a new plane: the Go values are current then, and the
// device's contact list is lost (the frame's later checks go one by one).
func (e *engine) ensureScene() {
	if e.sceneDirty {
		e.rebuild()
		e.batch = false
	}
}

// --- collision checks ---------------------------------------------------------

func (e *engine) check(one, two Collider, existing []*Contact) (bool, []*Contact) {
	e.flushIntegration()
	p := pair{e.index(one), e.index(two)}
	e.ensureScene()
	k := len(e.calls)
	e.calls = append(e.calls, p)
	if k == 0 && len(e.pred) > 0 {
		e.runBatch()
	}
	var rec contactRecords
	var lo, n int
	if e.batch && k < len(e.pred) && e.pred[k] == p {
		lo, n, rec = e.batchOff[k], int(e.batchCounts[k]), e.batchRec
	} else {
		// off the predicted list: this pair alone (the device's contact list
		// is no longer the batch's)
		e.batch = false
		counts := e.dev.narrowphasePairs([]int32{p.one}, []int32{p.two})
		n = int(counts[0])
		rec = e.dev.contacts(n)
	}
	for i := lo; i < lo+n; i++ {
		c := e.materialize(rec, i)
		existing = append(existing, c)
		e.handed = append(e.handed, c)
		e.handedVals = append(e.handedVals, *c)
	}
	return n > 0, existing
}

// runBatch runs the previous frame's checks as one narrowphase.
func (e *engine) runBatch() {
	one := make([]int32, len(e.pred))
	two := make([]int32, len(e.pred))
	for i, p := range e.pred {
		one[i], two[i] = p.one, p.two
	}
	e.batchCounts = e.dev.narrowphasePairs(one, two)
	e.batchOff = make([]int, len(e.pred))
	total := 0
	for i, c := range e.batchCounts {
		e.batchOff[i] = total
		total += int(c)
	}
	e.batchRec = e.dev.contacts(total)
	e.batch = true
}

func (e *engine) materialize(r contactRecords, i int) *Contact {
	c := &Contact{Friction: m.Real(r.fric[i]), Restitution: m.Real(r.rest[i]), Penetration: m.Real(r.pen[i])}
	c.Bodies[0] = e.bodies[r.body0[i]]
	if r.body1[i] >= 0 {
		c.Bodies[1] = e.bodies[r.body1[i]]
	}
	for d := 0; d < 3; d++ {
		c.ContactPoint[d] = m.Real(r.point[3*i+d])
		c.ContactNormal[d] = m.Real(r.normal[3*i+d])
	}
	return c
}

// --- resolve --------------------------------------------------------------------

func (e *engine) resolve(maxIterations int, contacts []*Contact, dt m.Real) {
	e.flushIntegration()
	if !e.deviceHolds(contacts) {
		e.dev.setContacts(e.records(contacts))
	}
	e.dev.resolve(float64(dt), maxIterations)
	e.download()
	pen := e.dev.penetrations(len(contacts))
	for i, c := range contacts {
		c.Penetration = m.Real(pen[i])
	}
	e.endFrame()
}

// deviceHolds: the caller resolves exactly the contacts the frame's batch
// appended, in order and unmodified, so the device's list is theirs.
func (e *engine) deviceHolds(contacts []*Contact) bool {
	if !e.batch || len(e.calls) != len(e.pred) || len(contacts) != len(e.handed) {
		return false
	}
	for i, c := range contacts {
		if c != e.handed[i] || *c != e.handedVals[i] {
			return false
		}
	}
	return true
}

func (e *engine) records(contacts []*Contact) contactRecords {
	n := len(contacts)
	r := contactRecords{
		body0: make([]int32, n), body1: make([]int32, n), point: make([]float64, 3*n),
		normal: make([]float64, 3*n), pen: make([]float64, n), fric: make([]float64, n), rest: make([]float64, n),
	}
	for i, c := range contacts {
		r.body0[i] = int32(c.Bodies[0].slot)
		r.body1[i] = -1
		if c.Bodies[1] != nil {
			r.body1[i] = int32(c.Bodies[1].slot)
		}
		for d := 0; d < 3; d++ {
			r.point[3*i+d] = float64(c.ContactPoint[d])
			r.normal[3*i+d] = float64(c.ContactNormal[d])
		}
		r.pen[i] = float64(c.Penetration)
		r.fric[i] = float64(c.Friction)
		r.rest[i] = float64(c.Restitution)
	}
	return r
}

// --- host <-> device fields ----------------------------------------------------

// The fields a caller may assign (directly or through a setter), and their
// widths: compared with the shadow and uploaded when changed.
var editable = []struct{ id, width int }{
	{fHalf, 3}, {fPos, 3}, {fQuat, 4}, {fVel, 3}, {fRot, 3}, {fAcc, 3}, {fInvMass, 1},
	{fIITBody, 9}, {fOffset, 12}, {fMotion, 1}, {fLinDamp, 1}, {fAngDamp, 1},
}

// The fields the device writes, downloaded after every launch that
// changes the bodies.
var written = []struct{ id, width int }{
	{fPos, 3}, {fQuat, 4}, {fVel, 3}, {fRot, 3}, {fIITWorld, 9}, {fTransform, 12}, {fCTransform, 12},
	{fLastAcc, 3}, {fMotion, 1},
}

func (e *engine) gather(id, width int) []float64 {
	out := make([]float64, width*len(e.bodies))
	for i, b := range e.bodies {
		v := out[width*i : width*(i+1)]
		switch id {
		case fHalf:
			switch c := b.collider.(type) {
			case *CollisionCube:
				v[0], v[1], v[2] = float64(c.HalfSize[0]), float64(c.HalfSize[1]), float64(c.HalfSize[2])
			case *CollisionSphere:
				v[0] = float64(c.Radius)
			}
		case fPos:
			copyReal(v, b.Position[:])
		case fQuat:
			copyReal(v, b.Orientation[:])
		case fVel:
			copyReal(v, b.Velocity[:])
		case fRot:
			copyReal(v, b.Rotation[:])
		case fAcc:
			copyReal(v, b.Acceleration[:])
		case fInvMass:
			v[0] = float64(b.inverseMass)
		case fIITBody:
			copyReal(v, b.InverseInertiaTensor[:])
		case fIITWorld:
			copyReal(v, b.inverseInertiaTensorWorld[:])
		case fTransform:
			copyReal(v, b.transform[:])
		case fCTransform:
			if t := colliderTransform(b); t != nil {
				copyReal(v, t[:])
			}
		case fOffset:
			if o := colliderOffset(b); o != nil {
				copyReal(v, o[:])
			} else {
				copyReal(v, identity3x4[:])
			}
		case fLastAcc:
			copyReal(v, b.lastFrameAccelleration[:])
		case fMotion:
			v[0] = float64(b.motion)
		case fLinDamp:
			v[0] = float64(b.LinearDamping)
		case fAngDamp:
			v[0] = float64(b.AngularDamping)
		}
	}
	return out
}

func (e *engine) scatter(id, width int, in []float64) {
	for i, b := range e.bodies {
		v := in[width*i : width*(i+1)]
		switch id {
		case fPos:
			copyToReal(b.Position[:], v)
		case fQuat:
			copyToReal(b.Orientation[:], v)
		case fVel:
			copyToReal(b.Velocity[:], v)
		case fRot:
			copyToReal(b.Rotation[:], v)
		case fIITWorld:
			copyToReal(b.inverseInertiaTensorWorld[:], v)
		case fTransform:
			copyToReal(b.transform[:], v)
		case fCTransform:
			if t := colliderTransform(b); t != nil {
				copyToReal(t[:], v)
			}
		case fLastAcc:
			copyToReal(b.lastFrameAccelleration[:], v)
		case fMotion:
			b.motion = m.Real(v[0])
		}
	}
}

func (e *engine) flags() []uint8 {
	out := make([]uint8, len(e.bodies))
	for i, b := range e.bodies {
		var f uint8
		if b.IsAwake {
			f |= flagAwake
		}
		if b.CanSleep {
			f |= flagCanSleep
		}
		if !e.integ[i] {
			f |= flagStatic // not integrated in this frame
		}
		out[i] = f
	}
	return out
}

func (e *engine) shapes() []int32 {
	out := make([]int32, len(e.bodies))
	for i, b := range e.bodies {
		out[i] = shapeCube
		if _, ok := b.collider.(*CollisionSphere); ok {
			out[i] = shapeSphere
		}
	}
	return out
}

func (e *engine) rebuild() {
	normals := make([]float64, 3*len(e.planes))
	offsets := make([]float64, len(e.planes))
	for i, p := range e.planes {
		copyReal(normals[3*i:3*i+3], p[:3])
		offsets[i] = float64(p[3])
	}
	e.dev.setScene(len(e.bodies), normals, offsets)
	if len(e.bodies) > 0 {
		sh := e.shapes()
		e.dev.setField(fShape, unsafe.Pointer(&sh[0]))
		all := append(append([]struct{ id, width int }{}, editable...), written...)
		for _, f := range all {
			v := e.gather(f.id, f.width)
			e.dev.setField(f.id, unsafe.Pointer(&v[0]))
			e.shadow[f.id] = v
		}
		fl := e.flags()
		e.dev.setField(fFlags, unsafe.Pointer(&fl[0]))
		e.shadowFlags = fl
	}
	e.sceneDirty = false
}

func (e *engine) uploadEdits() {
	if len(e.bodies) == 0 {
		return
	}
	for _, f := range editable {
		v := e.gather(f.id, f.width)
		if !equalFloats(v, e.shadow[f.id]) {
			e.dev.setField(f.id, unsafe.Pointer(&v[0]))
			e.shadow[f.id] = v
		}
	}
	fl := e.flags()
	if string(fl) != string(e.shadowFlags) {
		e.dev.setField(fFlags, unsafe.Pointer(&fl[0]))
		e.shadowFlags = fl
	}
}

func (e *engine) download() {
	if len(e.bodies) == 0 {
		return
	}
	for _, f := range written {
		v := make([]float64, f.width*len(e.bodies))
		e.dev.getField(f.id, unsafe.Pointer(&v[0]))
		e.scatter(f.id, f.width, v)
		e.shadow[f.id] = v
	}
	fl := make([]uint8, len(e.bodies))
	e.dev.getField(fFlags, unsafe.Pointer(&fl[0]))
	for i, b := range e.bodies {
		b.IsAwake = fl[i]&flagAwake != 0
	}
	e.shadowFlags = fl
}

// --- helpers -----------------------------------------------------------------

func colliderTransform(b *RigidBody) *m.Matrix3x4 {
	switch c := b.collider.(type) {
	case *CollisionCube:
		return &c.transform
	case *CollisionSphere:
		return &c.transform
	}
	return nil
}

func colliderOffset(b *RigidBody) *m.Matrix3x4 {
	switch c := b.collider.(type) {
	case *CollisionCube:
		return &c.Offset
	case *CollisionSphere:
		return &c.Offset
	}
	return nil
}

func copyReal(dst []float64, src []m.Real) {
	for i := range src {
		dst[i] = float64(src[i])
	}
}

func copyToReal(dst []m.Real, src []float64) {
	for i := range dst {
		dst[i] = m.Real(src[i])
	}
}

func equalFloats(a, b []float64) bool {
	if len(a) != len(b) {
		return false
	}
	for i := range a {
		// bitwise: a NaN the caller left in place is not an edit
		if a[i] != b[i] && !(a[i] != a[i] && b[i] != b[i]) {
			return false
		}
	}
	return true
}
