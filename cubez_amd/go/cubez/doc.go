// Package cubez is a drop-in for the simulation step of
// github.com/tbogdala/cubez that runs it on an MI355X through the C ABI of
// libcubez_amd (include/cubez_amd.h).
//
// It keeps the reference's package API: RigidBody with the same exported
// fields and methods (rigidbody.go), CollisionPlane, CollisionCube and
// CollisionSphere with the Collider interface (colliders.go:15-23),
// CheckForCollisions (colliders.go:720-747), Contact and ResolveContacts
// (contact.go:17-51, :208-222).  A caller changes its import of
// "github.com/tbogdala/cubez" to this package; its imports of
// "github.com/tbogdala/cubez/math" stay as they are.
//
// # How the calls reach the GPU: a lazy flush
//
// The reference is synchronous and per object: Integrate per body,
// CheckForCollisions per pair, ResolveContacts per frame.  One GPU launch
// per body or per pair would cost more than the arithmetic, so the calls of
// a frame are batched:
//
//   - Integrate records the body and the duration.  The device integrates
//     every recorded body at the frame's first CheckForCollisions (or
//     ResolveContacts, or Sync), after uploading the fields the caller
//     assigned since the last download (detected by comparison with the
//     last downloaded values).  Reading a body's fields between Integrate
//     and that point gives the values before the integration; call Sync to
//     have them now.
//   - CheckForCollisions answers from one batched narrowphase per frame
//     (cbz_narrowphase_pairs over the pair list the caller's previous frame
//     made, cbz_pair_contact_counts to split the contacts back into calls).
//     While the caller's calls follow that list, each call returns the
//     reference's bool and appends exactly the contacts the reference
//     would; from the first call that differs, and in the first frame, the
//     calls are answered one pair at a time, and the frame's list becomes
//     the next frame's prediction.
//   - ResolveContacts resolves on the device, uploading the caller's
//     contacts first unless they are the batch's own list in order, then
//     downloads every body's state and the contacts' Penetration into the
//     Go values.
//
// The arithmetic is the reference's, in FP64 (m.Real is float64): the
// library's exact path equals a statement-by-statement restatement of the
// reference bit for bit (its tests).
//
// # Differences
//
//   - CalculateDerivedData on a body the caller moved recomputes the derived
//     data of every body at the next flush (the device has no per-body
//     variant); Integrate includes it, as in the reference.
//   - Every body integrated in a frame must use the same duration.
//   - Contact's unexported resolver fields are the device's; the exported
//     ones are materialised when a call appends the contact and after
//     ResolveContacts.
//   - Not safe for concurrent use, like the reference.
package cubez
