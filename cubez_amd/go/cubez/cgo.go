package cubez

/*
#cgo CFLAGS: -I${SRCDIR}/../../../include
#cgo LDFLAGS: -L${SRCDIR}/../.. -lcubez_amd -Wl,-rpath,${SRCDIR}/../..
#include <stdlib.h>
#include "cubez_amd.h"
*/
import "C"

import (
	"fmt"
	"unsafe"
)

// device is one libcubez_amd context: one world holding every registered
// body and plane.
type device struct {
	ctx *C.cbz_ctx
}

func newDevice() *device {
	cfg := C.cbz_config{precision: C.CBZ_F64, device: 0, contact_capacity: 0}
	var ctx *C.cbz_ctx
	if rc := C.cbz_create(&cfg, &ctx); rc != C.CBZ_OK {
		panic(fmt.Sprintf("cubez: cbz_create failed (%d): no HIP device or libcubez_amd missing", int(rc)))
	}
	return &device{ctx: ctx}
}

// check panics with the library's message on a failed call: the reference
// has no error returns, and its failures are panics as well.
func (d *device) check(rc C.int, what string) {
	if rc != C.CBZ_OK {
		panic(fmt.Sprintf("cubez: %s: %s", what, C.GoString(C.cbz_last_error(d.ctx))))
	}
}

func (d *device) setScene(nBodies int, planeNormal []float64, planeOffset []float64) {
	bodyOff := []C.int32_t{0, C.int32_t(nBodies)}
	planeOff := []C.int32_t{0, C.int32_t(len(planeOffset))}
	var pn, po *C.double
	if len(planeOffset) > 0 {
		pn = (*C.double)(unsafe.Pointer(&planeNormal[0]))
		po = (*C.double)(unsafe.Pointer(&planeOffset[0]))
	}
	d.check(C.cbz_set_scene(d.ctx, 1, &bodyOff[0], &planeOff[0], pn, po), "cbz_set_scene")
}

func (d *device) setField(field int, p unsafe.Pointer) {
	d.check(C.cbz_set_field(d.ctx, C.int32_t(field), p), "cbz_set_field")
}

func (d *device) getField(field int, p unsafe.Pointer) {
	d.check(C.cbz_get_field(d.ctx, C.int32_t(field), p), "cbz_get_field")
}

func (d *device) calculateDerived() { d.check(C.cbz_calculate_derived(d.ctx), "cbz_calculate_derived") }

func (d *device) integrate(dt float64) { d.check(C.cbz_integrate(d.ctx, C.double(dt)), "cbz_integrate") }

// narrowphasePairs runs CheckForCollisions for the ordered pairs (body
// index, or -(p+1) for plane p) and returns each pair's contact count.
func (d *device) narrowphasePairs(one, two []int32) []int32 {
	n := len(one)
	counts := make([]int32, n)
	if n == 0 {
		return counts
	}
	var total C.int64_t
	d.check(C.cbz_narrowphase_pairs(d.ctx, C.int64_t(n), (*C.int32_t)(unsafe.Pointer(&one[0])),
		(*C.int32_t)(unsafe.Pointer(&two[0])), &total), "cbz_narrowphase_pairs")
	d.check(C.cbz_pair_contact_counts(d.ctx, C.int64_t(n), (*C.int32_t)(unsafe.Pointer(&counts[0]))),
		"cbz_pair_contact_counts")
	return counts
}

// contactRecords is the host copy of the device's contact list.
type contactRecords struct {
	body0, body1    []int32
	point, normal   []float64
	pen, fric, rest []float64
}

func (d *device) contacts(n int) contactRecords {
	r := contactRecords{
		body0: make([]int32, n), body1: make([]int32, n), point: make([]float64, 3*n),
		normal: make([]float64, 3*n), pen: make([]float64, n), fric: make([]float64, n), rest: make([]float64, n),
	}
	if n == 0 {
		return r
	}
	v := C.cbz_contact_view{
		body0:       (*C.int32_t)(unsafe.Pointer(&r.body0[0])),
		body1:       (*C.int32_t)(unsafe.Pointer(&r.body1[0])),
		point:       (*C.double)(unsafe.Pointer(&r.point[0])),
		normal:      (*C.double)(unsafe.Pointer(&r.normal[0])),
		pen:         (*C.double)(unsafe.Pointer(&r.pen[0])),
		friction:    (*C.double)(unsafe.Pointer(&r.fric[0])),
		restitution: (*C.double)(unsafe.Pointer(&r.rest[0])),
	}
	d.check(C.cbz_get_contacts(d.ctx, 0, &v), "cbz_get_contacts")
	return r
}

func (d *device) penetrations(n int) []float64 {
	pen := make([]float64, n)
	if n == 0 {
		return pen
	}
	v := C.cbz_contact_view{pen: (*C.double)(unsafe.Pointer(&pen[0]))}
	d.check(C.cbz_get_contacts(d.ctx, 0, &v), "cbz_get_contacts")
	return pen
}

// setContacts replaces the device's contact list with the caller's.
func (d *device) setContacts(r contactRecords) {
	n := len(r.body0)
	counts := []C.int64_t{C.int64_t(n)}
	if n == 0 {
		d.check(C.cbz_set_contacts(d.ctx, &counts[0], nil, nil, nil, nil, nil), "cbz_set_contacts")
		return
	}
	d.check(C.cbz_set_contacts(d.ctx, &counts[0], (*C.int32_t)(unsafe.Pointer(&r.body0[0])),
		(*C.int32_t)(unsafe.Pointer(&r.body1[0])), (*C.double)(unsafe.Pointer(&r.point[0])),
		(*C.double)(unsafe.Pointer(&r.normal[0])), (*C.double)(unsafe.Pointer(&r.pen[0]))), "cbz_set_contacts")
	d.setMaterials(r)
}

// setMaterials gives the device the contacts' Friction and Restitution: the
// plain resolver when every contact has the reference's 0.9 and 0.1
// (colliders.go:201-202), per-contact values otherwise.
func (d *device) setMaterials(r contactRecords) {
	plain := true
	for k := range r.fric {
		plain = plain && r.fric[k] == 0.9 && r.rest[k] == 0.1
	}
	if plain {
		d.check(C.cbz_set_materials(d.ctx, 0), "cbz_set_materials")
		return
	}
	d.check(C.cbz_set_materials(d.ctx, C.CBZ_MATERIALS), "cbz_set_materials")
	d.check(C.cbz_set_contact_materials(d.ctx, (*C.double)(unsafe.Pointer(&r.fric[0])),
		(*C.double)(unsafe.Pointer(&r.rest[0]))), "cbz_set_contact_materials")
}

// resolve is ResolveContacts(maxIterations, the device's contacts, dt).
func (d *device) resolve(dt float64, maxIterations int) {
	d.check(C.cbz_resolve(d.ctx, C.double(dt), 0, C.int64_t(maxIterations)), "cbz_resolve")
}

// field ids (include/cubez_amd.h)
const (
	fShape      = int(C.CBZ_SHAPE)
	fHalf       = int(C.CBZ_HALF)
	fPos        = int(C.CBZ_POS)
	fQuat       = int(C.CBZ_QUAT)
	fVel        = int(C.CBZ_VEL)
	fRot        = int(C.CBZ_ROT)
	fAcc        = int(C.CBZ_ACC)
	fInvMass    = int(C.CBZ_INV_MASS)
	fIITBody    = int(C.CBZ_IIT_BODY)
	fIITWorld   = int(C.CBZ_IIT_WORLD)
	fTransform  = int(C.CBZ_TRANSFORM)
	fCTransform = int(C.CBZ_CTRANSFORM)
	fOffset     = int(C.CBZ_OFFSET)
	fLastAcc    = int(C.CBZ_LAST_ACC)
	fMotion     = int(C.CBZ_MOTION)
	fLinDamp    = int(C.CBZ_LIN_DAMP)
	fAngDamp    = int(C.CBZ_ANG_DAMP)
	fFlags      = int(C.CBZ_FLAGS)

	shapeCube   = int32(C.CBZ_CUBE)
	shapeSphere = int32(C.CBZ_SPHERE)

	flagAwake    = uint8(C.CBZ_FLAG_AWAKE)
	flagCanSleep = uint8(C.CBZ_FLAG_CAN_SLEEP)
	flagStatic   = uint8(C.CBZ_FLAG_STATIC)
)
