module github.com/cubez-amd/cubez

go 1.21

// The math types (Real, Vector3, Quat, Matrix3, Matrix3x4) are the
// reference's own package, so values pass between callers, the reference
// and this package unchanged.  The reference has no tagged release: pin
// the commit the caller builds against with
//   go get github.com/tbogdala/cubez@<commit>
// which writes its pseudo-version here; the tests also import the
// reference itself, to compare the two step by step.
require github.com/tbogdala/cubez v0.0.0
