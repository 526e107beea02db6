"""Planted mistakes: each mutant is the oracle with one plausible
transcription error (a term, sign, index, operand order or tie rule changed),
compiled to a temporary library.  The oracle's own checks (known answers,
closed forms, brute force) must fail on every one of them; a mutant they
pass would be a part of the restatement they do not pin."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "cbz_oracle.c")
CFLAGS = ["-O2", "-std=c11", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-shared"]

# name -> (original line fragment, planted mistake)
MUTANTS = {
    "cross_product_term": ("r.v[1] = v->v[2] * v2->v[0] - v->v[0] * v2->v[2];",
                           "r.v[1] = v->v[0] * v2->v[2] - v->v[2] * v2->v[0];"),
    "position_tie_rule": ("if (contacts[i].Penetration > max) {", "if (contacts[i].Penetration >= max) {"),
    "velocity_tie_rule": ("if (contacts[i].desiredDeltaVelocity > max) {",
                          "if (contacts[i].desiredDeltaVelocity >= max) {"),
    "velocity_update_sign": ("if (b == 1) sign = -1.0;", "if (b == 0) sign = -1.0;"),
    "position_update_sign": ("if (b == 0) sign = -1.0;", "if (b == 1) sign = -1.0;"),
    "basis_not_transposed": ("Vector3 contactVelocity = m3_transform_transpose(&c->contactToWorld, &velocity);",
                             "Vector3 contactVelocity = m3_mul_vec(&c->contactToWorld, &velocity);"),
    "sat_tie_rule": ("if (penetration < *smallestPenetration) {", "if (penetration <= *smallestPenetration) {"),
    "restitution_limit": ("const Real velocityLimit = 0.25;", "const Real velocityLimit = 0.0;"),
    "sleep_average_weights": ("b->motion = bias * b->motion + (1.0 - bias) * currentMotion;",
                              "b->motion = (1.0 - bias) * b->motion + bias * currentMotion;"),
    "relative_velocity_operands": (
        "Vector3 velocity = v3_cross(&body->Rotation, &c->relativeContactPosition[bodyIndex]);",
        "Vector3 velocity = v3_cross(&c->relativeContactPosition[bodyIndex], &body->Rotation);"),
    "frictionless_linear_term": ("deltaVelocity += c->Bodies[0]->inverseMass;", "deltaVelocity += 0.0;"),
    "frictionless_angular_operands": ("dvw = v3_cross(&dvw, &c->relativeContactPosition[0]);",
                                      "dvw = v3_cross(&c->relativeContactPosition[0], &dvw);"),
    "friction_cone_coefficient": ("ic.v[2] *= c->Friction * ic.v[0];", "ic.v[2] *= ic.v[0];"),
    # box-box edge-edge contact (colliders.go:642-709)
    "edge_edge_normal_flip": ("if (v3_dot(&axis, &toCenter) > 0) v3_mul_with(&axis, -1.0);",
                              "if (v3_dot(&axis, &toCenter) < 0) v3_mul_with(&axis, -1.0);"),
    "edge_point_one_sign": ("if (v3_dot(&a, &axis) > 0) ptOnOneEdge.v[i] = -ptOnOneEdge.v[i];",
                            "if (v3_dot(&a, &axis) < 0) ptOnOneEdge.v[i] = -ptOnOneEdge.v[i];"),
    "edge_point_two_sign": ("if (v3_dot(&a, &axis) < 0) ptOnTwoEdge.v[i] = -ptOnTwoEdge.v[i];",
                            "if (v3_dot(&a, &axis) > 0) ptOnTwoEdge.v[i] = -ptOnTwoEdge.v[i];"),
    # the SAT's near-parallel threshold (colliders.go:460-462), either side
    "sat_parallel_threshold_high": ("if (v3_square_magnitude(&axis) < Epsilon) {",
                                    "if (v3_square_magnitude(&axis) < 1e-3) {"),
    "sat_parallel_threshold_low": ("if (v3_square_magnitude(&axis) < Epsilon) {",
                                   "if (v3_square_magnitude(&axis) < 1e-8) {"),
    # applyPositionChange (contact.go:286-386)
    "angular_limit": ("const Real angularLimit = 0.2;", "const Real angularLimit = 0.02;"),
    "angular_change_scale": ("v3_mul_with(&angularChange[i], angularMove[i] / angularInertia[i]);",
                             "v3_mul_with(&angularChange[i], angularMove[i]);"),
    "rederive_awake_bodies": ("if (body->IsAwake == 0) rb_calculate_derived(body);",
                              "rb_calculate_derived(body);"),
    # calculateDesiredDeltaVelocity: body one's acceleration term (contact.go:98-101)
    "velocity_from_acc_body_one_sign": ("        velocityFromAcc -= v3_dot(&tempVelocity, &c->ContactNormal);",
                                        "        velocityFromAcc += v3_dot(&tempVelocity, &c->ContactNormal);"),
    # matchAwakeState (contact.go:195-200)
    "wake_the_wrong_body": ("            rb_set_awake(c->Bodies[1], 1);\n        else\n            rb_set_awake(c->Bodies[0], 1);",
                            "            rb_set_awake(c->Bodies[0], 1);\n        else\n            rb_set_awake(c->Bodies[1], 1);"),
}
# Not planted: inverting useOne (colliders.go:687-690).  useOne only selects
# the point that contactPoint returns when the two edges are parallel or
# their closest points fall outside an edge (:538-563).  When an edge-edge
# axis is the minimal SAT axis, it is the Minkowski-difference face nearest
# the origin, whose foot point lies inside that face, i.e. the closest points
# lie within both edges; and an edge pair parallel enough for the first
# fallback has its axis skipped by tryAxis.  Random and constructed searches
# over 70,000 overlapping pairs (rods, plates, edge-on-face) reached neither
# fallback, so no input distinguishes the mutant.


@pytest.mark.parametrize("name", sorted(MUTANTS))
def test_oracle_checks_catch_planted_mistake(name, tmp_path):
    old, new = MUTANTS[name]
    src = open(SRC).read()
    assert src.count(old) == 1, "fragment not found once in the oracle: " + old
    mut = tmp_path / "mutant.c"
    mut.write_text(src.replace(old, new))
    lib = tmp_path / "libmutant.so"
    subprocess.run(["gcc", *CFLAGS, "-I", os.path.dirname(SRC), "-o", str(lib), str(mut), "-lm"], check=True)
    env = dict(os.environ, CBZ_ORACLE_LIB=str(lib))
    r = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider", "-m", "not gpu",
                        os.path.join(ROOT, "tests", "test_oracle_math.py"),
                        os.path.join(ROOT, "tests", "test_oracle_physics.py"),
                        os.path.join(ROOT, "tests", "test_oracle_pins.py"),
                        os.path.join(ROOT, "tests", "test_oracle_materials.py")],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=600)
    assert r.returncode != 0, f"mutant {name} passed every oracle check:\n{r.stdout[-2000:]}"
