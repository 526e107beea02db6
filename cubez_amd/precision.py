"""Error of the FP32 variant against the FP64 path.

The FP32 variant stores and resolves in float32, evaluates the narrowphase
(SAT penetration and overlap tests) in FP64 and builds the contact basis in
FP64.  ``compare`` runs both paths from the same scene, in lockstep on the
GPU, and reports the deviation at the requested step counts.
"""
import numpy as np

from .world import World


def _quat_angle(q1, q2):
    """Rotation angle between unit quaternions (rad), per body, as
    4 atan2(|q1 - q2|, |q1 + q2|) with q2's sign matched to q1 (arccos of the
    dot product loses half the digits near zero angle)."""
    q1 = np.asarray(q1, np.float64)
    q2 = np.asarray(q2, np.float64)
    s = np.where(np.sum(q1 * q2, axis=1) < 0, -1.0, 1.0)[:, None]
    q2 = q2 * s
    return 4.0 * np.arctan2(np.linalg.norm(q1 - q2, axis=1), np.linalg.norm(q1 + q2, axis=1))


def compare(scene, checkpoints=(1, 60, 600), iter_per_contact=None, per_step=False):
    """Run the FP64 and FP32 paths side by side.

    Returns a list of dicts, one per checkpoint (and, with per_step=True, a
    per-step list of max position errors), with max/mean position error (m),
    max orientation error (rad), max velocity error (m/s) and the number of
    worlds whose contact count differs at that step."""
    mult = iter_per_contact or scene.iter_mult
    out, steps_err = [], []
    with World(scene, precision="f64") as a, World(scene, precision="f32") as b:
        done = 0
        for cp in sorted(checkpoints):
            while done < cp:
                k = 1 if per_step else cp - done
                ca = a.step(scene.dt, k, mult, counts=True)
                cb = b.step(scene.dt, k, mult, counts=True)
                done += k
                if per_step:
                    steps_err.append(float(np.max(np.linalg.norm(a.get("pos") - b.get("pos"), axis=1))))
            pa, pb = a.get("pos"), b.get("pos")
            dp = np.linalg.norm(pa - pb, axis=1)
            out.append({
                "step": cp,
                "pos_err_max": float(dp.max()),
                "pos_err_mean": float(dp.mean()),
                "quat_err_max": float(_quat_angle(a.get("quat"), b.get("quat")).max()),
                "vel_err_max": float(np.linalg.norm(a.get("vel") - b.get("vel"), axis=1).max()),
                "worlds_count_differs": int(np.sum(ca[-1] != cb[-1])),
                "contacts_f64": int(ca[-1].sum()),
                "contacts_f32": int(cb[-1].sum()),
            })
    if per_step:
        return out, steps_err
    return out
