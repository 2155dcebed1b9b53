"""numpy view of the reference's checkpoint records, written from the struct
declarations in src/CUDA/IBM/structs/particleCenter.h:15-118 and
src/CUDA/structs/globalStructs.h (dfloat3 = x, y, z; dfloat4 = x, y, z, w;
dfloat6 = xx, yy, zz, xy, xz, yz), with the natural alignment of the host
compiler (numpy align=True follows the same C rules)."""
import numpy as np

MAX_ACTIVE_COLLISIONS = 13   # ibmBoundaryCondition.h:78


def particle_center_dtype(t) -> np.dtype:
    """ParticleCenter with dfloat = t (np.float64 or np.float32)."""
    collision = np.dtype([
        ("shape", np.int32),                                  # enum ParticleType, SPHERE = 0
        ("semiAxis", t, (3, 3)),                              # semiAxis, semiAxis2, semiAxis3
        ("partner", np.int32, (MAX_ACTIVE_COLLISIONS,)),      # collisionPartnerIDs
        ("tang", t, (MAX_ACTIVE_COLLISIONS, 3)),              # tangentialDisplacements
        ("last", np.int32, (MAX_ACTIVE_COLLISIONS,)),         # lastCollisionStep
    ], align=True)
    v3 = (t, (3,))
    return np.dtype([
        ("pos",) + v3, ("pos_old",) + v3, ("vel",) + v3, ("vel_old",) + v3,
        ("w",) + v3, ("w_avg",) + v3, ("w_old",) + v3, ("w_pos",) + v3,
        ("q_pos", t, (4,)), ("q_pos_old", t, (4,)),
        ("f",) + v3, ("f_old",) + v3, ("M",) + v3, ("M_old",) + v3,
        ("I", t, (6,)),
        ("dP",) + v3, ("dL",) + v3,
        ("S", t), ("radius", t), ("volume", t), ("density", t),
        ("movable", np.bool_),
        ("collision", collision),
    ], align=True)
