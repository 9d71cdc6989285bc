"""Ground-truth pose of a synthetic pair and how far a found model is from it.

A synthetic query B is a sub-chain of A moved by p_B = Rg (a - c) + c + tg
(a an A-frame position, c the sub-chain centroid).  A search model
(angle, translation) says that R p_B + t lands on A, with R =
euler2rot(angle) and t = convert_trans(...) (VESPER_1_prob_new.py:1549-1566).
The distance between the model's map of B and the truth is measured as the
RMSD over a 20 A cube of points around the sub-chain centroid."""
import numpy as np

from oracle import vesper_ref as ref


def on_grid_pose(angle):
    """Rg whose inverse is the rotation of Euler row `angle`: the search's
    own row then is the exact answer."""
    return ref.euler2rot(np.asarray(angle, np.float64)).as_matrix().T


def pose_rmsd(pair, angle, new_trans, rot=None):
    """RMSD of the model (euler2rot(angle), or the 3x3 `rot`, then
    new_trans) from the true placement of B on A."""
    c = np.asarray(pair.centroid, np.float64)
    g = np.linspace(-10.0, 10.0, 5)
    a = c + np.stack(np.meshgrid(g, g, g, indexing="ij"), axis=-1).reshape(-1, 3)
    pb = (a - c) @ np.asarray(pair.rotation).T + c + np.asarray(pair.translation)
    r = ref.euler2rot(np.asarray(angle, np.float64)).as_matrix() if rot is None else np.asarray(rot)
    found = pb @ r.T + np.asarray(new_trans, np.float64)
    return float(np.sqrt(np.mean(np.sum((found - a) ** 2, axis=1))))


# (seed, on-grid Euler row of the true pose, translation in A)
POSES = [(77, (30, 60, 90), (7.0, -6.5, 8.0)), (78, (120, 210, 150), (-8.0, 6.0, 7.0)),
         (79, (300, 30, 60), (6.0, 8.0, -7.0)), (80, (90, 330, 120), (-6.0, -7.0, -8.0))]
SMALL = dict(a_edge=48, natoms=150, b_edge=32, nsub=60)
