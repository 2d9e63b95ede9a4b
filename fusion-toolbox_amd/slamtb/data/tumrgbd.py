"""TUM RGB-D trajectory files (the ``read_trajectory`` part of
slamtb/data/tumrgbd.py): one pose per line, ``timestamp tx ty tz qx qy qz
qw``, '#' comments."""
import numpy as np

from slamtb.camera import RTCamera, quaternion_to_matrix


def read_trajectory(gt_filepath, rotate=False):
    """{timestamp: RTCamera} (tumrgbd.py:54-86). ``rotate`` swaps the y and
    z axes of every pose, as the reference does."""
    traj = {}
    with open(str(gt_filepath), 'r') as stream:
        for line in stream:
            line = line.strip()
            if not line or line.startswith("#"):
                continue
            timestamp, tx, ty, tz, qx, qy, qz, qw = map(float, line.split())
            m = np.eye(4)
            m[:3, :3] = quaternion_to_matrix(qw, qx, qy, qz).numpy()
            m[:3, 3] = [tx, ty, tz]
            if rotate:
                swap = np.array([[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0],
                                 [0, 0, 0, 1]], dtype=m.dtype)
                m = swap @ m
            traj[timestamp] = RTCamera(m)
    return traj


def write_trajectory(filepath, trajectory):
    """Inverse of :func:`read_trajectory` (quaternion with w >= 0)."""
    with open(str(filepath), 'w') as stream:
        for timestamp, cam in trajectory.items():
            m = np.asarray(cam.matrix, dtype=np.float64)
            qw, qx, qy, qz = _matrix_to_quaternion(m[:3, :3])
            stream.write(f"{timestamp} {m[0, 3]:.9f} {m[1, 3]:.9f} {m[2, 3]:.9f} "
                         f"{qx:.9f} {qy:.9f} {qz:.9f} {qw:.9f}\n")


def _matrix_to_quaternion(R):
    """Shepperd's method: branch on the largest of w, x, y, z."""
    t = np.trace(R)
    cand = [t, R[0, 0], R[1, 1], R[2, 2]]
    k = int(np.argmax(cand))
    if k == 0:
        s = np.sqrt(1.0 + t) * 2
        q = [s / 4, (R[2, 1] - R[1, 2]) / s, (R[0, 2] - R[2, 0]) / s,
             (R[1, 0] - R[0, 1]) / s]
    elif k == 1:
        s = np.sqrt(1.0 + R[0, 0] - R[1, 1] - R[2, 2]) * 2
        q = [(R[2, 1] - R[1, 2]) / s, s / 4, (R[0, 1] + R[1, 0]) / s,
             (R[0, 2] + R[2, 0]) / s]
    elif k == 2:
        s = np.sqrt(1.0 + R[1, 1] - R[0, 0] - R[2, 2]) * 2
        q = [(R[0, 2] - R[2, 0]) / s, (R[0, 1] + R[1, 0]) / s, s / 4,
             (R[1, 2] + R[2, 1]) / s]
    else:
        s = np.sqrt(1.0 + R[2, 2] - R[0, 0] - R[1, 1]) * 2
        q = [(R[1, 0] - R[0, 1]) / s, (R[0, 2] + R[2, 0]) / s,
             (R[1, 2] + R[2, 1]) / s, s / 4]
    q = np.array(q)
    return q if q[0] >= 0 else -q
