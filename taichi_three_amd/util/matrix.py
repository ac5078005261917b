"""Camera and affine helpers (tina/util/matrix.py:4-104), NumPy FP64.

lookat / perspective / frustum / ortho follow the reference's formulas; the
quaternion rotation step (RotationStep) is written from the textbook
quaternion product instead of the `transformations` package the reference
imports, which is not available here.
"""
import numpy as np

__all__ = ['camera_matrices', 'identity', 'affine', 'lookat', 'ortho', 'frustum', 'orthogonal', 'perspective',
           'scale', 'translate', 'quaternion', 'eularXYZ', 'RotationStep']


def identity():
    return np.eye(4)


def affine(lin, pos):
    """4x4 from a 3x3 linear part and a translation (matrix.py:8-12)."""
    lin = np.concatenate([np.asarray(lin, dtype=float), np.zeros((1, 3))], axis=0)
    pos = np.concatenate([np.asarray(pos, dtype=float), np.ones(1)])
    return np.concatenate([lin, pos[:, None]], axis=1)


def _quat_from_matrix(R):
    """Unit quaternion (w, x, y, z) of the rotation part of R (Shepperd's method)."""
    m = np.asarray(R, dtype=float)[:3, :3]
    t = np.trace(m)
    if t > 0:
        s = np.sqrt(t + 1.0) * 2
        q = [0.25 * s, (m[2, 1] - m[1, 2]) / s, (m[0, 2] - m[2, 0]) / s, (m[1, 0] - m[0, 1]) / s]
    elif m[0, 0] > m[1, 1] and m[0, 0] > m[2, 2]:
        s = np.sqrt(1.0 + m[0, 0] - m[1, 1] - m[2, 2]) * 2
        q = [(m[2, 1] - m[1, 2]) / s, 0.25 * s, (m[0, 1] + m[1, 0]) / s, (m[0, 2] + m[2, 0]) / s]
    elif m[1, 1] > m[2, 2]:
        s = np.sqrt(1.0 + m[1, 1] - m[0, 0] - m[2, 2]) * 2
        q = [(m[0, 2] - m[2, 0]) / s, (m[0, 1] + m[1, 0]) / s, 0.25 * s, (m[1, 2] + m[2, 1]) / s]
    else:
        s = np.sqrt(1.0 + m[2, 2] - m[0, 0] - m[1, 1]) * 2
        q = [(m[1, 0] - m[0, 1]) / s, (m[0, 2] + m[2, 0]) / s, (m[1, 2] + m[2, 1]) / s, 0.25 * s]
    return np.array(q)


def _quat_mul(a, b):
    w1, x1, y1, z1 = a
    w2, x2, y2, z2 = b
    return np.array([w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2,
                     w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2,
                     w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2,
                     w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2])


def _quat_matrix(q):
    w, x, y, z = q / np.linalg.norm(q)
    R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                  [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                  [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
    return affine(R, np.zeros(3))


def RotationStep(R, wx, wy, wz):
    """q += q * (0, w)/2, renormalised (matrix.py:14-19)."""
    q = _quat_from_matrix(R)
    qw = 0.5 * np.array([0, wx, wy, wz])
    q = q + _quat_mul(q, qw)
    q /= np.linalg.norm(q)
    return _quat_matrix(q)


def lookat(pos=(0, 0, 0), back=(0, 0, 3), up=(0, 1, 1e-12)):
    """View matrix of a camera at pos + back looking at pos (matrix.py:22-34)."""
    pos = np.array(pos, dtype=float)
    back = np.array(back, dtype=float)
    up = np.array(up, dtype=float)
    fwd = -back
    fwd /= np.linalg.norm(fwd)
    right = np.cross(fwd, up)
    right /= np.linalg.norm(right)
    up = np.cross(right, fwd)
    lin = np.transpose(np.stack([right, up, -fwd]))
    return np.linalg.inv(affine(lin, (pos + back)))


def ortho(left=-1, right=1, bottom=-1, top=1, near=-100, far=100):
    lin = np.eye(4)
    lin[0, 0] = 2 / (right - left)
    lin[1, 1] = 2 / (top - bottom)
    lin[2, 2] = -2 / (far - near)
    lin[0, 3] = -(right + left) / (right - left)
    lin[1, 3] = -(top + bottom) / (top - bottom)
    lin[2, 3] = -(far + near) / (far - near)
    return lin


def frustum(left=-1, right=1, bottom=-1, top=1, near=1, far=100):
    """OpenGL-style projection (matrix.py:48-58)."""
    lin = np.eye(4)
    lin[0, 0] = 2 * near / (right - left)
    lin[1, 1] = 2 * near / (top - bottom)
    lin[0, 2] = (right + left) / (right - left)
    lin[1, 2] = (top + bottom) / (top - bottom)
    lin[2, 2] = -(far + near) / (far - near)
    lin[2, 3] = -2 * far * near / (far - near)
    lin[3, 2] = -1
    lin[3, 3] = 0
    return lin


def orthogonal(size=1, aspect=1, near=-100, far=100):
    ax, ay = size * aspect, size
    return ortho(-ax, ax, -ay, ay, near, far)


def perspective(fov=60, aspect=1, near=0.05, far=500):
    """Symmetric frustum from a vertical field of view in degrees (matrix.py:66-69)."""
    fov = np.tan(np.radians(fov) / 2)
    ax, ay = fov * aspect, fov
    return frustum(-near * ax, near * ax, -near * ay, near * ay, near, far)


def scale(factor):
    return affine(np.eye(3) * np.array(factor), np.zeros(3))


def translate(offset):
    return affine(np.eye(3), np.array(offset) * np.ones(3))


def quaternion(q):
    """Rotation from (x, y, z, w) components (matrix.py:82-93)."""
    R = np.array([
        [1.0 - 2 * (q[1] * q[1] + q[2] * q[2]), 2 * (q[0] * q[1] - q[3] * q[2]), 2 * (q[3] * q[1] + q[0] * q[2])],
        [2 * (q[0] * q[1] + q[3] * q[2]), 1.0 - 2 * (q[0] * q[0] + q[2] * q[2]), 2 * (q[1] * q[2] - q[3] * q[0])],
        [2 * (q[0] * q[2] - q[3] * q[1]), 2 * (q[1] * q[2] + q[3] * q[0]), 1.0 - 2 * (q[0] * q[0] + q[1] * q[1])]])
    return affine(R, np.zeros(3))


def eularXYZ(theta):
    Rx = np.array([[1, 0, 0], [0, np.cos(theta[0]), -np.sin(theta[0])], [0, np.sin(theta[0]), np.cos(theta[0])]])
    Ry = np.array([[np.cos(theta[1]), 0, np.sin(theta[1])], [0, 1, 0], [-np.sin(theta[1]), 0, np.cos(theta[1])]])
    Rz = np.array([[np.cos(theta[2]), -np.sin(theta[2]), 0], [np.sin(theta[2]), np.cos(theta[2]), 0], [0, 0, 1]])
    return affine(Rz @ Ry @ Rx, np.zeros(3))


def camera_matrices(view, proj):
    """Engine.set_camera (tina/core/engine.py:72-76): W2V = proj @ view and
    V2W = inv(W2V) in FP64, both cast to f32."""
    W2V = np.asarray(proj, dtype=np.float64) @ np.asarray(view, dtype=np.float64)
    V2W = np.linalg.inv(W2V)
    return np.array(W2V, dtype=np.float32), np.array(V2W, dtype=np.float32)
