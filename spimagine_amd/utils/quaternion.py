"""Unit-quaternion helper used to build rotation matrices for the renderer.

Same interface as spimagine/utils/quaternion.py (Quaternion(w, x, y, z),
toRotation4/toRotation3, slerp), so poses built here are the poses a reference caller
builds: q = (w, x, y, z) with w the scalar part.
"""
import numpy as np


class Quaternion(object):
    def __init__(self, w=1., x=0., y=0., z=0.):
        self.data = np.array([w, x, y, z])

    @classmethod
    def copy(cls, rhs):
        return cls(*rhs.data)

    def __getitem__(self, i):
        return self.data[i]

    def __setitem__(self, i, val):
        self.data[i] = val

    def conj(self):
        """As quaternion.py:24 returns it: (w, -x, -y, -y) -- the last component repeats
        -y, not -z.  Kept for drop-in parity (nothing in the reference calls it);
        conjugate() is the mathematical conjugate."""
        return Quaternion(self.data[0], -self.data[1], -self.data[2], -self.data[2])

    def conjugate(self):
        """(w, -x, -y, -z)."""
        return Quaternion(self.data[0], -self.data[1], -self.data[2], -self.data[3])

    def norm(self):
        return np.linalg.norm(self.data)

    def __add__(self, q):
        return Quaternion(*(self.data + q.data))

    def __sub__(self, q):
        return Quaternion(*(self.data - q.data))

    def __mul__(self, q):
        if not isinstance(q, Quaternion):
            return Quaternion(*(q * self.data))
        w1, v1 = self.data[0], self.data[1:]
        w2, v2 = q.data[0], q.data[1:]
        w = w1 * w2 - np.dot(v1, v2)
        v = w1 * v2 + w2 * v1 + np.cross(v1, v2)
        return Quaternion(w, *v)

    def __repr__(self):
        return "Quaternion(%s,%s,%s,%s)" % tuple(self.data)

    def dot(self, q):
        return np.inner(self.data, q.data)

    def normalize(self):
        return Quaternion(*(self.data / self.norm()))

    def toRotation3(self):
        """3x3 rotation matrix of the (unit) quaternion, v' = R v."""
        w, x, y, z = self.data
        return np.array([
            [w * w + x * x - y * y - z * z, 2 * (x * y - w * z), 2 * (x * z + w * y)],
            [2 * (x * y + w * z), w * w - x * x + y * y - z * z, 2 * (y * z - w * x)],
            [2 * (x * z - w * y), 2 * (y * z + w * x), w * w - x * x - y * y + z * z],
        ])

    def toRotation4(self):
        R = np.zeros((4, 4), dtype=np.result_type(self.data.dtype, np.float64))
        R[:3, :3] = self.toRotation3()
        R[3, 3] = 1
        return R


def quaternion_slerp(q1, q2, t):
    """Spherical interpolation along the shorter arc (linear when nearly parallel)."""
    q1, q2 = q1.normalize(), q2.normalize()
    c = q1.dot(q2)
    if abs(c) > .9998:
        return q1 + (q2 - q1) * t
    if c < 0:
        q2 = q2 * (-1.)
        c = -c
    ang = np.arccos(c)
    s = np.sin(ang)
    return q1 * (np.sin((1. - t) * ang) / s) + q2 * (np.sin(t * ang) / s)
