"""OpenGL-convention 4x4 matrices that callers feed to VolumeRenderer.

Same function names, arguments and result dtypes as
spimagine/utils/transform_matrices.py:21-118 (the renderer inverts these in float32 or
float64 depending on the dtype, volumerender.py:310-316, so the dtype is part of the
interface):
  float32: mat4_scale, mat4_perspective, mat4_frustrum, mat4_ortho, mat4_stereo_perspective
  float64: mat4_identity, mat4_translate, mat4_rotation, mat4_rotation_euler, mat4_lookat
The formulas are the standard gluPerspective / glFrustum / glOrtho / gluLookAt ones.
"""
import numpy as np

from .quaternion import Quaternion


def mat4_scale(x=1., y=1., z=1.):
    return np.diag(np.array([x, y, z, 1.], np.float32))


def mat4_rotation(w=0, x=1, y=0, z=0):
    """Rotation by angle w (radians) about the axis (x, y, z)."""
    axis = np.array([x, y, z], np.float32)
    axis *= 1. / np.sqrt(1. * np.sum(axis ** 2))
    half = .5 * w
    return Quaternion(np.cos(half), *(np.sin(half) * axis)).toRotation4()


def mat4_rotation_euler(yaw=0, pitch=0, roll=0):
    """z-y'-x' convention: R = Rz(yaw) . Ry(pitch) . Rx(roll)."""
    return np.dot(mat4_rotation(yaw, 0, 0, 1),
                  np.dot(mat4_rotation(pitch, 0, 1, 0), mat4_rotation(roll, 1, 0, 0)))


def mat4_perspective(fovy=45, aspect=1., z1=0.1, z2=10):
    """gluPerspective(fovy [deg], aspect, zNear, zFar)."""
    f = 1. / np.tan(np.deg2rad(fovy) / 2.)
    M = np.zeros((4, 4), np.float32)
    M[0, 0] = f / aspect
    M[1, 1] = f
    M[2, 2] = -(z2 + z1) / (z2 - z1)
    M[2, 3] = -2. * z1 * z2 / (z2 - z1)
    M[3, 2] = -1
    return M


def mat4_frustrum(left, right, bottom, top, zNear, zFar):
    """glFrustum."""
    M = np.zeros((4, 4), np.float32)
    M[0, 0] = 2. * zNear / (right - left)
    M[0, 2] = 1. * (right + left) / (right - left)
    M[1, 1] = 2. * zNear / (top - bottom)
    M[1, 2] = (top + bottom) / (top - bottom)
    M[2, 2] = -1. * (zFar + zNear) / (zFar - zNear)
    M[2, 3] = -2. * zFar * zNear / (zFar - zNear)
    M[3, 2] = -1.
    return M


def mat4_stereo_perspective(fovy=45, aspect=1., z1=0.1, z2=10, eye_shift=0):
    h = z1 * np.tan(np.deg2rad(fovy) / 2.)
    w = h * aspect
    return mat4_frustrum(-w - eye_shift, w - eye_shift, -h, h, z1, z2)


def mat4_ortho(x1=-1, x2=1, y1=-1, y2=1, z1=-1, z2=1):
    """glOrtho(left, right, bottom, top, near, far)."""
    M = np.zeros((4, 4), np.float32)
    M[0, 0] = 2. / (x2 - x1)
    M[1, 1] = 2. / (y2 - y1)
    M[2, 2] = -2. / (z2 - z1)
    M[0, 3] = -1. * (x2 + x1) / (x2 - x1)
    M[1, 3] = -1. * (y2 + y1) / (y2 - y1)
    M[2, 3] = -1. * (z2 + z1) / (z2 - z1)
    M[3, 3] = 1.
    return M


def mat4_identity():
    return np.identity(4)


def mat4_translate(x=0, y=0, z=0):
    M = np.identity(4)
    M[:3, 3] = x, y, z
    return M


def mat4_lookat(eye, center, up):
    """gluLookAt."""
    eye = np.array(eye, np.float32)
    fwd = np.array(center, np.float32) - eye
    fwd *= 1. / np.sqrt(np.sum(fwd ** 2))
    side = np.cross(fwd, np.array(up, np.float32))
    side *= 1. / np.sqrt(np.sum(side ** 2))
    upv = np.cross(side, fwd)
    M = np.identity(4)
    M[0, :3] = side
    M[1, :3] = upv
    M[2, :3] = -fwd
    return np.dot(M, mat4_translate(*(-eye)))
