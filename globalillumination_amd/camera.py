"""Light/camera matrices through the C ABI (FP32 glm 0.9.3.1 restatement).

All matrices are numpy float32 arrays of 16 entries, column-major, exactly as
glm stores them (``m[col*4 + row]``).  These wrap the host functions of
include/gi.h; they need no GPU.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib


def _f16() -> np.ndarray:
    return np.zeros(16, dtype=np.float32)


def _ptr(a: np.ndarray) -> ctypes.c_void_p:
    assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"]
    return ctypes.c_void_p(a.ctypes.data)


def perspective(fovy_deg: float, aspect: float, znear: float, zfar: float) -> np.ndarray:
    """glm::perspective, fovy in degrees (glm/gtc/matrix_transform.inl:223-244)."""
    out = _f16()
    _lib.call("gi_perspective", fovy_deg, aspect, znear, zfar, _ptr(out))
    return out


def ortho(left: float, right: float, bottom: float, top: float, znear: float, zfar: float) -> np.ndarray:
    out = _f16()
    _lib.call("gi_ortho", left, right, bottom, top, znear, zfar, _ptr(out))
    return out


def lookat(eye, center, up) -> np.ndarray:
    """glm::lookAt (glm/gtc/matrix_transform.inl:383-411)."""
    e = np.ascontiguousarray(eye, dtype=np.float32)
    c = np.ascontiguousarray(center, dtype=np.float32)
    u = np.ascontiguousarray(up, dtype=np.float32)
    out = _f16()
    _lib.call("gi_lookat", _ptr(e), _ptr(c), _ptr(u), _ptr(out))
    return out


def mat4_mul(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.float32)
    b = np.ascontiguousarray(b, dtype=np.float32)
    out = _f16()
    _lib.call("gi_mat4_mul", _ptr(a), _ptr(b), _ptr(out))
    return out


def bias_matrix(zbias: float = 0.0) -> np.ndarray:
    """Texture-space bias B (MyGLGeometryViewer.cpp:139-145), B[3][2] = 0.5 - zbias."""
    out = _f16()
    _lib.call("gi_bias_matrix", zbias, _ptr(out))
    return out


def uniform_sampling(base, size: int, nsamples: int, index: int) -> np.ndarray:
    """UniformSampledLightSource::computeUniformSampling (:27-37)."""
    b = np.ascontiguousarray(base, dtype=np.float32)
    out = np.zeros(3, dtype=np.float32)
    _lib.call("gi_uniform_sampling", _ptr(b), int(size), int(nsamples), int(index), _ptr(out))
    return out


def soft_light_trans(biased_mvps: np.ndarray) -> np.ndarray:
    """Column 3 of each B*lightMVP_i (MyGLGeometryViewer.cpp:184-195) -> (S, 4)."""
    m = np.ascontiguousarray(biased_mvps, dtype=np.float32).reshape(-1, 16)
    out = np.zeros((m.shape[0], 4), dtype=np.float32)
    _lib.call("gi_soft_light_trans", _ptr(m), int(m.shape[0]), _ptr(out))
    return out


def normal_matrix(mv: np.ndarray) -> np.ndarray:
    """glm::inverseTranspose(mat3(MV)) as configurePhong uploads it
    (MyGLGeometryViewer.cpp:113-116); column-major mat3 (9 floats).

    The pre-evaluation only tests the sign of dot(N, L), so this matrix is an
    input like MV; it is computed in float64 and rounded once."""
    m = np.asarray(mv, dtype=np.float64).reshape(4, 4).T[:3, :3]  # row-major 3x3
    nm = np.linalg.inv(m).T
    return np.ascontiguousarray(nm.T.reshape(9), dtype=np.float32)  # back to column-major
