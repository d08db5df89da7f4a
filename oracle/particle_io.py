"""numpy restatement of the quantised particle file (engine/particle_io.py).

TEST INFRASTRUCTURE: this module is the oracle for the particle output path.
Only ``tests/`` may import it, as the checker; the product package never does.

``quantize`` follows ParticleIO.write_particles (engine/particle_io.py:12-76)
with the arithmetic of the NumPy 1.x the reference ran on (taichi 1.1.0
era):
  * lo, hi = min, max of the component (float32);
  * hi = max(lo + 1e-5, hi): ``ranges[c, d, 0] + 1e-5`` is a float32 scalar
    plus a Python float, which NumPy 1.x evaluates in float64; the winner is
    stored back into the float32 array (particle_io.py:45-47);
  * q = uint32(((c - lo) * f32(1 / (hi - lo))) * (2^bits - 1) + f32(0.499)),
    every operation a float32 array operation (particle_io.py:48-55); the
    reciprocal 1 / f32(hi - lo) is float64 in NumPy 1.x and rounded to
    float32 when it meets the float32 array, which equals the correctly
    rounded float32 quotient;
  * x_and_v = (qx << 8) + qv in uint32 (particle_io.py:56);
  * colour: int32 -> float32 staging -> uint32, bytes (c >> 16, c >> 8, c)
    & 255 (particle_io.py:60-71).
"""
import numpy as np

X_BITS, V_BITS = 24, 8


def _ranges(a):
    lo = np.float32(a.min())
    hi = np.float32(a.max())
    widened = float(lo) + 1e-5  # float64, as NumPy 1.x promotes the scalar sum
    if widened > float(hi):
        hi = np.float32(widened)
    return lo, hi


def _quantize(a, lo, hi, bits):
    width = np.float32(hi - lo)
    inv = np.float32(1.0 / float(width))
    t = (a.astype(np.float32) - lo) * inv
    t = t * np.float32(2**bits - 1)
    t = t + np.float32(0.499)
    return t.astype(np.uint32)


def quantize(x, v, color):
    """(ranges (2, d, 2) float32, x_and_v (n, d) uint32, color (n, 3) uint8)."""
    x = np.asarray(x, np.float32)
    v = np.asarray(v, np.float32)
    n, d = x.shape
    if n == 0:
        raise ValueError("zero-size array: no particles")
    ranges = np.zeros((2, d, 2), np.float32)
    x_and_v = np.zeros((n, d), np.uint32)
    for a in range(d):
        xl, xh = _ranges(x[:, a])
        vl, vh = _ranges(v[:, a])
        ranges[0, a] = (xl, xh)
        ranges[1, a] = (vl, vh)
        qx = _quantize(x[:, a], xl, xh, X_BITS)
        qv = _quantize(v[:, a], vl, vh, V_BITS)
        x_and_v[:, a] = (qx << np.uint32(V_BITS)) + qv
    c = np.asarray(color, np.int32).astype(np.float32).astype(np.int64).astype(np.uint32)
    rgb = np.stack([(c >> np.uint32(16)) & np.uint32(255), (c >> np.uint32(8)) & np.uint32(255),
                    c & np.uint32(255)], axis=1).astype(np.uint8)
    return ranges, x_and_v, rgb
