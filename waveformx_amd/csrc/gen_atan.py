"""The coefficients of ``fast_atan`` (csrc/xas_point.h), the FP32 arctangent
of the split and FP32 evaluations: atan(x) = x P(x^2) on [0, 1] with P of
degree 7, the near-minimax polynomial of the absolute error by Lawson's
iteratively reweighted least squares (C. L. Lawson, 1961: w <- w |e|, which
converges to the minimax solution's weights) on a Chebyshev-spaced grid,
in float64, rounded to float32 at the end; beyond 1 the kernel takes
pi/2 - atan(1/x).

``python gen_atan.py`` prints the coefficients, highest degree first, as
the C literals the kernel's Horner chain uses, and the fit's error: in
exact arithmetic with the float32 coefficients, and in float32 arithmetic
(each fused multiply-add rounded once).
"""

import numpy as np

DEGREE = 7
POINTS = 4001
ITERATIONS = 100


def fit(degree: int = DEGREE, points: int = POINTS, iterations: int = ITERATIONS) -> np.ndarray:
    """P's coefficients, constant first, in float64."""
    x = 0.5 * (1.0 - np.cos(np.pi * np.arange(points) / (points - 1)))
    basis = np.stack([x ** (2 * k + 1) for k in range(degree + 1)], axis=1)
    y = np.arctan(x)
    w = np.full(points, 1.0 / points)
    c = None
    for _ in range(iterations):
        sw = np.sqrt(w)
        c = np.linalg.lstsq(basis * sw[:, None], y * sw, rcond=None)[0]
        e = np.abs(basis @ c - y)
        w = w * e
        w /= w.sum()
    return c


def coefficients() -> np.ndarray:
    """The float32 coefficients, constant first."""
    return fit().astype(np.float32)


def fast_atan_f32(x: np.ndarray, c: np.ndarray) -> np.ndarray:
    """The kernel's arithmetic in float32 (each fmaf rounded once, the
    reciprocal exact): a reference for the bound, not the device's bits."""
    x = np.asarray(x, dtype=np.float32)
    a = np.abs(x)
    big = a > 1.0
    with np.errstate(divide="ignore"):
        r = np.where(big, np.float32(1.0) / a, a).astype(np.float32)
    r2 = (r * r).astype(np.float32)
    p = np.full_like(r, c[-1])
    for k in range(len(c) - 2, -1, -1):
        p = (p.astype(np.float64) * r2 + np.float64(c[k])).astype(np.float32)
    p = (p * r).astype(np.float32)
    p = np.where(big, (np.float32(np.pi / 2) - p).astype(np.float32), p)
    return np.copysign(p, x)


def literals(c: np.ndarray) -> list:
    """The C float literals, highest degree first."""
    return [repr(float(v)) + "f" for v in c[::-1]]


if __name__ == "__main__":
    c = coefficients()
    print("\n".join(literals(c)))
    x = np.linspace(0.0, 1.0, 200001)
    exact = np.max(np.abs(x * np.polyval(c[::-1].astype(np.float64), x * x) - np.arctan(x)))
    z = np.concatenate([x, np.geomspace(1.0, 1e30, 200001)])
    f32 = np.max(np.abs(fast_atan_f32(z, c).astype(np.float64) - np.arctan(z)))
    print(f"max |error|: {exact:.2e} exact arithmetic on [0, 1], {f32:.2e} float32 arithmetic on [0, 1e30]")
