"""Shared test helpers: error measures, FP64 references, fixture key parsing."""
from __future__ import annotations

import numpy as np

# Float pass criterion: rel-L2 against FP64 <= min(1e-6, 2 x the reference float build's
# own error on the same inputs); the reference errors are stored in golden_float.npz.
ABS_FLOAT_BOUND = 1e-6


def float_bound(golden_float, name: str) -> float:
    return min(ABS_FLOAT_BOUND, 2.0 * float(golden_float["err_" + name]))


def cplx(a) -> np.ndarray:
    a = np.asarray(a, np.float64)
    return a[..., 0] + 1j * a[..., 1]


def pairs(z, scalar=np.float32) -> np.ndarray:
    z = np.asarray(z)
    return np.stack([z.real, z.imag], axis=-1).astype(scalar)


def rel_l2(got, want) -> float:
    got = np.asarray(got, np.complex128)
    want = np.asarray(want, np.complex128)
    return float(np.linalg.norm((got - want).ravel()) / np.linalg.norm(want.ravel()))


def fixture_cases(golden, prefix: str):
    """Yield (key-stem, input, output) for every '<prefix>_..._in' / '_out' pair."""
    for k in sorted(golden.files):
        if k.startswith(prefix + "_") and k.endswith("_in"):
            stem = k[:-3]
            yield stem, golden[stem + "_in"], golden[stem + "_out"]


def parse_dims(s: str):
    return tuple(int(v) for v in s.split("x"))
