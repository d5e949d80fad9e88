"""Host helpers of the ray path (math_functions.py subset)."""
import numpy as np


def calc_dv_norm(source, target, direction=None, length=None):
    """Unit direction and length of target - source in float32
    (math_functions.py:148-156): float32 difference, squares and sum,
    sqrt, then a float32 division."""
    s = np.asarray(source, np.float32)
    t = np.asarray(target, np.float32)
    d = t - s
    ln = np.sqrt(d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2]).reshape(-1, 1)
    return (d / ln).astype(np.float32), ln.astype(np.float32)
