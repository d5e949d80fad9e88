import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "lyceanem-python_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X) and liblem.so")


def gpu_available() -> bool:
    try:
        import torch

        return torch.cuda.is_available()
    except Exception:  # pragma: no cover
        return False


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as orc

    orc.build()
    return orc


@pytest.fixture(scope="session")
def native():
    """The HIP library bound through ctypes; GPU tests fail (not skip) if it is missing."""
    from lyceanem import _native

    _native.load_library()
    import torch

    assert torch.cuda.is_available(), "GPU test run without a HIP device"
    return _native


@pytest.fixture(scope="session")
def term_bound():
    """Per sink, an upper bound of the sum over sources of |field term| of the
    line-of-sight kernel: |w_s| |n_k| |wk_k| e^{-alpha r} |alpha + 1/r + j beta|
    / (2 pi r) (the projector and the sink tangent step do not lengthen a
    vector): the scale on which field sums whose sources are grouped
    differently (other chunk counts) are compared -- they differ by float32
    summation rounding, not bit for bit."""
    import numpy as np
    import torch

    def bound(src, w, snk, nrm, wk=None, alpha=0.0, beta=2 * np.pi):
        t = [torch.as_tensor(np.asarray(a.cpu() if hasattr(a, "cpu") else a))
             for a in (src, w, snk, nrm)]
        src, w, snk, nrm = t
        wn = torch.linalg.vector_norm(w, dim=1).double()
        out = []
        for k0 in range(0, snk.shape[0], 4096):
            r = torch.cdist(snk[k0:k0 + 4096].double(), src.double())
            f = torch.exp(-alpha * r) * torch.sqrt((alpha + 1 / r) ** 2 + beta ** 2) / (
                2 * np.pi * r)
            out.append(f @ wn)
        b = torch.cat(out) * torch.linalg.vector_norm(nrm.double(), dim=1)
        if wk is not None:
            wk = torch.as_tensor(np.asarray(wk.cpu() if hasattr(wk, "cpu") else wk))
            b = b * torch.linalg.vector_norm(wk, dim=1).double()
        return b.numpy()

    return bound
