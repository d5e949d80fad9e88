"""ctypes binding of liblem.so (the gfx950 HIP library, declared in include/lem.h).

PyTorch-ROCm provides device memory and the current stream; every compute call
goes to a hand-written HIP kernel.  There is no CPU fallback: without the built
library or without a HIP device the calls raise.
"""
from __future__ import annotations

import ctypes
import os
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "liblem.so")
ABI_VERSION = 1

SEM_NUMBA = 0
SEM_NATIVE = 1
VIS_ANY_HIT = 1
VIS_SKIP_SELF = 2
VIS_SINK_MAJOR = 4
VIS_INPUT_ORDER = 8
VIS_FROM_SINK = 16
FIELD_SINK_TANGENT = 1
FIELD_ELEMENTS = 2
FIELD_BITS_SINK_MAJOR = 4
FIELD_INPUT_ORDER = 8

# Entry points and their ctypes signatures (mirrors include/lem.h).
_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_INT = ctypes.c_int
_D = ctypes.c_double
SIGNATURES = {
    "lem_abi_version": ([], _INT),
    "lem_last_error": ([], ctypes.c_char_p),
    "lem_device_arch": ([ctypes.c_char_p, _INT], _INT),
    "lem_visibility_dense": ([_P, _I64, _P, _I64, _P, _I64, _INT, _INT, _P, _P, _P], _INT),
    "lem_raycast_pairs": ([_P, _I64, _P, _I64, _P, _I64, _INT, _INT, _P, _P, _P], _INT),
    "lem_field_los_workspace": ([_I64, _I64, _INT], _I64),
    "lem_field_los": ([_P, _P, _I64, _P, _P, _P, _I64, _P, _D, _D, _INT, _P, _P, _I64, _P],
                      _INT),
    "lem_field_paths": ([_P, _P, _P, _P, _I64, _P, _INT, _P, _P, _P, _I64, _D, _D, _P, _P, _P],
                        _INT),
    "lem_grid_workspace": ([_I64, ctypes.c_float], _I64),
    "lem_grid_plan": ([_P, _I64, ctypes.c_float, _P, _P, _I64, _P], _INT),
    "lem_grid_fill": ([_P, _I64, _P, _P, _P, _P, _P, _P, _I64, ctypes.c_float, _P], _INT),
    "lem_visibility_grid_workspace": ([_I64, _I64, _INT], _I64),
    "lem_visibility_grid": ([_P, _I64, _P, _I64, _P, _P, _P, _P, _INT, _INT, _P, _P, _P, _I64,
                             _P], _INT),
    "lem_visibility_grid_stats": ([_P, _I64, _P, _I64, _P, _P, _P, _P, _INT, _P, _P], _INT),
    "lem_beam_workspace": ([_I64, _I64], _I64),
    "lem_beam_power": ([_P, _I64, _P, _P, _I64, _I64, _P, _P, _P, _P, _P, _I64, _P], _INT),
    "lem_projected_area": ([_P, _P, _P, _I64, _P, _I64, _P, _P, _P], _INT),
    "lem_time_paths": ([_P, _P, _P, _P, _I64, _P, _I64, _INT, _D, _D, _P, _P, _P, _P], _INT),
    "lem_time_wake": ([_P, _I64, _P, _P], _INT),
    "lem_time_bins": ([_P, _I64, _P, _D, _P, _P], _INT),
    "lem_time_map": ([_P, _P, _P, _P, _P, _I64, _P, _I64, _I64, _P, _P], _INT),
    "lem_time_trace_workspace": ([_I64, _I64], _I64),
    "lem_time_trace": ([_P, _P, _I64, _P, _I64, _I64, _P, _P, _I64, _P], _INT),
    "lem_field_los_grid_workspace": ([_I64, _I64, _INT], _I64),
    "lem_field_los_grid": ([_P, _P, _I64, _P, _P, _P, _I64, _P, _P, _P, _P, _D, _D, _INT, _P,
                            _P, _I64, _P], _INT),
    "lem_gw_mfma": ([_P, _I64, ctypes.c_float, ctypes.c_float, ctypes.c_float, _P, _P, _I64, _P,
                     _I64, _D, _P, _I64, _P, _P], _INT),
    "lem_field_los_noalloc_max_sinks": ([], _I64),
    "lem_field_los_noalloc": ([_P, _P, _I64, _P, _P, _P, _I64, _P, _D, _D, _INT, _P, _P], _INT),
    "lem_field_los_adjoint_workspace": ([_I64, _I64], _I64),
    "lem_db_mask_loss_workspace": ([_I64], _I64),
    "lem_db_mask_loss": ([_P, _P, _I64, _D, _D, _D, _P, _P, _P, _I64, _P], _INT),
    "lem_field_los_adjoint": ([_P, _P, _I64, _P, _P, _P, _I64, _P, _D, _D, _INT, _P, _P, _P, _P,
                               _I64, _P], _INT),
}


class GridInfo(ctypes.Structure):
    """lem_grid_info (include/lem.h)."""
    _fields_ = [("lo", ctypes.c_float * 3), ("cell", ctypes.c_float * 3),
                ("n", ctypes.c_int32 * 3), ("mshift", ctypes.c_int32),
                ("ncells", ctypes.c_int64), ("nrefs", ctypes.c_int64),
                ("eps", ctypes.c_float), ("nmacro", ctypes.c_int32)]

_lock = threading.Lock()
_lib = None


class LemError(RuntimeError):
    """A liblem.so entry point returned an error code."""


def _mapped_hip_runtimes():
    """Paths of every libamdhip64 mapped into this process."""
    paths = set()
    try:
        with open("/proc/self/maps") as f:
            for line in f:
                if "libamdhip64" in line:
                    paths.add(os.path.realpath(line.split()[-1]))
    except OSError:  # pragma: no cover
        pass
    return paths


def load_library(path: str = LIB_PATH):
    """Load liblem.so and bind every entry point (no GPU needed for this step).

    torch is imported first: torch-ROCm ships its own HIP runtime with the same
    soname (libamdhip64.so.7) as the system one liblem.so links against, so
    loading torch first makes liblem.so bind to that one runtime; the other
    order would put two HIP runtimes in the process."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(path):
                raise LemError(
                    f"{path} is missing: build it with "
                    "`python lyceanem-python_amd/build_lib.py` (there is no CPU fallback)")
            import torch  # noqa: F401  (loads torch's HIP runtime before liblem.so)

            lib = ctypes.CDLL(path)
            runtimes = _mapped_hip_runtimes()
            if len(runtimes) > 1:
                raise LemError(f"two HIP runtimes are loaded ({sorted(runtimes)}); import torch "
                               "before anything that loads libamdhip64")
            for name, (args, res) in SIGNATURES.items():
                fn = getattr(lib, name)
                fn.argtypes = args
                fn.restype = res
            if lib.lem_abi_version() != ABI_VERSION:
                raise LemError(f"liblem.so ABI {lib.lem_abi_version()} != {ABI_VERSION}")
            _lib = lib
    return _lib


def _torch():
    import torch

    if not torch.cuda.is_available():
        raise LemError("no HIP device visible: the LyceanEM MI355X path has no CPU fallback")
    return torch


def device():
    return _torch().device("cuda", _torch().cuda.current_device())


def stream_ptr():
    return ctypes.c_void_p(_torch().cuda.current_stream().cuda_stream)


def check(rc: int, what: str):
    if rc != 0:
        msg = load_library().lem_last_error().decode(errors="replace")
        raise LemError(f"{what} failed ({rc}): {msg}")


def ptr(t):
    """Device pointer of a contiguous tensor (None -> NULL)."""
    if t is None:
        return None
    if not t.is_contiguous():
        raise LemError("liblem.so needs contiguous tensors")
    return ctypes.c_void_p(t.data_ptr())


def to_dev(a, dtype=None):
    """numpy array or tensor -> contiguous tensor on the current HIP device."""
    torch = _torch()
    if isinstance(a, torch.Tensor):
        t = a
        if dtype is not None and t.dtype != dtype:
            t = t.to(dtype)
        if t.device.type != "cuda":
            t = t.to(device())
        return t.contiguous()
    arr = np.ascontiguousarray(a)
    t = torch.from_numpy(arr)
    if dtype is not None:
        t = t.to(dtype)
    return t.to(device()).contiguous()


# --------------------------------------------------------------------------
# typed wrappers (tensors in, tensors out; all on the current stream)
# --------------------------------------------------------------------------

def visibility_dense(src, snk, tris, semantics=SEM_NUMBA, any_hit=True, skip_self=False,
                     want_dist=False):
    """Dense S x K hit bits (uint64 words, 1 = blocked) and optional dist."""
    torch = _torch()
    lib = load_library()
    src = to_dev(src, torch.float32).reshape(-1, 3)
    snk = to_dev(snk, torch.float32).reshape(-1, 3)
    tris = to_dev(tris, torch.float32).reshape(-1, 9)
    S, K, T = src.shape[0], snk.shape[0], tris.shape[0]
    W = (K + 63) // 64
    bits = torch.zeros((S, W), dtype=torch.int64, device=src.device)
    dist = torch.empty((S, K), dtype=torch.float32, device=src.device) if want_dist else None
    flags = (VIS_ANY_HIT if any_hit else 0) | (VIS_SKIP_SELF if skip_self else 0)
    check(lib.lem_visibility_dense(ptr(src), S, ptr(snk), K, ptr(tris) if T else None, T,
                                   semantics, flags, ptr(bits), ptr(dist), stream_ptr()),
          "lem_visibility_dense")
    return bits, dist


def raycast_pairs(points, pairs, tris, semantics=SEM_NUMBA, any_hit=False, want_dist=True):
    """kernel1Dv3 over explicit 0-based (origin, target) rows of `points`."""
    torch = _torch()
    lib = load_library()
    pts = to_dev(points, torch.float32).reshape(-1, 3)
    pr = to_dev(pairs, torch.int32).reshape(-1, 2)
    tris = to_dev(tris, torch.float32).reshape(-1, 9)
    R, T = pr.shape[0], tris.shape[0]
    inter = torch.empty(R, dtype=torch.uint8, device=pts.device)
    dist = torch.empty(R, dtype=torch.float32, device=pts.device) if want_dist else None
    flags = VIS_ANY_HIT if any_hit else 0
    check(lib.lem_raycast_pairs(ptr(pts), pts.shape[0], ptr(pr), R, ptr(tris) if T else None, T,
                                semantics, flags, ptr(inter), ptr(dist), stream_ptr()),
          "lem_raycast_pairs")
    return inter, dist


def field_los(src, src_w, snk, snk_n, hit_bits=None, snk_w=None, alpha=0.0, beta=2 * np.pi,
              sink_tangent=False, elements=False, bits_sink_major=False):
    """Fused line-of-sight field: (K,3) complex64 source sum or (S,K,3) network.
    hit_bits: S x W64(K) words, or K x W64(S) with bits_sink_major (the layout
    of Grid.visibility(sink_major=True))."""
    torch = _torch()
    lib = load_library()
    src = to_dev(src, torch.float32).reshape(-1, 3)
    snk = to_dev(snk, torch.float32).reshape(-1, 3)
    snk_n = to_dev(snk_n, torch.float32).reshape(-1, 3)
    S, K = src.shape[0], snk.shape[0]
    src_w = to_dev(src_w, torch.complex64).reshape(S, 3)
    if snk_w is not None:
        snk_w = to_dev(snk_w, torch.complex64).reshape(K, 3)
    flags = (FIELD_SINK_TANGENT if sink_tangent else 0) | (FIELD_ELEMENTS if elements else 0) | \
        (FIELD_BITS_SINK_MAJOR if bits_sink_major else 0)
    out = torch.empty((S, K, 3) if elements else (K, 3), dtype=torch.complex64,
                      device=src.device)
    wsb = int(lib.lem_field_los_workspace(S, K, flags))
    ws = torch.empty(max(wsb, 1), dtype=torch.uint8, device=src.device)
    if hit_bits is not None:
        hit_bits = to_dev(hit_bits, torch.int64)
        want = (K, (S + 63) // 64) if bits_sink_major else (S, (K + 63) // 64)
        if tuple(hit_bits.shape) != want:
            raise LemError(f"hit_bits shape {tuple(hit_bits.shape)} != {want}")
    check(lib.lem_field_los(ptr(src), ptr(src_w), S, ptr(snk), ptr(snk_n), ptr(snk_w), K,
                            ptr(hit_bits), float(alpha), float(beta), flags, ptr(out), ptr(ws),
                            wsb, stream_ptr()), "lem_field_los")
    return out


def field_los_adjoint(src, src_w, snk, snk_n, grad_out, hit_bits=None, snk_w=None, alpha=0.0,
                      beta=2 * np.pi, sink_tangent=False, want_w=True, want_p=True,
                      input_order=False):
    """Adjoint of field_los (source sum): (grad_w (S,3) complex64, grad_p (S,3) float32)."""
    torch = _torch()
    lib = load_library()
    src = to_dev(src, torch.float32).reshape(-1, 3)
    snk = to_dev(snk, torch.float32).reshape(-1, 3)
    snk_n = to_dev(snk_n, torch.float32).reshape(-1, 3)
    S, K = src.shape[0], snk.shape[0]
    src_w = to_dev(src_w, torch.complex64).reshape(S, 3)
    g = to_dev(grad_out, torch.complex64).reshape(K, 3)
    if snk_w is not None:
        snk_w = to_dev(snk_w, torch.complex64).reshape(K, 3)
    if hit_bits is not None:
        hit_bits = to_dev(hit_bits, torch.int64)
    gw = torch.empty((S, 3), dtype=torch.complex64, device=src.device) if want_w else None
    gp = torch.empty((S, 3), dtype=torch.float32, device=src.device) if want_p else None
    wsb = int(lib.lem_field_los_adjoint_workspace(S, K))
    if wsb < 0:
        raise LemError("lem_field_los_adjoint_workspace: size query failed")
    ws = torch.empty(max(wsb, 1), dtype=torch.uint8, device=src.device)
    flags = (FIELD_SINK_TANGENT if sink_tangent else 0) | (FIELD_INPUT_ORDER if input_order else 0)
    check(lib.lem_field_los_adjoint(ptr(src), ptr(src_w), S, ptr(snk), ptr(snk_n), ptr(snk_w), K,
                                    ptr(hit_bits), float(alpha), float(beta), flags, ptr(g),
                                    ptr(gw), ptr(gp), ptr(ws), wsb, stream_ptr()),
          "lem_field_los_adjoint")
    return gw, gp


def db_mask_loss(E, in_beam, sidelobe_db=-25.0, ref_power=0.0, eps=1e-30, want_grad=True):
    """lem_db_mask_loss: (loss (0-dim float32), dloss/dE (K,3) complex64 or
    None).  ref_power <= 0: the current largest sink power."""
    torch = _torch()
    lib = load_library()
    E = to_dev(E, torch.complex64).reshape(-1, 3)
    K = E.shape[0]
    beam = to_dev(in_beam, torch.bool).reshape(-1)
    if beam.shape[0] != K:
        raise LemError("db_mask_loss: one in_beam flag per sink")
    loss = torch.empty((), dtype=torch.float32, device=E.device)
    grad = torch.empty((K, 3), dtype=torch.complex64, device=E.device) if want_grad else None
    wsb = int(lib.lem_db_mask_loss_workspace(K))
    if wsb < 0:
        raise LemError("lem_db_mask_loss_workspace: size query failed")
    ws = torch.empty(wsb, dtype=torch.uint8, device=E.device)
    check(lib.lem_db_mask_loss(ptr(E), ptr(beam), K, float(sidelobe_db), float(ref_power),
                               float(eps), ptr(loss), ptr(grad), ptr(ws), wsb, stream_ptr()),
          "lem_db_mask_loss")
    return loss, grad


def _eta(eta, N):
    """Per-point impedance (> 0: magnetic point) as an N float32 device tensor,
    or None when every point is electric."""
    if eta is None:
        return None
    torch = _torch()
    e = to_dev(eta, torch.float32).reshape(-1)
    if e.shape[0] != N:
        raise LemError("eta: one impedance per point")
    return e


def field_paths(pos, nrm, w, index, order, seg_start, seg_slot, S, K, alpha, beta, eta=None):
    """Network of an explicit path list -> (S,K,3) complex128 tensor.  eta:
    per-point impedance, > 0 for magnetic points (None: all electric)."""
    torch = _torch()
    lib = load_library()
    pos = to_dev(pos, torch.float32).reshape(-1, 3)
    nrm = to_dev(nrm, torch.float32).reshape(-1, 3)
    N = pos.shape[0]
    w = to_dev(w, torch.complex64).reshape(N, 3)
    eta = _eta(eta, N)
    idx = to_dev(index, torch.int32)
    order = to_dev(order, torch.int64)
    seg_start = to_dev(seg_start, torch.int64)
    seg_slot = to_dev(seg_slot, torch.int64)
    nseg = seg_slot.shape[0]
    out = torch.zeros((S, K, 3), dtype=torch.complex128, device=pos.device)
    err = torch.zeros(1, dtype=torch.int32, device=pos.device)
    check(lib.lem_field_paths(ptr(pos), ptr(nrm), ptr(w), ptr(eta), N, ptr(idx), idx.shape[1],
                              ptr(order),
                              ptr(seg_start), ptr(seg_slot), nseg, float(alpha), float(beta),
                              ptr(out), ptr(err), stream_ptr()), "lem_field_paths")
    if int(err.item()) != 0:
        raise LemError("lem_field_paths: path index points outside the point table")
    return out


def projected_area(src, nrm, area, snk, hit_bits):
    """(K,) float32 projected-area pattern of visiblespace; area None = unit
    portions."""
    torch = _torch()
    lib = load_library()
    src = to_dev(src, torch.float32).reshape(-1, 3)
    nrm = to_dev(nrm, torch.float32).reshape(-1, 3)
    snk = to_dev(snk, torch.float32).reshape(-1, 3)
    S, K = src.shape[0], snk.shape[0]
    a = None if area is None else to_dev(np.asarray(area, np.float32).reshape(-1), torch.float32)
    if a is not None and a.shape[0] != S:
        raise LemError("projected_area: one area per source")
    out = torch.empty(K, dtype=torch.float32, device=snk.device)
    bits = to_dev(hit_bits, torch.int64)
    check(lib.lem_projected_area(ptr(src), ptr(nrm), ptr(a), S, ptr(snk), K, ptr(bits),
                                 ptr(out), stream_ptr()), "lem_projected_area")
    return out


def beam_power(W, Et, Ep, q, pstar):
    """(power (nb,) float64, ustar (nb, 2) float32) of weight rows W (nb, ne)
    over element patterns Et, Ep (ne, np); pads to the kernel's tile sizes."""
    torch = _torch()
    lib = load_library()
    W = np.asarray(W, np.complex64)
    Et = np.asarray(Et, np.complex64)
    Ep = np.asarray(Ep, np.complex64)
    nb, ne = W.shape
    npts = Et.shape[1]
    if Et.shape != (ne, npts) or Ep.shape != (ne, npts):
        raise LemError("beam_power: patterns must be (elements, points)")
    NB, NE, NP = -(-nb // 32) * 32, -(-ne // 32) * 32, -(-npts // 64) * 64
    dev = device()
    Wd = torch.zeros((NB, NE), dtype=torch.complex64, device=dev)
    Wd[:nb, :ne] = torch.from_numpy(W).to(dev)
    Etd = torch.zeros((NE, NP), dtype=torch.complex64, device=dev)
    Etd[:ne, :npts] = torch.from_numpy(Et).to(dev)
    Epd = torch.zeros((NE, NP), dtype=torch.complex64, device=dev)
    Epd[:ne, :npts] = torch.from_numpy(Ep).to(dev)
    qd = torch.zeros(NP, dtype=torch.float32, device=dev)
    qd[:npts] = torch.from_numpy(np.asarray(q, np.float32)).to(dev)
    ps = torch.full((NB,), -1, dtype=torch.int32, device=dev)
    ps[:nb] = torch.from_numpy(np.asarray(pstar, np.int32)).to(dev)
    power = torch.empty(NB, dtype=torch.float64, device=dev)
    ustar = torch.zeros((NB, 2), dtype=torch.float32, device=dev)
    wsb = int(lib.lem_beam_workspace(NB, NP))
    ws = torch.empty(wsb, dtype=torch.uint8, device=dev)
    check(lib.lem_beam_power(ptr(Wd), NB, ptr(Etd), ptr(Epd), NE, NP, ptr(qd), ptr(ps),
                             ptr(power), ptr(ustar), ptr(ws), wsb, stream_ptr()), "lem_beam_power")
    return power[:nb].cpu().numpy(), ustar[:nb].cpu().numpy()


def time_paths(pos, nrm, w, index, alpha, beta, eta=None):
    """Per path (amp (V, 3), arrival (V,)) float64 tensors of a 1-based index;
    eta as for field_paths."""
    torch = _torch()
    lib = load_library()
    pos = to_dev(pos, torch.float32).reshape(-1, 3)
    nrm = to_dev(nrm, torch.float32).reshape(-1, 3)
    N = pos.shape[0]
    w = to_dev(w, torch.complex64).reshape(N, 3)
    eta = _eta(eta, N)
    idx = to_dev(index, torch.int32)
    if idx.dim() != 2:
        raise LemError("time_paths: index must be a 2-D array")
    V, ncol = idx.shape
    amp = torch.empty((V, 3), dtype=torch.float64, device=pos.device)
    arrival = torch.empty(V, dtype=torch.float64, device=pos.device)
    err = torch.zeros(1, dtype=torch.int32, device=pos.device)
    check(lib.lem_time_paths(ptr(pos), ptr(nrm), ptr(w), ptr(eta), N, ptr(idx), V, ncol,
                             float(alpha),
                             float(beta), ptr(amp), ptr(arrival), ptr(err), stream_ptr()),
          "lem_time_paths")
    if V and int(err.item()) != 0:
        raise LemError("lem_time_paths: path index points outside the point table")
    return amp, arrival


def time_wake(arrival):
    """min(1 s, earliest arrival) as a (1,) float64 device tensor."""
    torch = _torch()
    lib = load_library()
    arrival = to_dev(arrival, torch.float64).reshape(-1)
    wake = torch.empty(1, dtype=torch.float64, device=arrival.device)
    check(lib.lem_time_wake(ptr(arrival), arrival.shape[0], ptr(wake), stream_ptr()),
          "lem_time_wake")
    return wake


def time_bins(arrival, wake, sampling_freq):
    """Start sample of each path (int64 tensor)."""
    torch = _torch()
    lib = load_library()
    arrival = to_dev(arrival, torch.float64).reshape(-1)
    wake = to_dev(wake, torch.float64).reshape(1)
    b = torch.empty(arrival.shape[0], dtype=torch.int64, device=arrival.device)
    check(lib.lem_time_bins(ptr(arrival), arrival.shape[0], ptr(wake), float(sampling_freq),
                            ptr(b), stream_ptr()), "lem_time_bins")
    return b


def time_map(amp, bins, slot, nslots, excitation, N):
    """(nslots, N, 3) float64 time map: the paths of each slot (slot[v] in
    [0, nslots)) summed in path order."""
    torch = _torch()
    lib = load_library()
    amp = to_dev(amp, torch.float64).reshape(-1, 3)
    bins = to_dev(bins, torch.int64).reshape(-1)
    exc = to_dev(np.asarray(excitation, np.float64), torch.float64).reshape(-1)
    slot = np.asarray(slot, np.int64)
    out = torch.zeros((int(nslots), int(N), 3), dtype=torch.float64, device=amp.device)
    if slot.shape[0] == 0 or N == 0:
        return out
    if slot.min() < 0 or slot.max() >= nslots:
        raise LemError("time_map: slot outside [0, nslots)")
    order = np.argsort(slot, kind="stable")
    ss = slot[order]
    starts = np.flatnonzero(np.r_[True, ss[1:] != ss[:-1]])
    seg_start = to_dev(np.append(starts, len(ss)).astype(np.int64), torch.int64)
    seg_slot = to_dev(ss[starts].astype(np.int64), torch.int64)
    order = to_dev(order.astype(np.int64), torch.int64)
    check(lib.lem_time_map(ptr(amp), ptr(bins), ptr(order), ptr(seg_start), ptr(seg_slot),
                           seg_slot.shape[0], ptr(exc), exc.shape[0], int(N), ptr(out),
                           stream_ptr()), "lem_time_map")
    return out


def time_trace(amp, bins, excitation, N):
    """(N, 3) float64: every path's placed excitation, summed (per start sample in
    path order, then convolved with the excitation)."""
    torch = _torch()
    lib = load_library()
    amp = to_dev(amp, torch.float64).reshape(-1, 3)
    bins = to_dev(bins, torch.int64).reshape(-1)
    exc = to_dev(np.asarray(excitation, np.float64), torch.float64).reshape(-1)
    out = torch.zeros((int(N), 3), dtype=torch.float64, device=amp.device)
    V = amp.shape[0]
    wsb = int(lib.lem_time_trace_workspace(V, int(N)))
    if wsb < 0:
        raise LemError("time_trace: too many paths or samples, or no HIP device for the "
                       "sort's size query")
    ws = torch.empty(max(wsb, 1), dtype=torch.uint8, device=amp.device)
    check(lib.lem_time_trace(ptr(amp), ptr(bins), V, ptr(exc), exc.shape[0], int(N), ptr(out),
                             ptr(ws), wsb, stream_ptr()), "lem_time_trace")
    return out


class Grid:
    """Device uniform grid over a triangle table (lem_grid_plan + lem_grid_fill)."""

    def __init__(self, tris, cells_per_tri=2.0):
        torch = _torch()
        lib = load_library()
        self.tris = to_dev(tris, torch.float32).reshape(-1, 9)
        T = self.tris.shape[0]
        if T == 0:
            raise LemError("Grid needs at least one triangle")
        self.cells_per_tri = float(cells_per_tri)
        wsb = int(lib.lem_grid_workspace(T, self.cells_per_tri))
        self.ws = torch.empty(wsb, dtype=torch.uint8, device=self.tris.device)
        self.info = GridInfo()
        check(lib.lem_grid_plan(ptr(self.tris), T, self.cells_per_tri, ctypes.byref(self.info),
                                ptr(self.ws), wsb, stream_ptr()), "lem_grid_plan")
        dev = self.tris.device
        self.cell_start = torch.empty(self.info.ncells + 1, dtype=torch.int32, device=dev)
        self.tri_ids = torch.empty(max(self.info.nrefs, 1), dtype=torch.int32, device=dev)
        self.recs = torch.empty((max(self.info.nrefs, 1), 12), dtype=torch.float32, device=dev)
        # float boxes of cells and macro blocks, then the 8-byte quantised macro table
        nbox = 8 * (self.info.ncells + self.info.nmacro) + 2 * self.info.nmacro
        self.cell_box = torch.empty(nbox, dtype=torch.float32, device=dev)
        # traversal workspaces (task counters, plane-ray queue, sort), one per
        # stream: calls on one stream run in order, calls on two streams must
        # not share the counters
        self.vis_ws = {}
        check(lib.lem_grid_fill(ptr(self.tris), T, ctypes.byref(self.info), ptr(self.cell_start),
                                ptr(self.tri_ids), ptr(self.recs), ptr(self.cell_box),
                                ptr(self.ws), wsb, self.cells_per_tri, stream_ptr()),
              "lem_grid_fill")
        self.ws = None

    def vis_workspace(self, S, K, flags):
        """lem_visibility_grid's workspace for (S, K, flags) on the current
        stream, kept and grown per stream (the traversal's task counters and
        plane-ray queue live in it, so two streams never share one)."""
        torch = _torch()
        wsb = int(load_library().lem_visibility_grid_workspace(S, K, flags))
        if wsb < 0:
            raise LemError("lem_visibility_grid_workspace: size query failed")
        key = torch.cuda.current_stream().cuda_stream
        ws = self.vis_ws.get(key)
        if ws is None or ws.numel() < wsb:
            ws = torch.empty(wsb, dtype=torch.uint8, device=self.tris.device)
            self.vis_ws[key] = ws
        return ws

    def visibility(self, src, snk, any_hit=True, want_dist=False, sink_major=False,
                   semantics=SEM_NUMBA, skip_self=False, input_order=False, from_sink=False):
        """Hit bits (S x W64(K), or K x W64(S) sink-major) and optional dist.
        input_order: lanes take the lane-side points in input order instead of
        a Morton order.  from_sink (any-hit): walk each ray from its sink end."""
        torch = _torch()
        lib = load_library()
        src = to_dev(src, torch.float32).reshape(-1, 3)
        snk = to_dev(snk, torch.float32).reshape(-1, 3)
        S, K = src.shape[0], snk.shape[0]
        rows, cols = (K, S) if sink_major else (S, K)
        bits = torch.zeros((rows, (cols + 63) // 64), dtype=torch.int64, device=src.device)
        dist = torch.empty((rows, cols), dtype=torch.float32, device=src.device) \
            if want_dist else None
        flags = (VIS_ANY_HIT if any_hit else 0) | (VIS_SINK_MAJOR if sink_major else 0) | \
            (VIS_SKIP_SELF if skip_self else 0) | (VIS_INPUT_ORDER if input_order else 0) | \
            (VIS_FROM_SINK if from_sink else 0)
        ws = self.vis_workspace(S, K, flags)
        check(lib.lem_visibility_grid(ptr(src), S, ptr(snk), K, ctypes.byref(self.info),
                                      ptr(self.cell_start), ptr(self.recs), ptr(self.cell_box),
                                      semantics, flags, ptr(bits), ptr(dist), ptr(ws),
                                      ws.numel(), stream_ptr()),
              "lem_visibility_grid")
        return bits, dist


def field_los_grid(grid, src, src_w, snk, snk_n, snk_w=None, alpha=0.0, beta=2 * np.pi,
                   sink_tangent=False, input_order=False):
    """Source-summed line-of-sight field with visibility cast through `grid`
    inside the field kernel: (K,3) complex64.  input_order: lanes take the
    sinks in input order instead of a Morton order (same result)."""
    torch = _torch()
    lib = load_library()
    src = to_dev(src, torch.float32).reshape(-1, 3)
    snk = to_dev(snk, torch.float32).reshape(-1, 3)
    snk_n = to_dev(snk_n, torch.float32).reshape(-1, 3)
    S, K = src.shape[0], snk.shape[0]
    out = torch.zeros((K, 3), dtype=torch.complex64, device=snk.device)
    if S == 0 or K == 0:
        return out
    src_w = to_dev(src_w, torch.complex64).reshape(S, 3)
    if snk_w is not None:
        snk_w = to_dev(snk_w, torch.complex64).reshape(K, 3)
    flags = (FIELD_SINK_TANGENT if sink_tangent else 0) | (FIELD_INPUT_ORDER if input_order else 0)
    wsb = int(lib.lem_field_los_grid_workspace(S, K, flags))
    if wsb < 0:
        raise LemError("lem_field_los_grid_workspace: size query failed")
    ws = torch.empty(max(wsb, 1), dtype=torch.uint8, device=snk.device)
    check(lib.lem_field_los_grid(ptr(src), ptr(src_w), S, ptr(snk), ptr(snk_n), ptr(snk_w), K,
                                 ctypes.byref(grid.info), ptr(grid.cell_start), ptr(grid.recs),
                                 ptr(grid.cell_box), float(alpha), float(beta), flags, ptr(out),
                                 ptr(ws), wsb, stream_ptr()), "lem_field_los_grid")
    return out


MASK_CHUNK_BYTES = 1 << 31

_POPCOUNT8 = None


def count_blocked(bits):
    """Number of set bits (blocked rays) in a hit-bit tensor; the padding bits
    of each row's last word are never set."""
    global _POPCOUNT8
    torch = _torch()
    if _POPCOUNT8 is None or _POPCOUNT8.device != bits.device:
        _POPCOUNT8 = torch.tensor([bin(i).count("1") for i in range(256)], dtype=torch.int64,
                                  device=bits.device)
    return int(_POPCOUNT8[bits.contiguous().view(torch.uint8).long()].sum().item())


def block_same_point(bits, first, n):
    """Mark the rays i -> i blocked in rows [first, first + n) of a mask whose
    sources and sinks are one point set (row r holds point first + r; its
    column is the point's own index): the path list never pairs a scattering
    point with itself (create_model_index drops such rows)."""
    torch = _torch()
    i = torch.arange(first, first + n, device=bits.device, dtype=torch.int64)
    bits[i - first, i >> 6] |= torch.bitwise_left_shift(torch.ones_like(i), i & 63)
    return bits


def field_los_grid_masked(grid, src, src_w, snk, snk_n, snk_w=None, alpha=0.0, beta=2 * np.pi,
                          sink_tangent=False, sink_major=True, from_sink=False,
                          chunk_bytes=MASK_CHUNK_BYTES, stats=None, same_points=False):
    """field_los_grid's result, computed as an any-hit mask from the persistent
    grid traversal followed by the masked field kernel.

    sink_major=True: each wave casts 64 neighbouring sources to one sink -- for
    compact sources (array elements) and spread-out sinks (scatterers over a
    scene).  sink_major=False: each wave casts one source to 64 neighbouring
    sinks -- for spread-out sources and compact sinks.  from_sink: walk each
    ray from its sink end, which reaches first the blockers that sit next to
    sinks on surfaces (a scatterer behind its own face); the mask is the same
    either way.  The mask is built in chunks (of sinks, or of sources) that stay
    within chunk_bytes.  stats (a dict, optional) receives "blocked", the number
    of blocked rays.  same_points: src and snk are the same point set and the
    pairs i -> i are left out.  Returns (K,3) complex64."""
    torch = _torch()
    src = to_dev(src, torch.float32).reshape(-1, 3)
    snk = to_dev(snk, torch.float32).reshape(-1, 3)
    snk_n = to_dev(snk_n, torch.float32).reshape(-1, 3)
    S, K = src.shape[0], snk.shape[0]
    out = torch.zeros((K, 3), dtype=torch.complex64, device=snk.device)
    if S == 0 or K == 0:
        return out
    src_w = to_dev(src_w, torch.complex64).reshape(S, 3)
    if snk_w is not None:
        snk_w = to_dev(snk_w, torch.complex64).reshape(K, 3)
    blocked = 0
    if sink_major:
        kc = max(1, int(chunk_bytes) // (8 * ((S + 63) // 64)))
        for k0 in range(0, K, kc):
            k1 = min(K, k0 + kc)
            bits, _ = grid.visibility(src, snk[k0:k1], any_hit=True, sink_major=True,
                                      from_sink=from_sink)
            if same_points:
                block_same_point(bits, k0, k1 - k0)
            if stats is not None:
                blocked += count_blocked(bits)
            out[k0:k1] = field_los(src, src_w, snk[k0:k1], snk_n[k0:k1], hit_bits=bits,
                                   snk_w=None if snk_w is None else snk_w[k0:k1], alpha=alpha,
                                   beta=beta, sink_tangent=sink_tangent, bits_sink_major=True)
    else:
        sc = max(1, int(chunk_bytes) // (8 * ((K + 63) // 64)))
        for s0 in range(0, S, sc):
            s1 = min(S, s0 + sc)
            bits, _ = grid.visibility(src[s0:s1], snk, any_hit=True, from_sink=from_sink)
            if same_points:
                block_same_point(bits, s0, s1 - s0)
            if stats is not None:
                blocked += count_blocked(bits)
            part = field_los(src[s0:s1], src_w[s0:s1], snk, snk_n, hit_bits=bits, snk_w=snk_w,
                             alpha=alpha, beta=beta, sink_tangent=sink_tangent)
            out = part if s0 == 0 else out + part
    if stats is not None:
        stats["blocked"] = blocked
    return out


def field_los_grid_sink_major(grid, src, src_w, snk, snk_n, snk_w=None, alpha=0.0,
                              beta=2 * np.pi, sink_tangent=False, chunk_bytes=MASK_CHUNK_BYTES,
                              from_sink=True, stats=None):
    """field_los_grid_masked with a sink-major mask walked from the sink end
    (the one-bounce leg from an aperture to scatterers)."""
    return field_los_grid_masked(grid, src, src_w, snk, snk_n, snk_w=snk_w, alpha=alpha,
                                 beta=beta, sink_tangent=sink_tangent, sink_major=True,
                                 from_sink=from_sink, chunk_bytes=chunk_bytes, stats=stats)


def grid_stats(grid, src, snk, any_hit=False):
    """(MT tests, cells visited, blocked rays, macro blocks entered, conservative
    DDA tests up to the first hit) of a sink-major grid traversal."""
    torch = _torch()
    lib = load_library()
    src = to_dev(src, torch.float32).reshape(-1, 3)
    snk = to_dev(snk, torch.float32).reshape(-1, 3)
    stats = torch.zeros(5, dtype=torch.int64, device=src.device)
    check(lib.lem_visibility_grid_stats(ptr(src), src.shape[0], ptr(snk), snk.shape[0],
                                        ctypes.byref(grid.info), ptr(grid.cell_start),
                                        ptr(grid.recs), ptr(grid.cell_box),
                                        VIS_ANY_HIT if any_hit else 0, ptr(stats), stream_ptr()),
          "lem_visibility_grid_stats")
    return tuple(int(x) for x in stats.cpu().numpy())


def gw_mfma(src, pol, snk, snk_n, wavenumbers, W, alpha=0.0):
    """E[f, c, k, b] = sum_s G_c[k, s, f] W[f, s, b] on MFMA; W: (F, S, 64) complex64.
    Returns (F, 3, K, 64) complex64."""
    torch = _torch()
    lib = load_library()
    src = to_dev(src, torch.float32).reshape(-1, 3)
    snk = to_dev(snk, torch.float32).reshape(-1, 3)
    snk_n = to_dev(snk_n, torch.float32).reshape(-1, 3)
    kf = to_dev(np.asarray(wavenumbers, np.float64), torch.float64).reshape(-1)
    W = to_dev(W, torch.complex64)
    S, K, F = src.shape[0], snk.shape[0], kf.shape[0]
    if tuple(W.shape) != (F, S, 64):
        raise LemError(f"W shape {tuple(W.shape)} != {(F, S, 64)}")
    out = torch.empty((F, 3, K, 64), dtype=torch.complex64, device=src.device)
    pol = [float(x) for x in np.asarray(pol, np.float32).ravel()]
    check(lib.lem_gw_mfma(ptr(src), S, pol[0], pol[1], pol[2], ptr(snk), ptr(snk_n), K, ptr(kf),
                          F, float(alpha), ptr(W), 64, ptr(out), stream_ptr()), "lem_gw_mfma")
    return out


def unpack_bits(bits, K):
    """(S, ceil(K/64)) uint64 words -> (S, K) bool (host numpy)."""
    b = bits.detach().cpu().numpy().view(np.uint64)
    S = b.shape[0]
    by = b.view(np.uint8).reshape(S, -1)
    return np.unpackbits(by, axis=1, bitorder="little")[:, :K].astype(bool)
