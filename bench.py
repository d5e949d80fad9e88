#!/usr/bin/env python
"""Benchmark of the LyceanEM raycast + frequency-domain hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c1|c2|c3|c4|c5|td|bf]
                    [--no-cpu]

One "step" is one pass of the hot path over the configuration's synthetic scene
with every input already resident in HBM:
  c2  (default) 1000 x 500 quad (1,000,000 triangle) terrain, 1024 x 1024
      sources 30 m up, 16,384 sinks 1.5 m up: uniform-grid DDA visibility of all
      1.718e10 source -> sink rays, writing the hit mask and the kernel1Dv3
      closest-hit distance of every ray (grid built once, outside the step).
  c1  32x32 lambda/2 array (3 GHz), 2,000 random triangles, 91 x 181 far-field
      sinks: brute-force visibility with kernel1Dv3 semantics (closest hit,
      every ray tests every triangle) + the fused masked field (source sum).
  c3  one bounce + line of sight: 256 x 256 array, 50k triangles, 200k
      scatterers, 64 x 64 sinks (grid masks and masked fields, leg by leg).
  c4  64 x 64 conformal array: masked far field forward and its adjoint.
  c5  per-element network times 64 excitation sets on FP32 MFMA.
  td  time domain: 16.8M line-of-sight paths placed into a 65,536-sample trace.
  bf  MaximumDirectivityMap: every grid point as a command angle, three weight
      schemes, beamformed power on FP32 MFMA.

The dominant kernel's average duration comes from HIP events recorded on the
stream the kernels run on (torch's current stream).  The CPU baseline is the
oracle (C + OpenMP restatement of the reference algorithm) on a bounded sample
of the same workload, on this host's cores.  Rank 0 prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "lyceanem-python_amd")
for _p in (ROOT, PKG):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402

MT_FLOPS = 46          # FP32 flops of one Moller-Trumbore test (+1 divide)
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "k_visibility_grid_lds_c2_terrain_1M_traffic.md")


def _traffic_per_ray():
    """HBM bytes per ray of the grid traversal, from the rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary (copy-calibrated) kept with the profiles, or None."""
    try:
        for line in open(TRAFFIC_PROFILE):
            if line.startswith("traffic_bytes_per_ray:"):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None
PAIR_FLOPS = 92        # FP32 flops of one Green's-function pair
PEAK_FP32 = 157.3e12   # MI355X FP32 vector peak (packed)
PEAK_HBM = 8.0e12
PEAK_FP64 = 78.6e12    # MI355X FP64 vector peak


def _events(torch, n):
    return [torch.cuda.Event(enable_timing=True) for _ in range(n)]


def _bits_at(torch, bits, rows, cols):
    """Bits (rows[i], cols[i]) of a (rows, words) int64 device tensor -> bool array."""
    r = torch.from_numpy(np.asarray(rows, np.int64)).to(bits.device)
    c = torch.from_numpy(np.asarray(cols, np.int64)).to(bits.device)
    return ((bits[r, c >> 6] >> (c & 63)) & 1).bool().cpu().numpy()


def _field_errors(got, ref):
    """Normwise relative error against the float64 oracle, the worst relative
    error of the entries with |ref| >= 1e-3 max |ref| (the rest sit in pattern
    nulls) and the worst phase error of those entries (rad)."""
    got = np.asarray(got, np.complex128).reshape(-1)
    ref = np.asarray(ref, np.complex128).reshape(-1)
    big = np.abs(ref) >= 1e-3 * np.abs(ref).max()
    return {"normwise_rel": float(np.linalg.norm(got - ref) / np.linalg.norm(ref)),
            "max_rel_entry": float((np.abs(got[big] - ref[big]) / np.abs(ref[big])).max()),
            "max_phase_rad": float(np.abs(np.angle(got[big] / ref[big])).max()),
            "entries": int(got.size), "entries_compared_per_entry": int(big.sum())}


def _mismatch(name, got, ref):
    return {f"{name}_mismatches": int((np.asarray(got) != np.asarray(ref)).sum())}


# --------------------------------------------------------------------------
# config 1: brute-force visibility + line-of-sight field
# --------------------------------------------------------------------------

class C1:
    name = "c1"

    def __init__(self, torch):
        from lyceanem import _native, scenes

        self.torch, self.nat = torch, _native
        self.lib = _native.load_library()
        sc = scenes.planar_array_c1()
        self.sc = sc
        dev = _native.device()
        f32 = torch.float32
        self.src = torch.from_numpy(sc["src"]).to(dev)
        self.snk = torch.from_numpy(sc["snk"]).to(dev)
        self.snk_n = torch.from_numpy(sc["snk_n"]).to(dev)
        self.w = torch.from_numpy(sc["w"]).to(dev)
        self.tris = torch.from_numpy(sc["tris"]).to(dev)
        self.S, self.K, self.T = len(sc["src"]), len(sc["snk"]), len(sc["tris"])
        W = (self.K + 63) // 64
        self.bits = torch.zeros((self.S, W), dtype=torch.int64, device=dev)
        self.dist = torch.empty((self.S, self.K), dtype=f32, device=dev)
        self.out = torch.empty((self.K, 3), dtype=torch.complex64, device=dev)
        self.wsb = int(self.lib.lem_field_los_workspace(self.S, self.K, 0))
        self.ws = torch.empty(self.wsb, dtype=torch.uint8, device=dev)
        self.ev = None

    def rays(self):
        return self.S * self.K

    def step(self, ev=None):
        nat, lib, p = self.nat, self.lib, self.nat.ptr
        st = nat.stream_ptr()
        if ev is not None:
            ev[0].record()
        nat.check(lib.lem_visibility_dense(p(self.src), self.S, p(self.snk), self.K, p(self.tris),
                                           self.T, nat.SEM_NUMBA, 0, p(self.bits), p(self.dist),
                                           st), "visibility")
        if ev is not None:
            ev[1].record()
        nat.check(lib.lem_field_los(p(self.src), p(self.w), self.S, p(self.snk), p(self.snk_n),
                                    None, self.K, p(self.bits), 0.0, float(self.sc["beta"]), 0,
                                    p(self.out), p(self.ws), self.wsb, st), "field")
        if ev is not None:
            ev[2].record()

    def record(self, times_ms):
        vis = float(np.mean([a for a, _ in times_ms]))
        fld = float(np.mean([b for _, b in times_ms]))
        flops = self.rays() * self.T * MT_FLOPS
        achieved = flops / (vis * 1e-3)
        return {
            "kernel": "k_visibility_dense<numba, closest-hit>",
            "kernel_ms": vis,
            "roofline": {"bound": "valu", "achieved": achieved / 1e12, "peak": PEAK_FP32 / 1e12,
                         "unit": "TFLOP/s", "frac": achieved / PEAK_FP32, "traffic": None},
            "extra": {"field_kernel_ms": fld,
                      "field_gpairs_per_s": self.rays() / (fld * 1e-3) / 1e9,
                      "mt_tests": self.rays() * self.T},
        }

    def cpu_baseline(self, budget_s):
        from oracle import oracle as orc

        sc = self.sc
        nth = orc.nthreads_default()
        rows = 1
        t0 = time.perf_counter()
        orc.visibility_dense(sc["src"][:rows], sc["snk"], sc["tris"], want_dist=True,
                             nthreads=nth)
        dt = time.perf_counter() - t0
        rows = int(max(1, min(self.S, budget_s / max(dt, 1e-6))))
        t0 = time.perf_counter()
        hit, dist = orc.visibility_dense(sc["src"][:rows], sc["snk"], sc["tris"], want_dist=True,
                                         nthreads=nth)
        orc.field_los(sc["src"][:rows], sc["w"][:rows], sc["snk"], sc["snk_n"], 0.0,
                      sc["beta"], hit=hit, nthreads=nth)
        dt = time.perf_counter() - t0
        self.cpu_sample = (rows, hit, dist)
        return {"value": rows * self.K / dt / 1e6, "unit": "Mrays/s", "cores": nth,
                "kind": "port",
                "sample": f"first {rows} of {self.S} source rows x {self.K} sinks x {self.T} "
                          f"triangles (oracle closest-hit visibility + float64 field), "
                          f"{dt:.1f} s"}

    cpu_sample = None

    def parity(self):
        """Masks and distances of the last step against the oracle's brute force
        (the CPU baseline's source rows, else the first 8); the field against
        the float64 oracle over every pair, given the step's mask."""
        from oracle import oracle as orc

        sc = self.sc
        if self.cpu_sample is None:
            rows = 8
            hit, dist = orc.visibility_dense(sc["src"][:rows], sc["snk"], sc["tris"])
        else:
            rows, hit, dist = self.cpu_sample
        mask = self.nat.unpack_bits(self.bits, self.K)
        out = {"rays_vs_oracle": rows * self.K,
               **_mismatch("mask", mask[:rows], hit),
               **_mismatch("dist", self.dist[:rows].cpu().numpy(), dist),
               "blocked_fraction": float(mask.mean())}
        ref = orc.field_los(sc["src"], sc["w"], sc["snk"], sc["snk_n"], 0.0, sc["beta"],
                            hit=mask)
        out["field_vs_fp64_all_pairs"] = _field_errors(self.out.cpu().numpy(), ref)
        return out

    def config(self):
        return {"workload": "c1: 32x32 array, 2000 random triangles, 91x181 sinks at 1 km, 3 GHz",
                "sources": self.S, "sinks": self.K, "triangles": self.T, "rays": self.rays()}


# --------------------------------------------------------------------------
# config 2: 1M-triangle terrain, grid DDA visibility of 1.7e10 rays
# --------------------------------------------------------------------------

class C2:
    name = "c2"
    stats_stride = 8  # MT tests counted on every 8th sink, scaled

    def __init__(self, torch, n_sinks=16384, cells_per_tri=2.0, input_order=False):
        from lyceanem import _native, scenes

        self.torch, self.nat = torch, _native
        self.input_order = input_order
        self.lib = _native.load_library()
        t0 = time.perf_counter()
        sc = scenes.terrain_c2(n_snk=n_sinks)
        self.sc = sc
        self.gen_s = time.perf_counter() - t0
        dev = _native.device()
        self.src = torch.from_numpy(sc["src"]).to(dev)
        self.snk = torch.from_numpy(sc["snk"]).to(dev)
        self.tris_t = torch.from_numpy(sc["tris"]).to(dev)
        self.S, self.K, self.T = len(sc["src"]), len(sc["snk"]), len(sc["tris"])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        self.grid = _native.Grid(self.tris_t, cells_per_tri=cells_per_tri)
        torch.cuda.synchronize()
        self.build_s = time.perf_counter() - t0
        W = (self.S + 63) // 64
        self.bits = torch.empty((self.K, W), dtype=torch.int64, device=dev)
        self.dist = torch.empty((self.K, self.S), dtype=torch.float32, device=dev)

    def rays(self):
        return self.S * self.K

    def step(self, ev=None):
        nat, lib, p, g = self.nat, self.lib, self.nat.ptr, self.grid
        flags = nat.VIS_SINK_MAJOR | (nat.VIS_INPUT_ORDER if self.input_order else 0)
        ws = g.vis_workspace(self.S, self.K, flags)
        if ev is not None:
            ev[0].record()
        nat.check(lib.lem_visibility_grid(p(self.src), self.S, p(self.snk), self.K,
                                          nat.ctypes.byref(g.info), p(g.cell_start), p(g.recs),
                                          p(g.cell_box), nat.SEM_NUMBA, flags, p(self.bits),
                                          p(self.dist), p(ws), ws.numel(), nat.stream_ptr()),
                  "visibility_grid")
        if ev is not None:
            ev[1].record()
            ev[2].record()

    def record(self, times_ms):
        vis = float(np.mean([a for a, _ in times_ms]))
        sub = self.snk[:: self.stats_stride].contiguous()
        tests, cells, blocked, macros, conservative = self.nat.grid_stats(self.grid, self.src,
                                                                         sub)
        scale = self.K / sub.shape[0]
        # algorithmic work: the MT tests of a conservative front-to-back DDA up to
        # the first hit (every reference of every crossed cell), 46 flops each
        work_all = conservative * scale
        achieved = work_all * MT_FLOPS / (vis * 1e-3)
        executed = tests * scale * MT_FLOPS / (vis * 1e-3)
        bytes_out = self.bits.numel() * 8 + self.dist.numel() * 4
        per_ray = _traffic_per_ray()
        traffic = None if per_ray is None else per_ray * self.rays()
        return {
            "kernel": "k_visibility_grid_lds<closest-hit, sink-major>",
            "kernel_ms": vis,
            "roofline": {"bound": "valu", "achieved": achieved / 1e12, "peak": PEAK_FP32 / 1e12,
                         "unit": "TFLOP/s", "frac": achieved / PEAK_FP32,
                         "basis": "algorithmic: 46 FP32 flops x the MT tests a conservative "
                                  "front-to-back DDA performs up to the first hit (every "
                                  "triangle reference of every crossed cell)",
                         "achieved_executed": executed / 1e12,
                         "frac_executed": executed / PEAK_FP32,
                         "basis_executed": "46 FP32 flops x the MT tests the kernel performs",
                         "traffic": traffic,
                         "traffic_source": None if traffic is None else
                         "profiles/" + os.path.basename(TRAFFIC_PROFILE) + " (bytes per ray x "
                         "rays of this launch)"},
            "extra": {
                "grid_build_s": self.build_s,
                "lane_order": "input" if self.input_order else "morton",
                "grid_cells": [int(x) for x in self.grid.info.n],
                "grid_refs": int(self.grid.info.nrefs),
                "conservative_mt_tests_per_ray": conservative / (self.S * sub.shape[0]),
                "mt_tests_per_ray": tests / (self.S * sub.shape[0]),
                "cells_per_ray": cells / (self.S * sub.shape[0]),
                "macro_blocks_per_ray": macros / (self.S * sub.shape[0]),
                "blocked_fraction": blocked / (self.S * sub.shape[0]),
                "work_counted_on": f"every {self.stats_stride}th sink ({sub.shape[0]} sinks), "
                                   f"scaled x{scale:g}",
                "output_bytes": bytes_out,
                "output_write_GBps": bytes_out / (vis * 1e-3) / 1e9,
            },
        }

    def cpu_baseline(self, budget_s):
        from oracle import oracle as orc

        sc = self.sc
        nth = orc.nthreads_default()
        rng = np.random.default_rng(0)
        n = 64
        pairs = np.stack([rng.integers(0, self.S, n), rng.integers(0, self.K, n)], 1)
        t0 = time.perf_counter()
        orc.visibility_sampled(sc["src"], sc["snk"], pairs, sc["tris"], nthreads=nth)
        dt = time.perf_counter() - t0
        n = int(max(64, min(10 ** 6, n * budget_s / max(dt, 1e-6))))
        pairs = np.stack([rng.integers(0, self.S, n), rng.integers(0, self.K, n)], 1)
        t0 = time.perf_counter()
        hit, dist = orc.visibility_sampled(sc["src"], sc["snk"], pairs, sc["tris"], nthreads=nth)
        dt = time.perf_counter() - t0
        self.cpu_sample = (pairs, hit, dist)
        self.cpu_tiles = self._cpu_tiles(budget_s / 2, nth)
        return {"value": n / dt / 1e6, "unit": "Mrays/s", "cores": nth, "kind": "port",
                "sample": f"{n} random source->sink rays of the 1.718e10, brute force over all "
                          f"{self.T} triangles (kernel1Dv3 closest hit, the reference "
                          f"algorithm), {dt:.1f} s"}

    cpu_tiles = None

    def _cpu_tiles(self, budget_s, nth):
        """The reference's own acceleration structure on the CPU: the oracle's
        restatement of Tile_acceleration_structure (vertex binning into 2 m
        tiles over x, y; the tile rectangle between where each ray enters and
        leaves the terrain's height slab scanned until a triangle hits), any-hit
        masks with the numba predicate, on random rays of the launch; masks
        compared with the GPU's."""
        from oracle import oracle as orc

        sc, torch = self.sc, self.torch
        t0 = time.perf_counter()
        tiles = orc.Tiles(sc["tris"], up=2, n_cells=250)
        build_s = time.perf_counter() - t0
        rng = np.random.default_rng(1)
        n = 256
        pairs = np.stack([rng.integers(0, self.S, n), rng.integers(0, self.K, n)], 1)
        t0 = time.perf_counter()
        tiles.cast(sc["src"], sc["snk"], pairs, nthreads=nth)
        dt = time.perf_counter() - t0
        n = int(max(256, min(10 ** 7, n * budget_s / max(dt, 1e-6))))
        pairs = np.stack([rng.integers(0, self.S, n), rng.integers(0, self.K, n)], 1)
        t0 = time.perf_counter()
        hit = tiles.cast(sc["src"], sc["snk"], pairs, nthreads=nth)
        dt = time.perf_counter() - t0
        m_g = _bits_at(torch, self.bits, pairs[:, 1], pairs[:, 0])
        return {"value": n / dt / 1e6, "unit": "Mrays/s", "cores": nth, "kind": "port",
                "sample": f"{n} random rays of the launch, {dt:.1f} s; tiles "
                          f"{tiles.na} x {tiles.nb} of {tiles.tile:.3f} m, {tiles.refs} "
                          f"references, built in {build_s:.2f} s",
                "mask_mismatches_vs_gpu": int((m_g != hit).sum())}

    cpu_sample = None

    def parity(self, n=1_000_000):
        """The last step's mask and distances on n random rays against the
        brute-force kernel over every triangle (lem_raycast_pairs, bit-exact
        against the oracle in the tests), and on the CPU baseline's rays (or
        10,000 rays) against the oracle itself."""
        from oracle import oracle as orc

        torch, sc = self.torch, self.sc
        rng = np.random.default_rng(12345)
        s_i, k_i = rng.integers(0, self.S, n), rng.integers(0, self.K, n)
        m_g = _bits_at(torch, self.bits, k_i, s_i)
        dev = self.dist.device
        d_g = self.dist[torch.from_numpy(k_i).to(dev), torch.from_numpy(s_i).to(dev)]
        pts = torch.cat([self.src, self.snk])
        pairs = torch.from_numpy(np.stack([s_i, self.S + k_i], 1).astype(np.int32)).to(dev)
        inter, bd = self.nat.raycast_pairs(pts, pairs, self.tris_t, any_hit=False,
                                           want_dist=True)
        m_b = inter.cpu().numpy().astype(bool)
        out = {"gpu_brute_force": {"rays": n, **_mismatch("mask", m_g, m_b),
                                   **_mismatch("dist", d_g.cpu().numpy(), bd.cpu().numpy()),
                                   "blocked_fraction": float(m_b.mean())}}
        if self.cpu_sample is None:
            pairs_c = np.stack([s_i[:10000], k_i[:10000]], 1)
            hit, dist = orc.visibility_sampled(sc["src"], sc["snk"], pairs_c, sc["tris"])
        else:
            pairs_c, hit, dist = self.cpu_sample
        m_c = _bits_at(torch, self.bits, pairs_c[:, 1], pairs_c[:, 0])
        d_c = self.dist[torch.from_numpy(pairs_c[:, 1]).to(dev),
                        torch.from_numpy(pairs_c[:, 0]).to(dev)].cpu().numpy()
        out["oracle"] = {"rays": int(len(pairs_c)), **_mismatch("mask", m_c, hit),
                         **_mismatch("dist", d_c, dist)}
        return out

    def config(self):
        return {"workload": "c2: 1000x500-quad terrain (1,000,000 triangles, Gaussian heights "
                            "sigma 5 m, corr 50 m), 1024x1024 sources at +30 m, 16384 sinks at "
                            "+1.5 m; grid DDA, mask + closest-hit distance per ray",
                "sources": self.S, "sinks": self.K, "triangles": self.T, "rays": self.rays()}


# --------------------------------------------------------------------------
# config 3: one-bounce scattering network (fused grid legs)
# --------------------------------------------------------------------------

class C3:
    name = "c3"

    def __init__(self, torch, from_sink=True, cells_per_tri=2.0):
        from lyceanem import _native, scenes

        self.torch, self.nat = torch, _native
        self.from_sink = from_sink
        sc = scenes.clutter_c3()
        self.sc = sc
        dev = _native.device()
        t = {k: torch.from_numpy(np.ascontiguousarray(v)).to(dev)
             for k, v in sc.items() if isinstance(v, np.ndarray)}
        self.t = t
        self.grid = _native.Grid(t["tris"], cells_per_tri=cells_per_tri)
        self.S, self.P, self.K = len(sc["src"]), len(sc["scat"]), len(sc["snk"])
        self.T = len(sc["tris"])

    def rays(self):
        return self.S * self.P + self.P * self.K + self.S * self.K

    def step(self, ev=None):
        nat, t, g, b = self.nat, self.t, self.grid, self.sc["beta"]
        if ev is not None:
            ev[0].record()
        # leg 1 (array -> scatterers): sink-major any-hit mask, 64 neighbouring
        # elements to one scatterer per wave, each ray walked from the
        # scatterer end, then the masked field
        bits, _ = g.visibility(t["src"], t["scat"], any_hit=True, sink_major=True,
                               from_sink=self.from_sink)
        if ev is not None:
            ev[1].record()
        H = nat.field_los(t["src"], t["w"], t["scat"], t["scat_n"], hit_bits=bits,
                          snk_w=t["scat_w"], beta=b, sink_tangent=True, bits_sink_major=True)
        del bits
        # leg 2 and line of sight through the same persistent traversal
        E = nat.field_los_grid_masked(g, t["scat"], H, t["snk"], t["snk_n"], beta=b,
                                      sink_major=False)
        E += nat.field_los_grid_masked(g, t["src"], t["w"], t["snk"], t["snk_n"], beta=b,
                                       sink_tangent=True, sink_major=True)
        if ev is not None:
            ev[2].record()
        self.E = E

    stats_stride = 100

    def record(self, times_ms):
        vis = float(np.mean([a for a, _ in times_ms]))
        rest = float(np.mean([b for _, b in times_ms]))
        pairs1 = self.S * self.P
        sub = self.t["scat"][:: self.stats_stride].contiguous()
        tests, cells, blocked, macros, conservative = self.nat.grid_stats(
            self.grid, self.t["src"], sub, any_hit=True)
        n_sub = self.S * sub.shape[0]
        achieved = conservative * (self.P / sub.shape[0]) * MT_FLOPS / (vis * 1e-3)
        executed = tests * (self.P / sub.shape[0]) * MT_FLOPS / (vis * 1e-3)
        return {
            "kernel": "k_visibility_grid_lds<any-hit, sink-major> (leg 1 mask: array -> "
                      "scatterers)",
            "kernel_ms": vis,
            "roofline": {"bound": "valu", "achieved": achieved / 1e12, "peak": PEAK_FP32 / 1e12,
                         "unit": "TFLOP/s", "frac": achieved / PEAK_FP32,
                         "basis": "algorithmic: 46 FP32 flops x the MT tests a conservative "
                                  "front-to-back DDA performs up to the first hit",
                         "achieved_executed": executed / 1e12,
                         "frac_executed": executed / PEAK_FP32,
                         "basis_executed": "46 FP32 flops x the MT tests the kernel performs",
                         "traffic": None},
            "extra": {"gpairs_per_s": self.rays() / ((vis + rest) * 1e-3) / 1e9,
                      "leg1_walk": "from scatterer" if self.from_sink else "from array",
                      "leg1_mask_grays_per_s": pairs1 / (vis * 1e-3) / 1e9,
                      "leg1_field_legs2_los_ms": rest,
                      "grid_cells": [int(x) for x in self.grid.info.n],
                      "conservative_mt_tests_per_ray": conservative / n_sub,
                      "mt_tests_per_ray": tests / n_sub, "cells_per_ray": cells / n_sub,
                      "blocked_fraction": blocked / n_sub,
                      "work_counted_on": f"every {self.stats_stride}th scatterer "
                                         f"({sub.shape[0]}), scaled"},
        }

    def cpu_baseline(self, budget_s):
        from oracle import oracle as orc

        sc = self.sc
        nth = orc.nthreads_default()
        rng = np.random.default_rng(0)
        n = 32
        t0 = time.perf_counter()
        pairs = np.stack([rng.integers(0, self.S, n), rng.integers(0, self.P, n)], 1)
        orc.visibility_sampled(sc["src"], sc["scat"], pairs, sc["tris"], nthreads=nth)
        dt = time.perf_counter() - t0
        n = int(max(32, min(10 ** 6, n * budget_s / max(dt, 1e-6))))
        pairs = np.stack([rng.integers(0, self.S, n), rng.integers(0, self.P, n)], 1)
        t0 = time.perf_counter()
        hit, _ = orc.visibility_sampled(sc["src"], sc["scat"], pairs, sc["tris"], nthreads=nth)
        dt = time.perf_counter() - t0
        return {"value": n / dt / 1e9, "unit": "Gpairs/s", "cores": nth, "kind": "port",
                "sample": f"{n} random source->scatterer pairs, brute-force visibility over "
                          f"{self.T} triangles (the field term is negligible beside it), "
                          f"{dt:.1f} s"}

    def parity(self, n_rays=100_000, n_scat=512, n_sinks=256):
        """Outside the timed region: the three grid masks as the step casts them,
        on random rays against the oracle's brute force over every triangle; the
        scatterer fields H of leg 1 on n_scat random scatterers and the step's
        sink fields on n_sinks random sinks against the float64 oracle given the
        GPU masks (leg 2 takes the GPU's H as its input)."""
        from oracle import oracle as orc

        torch, nat, t, g, sc = self.torch, self.nat, self.t, self.grid, self.sc
        b = sc["beta"]
        S, P, K = self.S, self.P, self.K
        rng = np.random.default_rng(54321)
        out = {}
        bits1, _ = g.visibility(t["src"], t["scat"], any_hit=True, sink_major=True,
                                from_sink=self.from_sink)
        s_i, p_i = rng.integers(0, S, n_rays), rng.integers(0, P, n_rays)
        ref, _ = orc.visibility_sampled(sc["src"], sc["scat"], np.stack([s_i, p_i], 1),
                                        sc["tris"], any_hit_only=True)
        out["leg1_mask_vs_oracle"] = {"rays": n_rays, **_mismatch(
            "mask", _bits_at(torch, bits1, p_i, s_i), ref), "blocked_fraction": float(ref.mean())}
        H = nat.field_los(t["src"], t["w"], t["scat"], t["scat_n"], hit_bits=bits1,
                          snk_w=t["scat_w"], beta=b, sink_tangent=True, bits_sink_major=True)
        ps = np.sort(rng.choice(P, n_scat, replace=False))
        hit1 = nat.unpack_bits(bits1[torch.from_numpy(ps).to(bits1.device)], S).T
        del bits1
        H_ref = orc.field_los(sc["src"], sc["w"], sc["scat"][ps], sc["scat_n"][ps], 0.0, b,
                              hit=hit1, snk_w=sc["scat_w"][ps], sink_tangent=True)
        H_np = H.cpu().numpy()
        out["leg1_field_vs_fp64"] = {"scatterers": n_scat, **_field_errors(H_np[ps], H_ref)}
        bits2, _ = g.visibility(t["scat"], t["snk"], any_hit=True)
        q_i, k_i = rng.integers(0, P, n_rays), rng.integers(0, K, n_rays)
        ref, _ = orc.visibility_sampled(sc["scat"], sc["snk"], np.stack([q_i, k_i], 1),
                                        sc["tris"], any_hit_only=True)
        out["leg2_mask_vs_oracle"] = {"rays": n_rays, **_mismatch(
            "mask", _bits_at(torch, bits2, q_i, k_i), ref), "blocked_fraction": float(ref.mean())}
        bits3, _ = g.visibility(t["src"], t["snk"], any_hit=True, sink_major=True)
        m = n_rays // 5
        s3, k3 = rng.integers(0, S, m), rng.integers(0, K, m)
        ref, _ = orc.visibility_sampled(sc["src"], sc["snk"], np.stack([s3, k3], 1), sc["tris"],
                                        any_hit_only=True)
        out["los_mask_vs_oracle"] = {"rays": m, **_mismatch(
            "mask", _bits_at(torch, bits3, k3, s3), ref), "blocked_fraction": float(ref.mean())}
        ks = np.sort(rng.choice(K, n_sinks, replace=False))
        pp, kk = np.meshgrid(np.arange(P), ks, indexing="ij")
        hit2 = _bits_at(torch, bits2, pp.ravel(), kk.ravel()).reshape(P, n_sinks)
        hit3 = nat.unpack_bits(bits3[torch.from_numpy(ks).to(bits3.device)], S).T
        E_ref = orc.field_los(sc["scat"], H_np, sc["snk"][ks], sc["snk_n"][ks], 0.0, b, hit=hit2)
        E_ref = E_ref + orc.field_los(sc["src"], sc["w"], sc["snk"][ks], sc["snk_n"][ks], 0.0, b,
                                      hit=hit3, sink_tangent=True)
        out["sink_field_vs_fp64"] = {"sinks": n_sinks,
                                     **_field_errors(self.E.cpu().numpy()[ks], E_ref)}
        return out

    def config(self):
        return {"workload": "c3: 256x256 array at 5.8 GHz, 50k random triangles, 200k "
                            "scatterers, 64x64 sinks; one bounce + line of sight, visibility on "
                            "both legs", "sources": self.S, "scatterers": self.P,
                "sinks": self.K, "triangles": self.T, "pairs": self.rays()}


# --------------------------------------------------------------------------
# config 4: beam synthesis, adjoint vs forward
# --------------------------------------------------------------------------

class C4:
    name = "c4"

    def __init__(self, torch):
        from lyceanem import _native, scenes

        self.torch, self.nat = torch, _native
        sc = scenes.conformal_c4()
        self.sc = sc
        dev = _native.device()
        self.grid = _native.Grid(sc["tris"])
        self.bits, _ = self.grid.visibility(sc["src"], sc["snk"], any_hit=True)
        self.pos = torch.from_numpy(sc["src"]).to(dev)
        self.w = torch.from_numpy(sc["w"]).to(dev)
        self.snk = torch.from_numpy(sc["snk"]).to(dev)
        self.snk_n = torch.from_numpy(sc["snk_n"]).to(dev)
        self.in_beam = torch.from_numpy(sc["in_beam"]).to(dev)
        self.S, self.K = len(sc["src"]), len(sc["snk"])
        # the forward over a spatial order of the sinks (fixed sinks and mask)
        from lyceanem.differentiable import FarFieldPlan
        self.plan = FarFieldPlan(self.snk, self.snk_n, self.bits, center=sc["src"].mean(0),
                                 sources=self.pos)
        # the gradient step (forward, loss, adjoint: all on the device) captured
        # once in a HIP graph; each timed step copies the current weights and
        # positions into the captured leaves and replays it
        self.wg = self.w.clone().requires_grad_(True)
        self.pg = self.pos.clone().requires_grad_(True)
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            for _ in range(3):
                self._eager(self.wg, self.pg)
        torch.cuda.current_stream().wait_stream(side)
        self.wg.grad = None
        self.pg.grad = None
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self.g_loss = self._eager(self.wg, self.pg)

    def rays(self):
        return self.S * self.K

    def _eager(self, w, pos, ev=None):
        from lyceanem.differentiable import far_field, db_mask_loss

        if ev is not None:
            ev[0].record()
        E = far_field(w, pos, self.snk, self.snk_n, hit_bits=self.bits, beta=self.sc["beta"],
                      plan=self.plan)
        if ev is not None:
            ev[1].record()
        loss = db_mask_loss(E, self.in_beam)
        loss.backward()
        if ev is not None:
            ev[2].record()
        return loss.detach()

    def step(self, ev=None):
        if ev is not None:
            ev[0].record()
        with self.torch.no_grad():
            self.wg.copy_(self.w)
            self.pg.copy_(self.pos)
        self.graph.replay()
        if ev is not None:
            ev[1].record()
            ev[2].record()
        self.loss = self.g_loss

    def record(self, times_ms):
        torch = self.torch
        graph_ms = float(np.mean([a for a, _ in times_ms]))
        # forward / loss + adjoint split from eager steps timed with events
        # (after two untimed eager steps, which set up the allocator's blocks)
        for _ in range(2):
            self._eager(self.w.clone().requires_grad_(True),
                        self.pos.clone().requires_grad_(True))
        evs = [_events(torch, 3) for _ in range(20)]
        for e in evs:
            self._eager(self.w.clone().requires_grad_(True), self.pos.clone().requires_grad_(True),
                        e)
        torch.cuda.synchronize()
        fwd = float(np.mean([e[0].elapsed_time(e[1]) for e in evs]))
        bwd = float(np.mean([e[1].elapsed_time(e[2]) for e in evs]))
        pairs = self.rays()
        return {
            "kernel": "k_field_los_partial (forward) / k_adjoint_partial (backward)",
            "kernel_ms": fwd,
            "roofline": {"bound": "valu", "achieved": pairs * PAIR_FLOPS / (fwd * 1e-3) / 1e12,
                         "peak": PEAK_FP32 / 1e12, "unit": "TFLOP/s",
                         "frac": pairs * PAIR_FLOPS / (fwd * 1e-3) / PEAK_FP32, "traffic": None},
            "extra": {"graph_step_ms": graph_ms,
                      "forward_ms": fwd, "loss_and_adjoint_ms": bwd,
                      "adjoint_over_forward": bwd / fwd,
                      "gradient_step_over_forward": (fwd + bwd) / fwd,
                      "forward_gpairs_per_s": pairs / (fwd * 1e-3) / 1e9,
                      "plan_sink_order": self.plan.sink_order,
                      "plan_forward_ms": self.plan.forward_times,
                      "plan_source_order": self.plan.adjoint_order,
                      "plan_adjoint_ms": self.plan.adjoint_times},
        }

    def cpu_baseline(self, budget_s):
        from oracle import oracle as orc

        sc = self.sc
        nth = orc.nthreads_default()
        rows = 8
        t0 = time.perf_counter()
        orc.field_los(sc["src"][:rows], sc["w"][:rows], sc["snk"], sc["snk_n"], 0.0, sc["beta"],
                      nthreads=nth)
        dt = time.perf_counter() - t0
        rows = int(max(8, min(self.S, rows * budget_s / max(dt, 1e-6))))
        t0 = time.perf_counter()
        orc.field_los(sc["src"][:rows], sc["w"][:rows], sc["snk"], sc["snk_n"], 0.0, sc["beta"],
                      nthreads=nth)
        dt = time.perf_counter() - t0
        return {"value": rows * self.K / dt / 1e9, "unit": "Gpairs/s", "cores": nth,
                "kind": "port", "sample": f"float64 field of the first {rows} of {self.S} "
                                          f"elements x {self.K} sinks, {dt:.1f} s"}

    def parity(self, n_sinks=4096):
        """The forward field (through the plan, as timed) on n_sinks random
        sinks against the float64 oracle over every element, given the frozen
        mask."""
        from oracle import oracle as orc
        from lyceanem.differentiable import far_field

        sc = self.sc
        with self.torch.no_grad():
            E = far_field(self.w, self.pos, self.snk, self.snk_n, hit_bits=self.bits,
                          beta=sc["beta"], plan=self.plan).cpu().numpy()
        ks = np.sort(np.random.default_rng(777).choice(self.K, n_sinks, replace=False))
        hit = self.nat.unpack_bits(self.bits, self.K)[:, ks]
        ref = orc.field_los(sc["src"], sc["w"], sc["snk"][ks], sc["snk_n"][ks], 0.0, sc["beta"],
                            hit=hit)
        return {"forward_field_vs_fp64": {"sinks": n_sinks, **_field_errors(E[ks], ref)}}

    def config(self):
        return {"workload": "c4: 64x64 conformal array on a 0.5 m cylinder, 10 GHz, 181x361 "
                            "far field, cylinder+fin mask (20k triangles) frozen",
                "sources": self.S, "sinks": self.K, "pairs": self.rays()}


# --------------------------------------------------------------------------
# config 5: batched-excitation contraction on MFMA
# --------------------------------------------------------------------------

class C5:
    name = "c5"

    def __init__(self, torch):
        from lyceanem import _native, scenes

        self.torch, self.nat = torch, _native
        self.lib = _native.load_library()
        sc = scenes.batched_c5()
        self.sc = sc
        dev = _native.device()
        self.src = torch.from_numpy(sc["src"]).to(dev)
        self.snk = torch.from_numpy(sc["snk"]).to(dev)
        self.snk_n = torch.from_numpy(sc["snk_n"]).to(dev)
        self.kf = torch.from_numpy(sc["wavenumbers"]).to(dev)
        self.W = torch.from_numpy(sc["W"]).to(dev)
        self.S, self.K, self.F = len(sc["src"]), len(sc["snk"]), len(sc["wavenumbers"])
        self.out = torch.empty((self.F, 3, self.K, 64), dtype=torch.complex64, device=dev)

    def rays(self):
        return self.S * self.K * self.F * 64

    def step(self, ev=None):
        nat, p = self.nat, self.nat.ptr
        pol = self.sc["pol"]
        if ev is not None:
            ev[0].record()
        nat.check(self.lib.lem_gw_mfma(p(self.src), self.S, float(pol[0]), float(pol[1]),
                                       float(pol[2]), p(self.snk), p(self.snk_n), self.K,
                                       p(self.kf), self.F, 0.0, p(self.W), 64, p(self.out),
                                       nat.stream_ptr()), "gw")
        if ev is not None:
            ev[1].record()
            ev[2].record()

    def record(self, times_ms):
        t = float(np.mean([a for a, _ in times_ms]))
        flops = 8.0 * self.S * self.K * 64 * self.F * 3
        torch = self.torch
        # the plain per-excitation kernel on one (f, b): scaled to all F x 64
        w = (self.W[0, :, 0][:, None] * torch.tensor([0, 0, 1], dtype=torch.complex64,
                                                     device=self.W.device)).contiguous()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        self.nat.field_los(self.src, w, self.snk, self.snk_n, beta=float(self.sc["wavenumbers"][0]))
        e0.record()
        for _ in range(4):
            self.nat.field_los(self.src, w, self.snk, self.snk_n,
                               beta=float(self.sc["wavenumbers"][0]))
        e1.record()
        torch.cuda.synchronize()
        per = e0.elapsed_time(e1) / 4
        return {
            "kernel": "k_gw_mfma",
            "kernel_ms": t,
            "roofline": {"bound": "mfma", "achieved": flops / (t * 1e-3) / 1e12,
                         "peak": PEAK_FP32 / 1e12, "unit": "TFLOP/s",
                         "frac": flops / (t * 1e-3) / PEAK_FP32, "traffic": None},
            "extra": {"gemm_flops": flops, "g_evals": self.S * self.K * self.F,
                      "per_excitation_kernel_ms": per,
                      "per_excitation_total_ms": per * self.F * 64,
                      "speedup_vs_per_excitation": per * self.F * 64 / t},
        }

    def cpu_baseline(self, budget_s):
        from oracle import oracle as orc

        sc = self.sc
        nth = orc.nthreads_default()
        rows = 16
        t0 = time.perf_counter()
        G = orc.field_los(sc["src"][:rows], sc["pol"].astype(np.complex64), sc["snk"], sc["snk_n"],
                          0.0, sc["wavenumbers"][0], elements=True, nthreads=nth)
        np.einsum("skc,sb->ckb", G, sc["W"][0, :rows].astype(np.complex128))
        dt = time.perf_counter() - t0
        return {"value": rows * self.K * 64 / dt / 1e9, "unit": "G (pair x set)/s", "cores": nth,
                "kind": "port", "sample": f"float64 G of {rows} sources x {self.K} sinks at one "
                                          f"frequency, contracted with 64 sets, {dt:.1f} s"}

    def parity(self, n_sinks=64):
        """The last step's E at the lowest and highest frequency on n_sinks random
        sinks against float64 G . W (the oracle's per-element field times W)."""
        from oracle import oracle as orc

        sc = self.sc
        ks = np.sort(np.random.default_rng(99).choice(self.K, n_sinks, replace=False))
        out = {}
        for f in (0, self.F - 1):
            G = orc.field_los(sc["src"], sc["pol"].astype(np.complex64), sc["snk"][ks],
                              sc["snk_n"][ks], 0.0, float(sc["wavenumbers"][f]), elements=True)
            ref = np.einsum("skc,sb->ckb", G, sc["W"][f].astype(np.complex128))
            got = self.out[f][:, self.torch.from_numpy(ks).to(self.out.device)].cpu().numpy()
            out[f"freq{f}_vs_fp64"] = {"sinks": n_sinks, **_field_errors(got, ref)}
        return out

    def config(self):
        return {"workload": "c5: 128x128 sources x 128x128 far-field sinks x 32 frequencies "
                            "(2-4 GHz) x 64 steering sets, G generated on the fly, FP32 MFMA",
                "sources": self.S, "sinks": self.K, "frequencies": self.F, "sets": 64}


# --------------------------------------------------------------------------
# time domain: every line-of-sight path of a 64x64 array to 256 sinks
# --------------------------------------------------------------------------

TD_PATH_FLOPS = 75  # float64 work of one line-of-sight path in k_time_paths
# placement bytes per path: arrival read by the wake and bin passes (2 x 8 B),
# the start sample written and read (2 x 8 B), the amplitude read (24 B)
TD_PLACE_BYTES = 56


class TD:
    name = "td"

    def __init__(self, torch):
        from lyceanem import _native, scenes

        self.torch, self.nat = torch, _native
        self.lib = _native.load_library()
        sc = scenes.time_domain_td()
        self.sc = sc
        dev = _native.device()
        self.pos = torch.from_numpy(sc["pos"]).to(dev)
        self.nrm = torch.from_numpy(sc["nrm"]).to(dev)
        self.w = torch.from_numpy(sc["w"]).to(dev)
        self.index = torch.from_numpy(sc["index"]).to(dev)
        self.exc = torch.from_numpy(sc["exc"]).to(dev)
        self.V, self.N = sc["index"].shape[0], int(sc["n_samples"])
        f64 = torch.float64
        self.amp = torch.empty((self.V, 3), dtype=f64, device=dev)
        self.arr = torch.empty(self.V, dtype=f64, device=dev)
        self.wake = torch.empty(1, dtype=f64, device=dev)
        self.bins = torch.empty(self.V, dtype=torch.int64, device=dev)
        self.err = torch.zeros(1, dtype=torch.int32, device=dev)
        self.out = torch.empty((self.N, 3), dtype=f64, device=dev)
        self.wsb = int(self.lib.lem_time_trace_workspace(self.V, self.N))
        self.ws = torch.empty(self.wsb, dtype=torch.uint8, device=dev)

    def rays(self):
        return self.V

    def step(self, ev=None):
        nat, lib, p = self.nat, self.lib, self.nat.ptr
        st = nat.stream_ptr()
        if ev is not None:
            ev[0].record()
        nat.check(lib.lem_time_paths(p(self.pos), p(self.nrm), p(self.w), None, self.pos.shape[0],
                                     p(self.index), self.V, 2, 0.0, float(self.sc["beta"]),
                                     p(self.amp), p(self.arr), p(self.err), st), "time_paths")
        if ev is not None:
            ev[1].record()
        nat.check(lib.lem_time_wake(p(self.arr), self.V, p(self.wake), st), "time_wake")
        nat.check(lib.lem_time_bins(p(self.arr), self.V, p(self.wake), float(self.sc["fs"]),
                                    p(self.bins), st), "time_bins")
        nat.check(lib.lem_time_trace(p(self.amp), p(self.bins), self.V, p(self.exc),
                                     self.exc.shape[0], self.N, p(self.out), p(self.ws), self.wsb,
                                     st), "time_trace")
        if ev is not None:
            ev[2].record()

    def record(self, times_ms):
        tp = float(np.mean([a for a, _ in times_ms]))
        rest = float(np.mean([b for _, b in times_ms]))
        flops = self.V * TD_PATH_FLOPS
        place = self.V * TD_PLACE_BYTES + self.N * 24 + self.exc.shape[0] * 8
        extra = {"paths": self.V, "samples": self.N, "excitation_samples":
                 int(self.exc.shape[0]), "time_paths_ms": tp, "placement_ms": rest,
                 "time_paths_fp64_frac": flops / (tp * 1e-3) / PEAK_FP64,
                 "placement_hbm_frac": place / (rest * 1e-3) / PEAK_HBM}
        if tp >= rest:
            return {"kernel": "k_time_paths", "kernel_ms": tp, "extra": extra,
                    "roofline": {"bound": "valu-fp64", "achieved": flops / (tp * 1e-3) / 1e12,
                                 "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s",
                                 "frac": flops / (tp * 1e-3) / PEAK_FP64, "traffic": None}}
        return {"kernel": "placement: k_time_wake + k_time_bins + k_trace_keys + radix sort + "
                          "k_trace_starts/hist/conv", "kernel_ms": rest, "extra": extra,
                "roofline": {"bound": "hbm", "achieved": place / (rest * 1e-3) / 1e9,
                             "peak": PEAK_HBM / 1e9, "unit": "GB/s",
                             "frac": place / (rest * 1e-3) / PEAK_HBM, "traffic": None,
                             "note": f"algorithmic bytes: {TD_PLACE_BYTES} B per path + the "
                                     f"trace and excitation"}}

    def cpu_baseline(self, budget_s):
        from oracle import oracle as orc

        sc = self.sc
        n = 4096
        t0 = time.perf_counter()
        orc.time_paths(sc["pos"], sc["nrm"], sc["w"], sc["index"][:n], 0.0, sc["beta"])
        dt = time.perf_counter() - t0
        n = int(max(4096, min(self.V, n * budget_s / max(dt, 1e-6))))
        t0 = time.perf_counter()
        amp, arr = orc.time_paths(sc["pos"], sc["nrm"], sc["w"], sc["index"][:n], 0.0, sc["beta"])
        orc.time_place(amp, arr, np.zeros(n, np.int64), 1, sc["exc"], sc["fs"], self.N)
        dt = time.perf_counter() - t0
        return {"value": n / dt / 1e9, "unit": "Gpaths/s", "cores": 1, "kind": "port",
                "sample": f"first {n} of {self.V} paths: per-path amplitude/arrival and "
                          f"placement of the {sc['exc'].shape[0]}-sample pulse into a "
                          f"{self.N}-sample trace (the reference's per-path scatter), "
                          f"{dt:.1f} s"}

    def parity(self, n_paths=65536):
        """Per-path amplitude and arrival of the last step on random paths, and
        the trace of the first n_paths paths placed on their own, against the
        oracle."""
        from oracle import oracle as orc

        sc, nat = self.sc, self.nat
        rng = np.random.default_rng(4242)
        pick = np.sort(rng.choice(self.V, 4096, replace=False))
        amp_o, arr_o = orc.time_paths(sc["pos"], sc["nrm"], sc["w"], sc["index"][pick], 0.0,
                                      sc["beta"])
        amp = self.amp.cpu().numpy()[pick]
        arr = self.arr.cpu().numpy()[pick]
        out = {"paths_vs_oracle": {
            "paths": 4096,
            "amp_max_rel": float(np.abs(amp - amp_o).max() / np.abs(amp_o).max()),
            "arrival_max_rel": float(np.abs(arr - arr_o).max() / np.abs(arr_o).max())}}
        idx = sc["index"][:n_paths]
        a, r = nat.time_paths(sc["pos"], sc["nrm"], sc["w"], idx, 0.0, sc["beta"])
        wake = nat.time_wake(r)
        bins = nat.time_bins(r, wake, sc["fs"])
        tr = nat.time_trace(a, bins, sc["exc"], self.N).cpu().numpy()
        amp_o, arr_o = orc.time_paths(sc["pos"], sc["nrm"], sc["w"], idx, 0.0, sc["beta"])
        ref, _, bins_o = orc.time_place(amp_o, arr_o, np.zeros(len(idx), np.int64), 1, sc["exc"],
                                        sc["fs"], self.N)
        out["trace_vs_oracle"] = {
            "paths": int(len(idx)), **_mismatch("start_sample", bins.cpu().numpy(), bins_o),
            "max_abs_err_over_max": float(np.abs(tr - ref[0]).max() / np.abs(ref[0]).max())}
        return out

    def config(self):
        return {"workload": "td: 64x64 array at 3 GHz -> 4096 sinks 50-500 m, 16,777,216 "
                            "line-of-sight paths, 256-sample pulse at 40 GS/s into a "
                            "65,536-sample trace", "paths": self.V, "samples": self.N}


# --------------------------------------------------------------------------
# beamforming: MaximumDirectivityMap's contraction for every command angle
# --------------------------------------------------------------------------

class BF:
    name = "bf"

    def __init__(self, torch):
        from lyceanem import _native, scenes
        from lyceanem.electromagnetics import beamforming as B

        self.torch, self.nat = torch, _native
        self.lib = _native.load_library()
        sc = scenes.beamforming_bf()
        self.sc = sc
        Et, Ep = sc["Et"], sc["Ep"]
        ne, n_el, n_az = Et.shape
        ai, ei = np.arange(n_az), np.arange(n_el)
        W, pstar = B.command_weights(Et, Ep, sc["points"], sc["wavelength"], sc["az"], sc["el"],
                                     ai, ei)
        q = B._power_weights(sc["az"], sc["el"])
        nb, npts = W.shape[0] * W.shape[1] * 3, n_el * n_az
        NB, NE, NP = -(-nb // 32) * 32, -(-ne // 32) * 32, -(-npts // 64) * 64
        dev = _native.device()
        self.W = torch.zeros((NB, NE), dtype=torch.complex64, device=dev)
        self.W[:nb, :ne] = torch.from_numpy(W.reshape(nb, ne)).to(dev)
        self.Et = torch.zeros((NE, NP), dtype=torch.complex64, device=dev)
        self.Et[:ne, :npts] = torch.from_numpy(Et.reshape(ne, npts)).to(dev)
        self.Ep = torch.zeros((NE, NP), dtype=torch.complex64, device=dev)
        self.Ep[:ne, :npts] = torch.from_numpy(Ep.reshape(ne, npts)).to(dev)
        self.q = torch.zeros(NP, dtype=torch.float32, device=dev)
        self.q[:npts] = torch.from_numpy(q.astype(np.float32)).to(dev)
        self.ps = torch.full((NB,), -1, dtype=torch.int32, device=dev)
        self.ps[:nb] = torch.from_numpy(pstar.reshape(-1)).to(dev)
        self.power = torch.empty(NB, dtype=torch.float64, device=dev)
        self.ustar = torch.zeros((NB, 2), dtype=torch.float32, device=dev)
        self.wsb = int(self.lib.lem_beam_workspace(NB, NP))
        self.ws = torch.empty(self.wsb, dtype=torch.uint8, device=dev)
        self.NB, self.NE, self.NP = NB, NE, NP
        self.ncmd, self.ne, self.npts = n_el * n_az, ne, npts

    def rays(self):
        return self.ncmd

    def step(self, ev=None):
        nat, p = self.nat, self.nat.ptr
        if ev is not None:
            ev[0].record()
        nat.check(self.lib.lem_beam_power(p(self.W), self.NB, p(self.Et), p(self.Ep), self.NE,
                                          self.NP, p(self.q), p(self.ps), p(self.power),
                                          p(self.ustar), p(self.ws), self.wsb, nat.stream_ptr()),
                  "beam_power")
        if ev is not None:
            ev[1].record()
            ev[2].record()

    def record(self, times_ms):
        t = float(np.mean([a for a, _ in times_ms]))
        flops = 8.0 * self.NB * self.NE * self.NP * 2
        return {
            "kernel": "k_beam_power",
            "kernel_ms": t,
            "roofline": {"bound": "mfma", "achieved": flops / (t * 1e-3) / 1e12,
                         "peak": PEAK_FP32 / 1e12, "unit": "TFLOP/s",
                         "frac": flops / (t * 1e-3) / PEAK_FP32, "traffic": None},
            "extra": {"mfma_flops": flops, "weight_rows": self.ncmd * 3, "elements": self.ne,
                      "points": self.npts},
        }

    def cpu_baseline(self, budget_s):
        # the reference's numpy beamforming of one command: three weight sets x
        # np.sum(w.reshape(n, 1, 1) * E, axis=0) for both polarisations
        sc = self.sc
        Et, Ep = sc["Et"].astype(np.complex128), sc["Ep"].astype(np.complex128)
        rng = np.random.default_rng(0)
        n, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < budget_s / 2 or n < 2:
            w = np.exp(1j * rng.uniform(-np.pi, np.pi, (3, Et.shape[0])))
            for k in range(3):
                np.sum(w[k].reshape(-1, 1, 1) * Et, axis=0)
                np.sum(w[k].reshape(-1, 1, 1) * Ep, axis=0)
            n += 1
        dt = time.perf_counter() - t0
        return {"value": n / dt, "unit": "commands/s", "cores": 1, "kind": "port",
                "sample": f"{n} command angles of {self.ncmd}: the reference's numpy "
                          f"beamforming of {Et.shape[0]} element patterns on a "
                          f"{Et.shape[1]}x{Et.shape[2]} grid with 3 weight sets, {dt:.1f} s "
                          f"(the directivity integrals not included)"}

    def parity(self, n_rows=16):
        """Beamformed power of n_rows random weight rows against a float64 numpy
        contraction of the same rows."""
        rng = np.random.default_rng(31)
        rows = np.sort(rng.choice(self.ncmd * 3, n_rows, replace=False))
        W = self.W[:, : self.ne].cpu().numpy()[rows].astype(np.complex128)
        Et = self.Et[: self.ne, : self.npts].cpu().numpy().astype(np.complex128)
        Ep = self.Ep[: self.ne, : self.npts].cpu().numpy().astype(np.complex128)
        q = self.q[: self.npts].cpu().numpy().astype(np.float64)
        ref = (np.abs(W @ Et) ** 2 + np.abs(W @ Ep) ** 2) @ q
        got = self.power.cpu().numpy()[rows]
        return {"power_vs_fp64": {"rows": n_rows,
                                  "max_rel": float(np.abs(got - ref).max() / np.abs(ref).max())}}

    def config(self):
        return {"workload": "bf: 32x32 array at 3 GHz, 91x181 pattern grid, every grid point "
                            "a command angle x 3 weight schemes (MaximumDirectivityMap)",
                "commands": self.ncmd, "elements": self.ne, "points": self.npts}


CONFIGS = {"c1": C1, "c2": C2, "c3": C3, "c4": C4, "c5": C5, "td": TD, "bf": BF}
DEFAULT_CONFIG = "c2"


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default=DEFAULT_CONFIG, choices=sorted(CONFIGS))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-parity", action="store_true",
                    help="skip the parity checks that run after the timed region")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--c2-sinks", type=int, default=16384, help="c2 sink count (probe runs)")
    ap.add_argument("--cells-per-tri", type=float, default=2.0, help="grid cell budget")
    ap.add_argument("--c3-walk-from-source", action="store_true",
                    help="c3: walk the leg-1 rays from the array end (default: from the "
                         "scatterer end)")
    ap.add_argument("--input-order", action="store_true",
                    help="c2: wavefronts take sources in input order (64-source rows) instead "
                         "of the default Morton order")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="only run a plain device copy of this many MiB, 3 times (calibrates "
                         "the FETCH_SIZE / WRITE_SIZE counters)")
    a = ap.parse_args(argv)
    if a.copy_mib > 0:
        import torch

        n = a.copy_mib * (1 << 20) // 4
        x = torch.ones(n, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4, "launches": 3}))
        return 0
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(the sharding plan is design-only).", file=sys.stderr)
        return 2
    import torch

    if not torch.cuda.is_available():
        print("bench.py needs a HIP device", file=sys.stderr)
        return 2
    torch.cuda.set_device(0)
    if a.config == "c2":
        cfg = C2(torch, a.c2_sinks, a.cells_per_tri, a.input_order)
    elif a.config == "c3":
        cfg = C3(torch, from_sink=not a.c3_walk_from_source, cells_per_tri=a.cells_per_tri)
    else:
        cfg = CONFIGS[a.config](torch)
    print(f"[bench] config {a.config}: setup done", file=sys.stderr, flush=True)
    for _ in range(a.warmup):
        cfg.step()
    torch.cuda.synchronize()
    print(f"[bench] warmup {a.warmup} done", file=sys.stderr, flush=True)
    evs = [_events(torch, 3) for _ in range(a.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(a.steps):
        cfg.step(evs[i])
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms = (t1 - t0) / a.steps * 1e3
    times = [(e[0].elapsed_time(e[1]), e[1].elapsed_time(e[2])) for e in evs]
    rec = cfg.record(times)
    print(f"[bench] timed {a.steps} steps: {ms:.2f} ms/step", file=sys.stderr, flush=True)
    cpu = None
    if not a.no_cpu:
        cpu = cfg.cpu_baseline(a.cpu_seconds)
        print(f"[bench] cpu baseline: {cpu['value']:.3f} {cpu['unit']}", file=sys.stderr,
              flush=True)
        if getattr(cfg, "cpu_tiles", None) is not None:
            rec["extra"]["cpu_baseline_tiles"] = cfg.cpu_tiles
    parity = None
    if not a.no_parity:
        t0 = time.perf_counter()
        parity = cfg.parity()
        parity["checked_in_s"] = time.perf_counter() - t0
        print(f"[bench] parity checked in {parity['checked_in_s']:.1f} s", file=sys.stderr,
              flush=True)
    rec["extra"]["parity"] = parity
    unit = {"c3": "Gpairs/s", "c4": "Gpairs/s", "c5": "G (pair x set)/s",
            "td": "Gpaths/s", "bf": "commands/s"}.get(a.config, "Mrays/s")
    scale = {"Mrays/s": 1e6, "commands/s": 1.0}.get(unit, 1e9)
    line = {
        "metric": "Mrays/s (1M-tri DDA); Gpairs/s scattering + % roofline; adjoint/forward time",
        "value": cfg.rays() / (ms * 1e-3) / scale,
        "unit": unit,
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (f64 phase)",
        "data": "synthetic",
        "config": cfg.config(),
        "roofline": rec["roofline"],
        "cpu_baseline": cpu,
        "dominant_kernel": {"name": rec["kernel"], "avg_ms": rec["kernel_ms"]},
        "extra": rec["extra"],
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
