"""Measurement-driven choice of launch shapes (parameters that enter no single shift's arithmetic).

Two kinds of parameter are tuned on the device, with the caller's own operator:

* the window cap of psTriRecy's batched mode (``PsTriRecyControl.window``): how many
  consecutive shifts one recycled Rayleigh-Ritz launch evaluates against a recycle snapshot.
  Every shift before a window's first failing one takes the sequential result, so in exact
  arithmetic the trajectory (refined points, sigma) does not depend on the cap; the cap moves
  work between wider launches and speculative shifts that are thrown away.  In floating
  point, windows of different composition round the batched products differently, which can
  flip an acceptance whose residual sits at the threshold; the tuner reports how many
  decisions and how much sigma differ between the candidates.
* the sparse triangular solve's chunk width (``RLS_TUNE_SPTRSV_CW_SMALL`` / ``_LARGE``): the
  number of right-hand-side columns one workgroup carries through the dependency chain.  The
  arithmetic of each column is the same for every width.

Each candidate is timed (wall clock around device-synchronised calls), the fastest is kept,
and the choice is cached per device and problem signature in a JSON file
(``$RLS_TUNE_CACHE``, default ``~/.cache/rls_amd/autotune.json``).  No counterpart in the
reference, whose operators are host library calls.
"""
from __future__ import annotations

import dataclasses
import json
import os
import time
from typing import Dict, Iterable, Optional

import numpy as np

from . import _capi

WINDOW_CANDIDATES = (64, 128, 256, 512)
CHUNK_CANDIDATES = (1, 2, 4, 8)


def _cache_path() -> str:
    return os.environ.get("RLS_TUNE_CACHE") or os.path.join(os.path.expanduser("~"), ".cache", "rls_amd",
                                                            "autotune.json")


def _device_name(device: int) -> str:
    try:
        import torch
        return torch.cuda.get_device_name(device)
    except Exception:
        return f"device{device}"


def cache_get(key: str) -> Optional[dict]:
    try:
        with open(_cache_path()) as fh:
            return json.load(fh).get(key)
    except (OSError, ValueError):
        return None


def cache_put(key: str, value: dict) -> None:
    path = _cache_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        data = {}
        if os.path.exists(path):
            with open(path) as fh:
                data = json.load(fh)
        data[key] = value
        with open(path, "w") as fh:
            json.dump(data, fh, indent=1, sort_keys=True)
    except (OSError, ValueError):
        pass   # an unwritable cache only costs a re-tune


def tune_window(solver, Z, control, rec, candidates: Iterable[int] = WINDOW_CANDIDATES,
                key: Optional[str] = None, use_cache: bool = True) -> dict:
    """Window cap of the batched psTriRecy solve: time ``solver.solve(Z, control, rec)`` for
    every candidate cap and return ``{"window": best, "seconds": {cap: s}, "max_rel_diff": d,
    "max_rel_diff_point": i, "sigmas_at_point": {cap: sigma_i}, "decisions_differ": m}``.
    A warm-up solve comes first, so every candidate is timed warm.  The sigma grids and refine
    decisions of the candidates are compared with the first candidate's."""
    cands = [int(c) for c in candidates]
    if key and use_cache:
        hit = cache_get(key)
        if hit and hit.get("window") in cands:
            return dict(hit, cached=True)
    ctl0 = dataclasses.replace(control, mode="batched", window=cands[0])
    solver.solve(Z, ctl0, rec)                     # warm-up: buffers sized, kernels loaded
    secs: Dict[int, float] = {}
    ref_sig = ref_dec = None
    max_rel, mism, at = 0.0, 0, -1
    sigs = {}
    for c in cands:
        ctl = dataclasses.replace(control, mode="batched", window=c)
        t0 = time.perf_counter()
        sig, st = solver.solve(Z, ctl, rec)
        secs[c] = time.perf_counter() - t0
        dec = st.lobpcg_it > 0
        sigs[c] = sig
        if ref_sig is None:
            ref_sig, ref_dec = sig, dec
        else:
            rel = np.abs(sig - ref_sig) / ref_sig
            i = int(np.nanargmax(rel))
            if rel[i] > max_rel:
                max_rel, at = float(rel[i]), i
            mism = max(mism, int((dec != ref_dec).sum()))
    best = min(secs, key=secs.get)
    out = {"window": best, "seconds": {str(c): s for c, s in secs.items()}, "max_rel_diff": max_rel,
           "max_rel_diff_point": at, "decisions_differ": mism, "points": int(len(Z)),
           "sigmas_at_point": {str(c): float(v[at]) for c, v in sigs.items()} if at >= 0 else None}
    if key:
        cache_put(key, out)
    return out


def tune_sparse_chunks(ctx: "_capi.Context", n: int, op: int = _capi.RLS_OP_KINV, ks: Iterable[int] = (6, 40),
                       candidates: Iterable[int] = CHUNK_CANDIDATES, reps: int = 5, device: int = 0,
                       key: Optional[str] = None, use_cache: bool = True, apply: bool = True) -> dict:
    """Chunk widths of the sparse triangular solves of a context that holds a sparse factor
    (an IC factor for ``RLS_OP_KINV``): time ``reps`` applications of ``op`` to random n x k
    blocks for each width, for a k <= 8 and a k > 8 class, and set the fastest widths on the
    context (``apply``).  The outputs of all widths are compared bitwise."""
    import torch
    cands = [int(c) for c in candidates]
    if key and use_cache:
        hit = cache_get(key)
        if hit and apply:
            ctx.set_tuning(_capi.RLS_TUNE_SPTRSV_CW_SMALL, hit["small"])
            ctx.set_tuning(_capi.RLS_TUNE_SPTRSV_CW_LARGE, hit["large"])
        if hit:
            return dict(hit, cached=True)
    dev = torch.device("cuda", device)
    gen = torch.Generator(device="cpu").manual_seed(7)
    before = {t: ctx.get_tuning(t) for t in (_capi.RLS_TUNE_SPTRSV_CW_SMALL, _capi.RLS_TUNE_SPTRSV_CW_LARGE)}
    result = {}
    identical = True
    for k in ks:
        X = torch.complex(torch.randn(k, n, generator=gen, dtype=torch.float64),
                          torch.randn(k, n, generator=gen, dtype=torch.float64)).to(dev).T
        Y = torch.empty((k, n), dtype=torch.complex128, device=dev).T
        tkey = _capi.RLS_TUNE_SPTRSV_CW_SMALL if k <= 8 else _capi.RLS_TUNE_SPTRSV_CW_LARGE
        secs, ref = {}, None
        for c in cands:
            ctx.set_tuning(tkey, c)
            call = lambda: ctx.check(ctx.lib.rls_dev_apply(  # noqa: E731
                ctx.h, op, 0.0, 0.0, X.data_ptr(), n, k, Y.data_ptr(), n))
            call()
            ctx.check(ctx.lib.rls_dev_sync(ctx.h))
            t0 = time.perf_counter()
            for _ in range(reps):
                call()
            ctx.check(ctx.lib.rls_dev_sync(ctx.h))
            secs[c] = (time.perf_counter() - t0) / reps
            out = Y.cpu().numpy()
            if ref is None:
                ref = out
            elif not np.array_equal(out, ref):
                identical = False
        best = min(secs, key=secs.get)
        result["small" if k <= 8 else "large"] = best
        result[f"ms_k{k}"] = {str(c): 1e3 * s for c, s in secs.items()}
        ctx.set_tuning(tkey, best)
    result["identical_outputs"] = identical
    if not apply:
        for t, v in before.items():
            ctx.set_tuning(t, v)
    if key:
        cache_put(key, result)
    return result


def signature(*parts) -> str:
    """Cache key: the device name and the problem's defining sizes."""
    return "|".join([_device_name(0)] + [str(p) for p in parts])
