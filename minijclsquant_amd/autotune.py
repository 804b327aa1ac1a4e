"""Launch-shape autotuner with a persistent table.

The step kernel has 48 variants (rows per lane x non-temporal vector loads x
vectors per lane x non-temporal stores x non-temporal matrix loads) and a
work-group count per CU, set separately for the
plain recursion and for the doubling form; jkpm_autotune times them on the
caller's matrix and vectors.  ``tuned`` also chooses how many of the R
vectors one recursion pass carries (all of them, or halves down to 4), so a
pass's vectors and the matrix can stay in the 256 MB Infinity Cache.
Results are cached in a JSON table keyed by device, matrix shape (N, K),
value type, batch size R and form, so a shape is tuned once per machine.
``JKPM_LAUNCH="variant,chunk_mult"`` overrides the table (for both forms).

The reference has a single fixed launch (<<<ceil(N/1024), 1024>>>,
jclsquant/cuda_cython/cuda/recursions.cu:19-20).
"""
from __future__ import annotations

import json
import os
import threading
from pathlib import Path

from . import _native as nat

__all__ = ["tuned", "tuned_batch", "table_path", "pass_candidates", "apply_cached"]

_lock = threading.Lock()


def table_path() -> Path:
    env = os.environ.get("JKPM_TUNE_TABLE")
    if env:
        return Path(env)
    return Path(os.path.expanduser("~")) / ".cache" / "jclkpm" / "launch_table.json"


def _key(h: "nat.DeviceELL", R: int, doubling: bool) -> str:
    import torch

    # architecture and CU count: stable under profilers, unlike the marketing name
    p = torch.cuda.get_device_properties(torch.cuda.current_device())
    dev = f"{getattr(p, 'gcnArchName', p.name)}|cu={p.multi_processor_count}"
    form = "|doubling" if doubling else ""
    # the step kernels on the packed real values are other kernels
    rv = "|real" if h.is_real() and getattr(h, "_use_real", True) else ""
    # complex64 plain recursions with the left-vector phases turned off
    lph = ("|complex-left" if h.vtype == nat.JKPM_C64 and not doubling
           and not getattr(h, "_left_phases", True) else "")
    return f"{dev}|N={h.n}|K={h.k}|vtype={h.vtype}|R={R}{rv}{lph}{form}"


def _load():
    p = table_path()
    try:
        return json.loads(p.read_text())
    except (OSError, ValueError):
        return {}


def _store(tab):
    p = table_path()
    try:
        p.parent.mkdir(parents=True, exist_ok=True)
        tmp = p.with_suffix(".tmp")
        tmp.write_text(json.dumps(tab, indent=1, sort_keys=True))
        os.replace(tmp, p)
    except OSError:
        pass  # a read-only home just means re-tuning next time


_MIN_TIMED_US = 2000.0  # timed span per candidate the tuner aims for
_MAX_TUNE_STEPS = 256


_FINAL_US = 12000.0  # timed span per finalist pass
_FINAL_PASSES = 3


N_VARIANTS = nat.JKPM_STEP_VARIANTS


def _neighbours(v):
    """Variants one choice away from v: non-temporal matrix loads, stores or
    stream loads flipped."""
    def flip_ntm(u):  # non-temporal matrix loads: +-24
        return u - 24 if u >= 24 else u + 24

    def flip_ns(u):  # non-temporal stores: +-12 within the 24-block
        return u - 12 if u % 24 >= 12 else u + 12

    def flip_nt(u):  # non-temporal stream loads: bit 1 (12 and 24 are multiples of 4)
        return u ^ 2

    return [f(v) for f in (flip_ntm, flip_ns, flip_nt)]


def _finalists(h, v_r, v, m, us, doubling):
    """Second round over the winner of the sweep and its neighbours one
    variant bit away (non-temporal matrix loads, stores, stream loads), at
    the winner's chunk_mult: each timed over passes of about 12 ms, best of
    three.  The sweep times every candidate once over about 2 ms, where
    variants a few percent apart can swap places.  Returns (variant, us)."""
    import torch

    steps = max(16, min(1024, int(_FINAL_US / max(us, 1e-3))))
    M = steps + 1 if not doubling else 2 * steps
    need = nat.moments_ex_workspace_bytes(h, None, int(v_r.shape[1]), M, doubling)
    ws = torch.empty(need, dtype=torch.uint8, device=v_r.device)
    cands = [v] + _neighbours(v)
    best = (float("inf"), v)
    for c in cands:
        h.set_launch(c, m, doubling)
        nat.moments_ex(h, None, v_r, M, doubling, workspace=ws)
        t = float("inf")
        for _ in range(_FINAL_PASSES):
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            nat.moments_ex(h, None, v_r, M, doubling, workspace=ws)
            e1.record()
            e1.synchronize()
            t = min(t, e0.elapsed_time(e1))
        if t < best[0]:
            best = (t, c)
    h.set_launch(best[1], m, doubling)
    return best[1], best[0] * 1e3 / steps


def pass_candidates(R: int):
    """Vectors per recursion pass the tuner tries for R vectors: all of
    them, then halves down to 4 (0 stands for all at once)."""
    out, g = [0], R // 2
    while g >= 4:
        out.append(g)
        g //= 2
    return out


def _tune_shape(h, v_r, steps, doubling):
    us = h.autotune(v_r, steps, doubling=doubling)
    # Short steps: time again over about 2 ms per candidate, so that
    # shapes a few percent apart are told apart.
    if us * steps < _MIN_TIMED_US:
        steps2 = min(_MAX_TUNE_STEPS, int(_MIN_TIMED_US / max(us, 1e-3)) + 1)
        if steps2 > steps:
            us = h.autotune(v_r, steps2, doubling=doubling)
    v, m = h.get_launch(doubling)
    v, us = _finalists(h, v_r, v, m, us, doubling)
    return v, m, us


_SLOPE_US = 4000.0  # timed span of the shorter run of the pass comparison
_PASS_MARGIN = 0.02  # a pass size below R must beat all-at-once by this much


def _step_slope(h, v_r, doubling, us):
    """Per-step time (us) of the handle's current setting, as the slope
    between recursions of a and 3a steps (best of two each): the fixed cost
    of a call -- per pass a column copy, the left-phase packing and the
    first step -- cancels, as it does in a recursion of thousands of steps."""
    import torch

    R = int(v_r.shape[1])
    a = max(16, int(_SLOPE_US / max(us, 1e-3)))
    times = []
    for steps in (a, 3 * a):
        M = steps + 1 if not doubling else 2 * steps
        need = nat.moments_ex_workspace_bytes(h, None, R, M, doubling)
        ws = torch.empty(need, dtype=torch.uint8, device=v_r.device)
        nat.moments_ex(h, None, v_r, M, doubling, workspace=ws)
        t = float("inf")
        for _ in range(2):
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            nat.moments_ex(h, None, v_r, M, doubling, workspace=ws)
            e1.record()
            e1.synchronize()
            t = min(t, e0.elapsed_time(e1))
        times.append(t)
    return (times[1] - times[0]) * 1e3 / (2 * a)


def apply_cached(h: "nat.DeviceELL", v_r, doubling: bool = False) -> bool:
    """Applies the table's launch shape and vectors per pass for this shape
    and form, when the table has one, without timing anything; returns
    whether it did.  ``JKPM_LAUNCH`` applies as in tuned."""
    env = os.environ.get("JKPM_LAUNCH")
    if env:
        v, m = (int(x) for x in env.split(","))
        h.set_launch(v, m, doubling)
        return True
    key = _key(h, int(v_r.shape[1]), doubling)
    with _lock:
        ent = _load().get(key)
    if ent is None:
        return False
    h.set_launch(ent["variant"], ent["chunk_mult"], doubling)
    h.set_pass(ent.get("pass", 0), doubling)
    return True


def tuned(h: "nat.DeviceELL", v_r, steps: int = 8, retune: bool = False,
          doubling: bool = False, passes: bool = True):
    """Applies the cached (or freshly tuned) launch shape and vectors per
    pass to handle h for the batch v_r [N, R] and the chosen form; returns
    (variant, chunk_mult).  Each pass size of pass_candidates(R) is tuned
    on its own and the fastest whole step (all passes) wins, a pass size
    below R only when it is at least 2 % faster per step.  passes=False
    tunes all R vectors at once only (for the moments-with-stack form, which
    never runs in passes)."""
    env = os.environ.get("JKPM_LAUNCH")
    if env:
        v, m = (int(x) for x in env.split(","))
        h.set_launch(v, m, doubling)
        return v, m
    R = int(v_r.shape[1])
    key = _key(h, R, doubling) + ("" if passes else "|at-once")
    with _lock:
        tab = _load()
        if not retune and key in tab:
            v, m = tab[key]["variant"], tab[key]["chunk_mult"]
            h.set_launch(v, m, doubling)
            h.set_pass(tab[key].get("pass", 0), doubling)
            return v, m
        best = None
        cands = pass_candidates(R) if passes else [0]
        for g in cands:
            h.set_pass(g, doubling)
            v, m, us = _tune_shape(h, v_r, steps, doubling)
            if len(cands) > 1:
                us = _step_slope(h, v_r, doubling, us)
            # passes add copies and a schedule to the call: they must win clearly
            lim = us if g == 0 else us * (1 + _PASS_MARGIN)
            if best is None or lim < best[3]:
                best = (g, v, m, us)
        g, v, m, us = best
        h.set_pass(g, doubling)
        h.set_launch(v, m, doubling)
        tab[key] = {"variant": v, "chunk_mult": m, "pass": g, "step_us": us}
        _store(tab)
        return v, m


def tuned_batch(bk, steps: int = 16, retune: bool = False, variants=range(N_VARIANTS)):
    """Step-kernel variant of both forms for a BatchKPM whose chunk tables are
    fixed (chunks of one lattice stride): times every variant on the batch
    itself, where the tables, not chunk_mult, set the work-group count.
    Cached in the same table under the batch's own key (ΣN, K, R, the
    strides' sum).  Returns {"plain": v, "doubling": v}."""
    import torch

    env = os.environ.get("JKPM_LAUNCH")
    h, R = bk.H, bk.R
    out = {}
    for doubling in (False, True):
        cur_v, cur_m = h.get_launch(doubling)
        if env:
            out["doubling" if doubling else "plain"] = cur_v
            continue
        key = _key(h, R, doubling) + f"|batch-strided|G={bk.batch.G}|S={sum(bk.strides or [])}"
        with _lock:
            tab = _load()
            if not retune and key in tab:
                v = tab[key]["variant"]
                h.set_launch(v, cur_m, doubling)
                out["doubling" if doubling else "plain"] = v
                continue
            M = steps + 1 if not doubling else 2 * steps
            need = nat.moments_ex_workspace_bytes(None, bk.batch, R, M, doubling)
            ws = torch.empty(need, dtype=torch.uint8, device="cuda")
            mu = torch.empty((bk.batch.G, M, R), dtype=torch.complex128, device="cuda")
            best = (float("inf"), cur_v)
            for v in variants:
                h.set_launch(v, cur_m, doubling)
                nat.moments_ex(None, bk.batch, bk.vr, M, doubling, workspace=ws, out=mu)
                t = float("inf")
                for _ in range(2):
                    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                    e0.record()
                    nat.moments_ex(None, bk.batch, bk.vr, M, doubling, workspace=ws, out=mu)
                    e1.record()
                    e1.synchronize()
                    t = min(t, e0.elapsed_time(e1))
                if t < best[0]:
                    best = (t, v)
            h.set_launch(best[1], cur_m, doubling)
            tab[key] = {"variant": best[1], "chunk_mult": cur_m, "step_us": best[0] * 1e3 / steps}
            _store(tab)
            out["doubling" if doubling else "plain"] = best[1]
    return out
