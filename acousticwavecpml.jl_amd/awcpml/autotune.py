"""Launch-shape autotuner for the stencil kernel.

For a plan, time each candidate (block_y, march_len, prefetch, shot_group) with HIP
events around a few stencil launches (``Plan.time_stencil``), keep the
fastest, apply it to the plan and remember it in a JSON cache keyed by the
problem's shape, halo, precision and shot count.  The wavefield state is
reset afterwards.

    from awcpml.autotune import autotune
    best = autotune(plan)            # {"block_y": 4, "march_len": 148, ...}

Multi-shot plans also try shot groups (the waves of a block march the same
rows of several shots, sharing the model rows they read).  Two-step
launches (``Plan.set_time_block``: two time steps per launch, bit-identical
results) are timed per time step too, each shape with aligned 64-node x
tiles and with 62-node tiles (``Plan.set_two_step_align``), and kept when
they are faster; the choice is recorded as ``time_block["align"]``.
"""
import json
import os

DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "awcpml", "tuning.json")


def candidates(plan):
    """Launch shapes worth timing for this plan's grid."""
    n_march = plan.n[-1] - 2 if plan.ndim == 3 else plan.n[1] - 2
    lens = sorted({min(n_march, v) for v in (8, 16, 32, 64, (n_march + 1) // 2, n_march)})
    shapes = [(1, 1), (2, 1), (4, 1)]
    if plan.nshots > 1:
        shapes = [(by, g) for by in (1, 2, 4, 8) for g in (1, 2, 4, 8)
                  if by % g == 0 and g <= plan.nshots]
    out = []
    for by, g in shapes:
        for ml in lens:
            for pf in (1, 2, 3, 4, 5):
                if (1 < pf <= 3 and ml < 2 * pf) or (pf > 3 and ml < 8):
                    continue
                out.append((by, ml, pf, g))
    return out


def two_step_candidates(plan):
    """(block_rows, march_len, prefetch) of two-step launches worth timing."""
    if plan.ndim == 2:
        n_march = plan.n[1] - 2
        return [(w, min(n_march, ml), pf) for w in (4, 8) for ml in (16, 32, 64, 128)
                for pf in (1, 2, 3)]
    n_march = plan.n[2] - 2
    out = []
    for w in ((16, 8) if plan.dtype != 0 else (8, 4)):
        # FP64 on 298^3 (8 waves): chunk 40 measured 3 % faster than 48
        lens = (24, 48, 96, 144) if w == 16 else (24, 32, 40, 48, 96)
        pfs = (1, 2, 3) if w == 16 else (1, 2)
        for ml in sorted({min(n_march, v) for v in lens}):
            out += [(w, ml, pf) for pf in pfs]
    return out


def two_step_batch_candidates(plans):
    """(block_rows, march_len, prefetch) of two-step batch launches worth
    timing; march_len is applied to every plan (0: each plan's default,
    sized to its grid)."""
    if plans[0].ndim == 2:
        return [(w, ml, pf) for w in (4, 8) for ml in (0, 32, 64) for pf in (1, 2)]
    if plans[0].dtype == 0:
        return [(8, ml, pf) for ml in (0, 24, 48) for pf in (1, 2)] + [(4, 24, 2)]
    return ([(16, ml, pf) for ml in (0, 24, 48, 96, 144) for pf in (1, 2)]
            + [(8, ml, 2) for ml in (24, 48)])


def _apply(plan, best):
    plan.set_tiling(best["block_y"], best["march_len"], best["prefetch"])
    plan.set_shot_group(best.get("shot_group", 1))
    tb = best.get("time_block")
    if tb:
        if "align" in tb:
            plan.set_two_step_align(tb["align"])
        plan.set_time_block(2, tb["block_rows"], tb["march_len"], tb["prefetch"])
    else:
        plan.set_time_block(1)


def _key(plan):
    return "%dD:%s:h%d:%s:s%d" % (plan.ndim, "x".join(map(str, plan.n)), plan.halo,
                                   "f64" if plan.dtype == 0 else "f32", plan.nshots)


def _load(path):
    try:
        with open(path) as fh:
            return json.load(fh)
    except (OSError, ValueError):
        return {}


def autotune(plan, cands=None, launches=8, cache=DEFAULT_CACHE, use_cache=True):
    """Pick the fastest launch shape for ``plan`` and apply it."""
    key = _key(plan)
    db = _load(cache) if cache else {}
    if use_cache and key in db:
        best = db[key]
        _apply(plan, best)
        return dict(best, cached=True)
    plan.set_time_block(1)
    results = []
    for cand in (cands or candidates(plan)):
        by, ml, pf = cand[:3]
        g = cand[3] if len(cand) > 3 else 1
        plan.set_tiling(by, ml, pf)
        plan.set_shot_group(g)
        plan.reset()
        plan.time_stencil(2)  # warm-up
        ms = plan.time_stencil(launches) / launches
        results.append((ms, by, ml, pf, g))
    plan.reset()
    ms, by, ml, pf, g = min(results)
    plan.set_tiling(by, ml, pf)
    plan.set_shot_group(g)
    best = {"block_y": by, "march_len": ml, "prefetch": pf, "shot_group": g,
            "ms_per_launch": ms, "ms_per_step": ms, "time_block": None, "tried": len(results)}
    # two time steps per launch, compared per time step
    from . import _lib
    two = []
    # each shape with aligned 64-node tiles and with 62-node tiles
    aligns = (True, False)
    align0 = plan.two_step_align()
    for w, ml2, pf2 in (two_step_candidates(plan) if cands is None else []):
        try:
            for al in aligns:
                if al is not None:
                    plan.set_two_step_align(al)
                plan.set_time_block(2, w, ml2, pf2)
                plan.reset()
                plan.time_stencil2(1)  # warm-up
                n2 = max(1, launches // 2)
                two.append((plan.time_stencil2(n2) / n2 / 2.0, w, ml2, pf2, al))
        except _lib.AwcpmlError:
            break  # two-step launches do not apply to this plan
    if align0 is not None:
        plan.set_two_step_align(align0)
    if two and min(two)[0] < best["ms_per_step"]:
        ms2, w, ml2, pf2, al = min(two, key=lambda t: t[0])
        best["time_block"] = {"block_rows": w, "march_len": ml2, "prefetch": pf2}
        if al is not None:
            best["time_block"]["align"] = al
        best["ms_per_step"] = ms2
    best["tried"] += len(two)
    _apply(plan, best)
    plan.reset()
    if cache:
        db[key] = best
        try:
            os.makedirs(os.path.dirname(cache), exist_ok=True)
            with open(cache, "w") as fh:
                json.dump(db, fh, indent=1)
        except OSError:
            pass
    return dict(best, cached=False)


def autotune_batch(plans, cands=None, launches=3, two_step=True):
    """Pick the launch shape (block_y, march_len, prefetch) shared by a batch
    of plans: each candidate is applied to every plan, a Batch is built and
    its stencil launch timed.  Two-step batch launches (two time steps of
    every plan per launch) are then timed per time step and kept when they
    are faster.  Returns the best shape and leaves it applied to the plans:
    a Batch built from them afterwards uses it."""
    from . import _lib
    from .solver import Batch
    cands = cands or [(by, ml, pf) for by in (2, 4) for ml in (32, 64, 128) for pf in (2, 3, 5)]
    for P in plans:
        P.set_time_block(1)
        P.reset()  # the wavefields are overwritten by the timing anyway
    results = []
    for by, ml, pf in cands:
        for P in plans:
            P.set_tiling(by, ml, pf)
            P.set_shot_group(1)
        B = Batch(plans)
        B.reset()
        B.time_stencil(2)
        ms = B.time_stencil(launches) / launches
        B.close()
        results.append((ms, by, ml, pf))
    ms, by, ml, pf = min(results)
    for P in plans:
        P.set_tiling(by, ml, pf)
    best = {"block_y": by, "march_len": ml, "prefetch": pf, "ms_per_launch": ms,
            "ms_per_step": ms, "time_block": None, "tried": len(results)}
    two = []
    for w, ml2, pf2 in (two_step_batch_candidates(plans) if two_step else []):
        try:
            for P in plans:
                P.set_time_block(2, w, ml2, pf2)
            B = Batch(plans)
            try:
                if B.time_block()[0] != 2:
                    break  # two-step launches do not apply to some plan
                B.reset()
                B.time_stencil2(1)
                n2 = max(1, launches // 2)
                two.append((B.time_stencil2(n2) / n2 / 2.0, w, ml2, pf2))
            finally:
                B.close()
        except _lib.AwcpmlError:
            break
    best["tried"] += len(two)
    if two and min(two)[0] < best["ms_per_step"]:
        ms2, w, ml2, pf2 = min(two)
        best["time_block"] = {"block_rows": w, "march_len": ml2, "prefetch": pf2}
        best["ms_per_step"] = ms2
        for P in plans:
            P.set_time_block(2, w, ml2, pf2)
    else:
        for P in plans:
            P.set_time_block(1)
    for P in plans:
        P.reset()
    return best
