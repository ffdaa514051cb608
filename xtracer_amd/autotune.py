"""Launch-shape autotuner.

Searches the device's tuning space (kd short-stack depth, trace block size,
persistent grid size, wavefront size) by timing short renders of the caller's
own job, keeps the fastest shape on the device and returns the table. The
knobs change speed only: every shape gives the same hits and images (the
kd-tree traversal restarts from the root when its short stack overflows).
"""
import itertools

from ._native import TUNING_AUTO

DEFAULT_SPACE = {
    "kd_stack": (8, 12, 16, 24),
    "trace_block": (128, 256),
    "trace_blocks_per_cu": (TUNING_AUTO, 2, 8),
}


def _time_job(dev, job, out_ptr, repeats):
    best = float("inf")
    for _ in range(repeats):
        st = dev.render_device(job, out_ptr)
        best = min(best, st.ms_total)
    return best


def autotune(dev, job, out_ptr, space=None, samples=2, repeats=2, wave_paths=(), log=None):
    """Times `job` (cut to `samples` samples per pixel) under every shape of
    `space` on `dev`, then leaves `dev` set to the fastest. `wave_paths` adds
    wavefront sizes to try once the best kernel shape is known. Returns
    (best tuning dict, [(tuning dict, ms), ...])."""
    space = dict(DEFAULT_SPACE if space is None else space)
    probe = type(job).from_buffer_copy(job)
    probe.samples = min(job.samples, samples)
    probe.samples_per_pass = 0
    keys = list(space)
    table = []
    start = dev.tuning()
    _time_job(dev, probe, out_ptr, 1)  # first launch of each kernel loads its code object
    for combo in itertools.product(*(space[k] for k in keys)):
        knobs = dict(zip(keys, combo))
        dev.set_tuning(**knobs)
        _time_job(dev, probe, out_ptr, 1)
        ms = _time_job(dev, probe, out_ptr, repeats)
        table.append((dev.tuning(), ms))
        if log:
            log(f"autotune {knobs}: {ms:.2f} ms")
    best, _ = min(table, key=lambda e: e[1])
    shape = {k: (v if v else TUNING_AUTO) for k, v in best.items() if k != "wave_paths"}
    if wave_paths:
        # the wavefront size matters only across passes: time the full job
        full = type(job).from_buffer_copy(job)
        full.samples_per_pass = 0
        best_ms = float("inf")
        for wp in sorted(set(wave_paths) | {start["wave_paths"]}):
            dev.set_tuning(**shape, wave_paths=wp)
            ms = _time_job(dev, full, out_ptr, repeats)
            table.append((dev.tuning(), ms))
            if log:
                log(f"autotune wave_paths={wp}: {ms:.2f} ms")
            if ms < best_ms:
                best, best_ms = dev.tuning(), ms
    dev.set_tuning(**shape, wave_paths=best["wave_paths"] if wave_paths else start["wave_paths"])
    return dev.tuning(), table
