"""Launch-shape autotuner.

Searches the device's tuning space by timing renders of the caller's own
job, keeps the fastest shape on the device and returns the table. The
knobs change speed only: every shape gives the same hits and images (the
kd-tree traversal restarts from the root when its short stack overflows, and
ray order only permutes the queues).

The search is staged: the trace kernel's shape (kd short-stack depth, block
size, persistent grid) over its full product first, then the ray-sort
granularity (camera rays sample by sample over the image, by 8x8 tile, or by
pixel), then whether the trace queue is cut into per-XCD segments, then
whether kd leaves are read as per-reference copies or through the reference
list, then how many pixel groups a pass runs as on streams of their own,
each stage keeping the best of the one before; the wavefront size (paths in
flight per pass) last, on the full job.
"""
import itertools

from ._native import TUNING_AUTO

KERNEL_SPACE = {
    "kd_stack": (8, 12, 16, 24),
    "trace_block": (128, 256),
    "trace_blocks_per_cu": (TUNING_AUTO, 2, 8),
}
STAGES = [KERNEL_SPACE, {"ray_order": (1, 2, 3)}, {"trace_segments": (1, 8)}, {"kd_leaf": (1, 2)},
          {"frame_lanes": (1, 2)}]
DEFAULT_SPACE = KERNEL_SPACE  # the first stage alone (older callers)


def _time_job(dev, job, out_ptr, repeats):
    best = float("inf")
    for _ in range(repeats):
        st = dev.render_device(job, out_ptr)
        best = min(best, st.ms_total)
    return best


def _shape(t):
    """A tuning dict as xt_dev_set_tuning takes it (0: keep, so the resident
    grid is asked for by name)."""
    out = {k: v for k, v in t.items() if k != "wave_paths"}
    if not out.get("trace_blocks_per_cu"):
        out["trace_blocks_per_cu"] = TUNING_AUTO
    return out


def autotune(dev, job, out_ptr, space=None, samples=0, repeats=2, wave_paths=(), log=None, stages=None):
    """Times `job` (cut to `samples` samples per pixel when `samples` > 0;
    whole by default, since a cut job runs smaller query rounds than the
    whole one and can favour another shape) under the shapes of each stage
    (`stages`, default STAGES; `space` alone: one stage) on `dev`, then
    leaves `dev` set to the fastest. `wave_paths` adds wavefront sizes to try
    once the best shape is known. Returns (best tuning dict, [(tuning dict,
    ms), ...])."""
    if stages is None:
        stages = [dict(space)] if space is not None else STAGES
    probe = type(job).from_buffer_copy(job)
    probe.samples = min(job.samples, samples) if samples > 0 else job.samples
    probe.samples_per_pass = 0
    table = []
    start = dev.tuning()
    _time_job(dev, probe, out_ptr, 1)  # first launch of each kernel loads its code object
    best_shape, best_ms = _shape(start), float("inf")
    for stage in stages:
        keys = list(stage)
        stage_best = None
        for combo in itertools.product(*(stage[k] for k in keys)):
            knobs = dict(best_shape, **dict(zip(keys, combo)))
            dev.set_tuning(**knobs)
            _time_job(dev, probe, out_ptr, 1)
            ms = _time_job(dev, probe, out_ptr, repeats)
            table.append((dev.tuning(), ms))
            if log:
                log(f"autotune {dict(zip(keys, combo))}: {ms:.2f} ms")
            if stage_best is None or ms < stage_best[1]:
                stage_best = (_shape(dev.tuning()), ms)
        best_shape, best_ms = stage_best
    wp_best = start["wave_paths"]
    if wave_paths:
        # the wavefront size matters only across passes: time the full job
        full = type(job).from_buffer_copy(job)
        full.samples_per_pass = 0
        wp_ms = float("inf")
        for wp in sorted(set(wave_paths) | {start["wave_paths"]}):
            dev.set_tuning(**best_shape, wave_paths=wp)
            ms = _time_job(dev, full, out_ptr, repeats)
            table.append((dev.tuning(), ms))
            if log:
                log(f"autotune wave_paths={wp}: {ms:.2f} ms")
            if ms < wp_ms:
                wp_best, wp_ms = wp, ms
    dev.set_tuning(**best_shape, wave_paths=wp_best)
    return dev.tuning(), table
