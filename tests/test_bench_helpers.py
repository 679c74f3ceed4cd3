"""Host logic of bench.py's time-to-target search (no GPU): frames between
error checks are kept as snapshots, and a check that meets the target is
traced back to the first kept frame that met it."""
import numpy as np

import bench


class _Buf:
    def __init__(self, n):
        self.a = np.zeros(n)
        self.freed = False

    def download(self, dtype, count=None):
        return self.a.astype(dtype)

    def free(self):
        self.freed = True


class _Ctx:
    def buffer(self, nbytes):
        return _Buf(nbytes // 8)

    def copy(self, dst, src, nbytes):
        dst.a[:] = src.a


def _run(errors, checks, target):
    """Frames whose error is errors[f - 1]; the image holds that error times
    the frame count, so err(buffer, frame) reads it back. Returns the result
    of the search like the bench's loops do."""
    ctx, buf = _Ctx(), _Buf(4)
    err = lambda b, divisor: float(b.a[0]) / divisor
    s = bench._TargetSearch(ctx, buf, 4 * 8, err)
    ms = 0.0
    for f, e in enumerate(errors, start=1):
        ms += 10.0
        buf.a[:] = e * f
        if f not in checks:
            s.keep(f, ms, f)
            continue
        if err(buf, f) <= target:
            hit = s.first_hit(target) or (f, ms, err(buf, f))
            s.free()
            return hit, s
        s.clear()
    s.free()
    return None, s


def test_first_frame_between_checks():
    errors = [0.9, 0.7, 0.6, 0.55, 0.52, 0.49, 0.47, 0.45, 0.44]
    hit, s = _run(errors, checks={1, 2, 4, 8}, target=0.5)
    assert hit[0] == 6 and hit[1] == 60.0 and abs(hit[2] - 0.49) < 1e-12
    assert s.snaps == []


def test_target_met_at_the_check_itself():
    errors = [0.9, 0.8, 0.7, 0.45]
    hit, _ = _run(errors, checks={1, 2, 4}, target=0.5)
    assert hit[0] == 4 and hit[1] == 40.0


def test_first_kept_frame_wins_over_later_ones():
    # frames 3 and 4 are kept after the missed check at 2; frame 3 meets the
    # target and frame 4 does not: the time stops at frame 3
    errors = [0.9, 0.8, 0.49, 0.6, 0.4]
    hit, _ = _run(errors, checks={1, 2, 5}, target=0.5)
    assert hit[0] == 3 and hit[1] == 30.0


def test_snapshots_before_a_missed_check_are_not_searched():
    # frame 2 met the target but was checked only by the later check at 3,
    # which missed; the search restarts after it and stops at frame 5
    errors = [0.9, 0.45, 0.6, 0.7, 0.4]
    hit, _ = _run(errors, checks={1, 3, 5}, target=0.5)
    assert hit[0] == 5


def test_never_met():
    hit, _ = _run([0.9, 0.8, 0.7], checks={1, 2, 3}, target=0.5)
    assert hit is None
