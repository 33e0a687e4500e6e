"""The precondition of the traversal's walk without a containment test, checked on
the CPU: for theta <= 2/3 the opening test alone rejects every cell that holds the
target.

A target in a cell and the cell's centre of mass lie in the cell's box of side w,
so d2 <= 2 w^2 up to a few f32 roundings, while the kernels' flag
(`contained_rejected`) requires the threshold T >= 2.25 w^2 at every depth.  Here:

- T is computed from its definition, the smallest f32 d2 with
  fl(w / fl(sqrt(d2))) < theta (the criterion of nbody.wgsl:105-108; the GPU test
  test_opening_thresholds_are_the_criterion_boundary pins the kernels' thresholds to
  the same definition), and T >= 2.25 w^2 holds for theta up to 2/3 and fails above
  about 0.667;
- random cells, at offsets where the f32 spacing is coarse as well as fine, hold
  random f32 bodies; the centre of mass is summed in f64 in order and rounded once
  to f32 as the build does, and the distance of every body of the cell to it is
  formed as the walk forms it (dx = fl(c - x), fl(fl(dx dx) + fl(dy dy))).  It never
  exceeds 2 w^2 (1 + 2^-20), so it stays below T.
"""
import numpy as np

F32 = np.float32


def threshold(half, theta):
    """Smallest f32 d2 >= 0 with fl(2 half / fl(sqrt(d2))) < theta, by bisection on bit patterns."""
    half = np.asarray(half, F32)
    theta = np.asarray(theta, F32)
    w = F32(2.0) * half

    def passes(bits):
        d2 = bits.astype(np.uint32).view(F32)
        with np.errstate(divide="ignore", over="ignore", invalid="ignore"):
            return (w / np.sqrt(d2)) < theta

    lo = np.ones(half.shape, np.int64)                  # passes(lo) is false for these inputs
    hi = np.full(half.shape, 0x7F800000, np.int64)      # +inf passes
    assert passes(hi).all() and not passes(lo).any()
    while (hi - lo > 1).any():
        mid = (lo + hi) // 2
        p = passes(mid)
        hi = np.where(p, mid, hi)
        lo = np.where(p, lo, mid)
    return hi.astype(np.uint32).view(F32)


def test_flag_holds_up_to_two_thirds():
    rng = np.random.default_rng(11)
    half = np.exp(rng.uniform(np.log(1e-6), np.log(1e6), 4000)).astype(F32)
    for theta, want in [(0.5, True), (0.6, True), (0.66, True), (2.0 / 3.0 - 1e-4, True), (0.668, False),
                        (0.7, False), (0.75, False), (1.0, False)]:
        t = threshold(half, np.full(half.shape, theta, F32))
        w = 2.0 * half.astype(np.float64)
        ok = t.astype(np.float64) >= 2.25 * w * w
        assert ok.all() == want, theta


def test_cells_holding_the_target_fail_the_opening_test():
    rng = np.random.default_rng(12)
    theta = F32(0.5)
    worst = 0.0
    for trial in range(3000):
        side = float(np.exp(rng.uniform(np.log(1e-5), np.log(1e3))))
        # offsets up to 2^17 sides away: the f32 spacing of the coordinates reaches side / 64
        off = rng.uniform(-1.0, 1.0, 2) * side * float(2.0 ** rng.integers(0, 18))
        lo = off
        k = int(rng.integers(1, 40))
        pts = (lo + rng.uniform(0.0, 1.0, (k, 2)) * side).astype(F32)
        keep = ((pts.astype(np.float64) >= lo) & (pts.astype(np.float64) <= lo + side)).all(1)
        pts = pts[keep]
        if len(pts) == 0:
            continue
        m = rng.uniform(0.1, 10.0, len(pts)).astype(F32)
        if trial % 7 == 0:  # a few dominant masses pull the centre of mass to the box's edge
            m[rng.integers(0, len(pts))] = F32(1e6)
        mass = wx = wy = 0.0
        for (x, y), mi in zip(pts.astype(np.float64), m.astype(np.float64)):
            mass = mass + mi
            wx = wx + mi * x
            wy = wy + mi * y
        cx, cy = F32(wx / mass), F32(wy / mass)
        dx = cx - pts[:, 0]
        dy = cy - pts[:, 1]
        d2 = dx * dx + dy * dy  # f32 throughout, one rounding per operation as in the walk
        w = float(F32(2.0) * F32(side / 2.0))
        ratio = float(d2.astype(np.float64).max()) / (w * w)
        worst = max(worst, ratio)
        assert ratio <= 2.0 * (1.0 + 2.0 ** -20), (trial, ratio)
        t = threshold(np.array([side / 2.0], F32), np.array([theta], F32))[0]
        assert (d2 < t).all(), trial
    assert worst > 1.0  # the sample reaches well beyond the box's half diagonal
