"""The binning's pair rule restated on the host (the checker of bricks.hip's k_pairs_count and
k_pairs_fill): a sub-segment is paired with every periodic brick image its box meets, unless the
part of the box in that brick lies entirely outside the slab between the end planes (t0 >= 0,
t1 < 0 of tube.f90:806-813, corner extremes with a 1e-12 margin) or entirely beyond the reach of
its hits from ci0, sqrt(R^2 + (l + R tan(delta))^2) (l the chord length, delta the angle between
the two chords; no bound when cos(delta) <= 0.5)."""
import numpy as np


def _brick_images(lo, hi, N, bd=16):
    """(brick, image, clipped a, clipped b) of one axis met by the extended range [lo, hi]."""
    out = []
    for k in (-1, 0, 1):
        a, b = max(lo, 1 + k * N), min(hi, N + k * N)
        if a <= b:
            for br in range((a - k * N - 1) // bd, (b - k * N - 1) // bd + 1):
                org, ext = br * bd, min(bd, N - br * bd)
                out.append((br, k + 1, max(lo, org + 1 + k * N), min(hi, org + ext + k * N)))
    return out


def hit_reach2(t0v, t1v, ndci0, R):
    """Squared reach of a sub-segment's hits from ci0, with the binning's margins."""
    cd = (t0v[0] * t1v[0] + t0v[1] * t1v[1]) + t0v[2] * t1v[2]
    if not cd > 0.5:
        return np.inf
    tand = np.sqrt(max(1.0 - cd * cd, 0.0)) / cd
    e = ndci0 + R * tand
    return (R * R + e * e) * (1.0 + 1e-9) + 1e-12


def slab_pairs(geo, b, N, bd=(16, 16, 16)):
    """The binning's (sub-segment, brick, image) pairs, restated on the host from the oracle's
    boxes and chords: every brick image the box meets whose part of the box can hold a cell with
    t0 >= 0 and t1 < 0 (corner extremes, 1e-12 margin) and lies within the hits' reach from ci0;
    bd: the brick edge per axis. Returns (kept count, pruned list of (segment, per-axis clipped
    ranges))."""
    dx, start = 2 * np.pi / N, -np.pi
    kept, pruned = 0, []
    for s in range(geo.nseg):
        if b["R"][s] < 0:
            continue
        q = geo.sp_base[geo.seg_sec[s]] + geo.seg_j[s]
        t0v, t1v = geo.tci[q], geo.tci[q + 1]
        reach2 = hit_reach2(t0v, t1v, geo.ndci[q], b["R"][s])
        axes = [_brick_images(b["lo"][s][c], b["hi"][s][c], N, bd[c]) for c in range(3)]
        for ex in axes[0]:
            for ey in axes[1]:
                for ez in axes[2]:
                    rng = [(ex[2], ex[3]), (ey[2], ey[3]), (ez[2], ez[3])]
                    hi0 = [((rng[c][1] if t0v[c] > 0 else rng[c][0]) - 1) * dx + start for c in range(3)]
                    lo1 = [((rng[c][0] if t1v[c] > 0 else rng[c][1]) - 1) * dx + start for c in range(3)]
                    ci0, ci1 = b["ci0"][s], b["ci1"][s]
                    t0 = ((hi0[0] - ci0[0]) * t0v[0] + (hi0[1] - ci0[1]) * t0v[1]) + (hi0[2] - ci0[2]) * t0v[2]
                    t1 = ((lo1[0] - ci1[0]) * t1v[0] + (lo1[1] - ci1[1]) * t1v[1]) + (lo1[2] - ci1[2]) * t1v[2]
                    dd = [max(max(((rng[c][0] - 1) * dx + start) - ci0[c], ci0[c] - ((rng[c][1] - 1) * dx + start)), 0.0)
                          for c in range(3)]
                    d2 = [d * d for d in dd]
                    if t0 >= -1e-12 and t1 < 1e-12 and (d2[0] + d2[1]) + d2[2] < reach2:
                        kept += 1
                    else:
                        pruned.append((s, rng))
    return kept, pruned
