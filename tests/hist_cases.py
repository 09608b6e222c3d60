"""Edge and sample sets for the custom-edge histogram tests (CPU emulation
and GPU): spread edges, 200 edges inside one cell of the kernel's table,
duplicate edges, a float range whose width overflows to inf, ranges a few
ulps wide and denormal, integer ranges of every width.  Samples include
every edge, its neighbours, NaN and +-inf."""
import numpy as np


def range_level_sets():
    """Edge sets (float32, uint32) that stress the range histogram's cell table."""
    g = np.random.default_rng(77)
    f = np.float32
    sets = []
    sets.append(np.sort(g.random(257).astype(f) * f(1024)))                     # spread
    clustered = np.concatenate([np.linspace(10, 10.001, 200), np.linspace(-50, 900, 57)])
    sets.append(np.sort(clustered.astype(f)))                                    # 200 in a cell
    sets.append(np.sort(np.repeat(g.random(20).astype(f) * f(100), 3)))          # duplicates
    sets.append(np.array([-3.4e38, -1.0, 0.0, 1.0, 3.4e38], f))                 # upper-lower = inf
    one = f(1.0)
    sets.append(np.array([one, np.nextafter(one, f(2)),
                          np.nextafter(np.nextafter(one, f(2)), f(2))], f))     # 2 ulps wide
    sets.append(np.array([0.0, 1e-45, 2e-45, 1e-44], f))                         # denormal range
    sets.append(np.array([-1.0, 1.0], f))                                        # one bin
    isets = []
    isets.append(np.sort(g.integers(0, 1 << 32, 257, dtype=np.uint64).astype(np.uint32)))
    isets.append(np.array([0, 1 << 31, 0xFFFFFFFF], np.uint32))
    isets.append(np.concatenate([[0], np.arange(1000, 1200), [0xFFFFFFFF]]).astype(np.uint32))
    isets.append(np.array([5, 6], np.uint32))
    isets.append(np.array([5, 6, 7], np.uint32))
    isets.append(np.array([3, 3, 3, 9, 9, 4096 + 9, 4096 + 9, 1 << 20], np.uint32))
    return sets, isets


def range_samples(lv, n_rand, seed):
    g = np.random.default_rng(seed)
    if lv.dtype == np.float32:
        width = float(lv[-1]) - float(lv[0])
        body = ((g.random(n_rand) * width * 1.1 + float(lv[0])).astype(np.float32) if width < 1e38
                else g.standard_normal(n_rand).astype(np.float32) * np.float32(1e30))
        return np.concatenate([
            body, lv, np.nextafter(lv, np.float32(np.inf)), np.nextafter(lv, np.float32(-np.inf)),
            np.array([np.nan, np.inf, -np.inf, 0.0, -0.0], np.float32)]).astype(np.float32)
    return np.concatenate([g.integers(0, 1 << 32, n_rand, dtype=np.uint64).astype(np.uint32),
                           lv, lv + np.uint32(1), lv - np.uint32(1),
                           g.integers(0, 1 << 21, 4096, dtype=np.uint64).astype(np.uint32)])
