"""The bulk-window schedules of the tiled inside (messi_amd/csrc/inside.hip,
inside_step_kernel / inside_batch_run) and of the MEA fill (csrc/mea.hip,
mea_step_kernel / mea_run), restated on the CPU: the host's per-launch item
counts, the kernels' decode of a bulk item (binary search over the items of the
first k distances) and the launches each visit covers.  The window must not
change what is added to a tile or when its products become ready, so for every
tile (I, J) and split tile K in the pairing band the product is added exactly
once, in the launch max(K-I, J-K)+1 that made both operands final or later, and
no later than the tile's own finalisation in launch J-I; tiles outside the band
get no product.  The GPU tests (test_gpu_window.py, test_gpu_mea.py) check
that the kernels' results do not depend on the window."""
import pytest


def _items(NT, t, W, k):
    return k * (NT - t - 1) - W * (k * (k - 1) // 2)


def _decode(NT, t, W, nd, idx):
    lo, hi = 0, nd - 1
    while lo < hi:
        mid = (lo + hi + 1) // 2
        if _items(NT, t, W, mid) <= idx:
            lo = mid
        else:
            hi = mid - 1
    dist = t + 1 + lo * W
    I = idx - _items(NT, t, W, lo)
    return I, I + dist


def _prior(t, d, W, KB, fin):
    """The kernels' test for an earlier scheduled bulk visit of a tile at
    distance d (visits t' = d-1, d-1-W, ... >= (d+3)//2; a visit is scheduled
    unless t'-1 > KB and d > t'-1+KB)."""
    smin = (d + 3) // 2
    last = d - 1 if fin else t - W
    if last < smin:
        return False
    first = last - W * ((last - smin) // W)
    return first - 1 <= KB or d <= last - 1 + KB


def _inside_schedule(NT, KB, W):
    bp = 2 if W == 1 else 1
    done, written = {}, set()
    for t in range(NT):
        dmax = min(2 * t - 2, NT - 1)
        if t - 1 > KB:
            dmax = min(dmax, t - 1 + KB)
        host = sum(NT - d for d in range(t + 1, dmax + 1, W)) if dmax >= t + 1 else 0
        nd = (dmax - t - 1) // W + 1 if dmax >= t + 1 else 0
        assert _items(NT, t, W, nd) == host
        for idx in range(((host + bp - 1) // bp) * bp):
            if idx >= host:
                continue
            I, J = _decode(NT, t, W, nd, idx)
            assert 0 <= I < J < NT
            assert _prior(t, J - I, W, KB, False) == ((I, J) in written)
            written.add((I, J))  # every scheduled visit writes its accumulator tile
            smin = max(t - W + 1, (J - I + 3) // 2)
            for s in range(t - W + 1, t + 1):
                k1, k2 = I + s - 1, J - (s - 1)
                if s >= smin and k1 - I <= KB:
                    done.setdefault((I, J, k1), []).append((t, s))
                if s >= smin and k2 != k1 and k2 - I <= KB:
                    done.setdefault((I, J, k2), []).append((t, s))
        for I in range(NT - t):  # finalising workgroups: launch t's own products
            J = I + t
            assert _prior(t, t, W, KB, True) == ((I, J) in written)
            k1, k2 = I + t - 1, J - (t - 1)
            if t >= 2 and k1 - I <= KB:
                done.setdefault((I, J, k1), []).append((t, t))
            if t >= 2 and k2 != k1 and k2 - I <= KB:
                done.setdefault((I, J, k2), []).append((t, t))
    return done


@pytest.mark.parametrize("W", [1, 2, 4])
@pytest.mark.parametrize("NT", [1, 2, 3, 5, 8, 13, 21, 40, 94])
def test_inside_window_covers_each_product_once(NT, W):
    for KB in sorted({0, 1, 2, 5, NT - 1, NT}):
        done = _inside_schedule(NT, KB, W)
        for I in range(NT):
            for J in range(I + 2, NT):
                for K in range(I + 1, J):
                    got = done.get((I, J, K), [])
                    if K - I > KB:
                        assert not got
                        continue
                    assert len(got) == 1, (NT, KB, W, I, J, K, got)
                    t, s = got[0]
                    assert s == max(K - I, J - K) + 1 and s <= t <= J - I


def test_inside_window_one_matches_per_launch_schedule():
    """W = 1 is the schedule with a visit every launch: each bulk item gets the
    products of its own launch only."""
    done = _inside_schedule(30, 30, 1)
    assert all(t == s for v in done.values() for t, s in v)


@pytest.mark.parametrize("NT", list(range(1, 30)) + [94, 300])
def test_mea_window_covers_each_product_once(NT):
    W = 4
    done, visited = {}, set()
    for t in range(NT):
        dmax = min(2 * t - 2, NT - 1)
        nb = _items(NT, t, W, (dmax - t - 1) // W + 1) if dmax >= t + 1 else 0
        for item in range(NT - t + nb):
            if item < NT - t:
                fin, I, J = True, item, item + t
            else:
                fin = False
                I, J = _decode(NT, t, W, (dmax - t - 1) // W + 1, item - (NT - t))
                assert 0 <= I < J < NT
            smin = (J - I + 3) // 2
            prior = t >= 4 if fin else t - W >= smin
            assert prior == ((I, J) in visited)  # the accumulator is read only once written
            for st in range(t if fin else max(t - W + 1, smin), t + 1):
                if st < 2:
                    continue
                k1, k2 = I + st - 1, J - st + 1
                for K in ([k1] if k2 == k1 else [k1, k2]):
                    done.setdefault((I, J, K), []).append((t, st))
            if not fin:
                visited.add((I, J))
    for I in range(NT):
        for J in range(I + 2, NT):
            for K in range(I + 1, J):
                got = done.get((I, J, K), [])
                assert len(got) == 1, (NT, I, J, K, got)
                t, st = got[0]
                assert st == max(K - I, J - K) + 1 and st <= t <= J - I
