"""Bounds-checked CPU emulation of the set_points index math (no GPU needed).

Restates, in NumPy, what the set_points kernels of mripy_amd/csrc/gn_kernels.hip
compute for every sample -- k_prep (FP64 window start, offset, home gather tile),
the digit-wise stable counting sort (k_radix_count / k_radix_scatter /
k_bin_starts_sorted), k_pair_count / k_pair_keys / k_pair_records (axis_tiles, tap_box: the spread
tiles a window touches, the taps inside each, packed into 8- and 10-bit fields),
k_bank_order (LDS-bank permutation of each gather work item) -- and the LDS
addresses the gather and trajectory-gradient tap loops form from them, and asserts
that every index stays inside the buffer the kernel writes or reads.  Geometries:
the 3-D case (16, 20, 18) x 3 coils with 4000 samples on nf (32, 40, 36), the
default-tile 3-D and 2-D cases, wide windows, and grids smaller than one tile.
"""
import numpy as np
import pytest

from oracle import nufft_oracle as O

TWO_PI = 2 * np.pi
KGATHER_CHUNK = 512
KSPREAD_CHUNK = 512


def default_tiles(dim, ns, nf):
    """default_tiles + setup_tiles.fit (gn_kernels.hip, gn_api.hip)."""
    tile, gtile = ([4, 4, 8], [10, 10, 16]) if dim == 3 else ([8, 16], [24, 24])

    def staged():
        r = 16
        for a in range(dim):
            r *= min(gtile[a] + 2 * ns - 1, nf[a])
        return r
    while staged() > 160 * 1024:
        a = int(np.argmax(gtile[:dim]))
        if gtile[a] == 1:
            break
        gtile[a] -= 1

    def fit(t, a, tmax):
        t = min(min(max(t, 1), tmax), nf[a])
        if t + 2 * ns - 1 > nf[a]:
            t = nf[a] if nf[a] <= tmax else nf[a] - 2 * ns + 1
        return t
    tile = [fit(tile[a], a, 32) for a in range(dim)]
    gtile = [fit(gtile[a], a, 1 << 20) for a in range(dim)]
    return tile, gtile


def prep(coords, nf, ns, gtile):
    """k_prep: window start s (wrapped), t = xi/h - m, home gather tile key."""
    dim = len(nf)
    ng = [-(-nf[a] // gtile[a]) for a in range(dim)]
    s, t = [], []
    key = np.zeros(coords[0].shape[0], np.int64)
    for a in range(dim):
        xi, m = O.cell_index(coords[a], nf[a])
        h = TWO_PI / nf[a]
        t.append(((xi - h * m) / h).astype(np.float32))
        st = np.mod(m - ns, nf[a])
        assert (st >= 0).all() and (st < nf[a]).all()
        s.append(st)
        key = key * ng[a] + st // gtile[a]
    assert (key >= 0).all() and (key < np.prod(ng)).all()
    return s, t, key, ng


def axis_tiles(s, W, nf, T):
    out, last = [], -1
    for j in range(W):
        c = s + j
        if c >= nf:
            c -= nf
        q = c // T
        if q != last:
            out.append(q)
            last = q
    return out


def axis_tile_span(s, W, nf, T):
    """The kernels' closed form of axis_tiles: first tile and count (consecutive mod ntiles)."""
    if T == nf:
        return 0, 1
    q0 = s // T
    e = s + W - 1
    n = e // T - q0 + 1 if e < nf else (-(-nf // T) - q0) + (e - nf) // T + 1
    return q0, n


def tap_box(s, q, T, nf, W):
    org = q * T
    teff = min(T, nf - org)
    if T == nf:
        return 0, W, s
    v = org - s
    if v < 0:
        v += nf
    if v < W:
        return v, min(W - v, teff), 0
    return 0, min(W, teff - (s - org)), s - org


def bank_order(cells, org, region, nf, dim):
    """k_bank_order for one work item: position of each record, checked as a permutation."""
    n = cells.shape[0]
    assert n <= KGATHER_CHUNK
    res = np.zeros(n, np.int64)
    for i in range(n):
        addr = 0
        for a in range(dim):
            la = cells[i, a] - org[a]
            if la < 0:
                la += nf[a]
            assert 0 <= la < region[a]
            addr = addr * region[a] + la if a else la
        res[i] = addr & 15
    rank = np.zeros(n, np.int64)
    for v in range(16):
        idx = np.nonzero(res == v)[0]
        rank[idx] = np.arange(idx.size)
    sz = np.bincount(res, minlength=16)
    pos = np.array([sum(min(sz[v], k) + (1 if (v < r and sz[v] > k) else 0) for v in range(16))
                    for k, r in zip(rank, res)])
    assert sorted(pos.tolist()) == list(range(n))
    return pos


def emulate(coords, modes, ns, rat=2, check_pairs=True):
    dim = len(modes)
    nf = [max(rat * m, 2 * ns) for m in modes]
    tile, gtile = default_tiles(dim, ns, nf)
    W = 2 * ns
    s, t, key, ng = prep(coords, nf, ns, gtile)
    M = key.shape[0]
    order = np.argsort(key, kind="stable")          # the counting sort's result (pinned on the GPU)
    Kg = int(np.prod(ng))
    bin_start = np.searchsorted(key[order], np.arange(Kg + 1), side="left")
    cells = np.stack([s[a][order] for a in range(dim)], 1)
    # gather work items and the LDS addresses of the tap loop
    region = [min(gtile[a] + W - 1, nf[a]) for a in range(dim)]
    assert 16 * np.prod(region) <= 160 * 1024
    for k in range(Kg):
        b, e = bin_start[k], bin_start[k + 1]
        rr = k
        org = [0] * dim
        for a in reversed(range(dim)):
            org[a] = (rr % ng[a]) * gtile[a]
            rr //= ng[a]
        for s0 in range(b, e, KGATHER_CHUNK):
            c = cells[s0:min(e, s0 + KGATHER_CHUNK)]
            bank_order(c, org, region, nf, dim)
            for a in range(dim):
                la = np.mod(c[:, a] - org[a], nf[a])
                wrap = region[a] == nf[a]
                hi = la + W - 1
                if wrap:
                    hi = np.where(hi >= region[a], hi - nf[a], hi)
                assert (hi < region[a]).all() and (la >= 0).all()
    if not check_pairs:
        return
    # spread pairs: tiles per axis, taps inside each, packed fields
    ntiles = [-(-nf[a] // tile[a]) for a in range(dim)]
    P = 0
    keys = []
    for i in range(M):
        tl = [axis_tiles(int(cells[i, a]), W, nf[a], tile[a]) for a in range(dim)]
        for a in range(dim):
            q0, na = axis_tile_span(int(cells[i, a]), W, nf[a], tile[a])
            assert tl[a] == [(q0 + k) % ntiles[a] for k in range(na)]
        n = int(np.prod([len(x) for x in tl]))
        ns_ax = [len(x) for x in tl] + [1] * (3 - dim)
        for local, combo in enumerate(np.ndindex(*[len(x) for x in tl])):
            # k_pair_records: local index (i * n1 + k) * n2 + m of the emission order
            m_, r_ = local % ns_ax[2], local // ns_ax[2]
            assert (r_ // ns_ax[1], r_ % ns_ax[1], m_)[:dim] == tuple(combo)
            kk = 0
            for a in range(dim):
                q = tl[a][combo[a]]
                j, cnt, l0 = tap_box(int(cells[i, a]), q, tile[a], nf[a], W)
                assert 0 <= j < 256 and 1 <= cnt < 256 and 0 <= l0 < 1024
                assert j + cnt <= W and l0 + cnt <= tile[a] if tile[a] != nf[a] else True
                kk = kk * ntiles[a] + q
            assert 0 <= kk < np.prod(ntiles)
            keys.append(kk)
        P += n
    assert len(keys) == P and P < 2 ** 31
    pstart = np.searchsorted(np.sort(np.array(keys)), np.arange(np.prod(ntiles) + 1))
    assert pstart[-1] == P
    # split-tile slots
    sizes = np.diff(pstart)
    assert (sizes >= 0).all()
    return P


def _random(M, dim, seed, spread=1.0):
    rng = np.random.default_rng(seed)
    return [(spread * np.pi * (2 * rng.random(M) - 1)).astype(np.float32).astype(np.float64) for _ in range(dim)]


def test_traj_grad_geometry_16_20_18():
    # the geometry of tests/test_gpu_fft.py::test_traj_grad_engines_agree (default_rng(3))
    rng = np.random.default_rng(3)
    cs = [(np.pi * (2 * rng.random(4000) - 1)).astype(np.float32).astype(np.float64) for _ in range(3)]
    emulate(cs, (16, 20, 18), 3)


@pytest.mark.parametrize("modes,ns,M", [((32, 32, 32), 3, 3000), ((24, 20, 18), 5, 2000), ((17, 19, 21), 3, 2000),
                                        ((8, 8, 8), 6, 500), ((16, 16, 16), 8, 300)])
def test_3d_geometries(modes, ns, M):
    emulate(_random(M, 3, sum(modes) + ns), modes, ns)


@pytest.mark.parametrize("modes,ns,rat,M", [((128, 128), 3, 2, 4000), ((40, 36), 6, 3, 3000), ((16, 16), 14, 3, 2000),
                                            ((7, 9), 3, 2, 500)])
def test_2d_geometries(modes, ns, rat, M):
    emulate(_random(M, 2, sum(modes) + ns, spread=3.0), modes, ns, rat=rat)


def test_wide_window_default_gather_tile_fits_lds():
    for ns in (7, 8):
        nf = [256, 256, 256]
        tile, gtile = default_tiles(3, ns, nf)
        staged = 16 * np.prod([min(g + 2 * ns - 1, n) for g, n in zip(gtile, nf)])
        assert staged <= 160 * 1024, (ns, gtile)


def stage_rows_emulated(region, org, nf, nthreads):
    """stage_region_rows (gn_kernels.hip): the LDS cell and grid index every lane writes,
    stepping rows exactly as the kernel does (carry on the second axis, U rows per pass)."""
    dim = len(nf)
    R0, R1 = region[0], region[1]
    R2 = region[2] if dim == 3 else 1
    RL = R2 if dim == 3 else R1
    nrow = R0 * R1 if dim == 3 else R0
    U = 4
    writes = []
    for tid in range(nthreads):
        lane, wave = tid & 63, tid >> 6
        rpw = 64 // RL
        k, c = lane // RL, lane % RL
        if k >= rpw:
            continue
        step = (nthreads >> 6) * rpw
        ql = org[dim - 1] + c
        if ql >= nf[dim - 1]:
            ql -= nf[dim - 1]
        rho = wave * rpw + k
        if dim == 3:
            r0, r1 = rho // R1, rho % R1
            s0, s1 = step // R1, step % R1
        else:
            r0, r1, s0, s1 = rho, 0, step, 0
        while rho < nrow:
            for u in range(U):
                if rho + u * step < nrow:
                    q0 = org[0] + r0
                    if q0 >= nf[0]:
                        q0 -= nf[0]
                    assert 0 <= q0 < nf[0]
                    if dim == 3:
                        q1 = org[1] + r1
                        if q1 >= nf[1]:
                            q1 -= nf[1]
                        assert 0 <= q1 < nf[1]
                        off = (q0 * nf[1] + q1) * nf[2] + ql
                    else:
                        off = q0 * nf[1] + ql
                    assert 0 <= off < min(int(np.prod(nf)), 2 ** 31 - 1)
                    writes.append(((rho + u * step) * RL + c, off))
                if dim == 3:
                    r1 += s1
                    cc = 1 if r1 >= R1 else 0
                    r1 -= cc * R1
                    r0 += s0 + cc
                else:
                    r0 += step
            rho += U * step
    return writes


@pytest.mark.parametrize("modes,ns,rat", [((128, 128, 128), 3, 2), ((128, 128, 128), 5, 2), ((16, 20, 18), 3, 2),
                                          ((24, 20, 18), 5, 2), ((8, 8, 8), 6, 2), ((128, 128), 3, 2),
                                          ((40, 36), 6, 3), ((7, 9), 3, 2)])
def test_row_staging_covers_region_once(modes, ns, rat):
    """Every cell of a gather region is staged exactly once, at the wrapped grid index
    of its region coordinates, for tiles at the origin, in the interior and at the
    upper edge of every axis (wrap-around), with 512-thread workgroups."""
    dim = len(modes)
    nf = [max(rat * m, 2 * ns) for m in modes]
    _, gtile = default_tiles(dim, ns, nf)
    region = [min(gtile[a] + 2 * ns - 1, nf[a]) for a in range(dim)]
    RL = region[-1]
    if RL > 64 or (64 // RL) * RL < 56:
        pytest.skip("rows that fill less than 7/8 of a wave use the flat staging loop")
    ng = [-(-nf[a] // gtile[a]) for a in range(dim)]
    for pick in ("first", "middle", "last"):
        org = [{"first": 0, "middle": ng[a] // 2, "last": ng[a] - 1}[pick] * gtile[a] for a in range(dim)]
        w = stage_rows_emulated(region, org, nf, 512)
        RC = int(np.prod(region))
        cells = np.array([x[0] for x in w])
        assert len(cells) == RC and np.array_equal(np.sort(cells), np.arange(RC))
        want = {}
        for r in np.ndindex(*region):
            q = [(org[a] + r[a]) % nf[a] for a in range(dim)]
            lin = int(np.ravel_multi_index(r, region))
            want[lin] = int(np.ravel_multi_index(q, nf))
        assert all(want[cidx] == off for cidx, off in w)


def spread_items_host(pstart, chunk):
    """build_work_items (gn_host_plan.h), spread part: items (tiles with pairs, then the
    empty tiles), reduce items, slots."""
    items, empty, red, slots = [], [], [], 0
    for k in range(len(pstart) - 1):
        b, e = int(pstart[k]), int(pstart[k + 1])
        if e == b:
            empty.append((k, b, e, -1))
        elif e - b <= chunk:
            items.append((k, b, e, -1))
        else:
            nch = -(-(e - b) // chunk)
            red.append((k, slots, nch, 0))
            for c in range(nch):
                items.append((k, b + c * chunk, min(e, b + (c + 1) * chunk), slots + c))
            slots += nch
    return items + empty, red, slots


def spread_items_device(pstart, chunk):
    """k_item_counts, four exclusive scans and k_item_write (gn_kernels.hip)."""
    K = len(pstart) - 1
    n = np.diff(pstart)
    nch = -(-n // chunk)
    split = n > chunk
    ni = np.where(split, nch, np.where(n > 0, 1, 0))
    ne = (n == 0).astype(np.int64)
    ns = np.where(split, nch, 0)
    nr = split.astype(np.int64)
    ioff = np.concatenate([[0], np.cumsum(ni)])
    eoff = np.concatenate([[0], np.cumsum(ne)])
    soff = np.concatenate([[0], np.cumsum(ns)])
    roff = np.concatenate([[0], np.cumsum(nr)])
    nfull = int(ioff[-1])
    items = [None] * (nfull + int(eoff[-1]))
    red = [None] * int(roff[-1])
    for k in range(K):
        b, e = int(pstart[k]), int(pstart[k + 1])
        if e == b:
            items[nfull + int(eoff[k])] = (k, b, e, -1)
            continue
        o = int(ioff[k])
        if e - b <= chunk:
            items[o] = (k, b, e, -1)
            continue
        c_n, s0 = int(nch[k]), int(soff[k])
        red[int(roff[k])] = (k, s0, c_n, 0)
        for c in range(c_n):
            items[o + c] = (k, b + c * chunk, min(e, b + (c + 1) * chunk), s0 + c)
    return items, red, int(soff[-1])


@pytest.mark.parametrize("seed,K,P", [(0, 1, 0), (1, 7, 3000), (2, 4096, 70000), (3, 300, 0)])
def test_device_spread_items_follow_host_rule(seed, K, P):
    """The device build of the spread work items (per-tile counts, exclusive scans,
    writes) gives exactly the host planner's items, reduce items and slot count, for
    empty, uniform and clustered tile loads; and the occupancy masks (k_occ_masks)
    mark exactly the columns' tiles that hold pairs."""
    rng = np.random.default_rng(seed)
    cnt = np.zeros(K, np.int64)
    if P:
        k = np.where(rng.random(P) < 0.5, rng.integers(0, K, P), rng.integers(0, max(1, K // 50), P))
        np.add.at(cnt, k, 1)
    pstart = np.concatenate([[0], np.cumsum(cnt)])
    assert spread_items_device(pstart, 512) == spread_items_host(pstart, 512)
    nt = [16, 16, K // 256] if K % 256 == 0 else [K, 1, 1]
    ncolt = nt[1] * nt[2]
    mask_k = (nt[0] + 63) // 64
    for c in range(ncolt):
        m = 0
        for t0 in range(nt[0]):
            if pstart[t0 * ncolt + c + 1] > pstart[t0 * ncolt + c]:
                m |= 1 << (t0 // mask_k)
        want = 0
        for it in spread_items_host(pstart, 512)[0]:
            if it[2] > it[1] and it[0] % ncolt == c:
                want |= 1 << ((it[0] // ncolt) // mask_k)
        assert m == want
