"""CPU check of the two-sweep Jacobi pass's schedule (k_jacobi2 in
shg_focused_gaussian_beams_amd/csrc/thermal_kernels.hip), restated step for step: the
XCD-dealt block order, the tile / z-chunk decomposition, the LDS rings of four planes of
Ti (halo 2) and of the first sweep (halo 1), the staging at march step k, the Q values
loaded into registers one step ahead, the first sweep of plane k-2 and the second of
plane k-4, with one barrier per step: every ring entry a step reads was written by an
earlier step (never by the same one).  A toy stencil that reads every in-domain
neighbour stands in for the Jacobi formulas (which the kernel shares with the
single-sweep kernel); the schedule must give two plain sweeps exactly, read only staged
ring entries, and stay inside its rings — on grids that are not multiples of the tile,
with chunks that do not divide the depth, with both boundary kinds.  The interior tiles'
march (unrolled by four in the kernel: fixed ring slots, two register sets of loads two
steps ahead, the tile's Q held in registers) is restated and checked the same way.  The
bit-exactness of the real formulas is a GPU test (tests/test_gpu_jacobi2.py).
"""
import itertools

import numpy as np
import pytest

def sweep(T, Q, upd):
    hz, ny, nx = T.shape
    out = T.copy()
    for z, y, x in itertools.product(range(hz), range(ny), range(nx)):
        if not upd(x, y, nx, ny): continue
        out[z, y, x] = voxel(lambda dx, dy, dz: T[z+dz, y+dy, x+dx], x, y, z, nx, ny, hz) + Q[z, y, x]
    return out

W = {(1,0,0): .11, (-1,0,0): .13, (0,1,0): .17, (0,-1,0): .19, (0,0,1): .23, (0,0,-1): .29}
def voxel(ld, x, y, z, nx, ny, hz):
    v = 0.0
    for (dx, dy, dz), w in W.items():
        if 0 <= x+dx < nx and 0 <= y+dy < ny and 0 <= z+dz < hz:
            v += w * ld(dx, dy, dz)
    return v

def emulate(Ti, Q, TX, TY, zc, upd, ids=None):
    """Ti, Q: (count, hz, ny, nx); one launch whose slot j (gridDim.y = len(ids)) sweeps
    instance ids[j]; the other instances keep Tf's prior contents (here: NaN)."""
    count, hz, ny, nx = Ti.shape
    ids = list(range(count)) if ids is None else ids
    W2, H2, N2 = TX+4, TY+4, (TX+4)*(TY+4)
    W1, H1, N1 = TX+2, TY+2, (TX+2)*(TY+2)
    tiles_x, tiles_y = -(-nx//TX), -(-ny//TY)
    nch = -(-hz//zc)
    blocks = tiles_x*tiles_y*nch
    nb = blocks*len(ids)
    Tf_all = np.full_like(Ti, np.nan)
    for i in ids:
        for y in range(ny):
            for x in range(nx):
                if not upd(x, y, nx, ny): Tf_all[i, :, y, x] = Ti[i, :, y, x]
    Ti_all, Q_all = Ti, Q
    seen = set()
    for b in range(nb):                       # b = blockIdx.y * gridDim.x + blockIdx.x
        per, rem, xcd = nb >> 3, nb & 7, b & 7
        gb = xcd*per + min(xcd, rem) + (b >> 3)
        assert 0 <= gb < nb and gb not in seen; seen.add(gb)
        slot, lb = divmod(gb, blocks)
        assert 0 <= slot < len(ids)
        inst = ids[slot]
        Ti, Q, Tf = Ti_all[inst], Q_all[inst], Tf_all[inst]
        tile, chunk = divmod(lb, nch)
        ty, tx = divmod(tile, tiles_x)
        x0, y0 = tx*TX, ty*TY
        za, zb = chunk*zc, min(hz, chunk*zc + zc)
        Li = np.full((4, N2), np.nan); L1 = np.full((4, N1), np.nan)
        Li_at = [(None, None)] * 4; L1_at = [(None, None)] * 4   # (plane, step) of each ring slot
        # registers: Ti plane k over R2, Q plane k-2 at the R1 cells, Q plane k-4 at the
        # tile cells, each loaded one step ahead (NaN: not loaded)
        nxt = {}

        def fetch(k):
            pre = np.full(N2, np.nan); qn1 = np.full(N1, np.nan); qn0 = np.full(TX*TY, np.nan)
            if 0 <= k < hz:
                for t in range(N2):
                    row, col = divmod(t, W2); gx, gy = x0-2+col, y0-2+row
                    pre[t] = Ti[k, gy, gx] if (0 <= gx < nx and 0 <= gy < ny) else 0.0
            if 2 <= k <= hz + 1:
                for t in range(N1):
                    row, col = divmod(t, W1); gx, gy = x0-1+col, y0-1+row
                    qn1[t] = Q[k-2, gy, gx] if (0 <= gx < nx and 0 <= gy < ny) else 0.0
            if 4 <= k <= hz + 3:
                for t in range(TX*TY):
                    row, col = divmod(t, TX); gx, gy = x0+col, y0+row
                    qn0[t] = Q[k-4, gy, gx] if (gx < nx and gy < ny) else np.nan
            nxt.update(pre=pre, qn1=qn1, qn0=qn0)
        fetch(za-2)
        for k in range(za-2, zb+4):
            if 0 <= k < hz:
                Li[(k+4) % 4] = nxt["pre"]; Li_at[(k+4) % 4] = (k, k)
            q1c, q0c = nxt["qn1"], nxt["qn0"]
            if k+1 <= zb+3:
                fetch(k+1)
            p1 = k-2
            if max(za-1, 0) <= p1 <= min(zb, hz-1):
                for t in range(N1):
                    row, col = divmod(t, W1); gx, gy = x0-1+col, y0-1+row
                    c2 = (row+1)*W2 + col+1
                    v = 0.0
                    if 0 <= gx < nx and 0 <= gy < ny:
                        if upd(gx, gy, nx, ny):
                            def ld(dx, dy, dz):
                                i = c2 + dy*W2 + dx; assert 0 <= i < N2
                                slot = (p1+dz+4) % 4
                                assert Li_at[slot][0] == p1+dz and Li_at[slot][1] < k, (b, k, slot, Li_at[slot])
                                val = Li[slot, i]; assert not np.isnan(val), (b, k, gx, gy, dx, dy, dz); return val
                            q = q1c[t]; assert not np.isnan(q)
                            v = voxel(ld, gx, gy, p1, nx, ny, hz) + q
                        else:
                            slot = (p1+4) % 4
                            assert Li_at[slot][0] == p1 and Li_at[slot][1] < k
                            v = Li[slot, c2]
                    L1[(p1+4) % 4, t] = v
                L1_at[(p1+4) % 4] = (p1, k)
            p2 = k-4
            if za <= p2 < zb:
                for t in range(TX*TY):
                    row, col = divmod(t, TX); gx, gy = x0+col, y0+row
                    c1 = (row+1)*W1 + col+1
                    if gx < nx and gy < ny and upd(gx, gy, nx, ny):
                        def ld(dx, dy, dz):
                            i = c1 + dy*W1 + dx; assert 0 <= i < N1
                            slot = (p2+dz+4) % 4
                            assert L1_at[slot][0] == p2+dz and L1_at[slot][1] < k, (b, k, slot, L1_at[slot])
                            val = L1[slot, i]; assert not np.isnan(val); return val
                        q = q0c[t]; assert not np.isnan(q)
                        Tf[p2, gy, gx] = voxel(ld, gx, gy, p2, nx, ny, hz) + q
    assert len(seen) == nb
    return Tf_all

def _top(x, y, nx, ny):
    return y >= 1


def _oven(x, y, nx, ny):
    return 1 <= x < nx - 1 and 1 <= y < ny - 1


@pytest.mark.parametrize("nx,ny,hz,TX,TY,zc,bc", [(20, 13, 11, 8, 4, 3, "top"), (16, 9, 7, 8, 4, 4, "oven"),
                                                   (12, 12, 2, 16, 8, 5, "top"), (24, 10, 9, 8, 4, 2, "top")])
def test_two_sweep_schedule_equals_two_sweeps(nx, ny, hz, TX, TY, zc, bc):
    upd = _top if bc == "top" else _oven
    rng = np.random.default_rng(nx * 100 + hz)
    Ti = rng.random((hz, ny, nx))
    Q = 0.01 * rng.random((hz, ny, nx))
    ref = sweep(sweep(Ti, Q, upd), Q, upd)
    em = emulate(Ti[None], Q[None], TX, TY, zc, upd)[0]
    assert not np.isnan(em).any()
    assert np.array_equal(em, ref)


@pytest.mark.parametrize("count,ids,nx,ny,hz,TX,TY,zc", [(3, [0, 2], 20, 13, 7, 8, 4, 3),
                                                          (5, [1, 2, 4], 12, 9, 5, 8, 4, 2),
                                                          (2, [0, 1], 16, 8, 6, 8, 8, 4)])
def test_batched_two_sweep_schedule(count, ids, nx, ny, hz, TX, TY, zc):
    # one launch over a list of instances: the XCD dealing runs over every listed
    # instance's blocks (a bijection onto slot x block), each listed instance gets its two
    # sweeps exactly, and the others are not written
    rng = np.random.default_rng(count * 7 + nx)
    Ti = rng.random((count, hz, ny, nx))
    Q = 0.01 * rng.random((count, hz, ny, nx))
    em = emulate(Ti, Q, TX, TY, zc, _top, ids)
    for i in range(count):
        if i in ids:
            assert np.array_equal(em[i], sweep(sweep(Ti[i], Q[i], _top), Q[i], _top)), i
        else:
            assert np.isnan(em[i]).all(), i


def region_cell(i, W, H, ax, bx, ay, by):
    """region_cell of thermal_kernels.hip: cell i of a W x H region, inner rectangle first."""
    wi, hi = bx - ax, by - ay
    ni = wi * hi
    if i < ni:
        return ay + i // wi, ax + i % wi
    i -= ni
    if i < ay * W:
        return i // W, i % W
    i -= ay * W
    if i < (H - by) * W:
        return by + i // W, i % W
    i -= (H - by) * W
    if i < ax * hi:
        return ay + i // ax, i % ax
    i -= ax * hi
    wr = W - bx
    return ay + i // wr, bx + i % wr


@pytest.mark.parametrize("nx,ny,TX,TY", [(256, 256, 64, 8), (64, 64, 64, 16), (512, 512, 128, 8), (128, 96, 64, 8),
                                         (64, 64, 128, 8)])
def test_inner_first_cell_order(nx, ny, TX, TY):
    # the kernel's cell order over R1 (halo 1) and over the tile, at every tile position:
    # a bijection onto the region whose first cells are exactly the inner cells
    # (x in [1, nx-2], y in [1, ny-2]), so whole wavefronts share one formula
    W1, H1 = TX + 2, TY + 2
    for y0 in range(0, ny, TY):
        for x0 in range(0, nx, TX):
            for (W, H, ox, oy, ax, bx, ay, by) in (
                    (W1, H1, x0 - 1, y0 - 1, max(0, 2 - x0), None, max(0, 2 - y0), None),
                    (TX, TY, x0, y0, max(0, 1 - x0), None, max(0, 1 - y0), None)):
                if W == W1:
                    bx, by = max(ax, min(W1, nx - x0)), max(ay, min(H1, ny - y0))
                else:
                    bx, by = max(ax, min(TX, nx - 1 - x0)), max(ay, min(TY, ny - 1 - y0))
                cells = [region_cell(i, W, H, ax, bx, ay, by) for i in range(W * H)]
                assert sorted(cells) == [(r, c) for r in range(H) for c in range(W)]
                inner = [1 <= ox + c <= nx - 2 and 1 <= oy + r <= ny - 2 for r, c in cells]
                ni = sum(inner)
                assert all(inner[:ni]) and not any(inner[ni:]), (x0, y0, W)


def ring_cell_tile_first(i, TX, TY):
    """ring_cell_tile_first of thermal_kernels.hip, restated."""
    W = TX + 2
    if i < TX * TY:
        return 1 + i // TX, 1 + i % TX
    i -= TX * TY
    if i < 2 * W:
        return (0, i) if i < W else (TY + 1, i - W)
    i -= 2 * W
    return (1 + i, 0) if i < TY else (1 + i - TY, W - 1)


@pytest.mark.parametrize("TX,TY", [(64, 16), (64, 8), (128, 8)])
def test_interior_tile_first_order(TX, TY):
    # interior tiles visit R1 tile first: a bijection onto R1 whose first TX*TY entries are
    # the tile in the second sweep's (row-major) order, so slot r < R0 of a thread's R1
    # cells is its slot r of the tile and one Q value serves both sweeps
    W1, H1 = TX + 2, TY + 2
    cells = [ring_cell_tile_first(i, TX, TY) for i in range(W1 * H1)]
    assert sorted(cells) == [(r, c) for r in range(H1) for c in range(W1)]
    n0 = TX * TY
    assert n0 % 256 == 0                                   # the tile fills whole thread slots
    for tid in range(256):
        for r in range(n0 // 256):
            t = tid + 256 * r
            tile = region_cell(t, TX, TY, 0, TX, 0, TY)    # an interior tile's order
            assert cells[t] == (tile[0] + 1, tile[1] + 1)


def emulate_interior_march(Ti, Q, TX, TY, zc, upd):
    """The interior tiles' march (jacobi2_march<..., INNER>), restated for one instance with
    every tile marched that way (the schedule does not depend on the cell formulas): steps
    k = za-4 .. zb+3, unrolled by four in the kernel so that ring slots (plane & 3) and
    register sets (k & 1) are fixed; step k stages Ti plane k from set k&1, runs the first
    sweep on plane k-2 with Q plane k-2 from set k&1, holds that plane's tile values in
    qh[k&3], loads Ti plane k+2 and Q plane k into set k&1, and runs the second sweep on
    plane k-4 with qh[(k+2)&3].  Registers and ring slots carry (plane, step written) tags,
    so every read is checked to see the plane it needs, written at an earlier step (ring
    slots) or at least one step earlier (loads)."""
    hz, ny, nx = Ti.shape
    W2, N2, W1, N1 = TX+4, (TX+4)*(TY+4), TX+2, (TX+2)*(TY+2)
    tiles_x, tiles_y, nch = -(-nx//TX), -(-ny//TY), -(-hz//zc)
    Tf = np.where(np.array([[upd(x, y, nx, ny) for x in range(nx)] for y in range(ny)])[None], np.nan, Ti)
    for lb in range(tiles_x*tiles_y*nch):
        tile, chunk = divmod(lb, nch)
        ty, tx = divmod(tile, tiles_x)
        x0, y0 = tx*TX, ty*TY
        za, zb = chunk*zc, min(hz, chunk*zc + zc)
        Li = np.full((4, N2), np.nan); L1 = np.full((4, N1), np.nan)
        Li_at = [(None, None)] * 4; L1_at = [(None, None)] * 4
        pre = [(None, None, None)] * 2; qn1 = [(None, None, None)] * 2; qh = [(None, None)] * 4
        def plane_vals(A, k, xs, ys):
            kk = min(max(k, 0), hz-1)
            return np.array([A[kk, gy, gx] if (0 <= gx < nx and 0 <= gy < ny) else 0.0 for gx, gy in zip(xs, ys)])
        r2 = [divmod(t, W2) for t in range(N2)]
        r1 = [divmod(t, W1) for t in range(N1)]
        kb = (za - 4) & ~3
        for k in range(kb, zb + 4):
            if k < za - 4:
                continue
            S, P, live = k & 3, k & 1, k >= za - 2
            if live and 0 <= k < hz:
                vals, plane, at = pre[P]
                assert plane == k and at < k, (k, plane, at)
                Li[S] = vals; Li_at[S] = (k, k)
            p1 = k - 2
            if live and max(za-1, 0) <= p1 <= min(zb, hz-1):
                qv, qplane, qat = qn1[P]
                assert qplane == p1 and qat < k
                for t, (row, col) in enumerate(r1):
                    gx, gy = x0-1+col, y0-1+row
                    c2 = (row+1)*W2 + col+1
                    v = 0.0
                    if 0 <= gx < nx and 0 <= gy < ny:
                        def ld(dx, dy, dz):
                            slot = (p1+dz) & 3
                            assert Li_at[slot][0] == p1+dz and Li_at[slot][1] < k, (k, slot, Li_at[slot])
                            return Li[slot, c2 + dy*W2 + dx]
                        v = voxel(ld, gx, gy, p1, nx, ny, hz) + qv[t] if upd(gx, gy, nx, ny) else ld(0, 0, 0)
                    L1[p1 & 3, t] = v
                L1_at[p1 & 3] = (p1, k)
            assert S == (p1 + 2) & 3
            qh[S] = (qn1[P][0], qn1[P][1])                       # the tile's Q of plane k-2
            if k + 2 <= zb + 3:
                pre[P] = (plane_vals(Ti, k+2, [x0-2+c for _, c in r2], [y0-2+r for r, _ in r2]), k+2, k)
                qn1[P] = (plane_vals(Q, k, [x0-1+c for _, c in r1], [y0-1+r for r, _ in r1]), k, k)
            p2 = k - 4
            if live and za <= p2 < zb:
                qv, qplane = qh[(S+2) & 3]
                assert qplane == p2
                for row in range(TY):
                    for col in range(TX):
                        gx, gy = x0+col, y0+row
                        if not (gx < nx and gy < ny and upd(gx, gy, nx, ny)):
                            continue
                        c1 = (row+1)*W1 + col+1
                        def ld(dx, dy, dz):
                            slot = (p2+dz) & 3
                            assert L1_at[slot][0] == p2+dz and L1_at[slot][1] < k, (k, slot, L1_at[slot])
                            return L1[slot, c1 + dy*W1 + dx]
                        Tf[p2, gy, gx] = voxel(ld, gx, gy, p2, nx, ny, hz) + qv[c1]
    return Tf


@pytest.mark.parametrize("nx,ny,hz,TX,TY,zc,bc", [(20, 13, 11, 8, 4, 3, "top"), (16, 9, 7, 8, 4, 4, "oven"),
                                                   (12, 12, 2, 16, 8, 5, "top"), (24, 10, 9, 8, 4, 2, "top")])
def test_interior_march_schedule_equals_two_sweeps(nx, ny, hz, TX, TY, zc, bc):
    upd = _top if bc == "top" else _oven
    rng = np.random.default_rng(nx * 10 + hz)
    Ti = rng.random((hz, ny, nx))
    Q = 0.01 * rng.random((hz, ny, nx))
    ref = sweep(sweep(Ti, Q, upd), Q, upd)
    em = emulate_interior_march(Ti, Q, TX, TY, zc, upd)
    assert not np.isnan(em).any()
    assert np.array_equal(em, ref)
