"""Host-side builder of the tile neighbour list that the direct-space kernel
reads, in the reference's format (CudaNeighborListBuild.h:20-30,
CudaDirectForce_util.h:140-251): atoms in tiles of 32 consecutive atoms;
an i entry is (first atom of an i tile, periodic shift ish, first and last
index into the j-tile list); a j tile is its first atom plus 32 exclusion
words, bit t of word l excluding the pair (i = (l + t) mod 32, j = l).

This builds inputs for the tests and the benchmark (the reference builds
its list on the GPU, CudaNeighborList*, which is outside this engine).  Atoms
are wrapped into the box and sorted as the reference's list sorts them: x-y
columns about as wide as a tile of 32 atoms is tall, atoms of a column by z,
each column padded to whole tiles; every pair of tiles whose bounding boxes,
in one of the 27 periodic images, come closer than roff gets one (i tile,
shift, j tile) entry, each unordered atom pair and image once.  Exclusions:
padding atoms, the lower triangle of self tiles, and the listed topological
pairs (in every image).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

TILE = 32


@dataclass
class TileList:
    perm: np.ndarray        # sorted position -> original atom index (padding: -1)
    xyzq: np.ndarray        # (npad, 4) float32 sorted, wrapped coordinates (padding: q = 0)
    ientry: np.ndarray      # (n_ientry, 4) int32: iatomStart, ish, tileStart, tileEnd
    tile_indj: np.ndarray   # (n_tile,) int32: first atom of each j tile
    tile_excl: np.ndarray   # (n_tile, 32) uint32 exclusion words
    natom: int

    @property
    def npad(self) -> int:
        return len(self.xyzq)

    def sorted_index(self) -> np.ndarray:
        """original atom index -> sorted position"""
        inv = np.empty(self.natom, np.int64)
        m = self.perm >= 0
        inv[self.perm[m]] = np.nonzero(m)[0]
        return inv


def shift_index(s) -> int:
    """ish = (sx+1) + 3 (sy+1) + 9 (sz+1) (gpu_utils.h:308)"""
    return int(s[0] + 1 + 3 * (s[1] + 1) + 9 * (s[2] + 1))


SHIFTS = np.array([[s % 3 - 1, (s // 3) % 3 - 1, s // 9 - 1] for s in range(27)], np.int64)


def _diag_word(jl: int) -> int:
    """self tile without shift: each pair once (i < j), so exclude i >= j"""
    w = 0
    for t in range(32):
        if (jl + t) % 32 >= jl:
            w |= 1 << t
    return w


_DIAG = np.array([_diag_word(jl) for jl in range(32)], np.uint32)


def _column_order(pos, box, n, density_cell):
    """Sorted positions of the atoms: x-y columns of about density_cell A, atoms
    of a column by z, every column padded to whole tiles (padding: -1)."""
    ncx = max(1, int(box[0] // density_cell))
    ncy = max(1, int(box[1] // density_cell))
    cx = np.minimum((pos[:, 0] / box[0] * ncx).astype(np.int64), ncx - 1)
    cy = np.minimum((pos[:, 1] / box[1] * ncy).astype(np.int64), ncy - 1)
    col = cy * ncx + cx
    order = np.lexsort((pos[:, 2], col))
    counts = np.bincount(col, minlength=ncx * ncy)
    padded = (counts + TILE - 1) // TILE * TILE
    out = np.full(int(padded.sum()), -1, np.int64)
    starts = np.concatenate([[0], np.cumsum(padded)[:-1]])
    cstart = np.concatenate([[0], np.cumsum(counts)[:-1]])
    sc = col[order]
    rank = np.arange(n) - cstart[sc]
    out[starts[sc] + rank] = order
    return out


def build(xyzq, box, roff, excl_pairs=None, density_cell=None) -> TileList:
    x = np.ascontiguousarray(xyzq, np.float32)
    n = len(x)
    box = np.asarray(box, np.float64)
    pos = np.mod(x[:, :3].astype(np.float64), box)
    if density_cell is None:  # columns whose 32-atom tiles are about cubic
        density_cell = (TILE * np.prod(box) / max(n, 1)) ** (1.0 / 3.0)
    full_perm = _column_order(pos, box, n, density_cell)
    npad = len(full_perm)
    ntile = npad // TILE
    real = full_perm >= 0
    xs = np.zeros((npad, 4), np.float32)
    xs[real, :3] = pos[full_perm[real]].astype(np.float32)
    xs[real, 3] = x[full_perm[real], 3]
    # padding atoms sit on their tile's first atom, with no charge (and are excluded)
    first = np.repeat(np.arange(ntile) * TILE, TILE)
    xs[~real, :3] = xs[first[~real], :3]
    # tile bounding boxes (real atoms only)
    p = xs[:, :3].astype(np.float64).reshape(ntile, TILE, 3)
    valid = real.reshape(ntile, TILE)
    lo = np.where(valid[..., None], p, np.inf).min(1)
    hi = np.where(valid[..., None], p, -np.inf).max(1)
    cen = 0.5 * (lo + hi)
    hw = 0.5 * (hi - lo)
    shifts = SHIFTS * box[None, :]
    roff2 = roff * roff
    entries, indj, keys = [], [], []
    for I in range(ntile):
        J = np.arange(I, ntile)
        # box-box distance per shift: (nJ, 27, 3)
        d = np.abs((cen[I][None, None, :] + shifts[None, :, :]) - cen[J][:, None, :]) - (hw[I] + hw[J])[:, None, :]
        d2 = (np.maximum(d, 0.0) ** 2).sum(-1)
        near = d2 < roff2
        near[0, 14:] = False  # self tile: shifts 0..12 and 13 (with i < j) only
        for s in range(27):
            js = J[near[:, s]]
            if len(js) == 0:
                continue
            start = len(indj)
            indj.extend((js * TILE).tolist())
            keys.extend([(I, int(j), s) for j in js])
            entries.append((I * TILE, s, start, len(indj) - 1))
    ientry = np.array(entries, np.int32).reshape(-1, 4)
    tile_indj = np.array(indj, np.int32)
    excl = np.zeros((len(tile_indj), 32), np.uint32)
    key_arr = np.array(keys, np.int64).reshape(-1, 3)
    # self tiles without shift: i < j only
    selfm = (key_arr[:, 0] == key_arr[:, 1]) & (key_arr[:, 2] == 13)
    excl[selfm] |= _DIAG[None, :]
    # padding atoms: as j (whole word) and as i (bit t of word l: i = (l + t) mod 32)
    pad = (~real).reshape(ntile, TILE)
    t = np.arange(32)
    jword = np.where(pad, np.uint32(0xFFFFFFFF), np.uint32(0)).astype(np.uint32)
    iword = np.zeros((ntile, 32), np.uint32)
    for jl in range(32):
        bits = pad[:, (jl + t) % 32]  # (ntile, 32) over t
        iword[:, jl] = (bits.astype(np.uint64) << t.astype(np.uint64)).sum(1).astype(np.uint32)
    excl |= jword[key_arr[:, 1]]
    excl |= iword[key_arr[:, 0]]
    # topological exclusions, every image
    if excl_pairs is not None and len(excl_pairs):
        inv = np.empty(n, np.int64)
        inv[full_perm[real]] = np.nonzero(real)[0]
        lookup: dict = {}
        for k, (I, Jt, s) in enumerate(key_arr.tolist()):
            lookup.setdefault((I, Jt), []).append((k, s))
        ep = inv[np.asarray(excl_pairs, np.int64).reshape(-1, 2)]
        for a, b in ep.tolist():
            if a == b:
                continue
            for i, j in ((a, b), (b, a)):
                I, Jt = i // TILE, j // TILE
                if I > Jt:
                    continue
                jl = j % TILE
                tt = (i % TILE - jl) % TILE
                for k, s in lookup.get((I, Jt), ()):
                    excl[k, jl] |= np.uint32(1 << tt)
    return TileList(full_perm, xs, ientry, tile_indj, excl, n)


def water_exclusions(natom: int) -> np.ndarray:
    """The three intramolecular pairs of every O, H, H triplet (0,1), (0,2), (1,2)."""
    o = np.arange(0, natom - natom % 3, 3)
    return np.stack([np.concatenate([o, o, o + 1]), np.concatenate([o + 1, o + 2, o + 2])], 1)


def ex14_list(xyzq, box, pairs) -> np.ndarray:
    """(i, j, ishift) entries for pairs of atoms (indices as given), with the
    shift that makes x_i + shift - x_j the minimum image."""
    x = np.asarray(xyzq, np.float64)[:, :3]
    box = np.asarray(box, np.float64)
    pairs = np.asarray(pairs, np.int64).reshape(-1, 2)
    s = np.rint((x[pairs[:, 1]] - x[pairs[:, 0]]) / box).astype(np.int64)
    ish = (s[:, 0] + 1) + 3 * (s[:, 1] + 1) + 9 * (s[:, 2] + 1)
    return np.stack([pairs[:, 0], pairs[:, 1], ish], 1).astype(np.int32)


def count_pairs(tl: TileList, box, roff, chunk=2048):
    """(pair checks, interacting pairs) of the list: every (i, j) of every
    (i tile, shift, j tile) is a check; a pair is interacting when it is not
    excluded and closer than roff (what the direct kernel evaluates)."""
    box = np.asarray(box, np.float32)
    x = tl.xyzq[:, :3].astype(np.float32)
    ntp = len(tl.tile_indj)
    counts = tl.ientry[:, 3] - tl.ientry[:, 2] + 1
    ent = np.repeat(np.arange(len(tl.ientry)), counts)
    i0 = tl.ientry[ent, 0]
    sh = (SHIFTS[tl.ientry[ent, 1]] * box).astype(np.float32)
    lane = np.arange(32)
    rot = (lane[None, :] - lane[:, None]) % 32  # [j, i] -> bit (i - j) mod 32
    inter = 0
    for a in range(0, ntp, chunk):
        b = min(a + chunk, ntp)
        xi = x[i0[a:b, None] + lane[None, :]] + sh[a:b, None, :]       # (c, 32 i, 3)
        xj = x[tl.tile_indj[a:b, None] + lane[None, :]]                 # (c, 32 j, 3)
        d = xi[:, None, :, :] - xj[:, :, None, :]                      # (c, j, i, 3)
        r2 = (d * d).sum(-1)
        ex = (tl.tile_excl[a:b, :, None] >> rot[None, :, :].astype(np.uint32)) & 1  # (c, j, i)
        inter += int(((ex == 0) & (r2 < roff * roff)).sum())
    return ntp * 1024, inter


def water_box(xyzq, box, roff):
    """The direct-space inputs of a box of O, H, H triplets: the tile list with
    the intramolecular pairs excluded, and those pairs as 1-4 exclusion
    corrections in the list's sorted atom order (the reciprocal sum includes
    them, so the direct term subtracts them)."""
    water = water_exclusions(len(xyzq))
    tl = build(xyzq, box, roff, excl_pairs=water)
    ex14 = ex14_list(tl.xyzq, box, tl.sorted_index()[water])
    return tl, ex14
