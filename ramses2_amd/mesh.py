"""Oct meshes in the reference layout.

An oct holds 2^ndim cells; a level is a contiguous run of octs; hydro arrays
are u(1:twotondim, 1:nvar, 1:ngrid) in Fortran order, i.e. numpy arrays of
shape (ngrid, nvar, twotondim) in C order (amr/amr_commons.f90:816-819,
amr/init_amr.f90:143).  Oct keys (ckey) are integer oct coordinates at the
oct's level, 0 <= ckey < 2^(lev-1) in the full box.
"""
import numpy as np

NVAR = 5

# type(oct) of amr/oct_commons.f90:5-10 (nhilbert=1)
OCT3 = np.dtype([("hkey", "<i8"), ("ckey", "<i4", (3,)), ("refined", "u1", (8,)), ("lev", "<i4")])
OCT1 = np.dtype([("hkey", "<i8"), ("ckey", "<i4", (1,)), ("refined", "u1", (2,)), ("pad", "u1", (2,)),
                 ("lev", "<i4"), ("pad2", "<i4")])
assert OCT3.itemsize == 32 and OCT1.itemsize == 24


def _part1by2(x):
    x = x.astype(np.uint64) & np.uint64(0x1FFFFF)
    x = (x | (x << np.uint64(32))) & np.uint64(0x1F00000000FFFF)
    x = (x | (x << np.uint64(16))) & np.uint64(0x1F0000FF0000FF)
    x = (x | (x << np.uint64(8))) & np.uint64(0x100F00F00F00F00F)
    x = (x | (x << np.uint64(4))) & np.uint64(0x10C30C30C30C30C3)
    x = (x | (x << np.uint64(2))) & np.uint64(0x1249249249249249)
    return x


def morton3(ix, iy, iz):
    return _part1by2(ix) | (_part1by2(iy) << np.uint64(1)) | (_part1by2(iz) << np.uint64(2))


def level_params(nlevelmax, ndim, box_octs_at=None):
    """ckey_max(l), key_off(l), box_ckey_min/max(1:3,l) for l = 1..nlevelmax+1.

    key_off(1)=1, key_off(l)=key_off(l-1)+hkey_max(l-1) as metal_runner.f90:26-30
    with hkey_max(l) = (2^(l-1))^ndim keys at level l.
    """
    L = nlevelmax + 1
    ckey_max = np.array([2 ** (l - 1) for l in range(1, L + 1)], dtype=np.int32)
    key_off = np.zeros(L, dtype=np.int64)
    key_off[0] = 1
    for l in range(1, L):
        key_off[l] = key_off[l - 1] + int(ckey_max[l - 1]) ** ndim
    box_min = np.zeros((L, 3), dtype=np.int32)
    box_max = np.ones((L, 3), dtype=np.int32)
    for l in range(L):
        box_max[l, :ndim] = ckey_max[l]
    return ckey_max, key_off, box_min, box_max


class UniformMesh:
    """One uniform level of n octs per dimension, octs in Morton order.

    n defaults to the full level, 2^(level-1); a smaller n is a computational
    box [0, n) per dimension inside the level (box_ckey_min = 0,
    box_ckey_max = n, amr/init_amr.f90:250-280), which gives grids of any
    even number of cells per dimension."""

    def __init__(self, ndim, level, boxlen=1.0, n=None):
        assert ndim in (1, 3)
        self.ndim = ndim
        self.level = level
        self.boxlen = boxlen
        self.n = 2 ** (level - 1) if n is None else int(n)   # octs per dimension
        assert 1 <= self.n <= 2 ** (level - 1)
        self.full_box = self.n == 2 ** (level - 1)
        self.ncell = 2 * self.n               # cells per dimension
        self.ttd = 2 ** ndim
        self.dx = boxlen / 2 ** level
        n = self.n
        if ndim == 1:
            self.ckey = np.arange(n, dtype=np.int32)[:, None]
        else:
            ii = np.arange(n, dtype=np.int64)
            iz, iy, ix = np.meshgrid(ii, ii, ii, indexing="ij")
            ix, iy, iz = ix.ravel(), iy.ravel(), iz.ravel()
            order = np.argsort(morton3(ix, iy, iz), kind="stable")
            self.ckey = np.stack([ix[order], iy[order], iz[order]], axis=1).astype(np.int32)
        self.noct = self.ckey.shape[0]
        # cell (i[,j,k]) of the level -> (oct index 0-based, cell index in oct)
        self._build_maps()

    def _build_maps(self):
        n, nd = self.n, self.ndim
        if nd == 1:
            octmap = np.zeros(n, dtype=np.int64)
            octmap[self.ckey[:, 0]] = np.arange(self.noct)
            self.octmap = octmap
        else:
            octmap = np.zeros((n, n, n), dtype=np.int64)  # [iz,iy,ix]
            octmap[self.ckey[:, 2], self.ckey[:, 1], self.ckey[:, 0]] = np.arange(self.noct)
            self.octmap = octmap

    def grid_records(self):
        if self.ndim == 3:
            g = np.zeros(self.noct, dtype=OCT3)
            g["ckey"] = self.ckey
        else:
            g = np.zeros(self.noct, dtype=OCT1)
            g["ckey"][:, 0] = self.ckey[:, 0]
        g["lev"] = self.level
        return g

    def from_cells(self, cells):
        """cells: (nvar, N[,N,N]) array indexed [var, (k, j,) i] -> uold (noct, nvar, ttd)."""
        nv = cells.shape[0]
        u = np.zeros((self.noct, nv, self.ttd), dtype=np.float64)
        if self.ndim == 1:
            c = cells.reshape(nv, self.n, 2)                      # [v, ox, i2]
            u[self.octmap] = np.transpose(c, (1, 0, 2))
        else:
            n = self.n
            c = cells.reshape(nv, n, 2, n, 2, n, 2)              # [v, oz, k2, oy, j2, ox, i2]
            c = np.transpose(c, (1, 3, 5, 0, 2, 4, 6))             # [oz, oy, ox, v, k2, j2, i2]
            c = c.reshape(n, n, n, nv, 8)                          # ind = i2 + 2 j2 + 4 k2
            u[self.octmap.ravel()] = c.reshape(n ** 3, nv, 8)
        return u

    def to_cells(self, u):
        nv = u.shape[1]
        if self.ndim == 1:
            c = u[self.octmap]                                     # [ox, v, i2]
            return np.transpose(c, (1, 0, 2)).reshape(nv, self.ncell)
        n = self.n
        c = u[self.octmap.ravel()].reshape(n, n, n, nv, 2, 2, 2)   # [oz,oy,ox,v,k2,j2,i2]
        c = np.transpose(c, (3, 0, 4, 1, 5, 2, 6))                 # [v, oz, k2, oy, j2, ox, i2]
        return c.reshape(nv, self.ncell, self.ncell, self.ncell)

    def level_params(self, nlevelmax=None):
        ckey_max, key_off, box_min, box_max = level_params(nlevelmax or self.level, self.ndim)
        if not self.full_box:
            for l in range(1, len(ckey_max) + 1):
                sh = l - self.level
                b = self.n << sh if sh >= 0 else max(1, self.n >> -sh)
                box_max[l - 1, :self.ndim] = b
        return ckey_max, key_off, box_min, box_max

    def box(self):
        """(lev, ckey_min, ckey_max) of the computational box, or None for the full level."""
        if self.full_box:
            return None
        hi = [self.n if d < self.ndim else 1 for d in range(3)]
        return (self.level, [0, 0, 0], hi)

    @classmethod
    def for_cells(cls, ndim, ncell, boxlen=1.0):
        """The smallest level holding ncell (even) cells per dimension, with the
        box length scaled so that the box of ncell cells has size boxlen."""
        assert ncell % 2 == 0 and ncell >= 2
        n = ncell // 2
        level = 1
        while 2 ** (level - 1) < n:
            level += 1
        return cls(ndim, level, boxlen=boxlen * 2 ** level / ncell, n=n)


def prim_to_cons(rho, vel, p, gamma):
    """(rho, (vx,vy,vz), p) cell arrays -> conservative (5, ...) array."""
    vx, vy, vz = vel
    ekin = 0.5 * rho * (vx ** 2 + vy ** 2 + vz ** 2)
    return np.stack([rho, rho * vx, rho * vy, rho * vz, p / (gamma - 1.0) + ekin])


class AmrMesh:
    """Several levels of octs in the reference layout: levels stored
    contiguously from coarse to fine (head(l), noct(l)), Morton order inside a
    level, refined(ind) set where the child oct exists (an oct of level l+1
    with key K is the child of the level-l cell with coordinates K)."""

    def __init__(self, ndim, level_keys, boxlen=1.0, box=None):
        """box = (lev, ckey_min[3], ckey_max[3]): a computational box smaller
        than the full levels (the reference's box_ckey_min/max, given at level
        lev and scaled by 2 per level); None for the full levels."""
        assert ndim == 3, "AmrMesh builds 3D meshes"
        self.ndim = ndim
        self.boxlen = boxlen
        self.box = box
        self.ttd = 8
        self.levels = sorted(level_keys)
        self.levelmin, self.levelmax = self.levels[0], self.levels[-1]
        ck, lv = [], []
        self.head, self.noct = {}, {}
        pos = 0
        for l in self.levels:
            k = np.asarray(level_keys[l], dtype=np.int64).reshape(-1, 3)
            order = np.argsort(morton3(k[:, 0], k[:, 1], k[:, 2]), kind="stable")
            k = k[order]
            self.head[l] = pos + 1
            self.noct[l] = len(k)
            pos += len(k)
            ck.append(k)
            lv.append(np.full(len(k), l))
        self.ckey = np.concatenate(ck).astype(np.int32)
        self.lev = np.concatenate(lv).astype(np.int32)
        self.ngrid = pos
        # per level: encoded keys in sorted order and the matching oct indices
        self._sorted = {}
        for l in self.levels:
            i0 = self.head[l] - 1
            enc = self._encode(self.ckey[i0:i0 + self.noct[l]].astype(np.int64))
            order = np.argsort(enc, kind="stable")
            self._sorted[l] = (enc[order], order + i0)
        self.refined = np.zeros((pos, 8), dtype=np.uint8)
        for l in self.levels[1:]:
            if l - 1 not in self.head:
                continue
            i0 = self.head[l] - 1
            par, ind = self.parent_cells(l, self.ckey[i0:i0 + self.noct[l]].astype(np.int64))
            self.refined[par, ind] = 1

    @staticmethod
    def _encode(k):
        return k[:, 0] | (k[:, 1] << 21) | (k[:, 2] << 42)

    def find(self, l, keys):
        """0-based indices of the octs of level l with the given keys (-1 if absent)."""
        enc, idx = self._sorted[l]
        q = self._encode(np.asarray(keys, dtype=np.int64).reshape(-1, 3))
        pos = np.searchsorted(enc, q)
        pos = np.minimum(pos, len(enc) - 1)
        hit = enc[pos] == q
        return np.where(hit, idx[pos], -1)

    def parent_cells(self, l, keys):
        """(parent oct index 0-based, cell 0..7) at level l-1 of level-l octs with these keys."""
        k = np.asarray(keys, dtype=np.int64).reshape(-1, 3)
        par = self.find(l - 1, k >> 1)
        if np.any(par < 0):
            raise ValueError("level %d octs without a parent cell at level %d" % (l, l - 1))
        ind = (k[:, 0] & 1) + 2 * (k[:, 1] & 1) + 4 * (k[:, 2] & 1)
        return par, ind

    def dx(self, l):
        return self.boxlen / 2 ** l

    def level_params(self):
        """ckey_max, key_off, box_ckey_min/max per level 1..levelmax+1, the
        box scaled from its level by powers of two."""
        ckey_max, key_off, box_min, box_max = level_params(self.levelmax, self.ndim)
        if self.box is not None:
            bl, bmin, bmax = self.box
            for l in range(1, len(ckey_max) + 1):
                sh = l - bl
                for d in range(self.ndim):
                    box_min[l - 1, d] = bmin[d] << sh if sh >= 0 else bmin[d] >> -sh
                    box_max[l - 1, d] = bmax[d] << sh if sh >= 0 else max(1, bmax[d] >> -sh)
        return ckey_max, key_off, box_min, box_max

    def grid_records(self):
        g = np.zeros(self.ngrid, dtype=OCT3)
        g["ckey"] = self.ckey
        g["lev"] = self.lev
        g["refined"] = self.refined
        return g

    def heads(self):
        """head/noct arrays indexed by level-1 for levels 1..levelmax."""
        head = np.zeros(self.levelmax, dtype=np.int32)
        noct = np.zeros(self.levelmax, dtype=np.int32)
        for l in self.levels:
            head[l - 1] = self.head[l]
            noct[l - 1] = self.noct[l]
        return head, noct

    def cell_centres(self, l):
        """(noct(l), 8, 3) cell-centre coordinates of level l."""
        i0 = self.head[l] - 1
        k = self.ckey[i0:i0 + self.noct[l]].astype(np.float64)
        off = np.array([[(ind & 1), (ind >> 1) & 1, (ind >> 2) & 1] for ind in range(8)], dtype=np.float64)
        return (2 * k[:, None, :] + off[None, :, :] + 0.5) * self.dx(l)
