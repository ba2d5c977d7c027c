"""The device index tables against a brute-force (level, key) dictionary.

What "neighbour and oct indexing bit-exact" means here: for every oct of a
level and each of its 3^ndim stencil slots (slot 1+i+3j+9k,
gpu/gpu_refine.cuf:2006), the device neighbour table holds
- the index of the oct with that key (periodic directions wrapped as at
  hydro/godunov_fine.f90:463-468) when the level has one;
- for a key outside the box in one dimension, -(index of the oct one key back
  inside) when that oct is one of the level's own (the reference's boundary
  oct built by init_bound_refine from it, amr/nbors_utils.f90:580-640);
- otherwise a ghost oct (index above ngridmax) carrying that key: the
  reference interpolates those cells from the parent level (get_parent_cell,
  interpol_hydro), and ghost slot numbers are the implementation's own, so a
  ghost is compared by its key, never by its index.
The father table must give get_parent_cell (amr/nbors_utils.f90:222-249):
the parent oct of key k>>1 and cell 1 + sum_d 2^d (k_d & 1); each ghost's
parent cell likewise (floor division: a boundary oct of the parent level,
coded -(reference oct), for a key beside a non-periodic face).  The brick
windows must hold the same resolution for every slot of every brick.
"""
import numpy as np
import pytest

from ramses2_amd import HydroParams, UniformLevel, UniformMesh, ic
from ramses2_amd.autotune import brick_offset
from ramses2_amd.hydro import AmrHierarchy
from tests.amr_cases import low_face_mesh, smooth_state, three_level_mesh, two_level_mesh

pytestmark = pytest.mark.gpu


class Oracle:
    """Brute-force resolution of (level, key) over a mesh's key lists."""

    def __init__(self, ndim, keys_of_level, base_of_level, box_min, box_max, periodic, ngridmax):
        self.ndim = ndim
        self.real = {}
        for l, keys in keys_of_level.items():
            for i, k in enumerate(keys):
                self.real[(l, tuple(int(x) for x in k[:ndim]))] = base_of_level[l] + i
        self.box_min, self.box_max = box_min, box_max
        self.periodic = periodic
        self.ngridmax = ngridmax

    def wrap(self, l, k):
        k = list(k)
        for d in range(self.ndim):
            lo, hi = int(self.box_min[l - 1][d]), int(self.box_max[l - 1][d])
            if self.periodic[d]:
                k[d] = lo + (k[d] - lo) % (hi - lo)
        return tuple(k)

    def out_dims(self, l, k):
        return [d for d in range(self.ndim)
                if k[d] < self.box_min[l - 1][d] or k[d] >= self.box_max[l - 1][d]]

    def resolve(self, l, k):
        """('oct', index) | ('bnd', reference index) | ('ghost', key) | ('none', None)"""
        k = self.wrap(l, k)
        out = self.out_dims(l, k)
        if not out:
            if (l, k) in self.real:
                return "oct", self.real[(l, k)]
            return "ghost", k
        if len(out) > 1:
            return "none", None
        d = out[0]
        kr = list(k)
        kr[d] += 1 if k[d] < self.box_min[l - 1][d] else -1
        if (l, tuple(kr)) in self.real:
            return "bnd", self.real[(l, tuple(kr))]
        return "ghost", k


def check_value(kind, ref, v, ghost_key):
    if kind == "oct":
        assert v == ref, (kind, ref, v)
    elif kind == "bnd":
        assert v == -ref, (kind, ref, v)
    elif kind == "ghost":
        assert v in ghost_key, (kind, ref, v)
        assert ghost_key[v] == ref, (kind, ref, v, ghost_key[v])
    else:
        assert v == 0, (kind, v)


def check_level(hh, orc, l, keys, head, levelmin, bricks=True):
    ndim = orc.ndim
    nslot = 3 ** ndim
    g = hh.level_table(l, 4).reshape(-1, 4)
    ghost_key = {int(r[0]): tuple(int(x) for x in r[1:1 + ndim]) for r in g}
    assert all(idx > orc.ngridmax for idx in ghost_key)
    assert len(set(ghost_key.values())) == len(ghost_key), "two ghosts with one key"
    nb = hh.level_table(l, 0).reshape(len(keys), nslot)
    made = set()  # keys of the ghosts the level's stencils need (the library makes exactly these)
    for o, k in enumerate(keys):
        for s in range(nslot):
            off = (s % 3 - 1, (s // 3) % 3 - 1, s // 9 - 1)[:ndim]
            kind, ref = orc.resolve(l, tuple(int(k[d]) + off[d] for d in range(ndim)))
            check_value(kind, ref, int(nb[o, s]), ghost_key)
            if kind == "ghost":
                made.add(ref)
    assert made == set(ghost_key.values())
    # every ghost is referenced by some slot (none is made in vain)
    used = set(int(v) for v in nb.ravel() if v > orc.ngridmax)
    assert used == set(ghost_key), (len(used), len(ghost_key))
    if l > levelmin:
        fa = hh.level_table(l, 1).reshape(len(keys), 2)
        for o, k in enumerate(keys):
            kind, ref = orc.resolve(l - 1, tuple(int(k[d]) >> 1 for d in range(ndim)))
            assert kind == "oct" and fa[o, 0] == ref, (k, kind, ref, fa[o])
            assert fa[o, 1] == 1 + sum((int(k[d]) & 1) << d for d in range(ndim)), (k, fa[o])
        gp = hh.level_table(l, 5).reshape(-1, 2)
        for (idx, gk), (code, cell) in zip(sorted(ghost_key.items()), gp):
            kind, ref = orc.resolve(l - 1, tuple(x >> 1 for x in gk))
            assert kind in ("oct", "bnd"), (gk, kind)
            check_value(kind, ref, int(code), {})
            assert cell == 1 + sum((gk[d] & 1) << d for d in range(ndim)), (gk, cell)
    if bricks and ndim == 3:
        bx, by, bz = hh.brick_shape()
        org = hh.level_table(l, 2).reshape(-1, 3)
        # one lattice per level, at the offset the library's work model picks,
        # with every origin inside the box and one brick per occupied cell of it
        bmin = np.array([int(x) for x in orc.box_min[l - 1][:3]])
        off, _ = brick_offset(keys, (bx, by, bz), bmin, [int(x) for x in orc.box_max[l - 1][:3]])
        assert np.all(org >= bmin)
        assert np.all((org - bmin - np.array(off)) % np.array([bx, by, bz]) == 0), (off, org[:4])
        cells = (np.asarray(keys, dtype=np.int64)[:, :3] - bmin - np.array(off)) // np.array([bx, by, bz])
        assert len(org) == len(np.unique(cells, axis=0))
        win = hh.level_table(l, 3).reshape(len(org), bz + 2, by + 2, bx + 2)
        for b, o in enumerate(org):
            for z in range(bz + 2):
                for y in range(by + 2):
                    for x in range(bx + 2):
                        kind, ref = orc.resolve(l, (int(o[0]) + x - 1, int(o[1]) + y - 1, int(o[2]) + z - 1))
                        if kind == "ghost" and ref not in made:
                            kind, ref = "none", None  # a window slot no stencil of the level reaches
                        check_value(kind, ref, int(win[b, z, y, x]), ghost_key)


def amr_oracle(am, gpu, periodic):
    ckey_max, key_off, box_min, box_max = am.level_params()
    keys = {l: am.ckey[am.head[l] - 1:am.head[l] - 1 + am.noct[l]] for l in am.levels}
    base = {l: am.head[l] for l in am.levels}
    return Oracle(3, keys, base, box_min, box_max, periodic, gpu.hh.ngrid), keys


@pytest.mark.parametrize("case", ["two", "three", "wrapped", "low_face"])
def test_amr_tables_match_brute_force(case):
    periodic = (1, 1, 1)
    btype = (2,) * 6
    if case == "two":
        am = two_level_mesh()
    elif case == "three":
        am = three_level_mesh()
    elif case == "wrapped":
        am = two_level_mesh(region=(3, 10), offset=(13, 2, 9))
    else:
        am = low_face_mesh(0)
        periodic = (0, 1, 1)
    print("case", case, flush=True)
    u = smooth_state(am, 1)
    p = HydroParams(gamma=1.4, courant_factor=0.5, slope_type=2, riemann="hllc")
    gpu = AmrHierarchy(am, u, p, periodic=periodic, bound_type=btype)
    orc, keys = amr_oracle(am, gpu, periodic)
    for l in am.levels:
        # the bricks are rebuilt per level with the library's shape; every
        # level was built with the same one
        check_level(gpu.hh, orc, l, keys[l], am.head[l], am.levelmin)


def test_uniform_tables_match_brute_force():
    mesh = UniformMesh(3, 4)
    lev = UniformLevel(mesh, ic.sedov_3d(16), HydroParams())
    ckey_max, key_off, box_min, box_max = mesh.level_params()
    orc = Oracle(3, {4: mesh.ckey}, {4: 1}, box_min, box_max, (1, 1, 1), mesh.noct)
    check_level(lev.hh, orc, 4, mesh.ckey, 1, 4)


def test_sod_tables_match_brute_force():
    """1D tube, zero-gradient ends: the end octs' outer slots are boundary
    octs coded -(end oct)."""
    mesh = UniformMesh(1, 10)
    lev = UniformLevel(mesh, ic.sod_1d(1024), HydroParams(), periodic=False)
    ckey_max, key_off, box_min, box_max = mesh.level_params()
    orc = Oracle(1, {10: mesh.ckey}, {10: 1}, box_min, box_max, (0, 1, 1), mesh.noct)
    check_level(lev.hh, orc, 10, mesh.ckey, 1, 10, bricks=False)
