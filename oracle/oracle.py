"""ctypes wrapper of the CPU oracle (gvx_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product path never uses it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libgvx_oracle.so")

BITVECTOR_MAP = 0
PROB_MAP = 4
DIST_MAP = 9

_lib = None


def build():
    """Compile the oracle with its Makefile (gcc)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.c_void_p
        u32, u64, f32, i32 = ctypes.c_uint32, ctypes.c_uint64, ctypes.c_float, ctypes.c_int
        sig = {
            "gvo_quantize": (u32, [f32, f32]),
            "gvo_sincosf": (None, [f32, P, P]),
            "gvo_dh_matrix": (None, [f32, f32, f32, f32, i32, f32, P]),
            "gvo_affine_mul": (None, [P, P, P]),
            "gvo_transform_points": (None, [P, P, u64, P]),
            "gvo_float_to_probability": (ctypes.c_int8, [f32]),
            "gvo_prob_update": (ctypes.c_int8, [ctypes.c_int8, ctypes.c_int8]),
            "gvo_prob_insert": (ctypes.c_int8, [ctypes.c_int8, u32]),
            "gvo_bitmargin_collision": (i32, [P, P, P, u32, u32]),
            "gvo_dist_uninit": (u32, [P]),
            "gvo_map_clear": (None, [i32, P, P]),
            "gvo_map_insert_points": (None, [i32, P, P, f32, P, u64, u32, P]),
            "gvo_map_insert_points_mt": (None, [i32, P, P, f32, P, u64, u32, P, i32]),
            "gvo_map_insert_coords": (None, [i32, P, P, P, u64, u32, P]),
            "gvo_map_insert_meta": (None, [i32, P, P, f32, P, P, u32, P, u32, P]),
            "gvo_map_insert_meta_selfcoll": (i32, [P, P, f32, P, P, u32, P, P, P, P]),
            "gvo_map_collide": (u64, [i32, P, i32, P, P, f32, P]),
            "gvo_map_collide_mt": (u64, [i32, P, i32, P, P, f32, P, i32]),
            "gvo_map_collide_types": (u64, [i32, P, i32, P, P, f32, P]),
            "gvo_map_clear_bits": (None, [P, P, P]),
            "gvo_ray_cells": (u64, [P, P, P]),
            "gvo_map_collide_sv": (u64, [i32, P, i32, P, P, f32, u32, u32, P]),
            "gvo_perform_left_shift": (None, [P, u32]),
            "gvo_shift_sv": (i32, [P, u64, u32]),
            "gvo_list_collide_bitcheck": (u64, [P, P, u64, P, P, u64, u32]),
            "gvo_list_collide_per_meaning": (u64, [P, P, u64, P, u64, P]),
            "gvo_prob_insert_sensor": (None, [P, P, f32, P, P, u64, i32, P, u32]),
            "gvo_list_insert_points": (None, [P, P, P, P, P, f32, P, u64, P, u32, P, u32]),
            "gvo_list_insert_coords": (None, [P, P, P, P, P, P, u64, u32]),
            "gvo_list_make_unique": (None, [P, P, P, P]),
            "gvo_list_append_list": (None, [P, P, P, P, P, P, P, P, u64, P, i32, u32]),
            "gvo_list_append_map": (u64, [P, P, P, P, P, P, P, i32, u32]),
            "gvo_list_collide_map": (u64, [P, P, u64, i32, P, P, f32, P, i32, P, P]),
            "gvo_list_collide_list": (u64, [P, P, u64, P, P, u64, P]),
            "gvo_dist_merge_occupied": (None, [P, P, P, P, f32]),
            "gvo_pba": (None, [P, P]),
            "gvo_pba_mt": (None, [P, P, ctypes.c_int]),
            "gvo_edt_bruteforce": (None, [P, P, P]),
            "gvo_exact_distances": (None, [P, P, f32, P, u64]),
            "gvo_dist_squared": (None, [P, P, P, u64, P]),
            "gvo_fk_chain": (None, [P, u32, P, P, P]),
            "gvo_batch_collide": (None, [P, P, f32, P, P, u32, P, P, P, u32, P, P]),
            "gvo_batch_collide_mt": (None, [P, P, f32, P, P, u32, P, P, P, u32, P, P, ctypes.c_int]),
            "gvo_trilinear_d": (ctypes.c_double, [P, P, ctypes.c_double, P, P]),
            "gvo_fk_chain_d": (None, [P, u32, P, P, P]),
            "gvo_cost_d": (None, [P, P, ctypes.c_double, P, u32, P, P, u32, P, u32, u32,
                                  ctypes.c_double, ctypes.c_double, P, P, P]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _dims(d):
    return np.ascontiguousarray(d, dtype=np.uint32)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


# ---------------------------------------------------------------- scalars
def quantize(p, side):
    return lib().gvo_quantize(float(p), float(side))


def sincosf(x):
    s = np.zeros(1, np.float32)
    c = np.zeros(1, np.float32)
    lib().gvo_sincosf(float(x), _p(s), _p(c))
    return s[0], c[0]


def dh_matrix(theta, d, a, alpha, prismatic, value):
    m = np.zeros(12, np.float32)
    lib().gvo_dh_matrix(theta, d, a, alpha, int(prismatic), value, _p(m))
    return m


def transform_points(m, pts):
    pts = _f32(pts).reshape(-1, 3)
    out = np.empty_like(pts)
    lib().gvo_transform_points(_p(_f32(m)), _p(pts), pts.shape[0], _p(out))
    return out


def float_to_probability(v):
    return lib().gvo_float_to_probability(float(v))


def bitmargin(v1, v2, margin=0, sv_offset=0):
    out = np.zeros(32, np.uint8)
    r = lib().gvo_bitmargin_collision(_p(np.ascontiguousarray(v1, np.uint8)),
                                      _p(np.ascontiguousarray(v2, np.uint8)), _p(out), margin, sv_offset)
    return bool(r), out


def dist_uninit(dims):
    return lib().gvo_dist_uninit(_p(_dims(dims)))


# ---------------------------------------------------------------- dense maps
def new_map(mtype, dims):
    dims = _dims(dims)
    n = int(np.prod(dims.astype(np.uint64)))
    if mtype == BITVECTOR_MAP:
        m = np.zeros((n, 32), np.uint8)
    elif mtype == PROB_MAP:
        m = np.zeros(n, np.int8)
    else:
        m = np.zeros(n, np.uint32)
    lib().gvo_map_clear(mtype, _p(m), _p(dims))
    return m


def map_insert_points(mtype, m, dims, side, pts, meaning, threads=1):
    """threads > 1: points bucketed by voxel range over OpenMP threads (same map)."""
    out = np.zeros(1, np.int32)
    pts = _f32(pts).reshape(-1, 3)
    if threads > 1:
        lib().gvo_map_insert_points_mt(mtype, _p(m), _p(_dims(dims)), side, _p(pts), pts.shape[0], meaning, _p(out),
                                       int(threads))
    else:
        lib().gvo_map_insert_points(mtype, _p(m), _p(_dims(dims)), side, _p(pts), pts.shape[0], meaning, _p(out))
    return bool(out[0])


def ray_cells(frm, to):
    """cells visited by the ray caster from voxel frm to voxel to, (k, 3) uint32"""
    f = np.ascontiguousarray(frm, np.uint32)
    t = np.ascontiguousarray(to, np.uint32)
    k = lib().gvo_ray_cells(_p(f), _p(t), None)
    out = np.zeros((k, 3), np.uint32)
    lib().gvo_ray_cells(_p(f), _p(t), _p(out))
    return out


def prob_insert_sensor(m, dims, side, sensor, pts, raycast=True, robot=None, robot_bit=0):
    """insertSensorData into an int8 ProbVoxelMap array m (in place)"""
    pts = _f32(pts).reshape(-1, 3)
    s = _f32(sensor).reshape(3)
    rb = None if robot is None else np.ascontiguousarray(robot, np.uint8)
    lib().gvo_prob_insert_sensor(_p(m), _p(_dims(dims)), side, _p(s), _p(pts), pts.shape[0], int(bool(raycast)),
                                 None if rb is None else _p(rb), robot_bit)


def map_insert_coords(mtype, m, dims, coords, meaning):
    out = np.zeros(1, np.int32)
    c = np.ascontiguousarray(coords, np.uint32).reshape(-1, 3)
    lib().gvo_map_insert_coords(mtype, _p(m), _p(_dims(dims)), _p(c), c.shape[0], meaning, _p(out))
    return bool(out[0])


def map_insert_meta(mtype, m, dims, side, pts, sizes, meanings=None, meaning=0):
    out = np.zeros(1, np.int32)
    pts = _f32(pts).reshape(-1, 3)
    sizes = np.ascontiguousarray(sizes, np.uint32)
    mn = None if meanings is None else np.ascontiguousarray(meanings, np.uint32)
    lib().gvo_map_insert_meta(mtype, _p(m), _p(_dims(dims)), side, _p(pts), _p(sizes), len(sizes), _p(mn),
                              meaning, _p(out))
    return bool(out[0])


def map_insert_meta_selfcoll(m, dims, side, pts, sizes, meanings=None, masks=None):
    """insertMetaPointCloudWithSelfCollisionCheck on a BitVector map: (self collision, colliding
    meanings as 32 bytes, points outside)"""
    out = np.zeros(1, np.int32)
    col = np.zeros(32, np.uint8)
    pts = _f32(pts).reshape(-1, 3)
    sizes = np.ascontiguousarray(sizes, np.uint32)
    mn = None if meanings is None else np.ascontiguousarray(meanings, np.uint32)
    mk = None if masks is None else np.ascontiguousarray(masks, np.uint8).reshape(-1, 32)
    r = lib().gvo_map_insert_meta_selfcoll(_p(m), _p(_dims(dims)), side, _p(pts), _p(sizes), len(sizes), _p(mn),
                                           _p(mk), _p(col), _p(out))
    return bool(r), col, bool(out[0])


def map_collide(ta, a, tb, b, dims, thr, offset=(0, 0, 0), threads=1):
    """threads > 1: voxels over OpenMP threads (same count and marks)."""
    off = np.ascontiguousarray(offset, np.int32)
    if threads > 1:
        return int(lib().gvo_map_collide_mt(ta, _p(a), tb, _p(b), _p(_dims(dims)), thr, _p(off), int(threads)))
    return int(lib().gvo_map_collide(ta, _p(a), tb, _p(b), _p(_dims(dims)), thr, _p(off)))


def map_collide_types(ta, a, tb, b, dims, thr):
    mask = np.zeros(32, np.uint8)
    c = lib().gvo_map_collide_types(ta, _p(a), tb, _p(b), _p(_dims(dims)), thr, _p(mask))
    return int(c), mask


def map_collide_sv(ta, a, tb, b, dims, thr, window=0, sv_offset=0):
    mask = np.zeros(32, np.uint8)
    c = lib().gvo_map_collide_sv(ta, _p(a), tb, _p(b), _p(_dims(dims)), thr, window, sv_offset, _p(mask))
    return int(c), mask


def perform_left_shift(v, shift):
    out = np.array(v, np.uint8).copy()
    lib().gvo_perform_left_shift(_p(out), shift)
    return out


def shift_sv(voxels, shift):
    """in place on an (n, 32) uint8 array; False (untouched) for shifts above 56"""
    return bool(lib().gvo_shift_sv(_p(voxels), voxels.size // 32, shift))


def map_clear_bits(m, dims, bits):
    lib().gvo_map_clear_bits(_p(m), _p(_dims(dims)), _p(np.ascontiguousarray(bits, np.uint8)))


# ---------------------------------------------------------------- voxel lists
class List:
    """Oracle BitVectorVoxelList: sorted, unique ids with coords and 256-bit payloads."""

    def __init__(self, dims, side):
        self.dims = _dims(dims)
        self.side = float(side)
        self.ids = np.zeros(0, np.uint32)
        self.coords = np.zeros((0, 3), np.uint32)
        self.bits = np.zeros((0, 32), np.uint8)

    def __len__(self):
        return self.ids.shape[0]

    def _grow(self, extra):
        n = len(self)
        cap = n + extra
        ids = np.zeros(cap, np.uint32)
        coords = np.zeros((cap, 3), np.uint32)
        bits = np.zeros((cap, 32), np.uint8)
        ids[:n] = self.ids
        coords[:n] = self.coords
        bits[:n] = self.bits
        return ids, coords, bits, np.array([n], np.uint64)

    def _set(self, ids, coords, bits, ln):
        n = int(ln[0])
        self.ids, self.coords, self.bits = ids[:n].copy(), coords[:n].copy(), bits[:n].copy()

    def insert_points(self, pts, meaning=0, sizes=None, meanings=None):
        pts = _f32(pts).reshape(-1, 3)
        ids, coords, bits, ln = self._grow(pts.shape[0])
        sz = None if sizes is None else np.ascontiguousarray(sizes, np.uint32)
        mn = None if meanings is None else np.ascontiguousarray(meanings, np.uint32)
        lib().gvo_list_insert_points(_p(ids), _p(coords), _p(bits), _p(ln), _p(self.dims), self.side, _p(pts),
                                     pts.shape[0], _p(sz), 0 if sz is None else len(sz), _p(mn), meaning)
        self._set(ids, coords, bits, ln)

    def insert_coords(self, c, meaning):
        c = np.ascontiguousarray(c, np.uint32).reshape(-1, 3)
        ids, coords, bits, ln = self._grow(c.shape[0])
        lib().gvo_list_insert_coords(_p(ids), _p(coords), _p(bits), _p(ln), _p(self.dims), _p(c), c.shape[0],
                                     meaning)
        self._set(ids, coords, bits, ln)

    def merge_list(self, other, offset=(0, 0, 0), meaning=None):
        ids, coords, bits, ln = self._grow(len(other))
        off = np.ascontiguousarray(offset, np.int32)
        lib().gvo_list_append_list(_p(ids), _p(coords), _p(bits), _p(ln), _p(self.dims), _p(other.ids),
                                   _p(other.coords), _p(other.bits), len(other), _p(off),
                                   0 if meaning is None else 1, 0 if meaning is None else meaning)
        self._set(ids, coords, bits, ln)

    def merge_map(self, bitmap, offset=(0, 0, 0), meaning=None):
        occ = int((~_none_but_empty(bitmap)).sum())
        ids, coords, bits, ln = self._grow(occ)
        off = np.ascontiguousarray(offset, np.int32)
        lib().gvo_list_append_map(_p(ids), _p(coords), _p(bits), _p(ln), _p(self.dims), _p(bitmap), _p(off),
                                  0 if meaning is None else 1, 0 if meaning is None else meaning)
        self._set(ids, coords, bits, ln)

    def collide_map(self, mtype, m, thr, offset=(0, 0, 0), mode=0, typemask=None):
        mask = np.zeros(32, np.uint8)
        tm = np.zeros(32, np.uint8) if typemask is None else np.ascontiguousarray(typemask, np.uint8)
        off = np.ascontiguousarray(offset, np.int32)
        c = lib().gvo_list_collide_map(_p(self.ids), _p(self.bits), len(self), mtype, _p(m), _p(self.dims), thr,
                                       _p(off), mode, _p(tm), _p(mask))
        return int(c), mask

    def collide_bitcheck(self, other, margin=0):
        return int(lib().gvo_list_collide_bitcheck(_p(self.ids), _p(self.bits), len(self), _p(other.ids),
                                                   _p(other.bits), len(other), margin))

    def collide_per_meaning(self, other):
        counts = np.zeros(256, np.uint64)
        total = lib().gvo_list_collide_per_meaning(_p(self.ids), _p(self.bits), len(self), _p(other.ids), len(other),
                                                   _p(counts))
        return int(total), counts

    def shift_sv(self, shift):
        return shift_sv(self.bits, shift)

    def collide_list(self, other):
        mask = np.zeros(32, np.uint8)
        c = lib().gvo_list_collide_list(_p(self.ids), _p(self.bits), len(self), _p(other.ids), _p(other.bits),
                                        len(other), _p(mask))
        return int(c), mask


def _none_but_empty(bitmap):
    b = bitmap.reshape(-1, 32)
    return ((b[:, 0] & 0xFE) == 0) & (b[:, 1:] == 0).all(axis=1)


def occupancy_bits(bitmap):
    """1-bit-per-voxel occupancy (BitVoxel::isOccupied) packed little-endian."""
    occ = ~_none_but_empty(bitmap)
    return np.packbits(occ, bitorder="little")


# ---------------------------------------------------------------- distance maps
def dist_merge_occupied(dist, prob, dims, offset, thr):
    lib().gvo_dist_merge_occupied(_p(dist), _p(prob), _p(_dims(dims)),
                                  _p(np.ascontiguousarray(offset, np.uint32)), thr)


def pba(dist, dims, threads=1):
    """threads > 1: the lines of each pass over OpenMP threads (same result)."""
    if threads > 1:
        lib().gvo_pba_mt(_p(dist), _p(_dims(dims)), int(threads))
    else:
        lib().gvo_pba(_p(dist), _p(_dims(dims)))


def edt_bruteforce(dist, dims):
    out = np.empty_like(dist)
    lib().gvo_edt_bruteforce(_p(dist), _p(_dims(dims)), _p(out))
    return out


def exact_distances(dist, dims, side, obstacles):
    ob = _f32(obstacles).reshape(-1, 3)
    lib().gvo_exact_distances(_p(dist), _p(_dims(dims)), side, _p(ob), ob.shape[0])


def dist_squared(dist, dims, idx=None):
    dims = _dims(dims)
    if idx is None:
        idx = np.arange(dist.shape[0], dtype=np.uint32)
    idx = np.ascontiguousarray(idx, np.uint32)
    out = np.empty(idx.shape[0], np.int32)
    lib().gvo_dist_squared(_p(dist), _p(dims), _p(idx), idx.shape[0], _p(out))
    return out


# ---------------------------------------------------------------- robot
def fk_chain(dh, base, q):
    dh = _f32(dh).reshape(-1, 5)
    out = np.zeros((dh.shape[0], 12), np.float32)
    lib().gvo_fk_chain(_p(dh), dh.shape[0], _p(_f32(base)), _p(_f32(q)), _p(out))
    return out


def batch_collide(env_bits, dims, side, link_pts, link_sizes, dh, base, q, threads=1):
    """threads > 1: the OpenMP variant (same results; the all-cores CPU baseline)."""
    q = _f32(q)
    sizes = np.ascontiguousarray(link_sizes, np.uint32)
    dh = _f32(dh).reshape(-1, 5)
    npose = q.reshape(-1, dh.shape[0]).shape[0]
    counts = np.zeros(npose, np.uint32)
    masks = np.zeros((npose, 32), np.uint8)
    args = (_p(np.ascontiguousarray(env_bits, np.uint8)), _p(_dims(dims)), side, _p(_f32(link_pts)), _p(sizes),
            len(sizes), _p(dh), _p(_f32(base)), _p(q), npose, _p(counts), _p(masks))
    if threads > 1:
        lib().gvo_batch_collide_mt(*args, int(threads))
    else:
        lib().gvo_batch_collide(*args)
    return counts, masks


def fk_chain_d(dh, base, q):
    dh = np.ascontiguousarray(dh, np.float64).reshape(-1, 5)
    out = np.zeros((dh.shape[0], 12), np.float64)
    lib().gvo_fk_chain_d(_p(dh), dh.shape[0], _p(np.ascontiguousarray(base, np.float64)),
                         _p(np.ascontiguousarray(q, np.float64)), _p(out))
    return out


# ---------------------------------------------------------------- cost
def cost_d(grid, dims, side, dh, base, spheres, q, eps, w, want_grad=True):
    dh = np.ascontiguousarray(dh, np.float64).reshape(-1, 5)
    nl = dh.shape[0]
    sph = np.ascontiguousarray(spheres, np.float64).reshape(nl, -1, 4)
    spl = sph.shape[1]
    q = np.ascontiguousarray(q, np.float64)
    ntraj, nwp = q.shape[0], q.shape[1]
    cost = np.zeros(ntraj, np.float64)
    dq = np.zeros((ntraj, nwp, nl), np.float64) if want_grad else None
    dx = np.zeros((ntraj, nwp, nl, spl, 3), np.float64) if want_grad else None
    lib().gvo_cost_d(_p(_f32(grid)), _p(_dims(dims)), float(side), _p(dh), nl,
                     _p(np.ascontiguousarray(base, np.float64)), _p(sph), spl, _p(q), ntraj, nwp, float(eps),
                     float(w), _p(cost), _p(dq), _p(dx))
    return cost, dq, dx


def trilinear_d(grid, dims, side, x):
    g = np.zeros(3, np.float64)
    v = lib().gvo_trilinear_d(_p(_f32(grid)), _p(_dims(dims)), float(side),
                              _p(np.ascontiguousarray(x, np.float64)), _p(g))
    return v, g
