"""ctypes binding of libgvx.so, the MI355X-native GPU-Voxels core.

Thin Python classes over the C ABI declared in include/gvx.h.  The HIP
library is required: importing works without it, but every call raises
GvxError when libgvx.so cannot be loaded (there is no CPU fallback).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(os.path.dirname(_HERE))          # gpu-voxels_amd/
LIB_DIR = os.path.join(PKG, "lib")
LIB_PATH = os.environ.get("GVX_LIBRARY") or os.path.join(LIB_DIR, "libgvx.so")  # GVX_LIBRARY: another build

BITVECTOR_VOXELMAP = 0
BITVECTOR_VOXELLIST = 1
PROBAB_VOXELMAP = 4
DISTANCE_VOXELMAP = 9
MEM_HOST = 0
MEM_DEVICE = 1

eBVM_FREE, eBVM_OCCUPIED, eBVM_COLLISION, eBVM_UNKNOWN = 0, 1, 2, 3
eBVM_SWEPT_VOLUME_START, eBVM_SWEPT_VOLUME_END, eBVM_UNDEFINED = 4, 254, 255


class GvxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"gvx error {code}: {msg}")
        self.code = code


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise GvxError(-1, f"{LIB_PATH} not built (run gpu-voxels_amd/build.py)")
        # One HIP runtime per process: when PyTorch is present its bundled
        # libamdhip64.so.7 is loaded first, and libgvx.so binds to it by soname
        # instead of pulling a second runtime in from /opt/rocm.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        i32, u32, u64, f32 = ctypes.c_int, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_float
        sig = {
            "gvx_init": [i32], "gvx_last_error": [], "gvx_get_stream": [P],
            "gvx_profile_enable": [i32], "gvx_profile_get": [ctypes.c_char_p, P, P], "gvx_profile_reset": [],
            "gvx_map_create": [i32, u32, u32, u32, f32, P], "gvx_map_destroy": [P],
            "gvx_map_info": [P, P, P, P, P], "gvx_map_device_ptr": [P, P], "gvx_map_clear": [P],
            "gvx_map_clear_bits": [P, P], "gvx_map_read": [P, P, i32], "gvx_map_write": [P, P, i32],
            "gvx_map_insert_points": [P, P, u64, u32, i32, P],
            "gvx_map_insert_coords": [P, P, u64, u32, i32, P],
            "gvx_map_insert_meta": [P, P, P, u32, P, u32, i32, P],
            "gvx_map_insert_meta_selfcoll": [P, P, P, u32, P, P, i32, P, P, P],
            "gvx_map_collide": [P, P, f32, P, P], "gvx_map_collide_types": [P, P, f32, P, P],
            "gvx_dist_merge_occupied": [P, P, P, f32], "gvx_dist_pba": [P], "gvx_dist_pba_noalloc": [P],
            "gvx_dist_exact": [P, P, u64, i32], "gvx_dist_squared": [P, P, u64, P, i32],
            "gvx_dist_metric": [P, P, i32],
            "gvx_list_create": [u32, u32, u32, f32, P], "gvx_list_destroy": [P], "gvx_list_size": [P, P],
            "gvx_list_clear": [P], "gvx_list_read": [P, P, P, P, i32], "gvx_list_device_ptrs": [P, P, P, P],
            "gvx_list_insert_points": [P, P, u64, P, u32, P, u32, i32],
            "gvx_list_insert_coords": [P, P, u64, u32, i32],
            "gvx_list_merge_list": [P, P, P, i32], "gvx_list_merge_map": [P, P, P, i32],
            "gvx_list_collide_map": [P, P, f32, P, i32, P, P, P], "gvx_list_collide_list": [P, P, P, P],
            "gvx_dh_matrix": [P, f32, P],
            "gvx_robot_transform": [P, P, u32, P, P, P, i32, P, i32],
            "gvx_map_insert_robot": [P, P, P, u32, P, P, P, P, u32, i32, P],
            "gvx_map_insert_robot_selfcoll": [P, P, P, u32, P, P, P, P, P, i32, P, P, P],
            "gvx_list_insert_robot": [P, P, P, u32, P, P, P, P, u32, i32],
            "gvx_batch_create": [P, f32, P, P, u32, P, P, u32, P], "gvx_batch_destroy": [P],
            "gvx_batch_refresh_env": [P], "gvx_batch_collide": [P, P, u32, i32, P, P, i32],
            "gvx_batch_autotune": [P, u32, P], "gvx_batch_set_schedule": [P, i32],
            "gvx_batch_collide_per_pose": [P, P, P, u32, i32, P],
            "gvx_batch_collide_async": [P, P, u32, P, P], "gvx_sync": [],
            "gvx_batch_last_stats": [P, P, P, P],
            "gvx_cost_create": [P, P, f32, P, u32, P, P, u32, f32, f32, P], "gvx_cost_destroy": [P],
            "gvx_cost_forward": [P, P, u32, u32, P], "gvx_cost_backward": [P, P, u32, u32, P, P, P],
            "gvx_list_info": [P, P, P],
            "gvx_map_insert_sensor": [P, P, u64, P, i32, P, u32, i32],
            "gvx_map_collide_sv": [P, P, f32, u32, u32, P, P], "gvx_map_shift_sv": [P, u32],
            "gvx_list_shift_sv": [P, u32], "gvx_list_collide_bitcheck": [P, P, u32, P],
            "gvx_list_collide_per_meaning": [P, P, P, P],
            "gvx_map_write_file": [P, ctypes.c_char_p], "gvx_map_read_file": [P, ctypes.c_char_p, P],
            "gvx_list_write_file": [P, ctypes.c_char_p], "gvx_list_read_file": [P, ctypes.c_char_p],
            "gvx_load_point_cloud": [ctypes.c_char_p, i32, i32, P, f32, P, P], "gvx_free": [P],
        }
        for name, args in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = {"gvx_last_error": ctypes.c_char_p, "gvx_free": None}.get(name, ctypes.c_int)
        _lib = L
    return _lib


def check(code):
    if code != 0:
        raise GvxError(code, lib().gvx_last_error().decode(errors="replace"))


def _p(a):
    if a is None:
        return None
    if isinstance(a, int):
        return ctypes.c_void_p(a)
    if hasattr(a, "data_ptr"):           # torch tensor (device or host)
        return ctypes.c_void_p(a.data_ptr())
    return a.ctypes.data_as(ctypes.c_void_p)


def _mem(a):
    if hasattr(a, "is_cuda") and a.is_cuda:
        return MEM_DEVICE
    return MEM_HOST


def _host(a, dt):
    if hasattr(a, "data_ptr"):
        return a
    return np.ascontiguousarray(a, dtype=dt)


def init(device=0):
    check(lib().gvx_init(device))


def stream_handle():
    s = ctypes.c_void_p()
    check(lib().gvx_get_stream(ctypes.byref(s)))
    return s.value


def sync():
    """Wait for all work enqueued on the library stream."""
    check(lib().gvx_sync())


def profile_enable(on=True):
    check(lib().gvx_profile_enable(1 if on else 0))


def profile_get(name):
    avg = ctypes.c_double()
    n = ctypes.c_uint64()
    check(lib().gvx_profile_get(name.encode(), ctypes.byref(avg), ctypes.byref(n)))
    return avg.value, n.value


def profile_reset():
    check(lib().gvx_profile_reset())


def _i3(v):
    return None if v is None else np.ascontiguousarray(v, dtype=np.int32)


class Map:
    """BitVectorVoxelMap / ProbVoxelMap / DistanceVoxelMap."""

    def __init__(self, mtype, dims, side):
        h = ctypes.c_void_p()
        check(lib().gvx_map_create(mtype, int(dims[0]), int(dims[1]), int(dims[2]), float(side), ctypes.byref(h)))
        self.h = h
        self.type = mtype
        self.dims = tuple(int(d) for d in dims)
        self.side = float(side)
        self.n = self.dims[0] * self.dims[1] * self.dims[2]

    def close(self):
        if self.h:
            lib().gvx_map_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def dtype(self):
        return {BITVECTOR_VOXELMAP: np.uint8, PROBAB_VOXELMAP: np.int8, DISTANCE_VOXELMAP: np.uint32}[self.type]

    def device_ptr(self):
        p = ctypes.c_void_p()
        check(lib().gvx_map_device_ptr(self.h, ctypes.byref(p)))
        return p.value

    def clear(self):
        check(lib().gvx_map_clear(self.h))

    def clear_bits(self, bits):
        check(lib().gvx_map_clear_bits(self.h, _p(np.ascontiguousarray(bits, np.uint8))))

    def read(self):
        if self.type == BITVECTOR_VOXELMAP:
            out = np.empty((self.n, 32), np.uint8)
        else:
            out = np.empty(self.n, self.dtype)
        check(lib().gvx_map_read(self.h, _p(out), MEM_HOST))
        return out

    @property
    def nbytes(self):
        return self.n * (32 if self.type == BITVECTOR_VOXELMAP else np.dtype(self.dtype).itemsize)

    def write(self, arr):
        """Overwrite every voxel from a host array or a contiguous torch tensor (host or device) of
        the map's voxel bytes."""
        if hasattr(arr, "data_ptr"):
            if not arr.is_contiguous() or arr.numel() * arr.element_size() != self.nbytes:
                raise GvxError(f"write: need a contiguous tensor of {self.nbytes} bytes")
            if arr.is_cuda:
                import torch
                torch.cuda.current_stream().synchronize()  # the library works on its own stream
            check(lib().gvx_map_write(self.h, _p(arr), _mem(arr)))
            return
        arr = np.ascontiguousarray(arr)
        if arr.nbytes != self.nbytes:
            raise GvxError(f"write: need {self.nbytes} bytes, got {arr.nbytes}")
        check(lib().gvx_map_write(self.h, _p(arr), MEM_HOST))

    def collide_sv(self, other, threshold=1.0, window=0, sv_offset=0):
        """collisionCheckBitvector with SVCollider(threshold, window): (count, 32-byte mask)"""
        c = ctypes.c_uint64()
        mask = np.zeros(32, np.uint8)
        check(lib().gvx_map_collide_sv(self.h, other.h, float(threshold), int(window), int(sv_offset), ctypes.byref(c),
                                       _p(mask)))
        return c.value, mask

    def shift_sv(self, shift):
        """shiftLeftSweptVolumeIDs"""
        check(lib().gvx_map_shift_sv(self.h, int(shift)))

    def insert_sensor_data(self, pts, sensor_pose, raycast=True, robot=None, robot_meaning=0):
        """ProbVoxelMap::insertSensorData; robot: a BitVector Map of the same dims for cut_real_robot"""
        pts = _host(pts, np.float32)
        n = pts.shape[0] if not hasattr(pts, "numel") else pts.numel() // 3
        pose = np.ascontiguousarray(sensor_pose, np.float32).reshape(3)
        if robot is not None:
            # the kernel reads 32 bytes per voxel of this map at the point voxels' ids
            if not isinstance(robot, Map) or robot.type != BITVECTOR_VOXELMAP or robot.dims != self.dims:
                raise ValueError("insert_sensor_data: robot must be a BitVectorVoxelMap with the same dimensions")
        rb = ctypes.c_void_p(robot.device_ptr()) if robot is not None else None
        check(lib().gvx_map_insert_sensor(self.h, _p(pts), n, _p(pose), int(bool(raycast)), rb, int(robot_meaning),
                                          _mem(pts)))

    def write_to_disk(self, path):
        """TemplateVoxelMap::writeToDisk"""
        check(lib().gvx_map_write_file(self.h, os.fsencode(path)))

    def read_from_disk(self, path):
        """TemplateVoxelMap::readFromDisk; returns whether the file's side length matched"""
        ok = ctypes.c_int(0)
        check(lib().gvx_map_read_file(self.h, os.fsencode(path), ctypes.byref(ok)))
        return bool(ok.value)

    def insert_points(self, pts, meaning):
        pts = _host(pts, np.float32)
        n = pts.shape[0] if not hasattr(pts, "numel") else pts.numel() // 3
        out = ctypes.c_int(0)
        check(lib().gvx_map_insert_points(self.h, _p(pts), n, meaning, _mem(pts), ctypes.byref(out)))
        return bool(out.value)

    def insert_coords(self, coords, meaning):
        c = _host(coords, np.uint32)
        n = c.shape[0] if not hasattr(c, "numel") else c.numel() // 3
        out = ctypes.c_int(0)
        check(lib().gvx_map_insert_coords(self.h, _p(c), n, meaning, _mem(c), ctypes.byref(out)))
        return bool(out.value)

    def insert_meta(self, pts, sizes, meanings=None, meaning=0):
        pts = _host(pts, np.float32)
        sizes = np.ascontiguousarray(sizes, np.uint32)
        mn = None if meanings is None else np.ascontiguousarray(meanings, np.uint32)
        out = ctypes.c_int(0)
        check(lib().gvx_map_insert_meta(self.h, _p(pts), _p(sizes), len(sizes), _p(mn), meaning, _mem(pts),
                                        ctypes.byref(out)))
        return bool(out.value)

    def insert_meta_selfcoll(self, pts, sizes, meanings=None, masks=None):
        """insertMetaPointCloudWithSelfCollisionCheck: (self collision, colliding meanings as
        32 bytes, points outside)"""
        pts = _host(pts, np.float32)
        sizes = np.ascontiguousarray(sizes, np.uint32)
        mn = None if meanings is None else np.ascontiguousarray(meanings, np.uint32)
        mk = None if masks is None else np.ascontiguousarray(masks, np.uint8).reshape(-1, 32)
        col = np.zeros(32, np.uint8)
        sc, out = ctypes.c_int(0), ctypes.c_int(0)
        check(lib().gvx_map_insert_meta_selfcoll(self.h, _p(pts), _p(sizes), len(sizes), _p(mn), _p(mk), _mem(pts),
                                                 _p(col), ctypes.byref(sc), ctypes.byref(out)))
        return bool(sc.value), col, bool(out.value)

    def collide(self, other, threshold=1.0, offset=None):
        c = ctypes.c_uint64()
        check(lib().gvx_map_collide(self.h, other.h, float(threshold), _p(_i3(offset)), ctypes.byref(c)))
        return c.value

    def collide_types(self, other, threshold=1.0):
        c = ctypes.c_uint64()
        mask = np.zeros(32, np.uint8)
        check(lib().gvx_map_collide_types(self.h, other.h, float(threshold), ctypes.byref(c), _p(mask)))
        return c.value, mask

    # distance maps
    def merge_occupied(self, prob, offset=(0, 0, 0), threshold=0.5):
        check(lib().gvx_dist_merge_occupied(self.h, prob.h, _p(np.ascontiguousarray(offset, np.uint32)),
                                            float(threshold)))

    def pba(self):
        check(lib().gvx_dist_pba(self.h))

    def pba_noalloc(self):
        """The same EDT in the library's fixed 32 MiB workspace (no device allocation)."""
        check(lib().gvx_dist_pba_noalloc(self.h))

    def exact_distances(self, obstacles):
        ob = _host(obstacles, np.float32)
        n = ob.shape[0] if not hasattr(ob, "numel") else ob.numel() // 3
        check(lib().gvx_dist_exact(self.h, _p(ob), n, _mem(ob)))

    def squared_distances(self, idx=None):
        if idx is None:
            out = np.empty(self.n, np.int32)
            check(lib().gvx_dist_squared(self.h, None, self.n, _p(out), MEM_HOST))
            return out
        idx = np.ascontiguousarray(idx, np.uint32)
        out = np.empty(idx.shape[0], np.int32)
        check(lib().gvx_dist_squared(self.h, _p(idx), idx.shape[0], _p(out), MEM_HOST))
        return out

    def metric(self, out=None):
        """float32 metric distance per voxel; `out` may be a device tensor."""
        if out is None:
            out = np.empty(self.n, np.float32)
        check(lib().gvx_dist_metric(self.h, _p(out), _mem(out)))
        return out


def load_point_cloud(path, use_model_path=False, shift_to_zero=False, offset=(0.0, 0.0, 0.0), scaling=1.0):
    """PointcloudFileHandler::loadPointCloud: an (n, 3) float32 array from a .xyz, .pcd or .binvox
    file (host code; needs no GPU)."""
    ptr = ctypes.c_void_p()
    n = ctypes.c_uint64()
    off = np.ascontiguousarray(offset, np.float32)
    check(lib().gvx_load_point_cloud(os.fsencode(path), int(bool(use_model_path)), int(bool(shift_to_zero)),
                                     _p(off), float(scaling), ctypes.byref(ptr), ctypes.byref(n)))
    try:
        if n.value == 0:
            return np.zeros((0, 3), np.float32)
        buf = (ctypes.c_float * (3 * n.value)).from_address(ptr.value)
        return np.frombuffer(buf, np.float32).reshape(-1, 3).copy()
    finally:
        lib().gvx_free(ptr)


def robot_transform(link_pts, sizes, dh, q, base=None):
    """Posed link clouds (n x 3 float32) for joint values q."""
    pts = np.ascontiguousarray(link_pts, np.float32).reshape(-1, 3)
    sizes = np.ascontiguousarray(sizes, np.uint32)
    dh = np.ascontiguousarray(dh, np.float32).reshape(-1, 5)
    b = None if base is None else np.ascontiguousarray(base, np.float32).reshape(12)
    out = np.empty_like(pts)
    check(lib().gvx_robot_transform(_p(pts), _p(sizes), len(sizes), _p(dh), _p(b),
                                    _p(np.ascontiguousarray(q, np.float32)), MEM_HOST, _p(out), MEM_HOST))
    return out


def _robot_args(link_pts, sizes, dh, q, base, meanings):
    pts = np.ascontiguousarray(link_pts, np.float32).reshape(-1, 3)
    sizes = np.ascontiguousarray(sizes, np.uint32)
    dh = np.ascontiguousarray(dh, np.float32).reshape(-1, 5)
    b = None if base is None else np.ascontiguousarray(base, np.float32).reshape(12)
    mn = None if meanings is None else np.ascontiguousarray(meanings, np.uint32)
    return pts, sizes, dh, b, np.ascontiguousarray(q, np.float32), mn


def map_insert_robot(m, link_pts, sizes, dh, q, base=None, meanings=None, meaning=0):
    pts, sizes, dh, b, q, mn = _robot_args(link_pts, sizes, dh, q, base, meanings)
    out = ctypes.c_int(0)
    check(lib().gvx_map_insert_robot(m.h, _p(pts), _p(sizes), len(sizes), _p(dh), _p(b), _p(q), _p(mn), meaning,
                                     MEM_HOST, ctypes.byref(out)))
    return bool(out.value)


def map_insert_robot_selfcoll(m, link_pts, sizes, dh, q, base=None, meanings=None, masks=None):
    """insertRobotIntoMapSelfCollAware: (self collision, colliding meanings, points outside)"""
    pts, sizes, dh, b, q, mn = _robot_args(link_pts, sizes, dh, q, base, meanings)
    mk = None if masks is None else np.ascontiguousarray(masks, np.uint8).reshape(-1, 32)
    col = np.zeros(32, np.uint8)
    sc, out = ctypes.c_int(0), ctypes.c_int(0)
    check(lib().gvx_map_insert_robot_selfcoll(m.h, _p(pts), _p(sizes), len(sizes), _p(dh), _p(b), _p(q), _p(mn),
                                              _p(mk), MEM_HOST, _p(col), ctypes.byref(sc), ctypes.byref(out)))
    return bool(sc.value), col, bool(out.value)


def list_insert_robot(lst, link_pts, sizes, dh, q, base=None, meanings=None, meaning=0):
    pts, sizes, dh, b, q, mn = _robot_args(link_pts, sizes, dh, q, base, meanings)
    check(lib().gvx_list_insert_robot(lst.h, _p(pts), _p(sizes), len(sizes), _p(dh), _p(b), _p(q), _p(mn),
                                      meaning, MEM_HOST))


class VoxelList:
    """BitVectorVoxelList."""

    def __init__(self, dims, side):
        h = ctypes.c_void_p()
        check(lib().gvx_list_create(int(dims[0]), int(dims[1]), int(dims[2]), float(side), ctypes.byref(h)))
        self.h = h
        self.dims = tuple(int(d) for d in dims)
        self.side = float(side)

    def close(self):
        if self.h:
            lib().gvx_list_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __len__(self):
        n = ctypes.c_uint64()
        check(lib().gvx_list_size(self.h, ctypes.byref(n)))
        return n.value

    def clear(self):
        check(lib().gvx_list_clear(self.h))

    def shift_sv(self, shift):
        """shiftLeftSweptVolumeIDs"""
        check(lib().gvx_list_shift_sv(self.h, int(shift)))

    def collide_bitcheck(self, other, margin=0):
        """collideWithBitcheck"""
        c = ctypes.c_uint64()
        check(lib().gvx_list_collide_bitcheck(self.h, other.h, int(margin), ctypes.byref(c)))
        return c.value

    def collide_per_meaning(self, other):
        """collideCountingPerMeaning: (total, counts[256])"""
        counts = np.zeros(256, np.uint64)
        t = ctypes.c_uint64()
        check(lib().gvx_list_collide_per_meaning(self.h, other.h, _p(counts), ctypes.byref(t)))
        return t.value, counts

    def write_to_disk(self, path):
        """TemplateVoxelList::writeToDisk"""
        check(lib().gvx_list_write_file(self.h, os.fsencode(path)))

    def read_from_disk(self, path):
        """TemplateVoxelList::readFromDisk"""
        check(lib().gvx_list_read_file(self.h, os.fsencode(path)))

    def read(self):
        n = len(self)
        ids = np.empty(n, np.uint32)
        coords = np.empty((n, 3), np.uint32)
        bits = np.empty((n, 32), np.uint8)
        check(lib().gvx_list_read(self.h, _p(ids), _p(coords), _p(bits), MEM_HOST))
        return ids, coords, bits

    def insert_points(self, pts, meaning=0, sizes=None, meanings=None):
        pts = _host(pts, np.float32)
        n = pts.shape[0] if not hasattr(pts, "numel") else pts.numel() // 3
        sz = None if sizes is None else np.ascontiguousarray(sizes, np.uint32)
        mn = None if meanings is None else np.ascontiguousarray(meanings, np.uint32)
        check(lib().gvx_list_insert_points(self.h, _p(pts), n, _p(sz), 0 if sz is None else len(sz), _p(mn),
                                           meaning, _mem(pts)))

    def insert_coords(self, coords, meaning):
        c = _host(coords, np.uint32)
        n = c.shape[0] if not hasattr(c, "numel") else c.numel() // 3
        check(lib().gvx_list_insert_coords(self.h, _p(c), n, meaning, _mem(c)))

    def merge_list(self, other, offset=None, meaning=-1):
        check(lib().gvx_list_merge_list(self.h, other.h, _p(_i3(offset)), int(meaning)))

    def merge_map(self, m, offset=None, meaning=-1):
        check(lib().gvx_list_merge_map(self.h, m.h, _p(_i3(offset)), int(meaning)))

    def collide_map(self, m, threshold=1.0, offset=None, mode=0, typemask=None):
        c = ctypes.c_uint64()
        mask = np.zeros(32, np.uint8)
        tm = None if typemask is None else np.ascontiguousarray(typemask, np.uint8)
        check(lib().gvx_list_collide_map(self.h, m.h, float(threshold), _p(_i3(offset)), int(mode), _p(tm),
                                         ctypes.byref(c), _p(mask)))
        return c.value, mask

    def collide_list(self, other):
        c = ctypes.c_uint64()
        mask = np.zeros(32, np.uint8)
        check(lib().gvx_list_collide_list(self.h, other.h, ctypes.byref(c), _p(mask)))
        return c.value, mask


class Batch:
    """Batched multi-pose collision of a DH robot against one environment map."""

    def __init__(self, env, link_pts, link_sizes, dh, base=None, max_poses=1024, threshold=1.0):
        self.link_pts = np.ascontiguousarray(link_pts, np.float32).reshape(-1, 3)
        self.sizes = np.ascontiguousarray(link_sizes, np.uint32)
        self.dh = np.ascontiguousarray(dh, np.float32).reshape(-1, 5)
        self.base = None if base is None else np.ascontiguousarray(base, np.float32).reshape(12)
        self.n_links = self.dh.shape[0]
        self.max_poses = int(max_poses)
        h = ctypes.c_void_p()
        check(lib().gvx_batch_create(env.h, float(threshold), _p(self.link_pts), _p(self.sizes), len(self.sizes),
                                     _p(self.dh), _p(self.base), self.max_poses, ctypes.byref(h)))
        self.h = h
        self.env = env

    def close(self):
        if self.h:
            lib().gvx_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def refresh_env(self):
        check(lib().gvx_batch_refresh_env(self.h))

    def collide(self, q, counts=None, masks=None):
        """q: (n, n_links) float32 (numpy or device tensor).  Returns (counts, masks)."""
        q = _host(q, np.float32)
        n = q.shape[0]
        out_mem = MEM_HOST
        if counts is None:
            counts = np.empty(n, np.uint32)
        if masks is None:
            masks = np.empty((n, 32), np.uint8)
        if hasattr(counts, "is_cuda") and counts.is_cuda:
            out_mem = MEM_DEVICE
        check(lib().gvx_batch_collide(self.h, _p(q), n, _mem(q), _p(counts), _p(masks), out_mem))
        return counts, masks

    def collide_async(self, q_dev, counts_dev, masks_dev):
        """Device tensors only: q (n, n_links) float32, counts (n,) 32-bit, masks (n, 32) uint8.
        Enqueued on the library stream; gvx.sync() waits."""
        n = int(q_dev.shape[0])
        check(lib().gvx_batch_collide_async(self.h, _p(q_dev), n, _p(counts_dev), _p(masks_dev)))

    def last_stats(self):
        """(pairs tested, pairs kept by the broad phase, their point-poses) of the last call."""
        v = [ctypes.c_uint64() for _ in range(3)]
        check(lib().gvx_batch_last_stats(self.h, *[ctypes.byref(x) for x in v]))
        return tuple(int(x.value) for x in v)

    def autotune(self, n):
        s = ctypes.c_int()
        check(lib().gvx_batch_autotune(self.h, int(n), ctypes.byref(s)))
        return s.value

    def set_schedule(self, s):
        check(lib().gvx_batch_set_schedule(self.h, int(s)))

    def collide_per_pose(self, robot_map, q):
        q = np.ascontiguousarray(q, np.float32)
        counts = np.empty(q.shape[0], np.uint32)
        check(lib().gvx_batch_collide_per_pose(self.h, robot_map.h, _p(q), q.shape[0], MEM_HOST, _p(counts)))
        return counts


class Cost:
    """Hinge-squared clearance cost on a float EDT grid (device pointer)."""

    def __init__(self, grid, dims, side, dh, spheres, eps=0.05, weight=1.0, base=None):
        """grid: float32 CUDA tensor of dims[0] * dims[1] * dims[2] metric distances (kept alive by this
        object), or a raw device address the caller keeps alive"""
        n_grid = int(dims[0]) * int(dims[1]) * int(dims[2])
        if hasattr(grid, "data_ptr"):
            if not grid.is_cuda or str(grid.dtype) != "torch.float32" or grid.numel() != n_grid \
                    or not grid.is_contiguous():
                raise ValueError("Cost: grid must be a contiguous float32 CUDA tensor of prod(dims) values")
            self._grid = grid
            grid_ptr = grid.data_ptr()
        else:
            self._grid = None
            grid_ptr = int(grid)
        self.dh = np.ascontiguousarray(dh, np.float32).reshape(-1, 5)
        self.n_links = self.dh.shape[0]
        sph = np.ascontiguousarray(spheres, np.float32).reshape(self.n_links, -1, 4)
        self.spl = sph.shape[1]
        self.sph = sph
        self.base = None if base is None else np.ascontiguousarray(base, np.float32).reshape(12)
        d = np.ascontiguousarray(dims, np.uint32)
        h = ctypes.c_void_p()
        check(lib().gvx_cost_create(ctypes.c_void_p(grid_ptr), _p(d), float(side), _p(self.dh), self.n_links,
                                    _p(self.base), _p(sph), self.spl, float(eps), float(weight), ctypes.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            lib().gvx_cost_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, name, t, shape):
        if t is None:
            return
        if not (hasattr(t, "is_cuda") and t.is_cuda and str(t.dtype) == "torch.float32" and t.is_contiguous()):
            raise ValueError(f"Cost: {name} must be a contiguous float32 CUDA tensor")
        if tuple(t.shape) != tuple(shape):
            raise ValueError(f"Cost: {name} has shape {tuple(t.shape)}, expected {tuple(shape)}")

    def forward(self, q, out):
        """q: device float32 (T, W, L); out: device float32 (T,)"""
        self._check("q", q, (q.shape[0], q.shape[1], self.n_links) if q.dim() == 3 else (-1,))
        self._check("out", out, (q.shape[0],))
        check(lib().gvx_cost_forward(self.h, _p(q), q.shape[0], q.shape[1], _p(out)))
        return out

    def backward(self, q, dq, dx=None, upstream=None):
        """dq: (T, W, L); dx (optional): (T, W, L, spheres per link, 3); upstream (optional): (T,)"""
        self._check("q", q, (q.shape[0], q.shape[1], self.n_links) if q.dim() == 3 else (-1,))
        self._check("dq", dq, q.shape)
        self._check("dx", dx, (q.shape[0], q.shape[1], self.n_links, self.spl, 3))
        self._check("upstream", upstream, (q.shape[0],))
        check(lib().gvx_cost_backward(self.h, _p(q), q.shape[0], q.shape[1], _p(upstream), _p(dq), _p(dx)))
        return dq, dx
