"""ctypes binding of libfusion_amd.so (the C-ABI in include/libfusion_amd.h).

torch is used only as the device allocator: every call passes raw device
pointers (tensor.data_ptr()) and the current HIP stream.  Nothing here falls
back to a CPU path: if the library is missing, load() raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfusion_amd.so")
HEADER = os.path.join(os.path.dirname(_HERE), "include", "libfusion_amd.h")

LF_OK, LF_ERR_ARG, LF_ERR_HIP, LF_ERR_CAPACITY, LF_ERR_NOT_READY = 0, 1, 2, 3, 4
CNT_ALLOCATED, CNT_VISIBLE, CNT_OVERFLOW, CNT_TRIANGLES = 0, 1, 2, 3
CNT_VIS_SUM_LO, CNT_VIS_SUM_HI, CNT_UPDATES, CNT_RENDER_BLOCKS = 4, 5, 6, 7
CNT_FUSE_PASSES = 8  # block-selection passes since reset (seed of the mixed weights' dither)
OVF_TABLE, OVF_POOL, OVF_TRIANGLES = 1, 2, 4
SDF_UNSEEN = -32768  # sdf code of a voxel with weight 0
PRECISION_FP32 = 0   # int16 TSDF, float32 weights and maps
PRECISION_MIXED = 1  # FP16 TSDF, 8-bit weight codes, FP16 maps
OCC_WORDS = 4096  # occupancy words of a map (2^17 bits)
DIR_DIM = 256  # block directory: DIR_DIM^3 cells
DIR_CELLS = DIR_DIM ** 3
TIMING_INTEGRATE, TIMING_RAYCAST, TIMING_TRACKING, TIMING_FRAME = 0x1, 0x2, 0x4, 0x10
TIMING_ALL = 0x17
# batch stage ids (bit s of the mask brackets stage s)
BATCH_INTEGRATE, BATCH_RAYCAST, BATCH_TRACKING, BATCH_FUSION, BATCH_STEP = range(5)
BATCH_TIMING_ALL = 0x1F
BATCH_OPT_UPLOAD_OVERLAP, BATCH_OPT_TRACK_WORKGROUPS, BATCH_OPT_INTEGRATE_WORKGROUPS, BATCH_OPT_ZRANGE_WORKGROUPS = range(4)


class LfIntrinsics(ctypes.Structure):
    _fields_ = [("width", ctypes.c_int), ("height", ctypes.c_int),
                ("fx", ctypes.c_float), ("fy", ctypes.c_float), ("cx", ctypes.c_float),
                ("cy", ctypes.c_float), ("invfx", ctypes.c_float), ("invfy", ctypes.c_float)]


class LfMapParams(ctypes.Structure):
    _fields_ = [("num_buckets", ctypes.c_int), ("num_entries", ctypes.c_int),
                ("num_blocks", ctypes.c_int), ("max_triangles", ctypes.c_int),
                ("zmin_raycast", ctypes.c_float), ("zmax_raycast", ctypes.c_float),
                ("zmin_update", ctypes.c_float), ("zmax_update", ctypes.c_float),
                ("voxel_size", ctypes.c_float), ("truncation", ctypes.c_float),
                ("precision", ctypes.c_int)]


class LfMapStorage(ctypes.Structure):
    _fields_ = [("keys", ctypes.c_void_p), ("slot_block", ctypes.c_void_p), ("slot_flag", ctypes.c_void_p),
                ("block_pos", ctypes.c_void_p), ("sdf", ctypes.c_void_p), ("weight", ctypes.c_void_p),
                ("rgb", ctypes.c_void_p), ("visible", ctypes.c_void_p), ("counters", ctypes.c_void_p),
                ("zrange", ctypes.c_void_p), ("occupancy", ctypes.c_void_p),
                ("visible_pos", ctypes.c_void_p), ("block_dir", ctypes.c_void_p), ("tiles_x", ctypes.c_int),
                ("tiles_y", ctypes.c_int)]


class LfTrackingResult(ctypes.Structure):
    _fields_ = [("success", ctypes.c_int), ("icp_error", ctypes.c_float), ("update", ctypes.c_double * 12)]


class LfSystemConfig(ctypes.Structure):
    _fields_ = [("K", LfIntrinsics), ("map", LfMapParams), ("num_pyr", ctypes.c_int),
                ("iterations", ctypes.c_int * 8), ("depth_scale", ctypes.c_float),
                ("integrate_colour", ctypes.c_int)]


# (name, restype, argtypes) of every exported function; the CPU tests check
# that this list covers every declaration of the header.
P, I, F, U = ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_uint
SIGNATURES = [
    ("lf_version", ctypes.c_char_p, []),
    ("lf_intrinsics_make", None, [I, I, F, F, F, F, P]),
    ("lf_intrinsics_pyr_down", None, [P, P]),
    ("lf_compute_hash", I, [I, I, I, I]),
    ("lf_occupancy_bit", I, [I, I, I]),
    ("lf_dir_cell", I, [I, I, I]),
    ("lf_weight_code", I, [F]),
    ("lf_weight_value", F, [I]),
    ("lf_weight_code_dither", I, [F, U]),
    ("lf_mc_tables", None, [P, P]),
    ("lf_device_count", I, []),
    ("lf_last_error", ctypes.c_char_p, []),
    ("lf_map_create", I, [P, I, I, P]),
    ("lf_map_release", I, [P]),
    ("lf_map_reset", I, [P, P, P]),
    ("lf_map_read_counters", I, [P, P, P]),
    ("lf_update_weighted", I, [P, P, P, P, P, I, I, P, P, P]),
    ("lf_update", I, [P, P, P, P, I, I, P, P, P]),
    ("lf_raycast", I, [P, P, P, P, I, I, P, P, I, P]),
    ("lf_select_visible", I, [P, P, P, P, I, I, P, P]),
    ("lf_raycast_check_visibility", I, [P, P, P, P, I, I, P, P, I, P]),
    ("lf_create_mesh", I, [P, P, P, P, P, I, P, P]),
    ("lf_map_write", I, [P, P, ctypes.c_char_p, P]),
    ("lf_map_read", I, [P, P, ctypes.c_char_p, P]),
    ("lf_map_copy", I, [P, P, P, P]),
    ("lf_depth_to_float", I, [P, P, I, I, F, P]),
    ("lf_bilateral_filter", I, [P, P, I, I, P]),
    ("lf_pyr_down_depth", I, [P, P, I, I, P]),
    ("lf_pyr_down_vmap", I, [P, P, I, I, P]),
    ("lf_pyr_down_image", I, [P, P, I, I, P]),
    ("lf_rgb_to_intensity", I, [P, P, I, I, P]),
    ("lf_compute_derivative", I, [P, P, P, I, I, P]),
    ("lf_back_project_depth", I, [P, P, I, I, P, P]),
    ("lf_compute_nmap", I, [P, P, I, I, P]),
    ("lf_render_scene", I, [P, P, P, I, I, P, P]),
    ("lf_semi_dense", I, [P, P, P, I, I, F, F, P, P]),
    ("lf_warp_image", I, [P, P, I, I, P, P, P, P]),
    ("lf_image_create", I, [P, I, P, P]),
    ("lf_image_destroy", I, [P]),
    ("lf_image_upload", I, [P, P, P, F]),
    ("lf_image_resize_device_map", I, [P]),
    ("lf_image_plane", I, [P, I, I, P]),
    ("lf_image_render", I, [P, I, P]),
    ("lf_update_image", I, [P, P, P, P, I, P]),
    ("lf_deintegrate_weighted", I, [P, P, P, P, P, I, I, P, P, P]),
    ("lf_deintegrate_image", I, [P, P, P, P, I, P]),
    ("lf_refuse_keyframes", I, [P, P, P, P, P, I, I, P]),
    ("lf_tracker_create", I, [P, I, P, P]),
    ("lf_tracker_destroy", I, [P]),
    ("lf_tracker_set_source_vmap", I, [P, P]),
    ("lf_tracker_set_source_image", I, [P, P]),
    ("lf_tracker_set_source_depth", I, [P, P]),
    ("lf_tracker_set_source_intensity", I, [P, P]),
    ("lf_tracker_set_reference_image", I, [P, P]),
    ("lf_tracker_set_reference_intensity", I, [P, P]),
    ("lf_tracker_set_reference_vmap", I, [P, P]),
    ("lf_tracker_swap_intensity", I, [P]),
    ("lf_tracker_compute_transform", I, [P, P, P, I, P]),
    ("lf_tracker_compute_transform_images", I, [P, P, P, P, P, I, P]),
    ("lf_tracker_plane", I, [P, I, I, I, P]),
    ("lf_tracker_get_sums", I, [P, P]),
    ("lf_reduce_scratch_floats", ctypes.c_size_t, []),
    ("lf_icp_reduce", I, [P, P, P, P, I, I, P, P, P, P, P]),
    ("lf_icp_reduce_noalloc", I, [P, P, P, P, I, I, P, P, P, P]),
    ("lf_rgb_step", I, [P, P, P, P, P, I, I, F, P, P, P, P, P]),
    ("lf_unpack_jtjjtr", None, [P, P, P, P]),
    ("lf_system_create", I, [P, P]),
    ("lf_system_destroy", I, [P]),
    ("lf_system_process_images", I, [P, P, P]),
    ("lf_system_process_images_host", I, [P, P, P]),
    ("lf_system_synchronize", I, [P]),
    ("lf_system_get_pose", I, [P, P]),
    ("lf_system_get_tracking", I, [P, P, P]),
    ("lf_system_restart", I, [P]),
    ("lf_system_fetch_mesh", I, [P, P, P, P, I, P]),
    ("lf_system_map", I, [P, P, P]),
    ("lf_system_maps", I, [P, I, P, P]),
    ("lf_system_render", I, [P, I, P]),
    ("lf_system_write_map", I, [P, ctypes.c_char_p]),
    ("lf_system_read_map", I, [P, ctypes.c_char_p]),
    ("lf_system_stream", P, [P]),
    ("lf_system_set_timing", I, [P, I]),
    ("lf_system_set_graphs", I, [P, I]),
    ("lf_system_read_timing", I, [P, I, P, P]),
    ("lf_batch_create", I, [P, I, P]),
    ("lf_batch_destroy", I, [P]),
    ("lf_batch_process", I, [P, P, P]),
    ("lf_batch_process_host", I, [P, P, P]),
    ("lf_batch_synchronize", I, [P]),
    ("lf_batch_set_timing", I, [P, I]),
    ("lf_batch_set_option", I, [P, I, I]),
    ("lf_batch_set_graphs", I, [P, I]),
    ("lf_batch_autotune", I, [P, I, I]),
    ("lf_batch_tuning", I, [P]),
    ("lf_batch_tune_file", I, [ctypes.c_char_p]),
    ("lf_batch_map", I, [P, I, P, P]),
    ("lf_batch_get_option", I, [P, I, P]),
    ("lf_batch_read_timing", I, [P, I, P, P]),
    ("lf_batch_get_pose", I, [P, I, P]),
    ("lf_batch_get_tracking", I, [P, I, P, P]),
    ("lf_batch_read_counters", I, [P, I, P]),
]

_lib = None


def load():
    """Load libfusion_amd.so (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libfusion_amd.so is not built (run libfusion_amd.build.build_all()); "
                               "there is no fallback path")
        lib = ctypes.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


class LfError(RuntimeError):
    pass


def check(rc):
    if rc != LF_OK:
        msg = load().lf_last_error().decode(errors="replace")
        raise LfError("libfusion_amd error %d: %s" % (rc, msg))
    return rc


def intrinsics(width, height, fx, fy, cx, cy):
    K = LfIntrinsics()
    load().lf_intrinsics_make(int(width), int(height), float(fx), float(fy), float(cx), float(cy), ctypes.byref(K))
    return K


def pyr_down(K):
    out = LfIntrinsics()
    load().lf_intrinsics_pyr_down(ctypes.byref(K), ctypes.byref(out))
    return out


def map_params(num_buckets, num_blocks, voxel_size, truncation, excess_ratio=0.25, zmin=0.3, zmax=5.0,
               max_triangles=20_000_000, precision=PRECISION_FP32):
    return LfMapParams(int(num_buckets), int(num_buckets + int(num_buckets * excess_ratio)), int(num_blocks),
                       int(max_triangles), float(zmin), float(zmax), float(zmin), float(zmax),
                       float(voxel_size), float(truncation), int(precision))


def weight_values():
    """The 256 weights of the 8-bit codes of a mixed-precision map."""
    return np.array([load().lf_weight_value(c) for c in range(256)], np.float32)


def read_tsdf(params, storage):
    """(blocks[n,3], tsdf float32[n,512] in truncation units with NaN where the
    weight is 0, weight float32[n,512]) of a map's allocated blocks in pool
    order, for either precision."""
    torch = _torch()
    c = np.zeros(16, np.uint32)
    check(load().lf_map_read_counters(ctypes.byref(storage), c.ctypes.data_as(P), None))
    n = min(int(c[CNT_ALLOCATED]), params.num_blocks)
    pos = torch.empty(4 * max(n, 1), dtype=torch.int32, device="cuda")
    sdf = torch.empty(512 * max(n, 1), dtype=torch.int16, device="cuda")
    mixed = params.precision == PRECISION_MIXED
    w = torch.empty(512 * max(n, 1), dtype=torch.int32 if mixed else torch.float32, device="cuda")
    if n:
        copy_device(pos, storage.block_pos, 16 * n)
        copy_device(sdf, storage.sdf, 1024 * n)
        copy_device(w, storage.weight, 512 * n * 4)
    pos = pos.cpu().numpy()[: 4 * n].reshape(n, 4)[:, :3]
    sdf = sdf.cpu().numpy()[: 512 * n].reshape(n, 512)
    w = w.cpu().numpy()[: 512 * n].reshape(n, 512)
    if mixed:
        weight = weight_values()[w.view(np.uint32) & 0xFF]
        tsdf = sdf.view(np.float16).astype(np.float32)
    else:
        weight = w.astype(np.float32)
        tsdf = np.where(sdf == SDF_UNSEEN, np.nan, sdf.astype(np.float32) / np.float32(32767)).astype(np.float32)
    tsdf[weight == 0] = np.nan
    return pos, tsdf, weight


def mc_tables():
    e = np.zeros(256, np.int32)
    t = np.zeros((256, 16), np.int32)
    load().lf_mc_tables(e.ctypes.data_as(P), t.ctypes.data_as(P))
    return e, t


def pose24(T):
    """Float pose block (T_world_camera 3x4, then its inverse) from a 4x4 FP64 pose."""
    T = np.asarray(T, np.float64)
    Ti = np.eye(4)
    Ti[:3, :3] = T[:3, :3].T
    Ti[:3, 3] = -T[:3, :3].T @ T[:3, 3]
    return np.concatenate([T[:3, :].ravel(), Ti[:3, :].ravel()]).astype(np.float32)


def _torch():
    import torch
    return torch


def stream_ptr():
    torch = _torch()
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def ptr(t):
    return ctypes.c_void_p(0 if t is None else t.data_ptr())


def host_ptr(a):
    """Address of a contiguous host buffer (a CPU tensor or a numpy array)."""
    if isinstance(a, np.ndarray):
        assert a.flags["C_CONTIGUOUS"]
        return a.ctypes.data
    assert not a.is_cuda and a.is_contiguous()
    return a.data_ptr()


PLANE_DEPTH, PLANE_INTENSITY, PLANE_DX, PLANE_DY, PLANE_VMAP, PLANE_NMAP, PLANE_RAW_DEPTH, PLANE_RGB, PLANE_SAMPLES = range(9)

_hip = None


def copy_device(dst, src_ptr, nbytes):
    """Synchronous device-to-device copy from a raw device pointer into a tensor."""
    global _hip
    if _hip is None:
        _hip = ctypes.CDLL("libamdhip64.so")
        _hip.hipMemcpy.restype = ctypes.c_int
        _hip.hipMemcpy.argtypes = [P, P, ctypes.c_size_t, ctypes.c_int]
        _hip.hipDeviceSynchronize.restype = ctypes.c_int
    if dst is None:
        return
    assert dst.is_contiguous() and dst.numel() * dst.element_size() >= nbytes
    if _hip.hipDeviceSynchronize() != 0 or _hip.hipMemcpy(ptr(dst), ctypes.c_void_p(src_ptr), nbytes, 3) != 0:
        raise LfError("hipMemcpy failed")


def copy_device_to(dst_ptr, src):
    """Synchronous device-to-device copy from a tensor to a raw device pointer."""
    copy_device(None, 0, 0)  # loads the runtime
    assert src.is_contiguous()
    if _hip.hipDeviceSynchronize() != 0 or _hip.hipMemcpy(ctypes.c_void_p(dst_ptr), ptr(src),
                                                         src.numel() * src.element_size(), 3) != 0:
        raise LfError("hipMemcpy failed")


class Image:
    """DeviceImage: one RGB-D frame's device pyramid (lf_image_*), on the current stream."""

    def __init__(self, K, levels):
        self.K = [K]
        for _ in range(levels - 1):
            self.K.append(pyr_down(self.K[-1]))
        self.h = ctypes.c_void_p()
        check(load().lf_image_create(ctypes.byref(K), int(levels), stream_ptr(), ctypes.byref(self.h)))

    def close(self):
        if self.h:
            load().lf_image_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, depth, rgb, depth_scale):
        check(load().lf_image_upload(self.h, ptr(depth), ptr(rgb), float(np.float32(depth_scale))))

    def resize_device_map(self):
        check(load().lf_image_resize_device_map(self.h))

    def plane_ptr(self, level, plane):
        out = ctypes.c_void_p()
        check(load().lf_image_plane(self.h, int(level), int(plane), ctypes.byref(out)))
        return out.value

    def render(self, textured=False):
        """get_rendered_image / get_rendered_scene_textured: a copy of the RGBA8 render."""
        torch = _torch()
        out = ctypes.c_void_p()
        check(load().lf_image_render(self.h, 1 if textured else 0, ctypes.byref(out)))
        t = torch.empty((self.K[0].height, self.K[0].width, 4), dtype=torch.uint8, device="cuda")
        copy_device(t, out.value, t.numel())
        return t

    def plane(self, level, plane):
        """A copy of a plane as a tensor ([h, w], [h, w, 4] for maps, [h, w, 3] uint8 for colour)."""
        torch = _torch()
        w, h = self.K[level].width, self.K[level].height
        if plane in (PLANE_VMAP, PLANE_NMAP, PLANE_SAMPLES):
            t = torch.empty((h, w, 4), device="cuda")
        elif plane == PLANE_RGB:
            t = torch.empty((h, w, 3), dtype=torch.uint8, device="cuda")
        else:
            t = torch.empty((h, w), device="cuda")
        copy_device(t, self.plane_ptr(level, plane), t.numel() * t.element_size())
        return t


class Tracker:
    """DenseTracking(K, NUM_PYR) with source / reference pyramids (lf_tracker_*), on the current stream."""

    def __init__(self, K, levels):
        self.K = [K]
        for _ in range(levels - 1):
            self.K.append(pyr_down(self.K[-1]))
        self.h = ctypes.c_void_p()
        check(load().lf_tracker_create(ctypes.byref(K), int(levels), stream_ptr(), ctypes.byref(self.h)))

    def close(self):
        if self.h:
            load().lf_tracker_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set(self, which, plane):
        """which: set_source_vmap, set_source_image, ... (a device tensor at level-0 size)."""
        check(getattr(load(), "lf_tracker_" + which)(self.h, ptr(plane)))

    def swap_intensity(self):
        check(load().lf_tracker_swap_intensity(self.h))

    @staticmethod
    def _result(r):
        T = np.eye(4)
        T[:3, :] = np.array(r.update).reshape(3, 4)
        return bool(r.success), float(r.icp_error), T

    def compute_transform(self, iterations, init=None):
        """compute_transform(ctx): (success, icp_error, update 4x4)."""
        r = LfTrackingResult()
        its = (ctypes.c_int * len(iterations))(*iterations)
        ini = None if init is None else (ctypes.c_double * 12)(*np.asarray(init, np.float64)[:3, :].ravel())
        check(load().lf_tracker_compute_transform(self.h, ini, its, len(iterations), ctypes.byref(r)))
        return self._result(r)

    def sums(self):
        """The last Gauss-Newton step's FP64 sums (31 values, lf_tracker_get_sums)."""
        out = (ctypes.c_double * 31)()
        check(load().lf_tracker_get_sums(self.h, out))
        return np.array(out, np.float64)

    def compute_transform_images(self, reference, current, iterations, init=None):
        r = LfTrackingResult()
        its = (ctypes.c_int * len(iterations))(*iterations)
        ini = None if init is None else (ctypes.c_double * 12)(*np.asarray(init, np.float64)[:3, :].ravel())
        check(load().lf_tracker_compute_transform_images(self.h, reference.h, current.h, ini, its, len(iterations),
                                                         ctypes.byref(r)))
        return self._result(r)


class Map:
    """A device voxel-hash map whose arrays are torch tensors (layout of lf_map_storage)."""

    def __init__(self, params, width, height, device="cuda"):
        torch = _torch()
        self.params = params
        E, B = params.num_entries, params.num_blocks
        tx, ty = width // 8, height // 8
        self.t = {
            "keys": torch.empty(E, dtype=torch.int64, device=device),
            "slot_block": torch.empty(E, dtype=torch.int32, device=device),
            "slot_flag": torch.empty(E, dtype=torch.uint8, device=device),
            "block_pos": torch.empty(B * 4, dtype=torch.int32, device=device),
            "sdf": torch.empty(B * 512, dtype=torch.int16, device=device),
            # FP32: float32 weights; mixed: one word per voxel, weight code | R<<8 | G<<16 | B<<24
            "weight": torch.empty(B * 512, dtype=torch.int32 if params.precision == PRECISION_MIXED
                                  else torch.float32, device=device),
            # colour planes of an FP32 map (a mixed map packs colour into its weight words)
            "rgb": torch.empty(B * 1536 if params.precision != PRECISION_MIXED else 0, dtype=torch.uint8,
                               device=device),
            "visible": torch.empty(B, dtype=torch.int32, device=device),
            "counters": torch.zeros(16, dtype=torch.int32, device=device),
            "zrange": torch.empty(2 * tx * ty, dtype=torch.float32, device=device),
            "occupancy": torch.empty(OCC_WORDS, dtype=torch.int32, device=device),
            "visible_pos": torch.empty(B * 4, dtype=torch.int32, device=device),
            "block_dir": torch.empty(DIR_CELLS, dtype=torch.int32, device=device),
        }
        s = LfMapStorage()
        for k, v in self.t.items():
            setattr(s, k, v.data_ptr() if v.numel() else 0)
        s.tiles_x, s.tiles_y = tx, ty
        self.storage = s
        self.reset()

    def reset(self):
        check(load().lf_map_reset(ctypes.byref(self.params), ctypes.byref(self.storage), stream_ptr()))

    def counters(self):
        return self.t["counters"].cpu().numpy().view(np.uint32).copy()

    def update_weighted(self, depth, nmap, rgb, pose24_dev, K):
        h, w = depth.shape[:2]
        check(load().lf_update_weighted(ctypes.byref(self.params), ctypes.byref(self.storage), ptr(depth), ptr(nmap),
                                        ptr(rgb), w, h, ptr(pose24_dev), ctypes.byref(K), stream_ptr()))

    def write(self, path):
        """writeMapToDisk: the reference's binary layout plus <path>.txt."""
        check(load().lf_map_write(ctypes.byref(self.params), ctypes.byref(self.storage), str(path).encode(),
                                  stream_ptr()))

    def read(self, path):
        """readMapFromDisk: reset, then load a reference-layout file."""
        check(load().lf_map_read(ctypes.byref(self.params), ctypes.byref(self.storage), str(path).encode(),
                                 stream_ptr()))

    def copy_to(self, other):
        check(load().lf_map_copy(ctypes.byref(self.params), ctypes.byref(self.storage), ctypes.byref(other.storage),
                                 stream_ptr()))

    def update_image(self, image, pose24_dev, colour=True):
        """update(RgbdImagePtr): weighted update from an Image's planes (packed samples)."""
        check(load().lf_update_image(ctypes.byref(self.params), ctypes.byref(self.storage), image.h,
                                     ptr(pose24_dev), 1 if colour else 0, stream_ptr()))

    def deintegrate_image(self, image, pose24_dev, colour=True):
        """Remove a frame fused with update_image (inverse weighted update)."""
        check(load().lf_deintegrate_image(ctypes.byref(self.params), ctypes.byref(self.storage), image.h,
                                          ptr(pose24_dev), 1 if colour else 0, stream_ptr()))

    def refuse_keyframes(self, images, old_pose24_devs, new_pose24_devs, colour=True):
        """Keyframe re-fusion (lf_refuse_keyframes): every image leaves the map at
        its old pose and is fused at its new pose, in one batched pass."""
        n = len(images)
        if not (len(old_pose24_devs) == len(new_pose24_devs) == n):
            raise ValueError("one old and one new pose per keyframe")
        hs = (ctypes.c_void_p * n)(*[im.h for im in images])
        olds = (ctypes.c_void_p * n)(*[ptr(t) for t in old_pose24_devs])
        news = (ctypes.c_void_p * n)(*[ptr(t) for t in new_pose24_devs])
        check(load().lf_refuse_keyframes(ctypes.byref(self.params), ctypes.byref(self.storage), hs, olds, news, n,
                                         1 if colour else 0, stream_ptr()))

    def deintegrate_weighted(self, depth, nmap, rgb, pose24_dev, K):
        h, w = depth.shape[:2]
        check(load().lf_deintegrate_weighted(ctypes.byref(self.params), ctypes.byref(self.storage), ptr(depth),
                                             ptr(nmap), ptr(rgb), w, h, ptr(pose24_dev), ctypes.byref(K), stream_ptr()))

    def update(self, depth, rgb, pose24_dev, K):
        h, w = depth.shape[:2]
        check(load().lf_update(ctypes.byref(self.params), ctypes.byref(self.storage), ptr(depth), ptr(rgb), w, h,
                               ptr(pose24_dev), ctypes.byref(K), stream_ptr()))

    def raycast(self, vmap, pose24_dev, K, check_colour=True, rgb_out=None):
        h, w = vmap.shape[:2]
        check(load().lf_raycast(ctypes.byref(self.params), ctypes.byref(self.storage), ptr(vmap), ptr(rgb_out),
                                w, h, ptr(pose24_dev), ctypes.byref(K), 1 if check_colour else 0, stream_ptr()))

    def select_visible(self, pose24_dev, K, cols, rows):
        n = ctypes.c_uint(0)
        check(load().lf_select_visible(ctypes.byref(self.params), ctypes.byref(self.storage), ptr(pose24_dev),
                                       ctypes.byref(K), int(cols), int(rows), ctypes.byref(n), stream_ptr()))
        return n.value

    def raycast_check_visibility(self, vmap, pose24_dev, K, check_colour=True, rgb_out=None):
        h, w = vmap.shape[:2]
        check(load().lf_raycast_check_visibility(ctypes.byref(self.params), ctypes.byref(self.storage), ptr(vmap),
                                                 ptr(rgb_out), w, h, ptr(pose24_dev), ctypes.byref(K),
                                                 1 if check_colour else 0, stream_ptr()))

    def mesh(self, mode=1, max_triangles=None):
        """mode 0 vertices, 1 + normals, 2 + colours -> (tri[n,3,3], extra, count)."""
        torch = _torch()
        cap = int(max_triangles if max_triangles is not None else self.params.max_triangles)
        vtx = torch.empty((max(cap, 1), 3, 3), dtype=torch.float32, device="cuda")
        nrm = torch.empty((max(cap, 1), 3, 3), dtype=torch.float32, device="cuda") if mode == 1 else None
        col = torch.empty((max(cap, 1), 3, 3), dtype=torch.uint8, device="cuda") if mode == 2 else None
        n = ctypes.c_uint(0)
        rc = load().lf_create_mesh(ctypes.byref(self.params), ctypes.byref(self.storage), ptr(vtx), ptr(nrm),
                                   ptr(col), cap, ctypes.byref(n), stream_ptr())
        if rc not in (LF_OK, LF_ERR_CAPACITY):
            check(rc)
        n = n.value
        extra = nrm if mode == 1 else col
        return vtx[:n], (extra[:n] if extra is not None else None), int(self.counters()[CNT_TRIANGLES])

    def blocks(self):
        """Allocated block coordinates (n,3) in pool order."""
        c = self.counters()
        n = min(int(c[CNT_ALLOCATED]), self.params.num_blocks)
        return self.t["block_pos"][: 4 * n].view(n, 4)[:, :3].cpu().numpy()

    def visible_blocks(self):
        c = self.counters()
        n = int(c[CNT_VISIBLE])
        idx = self.t["visible"][:n].long()
        return self.t["block_pos"].view(-1, 4)[idx, :3].cpu().numpy()

    def voxels(self):
        """(blocks[n,3], sdf int16[n,512], weight[n,512], rgb uint8[n,3,512]) in pool order.
        The sdf of a voxel with weight 0 is undefined on the device (it holds the
        unseen code SDF_UNSEEN); it is returned as 0.  FP32 maps only (read_tsdf
        reads either precision)."""
        assert self.params.precision != PRECISION_MIXED, "voxels() reads FP32 maps; use read_tsdf"
        c = self.counters()
        n = min(int(c[CNT_ALLOCATED]), self.params.num_blocks)
        sdf = self.t["sdf"][: 512 * n].view(n, 512).cpu().numpy()
        sdf = np.where(sdf == SDF_UNSEEN, np.int16(0), sdf)
        return (self.blocks(), sdf, self.t["weight"][: 512 * n].view(n, 512).cpu().numpy(),
                self.t["rgb"][: 1536 * n].view(n, 3, 512).cpu().numpy())

    def colours(self):
        """uint8[n,3,512] colour of the allocated blocks in pool order, for either precision."""
        c = self.counters()
        n = min(int(c[CNT_ALLOCATED]), self.params.num_blocks)
        if self.params.precision == PRECISION_MIXED:
            words = self.t["weight"][: 512 * n].view(n, 512).cpu().numpy().view(np.uint32)
            return np.stack([(words >> (8 + 8 * k)) & 0xFF for k in range(3)], 1).astype(np.uint8)
        return self.t["rgb"][: 1536 * n].view(n, 3, 512).cpu().numpy()


def system_config(K, params, num_pyr, iterations, depth_scale, integrate_colour=True):
    cfg = LfSystemConfig()
    cfg.K = K
    cfg.map = params
    cfg.num_pyr = int(num_pyr)
    for i, v in enumerate(iterations):
        cfg.iterations[i] = int(v)
    cfg.depth_scale = float(depth_scale)
    cfg.integrate_colour = 1 if integrate_colour else 0
    return cfg


class Batch:
    """lf_batch_*: independent systems stepped together, one frame each per
    step, with one batched raycast launch per stage for all of them."""

    def __init__(self, configs):
        n = len(configs)
        arr = (LfSystemConfig * n)(*configs)
        h = ctypes.c_void_p()
        check(load().lf_batch_create(arr, n, ctypes.byref(h)))
        self.h, self.n = h, n

    def close(self):
        if self.h:
            load().lf_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def process(self, depths, rgbs):
        """One frame per system: device tensors (uint16 depth, uint8 RGB)."""
        assert len(depths) == len(rgbs) == self.n
        d = (ctypes.c_void_p * self.n)(*[t.data_ptr() for t in depths])
        c = (ctypes.c_void_p * self.n)(*[t.data_ptr() for t in rgbs])
        check(load().lf_batch_process(self.h, d, c))

    def process_host(self, depths, rgbs):
        """One frame per system from host memory: CPU tensors (pinned for
        asynchronous copies) or contiguous numpy arrays.  They must stay
        unchanged until the step has been synchronised."""
        assert len(depths) == len(rgbs) == self.n
        d = (ctypes.c_void_p * self.n)(*[host_ptr(t) for t in depths])
        c = (ctypes.c_void_p * self.n)(*[host_ptr(t) for t in rgbs])
        check(load().lf_batch_process_host(self.h, d, c))

    def synchronize(self):
        check(load().lf_batch_synchronize(self.h))

    def set_timing(self, stages):
        """Stage mask of BATCH_TIMING_* bits (0: off)."""
        check(load().lf_batch_set_timing(self.h, int(stages)))

    def read_timing(self, stage):
        n = ctypes.c_int(0)
        t = ctypes.c_double(0)
        check(load().lf_batch_read_timing(self.h, int(stage), ctypes.byref(n), ctypes.byref(t)))
        return n.value, t.value

    def set_graphs(self, on=True):
        """Replay steady-state steps from HIP graphs (lf_batch_set_graphs)."""
        check(load().lf_batch_set_graphs(self.h, 1 if on else 0))

    def autotune(self, steps_per_trial=3, rounds=2):
        """Tune the launch options online over the next steps (lf_batch_autotune)."""
        check(load().lf_batch_autotune(self.h, int(steps_per_trial), int(rounds)))

    def tuning(self):
        return bool(load().lf_batch_tuning(self.h))

    @staticmethod
    def tune_file(path):
        """Read tuned choices from path and append new ones to it (lf_batch_tune_file)."""
        check(load().lf_batch_tune_file(None if path is None else str(path).encode()))

    def set_option(self, option, value):
        """A launch option (BATCH_OPT_*)."""
        check(load().lf_batch_set_option(self.h, int(option), int(value)))

    def option(self, option):
        v = ctypes.c_int(0)
        check(load().lf_batch_get_option(self.h, int(option), ctypes.byref(v)))
        return v.value

    def pose(self, i):
        T = np.zeros((4, 4))
        check(load().lf_batch_get_pose(self.h, int(i), T.ctypes.data_as(P)))
        return T

    def tracking(self, i):
        s = ctypes.c_int(0)
        e = ctypes.c_float(0)
        check(load().lf_batch_get_tracking(self.h, int(i), ctypes.byref(s), ctypes.byref(e)))
        return bool(s.value), float(e.value)

    def counters(self, i):
        out = np.zeros(16, np.uint32)
        check(load().lf_batch_read_counters(self.h, int(i), out.ctypes.data_as(P)))
        return out

    def map_storage(self, i):
        """System i's (params, storage) (lf_batch_map)."""
        p = LfMapParams()
        st = LfMapStorage()
        check(load().lf_batch_map(self.h, int(i), ctypes.byref(p), ctypes.byref(st)))
        return p, st


class System:
    """lf_system_* (the System facade) over device-resident frames."""

    def __init__(self, K, params, num_pyr, iterations, depth_scale, integrate_colour=True):
        cfg = LfSystemConfig()
        cfg.K = K
        cfg.map = params
        cfg.num_pyr = int(num_pyr)
        for i, v in enumerate(iterations):
            cfg.iterations[i] = int(v)
        cfg.depth_scale = float(depth_scale)
        cfg.integrate_colour = 1 if integrate_colour else 0
        self.cfg = cfg
        h = ctypes.c_void_p()
        check(load().lf_system_create(ctypes.byref(cfg), ctypes.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            load().lf_system_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def process(self, depth_dev, rgb_dev):
        check(load().lf_system_process_images(self.h, ptr(depth_dev), ptr(rgb_dev)))

    def process_host(self, depth, rgb):
        """process_images from host memory (CPU tensors or contiguous numpy arrays)."""
        check(load().lf_system_process_images_host(self.h, host_ptr(depth), host_ptr(rgb)))

    def synchronize(self):
        check(load().lf_system_synchronize(self.h))

    def pose(self):
        T = np.zeros((4, 4))
        check(load().lf_system_get_pose(self.h, T.ctypes.data_as(P)))
        return T

    def tracking(self):
        s = ctypes.c_int(0)
        e = ctypes.c_float(0)
        check(load().lf_system_get_tracking(self.h, ctypes.byref(s), ctypes.byref(e)))
        return bool(s.value), float(e.value)

    def map_storage(self):
        p = LfMapParams()
        s = LfMapStorage()
        check(load().lf_system_map(self.h, ctypes.byref(p), ctypes.byref(s)))
        return p, s

    def counters(self):
        p, s = self.map_storage()
        out = np.zeros(16, np.uint32)
        check(load().lf_map_read_counters(ctypes.byref(s), out.ctypes.data_as(P), load().lf_system_stream(self.h)))
        return out

    def ref_maps(self, level):
        v = ctypes.c_void_p()
        n = ctypes.c_void_p()
        check(load().lf_system_maps(self.h, int(level), ctypes.byref(v), ctypes.byref(n)))
        return v.value, n.value

    def mesh(self, vtx, nrm=None, col=None, cap=None):
        n = ctypes.c_uint(0)
        rc = load().lf_system_fetch_mesh(self.h, ptr(vtx), ptr(nrm), ptr(col), int(cap or vtx.shape[0]),
                                         ctypes.byref(n))
        if rc not in (LF_OK, LF_ERR_CAPACITY):
            check(rc)
        return n.value

    def set_graphs(self, on=True):
        check(load().lf_system_set_graphs(self.h, 1 if on else 0))

    def set_timing(self, stages):
        """Stage mask (TIMING_* bits); True means every stage, False none."""
        if stages is True:
            stages = TIMING_ALL
        check(load().lf_system_set_timing(self.h, int(stages or 0)))

    def read_timing(self, stage):
        n = ctypes.c_int(0)
        t = ctypes.c_double(0)
        check(load().lf_system_read_timing(self.h, int(stage), ctypes.byref(n), ctypes.byref(t)))
        return n.value, t.value

    def restart(self):
        check(load().lf_system_restart(self.h))
