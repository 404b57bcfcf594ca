"""ctypes binding of libadds.so (the C ABI declared in include/adds.h).

The shared library is built in-tree (adds_amd/lib/libadds.so) by ``make -C adds_amd/csrc``.  Loading
fails loudly when it is missing: there is no Python or CPU fallback for any solver entry point.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# ADDS_LIB selects another build of the library (for example a variant built for comparison).
LIB_PATH = os.environ.get("ADDS_LIB") or os.path.join(_HERE, "lib", "libadds.so")

# adds_status
OK, EINVAL, ENOMEM, EOVERFLOW, ETIMEOUT, EHIP, EIO, ENODEV = range(8)
# adds_wtype
W_U32, W_F32 = 0, 1
# adds_wkind
WK_U32_UNIFORM, WK_F32_UNIFORM, WK_F32_LOGNORMAL = 0, 1, 2
# flags
F_HOST_OUT, F_NO_SYNC, F_ATOMIC_PROBE, F_NO_LOCAL, F_NO_WAIT, F_REPLICA_PARTS = 0x1, 0x2, 0x4, 0x8, 0x10, 0x20
F_NO_PULL, F_PULL_EAGER, F_PULL_ALL_LIGHT, F_PULL_SPLIT, F_LOCAL_SPARSE, F_FULL_FIRST = (0x40, 0x80, 0x100, 0x200,
                                                                                0x400, 0x800)

INF_U32 = 0xFFFFFFFF
INF_F32_BITS = 0x7F800000

c_u32p = ctypes.POINTER(ctypes.c_uint32)


class AddsError(RuntimeError):
    def __init__(self, status, what=""):
        self.status = status
        msg = _lib().adds_strerror(status).decode()
        detail = _lib().adds_last_error_detail().decode()
        super().__init__(f"{what}: {msg}" + (f" ({detail})" if detail else ""))


class HostCSR(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_uint32),
        ("m", ctypes.c_uint64),
        ("row_ptr", c_u32p),
        ("col", c_u32p),
        ("w", c_u32p),
        ("wtype", ctypes.c_int),
    ]


class Params(ctypes.Structure):
    _fields_ = [
        ("delta", ctypes.c_float),
        ("wg_per_cu", ctypes.c_uint32),
        ("chunk", ctypes.c_uint32),
        ("flags", ctypes.c_uint32),
        ("nbuckets", ctypes.c_uint32),
        ("parts", ctypes.c_uint32),
        ("hub_degree", ctypes.c_uint32),
        ("task_edges", ctypes.c_uint32),
        ("concurrency", ctypes.c_uint32),
        ("timeout_ms", ctypes.c_uint32),
        ("tail_items", ctypes.c_uint32),
        ("edge_windows", ctypes.c_uint32),
        ("light_delta", ctypes.c_float),
        ("pool_entries", ctypes.c_uint64),
        ("pull_delta", ctypes.c_float),
        ("order_degree", ctypes.c_uint32),
        ("pull_early", ctypes.c_uint32),
    ]

    @classmethod
    def make(cls, **kw):
        p = cls()
        for k, v in kw.items():
            if v is None:
                continue
            if not hasattr(p, k):
                raise TypeError(f"unknown solver parameter {k!r}")
            setattr(p, k, v)
        return p


class Stats(ctypes.Structure):
    _fields_ = [
        ("ms", ctypes.c_double),
        ("vertices_expanded", ctypes.c_uint64),
        ("edges_scanned", ctypes.c_uint64),
        ("relaxations", ctypes.c_uint64),
        ("pushes", ctypes.c_uint64),
        ("stale_skips", ctypes.c_uint64),
        ("hub_tasks", ctypes.c_uint64),
        ("rebuckets", ctypes.c_uint64),
        ("pool_chunks_used", ctypes.c_uint64),
        ("kernel_ms", ctypes.c_double),
        ("probes", ctypes.c_uint64),
        ("cas_failures", ctypes.c_uint64),
        ("idle_iterations", ctypes.c_uint64),
        ("late_items", ctypes.c_uint64),
        ("ticks_probe", ctypes.c_uint64),
        ("ticks_items", ctypes.c_uint64),
        ("ticks_edges", ctypes.c_uint64),
        ("ticks_idle", ctypes.c_uint64),
        ("local_items", ctypes.c_uint64),
        ("heavy_edges", ctypes.c_uint64),
        ("ticks_flush", ctypes.c_uint64),
        ("pulls", ctypes.c_uint64),
        ("pull_edges", ctypes.c_uint64),
        ("pull_improved", ctypes.c_uint64),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


# Every symbol include/adds.h declares (checked by the CPU tests).
EXPORTS = [
    "adds_strerror", "adds_set_device", "adds_device_cu_count",
    "adds_graph_create", "adds_graph_create_device", "adds_graph_set_weights_device",
    "adds_graph_destroy", "adds_graph_info", "adds_graph_device_arrays", "adds_graph_split_rows",
    "adds_graph_set_symmetric",
    "adds_graph_check_symmetric",
    "adds_sssp", "adds_sssp_noalloc", "adds_sssp_nearfar", "adds_sssp_multi", "adds_sssp_batched", "adds_last_status",
    "adds_solve_graph_create", "adds_solve_graph_create_batched", "adds_solve_graph_launch",
    "adds_solve_graph_status", "adds_solve_graph_destroy",
    "adds_sssp_backward_f32",
    "adds_host_csr_free", "adds_read_gr", "adds_write_gr", "adds_write_dist_text", "adds_read_mtx",
    "adds_read_dimacs",
    "adds_gen_rmat", "adds_gen_grid", "adds_gen_er", "adds_gen_weights", "adds_max_degree_vertex",
    "adds_sample_component", "adds_reached_counts", "adds_last_error_detail", "adds_autotune",
    "adds_autotune_batched",
    "adds_litmus", "adds_xcd_map", "adds_gather_ceiling",
]

_LIB = None


def _lib():
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `make -C adds_amd/csrc` "
            "(or __graft_entry__.build()); the ADDS solver has no fallback path")
    # One HIP runtime per process.  PyTorch for ROCm ships its own HIP and HSA runtimes; once torch is
    # loaded, libadds.so binds to those (same sonames).  Loaded first, libadds.so would bring in the
    # system runtime and torch's would then find no GPU.  So torch, when installed, is loaded first.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(LIB_PATH)
    vp, u32, u64, i32 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int
    hp = ctypes.POINTER(HostCSR)
    pp = ctypes.POINTER(Params)
    sp = ctypes.POINTER(Stats)
    sig = {
        "adds_strerror": (ctypes.c_char_p, [i32]),
        "adds_last_error_detail": (ctypes.c_char_p, []),
        "adds_set_device": (i32, [i32]),
        "adds_device_cu_count": (u32, []),
        "adds_graph_create": (i32, [c_u32p, c_u32p, vp, u32, u64, i32, ctypes.POINTER(vp)]),
        "adds_graph_create_device": (i32, [vp, vp, vp, u32, u64, i32, ctypes.POINTER(vp)]),
        "adds_graph_set_weights_device": (i32, [vp, vp, vp]),
        "adds_graph_destroy": (None, [vp]),
        "adds_graph_set_symmetric": (i32, [vp, i32]),
        "adds_graph_check_symmetric": (i32, [vp, ctypes.POINTER(i32)]),
        "adds_graph_info": (i32, [vp, ctypes.POINTER(u32), ctypes.POINTER(u64), ctypes.POINTER(i32)]),
        "adds_graph_device_arrays": (i32, [vp, ctypes.POINTER(vp), ctypes.POINTER(vp), ctypes.POINTER(vp)]),
        "adds_graph_split_rows": (i32, [vp, ctypes.c_float, ctypes.POINTER(vp), ctypes.POINTER(vp)]),
        "adds_sssp": (i32, [vp, u32, vp, pp, sp, vp]),
        "adds_sssp_noalloc": (i32, [vp, u32, vp, pp, sp, vp]),
        "adds_sssp_nearfar": (i32, [vp, u32, vp, ctypes.c_float, u32, sp, vp]),
        "adds_sssp_multi": (i32, [vp, c_u32p, u32, vp, pp, sp, vp]),
        "adds_sssp_batched": (i32, [ctypes.POINTER(vp), c_u32p, u32, ctypes.POINTER(vp), pp, sp, vp]),
        "adds_last_status": (i32, [vp, sp]),
        "adds_solve_graph_create": (i32, [vp, vp, u32, vp, pp, ctypes.POINTER(vp)]),
        "adds_solve_graph_create_batched": (i32, [ctypes.POINTER(vp), c_u32p, u32, ctypes.POINTER(vp), pp,
                                                  ctypes.POINTER(vp)]),
        "adds_solve_graph_launch": (i32, [vp, vp]),
        "adds_solve_graph_status": (i32, [vp, sp]),
        "adds_solve_graph_destroy": (None, [vp]),
        "adds_sssp_backward_f32": (i32, [vp, c_u32p, u32, vp, vp, vp, ctypes.POINTER(ctypes.c_double), vp]),
        "adds_autotune": (i32, [vp, c_u32p, u32, u32, pp, ctypes.POINTER(ctypes.c_double)]),
        "adds_autotune_batched": (i32, [vp, c_u32p, u32, u32, pp, ctypes.POINTER(ctypes.c_double)]),
        "adds_litmus": (i32, [u32, u32, u32, ctypes.POINTER(u64), ctypes.POINTER(u64), ctypes.POINTER(u64)]),
        "adds_xcd_map": (i32, [u32, u32, c_u32p]),
        "adds_gather_ceiling": (i32, [u32, u64, ctypes.POINTER(ctypes.c_double)]),
        "adds_host_csr_free": (None, [hp]),
        "adds_read_gr": (i32, [ctypes.c_char_p, i32, hp]),
        "adds_read_mtx": (i32, [ctypes.c_char_p, i32, hp]),
        "adds_read_dimacs": (i32, [ctypes.c_char_p, i32, hp]),
        "adds_write_gr": (i32, [ctypes.c_char_p, hp]),
        "adds_write_dist_text": (i32, [ctypes.c_char_p, c_u32p, u32, i32]),
        "adds_gen_rmat": (i32, [u32, u32, ctypes.c_double, ctypes.c_double, ctypes.c_double, u64, i32,
                                ctypes.c_double, ctypes.c_double, hp]),
        "adds_gen_grid": (i32, [u32, u32, u64, u64, i32, ctypes.c_double, ctypes.c_double, hp]),
        "adds_gen_er": (i32, [u32, ctypes.c_double, u64, i32, ctypes.c_double, ctypes.c_double, hp]),
        "adds_gen_weights": (i32, [hp, u64, i32, ctypes.c_double, ctypes.c_double]),
        "adds_max_degree_vertex": (u32, [hp]),
        "adds_sample_component": (u32, [hp, u32, u32, u64, c_u32p]),
        "adds_reached_counts": (None, [hp, c_u32p, ctypes.POINTER(u64), ctypes.POINTER(u64)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = L
    return L


def check(status, what):
    if status != OK:
        raise AddsError(status, what)
