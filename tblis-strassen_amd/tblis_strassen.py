"""ctypes binding of libtblis_strassen_amd.so (the C-ABI in include/tblis_strassen.h).

Mirrors the reference's operator interface for the Strassen tensor-contraction path:
    stra_mult(alpha, A, idx_A, B, idx_B, beta, C, idx_C)         src/stra_iface/3t/stra_mult.h:26-36
    stra_mult_2level(alpha, A, idx_A, B, idx_B, beta, C, idx_C)  src/stra_iface/3t/stra_mult_2level.h:26-36
    mult(alpha, A, idx_A, B, idx_B, beta, C, idx_C)              src/iface/3t/mult.h:26-36
    set_impl / get_impl                                         tblis::internal::impl, src/internal/3t/mult.cxx:26
with index strings as str, tensors as torch tensors (device or host) or numpy arrays (host).  Strides
are taken from the tensor object in elements, as tblis_tensor expects.  The HIP library must be built:
loading fails loudly otherwise (there is no CPU fallback).
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtblis_strassen_amd.so")

TYPE_FLOAT, TYPE_DOUBLE = 0, 1
IMPL = {"abc": 0, "blis": 0, "blas": 1, "ref": 2, "naive": 3, "ab": 4, "auto": -1}
PREC = {"native": 0, "bf16x3": 1, "ozaki_bf16": 2, "fp8_block": 3, "ozaki_int8": 4}
FLAG_REF_FRINGE = 1
FLAG_SYNC = 2
FLAG_GRAPH = 4
FLAG_FUSED_SUMS = 8
FLAG_OUTER_STAGED = 16
FLAG_ALL_STAGED = 32
FLAG_BATCH_GROUPED = 64
FLAG_STAGE_PERSIST = 128
FLAG_ATOMIC_C = 256


class MistraError(RuntimeError):
    pass


class TblisMatrix(ctypes.Structure):
    """Byte image of tblis_matrix (src/util/basic_types.h:264-272)."""

    _fields_ = [
        ("type", ctypes.c_int),
        ("conj", ctypes.c_int),
        ("scalar", ctypes.c_char * 16),
        ("data", ctypes.c_void_p),
        ("m", ctypes.c_ssize_t),
        ("n", ctypes.c_ssize_t),
        ("rs", ctypes.c_ssize_t),
        ("cs", ctypes.c_ssize_t),
    ]


class TblisTensor(ctypes.Structure):
    """Byte image of tblis_tensor (src/util/basic_types.h:341-419)."""

    _fields_ = [
        ("type", ctypes.c_int),
        ("conj", ctypes.c_int),
        ("scalar", ctypes.c_char * 16),
        ("data", ctypes.c_void_p),
        ("ndim", ctypes.c_uint),
        ("len", ctypes.POINTER(ctypes.c_ssize_t)),
        ("stride", ctypes.POINTER(ctypes.c_ssize_t)),
    ]


class Contraction(ctypes.Structure):
    _fields_ = [
        ("A", ctypes.POINTER(TblisTensor)), ("idx_A", ctypes.c_char_p),
        ("B", ctypes.POINTER(TblisTensor)), ("idx_B", ctypes.c_char_p),
        ("C", ctypes.POINTER(TblisTensor)), ("idx_C", ctypes.c_char_p),
        ("level", ctypes.c_int), ("impl", ctypes.c_int),
    ]


assert ctypes.sizeof(TblisTensor) == 56 and TblisTensor.data.offset == 24 and TblisTensor.len.offset == 40

_lib = None


def load(path=None):
    """Load the HIP library (idempotent).  Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    p = path or os.environ.get("MISTRA_LIB", LIB_PATH)
    if not os.path.exists(p):
        raise MistraError(f"{p} is missing: build it with `python tblis-strassen_amd/build.py`")
    lib = ctypes.CDLL(p, mode=ctypes.RTLD_GLOBAL)
    T = ctypes.POINTER(TblisTensor)
    s = ctypes.c_char_p
    for name in ("stra_tblis_tensor_mult", "stra_tblis_tensor_mult_2level", "tblis_tensor_mult"):
        f = getattr(lib, name)
        f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, T, s, T, s, T, s]
        f.restype = None
    lib.mistra_tensor_mult.argtypes = [T, s, T, s, T, s, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_uint, ctypes.c_void_p]
    lib.mistra_tensor_mult.restype = ctypes.c_int
    lib.mistra_last_error.restype = ctypes.c_char_p
    lib.mistra_set_impl.argtypes = [ctypes.c_int]
    lib.mistra_get_impl.restype = ctypes.c_int
    lib.mistra_kernel_timing.argtypes = [ctypes.c_int]
    lib.mistra_kernel_timing_read.argtypes = [ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_longlong),
                                              ctypes.POINTER(ctypes.c_double), ctypes.c_int]
    lib.mistra_fill_uniform.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_longlong, ctypes.c_ulonglong,
                                        ctypes.c_int, ctypes.c_void_p]
    lib.mistra_copy_bytes.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_longlong, ctypes.c_void_p]
    lib.mistra_rel_error.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_longlong,
                                     ctypes.POINTER(ctypes.c_double), ctypes.c_void_p]
    lib.mistra_batch_mult.argtypes = [ctypes.c_int, ctypes.POINTER(Contraction), ctypes.c_uint, ctypes.c_void_p]
    lib.mistra_tensor_mult_backward.argtypes = [T, s, T, s, T, s, T, T, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                ctypes.c_void_p]
    lib.mistra_autotune.argtypes = [T, s, T, s, T, s, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                    ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int), ctypes.c_void_p]
    _lib = lib
    return lib


def _check(rc):
    if rc != 0:
        raise MistraError(f"error {rc}: {_lib.mistra_last_error().decode()}")


def _dtype_code(x):
    name = str(x.dtype)
    if name.endswith("float64"):
        return TYPE_DOUBLE
    if name.endswith("float32"):
        return TYPE_FLOAT
    raise MistraError(f"unsupported dtype {x.dtype}")


def _shape_strides_ptr(x):
    """(shape, strides in elements, data pointer) of a torch tensor or numpy array."""
    if hasattr(x, "data_ptr"):  # torch
        return list(x.shape), list(x.stride()), x.data_ptr()
    import numpy as np

    assert isinstance(x, np.ndarray)
    item = x.dtype.itemsize
    return list(x.shape), [s // item for s in x.strides], x.ctypes.data


def make_tensor(x, scalar=1.0):
    """Build a tblis_tensor over x (keeps its len/stride arrays alive on the struct)."""
    shape, strides, ptr = _shape_strides_ptr(x)
    t = TblisTensor()
    t.type = _dtype_code(x)
    t.conj = 0
    buf = (ctypes.c_double(scalar) if t.type == TYPE_DOUBLE else ctypes.c_float(scalar))
    # (t.scalar would be a bytes copy: write through the struct's own address)
    ctypes.memmove(ctypes.addressof(t) + TblisTensor.scalar.offset, ctypes.addressof(buf), ctypes.sizeof(buf))
    t.data = ptr
    t.ndim = len(shape)
    t._len = (ctypes.c_ssize_t * max(1, len(shape)))(*shape)
    t._stride = (ctypes.c_ssize_t * max(1, len(strides)))(*strides)
    t.len = ctypes.cast(t._len, ctypes.POINTER(ctypes.c_ssize_t))
    t.stride = ctypes.cast(t._stride, ctypes.POINTER(ctypes.c_ssize_t))
    return t


def _stream_ptr(stream):
    if stream is None:
        return None
    if hasattr(stream, "cuda_stream"):
        return stream.cuda_stream
    return int(stream)


def tensor_mult(alpha, A, idx_A, B, idx_B, beta, C, idx_C, level=1, impl="abc", precision="native",
                stream=None, flags=0):
    """Extension entry point: explicit level / impl / precision / stream (asynchronous on the stream)."""
    lib = load()
    tA, tB, tC = make_tensor(A, alpha), make_tensor(B, 1.0), make_tensor(C, beta)
    impl_c = IMPL[impl] if isinstance(impl, str) else int(impl)
    prec_c = PREC[precision] if isinstance(precision, str) else int(precision)
    _check(lib.mistra_tensor_mult(ctypes.byref(tA), idx_A.encode(), ctypes.byref(tB), idx_B.encode(),
                                  ctypes.byref(tC), idx_C.encode(), level, impl_c, prec_c, flags,
                                  _stream_ptr(stream)))


def tensor_mult_noalloc(alpha, A, idx_A, B, idx_B, beta, C, idx_C, level=1, impl="abc", precision="native",
                        stream=None, flags=0):
    """mistra_tensor_mult_noalloc: the same contraction with no device allocation (tables and workspace in
    the library's fixed 64 MiB scratch block per GPU; device tensors, explicit level and impl)."""
    lib = load()
    fn = lib.mistra_tensor_mult_noalloc
    fn.argtypes = lib.mistra_tensor_mult.argtypes
    fn.restype = ctypes.c_int
    tA, tB, tC = make_tensor(A, alpha), make_tensor(B, 1.0), make_tensor(C, beta)
    impl_c = IMPL[impl] if isinstance(impl, str) else int(impl)
    prec_c = PREC[precision] if isinstance(precision, str) else int(precision)
    _check(fn(ctypes.byref(tA), idx_A.encode(), ctypes.byref(tB), idx_B.encode(), ctypes.byref(tC), idx_C.encode(),
              level, impl_c, prec_c, flags, _stream_ptr(stream)))


def tensor_mult_backward(alpha, A, idx_A, B, idx_B, dC, idx_C, dA=None, dB=None, beta=0.0, level=1, impl="abc",
                         precision="native", stream=None):
    """Adjoints of C = alpha*A.B: dA = alpha*dC.B + beta*dA and dB = alpha*A.dC + beta*dB over the same
    labels, each as a Strassen contraction (asynchronous on the stream).  dA or dB may be None."""
    lib = load()
    tA, tB, tC = make_tensor(A, alpha), make_tensor(B, 1.0), make_tensor(dC, 1.0)
    gA = make_tensor(dA, beta) if dA is not None else None
    gB = make_tensor(dB, beta) if dB is not None else None
    impl_c = IMPL[impl] if isinstance(impl, str) else int(impl)
    prec_c = PREC[precision] if isinstance(precision, str) else int(precision)
    _check(lib.mistra_tensor_mult_backward(ctypes.byref(tA), idx_A.encode(), ctypes.byref(tB), idx_B.encode(),
                                           ctypes.byref(tC), idx_C.encode(),
                                           ctypes.byref(gA) if gA is not None else None,
                                           ctypes.byref(gB) if gB is not None else None, level, impl_c, prec_c,
                                           _stream_ptr(stream)))


def autotune(A, idx_A, B, idx_B, C, idx_C, measure=False, stream=None):
    """(level, impl, tile) chosen for this contraction shape: the analytic model, or with measure=True
    every candidate timed on the device (device tensors; C is not modified).  Cached per shape; calls
    with level=-1 / impl="auto" use the cached choice."""
    lib = load()
    tA, tB, tC = make_tensor(A), make_tensor(B), make_tensor(C)
    lv, im, ti = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    _check(lib.mistra_autotune(ctypes.byref(tA), idx_A.encode(), ctypes.byref(tB), idx_B.encode(), ctypes.byref(tC),
                               idx_C.encode(), 1 if measure else 0, ctypes.byref(lv), ctypes.byref(im),
                               ctypes.byref(ti), _stream_ptr(stream)))
    return lv.value, im.value, ti.value


def autotune_candidates(A, idx_A, B, idx_B, C, idx_C):
    """[(level, impl, tile, model_ms)] the autotuner weighs for this shape, fastest modelled first
    (mistra_debug_tune_candidates; host-side model, no device work)."""
    lib = load()
    tA, tB, tC = make_tensor(A), make_tensor(B), make_tensor(C)
    cap = 64
    lv, im, ti = (ctypes.c_int * cap)(), (ctypes.c_int * cap)(), (ctypes.c_int * cap)()
    ms = (ctypes.c_double * cap)()
    n = lib.mistra_debug_tune_candidates(ctypes.byref(tA), idx_A.encode(), ctypes.byref(tB), idx_B.encode(),
                                         ctypes.byref(tC), idx_C.encode(), lv, im, ti, ms, cap)
    _check(min(n, 0))
    return [(lv[i], im[i], ti[i], ms[i]) for i in range(min(n, cap))]


class Batch:
    """Keeps the tblis_tensor images of a batch alive for as long as the batch is reused (graph replays
    key on the tensors' pointers, shapes, strides and scalars)."""

    def __init__(self, entries):
        self.tensors = []
        arr = (Contraction * max(1, len(entries)))()
        for i, e in enumerate(entries):
            tA = make_tensor(e["A"], e.get("alpha", 1.0))
            tB = make_tensor(e["B"], 1.0)
            tC = make_tensor(e["C"], e.get("beta", 0.0))
            self.tensors += [tA, tB, tC]
            arr[i].A, arr[i].B, arr[i].C = ctypes.pointer(tA), ctypes.pointer(tB), ctypes.pointer(tC)
            arr[i].idx_A, arr[i].idx_B, arr[i].idx_C = (e[k].encode() for k in ("idx_A", "idx_B", "idx_C"))
            impl = e.get("impl", "abc")
            arr[i].level = e.get("level", 1)
            arr[i].impl = IMPL[impl] if isinstance(impl, str) else int(impl)
        self.entries = entries
        self.arr = arr
        self.count = len(entries)

    def run(self, flags=0, stream=None):
        # C's scalar reads 1 after every call (as in tblis); restore each entry's beta first
        for i, e in enumerate(self.entries):
            t = self.tensors[3 * i + 2]
            beta = e.get("beta", 0.0)
            buf = (ctypes.c_double(beta) if t.type == TYPE_DOUBLE else ctypes.c_float(beta))
            ctypes.memmove(ctypes.addressof(t) + TblisTensor.scalar.offset, ctypes.addressof(buf), ctypes.sizeof(buf))
        _check(load().mistra_batch_mult(self.count, self.arr, flags, _stream_ptr(stream)))


def batch_mult(entries, flags=0, stream=None):
    """mistra_batch_mult over a list of dicts (A, idx_A, B, idx_B, C, idx_C[, alpha, beta, level, impl])."""
    b = Batch(entries)
    b.run(flags, stream)
    return b


def make_matrix(x, scalar=1.0):
    """tblis_matrix over a 2-d torch tensor or numpy array (strides in elements)."""
    shape, strides, ptr = _shape_strides_ptr(x)
    if len(shape) != 2:
        raise MistraError("a matrix needs two dimensions")
    t = TblisMatrix()
    t.type = _dtype_code(x)
    t.conj = 0
    buf = (ctypes.c_double(scalar) if t.type == TYPE_DOUBLE else ctypes.c_float(scalar))
    ctypes.memmove(ctypes.addressof(t) + TblisMatrix.scalar.offset, ctypes.addressof(buf), ctypes.sizeof(buf))
    t.data = ptr
    t.m, t.n = shape
    t.rs, t.cs = strides
    return t


def matrix_mult(alpha, A, B, beta, C, impl="abc"):
    """Drop-in stra_tblis_matrix_mult{,_naive,_ab}: C = alpha*A.B + beta*C, one Strassen level. Synchronous."""
    lib = load()
    symbol = {"abc": "stra_tblis_matrix_mult", "naive": "stra_tblis_matrix_mult_naive",
              "ab": "stra_tblis_matrix_mult_ab"}[impl]
    tA, tB, tC = make_matrix(A, alpha), make_matrix(B, 1.0), make_matrix(C, beta)
    fn = getattr(lib, symbol)
    fn.argtypes = [ctypes.c_void_p, ctypes.c_void_p] + [ctypes.POINTER(TblisMatrix)] * 3
    fn.restype = None
    fn(None, None, ctypes.byref(tA), ctypes.byref(tB), ctypes.byref(tC))
    msg = lib.mistra_last_error().decode()
    if msg:
        raise MistraError(msg)


def _dropin(symbol, alpha, A, idx_A, B, idx_B, beta, C, idx_C):
    lib = load()
    tA, tB, tC = make_tensor(A, alpha), make_tensor(B, 1.0), make_tensor(C, beta)
    getattr(lib, symbol)(None, None, ctypes.byref(tA), idx_A.encode(), ctypes.byref(tB), idx_B.encode(),
                         ctypes.byref(tC), idx_C.encode())
    msg = lib.mistra_last_error().decode()  # the void symbols report through the thread-local message
    if msg:
        raise MistraError(msg)


def stra_mult(alpha, A, idx_A, B, idx_B, beta, C, idx_C):
    """Drop-in stra_tblis_tensor_mult (1-level; impl from set_impl). Synchronous."""
    _dropin("stra_tblis_tensor_mult", alpha, A, idx_A, B, idx_B, beta, C, idx_C)


def stra_mult_2level(alpha, A, idx_A, B, idx_B, beta, C, idx_C):
    """Drop-in stra_tblis_tensor_mult_2level. Synchronous."""
    _dropin("stra_tblis_tensor_mult_2level", alpha, A, idx_A, B, idx_B, beta, C, idx_C)


def mult(alpha, A, idx_A, B, idx_B, beta, C, idx_C):
    """Drop-in tblis_tensor_mult (classical, L=0). Synchronous."""
    _dropin("tblis_tensor_mult", alpha, A, idx_A, B, idx_B, beta, C, idx_C)


def set_impl(impl):
    lib = load()
    _check(lib.mistra_set_impl(IMPL[impl] if isinstance(impl, str) else int(impl)))


def get_impl():
    return load().mistra_get_impl()


def kernel_timing(enable=True):
    _check(load().mistra_kernel_timing(1 if enable else 0))


def kernel_timing_read(reset=True):
    """(total ms, launches, flops) of the timed leaf-kernel launches since the last reset."""
    ms, n, fl = ctypes.c_double(), ctypes.c_longlong(), ctypes.c_double()
    _check(load().mistra_kernel_timing_read(ctypes.byref(ms), ctypes.byref(n), ctypes.byref(fl), 1 if reset else 0))
    return ms.value, n.value, fl.value


def fill_uniform(x, seed, emax=0, stream=None):
    """Device fill with the library's splitmix64 uniform[-1,1) generator (x: contiguous torch tensor)."""
    _check(load().mistra_fill_uniform(_dtype_code(x), x.data_ptr(), x.numel(), seed, emax, _stream_ptr(stream)))


def copy_bytes(dst, src, nbytes, stream=None):
    _check(load().mistra_copy_bytes(dst.data_ptr(), src.data_ptr(), nbytes, _stream_ptr(stream)))


def debug_set_lowp(bf16_slices=0, int8_slices=0, int8_fold_ksteps=0):
    """Test hook (include/tblis_strassen_debug.h): Ozaki digit counts and the INT8 fold period for the
    later calls of this process; 0 restores a default."""
    lib = load()
    lib.mistra_debug_set_lowp.argtypes = [ctypes.c_int] * 3
    rc = lib.mistra_debug_set_lowp(bf16_slices, int8_slices, int8_fold_ksteps)
    if rc != 0:
        raise MistraError(f"mistra_debug_set_lowp: error {rc}")


def debug_set_lowp_fusion(fused=True):
    """mistra_debug_set_lowp_fusion: 1-level low-precision operands form their sums and digits in one
    launch (True, the default) or stage the sums first (False); the digits are the same either way."""
    lib = load()
    lib.mistra_debug_set_lowp_fusion.argtypes = [ctypes.c_int]
    lib.mistra_debug_set_lowp_fusion(1 if fused else 0)


def debug_set_tile(tile=-1):
    """Register tile of later calls' plans (mistra_debug_set_tile): 0 128x128, 1 64x64, 2 128x64,
    3 64x128, -1 the plan's own choice."""
    lib = load()
    lib.mistra_debug_set_tile.restype = ctypes.c_int
    lib.mistra_debug_set_tile.argtypes = [ctypes.c_int]
    rc = lib.mistra_debug_set_tile(int(tile))
    if rc != 0:
        raise MistraError(f"mistra_debug_set_tile: error {rc}")


def debug_set_batch_wait(spin_limit=0, dep_bias=0):
    """Test hook (include/tblis_strassen_debug.h): polls before a wait of a persistent launch (batch, or
    a staged plan's products) gives up (0 = default) and a bias added to every wait's target (a positive
    one makes every wait unreachable, so each gives up)."""
    lib = load()
    lib.mistra_debug_set_batch_wait.argtypes = [ctypes.c_longlong, ctypes.c_uint]
    lib.mistra_debug_set_batch_wait(spin_limit, dep_bias)


def release_workspace():
    """mistra_release_workspace: drop the plan cache and the batch graphs (plans captured by a caller's
    graph stay resident)."""
    _check(load().mistra_release_workspace())


def rel_error(X, Y, stream=None):
    """||X-Y||_F / ||Y||_F over two contiguous device tensors of equal size (FP64 reduction on device)."""
    out = ctypes.c_double()
    _check(load().mistra_rel_error(_dtype_code(X), X.data_ptr(), Y.data_ptr(), X.numel(), ctypes.byref(out),
                                   _stream_ptr(stream)))
    return out.value
