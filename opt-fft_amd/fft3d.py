"""Python host side of the MI355X batched 3D FFT: ctypes bindings of
``lib/libfft3d_mi355x.so`` and a numpy-facing mirror of the reference's transform interface.

Mirrored interface (python/fft3d.py:360-423): ``fftn`` / ``ifftn`` over the last three
axes, ``fft3d`` / ``ifft3d`` for (B, n0, n1, n2) batches, the same sign and normalisation
conventions (forward unnormalised, ``ifftn`` carries 1/V, python/slow_dft.py:11-18) and the
same guard rails for bad input (ValueError).  Two deliberate differences:

* only the geometries the GPU kernels serve are accepted (``supports``); anything else
  raises ValueError instead of falling back to a slower route -- there is no CPU path;
* complex64 input is transformed in single precision on the GPU (the reference's Python
  module upcasts to double internally); complex128 stays complex128.

Every call goes through the C ABI of the shared library; nothing here computes a transform.
The library must be built (``make -C opt-fft_amd``); a missing library raises.
"""
import ctypes
import os

import numpy as np

from . import hip

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libfft3d_mi355x.so")

FORWARD, BACKWARD = -1, +1
C128, C64 = 0, 1
NORMALIZE, NO_TUNE, MFMA_BF16, MFMA_FP16 = 1, 2, 4, 8

# Every symbol the three public headers declare (include/*.h).
EXPORTED = [
    "fft3d_name", "fft3d_description", "fft3d_supports", "fft3d_create", "fft3d_execute",
    "fft3d_destroy",
    "fft3d_gpu_name", "fft3d_gpu_description", "fft3d_gpu_supports", "fft3d_gpu_create",
    "fft3d_gpu_execute", "fft3d_gpu_destroy",
    "fft3d_mx_last_error", "fft3d_mx_clear_error", "fft3d_mx_supports", "fft3d_mx_create",
    "fft3d_mx_create_named", "fft3d_mx_execute", "fft3d_mx_execute_n",
    "fft3d_mx_execute_chain", "fft3d_mx_destroy",
    "fft3d_mx_plan_get_info", "fft3d_mx_kernel_count", "fft3d_mx_kernel_name",
    "fft3d_mx_bank_image", "fft3d_mx_volume_bytes", "fft3d_mx_execute_noalloc",
    "fft3d_mx_multi_create", "fft3d_mx_multi_execute", "fft3d_mx_multi_get_info",
    "fft3d_mx_multi_destroy",
    "fft3d_srv_create", "fft3d_srv_submit", "fft3d_srv_wait", "fft3d_srv_run_trace",
    "fft3d_srv_now", "fft3d_srv_get_stats", "fft3d_srv_destroy",
]


class SrvConfig(ctypes.Structure):
    _fields_ = [("budget_bytes", ctypes.c_size_t), ("streams", ctypes.c_int),
                ("batched", ctypes.c_int), ("max_items", ctypes.c_int),
                ("window_us", ctypes.c_double)]


class SrvRequest(ctypes.Structure):
    _fields_ = [("n0", ctypes.c_int), ("n1", ctypes.c_int), ("n2", ctypes.c_int),
                ("batch", ctypes.c_int), ("direction", ctypes.c_int),
                ("precision", ctypes.c_int), ("inp", ctypes.c_void_p),
                ("out", ctypes.c_void_p)]


class SrvStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_long), ("completed", ctypes.c_long),
                ("launches", ctypes.c_long), ("mixed_launches", ctypes.c_long),
                ("plan_launches", ctypes.c_long),
                ("mean_items_per_mixed_launch", ctypes.c_double),
                ("peak_inflight_bytes", ctypes.c_size_t), ("busy_seconds", ctypes.c_double)]


class Item(ctypes.Structure):
    _fields_ = [("n0", ctypes.c_int), ("n1", ctypes.c_int), ("n2", ctypes.c_int),
                ("batch", ctypes.c_int), ("direction", ctypes.c_int)]


class PlanInfo(ctypes.Structure):
    _fields_ = [("n0", ctypes.c_int), ("n1", ctypes.c_int), ("n2", ctypes.c_int),
                ("batch", ctypes.c_int), ("direction", ctypes.c_int),
                ("precision", ctypes.c_int), ("kind", ctypes.c_int),
                ("volumes_per_wg", ctypes.c_int), ("threads", ctypes.c_int),
                ("variant", ctypes.c_int), ("lds_bytes", ctypes.c_size_t),
                ("tune_us", ctypes.c_double), ("kernel", ctypes.c_char * 96),
                ("small_max", ctypes.c_long), ("kernel_small", ctypes.c_char * 96)]


_lib = None


def load():
    """Load the shared library (raises OSError if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"{LIB_PATH} is missing: build it with `make -C opt-fft_amd`")
        L = ctypes.CDLL(LIB_PATH)
        vp, ci, cu, cl, cs = (ctypes.c_void_p, ctypes.c_int, ctypes.c_uint, ctypes.c_long,
                              ctypes.c_size_t)
        ip = ctypes.POINTER(ci)
        L.fft3d_name.restype = ctypes.c_char_p
        L.fft3d_description.restype = ctypes.c_char_p
        L.fft3d_supports.argtypes = [ci]
        L.fft3d_create.argtypes = [ci, ci]
        L.fft3d_create.restype = vp
        L.fft3d_execute.argtypes = [vp, vp, vp]
        L.fft3d_destroy.argtypes = [vp]
        L.fft3d_gpu_name.restype = ctypes.c_char_p
        L.fft3d_gpu_description.restype = ctypes.c_char_p
        L.fft3d_gpu_supports.argtypes = [ci]
        L.fft3d_gpu_create.argtypes = [ci, ci]
        L.fft3d_gpu_create.restype = vp
        L.fft3d_gpu_execute.argtypes = [vp, vp, vp]
        L.fft3d_gpu_destroy.argtypes = [vp]
        L.fft3d_mx_last_error.restype = ctypes.c_char_p
        L.fft3d_mx_supports.argtypes = [ci, ci, ci, ci]
        L.fft3d_mx_create.argtypes = [ci, ci, ci, ci, ci, ci, cu]
        L.fft3d_mx_create.restype = vp
        L.fft3d_mx_create_named.argtypes = [ci, ci, ci, ci, ci, ci, cu, ctypes.c_char_p]
        L.fft3d_mx_create_named.restype = vp
        L.fft3d_mx_execute.argtypes = [vp, vp, vp, vp]
        L.fft3d_mx_execute_n.argtypes = [vp, vp, vp, cl, vp]
        L.fft3d_mx_execute_chain.argtypes = [vp, vp, vp, vp, ci, ctypes.c_double, vp]
        L.fft3d_mx_destroy.argtypes = [vp]
        L.fft3d_mx_plan_get_info.argtypes = [vp, ctypes.POINTER(PlanInfo)]
        L.fft3d_mx_kernel_name.argtypes = [ci, ip, ip, ip, ip, ip]
        L.fft3d_mx_kernel_name.restype = ctypes.c_char_p
        L.fft3d_mx_bank_image.argtypes = [ci, vp, cs, ctypes.POINTER(cs)]
        L.fft3d_mx_execute_noalloc.argtypes = [ci, ci, ci, cl, ci, ci, cu, vp, vp, vp]
        L.fft3d_mx_volume_bytes.argtypes = [ci, ci, ci, ci]
        L.fft3d_mx_volume_bytes.restype = cs
        L.fft3d_mx_multi_create.argtypes = [ctypes.POINTER(Item), ci, ci, cu]
        L.fft3d_mx_multi_create.restype = vp
        L.fft3d_mx_multi_execute.argtypes = [vp, ctypes.POINTER(vp), ctypes.POINTER(vp), vp]
        L.fft3d_mx_multi_get_info.argtypes = [vp, ctypes.c_char_p, cs, ip, ip]
        L.fft3d_mx_multi_destroy.argtypes = [vp]
        L.fft3d_srv_create.argtypes = [ctypes.POINTER(SrvConfig)]
        L.fft3d_srv_create.restype = vp
        L.fft3d_srv_submit.argtypes = [vp, ctypes.POINTER(SrvRequest), ctypes.c_double]
        L.fft3d_srv_submit.restype = cl
        L.fft3d_srv_wait.argtypes = [vp, cl, ctypes.POINTER(ctypes.c_double)]
        L.fft3d_srv_run_trace.argtypes = [vp, ctypes.POINTER(SrvRequest),
                                          ctypes.POINTER(ctypes.c_double), ci,
                                          ctypes.POINTER(ctypes.c_double)]
        L.fft3d_srv_now.restype = ctypes.c_double
        L.fft3d_srv_get_stats.argtypes = [vp, ctypes.POINTER(SrvStats)]
        L.fft3d_srv_destroy.argtypes = [vp]
        _lib = L
    return _lib


def last_error():
    return load().fft3d_mx_last_error().decode()


def kernels():
    """[(name, (n0,n1,n2), precision, direction)] for every compiled kernel variant."""
    L = load()
    out = []
    a = [ctypes.c_int() for _ in range(5)]
    for i in range(L.fft3d_mx_kernel_count()):
        name = L.fft3d_mx_kernel_name(i, *[ctypes.byref(x) for x in a])
        out.append((name.decode(), (a[0].value, a[1].value, a[2].value), a[3].value,
                    a[4].value))
    return out


def supports(shape, precision=C128):
    n0, n1, n2 = (int(n) for n in shape)
    return bool(load().fft3d_mx_supports(n0, n1, n2, precision))


def bank_image(precision=C128):
    """The plan-time constant bank as a flat float array (see fft3d_mx_bank_image)."""
    L = load()
    size = ctypes.c_size_t(0)
    if L.fft3d_mx_bank_image(precision, None, 0, ctypes.byref(size)) != 0:
        raise ValueError(last_error())
    buf = np.empty(size.value, dtype=np.uint8)
    if L.fft3d_mx_bank_image(precision, buf.ctypes.data, buf.nbytes, ctypes.byref(size)) != 0:
        raise ValueError(last_error())
    return buf.view(np.float64 if precision == C128 else np.float32)


def _prec_of(dtype):
    dtype = np.dtype(dtype)
    if dtype == np.complex128:
        return C128
    if dtype == np.complex64:
        return C64
    raise ValueError(f"dtype {dtype} is not complex64/complex128")


class Plan:
    """One fft3d_mx plan: (n0,n1,n2), batch, direction, precision, flags."""

    def __init__(self, shape, batch, direction=FORWARD, precision=C128, flags=0, kernel=None):
        L = load()
        self.shape = tuple(int(n) for n in shape)
        self.batch, self.direction, self.precision = int(batch), int(direction), int(precision)
        self.flags = int(flags)
        if kernel is None:
            h = L.fft3d_mx_create(*self.shape, self.batch, self.direction, self.precision,
                                  self.flags)
        else:
            h = L.fft3d_mx_create_named(*self.shape, self.batch, self.direction,
                                        self.precision, self.flags, kernel.encode())
        if not h:
            raise ValueError(last_error())
        self.handle = ctypes.c_void_p(h)
        self.elem = 16 if self.precision == C128 else 8
        self.dtype = np.complex128 if self.precision == C128 else np.complex64
        self.volume = self.shape[0] * self.shape[1] * self.shape[2]

    def info(self):
        inf = PlanInfo()
        if load().fft3d_mx_plan_get_info(self.handle, ctypes.byref(inf)) != 0:
            raise ValueError(last_error())
        return inf

    @property
    def kernel(self):
        return self.info().kernel.decode()

    @property
    def kernel_small(self):
        """(kernel, max volumes) of the small-batch class, or None."""
        inf = self.info()
        return (inf.kernel_small.decode(), inf.small_max) if inf.small_max else None

    def execute(self, din, dout, batch=None, stream=None):
        """Enqueue on `stream` (hip.Stream, raw handle or None) with device addresses."""
        s = stream.handle if isinstance(stream, hip.Stream) else stream
        n = self.batch if batch is None else int(batch)
        rc = load().fft3d_mx_execute_n(self.handle, din, dout, n, s)
        if rc != 0:
            raise RuntimeError(last_error())

    def chain(self, din, dout, dpong, m, step_scale, stream=None):
        """m chained transforms (each scaled by step_scale); the end state lands in dout."""
        s = stream.handle if isinstance(stream, hip.Stream) else stream
        rc = load().fft3d_mx_execute_chain(self.handle, din, dout, dpong, int(m),
                                           float(step_scale), s)
        if rc != 0:
            raise RuntimeError(last_error())

    def run(self, x):
        """Host convenience: upload x (B, n0, n1, n2), transform, synchronize, download."""
        x = np.ascontiguousarray(x, dtype=self.dtype)
        if x.shape[1:] != self.shape or x.shape[0] > self.batch:
            raise ValueError(f"input shape {x.shape} does not fit plan {self.shape} x {self.batch}")
        nbytes = x.nbytes
        din, dout = hip.DeviceBuffer(nbytes), hip.DeviceBuffer(nbytes)
        try:
            din.upload(x)
            self.execute(din.addr, dout.addr, batch=x.shape[0])
            hip.synchronize()
            return dout.download(self.dtype, x.shape)
        finally:
            din.close()
            dout.close()

    def close(self):
        if getattr(self, "handle", None) is not None and self.handle.value:
            load().fft3d_mx_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def execute_noalloc(shape, batch, direction, precision, din, dout, stream=None, flags=0):
    """fft3d_mx_execute_noalloc: the transform of a plan, with no plan and no allocation."""
    s = stream.handle if isinstance(stream, hip.Stream) else stream
    rc = load().fft3d_mx_execute_noalloc(int(shape[0]), int(shape[1]), int(shape[2]),
                                         int(batch), int(direction), int(precision),
                                         int(flags), din, dout, s)
    if rc != 0:
        raise RuntimeError(last_error())


def _transform(x, direction, normalize):
    x = np.asarray(x)
    if x.ndim < 3:
        raise ValueError(f"need at least 3 dimensions, got shape {x.shape}")
    if np.iscomplexobj(x) and x.dtype == np.complex64:
        dtype = np.complex64
    else:
        dtype = np.complex128
    shape = x.shape[-3:]
    if min(shape) == 0:
        raise ValueError(f"cannot transform a length-0 axis: shape {x.shape}")
    lead = x.shape[:-3]
    batch = int(np.prod(lead)) if lead else 1
    if batch == 0:
        return np.zeros(x.shape, dtype=dtype)
    prec = _prec_of(dtype)
    if not supports(shape, prec):
        raise ValueError(f"no MI355X kernel for {shape} at {np.dtype(dtype).name}")
    flat = np.ascontiguousarray(x.reshape((batch,) + shape), dtype=dtype)
    plan = Plan(shape, batch, direction, prec, NORMALIZE if normalize else 0)
    try:
        y = plan.run(flat)
    finally:
        plan.close()
    return y.reshape(x.shape)


def fftn(x):
    """Forward, unnormalised 3D DFT over the last three axes (python/fft3d.py:360)."""
    return _transform(x, FORWARD, False)


def ifftn(x):
    """Backward 3D DFT over the last three axes with 1/V (python/fft3d.py:402)."""
    return _transform(x, BACKWARD, True)


def fft3d(x):
    """(B, n0, n1, n2) or (n0, n1, n2) forward transform."""
    x = np.asarray(x)
    if x.ndim not in (3, 4):
        raise ValueError(f"fft3d expects 3 or 4 dimensions, got {x.shape}")
    return fftn(x)


def ifft3d(x):
    x = np.asarray(x)
    if x.ndim not in (3, 4):
        raise ValueError(f"ifft3d expects 3 or 4 dimensions, got {x.shape}")
    return ifftn(x)


class MultiPlan:
    """One persistent launch over a mixed batch: items = [(shape, batch, direction), ...]."""

    def __init__(self, items, precision=C128, flags=0):
        L = load()
        self.items = [(tuple(int(n) for n in sh), int(b), int(d)) for sh, b, d in items]
        arr = (Item * len(self.items))(*[Item(*sh, b, d) for sh, b, d in self.items])
        h = L.fft3d_mx_multi_create(arr, len(self.items), int(precision), int(flags))
        if not h:
            raise ValueError(last_error())
        self.handle = ctypes.c_void_p(h)
        self.precision = int(precision)
        self.dtype = np.complex128 if precision == C128 else np.complex64

    def info(self):
        name = ctypes.create_string_buffer(96)
        nt, gr = ctypes.c_int(0), ctypes.c_int(0)
        if load().fft3d_mx_multi_get_info(self.handle, name, 96, ctypes.byref(nt),
                                          ctypes.byref(gr)) != 0:
            raise ValueError(last_error())
        return name.value.decode(), nt.value, gr.value

    def execute(self, ins, outs, stream=None):
        n = len(self.items)
        s = stream.handle if isinstance(stream, hip.Stream) else stream
        a = (ctypes.c_void_p * n)(*ins)
        b = (ctypes.c_void_p * n)(*outs)
        if load().fft3d_mx_multi_execute(self.handle, a, b, s) != 0:
            raise RuntimeError(last_error())

    def run(self, xs):
        """Host convenience: upload every item, one launch, download every result."""
        ins, outs = [], []
        try:
            for x in xs:
                x = np.ascontiguousarray(x, dtype=self.dtype)
                bi, bo = hip.DeviceBuffer(x.nbytes), hip.DeviceBuffer(x.nbytes)
                bi.upload(x)
                ins.append((bi, x.shape))
                outs.append(bo)
            self.execute([b.addr for b, _ in ins], [b.addr for b in outs])
            hip.synchronize()
            return [o.download(self.dtype, shp) for o, (_, shp) in zip(outs, ins)]
        finally:
            for b, _ in ins:
                b.close()
            for b in outs:
                b.close()

    def close(self):
        if getattr(self, "handle", None) is not None and self.handle.value:
            load().fft3d_mx_multi_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


SRV_PLAIN_SYNC, SRV_BATCHED, SRV_PLAIN_ASYNC = 0, 1, 2


class Server:
    """fft3d_srv: continuous batching (batched=True or SRV_BATCHED), the plain
    one-launch-per-request path waited for request by request (batched=False or
    SRV_PLAIN_SYNC), or the same with `streams` launches in flight (SRV_PLAIN_ASYNC)."""

    def __init__(self, budget_bytes=64 << 30, streams=8, batched=True, max_items=16,
                 window_us=0.0):
        mode = SRV_BATCHED if batched is True else SRV_PLAIN_SYNC if batched is False \
            else int(batched)
        cfg = SrvConfig(int(budget_bytes), int(streams), mode, int(max_items),
                        float(window_us))
        h = load().fft3d_srv_create(ctypes.byref(cfg))
        if not h:
            raise RuntimeError(last_error())
        self.handle = ctypes.c_void_p(h)

    def submit(self, shape, batch, direction, precision, din, dout=None, arrival=0.0):
        r = SrvRequest(*[int(n) for n in shape], int(batch), int(direction), int(precision),
                       din, dout)
        rid = load().fft3d_srv_submit(self.handle, ctypes.byref(r), float(arrival))
        if rid < 0:
            raise ValueError(last_error())
        return rid

    def wait(self, rid):
        lat = ctypes.c_double(0)
        if load().fft3d_srv_wait(self.handle, rid, ctypes.byref(lat)) != 0:
            raise RuntimeError(last_error())
        return lat.value

    def run_trace(self, requests, arrivals):
        """requests: list of SrvRequest; arrivals: seconds from the call.  Latencies (s)."""
        n = len(requests)
        reqs = (SrvRequest * n)(*requests)
        arr = (ctypes.c_double * n)(*[float(a) for a in arrivals])
        lat = (ctypes.c_double * n)()
        if load().fft3d_srv_run_trace(self.handle, reqs, arr, n, lat) != 0:
            raise RuntimeError(last_error())
        return np.frombuffer(lat, dtype=np.float64).copy()

    def stats(self):
        st = SrvStats()
        if load().fft3d_srv_get_stats(self.handle, ctypes.byref(st)) != 0:
            raise RuntimeError(last_error())
        return {k: getattr(st, k) for k, _ in SrvStats._fields_}

    def close(self):
        if getattr(self, "handle", None) is not None and self.handle.value:
            load().fft3d_srv_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
