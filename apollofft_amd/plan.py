"""Device NUFFT plans over the C ABI (torch tensors for device memory/streams).

``NufftPlan1D`` / ``NufftPlan3D`` mirror apollo-nufft's reusable plans
(crates/apollo-nufft/src/application/execution/transform/dimension_1d.rs:102-276,
dimension_3d.rs:163-602): construction from the uniform domain, oversampling
sigma and KB half-width, then ``type1`` / ``type2`` executions.  Inputs and
outputs are torch tensors on the plan's GPU (complex64, float32 coordinates);
work is queued on torch's current stream.  The workspace is a torch uint8
tensor owned by the plan and grown on demand.
"""
import ctypes
import math

from . import _native as N


def _meta_dict(m):
    return {
        "dim": m.dim,
        "n": tuple(m.n),
        "m": tuple(m.m),
        "length": tuple(m.length),
        "kernel_width": m.kernel_width,
        "sigma": m.sigma,
        "beta": m.beta,
        "i0_beta": m.i0_beta,
        "grid_policy": m.grid_policy,
        "compat_reference_type2_3d_scaling": bool(m.compat_reference_type2_3d_scaling),
        "bin_size": tuple(m.bin_size),
        "max_subproblem": m.max_subproblem,
        "interp_bin_size": tuple(m.interp_bin_size),
        "interp_max_subproblem": m.interp_max_subproblem,
        "poly_degree": m.poly_degree,
        "poly_max_err": m.poly_max_err,
        "kernel_method": {1: "valu", 2: "mfma"}.get(m.kernel_method, m.kernel_method),
        "sub_work": m.sub_work,
        "blocks_per_cu": m.blocks_per_cu,
        "fft_lines": m.fft_lines,
    }


def _policy(p):
    if p in (None, "reference", N.GRID_REFERENCE):
        return N.GRID_REFERENCE
    if p in ("exact", "smooth", N.GRID_EXACT):
        return N.GRID_EXACT
    raise ValueError(f"unknown grid_policy {p!r}")


def _check_tensor(t, name, dtype, device, shape=None, numel=None):
    """Raise ValueError unless t is a contiguous tensor of `dtype` on `device`
    with the given shape or element count (the kernels take raw pointers)."""
    import torch
    if not isinstance(t, torch.Tensor):
        raise ValueError(f"{name} must be a torch tensor")
    if t.dtype != dtype:
        raise ValueError(f"{name} must be {dtype} (got {t.dtype})")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    if t.device != device:
        raise ValueError(f"{name} is on {t.device}, the positions on {device}")
    if shape is not None and tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name} must have shape {tuple(shape)} (got {tuple(t.shape)})")
    if numel is not None and t.numel() != numel:
        raise ValueError(f"{name} must have {numel} elements (got {t.numel()})")


_TUNED = ("kernel_method", "max_subproblem", "sub_work", "blocks_per_cu", "fft_lines")
_METHODS = {None: N.SPREAD_AUTO, "auto": N.SPREAD_AUTO, "valu": N.SPREAD_VALU,
            "mfma": N.SPREAD_MFMA}


def _make_opts(grid_policy="reference", compat=False, device_id=None, bin_size=None,
               max_subproblem=None, kernel_method=None, sub_work=None, blocks_per_cu=None,
               fft_lines=None):
    o = N.default_opts()
    o.kernel_method = _METHODS[kernel_method]
    o.grid_policy = _policy(grid_policy)
    o.compat_reference_type2_3d_scaling = int(bool(compat))
    o.device_id = -1 if device_id is None else int(device_id)
    if bin_size is not None:
        b = (bin_size,) * 3 if isinstance(bin_size, int) else tuple(bin_size)
        for a in range(3):
            o.bin_size[a] = int(b[a])
    if max_subproblem:
        o.max_subproblem = int(max_subproblem)
    if sub_work:
        o.sub_work = int(sub_work)
    if blocks_per_cu:
        o.blocks_per_cu = int(blocks_per_cu)
    if fft_lines:
        o.fft_lines = int(fft_lines)
    return o


def autotune3d(nx, ny, nz, dx, dy, dz, sigma=2, kernel_width=6, num_points=1_000_000,
               positions=None, workload="both", cache=None, grid_policy="reference",
               kernel_method=None, device_id=None):
    """Launch settings of a 3D plan picked by timing (anufft_autotune3d).

    positions: optional (M, 3) float32 device tensor of the points to tune
    on (num_points is then M); otherwise num_points uniform synthetic
    points.  workload: "type1", "type2" or "both".  cache: path of the JSON
    cache (looked up first, written after tuning).  Returns a dict of the
    settings (pass it as ``tuned=`` to NufftPlan3D) and "ms", the time of
    one workload with them."""
    o = _make_opts(grid_policy, False, device_id, None, None, kernel_method)
    coords, stride = None, 0
    if positions is not None:
        if positions.dtype != _torch().float32 or positions.dim() != 2 or positions.shape[1] != 3:
            raise ValueError("positions must be an (M, 3) float32 tensor")
        if not positions.is_contiguous() or positions.device.type != "cuda":
            raise ValueError("positions must be a contiguous device tensor")
        num_points = positions.shape[0]
        b = positions.data_ptr()
        coords, stride = N.ptr_array([b, b + 4, b + 8]), 3
    wl = {"type1": 1, "type2": 2, "both": 3}[workload]
    out = N.Opts()
    ms = ctypes.c_double()
    N.check(N.lib().anufft_autotune3d(int(nx), int(ny), int(nz), float(dx), float(dy), float(dz),
                                      int(sigma), int(kernel_width), ctypes.byref(o),
                                      int(num_points), coords, int(stride), wl,
                                      None if cache is None else str(cache).encode(),
                                      ctypes.byref(out), ctypes.byref(ms)), "autotune3d")
    res = {k: int(getattr(out, k)) for k in _TUNED}
    res["ms"] = ms.value
    return res


def _torch():
    import torch
    return torch


class _PlanBase:
    dim = 0

    def __init__(self):
        self._h = None
        self._ws = None
        self._ws_cap = -1

    def _opts(self, grid_policy, compat, device_id, bin_size, max_subproblem, kernel_method=None,
              sub_work=None, blocks_per_cu=None, fft_lines=None, tuned=None):
        o = _make_opts(grid_policy, compat, device_id, bin_size, max_subproblem, kernel_method,
                       sub_work, blocks_per_cu, fft_lines)
        if tuned:
            for k in _TUNED:
                setattr(o, k, int(tuned[k]))
        return o

    @property
    def handle(self):
        return self._h

    @property
    def meta(self):
        m = N.Meta()
        N.check(N.lib().anufft_plan_get_meta(self._h, ctypes.byref(m)), "get_meta")
        return _meta_dict(m)

    def deconv(self, axis=0):
        """FP64 deconvolution factors of one axis (kaiser_bessel.rs:137-148)."""
        import numpy as np
        n = self.meta["n"][axis]
        out = np.empty(n, dtype=np.float64)
        N.check(N.lib().anufft_plan_get_deconv(self._h, axis, out.ctypes.data), "get_deconv")
        return out

    def workspace_size(self, num_points):
        b = ctypes.c_size_t()
        N.check(N.lib().anufft_workspace_size(self._h, int(num_points), ctypes.byref(b)),
                "workspace_size")
        return b.value

    def workspace(self, num_points, device, copies=1):
        import torch
        need = copies * self.workspace_size(num_points)
        if self._ws is None or self._ws.numel() < need or self._ws.device != device:
            self._ws = torch.empty(max(need, 256), dtype=torch.uint8, device=device)
        return self._ws

    def timing(self, on=True):
        """Enable/disable per-stage HIP-event timing (clears records)."""
        N.check(N.lib().anufft_timing_enable(self._h, int(bool(on))), "timing_enable")

    def timing_collect(self):
        """{stage: (total_ms, calls)} of the recorded executions; clears them."""
        ms = (ctypes.c_double * len(N.STAGES))()
        calls = (ctypes.c_int64 * len(N.STAGES))()
        N.check(N.lib().anufft_timing_collect(self._h, ms, calls), "timing_collect")
        return {name: (ms[i], calls[i]) for i, name in enumerate(N.STAGES)}

    def close(self):
        if self._h:
            N.lib().anufft_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- device execution ------------------------------------------------------
    def _coords(self, positions):
        """(ptr_x, ptr_y, ptr_z, stride, M, device) for (M, dim) or tuple inputs."""
        import torch
        if isinstance(positions, (tuple, list)):
            cs = list(positions)
            if len(cs) != self.dim:
                raise ValueError(f"expected {self.dim} coordinate arrays")
            for c in cs:
                if c.dtype != torch.float32 or not c.is_contiguous() or c.dim() != 1:
                    raise ValueError("coordinate arrays must be contiguous 1-D float32 tensors")
            M = cs[0].numel()
            if any(c.numel() != M for c in cs):
                raise ValueError("coordinate arrays differ in length")
            if any(c.device != cs[0].device for c in cs) or cs[0].device.type != "cuda":
                raise ValueError("coordinate arrays must be on one GPU")
            ptrs = [c.data_ptr() for c in cs] + [cs[0].data_ptr()] * (3 - len(cs))
            return ptrs, 1, M, cs[0].device
        t = positions
        if not isinstance(t, torch.Tensor) or t.dtype != torch.float32 or not t.is_contiguous():
            raise ValueError("positions must be a contiguous float32 tensor")
        if t.device.type != "cuda":
            raise ValueError("positions must be on a GPU")
        if self.dim == 1:
            if t.dim() != 1:
                raise ValueError("1D positions must have shape (M,)")
            return [t.data_ptr()] * 3, 1, t.numel(), t.device
        if t.dim() != 2 or t.shape[1] != 3:
            raise ValueError("3D positions must have shape (M, 3)")
        b = t.data_ptr()
        return [b, b + 4, b + 8], 3, t.shape[0], t.device

    @staticmethod
    def _stream(device):
        import torch
        return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)

    def type1(self, positions, values, out=None):
        """Type-1: non-uniform samples -> modes (dimension_3d.rs:256-330)."""
        import torch
        ptrs, stride, M, dev = self._coords(positions)
        _check_tensor(values, "values", torch.complex64, dev, numel=M)
        shape = self.modes_shape
        if out is None:
            out = torch.empty(shape, dtype=torch.complex64, device=dev)
        else:
            _check_tensor(out, "out", torch.complex64, dev, shape=shape)
        ws = self.workspace(M, dev)
        N.check(N.lib().anufft_execute_type1(
            self._h, M, N.ptr_array(ptrs), stride, values.data_ptr(), out.data_ptr(),
            ws.data_ptr(), ws.numel(), self._stream(dev)), "execute_type1")
        return out

    def type2(self, positions, modes, out=None):
        """Type-2: modes -> non-uniform samples (dimension_3d.rs:496-602)."""
        import torch
        ptrs, stride, M, dev = self._coords(positions)
        shape = self.modes_shape
        _check_tensor(modes, "modes", torch.complex64, dev, shape=shape)
        if out is None:
            out = torch.empty(M, dtype=torch.complex64, device=dev)
        else:
            _check_tensor(out, "out", torch.complex64, dev, numel=M)
        ws = self.workspace(M, dev)
        N.check(N.lib().anufft_execute_type2(
            self._h, M, N.ptr_array(ptrs), stride, modes.data_ptr(), out.data_ptr(),
            ws.data_ptr(), ws.numel(), self._stream(dev)), "execute_type2")
        return out

    def direct_type1(self, positions, values):
        """Exact Type-1 sum in FP64 on the device -> complex128 modes (the
        WGPU direct backend, nufft.wgsl:92-117, computed there in f32)."""
        import torch
        ptrs, stride, M, dev = self._coords(positions)
        _check_tensor(values, "values", torch.complex64, dev, numel=M)
        out = torch.empty(self.modes_shape, dtype=torch.complex128, device=dev)
        N.check(N.lib().anufft_direct_type1(self._h, M, N.ptr_array(ptrs), stride,
                                            values.data_ptr(), out.data_ptr(), self._stream(dev)),
                "direct_type1")
        return out

    def direct_type1_select(self, positions, values, mode_index):
        """Exact Type-1 sum in FP64 at chosen flat mode indices (int64 tensor)
        -> complex128, one entry per index (anufft_direct_type1_select)."""
        import torch
        ptrs, stride, M, dev = self._coords(positions)
        if values.dtype == torch.complex64:
            values = values.to(torch.complex128)
        _check_tensor(values, "values", torch.complex128, dev, numel=M)
        _check_tensor(mode_index, "mode_index", torch.int64, dev)
        nmodes = math.prod(self.modes_shape)
        if mode_index.numel() and (int(mode_index.min()) < 0 or int(mode_index.max()) >= nmodes):
            raise ValueError(f"mode_index must lie in [0, {nmodes})")
        out = torch.empty(mode_index.numel(), dtype=torch.complex128, device=dev)
        N.check(N.lib().anufft_direct_type1_select(
            self._h, M, N.ptr_array(ptrs), stride, values.data_ptr(), mode_index.numel(),
            mode_index.data_ptr(), out.data_ptr(), self._stream(dev)), "direct_type1_select")
        return out

    def direct_type2(self, positions, modes):
        """Exact Type-2 sum in FP64 on the device -> complex128 per point
        (the WGPU direct backend, nufft.wgsl:119-140)."""
        import torch
        ptrs, stride, M, dev = self._coords(positions)
        _check_tensor(modes, "modes", torch.complex64, dev, shape=self.modes_shape)
        out = torch.empty(M, dtype=torch.complex128, device=dev)
        N.check(N.lib().anufft_direct_type2(self._h, M, N.ptr_array(ptrs), stride,
                                            modes.data_ptr(), out.data_ptr(), self._stream(dev)),
                "direct_type2")
        return out

    def grad_points(self, positions, u, modes_h):
        """grad[a][j] = Re(conj(u_j) d/dx_j^(a) sum_k h_k e^{+i k.x_j}) -> (M, dim)."""
        import torch
        ptrs, stride, M, dev = self._coords(positions)
        _check_tensor(u, "u", torch.complex64, dev, numel=M)
        _check_tensor(modes_h, "modes_h", torch.complex64, dev, shape=self.modes_shape)
        g = torch.empty((self.dim, M), dtype=torch.float32, device=dev)
        gp = [g[a].data_ptr() for a in range(self.dim)] + [None] * (3 - self.dim)
        ws = self.workspace(M, dev)
        N.check(N.lib().anufft_execute_grad_points(
            self._h, M, N.ptr_array(ptrs), stride, u.data_ptr(), modes_h.data_ptr(),
            N.ptr_array(gp, 3), ws.data_ptr(), ws.numel(), self._stream(dev)), "grad_points")
        return g.t()

    @property
    def fused_backward(self):
        """True when anufft_execute_type{1,2}_backward serve this plan (3D,
        MFMA gradient kernel: w <= 7)."""
        return (self.dim == 3 and self.meta["kernel_method"] == "mfma"
                and self.meta["kernel_width"] <= 7)

    def type2_backward(self, positions, g, modes):
        """Backward of c = T2(modes, x) for the upstream gradient g on c:
        (T1(g) at the points, the coordinate gradient (M, 3)), both from one
        binning of the points (anufft_execute_type2_backward)."""
        import torch
        ptrs, stride, M, dev = self._coords(positions)
        _check_tensor(g, "g", torch.complex64, dev, numel=M)
        _check_tensor(modes, "modes", torch.complex64, dev, shape=self.modes_shape)
        gm = torch.empty(self.modes_shape, dtype=torch.complex64, device=dev)
        gx = torch.empty((self.dim, M), dtype=torch.float32, device=dev)
        gp = [gx[a].data_ptr() for a in range(self.dim)] + [None] * (3 - self.dim)
        ws = self.workspace(M, dev, copies=2)
        N.check(N.lib().anufft_execute_type2_backward(
            self._h, M, N.ptr_array(ptrs), stride, g.data_ptr(), modes.data_ptr(), gm.data_ptr(),
            N.ptr_array(gp, 3), ws.data_ptr(), ws.numel(), self._stream(dev)), "type2_backward")
        return gm, gx.t()

    def type1_backward(self, positions, values, G):
        """Backward of F = T1(values, x) for the upstream gradient G on F:
        (T2(G) at the points, the coordinate gradient (M, 3)), from one FFT of
        G and one binning (anufft_execute_type1_backward)."""
        import torch
        ptrs, stride, M, dev = self._coords(positions)
        _check_tensor(values, "values", torch.complex64, dev, numel=M)
        _check_tensor(G, "G", torch.complex64, dev, shape=self.modes_shape)
        gv = torch.empty(M, dtype=torch.complex64, device=dev)
        gx = torch.empty((self.dim, M), dtype=torch.float32, device=dev)
        gp = [gx[a].data_ptr() for a in range(self.dim)] + [None] * (3 - self.dim)
        ws = self.workspace(M, dev)
        N.check(N.lib().anufft_execute_type1_backward(
            self._h, M, N.ptr_array(ptrs), stride, values.data_ptr(), G.data_ptr(), gv.data_ptr(),
            N.ptr_array(gp, 3), ws.data_ptr(), ws.numel(), self._stream(dev)), "type1_backward")
        return gv, gx.t()

    # ---- host drop-in execution (numpy; copies in and out) -------------------------
    def _host_coords(self, positions):
        import numpy as np
        pos = np.asarray(positions, dtype=np.float64)
        if self.dim == 1:
            pos = np.ascontiguousarray(pos.reshape(-1))
            M = pos.size
            base = pos.ctypes.data
            return pos, [base] * 3, 1, M
        pos = np.ascontiguousarray(pos.reshape(-1, 3))
        base = pos.ctypes.data
        return pos, [base, base + 8, base + 16], 3, pos.shape[0]

    def direct_type1_host(self, positions, values):
        """Exact Type-1 sum, FP64 coordinates and arithmetic (anufft_direct_type1_host):
        numpy float64 positions, complex128 values -> complex128 modes."""
        import numpy as np
        keep, ptrs, stride, M = self._host_coords(positions)
        vals = np.ascontiguousarray(values, dtype=np.complex128).reshape(-1)
        if vals.size != M:
            raise N.ShapeMismatch(f"positions/value length mismatch: {M} vs {vals.size}")
        out = np.empty(self.modes_shape, dtype=np.complex128)
        N.check(N.lib().anufft_direct_type1_host(self._h, M, N.ptr_array(ptrs), stride,
                                                 vals.ctypes.data, out.ctypes.data),
                "direct_type1_host")
        return out

    def direct_type2_host(self, positions, modes):
        """Exact Type-2 sum, FP64 coordinates and arithmetic (anufft_direct_type2_host):
        complex128 modes of the plan's shape -> complex128 per point."""
        import numpy as np
        keep, ptrs, stride, M = self._host_coords(positions)
        md = np.ascontiguousarray(modes, dtype=np.complex128)
        if md.size != math.prod(self.modes_shape):
            raise N.ShapeMismatch(f"modes shape mismatch: expected {self.modes_shape}")
        out = np.empty(M, dtype=np.complex128)
        N.check(N.lib().anufft_direct_type2_host(self._h, M, N.ptr_array(ptrs), stride,
                                                 md.ctypes.data, out.ctypes.data),
                "direct_type2_host")
        return out

    def type1_host(self, positions, values, storage=N.STORAGE_C128, profile=None):
        import numpy as np
        keep, ptrs, stride, M = self._host_coords(positions)
        dt = {N.STORAGE_C128: np.complex128, N.STORAGE_C64: np.complex64, N.STORAGE_F16X2: np.float16}[storage]
        vals = np.ascontiguousarray(values, dtype=dt)
        if storage == N.STORAGE_F16X2 and vals.shape[-1:] != (2,):
            raise ValueError("f16 storage takes (..., 2) half arrays")
        nvals = vals.size // 2 if storage == N.STORAGE_F16X2 else vals.size
        if nvals != M:
            raise N.ShapeMismatch(f"positions/value length mismatch: {M} vs {nvals}")
        shape = self.modes_shape + ((2,) if storage == N.STORAGE_F16X2 else ())
        out = np.empty(shape, dtype=dt)
        N.check(N.lib().anufft_execute_type1_host(
            self._h, storage, storage if profile is None else profile, M, N.ptr_array(ptrs),
            stride, vals.ctypes.data, out.ctypes.data), "execute_type1_host")
        return out

    def type2_host(self, positions, modes, storage=N.STORAGE_C128, profile=None):
        import numpy as np
        keep, ptrs, stride, M = self._host_coords(positions)
        dt = {N.STORAGE_C128: np.complex128, N.STORAGE_C64: np.complex64, N.STORAGE_F16X2: np.float16}[storage]
        md = np.ascontiguousarray(modes, dtype=dt)
        nmodes = md.size // 2 if storage == N.STORAGE_F16X2 else md.size
        if nmodes != math.prod(self.modes_shape):
            raise N.ShapeMismatch(f"modes shape mismatch: expected {self.modes_shape}")
        out = np.empty((M, 2) if storage == N.STORAGE_F16X2 else M, dtype=dt)
        N.check(N.lib().anufft_execute_type2_host(
            self._h, storage, storage if profile is None else profile, M, N.ptr_array(ptrs),
            stride, md.ctypes.data, out.ctypes.data), "execute_type2_host")
        return out


class NufftPlan1D(_PlanBase):
    """1D plan: NufftPlan1D::new(UniformDomain1D(n, dx), sigma, kernel_width)
    (dimension_1d.rs:127-154)."""

    dim = 1

    def __init__(self, n, dx, sigma=2, kernel_width=6, grid_policy="reference",
                 device_id=None):
        super().__init__()
        h = ctypes.c_void_p()
        o = self._opts(grid_policy, False, device_id, None, None)
        N.check(N.lib().anufft_plan1d_create(int(n), float(dx), int(sigma), int(kernel_width),
                                             ctypes.byref(o), ctypes.byref(h)), "plan1d_create")
        self._h = h
        self.modes_shape = (int(n),)


class NufftPlan3D(_PlanBase):
    """3D plan: NufftPlan3D::new(UniformGrid3D(nx,ny,nz,dx,dy,dz), sigma, w)
    (dimension_3d.rs:193-232)."""

    dim = 3

    def __init__(self, nx, ny, nz, dx, dy, dz, sigma=2, kernel_width=6,
                 grid_policy="reference", compat_reference_type2_3d_scaling=False,
                 device_id=None, bin_size=None, max_subproblem=None, kernel_method=None,
                 sub_work=None, blocks_per_cu=None, fft_lines=None, tuned=None):
        super().__init__()
        h = ctypes.c_void_p()
        o = self._opts(grid_policy, compat_reference_type2_3d_scaling, device_id, bin_size,
                       max_subproblem, kernel_method, sub_work, blocks_per_cu, fft_lines, tuned)
        N.check(N.lib().anufft_plan3d_create(int(nx), int(ny), int(nz), float(dx), float(dy),
                                             float(dz), int(sigma), int(kernel_width),
                                             ctypes.byref(o), ctypes.byref(h)), "plan3d_create")
        self._h = h
        self.modes_shape = (int(nx), int(ny), int(nz))


_BATCH_WS = {}


def execute_type1_batched(plans, positions, values, outs=None):
    """Batched Type-1 over independent 3D plans on torch's current stream.
    The workspace is cached per device and grown on demand."""
    import torch
    n = len(plans)
    if not (len(positions) == len(values) == n):
        raise ValueError("plans, positions and values differ in length")
    coords, vals, Ms, dev = [], [], [], None
    for i, (p, pos, v) in enumerate(zip(plans, positions, values)):
        if p.dim != 3:
            raise ValueError("batched execution takes 3D plans")
        ptrs, stride, M, d = p._coords(pos)
        if stride != 3:
            raise ValueError("batched execution takes (M, 3) position tensors")
        if dev is not None and d != dev:
            raise ValueError("all problems of a batch must be on one GPU")
        dev = d
        _check_tensor(v, f"values[{i}]", torch.complex64, dev, numel=M)
        coords += ptrs
        vals.append(v.data_ptr())
        Ms.append(M)
    if n == 0:
        return [] if outs is None else outs
    if outs is None:
        outs = [torch.empty(p.modes_shape, dtype=torch.complex64, device=dev) for p in plans]
    elif len(outs) != n:
        raise ValueError("outs and plans differ in length")
    else:
        for i, (p, o) in enumerate(zip(plans, outs)):
            _check_tensor(o, f"outs[{i}]", torch.complex64, dev, shape=p.modes_shape)
    hs = N.ptr_array([p.handle.value for p in plans])
    Marr = (ctypes.c_int64 * n)(*Ms)
    b = ctypes.c_size_t()
    N.check(N.lib().anufft_batch_workspace_size(hs, n, Marr, ctypes.byref(b)), "batch_workspace_size")
    ws = _BATCH_WS.get(dev)
    if ws is None or ws.numel() < b.value:
        ws = torch.empty(max(b.value, 256), dtype=torch.uint8, device=dev)
        _BATCH_WS[dev] = ws
    N.check(N.lib().anufft_execute_type1_batched(
        hs, n, Marr, N.ptr_array(coords), 3, N.ptr_array(vals),
        N.ptr_array([o.data_ptr() for o in outs]), ws.data_ptr(), ws.numel(),
        _PlanBase._stream(dev)), "execute_type1_batched")
    return outs
