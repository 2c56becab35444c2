"""ctypes binding of librcwa_amd.so (the C ABI declared in include/rcwa_amd.h).

The library is the product: every call below runs the HIP kernels of
csrc/ on the GPU.  There is no fallback: if the shared library is not built,
or no GPU is present, calls raise.
"""
import ctypes
import os
import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librcwa_amd.so")

RCWA_FP64 = 0
RCWA_SPLIT = 1
STATUS = {0: "OK", 1: "E_ARG", 2: "E_NOMEM", 3: "E_EIG_NOCONV", 4: "E_SINGULAR",
          5: "E_NONFINITE", 6: "E_HIP", 7: "E_NOTREADY", 8: "E_CANCELLED"}


class RcwaError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s: %s" % (STATUS.get(code, code), msg))
        self.code = code


class c128(ctypes.Structure):
    _fields_ = [("re", ctypes.c_double), ("im", ctypes.c_double)]


def _c(z):
    z = complex(z)
    return c128(z.real, z.imag)


class Device(ctypes.Structure):
    _fields_ = [("Nx", ctypes.c_int), ("Ny", ctypes.c_int), ("num_layers", ctypes.c_int),
                ("Lx", ctypes.c_double), ("Ly", ctypes.c_double),
                ("er", ctypes.c_void_p), ("ur", ctypes.c_void_p), ("t", ctypes.c_void_p),
                ("grids_on_device", ctypes.c_int)]


class Source(ctypes.Structure):
    _fields_ = [("lambda0", ctypes.c_double), ("theta", ctypes.c_double),
                ("phi", ctypes.c_double), ("pte", ctypes.c_double), ("ptm", ctypes.c_double)]


class Params(ctypes.Structure):
    _fields_ = [("Nx_harmonics", ctypes.c_int), ("Ny_harmonics", ctypes.c_int),
                ("er_ref", c128), ("ur_ref", c128), ("er_trn", c128), ("ur_trn", c128)]


class Results(ctypes.Structure):
    _fields_ = [("R", ctypes.c_void_p), ("T", ctypes.c_void_p),
                ("R_tot", ctypes.c_double), ("T_tot", ctypes.c_double),
                ("r", ctypes.c_void_p), ("t", ctypes.c_void_p), ("cond", ctypes.c_double)]


SHAPE_ELLIPSE, SHAPE_RECT, SHAPE_CIRCLE = 0, 1, 2


class Shape(ctypes.Structure):
    _fields_ = [("layer", ctypes.c_int), ("kind", ctypes.c_int), ("cx", ctypes.c_double),
                ("cy", ctypes.c_double), ("ax", ctypes.c_double), ("ay", ctypes.c_double),
                ("value", c128)]


class DeviceBuffer:
    """Device memory for tests and callers of the HBM-resident entry points:
    .ptr, upload(array), download(shape)."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        self.ptr = _hip_malloc(max(self.nbytes, 16))

    def upload(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        rc = _hip().hipMemcpy(ctypes.c_void_p(self.ptr), ctypes.c_void_p(a.ctypes.data),
                              ctypes.c_size_t(a.nbytes), 1)
        if rc != 0:
            raise RuntimeError("hipMemcpy failed: %d" % rc)

    def download(self, shape, dtype=np.complex128):
        out = np.empty(shape, dtype)
        assert out.nbytes <= self.nbytes
        rc = _hip().hipMemcpy(ctypes.c_void_p(out.ctypes.data), ctypes.c_void_p(self.ptr),
                              ctypes.c_size_t(out.nbytes), 2)
        if rc != 0:
            raise RuntimeError("hipMemcpy failed: %d" % rc)
        return out

    def free(self):
        if self.ptr:
            _hip().hipFree(ctypes.c_void_p(self.ptr))
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DeviceGrid:
    """Device memory of rcwa_rasterize's output: num_layers x Ny x Nx
    complex128 ([layer][iy][ix]); .ptr for device_grids in Context.solve."""

    def __init__(self, shape):
        self.shape = tuple(shape)
        self.nbytes = int(np.prod(self.shape)) * 16
        self.ptr = _hip_malloc(self.nbytes)

    def download(self):
        out = np.empty(self.shape, np.complex128)
        rc = _hip().hipMemcpy(ctypes.c_void_p(out.ctypes.data), ctypes.c_void_p(self.ptr),
                              ctypes.c_size_t(self.nbytes), 2)
        if rc != 0:
            raise RuntimeError("hipMemcpy failed: %d" % rc)
        return out

    def free(self):
        if self.ptr:
            _hip().hipFree(ctypes.c_void_p(self.ptr))
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


_HIP = None


def _hip():
    global _HIP
    if _HIP is None:
        _HIP = ctypes.CDLL("libamdhip64.so")
    return _HIP


def _hip_malloc(nbytes):
    p = ctypes.c_void_p()
    rc = _hip().hipMalloc(ctypes.byref(p), ctypes.c_size_t(nbytes))
    if rc != 0:
        raise RuntimeError("hipMalloc failed: %d" % rc)
    return p.value


class SweepStructure(ctypes.Structure):
    _fields_ = [("Nx", ctypes.c_int), ("Ny", ctypes.c_int), ("num_layers", ctypes.c_int),
                ("num_materials", ctypes.c_int), ("Lx", ctypes.c_double), ("Ly", ctypes.c_double),
                ("material", ctypes.c_void_p), ("t", ctypes.c_void_p)]


class Problem(ctypes.Structure):
    _fields_ = [("device", ctypes.POINTER(Device)), ("params", ctypes.POINTER(Params)),
                ("src", ctypes.POINTER(Source)), ("nsrc", ctypes.c_int)]


class Stats(ctypes.Structure):
    _fields_ = [("ms_total", ctypes.c_double), ("ms_convmat", ctypes.c_double),
                ("ms_layers", ctypes.c_double), ("ms_eig", ctypes.c_double),
                ("ms_fold", ctypes.c_double), ("ms_out", ctypes.c_double),
                ("flops_algorithmic", ctypes.c_double), ("n_smatrices", ctypes.c_int),
                ("n_layers", ctypes.c_int), ("ms_eig_hess", ctypes.c_double),
                ("ms_eig_schur", ctypes.c_double), ("ms_eig_vec", ctypes.c_double),
                ("n_eig_launches", ctypes.c_int), ("n_eig_matrices", ctypes.c_int)]


class ServerConfig(ctypes.Structure):
    _fields_ = [("hbm_budget", ctypes.c_size_t), ("max_batch", ctypes.c_int),
                ("n_workers", ctypes.c_int), ("precision", ctypes.c_int),
                ("autotune", ctypes.c_int)]


class TicketInfo(ctypes.Structure):
    _fields_ = [("t_submit", ctypes.c_double), ("t_start", ctypes.c_double),
                ("t_finish", ctypes.c_double), ("batch_size", ctypes.c_int)]


# Every function declared in include/rcwa_amd.h.
EXPORTS = [
    "rcwa_last_error", "rcwa_version", "rcwa_ctx_create", "rcwa_ctx_destroy",
    "rcwa_ctx_set_eig_maxit", "rcwa_ctx_set_qr_log", "rcwa_ctx_set_refine", "rcwa_ctx_get_stats", "rcwa_ctx_reset_stats", "rcwa_ctx_set_chunk",
    "rcwa_ctx_set_timing", "rcwa_ctx_set_autotune", "rcwa_autotune_entries",
    "rcwa_zgemm_force_shape",
    "rcwa_solve", "rcwa_solve_batch", "rcwa_last_source_columns", "rcwa_convmat",
    "rcwa_wave_vectors", "rcwa_gap_medium", "rcwa_layer_smatrix", "rcwa_end_smatrix",
    "rcwa_star_product", "rcwa_eig", "rcwa_inverse", "rcwa_zgemm", "rcwa_zgemm_split",
    "rcwa_efficiencies", "rcwa_convmat_batched", "rcwa_layer_smatrix_batched",
    "rcwa_star_product_batched", "rcwa_star_product_batched_noalloc", "rcwa_rasterize", "rcwa_sweep_create",
    "rcwa_sweep_run", "rcwa_sweep_destroy", "rcwa_server_create",
    "rcwa_server_submit",
    "rcwa_server_poll", "rcwa_server_wait", "rcwa_server_cancel", "rcwa_server_clock",
    "rcwa_server_destroy",
]

_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("librcwa_amd.so is not built: run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        L.rcwa_last_error.restype = ctypes.c_char_p
        L.rcwa_version.restype = ctypes.c_char_p
        L.rcwa_server_clock.restype = ctypes.c_double
        L.rcwa_server_clock.argtypes = [ctypes.c_void_p]
        _LIB = L
    return _LIB


def _check(rc):
    if rc != 0:
        raise RcwaError(rc, lib().rcwa_last_error().decode())


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


def _cz(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.complex128)
    return a if shape is None else a.reshape(shape)


def _grids(p, on_device=False):
    """(er, ur, thick) of a workloads.Problem as the library reads them:
    C-contiguous complex128 grids [layer][iy][ix] and float64 thicknesses,
    with er.shape == ur.shape == (len(thick), Ny, Nx).  Checked here, at the
    call, because the caller may have replaced p.er after constructing p."""
    thick = np.ascontiguousarray(p.thick, dtype=np.float64).reshape(-1)
    if on_device:
        return None, None, thick
    er = _cz(p.er)
    ur = _cz(p.ur)
    if er.ndim != 3 or er.shape != ur.shape or er.shape[0] != thick.size:
        raise ValueError("er and ur must both have shape (len(thick), Ny, Nx); got %s, %s, %d"
                         % (er.shape, ur.shape, thick.size))
    return er, ur, thick


def _fz(a):
    """Column-major complex128 array (the library's matrix layout)."""
    return np.asfortranarray(a, dtype=np.complex128)


def _stack_f(mats):
    """Batch of matrices -> contiguous [b][col-major n*m] buffer."""
    mats = [np.asarray(m, np.complex128) for m in mats]
    return np.ascontiguousarray(np.stack([m.T for m in mats]))  # m.T row-major == m col-major


def _unstack_f(buf, b, r, c):
    return np.stack([buf[i].reshape(c, r).T for i in range(b)])


class Context:
    """A solve context on one GPU (rcwa_ctx)."""

    def __init__(self, device=0, hbm_budget=0):
        self.h = ctypes.c_void_p()
        _check(lib().rcwa_ctx_create(device, ctypes.c_size_t(hbm_budget), ctypes.byref(self.h)))

    def close(self):
        if self.h:
            lib().rcwa_ctx_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_eig_maxit(self, m):
        _check(lib().rcwa_ctx_set_eig_maxit(self.h, m))

    def set_qr_log(self, rotations_per_matrix):
        _check(lib().rcwa_ctx_set_qr_log(self.h, ctypes.c_longlong(rotations_per_matrix)))

    def set_refine(self, level):
        _check(lib().rcwa_ctx_set_refine(self.h, level))

    def set_autotune(self, on):
        _check(lib().rcwa_ctx_set_autotune(self.h, 1 if on else 0))

    def set_timing(self, on):
        _check(lib().rcwa_ctx_set_timing(self.h, 1 if on else 0))

    def set_chunk(self, m):
        _check(lib().rcwa_ctx_set_chunk(self.h, m))

    def stats(self):
        s = Stats()
        _check(lib().rcwa_ctx_get_stats(self.h, ctypes.byref(s)))
        return {f: getattr(s, f) for f, _ in Stats._fields_}

    def reset_stats(self):
        _check(lib().rcwa_ctx_reset_stats(self.h))

    # ---------------------------------------------------------------- solves
    def solve(self, probs, precision=RCWA_FP64, device_grids=None, grid_shape=None,
              fields=False):
        """Solve a list of workloads.Problem.  Returns a list (per problem) of
        (R[npol][Q][P], T[npol][Q][P], tot[npol][2]), with fields=True
        followed by the per-order complex amplitudes r, t [npol][3][PQ]
        (x, y, z components; harmonic index m = ix*Q + iy).  device_grids, if
        given, is a list of (er_dev_ptr, ur_dev_ptr) for HBM-resident grids of
        shape [len(p.thick)][grid_shape[0]][grid_shape[1]]."""
        keep = []
        cprobs = (Problem * len(probs))()
        outs = []
        total_src = sum(len(p.pols) for p in probs)
        res = (Results * total_src)()
        k = 0
        for i, p in enumerate(probs):
            P, Q = 2 * p.hx + 1, 2 * p.hy + 1
            npol = len(p.pols)
            R = np.zeros((npol, P * Q))
            T = np.zeros((npol, P * Q))
            ra = np.zeros((npol, 3 * P * Q), np.complex128) if fields else None
            ta = np.zeros((npol, 3 * P * Q), np.complex128) if fields else None
            outs.append((R, T, ra, ta))
            er, ur, thick = _grids(p, on_device=device_grids is not None)
            if device_grids is not None:
                er_p, ur_p = device_grids[i]
                ondev = 1
                if grid_shape is None:
                    raise ValueError("device_grids needs grid_shape")
            else:
                er_p, ur_p = er.ctypes.data, ur.ctypes.data
                ondev = 0
            Ny, Nx = grid_shape if grid_shape is not None else er.shape[1:]
            d = Device(Nx, Ny, thick.size, p.Lx, p.Ly, er_p, ur_p, thick.ctypes.data, ondev)
            keep += [er, ur, thick]
            par = Params(p.hx, p.hy, _c(p.er_ref), _c(p.ur_ref), _c(p.er_trn), _c(p.ur_trn))
            srcs = (Source * npol)(*[Source(p.lam0, p.theta, p.phi, a, b) for a, b in p.pols])
            keep += [d, par, srcs, p]
            cprobs[i] = Problem(ctypes.pointer(d), ctypes.pointer(par), srcs, npol)
            for j in range(npol):
                res[k] = Results(R[j].ctypes.data, T[j].ctypes.data, 0.0, 0.0,
                                 ra[j].ctypes.data if fields else None,
                                 ta[j].ctypes.data if fields else None)
                k += 1
        _check(lib().rcwa_solve_batch(self.h, cprobs, len(probs), precision, res))
        out = []
        # per problem, per source: the conditioning bound of rcwa_results.cond
        self.last_cond = []
        k = 0
        for i, p in enumerate(probs):
            P, Q = 2 * p.hx + 1, 2 * p.hy + 1
            npol = len(p.pols)
            R, T, ra, ta = outs[i]
            tot = np.array([[res[k + j].R_tot, res[k + j].T_tot] for j in range(npol)])
            self.last_cond.append(np.array([res[k + j].cond for j in range(npol)]))
            k += npol
            one = (R.reshape(npol, P, Q).transpose(0, 2, 1),
                   T.reshape(npol, P, Q).transpose(0, 2, 1), tot)
            if fields:
                one += (ra.reshape(npol, 3, P * Q), ta.reshape(npol, 3, P * Q))
            out.append(one)
        return out

    def rasterize(self, Nx, Ny, Lx, Ly, background, shapes):
        """Shapes (tuples (layer, kind, cx, cy, ax, ay, value)) over
        background[layer], sampled at cell centres on the device.  Returns a
        DeviceGrid of shape (len(background), Ny, Nx)."""
        bg = _cz(background)
        L = bg.size
        arr = (Shape * max(len(shapes), 1))(
            *[Shape(int(l), int(k), float(cx), float(cy), float(ax), float(ay), _c(v))
              for (l, k, cx, cy, ax, ay, v) in shapes])
        g = DeviceGrid((L, Ny, Nx))
        _check(lib().rcwa_rasterize(self.h, Nx, Ny, ctypes.c_double(Lx), ctypes.c_double(Ly), L,
                                    _ptr(bg), arr, len(shapes), ctypes.c_void_p(g.ptr)))
        return g

    def sweep(self, material, thick, Lx, Ly, num_materials, hx, hy, er_ref=1.0, ur_ref=1.0,
              er_trn=2.25, ur_trn=1.0):
        """A sweep plan (rcwa_sweep_create) of a structure given as material
        indices material[layer][iy][ix] (uint8, < num_materials)."""
        return Sweep(self, material, thick, Lx, Ly, num_materials, hx, hy, er_ref, ur_ref,
                     er_trn, ur_trn)

    def last_source_columns(self, count):
        buf = np.zeros(count, np.complex128)
        _check(lib().rcwa_last_source_columns(self.h, _ptr(buf), ctypes.c_size_t(count)))
        return buf

    # ---------------------------------------------------------------- stages
    def convmat(self, grids, layer, hx, hy):
        g = _cz(grids)
        L, Ny, Nx = g.shape
        PQ = (2 * hx + 1) * (2 * hy + 1)
        out = np.zeros(PQ * PQ, np.complex128)
        _check(lib().rcwa_convmat(self.h, _ptr(g), layer, Nx, Ny, hx, hy, _ptr(out)))
        return out.reshape(PQ, PQ).T

    def layer_smatrix(self, ercs, urcs, Kx, Ky, k0t):
        b = len(ercs)
        N = ercs[0].shape[0]
        e = _stack_f(ercs)
        u = _stack_f(urcs)
        kx = _cz(Kx)
        ky = _cz(Ky)
        kt = np.ascontiguousarray(k0t, np.float64)
        n = 2 * N
        S11 = np.zeros((b, n * n), np.complex128)
        S12 = np.zeros((b, n * n), np.complex128)
        _check(lib().rcwa_layer_smatrix(self.h, N, b, _ptr(e), _ptr(u), _ptr(kx), _ptr(ky),
                                        _ptr(kt), _ptr(S11), _ptr(S12)))
        return _unstack_f(S11, b, n, n), _unstack_f(S12, b, n, n)

    def star(self, A, B):
        """A, B: tuples of 4 arrays [batch][n][n]."""
        b, n = A[0].shape[0], A[0].shape[1]
        As = [_stack_f(list(x)) for x in A]
        Bs = [_stack_f(list(x)) for x in B]
        Ss = [np.zeros((b, n * n), np.complex128) for _ in range(4)]
        arr = ctypes.c_void_p * 4
        _check(lib().rcwa_star_product(self.h, n, b, arr(*[x.ctypes.data for x in As]),
                                       arr(*[x.ctypes.data for x in Bs]),
                                       arr(*[x.ctypes.data for x in Ss])))
        return tuple(_unstack_f(x, b, n, n) for x in Ss)

    def layer_smatrix_batched(self, erc_dev, urc_dev, PQ, batch, Kx, Ky, k0t, S11_dev, S12_dev):
        """rcwa_layer_smatrix_batched on device buffers (addresses): per-layer
        Kx, Ky [batch][PQ] and k0t [batch] on the host."""
        Kx = _cz(Kx).reshape(batch, PQ)
        Ky = _cz(Ky).reshape(batch, PQ)
        k0t = np.ascontiguousarray(k0t, np.float64).reshape(batch)
        _check(lib().rcwa_layer_smatrix_batched(
            self.h, PQ, batch, ctypes.c_void_p(erc_dev), ctypes.c_void_p(urc_dev), _ptr(Kx),
            _ptr(Ky), _ptr(k0t), ctypes.c_void_p(S11_dev), ctypes.c_void_p(S12_dev)))

    def star_batched(self, n, batch, A_dev, B_dev, S_dev, noalloc=False):
        """rcwa_star_product_batched (or its allocation-free form) on device
        blocks: A_dev, B_dev, S_dev are four device addresses each."""
        arr = ctypes.c_void_p * 4
        fn = (lib().rcwa_star_product_batched_noalloc if noalloc
              else lib().rcwa_star_product_batched)
        _check(fn(self.h, n, batch, arr(*A_dev), arr(*B_dev), arr(*S_dev)))

    def eig(self, mats):
        b, n = len(mats), mats[0].shape[0]
        A = _stack_f(mats)
        w = np.zeros((b, n), np.complex128)
        V = np.zeros((b, n * n), np.complex128)
        _check(lib().rcwa_eig(self.h, n, b, _ptr(A), _ptr(w), _ptr(V)))
        return w, _unstack_f(V, b, n, n)

    def inverse(self, mats):
        b, n = len(mats), mats[0].shape[0]
        A = _stack_f(mats)
        out = np.zeros((b, n * n), np.complex128)
        _check(lib().rcwa_inverse(self.h, n, b, _ptr(A), _ptr(out)))
        return _unstack_f(out, b, n, n)

    def zgemm(self, As, Bs, split=False):
        b = len(As)
        m, k = As[0].shape
        n = Bs[0].shape[1]
        A = _stack_f(As)
        B = _stack_f(Bs)
        C = np.zeros((b, m * n), np.complex128)
        fn = lib().rcwa_zgemm_split if split else lib().rcwa_zgemm
        _check(fn(self.h, m, n, k, b, _ptr(A), _ptr(B), _ptr(C)))
        return _unstack_f(C, b, m, n)


def autotune_entries():
    """The process's tuned ZGEMM tile shapes: (M, N, K, batch class, device,
    shape) tuples (rcwa_autotune_entries)."""
    n = lib().rcwa_autotune_entries(None, 0)
    buf = (ctypes.c_int * (6 * max(n, 1)))()
    n = lib().rcwa_autotune_entries(buf, n)
    return [tuple(buf[6 * i:6 * i + 6]) for i in range(n)]


def zgemm_force_shape(shape):
    """Forces ZGEMM tile shape `shape` (0 .. 4) for the process, -1 undoes."""
    _check(lib().rcwa_zgemm_force_shape(shape))


# ---------------------------------------------------------------- host stages
def wave_vectors(Lx, Ly, lam0, theta, phi, hx, hy, er_ref, ur_ref, er_trn, ur_trn):
    PQ = (2 * hx + 1) * (2 * hy + 1)
    src = Source(lam0, theta, phi, 1.0, 0.0)
    par = Params(hx, hy, _c(er_ref), _c(ur_ref), _c(er_trn), _c(ur_trn))
    k = np.zeros(3, np.complex128)
    o = [np.zeros(PQ, np.complex128) for _ in range(4)]
    _check(lib().rcwa_wave_vectors(ctypes.c_double(Lx), ctypes.c_double(Ly), ctypes.byref(src),
                                   ctypes.byref(par), _ptr(k), *[_ptr(x) for x in o]))
    return (k, *o)


def gap_medium(Kx, Ky):
    PQ = len(Kx)
    W0 = np.zeros(4 * PQ * PQ, np.complex128)
    V0 = np.zeros(4 * PQ * PQ, np.complex128)
    _check(lib().rcwa_gap_medium(PQ, _ptr(_cz(Kx)), _ptr(_cz(Ky)), _ptr(W0), _ptr(V0)))
    n = 2 * PQ
    return W0.reshape(n, n).T, V0.reshape(n, n).T


def end_smatrix(which, hx, hy, Kx, Ky, Kz, er_ref, ur_ref, er_trn, ur_trn):
    PQ = len(Kx)
    n = 2 * PQ
    par = Params(hx, hy, _c(er_ref), _c(ur_ref), _c(er_trn), _c(ur_trn))
    S = [np.zeros(n * n, np.complex128) for _ in range(4)]
    _check(lib().rcwa_end_smatrix(which, PQ, ctypes.byref(par), _ptr(_cz(Kx)), _ptr(_cz(Ky)),
                                  _ptr(_cz(Kz)), *[_ptr(x) for x in S]))
    return tuple(x.reshape(n, n).T for x in S)


class Sweep:
    """A structure made of materials, for sweeps over wavelength and angle:
    the convolution matrices of its (layer, material) masks are computed once
    on the device, and each point mixes them with its material values
    (rcwa_sweep_*)."""

    def __init__(self, ctx, material, thick, Lx, Ly, num_materials, hx, hy, er_ref, ur_ref,
                 er_trn, ur_trn):
        self.ctx = ctx  # the plan uses the context: keep it alive
        self.material = np.ascontiguousarray(material, dtype=np.uint8)
        L, Ny, Nx = self.material.shape
        self.thick = np.ascontiguousarray(thick, dtype=np.float64)
        self.M, self.hx, self.hy = int(num_materials), int(hx), int(hy)
        self.params = Params(hx, hy, _c(er_ref), _c(ur_ref), _c(er_trn), _c(ur_trn))
        st = SweepStructure(Nx, Ny, L, self.M, Lx, Ly, self.material.ctypes.data,
                            self.thick.ctypes.data)
        self.h = ctypes.c_void_p()
        _check(lib().rcwa_sweep_create(ctx.h, ctypes.byref(st), ctypes.byref(self.params),
                                       ctypes.byref(self.h)))

    def run(self, sources, eps, mu=None, precision=RCWA_FP64, fields=False):
        """sources: per point a list of (lam0, theta, phi, pte, ptm), the same
        count for every point; eps (and mu, None = 1): [points][materials].
        Returns per point what Context.solve returns per problem."""
        npt = len(sources)
        nsrc = len(sources[0]) if npt else 1
        if any(len(ss) != nsrc for ss in sources):
            raise ValueError("every point needs the same number of sources")
        P, Q = 2 * self.hx + 1, 2 * self.hy + 1
        src = (Source * max(npt * nsrc, 1))(*[Source(*x) for ss in sources for x in ss])
        e = _cz(eps, (npt, self.M))
        m = None if mu is None else _cz(mu, (npt, self.M))
        R = np.zeros((npt * nsrc, P * Q))
        T = np.zeros((npt * nsrc, P * Q))
        ra = np.zeros((npt * nsrc, 3 * P * Q), np.complex128) if fields else None
        ta = np.zeros((npt * nsrc, 3 * P * Q), np.complex128) if fields else None
        res = (Results * max(npt * nsrc, 1))(
            *[Results(R[k].ctypes.data, T[k].ctypes.data, 0.0, 0.0,
                      ra[k].ctypes.data if fields else None,
                      ta[k].ctypes.data if fields else None) for k in range(npt * nsrc)])
        _check(lib().rcwa_sweep_run(self.h, npt, src, nsrc, _ptr(e),
                                    None if m is None else _ptr(m), precision, res))
        out = []
        for p in range(npt):
            k = slice(p * nsrc, (p + 1) * nsrc)
            tot = np.array([[res[p * nsrc + j].R_tot, res[p * nsrc + j].T_tot]
                            for j in range(nsrc)])
            one = (R[k].reshape(nsrc, P, Q).transpose(0, 2, 1),
                   T[k].reshape(nsrc, P, Q).transpose(0, 2, 1), tot)
            if fields:
                one += (ra[k].reshape(nsrc, 3, P * Q), ta[k].reshape(nsrc, 3, P * Q))
            out.append(one)
        return out

    def close(self):
        if self.h:
            lib().rcwa_sweep_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """The serving form (rcwa_server): continuous batching of submitted
    solves on one GPU."""

    def __init__(self, device=0, hbm_budget=0, max_batch=0, n_workers=0,
                 precision=RCWA_FP64, autotune=False):
        self.h = ctypes.c_void_p()
        cfg = ServerConfig(hbm_budget, max_batch, n_workers, precision, 1 if autotune else 0)
        _check(lib().rcwa_server_create(device, ctypes.byref(cfg), ctypes.byref(self.h)))
        self._meta = {}

    def submit(self, p, sources=None):
        """Queues problem p; its sources are p's polarisations at p's
        incidence, or `sources`: (lam0, theta, phi, pte, ptm) tuples, which
        may name several incidences (each runs as its own flight)."""
        er, ur, thick = _grids(p)
        L, Ny, Nx = er.shape
        d = Device(Nx, Ny, L, p.Lx, p.Ly, er.ctypes.data, ur.ctypes.data, thick.ctypes.data, 0)
        par = Params(p.hx, p.hy, _c(p.er_ref), _c(p.ur_ref), _c(p.er_trn), _c(p.ur_trn))
        if sources is None:
            sources = [(p.lam0, p.theta, p.phi, a, b) for a, b in p.pols]
        npol = len(sources)
        srcs = (Source * npol)(*[Source(*s_) for s_ in sources])
        prob = Problem(ctypes.pointer(d), ctypes.pointer(par), srcs, npol)
        t = ctypes.c_uint64()
        _check(lib().rcwa_server_submit(self.h, ctypes.byref(prob), ctypes.byref(t)))
        self._meta[t.value] = (p.hx, p.hy, npol)
        return t.value

    def _collect(self, fn, ticket, fields=False):
        if ticket not in self._meta:
            raise RcwaError(1, "unknown ticket %d (never submitted, or already collected)" % ticket)
        hx, hy, npol = self._meta[ticket]
        P, Q = 2 * hx + 1, 2 * hy + 1
        R = np.zeros((npol, P * Q))
        T = np.zeros((npol, P * Q))
        ra = np.zeros((npol, 3 * P * Q), np.complex128)
        ta = np.zeros((npol, 3 * P * Q), np.complex128)
        res = (Results * npol)(*[Results(R[j].ctypes.data, T[j].ctypes.data, 0.0, 0.0,
                                         ra[j].ctypes.data if fields else None,
                                         ta[j].ctypes.data if fields else None)
                                 for j in range(npol)])
        info = TicketInfo()
        rc = fn(self.h, ctypes.c_uint64(ticket), res, ctypes.byref(info))
        if rc == 7:
            return None
        # any other status is final: the server has erased the ticket
        self._meta.pop(ticket, None)
        _check(rc)
        tot = np.array([[res[j].R_tot, res[j].T_tot] for j in range(npol)])
        info_d = {"t_submit": info.t_submit, "t_start": info.t_start,
                  "t_finish": info.t_finish, "batch_size": info.batch_size,
                  "cond": max(res[j].cond for j in range(npol))}
        out = (R.reshape(npol, P, Q).transpose(0, 2, 1), T.reshape(npol, P, Q).transpose(0, 2, 1),
               tot, info_d)
        if fields:
            out += (ra.reshape(npol, 3, P * Q), ta.reshape(npol, 3, P * Q))
        return out

    def wait(self, ticket, fields=False):
        """(R, T, tot, info), with fields=True followed by r, t [npol][3][PQ]."""
        return self._collect(lib().rcwa_server_wait, ticket, fields)

    def poll(self, ticket, fields=False):
        return self._collect(lib().rcwa_server_poll, ticket, fields)

    def cancel(self, ticket):
        _check(lib().rcwa_server_cancel(self.h, ctypes.c_uint64(ticket)))

    def clock(self):
        return lib().rcwa_server_clock(self.h)

    def close(self):
        if self.h:
            lib().rcwa_server_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
