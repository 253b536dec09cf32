"""Python mirror of the reference's HSS2D MATLAB entry points, backed by libhssc.so.

Function names and argument meaning follow HSSC_DirectSolver (DOCUMENTATION_HSS2D.txt):

    params    = HSS_tree_parameters(pot, kh, sym, Ng, np, acc, lay, n_cut, TI)
    INFO_TREE = HSS2D_build_tree_fsym(params)
    A_HSS     = HSS2D_invert_fsym(INFO_TREE, params)            # Fopts = 'dense'
    Y         = HSS2D_applyinv_TI(INFO_TREE, A_HSS, X, params)  # or _fsym for NTI
    M_HSS     = HSS2D_compress_fsym(INFO_TREE, params)
    Y         = HSS2D_apply_TI(INFO_TREE, M_HSS, X, params)     # or _fsym
    lHSS      = levHSS(INFO_TREE, n_cut, trinv)

All numerical work runs on the GPU through the C ABI.  Right-hand sides and solutions are
N x m column-major blocks in the X_source ordering (one column per right-hand side), as the
C ABI takes them: numpy arrays of shape (N,) or (N, m) for the host functions, and
contiguous torch tensors on the GPU of shape (N,) or (m, N) (each row one right-hand side,
which is the same N x m column-major memory) for the *_device helpers; float64 for
'SL_L_2D', complex128 for the Helmholtz kernel 'SL_H_2D'.  Shapes, dtypes, devices and
contiguity are checked before any pointer reaches the library.
Errors raise HsscError (the reference raises MATLAB runtime errors instead).

Handles keep what they read alive: an Inverse or Operator holds a reference to its Tree,
whose device data they use until they are destroyed.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import CONTRAST, PREC, POT, PROXY_RULE, HsscParams, check, lib

_DP = C.POINTER(C.c_double)


def _dptr(a):
    return a.ctypes.data_as(_DP)


class Params:
    """The params struct of Common/HSS_tree_parameters.m:116-139."""

    def __init__(self, raw: HsscParams):
        self.raw = raw

    def __getattr__(self, name):
        raw = object.__getattribute__(self, "raw")
        return getattr(raw, name)

    def __setattr__(self, name, value):
        # params.<field> = value sets the struct field (e.g. params.proxy = 1 as
        # HSS1D/Test/Run_HSS1D_tests.m:19 does)
        if name != "raw" and name in dict(HsscParams._fields_):
            setattr(self.raw, name, value)
        else:
            object.__setattr__(self, name, value)

    @property
    def is_complex(self) -> bool:
        return self.raw.flag_pot == POT["SL_H_2D"]

    @property
    def dtype(self):
        return np.complex128 if self.is_complex else np.float64

    @property
    def N(self) -> int:
        return int(lib().hssc_params_npoints(C.byref(self.raw)))

    @property
    def X_source(self) -> np.ndarray:
        """N x 2 grid of Common/HSS_tree_parameters.m:144-160."""
        X = np.empty((2, self.N))
        check(lib().hssc_source_points(C.byref(self.raw), _dptr(X)))
        return X.T


def HSS_tree_parameters(pot, kh, sym, Ng, np_, acc, lay, n_cut, TI, corr0=None,
                        proxy_layers=None, preset_all=None, contrast=None, kernel_scale=None,
                        length_scale=None, tree_precision=None, proxy=None, proxy_rule=None,
                        lam=None, skel_extra_pts=None) -> Params:
    """Common/HSS_tree_parameters.m:1.

    For kh > 0 the Duan-Rokhlin weight corr0 is computed by the library (``corr0``
    overrides it); ``proxy_layers`` replaces the proxy-ring depth rule at lev > 0;
    ``preset_all=True`` keeps the 'bdry' skeleton rule at the two bottom levels of a kh > 0
    tree, where the reference selects skeletons by an ID.  Options without a reference
    counterpart: ``contrast=("gaussian", amp, width)`` (kh > 0, NTI) instead of the tanh
    plateau; ``kernel_scale`` s and ``length_scale`` L (SL_L_2D) for the kernel
    s (L h)^2 log(L r); ``tree_precision="fp32"`` for FP32 interpolation operators (a
    preconditioner, refined against the exact grid operator).  ``proxy`` sets params.proxy
    (Kernel_Eval's b / c factors; HSS2D trees need 0); ``proxy_rule`` ('bdry' or 'rand'),
    ``lam`` (params.lambda) and ``skel_extra_pts`` are the reference's fields of those
    names."""
    raw = HsscParams()
    check(lib().hssc_params_init(C.byref(raw), POT[pot] if isinstance(pot, str) else int(pot),
                                 float(kh), int(sym), int(Ng), int(np_), float(acc), int(lay),
                                 int(n_cut), int(TI)))
    if corr0 is not None:
        check(lib().hssc_params_set_dr_weights(C.byref(raw), float(np.real(corr0)),
                                               float(np.imag(corr0))))
    if proxy_layers is not None:
        raw.proxy_layers = int(proxy_layers)
    if preset_all is not None:
        raw.preset_all = int(bool(preset_all))
    if contrast is not None:
        if contrast == "plateau":
            raw.contrast = CONTRAST["plateau"]
        else:
            kind, amp, width = contrast
            raw.contrast = CONTRAST[kind]
            raw.contrast_amp, raw.contrast_width = float(amp), float(width)
    if kernel_scale is not None:
        raw.kernel_scale = float(kernel_scale)
    if length_scale is not None:
        raw.length_scale = float(length_scale)
    if tree_precision is not None:
        raw.tree_precision = PREC[tree_precision]
    if proxy is not None:
        raw.proxy = int(proxy)
    if proxy_rule is not None:
        raw.proxy_rule = PROXY_RULE[proxy_rule]
    if lam is not None:
        raw.lambda_ = float(lam)
    if skel_extra_pts is not None:
        raw.skel_extra_pts = int(bool(skel_extra_pts))
    return Params(raw)


def Kernel_Eval(X1, X2, params: Params) -> np.ndarray:
    """Common/Kernel_Eval.m:1 (computed on the GPU)."""
    X1 = np.asfortranarray(X1, dtype=np.float64)
    X2 = np.asfortranarray(X2, dtype=np.float64)
    m, n = X1.shape[0], X2.shape[0]
    A = np.empty((m, n), order="F", dtype=params.dtype)
    check(lib().hssc_kernel_eval(C.byref(params.raw), _dptr(X1), m, _dptr(X2), n, _dptr(A),
                                 max(m, 1)))
    return A


class Tree:
    """INFO_TREE: owns the device-resident tree (hssc_tree*)."""

    def __init__(self, handle, params: Params):
        self.handle = handle
        self.params = params

    def __del__(self):
        h = getattr(self, "handle", None)
        if h:
            lib().hssc_tree_destroy(h)
            self.handle = None

    @property
    def depth(self) -> int:
        d = C.c_int32()
        check(lib().hssc_tree_depth(self.handle, C.byref(d)))
        return d.value

    def level_info(self, lev):
        vals = [C.c_int64() for _ in range(6)]
        check(lib().hssc_tree_level_info(self.handle, lev, *[C.byref(v) for v in vals]))
        return dict(zip(("n", "k", "M", "nbox", "nkey", "q"), (v.value for v in vals)))

    def box_ranks(self, lev):
        """(ranks, id_level): each key box's own skeleton count before the level's padding
        (the adaptive ID rank on ID levels, the 'bdry' count on preset levels)."""
        nkey = self.level_info(lev)["nkey"]
        out = np.empty(nkey, dtype=np.int32)
        idl = C.c_int32()
        check(lib().hssc_tree_box_ranks(self.handle, int(lev),
                                        out.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(idl)))
        return out, bool(idl.value)

    def id_stats(self):
        """(blocks, fallbacks to the plain ID, sketch rows, rows factored by the plain ID) of
        the skeleton IDs."""
        v = [C.c_int64() for _ in range(4)]
        check(lib().hssc_tree_id_stats(self.handle, *[C.byref(x) for x in v]))
        return dict(zip(("blocks", "fallbacks", "sketch_rows", "full_rows"), (x.value for x in v)))

    def box_features(self, lev):
        """Rank-network features (nkey x 8) of the key boxes of a level."""
        nkey = self.level_info(lev)["nkey"]
        X = np.empty((8, nkey), dtype=np.float32)  # column-major nkey x 8
        check(lib().hssc_tree_box_features(self.handle, int(lev),
                                           X.ctypes.data_as(C.POINTER(C.c_float))))
        return X.T.copy()

    def _i64(self, fn, box, size):
        out = np.empty(size, dtype=np.int64)
        check(fn(self.handle, int(box), out.ctypes.data_as(C.POINTER(C.c_int64))))
        return out

    def _box_level(self, box):
        return int(box).bit_length() - 1

    def I_src(self, box):
        return self._i64(lib().hssc_tree_get_isrc, box, self.level_info(self._box_level(box))["n"])

    def I_sk(self, box):
        return self._i64(lib().hssc_tree_get_isk, box, self.level_info(self._box_level(box))["k"])

    def J(self, box):
        return self._i64(lib().hssc_tree_get_J, box, self.level_info(self._box_level(box))["n"])

    def proxy(self, box):
        M = self.level_info(self._box_level(box))["M"]
        out = np.empty((2, M))
        check(lib().hssc_tree_get_proxy(self.handle, int(box), _dptr(out)))
        return out.T

    def T(self, box):
        info = self.level_info(self._box_level(box))
        k, r = info["k"], info["n"] - info["k"]
        out = np.empty((r, k), dtype=self.params.dtype).T  # k x r column-major view
        buf = np.empty(k * r, dtype=self.params.dtype)
        check(lib().hssc_tree_get_T(self.handle, int(box), _dptr(buf)))
        return buf.reshape((r, k)).T.copy() if k * r else out


RANK_MODES = {"id": 0, "table": 1, "mlp": 2}


def _tree(params: Params, index_only=False) -> Tree:
    h = C.c_void_p()
    fn = lib().hssc_hss2d_build_index_sets if index_only else lib().hssc_hss2d_build_tree
    check(fn(C.byref(params.raw), C.byref(h)))
    return Tree(h, params)


def HSS2D_build_tree_fsym(params: Params, rank="id", mlp=None, oversample=8) -> Tree:
    """HSS2D/HSS2D_build_tree_fsym.m:1 (index sets on the host, T on the GPU).

    ``rank`` sizes the skeleton IDs of a kh > 0 tree: 'id' runs the plain ID on every block
    (the reference's HSS2D_build_tree_fsym.m:686-729); 'table' sketches each level with the
    closed-form rank guess of HSS2D/predict_rank.m; 'mlp' with the prediction of ``mlp`` (a
    RankMLP trained on box_features).  Sketched IDs that fail the check on their full block
    fall back to the plain ID."""
    if rank == "id":
        return _tree(params)
    if rank == "mlp" and mlp is None:
        raise ValueError("rank='mlp' needs a trained RankMLP")
    h = C.c_void_p()
    check(lib().hssc_hss2d_build_tree_ranked(C.byref(params.raw), RANK_MODES[rank],
                                             mlp.handle if mlp is not None else None,
                                             int(oversample), C.byref(h)))
    return Tree(h, params)


def HSS2D_build_index_sets(params: Params) -> Tree:
    """Tree geometry and index sets only (no interpolation operators; the ID-selected skeleton
    levels run on the GPU)."""
    return _tree(params, index_only=True)


def levHSS(tree: Tree, n_cut, trinv=None) -> int:
    """HSS2D/levHSS.m:1."""
    out = C.c_int32()
    check(lib().hssc_levhss(tree.handle, int(n_cut), C.byref(out)))
    return out.value


class Inverse:
    """A_HSS of HSS2D_invert_fsym (dense F^-1 and E per box or per level)."""

    def __init__(self, handle, tree: Tree, precision: str):
        self.handle = handle
        self.tree = tree  # keeps the tree alive
        self.precision = precision

    def __del__(self):
        h = getattr(self, "handle", None)
        if h:
            lib().hssc_inverse_destroy(h)
            self.handle = None

    def FI(self, box):
        n = self.tree.level_info(self.tree._box_level(box))["n"]
        buf = np.empty(n * n, dtype=self.tree.params.dtype)
        check(lib().hssc_inverse_get_FI(self.handle, int(box), _dptr(buf)))
        return buf.reshape((n, n)).T.copy()

    def E(self, box):
        k = self.tree.level_info(self.tree._box_level(box))["k"]
        buf = np.empty(k * k, dtype=self.tree.params.dtype)
        check(lib().hssc_inverse_get_E(self.handle, int(box), _dptr(buf)))
        return buf.reshape((k, k)).T.copy()


def HSS2D_invert_fsym(tree: Tree, params: Params = None, precision: str = "fp64") -> Inverse:
    """HSS2D/HSS2D_invert_fsym.m:1 with Fopts = 'dense'."""
    h = C.c_void_p()
    check(lib().hssc_hss2d_invert(tree.handle, PREC[precision], C.byref(h)))
    return Inverse(h, tree, precision)


def dense_inv_device(A, sym=False):
    """MATLAB's inv on a batch of dense blocks, in place, on the GPU: A is a contiguous
    torch tensor on the device of shape (batch, n, n) or (n, n), float64 / float32 /
    complex128 / complex64, each block stored column-major (A[b] holds the transpose of
    block b, which is block b itself when it is symmetric).  sym=True: complex symmetric
    blocks (upper-tile sweep).  Raises HsscError (HSSC_ERR_SINGULAR) on a zero, non-finite
    or negligible pivot."""
    import torch
    if not (A.is_cuda and A.is_contiguous() and A.dim() in (2, 3) and A.shape[-1] == A.shape[-2]):
        raise ValueError("A must be a contiguous CUDA tensor of shape (batch, n, n) or (n, n)")
    kinds = {torch.float64: (0, "fp64"), torch.float32: (0, "fp32"),
             torch.complex128: (1, "fp64"), torch.complex64: (1, "fp32")}
    if A.dtype not in kinds:
        raise ValueError(f"unsupported dtype {A.dtype}")
    cx, prec = kinds[A.dtype]
    n = A.shape[-1]
    batch = A.shape[0] if A.dim() == 3 else 1
    torch.cuda.synchronize(A.device)
    check(lib().hssc_dense_invert_device(cx, PREC[prec], int(bool(sym)), n,
                                         C.c_void_p(A.data_ptr()), n, n * n, batch))
    return A


class Operator:
    """A_HSS of HSS2D_compress_fsym (dense Bself / B12 blocks)."""

    def __init__(self, handle, tree: Tree):
        self.handle = handle
        self.tree = tree

    def __del__(self):
        h = getattr(self, "handle", None)
        if h:
            lib().hssc_operator_destroy(h)
            self.handle = None


def HSS2D_compress_fsym(tree: Tree, params: Params = None) -> Operator:
    """HSS2D/HSS2D_compress_fsym.m:1 with dense blocks."""
    h = C.c_void_p()
    check(lib().hssc_hss2d_compress(tree.handle, C.byref(h)))
    return Operator(h, tree)


def _host_call(fn, handle, X, dtype, N):
    X = np.asarray(X, dtype=dtype)
    shape = X.shape
    if X.ndim not in (1, 2) or shape[0] != N or X.size == 0:
        raise ValueError(f"right-hand side must have shape (N,) or (N, m) with N = {N}, "
                         f"got {shape}")
    Xf = np.asfortranarray(X.reshape(shape[0], -1))
    Y = np.empty_like(Xf, order="F")
    check(fn(handle, _dptr(Xf), _dptr(Y), Xf.shape[1]))
    return Y.reshape(shape)


def HSS2D_applyinv_TI(tree: Tree, A_HSS: Inverse, X, params: Params = None):
    """HSS2D/HSS2D_applyinv_TI.m:1."""
    return _host_call(lib().hssc_hss2d_applyinv, A_HSS.handle, X, A_HSS.tree.params.dtype,
                      A_HSS.tree.params.N)


HSS2D_applyinv_fsym = HSS2D_applyinv_TI  # HSS2D/HSS2D_applyinv_fsym.m:1 (same C entry point)


def HSS2D_apply_TI(tree: Tree, M_HSS: Operator, X, params: Params = None):
    """HSS2D/HSS2D_apply_TI.m:1."""
    return _host_call(lib().hssc_hss2d_apply, M_HSS.handle, X, M_HSS.tree.params.dtype,
                      M_HSS.tree.params.N)


HSS2D_apply_fsym = HSS2D_apply_TI  # HSS2D/HSS2D_apply_fsym.m:1


# ---- device-resident variants (torch tensors on the GPU, float64 or complex128 for the
# Helmholtz kernel, shape (N,) or (m, N): N x m column-major) ----

def _stream_ptr(stream):
    return None if stream is None else C.c_void_p(stream.cuda_stream)


def _device_rhs(params: Params, *tensors):
    """Checks that every tensor is a contiguous GPU tensor of the tree's dtype holding N x m
    entries (shape (N,) or (m, N)), all of the same shape; returns m."""
    import torch
    want = torch.complex128 if params.is_complex else torch.float64
    N = params.N
    shape = tensors[0].shape
    for X in tensors:
        if not X.is_cuda:
            raise ValueError("device variants take GPU tensors")
        if X.dtype != want:
            raise ValueError(f"expected dtype {want}, got {X.dtype}")
        if not X.is_contiguous():
            raise ValueError("device variants take contiguous tensors")
        if X.shape != shape:
            raise ValueError(f"shape mismatch: {tuple(X.shape)} vs {tuple(shape)}")
    if len(shape) == 1 and shape[0] == N:
        return 1
    if len(shape) == 2 and shape[1] == N and shape[0] >= 1:
        return int(shape[0])
    raise ValueError(f"expected shape (N,) or (m, N) with N = {N}, got {tuple(shape)}")


def applyinv_device(A_HSS: Inverse, X, Y, stream=None):
    """Y = A^-1 X with X, Y torch tensors of shape (N,) or (m, N) (each row one RHS)."""
    m = _device_rhs(A_HSS.tree.params, X, Y)
    check(lib().hssc_hss2d_applyinv_device(A_HSS.handle, C.c_void_p(X.data_ptr()),
                                           C.c_void_p(Y.data_ptr()), m, _stream_ptr(stream)))


def applyinv_device_noalloc(A_HSS: Inverse, X, Y, stream=None):
    """applyinv_device with the library's fixed device workspace (no allocation; calls that
    overlap on one GPU race)."""
    m = _device_rhs(A_HSS.tree.params, X, Y)
    check(lib().hssc_hss2d_applyinv_device_noalloc(A_HSS.handle, C.c_void_p(X.data_ptr()),
                                                   C.c_void_p(Y.data_ptr()), m,
                                                   _stream_ptr(stream)))


def apply_device(M_HSS: Operator, X, Y, stream=None):
    """Y = A_HSS X with X, Y torch tensors of shape (N,) or (m, N)."""
    m = _device_rhs(M_HSS.tree.params, X, Y)
    check(lib().hssc_hss2d_apply_device(M_HSS.handle, C.c_void_p(X.data_ptr()),
                                        C.c_void_p(Y.data_ptr()), m, _stream_ptr(stream)))


def solve_device(A_HSS: Inverse, M_HSS: Operator, B, Y, tol=1e-10, maxit=10, stream=None):
    """Common/Iterative_Refinement.m on the GPU; returns (iterations, relative residual)."""
    if A_HSS.tree is not M_HSS.tree:
        raise ValueError("the inverse and the operator must come from the same tree")
    m = _device_rhs(A_HSS.tree.params, B, Y)
    it = C.c_int32()
    rel = C.c_double()
    check(lib().hssc_hss2d_solve_device(A_HSS.handle, M_HSS.handle, C.c_void_p(B.data_ptr()),
                                        C.c_void_p(Y.data_ptr()), m, float(tol), int(maxit),
                                        C.byref(it), C.byref(rel), _stream_ptr(stream)))
    return it.value, rel.value


class GridOperator:
    """The exact operator A = Kernel_Eval(X_source, X_source, params) of the uniform grid,
    applied by FFT convolution on the GPU (the dense products of
    HSS2D/Test/Test_HSS2D_InverseCompression.m:300-311 and Test_HSS2D_Matrix_Apply.m:227-235
    without forming A)."""

    def __init__(self, params: Params):
        self.params = params
        h = C.c_void_p()
        check(lib().hssc_grid_op_create(C.byref(params.raw), C.byref(h)))
        self.handle = h

    def __del__(self):
        h = getattr(self, "handle", None)
        if h:
            lib().hssc_grid_op_destroy(h)
            self.handle = None

    def apply(self, X):
        """A X for host arrays of shape (N,) or (N, m)."""
        return _host_call(lib().hssc_grid_op_apply, self.handle, X, self.params.dtype,
                          self.params.N)

    def apply_device(self, X, Y, stream=None):
        """Y = A X for GPU tensors of shape (N,) or (m, N)."""
        m = _device_rhs(self.params, X, Y)
        check(lib().hssc_grid_op_apply_device(self.handle, C.c_void_p(X.data_ptr()),
                                              C.c_void_p(Y.data_ptr()), m, _stream_ptr(stream)))


def solve_exact_device(A_HSS: Inverse, A: GridOperator, B, Y, tol=1e-10, maxit=10, stream=None):
    """Common/Iterative_Refinement.m on the GPU with the exact grid operator as the residual
    operator; returns (solves, relative residual ||B - A Y|| / ||B||)."""
    m = _device_rhs(A_HSS.tree.params, B, Y)
    it = C.c_int32()
    rel = C.c_double()
    check(lib().hssc_hss2d_solve_exact_device(A_HSS.handle, A.handle, C.c_void_p(B.data_ptr()),
                                              C.c_void_p(Y.data_ptr()), m, float(tol), int(maxit),
                                              C.byref(it), C.byref(rel), _stream_ptr(stream)))
    return it.value, rel.value


def bicgstab(A, b, tol=1e-10, maxit=200, M=None):
    """MATLAB's bicgstab(A, b, tol, maxit, M) on GPU tensors (the preconditioned iterative
    solve of HSS2D/Test/Test_HSS2D_InverseCompression.m:201-228: A the compressed NTI operator,
    M the HSS inverse of the translation-invariant operator).  A and M are callables
    y = A(x), y = M(x) (M applies the preconditioner's inverse) on tensors of shape (N,).
    Van der Vorst's iteration from x0 = 0, the preconditioner applied to the search and the
    intermediate directions; stops when the true relative residual ||b - A x|| / ||b|| is at
    most tol, after a half step (s) or a full one.  Returns (x, flag, relres, it) like MATLAB:
    flag 0 = converged, 1 = maxit reached, 4 = breakdown; it counts half steps as 0.5."""
    import torch
    nb = torch.linalg.vector_norm(b).item()
    x = torch.zeros_like(b)
    if nb == 0:
        return x, 0, 0.0, 0
    Mi = M if M is not None else (lambda v: v)
    r = b.clone()
    rt = r.clone()
    rho = alpha = omega = 1.0 + 0j
    p = torch.zeros_like(b)
    v = torch.zeros_like(b)

    def true_res(y):
        return torch.linalg.vector_norm(b - A(y)).item() / nb

    relres = 1.0
    for i in range(1, maxit + 1):
        rho1, rho = rho, torch.vdot(rt, r).item()
        if rho == 0:
            return x, 4, relres, i - 1
        if i == 1:
            p = r.clone()
        else:
            p = r + ((rho / rho1) * (alpha / omega)) * (p - omega * v)
        ph = Mi(p)
        v = A(ph)
        den = torch.vdot(rt, v).item()
        if den == 0:
            return x, 4, relres, i - 1
        alpha = rho / den
        xh = x + alpha * ph
        s = r - alpha * v
        if torch.linalg.vector_norm(s).item() / nb <= tol:
            relres = true_res(xh)
            if relres <= tol:
                return xh, 0, relres, i - 0.5
        sh = Mi(s)
        t = A(sh)
        tt = torch.vdot(t, t).real.item()
        if tt == 0:
            return xh, 4, relres, i - 0.5
        omega = torch.vdot(t, s).item() / tt
        x = xh + omega * sh
        r = s - omega * t
        if torch.linalg.vector_norm(r).item() / nb <= tol:
            relres = true_res(x)
            if relres <= tol:
                return x, 0, relres, i
        if omega == 0:
            return x, 4, relres, i
    return x, 1, true_res(x), maxit
