"""Device-resident interface over the C ABI, for column-major torch tensors in HBM.

A column-major m x n matrix is a 2-D tensor with stride (1, ld), ld >= max(m, 1) -- e.g.
`colmajor_empty(m, n)` or `to_device(numpy_array)`. Pivots are int32 swap destinations
(Elemental's Permutation::SwapDestinations()). torch provides memory and streams only; all
arithmetic runs in the HIP kernels of libelemental_amd.so.
"""
import numpy as np
import torch

from ._lib import IrStats, call, load

NORMAL, TRANSPOSE, ADJOINT = 0, 1, 2
LOWER, UPPER = 0, 1
LEFT, RIGHT = 0, 1
NON_UNIT, UNIT = 0, 1

FACTOR = {"fp64": 0, "fp32": 1, "fp16": 2, "bf16": 3}
REFINE = {"none": 0, "ir": 1, "gmres": 2, "auto": 3}


def _stream(stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes_void(s.cuda_stream)


def ctypes_void(x):
    import ctypes
    return ctypes.c_void_p(x)


def _mat(t, dtype=torch.float64):
    if t.dim() != 2 or t.dtype != dtype or not t.is_cuda:
        raise ValueError(f"expected a 2-D {dtype} CUDA tensor, got {t.dtype} {tuple(t.shape)}")
    m, n = t.shape
    if m > 1 and t.stride(0) != 1:
        raise ValueError("expected a column-major tensor (stride(0) == 1)")
    ld = t.stride(1) if n > 1 else max(m, 1)
    if ld < max(m, 1):
        raise ValueError("leading dimension smaller than the height")
    return t.data_ptr(), ld


def _pre(t):
    """Entry-point prefix for the tensor's field: elx_d (float64) or elx_s (float32)."""
    if t.dtype == torch.float64:
        return "elx_d"
    if t.dtype == torch.float32:
        return "elx_s"
    raise ValueError(f"expected float64 or float32, got {t.dtype}")


def _m(t):
    """(pointer, leading dimension) of a float64 or float32 column-major tensor."""
    return _mat(t, t.dtype if t.dtype in (torch.float32, torch.float64) else torch.float64)


def colmajor_empty(m, n, dtype=torch.float64, device="cuda"):
    return torch.empty((n, m), dtype=dtype, device=device).t()


def to_device(a, dtype=None, device="cuda"):
    """Column-major device copy of a numpy array (float32 stays float32, anything else becomes
    float64 unless dtype is given)."""
    a = np.asarray(a)
    if dtype is None:
        dtype = torch.float32 if a.dtype == np.float32 else torch.float64
    if a.ndim == 1:
        a = a[:, None]
    t = colmajor_empty(a.shape[0], a.shape[1], dtype=dtype, device=device)
    t.copy_(torch.from_numpy(np.ascontiguousarray(a)))
    return t


def to_host(t):
    return np.asfortranarray(t.detach().cpu().numpy())


def getrf(A, nb=128, stream=None):
    """In-place PA = LU (float64 or float32). Returns (dests int32 tensor, info int32 tensor)."""
    pA, lda = _m(A)
    m, n = A.shape
    dests = torch.zeros(max(min(m, n), 1), dtype=torch.int32, device=A.device)
    info = torch.zeros(1, dtype=torch.int32, device=A.device)
    call(_pre(A) + "getrf", _stream(stream), m, n, pA, lda, nb, dests.data_ptr(), info.data_ptr())
    return dests[: min(m, n)], info


def getrf_nopiv(A, nb=128, stream=None):
    """In-place A = LU without pivoting. Returns the info int32 tensor."""
    pA, lda = _m(A)
    m, n = A.shape
    info = torch.zeros(1, dtype=torch.int32, device=A.device)
    call(_pre(A) + "getrf_nopiv", _stream(stream), m, n, pA, lda, nb, info.data_ptr())
    return info


def getrf_full(A, stream=None):
    """In-place P A Q = L U with full pivoting. Returns (p, q int32 tensors, info)."""
    pA, lda = _m(A)
    m, n = A.shape
    k = max(min(m, n), 1)
    p = torch.zeros(k, dtype=torch.int32, device=A.device)
    q = torch.zeros(k, dtype=torch.int32, device=A.device)
    info = torch.zeros(1, dtype=torch.int32, device=A.device)
    call(_pre(A) + "getrf_full", _stream(stream), m, n, pA, lda, p.data_ptr(), q.data_ptr(),
         info.data_ptr())
    return p[: min(m, n)], q[: min(m, n)], info


def getrs_full(A, p, q, B, orient=NORMAL, stream=None):
    pA, lda = _m(A)
    pB, ldb = _m(B)
    call(_pre(A) + "getrs_full", _stream(stream), orient, A.shape[0], B.shape[1], pA, lda,
         p.data_ptr(), q.data_ptr(), pB, ldb)
    return B


def getrs(A, dests, B, orient=NORMAL, stream=None):
    """lu::SolveAfter; dests None for the factors of getrf_nopiv."""
    pA, lda = _m(A)
    pB, ldb = _m(B)
    call(_pre(A) + "getrs", _stream(stream), orient, A.shape[0], B.shape[1], pA, lda,
         None if dests is None else dests.data_ptr(), pB, ldb)
    return B


def gesv(A, B, nb=128, stream=None, want_dests=False):
    """Row-echelon LU on [A|B] and upper solve (A and B overwritten)."""
    pA, lda = _m(A)
    pB, ldb = _m(B)
    n = A.shape[0]
    dests = torch.zeros(max(n, 1), dtype=torch.int32, device=A.device)
    info = torch.zeros(1, dtype=torch.int32, device=A.device)
    call(_pre(A) + "gesv", _stream(stream), n, B.shape[1], pA, lda, pB, ldb, nb,
         dests.data_ptr(), info.data_ptr())
    return (dests[:n], info) if want_dests else info


def potrf(A, uplo=LOWER, nb=128, stream=None):
    pA, lda = _m(A)
    info = torch.zeros(1, dtype=torch.int32, device=A.device)
    call(_pre(A) + "potrf", _stream(stream), uplo, A.shape[0], pA, lda, nb, info.data_ptr())
    return info


def potrs(A, B, uplo=LOWER, orient=NORMAL, stream=None):
    pA, lda = _m(A)
    pB, ldb = _m(B)
    call(_pre(A) + "potrs", _stream(stream), uplo, orient, A.shape[0], B.shape[1], pA, lda, pB,
         ldb)
    return B


def posv(A, B, uplo=LOWER, orient=NORMAL, nb=128, stream=None):
    pA, lda = _m(A)
    pB, ldb = _m(B)
    info = torch.zeros(1, dtype=torch.int32, device=A.device)
    call(_pre(A) + "posv", _stream(stream), uplo, orient, A.shape[0], B.shape[1], pA, lda, pB,
         ldb, nb, info.data_ptr())
    return info


def gemm(transA, transB, alpha, A, B, beta, C, stream=None):
    pA, lda = _m(A)
    pB, ldb = _m(B)
    pC, ldc = _m(C)
    m, n = C.shape
    k = A.shape[1] if transA == NORMAL else A.shape[0]
    call(_pre(C) + "gemm", _stream(stream), transA, transB, m, n, k, alpha, pA, lda, pB, ldb,
         beta, pC, ldc)
    return C


def trsm(side, uplo, orient, diag, alpha, A, B, stream=None):
    pA, lda = _m(A)
    pB, ldb = _m(B)
    m, n = B.shape
    call(_pre(B) + "trsm", _stream(stream), side, uplo, orient, diag, m, n, alpha, pA, lda, pB,
         ldb)
    return B


def syrk(uplo, orient, alpha, A, beta, C, stream=None):
    pA, lda = _m(A)
    pC, ldc = _m(C)
    n = C.shape[0]
    k = A.shape[1] if orient == NORMAL else A.shape[0]
    call(_pre(C) + "syrk", _stream(stream), uplo, orient, n, k, alpha, pA, lda, beta, pC, ldc)
    return C


def ruiz_equil(A, stream=None):
    """El::RuizEquil in place on a column-major FP64 device matrix; returns (dRow, dCol)."""
    import torch
    pA, lda = _mat(A)
    m, n = A.shape
    dRow = torch.empty(m, dtype=torch.float64, device=A.device)
    dCol = torch.empty(n, dtype=torch.float64, device=A.device)
    call("elx_druiz_equil", _stream(stream), m, n, pA, lda, dRow.data_ptr(), dCol.data_ptr())
    return dRow, dCol


def geom_equil(A, stream=None):
    """El::GeomEquil in place on a column-major FP64 device matrix; returns (dRow, dCol)."""
    import torch
    pA, lda = _mat(A)
    m, n = A.shape
    dRow = torch.empty(m, dtype=torch.float64, device=A.device)
    dCol = torch.empty(n, dtype=torch.float64, device=A.device)
    call("elx_dgeom_equil", _stream(stream), m, n, pA, lda, dRow.data_ptr(), dCol.data_ptr())
    return dRow, dCol


def symmetric_ruiz_equil(A, maxiter=3, stream=None):
    """El::SymmetricRuizEquil in place on a square column-major FP64 device matrix; returns d."""
    import torch
    pA, lda = _mat(A)
    n = A.shape[0]
    d = torch.empty(n, dtype=torch.float64, device=A.device)
    call("elx_dsymmetric_ruiz_equil", _stream(stream), n, pA, lda, d.data_ptr(), maxiter)
    return d


def permute_rows(dests, A, inverse=False, stream=None):
    pA, lda = _m(A)
    call(_pre(A) + "permute_rows", _stream(stream), 1 if inverse else 0, dests.data_ptr(),
         dests.numel(), A.shape[0], A.shape[1], pA, lda)
    return A


def gesv_mixed(A, B, X, factor="fp16", refine="auto", tol=0.0, max_iters=0, stream=None,
               equilibrate=False):
    """Low-precision LU + FP64 refinement; returns the stats dict (synchronous). With
    equilibrate, the LU factors the Ruiz-equilibrated matrix (ELX_FACTOR_EQUILIBRATE)."""
    pA, lda = _mat(A)
    pB, ldb = _mat(B)
    pX, ldx = _mat(X)
    st = IrStats()
    import ctypes
    status = load().elx_dgesv_mixed(_stream(stream), A.shape[0], B.shape[1], pA, lda, pB, ldb,
                                     pX, ldx, FACTOR[factor] | (0x100 if equilibrate else 0),
                                     REFINE[refine], tol, max_iters, ctypes.byref(st))
    if status not in (0, 5):
        from ._lib import ElxError
        raise ElxError("elx_dgesv_mixed", status)
    out = st.as_dict()
    out["status"] = status
    return out


def posv_mixed(A, B, X, uplo=LOWER, factor="fp16", refine="auto", tol=0.0, max_iters=0,
               stream=None, equilibrate=False):
    """Low-precision Cholesky + FP64 refinement of an HPD A (its `uplo` triangle); returns the
    stats dict (synchronous), with "status" 0 or 5 (tolerance not reached). With equilibrate,
    the Cholesky factors the symmetrically equilibrated matrix (ELX_FACTOR_EQUILIBRATE)."""
    pA, lda = _mat(A)
    pB, ldb = _mat(B)
    pX, ldx = _mat(X)
    st = IrStats()
    import ctypes
    status = load().elx_dposv_mixed(_stream(stream), uplo, A.shape[0], B.shape[1], pA, lda, pB,
                                     ldb, pX, ldx, FACTOR[factor] | (0x100 if equilibrate else 0),
                                     REFINE[refine], tol, max_iters, ctypes.byref(st))
    if status not in (0, 5):
        from ._lib import ElxError
        raise ElxError("elx_dposv_mixed", status)
    out = st.as_dict()
    out["status"] = status
    return out


__all__ = [
    "load", "getrf", "getrf_nopiv", "getrf_full", "getrs_full", "getrs", "gesv", "potrf", "potrs", "posv", "gemm", "trsm", "syrk",
    "permute_rows", "ruiz_equil", "symmetric_ruiz_equil", "geom_equil", "gesv_mixed", "posv_mixed", "to_device", "to_host", "colmajor_empty",
    "Batch", "potrf_vbatched", "potrs_vbatched", "PosvGraph", "batched_variant", "gesv_mixed_batched",
]


class Batch:
    """A batch of column-major matrices (one torch tensor each) for the vbatched entry
    points: host arrays of sizes and leading dimensions, a device array of pointers."""

    def __init__(self, mats, dtype):
        self.mats = list(mats)
        self.dtype = dtype
        ptrs, ns, lds = [], [], []
        for t in self.mats:
            p, ld = _mat(t, dtype)
            ptrs.append(p)
            ns.append(t.shape[0])
            lds.append(ld)
        self.n = np.ascontiguousarray(ns, dtype=np.int32)
        self.ld = np.ascontiguousarray(lds, dtype=np.int32)
        self.cols = np.ascontiguousarray([t.shape[1] for t in self.mats], dtype=np.int32)
        self.dptrs = torch.tensor(ptrs, dtype=torch.int64, device="cuda") if ptrs else \
            torch.zeros(1, dtype=torch.int64, device="cuda")

    def __len__(self):
        return len(self.mats)


def _np_ptr(a):
    import ctypes
    return ctypes.c_void_p(a.ctypes.data)


def potrf_vbatched(batch, uplo=LOWER, stream=None):
    """Cholesky of every matrix of the batch in place; returns the device int32 info array."""
    info = torch.zeros(max(len(batch), 1), dtype=torch.int32, device="cuda")
    name = "elx_dpotrf_vbatched" if batch.dtype == torch.float64 else "elx_spotrf_vbatched"
    call(name, _stream(stream), uplo, len(batch), _np_ptr(batch.n), ctypes_void(batch.dptrs.data_ptr()),
         _np_ptr(batch.ld), ctypes_void(info.data_ptr()))
    return info[: len(batch)]


def potrs_vbatched(factors, rhs, uplo=LOWER, stream=None):
    """L L^T X = B for every (factor, right-hand sides) pair; B is overwritten with X."""
    if len(factors) != len(rhs):
        raise ValueError("factor and right-hand-side batches differ in length")
    name = "elx_dpotrs_vbatched" if factors.dtype == torch.float64 else "elx_spotrs_vbatched"
    call(name, _stream(stream), uplo, len(factors), _np_ptr(factors.n), _np_ptr(rhs.cols),
         ctypes_void(factors.dptrs.data_ptr()), _np_ptr(factors.ld),
         ctypes_void(rhs.dptrs.data_ptr()), _np_ptr(rhs.ld))


class PosvGraph:
    """elx_?posv_vbatched_graph: the batched factor + solve of fixed matrices and right-hand
    sides captured into a HIP graph; launch() replays it, close() frees it."""

    def __init__(self, factors, rhs, uplo=LOWER, stream=None):
        import ctypes
        if len(factors) != len(rhs) or len(factors) == 0:
            raise ValueError("factor and right-hand-side batches must have the same, non-zero length")
        self.factors, self.rhs = factors, rhs
        self.info = torch.zeros(len(factors), dtype=torch.int32, device="cuda")
        self.handle = ctypes.c_void_p()
        name = "elx_dposv_vbatched_graph" if factors.dtype == torch.float64 else "elx_sposv_vbatched_graph"
        call(name, _stream(stream), uplo, len(factors), _np_ptr(factors.n), _np_ptr(rhs.cols),
             ctypes_void(factors.dptrs.data_ptr()), _np_ptr(factors.ld),
             ctypes_void(rhs.dptrs.data_ptr()), _np_ptr(rhs.ld), ctypes_void(self.info.data_ptr()),
             ctypes.byref(self.handle))

    def launch(self, stream=None):
        call("elx_graph_launch", self.handle, _stream(stream))
        return self.info

    def close(self):
        if self.handle:
            call("elx_graph_destroy", self.handle)
            self.handle = None


def batched_variant(dtype, bin_index, nb=None, occ=None, kind=0):
    """Reads (nb, occ, kind) of a batched-Cholesky size bin; with nb given, sets it first (nb 0
    restores the built-in table; kind 1 = operands read straight into registers, kind 2 =
    64-column panels (nb 64), both in bins 1-2)."""
    import ctypes
    d = 0 if dtype == torch.float64 else 1
    if nb is not None:
        call("elx_batched_set_variant", d, bin_index, nb, occ or 0, kind)
    a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    call("elx_batched_get_variant", d, bin_index, ctypes.byref(a), ctypes.byref(b),
         ctypes.byref(c))
    return a.value, b.value, c.value


def gesv_mixed_batched(As, Bs, Xs, factor="fp16", refine="auto", tol=0.0, max_iters=0,
                       stream=None):
    """elx_dgesv_mixed over lists of column-major FP64 tensors; returns (status, list of stats
    dicts)."""
    import ctypes
    cnt = len(As)
    if not (len(Bs) == len(Xs) == cnt):
        raise ValueError("batch lengths differ")
    pa, la, pb, lb, px, lx, ns, nr = [], [], [], [], [], [], [], []
    for A, B, X in zip(As, Bs, Xs):
        p, l = _mat(A)
        pa.append(p)
        la.append(l)
        p, l = _mat(B)
        pb.append(p)
        lb.append(l)
        p, l = _mat(X)
        px.append(p)
        lx.append(l)
        ns.append(A.shape[0])
        nr.append(B.shape[1])
    arr64 = lambda v: np.ascontiguousarray(v, dtype=np.uint64)
    arr32 = lambda v: np.ascontiguousarray(v, dtype=np.int32)
    keep = [arr32(ns), arr32(nr), arr64(pa), arr32(la), arr64(pb), arr32(lb), arr64(px), arr32(lx)]
    stats = (IrStats * max(cnt, 1))()
    status = load().elx_dgesv_mixed_batched(_stream(stream), cnt, *[_np_ptr(a) for a in keep],
                                             FACTOR[factor], REFINE[refine], tol, max_iters,
                                             ctypes.cast(stats, ctypes.c_void_p))
    if status not in (0, 5):
        from ._lib import ElxError
        raise ElxError("elx_dgesv_mixed_batched", status)
    return status, [stats[i].as_dict() for i in range(cnt)]

