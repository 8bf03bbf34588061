"""Batched CG over many independent dense systems of uneven size (block-diagonal operators).

cola solves a block-diagonal system block by block: ``inv(BlockDiag) =
BlockDiag(inv(M_i))`` (cola/linalg/inverse/inv.py:135-136), i.e. one CG loop per
block (cg.py:94-178).  ``HipBatchedCG`` runs every block in the same launches
(``ck_bcg_*``): five kernels per iteration whatever the number of blocks, systems
that have converged skip their work, and runs of iterations are replayed from a
HIP graph.  ``BatchedMixedSolver`` adds FP64 refinement (residuals with the fp32
matrices as stored) around inner solves that may read bf16 copies of the
matrices, with a per-system fall back to fp32 storage when a refinement step does
not contract.  ``autotune`` picks the matvec tile height by timing the candidates.
"""
import ctypes
import time
from dataclasses import dataclass

import numpy as np
import torch

from .. import _native as N
from ..ops import Dense, LinearOperator, Product, Sum
from .algorithm_base import Algorithm

_ROWS = (4, 8, 16, 32)
_TUNE_CACHE = {}


def _round_up(v, m):
    return (v + m - 1) // m * m


class BatchedSystems:
    """Concatenated storage for systems (K_s + shift_s I), s = 0..S-1: fp32 matrices (and
    optional bf16 copies) row-major with a leading dimension that is a multiple of 8."""

    def __init__(self, sizes, shifts, device, bf16=False):
        self.sizes = np.asarray(sizes, dtype=np.int64)
        if self.sizes.ndim != 1 or self.sizes.size < 1 or self.sizes.min() < 1:
            raise ValueError("sizes must be a non-empty list of positive orders")
        self.nsys = int(self.sizes.size)
        self.device = torch.device(device)
        self.lda = _round_up(self.sizes, 8)
        msz = self.sizes * self.lda
        self.moff = np.concatenate([[0], np.cumsum(msz)[:-1]]).astype(np.int64)
        vsz = _round_up(self.sizes, 64)
        self.voff = np.concatenate([[0], np.cumsum(vsz)[:-1]]).astype(np.int64)
        self.total_v = int(vsz.sum())
        self.total_m = int(msz.sum())
        self.A32 = torch.zeros(self.total_m, dtype=torch.float32, device=self.device)
        self.A16 = torch.zeros(self.total_m, dtype=torch.bfloat16, device=self.device) if bf16 else None
        self.shift = torch.tensor(np.asarray(shifts, dtype=np.float32).reshape(-1).repeat(
            1 if np.size(shifts) == self.nsys else self.nsys), device=self.device)

    def matrix(self, s, dtype=torch.float32):
        n, ld, o = int(self.sizes[s]), int(self.lda[s]), int(self.moff[s])
        buf = self.A32 if dtype == torch.float32 else self.A16
        return buf[o:o + n * ld].view(n, ld)[:, :n]

    def segment(self, vec, s):
        o, n = int(self.voff[s]), int(self.sizes[s])
        return vec[o:o + n]

    def concat(self, vectors, dtype=torch.float32):
        """Concatenate per-system vectors into the padded layout."""
        out = torch.zeros(self.total_v, dtype=dtype, device=self.device)
        for s, v in enumerate(vectors):
            self.segment(out, s).copy_(v.reshape(-1))
        return out

    def split(self, vec):
        return [self.segment(vec, s) for s in range(self.nsys)]

    @classmethod
    def rbf(cls, points, lengthscales, shifts, device, bf16=False):
        """Dense RBF kernel systems exp(-|p_i - p_j|^2 / (2 l_s^2)) + shift_s I built on the engine."""
        sizes = [int(P.shape[0]) for P in points]
        out = cls(sizes, shifts, device, bf16=bf16)
        st = N.stream_ptr(out.device)
        for s, P in enumerate(points):
            Pd = P.to(device=out.device, dtype=torch.float32).contiguous()
            o = int(out.moff[s])
            kb = None if out.A16 is None else N.ptr(out.A16[o:])
            N.call("ck_rbf_dense_f32", sizes[s], Pd.shape[1], N.ptr(Pd), Pd.stride(0), float(lengthscales[s]),
                   N.ptr(out.A32[o:]), int(out.lda[s]), kb, int(out.lda[s]), st)
        return out

    @classmethod
    def dense(cls, mats, shifts, device, bf16=False):
        out = cls([int(M.shape[0]) for M in mats], shifts, device, bf16=bf16)
        for s, M in enumerate(mats):
            out.matrix(s).copy_(M)
            if bf16:
                out.matrix(s, torch.bfloat16).copy_(M.to(torch.bfloat16))
        return out


class BatchedCG:
    """Device state, tile lists and the ck_bcg_plan of one batched CG configuration."""

    def __init__(self, systems, rows_per_tile=16, max_iters=1000, precond=None):
        if rows_per_tile not in _ROWS:
            raise ValueError(f"rows_per_tile must be one of {_ROWS}")
        sy = self.systems = systems
        self.precond = precond
        dev = sy.device
        self.rows = rows_per_tile
        self.max_iters = int(max_iters)
        vt = int(N.q("ck_bcg_vec_tile"))
        tsys, trow, vsys, vi0 = [], [], [], []
        mtb, mte, vtb, vte = [], [], [], []
        for s, n in enumerate(sy.sizes.tolist()):
            mtb.append(len(tsys))
            rows = list(range(0, n, rows_per_tile))
            tsys += [s] * len(rows)
            trow += rows
            mte.append(len(tsys))
            vtb.append(len(vsys))
            idx = list(range(0, n, vt))
            vsys += [s] * len(idx)
            vi0 += idx
            vte.append(len(vsys))
        i32 = lambda a: torch.tensor(np.asarray(a, dtype=np.int32), device=dev)  # noqa: E731
        i64 = lambda a: torch.tensor(np.asarray(a, dtype=np.int64), device=dev)  # noqa: E731
        self.t = {"tile_sys": i32(tsys), "tile_row0": i32(trow), "vtile_sys": i32(vsys), "vtile_i0": i32(vi0),
                  "n": i32(sy.sizes), "lda": i32(sy.lda), "voff": i64(sy.voff), "moff": i64(sy.moff),
                  "mt_begin": i32(mtb), "mt_end": i32(mte), "vt_begin": i32(vtb), "vt_end": i32(vte)}
        S = sy.nsys
        f32 = lambda m: torch.zeros(m, dtype=torch.float32, device=dev)  # noqa: E731
        self.x, self.r, self.p, self.Ap = (f32(sy.total_v) for _ in range(4))
        self.part_mv, self.part_vec = f32(len(tsys)), f32(len(vsys))
        self.gamma, self.alpha, self.beta, self.thresh = (f32(S) for _ in range(4))
        self.k = torch.zeros(S, dtype=torch.int32, device=dev)
        self.done = torch.zeros(S, dtype=torch.int32, device=dev)
        self.prec = torch.zeros(S, dtype=torch.int32, device=dev)
        self.ntiles_mv, self.ntiles_vec = len(tsys), len(vsys)
        pl = N.CkBcgPlan()
        pl.nsys, pl.rows_per_tile, pl.ntiles_mv, pl.ntiles_vec = S, rows_per_tile, len(tsys), len(vsys)
        pl.max_iters = self.max_iters
        for name, tns in self.t.items():
            setattr(pl, name, tns.data_ptr())
        pl.shift, pl.prec = sy.shift.data_ptr(), self.prec.data_ptr()
        pl.A32 = sy.A32.data_ptr()
        pl.A16 = sy.A16.data_ptr() if sy.A16 is not None else None
        for name in ("x", "r", "p", "Ap", "part_mv", "part_vec", "gamma", "alpha", "beta", "thresh", "k", "done"):
            setattr(pl, name, getattr(self, name).data_ptr())
        if precond is not None:
            k = precond.rank
            self.z = f32(sy.total_v)
            self.part_rz = f32(len(vsys))
            self.part_ftr = torch.zeros(len(vsys) * k, dtype=torch.float64, device=dev)
            self.u = torch.zeros(S * k, dtype=torch.float64, device=dev)
            pl.rank = k
            pl.F, pl.foff, pl.Cinv = precond.F.data_ptr(), precond.foff_t.data_ptr(), precond.Cinv.data_ptr()
            for name in ("z", "part_rz", "part_ftr", "u"):
                setattr(pl, name, getattr(self, name).data_ptr())
        self.plan = pl
        self.stream = torch.cuda.Stream(device=dev)  # default priority; graphs need a non-legacy stream
        self.graphs = {}
        self._pinned = torch.empty(S, dtype=torch.int32, pin_memory=True)

    def set_precision(self, bf16_mask):
        """Per-system matvec storage: True reads the bf16 copy (needs systems built with bf16)."""
        m = np.asarray(bf16_mask, dtype=np.int32).reshape(-1)
        if m.any() and self.systems.A16 is None:
            raise ValueError("these systems have no bf16 copy")
        self.prec.copy_(torch.from_numpy(np.broadcast_to(m, (self.systems.nsys, )).copy()))

    def _graph(self, iters):
        g = self.graphs.get(iters)
        if g is None:
            h = ctypes.c_void_p()
            N.call("ck_bcg_graph_create", ctypes.byref(self.plan), int(iters), ctypes.c_void_p(self.stream.cuda_stream),
                   ctypes.byref(h))
            g = self.graphs[iters] = h
        return g

    def __del__(self):
        try:
            for h in self.graphs.values():
                N.q("ck_graph_destroy", h)
        except Exception:
            pass

    def matvec(self):
        N.call("ck_bcg_matvec", ctypes.byref(self.plan), ctypes.c_void_p(self.stream.cuda_stream))

    def solve(self, b, tol_rel=1e-6, tol_abs=0.0, active=None, chunk=8, use_graph=True):
        """CG on every active system with rhs b (concatenated, fp32) until ||r_s|| <= tol_rel ||b_s|| + tol_abs.
        Returns (x concatenated view, per-system iteration counts)."""
        S = self.systems.nsys
        st = ctypes.c_void_p(self.stream.cuda_stream)
        self.stream.wait_stream(torch.cuda.current_stream(self.systems.device))
        with torch.cuda.stream(self.stream):
            if active is None:
                self.done.zero_()
            else:
                self.done.copy_(torch.from_numpy((~np.asarray(active, dtype=bool)).astype(np.int32)))
            N.call("ck_bcg_start", ctypes.byref(self.plan), N.ptr(b), float(tol_rel), float(tol_abs), st)
            steps = 0
            while steps < self.max_iters:
                n_it = min(int(chunk), self.max_iters - steps)
                if use_graph:
                    N.call("ck_graph_launch", self._graph(n_it), st)
                else:
                    N.call("ck_bcg_iterate", ctypes.byref(self.plan), n_it, st)
                steps += n_it
                self._pinned.copy_(self.done, non_blocking=True)
                self.stream.synchronize()
                if bool(self._pinned.all()):
                    break
            ks = self.k.cpu().numpy().copy()
        torch.cuda.current_stream(self.systems.device).wait_stream(self.stream)
        return self.x, ks


def time_matvec(bcg, reps=10):
    """Average ck_bcg_matvec duration (HIP events on the solver's stream)."""
    bcg.done.zero_()
    bcg.p.normal_()
    torch.cuda.synchronize()
    with torch.cuda.stream(bcg.stream):
        bcg.matvec()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(bcg.stream)
        for _ in range(reps):
            bcg.matvec()
        e1.record(bcg.stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps * 1e-3


def autotune(systems, bf16_mask=None, candidates=_ROWS, reps=10, max_iters=1000):
    """Time the batched matvec for each tile height and return (best rows_per_tile, timings in s).
    Results are cached per (sizes, precision mask)."""
    mask = np.zeros(systems.nsys, dtype=bool) if bf16_mask is None else np.asarray(bf16_mask, dtype=bool)
    key = (systems.sizes.tobytes(), mask.tobytes())
    if key in _TUNE_CACHE:
        return _TUNE_CACHE[key]
    timings = {}
    for R in candidates:
        bcg = BatchedCG(systems, R, max_iters)
        bcg.set_precision(mask)
        timings[R] = time_matvec(bcg, reps)
        del bcg
    best = min(timings, key=timings.get)
    _TUNE_CACHE[key] = (best, timings)
    return best, timings


@dataclass
class HipBatchedCG(Algorithm):
    """Batched CG over the blocks of a block-diagonal operator (one launch set for all blocks).

    Args:
        tol: relative residual target per block (||r|| <= tol ||b||).
        max_iters: step cap per block.
        chunk: iterations per graph replay between reads of the done flags.
        rows_per_tile: matvec tile height, or None to autotune.
        use_graph: replay captured iterations from a HIP graph.
    """
    tol: float = 1e-6
    max_iters: int = 1000
    chunk: int = 8
    rows_per_tile: object = None
    use_graph: bool = True


class BatchedMixedSolver:
    """FP64 refinement around batched CG: r = b - A x with the fp32 matrices as stored and FP64
    arithmetic, inner solves on bf16 (or fp32) storage, per-system fallback to fp32 storage when a
    refinement step does not cut the residual by ``contraction``: an fp32 correction solve cuts it
    by about the inner tolerance, so a bf16 system that contracts by less than 0.05 per step costs
    more refinements than its halved matrix traffic saves."""

    def __init__(self, systems, rows_per_tile=None, inner_max_iters=1000, bf16=True, precond=None):
        self.systems = systems
        mask = np.full(systems.nsys, bool(bf16 and systems.A16 is not None))
        if rows_per_tile is None:
            rows_per_tile, self.tune = autotune(systems, mask, max_iters=inner_max_iters)
        else:
            self.tune = None
        self.cg = BatchedCG(systems, rows_per_tile, inner_max_iters, precond=precond)
        self.bf16 = mask
        dev = systems.device
        self.x64 = torch.zeros(systems.total_v, dtype=torch.float64, device=dev)
        self.b64 = torch.zeros_like(self.x64)
        self.r32 = torch.zeros(systems.total_v, dtype=torch.float32, device=dev)
        self.prr = torch.zeros(self.cg.ntiles_mv, dtype=torch.float64, device=dev)
        self.pbb = torch.zeros_like(self.prr)
        self.rel = torch.zeros(systems.nsys, dtype=torch.float64, device=dev)
        self.active = torch.ones(systems.nsys, dtype=torch.int32, device=dev)
        self.admissible = None

    def screen_bf16(self, iters=8, margin=0.05, seed=0):
        """A-priori bf16 admissibility per system.  Refinement with corrections solved on the
        bf16 copy converges at rate ||A^-1 E|| <= ||E|| / lambda_min(A), E = K_bf16 - K, and
        lambda_min(A) >= shift for PSD K.  ||E||_2 is estimated by `iters` power iterations on E
        (two batched matvecs each, the shift cancels); a system keeps bf16 storage only if twice
        the estimate (power iteration approaches ||E|| from below) is at most margin * shift.
        Returns the per-system estimates; the mask is applied to later solves."""
        sy, cg = self.systems, self.cg
        S = sy.nsys
        if sy.A16 is None or not self.bf16.any():
            self.admissible = np.zeros(S, dtype=bool)
            return np.zeros(S)
        dev = sy.device
        seg = getattr(self, "_seg", None)
        if seg is None:
            ids = np.full(sy.total_v, -1, dtype=np.int64)
            for k in range(S):
                o, n = int(sy.voff[k]), int(sy.sizes[k])
                ids[o:o + n] = k
            valid = torch.from_numpy(ids >= 0).to(dev)
            seg = self._seg = (torch.from_numpy(np.maximum(ids, 0)).to(dev), valid)
        ids_t, valid = seg
        g = torch.Generator(device=dev).manual_seed(seed)
        est = torch.zeros(S, dtype=torch.float64, device=dev)
        self.stream_sync()
        with torch.cuda.stream(cg.stream):
            v = torch.randn(sy.total_v, generator=g, device=dev) * valid

            def seg_norm(x):
                return torch.zeros(S, dtype=torch.float64, device=dev).index_add_(0, ids_t, x.double()**2).sqrt()

            v = v / seg_norm(v).clamp_min(1e-300).float()[ids_t]
            cg.done.zero_()
            for _ in range(iters):
                cg.p.copy_(v)
                cg.set_precision(np.ones(S, dtype=bool))
                cg.matvec()
                y16 = cg.Ap.clone()
                cg.set_precision(np.zeros(S, dtype=bool))
                cg.matvec()
                ev = (y16 - cg.Ap) * valid
                est = seg_norm(ev)  # ||E v|| with ||v|| = 1
                v = ev / est.clamp_min(1e-300).float()[ids_t]
            est_h = est.cpu().numpy()
        torch.cuda.current_stream(dev).wait_stream(cg.stream)
        shifts = sy.shift.cpu().numpy().astype(np.float64)
        self.admissible = 2.0 * est_h <= margin * shifts
        self.bf16_norm_estimate = est_h
        return est_h

    def residual(self, active):
        cg = self.cg
        st = ctypes.c_void_p(cg.stream.cuda_stream)
        with torch.cuda.stream(cg.stream):
            self.active.copy_(torch.from_numpy(active.astype(np.int32)))
            N.call("ck_bcg_residual_f64", ctypes.byref(cg.plan), N.ptr(self.active), N.ptr(self.x64), N.ptr(self.b64),
                   N.ptr(self.r32), N.ptr(self.prr), N.ptr(self.pbb), N.ptr(self.rel), st)
            rel = self.rel.cpu().numpy().copy()
        return rel

    def solve(self, b, tol=1e-8, inner_tol=1e-4, max_refinements=20, contraction=0.05, chunk=8, use_graph=True):
        """b: concatenated rhs (fp32 or fp64).  Returns (x64 concatenated, info)."""
        t0 = time.time()
        S = self.systems.nsys
        cg = self.cg
        self.stream_sync()
        self.b64.copy_(b.to(torch.float64))
        self.x64.zero_()
        if self.bf16.any() and self.admissible is None:
            self.screen_bf16()
        prec = self.bf16 & self.admissible if self.admissible is not None else self.bf16.copy()
        bf16_start = int(prec.sum())
        cg.set_precision(prec)
        everyone = np.ones(S, dtype=bool)
        rel = self.residual(everyone)
        rel_all = rel.copy()
        inner_total = np.zeros(S, dtype=np.int64)
        fallbacks = 0
        it = 0
        history = [float(rel.max())]
        per_refinement = []  # (active systems, inner steps summed over them, bf16 systems among them)
        while it < max_refinements:
            active = ~(rel_all <= tol)
            if not active.any():
                break
            if not np.all(np.isfinite(rel_all)):
                raise FloatingPointError("non-finite FP64 residual in batched refinement")
            _, ks = cg.solve(self.r32, tol_rel=inner_tol, tol_abs=0.0, active=active, chunk=chunk,
                             use_graph=use_graph)
            inner_total[active] += ks[active]
            per_refinement.append((int(active.sum()), int(ks[active].sum()), int((active & prec).sum())))
            with torch.cuda.stream(cg.stream):
                keep = self.x64.clone()
                self.active.copy_(torch.from_numpy(active.astype(np.int32)))
                N.call("ck_bcg_update_f64", ctypes.byref(cg.plan), N.ptr(self.active), N.ptr(self.x64),
                       ctypes.c_void_p(cg.stream.cuda_stream))
            new = self.residual(active)
            broken = active & ~np.isfinite(new)
            if broken.any():
                # a correction solve broke down (bf16 storage can make K_s + shift I indefinite):
                # drop its correction, recompute that residual and continue on fp32 storage
                with torch.cuda.stream(cg.stream):
                    for s_ in np.flatnonzero(broken):
                        o, n_ = int(self.systems.voff[s_]), int(self.systems.sizes[s_])
                        self.x64[o:o + n_].copy_(keep[o:o + n_])
                redo = self.residual(broken)
                new = np.where(broken, redo, new)
                if not np.all(np.isfinite(new[broken])):
                    raise FloatingPointError("non-finite FP64 residual in batched refinement")
                stuck = broken & ~prec
                if stuck.any():
                    raise FloatingPointError("an fp32 correction solve broke down")
                prec = prec & ~broken
                fallbacks += int(broken.sum())
                cg.set_precision(prec)
            stalled = active & prec & ~(new <= contraction * rel_all)
            if stalled.any():
                prec = prec & ~stalled
                fallbacks += int(stalled.sum())
                cg.set_precision(prec)
            rel_all = np.where(active, new, rel_all)
            history.append(float(rel_all.max()))
            it += 1
        torch.cuda.synchronize()
        info = {"refinements": it, "rel_residual_f64": rel_all, "inner_iterations": inner_total,
                "bf16_systems_final": int(prec.sum()), "precision_fallbacks": fallbacks, "history": history,
                "per_refinement": per_refinement, "bf16_systems_start": bf16_start,
                "converged": bool(np.all(rel_all <= tol)), "time": time.time() - t0}
        return self.x64, info

    def stream_sync(self):
        self.cg.stream.wait_stream(torch.cuda.current_stream(self.systems.device))


def _dense_block(M):
    """(matrix, scale, shift) of a block that lowers to scale * Dense + shift * I (fp32, device)."""
    got = M.hip_op()
    if got is None or got[0].kind != N.CK_OP_DENSE_F32:
        raise N.UnsupportedOperator("batched CG", N.CK_ERR_UNSUPPORTED,
                                    "every block must be a dense fp32 device matrix (plus c * I)")
    op = got[0]

    def find(X):
        if isinstance(X, Dense):
            return X
        if isinstance(X, (Sum, Product)):
            for Y in X.Ms:
                d = find(Y)
                if d is not None:
                    return d
        return getattr(X, "A", None) if isinstance(getattr(X, "A", None), LinearOperator) and find(X.A) else None

    D = find(M)
    if D is None:
        raise N.UnsupportedOperator("batched CG", N.CK_ERR_UNSUPPORTED, "no dense block found")
    return D.A, float(op.scale), float(op.shift)


class BatchedInverse(LinearOperator):
    """inv(BlockDiag(...)) applied by batched CG over all blocks (one rhs column at a time)."""

    def __init__(self, A, alg):
        self.device = A.device
        super().__init__(A.dtype, A.shape)
        self.A, self.alg = A, alg
        mats, shifts = [], []
        for M in A.blocks():
            K, scale, shift = _dense_block(M)
            mats.append(K if scale == 1.0 else K * scale)
            shifts.append(shift)
        self.systems = BatchedSystems.dense(mats, shifts, A.device)
        rows = alg.rows_per_tile
        self.tune = None
        if rows is None:
            rows, self.tune = autotune(self.systems, max_iters=alg.max_iters)
        self.cg = BatchedCG(self.systems, rows, alg.max_iters)
        self.info = {}

    def _matmat(self, X):
        sy = self.systems
        cols = []
        iters = []
        sizes = sy.sizes.tolist()
        bounds = np.concatenate([[0], np.cumsum(sizes)])
        for j in range(X.shape[1]):
            col = X[:, j].to(torch.float32)
            b = sy.concat([col[bounds[s]:bounds[s + 1]] for s in range(sy.nsys)])
            x, ks = self.cg.solve(b, tol_rel=self.alg.tol, chunk=self.alg.chunk, use_graph=self.alg.use_graph)
            cols.append(torch.cat(sy.split(x)).clone())
            iters.append(ks)
        self.info = {"iterations": np.stack(iters), "rows_per_tile": self.cg.rows, "autotune": self.tune}
        return torch.stack(cols, dim=1).to(X.dtype)


class BatchedPrecond:
    """Per-system low-rank preconditioners (shift_s I + F_s F_s^T)^-1 for batched CG.

    F_s (n_s x rank, row-major) sits at element offset foff[s] of one fp32 buffer, and
    Cinv_s = (shift_s I + F_s^T F_s)^-1 (FP64) at s * rank^2.  ``captured[s]`` is
    tr(F_s F_s^T), so tr(K_s) - captured[s] is the trace of the Nystrom error."""

    def __init__(self, systems, rank, F, Cinv, captured, kind):
        self.systems, self.rank, self.F, self.Cinv, self.kind = systems, int(rank), F, Cinv, kind
        self.foff = np.concatenate([[0], np.cumsum(systems.sizes * rank)[:-1]]).astype(np.int64)
        self.foff_t = torch.tensor(self.foff, device=systems.device)
        self.captured = captured

    @staticmethod
    def _alloc(systems, rank):
        if rank < 1 or rank > 64 or 256 % rank:
            raise ValueError("rank must divide 256 and be at most 64")
        if int(systems.sizes.min()) < rank:
            raise ValueError("every system needs order >= rank")
        return torch.zeros(int((systems.sizes * rank).sum()), dtype=torch.float32, device=systems.device)

    @staticmethod
    def _cores(systems, F, rank):
        """Cinv_s = (shift_s I + F_s^T F_s)^-1 for every system (one batched small-SPD launch)."""
        from .learned import gram
        S = systems.nsys
        foff = np.concatenate([[0], np.cumsum(systems.sizes * rank)[:-1]]).astype(np.int64)
        G = torch.empty((S, rank, rank), dtype=torch.float64, device=systems.device)
        for s in range(S):
            n = int(systems.sizes[s])
            Fs = F[int(foff[s]):int(foff[s]) + n * rank]
            G[s].copy_(gram(Fs, rank, 1, Fs, rank, 1, n, rank, rank))
        captured = torch.diagonal(G, dim1=1, dim2=2).sum(1).cpu().numpy()
        G.diagonal(dim1=1, dim2=2).add_(systems.shift.double()[:, None])
        linv = torch.empty_like(G)
        inv = torch.empty_like(G)
        st = torch.zeros(S, dtype=torch.int32, device=systems.device)
        N.call("ck_spd_inverse_f64", S, rank, N.ptr(G), rank, rank * rank, 0.0, 0.0, N.ptr(linv), N.ptr(inv), None,
               None, N.ptr(st), N.stream_ptr(systems.device))
        if int(st.max()) != 0:
            raise np.linalg.LinAlgError("a Woodbury core is not positive definite")
        return inv.reshape(-1).contiguous(), captured

    @classmethod
    def pivoted_cholesky(cls, systems, rank=32):
        """Rank-r pivoted Cholesky of every K_s (ck_pivchol_dense_f32 on the fp32 matrices)."""
        F = cls._alloc(systems, rank)
        dev = systems.device
        st = N.stream_ptr(dev)
        nmax = int(systems.sizes.max())
        ldm = _round_up(nmax, 64)
        L = torch.zeros((rank, ldm), dtype=torch.float32, device=dev)
        piv = torch.empty(rank, dtype=torch.int32, device=dev)
        dpiv = torch.empty(rank, dtype=torch.float32, device=dev)
        wb = N.q("ck_pivchol_workspace_bytes", nmax)
        work = torch.empty((wb + 3) // 4, dtype=torch.float32, device=dev)
        off = 0
        for s in range(systems.nsys):
            n = int(systems.sizes[s])
            A = systems.A32[int(systems.moff[s]):]
            N.call("ck_pivchol_dense_f32", n, N.ptr(A), int(systems.lda[s]), rank, N.ptr(L), ldm, N.ptr(piv),
                   N.ptr(dpiv), N.ptr(work), wb, st)
            N.call("ck_transpose_f32", n, rank, N.ptr(L), ldm, N.ptr(F[off:]), rank, 0, st)
            off += n * rank
        Cinv, captured = cls._cores(systems, F, rank)
        return cls(systems, rank, F, Cinv, captured, "pivoted_cholesky")

    @classmethod
    def learned(cls, systems, points, net, jitter_rel=1e-6):
        """Nystrom completion F_s = K_s G_s (G_s^T K_s G_s)^-1/2 of the network's sketch G_s per system."""
        from .learned import gram_blocks, rows_gemm, spd_inverse
        from ..ops import _op
        rank = net.dout
        F = cls._alloc(systems, rank)
        dev = systems.device
        st = N.stream_ptr(dev)
        off = 0
        for s in range(systems.nsys):
            n = int(systems.sizes[s])
            G = net.forward(points[s].to(dev))
            ld = G.stride(0)
            op = _op(N.CK_OP_DENSE_F32, n)
            op.a = systems.A32[int(systems.moff[s]):].data_ptr()
            op.lda = int(systems.lda[s])
            Y = torch.zeros_like(G)
            wb = N.q("ck_op_workspace_bytes", op, rank)
            work = torch.empty(max(wb, 4) // 4, dtype=torch.float32, device=dev)
            N.call("ck_op_apply", op, rank, N.ptr(G), ld, N.ptr(Y), ld, N.ptr(work), wb, st)
            C = gram_blocks(G, Y, n)
            fac = spd_inverse(C, jitter_rel=jitter_rel, linvT32=True)
            rows_gemm(Y, 1, ld, fac["linvT32"], F[off:], rank, 1, n)
            off += n * rank
        Cinv, captured = cls._cores(systems, F, rank)
        return cls(systems, rank, F, Cinv, captured, "learned")

    @classmethod
    def best_of(cls, a, b):
        """Per system, the factor with the larger captured trace (smaller Nystrom error trace)."""
        if a.rank != b.rank or a.systems is not b.systems:
            raise ValueError("candidates must share systems and rank")
        sy, k = a.systems, a.rank
        F = torch.empty_like(a.F)
        Cinv = torch.empty_like(a.Cinv)
        pick = np.where(b.captured > a.captured, 1, 0)
        for s in range(sy.nsys):
            src = b if pick[s] else a
            o, m = int(a.foff[s]), int(sy.sizes[s]) * k
            F[o:o + m].copy_(src.F[o:o + m])
            Cinv[s * k * k:(s + 1) * k * k].copy_(src.Cinv[s * k * k:(s + 1) * k * k])
        out = cls(sy, k, F, Cinv, np.maximum(a.captured, b.captured), f"best({a.kind},{b.kind})")
        out.pick = pick
        return out
