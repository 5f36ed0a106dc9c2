"""ctypes binding of the C ABI declared in include/idft.h.

The same prototypes bind any library that implements the header.  The
product package loads only its own HIP library (invdft_amd/_lib/libidft.so,
see ``invdft_amd.load``); the test suite binds the CPU oracle separately.
"""
import ctypes as C
import os

import numpy as np

IDFT_OK, IDFT_EINVAL, IDFT_EHIP, IDFT_ENOCONV, IDFT_ENAN, IDFT_ESTATE, IDFT_ENOMEM = 0, -1, -2, -3, -4, -5, -6
FP64, MIXED = 0, 1
GRADIENT_REFERENCE, GRADIENT_EXACT = 0, 1
LS_SECANT_LOSS, LS_SECANT_FORCE_NORM, LS_CP = 0, 1, 2
# cell-product kernels (include/idft_hip.h)
KERNEL_ROW_PASS, KERNEL_TILE_PASS, KERNEL_SUM_FACTOR = 0, 1, 2

_DP = C.POINTER(C.c_double)
_IP = C.POINTER(C.c_int32)


class Desc(C.Structure):
    _fields_ = [
        ("n_cells", C.c_int32), ("degree", C.c_int32), ("half_width", C.c_double),
        ("n_eig", C.c_int32), ("n_electrons", C.c_double), ("precision", C.c_int32),
        ("temperature", C.c_double), ("cheb_degree", C.c_int32), ("max_cheb_passes", C.c_int32),
        ("cheb_tol", C.c_double), ("initial_tol_cheb", C.c_double),
        ("adaptive_factor_cheb", C.c_double), ("adaptive_factor_adjoint", C.c_double),
        ("adjoint_initial_tol", C.c_double), ("adjoint_max_iter", C.c_int32),
        ("degeneracy_tol", C.c_double), ("frac_occ_tol", C.c_double),
        ("additional_states", C.c_int32), ("block_size", C.c_int32),
        ("rho_tol_constraints", C.c_double), ("gradient_mode", C.c_int32),
        ("lanczos_steps", C.c_int32), ("psi_checkpoint", C.c_int32), ("check_every", C.c_int32),
        ("seed", C.c_uint64), ("minres_graph", C.c_int32), ("relax_ratio", C.c_double),
        ("minres_fused", C.c_int32),
    ]


class Sizes(C.Structure):
    _fields_ = [("n_nodes", C.c_int64), ("n_cells", C.c_int64), ("nodes_per_cell", C.c_int64),
                ("quad_per_cell", C.c_int64), ("n_vxc", C.c_int64),
                ("child_quad_per_cell", C.c_int64)]


class Stats(C.Structure):
    _fields_ = [("t_eigen_ms", C.c_double), ("t_density_ms", C.c_double),
                ("t_adjoint_ms", C.c_double), ("t_force_ms", C.c_double),
                ("t_total_ms", C.c_double), ("cheb_passes", C.c_int32),
                ("minres_iters_max", C.c_int32), ("minres_iters_total", C.c_int64),
                ("refinements", C.c_int32), ("max_residual", C.c_double),
                ("fermi_energy", C.c_double), ("cheb_tol_used", C.c_double),
                ("adjoint_tol_used", C.c_double), ("loss_unweighted", C.c_double),
                ("l1_density_error", C.c_double), ("int_rho_veff", C.c_double),
                ("n_evaluations", C.c_int64), ("spectrum_upper_bound", C.c_double),
                ("applies_fp64", C.c_int64), ("applies_fp32", C.c_int64),
                ("apply_cols_fp64", C.c_int64), ("apply_cols_fp32", C.c_int64),
                ("adjoint_blocks", C.c_int32), ("psi_recomputes", C.c_int32),
                ("minres_compactions", C.c_int32), ("minres_column_steps", C.c_int64)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


_PROTOS = {
    "idft_desc_default": (None, [C.POINTER(Desc)]),
    "idft_geometry_sizes": (C.c_int, [C.POINTER(Desc), C.POINTER(Sizes)]),
    "idft_geometry_quad_points": (C.c_int, [C.POINTER(Desc), _DP]),
    "idft_geometry_node_points": (C.c_int, [C.POINTER(Desc), _DP]),
    "idft_geometry_vxc_points": (C.c_int, [C.POINTER(Desc), _DP]),
    "idft_basis_tables": (C.c_int, [C.c_int32, _DP, _DP, _DP, _DP]),
    "idft_create": (C.c_int, [C.POINTER(Desc), _DP, _DP, _DP, _DP, C.POINTER(C.c_void_p)]),
    "idft_destroy": (None, [C.c_void_p]),
    "idft_last_error": (C.c_char_p, [C.c_void_p]),
    "idft_get_sizes": (C.c_int, [C.c_void_p, C.POINTER(Sizes)]),
    "idft_get_stats": (C.c_int, [C.c_void_p, C.POINTER(Stats)]),
    "idft_backend": (C.c_char_p, []),
    "idft_get_force": (C.c_int, [C.c_void_p, _DP, _DP, _DP]),
    "idft_dot": (C.c_int, [C.c_void_p, _DP, _DP, C.c_int64, C.c_int32, _DP]),
    "idft_set_initial_guess": (C.c_int, [C.c_void_p, _DP]),
    "idft_get_initial_guess": (C.c_int, [C.c_void_p, _DP]),
    "idft_set_solution": (C.c_int, [C.c_void_p, _DP]),
    "idft_get_solution": (C.c_int, [C.c_void_p, _DP]),
    "idft_forward": (C.c_int, [C.c_void_p, _DP, _DP]),
    "idft_backward": (C.c_int, [C.c_void_p, _DP]),
    "idft_set_potential": (C.c_int, [C.c_void_p, _DP]),
    "idft_set_veff_quad": (C.c_int, [C.c_void_p, _DP]),
    "idft_get_veff_quad": (C.c_int, [C.c_void_p, _DP]),
    "idft_eigensolve": (C.c_int, [C.c_void_p, C.c_double, C.c_int32, _DP, _DP, _IP]),
    "idft_get_eigenvectors": (C.c_int, [C.c_void_p, _DP]),
    "idft_set_eigenvectors": (C.c_int, [C.c_void_p, _DP]),
    "idft_get_occupancies": (C.c_int, [C.c_void_p, _DP, _DP]),
    "idft_get_eigenvalues": (C.c_int, [C.c_void_p, _DP]),
    "idft_get_density": (C.c_int, [C.c_void_p, _DP]),
    "idft_apply_hamiltonian": (C.c_int, [C.c_void_p, _DP, _DP, C.c_int32, C.c_int32]),
    "idft_chebyshev_filter": (C.c_int, [C.c_void_p, _DP, C.c_int32, C.c_int32, C.c_double,
                                        C.c_double, C.c_double, C.c_int32]),
    "idft_mv_create": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_void_p)]),
    "idft_mv_destroy": (None, [C.c_void_p]),
    "idft_mv_set": (C.c_int, [C.c_void_p, C.c_void_p, _DP]),
    "idft_mv_get": (C.c_int, [C.c_void_p, C.c_void_p, _DP]),
    "idft_adj_update_input_psi": (C.c_int, [C.c_void_p, C.c_void_p, _DP, _DP, _IP, C.c_int32,
                                            C.c_double, _DP, C.c_int32]),
    "idft_adj_compute_rhs": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "idft_adj_vmult": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32]),
    "idft_adj_precondition_jacobi": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double]),
    "idft_adj_precondition_jacobi_sqrt": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double]),
    "idft_adj_distribute_x": (C.c_int, [C.c_void_p]),
    "idft_adj_dot_quad_wise": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, _DP]),
    "idft_minres_solve": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_int32,
                                    C.c_int32, _IP]),
    "idft_bfgs_solve": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_double, C.c_double,
                                  C.c_int32, C.c_int32, _DP, _DP, _DP, _IP]),
}

SYMBOLS = tuple(_PROTOS)


def bind(lib):
    """Attach argtypes/restype for every ABI symbol; raises if one is missing."""
    for name, (res, args) in _PROTOS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


class IdftError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"idft error {code}: {msg}")
        self.code = code


def _dp(a):
    return a.ctypes.data_as(_DP)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def default_desc(lib, **kw):
    d = Desc()
    lib.idft_desc_default(C.byref(d))
    for k, v in kw.items():
        if not hasattr(d, k):
            raise KeyError(k)
        setattr(d, k, v)
    return d


def geometry(lib, desc):
    s = Sizes()
    rc = lib.idft_geometry_sizes(C.byref(desc), C.byref(s))
    if rc:
        raise IdftError(rc, "invalid descriptor")
    return s


def quad_points(lib, desc):
    s = geometry(lib, desc)
    xyz = np.empty((s.n_cells * s.quad_per_cell, 3))
    lib.idft_geometry_quad_points(C.byref(desc), _dp(xyz))
    return xyz


def node_points(lib, desc):
    s = geometry(lib, desc)
    xyz = np.empty((s.n_nodes, 3))
    lib.idft_geometry_node_points(C.byref(desc), _dp(xyz))
    return xyz


def vxc_points(lib, desc):
    s = geometry(lib, desc)
    xyz = np.empty((s.n_vxc, 3))
    lib.idft_geometry_vxc_points(C.byref(desc), _dp(xyz))
    return xyz


def basis_tables(lib, p):
    gx, gw = np.empty(p + 1), np.empty(p + 1)
    qx, qw = np.empty(p + 3), np.empty(p + 3)
    rc = lib.idft_basis_tables(p, _dp(gx), _dp(gw), _dp(qx), _dp(qw))
    if rc:
        raise IdftError(rc, "invalid degree")
    return gx, gw, qx, qw


# include/idft_io.h (HIP library only)
_IO_PROTOS = {
    "idft_write_vxc_text": (C.c_int, [C.POINTER(Desc), C.c_char_p, _DP, C.c_int32]),
    "idft_read_vxc_text": (C.c_int, [C.POINTER(Desc), C.c_char_p, C.c_double, C.c_int32, _DP]),
    "idft_write_vxc_binary": (C.c_int, [C.POINTER(Desc), C.c_char_p, _DP, C.c_int32]),
    "idft_read_vxc_binary": (C.c_int, [C.POINTER(Desc), C.c_char_p, C.c_int32, _DP]),
    "idft_set_vxc_output": (C.c_int, [C.c_void_p, C.c_char_p, C.c_char_p, C.c_int32]),
    "idft_bfgs_set_checkpoint": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int32, C.c_int32]),
}


def _io(lib, name):
    fn = getattr(lib, name)
    fn.restype, fn.argtypes = _IO_PROTOS[name]
    return fn


def _path(p):
    return os.fsencode(p)


def write_vxc(lib, desc, path, pot, binary=False):
    """Write v_xc (n_vxc values, or (n_spin, n_vxc)) as text records or binary."""
    pot = _f64(pot)
    n_spin = 1 if pot.ndim == 1 else pot.shape[0]
    name = "idft_write_vxc_binary" if binary else "idft_write_vxc_text"
    rc = _io(lib, name)(C.byref(desc), _path(path), _dp(pot), n_spin)
    if rc:
        raise IdftError(rc, f"cannot write {path}")


def read_vxc(lib, desc, path, n_spin=1, binary=False, dist_tol=1e-3):
    """Read a v_xc file written by write_vxc (or by the reference's text writer)."""
    n = geometry(lib, desc).n_vxc
    pot = np.zeros(n_spin * n)
    if binary:
        rc = _io(lib, "idft_read_vxc_binary")(C.byref(desc), _path(path), n_spin, _dp(pot))
    else:
        rc = _io(lib, "idft_read_vxc_text")(C.byref(desc), _path(path), dist_tol, n_spin,
                                            _dp(pot))
    if rc:
        raise IdftError(rc, f"cannot read {path}")
    return pot if n_spin == 1 else pot.reshape(n_spin, n)


class MultiVector:
    """Context-owned node-major multivector handle (idft_mv)."""

    def __init__(self, ctx, ncols):
        self.ctx, self.ncols = ctx, ncols
        h = C.c_void_p()
        ctx._check(ctx.lib.idft_mv_create(ctx.h, ncols, C.byref(h)))
        self.h = h

    def set(self, a):
        a = _f64(a).reshape(-1)
        if a.size != self.ctx.sizes.n_nodes * self.ncols:
            raise ValueError("multivector size mismatch")
        self.ctx._check(self.ctx.lib.idft_mv_set(self.ctx.h, self.h, _dp(a)))
        return self

    def get(self):
        out = np.empty((self.ctx.sizes.n_nodes, self.ncols))
        self.ctx._check(self.ctx.lib.idft_mv_get(self.ctx.h, self.h, _dp(out)))
        return out

    def __del__(self):
        if getattr(self, "h", None) and self.ctx.h:
            self.ctx.lib.idft_mv_destroy(self.h)
            self.h = None


class Context:
    """One idft_ctx: the inverse-DFT objective plus its adjoint problem."""

    def __init__(self, lib, desc, pot_base_q, rho_target_q, weight_q, rho_target_nodal):
        self.lib, self.desc = lib, desc
        self.h = None
        self._keep = [_f64(pot_base_q), _f64(rho_target_q), _f64(weight_q), _f64(rho_target_nodal)]
        h = C.c_void_p()
        rc = lib.idft_create(C.byref(desc), *[_dp(a) for a in self._keep], C.byref(h))
        if rc:
            raise IdftError(rc, "idft_create failed")
        self.h = h
        self.sizes = Sizes()
        lib.idft_get_sizes(h, C.byref(self.sizes))
        self.nq_total = self.sizes.n_cells * self.sizes.quad_per_cell

    def close(self):
        if self.h:
            self.lib.idft_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != IDFT_OK:
            msg = self.lib.idft_last_error(self.h) if self.h else b""
            raise IdftError(rc, (msg or b"").decode())

    # ---- contract 1
    def get_force(self, pot):
        pot = _f64(pot).copy()
        force = np.zeros(self.sizes.n_vxc)
        loss = np.zeros(1)
        self._check(self.lib.idft_get_force(self.h, _dp(pot), _dp(force), _dp(loss)))
        return force, float(loss[0])

    def forward(self, pot):
        pot = _f64(pot)
        loss = np.zeros(1)
        self._check(self.lib.idft_forward(self.h, _dp(pot), _dp(loss)))
        return float(loss[0])

    def backward(self):
        force = np.zeros(self.sizes.n_vxc)
        self._check(self.lib.idft_backward(self.h, _dp(force)))
        return force

    def dot(self, v1, v2, block=1):
        v1, v2 = _f64(v1), _f64(v2)
        out = np.zeros(block)
        self._check(self.lib.idft_dot(self.h, _dp(v1), _dp(v2), v1.size, block, _dp(out)))
        return out

    def set_initial_guess(self, pot):
        self._check(self.lib.idft_set_initial_guess(self.h, _dp(_f64(pot))))

    def get_initial_guess(self):
        out = np.zeros(self.sizes.n_vxc)
        self._check(self.lib.idft_get_initial_guess(self.h, _dp(out)))
        return out

    def get_solution(self):
        out = np.zeros(self.sizes.n_vxc)
        self._check(self.lib.idft_get_solution(self.h, _dp(out)))
        return out

    # ---- forward pieces
    def set_potential(self, pot):
        self._check(self.lib.idft_set_potential(self.h, _dp(_f64(pot))))

    def set_veff_quad(self, v):
        self._check(self.lib.idft_set_veff_quad(self.h, _dp(_f64(v))))

    def veff_quad(self):
        out = np.zeros(self.nq_total)
        self._check(self.lib.idft_get_veff_quad(self.h, _dp(out)))
        return out

    def eigensolve(self, tol, max_passes=100):
        ev = np.zeros(self.desc.n_eig)
        mr = np.zeros(1)
        passes = C.c_int32(0)
        self._check(self.lib.idft_eigensolve(self.h, tol, max_passes, _dp(ev), _dp(mr),
                                             C.byref(passes)))
        return ev, float(mr[0]), passes.value

    def eigenvectors(self):
        out = np.zeros((self.sizes.n_nodes, self.desc.n_eig))
        self._check(self.lib.idft_get_eigenvectors(self.h, _dp(out)))
        return out

    def set_eigenvectors(self, psi):
        self._check(self.lib.idft_set_eigenvectors(self.h, _dp(_f64(psi))))

    def occupancies(self):
        occ = np.zeros(self.desc.n_eig)
        ef = np.zeros(1)
        self._check(self.lib.idft_get_occupancies(self.h, _dp(occ), _dp(ef)))
        return occ, float(ef[0])

    def eigenvalues(self):
        ev = np.zeros(self.desc.n_eig)
        self._check(self.lib.idft_get_eigenvalues(self.h, _dp(ev)))
        return ev

    def density(self):
        out = np.zeros(self.nq_total)
        self._check(self.lib.idft_get_density(self.h, _dp(out)))
        return out

    def apply_hamiltonian(self, x, precision=FP64):
        x = _f64(x)
        ncols = x.shape[1] if x.ndim == 2 else 1
        y = np.zeros_like(x)
        self._check(self.lib.idft_apply_hamiltonian(self.h, _dp(x), _dp(y), ncols, precision))
        return y

    def chebyshev_filter(self, x, degree, a, b, a0, precision=FP64):
        x = _f64(x).copy()
        ncols = x.shape[1] if x.ndim == 2 else 1
        self._check(self.lib.idft_chebyshev_filter(self.h, _dp(x), ncols, degree, a, b, a0,
                                                   precision))
        return x

    def stats(self):
        s = Stats()
        self._check(self.lib.idft_get_stats(self.h, C.byref(s)))
        return s.as_dict()

    # ---- contract 2
    def mv(self, ncols, data=None):
        m = MultiVector(self, ncols)
        if data is not None:
            m.set(data)
        return m

    def adj_update_input_psi(self, psi_mv, occ, u_q, pairs, fermi, eigenvalues):
        occ = _f64(occ)
        u_q = _f64(u_q)
        pairs = np.ascontiguousarray(pairs, dtype=np.int32).reshape(-1)
        ev = _f64(eigenvalues).copy()
        self._psi_keep = psi_mv
        self._check(self.lib.idft_adj_update_input_psi(
            self.h, psi_mv.h, _dp(occ), _dp(u_q), pairs.ctypes.data_as(_IP), pairs.size // 2,
            fermi, _dp(ev), psi_mv.ncols))
        return ev  # negated in place, as the reference does

    def adj_compute_rhs(self, ndbc, x, rhs):
        self._check(self.lib.idft_adj_compute_rhs(self.h, ndbc.h if ndbc else None, x.h, rhs.h))

    def adj_vmult(self, Ax, x):
        self._check(self.lib.idft_adj_vmult(self.h, Ax.h, x.h, x.ncols))

    def adj_precondition_jacobi(self, dst, src, omega=1.0):
        self._check(self.lib.idft_adj_precondition_jacobi(self.h, dst.h, src.h, omega))

    def adj_precondition_jacobi_sqrt(self, dst, src, omega=1.0):
        self._check(self.lib.idft_adj_precondition_jacobi_sqrt(self.h, dst.h, src.h, omega))

    def adj_distribute_x(self):
        self._check(self.lib.idft_adj_distribute_x(self.h))

    def adj_dot_quad_wise(self, v1, v2, n_pairs):
        out = np.zeros(n_pairs)
        self._check(self.lib.idft_adj_dot_quad_wise(self.h, v1.h, v2.h, _dp(out)))
        return out

    def minres_solve(self, x, ndbc, tol, max_iter, distribute=True):
        iters = np.zeros(x.ncols, dtype=np.int32)
        self._check(self.lib.idft_minres_solve(self.h, x.h, ndbc.h if ndbc else None, tol,
                                               max_iter, int(distribute),
                                               iters.ctypes.data_as(_IP)))
        return iters

    def set_vxc_output(self, folder, postfix, frequency):
        """Write every frequency-th evaluation's input v_xc as text records."""
        self._check(_io(self.lib, "idft_set_vxc_output")(self.h, _path(folder), _path(postfix),
                                                          frequency))

    def set_bfgs_checkpoint(self, path, every=1, resume=True):
        """L-BFGS state file written every `every` iterations (None: off)."""
        self._check(_io(self.lib, "idft_bfgs_set_checkpoint")(
            self.h, None if path is None else _path(path), every, int(resume)))

    def bfgs_solve(self, ls_type=LS_CP, max_iter=10, tol=1e-12, ls_tol=1e-6, history=100,
                   n_line_search=1):
        lh = np.zeros(max_iter + 1)
        gh = np.zeros(max_iter + 1)
        th = np.zeros(max_iter + 1)
        nd = C.c_int32(0)
        self._check(self.lib.idft_bfgs_solve(self.h, ls_type, max_iter, tol, ls_tol, history,
                                             n_line_search, _dp(lh), _dp(gh), _dp(th),
                                             C.byref(nd)))
        self.bfgs_times_ms = th[: nd.value + 1]
        return lh[: nd.value + 1], gh[: nd.value + 1], nd.value

    def set_apply_shape(self, precision, cells_per_wg, vectors_per_tile, gather_rows=0,
                        kernel=KERNEL_ROW_PASS):
        """Cell-product kernel and launch shape at one precision (HIP library
        only); for KERNEL_SUM_FACTOR vectors_per_tile is the number of columns
        per work-group (2, 4, 8) and cells_per_wg must be 1."""
        fn = self.lib.idft_set_apply_shape
        fn.restype = C.c_int
        fn.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32]
        self._check(fn(self.h, precision, cells_per_wg, vectors_per_tile, kernel, gather_rows))

    def autotune_apply(self, ncols, precision=FP64, reps=5):
        """Pick the fastest operator launch shape (HIP library only)."""
        cp, vt, tp, gr = C.c_int32(0), C.c_int32(0), C.c_int32(0), C.c_int32(0)
        ms = np.zeros(1)
        fn = self.lib.idft_autotune_apply
        fn.restype = C.c_int
        fn.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, _IP, _IP, _IP, _IP, _DP]
        self._check(fn(self.h, ncols, precision, reps, C.byref(cp), C.byref(vt), C.byref(tp),
                       C.byref(gr), _dp(ms)))
        return dict(cells_per_wg=cp.value, vectors_per_tile=vt.value, kernel=tp.value,
                    gather_rows=gr.value, ms=float(ms[0]))

    def profile_apply(self, ncols, precision=FP64, reps=10):
        """Device ms per Hamiltonian application and for its MFMA cell-product
        kernel alone (HIP library only)."""
        out = np.zeros(2)
        fn = self.lib.idft_profile_apply
        fn.restype = C.c_int
        fn.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, _DP, _DP]
        self._check(fn(self.h, ncols, precision, reps, _dp(out[:1]), _dp(out[1:])))
        return float(out[0]), float(out[1])
