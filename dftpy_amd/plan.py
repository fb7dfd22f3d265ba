"""Device plan: one grid, one precision, one stream (wraps ofdft_plan).

All arrays handed to a Plan are torch tensors resident on the GPU, in the
plan precision, C-contiguous.  The plan checks dtype, device, shape and
contiguity before every call so that a kernel never sees a buffer of the
wrong size.
"""
import ctypes

import numpy as np
import torch

from . import _lib as L

_PREC = {"fp64": L.F64, "fp32": L.F32}


def require_gpu():
    if not torch.cuda.is_available():
        raise RuntimeError("dftpy_amd runs on a ROCm GPU; torch.cuda.is_available() is False")


class Plan:
    """ofdft_plan for an n0 x n1 x n2 grid with lattice rows in bohr."""

    def __init__(self, nr, lattice, precision="fp64", full=False, device=None):
        require_gpu()
        self.lib = L.load()
        if precision not in _PREC:
            raise ValueError("precision must be 'fp64' or 'fp32'")
        self.precision = precision
        self.prec_code = _PREC[precision]
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        self.dtype = torch.float64 if precision == "fp64" else torch.float32
        self.cdtype = torch.complex128 if precision == "fp64" else torch.complex64
        self.nr = tuple(int(x) for x in nr)
        self.nh = self.nr[2] // 2 + 1
        self.N = int(np.prod(self.nr))
        self.shape_c = (self.nr[0], self.nr[1], self.nh)
        self.full = bool(full)
        self.lattice = np.ascontiguousarray(np.asarray(lattice, dtype=np.float64).reshape(3, 3))
        self.volume = abs(np.linalg.det(self.lattice))
        self.dV = self.volume / self.N
        self._h = ctypes.c_void_p()
        lat = (ctypes.c_double * 9)(*self.lattice.ravel().tolist())
        with torch.cuda.device(self.device):
            stream = torch.cuda.current_stream(self.device).cuda_stream
            L.check(self.lib.ofdft_plan_create(ctypes.byref(self._h), self.nr[0], self.nr[1], self.nr[2], lat,
                                               self.prec_code, 1 if self.full else 0, ctypes.c_void_p(stream)))
        self.functional = None
        self.bound = False
        # column tile widths measured fastest for these lengths (tuning.py)
        from . import tuning
        self.tiles = tuning.tiles_for(self.nr, precision)
        if self.tiles != (tuning.DEFAULT, tuning.DEFAULT):
            self.set_tiles(*self.tiles)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                self.lib.ofdft_plan_destroy(h)
            except Exception:
                pass
            self._h = None

    # ------------------------------------------------------------------
    def _stream(self):
        if self.bound:
            return
        s = torch.cuda.current_stream(self.device).cuda_stream
        L.check(self.lib.ofdft_plan_set_stream(self._h, ctypes.c_void_p(s)))

    def bind_stream(self, on=True):
        """Pin the plan to the current torch stream and skip the per-call
        stream lookup (inner loops); bind_stream(False) restores it."""
        self.bound = False
        if on:
            self._stream()
            self.bound = True

    def tn_inner(self, phi, p, W0, Ap, d, r, r0):
        """Fused inner TN step on buffers owned by the caller (no argument
        checks: the optimizer allocates them from this plan)."""
        out = (ctypes.c_double * 2)()
        self._stream()
        L.check(self.lib.ofdft_tn_inner(self._h, phi.data_ptr(), p.data_ptr(), W0.data_ptr(), Ap.data_ptr(),
                                        d.data_ptr(), r.data_ptr(), float(r0), out))
        return out[0], out[1]

    def tn_direction(self, phi, res, W0, d, work, maxfun, chunk=4):
        """get_direction_TN with the exact HVP as one device-resident loop.
        ``work`` holds 4 fields (r, p, Ap, Best).  Returns (number, status,
        pAp of the last iteration)."""
        n, st, pAp = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        self._stream()
        L.check(self.lib.ofdft_tn_direction(self._h, self._real(phi, "phi"), self._real(res, "res"),
                                            self._real(W0, "W0"), self._real(d, "d"),
                                            self._realn(work, "work", 4 * self.N), int(maxfun), int(chunk),
                                            ctypes.byref(n), ctypes.byref(st), ctypes.byref(pAp)))
        return n.value, st.value, pAp.value

    def _real(self, t, name):
        if t is None:
            return None
        if not isinstance(t, torch.Tensor):
            raise TypeError("%s must be a torch tensor" % name)
        if t.device != self.device or t.dtype != self.dtype:
            raise TypeError("%s must be %s on %s (got %s on %s)" % (name, self.dtype, self.device, t.dtype, t.device))
        if t.numel() != self.N or not t.is_contiguous():
            raise ValueError("%s must be a contiguous field of %d values" % (name, self.N))
        return L.ptr(t)

    def _realn(self, t, name, n):
        if t.device != self.device or t.dtype != self.dtype or t.numel() != n or not t.is_contiguous():
            raise ValueError("%s must be a contiguous %s tensor of %d values on %s" % (name, self.dtype, n, self.device))
        return L.ptr(t)

    def _cplx(self, t, name):
        if t.device != self.device or t.dtype != self.cdtype:
            raise TypeError("%s must be %s on %s" % (name, self.cdtype, self.device))
        if t.numel() != int(np.prod(self.shape_c)) or not t.is_contiguous():
            raise ValueError("%s must be a contiguous half spectrum %s" % (name, self.shape_c))
        return L.ptr(t)

    def empty(self):
        return torch.empty(self.nr, dtype=self.dtype, device=self.device)

    def zeros(self):
        return torch.zeros(self.nr, dtype=self.dtype, device=self.device)

    # ------------------------------------------------------------------
    def set_functional(self, tf=False, vw=False, wt=False, lda=False, hartree=False, pseudo=False, x=1.0, y=1.0,
                       alpha=5.0 / 6.0, beta=5.0 / 6.0, rho0=0.0):
        f = L.Functional(int(tf), int(vw), int(wt), int(lda), int(hartree), int(pseudo), float(x), float(y),
                         float(alpha), float(beta), float(rho0))
        self._stream()
        L.check(self.lib.ofdft_set_functional(self._h, ctypes.byref(f)))
        self.functional = dict(tf=tf, vw=vw, wt=wt, lda=lda, hartree=hartree, pseudo=pseudo, x=x, y=y, alpha=alpha,
                               beta=beta, rho0=rho0)

    def set_vloc(self, vloc):
        self._stream()
        L.check(self.lib.ofdft_set_vloc(self._h, self._real(vloc, "vloc")))

    def eval(self, rho, V=None):
        sums = L.darr(L.NSUM)
        self._stream()
        L.check(self.lib.ofdft_eval(self._h, self._real(rho, "rho"), self._real(V, "V"), sums))
        return L.as_numpy(sums)

    def eval_phi(self, phi, p=None, c=1.0, s=0.0, V=None, aux=None):
        sums = L.darr(L.NSUM)
        self._stream()
        auxp = self._realn(aux, "aux", 2 * self.N) if aux is not None else None
        L.check(self.lib.ofdft_eval_phi(self._h, self._real(phi, "phi"), self._real(p, "p"), float(c), float(s),
                                        self._real(V, "V"), auxp, sums))
        return L.as_numpy(sums)

    def gga(self, functional, params, rho=None, phi=None, p=None, c=1.0, s=0.0, V=None):
        """ofdft_gga: the GGA kinetic term at rho or (c phi + s p)^2; V += V_GGA
        when V is given.  Returns the OFDFT_GGA_NOUT sums."""
        out = L.darr(L.GGA_NOUT)
        prm = (ctypes.c_double * max(len(params), 1))(*params)
        self._stream()
        L.check(self.lib.ofdft_gga(self._h, int(functional), prm, len(params), self._real(rho, "rho"),
                                   self._real(phi, "phi"), self._real(p, "p"), float(c), float(s), self._real(V, "V"),
                                   out))
        return L.as_numpy(out)

    def lwt_set_kernel(self, kernel, deriv):
        """ofdft_lwt_set_kernel with two L.LwtTable (host arrays kept alive
        by the caller for the duration of the call)."""
        self._stream()
        L.check(self.lib.ofdft_lwt_set_kernel(self._h, ctypes.byref(kernel), ctypes.byref(deriv)))

    def lwt_range(self, rho=None, phi=None, p=None, c=1.0, s=0.0):
        """(min, max) of rho or (c phi + s p)^2."""
        out = L.darr(2)
        self._stream()
        L.check(self.lib.ofdft_lwt_range(self._h, self._real(rho, "rho"), self._real(phi, "phi"),
                                         self._real(p, "p"), float(c), float(s), out))
        return L.as_numpy(out)

    def lwt(self, params, kflist, rho=None, phi=None, p=None, c=1.0, s=0.0, V=None):
        """ofdft_lwt: the LWT term with the node list kflist; V += V_LWT when
        V is given.  Returns [E, 2 sum V phi dphi, sum V rho] (x dV)."""
        out = L.darr(3)
        kfl = np.ascontiguousarray(kflist, dtype=np.float64)
        self._stream()
        L.check(self.lib.ofdft_lwt(self._h, ctypes.byref(params), kfl.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                   int(kfl.size), self._real(rho, "rho"), self._real(phi, "phi"), self._real(p, "p"),
                                   float(c), float(s), self._real(V, "V"), out))
        return L.as_numpy(out)

    def hvp(self, phi, p, V, aux, mu, Ap):
        out = ctypes.c_double()
        self._stream()
        auxp = self._realn(aux, "aux", 2 * self.N) if aux is not None else None
        L.check(self.lib.ofdft_hvp(self._h, self._real(phi, "phi"), self._real(p, "p"), self._real(V, "V"), auxp,
                                   float(mu), self._real(Ap, "Ap"), ctypes.byref(out)))
        return out.value

    def potential_jvp(self, phi, p, V, aux, dV):
        """dV = (d V / d rho)[2 phi p] at rho = phi^2."""
        self._stream()
        auxp = self._realn(aux, "aux", 2 * self.N) if aux is not None else None
        L.check(self.lib.ofdft_potential_jvp(self._h, self._real(phi, "phi"), self._real(p, "p"), self._real(V, "V"),
                                             auxp, self._real(dV, "dV")))
        return dV

    def hvp_prepare(self, phi, V, aux, mu, res, W0):
        out = ctypes.c_double()
        self._stream()
        auxp = self._realn(aux, "aux", 2 * self.N) if aux is not None else None
        L.check(self.lib.ofdft_hvp_prepare(self._h, self._real(phi, "phi"), self._real(V, "V"), auxp, float(mu),
                                           self._real(res, "res"), self._real(W0, "W0"), ctypes.byref(out)))
        return out.value

    def hvp_fast(self, phi, p, W0, Ap):
        out = ctypes.c_double()
        self._stream()
        L.check(self.lib.ofdft_hvp_fast(self._h, self._real(phi, "phi"), self._real(p, "p"), self._real(W0, "W0"),
                                        self._real(Ap, "Ap"), ctypes.byref(out)))
        return out.value

    def r2c(self, f, out=None):
        if out is None:
            out = torch.empty(self.shape_c, dtype=self.cdtype, device=self.device)
        self._stream()
        L.check(self.lib.ofdft_fft_r2c(self._h, self._real(f, "in"), self._cplx(out, "out")))
        return out

    def c2r(self, F, out=None):
        if out is None:
            out = self.empty()
        self._stream()
        L.check(self.lib.ofdft_fft_c2r(self._h, self._cplx(F, "in"), self._real(out, "out")))
        return out

    def convolve(self, kind, f, out=None):
        if out is None:
            out = self.empty()
        self._stream()
        L.check(self.lib.ofdft_convolve(self._h, int(kind), self._real(f, "in"), self._real(out, "out")))
        return out

    # vector kernels --------------------------------------------------
    def dot3(self, a, b):
        out = L.darr(3)
        self._stream()
        L.check(self.lib.ofdft_dot3(self._h, self._real(a, "a"), self._real(b, "b"), out))
        return out[0], out[1], out[2]

    def lincomb(self, alpha, x, beta, y, z):
        self._stream()
        L.check(self.lib.ofdft_lincomb(self._h, float(alpha), self._real(x, "x"), float(beta), self._real(y, "y"),
                                       self._real(z, "z")))
        return z

    def residual(self, V, phi, mu, res):
        out = ctypes.c_double()
        self._stream()
        L.check(self.lib.ofdft_residual(self._h, self._real(V, "V"), self._real(phi, "phi"), float(mu),
                                        self._real(res, "res"), ctypes.byref(out)))
        return out.value

    def cg_update(self, a, p, Ap, d, r):
        out = ctypes.c_double()
        self._stream()
        L.check(self.lib.ofdft_cg_update(self._h, float(a), self._real(p, "p"), self._real(Ap, "Ap"),
                                         self._real(d, "d"), self._real(r, "r"), ctypes.byref(out)))
        return out.value

    def dot_vphip(self, V, phi, p):
        out = ctypes.c_double()
        self._stream()
        L.check(self.lib.ofdft_dot_vphip(self._h, self._real(V, "V"), self._real(phi, "phi"), self._real(p, "p"),
                                         ctypes.byref(out)))
        return out.value

    def orth_sums(self, p, phi, V):
        """(sum p phi, sum phi^2, sum p^2, sum V phi p, sum V phi phi) in one pass."""
        out = L.darr(5)
        self._stream()
        L.check(self.lib.ofdft_orth_sums(self._h, self._real(p, "p"), self._real(phi, "phi"), self._real(V, "V"), out))
        return tuple(out[:5])

    def orth_normalize(self, p, phi, V):
        """p = s (p - f phi) on the device (ofdft_orth_normalize); the five
        sums come back with the next deferred_sums().  No synchronisation."""
        self._stream()
        L.check(self.lib.ofdft_orth_normalize(self._h, self._real(p, "p"), self._real(phi, "phi"),
                                              self._real(V, "V")))

    def hvp_prepare_async(self, phi, V, aux, mu, res, W0):
        """hvp_prepare whose sum res^2 comes back with the next deferred_sums()."""
        self._stream()
        auxp = self._realn(aux, "aux", 2 * self.N) if aux is not None else None
        L.check(self.lib.ofdft_hvp_prepare_async(self._h, self._real(phi, "phi"), self._real(V, "V"), auxp,
                                                 float(mu), self._real(res, "res"), self._real(W0, "W0")))

    def deferred_sums(self):
        """(sum res^2 of the last hvp_prepare_async, the five sums of the last
        orth_normalize); synchronises the stream."""
        out = L.darr(6)
        self._stream()
        L.check(self.lib.ofdft_plan_deferred_sums(self._h, out))
        return tuple(out[:6])

    def convert_from(self, src, dst=None):
        """dst (this plan's precision) = src (any float precision), same size."""
        if dst is None:
            dst = self.empty()
        sp = L.F64 if src.dtype == torch.float64 else L.F32
        if not src.is_contiguous() or src.device != self.device or src.numel() != dst.numel():
            raise ValueError("convert: source must be a contiguous field on the plan device")
        self._stream()
        L.check(self.lib.ofdft_convert(self._h, L.ptr(dst), sp, L.ptr(src), dst.numel()))
        return dst

    PROFILE_CLASSES = ("rows_fwd", "rows_fwd_hvp", "axis_c2c", "pencil_conv", "rows_inv", "rows_inv_hvp", "vector")

    def set_profiling(self, on=True, classes=None, stride=1):
        """Time kernel launches with HIP events: all classes, or only the
        named ones (PROFILE_CLASSES), one launch in every ``stride`` of
        each (1..256); resets the counters."""
        if not on:
            mask = 0
        elif classes is None:
            mask = (1 << len(self.PROFILE_CLASSES)) - 1
        else:
            mask = 0
            for c in classes:
                mask |= 1 << self.PROFILE_CLASSES.index(c)
        if not 1 <= int(stride) <= 256:
            raise ValueError("stride must be in 1..256")
        if mask:
            mask |= (int(stride) - 1) << 8
        L.check(self.lib.ofdft_plan_set_profiling(self._h, mask))

    def get_profile(self):
        """{class: (ms_total, launches, algorithmic_bytes_total)}"""
        n = len(self.PROFILE_CLASSES)
        ms = L.darr(n)
        cnt = (ctypes.c_longlong * n)()
        by = L.darr(n)
        L.check(self.lib.ofdft_plan_get_profile(self._h, ms, cnt, by))
        return {k: (ms[i], int(cnt[i]), by[i]) for i, k in enumerate(self.PROFILE_CLASSES)}

    def set_tiles(self, tile_axis, tile_pencil):
        """Columns per workgroup of the axis pass and of the pencils
        (ofdft_plan_set_tiles)."""
        L.check(self.lib.ofdft_plan_set_tiles(self._h, int(tile_axis), int(tile_pencil)))
        self.tiles = (int(tile_axis), int(tile_pencil))

    CG_SPECTRA = {"plan": 0, "bf16": 1, "padded": 2}

    def set_cg_spectra(self, mode="plan", tiles=None):
        """Layout and storage of the spectra inside the device TN loop's
        Hessian products: "plan" (the plan precision in split rows, the
        Nyquist column apart, where the lengths allow; the default), "bf16"
        (split rows of bf16 pairs on FP32 plans; every transform still
        computes in FP32) or "padded" (the padded rows of the other passes).
        tiles: (axis, pencil) columns per workgroup of the bf16 passes; None
        takes the autotuned widths for these lengths, 0 the widest that
        divides n2/2 (ofdft_plan_set_cg_spectra)."""
        if mode not in self.CG_SPECTRA:
            raise ValueError("cg spectra mode must be one of %s, not %r" % (sorted(self.CG_SPECTRA), mode))
        if tiles is None:
            from . import tuning
            tiles = tuning.bf16_tiles_for(self.nr) if mode == "bf16" else (0, 0)
        L.check(self.lib.ofdft_plan_set_cg_spectra(self._h, self.CG_SPECTRA[mode], int(tiles[0]), int(tiles[1])))

    def cg_spectra(self):
        """(layout, (axis, pencil) tile widths) the next device TN loop uses,
        after the fall-backs: layout "plan", "bf16" or "padded"
        (ofdft_plan_cg_spectra)."""
        a, p = ctypes.c_int(), ctypes.c_int()
        mode = self.lib.ofdft_plan_cg_spectra(self._h, ctypes.byref(a), ctypes.byref(p))
        L.check(min(mode, 0))
        name = {v: k for k, v in self.CG_SPECTRA.items()}[mode]
        return name, (a.value, p.value)

    @property
    def cg_spectra_tiles(self):
        """(axis, pencil) tile widths of the bf16 spectra the next device TN
        loop uses, (0, 0) when it keeps plan-precision spectra."""
        name, tiles = self.cg_spectra()
        return tiles if name == "bf16" else (0, 0)

    def set_graph(self, on=True):
        """Run the device TN loop's CG iterations as replays of one captured
        hipGraph (ofdft_plan_set_graph)."""
        L.check(self.lib.ofdft_plan_set_graph(self._h, 1 if on else 0))

    def workspace_bytes(self):
        return int(self.lib.ofdft_plan_workspace_bytes(self._h))
