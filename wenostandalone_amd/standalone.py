"""Python mirror of the reference module's call surface.

Same procedure names, argument meanings and sequences as module
m_weno_standalone (inputs/weno_standalone.f90:48-1732); arrays live on the
device (DeviceArray) and every call goes through the C-ABI of
libweno_mi355.  The Fortran shim (fortran/m_weno_standalone.f90) is the
primary drop-in; this mirror serves Python callers and the tests.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import fields
from ._lib import (F64, FUSE_COPY, SRC_FIELD, SRC_RESHAPED, Coeffs, check, lib, make_grid)
from .device import DeviceArray


class WenoStandalone:
    """State of m_weno_standalone (:56-107) plus its procedures (:111-1730)."""

    def __init__(self, m_cells=199, n_cells=199, p_cells=199, v_size=8, buff_size=6,
                 weno_eps=1e-16, wenojs=True, teno_CT=1e-5, teno=False,
                 weno5_split_kernels=False, weno5_specialized_combined=False,
                 mapped_weno=False, wenoz=False):
        self.m_cells, self.n_cells, self.p_cells = m_cells, n_cells, p_cells
        self.v_size, self.buff_size = v_size, buff_size
        self.weno_eps, self.wenojs, self.teno_CT, self.teno = weno_eps, wenojs, teno_CT, teno
        self.mapped_weno, self.wenoz = mapped_weno, wenoz
        self.weno5_split_kernels = weno5_split_kernels
        self.weno5_specialized_combined = weno5_specialized_combined
        check(lib.weno_init(-1))

    # -- bounds and data (:146-279) -------------------------------------------
    @property
    def grid(self):
        return make_grid(self.m_cells, self.n_cells, self.p_cells, self.v_size, self.buff_size,
                         self.weno_eps, self.wenojs, self.teno, self.teno_CT, self.mapped_weno, self.wenoz)

    def face_shape(self, d):
        ns, nk, nl = fields.dir_extents(self.m_cells, self.n_cells, self.p_cells, self.buff_size, d)
        return (ns, nk, nl, self.v_size)

    def s_initialize_data(self, field: np.ndarray | None = None, coeffs=None):
        """Allocates and fills v_vf, the reshape buffers and the coefficient
        sets (:146-205); `field` replaces s_fill_input_field's data."""
        m, n, p, v, b = self.m_cells, self.n_cells, self.p_cells, self.v_size, self.buff_size
        if field is None:
            field = fields.harness_field(m, n, p, v, b)
        self.v_vf = DeviceArray.from_host(field, np.float64)
        self.v_rs_ws = []
        self.coef = []
        for d in range(3):
            ns, nk, nl, _ = self.face_shape(d)
            self.v_rs_ws.append(DeviceArray((ns + 6, nk, nl, v)))
            c = coeffs[d] if coeffs is not None else fields.harness_coefficients(ns - 6, fields.DIR_OFFSETS[d])
            dev = [DeviceArray.from_host(a, np.float64) for a in c]
            self.coef.append(dev)
        self.v_rs_ws_x, self.v_rs_ws_y, self.v_rs_ws_z = self.v_rs_ws

    def s_finalize_data(self):
        for a in [self.v_vf, *self.v_rs_ws] + [x for c in self.coef for x in c]:
            a.free()

    def _coeffs(self, d) -> Coeffs:
        pl, pr, dl, dr, be = self.coef[d]
        return Coeffs(pl.ptr, pr.ptr, dl.ptr, dr.ptr, be.ptr)

    def _dir_of(self, v_rs_ws) -> int:
        for d, a in enumerate(self.v_rs_ws):
            if a is v_rs_ws:
                return d
        raise ValueError("v_rs_ws must be one of v_rs_ws_x/y/z")

    # -- reshape (:281-330) ---------------------------------------------------
    def _reshape(self, d):
        check(lib.weno_reshape(C.byref(self.grid), d, F64, self.v_vf.vp, self.v_rs_ws[d].vp))

    def s_initialize_weno_x(self):
        self._reshape(0)

    def s_initialize_weno_y(self):
        self._reshape(1)

    def s_initialize_weno_z(self):
        self._reshape(2)

    # -- order-1 copy (:332-390) ----------------------------------------------
    def _copy(self, d, vL, vR):
        check(lib.weno_order1_copy(C.byref(self.grid), d, F64, self.v_vf.vp, vL.vp, vR.vp))

    def s_order1_copy_x(self, vL_x, vR_x):
        self._copy(0, vL_x, vR_x)

    def s_order1_copy_y(self, vL_y, vR_y):
        self._copy(1, vL_y, vR_y)

    def s_order1_copy_z(self, vL_z, vR_z):
        self._copy(2, vL_z, vR_z)

    # -- the 8-argument family (:392, :477, :1399, :1419) ---------------------
    def _recon(self, order, v_rs_ws, vL, vR, flags=0):
        d = self._dir_of(v_rs_ws)
        cf = self._coeffs(d)
        check(lib.weno_reconstruct(C.byref(self.grid), order, d, SRC_RESHAPED, F64, v_rs_ws.vp,
                                   C.byref(cf), vL.vp, vR.vp, flags))

    def s_weno3_kernel(self, v_rs_ws, polyL, polyR, dL, dR, beta_coef, vL, vR):
        self._recon(3, v_rs_ws, vL, vR)

    def s_weno5_kernel(self, v_rs_ws, polyL, polyR, dL, dR, beta_coef, vL, vR):
        self._recon(5, v_rs_ws, vL, vR)

    def s_weno5_kernel_dispatch(self, v_rs_ws, polyL, polyR, dL, dR, beta_coef, vL, vR):
        # split and combined kernels produce the same outputs (:1410-1415)
        self._recon(5, v_rs_ws, vL, vR)

    def s_weno7_kernel(self, v_rs_ws, polyL, polyR, dL, dR, beta_coef, vL, vR):
        self._recon(7, v_rs_ws, vL, vR)

    # -- specialised WENO5 (:671-1143): direction-aware loads from v_vf -------
    def _weno5_field(self, d, vL, vR, flags=0):
        cf = self._coeffs(d)
        check(lib.weno_reconstruct(C.byref(self.grid), 5, d, SRC_FIELD, F64, self.v_vf.vp,
                                   C.byref(cf), vL.vp, vR.vp, flags))

    def s_weno5_kernel_x(self, vL, vR):
        self._weno5_field(0, vL, vR)

    def s_weno5_kernel_y(self, vL, vR):
        self._weno5_field(1, vL, vR)

    def s_weno5_kernel_z(self, vL, vR):
        self._weno5_field(2, vL, vR)

    # -- MP (:1544-1615) -------------------------------------------------------
    def s_preserve_monotonicity(self, v_rs_ws, vL, vR):
        d = self._dir_of(v_rs_ws)
        check(lib.weno_mp_limit(C.byref(self.grid), d, SRC_RESHAPED, F64, v_rs_ws.vp, vL.vp, vR.vp))

    # -- utilities and sequences (:1617-1730) ---------------------------------
    @staticmethod
    def f_checksum(arr) -> float:
        a = arr.download() if isinstance(arr, DeviceArray) else np.asarray(arr)
        return float(np.abs(a).sum(dtype=np.float64))

    def new_faces(self, d) -> DeviceArray:
        a = DeviceArray(self.face_shape(d))
        a.fill(0.0)
        return a

    def s_run_all_kernels(self, vL_x, vR_x, vL_y, vR_y, vL_z, vR_z):
        self.s_initialize_weno_x(); self.s_initialize_weno_y(); self.s_initialize_weno_z()
        self.s_order1_copy_x(vL_x, vR_x); self.s_order1_copy_y(vL_y, vR_y); self.s_order1_copy_z(vL_z, vR_z)
        L, R = (vL_x, vL_y, vL_z), (vR_x, vR_y, vR_z)
        for d in range(3):
            self.s_weno3_kernel(self.v_rs_ws[d], *self.coef[d], L[d], R[d])
        for d in range(3):
            self.s_weno5_kernel_dispatch(self.v_rs_ws[d], *self.coef[d], L[d], R[d])
        for d in range(3):
            self.s_weno7_kernel(self.v_rs_ws[d], *self.coef[d], L[d], R[d])
        for d in range(3):
            self.s_preserve_monotonicity(self.v_rs_ws[d], L[d], R[d])

    def s_prepare_weno5_inputs(self, vL_x, vR_x, vL_y, vR_y, vL_z, vR_z):
        self.s_initialize_weno_x(); self.s_initialize_weno_y(); self.s_initialize_weno_z()
        self.s_order1_copy_x(vL_x, vR_x); self.s_order1_copy_y(vL_y, vR_y); self.s_order1_copy_z(vL_z, vR_z)

    def s_run_weno5_reconstruction(self, vL_x, vR_x, vL_y, vR_y, vL_z, vR_z):
        if self.weno5_specialized_combined and not self.weno5_split_kernels:
            self.s_weno5_kernel_x(vL_x, vR_x)
            self.s_weno5_kernel_y(vL_y, vR_y)
            self.s_weno5_kernel_z(vL_z, vR_z)
        else:
            L, R = (vL_x, vL_y, vL_z), (vR_x, vR_y, vR_z)
            for d in range(3):
                self.s_weno5_kernel_dispatch(self.v_rs_ws[d], *self.coef[d], L[d], R[d])

    def s_run_weno5_kernels(self, vL_x, vR_x, vL_y, vR_y, vL_z, vR_z):
        self.s_prepare_weno5_inputs(vL_x, vR_x, vL_y, vR_y, vL_z, vR_z)
        self.s_run_weno5_reconstruction(vL_x, vR_x, vL_y, vR_y, vL_z, vR_z)

    def s_weno5_fused(self, d, vL, vR, flags=FUSE_COPY):
        """One pass per sweep: order-1 halo + WENO5 core from v_vf."""
        self._weno5_field(d, vL, vR, flags)
