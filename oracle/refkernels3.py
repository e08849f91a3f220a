"""ctypes calls into the reference's own serial 3D kernels compiled from source (oracle/_ref).

TEST INFRASTRUCTURE ONLY.  Each wrapper passes exactly what the reference's C++ adapter passes
(cited per function), with Dirichlet boundaries (BMG_BCs_definite = 0) unless a boundary id
jpn is given, and symmetric relaxation (BMG_RELAX_SYM = 1).  Same numpy layout as oracle3.py.
"""
import ctypes

import numpy as np

from . import oracle as _o
from .oracle3 import coarse_dims3

_i = ctypes.c_int
_p = _o._ptr


def residual(so, qf, q):
    """include/cedar/3d/residual.h:39-61 (kg = nog = 1, ifd = 1 for 7-point)."""
    kk, jj, ii = q.shape
    ns = so.shape[0]
    r = np.zeros_like(q)
    _o.ref().BMG3_SymStd_residual(_i(1), _i(1), _i(1 if ns == 4 else 0), _p(q), _p(qf), _p(so),
                                  _p(r), _i(ii), _i(jj), _i(kk), _i(ns))
    return r


def setup_recip(so):
    """include/cedar/3d/relax.h:51-64 (nsorv = 2)."""
    ns, kk, jj, ii = so.shape
    sor = np.zeros((2, kk, jj, ii))
    _o.ref().BMG3_SymStd_SETUP_recip(_p(so), _p(sor), _i(ii), _i(jj), _i(kk), _i(ns), _i(2))
    return sor


def relax_gs(so, qf, q, sor, updown, jpn=0):
    """include/cedar/3d/relax.h:67-97 (k = 1; jpn from BMG_get_bc, :94)."""
    ns, kk, jj, ii = so.shape
    q = q.copy()
    _o.ref().BMG3_SymStd_relax_GS(_i(1), _p(so), _p(qf), _p(q), _p(sor), _i(ii), _i(jj), _i(kk),
                                  _i(1 if ns == 4 else 0), _i(ns), _i(2), _i(1), _i(updown),
                                  _i(jpn))
    return q


def restrict(ci, q, jpn=0, with_q=False):
    """src/3d/restrict.cc:16-35 (jpn from BMG_get_bc, :27; the periodic branch refreshes the
    images of q in place; with_q returns that q too)."""
    kkf, jjf, iif = q.shape
    _, kkc, jjc, iic = ci.shape
    q = q.copy()
    qc = np.zeros((kkc, jjc, iic))
    _o.ref().BMG3_SymStd_restrict(_p(q), _p(qc), _p(ci), _i(iif), _i(jjf), _i(kkf), _i(iic),
                                  _i(jjc), _i(kkc), _i(jpn))
    return (qc, q) if with_q else qc


def interp_add(q, qc, res, so, ci):
    """src/3d/interp.cc:18-45 (res divided by SO_p in place)."""
    kkf, jjf, iif = q.shape
    kkc, jjc, iic = qc.shape
    q = q.copy()
    res = res.copy()
    _o.ref().BMG3_SymStd_interp_add(_p(q), _p(qc), _p(so), _p(res), _p(ci), _i(iic), _i(jjc),
                                    _i(kkc), _i(iif), _i(jjf), _i(kkf), _i(so.shape[0]), _i(0))
    return q, res


def setup_interp(so, cdims=None, jpn=0):
    """include/cedar/3d/interp.h:40-73 (ifd = 1 for a 7-point fine operator, scratch yo; jpn from
    BMG_get_bc, :66)."""
    ns, kkf, jjf, iif = so.shape
    iic, jjc, kkc = cdims or coarse_dims3(iif, jjf, kkf)
    ci = np.zeros((26, kkc, jjc, iic))
    soc = np.zeros((14, kkc, jjc, iic))
    yo = np.zeros((14, 2, jjf, iif))
    _o.ref().BMG3_SymStd_SETUP_interp_OI(_p(so), _p(soc), _p(ci), _i(iif), _i(jjf), _i(kkf),
                                         _i(iic), _i(jjc), _i(kkc), _i(1 if ns == 4 else 0),
                                         _i(ns), _i(1), _i(jpn), _p(yo))
    return ci


def setup_galerkin(so, ci, jpn=0):
    """include/cedar/3d/coarsen.h:41-61 (ITLI07 for 7-point, ITLI27 for 27-point; ipn from
    BMG_get_bc, :48)."""
    ns, kkf, jjf, iif = so.shape
    _, kkc, jjc, iic = ci.shape
    soc = np.zeros((14, kkc, jjc, iic))
    f = _o.ref().BMG3_SymStd_SETUP_ITLI07_ex if ns == 4 else _o.ref().BMG3_SymStd_SETUP_ITLI27_ex
    f(_p(so), _p(soc), _p(ci), _i(iif), _i(jjf), _i(kkf), _i(iic), _i(jjc), _i(kkc), _i(jpn))
    return soc
