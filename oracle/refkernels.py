"""ctypes calls into the reference's own serial kernels compiled from source (oracle/_ref).

TEST INFRASTRUCTURE ONLY.  Each wrapper passes exactly what the reference's C++ adapter passes
(cited per function), with Dirichlet boundaries (BMG_BCs_definite = 0) and symmetric relaxation
(BMG_RELAX_SYM = 1).  Same numpy layout as oracle.py.
"""
import ctypes

import numpy as np

from . import oracle as _o

_i = ctypes.c_int
_p = _o._ptr


def _byref_int(v):
    return ctypes.byref(ctypes.c_int(v))


def residual(so, qf, q):
    """include/cedar/2d/residual.h:33-61 (k=kf=0, by reference)."""
    jj, ii = q.shape
    ns = so.shape[0]
    r = np.zeros_like(q)
    ifd = 1 if ns == 3 else 0
    _o.ref().BMG2_SymStd_residual(_byref_int(0), _p(so), _p(qf), _p(q), _p(r), _byref_int(ii),
                                  _byref_int(jj), _byref_int(0), _byref_int(ifd), _byref_int(ns),
                                  _byref_int(0), _byref_int(0), _byref_int(0), _byref_int(0))
    return r


def setup_recip(so):
    """src/2d/relax.cc:23-33."""
    ns, jj, ii = so.shape
    sor = np.zeros((2, jj, ii))
    _o.ref().BMG2_SymStd_SETUP_recip(_p(so), _p(sor), _i(ii), _i(jj), _i(ns), _i(2))
    return sor


def relax_gs(so, qf, q, sor, updown, jpn=0):
    """src/2d/relax.cc:38-68 (k = kf = 1)."""
    ns, jj, ii = so.shape
    q = q.copy()
    ifd = 1 if ns == 3 else 0
    _o.ref().BMG2_SymStd_relax_GS(_i(1), _p(so), _p(qf), _p(q), _p(sor), _i(ii), _i(jj), _i(1),
                                  _i(ifd), _i(ns), _i(2), _i(1), _i(updown), _i(jpn))
    return q


def restrict(ci, q, jpn=0):
    """src/2d/restrict.cc:14-29 (q is refreshed in place when periodic: a copy is passed)."""
    jjf, iif = q.shape
    q = q.copy()
    _, jjc, iic = ci.shape
    qc = np.zeros((jjc, iic))
    _o.ref().BMG2_SymStd_restrict(_p(q), _p(qc), _p(ci), _i(iif), _i(jjf), _i(iic), _i(jjc), _i(jpn))
    return qc


def interp_add(q, qc, res, so, ci, jpn=0):
    """src/2d/interp.cc:20-46 (res is divided by SO_c in place)."""
    jjf, iif = q.shape
    jjc, iic = qc.shape
    q = q.copy()
    res = res.copy()
    _o.ref().BMG2_SymStd_interp_add(_p(q), _p(qc), _p(res), _p(so), _p(ci), _i(iic), _i(jjc),
                                    _i(iif), _i(jjf), _i(so.shape[0]), _i(jpn))
    return q, res


def setup_interp(so, iic=None, jjc=None, jpn=0):
    """src/2d/interp.cc:49-106 (ifd = 1 for a 5-point fine operator)."""
    ns, jjf, iif = so.shape
    if iic is None:
        iic, jjc = _o.coarse_dims(iif, jjf)
    ci = np.zeros((8, jjc, iic))
    soc = np.zeros((5, jjc, iic))
    ifd = 1 if ns == 3 else 0
    _o.ref().BMG2_SymStd_SETUP_interp_OI(_p(so), _p(soc), _p(ci), _i(iif), _i(jjf), _i(iic),
                                         _i(jjc), _i(ifd), _i(ns), _i(jpn), _i(0))
    return ci


def setup_galerkin(so, ci, jpn=0):
    """include/cedar/2d/coarsen.h:34-61."""
    ns, jjf, iif = so.shape
    _, jjc, iic = ci.shape
    soc = np.zeros((5, jjc, iic))
    ifd = 1 if ns == 3 else 0
    _o.ref().BMG2_SymStd_SETUP_ITLI_ex(_p(so), _p(soc), _p(ci), _i(iif), _i(jjf), _i(iic),
                                       _i(jjc), _i(ifd), _i(ns), _i(jpn))
    return soc
