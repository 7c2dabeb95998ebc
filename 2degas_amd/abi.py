"""ctypes mirror of the C-ABI structs in include/qmc_hip.h.

Only plumbing: the structs are filled from a :class:`Deck` and handed to the
HIP library (and, in tests, to the CPU oracle, which shares the layout).
"""
import ctypes as C

import numpy as np

c_double_p = C.POINTER(C.c_double)
c_int32_p = C.POINTER(C.c_int32)

QMC_OK = 0
QMC_ERR_ARG = 1
QMC_ERR_LIMIT = 2
QMC_ERR_SINGULAR = 3
QMC_ERR_HIP = 4
QMC_ERR_NOTWIST = 5
QMC_ERR_STATE = 6

QMC_PREC_FP64 = 0
QMC_PREC_MIXED = 1


class QmcTables(C.Structure):
    _fields_ = [
        ("ndim", C.c_int32),
        ("ntypes", C.c_int32),
        ("np", C.c_int32 * 2),
        ("hbs2m", C.c_double * 2),
        ("el", C.c_double * 2),
        ("drt", C.c_double),
        ("ntab", C.c_int32),
        ("ldut", C.c_int32),
        ("ut", c_double_p),
        ("ngrid", c_int32_p),
        ("tail", c_double_p),
        ("iexp", c_int32_p),
        ("nk_ewald", c_int32_p),
        ("iu2", C.c_int32 * 4),
        ("iu3", C.c_int32 * 4),
        ("iub", C.c_int32 * 4),
        ("iv2", C.c_int32 * 4),
        ("v2value", C.c_double * 4),
        ("v0", C.c_double),
        ("nk", C.c_int32),
        ("kvec", c_double_p),
        ("irhok", C.c_int32 * 2),
        ("nrhok", C.c_int32),
        ("ntheta", C.c_int32),
        ("adrift", C.c_double),
        ("delta", C.c_double),
    ]


class QmcInfo(C.Structure):
    _fields_ = [
        ("nptot", C.c_int32),
        ("npnorm", C.c_int32),
        ("n_scalar_props", C.c_int32),
        ("n_props", C.c_int32),
        ("jetot", C.c_int32),
        ("jltf", C.c_int32),
        ("jacc", C.c_int32),
        ("jpot0", C.c_int32),
        ("jkin0", C.c_int32),
        ("je2", C.c_int32),
        ("jgg", C.c_int32),
        ("jefp", C.c_int32),
        ("jrhok", C.c_int32),
        ("jpot", C.c_int32 * 2),
        ("jkin", C.c_int32 * 2),
        ("jun", C.c_int32 * 2),
        ("ng", C.c_int32),
        ("has_det", C.c_int32),
        ("reserved0", C.c_int32),
        ("dev_bytes", C.c_uint64),
        ("eval_bytes_per_walker", C.c_uint64),
        ("vmc_bytes_per_walker", C.c_uint64),
        ("stage_bytes_per_walker", C.c_uint64),
    ]

    def as_dict(self):
        d = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            d[name] = list(v) if hasattr(v, "__len__") else int(v)
        return d


def dptr(a):
    """Pointer to a C-contiguous float64 numpy array."""
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(c_double_p)


def iptr(a):
    """Pointer to a C-contiguous int32 numpy array."""
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(c_int32_p)


class QmcWalkers(C.Structure):
    _fields_ = [
        ("nw", C.c_int32),
        ("x", C.c_void_p),
        ("g", C.c_void_p),
        ("h", C.c_void_p),
        ("props", C.c_void_p),
        ("grad2_ph", C.c_void_p),
        ("age", C.c_void_p),
        ("rng", C.c_void_p),
        ("blk", C.c_void_p),
    ]


class QmcReptiles(C.Structure):
    _fields_ = [
        ("nr", C.c_int32),
        ("ntau", C.c_int32),
        ("nbuf", C.c_int32),
        ("x", C.c_void_p),
        ("g", C.c_void_p),
        ("props", C.c_void_p),
        ("age", C.c_void_p),
        ("w", C.c_void_p),
        ("ends", C.c_void_p),
        ("rage", C.c_void_p),
        ("pp_old", C.c_void_p),
        ("grad2_ph", C.c_void_p),
        ("rng", C.c_void_p),
        ("blk", C.c_void_p),
    ]


class QmcRmcParams(C.Structure):
    _fields_ = [
        ("delta", C.c_double),
        ("etrial", C.c_double),
        ("mage", C.c_double),
        ("mage_r", C.c_double),
        ("ntauskip", C.c_int32),
        ("mdelta", C.c_int32),
        ("rejection", C.c_int32),
        ("wpiu", C.c_int32),
        ("gstorto", C.c_int32),
        ("nodalaction", C.c_int32),
        ("ecut", C.c_int32),
        ("bm_order", C.c_int32),
        ("value_ecut", C.c_double),
    ]


class QmcDmcParams(C.Structure):
    _fields_ = [
        ("delta", C.c_double),
        ("etrial", C.c_double),
        ("mage", C.c_double),
        ("mmstep", C.c_int32),
        ("mmult", C.c_int32),
        ("bm_order", C.c_int32),
        ("reserved0", C.c_int32),
    ]


class QmcDmcState(C.Structure):
    _fields_ = [
        ("wate", C.c_void_p),
        ("wsum", C.c_void_p),
        ("src", C.c_void_p),
        ("rng2", C.c_void_p),
        ("counts", C.c_void_p),
    ]


class QmcSemState(C.Structure):
    _fields_ = [
        ("ainv", C.c_void_p),
        ("dlogpsi", C.c_void_p),
        ("nacc", C.c_void_p),
    ]


class QmcSemJob(C.Structure):
    _fields_ = [
        ("ctx", C.c_void_p),
        ("w", C.c_void_p),
        ("d", C.c_void_p),
        ("j0", C.c_int32),
        ("j1", C.c_int32),
        ("sigma", C.c_double),
    ]


def limbs_to_state(limbs):
    """48-bit LCG state from the four 12-bit limbs savern writes."""
    l1, l2, l3, l4 = (int(v) for v in limbs)
    return ((l1 * 4096 + l2) * 4096 + l3) * 4096 + l4


def state_to_limbs(x):
    x = int(x)
    return [(x >> 36) & 4095, (x >> 24) & 4095, (x >> 12) & 4095, x & 4095]
