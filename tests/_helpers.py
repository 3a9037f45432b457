"""Shared test helpers: build the same normalised problem in the oracle and on the device."""
import numpy as np

from oracle import pyoracle as po
from qiotoolkit_amd import _qmc


def pair(ctx, kind, terms, lanes=0, visit=0):
    """(oracle problem, device problem) with identical internal ids."""
    op = po.Problem(kind, terms)
    off, ids, c = op.edges()
    code = _qmc.QMC_ISING if kind == "ising" else _qmc.QMC_PUBO
    qp = _qmc.Problem(ctx, code, op.n, off, ids, c, lanes=lanes, visit=visit)
    return op, qp


def arith_of(qp):
    return po.ARITH_F32 if qp.info.coupling == _qmc.QMC_COUPLING_F32 else po.ARITH_F64


def oracle_initial(op, opts, R):
    return np.stack([op.initial_state(opts, r) for r in range(R)])
