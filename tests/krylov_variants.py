"""The reference CGS recurrence (Sources/Process/Solvers_Mod/Cgs.f90:104-226)
with a pluggable dot-product order, for measuring how far rounding alone moves
CGS's iteration count.  Test infrastructure: the preconditioner comes from the
oracle, the products from scipy."""
import numpy as np
import scipy.sparse as sp


def _seq(a, b):
    return float(sum((a * b).tolist()))


def _rev(a, b):
    return float(sum((a * b)[::-1].tolist()))


def _blas(a, b):
    return float(np.dot(a, b))


ORDERS = {"sequential": _seq, "reversed": _rev, "blas": _blas}


def cgs_iterations(oracle, c, prec, tol, dot, cap=3000):
    A = sp.csr_matrix((c.val, c.col, c.row), shape=(c.n, c.n))
    d = oracle.prec_form(c.row, c.col, c.dia, c.val, prec)

    def M(v):
        return oracle.prec_solve(c.row, c.col, c.dia, c.val, d, prec, np.ascontiguousarray(v))
    x = np.zeros(c.n)
    r1 = c.b - A @ x
    r2 = r1.copy()
    rhoold = 1.0
    q1 = u2 = None
    for it in range(1, cap + 1):
        rho = dot(r1, r2)
        if it == 1:
            u1 = r1.copy()
            u2 = u1.copy()
        else:
            beta = rho / rhoold
            u1 = r1 + beta * q1
            u2 = u1 + beta * (q1 + beta * u2)
        p2 = M(u2)
        v2 = A @ p2
        alfa = rho / dot(r2, v2)
        q1 = u1 - alfa * v2
        p1 = M(u1 + q1)
        x = x + alfa * p1
        r1 = r1 - alfa * (A @ p1)
        if np.sqrt(dot(r1, r1)) < tol:
            return it
        rhoold = rho
    return cap + 1


def cgs_spread(oracle, c, prec, tol, cap=3000):
    """(min, max) iteration count over the summation orders."""
    its = [cgs_iterations(oracle, c, prec, tol, f, cap) for f in ORDERS.values()]
    return min(its), max(its)
