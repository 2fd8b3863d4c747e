"""Shared test cases: registered (wavelet, bc) combinations and their line parity."""
REGISTERED = [
    ("cdf3.5", "n"), ("cdf3.5", "d"), ("cdf3.5", "z"),
    ("cdf3.7", "n"), ("cdf3.7", "d"), ("cdf3.7", "z"),
    ("cdf2.6", "n"), ("cdf2.6", "d"), ("cdf2.6", "z"),
    ("cdf2.8", "n"), ("cdf2.8", "d"), ("cdf2.8", "z"),
    ("cdf4.4", "n"), ("cdf4.4", "d"), ("cdf4.4", "z"),
    ("cdf4.6", "n"), ("cdf4.6", "d"), ("cdf4.6", "z"),
    ("cdf2.0", "n"), ("cdf2.0", "d"),
    ("cdf1.1", "n"), ("cdf1.1", "d"),
    ("cdf1.7", "n"), ("cdf1.7", "d"),
]
ODD_FAMILIES = {"cdf2.6", "cdf2.8", "cdf4.4", "cdf4.6", "cdf2.0"}
# (primal, dual) of CDF(p, q): detail coefficients of polynomials of degree < p vanish,
# the synthesis wavelet has q vanishing moments
CDF_PQ = {"cdf3.5": (3, 5), "cdf3.7": (3, 7), "cdf1.1": (1, 1), "cdf1.7": (1, 7),
          "cdf2.6": (2, 6), "cdf2.8": (2, 8), "cdf4.4": (4, 4), "cdf4.6": (4, 6), "cdf2.0": (2, 0)}


def line_len(wt, levels):
    return (1 << levels) + (1 if wt in ODD_FAMILIES else 0)


def bc_consistent(x, wt, bc):
    """Zero the boundary nodes where the antisymmetric whole-sample reflection
    requires them (they are the Dirichlet nodes of the staggered axes)."""
    if wt in ODD_FAMILIES and bc == "d":
        x = x.copy()
        x[..., 0] = 0.0
        x[..., -1] = 0.0
    return x
