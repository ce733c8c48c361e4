"""50-digit solves of the reference's second-order Feautrier system of one ray
(the rows of solve_feautrier_order_2, solver.tpp:2194-2318), built from the
optical depths, sources and boundary intensities the oracle reports
(oracle.feautrier_system).  Test helper."""
import mpmath as mp


def system(sys_):
    """(a, b, c, r): sub-diagonal, diagonal, super-diagonal and right-hand side."""
    mp.mp.dps = 50
    n = sys_["n"]
    dt = [mp.mpf(float(v)) for v in sys_["dtau"]]
    S = [mp.mpf(float(v)) for v in sys_["S"]]
    If, Il = (mp.mpf(float(v)) for v in sys_["I_bdy"])
    a = [mp.mpf(0)] * n
    b = [mp.mpf(0)] * n
    c = [mp.mpf(0)] * n
    r = [mp.mpf(0)] * n
    Cf = 2 / dt[0] ** 2
    b[0] = 1 + 2 / dt[0] + Cf
    c[0] = -Cf
    r[0] = S[0] + 2 * If / dt[0]
    for j in range(1, n - 1):
        avg = (dt[j - 1] + dt[j]) / 2
        A = 1 / (avg * dt[j - 1])
        C = 1 / (avg * dt[j])
        a[j], b[j], c[j], r[j] = -A, 1 + A + C, -C, S[j]
    Al = 2 / dt[n - 2] ** 2
    a[n - 1] = -Al
    b[n - 1] = 1 + 2 / dt[n - 2] + Al
    r[n - 1] = S[n - 1] + 2 * Il / dt[n - 2]
    return a, b, c, r


def solve(abc, rhs):
    """Thomas algorithm in 50 digits."""
    a, b, c = abc
    n = len(b)
    cp = [mp.mpf(0)] * n
    dp = [mp.mpf(0)] * n
    cp[0] = c[0] / b[0]
    dp[0] = rhs[0] / b[0]
    for j in range(1, n):
        den = b[j] - a[j] * cp[j - 1]
        cp[j] = c[j] / den if j < n - 1 else 0
        dp[j] = (rhs[j] - a[j] * dp[j - 1]) / den
    x = [mp.mpf(0)] * n
    x[n - 1] = dp[n - 1]
    for j in range(n - 2, -1, -1):
        x[j] = dp[j] - cp[j] * x[j + 1]
    return x


def inverse_row(sys_, row):
    """Row `row` of T^-1: d u_row / d S_n for every n."""
    a, b, c, _ = system(sys_)
    n = len(b)
    out = []
    for j in range(n):
        e = [mp.mpf(0)] * n
        e[j] = mp.mpf(1)
        out.append(solve((a, b, c), e)[row])
    return out
