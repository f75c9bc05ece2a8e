"""Independent pure-Python model of the mathematics the oracle restates.

Used only to pin the oracle: field arithmetic is written as polynomial
arithmetic over M31 (schoolbook, not the Karatsuba/addition-chain forms the
reference uses), and polynomial evaluations use the explicit product form of
the circle / line FFT bases instead of the recursive fold:

  circle basis, n = 2^k coefficients, index bits (b_0 .. b_{k-1}):
      B(x, y) = y^b_0 * x^b_1 * pi(x)^b_2 * ... * pi^{k-2}(x)^b_{k-1},
      pi(x) = 2 x^2 - 1
  line basis (coefficients stored bit-reversed), n = 2^L:
      bit L-1 <-> x, bit L-2 <-> pi(x), ..., bit 0 <-> pi^{L-1}(x)
"""
P = (1 << 31) - 1
G = (2, 1268011823)  # generator of the circle group of order 2^31


# ---------------------------------------------------------------- fields
def cm_mul(p, q):
    # (a + b i)(c + d i), i^2 = -1
    a, b = p
    c, d = q
    return ((a * c - b * d) % P, (a * d + b * c) % P)


def qm_mul(p, q):
    # QM31 = CM31[u] / (u^2 - (2 + i)); elements (a, b, c, d) = (a + bi) + (c + di) u
    x1, y1 = (p[0], p[1]), (p[2], p[3])
    x2, y2 = (q[0], q[1]), (q[2], q[3])
    xx = cm_mul(x1, x2)
    yy = cm_mul(cm_mul(y1, y2), (2, 1))
    xy1, xy2 = cm_mul(x1, y2), cm_mul(y1, x2)
    return ((xx[0] + yy[0]) % P, (xx[1] + yy[1]) % P, (xy1[0] + xy2[0]) % P, (xy1[1] + xy2[1]) % P)


def qm_add(p, q): return tuple((a + b) % P for a, b in zip(p, q))
def qm_sub(p, q): return tuple((a - b) % P for a, b in zip(p, q))
def qm_scale(p, s): return tuple((a * s) % P for a in p)
def qm(v): return (v % P, 0, 0, 0)


def qm_pow(p, e):
    r = (1, 0, 0, 0)
    while e:
        if e & 1:
            r = qm_mul(r, p)
        p = qm_mul(p, p)
        e >>= 1
    return r


def qm_inv(p):
    # Fermat in the degree-4 extension: |QM31^*| = P^4 - 1
    return qm_pow(p, P ** 4 - 2)


def m31_inv(a): return pow(a, P - 2, P)


# ---------------------------------------------------------------- circle group
def pt_mul(p, q):
    return ((p[0] * q[0] - p[1] * q[1]) % P, (p[0] * q[1] + p[1] * q[0]) % P)


def pt_pow(p, e):
    r = (1, 0)
    while e:
        if e & 1:
            r = pt_mul(r, p)
        p = pt_mul(p, p)
        e >>= 1
    return r


def point_at_index(i): return pt_pow(G, i % (1 << 31))


def coset_points(initial_index, step_size, log_size):
    return [point_at_index(initial_index + step_size * i) for i in range(1 << log_size)]


def circle_domain_points(half_initial, half_step, half_log):
    """CircleDomain(half_coset): half_coset points, then their conjugates."""
    half = coset_points(half_initial, half_step, half_log)
    return half + [(x, (-y) % P) for x, y in half]


def canonic_half_coset(log_n):
    return 1 << (30 - log_n), (1 << (32 - log_n)) % (1 << 31), log_n - 1


def brev(i, bits):
    return int(format(i, f"0{bits}b")[::-1], 2) if bits else 0


# ---------------------------------------------------------------- polynomial bases
def _factors_circle(x, y, k, mul, sub2):
    """Factors [y, x, pi(x), ..., pi^{k-2}(x)] for index bits 0..k-1."""
    f = [y]
    if k >= 2:
        f.append(x)
        cur = x
        for _ in range(k - 2):
            cur = sub2(cur)
            f.append(cur)
    return f[:k]


def circle_poly_eval_m31(coeffs, point):
    """Evaluate a circle polynomial with M31 coefficients at an M31 point."""
    n = len(coeffs)
    k = n.bit_length() - 1
    x, y = point
    f = _factors_circle(x, y, k, None, lambda t: (2 * t * t - 1) % P)
    total = 0
    for idx, c in enumerate(coeffs):
        term = c
        for b in range(k):
            if (idx >> b) & 1:
                term = term * f[b] % P
        total += term
    return total % P


def circle_poly_eval_qm31(coeffs, px, py):
    n = len(coeffs)
    k = n.bit_length() - 1
    f = _factors_circle(px, py, k, None,
                        lambda t: qm_sub(qm_scale(qm_mul(t, t), 2), qm(1)))
    total = qm(0)
    for idx, c in enumerate(coeffs):
        term = c if isinstance(c, tuple) else qm(c)
        for b in range(k):
            if (idx >> b) & 1:
                term = qm_mul(term, f[b])
        total = qm_add(total, term)
    return total


def line_poly_eval(coeffs, x):
    """LinePoly::eval_at_point with coefficients in bit-reversed storage order:
    bit (L-1-j) of the index selects the factor pi^j(x)."""
    n = len(coeffs)
    L = n.bit_length() - 1
    fac = []
    cur = x
    for _ in range(L):
        fac.append(cur)
        cur = qm_sub(qm_scale(qm_mul(cur, cur), 2), qm(1))
    total = qm(0)
    for idx, c in enumerate(coeffs):
        term = c if isinstance(c, tuple) else qm(c)
        for j in range(L):
            if (idx >> (L - 1 - j)) & 1:
                term = qm_mul(term, fac[j])
        total = qm_add(total, term)
    return total


def random_qm31_point(rng):
    """Stereographic parametrisation: t -> ((1 - t^2)/(1 + t^2), 2t/(1 + t^2))."""
    while True:
        t = tuple(rng.randrange(P) for _ in range(4))
        t2 = qm_mul(t, t)
        den = qm_add(qm(1), t2)
        if den == qm(0):
            continue
        inv = qm_inv(den)
        return qm_mul(qm_sub(qm(1), t2), inv), qm_mul(qm_scale(t, 2), inv)


# ---------------------------------------------------------------- DEEP quotient pieces
def qm_conj(p):
    """Conjugation over CM31: x + y u -> x - y u."""
    return (p[0], p[1], (-p[2]) % P, (-p[3]) % P)


def qm_div(p, q): return qm_mul(p, qm_inv(q))


U = (0, 0, 1, 0)


def pair_vanishing(e0, e1, p):
    """Line through the circle points e0 and e1, as the determinant
        | p.x  p.y  1 |
        | e0.x e0.y 1 |
        | e1.x e1.y 1 |
    (zero exactly on the line through e0 and e1); every argument is a
    QM31 point ((x4), (y4))."""
    px, py = p
    t0 = qm_mul(px, qm_sub(e0[1], e1[1]))
    t1 = qm_mul(py, qm_sub(e0[0], e1[0]))
    t2 = qm_sub(qm_mul(e0[0], e1[1]), qm_mul(e0[1], e1[0]))
    return qm_add(qm_sub(t0, t1), t2)


def conj_line(point, value, y):
    """The line through (P.y, v) and (conj P.y, conj v), evaluated at y."""
    c = qm_sub(qm_conj(point[1]), point[1])
    slope = qm_div(qm_sub(qm_conj(value), value), c)
    return qm_add(value, qm_mul(slope, qm_sub(y, point[1])))


# ---------------------------------------------------------------- multilinear extensions
def eq_poly(x_bits, r):
    """eq(x, r) = prod_k (x_k r_k + (1 - x_k)(1 - r_k)) for boolean x."""
    e = qm(1)
    for xb, rk in zip(x_bits, r):
        e = qm_mul(e, rk if xb else qm_sub(qm(1), rk))
    return e


def index_bits(x, n):
    """Variable k of index x is bit (n-1-k): variable 0 is the most significant bit."""
    return [(x >> (n - 1 - k)) & 1 for k in range(n)]


def mle_eval(evals, r):
    """Multilinear extension of evals (length 2^n) at r: sum_x f(x) eq(x, r)."""
    n = len(r)
    total = qm(0)
    for x, f in enumerate(evals):
        fx = f if isinstance(f, tuple) else qm(int(f))
        total = qm_add(total, qm_mul(fx, eq_poly(index_bits(x, n), r)))
    return total
