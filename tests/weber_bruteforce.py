"""Weber's modular GCD run on exact integers (test infrastructure).

A second, independent statement of what the reference kernel computes
(getGcdKernel.cu:343-444), written on the integers U, V themselves instead
of their residues: every step divides exactly, so a slip in the oracle's
residue arithmetic, in the pivot key or its tie-break, or in the digit rule
shows up as a different sequence.  Python integers only; small inputs.

Pivot rule (getGcdKernel.cu:143-220): over the active moduli q, t_q =
toSigned(U * V^-1 mod q) (infinity when V == 0 mod q); the pivot minimises
(|t_q| << 32) | (q - [t_q >= 0]) with |infinity| = 2^31.  Step:
(U, V) <- (V, (U - b V) / p).  The loop stops when the next pivot is
infinity.  Digits (:404-433): the lowest-index active modulus with
U mod q != 0 gives g = toSigned(U mod q); U <- (U - g) / q.
"""

INF = -(1 << 31)


def to_signed(x, q):
    return x - q if x >= q // 2 else x


def key(t, q):
    a = 1 << 31 if t == INF else abs(t)
    return (a << 32) | (q - (1 if t >= 0 else 0))


def t_value(U, V, q):
    vq = V % q
    if vq == 0:
        return INF
    return to_signed(U * pow(vq, -1, q) % q, q)


def eliminate(U, V, moduli):
    """Returns (pivots, U_final, active set) following the reference loop,
    or raises ValueError when a step is not exact (should never happen)."""
    active = list(range(len(moduli)))
    best = min(active, key=lambda i: key(t_value(U, V, moduli[i]), moduli[i]))
    pivot = (moduli[best], t_value(U, V, moduli[best]))
    pivots = []
    while True:
        p, b = pivot
        pivots.append(pivot)
        active = [i for i in active if moduli[i] != p]
        num = U - b * V
        if num % p:
            raise ValueError("inexact elimination step")
        U, V = V, num // p
        cand = [(key(t_value(U, V, moduli[i]), moduli[i]), i) for i in active]
        k, i = min(cand)
        pivot = (moduli[i], t_value(U, V, moduli[i]))
        if pivot[1] == INF:
            break
    return pivots, U, active


def recover(U, moduli, active):
    """Signed mixed-radix digits of U over the active moduli."""
    digits = []
    active = list(active)
    while True:
        nz = [i for i in active if (U % moduli[i]) != 0]
        if not nz:
            break
        i = nz[0]
        q = moduli[i]
        g = to_signed(U % q, q)
        digits.append((q, g))
        active.remove(i)
        if (U - g) % q:
            raise ValueError("inexact recovery step")
        U = (U - g) // q
    return digits


def weber(u, v, moduli):
    pivots, U, active = eliminate(u, v, moduli)
    return pivots, recover(U, moduli, active)
