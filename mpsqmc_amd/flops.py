"""Algorithmic FP32 work of one walker-step of the environment-sweep algorithm
(the numerator of the roofline figures).

Per walker-step: the right-environment sweep and, per coupling (i, j), the
two-site mixed tensor (open at site i, carry through i < k < j, close at site
j) plus the left-environment advance over [i, j); all of it twice, once for
the propagation and once for the bond-local energy pass.  d = 2, t_k / w_k the
trial / walker bonds, one multiply-add = 2 flops:

    tau(k)  = d (t_k w_k w_{k+1} + t_k t_{k+1} w_{k+1})
    open    = d t_i w_i w_{i+1} + d^2 t_i t_{i+1} w_{i+1}
    carry   = d^2 tau(k),  i < k < j
    close   = d t_j t_{j+1} w_{j+1} + d^2 t_j w_{j+1} w_j + d^4 t_j w_j

The reference's own algorithm (16 + 1 full-chain overlaps per coupling,
MPSQMC2.cpp:540-549) does 9-45x more; the roofline charges only this count.
"""


def bond_profile(L, D, d=2):
    """w_k = min(d^k, D, d^(L-k)), k = 0..L."""
    return [min(d ** k, D, d ** (L - k)) for k in range(L + 1)]


def env_flops_fp32(first, second, t, w, d=2):
    """FP32 flops per walker-step for couplings (first[c], second[c]) with
    trial bonds t and walker bonds w (lists of length L+1)."""
    L = len(t) - 1

    def tau(k):
        return d * (t[k] * w[k] * w[k + 1] + t[k] * t[k + 1] * w[k + 1])

    right = sum(tau(k) for k in range(L))
    pairs = 0
    for i, j in zip(first, second):
        i, j = int(i), int(j)
        opening = d * t[i] * w[i] * w[i + 1] + d * d * t[i] * t[i + 1] * w[i + 1]
        closing = d * t[j] * t[j + 1] * w[j + 1] + d * d * t[j] * w[j + 1] * w[j] + d ** 4 * t[j] * w[j]
        carry = sum(d * d * tau(k) for k in range(i + 1, j))
        advance = sum(tau(k) for k in range(i, j))
        pairs += opening + closing + carry + advance
    return 2.0 * (2 * right + 2 * pairs)
