"""Algorithmic work per unit for the roofline figures.

Flop counts follow the loop structure of the kernels, counting each real
add and multiply once (a complex multiply is 6 flops):
  * U/dU recursion, per computed entry of level j (compute_ui.jl:71-100,
    227-267): one or two terms r * conj(c) * u -> u: 8 flops per term plus
    2 to add them; du per component: 16 per term plus 2.
  * Adjoint dot per emitted entry: Re(conj u Y) and three Re(conj du_k Y):
    16 flops.  ulisttot accumulation: 4 flops per emitted entry.
  * Z (compute_z.jl): 10 flops per inner term, 4 per outer term; B: 4 per
    half-sum term; Y: 4 per contributing Z entry.
  * Plain dB (compute_b.jl:104-342): 4 flops per half-sum term per component,
    plus the switching-function scaling (24 per left-half entry).
Bytes for the pair kernel (HBM-bound): per step N(3w + 4) positions+species
read, 3Nw forces written, 4 bytes per neighbor index read, 56 bytes of
energy/virial (w = 8 for the FP64 I/O used by both precisions).
"""


def _chain(J):
    """(computed entries as (j, ma), emitted entries) of the row-chain order."""
    comp, emit = [], []
    for mb in range(J // 2 + 1):
        if mb == 0:
            emit.append((0, 0, 0))
        else:
            j = 2 * mb
            comp.extend((j, ma) for ma in range(j + 1))
            emit.extend((j, mb, ma) for ma in range(mb + 1))
        for j in range(2 * mb + 1, J + 1):
            comp.extend((j, ma) for ma in range(j + 1))
            emit.extend((j, mb, ma) for ma in range(j + 1))
    return comp, emit


def _terms(j, ma):
    return int(ma >= 1) + int(ma <= j - 1)


def u_flops_entry(j, ma):
    t = _terms(j, ma)
    return 8 * t + 2 * (t - 1)


def du_flops_entry(j, ma):
    t = _terms(j, ma)
    return 3 * (16 * t + 2 * (t - 1))


def flops_ui_per_pair(J):
    comp, emit = _chain(J)
    return sum(u_flops_entry(j, ma) for j, ma in comp) + 4 * len(emit)


def flops_dedr_per_pair(J):
    comp, emit = _chain(J)
    return sum(u_flops_entry(j, ma) + du_flops_entry(j, ma) for j, ma in comp) + 16 * len(emit)


def _half_terms(l):
    n = sum(l + 1 for mb in range(l // 2 + 1) if 2 * mb < l)
    if l % 2 == 0:
        n += l // 2 + 1
    return n


def zby_tables(J):
    """(idxz list of (j1, j2, j, na, nb, jju), idxb list, half positions)."""
    z, b = [], []
    ub = [sum((l + 1) ** 2 for l in range(j)) for j in range(J + 2)]
    for j1 in range(J + 1):
        for j2 in range(j1 + 1):
            for j in range(j1 - j2, min(J, j1 + j2) + 1, 2):
                if j >= j1:
                    b.append((j1, j2, j))
                for mb in range(j // 2 + 1):
                    for ma in range(j + 1):
                        ma1min = max(0, int((2 * ma - j - j2 + j1) / 2))
                        na = min(j1, int((2 * ma - j + j2 + j1) / 2)) - ma1min + 1
                        mb1min = max(0, int((2 * mb - j - j2 + j1) / 2))
                        nb = min(j1, int((2 * mb - j + j2 + j1) / 2)) - mb1min + 1
                        z.append((j1, j2, j, na, nb, ub[j] + (j + 1) * mb + ma))
    return z, b


def flops_zby_per_atom(J, with_y=True):
    z, b = zby_tables(J)
    f = sum(nb * (na * 10 + 4) for (_, _, _, na, nb, _) in z)
    f += sum(4 * _half_terms(j) for (_, _, j) in b)
    if with_y:
        f += 4 * len(z)
    return f


def flops_dbdr_plain_per_pair(J):
    z, b = zby_tables(J)
    nleft = sum((j // 2 + 1) * (j + 1) for j in range(J + 1))
    f = 0
    for j in range(1, J + 1):
        for mb in range(j // 2 + 1):
            for ma in range(j + 1):
                f += u_flops_entry(j, ma) + du_flops_entry(j, ma)
    f += 24 * nleft
    for (j1, j2, j) in b:
        f += 3 * 4 * (_half_terms(j) + _half_terms(j1) + _half_terms(j2))
    return f


def pair_bytes(n, half_pairs, w=8):
    """Algorithmic HBM bytes of one pair-kernel launch over a half-pair list."""
    return n * (3 * w + 4) + 3 * n * w + 4 * half_pairs + 56
