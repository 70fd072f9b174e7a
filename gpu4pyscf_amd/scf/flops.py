'''Algorithmic cost model of the direct J/K build.

Per surviving contracted shell quartet of class (li, lj, lk, ll) with
primitive counts (pi, pj, pk, pl), n = (li+lj+lk+ll)//2 + 1 Rys roots,
Nf = prod (l+1)(l+2)/2 Cartesian components and G2 = (li+lj+1)(lk+ll+1):

    F = pi pj pk pl * n * (3 Nf + 9 G2) + 12 n_dm Nf      [FP64 flops]

3 Nf is the triple product and its accumulation per root, 9 G2 about three
flops per 2D-integral element in three directions, 12 n_dm Nf the six
digestion contractions (multiply + add) per density matrix.  Root
evaluation and pair prefactors are not counted.

Digestion bytes per quartet: 16 n_dm (nfi nfj + nfk nfl + nfi nfk + nfi nfl
+ nfj nfk + nfj nfl): an 8-byte density read and an 8-byte atomic update per
element of the six blocks.
'''


def ncart(l):
    return (l + 1) * (l + 2) // 2


def quartet_flops(ls, prims, n_dm=1):
    li, lj, lk, ll = ls
    pi, pj, pk, pl = prims
    n = (li + lj + lk + ll) // 2 + 1
    nf = ncart(li) * ncart(lj) * ncart(lk) * ncart(ll)
    g2 = (li + lj + 1) * (lk + ll + 1)
    return pi * pj * pk * pl * n * (3 * nf + 9 * g2) + 12 * n_dm * nf


def quartet_digest_bytes(ls, n_dm=1):
    fi, fj, fk, fl = (ncart(l) for l in ls)
    return 16 * n_dm * (fi*fj + fk*fl + fi*fk + fi*fl + fj*fk + fj*fl)
