'''Superposition-of-atomic-densities initial guess.

Each unique element is solved once by a closed-shell SCF of the isolated
atom in the molecule's basis, with the electrons of the highest occupied
level spread evenly over its degenerate orbitals (a spherically averaged
occupation); the molecular guess is the block-diagonal sum of the atomic
densities.  This plays the role of init_guess_by_atom
(gpu4pyscf/scf/hf.py:561-607).
'''

import numpy as np

_cache = {}


def _frac_occ(mo_energy, nelec, tol=1e-4):
    import torch
    e = mo_energy
    order = torch.argsort(e)
    occ = torch.zeros_like(e)
    left = float(nelec)
    i = 0
    n = len(e)
    while left > 1e-12 and i < n:
        # degenerate group starting at order[i]
        e0 = float(e[order[i]])
        j = i
        while j < n and abs(float(e[order[j]]) - e0) < tol:
            j += 1
        g = j - i
        take = min(left, 2. * g)
        for k in range(i, j):
            occ[order[k]] = take / g
        left -= take
        i = j
    return occ


def atomic_density(symbol, basis, conv_tol=1e-8):
    '''Density (nao_atom x nao_atom, numpy) of the isolated atom.'''
    from gpu4pyscf_amd.gto.mole import M, CHARGES
    from gpu4pyscf_amd.scf.hf import RHF
    key = (symbol, repr(basis))
    if key in _cache:
        return _cache[key]
    z = CHARGES[symbol]
    atom = M(atom=f'{symbol} 0 0 0', basis={symbol: basis} if not isinstance(basis, dict)
             else basis, spin=z % 2)
    mf = RHF(atom)
    mf.conv_tol = conv_tol
    mf.init_guess = '1e'
    mf.max_cycle = 60
    mf.get_occ = lambda mo_energy: _frac_occ(mo_energy, z)
    mf.kernel()
    dm = mf.dm.cpu().numpy()
    _cache[key] = dm
    return dm


def atom_guess(mol):
    nao = mol.nao
    dm = np.zeros((nao, nao))
    slices = mol.aoslice_by_atom()
    for ia in range(mol.natm):
        p0, p1 = slices[ia, 2], slices[ia, 3]
        if p1 == p0:
            continue
        sym = mol.atom_symbol(ia)
        basis = mol.basis if not isinstance(mol.basis, dict) else mol.basis
        dm[p0:p1, p0:p1] = atomic_density(sym, basis)
    return dm
