'''A PySCF-compatible molecule object.

The J/K engine consumes molecules through the integer/double tables PySCF and
libcint use (``_atm``, ``_bas``, ``_env``), so any object that carries these
tables -- a real ``pyscf.gto.Mole`` included -- can be handed to
:func:`gpu4pyscf_amd.scf.jk.get_jk`.  PySCF is not a dependency, so this module
builds the same tables from an atom string and a basis specification.

Conventions restated from PySCF (gpu4pyscf/gto/mole.py imports them from
``pyscf.gto``; the libcint slot layout is also spelled out in
gpu4pyscf/lib/gvhf-rys/vhf.cuh:5-24):

* ``_atm`` has 6 int32 slots per atom, ``_bas`` 8 int32 slots per shell,
  ``_env`` starts user data at offset 20 and holds the range-separation
  parameter omega at offset 8.
* Coordinates are stored in Bohr, with 1 Bohr = 0.52917721092 Angstrom.
* Primitive rows of a shell are sorted by exponent, largest first; the
  contraction coefficients are multiplied by the radial normalisation
  ``gto_norm`` and each contracted function is normalised radially.
* Shells of one atom are ordered by angular momentum (stable).
* Spherical AOs are ordered m = -l..l, except p which is (x, y, z);
  Cartesian AOs are ordered lexically (xx, xy, xz, yy, yz, zz, ...).
'''

import math
import re
import numpy as np
from scipy.special import gamma

from gpu4pyscf_amd.gto import basis as basis_mod

BOHR = 0.52917721092

# _atm slots
CHARGE_OF = 0
PTR_COORD = 1
NUC_MOD_OF = 2
PTR_ZETA = 3
PTR_FRAC_CHARGE = 4
ATM_SLOTS = 6
# _bas slots
ATOM_OF = 0
ANG_OF = 1
NPRIM_OF = 2
NCTR_OF = 3
KAPPA_OF = 4
PTR_EXP = 5
PTR_COEFF = 6
PTR_BAS_COORD = 7
BAS_SLOTS = 8
PTR_RANGE_OMEGA = 8
PTR_ENV_START = 20

ELEMENTS = ['X',
    'H', 'He', 'Li', 'Be', 'B', 'C', 'N', 'O', 'F', 'Ne',
    'Na', 'Mg', 'Al', 'Si', 'P', 'S', 'Cl', 'Ar', 'K', 'Ca',
    'Sc', 'Ti', 'V', 'Cr', 'Mn', 'Fe', 'Co', 'Ni', 'Cu', 'Zn',
    'Ga', 'Ge', 'As', 'Se', 'Br', 'Kr']
CHARGES = {s: z for z, s in enumerate(ELEMENTS)}
ANGULAR = 'spdfghik'


def gaussian_int(n, alpha):
    r'''int_0^inf x^n exp(-alpha x^2) dx'''
    n1 = (n + 1) * .5
    return gamma(n1) / (2. * alpha**n1)


def gto_norm(l, expnt):
    '''Radial normalisation of r^l exp(-a r^2): 1/sqrt(int r^2 (r^l e^{-a r^2})^2 dr)'''
    return 1. / np.sqrt(gaussian_int(l*2+2, 2*np.asarray(expnt, dtype=float)))


def normalize_contracted_ao(l, es, cs):
    '''Scale each contracted radial function sum_p c_p g_p(r) to unit norm.'''
    ee = es.reshape(-1, 1) + es.reshape(1, -1)
    ee = gaussian_int(l*2+2, ee)
    s1 = 1. / np.sqrt(np.einsum('pi,pq,qi->i', cs, ee, cs))
    return np.einsum('pi,i->pi', cs, s1)


def nf_cart(l):
    return (l + 1) * (l + 2) // 2


def nf_sph(l):
    return 2 * l + 1


def fp(a):
    '''Fingerprint of an array: dot(cos(arange(n)), a.ravel()), the checksum
    PySCF's tests (``lib.fp``) quote for J/K matrices.'''
    a = np.asarray(a)
    return np.dot(np.cos(np.arange(a.size)), a.ravel())


def _std_symbol(symb):
    symb = re.sub(r'[^A-Za-z]', '', symb)
    if symb.upper().startswith('GHOST'):
        raise NotImplementedError('ghost atoms')
    return symb[0].upper() + symb[1:].lower() if len(symb) > 1 else symb.upper()


def _parse_atoms(atom):
    if isinstance(atom, str):
        lines = atom.replace(';', '\n').splitlines()
        out = []
        for line in lines:
            line = line.replace(',', ' ').strip()
            if not line or line.startswith('#'):
                continue
            fields = line.split()
            out.append((fields[0], [float(x) for x in fields[1:4]]))
        return out
    out = []
    for a in atom:
        if isinstance(a, str):
            out.extend(_parse_atoms(a))
        else:
            out.append((a[0], [float(x) for x in a[1]]))
    return out


class Mole:
    '''Molecule with a Gaussian basis, laid out as PySCF's Mole.

    Parameters mirror ``pyscf.M``: ``atom`` (string "Sym x y z; ..." or list
    of (symbol, xyz)), ``basis`` (name, NWChem text, internal shell list, a
    dict per element with optional 'default', or a tuple of such specs to be
    concatenated), ``unit`` ('Angstrom' or 'Bohr'), ``cart``, ``charge``,
    ``spin``.
    '''

    def __init__(self, atom=None, basis='sto-3g', unit='Angstrom', cart=False,
                 charge=0, spin=0, verbose=0, output=None, **kwargs):
        self.atom = atom
        self.basis = basis
        self.unit = unit
        self.cart = cart
        self.charge = charge
        self.spin = spin
        self.verbose = verbose
        self.output = output
        self._omega = 0.
        self._atm = self._bas = self._env = None
        self._atom = None
        self._basis = None
        if atom is not None:
            self.build()

    # -- construction --------------------------------------------------
    def build(self):
        atoms = _parse_atoms(self.atom)
        unit = self.unit
        if isinstance(unit, str):
            if unit.upper().startswith(('B', 'AU')):
                fac = 1.
            else:
                fac = 1. / BOHR
        else:
            fac = 1. / float(unit)
        self._atom = [(_std_symbol(s), np.asarray(xyz) * fac) for s, xyz in atoms]
        symbols = sorted(set(s for s, _ in self._atom))
        self._basis = {s: basis_mod.shells_for(self.basis, s) for s in symbols}

        env = [0.] * PTR_ENV_START
        atm = []
        for ia, (symb, xyz) in enumerate(self._atom):
            ptr = len(env)
            env.extend(list(xyz))
            env.append(0.)  # zeta for point nuclei
            atm.append([CHARGES[symb], ptr, 1, ptr+3, 0, 0])
        bas = []
        # one env copy of each element's basis, shared by its atoms
        basis_ptr = {}
        for symb in symbols:
            rows = []
            for shell in self._basis[symb]:
                l = shell[0]
                prim = np.array(sorted([list(p) for p in shell[1:]], reverse=True),
                                dtype=float)
                es = prim[:, 0]
                cs = prim[:, 1:]
                nprim, nctr = cs.shape
                cs = np.einsum('pi,p->pi', cs, gto_norm(l, es))
                cs = normalize_contracted_ao(l, es, cs)
                ptr_exp = len(env)
                env.extend(es.tolist())
                ptr_coeff = len(env)
                env.extend(cs.T.ravel().tolist())
                rows.append([l, nprim, nctr, ptr_exp, ptr_coeff])
            basis_ptr[symb] = rows
        for ia, (symb, _) in enumerate(self._atom):
            for l, nprim, nctr, pexp, pcoeff in basis_ptr[symb]:
                bas.append([ia, l, nprim, nctr, 0, pexp, pcoeff, 0])
        self._atm = np.asarray(atm, dtype=np.int32).reshape(-1, ATM_SLOTS)
        self._bas = np.asarray(bas, dtype=np.int32).reshape(-1, BAS_SLOTS)
        self._env = np.asarray(env, dtype=np.float64)
        self._env[PTR_RANGE_OMEGA] = self._omega
        nelec = int(self._atm[:, CHARGE_OF].sum()) - self.charge
        if (nelec - self.spin) % 2 != 0:
            raise RuntimeError(f'electron number {nelec} and spin {self.spin} are inconsistent')
        return self

    def copy(self):
        import copy
        return copy.copy(self)

    # -- PySCF-like accessors -------------------------------------------
    @property
    def omega(self):
        return self._omega

    @omega.setter
    def omega(self, v):
        self._omega = float(v)
        if self._env is not None:
            self._env[PTR_RANGE_OMEGA] = self._omega

    @property
    def natm(self):
        return len(self._atm)

    @property
    def nbas(self):
        return len(self._bas)

    @property
    def nelectron(self):
        return int(self._atm[:, CHARGE_OF].sum()) - self.charge

    @property
    def nelec(self):
        ne = self.nelectron
        na = (ne + self.spin) // 2
        return na, ne - na

    @property
    def nao(self):
        return self.nao_nr()

    def nao_nr(self, cart=None):
        return int(self.ao_loc_nr(cart)[-1])

    def nao_cart(self):
        return self.nao_nr(cart=True)

    def ao_loc_nr(self, cart=None):
        if cart is None:
            cart = self.cart
        l = self._bas[:, ANG_OF]
        if cart:
            dims = (l + 1) * (l + 2) // 2 * self._bas[:, NCTR_OF]
        else:
            dims = (l * 2 + 1) * self._bas[:, NCTR_OF]
        ao_loc = np.zeros(len(dims) + 1, dtype=np.int32)
        ao_loc[1:] = np.cumsum(dims)
        return ao_loc

    @property
    def ao_loc(self):
        return self.ao_loc_nr()

    def atom_coords(self, unit='Bohr'):
        ptr = self._atm[:, PTR_COORD]
        c = self._env[np.vstack([ptr, ptr+1, ptr+2]).T].copy()
        if unit.upper().startswith('A'):
            c *= BOHR
        return c

    def atom_charges(self):
        return self._atm[:, CHARGE_OF].copy()

    def atom_symbol(self, ia):
        return self._atom[ia][0]

    def elements(self):
        return [s for s, _ in self._atom]

    def bas_angular(self, i):
        return int(self._bas[i, ANG_OF])

    def bas_nprim(self, i):
        return int(self._bas[i, NPRIM_OF])

    def bas_nctr(self, i):
        return int(self._bas[i, NCTR_OF])

    def bas_atom(self, i):
        return int(self._bas[i, ATOM_OF])

    def bas_exp(self, i):
        p = self._bas[i, PTR_EXP]
        return self._env[p:p+self._bas[i, NPRIM_OF]].copy()

    def bas_ctr_coeff(self, i):
        '''Contraction coefficients (nprim, nctr) with the radial normalisation
        folded in, as stored in ``_env``.'''
        p = self._bas[i, PTR_COEFF]
        nprim = self._bas[i, NPRIM_OF]
        nctr = self._bas[i, NCTR_OF]
        return self._env[p:p+nprim*nctr].reshape(nctr, nprim).T.copy()

    def bas_coord(self, i):
        ia = self._bas[i, ATOM_OF]
        p = self._atm[ia, PTR_COORD]
        return self._env[p:p+3].copy()

    def aoslice_by_atom(self, ao_loc=None):
        if ao_loc is None:
            ao_loc = self.ao_loc_nr()
        out = np.zeros((self.natm, 4), dtype=int)
        atom_of = self._bas[:, ATOM_OF]
        for ia in range(self.natm):
            idx = np.where(atom_of == ia)[0]
            if len(idx) == 0:
                b0 = b1 = out[ia-1, 1] if ia > 0 else 0
            else:
                b0, b1 = idx[0], idx[-1] + 1
            out[ia] = [b0, b1, ao_loc[b0], ao_loc[b1]]
        return out

    def energy_nuc(self):
        z = self.atom_charges().astype(float)
        r = self.atom_coords()
        e = 0.
        for i in range(len(z)):
            for j in range(i):
                e += z[i] * z[j] / np.linalg.norm(r[i] - r[j])
        return e


def M(**kwargs):
    '''Build a molecule, as ``pyscf.M``.'''
    return Mole(**kwargs)
