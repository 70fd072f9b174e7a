'''Basis specification parsing (NWChem text, names, PySCF shell lists).

Mirrors how ``pyscf.M(basis=...)`` interprets its ``basis`` argument, which
gpu4pyscf takes from PySCF unchanged (gpu4pyscf/gto/mole.py:42-47):

* a name ('def2-tzvp', 'sto-3g', ...), matched case- and dash-insensitively;
* NWChem text with "Sym  L" headers followed by "exponent coeff..." rows; a
  text prefixed by 'unc' is uncontracted into one shell per primitive;
* a shell list ``[[l, [e, c1, c2, ...], ...], ...]``;
* a dict ``{symbol or 'default': spec}``;
* a tuple/list whose items are specs, concatenated in order.

Shells of an element are returned sorted by angular momentum (stable).
'''

import re
from gpu4pyscf_amd.gto import basis_data

_L = {'S': 0, 'P': 1, 'D': 2, 'F': 3, 'G': 4, 'H': 5, 'I': 6, 'K': 7}


def _norm_name(name):
    return re.sub(r'[-_ ()]', '', name.lower())


def parse_nwchem(text, symb=None):
    '''Parse NWChem-format text into shell lists.

    Returns the shells of ``symb``; when ``symb`` is None, or no header of the
    text names ``symb``, the shells of every header are returned (the text is
    then taken as the basis of any element).'''
    blocks = []   # (symbol, l-string, rows)
    cur = None
    for raw in text.splitlines():
        line = raw.split('#')[0].strip()
        if not line or line.upper().startswith(('BASIS', 'END')):
            continue
        fields = line.split()
        head = fields[0]
        if head[0].isalpha() and len(fields) >= 2 and fields[1].upper().isalpha():
            cur = (head, fields[1].upper(), [])
            blocks.append(cur)
            continue
        vals = [float(x.replace('D', 'E').replace('d', 'e')) for x in fields]
        cur[2].append(vals)

    def to_shells(bl):
        out = []
        for _, lstr, rows in bl:
            if lstr == 'SP':
                out.append([0] + [[r[0], r[1]] for r in rows])
                out.append([1] + [[r[0], r[2]] for r in rows])
            else:
                out.append([_L[lstr]] + [list(r) for r in rows])
        return out

    if symb is not None:
        mine = [b for b in blocks if b[0].upper() == symb.upper()]
        if mine:
            return to_shells(mine)
    return to_shells(blocks)


def uncontract(shells):
    '''One shell per distinct (l, exponent), coefficient 1.'''
    out = []
    seen = set()
    for sh in shells:
        l = sh[0]
        for prim in sh[1:]:
            key = (l, prim[0])
            if key in seen:
                continue
            seen.add(key)
            out.append([l, [prim[0], 1.]])
    return out


def _from_spec(spec, symb):
    if isinstance(spec, str):
        s = spec.strip()
        unc = False
        if s.lower().startswith('unc'):
            unc = True
            s = s[3:].lstrip('-_ \n')
        key = _norm_name(s)
        if key in basis_data.NAMED:
            shells = parse_nwchem(basis_data.NAMED[key], symb)
            if not any(True for _ in shells):
                raise KeyError(f'basis {spec} has no data for {symb}')
            if not _has_symbol(basis_data.NAMED[key], symb):
                raise KeyError(f'basis {spec} has no data for {symb}')
        elif '\n' in s or len(s.split()) > 2:
            shells = parse_nwchem(s, symb)
        else:
            raise KeyError(f'basis {spec!r} is not available')
        if unc:
            shells = uncontract(shells)
        return shells
    if isinstance(spec, dict):
        if symb in spec:
            return _from_spec(spec[symb], symb)
        if 'default' in spec:
            return _from_spec(spec['default'], symb)
        raise KeyError(f'no basis for {symb}')
    if isinstance(spec, tuple):
        out = []
        for s in spec:
            out.extend(_from_spec(s, symb))
        return out
    if isinstance(spec, list):
        if spec and isinstance(spec[0], (int,)) :
            return [spec]
        if spec and all(isinstance(x, list) and x and isinstance(x[0], int) for x in spec):
            return [list(x) for x in spec]
        out = []
        for s in spec:
            out.extend(_from_spec(s, symb))
        return out
    raise TypeError(f'cannot interpret basis {spec!r}')


def _has_symbol(text, symb):
    for raw in text.splitlines():
        f = raw.split()
        if len(f) >= 2 and f[0].upper() == symb.upper() and f[1].isalpha():
            return True
    return False


def shells_for(spec, symb):
    '''Shell list of element ``symb`` for a basis ``spec``, sorted by l.'''
    shells = _from_spec(spec, symb)
    shells = [s for s in shells if len(s) > 1]
    return sorted(shells, key=lambda s: s[0])
