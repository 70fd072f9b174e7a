'''Pins the CPU oracle (and the molecule/basis layer it reads) against the
golden fingerprints held by the reference's own tests.'''
import numpy as np
import pytest

from oracle import oracle as O
from tests.molecules import case, GOLDEN_CASES


@pytest.mark.parametrize('name', GOLDEN_CASES)
def test_oracle_fingerprints(name):
    mol, dm, hermi, fpj, fpk, tol = case(name)
    vj, vk = O.get_jk(mol, dm, hermi=hermi)
    if fpj is not None:
        assert abs(O.fp(vj) - fpj) < tol
    if fpk is not None:
        assert abs(O.fp(vk) - fpk) < tol
    if hermi == 2:
        assert abs(vj).max() == 0
