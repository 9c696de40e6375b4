"""GPU: the reference's own SCF test cases (Tests/he.py, ne.py, h2.py,
ohm.py, he_mp2.py, lih_mp2.py), run through the drop-in rhf and MP2 with the
reference's defaults (6-31G**, ConvCriteria 1e-4, MaxIter 20, no DIIS), and
the HF energy of H2 in Tests/MiniSweet.py, each
against the golden energy and tolerance its file states (1e-4 where it states
none): a single atom (every shell pair on one centre), a two-centre s/p
molecule, an anion, and MP2 on top of the SCF."""
import pytest

from pyquante_amd.constants import ang2bohr  # Constants.py:20-21

pytestmark = pytest.mark.gpu



def _rhf(mol):
    from pyquante_amd.hartree_fock import rhf
    en, orbe, orbs = rhf(mol)
    return en


def test_he_rhf():
    from pyquante_amd.molecule import Molecule
    en = _rhf(Molecule("He", atomlist=[(2, (0, 0, 0))]))
    print("He %.10f" % en, flush=True)
    assert abs(en - (-2.85516040352)) < 1e-7     # Tests/he.py:16,25


def test_ne_rhf():
    from pyquante_amd.molecule import Molecule
    en = _rhf(Molecule("Ne", atomlist=[(10, (0, 0, 0))]))
    print("Ne %.10f" % en, flush=True)
    assert abs(en - (-128.474406)) < 1e-4        # Tests/ne.py:15,24


def test_h2_rhf():
    from pyquante_amd.molecule import Molecule
    en = _rhf(Molecule("h2", atomlist=[(1, (0, 0, 0)), (1, (1.0, 0, 0))]))
    print("H2 %.10f" % en, flush=True)
    assert abs(en - (-1.082098)) < 1e-6          # Tests/h2.py:15,33


def test_h2_rhf_minisweet():
    from pyquante_amd.molecule import Molecule
    en = _rhf(Molecule("H2", atomlist=[(1, (0.35, 0, 0)), (1, (-0.35, 0, 0))], units="Angs"))
    print("H2 (0.7 A) %.10f" % en, flush=True)
    assert round(en - (-1.130501), 4) == 0       # Tests/MiniSweet.py:25 (assertAlmostEqual, 4 places)


def test_ohm_rhf():
    from pyquante_amd.molecule import Molecule
    en = _rhf(Molecule("OH-", atomlist=[(8, (0, 0, 0)), (1, (0.957939 * ang2bohr, 0, 0))], charge=-1))
    print("OH- %.10f" % en, flush=True)
    assert abs(en - (-75.332606)) < 1e-4         # Tests/ohm.py:16


def _rhf_mp2(mol):
    from pyquante_amd.Ints import getbasis, getints
    from pyquante_amd.MP import MP2
    from pyquante_amd.hartree_fock import rhf
    bfs = getbasis(mol)
    nbf = len(bfs)
    nocc, nopen = mol.get_closedopen()
    assert nopen == 0
    S, h, Ints = getints(bfs, mol)
    en, orbe, orbs = rhf(mol, integrals=(S, h, Ints))
    return en + MP2(Ints, orbs, orbe, nocc, nbf - nocc)


def test_he_mp2():
    from pyquante_amd.molecule import Molecule
    e = _rhf_mp2(Molecule("He", [(2, (0.0, 0.0, 0.0))], units="Angstroms"))
    print("He MP2 %.10f" % e, flush=True)
    assert abs(e - (-2.880637)) < 1e-4           # Tests/he_mp2.py:14,28


def test_lih_mp2():
    from pyquante_amd.molecule import Molecule
    e = _rhf_mp2(Molecule("lih", [(3, (0.0, 0.0, 0.0)), (1, (0.0, 0.0, 1.629912))], units="Angstroms"))
    print("LiH MP2 %.10f" % e, flush=True)
    assert abs(e - (-8.0021911)) < 1e-4          # Tests/lih_mp2.py:15
