"""pyquante_amd: PyQuante's two-electron-integral and Fock-build path on
AMD MI355X (gfx950), through hand-written HIP kernels behind a C ABI
(include/pq_abi.h).

Modules mirroring the reference's interface:
  Ints          getbasis/getints/get1ints/get2ints/getJ/getK/get2JmK/ijkl2intindex
  cints, crys, chgp   contr_coulomb / coulomb_repulsion / ijkl2intindex / Fgamma
  hartree_fock  rhf on the GPU engine
  force         RHF nuclear gradients (backward kernels)
"""
__version__ = "0.1.0"
